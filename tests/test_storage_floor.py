"""What a storage form costs in accuracy, shown on the CPU with the oracle (no GPU).

The oracle can round E, B and the CPML auxiliaries after every update as a narrower storage
form would hold them (float32, or float32 plus a 10-bit residual word).  On the single-pole
Lorentz + Kerr workload with an x-polarised pulse:

* float32 storage alone leaves the main components (Ex, By) near 2e-7 but the
  cross-polarised ones (Ey, Bx), three to four orders smaller, at 1e-4 to 1e-3: the float32
  rounding of the large components enters their curl every step.  That is the floor of any
  float32-storage kernel, whatever its arithmetic (the GPU's FP32-bulk form matches it;
  tests/test_gpu_fp32_bulk.py).
* float32 + 10-bit residual storage brings every component about 2^10 lower, below 1e-5.
"""
import numpy as np
import pytest

from mpulse_amd import configs
from oracle_lib import Oracle, store_round

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def test_store_round_float32_is_the_float32_cast():
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.standard_normal(200) * 10.0 ** rng.integers(-30, 30, 200),
                         [0.0, -0.0, 1.0, -2.0, 3e-39, -1e-44, 2.0 ** -126, 1.0 - 2.0 ** -30]])
    for x in xs:
        assert store_round(x, 1) == float(np.float32(x))


def test_store_round_residual_error_bound():
    """float32 + residual: the float32 word's rounding error is at most ulp/2 = 512 residual
    units (2^19 steps of the float64 bit pattern each); r is that error rounded to a unit
    (|x - stored| <= 2^-11 ulp) and clamped to 511 when it rounds to +512 (then |x - stored| <=
    2^-10 ulp), so a normal word's stored value is at least 512x closer than the float32 word
    alone; zero and denormal words keep no residual (|x - stored| <= ulp/2)."""
    rng = np.random.default_rng(6)
    xs = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-20, 20, 2000),
                         rng.uniform(-1, 1, 2000),
                         # just below and above powers of two, where the float32 ulp changes
                         [2.0 ** e * (1 - 2.0 ** -26) for e in range(-5, 6)],
                         [2.0 ** e * (1 + 2.0 ** -26) for e in range(-5, 6)],
                         [0.0, 1e-42, -3e-40]])
    worst32, worst = 0.0, 0.0
    for x in xs:
        hi = np.float32(x)
        e0 = int((hi.view(np.uint32) >> 23) & 0xFF)
        e = max(e0, 1)
        ulp = 2.0 ** (e - 150)
        r = store_round(x, 2)
        if e0 == 0:
            assert r == float(hi), (x, r)
            continue
        assert abs(r - x) <= ulp * 2.0 ** -10 * (1 + 1e-12), (x, r)
        worst32 = max(worst32, abs(float(hi) - x) / abs(x))
        worst = max(worst, abs(r - x) / abs(x))
    assert worst * 512 < worst32


def run_modes(n, steps, tshift):
    out = {}
    for mode in (0, 1, 2):
        d, _ = configs.config_roofline(precision=0, n=n)
        d.factories[1].tshift = tshift
        o = Oracle(d, omp=True)
        o.set_storage(mode)
        o.step(steps)
        out[mode] = {f: o.get(f) for f in FIELDS}
        o.close()
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def runs():
    return run_modes(48, 200, 20.0)


def test_float32_storage_misses_the_per_component_bar(runs):
    ref = runs[0]
    e32 = {f: rel(runs[1][f], ref[f]) for f in FIELDS}
    print("float32 storage:", {k: f"{v:.2e}" for k, v in e32.items()})
    # the main components are at the float32 level ...
    assert e32["Ex"] < 1e-6 and e32["By"] < 1e-6
    # ... the cross-polarised ones are not
    assert e32["Ey"] > 1e-4 and e32["Bx"] > 1e-4
    # E and B as vector fields stay within 1e-5 (what the vector-norm check saw)
    for g in "EB":
        num = sum(np.sum((runs[1][f] - ref[f]) ** 2) for f in FIELDS if f[0] == g)
        den = sum(np.sum(ref[f] ** 2) for f in FIELDS if f[0] == g)
        assert (num / den) ** 0.5 < 1e-6


def test_residual_storage_meets_the_per_component_bar(runs):
    ref = runs[0]
    e32 = {f: rel(runs[1][f], ref[f]) for f in FIELDS}
    ec = {f: rel(runs[2][f], ref[f]) for f in FIELDS}
    print("float32 + residual storage:", {k: f"{v:.2e}" for k, v in ec.items()})
    for f in FIELDS:
        assert ec[f] <= 1e-5, (f, ec[f])
        # the error is the storage rounding: it scales with the word's precision
        assert 500 < e32[f] / ec[f] < 2000, (f, e32[f], ec[f])


def test_library_store_round_matches_the_oracle():
    """The library's host encoder (what mpx_set_field stores) holds the same value as the
    oracle's storage rounding, bit for bit, in both narrow forms (the device kernels share the
    encoder)."""
    from mpulse_amd import abi
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.standard_normal(3000) * 10.0 ** rng.integers(-30, 30, 3000),
                         [2.0 ** e * (1 + s * 2.0 ** -25 * f) for e in range(-8, 8) for s in (-1, 1)
                          for f in (0.9, 1.0, 1.1)],
                         [0.0, -0.0, 1e-42, -3e-40, 3.4e38, -3.4e38]])
    for x in xs:
        assert abi.store_round(x, abi.PREC_FP64) == x
        assert abi.store_round(x, abi.PREC_FP32_BULK) == store_round(x, 1)
        assert abi.store_round(x, abi.PREC_FP32_RES) == store_round(x, 2), x


def test_cpml_auxiliaries_need_the_residual_too():
    """E and B with residual words but the CPML auxiliaries in plain float32 (oracle mode 3)
    leave the cross-polarised components above 1e-5 at 96^3 after 300 steps; with residual
    auxiliaries (mode 2) they are three orders lower.  Run on 48^3 here: the same ordering."""
    out = {}
    for mode in (0, 2, 3):
        d, _ = configs.config_roofline(precision=0, n=48)
        d.factories[1].tshift = 20.0
        o = Oracle(d, omp=True)
        o.set_storage(mode)
        o.step(200)
        out[mode] = {f: o.get(f) for f in ("Ey", "Bx")}
        o.close()
    for f in ("Ey", "Bx"):
        e2, e3 = rel(out[2][f], out[0][f]), rel(out[3][f], out[0][f])
        print(f, f"residual aux {e2:.2e}, float32 aux {e3:.2e}")
        assert e3 > 20 * e2
