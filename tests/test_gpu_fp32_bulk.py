"""The float32 storage forms against float64 on the same workload (GPU).

Bar (north star): after 1000 steps of the 512^3 single-pole Lorentz + Kerr workload (16-cell
CPML, 20 fs pulse at 5e9 V/m), the relative L2 error of every field component against
float64 is <= 1e-5.

* FP32-res (E, B and the CPML auxiliaries as float32 words each backed by a 10-bit residual
  slot; the PLRC accumulators and all arithmetic, the Kerr solve included, in float64) meets
  it for every component.
* FP32-bulk (plain float32 words) does not for the cross-polarised components Ey and Bx of
  the x-polarised pulse: their error is the float32 storage floor -- the oracle, with every
  stored word rounded to float32 (tests/test_storage_floor.py), gives the same per-component
  errors as the GPU -- so no float32-storage kernel can meet the bar.  That test asserts the
  agreement with the floor, and the per-component bar for FP32-bulk is an expected failure.
"""
import os

import numpy as np
import pytest

from mpulse_amd import abi, configs

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def compare(n, steps, prec, tshift=None, path=abi.PATH_FUSED):
    print(f"case: precision {prec} vs fp64, {n}^3, {steps} steps", flush=True)
    out = {}
    num = {"E": 0.0, "B": 0.0}
    den = {"E": 0.0, "B": 0.0}
    d64, _ = configs.config_roofline(precision=abi.PREC_FP64, n=n)
    d32, _ = configs.config_roofline(precision=prec, n=n, path=path)
    if tshift is not None:
        d64.factories[1].tshift = d32.factories[1].tshift = tshift
    with abi.Solver(d64) as a, abi.Solver(d32) as b:
        a.step(steps)
        b.step(steps)
        for f in FIELDS:
            x, y = b.get(f), a.get(f)
            out[f] = rel_l2(x, y)
            num[f[0]] += float(np.sum((x - y) ** 2))
            den[f[0]] += float(np.sum(y ** 2))
            del x, y
    for g in ("E", "B"):
        out[g] = (num[g] / den[g]) ** 0.5
    print({k: f"{v:.2e}" for k, v in out.items()}, flush=True)
    return out


def check_every_field(errs, bar=1e-5):
    bad = {f: errs[f] for f in FIELDS if not errs[f] <= bar}
    assert not bad, (bad, errs)


def test_fp32_res_small():
    check_every_field(compare(96, 300, abi.PREC_FP32_RES, tshift=30.0))


@pytest.mark.skipif(os.environ.get("MPX_SKIP_LARGE") == "1", reason="large case disabled by MPX_SKIP_LARGE=1")
def test_fp32_res_512_1000_steps():
    check_every_field(compare(512, 1000, abi.PREC_FP32_RES))


def test_fp32_bulk_is_at_the_float32_storage_floor():
    """Per component, the GPU FP32-bulk error equals (within a factor 1.5) the error of the
    oracle with every stored word rounded to float32 -- the errors are the storage rounding,
    not the kernels' arithmetic."""
    from oracle_lib import Oracle
    gpu = compare(96, 300, abi.PREC_FP32_BULK, tshift=30.0)
    ref = {}
    for mode in (0, 1):
        d, _ = configs.config_roofline(precision=0, n=96)
        d.factories[1].tshift = 30.0
        o = Oracle(d, omp=True)
        o.set_storage(mode)
        o.step(300)
        ref[mode] = {f: o.get(f) for f in FIELDS}
        o.close()
    floor = {f: rel_l2(ref[1][f], ref[0][f]) for f in FIELDS}
    print("oracle float32-storage floor:", {k: f"{v:.2e}" for k, v in floor.items()}, flush=True)
    for f in FIELDS:
        assert floor[f] / 1.5 <= gpu[f] <= floor[f] * 1.5, (f, gpu[f], floor[f])
    # the vector fields stay within the bar; the cross-polarised components do not
    assert gpu["E"] <= 1e-5 and gpu["B"] <= 1e-5
    assert gpu["Ey"] > 1e-5 and gpu["Bx"] > 1e-5


@pytest.mark.xfail(strict=True, reason="float32 storage floor: Ey 5.8e-5 and Bx 6.8e-5 at 96^3 / 300 steps "
                                       "(4.4e-4 and 1.2e-4 at 512^3 / 1000 steps); FP32-res meets the bar")
def test_fp32_bulk_per_field_bar():
    check_every_field(compare(96, 300, abi.PREC_FP32_BULK, tshift=30.0))


def test_fp32_res_focused_two_poles():
    """The two-pole (silica-like) medium with a focused pulse: FP32-res against float64,
    every component within 1e-5."""
    out = {}
    runs = []
    for prec in (abi.PREC_FP64, abi.PREC_FP32_RES):
        d, _ = configs.config_focus(precision=prec, nx=64, nz=96, steps=300, d_cpml=8, d_src=12, na=0.3)
        d.factories[1].tshift = 20.0
        d.factories[1].length = 8.0
        runs.append(d)
    with abi.Solver(runs[0]) as a, abi.Solver(runs[1]) as b:
        a.step(300)
        b.step(300)
        for f in FIELDS:
            out[f] = rel_l2(b.get(f), a.get(f))
    print({k: f"{v:.2e}" for k, v in out.items()}, flush=True)
    check_every_field(out)


def test_fp32_res_field_access():
    """get/set/fetch/layers/energy/probes of the residual form: a set value reads back as the
    host encoder stores it; the device decoders (fetch, energy, probes) agree with get."""
    d, meta = configs.config_roofline(precision=abi.PREC_FP32_RES, n=40)
    d.factories[1].tshift = 10.0
    rng = np.random.default_rng(3)
    with abi.Solver(d) as s:
        shape = s.shape
        x = rng.standard_normal(shape) * 1e3
        s.set("Ey", x)
        got = s.get("Ey")
        want = np.vectorize(lambda v: abi.store_round(v, abi.PREC_FP32_RES))(x)
        assert np.array_equal(got[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1])
        inner = (slice(1, -1),) * 3
        assert np.max(np.abs(got[inner] - x[inner]) / np.abs(x[inner])) < 2.0 ** -33
        # the other components keep their (zero) values: a set touches only its own slot
        assert not np.any(s.get("Ex")) and not np.any(s.get("Ez"))
        s.set_probes(meta["probes"][:4], [1, 1, 4, 0], 16)
        s.step(12)
        E = {f: s.get(f) for f in FIELDS}
        assert np.array_equal(s.fetch("Bz"), E["Bz"].transpose(2, 1, 0))
        e = sum(float(np.sum(E[f][1:-1, 1:-1, 1:-1] ** 2)) for f in FIELDS) * d.dx * d.dy * d.dz
        assert abs(s.energy() - e) <= 1e-12 * e
        pr = s.get_probes()
        c = meta["probes"][:4]
        assert pr[-1, 0] == E["Ey"][tuple(c[0])] and pr[-1, 2] == E["By"][tuple(c[2])]
        lay, lo, hi = s.get_layer("CPMLPsiEx", abi.NORTH)
        assert np.any(lay) and np.all(np.isfinite(lay))
        # the accumulators stay float64: a set value reads back exactly and leaves the other
        # five values of the cells alone
        names = [f"Psi{p}1{c}" for p in "RI" for c in "xyz"]
        before = {q: s.get(q) for q in names}
        assert any(np.any(v) for v in before.values())
        y = rng.standard_normal(shape) * 1e2
        s.set("PsiI1y", y)
        assert np.array_equal(s.get("PsiI1y"), y)
        for q in names:
            if q != "PsiI1y":
                assert np.array_equal(s.get(q), before[q]), q
