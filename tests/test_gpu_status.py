"""Device status of the E updates (GPU): MPX_ERR_NUMERIC from mpx_sync when a Kerr Newton
solve stopped at its iteration cap or an E value is non-finite since the last check.

The reference iterates the Newton solve until |E - Eold| <= 1e-9 with no bound
(src/fdtd_plrc.cpp:406-425) and goes on silently past a NaN (with Kerr it stores E = 0 where
the update's C is NaN, and the accumulators keep the NaN); the kernels stop the solve at a cap
(default 200), compute what the reference computes otherwise, and say so.  The conditions are forced from the host -- a cap of 1 iteration
(tuning key newton_cap), a NaN written into the state -- so every kernel still runs to its
end normally.
"""
import numpy as np
import pytest

from mpulse_amd import abi, configs

pytestmark = pytest.mark.gpu


def kerr(path, prec=abi.PREC_FP64):
    d, _ = configs.plrc_kerr_pulse(32, 6, 5.0, 6, 3e9, 60, d_src=9, precision=prec, path=path, nz=40)
    d.factories[1].tshift = 6.0
    return d


@pytest.mark.parametrize("path", [abi.PATH_FUSED, abi.PATH_REFERENCE, abi.PATH_PASS, abi.PATH_SWEEP])
def test_newton_cap_is_reported(path):
    print(f"case: newton cap, path {path}", flush=True)
    with abi.Solver(kerr(path)) as s:
        s.step(30)
        s.sync()  # the default cap is never reached
        s.tune("newton_cap", 1)
        s.step(5)
        with pytest.raises(abi.MpxError) as e:
            s.sync()
        assert e.value.code == abi.ERR_NUMERIC and "iteration cap (1)" in str(e.value)
        s.sync()  # the status was cleared by the report
        s.tune("newton_cap", 200)
        s.step(5)
        s.sync()


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_BULK, abi.PREC_FP32_RES])
def test_non_finite_e_is_reported(prec):
    print(f"case: NaN, precision {prec}", flush=True)
    with abi.Solver(kerr(abi.PATH_FUSED, prec)) as s:
        s.step(3)
        ex = s.get("Ex")
        ex[10, 11, 12] = np.nan
        s.set("Ex", ex)
        s.step(2)
        with pytest.raises(abi.MpxError) as e:
            s.sync()
        assert e.value.code == abi.ERR_NUMERIC and "non-finite" in str(e.value)
        # as in the reference: E is zeroed where C is NaN, the accumulator keeps the NaN
        assert s.get("Ex")[10, 11, 12] == 0.0 and np.isnan(s.get("PsiR1x")[10, 11, 12])


def test_fdtd_nl_nan_at_zero_field_is_reported():
    """fdtd-nl divides by |D| (src/fdtd_nl.cpp:86-89): from an all-zero state every E becomes
    NaN, which the reference leaves unnoticed and the kernels report."""
    d = abi.desc_defaults()
    d.nx = d.ny = d.nz = 12
    d.solver = abi.SOLVER_NL
    with abi.Solver(d) as s:
        s.step(1)
        with pytest.raises(abi.MpxError) as e:
            s.sync()
        assert e.value.code == abi.ERR_NUMERIC
        assert np.isnan(s.get("Ex")[3, 3, 3])
