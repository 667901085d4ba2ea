"""Batched independent runs (mpx_batch_*): every run of a batch must evolve exactly as the
same descriptor does on its own (sweep path), whatever the tiling, graph replays and
remainder steps, and the autotuner must leave the batch in its initial state."""
import numpy as np
import pytest

from mpulse_amd import abi, configs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "PsiR1x", "PsiI1y")


def small_runs(precision):
    """Three uneven periodic grids with CPML and a pulse (partial tiles in every dimension)."""
    out = []
    for (nx, ny, nz), chi3, peak in (((37, 33, 40), 2e-22, 3e9), ((24, 24, 24), 1e-21, 5e9),
                                     ((45, 30, 70), 5e-23, 1e9)):
        d, _ = configs.plrc_kerr_pulse(nx, 6, 6.0, 5, peak, 0, chi3_si=chi3, d_src=9, precision=precision,
                                       path=abi.PATH_SWEEP, nz=nz)
        d.ny = ny
        d.factories[1].tshift = 12.0
        out.append(d)
    return out


def singles(descs, steps, path=abi.PATH_SWEEP):
    res = []
    for d in descs:
        d.path = path
        with abi.Solver(d) as s:
            s.step(steps)
            res.append({n: s.get(n) for n in FIELDS})
        d.path = abi.PATH_SWEEP
    return res


# (structure, tuning, graph steps): the fused structure is compared with the fused path run
# on its own, the sweep structure with the sweep path, the x-march structure with the pass path
CASES = [(0, {"march_b": 4, "fused_blocks_per_cu": 8}, 7), (0, {"march_b": 1, "fused_blocks_per_cu": 1}, 0),
         (1, {"variant": 2, "seg": 32}, 7), (1, {"variant": 0, "seg": 5}, 4), (1, {"variant": 1, "seg": 16}, 0),
         (2, {"pass_seg": 16, "pass_rows": 7}, 7), (2, {"pass_seg": 5, "pass_rows": 3}, 4),
         (2, {"pass_seg": 32, "pass_rows": 0}, 0)]
SINGLE_PATH = {0: abi.PATH_FUSED, 1: abi.PATH_SWEEP, 2: abi.PATH_PASS}
STRUCTURE = {"fused": 0, "sweep": 1, "pass": 2}


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_BULK])
@pytest.mark.parametrize("structure,tuning,graph", CASES)
def test_batch_matches_single_runs(prec, structure, tuning, graph):
    descs = small_runs(prec)
    steps = 45
    print(f"case: batch prec {prec} structure {structure} {tuning} graph {graph}", flush=True)
    want = singles(descs, steps, SINGLE_PATH[structure])
    with abi.Batch(descs) as b:
        b.tune("structure", structure)
        for k, v in tuning.items():
            b.tune(k, v)
        b.tune("graph_steps", graph)
        b.step(1)  # the x-march structure: a call of one step takes the two fused halves only
        b.step(19)
        b.step(steps - 20)
        assert b.time == steps
        for r in range(len(descs)):
            v = b.run(r)
            assert v.time == steps
            for n in FIELDS:
                got = v.get(n)
                assert np.array_equal(got, want[r][n]), (r, n, np.abs(got - want[r][n]).max())


def test_batch_run_against_oracle():
    descs = small_runs(abi.PREC_FP64)
    steps = 40
    for structure in (0, 1, 2):
        with abi.Batch(descs) as b:
            b.tune("structure", structure)
            b.tune("graph_steps", 10)
            b.step(steps)
            got = {n: b.run(2).get(n) for n in FIELDS[:6]}
        o = Oracle(descs[2])
        o.step(steps)
        for n in FIELDS[:6]:
            r = o.get(n)
            err = np.linalg.norm(got[n] - r) / np.linalg.norm(r)
            print(structure, n, err)
            assert err < 1e-12, (structure, n, err)


def test_batch_autotune_restores_initial_state():
    descs = small_runs(abi.PREC_FP64)
    with abi.Batch(descs) as b:
        best = b.autotune(steps_per_trial=2)
        print("autotune", best)
        assert best["structure"] in STRUCTURE
        want = singles(descs, 30, SINGLE_PATH[STRUCTURE[best["structure"]]])
        assert b.time == 0
        assert not np.any(b.run(0).get("Ex"))
        b.step(30)
        for r in range(len(descs)):
            for n in FIELDS:
                assert np.array_equal(b.run(r).get(n), want[r][n]), (r, n)


def test_batch_handles_refuse_direct_stepping():
    descs = small_runs(abi.PREC_FP64)[:1]
    with abi.Batch(descs) as b:
        v = b.run(0)
        with pytest.raises(abi.MpxError, match="belongs to a batch"):
            v.step(1)
        with pytest.raises(abi.MpxError, match="autotune runs before the first step"):
            b.step(1)
            b.autotune()


def test_batch_rejects_mixed_runs():
    a, _ = configs.plrc_kerr_pulse(24, 6, 6.0, 5, 1e9, 0, d_src=9, path=abi.PATH_SWEEP)
    b_, _ = configs.plrc_kerr_pulse(24, 6, 6.0, 5, 1e9, 0, d_src=9, precision=abi.PREC_FP32_BULK,
                                    path=abi.PATH_SWEEP)
    with pytest.raises(abi.MpxError, match="must share precision"):
        abi.Batch([a, b_])
