"""FP32-bulk form against the float64 form on the same workload (GPU).

Bar: after 1000 steps of the 512^3 single-pole Lorentz + Kerr workload (16-cell CPML,
20 fs pulse at 5e9 V/m), the relative L2 error of the electric and of the magnetic field
(all three components together) against float64 is <= 1e-5.  The error of every component
is printed as well.  The weak components (E_y, E_z, B_x, B_z of an x-polarised pulse, three
to four orders below the main ones) carry float32 rounding of the main components coupled in
through the curl, so their own relative errors are larger; they are bounded at 1e-3.
Smaller grids are checked to the same bars in a short run first.
"""
import os

import numpy as np
import pytest

from mpulse_amd import abi, configs

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def compare(n, steps, tshift=None):
    print(f"case: fp32-bulk vs fp64, {n}^3, {steps} steps", flush=True)
    out = {}
    num = {"E": 0.0, "B": 0.0}
    den = {"E": 0.0, "B": 0.0}
    d64, meta = configs.config_roofline(precision=abi.PREC_FP64, n=n)
    d32, _ = configs.config_roofline(precision=abi.PREC_FP32_BULK, n=n)
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


def check(errs):
    assert errs["E"] <= 1e-5 and errs["B"] <= 1e-5, errs
    assert max(errs[f] for f in FIELDS) <= 1e-3, errs


def test_fp32_bulk_small():
    check(compare(96, 300, tshift=30.0))


@pytest.mark.skipif(os.environ.get("MPX_SKIP_LARGE") == "1", reason="large case disabled by MPX_SKIP_LARGE=1")
def test_fp32_bulk_512_1000_steps():
    check(compare(512, 1000))


def test_fp32_bulk_focused_two_poles():
    """The two-pole (silica-like) medium with a focused pulse: FP32-bulk against float64."""
    out = {}
    num = {"E": 0.0, "B": 0.0}
    den = {"E": 0.0, "B": 0.0}
    runs = []
    for prec in (abi.PREC_FP64, abi.PREC_FP32_BULK):
        d, _ = configs.config_focus(precision=prec, nx=64, nz=96, steps=300, d_cpml=8, d_src=12, na=0.3)
        d.factories[1].tshift = 20.0
        d.factories[1].length = 8.0
        runs.append(d)
    with abi.Solver(runs[0]) as a, abi.Solver(runs[1]) as b:
        a.step(300)
        b.step(300)
        for f in FIELDS:
            x, y = b.get(f), a.get(f)
            out[f] = rel_l2(x, y)
            num[f[0]] += float(np.sum((x - y) ** 2))
            den[f[0]] += float(np.sum(y ** 2))
    for g in ("E", "B"):
        out[g] = (num[g] / den[g]) ** 0.5
    print({k: f"{v:.2e}" for k, v in out.items()}, flush=True)
    check(out)

