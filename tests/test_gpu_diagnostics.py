"""Per-step diagnostics on the GPU: mpx_sampler (the per-step grid reads of FrequencyDiag and
FollowMax) against mpx_get_field, and the host layer's FrequencyDiag (src/freqdiag.cpp:5-95) and
FollowMax (src/followmax.cpp:5-125) through the host driver.

The expected outputs come from the reference's calculate()/write() steps, restated below and
run over the grids mpx_get_field returns after every step, line for line in the reference's
text format.  (The zero-crossing rate at a fixed point is not a clean pin of the carrier: a
weak precursor ahead of the pulse, the Gouy phase of the diffracting beam and the medium's
dispersion all shift the crossings.)"""
import os

import numpy as np
import pytest

from mpulse_amd import abi, configs

pytestmark = pytest.mark.gpu


def small(prec, path, boundary):
    d, _ = configs.plrc_kerr_pulse(20, 4, 10.0, 6, 1e9, 30, d_src=7, nz=26, precision=prec, path=path,
                                   boundary=boundary)
    d.factories[1].tshift = 12.0
    return d


@pytest.mark.parametrize("prec,path,boundary", [
    (abi.PREC_FP64, abi.PATH_FUSED, abi.BOUNDARY_PERIODIC),
    (abi.PREC_FP32_BULK, abi.PATH_FUSED, abi.BOUNDARY_XY_PERIODIC),
    (abi.PREC_FP64, abi.PATH_SWEEP, abi.BOUNDARY_PERIODIC),
    (abi.PREC_FP64, abi.PATH_REFERENCE, abi.BOUNDARY_XY_PERIODIC)])
def test_sampler_records_equal_the_grid(prec, path, boundary):
    """Cells and lines through ghost cells of every dimension (the periodic images on the paths
    that leave the ghosts stale), a grid read on the host (PsiR1y), records in step order across
    drains, and the full-sampler error."""
    print(f"case: sampler prec {prec} path {path} boundary {boundary}", flush=True)
    with abi.Solver(small(prec, path, boundary)) as s:
        nx, ny, nz = 20, 20, 26
        specs = [("Ex", -1, (0, ny + 1, 3)), ("By", -1, (7, 8, 13)), ("Ez", 0, (0, 0, nz + 1)),
                 ("Bx", 1, (nx + 1, 0, 5)), ("Ey", 2, (5, 6, 0)), ("Bz", 2, (0, ny + 1, 0)),
                 ("PsiR1y", 2, (4, 5, 0))]
        samplers = [s.sampler(name, dim, pos, capacity=4) for name, dim, pos in specs]
        try:
            expect = [[] for _ in specs]
            got = [[] for _ in specs]
            for n in range(10):
                s.step(3)
                for q, (name, dim, pos) in enumerate(specs):
                    samplers[q].record()
                    g = s.get(name)
                    idx = list(pos)
                    if dim < 0:
                        expect[q].append([g[tuple(idx)]])
                    else:
                        idx[dim] = slice(None)
                        expect[q].append(g[tuple(idx)])
                if n % 3 == 2:
                    for q in range(len(specs)):
                        got[q].extend(samplers[q].drain())
            with pytest.raises(abi.MpxError):
                for _ in range(5):
                    samplers[0].record()
            for q in range(len(specs)):
                got[q].extend(samplers[q].drain()[:1])
                assert np.array_equal(np.array(got[q]), np.array(expect[q])), specs[q]
            with pytest.raises(abi.MpxError):
                s.sampler("Ex", -1, (0, 0, nz + 2), 4)
            with pytest.raises(abi.MpxError):
                s.sampler("NoSuchGrid", -1, (1, 1, 1), 4)
        finally:
            for p in samplers:
                p.close()


def frequency_reference(values, times, dt, interval, append=True):
    """FrequencyDiag::calculate / write (src/freqdiag.cpp:20-70) and DiagnosticInterface::execute
    (src/diagnostic.cpp:12-31): the lines written, in order."""
    count, lastcount, lastval, firstzero, frequency = -1, 0, 0.0, 0.0, None
    lines = []
    for t, (tm, v) in enumerate(zip(times, values)):
        val = 0.15 * v
        if val * lastval <= 0 and abs(val) * abs(lastval) > 0:
            d = lastval / (lastval - val)
            time = (tm - 1 + d) * dt
            if count < 0:
                firstzero, count = time, 0
            else:
                count += 1
                frequency = (0.5 * count) / (time - firstzero)
        lastval = val
        if t % interval == 0 and count > lastcount:
            lastcount = count
            lines.append("%d %s" % (tm, "%g" % frequency))
    return lines


def followmax_reference(lines, times, dim_pos, start, interval, n):
    """FollowMax::init / calculate / write (src/followmax.cpp:20-113) on storage lines 0..n+1."""
    low, high = 0, n + 1
    ipos = dim_pos
    position = float(ipos)
    out = []
    for t, (tm, line) in enumerate(zip(times, lines)):
        if tm >= start:
            while True:
                pm, pp = ipos - 1, ipos + 1
                if pm <= low:
                    pm = high - 1
                if pp >= high:
                    pp = low + 1
                mid, left, right = line[ipos], line[pm], line[pp]
                if right > mid:
                    ipos = pp
                elif left > mid:
                    ipos = pm
                else:
                    break
            position = ipos + (left - right) / (2 * (left - 2 * mid + right))
        if t % interval == 0 and tm >= start:
            out.append("%d %s" % (tm, "%.10g" % position))
    return out


SETUP = """
globals { grid-x 12 grid-y 14 grid-z 64 grid-dx 0.3141592653589793 grid-dy 0.3141592653589793
          grid-dz 0.3141592653589793 dt 0.09068996821171089 T-total 420 }
fieldsim {
  fdtd-plrc-nonlin { eps 1 chi 2e-4 L1eps 1.1 L1Om2 4 L1delta 0.005
                     cpml_border { d 4 sigmaMax 8.795 }
                     short_pulse_inject { d 8 length 20 width 4 om0 1 amp 1 eps 2.4666666666666668 TShift 30 } }
  single_xy_periodic
  frequency { field Ex file DIR/freq.dat x 6 y 7 z 40 interval 10 append y }
  frequency { field By file DIR/freq_#t.dat x 5 y 6 z 30 interval 150 }
  followmax { field Ex file DIR/fm.dat dim z posx 6 posy 7 posz 30 start 200 interval 7 append y }
  followmax { field Ey file DIR/fmx.dat dim x posx 3 posy 5 posz 36 start 150 interval 25 append y }
  frequency { field NoSuchGrid file DIR/none.dat interval 1 append y }
}
"""


def test_frequency_and_followmax_diagnostics(tmp_path):
    print("case: frequency / followmax through the host driver", flush=True)
    text = SETUP.replace("DIR", str(tmp_path))
    sim = abi.HostSimulation(text)
    try:
        sim.run(-1)
    finally:
        sim.close()
    d, total = abi.parse_setup(text)
    times = list(range(total))
    pts = {"Ex": [], "By": []}
    zline, xline = [], []
    with abi.Solver(d) as s:
        for _ in times:
            s.step(1)
            ex, by, ey = s.get("Ex"), s.get("By"), s.get("Ey")
            pts["Ex"].append(ex[6, 7, 40])
            pts["By"].append(by[5, 6, 30])
            zline.append(ex[6, 7, :].copy())
            xline.append(ey[:, 5, 36].copy())
    read = lambda name: open(tmp_path / name).read().splitlines()
    freq = read("freq.dat")
    assert freq == frequency_reference(pts["Ex"], times, d.dt, 10)
    assert len(freq) > 5
    # append n: every output reopens freq_#t.dat; a file exists for each output step
    names = sorted(os.listdir(tmp_path))
    assert names == sorted(["freq.dat", "fm.dat", "fmx.dat", "none.dat", "freq_0.dat", "freq_150.dat",
                            "freq_300.dat"])
    by_lines = frequency_reference(pts["By"], times, d.dt, 150)
    for tm in (0, 150, 300):
        want = [l for l in by_lines if int(l.split()[0]) == tm]
        assert read(f"freq_{tm}.dat") == want, tm
    assert read("none.dat") == []
    fm = read("fm.dat")
    assert fm == followmax_reference(zline, times, 30, 200, 7, 64)
    assert len(fm) == len([t for t in times if t % 7 == 0 and t >= 200])
    # the vertex of the parabola through a maximum and its neighbours lies within half a cell
    assert all(0.5 <= float(l.split()[1]) <= 64.5 for l in fm)
    assert read("fmx.dat") == followmax_reference(xline, times, 3, 150, 25, 12)
