"""Field output on the GPU: mpx_fetch_field (the reference's Grid order, x fastest, float64)
and the host layer's FieldDiag writing the reference's HDF5 form (src/fielddiag.cpp:9-24,
src/diagnostic.cpp:12-72, src/hdfstream.t:88-187)."""
import os

import numpy as np
import pytest

from hdf5_read import read_file
from mpulse_amd import abi, configs

pytestmark = pytest.mark.gpu


def small(prec, path=abi.PATH_FUSED):
    d, _ = configs.plrc_kerr_pulse(24, 5, 10.0, 6, 1e9, 40, d_src=7, nz=30, precision=prec, path=path)
    d.factories[1].tshift = 12.0
    return d


@pytest.mark.parametrize("prec,path", [(abi.PREC_FP64, abi.PATH_FUSED), (abi.PREC_FP32_BULK, abi.PATH_FUSED),
                                       (abi.PREC_FP64, abi.PATH_REFERENCE)])
def test_fetch_is_the_transposed_grid(prec, path):
    print(f"case: fetch prec {prec} path {path}", flush=True)
    with abi.Solver(small(prec, path)) as s:
        s.step(25)
        for name in ("Ex", "Bz", "PsiR1y", "Eps"):
            got = s.fetch(name)
            ref = s.get(name)
            assert got.shape == ref.shape[::-1]
            assert np.array_equal(got, ref.transpose(2, 1, 0)), name


def test_fetch_overlaps_later_steps():
    """A copy queued after 10 steps holds the state of step 10 although 10 more steps are
    queued before the caller waits for it (page-locked destination)."""
    print("case: async fetch", flush=True)
    d = small(abi.PREC_FP64)
    with abi.Solver(d) as a, abi.PinnedArray((32, 26, 26)) as buf:
        a.step(10)
        h = a.fetch_async("Ey", buf)
        a.step(10)
        h.wait()
        got = buf.array.copy()
        later = a.fetch("Ey")
        # a fetch still in flight when the buffer is closed: close() waits for it
        h2 = a.fetch_async("Ex", buf)
        buf.close()
        assert h2.h is None
    with abi.Solver(d) as b:
        b.step(10)
        ref = b.get("Ey").transpose(2, 1, 0)
    assert np.array_equal(got, ref)
    assert not np.array_equal(later, ref)


SETUP = """
globals { grid-x 20 grid-y 22 grid-z 26 grid-dx 0.3141592653589793 grid-dy 0.3141592653589793
          grid-dz 0.3141592653589793 dt 0.09068996821171089 T-total 12 }
fieldsim {
  fdtd-plrc-nonlin { eps 1 chi 2e-4 L1eps 1.1 L1Om2 4 L1delta 0.005
                     cpml_border { d 5 sigmaMax 8.795 }
                     short_pulse_inject { d 7 length 5 width 1.5 om0 1 amp 1 eps 2.4666666666666668 TShift 3 } }
  single_periodic
  fielddiag { field Ex file DIR/ex_#t.h5 interval 5 }
  fielddiag { field By file DIR/by_#p.h5 interval 4 append y }
}
"""


def test_fielddiag_writes_the_reference_hdf5(tmp_path):
    """Outputs after the steps of cycles 0, 5, 10 (Ex, one file each, #t = cycle) and 0, 4, 8
    (By, appended as data0..data2 to one file, #p = 0); each dataset holds the storage box,
    HDF5 dims (Nz+2, Ny+2, Nx+2), x fastest, equal to the grid after that many steps."""
    print("case: fielddiag through the host driver", flush=True)
    text = SETUP.replace("DIR", str(tmp_path))
    sim = abi.HostSimulation(text)
    try:
        sim.run(-1)
    finally:
        sim.close()
    d, _ = abi.parse_setup(text)
    expect = {}
    with abi.Solver(d) as s:
        for n in range(1, 12):
            s.step(1)
            expect[n] = {"Ex": s.get("Ex").transpose(2, 1, 0), "By": s.get("By").transpose(2, 1, 0)}
    for cycle in (0, 5, 10):
        f = read_file(os.path.join(tmp_path, f"ex_{cycle}.h5"))
        assert f.names == ["data0"]
        a = f.read("data0")
        assert a.shape == (28, 24, 22)
        assert np.array_equal(a, expect[cycle + 1]["Ex"]), cycle
    assert sorted(os.listdir(tmp_path)) == ["by_0.h5", "ex_0.h5", "ex_10.h5", "ex_5.h5"]
    f = read_file(os.path.join(tmp_path, "by_0.h5"))
    assert f.names == ["data0", "data1", "data2"]
    for q, cycle in enumerate((0, 4, 8)):
        assert np.array_equal(f.read(f"data{q}"), expect[cycle + 1]["By"]), cycle


SETUP2 = """
globals { grid-x 12 grid-y 14 grid-z 16 grid-dx 0.3141592653589793 grid-dy 0.3141592653589793
          grid-dz 0.3141592653589793 dt 0.09068996821171089 T-total 8 }
fieldsim {
  fdtd-plrc-nonlin { eps 1 chi 2e-4 L1eps 1.1 L1Om2 4 L1delta 0.005
                     cpml_border { d 3 sigmaMax 8.795 }
                     short_pulse_inject { d 5 length 4 width 1.5 om0 1 amp 1 eps 2.4666666666666668 TShift 2 } }
  single_periodic
  energy { file DIR/energy.dat interval 3 append y }
  linediag { field Ex file DIR/line_#t.dat dim z posx 6 posy 7 interval 4 }
  linediag { field By file DIR/gnu.dat dim x posy 7 posz 6 format gnu interval 7 append y }
  linediag { field Ez file DIR/never.dat dim y posx 0 posz 3 interval 1 }
  energy { file DIR/missing_dir/e.dat interval 2 }
  energy { interval 2 }
  slicediag { field Ex file DIR/slice_#t.h5 plane y pos 7 interval 4 }
  slicediag { field Bz file DIR/slices.h5 plane z pos 6 interval 3 append y }
}
"""


def test_energy_line_and_slice_diagnostics(tmp_path):
    """energy (src/fielddiag.cpp:219-286), linediag (:122-217) and slicediag (:26-118) through the
    host driver: files per output or appended, '#t' names, inactive lines write nothing, a text
    output that cannot be opened (no "file", a missing directory) writes nothing and the run goes on;
    contents equal the grids after those steps (cycles 0, 3, 6 / 0, 4 / 0, 7)."""
    print("case: energy / line / slice diagnostics", flush=True)
    text = SETUP2.replace("DIR", str(tmp_path))
    sim = abi.HostSimulation(text)
    try:
        sim.run(-1)
    finally:
        sim.close()
    d, _ = abi.parse_setup(text)
    grids, energy = {}, {}
    with abi.Solver(d) as s:
        for n in range(1, 9):
            s.step(1)
            grids[n] = {f: s.get(f) for f in ("Ex", "By", "Bz")}
            energy[n] = s.energy()
    assert sorted(os.listdir(tmp_path)) == ["energy.dat", "gnu.dat", "line_0.dat", "line_4.dat", "slice_0.h5",
                                            "slice_4.h5", "slices.h5"]
    rows = [l.split() for l in open(tmp_path / "energy.dat").read().splitlines()]
    assert [int(r[0]) for r in rows] == [0, 3, 6]
    for r in rows:
        e = energy[int(r[0]) + 1]
        assert abs(float(r[1]) - e) <= 1e-5 * abs(e)  # six significant digits, as the stream prints
    for cycle in (0, 4):
        vals = np.array([float(v) for v in open(tmp_path / f"line_{cycle}.dat").read().split()])
        ref = grids[cycle + 1]["Ex"][6, 7, :]
        assert vals.shape == (18,) and np.allclose(vals, ref, rtol=1e-9, atol=1e-14 * np.abs(ref).max())
    blocks = open(tmp_path / "gnu.dat").read().strip().split("\n\n")
    assert len(blocks) == 2
    for blk, cycle in zip(blocks, (0, 7)):
        pts = np.array([[float(x) for x in l.split()] for l in blk.splitlines()])
        assert list(pts[:, 0]) == list(range(14))
        ref = grids[cycle + 1]["By"][:, 7, 6]
        assert np.allclose(pts[:, 1], ref, rtol=1e-9, atol=1e-14 * np.abs(ref).max())
    for cycle in (0, 4):
        f = read_file(os.path.join(tmp_path, f"slice_{cycle}.h5"))
        a = f.read("data0")
        assert a.shape == (18, 14)  # (z, x) of the y = 7 plane, x fastest
        assert np.array_equal(a, grids[cycle + 1]["Ex"][:, 7, :].T)
    f = read_file(os.path.join(tmp_path, "slices.h5"))
    assert f.names == ["data0", "data1", "data2"]
    for q, cycle in enumerate((0, 3, 6)):
        assert np.array_equal(f.read(f"data{q}"), grids[cycle + 1]["Bz"][:, :, 6].T)
