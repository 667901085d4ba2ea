"""Field output in the reference's HDF5 on-disk form (CPU).

The fixture tests/golden/ref_testhdf.h5 is the file the reference's own test program
(test/testhdf.cpp) wrote with the HDF5 library through HDFostream.  The test reader
(tests/hdf5_read.py, a restatement of the HDF5 file-format specification) is pinned on it:
data0 is A(i, j) = exp(j - i) sqrt(j) bit for bit, i fastest in memory.  The writer of the host
layer is pinned on it byte for byte: the same two datasets, their "extent" attributes and the
fixture's modification time give the fixture's 3288 bytes.
"""
import os

import numpy as np
import pytest

from hdf5_read import H5Error, H5File, read_file
from mpulse_amd import abi

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "ref_testhdf.h5")


def fixture_matrices():
    """A and the parts of B that testhdf.cpp defines (it reads A outside its index range for
    the rest of B), both in the memory order of the reference's Matrix (first index fastest)."""
    A = np.array([[np.exp(j - i) * np.sqrt(j) for j in range(3, 9)] for i in range(2, 7)])
    return A


def test_reader_on_the_reference_file():
    f = read_file(FIXTURE)
    assert f.names == ["data0", "data1"]
    assert (f.leaf_k, f.internal_k) == (4, 16)
    a = f.read("data0")
    assert a.shape == (5, 6) and a.dtype == np.float64
    assert np.array_equal(a.ravel(), fixture_matrices().ravel(order="F"))
    assert list(f.attrs("data0")["extent"]) == [2, 6, 3, 8]  # low0, high0, low1, high1
    b = f.read("data1").ravel().reshape((5, 5, 5), order="F")
    assert list(f.attrs("data1")["extent"]) == [4, 8, 5, 9, 6, 10]
    Af = lambda p, q: np.exp(q - p) * np.sqrt(q)
    for i in range(4, 7):  # entries whose factors lie inside A's index range
        for j in range(5, 7):
            for k in range(6, 9):
                v = Af(i, j) * Af(j, k) * np.sin(k)
                assert abs(b[i - 4, j - 5, k - 6] - v) <= 2 * np.spacing(abs(v)), (i, j, k)
    assert f.mtime("data0") == f.mtime("data1") == 1177340639


def test_writer_reproduces_the_reference_file(tmp_path):
    f = read_file(FIXTURE)
    out = tmp_path / "w.h5"
    with abi.HDFostream(str(out), capacity=2) as h:
        h.set_mtime(f.mtime("data0"))
        for n in ("data0", "data1"):
            h.write(f.read(n), extent=f.attrs(n)["extent"])
    mine, ref = out.read_bytes(), open(FIXTURE, "rb").read()
    assert len(mine) == len(ref)
    diff = [i for i in range(len(ref)) if mine[i] != ref[i]]
    assert not diff, f"first differing byte at 0x{diff[0]:x}"


def test_writer_round_trip_many_datasets(tmp_path):
    """20 datasets (three symbol-table nodes, names out of numeric order: "Ex", "Ex1", ...,
    "Ex10" sorts before "Ex2"), ranks 1-3, without attributes as the current HDFostream
    writes them; the file is complete after every write."""
    rng = np.random.default_rng(3)
    out = tmp_path / "m.h5"
    arrays = []
    with abi.HDFostream(str(out), capacity=24) as h:
        h.set_block_name("Ex")
        for q in range(20):
            shape = [(7,), (3, 5), (2, 3, 4)][q % 3]
            a = rng.standard_normal(shape)
            arrays.append(a)
            h.write(a)
            f = read_file(str(out))  # a valid file after each dataset
            assert len(f.names) == q + 1
    f = read_file(str(out))
    names = ["Ex"] + [f"Ex{q}" for q in range(1, 20)]
    assert f.names == sorted(names)
    for n, a in zip(names, arrays):
        got = f.read(n)
        assert got.shape == a.shape and np.array_equal(got, a), n
        assert f.attrs(n) == {}


def test_writer_limits(tmp_path):
    with abi.HDFostream(str(tmp_path / "l.h5"), capacity=1) as h:
        h.write(np.zeros(3))
        with pytest.raises(abi.MpxError):
            h.write(np.zeros(3))
        with pytest.raises(abi.MpxError):
            h.write(np.zeros((1, 1, 1, 1)))
    with pytest.raises(H5Error):
        H5File(b"not an hdf5 file")


def test_fielddiag_keys_parse(capfd):
    """The fielddiag block and its keys (src/fielddiag.cpp:9-15, src/diagnostic.cpp:34-47) are
    part of the setup grammar; they leave the solver descriptor unchanged."""
    base = ("globals { grid-x 8 grid-y 8 grid-z 8 T-total 3 } fieldsim { fdtd-plrc-nonlin { } single_periodic "
            "%s }")
    d0, t0 = abi.parse_setup(base % "")
    d1, t1 = abi.parse_setup(base % "fielddiag { field Ex file out_#p_#t.h5 append n interval 2 }")
    assert t0 == t1 == 3
    assert bytes(d0) == bytes(d1)
    d2, _ = abi.parse_setup(base % ("energy { file e.dat interval 5 append y } "
                                    "linediag { field Ey file l.dat dim y posx 3 posz 4 format gnu } "
                                    "slicediag { field Bz file s_#t.h5 plane z pos 2 }"))
    assert bytes(d0) == bytes(d2)
    # FrequencyDiag (src/freqdiag.cpp:72-95) and FollowMax (src/followmax.cpp:115-125) keys
    d3, _ = abi.parse_setup(base % ("frequency { field Ey file f.dat x 2 y 3 z 4 interval 5 } "
                                    "followmax { field Ex file m.dat dim z posx 2 posy 3 posz 4 start 7 append y }"))
    assert bytes(d0) == bytes(d3)
    # an unknown key is reported and skipped, as the reference's parser does (src/rebuild.cpp:48-51)
    capfd.readouterr()
    d4, _ = abi.parse_setup(base % "followmax { field Ex nosuchkey start 3 }")
    assert "Unknown Parameter nosuchkey" in capfd.readouterr().err
    assert bytes(d0) == bytes(d4)
