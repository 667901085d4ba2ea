"""mpx_energy_to, the entry point that allocates nothing (GPU).

It returns mpx_energy's value bit for bit, through device memory the caller owns, with the
partial sums in the library's file-scope scratch array: two calls in a row, the second on a
smaller grid (the scratch keeps the first call's partials, which the second must not see), in
float64 and in FP32-res; a grid needing more partial sums than the scratch holds is refused
before any GPU work.
"""
import ctypes as C

import numpy as np
import pytest

from mpulse_amd import abi, configs

pytestmark = pytest.mark.gpu


class DeviceDouble:
    """One float64 of device memory owned by the test (the caller), through the HIP runtime
    the library links.  The library is loaded first and the runtime is opened by its soname,
    so that the process's one runtime copy is used: torch ships its own libamdhip64.so, and
    opening the other file by its plain name would start a second runtime in the process."""

    def __init__(self):
        abi.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(8)) == 0
        self.write(-1.0)

    def write(self, v):
        x = C.c_double(v)
        assert self.hip.hipMemcpy(self.p, C.byref(x), C.c_size_t(8), 1) == 0  # host to device

    def read(self):
        x = C.c_double()
        assert self.hip.hipMemcpy(C.byref(x), self.p, C.c_size_t(8), 2) == 0  # device to host
        return x.value

    def close(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def energies(n, prec, path=abi.PATH_FUSED):
    d, _ = configs.plrc_kerr_pulse(n, 4, 5.0, 4, 2e9, 20, d_src=6, precision=prec, path=path)
    d.factories[1].tshift = 5.0
    out = DeviceDouble()
    with abi.Solver(d) as s:
        s.step(15)
        ref = s.energy()
        s.energy_to(out.p.value)
        s.sync()
        got = out.read()
    out.close()
    print(f"case: energy_to {n}^3 precision {prec}: {got!r} vs {ref!r}", flush=True)
    return ref, got


def test_energy_to_equals_energy_on_two_calls():
    for n, prec in ((48, abi.PREC_FP64), (20, abi.PREC_FP64), (40, abi.PREC_FP32_RES), (16, abi.PREC_FP32_RES)):
        ref, got = energies(n, prec)
        assert ref > 0 and got == ref, (n, prec, ref, got)


def test_energy_to_on_the_pass_path():
    ref, got = energies(36, abi.PREC_FP32_BULK, abi.PATH_PASS)
    assert got == ref


def test_energy_to_refuses_a_grid_beyond_the_scratch():
    d = abi.desc_defaults()
    d.solver = abi.SOLVER_PLAIN
    d.nx, d.ny, d.nz = (1 << 19) + 1, 1, 1  # one partial sum per x plane: one more than the scratch holds
    out = DeviceDouble()
    with abi.Solver(d) as s:
        with pytest.raises(abi.MpxError) as e:
            s.energy_to(out.p.value)
        assert e.value.code == abi.ERR_UNSUPPORTED and "partial sums" in str(e.value)
        assert np.isfinite(s.energy())
    assert out.read() == -1.0  # nothing was written
    out.close()
