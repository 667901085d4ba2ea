"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) -- test infrastructure.

The oracle is the checker only: tests compare the HIP library against it.  It shares the
mpx_desc layout with the product so both run the same workload description.
"""
import ctypes as C
import os

import numpy as np

from mpulse_amd.abi import Desc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
LIB_OMP = os.path.join(ROOT, "oracle", "_build", "liboracle_omp.so")
REF_SPECFUNC = os.path.join(ROOT, "oracle", "_ref", "libref_specfunc.so")
REF_PULSE = os.path.join(ROOT, "oracle", "_ref", "refpulse")

_libs = {}


def mutant_path(m):
    """A build of the oracle with planted mistake `m` (`make -C oracle mutants`)."""
    return os.path.join(ROOT, "oracle", "_build", f"liboracle_mut{m}.so")


def load(omp=False, path=None):
    path = path or (LIB_OMP if omp else LIB)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run `make -C oracle`")
    lib = C.CDLL(path)
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    I3 = C.POINTER(C.c_int)
    sig = {
        "or_create": (P, [C.POINTER(Desc), C.c_char_p, C.c_size_t]),
        "or_destroy": (None, [P]),
        "or_step": (C.c_int, [P, C.c_int]),
        "or_step_init": (C.c_int, [P]),
        "or_get_time": (C.c_int, [P]),
        "or_set_storage": (C.c_int, [P, C.c_int]),
        "or_store_round": (C.c_double, [C.c_double, C.c_int]),
        "or_set_time": (None, [P, C.c_int]),
        "or_get_grid": (C.c_int, [P, C.c_char_p, D, C.c_size_t]),
        "or_set_grid": (C.c_int, [P, C.c_char_p, D, C.c_size_t]),
        "or_get_layer": (C.c_int, [P, C.c_char_p, C.c_int, D, C.c_size_t, I3, I3]),
        "or_get_line": (C.c_int, [P, C.c_char_p, D, C.c_size_t]),
        "or_energy": (C.c_double, [P]),
        "or_faddeeva": (None, [C.c_double, C.c_double, D]),
        "or_pole_constants": (None, [D, D, D, C.c_double, D, D, D, D]),
        "or_kerr_newton": (C.c_double, [C.c_double, C.c_double, C.c_double, I3]),
        "or_pulse_efunc": (None, [D, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, D]),
        "or_pulse_bfunc": (None, [D, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, D]),
        "or_pulse_amp": (C.c_double, [D]),
        "or_cpml_coeff": (C.c_int, [P, C.c_int, C.c_int, D, D, I3, I3]),
        "or_gauss_block": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, D, D, C.POINTER(C.c_int)]),
        "or_focused_plane": (None, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                    C.c_double, C.c_double, D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """Same surface as mpulse_amd.abi.Solver, computed on the CPU by the oracle."""

    def __init__(self, desc, omp=False, lib=None):
        self.lib = load(omp, lib)
        self.desc = desc
        self.shape = (desc.nx + 2, desc.ny + 2, desc.nz + 2)
        err = C.create_string_buffer(256)
        self.h = self.lib.or_create(C.byref(desc), err, 256)
        if not self.h:
            raise RuntimeError("oracle: " + err.value.decode())

    def close(self):
        if self.h:
            self.lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        self.lib.or_step(self.h, n)

    def set_storage(self, mode):
        """Round E, B and the CPML auxiliaries after every update: 0 float64 (the reference),
        1 float32, 2 float32 + 10-bit residual word."""
        if self.lib.or_set_storage(self.h, int(mode)):
            raise ValueError(mode)

    def step_init(self):
        self.lib.or_step_init(self.h)

    @property
    def time(self):
        return self.lib.or_get_time(self.h)

    def get(self, name):
        out = np.empty(self.shape)
        rc = self.lib.or_get_grid(self.h, name.encode(), _dp(out), out.size)
        if rc == -5:
            return np.zeros(self.shape)
        if rc:
            raise KeyError(name)
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if self.lib.or_set_grid(self.h, name.encode(), _dp(a), a.size):
            raise KeyError(name)

    def get_layer(self, name, direction):
        lo = (C.c_int * 3)()
        hi = (C.c_int * 3)()
        n = self.lib.or_get_layer(self.h, name.encode(), direction, None, 0, lo, hi)
        if n < 0:
            raise KeyError(name)
        shape = tuple(hi[d] - lo[d] + 1 for d in range(3))
        out = np.empty(shape)
        self.lib.or_get_layer(self.h, name.encode(), direction, _dp(out), out.size, lo, hi)
        return out, tuple(lo), tuple(hi)

    def get_line(self, name, n):
        out = np.empty(n)
        if self.lib.or_get_line(self.h, name.encode(), _dp(out), n):
            raise KeyError(name)
        return out

    def energy(self):
        return self.lib.or_energy(self.h)

    def cpml_coeff(self, is_h, direction):
        lo = (C.c_int * 1)()
        hi = (C.c_int * 1)()
        if self.lib.or_cpml_coeff(self.h, int(is_h), direction, None, None, lo, hi):
            raise KeyError(direction)
        n = hi[0] - lo[0] + 1
        b, c = np.empty(n), np.empty(n)
        self.lib.or_cpml_coeff(self.h, int(is_h), direction, _dp(b), _dp(c), lo, hi)
        return b, c, lo[0], hi[0]


def faddeeva(z):
    lib = load()
    out = np.empty(2)
    lib.or_faddeeva(float(z.real), float(z.imag), _dp(out))
    return complex(out[0], out[1])


def pole_constants(l_eps, l_delta, l_om2, dt):
    lib = load()
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (l_eps, l_delta, l_om2)]
    crec, dchi0, dxi0, sums = np.empty(6), np.empty(6), np.empty(6), np.empty(2)
    lib.or_pole_constants(_dp(a[0]), _dp(a[1]), _dp(a[2]), dt, _dp(crec), _dp(dchi0), _dp(dxi0), _dp(sums))
    cx = lambda v: v[0::2] + 1j * v[1::2]  # noqa: E731
    return cx(crec), cx(dchi0), cx(dxi0), sums[0], sums[1]


def kerr_newton(A, Cm, E0):
    lib = load()
    it = C.c_int()
    E = lib.or_kerr_newton(A, Cm, E0, C.byref(it))
    return E, it.value


def pulse_efunc(params, x, y, z, t, normalise=False):
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.empty(2)
    lib.or_pulse_efunc(_dp(p), int(normalise), x, y, z, t, _dp(out))
    return complex(out[0], out[1])


def pulse_bfunc(params, x, y, z, t, bx, normalise=False):
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.empty(2)
    lib.or_pulse_bfunc(_dp(p), int(normalise), x, y, z, t, int(bx), _dp(out))
    return complex(out[0], out[1])


def pulse_amp(params):
    lib = load()
    p = np.ascontiguousarray(params, dtype=np.float64)
    return lib.or_pulse_amp(_dp(p))


def focused_plane(fac, nx, ny, dx, dy, z, comp, sx, sy):
    lib = load()
    out = np.empty((nx + 2) * (ny + 2) * 2)
    lib.or_focused_plane(C.byref(fac), nx, ny, dx, dy, z, comp, sx, sy, _dp(out))
    v = out[0::2] + 1j * out[1::2]
    return v.reshape(nx + 2, ny + 2)


def gauss_block(fac, isH, nx, ny, dx, dy, dt, lowz, highz, time):
    """GaussInject sample block in force at step `time`: (F1, F2 of shape (Nt, nx+2, ny+2),
    first step of the block), or None when the source is inactive then."""
    lib = load()
    nt = lib.or_gauss_block(C.byref(fac), isH, nx, ny, dx, dy, dt, lowz, highz, time, None, None, None)
    if nt == 0:
        return None
    F1 = np.empty((nt, nx + 2, ny + 2))
    F2 = np.empty((nt, nx + 2, ny + 2))
    bt = C.c_int()
    lib.or_gauss_block(C.byref(fac), isH, nx, ny, dx, dy, dt, lowz, highz, time, _dp(F1), _dp(F2), C.byref(bt))
    return F1, F2, bt.value


def store_round(x, mode):
    """The oracle's storage rounding of one value (0 float64, 1 float32, 2 float32 + residual)."""
    return load().or_store_round(float(x), int(mode))
