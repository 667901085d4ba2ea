"""ctypes mirror of include/mpulse_hip.h and a thin Python handle over the C ABI.

The HIP library is the only compute path: if libmpulse_hip.so is missing this module
raises on load instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpulse_hip.so")

ABI_VERSION = 3
MAX_FACTORIES = 8

OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED, ERR_NAME, ERR_NUMERIC = 0, -1, -2, -3, -4, -5, -6

SOLVER_PLRC_NONLIN, SOLVER_PLRC_LIN, SOLVER_PLAIN, SOLVER_NL, SOLVER_DISP = 0, 1, 2, 3, 4
BOUNDARY_PERIODIC, BOUNDARY_XY_PERIODIC = 0, 1
PREC_FP64, PREC_FP32_BULK, PREC_FP32_RES = 0, 1, 2
PATH_FUSED, PATH_REFERENCE, PATH_SWEEP, PATH_PASS = 0, 1, 2, 3
FACTORY_CPML, FACTORY_CPML_1D, FACTORY_SHORT_PULSE, FACTORY_PLANE_WAVE, FACTORY_PLANE_GAUSS = 1, 2, 3, 4, 5
FACTORY_FOCUSED_PULSE = 6
FACTORY_PLASMA_CURRENT = 7
FACTORY_GAUSS_INJECT = 8

# Direction enum of the reference (src/mpulse.h:28)
NORTH, SOUTH, WEST, EAST, UP, DOWN = 0, 1, 2, 3, 4, 5

SOLVER_KEYS = {"fdtd-plrc-nonlin": SOLVER_PLRC_NONLIN, "fdtd-plrc-lin": SOLVER_PLRC_LIN,
               "fdtd-plain": SOLVER_PLAIN, "fdtd-nl": SOLVER_NL, "fdtd-disp": SOLVER_DISP}


class Factory(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("d", C.c_int),
        ("kappa_max", C.c_double), ("a_max", C.c_double), ("sigma_max", C.c_double),
        ("length", C.c_double), ("width", C.c_double), ("om0", C.c_double),
        ("tshift", C.c_double), ("zshift", C.c_double), ("phase", C.c_double),
        ("amp", C.c_double), ("eps", C.c_double),
        ("k", C.c_double * 3), ("h", C.c_double * 3), ("h_bg", C.c_double * 3), ("e_bg", C.c_double * 3),
        ("ramp", C.c_double), ("gwidth", C.c_double), ("goffset", C.c_double),
        ("em", C.c_double), ("gamma", C.c_double), ("rho_bg", C.c_double),
        ("gsize_t", C.c_int), ("toff", C.c_int),
        ("k0", C.c_double), ("gshift", C.c_double), ("gtime", C.c_double),
    ]


class Desc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("solver", C.c_int), ("boundary", C.c_int), ("precision", C.c_int), ("path", C.c_int),
        ("eps", C.c_double), ("chi", C.c_double),
        ("l_eps", C.c_double * 3), ("l_delta", C.c_double * 3), ("l_om2", C.c_double * 3),
        ("d_eps2", C.c_double * 3), ("d_om2", C.c_double * 3),
        ("n_factories", C.c_int),
        ("factories", Factory * MAX_FACTORIES),
        ("plasma", C.c_int), ("plasma_k", C.c_int),
        ("plasma_nu", C.c_double), ("plasma_mpa", C.c_double), ("plasma_wion", C.c_double),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("n_classes", C.c_int),
        ("name", (C.c_char * 32) * 16),
        ("mean_ms", C.c_double * 16),
        ("launches", C.c_int * 16),
        ("step_ms", C.c_double),
    ]


def desc_defaults():
    """Descriptor with the reference's MakeParamMap defaults (src/globals.cpp:21-36,
    src/fdtd_plrc.t:411-473)."""
    d = Desc()
    d.abi_version = ABI_VERSION
    d.nx = d.ny = d.nz = 128
    d.dx = d.dy = d.dz = 0.1
    d.dt = 1e-2
    d.solver = SOLVER_PLRC_NONLIN
    d.eps, d.chi = 1.0, 0.1
    for n in range(3):
        d.l_om2[n] = 1.0
    d.plasma_k = 5  # src/plasmadensity.cpp:66-69
    d.plasma_nu = d.plasma_mpa = d.plasma_wion = 1.0
    return d


def factory_defaults(kind):
    f = Factory()
    f.kind = kind
    if kind in (FACTORY_CPML, FACTORY_CPML_1D):
        f.d, f.kappa_max, f.a_max, f.sigma_max = 8, 15.0, 0.25, 3.0
    else:
        f.d = 15
        f.length, f.width, f.om0, f.amp, f.eps = 1.0, 1.0, 2 * np.pi, 1.0, 1.0
        f.k[2] = 1.0
        f.ramp, f.gwidth, f.goffset = 0.5, 10.0, 40.0
    f.em, f.gamma, f.rho_bg = 1.0, 0.1, 0.0  # src/plasmacurrent.cpp:10-12
    f.gsize_t, f.toff, f.k0, f.gshift, f.gtime = 4096, 0, 2 * np.pi, 0.0, 0.0  # src/gaussinject.cpp:37-47
    return f


def add_factory(desc, fac):
    if desc.n_factories >= MAX_FACTORIES:
        raise ValueError("too many factories")
    desc.factories[desc.n_factories] = fac
    desc.n_factories += 1
    return desc


class Objective(C.Structure):
    _fields_ = [("plane_k", C.c_int), ("weight", C.c_double)]


class MpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mpx error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libmpulse_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(there is no fallback compute path)")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "mpx_create": (C.c_int, [C.POINTER(Desc), C.POINTER(P)]),
        "mpx_destroy": (C.c_int, [P]),
        "mpx_step": (C.c_int, [P, C.c_int]),
        "mpx_step_init": (C.c_int, [P]),
        "mpx_get_time": (C.c_int, [P, C.POINTER(C.c_int)]),
        "mpx_set_time": (C.c_int, [P, C.c_int]),
        "mpx_get_field": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_double), C.c_size_t]),
        "mpx_set_field": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_double), C.c_size_t]),
        "mpx_fetch_field": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_double), C.c_size_t, C.POINTER(P)]),
        "mpx_fetch_wait": (C.c_int, [P]),
        "mpx_sampler_create": (C.c_int, [P, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(P)]),
        "mpx_sampler_record": (C.c_int, [P]),
        "mpx_sampler_drain": (C.c_int, [P, C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_int)]),
        "mpx_sampler_destroy": (C.c_int, [P]),
        "mpx_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(P)]),
        "mpx_host_free": (C.c_int, [P]),
        "mpx_get_layer": (C.c_int, [P, C.c_char_p, C.c_int, C.POINTER(C.c_double), C.c_size_t,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "mpx_set_probes": (C.c_int, [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]),
        "mpx_get_probes": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "mpx_energy": (C.c_int, [P, C.POINTER(C.c_double)]),
        "mpx_energy_to": (C.c_int, [P, C.c_void_p]),
        "mpx_sync": (C.c_int, [P]),
        "mpx_time_steps": (C.c_int, [P, C.c_int, C.POINTER(Timing)]),
        "mpx_state_bytes": (C.c_int, [P, C.POINTER(C.c_size_t)]),
        "mpx_copy_probe": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
        "mpx_set_tuning": (C.c_int, [P, C.c_char_p, C.c_int]),
        "mpx_batch_create": (C.c_int, [C.POINTER(Desc), C.c_int, C.POINTER(P)]),
        "mpx_batch_destroy": (C.c_int, [P]),
        "mpx_batch_size": (C.c_int, [P, C.POINTER(C.c_int)]),
        "mpx_batch_run": (C.c_int, [P, C.c_int, C.POINTER(P)]),
        "mpx_batch_step": (C.c_int, [P, C.c_int]),
        "mpx_batch_sync": (C.c_int, [P]),
        "mpx_batch_get_time": (C.c_int, [P, C.POINTER(C.c_int)]),
        "mpx_batch_set_tuning": (C.c_int, [P, C.c_char_p, C.c_int]),
        "mpx_batch_autotune": (C.c_int, [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "mpx_batch_time_steps": (C.c_int, [P, C.c_int, C.POINTER(C.c_double)]),
        "mpx_adjoint_grad": (C.c_int, [P, C.POINTER(Objective), C.c_int, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "mpx_last_error": (C.c_char_p, []),
        "mpx_store_round": (C.c_double, [C.c_double, C.c_int]),
        "mpx_abi_version": (C.c_int, []),
        "mpx_desc_defaults": (None, [C.POINTER(Desc)]),
        "mpx_factory_defaults": (None, [C.POINTER(Factory), C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.mpx_abi_version() != ABI_VERSION:
        raise ImportError("libmpulse_hip.so ABI version mismatch")
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class FetchHandle:
    """An mpx_fetch_field copy in flight; wait() returns once the data has landed.  The handle
    holds the destination (array or PinnedArray) until then, so the copy never lands in freed
    host memory."""

    def __init__(self, solver, h, out):
        self.solver, self.h, self.out = solver, h, out

    def wait(self):
        if self.h is not None:
            rc = self.solver.lib.mpx_fetch_wait(self.h)
            self.h = None
            self.solver._pending.discard(self)
            if isinstance(self.out, PinnedArray):
                self.out._pending.discard(self)
            self.out = None
            self.solver._check(rc)


class Sampler:
    """An mpx_sampler: records grid `name` at one storage cell (dim -1) or along the storage line
    through `pos` in dimension dim, once per record(); drain() returns the records made since
    the last drain, shape (records, values per record)."""

    def __init__(self, solver, name, dim, pos, capacity):
        self.solver, self.h = solver, None
        self.len = 1 if dim < 0 else solver.shape[dim]
        self.capacity = capacity
        p = (C.c_int * 3)(*[int(v) for v in pos])
        h = C.c_void_p()
        solver._check(solver.lib.mpx_sampler_create(solver.h, name.encode(), dim, p, capacity, C.byref(h)))
        self.h = h
        solver._samplers.add(self)

    def record(self):
        self.solver._check(self.solver.lib.mpx_sampler_record(self.h))

    def drain(self):
        out = np.empty(self.capacity * self.len, dtype=np.float64)
        n = C.c_int()
        self.solver._check(self.solver.lib.mpx_sampler_drain(self.h, _dptr(out), out.size, C.byref(n)))
        return out[: n.value * self.len].reshape(n.value, self.len)

    def close(self):
        if self.h is not None:
            rc = self.solver.lib.mpx_sampler_destroy(self.h)
            self.h = None
            self.solver._samplers.discard(self)
            self.solver._check(rc)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class PinnedArray:
    """A float64 array in page-locked host memory (mpx_host_alloc), freed by close()."""

    def __init__(self, shape):
        self.lib = load()
        n = int(np.prod(shape))
        p = C.c_void_p()
        rc = self.lib.mpx_host_alloc(max(1, n) * 8, C.byref(p))
        if rc != OK:
            raise MpxError(rc, self.lib.mpx_last_error().decode())
        self.p = p
        self.array = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(n,)).reshape(shape)
        self._pending = set()  # fetches still copying into this buffer

    def close(self):
        if self.p is not None:
            for f in list(self._pending):
                f.wait()
            self.array = None
            self.lib.mpx_host_free(self.p)
            self.p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Solver:
    """Owner of one mpx_solver handle (device state on the current GPU)."""

    def __init__(self, desc):
        self.lib = load()
        self.desc = desc
        self.shape = (desc.nx + 2, desc.ny + 2, desc.nz + 2)
        self._own = True
        self._samplers, self._pending = set(), set()
        h = C.c_void_p()
        self._check(self.lib.mpx_create(C.byref(desc), C.byref(h)))
        self.h = h

    @classmethod
    def _view(cls, lib, desc, handle):
        """A non-owning handle on one run of a Batch (fields, energy; stepping is refused)."""
        v = cls.__new__(cls)
        v.lib, v.desc, v.h, v._own = lib, desc, handle, False
        v._samplers, v._pending = set(), set()
        v.shape = (desc.nx + 2, desc.ny + 2, desc.nz + 2)
        return v

    def _check(self, rc):
        if rc != OK:
            raise MpxError(rc, self.lib.mpx_last_error().decode())

    def close(self):
        """Wait for the fetches in flight and destroy the open samplers, then the solver."""
        if getattr(self, "h", None):
            for f in list(getattr(self, "_pending", ())):
                try:
                    f.wait()
                except MpxError:
                    pass
            for smp in list(getattr(self, "_samplers", ())):
                try:
                    smp.close()
                except MpxError:
                    pass
            if getattr(self, "_own", True):
                self.lib.mpx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n=1):
        self._check(self.lib.mpx_step(self.h, n))

    def step_init(self):
        self._check(self.lib.mpx_step_init(self.h))

    @property
    def time(self):
        t = C.c_int()
        self._check(self.lib.mpx_get_time(self.h, C.byref(t)))
        return t.value

    def set_time(self, t):
        """The time index the sources read (Process::time); unchanged, it leaves the state as is."""
        self._check(self.lib.mpx_set_time(self.h, int(t)))

    def sync(self):
        self._check(self.lib.mpx_sync(self.h))

    def get(self, name):
        out = np.empty(self.shape, dtype=np.float64)
        self._check(self.lib.mpx_get_field(self.h, name.encode(), _dptr(out), out.size))
        return out

    def fetch_async(self, name, out):
        """Queue a copy of grid `name` into `out` (a float64 array or a PinnedArray, shape
        (nz+2, ny+2, nx+2): the reference's Grid order, x fastest) behind the steps queued so
        far; returns a handle whose wait() returns once the data is in `out`.  The handle keeps
        `out` alive until then; a PinnedArray's close() and this solver's close() wait for it."""
        arr = out.array if isinstance(out, PinnedArray) else out
        if arr is None or arr.dtype != np.float64 or not arr.flags.c_contiguous or arr.shape != self.shape[::-1]:
            raise ValueError(f"expected a C-contiguous float64 array of shape {self.shape[::-1]}")
        h = C.c_void_p()
        self._check(self.lib.mpx_fetch_field(self.h, name.encode(), _dptr(arr), arr.size, C.byref(h)))
        fh = FetchHandle(self, h, out)
        self._pending.add(fh)
        if isinstance(out, PinnedArray):
            out._pending.add(fh)
        return fh

    def fetch(self, name):
        """Grid `name` in the reference's Grid order (x fastest), shape (nz+2, ny+2, nx+2)."""
        out = np.empty(self.shape[::-1], dtype=np.float64)
        self.fetch_async(name, out).wait()
        return out

    def sampler(self, name, dim, pos, capacity):
        """A Sampler of grid `name` (close it before the solver)."""
        return Sampler(self, name, dim, pos, capacity)

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"expected shape {self.shape}")
        self._check(self.lib.mpx_set_field(self.h, name.encode(), _dptr(a), a.size))

    def get_layer(self, name, direction):
        lo = (C.c_int * 3)()
        hi = (C.c_int * 3)()
        self._check(self.lib.mpx_get_layer(self.h, name.encode(), direction, None, 0, lo, hi))
        shape = tuple(hi[d] - lo[d] + 1 for d in range(3))
        out = np.empty(shape, dtype=np.float64)
        self._check(self.lib.mpx_get_layer(self.h, name.encode(), direction, _dptr(out), out.size, lo, hi))
        return out, tuple(lo), tuple(hi)

    def set_probes(self, ijk, comps, capacity):
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        comps = np.ascontiguousarray(comps, dtype=np.int32).reshape(-1)
        self._check(self.lib.mpx_set_probes(self.h, len(comps), ijk.ctypes.data_as(C.POINTER(C.c_int)),
                                            comps.ctypes.data_as(C.POINTER(C.c_int)), capacity))
        self._nprobe = len(comps)

    def get_probes(self):
        n = C.c_int()
        self._check(self.lib.mpx_get_probes(self.h, None, C.byref(n)))
        out = np.empty((n.value, self._nprobe), dtype=np.float64)
        self._check(self.lib.mpx_get_probes(self.h, _dptr(out), C.byref(n)))
        return out

    def energy(self):
        e = C.c_double()
        self._check(self.lib.mpx_energy(self.h, C.byref(e)))
        return e.value

    def energy_to(self, device_ptr):
        """mpx_energy_to: the energy written to device memory at device_ptr (one float64) on the
        solver's stream, without allocating or waiting; sync() before reading it."""
        self._check(self.lib.mpx_energy_to(self.h, C.c_void_p(int(device_ptr))))

    def time_steps(self, nsteps):
        t = Timing()
        self._check(self.lib.mpx_time_steps(self.h, nsteps, C.byref(t)))
        res = {"step_ms": t.step_ms, "kernels": {}}
        for q in range(t.n_classes):
            res["kernels"][t.name[q].value.decode()] = (t.mean_ms[q], t.launches[q])
        return res

    def tune(self, key, value):
        self._check(self.lib.mpx_set_tuning(self.h, key.encode(), int(value)))

    def state_bytes(self):
        b = C.c_size_t()
        self._check(self.lib.mpx_state_bytes(self.h, C.byref(b)))
        return b.value

    def adjoint_grad(self, plane_k, weight, nsteps, checkpoint_every=50, grad=True):
        """Step nsteps forward, accumulating J = weight * sum_steps sum_ij |E(i, j, plane_k)|^2,
        and (grad) return (J, dJ/dEps full-storage array, dJ/dchi) from the discrete adjoint;
        without grad return (J, None, None)."""
        obj = Objective(int(plane_k), float(weight))
        J = C.c_double()
        if not grad:
            self._check(self.lib.mpx_adjoint_grad(self.h, C.byref(obj), nsteps, checkpoint_every, C.byref(J),
                                                  None, None))
            return J.value, None, None
        deps = np.empty(self.shape, dtype=np.float64)
        dchi = C.c_double()
        self._check(self.lib.mpx_adjoint_grad(self.h, C.byref(obj), nsteps, checkpoint_every, C.byref(J),
                                              _dptr(deps), C.byref(dchi)))
        return J.value, deps, dchi.value


class Batch:
    """Owner of one mpx_batch handle: independent runs stepped together (see mpx_batch_create)."""

    def __init__(self, descs):
        self.lib = load()
        self.descs = list(descs)
        arr = (Desc * len(self.descs))(*self.descs)
        h = C.c_void_p()
        self._check(self.lib.mpx_batch_create(arr, len(self.descs), C.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != OK:
            raise MpxError(rc, self.lib.mpx_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.descs)

    def run(self, i):
        r = C.c_void_p()
        self._check(self.lib.mpx_batch_run(self.h, i, C.byref(r)))
        return Solver._view(self.lib, self.descs[i], r)

    def step(self, n=1):
        self._check(self.lib.mpx_batch_step(self.h, n))

    def sync(self):
        self._check(self.lib.mpx_batch_sync(self.h))

    @property
    def time(self):
        t = C.c_int()
        self._check(self.lib.mpx_batch_get_time(self.h, C.byref(t)))
        return t.value

    def tune(self, key, value):
        self._check(self.lib.mpx_batch_set_tuning(self.h, key.encode(), int(value)))

    def autotune(self, steps_per_trial=3):
        res = (C.c_int * 4)()
        ms = C.c_double()
        self._check(self.lib.mpx_batch_autotune(self.h, steps_per_trial, res, C.byref(ms)))
        if res[0] == 0:
            return {"structure": "fused", "march_b": res[1], "blocks_per_cu": res[2], "ms_per_step": ms.value}
        if res[0] == 2:
            return {"structure": "pass", "pass_seg": res[1], "pass_rows": res[2], "ms_per_step": ms.value}
        return {"structure": "sweep", "variant": res[1], "seg": res[2], "blocks_per_cu": res[3],
                "ms_per_step": ms.value}

    def time_steps(self, nsteps):
        ms = C.c_double()
        self._check(self.lib.mpx_batch_time_steps(self.h, nsteps, C.byref(ms)))
        return ms.value


def copy_probe(nbytes, reps=20):
    """Mean ms of one device-to-device streaming copy of nbytes (read + write)."""
    lib = load()
    ms = C.c_double()
    rc = lib.mpx_copy_probe(nbytes, reps, C.byref(ms))
    if rc != OK:
        raise MpxError(rc, lib.mpx_last_error().decode())
    return ms.value


# ---------------------------------------------------------------------------------------
# C++ host layer (setup files, FieldSolver plugins): libmpulse_host.so
# ---------------------------------------------------------------------------------------

HOST_LIB_PATH = os.path.join(HERE, "libmpulse_host.so")
_host = None


def load_host():
    global _host
    if _host is not None:
        return _host
    load()  # the host layer links the HIP library
    if not os.path.exists(HOST_LIB_PATH):
        raise ImportError(f"{HOST_LIB_PATH} is missing: build it with __graft_entry__.build()")
    h = C.CDLL(HOST_LIB_PATH)
    P = C.c_void_p
    h.mph_parse_setup.restype = C.c_int
    h.mph_parse_setup.argtypes = [C.c_char_p, C.POINTER(Desc), C.POINTER(C.c_int)]
    h.mph_create.restype = C.c_int
    h.mph_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(P)]
    h.mph_run.restype = C.c_int
    h.mph_run.argtypes = [P, C.c_int]
    h.mph_solver.restype = P
    h.mph_solver.argtypes = [P]
    h.mph_destroy.restype = C.c_int
    h.mph_destroy.argtypes = [P]
    h.mph_last_error.restype = C.c_char_p
    h.mph_last_error.argtypes = []
    for name, args in (("mph_h5_open", [C.c_char_p, C.c_int, C.POINTER(P)]),
                       ("mph_h5_set_block_name", [P, C.c_char_p]),
                       ("mph_h5_set_mtime", [P, C.c_uint]),
                       ("mph_h5_write", [P, C.c_int, C.POINTER(C.c_ulonglong), C.POINTER(C.c_double),
                                         C.POINTER(C.c_int), C.c_int]),
                       ("mph_h5_close", [P])):
        fn = getattr(h, name)
        fn.restype = C.c_int
        fn.argtypes = args
    _host = h
    return h


class HDFostream:
    """MPulse's HDFostream (src/hdfstream.{h,t,cpp}) through the host library: float64
    datasets named data0, data1, ... (or block name, name1, ...), written in the reference's
    on-disk form without the HDF5 library."""

    def __init__(self, path, capacity=64):
        self.h = load_host()
        p = C.c_void_p()
        rc = self.h.mph_h5_open(os.fsencode(path), capacity, C.byref(p))
        if rc != OK:
            raise MpxError(rc, self.h.mph_last_error().decode())
        self.p = p

    def _check(self, rc):
        if rc != OK:
            raise MpxError(rc, self.h.mph_last_error().decode())

    def set_block_name(self, name):
        self._check(self.h.mph_h5_set_block_name(self.p, name.encode()))

    def set_mtime(self, t):
        self._check(self.h.mph_h5_set_mtime(self.p, int(t)))

    def write(self, array, extent=None):
        """One dataset: the HDF5 dims are array.shape, the data array in C order."""
        a = np.ascontiguousarray(array, dtype=np.float64)
        dims = (C.c_ulonglong * a.ndim)(*a.shape)
        ext = None if extent is None else np.ascontiguousarray(extent, dtype=np.int32)
        self._check(self.h.mph_h5_write(self.p, a.ndim, dims, _dptr(a),
                                        None if ext is None else ext.ctypes.data_as(C.POINTER(C.c_int)),
                                        0 if ext is None else ext.size))

    def close(self):
        if self.p is not None:
            self._check(self.h.mph_h5_close(self.p))
            self.p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def parse_setup(text):
    """Descriptor and T-total of an MPulse setup file (no GPU work)."""
    h = load_host()
    d = Desc()
    steps = C.c_int()
    rc = h.mph_parse_setup(text.encode(), C.byref(d), C.byref(steps))
    if rc != OK:
        raise MpxError(rc, h.mph_last_error().decode())
    return d, steps.value


class HostSimulation:
    """A setup file driven through the C++ Process / FieldSimulation / FieldSolver layer."""

    def __init__(self, text, precision=PREC_FP64, path=PATH_FUSED):
        self.h = load_host()
        self.desc, self.total = parse_setup(text)
        sim = C.c_void_p()
        rc = self.h.mph_create(text.encode(), precision, path, C.byref(sim))
        if rc != OK:
            raise MpxError(rc, self.h.mph_last_error().decode())
        self.sim = sim
        self.shape = (self.desc.nx + 2, self.desc.ny + 2, self.desc.nz + 2)

    def run(self, steps=-1):
        rc = self.h.mph_run(self.sim, steps)
        if rc != OK:
            raise MpxError(rc, self.h.mph_last_error().decode())

    def get(self, name):
        lib = load()
        out = np.empty(self.shape, dtype=np.float64)
        rc = lib.mpx_get_field(self.h.mph_solver(self.sim), name.encode(), _dptr(out), out.size)
        if rc != OK:
            raise MpxError(rc, lib.mpx_last_error().decode())
        return out

    def close(self):
        if getattr(self, "sim", None):
            self.h.mph_destroy(self.sim)
            self.sim = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def store_round(x, precision):
    """The value a field word holds once stored in `precision` (PREC_*)."""
    return load().mpx_store_round(float(x), int(precision))
