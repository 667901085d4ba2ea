"""C ABI checks that need no GPU: the library loads, exports every entry point that
include/*.h declares, and the ctypes mirror has the C layout of the descriptors."""
import ctypes as C
import os
import re
import subprocess

import pytest

from mpulse_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if fn.endswith(".h"):
            src = open(os.path.join(INCLUDE, fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(mpx_[a-z_0-9]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_defaults():
    lib = abi.load()
    assert lib.mpx_abi_version() == abi.ABI_VERSION
    d = abi.Desc()
    lib.mpx_desc_defaults(C.byref(d))
    py = abi.desc_defaults()
    for f, _ in abi.Desc._fields_:
        if f != "factories":
            a, b = getattr(d, f), getattr(py, f)
            if hasattr(a, "__len__"):
                assert list(a) == list(b), f
            else:
                assert a == b, f
    for kind in (abi.FACTORY_CPML, abi.FACTORY_SHORT_PULSE):
        fc = abi.Factory()
        lib.mpx_factory_defaults(C.byref(fc), kind)
        fp = abi.factory_defaults(kind)
        for name in ("d", "kappa_max", "a_max", "sigma_max", "length", "width", "om0", "eps"):
            assert getattr(fc, name) == getattr(fp, name), name


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "mpulse_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(mpx_desc), sizeof(mpx_factory), sizeof(mpx_timing),
         offsetof(mpx_desc, factories), offsetof(mpx_desc, n_factories), offsetof(mpx_desc, l_eps),
         offsetof(mpx_factory, k), offsetof(mpx_timing, step_ms));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    vals = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    want = [C.sizeof(abi.Desc), C.sizeof(abi.Factory), C.sizeof(abi.Timing), abi.Desc.factories.offset,
            abi.Desc.n_factories.offset, abi.Desc.l_eps.offset, abi.Factory.k.offset, abi.Timing.step_ms.offset]
    assert vals == want


def test_invalid_descriptors_rejected_without_gpu_work():
    """Argument validation happens before any device call (src/fdtd_plrc.t:20-24 for Omega)."""
    lib = abi.load()
    h = C.c_void_p()
    d = abi.desc_defaults()
    d.abi_version = 99
    assert lib.mpx_create(C.byref(d), C.byref(h)) == abi.ERR_INVALID
    d = abi.desc_defaults()
    d.nx = 0
    assert lib.mpx_create(C.byref(d), C.byref(h)) == abi.ERR_INVALID
    d = abi.desc_defaults()
    d.l_om2[1] = 0.0
    assert lib.mpx_create(C.byref(d), C.byref(h)) == abi.ERR_INVALID
    assert b"Omega should not be zero" in lib.mpx_last_error()
    d = abi.desc_defaults()
    d.solver = abi.SOLVER_PLAIN
    abi.add_factory(d, abi.factory_defaults(abi.FACTORY_CPML))
    assert lib.mpx_create(C.byref(d), C.byref(h)) == abi.ERR_INVALID
    assert b"Additional currents cannot be used" in lib.mpx_last_error()
    d = abi.desc_defaults()
    d.precision = abi.PREC_FP32_BULK
    d.path = abi.PATH_REFERENCE
    assert lib.mpx_create(C.byref(d), C.byref(h)) == abi.ERR_UNSUPPORTED
    assert lib.mpx_step(None, 1) == abi.ERR_INVALID
