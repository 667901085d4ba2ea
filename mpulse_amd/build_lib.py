"""Build the gfx950 shared library libmpulse_hip.so in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmpulse_hip.so")
SOURCES = ["mpx_kernels.hip", "mpx_pass.hip", "mpx_adjoint.hip", "mpx_gauss.hip", "mpx_solver.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # float64 parity with the CPU reference: no contraction into fused multiply-adds
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build(verbose=False):
    objs, procs = [], []
    for src in SOURCES:  # the translation units compile in parallel
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
    for p, cmd in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, cmd)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-L/opt/rocm/lib", "-lhipfft",
           "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


HOST_LIB = os.path.join(HERE, "libmpulse_host.so")
HOST_EXE = os.path.join(HERE, "mpulse_hip_run")
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", "-I", os.path.join(HERE, "..", "include")]
LINK = ["-L", HERE, "-lmpulse_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]


def build_host(verbose=False):
    """The C++ host layer (setup front-end, FieldSolver plugins, driver) and its CLI."""
    host = os.path.join(HERE, "host")
    cmd = [CXX, *HOST_FLAGS, "-shared", os.path.join(host, "mpulse_host.cpp"), os.path.join(host, "mpx_hdf5.cpp"),
           "-o", HOST_LIB + ".tmp", *LINK]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    cmd = [CXX, *HOST_FLAGS, os.path.join(host, "mpulse_main.cpp"), "-o", HOST_EXE + ".tmp", "-L", HERE,
           "-lmpulse_host", *LINK]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(HOST_EXE + ".tmp", HOST_EXE)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
