"""Generate the golden vectors under tests/golden/ from reference code compiled in place.

Run in a container where the MPulse tree is mounted at /root/reference, after
`make -C oracle ref`, which compiles (unmodified, outputs only under oracle/_ref/):
  * src/specfunc.cpp                -> oracle/_ref/libref_specfunc.so  (Faddeeva_2)
  * test/testshortpulse.cpp         -> oracle/_ref/refpulse            (its Efunc/Bfunc)
The reference's test harness defines Efunc with an extra 1/c factor relative to
src/shortpulsefunctions.cpp (test/testshortpulse.cpp:264 vs src line 93); Bfunc is identical.
The fixtures store the harness output as printed; the tests apply the factor.

Fixtures written:
  faddeeva.json   w(z) of Faddeeva_2 at seeded and structured points
  shortpulse.json Efunc / Bfunc at the reference test's own sampling grid
                  (test/testshortpulse.cpp:398-426: setParam(1e6, 2, 2, 0, 0, 1, 1), y = z = 0,
                  x in [-5, 5] step 0.05, t in [0, 8] step 0.025; every 7th point kept) and at
                  seeded points for the pulse parameters of the benchmark workloads.
"""
import ctypes as C
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")


def faddeeva_points():
    rng = np.random.default_rng(1234)
    pts = []
    # structured grid over the upper half plane used by the pulse arguments
    for x in np.linspace(-8, 8, 33):
        for y in [0.0, 1e-3, 0.005, 0.02, 0.3, 1.0, 2.5, 6.0, 6.2, 6.3, 8.0, 15.0]:
            pts.append((x, y))
    # near the poles of the h = 0.5 rule: |Im z| < 0.01, 2|Re z| within 0.02 of an integer
    for m in range(0, 13):
        for dx in (-0.009, -0.004, 0.0, 0.004, 0.009, 0.011):
            for y in (0.0, 0.003, 0.0099):
                pts.append((0.5 * m + dx, y))
    # tiny |z| (series / Pade branch) and the branch boundary |z|^2 = 1e-7
    for r in (0.0, 1e-6, 1e-5, 1e-4, 3.0e-4, 3.2e-4):
        for ang in (0.0, 0.7, 1.5707963, 2.5, 3.14159):
            pts.append((r * np.cos(ang), r * np.sin(ang)))
    # seeded scatter, including large arguments met by the few-cycle pulses
    for _ in range(400):
        pts.append((rng.uniform(-30, 30), rng.uniform(0, 30)))
    for _ in range(100):
        pts.append((rng.uniform(-3, 3), rng.uniform(-0.5, 0.0)))
    return pts


def make_faddeeva():
    lib = C.CDLL(os.path.join(REF, "libref_specfunc.so"))
    lib.ref_faddeeva.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
    out = (C.c_double * 2)()
    rows = []
    for x, y in faddeeva_points():
        lib.ref_faddeeva(x, y, out)
        rows.append([x, y, out[0], out[1]])
    return {"source": "src/specfunc.cpp:57-157 Faddeeva_2, compiled in place",
            "columns": ["re_z", "im_z", "re_w", "im_w"], "rows": rows}


def pulse_queries():
    q = []
    # the reference test's own sampling (every 7th point of its 201 x 321 grid)
    xs = np.arange(-5, 5.01, 0.05)
    ts = np.arange(0, 8.0001, 0.025)
    idx = 0
    for x in xs:
        for t in ts:
            if idx % 7 == 0:
                for kind in "EXY":
                    q.append((kind, 1e6, 2.0, 2.0, 0.0, 1.0, x, 0.0, 0.0, t))
            idx += 1
    # workload pulses: T, width, om0, phase, eps as used by the benchmark configurations
    rng = np.random.default_rng(99)
    dx = 2 * np.pi / 20
    params = [(20.0, 15 * dx, 1.0, 0.0, 2.4666666666666668), (40.0, 60 * dx, 1.0, 0.0, 2.4666666666666668),
              (20.0, 15 * dx, 1.0, 0.25, 1.0), (3.0, 4 * dx, 2.0, 0.1, 1.5), (8.0, 2.0, 1.0, 0.0, 1.0)]
    for (T, w, om0, ph, eps) in params:
        for _ in range(60):
            x, y = rng.uniform(-3 * w, 3 * w, 2)
            z = rng.uniform(-20, 20)
            t = rng.uniform(-3 * T, 3 * T)
            for kind in "EXY":
                q.append((kind, T, w, om0, ph, eps, x, y, z, t))
        # the normalisation point of setParam: Efunc(0, 0, 0, 0)
        q.append(("E", T, w, om0, 0.0, eps, 0.0, 0.0, 0.0, 0.0))
    return q


def make_pulse():
    q = pulse_queries()
    lines = "\n".join(" ".join(repr(float(v)) if i else v for i, v in enumerate(row)) for row in q) + "\n"
    with tempfile.TemporaryDirectory() as tmp:  # the harness opens debug*.out in its cwd
        res = subprocess.run([os.path.join(REF, "refpulse")], input=lines, capture_output=True, text=True,
                             cwd=tmp, check=True)
    vals = [list(map(float, ln.split())) for ln in res.stdout.strip().splitlines()]
    assert len(vals) == len(q)
    rows = [list(r) + v for r, v in zip(q, vals)]
    return {"source": "test/testshortpulse.cpp Efunc/Bfunc (Efunc carries an extra 1/c), compiled in place",
            "columns": ["kind", "length", "width", "om0", "phase", "eps", "x", "y", "z", "t", "re", "im"],
            "rows": rows}


def main():
    with open(os.path.join(HERE, "faddeeva.json"), "w") as f:
        json.dump(make_faddeeva(), f)
    with open(os.path.join(HERE, "shortpulse.json"), "w") as f:
        json.dump(make_pulse(), f)


if __name__ == "__main__":
    main()
