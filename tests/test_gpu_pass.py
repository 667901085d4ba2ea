"""The single-pass leapfrog path (MPX_PATH_PASS) against the fused path and the oracle (GPU).

The pass kernel computes B of step n and E of step n + 1 with the fused kernels' arithmetic
term for term, so after any number of steps -- with the half step it owes settled whenever
the state is read -- every field, accumulator and CPML auxiliary equals the fused path's bit
for bit, in float64, FP32-bulk and FP32-res.  Grids have a partial last z wave and a partial
last row group, and the banded tile order runs with several band heights (tile_band), so the
recomputed neighbours of the kernel (x - 1, y - 1, lane 0's z - 1, the periodic wraps) lie on
tile and band boundaries.  The x-march kernel runs with several tile heights and segment
lengths, so its halo row, its z - 1 column and its recomputed segment-start plane lie on
partial tiles and segments.
"""
import numpy as np
import pytest

from mpulse_amd import abi, configs

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def pair(make, prec, band=0, tuning=None):
    da = make(prec, abi.PATH_FUSED)
    db = make(prec, abi.PATH_PASS)
    a, b = abi.Solver(da), abi.Solver(db)
    if band:
        b.tune("tile_band", band)
    for key, value in (tuning or {}).items():
        b.tune(key, value)
    return a, b


# the pass kernels and launch shapes: one cell per thread (two band heights), the x march with
# its default 7-row tile and 32-plane segments (one partial segment), 3-row tiles over 5-plane
# segments (a partial last segment), and one-plane segments (every plane a segment start)
KERNELS = {
    "cell": (0, {"pass_kernel": 0}),
    "cell_band3": (3, {"pass_kernel": 0}),
    "march": (0, {}),
    "march_r3_s5": (0, {"march_rows": 3, "march_seg": 5}),
    "march_r7_s1": (0, {"march_rows": 7, "march_seg": 1}),
}


def same_state(a, b, names=FIELDS + ("PsiR1x", "PsiI1z")):
    for n in names:
        x, y = a.get(n), b.get(n)
        assert np.array_equal(x, y), (n, float(np.max(np.abs(x - y))))


def pulse(n=37, nz=70, poles=((1.1, 4.0, 0.005),), d_cpml=6):
    def make(prec, path):
        d, _ = configs.plrc_kerr_pulse(n, d_cpml, 5.0, 6, 2e9, 40, d_src=10, precision=prec, path=path, nz=nz,
                                       poles=poles)
        d.ny = n + 6  # a partial last row group
        d.factories[1].tshift = 8.0
        return d
    return make


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_BULK, abi.PREC_FP32_RES])
@pytest.mark.parametrize("kernel", list(KERNELS))
def test_pass_equals_fused(prec, kernel):
    band, tuning = KERNELS[kernel]
    print(f"case: pass vs fused, precision {prec}, kernel {kernel}", flush=True)
    a, b = pair(pulse(), prec, band, tuning)
    with a, b:
        for n in (1, 3, 17, 9):  # settled between the calls by the reads
            a.step(n)
            b.step(n)
            same_state(a, b)
        for name, d in (("CPMLPsiBx", abi.NORTH), ("CPMLPsiBy", abi.UP), ("CPMLPsiEz", abi.EAST)):
            x, y = a.get_layer(name, d)[0], b.get_layer(name, d)[0]
            assert np.array_equal(x, y), name
        assert a.energy() == b.energy()
        assert np.array_equal(a.fetch("Ez"), b.fetch("Ez"))


def test_pass_two_poles_lin_core_and_plane_wave():
    """Two active poles with the linear core, and a tilted plane wave with background fields."""
    def make2(prec, path):
        d = pulse(n=30, nz=66, poles=((0.69, 25.0, 0.0), (0.87, 0.2, 0.01)))(prec, path)
        d.solver = abi.SOLVER_PLRC_LIN
        return d

    def make_plane(prec, path):
        d, _ = configs.plrc_kerr_pulse(28, 5, 5.0, 6, 1e9, 30, d_src=9, precision=prec, path=path, nz=40)
        pw = abi.factory_defaults(abi.FACTORY_PLANE_WAVE)
        pw.d = 9
        pw.k[0], pw.k[1], pw.k[2] = 0.3, 0.1, 1.0
        pw.h[0], pw.h[1] = 0.2, 0.5
        pw.h_bg[0], pw.e_bg[1] = 0.01, -0.02
        d.factories[1] = pw
        return d
    for make in (make2, make_plane):
        for prec in (abi.PREC_FP64, abi.PREC_FP32_RES):
            print(f"case: {make.__name__} precision {prec}", flush=True)
            a, b = pair(make, prec, band=2)
            with a, b:
                a.step(13)
                b.step(13)
                same_state(a, b, FIELDS)


def test_pass_cpml_1d_and_per_cell_eps():
    def make(prec, path):
        d, _ = configs.plrc_kerr_pulse(26, 6, 5.0, 6, 2e9, 30, d_src=10, precision=prec, path=path, nz=50)
        d.factories[0].kind = abi.FACTORY_CPML_1D
        d.factories[1].tshift = 8.0
        return d
    rng = np.random.default_rng(4)
    for prec in (abi.PREC_FP64, abi.PREC_FP32_BULK):
        print(f"case: cpml_border_1d + Eps, precision {prec}", flush=True)
        a, b = pair(make, prec, band=5)
        with a, b:
            eps = 1.0 + rng.random(a.shape)
            a.set("Eps", eps)
            b.set("Eps", eps)
            a.step(11)
            b.step(11)
            same_state(a, b, FIELDS)


def test_pass_set_field_mid_run():
    """A write between steps settles the owed half step first, then replaces the grid."""
    a, b = pair(pulse(n=24, nz=40), abi.PREC_FP64, band=1)
    rng = np.random.default_rng(8)
    with a, b:
        a.step(5)
        b.step(5)
        x = rng.standard_normal(a.shape)
        a.set("By", x)
        b.set("By", x)
        a.step(6)
        b.step(6)
        same_state(a, b)
        assert a.time == b.time == 11


def test_pass_against_oracle():
    """64^3, one pole + Kerr, 10-cell CPML, short pulse, 400 steps (float64): every field
    within 1e-12 of the oracle."""
    from oracle_lib import Oracle
    d, meta = configs.config_small(precision=abi.PREC_FP64, path=abi.PATH_PASS)
    steps = meta["steps"]
    with abi.Solver(d) as s:
        s.step(steps)
        got = {f: s.get(f) for f in FIELDS}
    o = Oracle(d, omp=True)
    o.step(steps)
    for f in FIELDS:
        r = o.get(f)
        rel = np.linalg.norm(got[f] - r) / np.linalg.norm(r)
        print(f, f"{rel:.2e}")
        assert rel < 1e-12, f
    o.close()


def test_pass_probes_fall_back_to_fused_halves():
    d, meta = configs.config_small(precision=abi.PREC_FP64, path=abi.PATH_PASS)
    d2, _ = configs.config_small(precision=abi.PREC_FP64, path=abi.PATH_FUSED)
    with abi.Solver(d) as a, abi.Solver(d2) as b:
        a.step(4)
        b.step(4)
        for s in (a, b):
            s.set_probes(meta["probes"], [0] * len(meta["probes"]), 32)
        a.step(20)
        b.step(20)
        assert np.array_equal(a.get_probes(), b.get_probes())
        same_state(a, b, FIELDS)


def test_pass_gauss_inject_equals_fused():
    """gauss_inject over three sample blocks (src/gaussinject.cpp): the H sheet of step n and
    the E sheet of step n + 1 are scheduled before the one launch; fields equal the fused path."""
    import numpy as np_

    def make(prec, path):
        d, _ = configs.plrc_kerr_pulse(14, 4, 10.0, 4, 1e9, 60, d_src=6, nz=20, precision=prec, path=path)
        f = abi.factory_defaults(abi.FACTORY_GAUSS_INJECT)
        f.d, f.gsize_t, f.length, f.width, f.k0, f.gshift, f.gtime, f.amp, f.eps, f.toff = \
            6, 48, 1.2, 1.1, 2 * np_.pi, 0.4, 0.3, 0.8, 1.3, 2
        d.factories[1] = f
        return d
    for prec in (abi.PREC_FP64, abi.PREC_FP32_RES):
        print(f"case: gauss_inject pass vs fused, precision {prec}", flush=True)
        a, b = pair(make, prec)
        with a, b:
            for n in (7, 40, 13):
                a.step(n)
                b.step(n)
                same_state(a, b, FIELDS)


def test_pass_with_the_drivers_set_time_every_step():
    """The host driver sets the time index before every step (src/incsource.t:46); with the
    solver's own time that changes nothing, so the pending half step stays pending and the
    fields equal the fused path's."""
    a, b = pair(pulse(n=24, nz=40), abi.PREC_FP32_RES)
    with a, b:
        for t in range(9):
            a.set_time(t)
            b.set_time(t)
            a.step(1)
            b.step(1)
        same_state(a, b)


def test_pass_focused_pulse_two_poles_equals_fused():
    """A focused pulse (NA 0.4) into the two-pole silica medium + Kerr behind a CPML: the pass
    path equals the fused path in float64 and FP32-res (the focused-pulse workload's form)."""
    def make(prec, path):
        d = abi.desc_defaults()
        d.nx, d.ny, d.nz = 40, 36, 70
        d.dx = d.dy = d.dz = configs.DX
        d.dt = configs.DT
        d.solver = abi.SOLVER_PLRC_NONLIN
        d.precision = prec
        d.path = path
        d.eps, d.chi = 1.0, 2e-3
        for q, (deps, om2, delta) in enumerate(configs.SILICA_POLES):
            d.l_eps[q], d.l_om2[q], d.l_delta[q] = deps, om2, delta
        cp = abi.factory_defaults(abi.FACTORY_CPML)
        cp.d = 6
        cp.sigma_max = configs.sigma_max(6)
        abi.add_factory(d, cp)
        f = abi.factory_defaults(abi.FACTORY_FOCUSED_PULSE)
        f.d, f.width, f.om0, f.length, f.tshift, f.amp = 9, 5.0, 1.0, 4.0, 6.0, 2.0
        f.eps = configs.medium_eps_at_laser(1.0, configs.SILICA_POLES)
        f.zshift = -6 * configs.DX
        f.phase = 0.3
        abi.add_factory(d, f)
        return d
    for prec in (abi.PREC_FP64, abi.PREC_FP32_RES):
        print(f"case: focused pulse pass vs fused, precision {prec}", flush=True)
        a, b = pair(make, prec)
        with a, b:
            for n in (31, 60):
                a.step(n)
                b.step(n)
                same_state(a, b, FIELDS + ("PsiR1x", "PsiI2z"))
            assert np.abs(a.get("Ex")).max() > 0.1


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_FP32_RES])
def test_pass_xy_periodic_equals_fused(prec):
    """single_xy_periodic (src/periodic_bound.cpp:75-140): z is not wrapped, B is updated at
    k = 0 too, the z ghosts of E and B keep what is stored in them.  A set state with non-zero
    z ghosts, several grids with partial z tiles, the pass path equal to the fused path."""
    rng = np.random.default_rng(11)
    for nz in (70, 64, 20):
        def make(p, path, nz=nz):
            d = pulse(n=26, nz=nz)(p, path)
            d.boundary = abi.BOUNDARY_XY_PERIODIC
            return d
        print(f"case: xy-periodic pass vs fused, precision {prec}, nz {nz}", flush=True)
        a, b = pair(make, prec)
        with a, b:
            a.step(5)
            b.step(5)
            same_state(a, b, FIELDS)
            for f in ("Ex", "By"):
                x = a.get(f)
                x[:, :, 0] = 0.01 * rng.standard_normal(x[:, :, 0].shape)
                x[:, :, -1] = 0.01 * rng.standard_normal(x[:, :, -1].shape)
                a.set(f, x)
                b.set(f, x)
            for n in (3, 12, 7):
                a.step(n)
                b.step(n)
                same_state(a, b, FIELDS + ("PsiR1x",))


@pytest.mark.parametrize("order", [(0, 1, 2), (1, 0, 2), (0, 2, 1)])
def test_pass_plasma_equals_fused(order):
    """Multiphoton ionisation, the resistivity Sigma and a plasma current (src/plasmadensity.cpp,
    src/plasmacurrent.cpp) in the one-pole Kerr medium, with the plasma current listed after the
    CPML (between the current groups), first, and last: fields, the plasma grids and Sigma equal
    the fused path's, in float64 and FP32-res."""
    def make(prec, path):
        d, _ = configs.plrc_kerr_pulse(30, 6, 6.0, 6, 4e9, 0, d_src=9, precision=prec, path=path, nz=70,
                                       poles=((1.1, 4.0, 0.005),))
        d.ny = 33  # a partial last row group
        d.factories[1].tshift = 12.0
        f = abi.factory_defaults(abi.FACTORY_PLASMA_CURRENT)
        f.em, f.gamma, f.rho_bg = 0.8, 0.2, 0.01
        facs = [abi.Factory.from_buffer_copy(d.factories[0]), f, abi.Factory.from_buffer_copy(d.factories[1])]
        for q, n in enumerate(order):
            d.factories[q] = facs[n]
        d.n_factories = 3
        d.plasma, d.plasma_k, d.plasma_nu, d.plasma_mpa, d.plasma_wion = 1, 3, 0.05, 0.003, 0.5
        return d
    for prec in (abi.PREC_FP64, abi.PREC_FP32_RES):
        print(f"case: plasma pass vs fused, factory order {order}, precision {prec}", flush=True)
        a, b = pair(make, prec)
        with a, b:
            for n in (20, 35):
                a.step(n)
                b.step(n)
                same_state(a, b, FIELDS + ("PlasmaDensity", "PlasmaJx", "PlasmaJz", "Sigma"))
            assert np.abs(a.get("PlasmaDensity")).max() > 0
