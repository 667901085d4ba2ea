"""Workload descriptors in the reference's normalised units (eps0 = mu0 = c = 1).

Unit mapping (one choice, used everywhere in this package):
  length unit  L0 = lambda / (2 pi) with lambda = 800 nm, so the laser frequency omega_L = 1;
  time unit    t0 = L0 / c = 0.42471 fs;
  field unit   E0 = 1e9 V/m, so chi3 [m^2/V^2] maps to chi = chi3 * E0^2 and a peak field of
               1e9 V/m is amp = 1.
  grid         dx = dy = dz = lambda/20 = 2 pi / 20, dt = 0.5 dx / sqrt(3).
  Lorentz pole chi(w) = deps Om^2 / (Om^2 - w^2 - 2 i delta w), the form of the PLRC solver's
               kernel gamma e^{-delta t} sin(beta t); a damping "gamma" quoted in the
               Om^2 - w^2 - i gamma w convention is delta = gamma / 2.
  pulse        the FWHM is the intensity FWHM of the exp(-t^2/T^2) field envelope,
               T = FWHM / sqrt(2 ln 2); the peak arrives at the source plane at TShift = 2 T;
               the focus sits on the source plane (ZShift = (d - c_z) dz); the pulse's own
               eps is the medium's eps at omega_L (the real part of the Lorentz response).
  CPML         sigma_max = -(m + 1) ln(R0) / (2 d dx) with m = 3 (the cubic grading of the
               reference), kappaMax and aMax at the reference defaults 15 and 0.25.
"""
import math

import numpy as np

from . import abi

LAMBDA = 800e-9
C0 = 299792458.0
L0 = LAMBDA / (2 * math.pi)
T0 = L0 / C0
E0 = 1e9
DX = 2 * math.pi / 20
DT = 0.5 * DX / math.sqrt(3)


def fs_to_t(fs):
    return fs * 1e-15 / T0


def sigma_max(d, r0=1e-6, m=3, dx=DX):
    return -(m + 1) * math.log(r0) / (2 * d * dx)


def splitmix64(seed):
    """SplitMix64 (S. Vigna): a language-independent stream of 64-bit integers."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def probe_cells(n, d_cpml, sheet_planes, count=8, seed=0):
    """count interior cells drawn uniformly in [d+1, N-d]^3 (SplitMix64, seed), avoiding the
    source-sheet planes."""
    rng = splitmix64(seed)
    lo, hi = d_cpml + 1, n - d_cpml
    cells = []
    while len(cells) < count:
        c = [lo + int(next(rng) % (hi - lo + 1)) for _ in range(3)]
        if c[2] in sheet_planes or c in cells:
            continue
        cells.append(c)
    return np.array(cells, dtype=np.int32)


def medium_eps_at_laser(eps_inf, poles):
    e = eps_inf
    for deps, om2, _delta in poles:
        e += deps * om2 / (om2 - 1.0)
    return e


def plrc_kerr_pulse(n, d_cpml, fwhm_fs, waist_cells, peak_field, steps, chi3_si=2e-22, d_src=12,
                    precision=abi.PREC_FP64, path=abi.PATH_FUSED, nz=None, boundary=abi.BOUNDARY_PERIODIC,
                    poles=((1.1, 4.0, 0.005),)):
    """Single-pole Lorentz + Kerr medium, CPML on every face, short pulse injected upward from
    the plane k = d_src (E sheet at d_src, H sheet at d_src - 1)."""
    nz = n if nz is None else nz
    d = abi.desc_defaults()
    d.nx, d.ny, d.nz = n, n, nz
    d.dx = d.dy = d.dz = DX
    d.dt = DT
    d.solver = abi.SOLVER_PLRC_NONLIN
    d.boundary = boundary
    d.precision = precision
    d.path = path
    d.eps = 1.0
    d.chi = chi3_si * E0 * E0
    for q in range(3):
        d.l_eps[q], d.l_om2[q], d.l_delta[q] = 0.0, 1.0, 0.0
    for q, (deps, om2, delta) in enumerate(poles):
        d.l_eps[q], d.l_om2[q], d.l_delta[q] = deps, om2, delta
    cp = abi.factory_defaults(abi.FACTORY_CPML)
    cp.d = d_cpml
    cp.sigma_max = sigma_max(d_cpml)
    abi.add_factory(d, cp)
    sp = abi.factory_defaults(abi.FACTORY_SHORT_PULSE)
    T = fs_to_t(fwhm_fs) / math.sqrt(2 * math.log(2))
    sp.d = d_src
    sp.length = T
    sp.width = waist_cells * DX
    sp.om0 = 1.0
    sp.amp = peak_field / E0
    sp.eps = medium_eps_at_laser(1.0, poles)
    sp.tshift = 2 * T
    sp.zshift = (d_src - 0.5 * (nz + 1)) * DX
    sp.phase = 0.0
    abi.add_factory(d, sp)
    meta = {"steps": steps, "d_cpml": d_cpml, "d_src": d_src,
            "probes": probe_cells(n, d_cpml, (d_src - 1, d_src), seed=0)}
    return d, meta


def config_small(precision=abi.PREC_FP64, path=abi.PATH_FUSED):
    """64^3, one pole + Kerr, 10-cell CPML, 10 fs pulse of waist 15 cells at 1e9 V/m, 400 steps."""
    return plrc_kerr_pulse(64, 10, 10.0, 15, 1e9, 400, precision=precision, path=path)


def config_roofline(precision=abi.PREC_FP32_BULK, path=abi.PATH_FUSED, n=512, steps=1000):
    """512^3, one pole + Kerr, 16-cell CPML, 20 fs pulse of waist 60 cells at 5e9 V/m, 1000 steps."""
    return plrc_kerr_pulse(n, 16, 20.0, 60, 5e9, steps, precision=precision, path=path)


def desc_summary(d):
    fac = [d.factories[q] for q in range(d.n_factories)]
    return {"grid": [d.nx, d.ny, d.nz], "dx": d.dx, "dt": d.dt, "solver": d.solver, "precision": d.precision,
            "path": d.path, "chi": d.chi, "poles": [(d.l_eps[q], d.l_om2[q], d.l_delta[q]) for q in range(3)],
            "factories": [f.kind for f in fac]}


# fused-silica-like medium of the self-focusing workload: two undamped Lorentz poles
# (Sellmeier terms at 0.068 um and 9.9 um), Kerr n2 = 2.5e-20 m^2/W
SILICA_POLES = ((0.69, (0.8 / 0.068) ** 2, 0.0), (0.87, (0.8 / 9.9) ** 2, 0.0))
EPS0_SI = 8.8541878128e-12


def chi3_from_n2(n2, n0):
    """chi3 [m^2/V^2] of the instantaneous Kerr term chi E^2 E for intensity index n2 [m^2/W]:
    n2 = 3 chi3 / (4 n0^2 eps0 c)."""
    return 4.0 * n0 * n0 * EPS0_SI * C0 * n2 / 3.0


def field_from_intensity(intensity, n0):
    """peak field [V/m] of a linearly polarised wave of cycle-averaged intensity [W/m^2]:
    I = n0 eps0 c E^2 / 2."""
    return math.sqrt(2.0 * intensity / (n0 * EPS0_SI * C0))


def config_focus(precision=abi.PREC_FP32_BULK, path=abi.PATH_FUSED, nx=384, nz=768, steps=2000, d_cpml=16,
                 d_src=20, na=0.1, fwhm_fs=30.0, intensity=5e16, n2=2.5e-20, poles=SILICA_POLES):
    """Focused-pulse self-focusing: nx^2 x nz, two Lorentz poles, Kerr from n2, focused pulse of
    numerical aperture na (waist w0 = lambda / (pi NA) = 2 / NA length units) whose focus lies
    on the middle z plane, FWHM 30 fs, peak intensity 5e16 W/m^2 at the focus, 16-cell CPML."""
    eps_l = medium_eps_at_laser(1.0, poles)
    n0 = math.sqrt(eps_l)
    d = abi.desc_defaults()
    d.nx, d.ny, d.nz = nx, nx, nz
    d.dx = d.dy = d.dz = DX
    d.dt = DT
    d.solver = abi.SOLVER_PLRC_NONLIN
    d.precision = precision
    d.path = path
    d.eps = 1.0
    d.chi = chi3_from_n2(n2, n0) * E0 * E0
    for q in range(3):
        d.l_eps[q], d.l_om2[q], d.l_delta[q] = 0.0, 1.0, 0.0
    for q, (deps, om2, delta) in enumerate(poles):
        d.l_eps[q], d.l_om2[q], d.l_delta[q] = deps, om2, delta
    cp = abi.factory_defaults(abi.FACTORY_CPML)
    cp.d = d_cpml
    cp.sigma_max = sigma_max(d_cpml)
    abi.add_factory(d, cp)
    f = abi.factory_defaults(abi.FACTORY_FOCUSED_PULSE)
    T = fs_to_t(fwhm_fs) / math.sqrt(2 * math.log(2))
    f.d = d_src
    f.length = T
    f.width = 2.0 / na
    f.om0 = 1.0
    f.amp = field_from_intensity(intensity, n0) / E0
    f.eps = eps_l
    f.tshift = 2 * T
    f.zshift = 0.0  # the focus lies on the middle z plane
    f.phase = 0.0
    abi.add_factory(d, f)
    meta = {"steps": steps, "d_cpml": d_cpml, "d_src": d_src, "n0": n0,
            "probes": probe_cells(min(nx, nz), d_cpml, (d_src - 1, d_src), seed=0)}
    return d, meta


def config_sweep(nruns=64, steps=500, seed=1, precision=abi.PREC_FP64, path=abi.PATH_SWEEP, nmin=48, nmax=160):
    """Batched parameter sweep: nruns independent cubes with sizes uniform in [nmin, nmax]
    (SplitMix64, seed), the one-pole Lorentz + Kerr medium, chi3 log-uniform in
    [1e-23, 1e-21] m^2/V^2, peak field uniform in [2e8, 5e9] V/m, 10-cell CPML, the 10 fs pulse
    from the plane k = 12 with waist min(15, n/4) cells, `steps` steps each.  Draw order per
    run: size, chi3, peak field."""
    rng = splitmix64(seed)
    out = []
    for _ in range(nruns):
        n = nmin + int(next(rng) % (nmax - nmin + 1))
        u = next(rng) / 2.0 ** 64
        chi3 = 10.0 ** (-23.0 + 2.0 * u)
        u = next(rng) / 2.0 ** 64
        peak = 2e8 + (5e9 - 2e8) * u
        d, meta = plrc_kerr_pulse(n, 10, 10.0, min(15, n // 4), peak, steps, chi3_si=chi3, d_src=12,
                                  precision=precision, path=path)
        meta.update({"n": n, "chi3": chi3, "peak_field": peak})
        out.append((d, meta))
    return out


def splitmix64_array(seed, n):
    """The first n outputs of SplitMix64(seed) as a numpy uint64 array (the generator's state
    advances by a constant, so output i depends on seed + (i + 1) * gamma only)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def smooth_permittivity(shape, seed=2, sigma=6.0, lo=1.0, hi=2.25):
    """N(0,1) noise (Box-Muller on SplitMix64(seed) uniforms, cell order x, y, z with z fastest)
    filtered by a periodic Gaussian of sigma cells (in Fourier space) and mapped affinely onto
    [lo, hi]."""
    n = int(np.prod(shape))
    u = (splitmix64_array(seed, 2 * n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    noise = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    f = noise.reshape(shape)
    k2 = sum(np.meshgrid(*[np.fft.fftfreq(m) ** 2 for m in shape], indexing="ij"))
    f = np.real(np.fft.ifftn(np.fft.fftn(f) * np.exp(-2.0 * np.pi ** 2 * sigma ** 2 * k2)))
    f = (f - f.min()) / (f.max() - f.min())
    return lo + (hi - lo) * f


def config_adjoint(nx=128, nz=256, steps=2000, d_cpml=10, d_src=16, k_obj=232, chi3_si=1e-22,
                   checkpoint_every=50, fd_cells=16):
    """Adjoint gradient: nx^2 x nz float64, no Lorentz pole, per-cell permittivity from
    smooth_permittivity (seed 2, sigma 6 cells, [1, 2.25]), Kerr chi3 = 1e-22 m^2/V^2, 10-cell
    CPML, the 10 fs / 15-cell-waist pulse at 1e9 V/m injected upward from the plane k = 16,
    objective J = dx dy dt sum_steps sum_ij |E(i, j, 232)|^2, checkpoints every 50 steps.
    Returns (desc, full-storage Eps, meta) with meta holding the objective, the
    finite-difference cells (SplitMix64 seed 2, inside the pulse's path between the source and
    the objective plane) and the step h = 1e-4."""
    d, _ = plrc_kerr_pulse(nx, d_cpml, 10.0, 15, 1e9, steps, chi3_si=chi3_si, d_src=d_src, nz=nz,
                           precision=abi.PREC_FP64, path=abi.PATH_FUSED, poles=())
    eps = smooth_permittivity((nx, nx, nz))
    full = np.full((nx + 2, nx + 2, nz + 2), 1.0)
    full[1:-1, 1:-1, 1:-1] = eps
    d.factories[1].eps = float(eps[:, :, d_src - 1].mean())
    rng = splitmix64(2)
    cells = []
    lo, hi = nx // 2 - 12, nx // 2 + 12
    while len(cells) < fd_cells:
        c = (lo + int(next(rng) % (hi - lo + 1)), lo + int(next(rng) % (hi - lo + 1)),
             d_src + 8 + int(next(rng) % (k_obj - d_src - 16)))
        if c not in cells:
            cells.append(c)
    meta = {"steps": steps, "k_obj": k_obj, "weight": d.dx * d.dy * d.dt, "checkpoint_every": checkpoint_every,
            "fd_cells": cells, "fd_h": 1e-4}
    return d, full, meta
