/*
 * mpulse_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C, float64 restatement of the MPulse time-stepping path, written to follow the
 * reference source loop by loop and operation by operation, so that it can serve as the
 * checker for the HIP kernels.  Each function names the reference lines it restates
 * (paths relative to the MPulse tree).  Nothing in the product links or calls this file;
 * only tests/, the smoke entry and the CPU-baseline leg of bench.py load it.
 *
 * Storage conventions follow Storage/Globals: interior 1..N, storage 0..N+1
 * (src/globals.cpp:44-45, src/storage.cpp:58-76); every grid starts at zero.
 * Grids are stored z fastest.  Complex arithmetic uses C99 complex double, which GCC
 * lowers exactly as libstdc++ lowers std::complex<double> (inline products, __divdc3
 * divisions, glibc cexp).
 */
#include "mpulse_oracle.h"

/* ORACLE_MUTANT (test builds only, `make mutants`): a planted mistake, so that the tests can
   show that each of their physics checks fails when the part of the path it pins is wrong.
   0 = the oracle as it is.  1: the H sheet of a low face at `distance` (no decrement);
   2: no half-cell offset of the H-side CPML profile; 3: the outermost H-side CPML cell clamped
   to x = 1; 4: the CPML b/c tables one cell toward the interior; 5: sign of the E-side CPML
   psi0 update flipped; 6: the low-side kappaE line one cell toward the interior; 7: fdtd-nl
   without the chi E^2 term of D; 8: fdtd-nl with the curl's sign flipped; 9: fdtd-disp with
   b = dt^2 Om^2 eps instead of eps2; 10: fdtd-disp without the previous-P term of E. */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#undef I /* the complex.h macro; local constants named I are used below */

typedef double complex cplx;
#define CX(re, im) CMPLX((re), (im))

/* Direction enum of src/mpulse.h:28 */
enum { NORTH = 0, SOUTH = 1, WEST = 2, EAST = 3, UP = 4, DOWN = 5 };

/* ------------------------------------------------------------------------------------ */
/* grids                                                                                */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int lo[3], hi[3];
    int n[3];
    double *v;
} grid_t;

static int grid_alloc(grid_t *g, const int lo[3], const int hi[3])
{
    size_t cnt = 1;
    for (int d = 0; d < 3; ++d) {
        g->lo[d] = lo[d];
        g->hi[d] = hi[d];
        g->n[d] = hi[d] - lo[d] + 1;
        cnt *= (size_t)g->n[d];
    }
    g->v = (double *)calloc(cnt, sizeof(double));
    return g->v != NULL;
}

static inline size_t gidx(const grid_t *g, int i, int j, int k)
{
    return ((size_t)(i - g->lo[0]) * g->n[1] + (size_t)(j - g->lo[1])) * g->n[2] + (size_t)(k - g->lo[2]);
}
#define AT(g, i, j, k) ((g)->v[gidx((g), (i), (j), (k))])

static size_t grid_count(const grid_t *g) { return (size_t)g->n[0] * g->n[1] * g->n[2]; }

static void grid_fill(grid_t *g, double val)
{
    size_t n = grid_count(g);
    for (size_t q = 0; q < n; ++q) g->v[q] = val;
}

/* Storage rounding (not part of the reference): the value a field word holds after it is
   stored in a narrower form.  Mode 1: float32 (round to nearest even).  Mode 2: float32 plus a
   10-bit two's-complement residual r = round((bits(x) - bits(hi)) / 2^19), the difference of
   the float64 bit patterns of x and of hi = fl32(x), clamped to [-512, 511] and added back to
   the bit pattern of hi (in the binade of hi a unit of r is 2^-10 of the float32 word's ulp);
   zero and denormal words keep no residual.  The checker uses it to show what a storage form
   costs in accuracy against float64 storage. */
static double store_round(double x, int mode)
{
    float hi = (float)x;
    if (mode == 1 || !isfinite(hi)) return (double)hi;
    union { float f; unsigned u; } b = {hi};
    if (((b.u >> 23) & 0xffu) == 0) return (double)hi;
    union { double d; unsigned long long u; } X = {x}, H = {(double)hi};
    long long d = (long long)X.u - (long long)H.u; /* same sign bit: the magnitude difference */
    long long r = (d + (1LL << 18)) / (1LL << 19);
    if (d + (1LL << 18) < 0 && (d + (1LL << 18)) % (1LL << 19)) r -= 1; /* floor division */
    if (r > 511) r = 511;
    if (r < -512) r = -512;
    H.u += (unsigned long long)(r * (1LL << 19));
    return H.d;
}

static void grid_round(grid_t *g, int mode)
{
    if (!mode) return;
    size_t n = grid_count(g);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < n; ++q) g->v[q] = store_round(g->v[q], mode);
}

typedef struct {
    int lo, hi;
    double *v;
} line_t;

static int line_alloc(line_t *l, int lo, int hi, double def)
{
    l->lo = lo;
    l->hi = hi;
    l->v = (double *)malloc(sizeof(double) * (size_t)(hi - lo + 1));
    if (!l->v) return 0;
    for (int q = 0; q <= hi - lo; ++q) l->v[q] = def;
    return 1;
}
#define LN(l, i) ((l)->v[(i) - (l)->lo])

/* ------------------------------------------------------------------------------------ */
/* Faddeeva function w(z) = exp(-z^2) erfc(-iz)                                         */
/* ------------------------------------------------------------------------------------ */

/* Restates Faddeeva_2 (src/specfunc.cpp:57-157): the trapezoid-rule evaluation of
 * Matta & Reichel (Math. Comp. 25, 1971) in Gammel's arrangement, step h = 0.5, or
 * h = 0.53 when z lies near a pole of the h = 0.5 rule (|Im z| < 0.01, |Re z| < 6.01 and
 * 2|Re z| within 0.02 of an integer), with the pole correction for Im z < 2 pi.
 * The node tables hold n_k = (k h)^2 and e_k = exp(-n_k), k = 1..12, computed here.
 * For |z|^2 < 1e-7 the reference uses a Pade form accurate to ~1e-11; this restatement
 * uses the Maclaurin series w(z) = sum_n (iz)^n / Gamma(n/2 + 1), which is exact to
 * rounding there, so the two agree to the reference's own stated accuracy. */
static double fad_n[2][12], fad_e[2][12];
static int fad_ready = 0;

static void fad_init(void)
{
    if (fad_ready) return;
    const double hs[2] = {0.5, 0.53};
    for (int t = 0; t < 2; ++t)
        for (int q = 0; q < 12; ++q) {
            double kh = hs[t] * (double)(q + 1);
            fad_n[t][q] = kh * kh;
            fad_e[t][q] = exp(-fad_n[t][q]);
        }
    fad_ready = 1;
}

static cplx faddeeva(cplx z)
{
    double x = creal(z), y = cimag(z);
    double s = x * x + y * y;
    if (s < 1e-7) {
        /* sum_{n=0..7} (iz)^n / Gamma(n/2+1) */
        static const double inv_gamma[8] = {
            1.0, 1.1283791670955126, 1.0, 0.75225277806367508, 0.5, 0.30090111122547003,
            0.16666666666666666, 0.085971746064420009};
        cplx iz = CX(-y, x);
        cplx term = CX(1.0, 0.0), acc = CX(0.0, 0.0);
        for (int n = 0; n < 8; ++n) {
            acc += inv_gamma[n] * term;
            term *= iz;
        }
        return acc;
    }
    if (!fad_ready) fad_init();
    int tab = 0;
    double r = M_1_PI * 0.5;
    if (fabs(y) < 0.01 && fabs(x) < 6.01) {
        double ip;
        double fr = modf(2 * fabs(x), &ip);
        if (fr < 0.02 || fr > 0.98) {
            tab = 1;
            r = M_1_PI * 0.53;
        }
    }
    const double *n = fad_n[tab], *e = fad_e[tab];
    double d = (y - x) * (y + x);
    double f = 4 * x * x * y * y;
    double g = 0.0, h = 0.0;
    for (int q = 0; q < 12; ++q) {
        double t = d + n[q];
        double u = e[q] / (t * t + f);
        g += (s + n[q]) * u;
        h += (s - n[q]) * u;
    }
    double u = 1 / s;
    cplx c = r * CX(y * (u + 2.0 * g), x * (u + 2.0 * h));
    if (y < 2.0 * M_PI) {
        double s2 = 2.0 / r;
        double t = s2 * x;
        double uu = s2 * y;
        double sn = sin(t);
        double cs = cos(t);
        double ff = exp(-uu) - cs;
        double gg = 2.0 * exp(d - uu) / (sn * sn + ff * ff);
        double u2 = 2.0 * x * y;
        double h2 = cos(u2);
        double t2 = sin(u2);
        c += gg * CX(h2 * ff - t2 * sn, -(h2 * sn + t2 * ff));
    }
    return c;
}

void or_faddeeva(double re, double im, double out[2])
{
    cplx w = faddeeva(CX(re, im));
    out[0] = creal(w);
    out[1] = cimag(w);
}

/* ------------------------------------------------------------------------------------ */
/* Short pulse source function (paraxial few-cycle Gaussian pulse)                       */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    double DX, DY, DZ, DT;
    double lightspeed;
    double length, width, om0, TShift, ZShift, Phase;
    double ZRl;
    double centrex, centrey, centrez;
    double amp, eps;
} pulse_t;

/* ShortPulseInjectSourceFunc::Efunc, src/shortpulsefunctions.cpp:28-94 */
static cplx pulse_efunc(const pulse_t *P, double x, double y, double z, double t)
{
    const cplx I = CX(0, 1), I2 = CX(0, 2), nI = CX(0, -1);
    double c = P->lightspeed;
    double p = P->Phase;
    double T = P->length;
    double T2 = T * T;
    double T4 = T2 * T2;
    double tb = 0.5 * P->om0 * T;
    double r2 = x * x + y * y;
    double ZRl = P->ZRl;

    cplx q = CX(z, ZRl);
    cplx rT = ZRl / q;
    cplx rpiT = sqrt(M_PI) * rT;
    cplx rpT = -ZRl / (q * q);
    cplx eiphi = cexp((2.0 * I) * p);
    cplx tpT = t - z / c - r2 / ((2. * c) * q);
    cplx tpT2 = tpT * tpT;
    cplx tpzT = -1 / c - r2 / (((2. * c) * q) * q);
    cplx an = nI * tb + tpT / T;
    cplx ap = I * tb + tpT / T;

    cplx A = cexp(nI * p) / (rpiT * T4);
    cplx B = (rT * T) * ((((I * (1.0 + eiphi)) * rpT) * T2) -
                         (rT * (T * ((nI - 2 * tb) + eiphi * (nI + 2 * tb)) + (I2 * (1.0 + eiphi)) * tpT)) * tpzT);
    cplx C = (((I * eiphi) * rpiT) * (T * tb + I * tpT)) *
             ((rpT * T2 + ((((I2 * rT) * T) * tb) * tpzT)) - ((2.0 * rT) * tpT) * tpzT);
    cplx D = (rpiT * (T * tb - I * tpT)) * ((I * rpT) * T2 + ((2.0 * rT) * (T * tb - I * tpT)) * tpzT);

    cplx S = 0;
    if (cimag(an) < 0) {
        S += ((2.0 * A) * C) * cexp(((((2.0 * I) * tpT) * tb) / T) - tpT2 / T2);
        an = -an;
        C = -C;
    }
    if (cimag(ap) < 0) {
        S += ((2.0 * A) * D) * cexp(((((-2.0 * I) * tpT) * tb) / T) - tpT2 / T2);
        ap = -ap;
        D = -D;
    }
    cplx Ef = S + (A * exp(-tb * tb)) * ((B + C * faddeeva(an)) - D * faddeeva(ap));
    return (-I) * Ef;
}

/* ShortPulseInjectSourceFunc::Bfunc, src/shortpulsefunctions.cpp:96-226.
 * YComp is fixed to 0 by setParam (src/shortpulseinject.cpp:96), so the YComp terms of the
 * return statements vanish and the C-group (used only through YComp) is not needed. */
static cplx pulse_bfunc(const pulse_t *P, double x, double y, double z, double t, int bx)
{
    const cplx I = CX(0, 1), I2 = CX(0, 2), nI = CX(0, -1);
    double c = P->lightspeed;
    double p = P->Phase;
    double T = P->length;
    double T2 = T * T;
    double T4 = T2 * T2;
    double tb = 0.5 * P->om0 * T;
    double r2 = x * x + y * y;
    double ZRl = P->ZRl;
    double sqpi = sqrt(M_PI);

    cplx q = CX(z, ZRl);
    cplx rT = ZRl / q;
    cplx rT2 = rT * rT;
    cplx rT4T = rT * T4;
    cplx rpiT = sqpi * rT;
    cplx rpT = -ZRl / (q * q);
    cplx rppT = (2. * ZRl) / ((q * q) * q);
    cplx eiphi = cexp((2.0 * I) * p);
    cplx tpT = t - z / c - r2 / ((2. * c) * q);
    cplx tpT2 = tpT * tpT;
    cplx tpzT = -1 / c + r2 / (((2. * c) * q) * q);
    cplx tpzT2 = tpzT * tpzT;
    cplx tpxT = -x / (c * q);
    cplx tpyT = -y / (c * q);
    cplx tpxyT = 0.0;
    cplx tpxxT = -1. / (c * q);
    cplx tpzzT = -r2 / (((c * q) * q) * q);
    cplx tpxT2 = tpxT * tpxT;
    cplx aip = T * tb + I * tpT;
    cplx ain = T * tb - I * tpT;
    cplx aip2 = aip * aip;
    cplx ain2 = ain * ain;
    cplx an = nI * tb + tpT / T;
    cplx ap = I * tb + tpT / T;

    cplx bracket = T * ((nI - 2 * tb) + eiphi * (nI + 2 * tb)) + (I2 * (1.0 + eiphi)) * tpT;
    cplx e2 = cexp(((((2.0 * I) * tpT) * tb) / T) - tpT2 / T2);
    cplx e2m = cexp(((((-2.0 * I) * tpT) * tb) / T) - tpT2 / T2);
    int flip_n = cimag(an) < 0;
    int flip_p = cimag(ap) < 0;
    if (flip_n) an = -an;
    if (flip_p) ap = -ap;
    cplx wn = faddeeva(an);
    cplx wp = faddeeva(ap);

    if (bx) {
        /* A group: BfA */
        cplx AA = cexp(nI * p) / (sqpi * T4);
        cplx AB = ((((-tpxT) * tpyT) * rT) * T) * bracket +
                  (((((I * T2) * tpxyT) * (1.0 + eiphi)) * rT) * T);
        cplx AC = (((((-tpxT) * tpyT) * 2.0) * eiphi) * rpiT) * aip2 +
                  ((((I * T2) * tpxyT) * eiphi) * rpiT) * aip;
        cplx AD = ((((-tpxT) * tpyT) * 2.0) * rpiT) * ain2 +
                  (((I * T2) * tpxyT) * rpiT) * (-(T * tb) + I * tpT);
        cplx AS = 0;
        if (flip_n) {
            AS += ((2.0 * AA) * AC) * e2;
            AC = -AC;
        }
        if (flip_p) {
            AS += ((2.0 * AA) * AD) * e2m;
            AD = -AD;
        }
        cplx BfA = AS + (AA * exp(-tb * tb)) * ((AB + AC * wn) - AD * wp);
        return I * BfA;
    }
    /* B group: BfB */
    cplx BA = cexp(nI * p) / (sqpi * T4);
    cplx BB = (-tpxT2) * ((rT * T) * bracket) + (((((I * T2) * tpxxT) * (1.0 + eiphi)) * rT) * T);
    cplx BC = ((((-tpxT2) * 2.0) * eiphi) * rpiT) * aip2 + ((((I * T2) * tpxxT) * eiphi) * rpiT) * aip;
    cplx BD = (((-tpxT2) * 2.0) * rpiT) * ain2 + (((I * T2) * tpxxT) * rpiT) * (-(T * tb) + I * tpT);
    /* D group: BfD */
    cplx DA = cexp(nI * p) / ((2. * sqpi) * T4);
    cplx DB = (2 * T) * (((((I2 * (1.0 + eiphi)) * rpT) * T2) * tpzT) -
                         rT * ((T * ((nI - 2 * tb) + eiphi * (nI + 2 * tb))) * tpzT2 +
                               (I * (1.0 + eiphi)) * (((2.0 * tpT) * tpzT2) - T2 * tpzzT)));
    cplx DC = (eiphi * rpiT *
               (rppT * rT4T + ((I2 * rT) * aip) *
                                  ((((2.0 * rpT) * T2) * tpzT) +
                                   rT * (((((I2 * T) * tb) * tpzT2) - ((2.0 * tpT) * tpzT2)) + T2 * tpzzT)))) /
              rT2;
    cplx DD = (rpiT * (rppT * rT4T - ((2.0 * rT) * ain) *
                                         ((((I2 * rpT) * T2) * tpzT) +
                                          rT * (((((2 * T) * tb) * tpzT2) - ((I2 * tpT) * tpzT2)) + (I * T2) * tpzzT)))) /
              rT2;
    cplx BS = 0, DS = 0;
    if (flip_n) {
        BS += ((2.0 * BA) * BC) * e2;
        DS += ((2.0 * DA) * DC) * e2;
        BC = -BC;
        DC = -DC;
    }
    if (flip_p) {
        BS += ((2.0 * BA) * BD) * e2m;
        DS += ((2.0 * DA) * DD) * e2m;
        BD = -BD;
        DD = -DD;
    }
    cplx BfB = BS + (BA * exp(-tb * tb)) * ((BB + BC * wn) - BD * wp);
    cplx BfD = DS + (DA * exp(-tb * tb)) * ((DB + DC * wn) - DD * wp);
    return I * (BfD - BfB);
}

/* ShortPulseInjectSourceFunc::setParam, src/shortpulseinject.cpp:56-113 */
static void pulse_setparam(pulse_t *P, const double prm[8], int nx, int ny, int nz, double dx,
                           double dy, double dz, double dt, int normalise)
{
    P->DX = dx;
    P->DY = dy;
    P->DZ = dz;
    P->DT = dt;
    /* gridLow = 0, gridHigh = N+1 */
    P->centrex = 0.5 * (double)(nx + 1 + 0);
    P->centrey = 0.5 * (double)(ny + 1 + 0);
    P->centrez = 0.5 * (double)(nz + 1 + 0);
    P->length = prm[0];
    P->width = prm[1];
    P->om0 = prm[2];
    P->eps = prm[7];
    P->lightspeed = sqrt(1 / P->eps);
    P->ZRl = 0.5 * P->om0 * P->width * P->width / P->lightspeed;
    P->Phase = 0.0;
    P->TShift = 0.0;
    P->ZShift = 0.0;
    double Exmax = creal(pulse_efunc(P, 0, 0, 0, 0));
    P->amp = normalise ? prm[6] / Exmax : prm[6];
    P->Phase = 2 * M_PI * prm[5];
    P->TShift = prm[3];
    P->ZShift = prm[4];
}

void or_pulse_efunc(const double prm[8], int normalise, double x, double y, double z, double t,
                    double out[2])
{
    pulse_t P;
    pulse_setparam(&P, prm, 1, 1, 1, 1, 1, 1, 1, normalise);
    cplx e = pulse_efunc(&P, x, y, z, t);
    out[0] = creal(e);
    out[1] = cimag(e);
}

void or_pulse_bfunc(const double prm[8], int normalise, double x, double y, double z, double t,
                    int bx, double out[2])
{
    pulse_t P;
    pulse_setparam(&P, prm, 1, 1, 1, 1, 1, 1, 1, normalise);
    cplx e = pulse_bfunc(&P, x, y, z, t, bx);
    out[0] = creal(e);
    out[1] = cimag(e);
}

double or_pulse_amp(const double prm[8])
{
    pulse_t P;
    pulse_setparam(&P, prm, 1, 1, 1, 1, 1, 1, 1, 1);
    return P.amp;
}

/* getEField, src/shortpulseinject.cpp:126-149 (YComp = 0, so Ey = 0) */
static void pulse_getE(const pulse_t *P, int i, int j, int k, int time, double out[3])
{
    double posxh = (i + 0.5 - P->centrex) * P->DX;
    double posyo = (j - P->centrey) * P->DY;
    double poszo = (k - P->centrez) * P->DZ - P->ZShift;
    double posTime = time * P->DT - P->TShift;
    cplx Exc = pulse_efunc(P, posxh, posyo, poszo, posTime);
    out[0] = P->amp * creal(Exc);
    out[1] = 0;
    out[2] = 0;
}

/* getHField, src/shortpulseinject.cpp:151-178, with the k-1/2 position of line 163 */
static void pulse_getH(const pulse_t *P, int i, int j, int k, int time, double out[3])
{
    double posxo = (i - P->centrex) * P->DX;
    double posxh = (i + 0.5 - P->centrex) * P->DX;
    double posyo = (j - P->centrey) * P->DY;
    double posyh = (j + 0.5 - P->centrey) * P->DY;
    double poszh = (k - 0.5 - P->centrez) * P->DZ - P->ZShift;
    double posTime = (time - 0.5) * P->DT - P->TShift;
    cplx Bxc = pulse_bfunc(P, posxo, posyh, poszh, posTime, 1);
    cplx Byc = pulse_bfunc(P, posxh, posyo, poszh, posTime, 0);
    out[0] = P->amp * creal(Bxc);
    out[1] = P->amp * creal(Byc);
    out[2] = 0;
}

/* ------------------------------------------------------------------------------------ */
/* Plane-wave and plane-Gauss source functions, src/sources.cpp                          */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int gauss;
    double k[3], E[3], H[3], E_bg[3], H_bg[3];
    double ramp, eps, dt, om, dx, dy, dz;
    double width, offset;
} plane_t;

/* makeECurrent/makeHCurrent field set-up, src/sources.cpp:14-56 and 185-223 */
static void plane_setparam(plane_t *W, const mpx_factory *f, int gauss, double dx, double dy,
                           double dz, double dt)
{
    W->gauss = gauss;
    double kx = f->k[0], ky = f->k[1], kz = f->k[2];
    double Hx = f->h[0], Hy = f->h[1], Hz = f->h[2];
    double E[3] = {ky * Hz - kz * Hy, kz * Hx - kx * Hz, kx * Hy - ky * Hx};
    double bmag = sqrt(Hx * Hx + Hy * Hy + Hz * Hz);
    double factor = -bmag / sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    for (int d = 0; d < 3; ++d) {
        E[d] *= factor / sqrt(f->eps);
        W->k[d] = f->k[d];
        W->H[d] = f->h[d];
        W->E[d] = E[d];
        W->H_bg[d] = gauss ? 0 : f->h_bg[d];
        W->E_bg[d] = gauss ? 0 : f->e_bg[d];
    }
    W->ramp = f->ramp;
    W->eps = f->eps;
    W->dt = dt;
    W->om = sqrt(W->k[0] * W->k[0] + W->k[1] * W->k[1] + W->k[2] * W->k[2]) / sqrt(W->eps);
    W->dx = dx;
    W->dy = dy;
    W->dz = dz;
    W->width = f->gwidth;
    W->offset = f->goffset;
}

static double sqr(double x) { return x * x; }

/* PlaneWaveSourceEFunc::getHField src/sources.cpp:107-125;
   PlaneGaussSourceEFunc::getHField src/sources.cpp:266-284 */
static void plane_getH(const plane_t *W, int i, int j, int l, int time, double out[3])
{
    double realtime = W->dt * (time - 0.5);
    const double *k = W->k;
    double posx = k[0] * i * W->dx + k[1] * (j + 0.5) * W->dy + k[2] * (l + 0.5) * W->dz + W->om * realtime;
    double posy = k[0] * (i + 0.5) * W->dx + k[1] * j * W->dy + k[2] * (l + 0.5) * W->dz + W->om * realtime;
    double posz = k[0] * (i + 0.5) * W->dx + k[1] * (j + 0.5) * W->dy + k[2] * l * W->dz + W->om * realtime;
    if (!W->gauss) {
        double hx = W->H[0] * sin(posx) + W->H_bg[0];
        double hy = W->H[1] * sin(posy) + W->H_bg[1];
        double hz = W->H[2] * sin(posz) + W->H_bg[2];
        if (posx < W->ramp) hx *= posx / W->ramp;
        if (posy < W->ramp) hy *= posy / W->ramp;
        if (posz < W->ramp) hz *= posz / W->ramp;
        out[0] = hx;
        out[1] = hy;
        out[2] = hz;
    } else {
        double ampx = W->H[0] * exp(-sqr((posx - W->offset) / W->width));
        double ampy = W->H[1] * exp(-sqr((posy - W->offset) / W->width));
        double ampz = W->H[2] * exp(-sqr((posz - W->offset) / W->width));
        out[0] = ampx * sin(posx);
        out[1] = ampy * sin(posy);
        out[2] = ampz * sin(posz);
    }
}

/* PlaneWaveSourceHFunc::getEField src/sources.cpp:150-173;
   PlaneGaussSourceHFunc::getEField src/sources.cpp:308-331 */
static void plane_getE(const plane_t *W, int i, int j, int l, int time, double out[3])
{
    double realtime = W->dt * time;
    const double *k = W->k;
    double posx = k[0] * (i + 0.5) * W->dx + k[1] * j * W->dy + k[2] * l * W->dz + W->om * realtime;
    double posy = k[0] * i * W->dx + k[1] * (j + 0.5) * W->dy + k[2] * l * W->dz + W->om * realtime;
    double posz = k[0] * i * W->dx + k[1] * j * W->dy + k[2] * (l + 0.5) * W->dz + W->om * realtime;
    if (!W->gauss) {
        double ex = W->E[0] * sin(posx) + W->E_bg[0];
        double ey = W->E[1] * sin(posy) + W->E_bg[1];
        double ez = W->E[2] * sin(posz) + W->E_bg[2];
        if (posx < W->ramp) ex *= posx / W->ramp;
        if (posy < W->ramp) ey *= posy / W->ramp;
        if (posz < W->ramp) ez *= posz / W->ramp;
        out[0] = ex;
        out[1] = ey;
        out[2] = ez;
    } else {
        double ampx = W->E[0] * exp(-sqr((posx - W->offset) / W->width));
        double ampy = W->E[1] * exp(-sqr((posy - W->offset) / W->width));
        double ampz = W->E[2] * exp(-sqr((posz - W->offset) / W->width));
        out[0] = ampx * sin(posx);
        out[1] = ampy * sin(posy);
        out[2] = ampz * sin(posz);
    }
}

/* ------------------------------------------------------------------------------------ */
/* GaussInject source function, src/gaussinject.cpp:56-482 (FFTW transforms restated as   */
/* direct discrete Fourier sums with the same sign conventions and index shifts)          */
/* ------------------------------------------------------------------------------------ */

enum { GAUSS_EX = 0, GAUSS_EY = 1, GAUSS_BX = 2 };

typedef struct {
    int isH;
    int SizeX, SizeY, SizeT, SizeXH, SizeYH, SizeTH, Nt;
    double BoxSizeX, BoxSizeY, BoxSizeT, DX, DY, DT;
    double length, width, k0, Shift, Time, amp, eps;
    int toff, blockCount, blockToff, active;
    double sx, sy, sz; /* GaussFunc::setShift */
    double *F1, *F2;   /* [kh][i][j]: Nt x SizeX x SizeY */
} gauss_t;

/* GaussInjectSourceFunc::setParam src/gaussinject.cpp:56-106, initSourceFunc :108-150
   (Nt from the storage extent along z, lowz..highz) */
static int gauss_setup(gauss_t *G, const mpx_factory *f, int isH, int nx, int ny, double dx, double dy,
                       double dt, int lowz, int highz)
{
    memset(G, 0, sizeof(*G));
    G->isH = isH;
    G->SizeT = f->gsize_t;
    G->length = f->length;
    G->width = f->width;
    G->k0 = f->k0;
    G->Shift = f->gshift;
    G->Time = f->gtime;
    G->amp = f->amp;
    G->eps = f->eps;
    G->toff = f->toff;
    double lightspeed = sqrt(1 / G->eps);
    G->SizeX = nx + 2;
    G->SizeY = ny + 2;
    G->SizeTH = G->SizeT / 2;
    G->SizeXH = G->SizeX / 2;
    G->SizeYH = G->SizeY / 2;
    G->DX = dx;
    G->DY = dy;
    G->DT = dt * lightspeed;
    G->BoxSizeX = G->SizeX * G->DX;
    G->BoxSizeY = G->SizeY * G->DY;
    G->BoxSizeT = G->SizeT * G->DT;
    G->Nt = 1 << (int)round(log((double)(highz - lowz)) / log(2.0) + 1e-6);
    G->blockToff = G->toff;
    size_t n = (size_t)G->Nt * G->SizeX * G->SizeY;
    G->F1 = (double *)calloc(n, sizeof(double));
    G->F2 = (double *)calloc(n, sizeof(double));
    return G->F1 && G->F2;
}

/* GaussFunc, src/gaussinject.h:150-200 */
static double gauss_comp(int comp, double x, double y, double z)
{
    double kp = sqrt(x * x + z * z), k = sqrt(x * x + y * y + z * z);
    if (comp == GAUSS_EX) {
        double D = k * kp;
        return D == 0 ? 0 : (-x * y / D);
    }
    if (comp == GAUSS_EY) return k == 0 ? 0 : (kp / k);
    return kp == 0 ? 0 : (-z / kp);
}

/* direct sum  sum_m in[m] exp(sgn 2 pi i m n / N)  with exact twiddle indices (m n mod N) */
static cplx gauss_dft(const cplx *in, int N, int n, int sgn, const cplx *tw)
{
    cplx acc = 0;
    for (int m = 0; m < N; ++m) {
        long long q = ((long long)m * n) % N;
        acc += in[m] * (sgn > 0 ? tw[q] : conj(tw[q]));
    }
    return acc;
}

/* fillEx / fillEy / fillBx with transformT, transformY, transformX, src/gaussinject.cpp:235-482 */
static void gauss_fill(gauss_t *G, int comp, double *F)
{
    const int SX = G->SizeX, SY = G->SizeY, ST = G->SizeT, Nt = G->Nt;
    const int blockmin = G->blockCount * Nt;
    const double ft = 2 * M_PI * G->length / G->BoxSizeT, fy = G->width / G->BoxSizeY, fx = G->width / G->BoxSizeX;
    int NM = ST > SX ? ST : SX;
    if (SY > NM) NM = SY;
    cplx *tw = (cplx *)malloc(sizeof(cplx) * NM), *A = (cplx *)malloc(sizeof(cplx) * ST);
    cplx *uo = (cplx *)malloc(sizeof(cplx) * ST), *row = (cplx *)malloc(sizeof(cplx) * NM);
    cplx *TY = (cplx *)malloc(sizeof(cplx) * (size_t)SY * Nt);
    cplx *TX = (cplx *)malloc(sizeof(cplx) * (size_t)SX * SY * Nt);
    for (int i = 0; i < SX; ++i) {
        for (int j = 0; j < SY; ++j) {
            for (int k = 0; k < ST; ++k) {
                double kx = 2 * M_PI * (i - G->SizeXH) / G->BoxSizeX;
                double ky = 2 * M_PI * (j - G->SizeYH) / G->BoxSizeY;
                double om = 2 * M_PI * (k - G->SizeTH) / G->BoxSizeT;
                double K = om * om - ky * ky - kx * kx;
                if (fabs(K) < 1e-9) K = 0;
                if (K < 0)
                    A[k] = 0;
                else {
                    double kz = sqrt(K) * (om > 0 ? 1 : (om < 0 ? -1 : 0));
                    double zoff = kz - G->k0;
                    double g = exp(-0.25 * zoff * zoff * G->length * G->length -
                                   0.25 * (kx * kx + ky * ky) * G->width * G->width);
                    double a = gauss_comp(comp, kx, ky, kz) * g;
                    double ph = -(kx * G->sx + ky * G->sy + kz * G->sz) + 0.5 * M_PI;
                    double th = ph + G->Time * om; /* a exp(-i th) */
                    A[k] = a * CX(cos(th), -sin(th));
                }
            }
            /* transformT: backward transform over omega, keep this block's Nt samples */
            for (int k = 0; k < ST; ++k) uo[k] = A[(k + G->SizeTH) % ST];
            for (int q = 0; q < ST; ++q) tw[q] = CX(cos(2 * M_PI * q / ST), sin(2 * M_PI * q / ST));
            for (int kh = 0; kh < Nt; ++kh) {
                int n = (kh + blockmin + G->SizeTH) % ST;
                TY[(size_t)j * Nt + kh] = ft * gauss_dft(uo, ST, n, +1, tw);
            }
        }
        /* transformY(i): forward transform over y */
        for (int q = 0; q < SY; ++q) tw[q] = CX(cos(2 * M_PI * q / SY), sin(2 * M_PI * q / SY));
        for (int kh = 0; kh < Nt; ++kh) {
            for (int j = 0; j < SY; ++j) row[(j + G->SizeYH) % SY] = TY[(size_t)j * Nt + kh];
            for (int j = 0; j < SY; ++j) {
                int jh = (j + G->SizeYH) % SY;
                TX[((size_t)i * SY + j) * Nt + kh] = fy * gauss_dft(row, SY, jh, -1, tw);
            }
        }
    }
    /* transformX: forward transform over x, real part */
    for (int q = 0; q < SX; ++q) tw[q] = CX(cos(2 * M_PI * q / SX), sin(2 * M_PI * q / SX));
    for (int j = 0; j < SY; ++j)
        for (int kh = 0; kh < Nt; ++kh) {
            for (int i = 0; i < SX; ++i) row[(i + G->SizeXH) % SX] = TX[((size_t)i * SY + j) * Nt + kh];
            for (int i = 0; i < SX; ++i) {
                int ih = (i + G->SizeXH) % SX;
                F[((size_t)kh * SX + i) * SY + j] = fx * creal(gauss_dft(row, SX, ih, -1, tw));
            }
        }
    free(tw);
    free(A);
    free(uo);
    free(row);
    free(TY);
    free(TX);
}

/* GaussInjectSourceFunc::setTime, src/gaussinject.cpp:152-182 */
static void gauss_settime(gauss_t *G, int time)
{
    if (time < G->toff + G->blockCount * G->Nt) return;
    if (time >= G->toff + G->SizeT) {
        G->active = 0;
        return;
    }
    if (!G->isH) {
        G->sx = 0, G->sy = 0.5 * G->DX, G->sz = G->Shift + 0.5 * G->DX;
        gauss_fill(G, GAUSS_BX, G->F1);
    } else {
        G->Time -= 0.5 * G->DT;
        G->sx = 0.5 * G->DX, G->sy = 0, G->sz = G->Shift;
        gauss_fill(G, GAUSS_EX, G->F1);
        G->sx = 0, G->sy = 0.5 * G->DX, G->sz = G->Shift;
        gauss_fill(G, GAUSS_EY, G->F2);
        G->Time += 0.5 * G->DT;
    }
    G->blockToff = G->toff + G->blockCount * G->Nt;
    G->blockCount++;
    G->active = 1;
}

/* getHField / getEField, src/gaussinject.cpp:203-224 */
static double gauss_at(const gauss_t *G, const double *F, int i, int j, int time)
{
    int t = time - G->blockToff;
    return F[((size_t)t * G->SizeX + i) * G->SizeY + j];
}

/* ------------------------------------------------------------------------------------ */
/* Focused pulse (focused_pulse_inject keys; the field model is the library's own design)  */
/* ------------------------------------------------------------------------------------ */

/* x-polarised Gaussian pupil A(k) = exp(-(kx^2+ky^2) w0^2/4) at the carrier om0, vector plane
 * waves E(k) = A (1, 0, -kx/kz), B(k) = k x E / om0 (kz = sqrt(eps om0^2 - k_perp^2), evanescent
 * waves dropped), summed on the periodic transverse grid of the domain; the field is
 * amp Re[ env(t - z/v) U(x, y, z) exp(-i (om0 t - phase)) ] with env(s) = exp(-s^2/T^2),
 * z measured from the focus, v = 1/sqrt(eps) and U normalised to |U| = 1 at the focus.
 * The planes are sampled where the short-pulse source samples its fields (E_x at
 * (i+1/2, j, k), B_x at (i, j+1/2, k-1/2), B_y at (i+1/2, j, k-1/2), z index d-1). */
typedef struct {
    int nx2, ny2;
    double *ex, *bx, *by; /* interleaved re, im per (i, j) */
    double zE, zB, T, om0, v, TShift, phase, amp, DT;
} focused_t;

static void focused_plane(const mpx_factory *f, int nx, int ny, double dx, double dy, double z, int comp,
                          double sx, double sy, double *out)
{
    double n = sqrt(f->eps), k = n * f->om0, w0 = f->width;
    double cx = 0.5 * (nx + 1), cy = 0.5 * (ny + 1);
    int nx2 = nx + 2, ny2 = ny + 2;
    /* inner sum over ky for every (kx, y_j) */
    double complex *tmp = (double complex *)calloc((size_t)nx * ny2, sizeof(double complex));
    for (int mx = 0; mx < nx; ++mx) {
        double kx = 2 * M_PI * (mx - nx / 2) / (nx * dx);
        for (int j = 0; j < ny2; ++j) {
            double y = (j + sy - cy) * dy;
            double complex acc = 0;
            for (int my = 0; my < ny; ++my) {
                double ky = 2 * M_PI * (my - ny / 2) / (ny * dy);
                double kp2 = kx * kx + ky * ky;
                if (kp2 >= k * k) continue;
                double kz = sqrt(k * k - kp2);
                double A = exp(-kp2 * w0 * w0 / 4);
                double g = comp == 0 ? A : (comp == 1 ? -kx * ky * A / (kz * f->om0) : A * (kz * kz + kx * kx) / (kz * f->om0));
                acc += g * cexp(CX(0, ky * y + kz * z));
            }
            tmp[(size_t)mx * ny2 + j] = acc;
        }
    }
    double norm = 0;
    for (int mx = 0; mx < nx; ++mx)
        for (int my = 0; my < ny; ++my) {
            double kx = 2 * M_PI * (mx - nx / 2) / (nx * dx), ky = 2 * M_PI * (my - ny / 2) / (ny * dy);
            if (kx * kx + ky * ky < k * k) norm += exp(-(kx * kx + ky * ky) * w0 * w0 / 4);
        }
    for (int i = 0; i < nx2; ++i) {
        double x = (i + sx - cx) * dx;
        for (int j = 0; j < ny2; ++j) {
            double complex acc = 0;
            for (int mx = 0; mx < nx; ++mx) {
                double kx = 2 * M_PI * (mx - nx / 2) / (nx * dx);
                acc += tmp[(size_t)mx * ny2 + j] * cexp(CX(0, kx * x));
            }
            acc /= norm;
            out[2 * ((size_t)i * ny2 + j)] = creal(acc);
            out[2 * ((size_t)i * ny2 + j) + 1] = cimag(acc);
        }
    }
    free(tmp);
}

static int focused_setup(focused_t *F, const mpx_factory *f, int nx, int ny, int nz, double dx, double dy,
                         double dz, double dt)
{
    F->nx2 = nx + 2;
    F->ny2 = ny + 2;
    double cz = 0.5 * (nz + 1);
    int kk = f->d - 1; /* the plane both currents sample (layers at d and d-1, down face) */
    F->zE = (kk - cz) * dz - f->zshift;
    F->zB = (kk - 0.5 - cz) * dz - f->zshift;
    F->T = f->length;
    F->om0 = f->om0;
    F->v = sqrt(1 / f->eps);
    F->TShift = f->tshift;
    F->phase = 2 * M_PI * f->phase;
    F->amp = f->amp;
    F->DT = dt;
    size_t cnt = 2 * (size_t)F->nx2 * F->ny2;
    F->ex = (double *)malloc(cnt * sizeof(double));
    F->bx = (double *)malloc(cnt * sizeof(double));
    F->by = (double *)malloc(cnt * sizeof(double));
    if (!F->ex || !F->bx || !F->by) return 0;
    focused_plane(f, nx, ny, dx, dy, F->zE, 0, 0.5, 0.0, F->ex);
    focused_plane(f, nx, ny, dx, dy, F->zB, 1, 0.0, 0.5, F->bx);
    focused_plane(f, nx, ny, dx, dy, F->zB, 2, 0.5, 0.0, F->by);
    return 1;
}

static double focused_value(const focused_t *F, const double *U, int i, int j, double t, double z)
{
    size_t q = 2 * ((size_t)i * F->ny2 + j);
    double s = t - z / F->v;
    double env = exp(-(s * s) / (F->T * F->T));
    double ph = -(F->om0 * t - F->phase);
    return F->amp * env * (U[q] * cos(ph) - U[q + 1] * sin(ph));
}

int or_gauss_block(const mpx_factory *f, int isH, int nx, int ny, double dx, double dy, double dt, int lowz,
                   int highz, int time, double *F1, double *F2, int *block_toff)
{
    gauss_t G;
    if (f->gsize_t < 2 || !gauss_setup(&G, f, isH, nx, ny, dx, dy, dt, lowz, highz)) return 0;
    /* the sheet calls setTime every step: replay the block boundaries up to `time` */
    for (int t = 0; t <= time; ++t) gauss_settime(&G, t);
    int nt = G.active ? G.Nt : 0;
    if (nt) {
        size_t n = (size_t)G.Nt * G.SizeX * G.SizeY;
        if (F1) memcpy(F1, G.F1, n * sizeof(double));
        if (F2) memcpy(F2, G.F2, n * sizeof(double));
        if (block_toff) *block_toff = G.blockToff;
    }
    free(G.F1);
    free(G.F2);
    return nt;
}

void or_focused_plane(const mpx_factory *f, int nx, int ny, double dx, double dy, double z, int comp, double sx,
                      double sy, double *out)
{
    focused_plane(f, nx, ny, dx, dy, z, comp, sx, sy, out);
}

/* ------------------------------------------------------------------------------------ */
/* Currents (CPML borders and incident-source sheets)                                   */
/* ------------------------------------------------------------------------------------ */

enum { CUR_CPML_E = 1, CUR_CPML_H = 2, CUR_INC_E = 3, CUR_INC_H = 4, CUR_PLASMA = 5 };
enum { SRC_PULSE = 1, SRC_PLANE = 2, SRC_FOCUSED = 3, SRC_GAUSS = 4 };

typedef struct {
    char name[24];
    int dir;
    grid_t g;
} layer_t;

#define MAX_LAYERS 128

typedef struct {
    int type;
    int dir, dim, t1, t2;
    int valid;
    grid_t *J[3];
    /* CPML */
    int thickness;
    double kappaMax, aMax, sigmaMax, eps;
    int reverse;
    line_t b, c;
    int psi0, psi1;       /* component index of pPsi[0], pPsi[1] */
    int f0, f1;           /* field component index of pB[0]/pE[0], pB[1]/pE[1] */
    double dxn;
    /* incident */
    int distance;
    double dX[3];
    int src_kind;
    pulse_t pulse;
    plane_t plane;
    focused_t focused;
    gauss_t gauss;
    /* plasma current */
    double em, gamma, rho_bg;
} current_t;

#define MAX_CURRENTS 64

struct or_sim {
    mpx_desc d;
    int N[3];
    int low[3], high[3];
    int glow[3], ghigh[3];
    double dx, dy, dz, dt;
    int time;
    int solver;
    /* grids */
    grid_t E[3], B[3], Sigma;
    grid_t Eps; /* per-cell permittivity (allocated when set; not a reference grid) */
    grid_t Rho, Jp[3]; /* "PlasmaDensity", "PlasmaJx".."PlasmaJz" */
    grid_t PsiR[3][3], PsiI[3][3]; /* [pole][component] */
    grid_t P[3][3], Pp[3][3];      /* fdtd-disp */
    grid_t J[3], M[3];
    int haveJ, haveM;
    line_t kappaE[3], kappaH[3];
    int haveKappa;
    int store_mode; /* storage rounding of E, B and the CPML auxiliaries (store_round) */
    /* PLRCData, src/fdtd_plrc.h:71-78 */
    cplx dchi0[3], dxi0[3], Crec[3];
    double sumChi0, sumXi0;
    /* layers and currents */
    layer_t layers[MAX_LAYERS];
    int nlayers;
    current_t cur[MAX_CURRENTS];
    int ncur;
    int elist[MAX_CURRENTS], hlist[MAX_CURRENTS];
    int nE, nH;
};

/* Storage::getBorderExtent, src/storage.cpp:208-318 (bordercells = 0 on every path) */
static int border_extent(const or_sim *s, int dir, int thickness, int distance, int blow[3],
                         int bhigh[3])
{
    int have = 0;
    for (int d = 0; d < 3; ++d) {
        blow[d] = s->low[d];
        bhigh[d] = s->high[d];
    }
    const int *glow = s->glow, *ghigh = s->ghigh, *low = s->low, *high = s->high;
    switch (dir) {
    case WEST:
        if (low[0] < glow[0] + thickness + distance) {
            bhigh[0] = glow[0] + thickness - 1 + distance;
            blow[0] = glow[0] + distance;
            have = 1;
        }
        break;
    case SOUTH:
        if (low[1] < glow[1] + thickness + distance) {
            bhigh[1] = glow[1] + thickness - 1 + distance;
            blow[1] = glow[1] + distance;
            have = 1;
        }
        break;
    case DOWN:
        if ((low[2] < glow[2] + thickness + distance) && (high[2] >= glow[2] + thickness)) {
            bhigh[2] = glow[2] + thickness - 1 + distance;
            blow[2] = glow[2] + distance;
            have = 1;
        }
        break;
    case EAST:
        if (high[0] > ghigh[0] - thickness - distance) {
            blow[0] = ghigh[0] - thickness + 1 - distance;
            bhigh[0] = ghigh[0] - distance;
            have = 1;
        }
        break;
    case NORTH:
        if (high[1] > ghigh[1] - thickness - distance) {
            blow[1] = ghigh[1] - thickness + 1 - distance;
            bhigh[1] = ghigh[1] - distance;
            have = 1;
        }
        break;
    case UP:
        if (high[2] > ghigh[2] - thickness - distance) {
            blow[2] = ghigh[2] - thickness + 1 - distance;
            bhigh[2] = ghigh[2] - distance;
            have = 1;
        }
        break;
    }
    return have;
}

/* Storage::addBorderLayer, src/storage.cpp:106-142: one zeroed grid per (name, direction),
   shared by every caller asking for the same pair; NULL where the face has no layer. */
static grid_t *add_border_layer(or_sim *s, const char *name, int dir, int thickness, int distance)
{
    for (int q = 0; q < s->nlayers; ++q)
        if (s->layers[q].dir == dir && strcmp(s->layers[q].name, name) == 0) return &s->layers[q].g;
    int blow[3], bhigh[3];
    if (!border_extent(s, dir, thickness, distance, blow, bhigh)) return NULL;
    if (s->nlayers >= MAX_LAYERS) return NULL;
    layer_t *L = &s->layers[s->nlayers++];
    snprintf(L->name, sizeof(L->name), "%s", name);
    L->dir = dir;
    if (!grid_alloc(&L->g, blow, bhigh)) return NULL;
    return &L->g;
}

/* CPMLBorderCurrent constructor, src/cpml_border.cpp:262-285 */
static void cpml_dims(current_t *c)
{
    switch (c->dir) {
    case EAST:
    case WEST:
        c->dim = 0;
        c->t1 = 1;
        c->t2 = 2;
        break;
    case NORTH:
    case SOUTH:
        c->dim = 1;
        c->t1 = 0;
        c->t2 = 2;
        break;
    default:
        c->dim = 2;
        c->t1 = 0;
        c->t2 = 1;
        break;
    }
}

/* CPMLBorderCurrent::makeCoeff, src/cpml_border.cpp:287-339 */
static int cpml_make_coeff(or_sim *s, current_t *c, int isH)
{
    double dt = s->dt;
    int lowk = c->J[0]->lo[c->dim];
    int highk = c->J[0]->hi[c->dim];
    switch (c->dir) {
    case EAST:
    case NORTH:
    case UP:
        c->reverse = 0;
        break;
    default:
        c->reverse = 1;
        break;
    }
    if (!line_alloc(&c->b, lowk, highk, 0.0) || !line_alloc(&c->c, lowk, highk, 0.0)) return 0;
    double offset = (isH && ORACLE_MUTANT != 2) ? 0.5 : 0.0;
    for (int k = 0; k < c->thickness; ++k) {
        double x = 1 - ((double)k - offset) / (double)c->thickness;
        if (ORACLE_MUTANT == 3 && x > 1) x = 1;
        double x3 = x * x * x;
        int pos = c->reverse ? (lowk + k) : (highk - k);
        if (ORACLE_MUTANT == 4) {
            pos += c->reverse ? 1 : -1;
            if (pos < lowk || pos > highk) continue;
        }
        double sigma = x3 * c->sigmaMax;
        double kappa = 1 + (c->kappaMax - 1) * x3;
        double a = c->aMax * (1 - x);
        double b = exp(-(sigma / kappa + a) * dt / c->eps);
        double cc = sigma * (b - 1) / (kappa * (sigma + kappa * a));
        LN(&c->b, pos) = b;
        LN(&c->c, pos) = cc;
    }
    return 1;
}

/* CPMLBorderECurrent::initStorage src/cpml_border.cpp:350-393,
   CPMLBorderHCurrent::initStorage src/cpml_border.cpp:437-492 */
static int cpml_init_storage(or_sim *s, current_t *c, int isH)
{
    int distance = 1;
    if (isH) distance = (c->dir == EAST || c->dir == NORTH || c->dir == UP) ? 1 : 0;
    const char *nm[3];
    if (!isH) {
        nm[0] = "CPMLPsiEx";
        nm[1] = "CPMLPsiEy";
        nm[2] = "CPMLPsiEz";
    } else {
        nm[0] = "CPMLPsiBx";
        nm[1] = "CPMLPsiBy";
        nm[2] = "CPMLPsiBz";
    }
    for (int q = 0; q < 3; ++q) c->J[q] = add_border_layer(s, nm[q], c->dir, c->thickness, distance);
    c->valid = c->J[0] && c->J[1] && c->J[2];
    switch (c->dir) {
    case EAST:
    case WEST:
        c->psi0 = 1;
        c->psi1 = 2;
        c->f0 = 1;
        c->f1 = 2;
        c->dxn = s->dx;
        break;
    case NORTH:
    case SOUTH:
        c->psi0 = 2;
        c->psi1 = 0;
        c->f0 = 2;
        c->f1 = 0;
        c->dxn = s->dy;
        break;
    default:
        c->psi0 = 0;
        c->psi1 = 1;
        c->f0 = 0;
        c->f1 = 1;
        c->dxn = s->dz;
        break;
    }
    if (c->J[0]) return cpml_make_coeff(s, c, isH);
    return 1;
}

/* CPMLBorderECurrent::stepScheme, src/cpml_border.cpp:398-425 */
static void cpml_step_E(or_sim *s, current_t *c)
{
    grid_t *Psi0 = c->J[c->psi0], *Psi1 = c->J[c->psi1];
    grid_t *B0 = &s->B[c->f0], *B1 = &s->B[c->f1];
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = Psi0->lo[d];
        hi[d] = Psi0->hi[d];
    }
    int dim = c->dim;
    double dx = c->dxn;
#pragma omp parallel for schedule(static)
    for (int i0 = lo[0]; i0 <= hi[0]; ++i0)
        for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
            for (int i2 = lo[2]; i2 <= hi[2]; ++i2) {
                int ind[3] = {i0, i1, i2};
                int j = ind[dim];
                int m[3] = {i0, i1, i2};
                --m[dim];
                double *p0 = &AT(Psi0, i0, i1, i2);
                double *p1 = &AT(Psi1, i0, i1, i2);
                *p0 = LN(&c->b, j) * (*p0) - (ORACLE_MUTANT == 5 ? -1.0 : 1.0) *
                      LN(&c->c, j) * (AT(B1, i0, i1, i2) - AT(B1, m[0], m[1], m[2])) / dx;
                *p1 = LN(&c->b, j) * (*p1) +
                      LN(&c->c, j) * (AT(B0, i0, i1, i2) - AT(B0, m[0], m[1], m[2])) / dx;
            }
}

/* CPMLBorderHCurrent::stepScheme, src/cpml_border.cpp:499-527 */
static void cpml_step_H(or_sim *s, current_t *c)
{
    grid_t *Psi0 = c->J[c->psi0], *Psi1 = c->J[c->psi1];
    grid_t *E0 = &s->E[c->f0], *E1 = &s->E[c->f1];
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = Psi0->lo[d];
        hi[d] = Psi0->hi[d];
    }
    int dim = c->dim;
    double dx = c->dxn;
#pragma omp parallel for schedule(static)
    for (int i0 = lo[0]; i0 <= hi[0]; ++i0)
        for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
            for (int i2 = lo[2]; i2 <= hi[2]; ++i2) {
                int ind[3] = {i0, i1, i2};
                int j = ind[dim];
                int p[3] = {i0, i1, i2};
                ++p[dim];
                double *p0 = &AT(Psi0, i0, i1, i2);
                double *p1 = &AT(Psi1, i0, i1, i2);
                *p0 = LN(&c->b, j) * (*p0) +
                      LN(&c->c, j) * (AT(E1, p[0], p[1], p[2]) - AT(E1, i0, i1, i2)) / dx;
                *p1 = LN(&c->b, j) * (*p1) -
                      LN(&c->c, j) * (AT(E0, p[0], p[1], p[2]) - AT(E0, i0, i1, i2)) / dx;
            }
}

/* CPMLBorder::initCoefficients, src/cpml_border.cpp:58-140 (kappa profiles) */
static void cpml_init_kappa(or_sim *s, int thickness, double kappaMax)
{
    for (int dim = 0; dim < 3; ++dim) {
        line_t *KE = &s->kappaE[dim], *KH = &s->kappaH[dim];
        if (s->low[dim] < s->glow[dim] + thickness) {
            for (int q = KE->lo; q <= KE->hi; ++q) LN(KE, q) = 1.0;
            for (int q = KH->lo; q <= KH->hi; ++q) LN(KH, q) = 1.0;
            for (int i = 0; i <= thickness; ++i) {
                double x = 1 - (double)i / (double)thickness;
                double x3 = x * x * x;
                LN(KE, s->low[dim] + i + (ORACLE_MUTANT == 6)) = 1 + (kappaMax - 1) * x3;
            }
            for (int i = 0; i < thickness; ++i) {
                double x = 1 - ((double)i + 0.5) / (double)thickness;
                double x3 = x * x * x;
                LN(KH, s->low[dim] + i) = 1 + (kappaMax - 1) * x3;
            }
        }
        if (s->high[dim] > s->ghigh[dim] - thickness) {
            for (int i = 0; i <= thickness; ++i) {
                double x = 1 - (double)i / (double)thickness;
                double x3 = x * x * x;
                LN(KH, s->high[dim] - i) = 1 + (kappaMax - 1) * x3;
            }
            for (int i = 0; i < thickness; ++i) {
                double x = 1 - ((double)i + 0.5) / (double)thickness;
                double x3 = x * x * x;
                LN(KE, s->high[dim] - i) = 1 + (kappaMax - 1) * x3;
            }
        }
    }
}

/* CPMLBorderOneD::initCoefficients, src/cpml_border.cpp:179-244 */
static void cpml1d_init_kappa(or_sim *s, int thickness, double kappaMax)
{
    line_t *KE = &s->kappaE[2], *KH = &s->kappaH[2];
    for (int q = s->kappaE[1].lo; q <= s->kappaE[1].hi; ++q) LN(&s->kappaE[1], q) = 1;
    for (int q = s->kappaE[0].lo; q <= s->kappaE[0].hi; ++q) LN(&s->kappaE[0], q) = 1;
    for (int q = s->kappaH[1].lo; q <= s->kappaH[1].hi; ++q) LN(&s->kappaH[1], q) = 1;
    for (int q = s->kappaH[0].lo; q <= s->kappaH[0].hi; ++q) LN(&s->kappaH[0], q) = 1;
    if (s->low[2] < s->glow[2] + thickness) {
        for (int q = KE->lo; q <= KE->hi; ++q) LN(KE, q) = 1.0;
        for (int q = KH->lo; q <= KH->hi; ++q) LN(KH, q) = 1.0;
        for (int i = 0; i < thickness; ++i) {
            double x = 1 - (double)i / (double)thickness;
            double x3 = x * x * x;
            LN(KE, s->low[2] + i + 1) = 1 + (kappaMax - 1) * x3;
        }
        for (int i = 0; i < thickness; ++i) {
            double x = 1 - ((double)i + 0.5) / (double)thickness;
            double x3 = x * x * x;
            LN(KH, s->low[2] + i) = 1 + (kappaMax - 1) * x3;
        }
    }
    if (s->high[2] > s->ghigh[2] - thickness) {
        for (int i = 0; i < thickness; ++i) {
            double x = 1 - ((double)i) / (double)thickness;
            double x3 = x * x * x;
            LN(KE, s->high[2] - i - 1) = 1 + (kappaMax - 1) * x3;
        }
        for (int i = 0; i < thickness; ++i) {
            double x = 1 - (double)(i + 0.5) / (double)thickness;
            double x3 = x * x * x;
            LN(KH, s->high[2] - i - 1) = 1 + (kappaMax - 1) * x3;
        }
    }
}

/* IncidentSourceCurrent constructor, src/incsource.cpp:93-136 */
static void inc_dims(current_t *c, double dx, double dy, double dz)
{
    switch (c->dir) {
    case EAST:
    case WEST:
        c->dim = 0;
        c->t1 = 1;
        c->t2 = 2;
        c->dX[0] = dy;
        c->dX[1] = dz;
        c->dX[2] = dx;
        break;
    case NORTH:
    case SOUTH:
        c->dim = 1;
        c->t1 = 2;
        c->t2 = 0;
        c->dX[0] = dz;
        c->dX[2] = dx;
        c->dX[1] = dy;
        break;
    default:
        c->dim = 2;
        c->t1 = 0;
        c->t2 = 1;
        c->dX[0] = dx;
        c->dX[1] = dy;
        c->dX[2] = dz;
        break;
    }
    c->reverse = (c->dir == EAST) || (c->dir == NORTH) || (c->dir == UP);
}

/* IncidentSourceECurrent::initStorage src/incsource.t:15-31;
   IncidentSourceHCurrent::initStorage src/incsource.t:87-103 */
static void inc_init_storage(or_sim *s, current_t *c, int isH)
{
    if (isH && !c->reverse && ORACLE_MUTANT != 1) c->distance = c->distance - 1;
    const char *nm[3];
    if (!isH) {
        nm[0] = "IncidentEJx";
        nm[1] = "IncidentEJy";
        nm[2] = "IncidentEJz";
    } else {
        nm[0] = "IncidentHJx";
        nm[1] = "IncidentHJy";
        nm[2] = "IncidentHJz";
    }
    for (int q = 0; q < 3; ++q) c->J[q] = add_border_layer(s, nm[q], c->dir, 1, c->distance);
    c->valid = c->J[0] && c->J[1] && c->J[2];
}

static void src_getH(const current_t *c, int i, int j, int k, int time, double out[3])
{
    if (c->src_kind == SRC_GAUSS) {
        const gauss_t *G = &c->gauss;
        out[0] = G->active ? sqrt(G->eps) * G->amp * gauss_at(G, G->F1, i, j, time) : 0;
        out[1] = out[2] = 0;
        (void)k;
    } else if (c->src_kind == SRC_FOCUSED) {
        const focused_t *F = &c->focused;
        double t = (time - 0.5) * F->DT - F->TShift;
        out[0] = focused_value(F, F->bx, i, j, t, F->zB);
        out[1] = focused_value(F, F->by, i, j, t, F->zB);
        out[2] = 0;
        (void)k;
    } else if (c->src_kind == SRC_PULSE)
        pulse_getH(&c->pulse, i, j, k, time, out);
    else
        plane_getH(&c->plane, i, j, k, time, out);
}

static void src_getE(const current_t *c, int i, int j, int k, int time, double out[3])
{
    if (c->src_kind == SRC_GAUSS) {
        const gauss_t *G = &c->gauss;
        out[0] = G->active ? G->amp * gauss_at(G, G->F1, i, j, time) : 0;
        out[1] = G->active ? G->amp * gauss_at(G, G->F2, i, j, time) : 0;
        out[2] = 0;
        (void)k;
    } else if (c->src_kind == SRC_FOCUSED) {
        const focused_t *F = &c->focused;
        double t = time * F->DT - F->TShift;
        out[0] = focused_value(F, F->ex, i, j, t, F->zE);
        out[1] = 0;
        out[2] = 0;
        (void)k;
    } else if (c->src_kind == SRC_PULSE)
        pulse_getE(&c->pulse, i, j, k, time, out);
    else
        plane_getE(&c->plane, i, j, k, time, out);
}

/* IncidentSourceECurrent::stepScheme, src/incsource.t:40-73 */
static void inc_step_E(or_sim *s, current_t *c)
{
    grid_t *J0 = c->J[c->t1], *J1 = c->J[c->t2];
    int Time = s->time;
    if (c->src_kind == SRC_GAUSS) gauss_settime(&c->gauss, Time);
    int off[3] = {0, 0, 0};
    off[c->dim] = -1;
    double DX = c->dX[2];
#pragma omp parallel for schedule(static)
    for (int i0 = J0->lo[0]; i0 <= J0->hi[0]; ++i0)
        for (int i1 = J0->lo[1]; i1 <= J0->hi[1]; ++i1)
            for (int i2 = J0->lo[2]; i2 <= J0->hi[2]; ++i2) {
                double H[3];
                src_getH(c, i0 + off[0], i1 + off[1], i2 + off[2], Time, H);
                AT(J0, i0, i1, i2) = H[c->t2] / DX;
                AT(J1, i0, i1, i2) = H[c->t1] / DX;
            }
}

/* IncidentSourceHCurrent::stepScheme, src/incsource.t:112-145 */
static void inc_step_H(or_sim *s, current_t *c)
{
    grid_t *J0 = c->J[c->t1], *J1 = c->J[c->t2];
    int Time = s->time;
    if (c->src_kind == SRC_GAUSS) gauss_settime(&c->gauss, Time);
    int off[3] = {0, 0, 0};
    off[c->dim] = c->reverse ? +1 : 0;
    double DX = c->dX[2];
#pragma omp parallel for schedule(static)
    for (int i0 = J0->lo[0]; i0 <= J0->hi[0]; ++i0)
        for (int i1 = J0->lo[1]; i1 <= J0->hi[1]; ++i1)
            for (int i2 = J0->lo[2]; i2 <= J0->hi[2]; ++i2) {
                double E[3];
                src_getE(c, i0 + off[0], i1 + off[1], i2 + off[2], Time, E);
                AT(J0, i0, i1, i2) = E[c->t2] / DX;
                AT(J1, i0, i1, i2) = E[c->t1] / DX;
            }
}

/* PlasmaDensity::stepScheme, src/plasmadensity.cpp:18-59: over the whole storage box,
   rho <- (rho (1 + n I) + m I^K) / (1 - n I), Sigma += w I^(K-1), I = |E|^2, with
   n = dt nu / 2, m = dt mpa, w = mpa Wion (the absorption warning is not reproduced) */
static void plasma_density_step(or_sim *s, double dt)
{
    const double n = 0.5 * dt * s->d.plasma_nu, m = dt * s->d.plasma_mpa, w = s->d.plasma_mpa * s->d.plasma_wion;
    const int Km = s->d.plasma_k - 1;
#pragma omp parallel for schedule(static)
    for (int i = s->low[0]; i <= s->high[0]; ++i)
        for (int j = s->low[1]; j <= s->high[1]; ++j)
            for (int k = s->low[2]; k <= s->high[2]; ++k) {
                double ex = AT(&s->E[0], i, j, k), ey = AT(&s->E[1], i, j, k), ez = AT(&s->E[2], i, j, k);
                double I = ex * ex + ey * ey + ez * ez;
                double IKm = 1.0;
                for (int q = 0; q < Km; ++q) IKm *= I; /* ipow, src/util.h:67-73 */
                double IK = I * IKm;
                double *rho = &AT(&s->Rho, i, j, k);
                *rho = (*rho * (1 + n * I) + m * IK) / (1 - n * I);
                AT(&s->Sigma, i, j, k) += w * IKm;
            }
}

/* PlasmaCurrent::stepScheme, src/plasmacurrent.cpp:41-74: over low..high-1 of the box,
   J <- (J (1 - gamma dt / 2) - em dt E (rho + rho_bg)) / (1 + gamma dt / 2) */
static void plasma_current_step(or_sim *s, current_t *c)
{
    const double dt = s->dt;
    const double gdtn = 1 - 0.5 * c->gamma * dt, gdtd = 1 + 0.5 * c->gamma * dt, emdt = c->em * dt;
#pragma omp parallel for schedule(static)
    for (int i = s->low[0]; i < s->high[0]; ++i)
        for (int j = s->low[1]; j < s->high[1]; ++j)
            for (int k = s->low[2]; k < s->high[2]; ++k) {
                double rho = AT(&s->Rho, i, j, k) + c->rho_bg;
                for (int q = 0; q < 3; ++q) {
                    double *jq = &AT(&s->Jp[q], i, j, k);
                    *jq = (*jq * gdtn - emdt * AT(&s->E[q], i, j, k) * rho) / gdtd;
                }
            }
}

static void current_step(or_sim *s, current_t *c)
{
    switch (c->type) {
    case CUR_CPML_E:
        cpml_step_E(s, c);
        grid_round(c->J[c->psi0], s->store_mode == 3 ? 1 : s->store_mode);
        grid_round(c->J[c->psi1], s->store_mode == 3 ? 1 : s->store_mode);
        break;
    case CUR_CPML_H:
        cpml_step_H(s, c);
        grid_round(c->J[c->psi0], s->store_mode == 3 ? 1 : s->store_mode);
        grid_round(c->J[c->psi1], s->store_mode == 3 ? 1 : s->store_mode);
        break;
    case CUR_INC_E: inc_step_E(s, c); break;
    case CUR_INC_H: inc_step_H(s, c); break;
    case CUR_PLASMA: plasma_current_step(s, c); break;
    }
}

/* ------------------------------------------------------------------------------------ */
/* boundary                                                                              */
/* ------------------------------------------------------------------------------------ */

/* SinglePeriodicBoundary::exchangeX/Y/Z, src/periodic_bound.cpp:10-71;
   SingleXYPeriodicBoundary::exchangeZ is empty, src/periodic_bound.cpp:129-130 */
static void exchange(or_sim *s, grid_t *f, int dim)
{
    int U[3] = {f->hi[0], f->hi[1], f->hi[2]}, L[3] = {f->lo[0], f->lo[1], f->lo[2]};
    if (dim == 2 && s->d.boundary == MPX_BOUNDARY_XY_PERIODIC) return;
    if (dim == 0) {
        int mx0 = U[0], mx1 = mx0 - 1, lx0 = L[0], lx1 = lx0 + 1;
        for (int y = L[1]; y <= U[1]; ++y)
            for (int z = L[2]; z <= U[2]; ++z) {
                AT(f, lx0, y, z) = AT(f, mx1, y, z);
                AT(f, mx0, y, z) = AT(f, lx1, y, z);
            }
    } else if (dim == 1) {
        int my0 = U[1], my1 = my0 - 1, ly0 = L[1], ly1 = ly0 + 1;
        for (int x = L[0]; x <= U[0]; ++x)
            for (int z = L[2]; z <= U[2]; ++z) {
                AT(f, x, ly0, z) = AT(f, x, my1, z);
                AT(f, x, my0, z) = AT(f, x, ly1, z);
            }
    } else {
        int mz0 = U[2], mz1 = mz0 - 1, lz0 = L[2], lz1 = lz0 + 1;
        for (int x = L[0]; x <= U[0]; ++x)
            for (int y = L[1]; y <= U[1]; ++y) {
                AT(f, x, y, lz0) = AT(f, x, y, mz1);
                AT(f, x, y, mz0) = AT(f, x, y, lz1);
            }
    }
}

/* Storage::applyBoundary, src/storage.cpp:194-204: per component of the group, X, Y, Z */
static void apply_boundary(or_sim *s, grid_t *group)
{
    for (int c = 0; c < 3; ++c) {
        exchange(s, &group[c], 0);
        exchange(s, &group[c], 1);
        exchange(s, &group[c], 2);
    }
}

/* ------------------------------------------------------------------------------------ */
/* PLRC solver                                                                           */
/* ------------------------------------------------------------------------------------ */

/* pole constants of FDTD_PLRCSolver::stepD, src/fdtd_plrc.t:181-216 */
static void plrc_constants(const double LEps[3], const double LDelta[3], const double LOm2[3],
                           double dt, cplx Crec[3], cplx dchi0[3], cplx dxi0[3], double *sumChi0,
                           double *sumXi0)
{
    const cplx I = CX(0, 1);
    cplx sumChi = 0.0, sumXi = 0.0;
    for (int n = 0; n < 3; ++n) {
        double beta = sqrt(LOm2[n] - LDelta[n] * LDelta[n]);
        double gamma = LEps[n] * LOm2[n] / beta;
        cplx phasor = LDelta[n] - I * beta;
        Crec[n] = cexp(-phasor * dt);
        cplx chi0 = ((I * gamma) * (Crec[n] - 1.0) / phasor);
        cplx xi0 = (I * gamma) * (Crec[n] * (phasor * dt + 1.0) - 1.0) / ((phasor * phasor) * dt);
        dchi0[n] = chi0 * (1.0 - Crec[n]);
        dxi0[n] = xi0 * (1.0 - Crec[n]);
        sumChi += chi0;
        sumXi += xi0;
    }
    *sumChi0 = creal(sumChi);
    *sumXi0 = creal(sumXi);
}

void or_pole_constants(const double l_eps[3], const double l_delta[3], const double l_om2[3],
                       double dt, double crec[6], double dchi0[6], double dxi0[6], double sums[2])
{
    cplx C[3], dc[3], dx[3];
    plrc_constants(l_eps, l_delta, l_om2, dt, C, dc, dx, &sums[0], &sums[1]);
    for (int n = 0; n < 3; ++n) {
        crec[2 * n] = creal(C[n]);
        crec[2 * n + 1] = cimag(C[n]);
        dchi0[2 * n] = creal(dc[n]);
        dchi0[2 * n + 1] = cimag(dc[n]);
        dxi0[2 * n] = creal(dx[n]);
        dxi0[2 * n + 1] = cimag(dx[n]);
    }
}

/* Newton solve of A E^3 + E = C, src/fdtd_plrc.cpp:402-419 */
double or_kerr_newton(double A, double C, double E, int *iterations)
{
    double Eold;
    int it = 0;
    do {
        Eold = E;
        double E2 = E * E;
        E = (2 * A * E2 * E + C) / (3 * A * E2 + 1);
        ++it;
    } while (fabs(E - Eold) > 1e-9 && it < 1000000);
    if (iterations) *iterations = it;
    return E;
}

/* FDTD_PLRCSolver::sumCurrents / sumMagCurrents, src/fdtd_plrc.t:342-409 */
static void sum_currents(or_sim *s, grid_t *T, const int *list, int n)
{
    grid_fill(&T[0], 0);
    grid_fill(&T[1], 0);
    grid_fill(&T[2], 0);
    for (int q = 0; q < n; ++q) {
        current_t *c = &s->cur[list[q]];
        grid_t *jx = c->J[0], *jy = c->J[1], *jz = c->J[2];
#pragma omp parallel for schedule(static)
        for (int i = jx->lo[0]; i <= jx->hi[0]; ++i)
            for (int j = jx->lo[1]; j <= jx->hi[1]; ++j)
                for (int k = jx->lo[2]; k <= jx->hi[2]; ++k) {
                    AT(&T[0], i, j, k) += AT(jx, i, j, k);
                    AT(&T[1], i, j, k) += AT(jy, i, j, k);
                    AT(&T[2], i, j, k) += AT(jz, i, j, k);
                }
    }
}

/* FDTD_PLRCNonlinCore::plrcStepD, src/fdtd_plrc.cpp:304-427, and
   FDTD_PLRCLinCore::plrcStepD, src/fdtd_plrc.cpp:94-261 (without the debug checks) */
static inline void plrc_step_d_cell(or_sim *s, int nonlin, double dt, int i, int j, int k,
                                    double dx, double dy, double dz, double Jx, double Jy, double Jz)
{
    double *pex = &AT(&s->E[0], i, j, k), *pey = &AT(&s->E[1], i, j, k), *pez = &AT(&s->E[2], i, j, k);
    double ex = *pex, ey = *pey, ez = *pez;
    double kappaEdx = LN(&s->kappaE[0], i) * dx;
    double kappaEdy = LN(&s->kappaE[1], j) * dy;
    double kappaEdz = LN(&s->kappaE[2], k) * dz;
    double Psix = 0, Psiy = 0, Psiz = 0;
    for (int n = 0; n < 3; ++n) {
        double *pxr = &AT(&s->PsiR[n][0], i, j, k), *pyr = &AT(&s->PsiR[n][1], i, j, k),
               *pzr = &AT(&s->PsiR[n][2], i, j, k);
        double *pxi = &AT(&s->PsiI[n][0], i, j, k), *pyi = &AT(&s->PsiI[n][1], i, j, k),
               *pzi = &AT(&s->PsiI[n][2], i, j, k);
        cplx D = s->dchi0[n] - s->dxi0[n];
        cplx px = D * ex + CX(*pxr, *pxi);
        cplx py = D * ey + CX(*pyr, *pyi);
        cplx pz = D * ez + CX(*pzr, *pzi);
        Psix += creal(px);
        Psiy += creal(py);
        Psiz += creal(pz);
        px = s->dxi0[n] * ex + s->Crec[n] * px;
        py = s->dxi0[n] * ey + s->Crec[n] * py;
        pz = s->dxi0[n] * ez + s->Crec[n] * pz;
        *pxr = creal(px);
        *pyr = creal(py);
        *pzr = creal(pz);
        *pxi = cimag(px);
        *pyi = cimag(py);
        *pzi = cimag(pz);
    }
    grid_t *Bx = &s->B[0], *By = &s->B[1], *Bz = &s->B[2];
    double sigma = 0.5 * AT(&s->Sigma, i, j, k) * dt;
    /* a per-cell permittivity replaces the uniform eps where it has been set (the reference
       has no such grid: with Eps = eps everywhere the update is the reference's) */
    double eps = s->Eps.v ? AT(&s->Eps, i, j, k) : s->d.eps;
    if (nonlin) {
        double chi = s->d.chi;
        double denom = eps + s->sumChi0 - s->sumXi0 + sigma;
        double numer = eps - s->sumXi0 - sigma;
        double E2 = ex * ex + ey * ey + ez * ez;
        double cx = ((numer + chi * E2) * ex +
                     dt * ((AT(Bz, i, j, k) - AT(Bz, i, j - 1, k)) / kappaEdy -
                           (AT(By, i, j, k) - AT(By, i, j, k - 1)) / kappaEdz + Jx) +
                     Psix) / denom;
        double cy = ((numer + chi * E2) * ey +
                     dt * ((AT(Bx, i, j, k) - AT(Bx, i, j, k - 1)) / kappaEdz -
                           (AT(Bz, i, j, k) - AT(Bz, i - 1, j, k)) / kappaEdx + Jy) +
                     Psiy) / denom;
        double cz = ((numer + chi * E2) * ez +
                     dt * ((AT(By, i, j, k) - AT(By, i - 1, j, k)) / kappaEdx -
                           (AT(Bx, i, j, k) - AT(Bx, i, j - 1, k)) / kappaEdy + Jz) +
                     Psiz) / denom;
        double E = sqrt(E2);
        double A = chi / denom;
        double C = sqrt(cx * cx + cy * cy + cz * cz);
        if (C > 0) {
            double Eold;
            do {
                Eold = E;
                E2 = E * E;
                E = (2 * A * E2 * E + C) / (3 * A * E2 + 1);
            } while (fabs(E - Eold) > 1e-9);
            *pex = E * cx / C;
            *pey = E * cy / C;
            *pez = E * cz / C;
        } else {
            *pex = 0.0;
            *pey = 0.0;
            *pez = 0.0;
        }
    } else {
        double denomX = eps + (s->sumChi0 - s->sumXi0) + sigma;
        double numerX = eps - s->sumXi0 - sigma;
        double exn = (numerX * ex + (dt * ((AT(Bz, i, j, k) - AT(Bz, i, j - 1, k)) / kappaEdy -
                                           (AT(By, i, j, k) - AT(By, i, j, k - 1)) / kappaEdz + Jx) +
                                     Psix)) / denomX;
        double eyn = (numerX * ey + (dt * ((AT(Bx, i, j, k) - AT(Bx, i, j, k - 1)) / kappaEdz -
                                           (AT(Bz, i, j, k) - AT(Bz, i - 1, j, k)) / kappaEdx + Jy) +
                                     Psiy)) / denomX;
        double ezn = (numerX * ez + (dt * ((AT(By, i, j, k) - AT(By, i - 1, j, k)) / kappaEdx -
                                           (AT(Bx, i, j, k) - AT(Bx, i, j - 1, k)) / kappaEdy + Jz) +
                                     Psiz)) / denomX;
        *pex = exn;
        *pey = eyn;
        *pez = ezn;
    }
}

/* FDTD_PLRCSolver::stepD, src/fdtd_plrc.t:171-237 */
static void plrc_stepD(or_sim *s, double dt)
{
    int nonlin = (s->solver == MPX_SOLVER_PLRC_NONLIN);
    plrc_constants(s->d.l_eps, s->d.l_delta, s->d.l_om2, dt, s->Crec, s->dchi0, s->dxi0,
                   &s->sumChi0, &s->sumXi0);
    if (s->haveJ) sum_currents(s, s->J, s->elist, s->nE);
    double dx = s->dx, dy = s->dy, dz = s->dz;
#pragma omp parallel for schedule(static)
    for (int i = s->low[0] + 1; i < s->high[0]; ++i)
        for (int j = s->low[1] + 1; j < s->high[1]; ++j)
            for (int k = s->low[2] + 1; k < s->high[2]; ++k) {
                double jx = 0, jy = 0, jz = 0;
                if (s->haveJ) {
                    jx = AT(&s->J[0], i, j, k);
                    jy = AT(&s->J[1], i, j, k);
                    jz = AT(&s->J[2], i, j, k);
                }
                plrc_step_d_cell(s, nonlin, dt, i, j, k, dx, dy, dz, jx, jy, jz);
            }
    for (int c = 0; c < 3; ++c) grid_round(&s->E[c], s->store_mode == 3 ? 2 : s->store_mode);
    apply_boundary(s, s->E);
}

/* FDTD_PLRC{Lin,Nonlin}Core::plrcStepB, src/fdtd_plrc.cpp:263-293, 429-459, with the
   loop of FDTD_PLRCSolver::stepB, src/fdtd_plrc.t:312-339 */
static void plrc_stepB(or_sim *s, double dt)
{
    if (s->haveM) sum_currents(s, s->M, s->hlist, s->nH);
    double dx = s->dx, dy = s->dy, dz = s->dz;
    grid_t *Ex = &s->E[0], *Ey = &s->E[1], *Ez = &s->E[2];
    grid_t *Bx = &s->B[0], *By = &s->B[1], *Bz = &s->B[2];
#pragma omp parallel for schedule(static)
    for (int i = s->low[0]; i < s->high[0]; ++i)
        for (int j = s->low[1]; j < s->high[1]; ++j)
            for (int k = s->low[2]; k < s->high[2]; ++k) {
                double jx = 0, jy = 0, jz = 0;
                if (s->haveM) {
                    jx = AT(&s->M[0], i, j, k);
                    jy = AT(&s->M[1], i, j, k);
                    jz = AT(&s->M[2], i, j, k);
                }
                double kappaHdx = LN(&s->kappaH[0], i) * dx;
                double kappaHdy = LN(&s->kappaH[1], j) * dy;
                double kappaHdz = LN(&s->kappaH[2], k) * dz;
                AT(Bx, i, j, k) = AT(Bx, i, j, k) +
                                  dt * ((AT(Ey, i, j, k + 1) - AT(Ey, i, j, k)) / kappaHdz -
                                        (AT(Ez, i, j + 1, k) - AT(Ez, i, j, k)) / kappaHdy + jx);
                AT(By, i, j, k) = AT(By, i, j, k) +
                                  dt * ((AT(Ez, i + 1, j, k) - AT(Ez, i, j, k)) / kappaHdx -
                                        (AT(Ex, i, j, k + 1) - AT(Ex, i, j, k)) / kappaHdz + jy);
                AT(Bz, i, j, k) = AT(Bz, i, j, k) +
                                  dt * ((AT(Ex, i, j + 1, k) - AT(Ex, i, j, k)) / kappaHdy -
                                        (AT(Ey, i + 1, j, k) - AT(Ey, i, j, k)) / kappaHdx + jz);
            }
    for (int c = 0; c < 3; ++c) grid_round(&s->B[c], s->store_mode == 3 ? 2 : s->store_mode);
    apply_boundary(s, s->B);
}

/* FDTD_PLRCSolver::initAccumulator, src/fdtd_plrc.t:240-309 */
static void plrc_init_accumulator(or_sim *s, double dt)
{
    const cplx I = CX(0, 1);
    for (int n = 0; n < 3; ++n) {
        double beta = sqrt(s->d.l_om2[n] - s->d.l_delta[n] * s->d.l_delta[n]);
        double gamma = s->d.l_eps[n] * s->d.l_om2[n] / beta;
        cplx phasor = s->d.l_delta[n] - I * beta;
        s->Crec[n] = cexp(-phasor * dt);
        cplx xi0 = (I * gamma) * (s->Crec[n] * (phasor * dt + 1.0) - 1.0) / ((phasor * phasor) * dt);
        s->dxi0[n] = xi0 * (1.0 - s->Crec[n]);
    }
    for (int i = s->low[0] + 1; i < s->high[0]; ++i)
        for (int j = s->low[1] + 1; j < s->high[1]; ++j)
            for (int k = s->low[2] + 1; k < s->high[2]; ++k) {
                double e[3] = {AT(&s->E[0], i, j, k), AT(&s->E[1], i, j, k), AT(&s->E[2], i, j, k)};
                for (int n = 0; n < 3; ++n)
                    for (int c = 0; c < 3; ++c) {
                        double *pr = &AT(&s->PsiR[n][c], i, j, k), *pi = &AT(&s->PsiI[n][c], i, j, k);
                        cplx p = s->dxi0[n] * e[c] + s->Crec[n] * CX(*pr, *pi);
                        *pr = creal(p);
                        *pi = cimag(p);
                    }
            }
}

/* FDTD_PLRCSolver::stepScheme, src/fdtd_plrc.t:126-168 (no OptFields on this path) */
static void plrc_step_scheme(or_sim *s, double dt)
{
    grid_fill(&s->Sigma, 0);
    if (s->d.plasma) plasma_density_step(s, dt); /* the E-side OptField */
    for (int q = 0; q < s->nE; ++q) current_step(s, &s->cur[s->elist[q]]);
    plrc_stepD(s, dt);
    for (int q = 0; q < s->nH; ++q) current_step(s, &s->cur[s->hlist[q]]);
    plrc_stepB(s, dt);
}

/* FDTD_PLRCSolver::stepSchemeInit, src/fdtd_plrc.t:90-124 */
static void plrc_step_scheme_init(or_sim *s, double dt)
{
    plrc_init_accumulator(s, dt);
    plrc_stepB(s, 0.5 * dt);
    /* PlasmaDensity::stepSchemeInit = stepScheme(dt / 2), src/plasmadensity.h:35 */
    if (s->d.plasma) plasma_density_step(s, 0.5 * dt);
    for (int q = 0; q < s->nE; ++q) {
        current_t *c = &s->cur[s->elist[q]];
        if (c->type == CUR_INC_E) current_step(s, c); /* CPML E, plasma: empty stepSchemeInit */
    }
    for (int q = 0; q < s->nH; ++q) current_step(s, &s->cur[s->hlist[q]]);
}

/* ------------------------------------------------------------------------------------ */
/* Plain, Kerr-only and ADE-Lorentz solvers                                              */
/* ------------------------------------------------------------------------------------ */

/* FDTD_Plain/Nonlinear/Dispersion::stepB share this loop, including the /dx of the two Bz
   terms: src/fdtd_plain.cpp:80-120, src/fdtd_nl.cpp:103-143, src/fdtd_disp.cpp:209-249 */
static void simple_stepB(or_sim *s, double dt)
{
    grid_t *Ex = &s->E[0], *Ey = &s->E[1], *Ez = &s->E[2];
    grid_t *Bx = &s->B[0], *By = &s->B[1], *Bz = &s->B[2];
    double dx = s->dx, dy = s->dy, dz = s->dz;
#pragma omp parallel for schedule(static)
    for (int i = s->low[0]; i < s->high[0]; ++i)
        for (int j = s->low[1]; j < s->high[1]; ++j)
            for (int k = s->low[2]; k < s->high[2]; ++k) {
                AT(Bx, i, j, k) = AT(Bx, i, j, k) + dt * ((AT(Ey, i, j, k + 1) - AT(Ey, i, j, k)) / dz -
                                                          (AT(Ez, i, j + 1, k) - AT(Ez, i, j, k)) / dy);
                AT(By, i, j, k) = AT(By, i, j, k) + dt * ((AT(Ez, i + 1, j, k) - AT(Ez, i, j, k)) / dx -
                                                          (AT(Ex, i, j, k + 1) - AT(Ex, i, j, k)) / dz);
                AT(Bz, i, j, k) = AT(Bz, i, j, k) + dt * ((AT(Ex, i, j + 1, k) - AT(Ex, i, j, k)) / dx -
                                                          (AT(Ey, i + 1, j, k) - AT(Ey, i, j, k)) / dx);
            }
    apply_boundary(s, s->B);
}

/* FDTD_Plain::stepD, src/fdtd_plain.cpp:37-77 (loop includes the upper ghost N+1) */
static void plain_stepD(or_sim *s, double dt)
{
    grid_t *Ex = &s->E[0], *Ey = &s->E[1], *Ez = &s->E[2];
    grid_t *Bx = &s->B[0], *By = &s->B[1], *Bz = &s->B[2];
    double dx = s->dx, dy = s->dy, dz = s->dz;
#pragma omp parallel for schedule(static)
    for (int i = s->low[0] + 1; i <= s->high[0]; ++i)
        for (int j = s->low[1] + 1; j <= s->high[1]; ++j)
            for (int k = s->low[2] + 1; k <= s->high[2]; ++k) {
                AT(Ex, i, j, k) = AT(Ex, i, j, k) + dt * ((AT(Bz, i, j, k) - AT(Bz, i, j - 1, k)) / dy -
                                                          (AT(By, i, j, k) - AT(By, i, j, k - 1)) / dz);
                AT(Ey, i, j, k) = AT(Ey, i, j, k) + dt * ((AT(Bx, i, j, k) - AT(Bx, i, j, k - 1)) / dz -
                                                          (AT(Bz, i, j, k) - AT(Bz, i - 1, j, k)) / dx);
                AT(Ez, i, j, k) = AT(Ez, i, j, k) + dt * ((AT(By, i, j, k) - AT(By, i - 1, j, k)) / dx -
                                                          (AT(Bx, i, j, k) - AT(Bx, i, j - 1, k)) / dy);
            }
    apply_boundary(s, s->E);
}

/* FDTD_Nonlinear::stepD, src/fdtd_nl.cpp:38-100 (E*Dx/D is NaN where D == 0) */
static void nl_stepD(or_sim *s, double dt)
{
    grid_t *Ex = &s->E[0], *Ey = &s->E[1], *Ez = &s->E[2];
    grid_t *Bx = &s->B[0], *By = &s->B[1], *Bz = &s->B[2];
    double dx = s->dx, dy = s->dy, dz = s->dz, eps = s->d.eps, chi = s->d.chi;
#pragma omp parallel for schedule(static)
    for (int i = s->low[0] + 1; i <= s->high[0]; ++i)
        for (int j = s->low[1] + 1; j <= s->high[1]; ++j)
            for (int k = s->low[2] + 1; k <= s->high[2]; ++k) {
                double ex = AT(Ex, i, j, k), ey = AT(Ey, i, j, k), ez = AT(Ez, i, j, k);
                double E2 = ex * ex + ey * ey + ez * ez;
                const double f = eps + (ORACLE_MUTANT == 7 ? 0.0 : chi * E2);
                const double sg = ORACLE_MUTANT == 8 ? -1.0 : 1.0;
                double Dx = f * ex + sg * dt * ((AT(Bz, i, j, k) - AT(Bz, i, j - 1, k)) / dy -
                                                (AT(By, i, j, k) - AT(By, i, j, k - 1)) / dz);
                double Dy = f * ey + sg * dt * ((AT(Bx, i, j, k) - AT(Bx, i, j, k - 1)) / dz -
                                                (AT(Bz, i, j, k) - AT(Bz, i - 1, j, k)) / dx);
                double Dz = f * ez + sg * dt * ((AT(By, i, j, k) - AT(By, i - 1, j, k)) / dx -
                                                (AT(Bx, i, j, k) - AT(Bx, i, j - 1, k)) / dy);
                double E = sqrt(E2);
                double D = sqrt(Dx * Dx + Dy * Dy + Dz * Dz);
                double Eold;
                do {
                    Eold = E;
                    E2 = E * E;
                    E = (2 * chi * E2 * E + D) / (3 * chi * E2 + eps);
                } while (fabs(E - Eold) > 1e-9);
                AT(Ex, i, j, k) = E * Dx / D;
                AT(Ey, i, j, k) = E * Dy / D;
                AT(Ez, i, j, k) = E * Dz / D;
            }
    apply_boundary(s, s->E);
}

/* FDTD_Dispersion::stepD, src/fdtd_disp.cpp:94-206 */
static void disp_stepD(or_sim *s, double dt)
{
    grid_t *Ex = &s->E[0], *Ey = &s->E[1], *Ez = &s->E[2];
    grid_t *Bx = &s->B[0], *By = &s->B[1], *Bz = &s->B[2];
    double dx = s->dx, dy = s->dy, dz = s->dz;
    double dt2 = dt * dt;
#pragma omp parallel for schedule(static)
    for (int i = s->low[0] + 1; i <= s->high[0]; ++i)
        for (int j = s->low[1] + 1; j <= s->high[1]; ++j)
            for (int k = s->low[2] + 1; k <= s->high[2]; ++k) {
                double ex = AT(Ex, i, j, k), ey = AT(Ey, i, j, k), ez = AT(Ez, i, j, k);
                double Px = 0, Py = 0, Pz = 0, Pxp = 0, Pyp = 0, Pzp = 0;
                for (int n = 0; n < 3; ++n) {
                    double *px = &AT(&s->P[n][0], i, j, k), *py = &AT(&s->P[n][1], i, j, k),
                           *pz = &AT(&s->P[n][2], i, j, k);
                    double *pxp = &AT(&s->Pp[n][0], i, j, k), *pyp = &AT(&s->Pp[n][1], i, j, k),
                           *pzp = &AT(&s->Pp[n][2], i, j, k);
                    double a = dt2 * s->d.d_om2[n];
                    double b = a * (ORACLE_MUTANT == 9 ? s->d.eps : s->d.d_eps2[n]);
                    double pxn = (2 - a) * (*px) - (*pxp) + b * ex;
                    double pyn = (2 - a) * (*py) - (*pyp) + b * ey;
                    double pzn = (2 - a) * (*pz) - (*pzp) + b * ez;
                    *pxp = *px;
                    *pyp = *py;
                    *pzp = *pz;
                    *px = pxn;
                    *py = pyn;
                    *pz = pzn;
                    Px += *px;
                    Py += *py;
                    Pz += *pz;
                    Pxp += *pxp;
                    Pyp += *pyp;
                    Pzp += *pzp;
                }
                if (ORACLE_MUTANT == 10) Pxp = Pyp = Pzp = 0;
                double Dx = ex + Pxp - Px + dt * ((AT(Bz, i, j, k) - AT(Bz, i, j - 1, k)) / dy -
                                                  (AT(By, i, j, k) - AT(By, i, j, k - 1)) / dz);
                double Dy = ey + Pyp - Py + dt * ((AT(Bx, i, j, k) - AT(Bx, i, j, k - 1)) / dz -
                                                  (AT(Bz, i, j, k) - AT(Bz, i - 1, j, k)) / dx);
                double Dz = ez + Pzp - Pz + dt * ((AT(By, i, j, k) - AT(By, i - 1, j, k)) / dx -
                                                  (AT(Bx, i, j, k) - AT(Bx, i, j - 1, k)) / dy);
                AT(Ex, i, j, k) = Dx;
                AT(Ey, i, j, k) = Dy;
                AT(Ez, i, j, k) = Dz;
            }
    apply_boundary(s, s->E);
}

/* ------------------------------------------------------------------------------------ */
/* construction                                                                          */
/* ------------------------------------------------------------------------------------ */

static void seterr(char *err, size_t len, const char *msg)
{
    if (err && len) snprintf(err, len, "%s", msg);
}

static current_t *new_current(or_sim *s, int type, int dir)
{
    if (s->ncur >= MAX_CURRENTS) return NULL;
    current_t *c = &s->cur[s->ncur++];
    memset(c, 0, sizeof(*c));
    c->type = type;
    c->dir = dir;
    return c;
}

/* face order used by CPMLBorder::initCurrents (src/cpml_border.cpp:16-51) and
   IncidentSource::initCurrents (src/incsource.cpp:16-77) */
static const int face_order[6] = {NORTH, SOUTH, EAST, WEST, UP, DOWN};

or_sim *or_create(const mpx_desc *desc, char *err, size_t errlen)
{
    if (!desc || desc->nx < 1 || desc->ny < 1 || desc->nz < 1) {
        seterr(err, errlen, "invalid grid size");
        return NULL;
    }
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    if (!s) return NULL;
    s->d = *desc;
    s->solver = desc->solver;
    s->N[0] = desc->nx;
    s->N[1] = desc->ny;
    s->N[2] = desc->nz;
    for (int d = 0; d < 3; ++d) {
        s->glow[d] = 0;
        s->ghigh[d] = s->N[d] + 1;
        s->low[d] = 0;
        s->high[d] = s->N[d] + 1;
    }
    s->dx = desc->dx;
    s->dy = desc->dy;
    s->dz = desc->dz;
    s->dt = desc->dt;

    int isPLRC = (s->solver == MPX_SOLVER_PLRC_NONLIN || s->solver == MPX_SOLVER_PLRC_LIN);
    if (!isPLRC && desc->n_factories > 0) {
        /* FieldSolver::addCurrent default, src/fieldsolver.h:15-23 */
        seterr(err, errlen, "Additional currents cannot be used when using this Field Solver");
        free(s);
        return NULL;
    }
    if (!isPLRC && desc->plasma) {
        seterr(err, errlen, "the plasma block belongs to the FDTD_PLRC solvers");
        free(s);
        return NULL;
    }
    if (isPLRC && (desc->l_om2[0] == 0 || desc->l_om2[1] == 0 || desc->l_om2[2] == 0)) {
        /* src/fdtd_plrc.t:20-24 */
        seterr(err, errlen, "FDTD_PLRCSolver: Omega should not be zero!");
        free(s);
        return NULL;
    }
    for (int c = 0; c < 3; ++c) {
        if (!grid_alloc(&s->E[c], s->low, s->high) || !grid_alloc(&s->B[c], s->low, s->high)) {
            seterr(err, errlen, "out of memory");
            or_destroy(s);
            return NULL;
        }
    }
    if (isPLRC) {
        /* lines are created by the CPML factory or by coreInitStorage, default 1.0 */
        for (int d = 0; d < 3; ++d) {
            line_alloc(&s->kappaE[d], s->low[d], s->high[d], 1.0);
            line_alloc(&s->kappaH[d], s->low[d], s->high[d], 1.0);
        }
        /* currentFactories->initCurrents in setup-file order, src/fdtd_plrc.t:39-45 */
        for (int fi = 0; fi < desc->n_factories; ++fi) {
            const mpx_factory *f = &desc->factories[fi];
            if (f->kind == MPX_FACTORY_CPML || f->kind == MPX_FACTORY_CPML_1D) {
                int one_d = (f->kind == MPX_FACTORY_CPML_1D);
                int faces1d[2] = {UP, DOWN};
                int nf = one_d ? 2 : 6;
                for (int h = 0; h < 2; ++h)
                    for (int q = 0; q < nf; ++q) {
                        int dir = one_d ? faces1d[q] : face_order[q];
                        current_t *c = new_current(s, h ? CUR_CPML_H : CUR_CPML_E, dir);
                        if (!c) {
                            seterr(err, errlen, "too many currents");
                            or_destroy(s);
                            return NULL;
                        }
                        c->thickness = f->d;
                        c->kappaMax = f->kappa_max;
                        c->aMax = f->a_max;
                        c->sigmaMax = f->sigma_max;
                        c->eps = 1;
                        cpml_dims(c);
                        if (h)
                            s->hlist[s->nH++] = (int)(c - s->cur);
                        else
                            s->elist[s->nE++] = (int)(c - s->cur);
                    }
                if (one_d)
                    cpml1d_init_kappa(s, f->d, f->kappa_max);
                else
                    cpml_init_kappa(s, f->d, f->kappa_max);
            } else if (f->kind == MPX_FACTORY_SHORT_PULSE || f->kind == MPX_FACTORY_PLANE_WAVE ||
                       f->kind == MPX_FACTORY_PLANE_GAUSS || f->kind == MPX_FACTORY_FOCUSED_PULSE ||
                       f->kind == MPX_FACTORY_GAUSS_INJECT) {
                /* needCurrent: down face only, src/shortpulseinject.cpp:28-31, src/sources.cpp:8-11 */
                int dir = DOWN;
                for (int h = 0; h < 2; ++h) {
                    current_t *c = new_current(s, h ? CUR_INC_H : CUR_INC_E, dir);
                    if (!c) {
                        seterr(err, errlen, "too many currents");
                        or_destroy(s);
                        return NULL;
                    }
                    c->distance = f->d;
                    inc_dims(c, s->dx, s->dy, s->dz);
                    if (f->kind == MPX_FACTORY_GAUSS_INJECT) {
                        c->src_kind = SRC_GAUSS;
                        if (f->gsize_t < 2 || !gauss_setup(&c->gauss, f, h, s->N[0], s->N[1], s->dx, s->dy, s->dt,
                                                           s->low[2], s->high[2])) {
                            seterr(err, errlen, "gauss_inject: SizeT must be >= 2 (or out of memory)");
                            or_destroy(s);
                            return NULL;
                        }
                    } else if (f->kind == MPX_FACTORY_FOCUSED_PULSE) {
                        c->src_kind = SRC_FOCUSED;
                        if (!focused_setup(&c->focused, f, s->N[0], s->N[1], s->N[2], s->dx, s->dy, s->dz, s->dt)) {
                            seterr(err, errlen, "out of memory");
                            or_destroy(s);
                            return NULL;
                        }
                    } else if (f->kind == MPX_FACTORY_SHORT_PULSE) {
                        double prm[8] = {f->length, f->width, f->om0, f->tshift,
                                         f->zshift, f->phase, f->amp, f->eps};
                        c->src_kind = SRC_PULSE;
                        pulse_setparam(&c->pulse, prm, s->N[0], s->N[1], s->N[2], s->dx, s->dy, s->dz,
                                       s->dt, 1);
                    } else {
                        c->src_kind = SRC_PLANE;
                        plane_setparam(&c->plane, f, f->kind == MPX_FACTORY_PLANE_GAUSS, s->dx, s->dy,
                                       s->dz, s->dt);
                    }
                    if (h)
                        s->hlist[s->nH++] = (int)(c - s->cur);
                    else
                        s->elist[s->nE++] = (int)(c - s->cur);
                }
            } else if (f->kind == MPX_FACTORY_PLASMA_CURRENT) {
                /* PlasmaCurrentFactory::initCurrents, src/plasmacurrent.cpp:18-21 */
                current_t *c = new_current(s, CUR_PLASMA, -1);
                if (!c) {
                    seterr(err, errlen, "too many currents");
                    or_destroy(s);
                    return NULL;
                }
                c->em = f->em;
                c->gamma = f->gamma;
                c->rho_bg = f->rho_bg;
                c->valid = 1;
                for (int q = 0; q < 3; ++q) {
                    if (!s->Jp[q].v && !grid_alloc(&s->Jp[q], s->low, s->high)) {
                        seterr(err, errlen, "out of memory");
                        or_destroy(s);
                        return NULL;
                    }
                    c->J[q] = &s->Jp[q];
                }
                s->elist[s->nE++] = (int)(c - s->cur);
            } else {
                seterr(err, errlen, "unsupported current factory");
                or_destroy(s);
                return NULL;
            }
        }
        /* the plasma density grid, shared by the OptField and the plasma current */
        int needRho = desc->plasma;
        for (int q = 0; q < s->nE; ++q) needRho |= s->cur[s->elist[q]].type == CUR_PLASMA;
        if (needRho && !grid_alloc(&s->Rho, s->low, s->high)) {
            seterr(err, errlen, "out of memory");
            or_destroy(s);
            return NULL;
        }
        /* coreInitStorage, src/fdtd_plrc.cpp:13-86 */
        if (!grid_alloc(&s->Sigma, s->low, s->high)) {
            seterr(err, errlen, "out of memory");
            or_destroy(s);
            return NULL;
        }
        for (int n = 0; n < 3; ++n)
            for (int c = 0; c < 3; ++c)
                if (!grid_alloc(&s->PsiR[n][c], s->low, s->high) || !grid_alloc(&s->PsiI[n][c], s->low, s->high)) {
                    seterr(err, errlen, "out of memory");
                    or_destroy(s);
                    return NULL;
                }
        /* currents' initStorage in list order, then drop invalid ones */
        int ne = 0, nh = 0;
        for (int q = 0; q < s->nE; ++q) {
            current_t *c = &s->cur[s->elist[q]];
            if (c->type == CUR_CPML_E)
                cpml_init_storage(s, c, 0);
            else if (c->type == CUR_INC_E)
                inc_init_storage(s, c, 0);
            if (c->valid) s->elist[ne++] = s->elist[q];
        }
        for (int q = 0; q < s->nH; ++q) {
            current_t *c = &s->cur[s->hlist[q]];
            if (c->type == CUR_CPML_H)
                cpml_init_storage(s, c, 1);
            else
                inc_init_storage(s, c, 1);
            if (c->valid) s->hlist[nh++] = s->hlist[q];
        }
        s->nE = ne;
        s->nH = nh;
        /* JxTotal.. / MxTotal.., src/fdtd_plrc.t:49-75 */
        if (s->nE > 0) {
            s->haveJ = 1;
            for (int c = 0; c < 3; ++c) grid_alloc(&s->J[c], s->low, s->high);
        }
        if (s->nH > 0) {
            s->haveM = 1;
            for (int c = 0; c < 3; ++c) grid_alloc(&s->M[c], s->low, s->high);
        }
        s->haveKappa = 1;
    } else if (s->solver == MPX_SOLVER_DISP) {
        for (int n = 0; n < 3; ++n)
            for (int c = 0; c < 3; ++c)
                if (!grid_alloc(&s->P[n][c], s->low, s->high) || !grid_alloc(&s->Pp[n][c], s->low, s->high)) {
                    seterr(err, errlen, "out of memory");
                    or_destroy(s);
                    return NULL;
                }
    } else if (s->solver != MPX_SOLVER_PLAIN && s->solver != MPX_SOLVER_NL) {
        seterr(err, errlen, "unknown solver");
        or_destroy(s);
        return NULL;
    }
    s->time = 0;
    return s;
}

void or_destroy(or_sim *s)
{
    if (!s) return;
    for (int c = 0; c < 3; ++c) {
        free(s->E[c].v);
        free(s->B[c].v);
        free(s->J[c].v);
        free(s->M[c].v);
        free(s->kappaE[c].v);
        free(s->kappaH[c].v);
        for (int n = 0; n < 3; ++n) {
            free(s->PsiR[n][c].v);
            free(s->PsiI[n][c].v);
            free(s->P[n][c].v);
            free(s->Pp[n][c].v);
        }
    }
    free(s->Sigma.v);
    free(s->Eps.v);
    free(s->Rho.v);
    for (int q = 0; q < 3; ++q) free(s->Jp[q].v);
    for (int q = 0; q < s->nlayers; ++q) free(s->layers[q].g.v);
    for (int q = 0; q < s->ncur; ++q) {
        free(s->cur[q].b.v);
        free(s->cur[q].c.v);
        free(s->cur[q].focused.ex);
        free(s->cur[q].focused.bx);
        free(s->cur[q].focused.by);
        free(s->cur[q].gauss.F1);
        free(s->cur[q].gauss.F2);
    }
    free(s);
}

/* Process::run, src/process.cpp:17-31: the source time index is the step counter */
int or_step(or_sim *s, int nsteps)
{
    double dt = s->dt;
    for (int n = 0; n < nsteps; ++n) {
        switch (s->solver) {
        case MPX_SOLVER_PLRC_NONLIN:
        case MPX_SOLVER_PLRC_LIN:
            plrc_step_scheme(s, dt);
            break;
        case MPX_SOLVER_PLAIN:
            plain_stepD(s, dt);
            simple_stepB(s, dt);
            break;
        case MPX_SOLVER_NL:
            nl_stepD(s, dt);
            simple_stepB(s, dt);
            break;
        case MPX_SOLVER_DISP:
            disp_stepD(s, dt);
            simple_stepB(s, dt);
            break;
        }
        s->time++;
    }
    return 0;
}

/* FieldSolver::stepSchemeInit for each solver */
int or_step_init(or_sim *s)
{
    double dt = s->dt;
    if (s->solver == MPX_SOLVER_PLRC_NONLIN || s->solver == MPX_SOLVER_PLRC_LIN)
        plrc_step_scheme_init(s, dt);
    else
        simple_stepB(s, 0.5 * dt);
    return 0;
}

int or_get_time(or_sim *s) { return s->time; }

int or_set_storage(or_sim *s, int mode)
{
    if (mode < 0 || mode > 3) return -1;
    s->store_mode = mode;
    return 0;
}

double or_store_round(double x, int mode) { return store_round(x, mode); }
void or_set_time(or_sim *s, int t) { s->time = t; }

static grid_t *find_grid(or_sim *s, const char *name)
{
    static const char *eb[6] = {"Ex", "Ey", "Ez", "Bx", "By", "Bz"};
    for (int q = 0; q < 6; ++q)
        if (strcmp(name, eb[q]) == 0) return q < 3 ? &s->E[q] : &s->B[q - 3];
    const char comp[3] = {'x', 'y', 'z'};
    for (int n = 0; n < 3; ++n)
        for (int c = 0; c < 3; ++c) {
            char nm[16];
            snprintf(nm, sizeof nm, "PsiR%d%c", n + 1, comp[c]);
            if (strcmp(name, nm) == 0) return s->PsiR[n][c].v ? &s->PsiR[n][c] : NULL;
            snprintf(nm, sizeof nm, "PsiI%d%c", n + 1, comp[c]);
            if (strcmp(name, nm) == 0) return s->PsiI[n][c].v ? &s->PsiI[n][c] : NULL;
            snprintf(nm, sizeof nm, "P%d%c", n + 1, comp[c]);
            if (strcmp(name, nm) == 0) return s->P[n][c].v ? &s->P[n][c] : NULL;
            snprintf(nm, sizeof nm, "P%d%cp", n + 1, comp[c]);
            if (strcmp(name, nm) == 0) return s->Pp[n][c].v ? &s->Pp[n][c] : NULL;
        }
    static const char *jm[6] = {"JxTotal", "JyTotal", "JzTotal", "MxTotal", "MyTotal", "MzTotal"};
    for (int q = 0; q < 6; ++q)
        if (strcmp(name, jm[q]) == 0) {
            grid_t *g = q < 3 ? &s->J[q] : &s->M[q - 3];
            return g->v ? g : NULL;
        }
    if (strcmp(name, "Sigma") == 0) return s->Sigma.v ? &s->Sigma : NULL;
    if (strcmp(name, "PlasmaDensity") == 0) return s->Rho.v ? &s->Rho : NULL;
    static const char *pj[3] = {"PlasmaJx", "PlasmaJy", "PlasmaJz"};
    for (int q = 0; q < 3; ++q)
        if (strcmp(name, pj[q]) == 0) return s->Jp[q].v ? &s->Jp[q] : NULL;
    if (strcmp(name, "Eps") == 0) {
        if (!s->Eps.v) {
            if (!grid_alloc(&s->Eps, s->low, s->high)) return NULL;
            grid_fill(&s->Eps, s->d.eps);
        }
        return &s->Eps;
    }
    return NULL;
}

int or_get_grid(or_sim *s, const char *name, double *out, size_t count)
{
    grid_t *g = find_grid(s, name);
    if (!g) return -5;
    if (count != grid_count(g)) return -1;
    memcpy(out, g->v, count * sizeof(double));
    return 0;
}

int or_set_grid(or_sim *s, const char *name, const double *in, size_t count)
{
    grid_t *g = find_grid(s, name);
    if (!g) return -5;
    if (count != grid_count(g)) return -1;
    memcpy(g->v, in, count * sizeof(double));
    return 0;
}

int or_get_layer(or_sim *s, const char *name, int dir, double *out, size_t count, int lo[3], int hi[3])
{
    for (int q = 0; q < s->nlayers; ++q)
        if (s->layers[q].dir == dir && strcmp(s->layers[q].name, name) == 0) {
            grid_t *g = &s->layers[q].g;
            for (int d = 0; d < 3; ++d) {
                lo[d] = g->lo[d];
                hi[d] = g->hi[d];
            }
            if (out == NULL) return (int)grid_count(g);
            if (count != grid_count(g)) return -1;
            memcpy(out, g->v, count * sizeof(double));
            return 0;
        }
    return -5;
}

int or_get_line(or_sim *s, const char *name, double *out, size_t count)
{
    static const char *nm[6] = {"KappaEdx", "KappaEdy", "KappaEdz", "KappaHdx", "KappaHdy", "KappaHdz"};
    for (int q = 0; q < 6; ++q)
        if (strcmp(name, nm[q]) == 0) {
            line_t *l = q < 3 ? &s->kappaE[q] : &s->kappaH[q - 3];
            if (!l->v) return -5;
            if (count != (size_t)(l->hi - l->lo + 1)) return -1;
            memcpy(out, l->v, count * sizeof(double));
            return 0;
        }
    return -5;
}

/* FieldEnergyDiag::calculate, src/fielddiag.cpp:255-286 */
double or_energy(or_sim *s)
{
    double energy = 0;
    for (int i = s->low[0] + 1; i < s->high[0]; ++i)
        for (int j = s->low[1] + 1; j < s->high[1]; ++j)
            for (int k = s->low[2] + 1; k < s->high[2]; ++k) {
                double ex = AT(&s->E[0], i, j, k), ey = AT(&s->E[1], i, j, k), ez = AT(&s->E[2], i, j, k);
                double bx = AT(&s->B[0], i, j, k), by = AT(&s->B[1], i, j, k), bz = AT(&s->B[2], i, j, k);
                energy += ex * ex + ey * ey + ez * ez + bx * bx + by * by + bz * bz;
            }
    return energy * (s->dx * s->dy * s->dz);
}

int or_cpml_coeff(or_sim *s, int isH, int dir, double *b, double *c, int *lo, int *hi)
{
    const int *list = isH ? s->hlist : s->elist;
    int n = isH ? s->nH : s->nE;
    for (int q = 0; q < n; ++q) {
        current_t *cu = &s->cur[list[q]];
        if (cu->type == (isH ? CUR_CPML_H : CUR_CPML_E) && cu->dir == dir) {
            *lo = cu->b.lo;
            *hi = cu->b.hi;
            if (b)
                for (int k = cu->b.lo; k <= cu->b.hi; ++k) {
                    b[k - cu->b.lo] = LN(&cu->b, k);
                    c[k - cu->b.lo] = LN(&cu->c, k);
                }
            return 0;
        }
    }
    return -5;
}
