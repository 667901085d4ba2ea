// Device math for the MPulse time-stepping core on gfx950: complex float64 arithmetic,
// the Faddeeva function and the closed-form short-pulse fields.  Everything here is
// float64 and evaluated per source-sheet cell, never inside the bulk stencil.
#pragma once
#include <hip/hip_runtime.h>

namespace mpx {

struct cd {
    double re, im;
};

__device__ __forceinline__ cd mk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator-(cd a) { return cd{-a.re, -a.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b)
{
    return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd operator*(cd a, double b) { return cd{a.re * b, a.im * b}; }
__device__ __forceinline__ cd operator*(double a, cd b) { return cd{a * b.re, a * b.im}; }
__device__ __forceinline__ cd operator/(cd a, double b) { return cd{a.re / b, a.im / b}; }
__device__ __forceinline__ cd operator+(cd a, double b) { return cd{a.re + b, a.im}; }
__device__ __forceinline__ cd operator+(double a, cd b) { return cd{a + b.re, b.im}; }
__device__ __forceinline__ cd operator-(cd a, double b) { return cd{a.re - b, a.im}; }
__device__ __forceinline__ cd operator-(double a, cd b) { return cd{a - b.re, -b.im}; }
// Complex quotient with Smith's scaling (R. L. Smith, CACM 5(8), 1962): avoids the overflow
// and most of the cancellation of the textbook formula.
__device__ __forceinline__ cd operator/(cd a, cd b)
{
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re;
        double den = b.re + b.im * r;
        return cd{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
    }
    double r = b.re / b.im;
    double den = b.re * r + b.im;
    return cd{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}
__device__ __forceinline__ cd operator/(double a, cd b) { return cd{a, 0.0} / b; }
__device__ __forceinline__ cd cexp_(cd z)
{
    double e = exp(z.re);
    double s, c;
    sincos(z.im, &s, &c);
    return cd{e * c, e * s};
}

// Faddeeva-function node tables: n_k = (k h)^2 and e_k = exp(-n_k) for h = 0.5 and 0.53,
// k = 1..12, filled once from the host.
struct FaddeevaTables {
    double n[2][12];
    double e[2][12];
};

// w(z) = exp(-z^2) erfc(-iz) by the trapezoid rule of Matta & Reichel (Math. Comp. 25, 1971)
// in Gammel's arrangement, the scheme of Faddeeva_2 (src/specfunc.cpp:57-157): step h = 0.5,
// or h = 0.53 when z is close to a pole of the h = 0.5 rule, plus the pole correction for
// Im z < 2 pi.  For |z|^2 < 1e-7 a Maclaurin series is used.
__device__ inline cd faddeeva(cd z, const FaddeevaTables &T)
{
    double x = z.re, y = z.im;
    double s = x * x + y * y;
    if (s < 1e-7) {
        const double ig[8] = {1.0, 1.1283791670955126, 1.0, 0.75225277806367508, 0.5,
                              0.30090111122547003, 0.16666666666666666, 0.085971746064420009};
        cd iz = cd{-y, x};
        cd term = cd{1.0, 0.0}, acc = cd{0.0, 0.0};
        for (int n = 0; n < 8; ++n) {
            acc = acc + ig[n] * term;
            term = term * iz;
        }
        return acc;
    }
    int tab = 0;
    double r = 0.5 / M_PI;
    if (fabs(y) < 0.01 && fabs(x) < 6.01) {
        double ip;
        double fr = modf(2 * fabs(x), &ip);
        if (fr < 0.02 || fr > 0.98) {
            tab = 1;
            r = 0.53 / M_PI;
        }
    }
    double d = (y - x) * (y + x);
    double f = 4 * x * x * y * y;
    double g = 0.0, h = 0.0;
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        double nq = T.n[tab][q];
        double t = d + nq;
        double u = T.e[tab][q] / (t * t + f);
        g += (s + nq) * u;
        h += (s - nq) * u;
    }
    double u = 1 / s;
    cd c = r * cd{y * (u + 2.0 * g), x * (u + 2.0 * h)};
    if (y < 2.0 * M_PI) {
        double s2 = 2.0 / r;
        double t = s2 * x;
        double uu = s2 * y;
        double sn, cs;
        sincos(t, &sn, &cs);
        double ff = exp(-uu) - cs;
        double gg = 2.0 * exp(d - uu) / (sn * sn + ff * ff);
        double u2 = 2.0 * x * y;
        double t2, h2;
        sincos(u2, &t2, &h2);
        c = c + gg * cd{h2 * ff - t2 * sn, -(h2 * sn + t2 * ff)};
    }
    return c;
}

// Parameters of one short-pulse source after ShortPulseInjectSourceFunc::setParam
// (src/shortpulseinject.cpp:56-113); the amplitude normalisation is done on the host.
struct PulseParams {
    double DX, DY, DZ, DT;
    double lightspeed, length, om0, TShift, ZShift, Phase, ZRl;
    double centrex, centrey, centrez;
    double amp;
};

// Efunc, src/shortpulsefunctions.cpp:28-94
__device__ inline cd pulse_efunc(const PulseParams &P, const FaddeevaTables &FT, double x, double y,
                                 double z, double t)
{
    const cd I = mk(0, 1), I2 = mk(0, 2), nI = mk(0, -1);
    double c = P.lightspeed, p = P.Phase, T = P.length;
    double T2 = T * T, T4 = T2 * T2, tb = 0.5 * P.om0 * T;
    double r2 = x * x + y * y, ZRl = P.ZRl;
    cd q = mk(z, ZRl);
    cd rT = ZRl / q;
    cd rpiT = sqrt(M_PI) * rT;
    cd rpT = -ZRl / (q * q);
    cd eiphi = cexp_((2.0 * I) * p);
    cd tpT = (t - z / c) - r2 / ((2. * c) * q);
    cd tpT2 = tpT * tpT;
    cd tpzT = -1 / c - r2 / (((2. * c) * q) * q);
    cd an = nI * tb + tpT / T;
    cd ap = I * tb + tpT / T;
    cd A = cexp_(nI * p) / (rpiT * T4);
    cd B = (rT * T) * ((((I * (1.0 + eiphi)) * rpT) * T2) -
                       (rT * (T * ((nI - 2 * tb) + eiphi * (nI + 2 * tb)) + (I2 * (1.0 + eiphi)) * tpT)) * tpzT);
    cd C = (((I * eiphi) * rpiT) * (T * tb + I * tpT)) *
           ((rpT * T2 + ((((I2 * rT) * T) * tb) * tpzT)) - ((2.0 * rT) * tpT) * tpzT);
    cd D = (rpiT * (T * tb - I * tpT)) * ((I * rpT) * T2 + ((2.0 * rT) * (T * tb - I * tpT)) * tpzT);
    cd S = mk(0, 0);
    if (an.im < 0) {
        S = S + ((2.0 * A) * C) * cexp_(((((2.0 * I) * tpT) * tb) / T) - tpT2 / T2);
        an = -an;
        C = -C;
    }
    if (ap.im < 0) {
        S = S + ((2.0 * A) * D) * cexp_(((((-2.0 * I) * tpT) * tb) / T) - tpT2 / T2);
        ap = -ap;
        D = -D;
    }
    cd Ef = S + (A * exp(-tb * tb)) * ((B + C * faddeeva(an, FT)) - D * faddeeva(ap, FT));
    return (-I) * Ef;
}

// Bfunc, src/shortpulsefunctions.cpp:96-226.  YComp is 0 (src/shortpulseinject.cpp:96), so the
// x component needs the A group only and the y component the B and D groups.
__device__ inline cd pulse_bfunc(const PulseParams &P, const FaddeevaTables &FT, double x, double y,
                                 double z, double t, bool bx)
{
    const cd I = mk(0, 1), I2 = mk(0, 2), nI = mk(0, -1);
    double c = P.lightspeed, p = P.Phase, T = P.length;
    double T2 = T * T, T4 = T2 * T2, tb = 0.5 * P.om0 * T;
    double r2 = x * x + y * y, ZRl = P.ZRl;
    double sqpi = sqrt(M_PI);
    cd q = mk(z, ZRl);
    cd rT = ZRl / q;
    cd rT2 = rT * rT;
    cd rT4T = rT * T4;
    cd rpiT = sqpi * rT;
    cd rpT = -ZRl / (q * q);
    cd rppT = (2. * ZRl) / ((q * q) * q);
    cd eiphi = cexp_((2.0 * I) * p);
    cd tpT = (t - z / c) - r2 / ((2. * c) * q);
    cd tpT2 = tpT * tpT;
    cd tpzT = -1 / c + r2 / (((2. * c) * q) * q);
    cd tpzT2 = tpzT * tpzT;
    cd tpxT = -x / (c * q);
    cd tpyT = -y / (c * q);
    cd tpxyT = mk(0, 0);
    cd tpxxT = -1. / (c * q);
    cd tpzzT = -r2 / (((c * q) * q) * q);
    cd tpxT2 = tpxT * tpxT;
    cd aip = T * tb + I * tpT;
    cd ain = T * tb - I * tpT;
    cd aip2 = aip * aip;
    cd ain2 = ain * ain;
    cd an = nI * tb + tpT / T;
    cd ap = I * tb + tpT / T;
    cd bracket = T * ((nI - 2 * tb) + eiphi * (nI + 2 * tb)) + (I2 * (1.0 + eiphi)) * tpT;
    bool flip_n = an.im < 0, flip_p = ap.im < 0;
    cd e2 = mk(0, 0), e2m = mk(0, 0);
    if (flip_n) {
        e2 = cexp_(((((2.0 * I) * tpT) * tb) / T) - tpT2 / T2);
        an = -an;
    }
    if (flip_p) {
        e2m = cexp_(((((-2.0 * I) * tpT) * tb) / T) - tpT2 / T2);
        ap = -ap;
    }
    cd wn = faddeeva(an, FT);
    cd wp = faddeeva(ap, FT);
    if (bx) {
        cd AA = cexp_(nI * p) / (sqpi * T4);
        cd AB = ((((-tpxT) * tpyT) * rT) * T) * bracket + (((((I * T2) * tpxyT) * (1.0 + eiphi)) * rT) * T);
        cd AC = (((((-tpxT) * tpyT) * 2.0) * eiphi) * rpiT) * aip2 + ((((I * T2) * tpxyT) * eiphi) * rpiT) * aip;
        cd AD = ((((-tpxT) * tpyT) * 2.0) * rpiT) * ain2 + (((I * T2) * tpxyT) * rpiT) * (-(T * tb) + I * tpT);
        cd AS = mk(0, 0);
        if (flip_n) {
            AS = AS + ((2.0 * AA) * AC) * e2;
            AC = -AC;
        }
        if (flip_p) {
            AS = AS + ((2.0 * AA) * AD) * e2m;
            AD = -AD;
        }
        cd BfA = AS + (AA * exp(-tb * tb)) * ((AB + AC * wn) - AD * wp);
        return I * BfA;
    }
    cd BA = cexp_(nI * p) / (sqpi * T4);
    cd BB = (-tpxT2) * ((rT * T) * bracket) + (((((I * T2) * tpxxT) * (1.0 + eiphi)) * rT) * T);
    cd BC = ((((-tpxT2) * 2.0) * eiphi) * rpiT) * aip2 + ((((I * T2) * tpxxT) * eiphi) * rpiT) * aip;
    cd BD = (((-tpxT2) * 2.0) * rpiT) * ain2 + (((I * T2) * tpxxT) * rpiT) * (-(T * tb) + I * tpT);
    cd DA = cexp_(nI * p) / ((2. * sqpi) * T4);
    cd DB = (2 * T) * (((((I2 * (1.0 + eiphi)) * rpT) * T2) * tpzT) -
                       rT * ((T * ((nI - 2 * tb) + eiphi * (nI + 2 * tb))) * tpzT2 +
                             (I * (1.0 + eiphi)) * (((2.0 * tpT) * tpzT2) - T2 * tpzzT)));
    cd DC = ((eiphi * rpiT) *
             (rppT * rT4T + ((I2 * rT) * aip) *
                                ((((2.0 * rpT) * T2) * tpzT) +
                                 rT * (((((I2 * T) * tb) * tpzT2) - ((2.0 * tpT) * tpzT2)) + T2 * tpzzT)))) /
            rT2;
    cd DD = (rpiT * (rppT * rT4T - ((2.0 * rT) * ain) *
                                       ((((I2 * rpT) * T2) * tpzT) +
                                        rT * (((((2 * T) * tb) * tpzT2) - ((I2 * tpT) * tpzT2)) + (I * T2) * tpzzT)))) /
            rT2;
    cd BS = mk(0, 0), DS = mk(0, 0);
    if (flip_n) {
        BS = BS + ((2.0 * BA) * BC) * e2;
        DS = DS + ((2.0 * DA) * DC) * e2;
        BC = -BC;
        DC = -DC;
    }
    if (flip_p) {
        BS = BS + ((2.0 * BA) * BD) * e2m;
        DS = DS + ((2.0 * DA) * DD) * e2m;
        BD = -BD;
        DD = -DD;
    }
    cd BfB = BS + (BA * exp(-tb * tb)) * ((BB + BC * wn) - BD * wp);
    cd BfD = DS + (DA * exp(-tb * tb)) * ((DB + DC * wn) - DD * wp);
    return I * (BfD - BfB);
}

}  // namespace mpx
