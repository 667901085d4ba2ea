// gfx950 kernels of the MPulse time-stepping core.
//
// Two kernel structures advance the same state:
//  * the reference-structured path, one kernel per loop of FDTD_PLRCSolver::stepScheme
//    (src/fdtd_plrc.t:127-168): CPML currents, source sheets, current totals, E update,
//    ghost wrap, and the same for the magnetic side;
//  * the fused path, one E-side kernel (Yee curl + PLRC pole recursion + Kerr Newton solve +
//    CPML E auxiliaries + source sheet) and one B-side kernel (curl + CPML H auxiliaries +
//    source sheet), with the periodic ghost wrap done by index arithmetic.
// Threads run along z (the contiguous dimension), 64 per wavefront, so every field load of a
// wavefront is one coalesced 256/512-byte segment.
#include <hip/hip_runtime.h>

#include "mpx_device.h"
#include "mpx_launch.h"

namespace mpx {

// ------------------------------------------------------------------------------------
// source sheets: IncidentSource{E,H}Current::stepScheme, src/incsource.t:40-73, 112-145
// ------------------------------------------------------------------------------------

__device__ inline void plane_getH(const PlaneParams &W, int i, int j, int l, int time, double out[3])
{
    // PlaneWaveSourceEFunc::getHField src/sources.cpp:107-125 / PlaneGaussSourceEFunc src/sources.cpp:266-284
    double realtime = W.dt * (time - 0.5);
    double posx = W.k[0] * i * W.dx + W.k[1] * (j + 0.5) * W.dy + W.k[2] * (l + 0.5) * W.dz + W.om * realtime;
    double posy = W.k[0] * (i + 0.5) * W.dx + W.k[1] * j * W.dy + W.k[2] * (l + 0.5) * W.dz + W.om * realtime;
    double posz = W.k[0] * (i + 0.5) * W.dx + W.k[1] * (j + 0.5) * W.dy + W.k[2] * l * W.dz + W.om * realtime;
    double pos[3] = {posx, posy, posz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!W.gauss) {
            double h = W.H[d] * sin(pos[d]) + W.H_bg[d];
            if (pos[d] < W.ramp) h *= pos[d] / W.ramp;
            out[d] = h;
        } else {
            double q = (pos[d] - W.offset) / W.width;
            out[d] = (W.H[d] * exp(-(q * q))) * sin(pos[d]);
        }
    }
}

__device__ inline void plane_getE(const PlaneParams &W, int i, int j, int l, int time, double out[3])
{
    // PlaneWaveSourceHFunc::getEField src/sources.cpp:150-173 / PlaneGaussSourceHFunc src/sources.cpp:308-331
    double realtime = W.dt * time;
    double posx = W.k[0] * (i + 0.5) * W.dx + W.k[1] * j * W.dy + W.k[2] * l * W.dz + W.om * realtime;
    double posy = W.k[0] * i * W.dx + W.k[1] * (j + 0.5) * W.dy + W.k[2] * l * W.dz + W.om * realtime;
    double posz = W.k[0] * i * W.dx + W.k[1] * j * W.dy + W.k[2] * (l + 0.5) * W.dz + W.om * realtime;
    double pos[3] = {posx, posy, posz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!W.gauss) {
            double e = W.E[d] * sin(pos[d]) + W.E_bg[d];
            if (pos[d] < W.ramp) e *= pos[d] / W.ramp;
            out[d] = e;
        } else {
            double q = (pos[d] - W.offset) / W.width;
            out[d] = (W.E[d] * exp(-(q * q))) * sin(pos[d]);
        }
    }
}

// one sheet cell (i, j) at step `time`; KIND > 0 fixes the source kind at compile time (each
// kind then gets only the registers it needs), KIND = 0 reads it from the descriptor
template <int KIND = 0>
__device__ __forceinline__ void sheet_eval_point(const SheetEval &a, const FaddeevaTables &ft, int i, int j, int time)
{
    const int kind = KIND > 0 ? KIND : a.kind;
    int z = a.sheet.kp + a.off;
    double f[3] = {0, 0, 0};
    const PulseParams &P = a.pulse;
    if (kind == 1) {
        if (!a.isH) {
            // getHField, src/shortpulseinject.cpp:151-178
            double posxo = (i - P.centrex) * P.DX;
            double posxh = (i + 0.5 - P.centrex) * P.DX;
            double posyo = (j - P.centrey) * P.DY;
            double posyh = (j + 0.5 - P.centrey) * P.DY;
            double poszh = (z - 0.5 - P.centrez) * P.DZ - P.ZShift;
            double posTime = (time - 0.5) * P.DT - P.TShift;
            f[0] = P.amp * pulse_bfunc(P, ft, posxo, posyh, poszh, posTime, true).re;
            f[1] = P.amp * pulse_bfunc(P, ft, posxh, posyo, poszh, posTime, false).re;
        } else {
            // getEField, src/shortpulseinject.cpp:126-149
            double posxh = (i + 0.5 - P.centrex) * P.DX;
            double posyo = (j - P.centrey) * P.DY;
            double poszo = (z - P.centrez) * P.DZ - P.ZShift;
            double posTime = time * P.DT - P.TShift;
            f[0] = P.amp * pulse_efunc(P, ft, posxh, posyo, poszo, posTime).re;
        }
    } else if (kind == 4) {
        // GaussInjectSourceFunc::getHField / getEField, src/gaussinject.cpp:203-224
        const GaussSheetParams &G = a.gauss;
        if (G.active) {
            const size_t q = ((size_t)(time - G.blockToff) * a.nx2 + i) * a.ny2 + j;
            if (!a.isH) {
                f[0] = G.sqeps * G.amp * G.F1[q];
            } else {
                f[0] = G.amp * G.F1[q];
                f[1] = G.amp * G.F2[q];
            }
        }
        (void)z;
    } else if (kind == 3) {
        // amp Re[env(t - z/v) U exp(-i(om0 t - phase))], env(s) = exp(-s^2/T^2)
        const FocusedParams &F = a.focused;
        size_t p = (size_t)i * a.ny2 + j;
        double t = a.isH ? time * F.DT - F.TShift : (time - 0.5) * F.DT - F.TShift;
        double z = a.isH ? F.zE : F.zB;
        double s = t - z / F.v;
        double env = F.amp * exp(-(s * s) / (F.T * F.T));
        double sn, cs;
        sincos(-(F.om0 * t - F.phase), &sn, &cs);
        if (a.isH) {
            double2 u = F.ex[p];
            f[0] = env * (u.x * cs - u.y * sn);
        } else {
            double2 u = F.bx[p], w = F.by[p];
            f[0] = env * (u.x * cs - u.y * sn);
            f[1] = env * (w.x * cs - w.y * sn);
        }
        (void)z;
    } else {
        if (!a.isH)
            plane_getH(a.plane, i, j, z, time, f);
        else
            plane_getE(a.plane, i, j, z, time, f);
    }
    size_t q = (size_t)i * a.ny2 + j;
    a.sheet.J[a.t1][q] = f[a.t2] / a.DX;
    a.sheet.J[a.t2][q] = f[a.t1] / a.DX;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_sheet_eval(SheetEval a, FaddeevaTables ft)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= a.ny2 || i >= a.nx2) return;
    sheet_eval_point<KIND>(a, ft, i, j, a.time);
}

// The sheets of many runs in one launch.  Sheet s owns blocks first[s] .. first[s+1]-1, one
// block per 64 cells of a row; the step is *time + tstep (the counter lives on the device so
// that a captured graph can be replayed at later steps), hshift less for the H currents.  Only
// the sheets of the sides in `sides` (bit 0: E currents, bit 1: H currents) are evaluated.
__global__ __launch_bounds__(64) void k_sheet_eval_batch(const SheetEval *__restrict__ list,
                                                         const int *__restrict__ first, int n, FaddeevaTables ft,
                                                         const int *__restrict__ time, int tstep, int hshift,
                                                         int sides)
{
    const int b = blockIdx.x;
    int lo = 0, hi = n - 1;  // last sheet with first[s] <= b
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const SheetEval &a = list[lo];
    if (!(sides >> (a.isH ? 1 : 0) & 1)) return;
    const int nyb = (a.ny2 + 63) >> 6;
    const int local = b - first[lo];
    const int i = local / nyb, j = (local - i * nyb) * 64 + threadIdx.x;
    if (i >= a.nx2 || j >= a.ny2) return;
    sheet_eval_point(a, ft, i, j, *time + tstep - (a.isH ? hshift : 0));
}

__global__ void k_advance_time(int *time, int k)
{
    if (threadIdx.x == 0) *time += k;
}

void launch_sheet_eval_batch(const SheetEval *list, const int *first, int n, int blocks, const FaddeevaTables &ft,
                             const int *time, int tstep, hipStream_t s, int hshift, int sides)
{
    hipLaunchKernelGGL(k_sheet_eval_batch, dim3(blocks), dim3(64), 0, s, list, first, n, ft, time, tstep, hshift,
                       sides);
}

void launch_advance_time(int *time, int k, hipStream_t s)
{
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(64), 0, s, time, k);
}

void launch_sheet_eval(const SheetEval &a, const FaddeevaTables &ft, hipStream_t s)
{
    dim3 blk(64);
    dim3 grd((a.ny2 + 63) / 64, a.nx2);
    switch (a.kind) {
    case 1: hipLaunchKernelGGL(k_sheet_eval<1>, grd, blk, 0, s, a, ft); break;
    case 3: hipLaunchKernelGGL(k_sheet_eval<3>, grd, blk, 0, s, a, ft); break;
    case 4: hipLaunchKernelGGL(k_sheet_eval<4>, grd, blk, 0, s, a, ft); break;
    default: hipLaunchKernelGGL(k_sheet_eval<2>, grd, blk, 0, s, a, ft); break;
    }
}

// ------------------------------------------------------------------------------------
// reference-structured path
// ------------------------------------------------------------------------------------

// CPMLBorderECurrent::stepScheme, src/cpml_border.cpp:398-425
__global__ __launch_bounds__(256) void k_ref_cpml_E(StepArgs<double> a, int fi)
{
    const CpmlFace &f = a.face[fi];
    int k = f.lo[2] + blockIdx.x * 64 + threadIdx.x;
    int j = f.lo[1] + blockIdx.y * 4 + threadIdx.y;
    int i = f.lo[0] + blockIdx.z;
    if (k > f.hi[2] || j > f.hi[1]) return;
    int ind[3] = {i, j, k};
    int m[3] = {i, j, k};
    m[f.dim] -= 1;
    int q = ind[f.dim] - f.lo[f.dim];
    const GridDims &g = a.g;
    size_t c = g.at(i, j, k), cm = g.at(m[0], m[1], m[2]);
    const double *B0 = a.B[f.f0], *B1 = a.B[f.f1];
    double *p0 = (double *)f.psi0, *p1 = (double *)f.psi1;
    size_t l = f.at(i, j, k);
    p0[l] = f.b[q] * p0[l] - f.c[q] * (B1[c] - B1[cm]) / f.dxn;
    p1[l] = f.b[q] * p1[l] + f.c[q] * (B0[c] - B0[cm]) / f.dxn;
}

// CPMLBorderHCurrent::stepScheme, src/cpml_border.cpp:499-527
__global__ __launch_bounds__(256) void k_ref_cpml_H(StepArgs<double> a, int fi)
{
    const CpmlFace &f = a.face[fi];
    int k = f.lo[2] + blockIdx.x * 64 + threadIdx.x;
    int j = f.lo[1] + blockIdx.y * 4 + threadIdx.y;
    int i = f.lo[0] + blockIdx.z;
    if (k > f.hi[2] || j > f.hi[1]) return;
    int ind[3] = {i, j, k};
    int p[3] = {i, j, k};
    p[f.dim] += 1;
    int q = ind[f.dim] - f.lo[f.dim];
    const GridDims &g = a.g;
    size_t c = g.at(i, j, k), cp = g.at(p[0], p[1], p[2]);
    const double *E0 = a.E[f.f0], *E1 = a.E[f.f1];
    double *p0 = (double *)f.psi0, *p1 = (double *)f.psi1;
    size_t l = f.at(i, j, k);
    p0[l] = f.b[q] * p0[l] + f.c[q] * (E1[cp] - E1[c]) / f.dxn;
    p1[l] = f.b[q] * p1[l] - f.c[q] * (E0[cp] - E0[c]) / f.dxn;
}

void launch_ref_cpml(const StepArgs<double> &a, int fi, bool isH, hipStream_t s)
{
    const CpmlFace &f = a.face[fi];
    dim3 blk(64, 4);
    dim3 grd((f.hi[2] - f.lo[2] + 64) / 64, (f.hi[1] - f.lo[1] + 4) / 4, f.hi[0] - f.lo[0] + 1);
    if (isH)
        hipLaunchKernelGGL(k_ref_cpml_H, grd, blk, 0, s, a, fi);
    else
        hipLaunchKernelGGL(k_ref_cpml_E, grd, blk, 0, s, a, fi);
}

// FDTD_PLRCSolver::sumCurrents / sumMagCurrents, src/fdtd_plrc.t:342-409: totals := 0, then
// += every current's layer in list order, cell by cell.
__global__ __launch_bounds__(256) void k_ref_sum(GridDims g, SumList L, double *T0, double *T1, double *T2)
{
    int k = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int i = blockIdx.z;
    if (k > g.n[2] + 1 || j > g.n[1] + 1) return;
    double t[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < L.n; ++v) {
        const SumView &w = L.v[v];
        if (i < w.lo[0] || i > w.hi[0] || j < w.lo[1] || j > w.hi[1] || k < w.lo[2] || k > w.hi[2]) continue;
        size_t l = w.padded ? g.at(i, j, k)
                            : ((size_t)(i - w.lo[0]) * w.n1 + (size_t)(j - w.lo[1])) * w.n2 + (size_t)(k - w.lo[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (w.comp[c]) t[c] += w.comp[c][l];
    }
    size_t c = g.at(i, j, k);
    T0[c] = t[0];
    T1[c] = t[1];
    T2[c] = t[2];
}

void launch_ref_sum(const GridDims &g, const SumList &L, double *const T[3], hipStream_t s)
{
    dim3 blk(64, 4);
    dim3 grd((g.n[2] + 2 + 63) / 64, (g.n[1] + 2 + 3) / 4, g.n[0] + 2);
    hipLaunchKernelGGL(k_ref_sum, grd, blk, 0, s, g, L, T[0], T[1], T[2]);
}

// Storage::applyBoundary + SinglePeriodicBoundary::exchange{X,Y,Z}, src/storage.cpp:194-204,
// src/periodic_bound.cpp:10-71: f(0) = f(N), f(N+1) = f(1) over the full planes.
template <typename T>
__global__ __launch_bounds__(256) void k_wrap(GridDims g, T *f0, T *f1, T *f2, int dim)
{
    T *f = blockIdx.z == 0 ? f0 : (blockIdx.z == 1 ? f1 : f2);
    if (!f) return;  // a group of one array (the residual words of float32 + residual storage)
    int a = blockIdx.x * 64 + threadIdx.x;  // fastest remaining dimension
    int b = blockIdx.y * 4 + threadIdx.y;
    int N = g.n[dim];
    if (dim == 0) {
        if (a > g.n[2] + 1 || b > g.n[1] + 1) return;
        f[g.at(0, b, a)] = f[g.at(N, b, a)];
        f[g.at(N + 1, b, a)] = f[g.at(1, b, a)];
    } else if (dim == 1) {
        if (a > g.n[2] + 1 || b > g.n[0] + 1) return;
        f[g.at(b, 0, a)] = f[g.at(b, N, a)];
        f[g.at(b, N + 1, a)] = f[g.at(b, 1, a)];
    } else {
        if (a > g.n[1] + 1 || b > g.n[0] + 1) return;
        f[g.at(b, a, 0)] = f[g.at(b, a, N)];
        f[g.at(b, a, N + 1)] = f[g.at(b, a, 1)];
    }
}

template <typename T>
void launch_wrap(const GridDims &g, T *const f[3], int dim, hipStream_t s)
{
    int na = dim == 2 ? g.n[1] + 2 : g.n[2] + 2;
    int nb = dim == 0 ? g.n[1] + 2 : g.n[0] + 2;
    dim3 blk(64, 4);
    dim3 grd((na + 63) / 64, (nb + 3) / 4, 3);
    hipLaunchKernelGGL(k_wrap<T>, grd, blk, 0, s, g, f[0], f[1], f[2], dim);
}
template void launch_wrap<double>(const GridDims &, double *const[3], int, hipStream_t);
template void launch_wrap<float>(const GridDims &, float *const[3], int, hipStream_t);
template void launch_wrap<unsigned>(const GridDims &, unsigned *const[3], int, hipStream_t);

// FDTD_PLRCSolver::stepD loop, src/fdtd_plrc.t:221-233, reading the current totals
template <int NP, bool NONLIN>
__global__ __launch_bounds__(256) void k_ref_stepD(StepArgs<double> a)
{
    int k = blockIdx.x * 64 + threadIdx.x + 1;
    int j = blockIdx.y * 4 + threadIdx.y + 1;
    int i = blockIdx.z + 1;
    const GridDims &g = a.g;
    if (k > g.n[2] || j > g.n[1]) return;
    size_t c = g.at(i, j, k);
    const double *Bx = a.B[0], *By = a.B[1], *Bz = a.B[2];
    double bx = Bx[c], by = By[c], bz = Bz[c];
    double dBz_y = bz - Bz[g.at(i, j - 1, k)];
    double dBy_z = by - By[g.at(i, j, k - 1)];
    double dBx_z = bx - Bx[g.at(i, j, k - 1)];
    double dBz_x = bz - Bz[g.at(i - 1, j, k)];
    double dBy_x = by - By[g.at(i - 1, j, k)];
    double dBx_y = bx - Bx[g.at(i, j - 1, k)];
    double jx = 0, jy = 0, jz = 0;
    if (a.haveJ) {
        jx = a.Jt[0][c];
        jy = a.Jt[1][c];
        jz = a.Jt[2][c];
    }
    double ex = a.E[0][c], ey = a.E[1][c], ez = a.E[2][c];
    double denom, numer;
    cell_eps<NONLIN>(a, c, a.plasma ? 0.5 * a.sigma[c] * a.dt : 0.0, denom, numer);
    unsigned fl = 0;
    plrc_e_update<NP, NONLIN, false>(ex, ey, ez, dBz_y, dBy_z, dBx_z, dBz_x, dBy_x, dBx_y, a.kdE[0][i], a.kdE[1][j],
                                     a.kdE[2][k], jx, jy, jz, a.psi, c, a.pc, a.dt, denom, numer, a.chi, a.ncap, &fl);
    report(a.status, fl | finite_flag(ex, ey, ez));
    a.E[0][c] = ex;
    a.E[1][c] = ey;
    a.E[2][c] = ez;
}

// FDTD_PLRCSolver::stepB + plrcStepB, src/fdtd_plrc.t:312-339, src/fdtd_plrc.cpp:429-459;
// Jt holds the magnetic current totals here.
__global__ __launch_bounds__(256) void k_ref_stepB(StepArgs<double> a)
{
    int k = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int i = blockIdx.z;
    const GridDims &g = a.g;
    if (k > g.n[2] || j > g.n[1]) return;
    size_t c = g.at(i, j, k);
    const double *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
    double jx = 0, jy = 0, jz = 0;
    if (a.haveJ) {
        jx = a.Jt[0][c];
        jy = a.Jt[1][c];
        jz = a.Jt[2][c];
    }
    double kdx = a.kdH[0][i], kdy = a.kdH[1][j], kdz = a.kdH[2][k];
    double ex = Ex[c], ey = Ey[c], ez = Ez[c];
    a.B[0][c] = a.B[0][c] + a.dt * ((Ey[g.at(i, j, k + 1)] - ey) / kdz - (Ez[g.at(i, j + 1, k)] - ez) / kdy + jx);
    a.B[1][c] = a.B[1][c] + a.dt * ((Ez[g.at(i + 1, j, k)] - ez) / kdx - (Ex[g.at(i, j, k + 1)] - ex) / kdz + jy);
    a.B[2][c] = a.B[2][c] + a.dt * ((Ex[g.at(i, j + 1, k)] - ex) / kdy - (Ey[g.at(i + 1, j, k)] - ey) / kdx + jz);
}

void launch_ref_stepD(const StepArgs<double> &a, int np, bool nonlin, hipStream_t s)
{
    const GridDims &g = a.g;
    dim3 blk(64, 4);
    dim3 grd((g.n[2] + 63) / 64, (g.n[1] + 3) / 4, g.n[0]);
#define MPX_REF_D(NPV, NL) hipLaunchKernelGGL((k_ref_stepD<NPV, NL>), grd, blk, 0, s, a)
    if (nonlin) {
        switch (np) {
        case 0: MPX_REF_D(0, true); break;
        case 1: MPX_REF_D(1, true); break;
        case 2: MPX_REF_D(2, true); break;
        default: MPX_REF_D(3, true); break;
        }
    } else {
        switch (np) {
        case 0: MPX_REF_D(0, false); break;
        case 1: MPX_REF_D(1, false); break;
        case 2: MPX_REF_D(2, false); break;
        default: MPX_REF_D(3, false); break;
        }
    }
#undef MPX_REF_D
}

// PlasmaDensity over the whole storage box 0..N+1 (Sigma was cleared: it becomes w I^(K-1))
__global__ __launch_bounds__(256) void k_ref_plasma_density(StepArgs<double> a)
{
    int k = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int i = blockIdx.z;
    const GridDims &g = a.g;
    if (k > g.n[2] + 1 || j > g.n[1] + 1) return;
    size_t c = g.at(i, j, k);
    double r = a.rho[c];
    const double ikm = plasma_density_cell(a.E[0][c], a.E[1][c], a.E[2][c], r, a.pn, a.pm, a.pk - 1);
    a.rho[c] = r;
    a.sigma[c] = 0.0 + a.pw * ikm;
}

// PlasmaCurrent over low..high-1 = 0..N of the storage box
__global__ __launch_bounds__(256) void k_ref_plasma_current(StepArgs<double> a)
{
    int k = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int i = blockIdx.z;
    const GridDims &g = a.g;
    if (k > g.n[2] || j > g.n[1]) return;
    size_t c = g.at(i, j, k);
    const double rt = a.rho[c] + a.rho_bg;
    for (int q = 0; q < 3; ++q) a.jp[q][c] = plasma_current_comp(a.jp[q][c], a.E[q][c], rt, a.gdtn, a.gdtd, a.emdt);
}

void launch_ref_plasma_density(const StepArgs<double> &a, hipStream_t s)
{
    const GridDims &g = a.g;
    hipLaunchKernelGGL(k_ref_plasma_density, dim3((g.n[2] + 2 + 63) / 64, (g.n[1] + 2 + 3) / 4, g.n[0] + 2),
                       dim3(64, 4), 0, s, a);
}

void launch_ref_plasma_current(const StepArgs<double> &a, hipStream_t s)
{
    const GridDims &g = a.g;
    hipLaunchKernelGGL(k_ref_plasma_current, dim3((g.n[2] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[0] + 1),
                       dim3(64, 4), 0, s, a);
}

void launch_ref_stepB(const StepArgs<double> &a, hipStream_t s)
{
    const GridDims &g = a.g;
    dim3 blk(64, 4);
    dim3 grd((g.n[2] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[0] + 1);
    hipLaunchKernelGGL(k_ref_stepB, grd, blk, 0, s, a);
}

// FDTD_PLRCSolver::initAccumulator, src/fdtd_plrc.t:273-308 (stepSchemeInit only)
template <int NP>
__global__ __launch_bounds__(256) void k_init_accumulator(StepArgs<double> a)
{
    int k = blockIdx.x * 64 + threadIdx.x + 1;
    int j = blockIdx.y * 4 + threadIdx.y + 1;
    int i = blockIdx.z + 1;
    const GridDims &g = a.g;
    if (k > g.n[2] || j > g.n[1]) return;
    size_t c = g.at(i, j, k);
    double e[3] = {a.E[0][c], a.E[1][c], a.E[2][c]};
#pragma unroll
    for (int n = 0; n < NP; ++n)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double2 p = a.psi[n][d][psi_at(c)];
            const double2 X = a.pc.dxi0[n], C = a.pc.crec[n];
            a.psi[n][d][psi_at(c)] = make_double2(X.x * e[d] + (C.x * p.x - C.y * p.y), X.y * e[d] + (C.x * p.y + C.y * p.x));
        }
}

void launch_init_accumulator(const StepArgs<double> &a, int np, hipStream_t s)
{
    const GridDims &g = a.g;
    dim3 blk(64, 4);
    dim3 grd((g.n[2] + 63) / 64, (g.n[1] + 3) / 4, g.n[0]);
    switch (np) {
    case 0: break;
    case 1: hipLaunchKernelGGL(k_init_accumulator<1>, grd, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_init_accumulator<2>, grd, blk, 0, s, a); break;
    default: hipLaunchKernelGGL(k_init_accumulator<3>, grd, blk, 0, s, a); break;
    }
}


// Fused E side of one step.  A thread owns one (j, k) column segment of L cells along x and
// walks it plane by plane, keeping B of the previous plane in registers for the x difference.
template <typename T, int NP, bool NONLIN, bool CPML, int L, bool LIN = false, bool RES = false>
__device__ __forceinline__ void fused_E_tile(const StepArgs<T> &a, int tbx, int tby, int tbz)
{
    using V = typename Word<T, RES>::V;
    const GridDims &g = a.g;
    const int one[3] = {1, 1, 1};
    int j, k;
    tile_jk<LIN>(g, one, tbx, tby, j, k);
    const int i0 = tbz * L + 1;
    if (k > g.n[2] || j > g.n[1] || i0 > g.n[0]) return;
    const int i1 = min(i0 + L - 1, g.n[0]);
    const int sx = (int)g.sx, pz = g.pz;
    const int ojm = (j == 1 && a.periodic[1]) ? (g.n[1] - 1) * pz : -pz;
    const int okm = (k == 1 && a.periodic[2]) ? (g.n[2] - 1) : -1;
    int c = i0 * sx + j * pz + k + g.z0;
    const T *Bx = a.B[0], *By = a.B[1], *Bz = a.B[2];
    const int cim = (i0 == 1 && a.periodic[0]) ? c + (g.n[0] - 1) * sx : c - sx;
    const unsigned wim = lo_ld<RES>(a.Blo, cim);
    double bz_im = (double)rd<T, RES>(Bz, wim, 2, cim), by_im = (double)rd<T, RES>(By, wim, 1, cim);
    bool iny[2] = {false, false}, inz[2] = {false, false};
    if (CPML) {
        in_faces(a.face_by_dim, a.face, 1, j, iny);
        in_faces(a.face_by_dim, a.face, 2, k, inz);
    }
    const double kdy = gld(a.ikdE[1] + j), kdz = gld(a.ikdE[2] + k);
#pragma unroll
    for (int u = 0; u < L; ++u, c += sx) {
        const int i = i0 + u;
        if (i > i1) break;
        // z neighbours: the first lane of a wave and the first cell of a row load them, every
        // other lane takes them from the previous lane (the loads are issued first, so that one
        // wait covers them and the cell's own loads)
        const bool zload = threadIdx.x == 0 || k == 1;
        V bym = (V)0, bxm = (V)0;
        if (zload) {
            const unsigned wm = lo_ld<RES>(a.Blo, c + okm);
            bym = rd<T, RES>(By, wm, 1, c + okm);
            bxm = rd<T, RES>(Bx, wm, 0, c + okm);
        }
        const unsigned wb = lo_ld<RES>(a.Blo, c);
        const V bxt = rd<T, RES>(Bx, wb, 0, c), byt = rd<T, RES>(By, wb, 1, c);
        const double bx = bxt, by = byt, bz = rd<T, RES>(Bz, wb, 2, c);
        // E and the pole accumulators are read before any store of the cell (the CPML
        // auxiliaries), so that all its reads leave together
        const unsigned we = lo_once<RES>(a.Elo, c);
        double ex = (double)rd_once<T, RES>(a.E[0], we, 0, c), ey = (double)rd_once<T, RES>(a.E[1], we, 1, c);
        double ez = (double)rd_once<T, RES>(a.E[2], we, 2, c);
        PsiRegs<NP> pr;
#pragma unroll
        for (int n = 0; n < NP; ++n)
#pragma unroll
            for (int q = 0; q < 3; ++q) pr.v[n][q] = load_once(&a.psi[n][q][psi_at(c)]);
        const unsigned wj = lo_ld<RES>(a.Blo, c + ojm);
        const double dBz_y = bz - (double)rd<T, RES>(Bz, wj, 2, c + ojm);
        const double dBx_y = bx - (double)rd<T, RES>(Bx, wj, 0, c + ojm);
        const V bys = lane_prev(byt), bxs = lane_prev(bxt);
        const double dBy_z = by - (double)(zload ? bym : bys);
        const double dBx_z = bx - (double)(zload ? bxm : bxs);
        const double dBz_x = bz - bz_im;
        const double dBy_x = by - by_im;
        bz_im = bz;
        by_im = by;
        // the plasma OptField and the plasma current run before the E update, on E of this step
        double sg = 0.0, jpv[3] = {0.0, 0.0, 0.0};
        if (a.plasma || a.pcur) {
            double r = gld(a.rho + c);
            if (a.plasma) {
                sg = a.pw * plasma_density_cell(ex, ey, ez, r, a.pn, a.pm, a.pk - 1);
                gst(a.rho + c, r);
                gst(a.sigma + c, sg);
            }
            if (a.pcur) {
                const double rt = r + a.rho_bg, ev[3] = {ex, ey, ez};
                for (int q = 0; q < 3; ++q) {
                    jpv[q] = plasma_current_comp(gld(a.jp[q] + c), ev[q], rt, a.gdtn, a.gdtd, a.emdt);
                    gst(a.jp[q] + c, jpv[q]);
                }
            }
        }
        // current totals in list order: the CPML and sheet groups, the plasma current at its place
        double jx = 0.0, jy = 0.0, jz = 0.0;
        auto add_cpml = [&]() {
            if (CPML) {
                // current totals in the reference's list order N,S (y), E,W (x), U,D (z); within a
                // dimension the high face comes first in that list
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    if (iny[s]) {
                        const CpmlFace &f = a.face[a.face_by_dim[1][s]];
                        double v0, v1;
                        cpml_face<T, false, RES>(f, j, layer_index(f, i, j, k), dBx_y, dBz_y, v0, v1);
                        jz += v0;
                        jx += v1;
                    }
                bool inx[2];
                in_faces(a.face_by_dim, a.face, 0, i, inx);
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    if (inx[s]) {
                        const CpmlFace &f = a.face[a.face_by_dim[0][s]];
                        double v0, v1;
                        cpml_face<T, false, RES>(f, i, layer_index(f, i, j, k), dBz_x, dBy_x, v0, v1);
                        jy += v0;
                        jz += v1;
                    }
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    if (inz[s]) {
                        const CpmlFace &f = a.face[a.face_by_dim[2][s]];
                        double v0, v1;
                        cpml_face<T, false, RES>(f, k, layer_index(f, i, j, k), dBy_z, dBx_z, v0, v1);
                        jx += v0;
                        jy += v1;
                    }
            }
        };
        auto add_plasma = [&](int pos) {
            if (a.pcur && a.pcur_pos == pos) {
                jx += jpv[0];
                jy += jpv[1];
                jz += jpv[2];
            }
        };
        add_plasma(0);
        if (a.sheets_first)
            add_sheets(a, i, j, k, jx, jy, jz);
        else
            add_cpml();
        add_plasma(1);
        if (a.sheets_first)
            add_cpml();
        else
            add_sheets(a, i, j, k, jx, jy, jz);
        add_plasma(2);
        double idenom = a.idenom, numer = a.numer;
        if (a.epsf || a.plasma) {
            double denom;
            cell_eps<NONLIN>(a, (size_t)c, 0.5 * sg * a.dt, denom, numer);
            idenom = 1.0 / denom;
        }
        unsigned fl = 0;
        plrc_e_update_pp<NP, NONLIN, true>(ex, ey, ez, dBz_y, dBy_z, dBx_z, dBz_x, dBy_x, dBx_y, gld(a.ikdE[0] + i), kdy,
                                           kdz, jx, jy, jz, pr, a.psi, (size_t)c, true, a.pc, a.dt, idenom, numer,
                                           a.chi, a.ncap, &fl);
        report(a.status, fl | finite_flag(ex, ey, ez));
        unsigned wn = 0;
        const T hx = put<T, RES>(ex, wn, 0), hy = put<T, RES>(ey, wn, 1), hz = put<T, RES>(ez, wn, 2);
        store_once(&a.E[0][c], hx);
        store_once(&a.E[1][c], hy);
        store_once(&a.E[2][c], hz);
        if (RES) store_once(&a.Elo[c], wn);
    }
}

// Fused B side of one step: a thread walks L planes along x keeping E of the next plane in
// registers as the centre of the following iteration.
template <typename T, bool CPML, int L, bool LIN = false, bool RES = false>
__device__ __forceinline__ void fused_B_tile(const StepArgs<T> &a, int tbx, int tby, int tbz)
{
    using V = typename Word<T, RES>::V;
    const GridDims &g = a.g;
    int j, k;
    tile_jk<LIN>(g, a.blo, tbx, tby, j, k);
    const int i0 = tbz * L + a.blo[0];
    if (k > g.n[2] || j > g.n[1] || i0 > g.n[0]) return;
    const int i1 = min(i0 + L - 1, g.n[0]);
    const int sx = (int)g.sx, pz = g.pz;
    const int ojp = (j == g.n[1] && a.periodic[1]) ? -(g.n[1] - 1) * pz : pz;
    const int okp = (k == g.n[2] && a.periodic[2]) ? -(g.n[2] - 1) : 1;
    int c = i0 * sx + j * pz + k + g.z0;
    const T *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
    T *Bx = a.B[0], *By = a.B[1], *Bz = a.B[2];
    // z neighbours: the last lane of a wave and the last cell of a row load them, every other
    // lane takes them from the next lane
    const bool zload = threadIdx.x == 63 || k == g.n[2];
    V ezl[L][2];
    if (zload) {
#pragma unroll
        for (int u = 0; u < L; ++u) {
            const int cu = c + u * sx;
            ezl[u][0] = ezl[u][1] = (V)0;
            if (i0 + u <= i1) {
                const unsigned w = lo_ld<RES>(a.Elo, cu + okp);
                ezl[u][0] = rd<T, RES>(Ey, w, 1, cu + okp);
                ezl[u][1] = rd<T, RES>(Ex, w, 0, cu + okp);
            }
        }
    }
    const unsigned w0 = lo_ld<RES>(a.Elo, c);
    const V ex0 = rd<T, RES>(Ex, w0, 0, c), ey0 = rd<T, RES>(Ey, w0, 1, c), ez0 = rd<T, RES>(Ez, w0, 2, c);
    double ex = ex0, ey = ey0, ez = ez0;
    bool iny[2] = {false, false}, inz[2] = {false, false};
    if (CPML) {
        in_faces(a.face_by_dim, a.face, 1, j, iny);
        in_faces(a.face_by_dim, a.face, 2, k, inz);
    }
    const double kdy = gld(a.ikdH[1] + j), kdz = gld(a.ikdH[2] + k);
    // the E and B values of all L planes are read before the first store, so that they leave
    // together (per plane: E at x+1, E_y and E_x at z+1, E_z and E_x at y+1, B)
    V pe[L][7], pb[L][3];
#pragma unroll
    for (int u = 0; u < L; ++u) {
        const int i = i0 + u;
        if (i > i1) break;
        const int cu = c + u * sx;
        const int cip = (i == g.n[0] && a.periodic[0]) ? cu - (g.n[0] - 1) * sx : cu + sx;
        const unsigned wx = lo_ld<RES>(a.Elo, cip), wy = lo_ld<RES>(a.Elo, cu + ojp);
        const unsigned wbu = lo_once<RES>(a.Blo, cu);
        pe[u][0] = rd<T, RES>(Ex, wx, 0, cip);
        pe[u][1] = rd<T, RES>(Ey, wx, 1, cip);
        pe[u][2] = rd<T, RES>(Ez, wx, 2, cip);
        pe[u][5] = rd<T, RES>(Ez, wy, 2, cu + ojp);
        pe[u][6] = rd<T, RES>(Ex, wy, 0, cu + ojp);
        pb[u][0] = rd_once<T, RES>(Bx, wbu, 0, cu);
        pb[u][1] = rd_once<T, RES>(By, wbu, 1, cu);
        pb[u][2] = rd_once<T, RES>(Bz, wbu, 2, cu);
    }
#pragma unroll
    for (int u = 0; u < L; ++u) {
        if (i0 + u > i1) break;
        // E at z+1: the next lane's centre value of the same plane
        const V sy = lane_next(u == 0 ? ey0 : pe[u - 1][1]), sx_ = lane_next(u == 0 ? ex0 : pe[u - 1][0]);
        pe[u][3] = zload ? ezl[u][0] : sy;
        pe[u][4] = zload ? ezl[u][1] : sx_;
    }
#pragma unroll
    for (int u = 0; u < L; ++u, c += sx) {
        const int i = i0 + u;
        if (i > i1) break;
        const double ex1 = pe[u][0], ey1 = pe[u][1], ez1 = pe[u][2];
        const double dEy_z = (double)pe[u][3] - ey;
        const double dEx_z = (double)pe[u][4] - ex;
        const double dEz_y = (double)pe[u][5] - ez;
        const double dEx_y = (double)pe[u][6] - ex;
        const double dEz_x = ez1 - ez;
        const double dEy_x = ey1 - ey;
        double mx = 0.0, my = 0.0, mz = 0.0;
        if (a.sheets_first) add_sheets(a, i, j, k, mx, my, mz);
        if (CPML) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (iny[s]) {
                    const CpmlFace &f = a.face[a.face_by_dim[1][s]];
                    double v0, v1;  // psi0 -> Mz from dEx_y, psi1 -> Mx from dEz_y
                    cpml_face<T, true, RES>(f, j, layer_index(f, i, j, k), dEx_y, dEz_y, v0, v1);
                    mz += v0;
                    mx += v1;
                }
            bool inx[2];
            in_faces(a.face_by_dim, a.face, 0, i, inx);
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (inx[s]) {
                    const CpmlFace &f = a.face[a.face_by_dim[0][s]];
                    double v0, v1;  // psi0 -> My from dEz_x, psi1 -> Mz from dEy_x
                    cpml_face<T, true, RES>(f, i, layer_index(f, i, j, k), dEz_x, dEy_x, v0, v1);
                    my += v0;
                    mz += v1;
                }
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (inz[s]) {
                    const CpmlFace &f = a.face[a.face_by_dim[2][s]];
                    double v0, v1;  // psi0 -> Mx from dEy_z, psi1 -> My from dEx_z
                    cpml_face<T, true, RES>(f, k, layer_index(f, i, j, k), dEy_z, dEx_z, v0, v1);
                    mx += v0;
                    my += v1;
                }
        }
        if (!a.sheets_first) add_sheets(a, i, j, k, mx, my, mz);
        const double kdx = gld(a.ikdH[0] + i);
        unsigned wn = 0;
        const T hx = put<T, RES>((double)pb[u][0] + a.dt * ((dEy_z * kdz - dEz_y * kdy) + mx), wn, 0);
        const T hy = put<T, RES>((double)pb[u][1] + a.dt * ((dEz_x * kdx - dEx_z * kdz) + my), wn, 1);
        const T hz = put<T, RES>((double)pb[u][2] + a.dt * ((dEx_y * kdy - dEy_x * kdx) + mz), wn, 2);
        store_once(&Bx[c], hx);
        store_once(&By[c], hy);
        store_once(&Bz[c], hz);
        if (RES) store_once(&a.Blo[c], wn);
        ex = ex1;
        ey = ey1;
        ez = ez1;
    }
}

// Walk this block's share of a batch's tile list: the tiles are numbered run by run (z tile
// fastest, then y, then x), the 8 XCDs take contiguous eighths of the list so that tiles
// sharing halo cells meet in one L2, and each block strides its XCD's eighth by
// gridDim.x / 8 (gridDim.x is a multiple of 8; the hardware deals blocks to the XCDs round
// robin).  Each XCD starts its eighth at a different point (XCD x at x/8 of it, wrapping
// round), so that the XCDs do not walk equal runs at equal offsets in lock step; that
// measured 2-5 % faster on batches of equal-size runs.  f(run, tz, ty, tx) processes one tile.
template <typename F>
__device__ __forceinline__ void batch_tiles(const BatchRun *__restrict__ runs, int nruns, int total, F &&f)
{
    const int per_xcd = gridDim.x >> 3;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int chunk = (total + 7) >> 3;
    const int t0 = xcd * chunk, t1 = min(total, t0 + chunk);
    const int len = max(0, t1 - t0);
    const int rot = (int)(((long long)len * xcd) >> 3);
    int r = 0;
    for (int u = slot; u < len; u += per_xcd) {
        int v = u + rot;
        if (v >= len) v -= len;
        const int t = t0 + v;
        if (t < runs[r].first) r = 0;  // wrapped round to the start of the eighth
        while (r + 1 < nruns && t >= runs[r + 1].first) ++r;
        const int ntz = runs[r].ntz, nty = runs[r].nty;
        int local = t - runs[r].first;
        const int tz = local % ntz;
        local /= ntz;
        f(r, tz, local % nty, local / nty);
    }
}

// The argument block is read from device memory (the solver keeps a copy there): the face
// tables are indexed at run time, which a by-value kernel argument would have to copy into
// per-lane scratch.
// Waves per SIMD the E kernel is compiled for: 6 (at most 80 registers) for the float32 Kerr
// kernel with one pole and one plane per thread, measured 0.3 % faster than the 5 its registers
// allow; no constraint otherwise (other variants would spill).
template <typename T, int NP, bool NONLIN, int L, bool RES>
struct FusedEWaves {
    static constexpr int v = (sizeof(T) == 4 && !RES && NP <= 1 && NONLIN && L == 1) ? 6 : 1;
};

template <typename T, int NP, bool NONLIN, bool CPML, int L, bool RES = false>
__global__ __attribute__((amdgpu_waves_per_eu(FusedEWaves<T, NP, NONLIN, L, RES>::v))) __launch_bounds__(256) void k_fused_E(
    const StepArgs<T> *__restrict__ pa, int nbx, int nby, int nbz)
{
    int tbx, tby, tbz;
    fused_tile_of(pa->order, pa->band, pa->rev, nbx, nby, nbz, tbx, tby, tbz);
    fused_E_tile<T, NP, NONLIN, CPML, L, false, RES>(*pa, tbx, tby, tbz);
}

template <typename T, bool CPML, int L, bool RES = false>
__global__ __launch_bounds__(256) void k_fused_B(const StepArgs<T> *__restrict__ pa, int nbx, int nby, int nbz)
{
    int tbx, tby, tbz;
    fused_tile_of(pa->order, pa->band, pa->rev, nbx, nby, nbz, tbx, tby, tbz);
    fused_B_tile<T, CPML, L, false, RES>(*pa, tbx, tby, tbz);
}

// the fused kernels over the tiles of many independent runs, one persistent launch
template <typename T, int NP, bool NONLIN, bool CPML, int L>
__global__ __launch_bounds__(256) void k_fused_E_batch(const StepArgs<T> *__restrict__ args,
                                                       const BatchRun *__restrict__ runs, int nruns, int total)
{
    batch_tiles(runs, nruns, total, [&](int r, int tz, int ty, int tx) {
        fused_E_tile<T, NP, NONLIN, CPML, L, true>(args[r], tz, ty, tx);
    });
}

template <typename T, bool CPML, int L>
__global__ __launch_bounds__(256) void k_fused_B_batch(const StepArgs<T> *__restrict__ args,
                                                       const BatchRun *__restrict__ runs, int nruns, int total)
{
    batch_tiles(runs, nruns, total,
                [&](int r, int tz, int ty, int tx) { fused_B_tile<T, CPML, L, true>(args[r], tz, ty, tx); });
}

// tile grid of the batched fused kernels of one run: 256 consecutive y-z cells per tile
// (tile_jk with LIN), L x planes per thread
void fused_tiles(const GridDims &g, const int blo[3], bool hside, int L, int &ntz, int &nty, int &nseg)
{
    const int one[3] = {1, 1, 1};
    const int *lo = hside ? blo : one;
    long long yz = (long long)(g.n[1] - lo[1] + 1) * (g.n[2] - lo[2] + 1);
    ntz = (int)((yz + 255) / 256);
    nty = 1;
    nseg = (g.n[0] - lo[0] + 1 + L - 1) / L;
}

template <typename T, int L>
static void launch_fused_E_batch_L(const StepArgs<T> *args, const BatchRun *runs, int nruns, int total, int blocks,
                                   int np, bool nonlin, bool cpml, hipStream_t s)
{
#define MPX_FEB(NPV, NL, CP)                                                                                    \
    hipLaunchKernelGGL((k_fused_E_batch<T, NPV, NL, CP, L>), dim3(blocks), dim3(64, 4), 0, s, args, runs, nruns, \
                       total)
#define MPX_FEB_NP(NL, CP)             \
    switch (np) {                      \
    case 0: MPX_FEB(0, NL, CP); break; \
    case 1: MPX_FEB(1, NL, CP); break; \
    case 2: MPX_FEB(2, NL, CP); break; \
    default: MPX_FEB(3, NL, CP); break; \
    }
    if (nonlin) {
        if (cpml) { MPX_FEB_NP(true, true) } else { MPX_FEB_NP(true, false) }
    } else {
        if (cpml) { MPX_FEB_NP(false, true) } else { MPX_FEB_NP(false, false) }
    }
#undef MPX_FEB_NP
#undef MPX_FEB
}

template <typename T>
void launch_fused_E_batch(const StepArgs<T> *args, const BatchRun *runs, int nruns, int total, int blocks, int L,
                          int np, bool nonlin, bool cpml, hipStream_t s)
{
    (void)L;  // the batch walks one plane per thread on the E side
    launch_fused_E_batch_L<T, 1>(args, runs, nruns, total, blocks, np, nonlin, cpml, s);
}

template <typename T, int L>
static void launch_fused_B_batch_L(const StepArgs<T> *args, const BatchRun *runs, int nruns, int total, int blocks,
                                   bool cpml, hipStream_t s)
{
    if (cpml)
        hipLaunchKernelGGL((k_fused_B_batch<T, true, L>), dim3(blocks), dim3(64, 4), 0, s, args, runs, nruns, total);
    else
        hipLaunchKernelGGL((k_fused_B_batch<T, false, L>), dim3(blocks), dim3(64, 4), 0, s, args, runs, nruns, total);
}

template <typename T>
void launch_fused_B_batch(const StepArgs<T> *args, const BatchRun *runs, int nruns, int total, int blocks, int L,
                          bool cpml, hipStream_t s)
{
    switch (L) {
    case 2: launch_fused_B_batch_L<T, 2>(args, runs, nruns, total, blocks, cpml, s); break;
    case 4: launch_fused_B_batch_L<T, 4>(args, runs, nruns, total, blocks, cpml, s); break;
    default: launch_fused_B_batch_L<T, 1>(args, runs, nruns, total, blocks, cpml, s); break;
    }
}
template void launch_fused_E_batch<double>(const StepArgs<double> *, const BatchRun *, int, int, int, int, int, bool,
                                           bool, hipStream_t);
template void launch_fused_E_batch<float>(const StepArgs<float> *, const BatchRun *, int, int, int, int, int, bool,
                                          bool, hipStream_t);
template void launch_fused_B_batch<double>(const StepArgs<double> *, const BatchRun *, int, int, int, int, bool,
                                           hipStream_t);
template void launch_fused_B_batch<float>(const StepArgs<float> *, const BatchRun *, int, int, int, int, bool,
                                          hipStream_t);

template <typename T, int L, bool RES = false>
static void launch_fused_E_L(const StepArgs<T> &a, const StepArgs<T> *dev, int np, bool nonlin, hipStream_t s)
{
    const GridDims &g = a.g;
    int nbx = (g.n[2] + 63) / 64, nby = (g.n[1] + 3) / 4, nbz = (g.n[0] + L - 1) / L;
    dim3 blk(64, 4);
    dim3 grd(nbx, nby, nbz);
    bool cp = a.nfaces > 0;
#define MPX_FE(NPV, NL, CP) hipLaunchKernelGGL((k_fused_E<T, NPV, NL, CP, L, RES>), grd, blk, 0, s, dev, nbx, nby, nbz)
#define MPX_FE_NP(NL, CP)            \
    switch (np) {                    \
    case 0: MPX_FE(0, NL, CP); break; \
    case 1: MPX_FE(1, NL, CP); break; \
    case 2: MPX_FE(2, NL, CP); break; \
    default: MPX_FE(3, NL, CP); break; \
    }
    if (nonlin) {
        if (cp) { MPX_FE_NP(true, true) } else { MPX_FE_NP(true, false) }
    } else {
        if (cp) { MPX_FE_NP(false, true) } else { MPX_FE_NP(false, false) }
    }
#undef MPX_FE_NP
#undef MPX_FE
}

template <typename T>
void launch_fused_E(const StepArgs<T> &a, const StepArgs<T> *dev, int np, bool nonlin, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        // float32 + residual storage: one plane per thread (the march of the float32 default)
        if (a.Elo) {
            launch_fused_E_L<T, 1, true>(a, dev, np, nonlin, s);
            return;
        }
    }
    switch (a.march) {
    case 2: launch_fused_E_L<T, 2>(a, dev, np, nonlin, s); break;
    case 4: launch_fused_E_L<T, 4>(a, dev, np, nonlin, s); break;
    default: launch_fused_E_L<T, 1>(a, dev, np, nonlin, s); break;
    }
}

template <typename T, int L, bool RES = false>
static void launch_fused_B_L(const StepArgs<T> &a, const StepArgs<T> *dev, hipStream_t s)
{
    const GridDims &g = a.g;
    int nk = g.n[2] - a.blo[2] + 1, nj = g.n[1] - a.blo[1] + 1, ni = g.n[0] - a.blo[0] + 1;
    int nbx = (nk + 63) / 64, nby = (nj + 3) / 4, nbz = (ni + L - 1) / L;
    dim3 blk(64, 4);
    dim3 grd(nbx, nby, nbz);
    if (a.nfaces > 0)
        hipLaunchKernelGGL((k_fused_B<T, true, L, RES>), grd, blk, 0, s, dev, nbx, nby, nbz);
    else
        hipLaunchKernelGGL((k_fused_B<T, false, L, RES>), grd, blk, 0, s, dev, nbx, nby, nbz);
}

template <typename T>
void launch_fused_B(const StepArgs<T> &a, const StepArgs<T> *dev, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        // float32 + residual storage: one or two planes per thread
        if (a.Blo) {
            if (a.march == 1)
                launch_fused_B_L<T, 1, true>(a, dev, s);
            else
                launch_fused_B_L<T, 2, true>(a, dev, s);
            return;
        }
    }
    switch (a.march) {
    case 2: launch_fused_B_L<T, 2>(a, dev, s); break;
    case 4: launch_fused_B_L<T, 4>(a, dev, s); break;
    default: launch_fused_B_L<T, 1>(a, dev, s); break;
    }
}

template void launch_fused_E<double>(const StepArgs<double> &, const StepArgs<double> *, int, bool, hipStream_t);
template void launch_fused_E<float>(const StepArgs<float> &, const StepArgs<float> *, int, bool, hipStream_t);
template void launch_fused_B<double>(const StepArgs<double> &, const StepArgs<double> *, hipStream_t);
template void launch_fused_B<float>(const StepArgs<float> &, const StepArgs<float> *, hipStream_t);

// ------------------------------------------------------------------------------------
// fdtd-plain, fdtd-nl, fdtd-disp (periodic only; no currents)
// ------------------------------------------------------------------------------------

// E loops of FDTD_Plain / FDTD_Nonlinear / FDTD_Dispersion over [1, N+1]^3
// (src/fdtd_plain.cpp:54-75, src/fdtd_nl.cpp:55-98, src/fdtd_disp.cpp:113-204)
__global__ __launch_bounds__(256) void k_simple_stepD(SimpleArgs a)
{
    int k = blockIdx.x * 64 + threadIdx.x + 1;
    int j = blockIdx.y * 4 + threadIdx.y + 1;
    int i = blockIdx.z + 1;
    const GridDims &g = a.g;
    if (k > g.n[2] + 1 || j > g.n[1] + 1) return;
    size_t c = g.at(i, j, k);
    double *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
    const double *Bx = a.B[0], *By = a.B[1], *Bz = a.B[2];
    double dx = a.dx, dy = a.dy, dz = a.dz, dt = a.dt;
    double cx = (Bz[c] - Bz[g.at(i, j - 1, k)]) / dy - (By[c] - By[g.at(i, j, k - 1)]) / dz;
    double cy = (Bx[c] - Bx[g.at(i, j, k - 1)]) / dz - (Bz[c] - Bz[g.at(i - 1, j, k)]) / dx;
    double cz = (By[c] - By[g.at(i - 1, j, k)]) / dx - (Bx[c] - Bx[g.at(i, j - 1, k)]) / dy;
    double ex = Ex[c], ey = Ey[c], ez = Ez[c];
    if (a.kind == 0) {
        Ex[c] = ex + dt * cx;
        Ey[c] = ey + dt * cy;
        Ez[c] = ez + dt * cz;
    } else if (a.kind == 1) {
        double E2 = ex * ex + ey * ey + ez * ez;
        double Dx = (a.eps + a.chi * E2) * ex + dt * cx;
        double Dy = (a.eps + a.chi * E2) * ey + dt * cy;
        double Dz = (a.eps + a.chi * E2) * ez + dt * cz;
        double E = sqrt(E2);
        double D = sqrt(Dx * Dx + Dy * Dy + Dz * Dz);
        double Eold;
        int it = 0;
        do {
            Eold = E;
            E2 = E * E;
            E = (2 * a.chi * E2 * E + D) / (3 * a.chi * E2 + a.eps);
        } while (fabs(E - Eold) > 1e-9 && ++it < a.ncap);
        Ex[c] = E * Dx / D;
        Ey[c] = E * Dy / D;
        Ez[c] = E * Dz / D;
        // a zero D gives NaN, as in the reference (src/fdtd_nl.cpp:86-89): reported, not avoided
        report(a.status, (it >= a.ncap ? 1u : 0u) | finite_flag(Ex[c], Ey[c], Ez[c]));
        return;
    } else {
        double e[3] = {ex, ey, ez};
        double P[3] = {0, 0, 0}, Pp[3] = {0, 0, 0};
        double dt2 = dt * dt;
        for (int n = 0; n < 3; ++n) {
            double aa = dt2 * a.om2[n];
            double b = aa * a.eps2[n];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double p = a.P[n][d][c], pp = a.Pp[n][d][c];
                double pn = (2 - aa) * p - pp + b * e[d];
                a.Pp[n][d][c] = p;
                a.P[n][d][c] = pn;
                P[d] += pn;
                Pp[d] += p;
            }
        }
        Ex[c] = ex + Pp[0] - P[0] + dt * cx;
        Ey[c] = ey + Pp[1] - P[1] + dt * cy;
        Ez[c] = ez + Pp[2] - P[2] + dt * cz;
    }
}

// stepB shared by the three solvers, including the /dx of both Bz terms
// (src/fdtd_plain.cpp:97-118)
__global__ __launch_bounds__(256) void k_simple_stepB(SimpleArgs a)
{
    int k = blockIdx.x * 64 + threadIdx.x;
    int j = blockIdx.y * 4 + threadIdx.y;
    int i = blockIdx.z;
    const GridDims &g = a.g;
    if (k > g.n[2] || j > g.n[1]) return;
    size_t c = g.at(i, j, k);
    const double *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
    double dx = a.dx, dy = a.dy, dz = a.dz, dt = a.dt;
    double ex = Ex[c], ey = Ey[c], ez = Ez[c];
    a.B[0][c] = a.B[0][c] + dt * ((Ey[g.at(i, j, k + 1)] - ey) / dz - (Ez[g.at(i, j + 1, k)] - ez) / dy);
    a.B[1][c] = a.B[1][c] + dt * ((Ez[g.at(i + 1, j, k)] - ez) / dx - (Ex[g.at(i, j, k + 1)] - ex) / dz);
    a.B[2][c] = a.B[2][c] + dt * ((Ex[g.at(i, j + 1, k)] - ex) / dx - (Ey[g.at(i + 1, j, k)] - ey) / dx);
}

void launch_simple_stepD(const SimpleArgs &a, hipStream_t s)
{
    const GridDims &g = a.g;
    dim3 blk(64, 4);
    dim3 grd((g.n[2] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[0] + 1);
    hipLaunchKernelGGL(k_simple_stepD, grd, blk, 0, s, a);
}

void launch_simple_stepB(const SimpleArgs &a, hipStream_t s)
{
    const GridDims &g = a.g;
    dim3 blk(64, 4);
    dim3 grd((g.n[2] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[0] + 1);
    hipLaunchKernelGGL(k_simple_stepB, grd, blk, 0, s, a);
}

// ------------------------------------------------------------------------------------
// diagnostics: probes and energy
// ------------------------------------------------------------------------------------

template <typename T>
__global__ void k_probe(ProbeArgs<T> a, double *row)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const int q = a.comp[p];
    const size_t c = a.idx[p];
    const unsigned *lo = a.lo[q / 3];
    row[p] = lo ? res_decode((float)a.f[q][c], lo[c], q % 3) : (double)a.f[q][c];
}

template <typename T>
void launch_probe(const ProbeArgs<T> &a, double *row, hipStream_t s)
{
    hipLaunchKernelGGL(k_probe<T>, dim3((a.n + 63) / 64), dim3(64), 0, s, a, row);
}
template void launch_probe<double>(const ProbeArgs<double> &, double *, hipStream_t);
template void launch_probe<float>(const ProbeArgs<float> &, double *, hipStream_t);

// FieldEnergyDiag::calculate, src/fielddiag.cpp:255-286: per-block partial sums, then one
// block adds the partials in a fixed order.
template <typename T>
__device__ __forceinline__ void cell3(const T *f0, const T *f1, const T *f2, const unsigned *lo, size_t c, double v[3])
{
    if (lo) {
        const unsigned w = lo[c];
        v[0] = res_decode((float)f0[c], w, 0);
        v[1] = res_decode((float)f1[c], w, 1);
        v[2] = res_decode((float)f2[c], w, 2);
    } else {
        v[0] = f0[c];
        v[1] = f1[c];
        v[2] = f2[c];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_energy_partial(GridDims g, const T *e0, const T *e1, const T *e2,
                                                        const T *b0, const T *b1, const T *b2, const unsigned *elo,
                                                        const unsigned *blo, double *part)
{
    __shared__ double red[256];
    int k = blockIdx.x * 64 + threadIdx.x + 1;
    int j = blockIdx.y * 4 + threadIdx.y + 1;
    int i = blockIdx.z + 1;
    double v = 0;
    if (k <= g.n[2] && j <= g.n[1]) {
        size_t c = g.at(i, j, k);
        double e[3], b[3];
        cell3(e0, e1, e2, elo, c, e);
        cell3(b0, b1, b2, blo, c, b);
        v = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    }
    int t = threadIdx.y * 64 + threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *part, size_t n, double *out, double scale)
{
    __shared__ double red[256];
    double v = 0;
    for (size_t q = threadIdx.x; q < n; q += 256) v += part[q];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

size_t energy_partials(const GridDims &g)
{
    return (size_t)((g.n[2] + 63) / 64) * ((g.n[1] + 3) / 4) * g.n[0];
}

template <typename T>
void launch_energy(const GridDims &g, T *const e[3], T *const b[3], double *part, double *out, hipStream_t s,
                   const unsigned *elo, const unsigned *blo, double scale)
{
    dim3 blk(64, 4);
    dim3 grd((g.n[2] + 63) / 64, (g.n[1] + 3) / 4, g.n[0]);
    hipLaunchKernelGGL(k_energy_partial<T>, grd, blk, 0, s, g, e[0], e[1], e[2], b[0], b[1], b[2], elo, blo, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, part, energy_partials(g), out, scale);
}
template void launch_energy<double>(const GridDims &, double *const[3], double *const[3], double *, double *,
                                    hipStream_t, const unsigned *, const unsigned *, double);
template void launch_energy<float>(const GridDims &, float *const[3], float *const[3], double *, double *,
                                   hipStream_t, const unsigned *, const unsigned *, double);

}  // namespace mpx

namespace mpx {

// Re Efunc(0,0,0,0) with Phase = TShift = ZShift = 0: the amplitude normalisation of
// ShortPulseInjectSourceFunc::setParam, src/shortpulseinject.cpp:101-112
__global__ void k_pulse_norm(PulseParams P, FaddeevaTables ft, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = pulse_efunc(P, ft, 0, 0, 0, 0).re;
}

double pulse_norm_device(const PulseParams &P, const FaddeevaTables &ft, hipStream_t s)
{
    double *d = nullptr, h = 0;
    if (hipMalloc(&d, sizeof(double)) != hipSuccess) return NAN;
    hipLaunchKernelGGL(k_pulse_norm, dim3(1), dim3(64), 0, s, P, ft, d);
    if (hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) h = NAN;
    if (hipStreamSynchronize(s) != hipSuccess) h = NAN;
    (void)hipFree(d);
    return h;
}

}  // namespace mpx

namespace mpx {

// Plain streaming copy, one 16-byte vector per thread over a grid that covers the buffer
// (the fastest copy shape measured on gfx950): the calibration kernel for HBM byte counters
// and the practical bandwidth ceiling of the device.
__global__ __launch_bounds__(256) void k_hbm_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = src[q];
}

void launch_hbm_copy(const void *src, void *dst, size_t bytes, hipStream_t s)
{
    size_t n = bytes / 16;
    hipLaunchKernelGGL(k_hbm_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const double2 *)src,
                       (double2 *)dst, n);
}

// A grid over the storage box 0..N+1 in the order of the reference's Grid (x fastest), as
// float64: dst[(k (ny+2) + j) (nx+2) + i] = src(i, j, k), the form FieldDiag hands to its HDF5
// stream.  A block transposes a 32 (z) x 32 (x) tile of one y row through LDS, so that both
// the reads (z contiguous) and the writes (x contiguous) are coalesced.
template <typename T>
__global__ __launch_bounds__(256) void k_pack_xfast(GridDims g, const T *__restrict__ src, const unsigned *lo, int q,
                                                    double *__restrict__ dst)
{
    __shared__ double tile[32][33];
    const int nx2 = g.n[0] + 2, ny2 = g.n[1] + 2, nz2 = g.n[2] + 2;
    const int k0 = blockIdx.x * 32, i0 = blockIdx.y * 32, j = blockIdx.z;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int i = i0 + r, k = k0 + threadIdx.x;
        if (i < nx2 && k < nz2) {
            const size_t c = g.at(i, j, k);
            tile[r][threadIdx.x] = lo ? res_decode((float)src[c], lo[c], q) : (double)src[c];
        }
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int k = k0 + r, i = i0 + threadIdx.x;
        if (i < nx2 && k < nz2) dst[((size_t)k * ny2 + j) * nx2 + i] = tile[threadIdx.x][r];
    }
}

template <typename T>
void launch_pack_xfast(const GridDims &g, const T *src, double *dst, hipStream_t s, const unsigned *lo, int q)
{
    dim3 grd((g.n[2] + 2 + 31) / 32, (g.n[0] + 2 + 31) / 32, g.n[1] + 2);
    hipLaunchKernelGGL(k_pack_xfast<T>, grd, dim3(32, 8), 0, s, g, src, lo, q, dst);
}
template void launch_pack_xfast<float>(const GridDims &, const float *, double *, hipStream_t, const unsigned *, int);
template void launch_pack_xfast<double>(const GridDims &, const double *, double *, hipStream_t, const unsigned *, int);

// One record of a sampler: the value at storage cell c (dim < 0) or the len = N[dim] + 2 values
// of the storage line through c along dim, as float64.  A coordinate of a dimension in
// wrapmask that names a ghost reads its periodic image (0 -> N, N + 1 -> 1), which is what the
// ghost holds once the wrap has run; the caller checks 0 <= c[d] <= N[d] + 1.
template <typename T>
__global__ __launch_bounds__(256) void k_sample_gather(GridDims g, const T *__restrict__ src, const unsigned *lo, int qc,
                                                      int dim, int c0, int c1, int c2, int len, int wrapmask,
                                                      double *__restrict__ dst)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= len) return;
    int c[3] = {c0, c1, c2};
    if (dim >= 0) c[dim] = q;
    for (int d = 0; d < 3; ++d)
        if ((wrapmask >> d) & 1) c[d] = c[d] == 0 ? g.n[d] : (c[d] == g.n[d] + 1 ? 1 : c[d]);
    const size_t at = g.at(c[0], c[1], c[2]);
    dst[q] = lo ? res_decode((float)src[at], lo[at], qc) : (double)src[at];
}

template <typename T>
void launch_sample_gather(const GridDims &g, const T *src, int dim, const int *c, int len, int wrapmask, double *dst,
                         hipStream_t s, const unsigned *lo, int q)
{
    hipLaunchKernelGGL(k_sample_gather<T>, dim3((len + 255) / 256), dim3(256), 0, s, g, src, lo, q, dim, c[0], c[1],
                       c[2], len, wrapmask, dst);
}
template void launch_sample_gather<float>(const GridDims &, const float *, int, const int *, int, int, double *,
                                         hipStream_t, const unsigned *, int);
template void launch_sample_gather<double>(const GridDims &, const double *, int, const int *, int, int, double *,
                                          hipStream_t, const unsigned *, int);

}  // namespace mpx

namespace mpx {


// ------------------------------------------------------------------------------------
// single-pass sweep path (periodic boundaries): E and B of a step in one kernel
// ------------------------------------------------------------------------------------
//
// Block = 64 lanes along z x BY thread rows along y; each thread row handles RPT tile rows,
// so a tile is 64 x TYR cells (TYR = BY * RPT).  Lane 63 and tile row TYR-1 are a one-cell
// halo: E is computed there (identically to the owning block) but never stored, so the
// B update of the owned 63 x (TYR-1) cells finds E_new at k+1 and j+1 inside the block.
// The block walks the x planes i0..i1 of its segment: at plane p it computes E_new(p), puts
// it in a triple-buffered LDS tile, synchronises once, and updates B at plane p-1 from
// E_new(p-1) and E_new(p).  Every array the step reads is an "old" buffer and every array it
// writes a "new" buffer (CPML H auxiliaries excepted: only owners touch them), so halo
// readers never see another block's results.
//
// sweep_tile is one tile's walk; k_sweep maps the blocks of one grid onto tiles, k_sweep_batch
// walks the tiles of many independent grids from one persistent launch.
template <typename T, int NP, bool NONLIN, bool CPML, int BY, int RPT>
__device__ __forceinline__ void sweep_tile(const SweepArgs<T> &a, int tz, int ty, int tx,
                                           T (&lds)[3][3][BY * RPT][64])
{
    constexpr int TYR = BY * RPT;
    const GridDims &g = a.g;
    const int Nx = g.n[0], Ny = g.n[1], Nz = g.n[2];
    const int sx = (int)g.sx, pz = g.pz;
    const int lane = threadIdx.x;
    const int k = 1 + 63 * tz + lane;
    const bool kact = k <= Nz + 1;
    // physical z index (periodic image of the halo lane); idle lanes get 1 so that every
    // index they form stays inside the arrays, although they load nothing from the grids
    const int kk = (!kact || k == Nz + 1) ? 1 : k;
    const int okm = (kk == 1) ? (Nz - 1) : -1;
    const int i0 = 1 + a.seg * tx;
    const int i1 = min(i0 + a.seg - 1, Nx);
    int jj[RPT], ojm[RPT], rowi[RPT];
    bool eact[RPT], own[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        rowi[r] = threadIdx.y + BY * r;
        int j = 1 + (TYR - 1) * ty + rowi[r];
        eact[r] = kact && j <= Ny + 1;
        own[r] = eact[r] && lane < 63 && k <= Nz && rowi[r] < TYR - 1 && j <= Ny;
        jj[r] = (j == Ny + 1) ? 1 : j;
        ojm[r] = (jj[r] == 1) ? (Ny - 1) * pz : -pz;
        if (!eact[r]) jj[r] = 1;  // keeps idle threads' addresses in range; they load nothing
    }
    const T *Bxo = a.Bo[0], *Byo = a.Bo[1], *Bzo = a.Bo[2];
    // B_old at plane i0-1 (the x neighbour of the first E plane)
    const int im0 = (i0 == 1) ? Nx : i0 - 1;
    double bprev[RPT][3], eprev[RPT][3];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        eprev[r][0] = eprev[r][1] = eprev[r][2] = 0.0;
        bprev[r][0] = bprev[r][1] = bprev[r][2] = 0.0;
        if (eact[r]) {
            int c = im0 * sx + jj[r] * pz + kk + g.z0;
            bprev[r][0] = Bxo[c];
            bprev[r][1] = Byo[c];
            bprev[r][2] = Bzo[c];
        }
    }
    const double kdz_e = a.ikdE[2][kk], kdz_h = a.ikdH[2][kk];  // reciprocals
    // the grid inputs of a plane (B and its y / z neighbours, E, the pole accumulators) are
    // loaded one plane ahead, so that their latency overlaps the work and the barrier of the
    // plane before
    T nb[RPT][7], ne[RPT][3];
    PsiRegs<NP> npsi[RPT];
    auto load_plane = [&](int p) {
        const int pp = (p == Nx + 1) ? 1 : p;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            if (!eact[r]) continue;
            const int c = pp * sx + jj[r] * pz + kk + g.z0;
            nb[r][0] = Bxo[c];
            nb[r][1] = Byo[c];
            nb[r][2] = Bzo[c];
            nb[r][3] = Bzo[c + ojm[r]];
            nb[r][4] = Bxo[c + ojm[r]];
            nb[r][5] = Byo[c + okm];
            nb[r][6] = Bxo[c + okm];
            ne[r][0] = a.Eo[0][c];
            ne[r][1] = a.Eo[1][c];
            ne[r][2] = a.Eo[2][c];
#pragma unroll
            for (int n = 0; n < NP; ++n)
#pragma unroll
                for (int q = 0; q < 3; ++q) npsi[r].v[n][q] = load_once(&a.psio[n][q][psi_at(c)]);
        }
    };
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int q = 0; q < 7; ++q) nb[r][q] = (T)0;
        ne[r][0] = ne[r][1] = ne[r][2] = (T)0;
#pragma unroll
        for (int n = 0; n < NP; ++n)
#pragma unroll
            for (int q = 0; q < 3; ++q) npsi[r].v[n][q] = make_double2(0.0, 0.0);
    }
    load_plane(i0);
    for (int p = i0; p <= i1 + 1; ++p) {
        const int pp = (p == Nx + 1) ? 1 : p;
        const int buf = p % 3;
        T cb[RPT][7], ce[RPT][3];
        PsiRegs<NP> cpsi[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
#pragma unroll
            for (int q = 0; q < 7; ++q) cb[r][q] = nb[r][q];
            ce[r][0] = ne[r][0];
            ce[r][1] = ne[r][1];
            ce[r][2] = ne[r][2];
            cpsi[r] = npsi[r];
        }
        if (p <= i1) load_plane(p + 1);
        double ecur[RPT][3], bcur[RPT][3];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            ecur[r][0] = ecur[r][1] = ecur[r][2] = 0.0;
            bcur[r][0] = bcur[r][1] = bcur[r][2] = 0.0;
            if (!eact[r]) continue;
            const int j = jj[r];
            const int c = pp * sx + j * pz + kk + g.z0;
            const double bx = cb[r][0], by = cb[r][1], bz = cb[r][2];
            const double dBz_y = bz - (double)cb[r][3];
            const double dBx_y = bx - (double)cb[r][4];
            const double dBy_z = by - (double)cb[r][5];
            const double dBx_z = bx - (double)cb[r][6];
            const double dBz_x = bz - bprev[r][2];
            const double dBy_x = by - bprev[r][1];
            bcur[r][0] = bx;
            bcur[r][1] = by;
            bcur[r][2] = bz;
            const bool store = own[r] && p <= i1;
            double jx = 0.0, jy = 0.0, jz = 0.0;
            if (a.sheets_first) add_sheet_planes(a.sheet, a.nsheets, g.py, pp, j, kk, jx, jy, jz);
            if (CPML) {
                // E auxiliaries, list order N,S (y), E,W (x), U,D (z), high face first
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int dim = d == 0 ? 1 : (d == 1 ? 0 : 2);
                    const int coord = dim == 0 ? pp : (dim == 1 ? j : kk);
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        const int fi = a.efbd[dim][s];
                        if (fi < 0) continue;
                        const CpmlFace &f = a.eface[fi];
                        if (coord < f.lo[dim] || coord > f.hi[dim]) continue;
                        const int l = layer_index(f, pp, j, kk);
                        const int q = coord - f.lo[dim];
                        const T *o0 = (const T *)a.epsi_old[fi][0], *o1 = (const T *)a.epsi_old[fi][1];
                        double d1 = dim == 0 ? dBz_x : (dim == 1 ? dBx_y : dBy_z);
                        double d0 = dim == 0 ? dBy_x : (dim == 1 ? dBz_y : dBx_z);
                        T n0 = (T)(f.b[q] * (double)o0[l] - f.c[q] * d1 * f.idxn);
                        T n1 = (T)(f.b[q] * (double)o1[l] + f.c[q] * d0 * f.idxn);
                        if (store) {
                            ((T *)f.psi0)[l] = n0;
                            ((T *)f.psi1)[l] = n1;
                        }
                        if (dim == 0) {
                            jy += (double)n0;
                            jz += (double)n1;
                        } else if (dim == 1) {
                            jz += (double)n0;
                            jx += (double)n1;
                        } else {
                            jx += (double)n0;
                            jy += (double)n1;
                        }
                    }
                }
            }
            if (!a.sheets_first) add_sheet_planes(a.sheet, a.nsheets, g.py, pp, j, kk, jx, jy, jz);
            double ex = (double)ce[r][0], ey = (double)ce[r][1], ez = (double)ce[r][2];
            unsigned fl = 0;
            plrc_e_update_pp<NP, NONLIN, true>(ex, ey, ez, dBz_y, dBy_z, dBx_z, dBz_x, dBy_x, dBx_y,
                                               a.ikdE[0][pp], a.ikdE[1][j], kdz_e, jx, jy, jz, cpsi[r], a.psin, c,
                                               store, a.pc, a.dt, a.idenom, a.numer, a.chi, a.ncap, &fl);
            if (store) fl |= finite_flag(ex, ey, ez);
            report(a.status, fl);
            const T tx_ = (T)ex, ty_ = (T)ey, tz_ = (T)ez;
            if (store) {
                a.En[0][c] = tx_;
                a.En[1][c] = ty_;
                a.En[2][c] = tz_;
            }
            lds[buf][0][rowi[r]][lane] = tx_;
            lds[buf][1][rowi[r]][lane] = ty_;
            lds[buf][2][rowi[r]][lane] = tz_;
            ecur[r][0] = (double)tx_;
            ecur[r][1] = (double)ty_;
            ecur[r][2] = (double)tz_;
        }
        __syncthreads();
        if (p > i0) {
            const int pb = p - 1;  // B plane, inside [i0, i1]
            const int pbuf = pb % 3;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                if (!own[r]) continue;
                const int j = jj[r];
                const int c = pb * sx + j * pz + kk + g.z0;
                const int row = rowi[r];
                const double ex = eprev[r][0], ey = eprev[r][1], ez = eprev[r][2];
                const double dEy_z = (double)lds[pbuf][1][row][lane + 1] - ey;
                const double dEx_z = (double)lds[pbuf][0][row][lane + 1] - ex;
                const double dEz_y = (double)lds[pbuf][2][row + 1][lane] - ez;
                const double dEx_y = (double)lds[pbuf][0][row + 1][lane] - ex;
                const double dEz_x = ecur[r][2] - ez;
                const double dEy_x = ecur[r][1] - ey;
                double mx = 0.0, my = 0.0, mz = 0.0;
                if (a.sheets_first) add_sheet_planes(a.hsheet, a.nhsheets, g.py, pb, j, kk, mx, my, mz);
                if (CPML) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const int dim = d == 0 ? 1 : (d == 1 ? 0 : 2);
                        const int coord = dim == 0 ? pb : (dim == 1 ? j : kk);
#pragma unroll
                        for (int s = 0; s < 2; ++s) {
                            const int fi = a.hfbd[dim][s];
                            if (fi < 0) continue;
                            const CpmlFace &f = a.hface[fi];
                            if (coord < f.lo[dim] || coord > f.hi[dim]) continue;
                            const int l = layer_index(f, pb, j, kk);
                            const int q = coord - f.lo[dim];
                            T *p0 = (T *)f.psi0, *p1 = (T *)f.psi1;
                            double d1 = dim == 0 ? dEz_x : (dim == 1 ? dEx_y : dEy_z);
                            double d0 = dim == 0 ? dEy_x : (dim == 1 ? dEz_y : dEx_z);
                            T n0 = (T)(f.b[q] * (double)p0[l] + f.c[q] * d1 * f.idxn);
                            T n1 = (T)(f.b[q] * (double)p1[l] - f.c[q] * d0 * f.idxn);
                            p0[l] = n0;
                            p1[l] = n1;
                            if (dim == 0) {
                                my += (double)n0;
                                mz += (double)n1;
                            } else if (dim == 1) {
                                mz += (double)n0;
                                mx += (double)n1;
                            } else {
                                mx += (double)n0;
                                my += (double)n1;
                            }
                        }
                    }
                }
                if (!a.sheets_first) add_sheet_planes(a.hsheet, a.nhsheets, g.py, pb, j, kk, mx, my, mz);
                const double kdx = a.ikdH[0][pb], kdy = a.ikdH[1][j];
                a.Bn[0][c] = (T)(bprev[r][0] + a.dt * ((dEy_z * kdz_h - dEz_y * kdy) + mx));
                a.Bn[1][c] = (T)(bprev[r][1] + a.dt * ((dEz_x * kdx - dEx_z * kdz_h) + my));
                a.Bn[2][c] = (T)(bprev[r][2] + a.dt * ((dEx_y * kdy - dEy_x * kdx) + mz));
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                eprev[r][q] = ecur[r][q];
                bprev[r][q] = bcur[r][q];
            }
    }
}

template <typename T, int NP, bool NONLIN, bool CPML, int BY, int RPT>
__global__ __launch_bounds__(64 * BY) void k_sweep(SweepArgs<T> a, int ntz, int nty, int nseg)
{
    __shared__ T lds[3][3][BY * RPT][64];
    int tz, ty, tx;
    xcd_tile(ntz, nty, nseg, tz, ty, tx);
    sweep_tile<T, NP, NONLIN, CPML, BY, RPT>(a, tz, ty, tx, lds);
}

// Persistent batched sweep over the tiles of all runs (see batch_tiles).
template <typename T, int NP, bool NONLIN, bool CPML, int BY, int RPT>
__global__ __launch_bounds__(64 * BY) void k_sweep_batch(const SweepArgs<T> *__restrict__ args,
                                                         const BatchRun *__restrict__ runs, int nruns, int total)
{
    __shared__ T lds[3][3][BY * RPT][64];
    batch_tiles(runs, nruns, total, [&](int r, int tz, int ty, int tx) {
        sweep_tile<T, NP, NONLIN, CPML, BY, RPT>(args[r], tz, ty, tx, lds);
        __syncthreads();  // the next tile rewrites the LDS planes
    });
}

template <typename T, int NP, bool NONLIN, bool CPML>
static void launch_sweep_np(const SweepArgs<T> &a, hipStream_t s)
{
    int ntz, nty, nseg;
    sweep_tiles(a.g, a.variant, a.seg, ntz, nty, nseg);
    if (a.variant == 1)  // 64 x 4 threads: 4-row tiles
        hipLaunchKernelGGL((k_sweep<T, NP, NONLIN, CPML, 4, 1>), dim3(ntz, nty, nseg), dim3(64, 4), 0, s, a, ntz, nty,
                           nseg);
    else if (a.variant == 2)  // 64 x 16 threads, 1 row each
        hipLaunchKernelGGL((k_sweep<T, NP, NONLIN, CPML, 16, 1>), dim3(ntz, nty, nseg), dim3(64, 16), 0, s, a, ntz,
                           nty, nseg);
    else  // 64 x 8 threads, 1 row each: 8-row tiles
        hipLaunchKernelGGL((k_sweep<T, NP, NONLIN, CPML, 8, 1>), dim3(ntz, nty, nseg), dim3(64, 8), 0, s, a, ntz, nty,
                           nseg);
}

// tile grid of one run for the block shape `variant` and segment length `seg`
void sweep_tiles(const GridDims &g, int variant, int seg, int &ntz, int &nty, int &nseg)
{
    ntz = (g.n[2] + 62) / 63;
    nty = variant == 0 ? (g.n[1] + 6) / 7 : (variant == 1 ? (g.n[1] + 2) / 3 : (g.n[1] + 14) / 15);
    nseg = (g.n[0] + seg - 1) / seg;
}

template <typename T, int NP, bool NONLIN, bool CPML>
static void launch_sweep_batch_np(const SweepArgs<T> *args, const BatchRun *runs, int nruns, int total, int variant,
                                  int blocks, hipStream_t s)
{
    if (variant == 1)
        hipLaunchKernelGGL((k_sweep_batch<T, NP, NONLIN, CPML, 4, 1>), dim3(blocks), dim3(64, 4), 0, s, args, runs,
                           nruns, total);
    else if (variant == 2)
        hipLaunchKernelGGL((k_sweep_batch<T, NP, NONLIN, CPML, 16, 1>), dim3(blocks), dim3(64, 16), 0, s, args, runs,
                           nruns, total);
    else
        hipLaunchKernelGGL((k_sweep_batch<T, NP, NONLIN, CPML, 8, 1>), dim3(blocks), dim3(64, 8), 0, s, args, runs,
                           nruns, total);
}

template <typename T>
void launch_sweep_batch(const SweepArgs<T> *args, const BatchRun *runs, int nruns, int total, int variant,
                        int blocks, int np, bool nonlin, bool cpml, hipStream_t s)
{
#define MPX_SB(NPV, NL, CP) launch_sweep_batch_np<T, NPV, NL, CP>(args, runs, nruns, total, variant, blocks, s)
#define MPX_SB_NP(NL, CP)             \
    switch (np) {                     \
    case 0: MPX_SB(0, NL, CP); break; \
    case 1: MPX_SB(1, NL, CP); break; \
    case 2: MPX_SB(2, NL, CP); break; \
    default: MPX_SB(3, NL, CP); break; \
    }
    if (nonlin) {
        if (cpml) { MPX_SB_NP(true, true) } else { MPX_SB_NP(true, false) }
    } else {
        if (cpml) { MPX_SB_NP(false, true) } else { MPX_SB_NP(false, false) }
    }
#undef MPX_SB_NP
#undef MPX_SB
}
template void launch_sweep_batch<double>(const SweepArgs<double> *, const BatchRun *, int, int, int, int, int, bool,
                                         bool, hipStream_t);
template void launch_sweep_batch<float>(const SweepArgs<float> *, const BatchRun *, int, int, int, int, int, bool,
                                        bool, hipStream_t);

template <typename T>
void launch_sweep(const SweepArgs<T> &a, int np, bool nonlin, hipStream_t s)
{
    bool cp = a.nefaces > 0;
#define MPX_SW(NPV, NL, CP) launch_sweep_np<T, NPV, NL, CP>(a, s)
#define MPX_SW_NP(NL, CP)            \
    switch (np) {                    \
    case 0: MPX_SW(0, NL, CP); break; \
    case 1: MPX_SW(1, NL, CP); break; \
    case 2: MPX_SW(2, NL, CP); break; \
    default: MPX_SW(3, NL, CP); break; \
    }
    if (nonlin) {
        if (cp) { MPX_SW_NP(true, true) } else { MPX_SW_NP(true, false) }
    } else {
        if (cp) { MPX_SW_NP(false, true) } else { MPX_SW_NP(false, false) }
    }
#undef MPX_SW_NP
#undef MPX_SW
}
template void launch_sweep<double>(const SweepArgs<double> &, int, bool, hipStream_t);
template void launch_sweep<float>(const SweepArgs<float> &, int, bool, hipStream_t);

}  // namespace mpx
