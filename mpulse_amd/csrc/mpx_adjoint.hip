// Discrete adjoint of one time step of the fused float64 PLRC update (the forward kernels are
// fused_E_tile / fused_B_tile in mpx_kernels.hip; the reference update is
// FDTD_PLRCNonlinCore::plrcStepD, src/fdtd_plrc.cpp:304-427, with the pole recursion of
// src/fdtd_plrc.cpp:321-353, plrcStepB, src/fdtd_plrc.cpp:429-459, and the CPML auxiliaries
// of src/cpml_border.cpp:398-425, 499-527).  Periodic boundary in every dimension: E and B
// cells are 1..N, neighbours wrap.
//
// Forward step n (per cell; D_F[d][q] = difference along d of component q of field F):
//   E kernel: psiE' = b psiE -/+ c D_B / delta (face layers), J = sheets + psiE',
//             R = curl_kappa(D_B) + J, poles p: P_p = Dd_p E + Psi_p (complex, Dd = dchi0 -
//             dxi0), Psi_p' = dxi0_p E + Crec_p P_p,
//             C = ((numer + chi |E|^2) E + dt R + sum_p Re P_p) / denom with
//             denom = eps + sumChi0 - sumXi0 and numer = eps - sumXi0,
//             E' = (e / s) C with s = |C| and (chi / denom) e^3 + e = s (the linear solver:
//             chi = 0, E' = C);
//   B kernel: psiH' = b psiH +/- c D_E' / delta, B' = B + dt (curl_kappa(D_E') + M),
//             M = sheets + psiH'.
// The reverse of step n runs three kernels: k_adj_bside (adjoint of the B kernel: the
// adjoints of its E' differences and of psiH), k_adj_eside (gathers the adjoint of E', adds
// the objective, takes the Newton solve, eps, chi and the pole recursion back, then the
// adjoints of the E kernel's B differences and of psiE) and k_adj_bgather (adds those to the
// adjoint of B).  A complex accumulator's adjoint is held as (dL/dRe, dL/dIm); through
// w = K z it goes back as conj(K) lw, and a real input x of w = G x receives Re(conj(G) lw).
#include <hip/hip_runtime.h>

#include "mpx_launch.h"

namespace mpx {

namespace {

__device__ __forceinline__ int slot(int d, int q) { return d * 2 + (q > d ? q - 1 : q); }

// periodic neighbour index along one dimension
__device__ __forceinline__ int wrapm(int v, int n) { return v == 1 ? n : v - 1; }
__device__ __forceinline__ int wrapp(int v, int n) { return v == n ? 1 : v + 1; }

// whether (coord along dim) lies in face f's layer
__device__ __forceinline__ bool in_layer(const CpmlFace &f, int dim, int coord)
{
    return coord >= f.lo[dim] && coord <= f.hi[dim];
}

}  // namespace

// Adjoint of the B kernel at B cell (i, j, k): the E' differences get dt lB through the
// curl and through the H auxiliaries; the adjoint of each auxiliary is carried to step n.
__global__ __launch_bounds__(256) void k_adj_bside(AdjArgs a)
{
    const int k = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    const GridDims &g = a.g;
    if (k > g.n[2] || j > g.n[1]) return;
    const size_t c = g.at(i, j, k);
    const double lb[3] = {a.lB[0][c], a.lB[1][c], a.lB[2][c]};
    const double ikx = a.ikdH[0][i], iky = a.ikdH[1][j], ikz = a.ikdH[2][k];
    double D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    // B'_x += dt (D[z][y] ikz - D[y][z] iky), B'_y += dt (D[x][z] ikx - D[z][x] ikz),
    // B'_z += dt (D[y][x] iky - D[x][y] ikx)
    D[2][1] = a.dt * lb[0] * ikz;
    D[1][2] = -a.dt * lb[0] * iky;
    D[0][2] = a.dt * lb[1] * ikx;
    D[2][0] = -a.dt * lb[1] * ikz;
    D[1][0] = a.dt * lb[2] * iky;
    D[0][1] = -a.dt * lb[2] * ikx;
    const int coord[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        const int c0 = (d + 1) % 3, c1 = (d + 2) % 3;
        for (int s = 0; s < 2; ++s) {
            const int fi = a.hfbd[d][s];
            if (fi < 0) continue;
            const CpmlFace &f = a.hface[fi];
            if (!in_layer(f, d, coord[d])) continue;
            const size_t l = f.at(i, j, k);
            const int q = coord[d] - f.lo[d];
            const double bq = f.b[q], cq = f.c[q] * f.idxn;
            // n0 = b psi0 + cq D[d][c1] feeds M[c0]; n1 = b psi1 - cq D[d][c0] feeds M[c1]
            const double a0 = a.lpsiH[fi][0][l] + a.dt * lb[c0];
            const double a1 = a.lpsiH[fi][1][l] + a.dt * lb[c1];
            D[d][c1] += a0 * cq;
            D[d][c0] -= a1 * cq;
            a.lpsiH[fi][0][l] = bq * a0;
            a.lpsiH[fi][1][l] = bq * a1;
        }
    }
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < 3; ++q)
            if (q != d) a.dEb[slot(d, q)][c] = D[d][q];
}

// Adjoint of the E kernel at E cell (i, j, k).
__global__ __launch_bounds__(256) void k_adj_eside(AdjArgs a)
{
    const int k = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    const GridDims &g = a.g;
    if (k > g.n[2] || j > g.n[1]) return;
    const size_t c = g.at(i, j, k);
    const size_t cm[3] = {g.at(wrapm(i, g.n[0]), j, k), g.at(i, wrapm(j, g.n[1]), k), g.at(i, j, wrapm(k, g.n[2]))};
    // adjoint of E' at c: its own adjoint plus the differences D_E'[d][q] = E'_q(c + e_d) - E'_q(c)
    double le[3];
    for (int q = 0; q < 3; ++q) {
        double v = a.lE[q][c];
        for (int d = 0; d < 3; ++d)
            if (d != q) v += a.dEb[slot(d, q)][cm[d]] - a.dEb[slot(d, q)][c];
        le[q] = v;
    }
    const double e0[3] = {a.E0[0][c], a.E0[1][c], a.E0[2][c]};
    const double e1[3] = {a.E1[0][c], a.E1[1][c], a.E1[2][c]};
    if (k == a.kobj)
        for (int q = 0; q < 3; ++q) le[q] += 2.0 * a.wobj * e1[q];
    const double eps = a.epsf ? a.epsf[c] : a.eps;
    double S = 0.0;  // sum_p Re Dd_p: the part of sum_p Re P_p proportional to E
    for (int p = 0; p < a.np; ++p) S += a.pc.dd[p].x;
    const double denom = eps + a.sumChi0 - a.sumXi0;
    const double numer = eps - a.sumXi0;
    const double A = a.chi / denom;
    // the Kerr solve E' = (e / s) C: C is recovered from E' (e = |E'|, s = A e^3 + e)
    const double e = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    double cv[3] = {0, 0, 0}, cb[3], Ab = 0.0;
    if (e > 0) {
        const double s = A * e * e * e + e;
        const double phi = e / s;
        const double des = 1.0 / (3.0 * A * e * e + 1.0);
        const double dphi = (des - phi) / s;
        for (int q = 0; q < 3; ++q) cv[q] = e1[q] / phi;
        const double cl = cv[0] * le[0] + cv[1] * le[1] + cv[2] * le[2];
        for (int q = 0; q < 3; ++q) cb[q] = phi * le[q] + (dphi / s) * cl * cv[q];
        Ab = (cl / s) * (-e * e * e * des);
    } else {
        for (int q = 0; q < 3; ++q) cb[q] = le[q];  // the limit of dE'/dC at C = 0 is the identity
    }
    // C denom = (numer + S + chi |E|^2) E + dt R + sum_p Re Psi_p; denom and numer both
    // grow by d eps, so dC/deps = (E - C) / denom
    const double E2 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    const double Ecb = e0[0] * cb[0] + e0[1] * cb[1] + e0[2] * cb[2];
    const double f = numer + S + a.chi * E2;
    double lnew[3];
    for (int q = 0; q < 3; ++q) lnew[q] = (f * cb[q] + 2.0 * a.chi * e0[q] * Ecb) / denom;
    // poles: Psi_p' = G_p E + Crec_p Psi_p with G_p = dxi0_p + Crec_p Dd_p, and C reads Re Psi_p
    for (int p = 0; p < a.np; ++p) {
        const double2 K = a.pc.crec[p], X = a.pc.dxi0[p], Dd = a.pc.dd[p];
        const double Gr = X.x + (K.x * Dd.x - K.y * Dd.y), Gi = X.y + (K.x * Dd.y + K.y * Dd.x);
        for (int q = 0; q < 3; ++q) {
            const double2 l = a.lpsi[p][q][c];
            lnew[q] += Gr * l.x + Gi * l.y;
            a.lpsi[p][q][c] = make_double2(K.x * l.x + K.y * l.y + cb[q] / denom, K.x * l.y - K.y * l.x);
        }
    }
    for (int q = 0; q < 3; ++q) a.lE[q][c] = lnew[q];
    double ecb = 0.0;
    for (int q = 0; q < 3; ++q) ecb += (e0[q] - cv[q]) * cb[q];
    a.epsb[c] += -Ab * a.chi / (denom * denom) + ecb / denom;
    a.chib[c] += Ab / denom + E2 * Ecb / denom;
    double rb[3];
    for (int q = 0; q < 3; ++q) rb[q] = a.dt * cb[q] / denom;
    // R_x = D[y][z] iky - D[z][y] ikz + J_x, R_y = D[z][x] ikz - D[x][z] ikx + J_y,
    // R_z = D[x][y] ikx - D[y][x] iky + J_z
    const double ikx = a.ikdE[0][i], iky = a.ikdE[1][j], ikz = a.ikdE[2][k];
    double D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    D[1][2] = rb[0] * iky;
    D[2][1] = -rb[0] * ikz;
    D[2][0] = rb[1] * ikz;
    D[0][2] = -rb[1] * ikx;
    D[0][1] = rb[2] * ikx;
    D[1][0] = -rb[2] * iky;
    const int coord[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        const int c0 = (d + 1) % 3, c1 = (d + 2) % 3;
        for (int s = 0; s < 2; ++s) {
            const int fi = a.efbd[d][s];
            if (fi < 0) continue;
            const CpmlFace &fc = a.eface[fi];
            if (!in_layer(fc, d, coord[d])) continue;
            const size_t l = fc.at(i, j, k);
            const int qq = coord[d] - fc.lo[d];
            const double bq = fc.b[qq], cq = fc.c[qq] * fc.idxn;
            // n0 = b psi0 - cq D[d][c1] feeds J[c0]; n1 = b psi1 + cq D[d][c0] feeds J[c1]
            const double a0 = a.lpsiE[fi][0][l] + rb[c0];
            const double a1 = a.lpsiE[fi][1][l] + rb[c1];
            D[d][c1] -= a0 * cq;
            D[d][c0] += a1 * cq;
            a.lpsiE[fi][0][l] = bq * a0;
            a.lpsiE[fi][1][l] = bq * a1;
        }
    }
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < 3; ++q)
            if (q != d) a.dBb[slot(d, q)][c] = D[d][q];
}

// Adjoint of B: the E kernel's differences D_B[d][q] = B_q(c) - B_q(c - e_d) give B_q(c)
// +dBb(c) and -dBb(c + e_d).
__global__ __launch_bounds__(256) void k_adj_bgather(AdjArgs a)
{
    const int k = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    const GridDims &g = a.g;
    if (k > g.n[2] || j > g.n[1]) return;
    const size_t c = g.at(i, j, k);
    const size_t cp[3] = {g.at(wrapp(i, g.n[0]), j, k), g.at(i, wrapp(j, g.n[1]), k), g.at(i, j, wrapp(k, g.n[2]))};
    for (int q = 0; q < 3; ++q) {
        double v = a.lB[q][c];
        for (int d = 0; d < 3; ++d)
            if (d != q) v += a.dBb[slot(d, q)][c] - a.dBb[slot(d, q)][cp[d]];
        a.lB[q][c] = v;
    }
}

// J += weight * sum_{i,j} |E(i, j, kobj)|^2 over the interior of the plane; one block, a fixed
// reduction order
__global__ __launch_bounds__(1024) void k_plane_objective(GridDims g, const double *Ex, const double *Ey,
                                                          const double *Ez, int kobj, double w, double *acc)
{
    __shared__ double red[1024];
    double v = 0.0;
    const int n = g.n[0] * g.n[1];
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int i = t / g.n[1] + 1, j = t % g.n[1] + 1;
        const size_t c = g.at(i, j, kobj);
        v += Ex[c] * Ex[c] + Ey[c] * Ey[c] + Ez[c] * Ez[c];
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += w * red[0];
}

static dim3 adj_grid(const GridDims &g) { return dim3((g.n[2] + 63) / 64, (g.n[1] + 3) / 4, g.n[0]); }

void launch_adjoint_step(const AdjArgs &a, hipStream_t s)
{
    const dim3 grd = adj_grid(a.g), blk(64, 4);
    hipLaunchKernelGGL(k_adj_bside, grd, blk, 0, s, a);
    hipLaunchKernelGGL(k_adj_eside, grd, blk, 0, s, a);
    hipLaunchKernelGGL(k_adj_bgather, grd, blk, 0, s, a);
}

void launch_plane_objective(const GridDims &g, const double *const E[3], int kobj, double w, double *acc,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_plane_objective, dim3(1), dim3(1024), 0, s, g, E[0], E[1], E[2], kobj, w, acc);
}

}  // namespace mpx
