// Single-pass leapfrog step on gfx950: B of step n and E of step n + 1 in one kernel.
//
// The reference advances E then B inside a step (FDTD_PLRCSolver::stepScheme,
// src/fdtd_plrc.t:127-168).  Shifting the step boundary by half a step puts B^(n+1) = B^n +
// dt (curl E^(n+1) + M) and then E^(n+2) = F(E^(n+1), B^(n+1), Psi, J) in one pass over
// HBM, so E, B and the pole accumulators are each read and written once per step -- the
// algorithmic byte count of the step -- instead of the two fused kernels' extra read of B (E
// side) and of E (B side).  The solver runs the E half of the first step and the B half owed
// at a read with the fused kernels; the arithmetic of every update is theirs, term for term,
// so the fields agree with the fused path bit for bit.
//
// The E update at a cell needs B^(n+1) of the cell and of its x - 1, y - 1 and z - 1
// neighbours.  Each thread updates one cell and computes the B components it needs at the
// three neighbours again from old values (their owners compute and store them too).  The
// extra reads are E and B words that neighbouring threads read as well (L1 / L2 hits in the
// banded tile order); the pole accumulators -- two thirds of the bytes -- are read once.
// Threads never wait for each other.  E, B and the CPML H auxiliaries, which the recomputations read, are
// ping-ponged ("o" read, "n" written); the accumulators and the CPML E auxiliaries, touched
// only by their owner, are updated in place.  Periodic boundary only.
#include <hip/hip_runtime.h>

#include "mpx_device.h"
#include "mpx_launch.h"

namespace mpx {

// One CPML H face at cell l: the old auxiliaries come from the step's old buffers; the new
// ones go to the face's (new) buffers when the cell is this wave's (store).  Returns the
// stored values (CPMLBorderHCurrent::stepScheme, src/cpml_border.cpp:499-527).
template <typename T, bool RES>
__device__ __forceinline__ void cpml_hface_pp(const CpmlFace &f, const T *o0, const T *o1, const unsigned *orw,
                                              int coord, int l, double d1, double d0, bool store, double &v0,
                                              double &v1)
{
    const int q = coord - f.lo[f.dim];
    const double bq = gld(f.b + q), cq = gld(f.c + q);
    const unsigned w = lo_ld<RES>(orw, l);
    const double a0 = (double)rd<T, RES>(o0, w, 0, l), a1 = (double)rd<T, RES>(o1, w, 1, l);
    const double n0 = bq * a0 + cq * d1 * f.idxn;
    const double n1 = bq * a1 - cq * d0 * f.idxn;
    unsigned wn = 0;
    const T h0 = put<T, RES>(n0, wn, 0), h1 = put<T, RES>(n1, wn, 1);
    if (store) {
        gst((T *)f.psi0 + l, h0);
        gst((T *)f.psi1 + l, h1);
        if (RES) gst(f.rw + l, wn);
    }
    v0 = got<T, RES>(h0, wn, 0);
    v1 = got<T, RES>(h1, wn, 1);
}

// Magnetic current total at (i, j, k) in the reference's list order (sheets before or after
// the CPML faces; faces N,S (y), E,W (x), U,D (z), high face first), from the six E
// differences of the cell.
template <typename T, bool CPML, bool RES>
__device__ __forceinline__ void pass_m(const PassArgs<T> &a, int i, int j, int k, bool store, double dEy_z,
                                       double dEx_z, double dEz_y, double dEx_y, double dEz_x, double dEy_x,
                                       double &mx, double &my, double &mz, bool cp, bool sh)
{
    mx = my = mz = 0.0;
    if (sh && a.sheets_first) add_sheet_planes(a.hsheet, a.nhsheets, a.g.py, i, j, k, mx, my, mz);
    if (CPML && cp) {
        bool in[2];
        in_faces(a.hfbd, a.hface, 1, j, in);
#pragma unroll
        for (int s = 0; s < 2; ++s)
            if (in[s]) {
                const int fi = a.hfbd[1][s];
                const CpmlFace &f = a.hface[fi];
                double v0, v1;  // psi0 -> Mz from dEx_y, psi1 -> Mx from dEz_y
                cpml_hface_pp<T, RES>(f, (const T *)a.hold[fi][0], (const T *)a.hold[fi][1], a.hold_rw[fi], j,
                                      layer_index(f, i, j, k), dEx_y, dEz_y, store, v0, v1);
                mz += v0;
                mx += v1;
            }
        in_faces(a.hfbd, a.hface, 0, i, in);
#pragma unroll
        for (int s = 0; s < 2; ++s)
            if (in[s]) {
                const int fi = a.hfbd[0][s];
                const CpmlFace &f = a.hface[fi];
                double v0, v1;  // psi0 -> My from dEz_x, psi1 -> Mz from dEy_x
                cpml_hface_pp<T, RES>(f, (const T *)a.hold[fi][0], (const T *)a.hold[fi][1], a.hold_rw[fi], i,
                                      layer_index(f, i, j, k), dEz_x, dEy_x, store, v0, v1);
                my += v0;
                mz += v1;
            }
        in_faces(a.hfbd, a.hface, 2, k, in);
#pragma unroll
        for (int s = 0; s < 2; ++s)
            if (in[s]) {
                const int fi = a.hfbd[2][s];
                const CpmlFace &f = a.hface[fi];
                double v0, v1;  // psi0 -> Mx from dEy_z, psi1 -> My from dEx_z
                cpml_hface_pp<T, RES>(f, (const T *)a.hold[fi][0], (const T *)a.hold[fi][1], a.hold_rw[fi], k,
                                      layer_index(f, i, j, k), dEy_z, dEx_z, store, v0, v1);
                mx += v0;
                my += v1;
            }
    }
    if (sh && !a.sheets_first) add_sheet_planes(a.hsheet, a.nhsheets, a.g.py, i, j, k, mx, my, mz);
}

// Electric current total at (i, j, k) (CPML E auxiliaries updated in place; owned cells only).
template <typename T, bool CPML, bool RES>
__device__ __forceinline__ void pass_j(const PassArgs<T> &a, int i, int j, int k, double dBz_y, double dBy_z,
                                       double dBx_z, double dBz_x, double dBy_x, double dBx_y, double &jx,
                                       double &jy, double &jz, bool cp, bool sh)
{
    jx = jy = jz = 0.0;
    auto add_cpml = [&]() {
        if (CPML && cp) {
            bool in[2];
            in_faces(a.efbd, a.eface, 1, j, in);
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (in[s]) {
                    const CpmlFace &f = a.eface[a.efbd[1][s]];
                    double v0, v1;
                    cpml_face<T, false, RES>(f, j, layer_index(f, i, j, k), dBx_y, dBz_y, v0, v1);
                    jz += v0;
                    jx += v1;
                }
            in_faces(a.efbd, a.eface, 0, i, in);
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (in[s]) {
                    const CpmlFace &f = a.eface[a.efbd[0][s]];
                    double v0, v1;
                    cpml_face<T, false, RES>(f, i, layer_index(f, i, j, k), dBz_x, dBy_x, v0, v1);
                    jy += v0;
                    jz += v1;
                }
            in_faces(a.efbd, a.eface, 2, k, in);
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (in[s]) {
                    const CpmlFace &f = a.eface[a.efbd[2][s]];
                    double v0, v1;
                    cpml_face<T, false, RES>(f, k, layer_index(f, i, j, k), dBy_z, dBx_z, v0, v1);
                    jx += v0;
                    jy += v1;
                }
        }
    };
    if (a.sheets_first) {
        if (sh) add_sheet_planes(a.sheet, a.nsheets, a.g.py, i, j, k, jx, jy, jz);
        add_cpml();
    } else {
        add_cpml();
        if (sh) add_sheet_planes(a.sheet, a.nsheets, a.g.py, i, j, k, jx, jy, jz);
    }
}

// B^(n+1) at one cell as the B update stores it: the stored (rounded) value of each component
template <typename T, bool RES>
__device__ __forceinline__ void b_update(double bx0, double by0, double bz0, double dEy_z, double dEx_z,
                                         double dEz_y, double dEx_y, double dEz_x, double dEy_x, double kdx,
                                         double kdy, double kdz, double dt, double mx, double my, double mz,
                                         T h[3], unsigned &w, double b[3])
{
    w = 0;
    h[0] = put<T, RES>(bx0 + dt * ((dEy_z * kdz - dEz_y * kdy) + mx), w, 0);
    h[1] = put<T, RES>(by0 + dt * ((dEz_x * kdx - dEx_z * kdz) + my), w, 1);
    h[2] = put<T, RES>(bz0 + dt * ((dEx_y * kdy - dEy_x * kdx) + mz), w, 2);
    b[0] = got<T, RES>(h[0], w, 0);
    b[1] = got<T, RES>(h[1], w, 1);
    b[2] = got<T, RES>(h[2], w, 2);
}

// E^(n+1) or B^n of a cell as stored, decoded
template <typename T, bool RES>
struct Cell3 {
    typename Word<T, RES>::V x, y, z;
};
template <typename T, bool RES>
__device__ __forceinline__ Cell3<T, RES> ld3(const T *const f[3], const unsigned *lo, int c)
{
    const unsigned w = lo_ld<RES>(lo, c);
    return {rd<T, RES>(f[0], w, 0, c), rd<T, RES>(f[1], w, 1, c), rd<T, RES>(f[2], w, 2, c)};
}

// B^(n+1) at a cell from its B^n, the six E differences (dEy_z, dEx_z, dEz_y, dEx_y, dEz_x,
// dEy_x) and its position; the CPML H auxiliaries are written only with store
template <typename T, bool CPML, bool RES>
__device__ __forceinline__ void b_cell(const PassArgs<T> &a, int i, int j, int k, const Cell3<T, RES> &b0,
                                       const double d[6], bool store, T h[3], unsigned &w, double b[3], bool cp,
                                       bool sh)
{
    double mx, my, mz;
    pass_m<T, CPML, RES>(a, i, j, k, store, d[0], d[1], d[2], d[3], d[4], d[5], mx, my, mz, cp, sh);
    b_update<T, RES>((double)b0.x, (double)b0.y, (double)b0.z, d[0], d[1], d[2], d[3], d[4], d[5], gld(a.ikdH[0] + i),
                     gld(a.ikdH[1] + j), gld(a.ikdH[2] + k), a.dt, mx, my, mz, h, w, b);
}

// One cell per thread, in the fused kernels' banded tile order (64 z cells x 4 rows per block).
// The E update at (i, j, k) takes B^(n+1) at the cell itself (computed and stored here) and at
// x - 1, y - 1 and z - 1 (computed here again from old values and not stored: their owners
// store them).  Every grid word the thread needs is loaded before anything is computed or
// stored, so that one wait covers them; the neighbours' B is computed before the cell's own
// (whose CPML H auxiliaries are stored), and all field stores come last.
template <typename T, int NP, bool NONLIN, bool CPML, bool RES>
__global__ __launch_bounds__(256) void k_pass(const PassArgs<T> *__restrict__ pa, int nbx, int nby, int nbz)
{
    using V = typename Word<T, RES>::V;
    using C3 = Cell3<T, RES>;
    const PassArgs<T> &a = *pa;
    const GridDims &g = a.g;
    const int Nx = g.n[0], Ny = g.n[1], Nz = g.n[2];
    const int sx = (int)g.sx, pz = g.pz;
    int tbx, tby, tbz;
    fused_tile_of(2, a.band, 0, nbx, nby, nbz, tbx, tby, tbz);
    const int lane = threadIdx.x;
    const int k = tbx * 64 + lane + 1, j = tby * 4 + threadIdx.y + 1, i = tbz + 1;
    if (k > Nz || j > Ny || i > Nx) return;
    const int c = i * sx + j * pz + k + g.z0;
    const int ox_m = i == 1 ? (Nx - 1) * sx : -sx, ox_p = i == Nx ? -(Nx - 1) * sx : sx;
    const int oy_m = j == 1 ? (Ny - 1) * pz : -pz, oy_p = j == Ny ? -(Ny - 1) * pz : pz;
    const int oz_m = k == 1 ? Nz - 1 : -1, oz_p = k == Nz ? -(Nz - 1) : 1;
    const int im = i == 1 ? Nx : i - 1, jm = j == 1 ? Ny : j - 1, km = k == 1 ? Nz : k - 1;
    const bool zload = lane == 63 || k == Nz;  // the z + 1 neighbour is not in the next lane
    const T *const *Eo = a.Eo, *const *Bo = a.Bo;
    // Whether a CPML face (cp) or a source sheet (sh) holds any cell this wave updates or
    // recomputes (x: i - 1, i; y: j - 1, j; z: k - 1, k): most waves meet neither and skip the
    // face and sheet logic as a whole (wave-uniform branches).
    bool cp = false, sh = false;
    if (CPML) {
        const int lo[3] = {im, jm, km}, hi[3] = {i, j, k};
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int fh = a.hfbd[d][q], fe = a.efbd[d][q];
                // a periodic x - 1 / y - 1 / z - 1 of index 1 is N: both ends are tested
                if (fh >= 0)
                    cp |= (lo[d] >= a.hface[fh].lo[d] && lo[d] <= a.hface[fh].hi[d]) ||
                          (hi[d] >= a.hface[fh].lo[d] && hi[d] <= a.hface[fh].hi[d]);
                if (fe >= 0) cp |= hi[d] >= a.eface[fe].lo[d] && hi[d] <= a.eface[fe].hi[d];
            }
        cp = __ballot(cp) != 0;
    }
    for (int q = 0; q < a.nhsheets; ++q) sh |= a.hsheet[q].kp == k || a.hsheet[q].kp == km;
    for (int q = 0; q < a.nsheets; ++q) sh |= a.sheet[q].kp == k;
    sh = __ballot(sh) != 0;

    // E^(n+1) of the cell and of the cells the four B values need
    const C3 e0 = ld3<T, RES>(Eo, a.Eolo, c);
    const C3 exp_ = ld3<T, RES>(Eo, a.Eolo, c + ox_p), eyp = ld3<T, RES>(Eo, a.Eolo, c + oy_p);
    const C3 exm = ld3<T, RES>(Eo, a.Eolo, c + ox_m), eym = ld3<T, RES>(Eo, a.Eolo, c + oy_m);
    const C3 ezm = ld3<T, RES>(Eo, a.Eolo, c + oz_m);
    const C3 exm_yp = ld3<T, RES>(Eo, a.Eolo, c + ox_m + oy_p), eym_xp = ld3<T, RES>(Eo, a.Eolo, c + oy_m + ox_p);
    const C3 ezm_yp = ld3<T, RES>(Eo, a.Eolo, c + oz_m + oy_p), ezm_xp = ld3<T, RES>(Eo, a.Eolo, c + oz_m + ox_p);
    C3 z0{}, zx{}, zy{};  // z + 1 of the cell and of its x - 1 and y - 1 neighbours
    if (zload) {
        z0 = ld3<T, RES>(Eo, a.Eolo, c + oz_p);
        zx = ld3<T, RES>(Eo, a.Eolo, c + ox_m + oz_p);
        zy = ld3<T, RES>(Eo, a.Eolo, c + oy_m + oz_p);
    }
    // B^n of the cell and of its x - 1, y - 1, z - 1 neighbours
    const unsigned wb = lo_once<RES>(a.Bolo, c);
    const C3 b0{rd_once<T, RES>(Bo[0], wb, 0, c), rd_once<T, RES>(Bo[1], wb, 1, c), rd_once<T, RES>(Bo[2], wb, 2, c)};
    const C3 bxm = ld3<T, RES>(Bo, a.Bolo, c + ox_m), bym = ld3<T, RES>(Bo, a.Bolo, c + oy_m),
             bzm = ld3<T, RES>(Bo, a.Bolo, c + oz_m);
    PsiRegs<NP> pr;
#pragma unroll
    for (int n = 0; n < NP; ++n)
#pragma unroll
        for (int q = 0; q < 3; ++q) pr.v[n][q] = load_once(&a.psi[n][q][psi_at(c)]);

    // z + 1 neighbours from the next lane
    {
        const V n0x = lane_next(e0.x), n0y = lane_next(e0.y), nxx = lane_next(exm.x), nxy = lane_next(exm.y),
                nyx = lane_next(eym.x), nyy = lane_next(eym.y);
        if (!zload) {
            z0.x = n0x;
            z0.y = n0y;
            zx.x = nxx;
            zx.y = nxy;
            zy.x = nyx;
            zy.y = nyy;
        }
    }
    const double ex = e0.x, ey = e0.y, ez = e0.z;
    T hh[3];
    unsigned hw;
    // B^(n+1) at x - 1 (By, Bz are used), y - 1 (Bx, Bz) and z - 1 (Bx, By): not stored
    double bx_m[3], by_m[3], bz_m[3];
    {
        const double x_ = exm.x, y_ = exm.y, z_ = exm.z;
        const double d[6] = {(double)zx.y - y_, (double)zx.x - x_, (double)exm_yp.z - z_, (double)exm_yp.x - x_,
                             ez - z_, ey - y_};
        b_cell<T, CPML, RES>(a, im, j, k, bxm, d, false, hh, hw, bx_m, cp, sh);
    }
    {
        const double x_ = eym.x, y_ = eym.y, z_ = eym.z;
        const double d[6] = {(double)zy.y - y_, (double)zy.x - x_, ez - z_, ex - x_, (double)eym_xp.z - z_,
                             (double)eym_xp.y - y_};
        b_cell<T, CPML, RES>(a, i, jm, k, bym, d, false, hh, hw, by_m, cp, sh);
    }
    {
        const double x_ = ezm.x, y_ = ezm.y, z_ = ezm.z;
        const double d[6] = {ey - y_, ex - x_, (double)ezm_yp.z - z_, (double)ezm_yp.x - x_, (double)ezm_xp.z - z_,
                             (double)ezm_xp.y - y_};
        b_cell<T, CPML, RES>(a, i, j, km, bzm, d, false, hh, hw, bz_m, cp, sh);
    }
    // B^(n+1) at the cell (its CPML H auxiliaries stored)
    T bh[3];
    unsigned bw;
    double b[3];
    {
        const double d[6] = {(double)z0.y - ey, (double)z0.x - ex, (double)eyp.z - ez, (double)eyp.x - ex,
                             (double)exp_.z - ez, (double)exp_.y - ey};
        b_cell<T, CPML, RES>(a, i, j, k, b0, d, true, bh, bw, b, cp, sh);
    }
    // E^(n+2) at the cell
    const double dBz_y = b[2] - by_m[2], dBx_y = b[0] - by_m[0];
    const double dBy_z = b[1] - bz_m[1], dBx_z = b[0] - bz_m[0];
    const double dBz_x = b[2] - bx_m[2], dBy_x = b[1] - bx_m[1];
    double jx, jy, jz;
    pass_j<T, CPML, RES>(a, i, j, k, dBz_y, dBy_z, dBx_z, dBz_x, dBy_x, dBx_y, jx, jy, jz, cp, sh);
    double idenom = a.idenom, numer = a.numer;
    if (a.epsf) {
        double denom;
        cell_eps<NONLIN>(a, (size_t)c, 0.0, denom, numer);
        idenom = 1.0 / denom;
    }
    double nex = ex, ney = ey, nez = ez;
    unsigned fl = 0;
    plrc_e_update_pp<NP, NONLIN, true>(nex, ney, nez, dBz_y, dBy_z, dBx_z, dBz_x, dBy_x, dBx_y, gld(a.ikdE[0] + i),
                                       gld(a.ikdE[1] + j), gld(a.ikdE[2] + k), jx, jy, jz, pr, a.psi, (size_t)c, true,
                                       a.pc, a.dt, idenom, numer, a.chi, a.ncap, &fl);
    report(a.status, fl | finite_flag(nex, ney, nez));
    unsigned ew = 0;
    const T hx = put<T, RES>(nex, ew, 0), hy = put<T, RES>(ney, ew, 1), hz = put<T, RES>(nez, ew, 2);
    store_once(&a.Bn[0][c], bh[0]);
    store_once(&a.Bn[1][c], bh[1]);
    store_once(&a.Bn[2][c], bh[2]);
    if (RES) store_once(&a.Bnlo[c], bw);
    store_once(&a.En[0][c], hx);
    store_once(&a.En[1][c], hy);
    store_once(&a.En[2][c], hz);
    if (RES) store_once(&a.Enlo[c], ew);
}

// tile grid: 64 z cells x 4 rows x 1 plane per block
void pass_tiles(const GridDims &g, int &nbx, int &nby, int &nbz)
{
    nbx = (g.n[2] + 63) / 64;
    nby = (g.n[1] + 3) / 4;
    nbz = g.n[0];
}

template <typename T, bool RES>
static void launch_pass_t(const PassArgs<T> &a, const PassArgs<T> *dev, int np, bool nonlin, hipStream_t s)
{
    int nbx, nby, nbz;
    pass_tiles(a.g, nbx, nby, nbz);
    dim3 grd(nbx, nby, nbz), blk(64, 4);
    const bool cp = a.nefaces > 0 || a.nhfaces > 0;
#define MPX_PS(NPV, NL, CP) hipLaunchKernelGGL((k_pass<T, NPV, NL, CP, RES>), grd, blk, 0, s, dev, nbx, nby, nbz)
#define MPX_PS_NP(NL, CP)            \
    switch (np) {                    \
    case 0: MPX_PS(0, NL, CP); break; \
    case 1: MPX_PS(1, NL, CP); break; \
    case 2: MPX_PS(2, NL, CP); break; \
    default: MPX_PS(3, NL, CP); break; \
    }
    if (nonlin) {
        if (cp) { MPX_PS_NP(true, true) } else { MPX_PS_NP(true, false) }
    } else {
        if (cp) { MPX_PS_NP(false, true) } else { MPX_PS_NP(false, false) }
    }
#undef MPX_PS_NP
#undef MPX_PS
}

template <typename T>
void launch_pass(const PassArgs<T> &a, const PassArgs<T> *dev, int np, bool nonlin, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        if (a.Eolo) {
            launch_pass_t<T, true>(a, dev, np, nonlin, s);
            return;
        }
    }
    launch_pass_t<T, false>(a, dev, np, nonlin, s);
}
template void launch_pass<double>(const PassArgs<double> &, const PassArgs<double> *, int, bool, hipStream_t);
template void launch_pass<float>(const PassArgs<float> &, const PassArgs<float> *, int, bool, hipStream_t);

}  // namespace mpx
