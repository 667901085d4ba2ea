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
// neighbours.  Two kernels provide them.  k_march (the default, below) walks each tile along x
// and passes the neighbours' values between the waves of a block through LDS.  k_pass updates
// one cell per thread and computes the B components it needs at the three neighbours again
// from old values (their owners compute and store them too; single_periodic boundary, no
// plasma).  E, B and the CPML H auxiliaries, which neighbours read, are ping-ponged ("o" read,
// "n" written); the accumulators, the CPML E auxiliaries and the plasma grids, touched only by
// their owner, are updated in place.
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
                                       double &mx, double &my, double &mz, bool cp, bool sh, bool live = true)
{
    mx = my = mz = 0.0;
    if (sh && a.sheets_first) add_sheet_planes(a.hsheet, a.nhsheets, a.g.py, i, j, k, mx, my, mz);
    if (CPML && cp) {
        bool in[2];
        in_faces(a.hfbd, a.hface, 1, j, in);
        in[0] = in[0] && live;
        in[1] = in[1] && live;
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
        in[0] = in[0] && live;
        in[1] = in[1] && live;
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
        in[0] = in[0] && live;
        in[1] = in[1] && live;
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
                                       double &jy, double &jz, bool cp, bool sh, bool live = true,
                                       const double *jpv = nullptr)
{
    jx = jy = jz = 0.0;
    // the plasma current (jpv) joins the totals at its place in the list (pcur_pos: before the
    // first group, between the two, after both)
    auto add_plasma = [&](int pos) {
        if (jpv && a.pcur && a.pcur_pos == pos) {
            jx += jpv[0];
            jy += jpv[1];
            jz += jpv[2];
        }
    };
    auto add_cpml = [&]() {
        if (CPML && cp) {
            bool in[2];
            in_faces(a.efbd, a.eface, 1, j, in);
            in[0] = in[0] && live;
            in[1] = in[1] && live;
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
            in[0] = in[0] && live;
            in[1] = in[1] && live;
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
            in[0] = in[0] && live;
            in[1] = in[1] && live;
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
    add_plasma(0);
    if (a.sheets_first) {
        if (sh) add_sheet_planes(a.sheet, a.nsheets, a.g.py, i, j, k, jx, jy, jz);
        add_plasma(1);
        add_cpml();
    } else {
        add_cpml();
        add_plasma(1);
        if (sh) add_sheet_planes(a.sheet, a.nsheets, a.g.py, i, j, k, jx, jy, jz);
    }
    add_plasma(2);
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
// (kdx, kdy, kdz: the reciprocal kappa * delta lines of the H side at i, j, k)
template <typename T, bool CPML, bool RES>
__device__ __forceinline__ void b_cell_k(const PassArgs<T> &a, int i, int j, int k, const Cell3<T, RES> &b0,
                                         const double d[6], bool store, T h[3], unsigned &w, double b[3], bool cp,
                                         bool sh, bool live, double kdx, double kdy, double kdz)
{
    double mx, my, mz;
    pass_m<T, CPML, RES>(a, i, j, k, store, d[0], d[1], d[2], d[3], d[4], d[5], mx, my, mz, cp, sh, live);
    b_update<T, RES>((double)b0.x, (double)b0.y, (double)b0.z, d[0], d[1], d[2], d[3], d[4], d[5], kdx, kdy, kdz,
                     a.dt, mx, my, mz, h, w, b);
}
template <typename T, bool CPML, bool RES>
__device__ __forceinline__ void b_cell(const PassArgs<T> &a, int i, int j, int k, const Cell3<T, RES> &b0,
                                       const double d[6], bool store, T h[3], unsigned &w, double b[3], bool cp,
                                       bool sh, bool live = true)
{
    b_cell_k<T, CPML, RES>(a, i, j, k, b0, d, store, h, w, b, cp, sh, live, gld(a.ikdH[0] + i), gld(a.ikdH[1] + j),
                           gld(a.ikdH[2] + k));
}

// k_pass: one cell per thread, in the fused kernels' banded tile order (64 z cells x 4 rows per
// block).  The extra reads of the recomputations are E and B words that neighbouring threads read
// as well (L1 / L2 hits); the pole accumulators -- two thirds of the bytes -- are read once.
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

// ---------------------------------------------------------------------------------------
// The same pass with each block marching over a segment of x planes (k_march).
//
// A block owns 64 z cells x R rows and walks `seg` x planes.  It runs R + 1 waves: waves 1..R
// hold the rows j0 .. j0 + R - 1, wave 0 (the halo wave) the row j0 - 1 below the tile, whose
// B^(n+1) the first row's E update needs.  Per plane every wave computes B^(n+1) of its row;
// the E update then takes B of the x - 1 plane from the wave's own result of the previous plane
// (registers), B of the y - 1 row from the wave below and B of the z - 1 cell from the previous
// lane (DPP) or, for lane 0, from the column k0 - 1, which the halo wave's first R lanes
// compute beside its row.  The B curl's E of the y + 1 row comes from the wave above, and the
// z + 1 cell's E from the next lane or, at the end of the row segment, from the right column
// the halo wave reads.  Everything exchanged goes through LDS as decoded float64 values, so
// every word of E, B and the accumulators crosses HBM once per step apart from the halo (one
// row and one column cell per tile row, one x plane per segment, recomputed and never stored)
// and is decoded once.  One barrier per plane; the LDS rows are double-buffered by plane
// parity, so a row written for plane p + 1 never overwrites one still being read for plane p.
//
// A row wave requests the words of plane p + 1 as soon as it has used those of plane p, and
// the accumulators of plane p + 1 after the E update of plane p.  Every such load and every
// field store is issued unconditionally, so that the number of memory operations in flight
// is the same on every path and a wave waits only for the plane it computes, not for the one
// it prefetched or for its own stores: lanes outside the grid (a partial last row of tiles or
// z tile) read and write the front padding of their row (storage positions 0 .. 30 of a row,
// before the ghost cell k = 0, hold no cell and are read by no kernel), the last plane of a
// segment prefetches position 0 of the plane (padding), and the CPML auxiliaries, which only
// some cells have, are updated under a per-lane mask.
// ---------------------------------------------------------------------------------------

// a cell of E or B as stored: three words and the residual word (zero without residuals)
template <typename T, bool RES>
struct Raw3 {
    T x, y, z;
    unsigned w;
};
template <typename T, bool RES>
__device__ __forceinline__ Cell3<T, RES> cell_of(const Raw3<T, RES> &r)
{
    if constexpr (RES) return {res_decode(r.x, r.w, 0), res_decode(r.y, r.w, 1), res_decode(r.z, r.w, 2)};
    else return {r.x, r.y, r.z};
}
// two decoded components of the 64 cells of a row segment, staged in LDS
struct RowD {
    double a[64], b[64];
};

// Word at byte offset bo (per lane, 32 bits: a plane holds (ny + 2) pz words < 2^32 bytes)
// from the plane base `base` (wave-uniform)
template <typename T>
__device__ __forceinline__ T ldp(const T *base, unsigned bo)
{
    return gld((const T *)((const char *)base + bo));
}
template <typename T>
__device__ __forceinline__ void stp_once(T *base, unsigned bo, T v)
{
    store_once((T *)((char *)base + bo), v);
}
// A read-only table entry at a wave-uniform address, through the constant address space: a
// scalar load into SGPRs, counted apart from the vector loads in flight
__device__ __forceinline__ double kload(const double *p)
{
    return *(const __attribute__((address_space(4))) double *)p;
}
// three components and the residual word of the cell at bo in plane `pl` (storage offset)
template <typename T, bool RES>
__device__ __forceinline__ Raw3<T, RES> ldp3(const T *const f[3], const unsigned *lo, size_t pl, unsigned bo)
{
    return {ldp(f[0] + pl, bo), ldp(f[1] + pl, bo), ldp(f[2] + pl, bo), RES ? ldp(lo + pl, bo) : 0u};
}

// Whether a CPML face of the E side or of the H side holds cell (i, j, k)
template <typename T>
__device__ __forceinline__ bool in_cpml(const PassArgs<T> &a, bool eside, int i, int j, int k)
{
    const int x[3] = {i, j, k};
    bool in = false;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int f = eside ? a.efbd[d][q] : a.hfbd[d][q];
            if (f >= 0) {
                const CpmlFace &F = eside ? a.eface[f] : a.hface[f];
                in |= x[d] >= F.lo[d] && x[d] <= F.hi[d];
            }
        }
    return in;
}

// accumulators of the active poles through one base pointer per pole: component q of a pole
// sits 64 elements after component q - 1 (psi_at), so a lane needs one address per pole
template <int NP>
struct PsiBase {
    double2 *b[NP > 0 ? NP : 1];
    struct Pole {
        double2 *p;
        __device__ double2 *operator[](int q) const { return p + 64 * q; }
    };
    __device__ Pole operator[](int n) const { return Pole{b[n]}; }
};

// A block barrier that orders the waves' LDS accesses only.  __syncthreads() also waits for
// every global load and store the wave has in flight; the march kernel exchanges data between
// waves through LDS alone, so the loads of the next plane and the stores of this one stay in
// flight across the barrier.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One tile (64 z cells x R rows from (k0, j0) = (64 tz + 1, R ty + 1), x planes
// seg ts + 1 .. seg (ts + 1)) of the march; the kernels below map their blocks to tiles.
template <typename T, int NP, bool NONLIN, bool CPML, bool RES, int R, bool PL>
__device__ __forceinline__ void march_tile(const PassArgs<T> &a, int tz, int ty, int ts)
{
    using RW = Raw3<T, RES>;
    using C3 = Cell3<T, RES>;
    const GridDims &g = a.g;
    const int Nx = g.n[0], Ny = g.n[1], Nz = g.n[2];
    const size_t sx = (size_t)g.sx;
    const int pz = g.pz, z0 = g.z0;
    constexpr unsigned SZ = sizeof(T);
    // erow[b][s]: E (x, z) of tile slot s (row j0 - 1 + s, wrapped) at the plane of parity b;
    // brow[b][s]: B^(n+1) (x, z) of slot s; bcol[b][l]: B^(n+1) (x, y) of the column k0 - 1 of
    // tile row l; ecol[b][l]: E (x, y) of the right-end neighbour of tile row l
    __shared__ RowD erow[2][R + 2], brow[2][R + 1], bcol[2], ecol[2];

    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // wave-uniform role
    const int k0 = tz * 64 + 1, j0 = ty * R + 1;
    const int i0 = ts * a.seg + 1, i1 = min(i0 + a.seg - 1, Nx);
    const int pm = i0 == 1 ? Nx : i0 - 1;  // the x - 1 plane of the segment, recomputed
    const bool kok = k0 + lane <= Nz;
    const int k = kok ? k0 + lane : Nz;  // lanes past the row end compute a copy of cell Nz
    // the last lane of the segment and the cell after it (at the row end: the periodic image,
    // or with a non-periodic z the ghost cell as stored)
    const int lr = k0 + 63 <= Nz ? 63 : Nz - k0;
    const int kr = k0 + lr == Nz ? (a.zper ? 1 : Nz + 1) : k0 + 64;
    // the column before the segment: the periodic image of the row end, or with a non-periodic
    // z the cell k = 0, whose B the update covers (FDTD_PLRCSolver::stepB, B over [0, N] in z)
    const int kc = k0 == 1 ? (a.zper ? Nz : 0) : k0 - 1;
    const bool kzero = !a.zper && k0 == 1;    // this segment's column is the cell k = 0
    const bool kend = !a.zper && k0 + lr == Nz;  // this segment ends at the row end
    auto wrap_row = [&](int r) { return (r - 1) % Ny + 1; };
    bool sh = false;  // a source sheet lies on a z plane this wave touches (wave-uniform)
    unsigned fl = 0;

    if (w == 0) {
        // ---- halo wave: B^(n+1) of the row j0 - 1 and of the column k0 - 1 of each tile row;
        // E of the row above the tile and of the right-end cells, for the row waves
        const int j = j0 == 1 ? Ny : j0 - 1;
        const bool ok = kok;
        const unsigned pad = (unsigned)(j * pz + lane % 31);
        const unsigned o = ok ? (unsigned)(j * pz + k + z0) : pad;
        const unsigned ozp = ok ? (unsigned)(j * pz + (k == Nz ? (a.zper ? 1 : Nz + 1) : k + 1) + z0) : pad;
        const int ju = wrap_row(j0 + R);  // the row above the tile
        const unsigned ou = ok ? (unsigned)(ju * pz + k + z0) : pad;
        const int jc = j0 + lane;
        const bool col = lane < R && jc <= Ny;
        const unsigned oc = col ? (unsigned)(jc * pz + kc + z0) : 0u;
        const unsigned ocz = col ? (unsigned)(jc * pz + k0 + z0) : 0u;  // kc + 1 is k0, also across the wrap
        const unsigned ocy = col ? (unsigned)((jc == Ny ? 1 : jc + 1) * pz + kc + z0) : 0u;
        const unsigned ocr = col ? (unsigned)(jc * pz + kr + z0) : 0u;
        const unsigned oce = col ? (unsigned)(jc * pz + Nz + 1 + z0) : 0u;  // the row's ghost k = Nz + 1
        for (int q = 0; q < a.nhsheets; ++q) sh |= a.hsheet[q].kp == k || (col && a.hsheet[q].kp == kc);
        sh = __ballot(sh) != 0;
        const double khy = kload(a.ikdH[1] + j), khz = gld(a.ikdH[2] + k);
        const double khyc = gld(a.ikdH[1] + (col ? jc : j)), khzc = kload(a.ikdH[2] + kc);
        // E of the row above the tile and of the right-end cells at plane p, into buffer b
        auto provide = [&](int p, int b) {
            const size_t bp = (size_t)p * sx;
            const RW u = ldp3<T, RES>(a.Eo, a.Eolo, bp, ou * SZ);
            const RW rc = ldp3<T, RES>(a.Eo, a.Eolo, bp, ocr * SZ);
            const C3 ud = cell_of(u), rd_ = cell_of(rc);
            erow[b][R + 1].a[lane] = ud.x;
            erow[b][R + 1].b[lane] = ud.z;
            ecol[b].a[lane] = rd_.x;
            ecol[b].b[lane] = rd_.y;
        };
        provide(pm, 0);
        lds_barrier();
        provide(i0, 1);
        lds_barrier();
        for (int p = i0; p <= i1; ++p) {
            const int pn = p == Nx ? 1 : p + 1;
            const size_t bp = (size_t)p * sx, bn = (size_t)pn * sx;
            const int be = (p - i0 + 1) & 1, bb = (p - i0) & 1;  // parities of the E and B buffers
            const double kxh = kload(a.ikdH[0] + p);
            // this plane's words: the halo row, its z + 1 cell, the column cells
            const RW e0 = ldp3<T, RES>(a.Eo, a.Eolo, bp, o * SZ), en = ldp3<T, RES>(a.Eo, a.Eolo, bn, o * SZ);
            const RW b0 = ldp3<T, RES>(a.Bo, a.Bolo, bp, o * SZ), ez = ldp3<T, RES>(a.Eo, a.Eolo, bp, ozp * SZ);
            const RW ce0 = ldp3<T, RES>(a.Eo, a.Eolo, bp, oc * SZ), cb0 = ldp3<T, RES>(a.Bo, a.Bolo, bp, oc * SZ);
            const RW cez = ldp3<T, RES>(a.Eo, a.Eolo, bp, ocz * SZ), ceu = ldp3<T, RES>(a.Eo, a.Eolo, bp, ocy * SZ);
            const RW cen = ldp3<T, RES>(a.Eo, a.Eolo, bn, oc * SZ);
            bool cp = false;
            if (CPML) cp = __ballot((ok && in_cpml(a, false, p, j, k)) || (col && in_cpml(a, false, p, jc, kc))) != 0;
            {  // the halo row; E of its y + 1 row is the tile's first row (slot 1)
                const C3 e = cell_of(e0), n = cell_of(en), z = cell_of(ez);
                const double d[6] = {(double)z.y - e.y, (double)z.x - e.x, erow[be][1].b[lane] - e.z,
                                     erow[be][1].a[lane] - e.x, (double)n.z - e.z, (double)n.y - e.y};
                T h[3];
                unsigned hw;
                double b[3];
                b_cell_k<T, CPML, RES>(a, p, j, k, cell_of(b0), d, false, h, hw, b, cp, sh, ok, kxh, khy, khz);
                brow[bb][0].a[lane] = b[0];
                brow[bb][0].b[lane] = b[2];
            }
            {  // the column
                const C3 e = cell_of(ce0), z = cell_of(cez), u = cell_of(ceu), n = cell_of(cen);
                const double d[6] = {(double)z.y - e.y, (double)z.x - e.x, (double)u.z - e.z, (double)u.x - e.x,
                                     (double)n.z - e.z, (double)n.y - e.y};
                T h[3];
                unsigned hw;
                double b[3];
                b_cell_k<T, CPML, RES>(a, p, col ? jc : j, col ? kc : k, cell_of(cb0), d, kzero, h, hw, b, cp, sh,
                                       col, kxh, khyc, khzc);
                bcol[bb].a[lane] = b[0];
                bcol[bb].b[lane] = b[1];
                if (kzero && col) {
                    // B of the cell k = 0 is this segment's to store; E there is a ghost the
                    // update never writes: carried into the new set as it stands
                    const unsigned ob = oc * SZ;
                    stp_once(a.Bn[0] + bp, ob, h[0]);
                    stp_once(a.Bn[1] + bp, ob, h[1]);
                    stp_once(a.Bn[2] + bp, ob, h[2]);
                    if (RES) stp_once(a.Bnlo + bp, ob, hw);
                    stp_once(a.En[0] + bp, ob, ce0.x);
                    stp_once(a.En[1] + bp, ob, ce0.y);
                    stp_once(a.En[2] + bp, ob, ce0.z);
                    if (RES) stp_once(a.Enlo + bp, ob, ce0.w);
                }
            }
            if (kend && col) {  // the ghosts k = Nz + 1 of E and B: carried into the new set
                const unsigned ob = oce * SZ;
                const RW ge = ldp3<T, RES>(a.Eo, a.Eolo, bp, ob), gb = ldp3<T, RES>(a.Bo, a.Bolo, bp, ob);
                stp_once(a.En[0] + bp, ob, ge.x);
                stp_once(a.En[1] + bp, ob, ge.y);
                stp_once(a.En[2] + bp, ob, ge.z);
                if (RES) stp_once(a.Enlo + bp, ob, ge.w);
                stp_once(a.Bn[0] + bp, ob, gb.x);
                stp_once(a.Bn[1] + bp, ob, gb.y);
                stp_once(a.Bn[2] + bp, ob, gb.z);
                if (RES) stp_once(a.Bnlo + bp, ob, gb.w);
            }
            if (p < i1) provide(pn, be ^ 1);
            lds_barrier();
        }
    } else {
        // ---- row wave: B^(n+1) and E^(n+2) of row j0 + w - 1 (slot w)
        const int jr = j0 + w - 1;
        const bool rok = jr <= Ny;
        const bool ok = kok && rok;
        // a wave past the last row only provides E of its (wrapped) row to the wave below
        const int j = rok ? jr : wrap_row(jr);
        const unsigned pad = (unsigned)(j * pz + lane % 31);
        const unsigned o = kok ? (unsigned)(j * pz + k + z0) : pad;  // loads
        const unsigned os = ok ? o : pad;  // stores: only the tile's own cells are written
        for (int q = 0; q < a.nhsheets; ++q) sh |= a.hsheet[q].kp == k;
        for (int q = 0; q < a.nsheets; ++q) sh |= a.sheet[q].kp == k;
        sh = __ballot(sh) != 0;
        PsiBase<NP> psb;
#pragma unroll
        for (int n = 0; n < NP; ++n) psb.b[n] = a.psi[n][0];
        const double khy = kload(a.ikdH[1] + j), khz = gld(a.ikdH[2] + k);
        const double key = kload(a.ikdE[1] + j), kez = gld(a.ikdE[2] + k);
        // E^(n+1) of the row at p + 1 and B^n at p (real: at the lane's cell, else at the
        // padding position 0), the cell's permittivity
        struct Stage {
            RW en, b0;
            double eps;
        };
        auto load = [&](int p, bool real, Stage &st) {
            const int pn = p == Nx ? 1 : p + 1;
            const unsigned m = real ? 0xffffffffu : 0u;
            st.en = ldp3<T, RES>(a.Eo, a.Eolo, (size_t)pn * sx, (o & m) * SZ);
            st.b0 = ldp3<T, RES>(a.Bo, a.Bolo, (size_t)p * sx, (o & m) * SZ);
            if (a.epsf) st.eps = ldp(a.epsf + (size_t)p * sx, (o & m) * 8u);
        };
        // the accumulators of the cell at plane p, or (real false) element 0 of each block
        auto load_psi = [&](int p, bool real, PsiRegs<NP> &pr) {
            const size_t pq = real ? psi_at((size_t)p * sx + o) : 0;
#pragma unroll
            for (int n = 0; n < NP; ++n)
#pragma unroll
                for (int r = 0; r < 3; ++r) pr.v[n][r] = load_once(a.psi[n][0] + pq + 64 * r);
        };
        // B^(n+1) of the row at plane p from E^(n+1) at p (e, decoded), at p + 1 (n) and the
        // E rows above and to the right (LDS buffer be); CPML H auxiliaries stored with `store`
        auto bstep = [&](int p, int be, const C3 &e, const C3 &n, const RW &b0w, bool store, bool cp, double kxh,
                         T h[3], unsigned &hw, double b[3]) {
            double zx = lane_next((double)e.x), zy = lane_next((double)e.y);
            if (lane == lr) {
                zx = ecol[be].a[w - 1];
                zy = ecol[be].b[w - 1];
            }
            const double d[6] = {zy - e.y, zx - e.x, erow[be][w + 1].b[lane] - e.z, erow[be][w + 1].a[lane] - e.x,
                                 (double)n.z - e.z, (double)n.y - e.y};
            b_cell_k<T, CPML, RES>(a, p, j, k, cell_of(b0w), d, store, h, hw, b, cp, sh, ok, kxh, khy, khz);
        };
        Stage cur;
        C3 e = cell_of(ldp3<T, RES>(a.Eo, a.Eolo, (size_t)pm * sx, o * SZ));  // E^(n+1) at plane pm
        load(pm, true, cur);
        erow[0][w].a[lane] = e.x;
        erow[0][w].b[lane] = e.z;
        lds_barrier();
        // the segment's x - 1 plane: B^(n+1) recomputed there (not stored), for the first plane's E
        double bp1, bp2;  // By, Bz of B^(n+1) at plane p - 1
        {
            const bool cp = CPML && __ballot(ok && in_cpml(a, false, pm, j, k)) != 0;
            const C3 n = cell_of(cur.en);
            T h[3];
            unsigned hw;
            double b[3];
            bstep(pm, 0, e, n, cur.b0, false, cp, kload(a.ikdH[0] + pm), h, hw, b);
            bp1 = b[1];
            bp2 = b[2];
            e = n;
            erow[1][w].a[lane] = e.x;
            erow[1][w].b[lane] = e.z;
        }
        load(i0, true, cur);
        PsiRegs<NP> pr;
        load_psi(i0, true, pr);
        lds_barrier();
        for (int p = i0; p <= i1; ++p) {
            const bool last = p == i1;
            const int be = (p - i0 + 1) & 1, bb = (p - i0) & 1;  // parities of the E and B buffers
            const size_t bp = (size_t)p * sx;
            const unsigned ob = os * SZ;
            const double kxh = kload(a.ikdH[0] + p), kxe = kload(a.ikdE[0] + p);
            const bool cp = CPML && __ballot(ok && (in_cpml(a, false, p, j, k) || in_cpml(a, true, p, j, k))) != 0;
            const C3 n = cell_of(cur.en);  // E^(n+1) at p + 1
            const double ee = cur.eps;
            T bh[3];
            unsigned bw;
            double b[3];
            bstep(p, be, e, n, cur.b0, true, cp, kxh, bh, bw, b);
            // the next plane's words, once this plane's are consumed
            load(last ? p : p + 1, !last, cur);
            stp_once(a.Bn[0] + bp, ob, bh[0]);
            stp_once(a.Bn[1] + bp, ob, bh[1]);
            stp_once(a.Bn[2] + bp, ob, bh[2]);
            if (RES) stp_once(a.Bnlo + bp, ob, bw);
            brow[bb][w].a[lane] = b[0];
            brow[bb][w].b[lane] = b[2];
            erow[be ^ 1][w].a[lane] = n.x;
            erow[be ^ 1][w].b[lane] = n.z;
            lds_barrier();
            // B^(n+1) of the y - 1 row (the wave below) and of the z - 1 cell (previous lane,
            // or the column for lane 0)
            double bzm0 = lane_prev(b[0]), bzm1 = lane_prev(b[1]);
            if (lane == 0) {
                bzm0 = bcol[bb].a[w - 1];
                bzm1 = bcol[bb].b[w - 1];
            }
            const double bym0 = brow[bb][w - 1].a[lane], bym2 = brow[bb][w - 1].b[lane];
            // E^(n+2) at the cell
            const double dBz_y = b[2] - bym2, dBx_y = b[0] - bym0;
            const double dBy_z = b[1] - bzm1, dBx_z = b[0] - bzm0;
            const double dBz_x = b[2] - bp2, dBy_x = b[1] - bp1;
            // the plasma OptField and the plasma current run before the E update, on E of this
            // step (the fused E kernel's order; each cell's owner updates them in place)
            double sg = 0.0, jpv[3] = {0.0, 0.0, 0.0};
            if (PL) {
                const size_t cl = bp + o, cs = bp + os;
                double r = gld(a.rho + cl);
                if (a.plasma) {
                    sg = a.pw * plasma_density_cell(e.x, e.y, e.z, r, a.pn, a.pm, a.pk - 1);
                    gst(a.rho + cs, r);
                    gst(a.sigma + cs, sg);
                }
                if (a.pcur) {
                    const double rt = r + a.rho_bg, ev[3] = {e.x, e.y, e.z};
                    for (int q = 0; q < 3; ++q) {
                        jpv[q] = plasma_current_comp(gld(a.jp[q] + cl), ev[q], rt, a.gdtn, a.gdtd, a.emdt);
                        gst(a.jp[q] + cs, jpv[q]);
                    }
                }
            }
            double jx, jy, jz;
            pass_j<T, CPML, RES>(a, p, j, k, dBz_y, dBy_z, dBx_z, dBz_x, dBy_x, dBx_y, jx, jy, jz, cp, sh, ok,
                                 PL ? jpv : nullptr);
            double idenom = a.idenom, numer = a.numer;
            if (a.epsf || (PL && a.plasma)) {  // cell_eps with the prefetched permittivity and Sigma dt / 2
                const double e0 = a.epsf ? ee : a.eps, sgh = 0.5 * sg * a.dt;
                const double denom = NONLIN ? e0 + a.sumChi0 - a.sumXi0 + sgh : e0 + (a.sumChi0 - a.sumXi0) + sgh;
                numer = e0 - a.sumXi0 - sgh;
                idenom = 1.0 / denom;
            }
            double nex = e.x, ney = e.y, nez = e.z;
            unsigned lf = 0;
            plrc_e_update_pp<NP, NONLIN, true>(nex, ney, nez, dBz_y, dBy_z, dBx_z, dBz_x, dBy_x, dBx_y, kxe, key,
                                               kez, jx, jy, jz, pr, psb, bp + os, true, a.pc, a.dt, idenom, numer,
                                               a.chi, a.ncap, &lf);
            lf |= finite_flag(nex, ney, nez);
            fl |= ok ? lf : 0u;
            unsigned ew = 0;
            const T hx = put<T, RES>(nex, ew, 0), hy = put<T, RES>(ney, ew, 1), hz = put<T, RES>(nez, ew, 2);
            stp_once(a.En[0] + bp, ob, hx);
            stp_once(a.En[1] + bp, ob, hy);
            stp_once(a.En[2] + bp, ob, hz);
            if (RES) stp_once(a.Enlo + bp, ob, ew);
            load_psi(last ? p : p + 1, !last, pr);
            bp1 = b[1];
            bp2 = b[2];
            e = n;
        }
    }
    report(a.status, fl);
}

// Block -> tile: each XCD (blocks are dealt round-robin over the 8 XCDs) takes a contiguous
// share of the tile list, so the tiles whose rows border each other run at the same time on
// the same L2
__device__ __forceinline__ unsigned xcd_share(unsigned lin, unsigned total)
{
    const unsigned per = total / 8u, xcd = lin & 7u, idx = lin >> 3;
    return (lin < per * 8u) ? xcd * per + idx : lin;
}

// one grid: tiles in the order z tiles fastest, then tile rows, then x segments
template <typename T, int NP, bool NONLIN, bool CPML, bool RES, int R, bool PL>
__global__ __launch_bounds__(64 * (R + 1)) void k_march(const PassArgs<T> *__restrict__ pa, int nbz, int nby,
                                                        int nseg)
{
    unsigned t = xcd_share(blockIdx.x, (unsigned)nbz * nby * nseg);
    const int tz = (int)(t % (unsigned)nbz);
    t /= (unsigned)nbz;
    march_tile<T, NP, NONLIN, CPML, RES, R, PL>(*pa, tz, (int)(t % (unsigned)nby), (int)(t / (unsigned)nby));
}

// Independent runs in one launch: run r owns the tiles runs[r].first .. runs[r+1].first - 1
// of the list, in its own grid's order; its argument block is pa[r].
template <typename T, int NP, bool NONLIN, bool CPML, bool RES, int R, bool PL>
__global__ __launch_bounds__(64 * (R + 1)) void k_march_batch(const PassArgs<T> *__restrict__ pa,
                                                              const BatchRun *__restrict__ runs, int nruns,
                                                              int total)
{
    const int t = (int)xcd_share(blockIdx.x, (unsigned)total);
    int lo = 0, hi = nruns - 1;  // last run with first <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (runs[mid].first <= t) lo = mid; else hi = mid - 1;
    }
    const BatchRun &B = runs[lo];
    const int local = t - B.first, tz = local % B.ntz, rest = local / B.ntz;
    march_tile<T, NP, NONLIN, CPML, RES, R, PL>(pa[lo], tz, rest % B.nty, rest / B.nty);
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

// x march: 64 z cells x R rows x `seg` planes per block, R + 1 waves
template <typename T, bool RES, int R, bool PL>
static void launch_march_r(const PassArgs<T> &a, const PassArgs<T> *dev, int np, bool nonlin, hipStream_t s)
{
    const int nbz = (a.g.n[2] + 63) / 64, nby = (a.g.n[1] + R - 1) / R, nseg = (a.g.n[0] + a.seg - 1) / a.seg;
    dim3 grd(nbz * nby * nseg), blk(64, R + 1);
    const bool cp = a.nefaces > 0 || a.nhfaces > 0;
#define MPX_MS(NPV, NL, CP) \
    hipLaunchKernelGGL((k_march<T, NPV, NL, CP, RES, R, PL>), grd, blk, 0, s, dev, nbz, nby, nseg)
#define MPX_MS_NP(NL, CP)             \
    switch (np) {                     \
    case 0: MPX_MS(0, NL, CP); break; \
    case 1: MPX_MS(1, NL, CP); break; \
    case 2: MPX_MS(2, NL, CP); break; \
    default: MPX_MS(3, NL, CP); break; \
    }
    if (nonlin) {
        if (cp) { MPX_MS_NP(true, true) } else { MPX_MS_NP(true, false) }
    } else {
        if (cp) { MPX_MS_NP(false, true) } else { MPX_MS_NP(false, false) }
    }
#undef MPX_MS_NP
#undef MPX_MS
}

// Tile height when none is set: 7 rows (8-wave blocks, two per CU) when the variant fits four
// waves per SIMD (<= 128 VGPRs), else 3 rows (4-wave blocks, three per CU at three waves per
// SIMD; the two- and three-pole float32 + residual variants, the plasma variants).  Read once
// per variant from the compiled kernel.
template <typename T, bool RES, bool PL>
static int march_auto_rows(int np, bool nonlin, bool cp)
{
    static int cache[4][2][2] = {};
    int &r = cache[np < 0 ? 0 : (np > 3 ? 3 : np)][nonlin][cp];
    if (!r) {
        const void *f = nullptr;
#define MPX_MF(NPV, NL, CP) f = (const void *)&k_march<T, NPV, NL, CP, RES, 7, PL>
#define MPX_MF_NP(NL, CP)             \
    switch (np) {                     \
    case 0: MPX_MF(0, NL, CP); break; \
    case 1: MPX_MF(1, NL, CP); break; \
    case 2: MPX_MF(2, NL, CP); break; \
    default: MPX_MF(3, NL, CP); break; \
    }
        if (nonlin) {
            if (cp) { MPX_MF_NP(true, true) } else { MPX_MF_NP(true, false) }
        } else {
            if (cp) { MPX_MF_NP(false, true) } else { MPX_MF_NP(false, false) }
        }
#undef MPX_MF_NP
#undef MPX_MF
        hipFuncAttributes at{};
        r = (hipFuncGetAttributes(&at, f) == hipSuccess && at.numRegs > 128) ? 3 : 7;
    }
    return r;
}

template <typename T, bool RES, bool PL>
static void launch_march_p(const PassArgs<T> &a, const PassArgs<T> *dev, int np, bool nonlin, hipStream_t s)
{
    const int rows = a.rows ? a.rows : march_auto_rows<T, RES, PL>(np, nonlin, a.nefaces > 0 || a.nhfaces > 0);
    if (rows == 3)
        launch_march_r<T, RES, 3, PL>(a, dev, np, nonlin, s);
    else
        launch_march_r<T, RES, 7, PL>(a, dev, np, nonlin, s);
}

// the plasma updates are a separate instantiation: they raise the register count
template <typename T, bool RES>
static void launch_march_t(const PassArgs<T> &a, const PassArgs<T> *dev, int np, bool nonlin, hipStream_t s)
{
    if (a.plasma || a.pcur)
        launch_march_p<T, RES, true>(a, dev, np, nonlin, s);
    else
        launch_march_p<T, RES, false>(a, dev, np, nonlin, s);
}

template <typename T>
void launch_pass(const PassArgs<T> &a, const PassArgs<T> *dev, int np, bool nonlin, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        if (a.Eolo) {
            if (a.kernel == 1)
                launch_march_t<T, true>(a, dev, np, nonlin, s);
            else
                launch_pass_t<T, true>(a, dev, np, nonlin, s);
            return;
        }
    }
    if (a.kernel == 1)
        launch_march_t<T, false>(a, dev, np, nonlin, s);
    else
        launch_pass_t<T, false>(a, dev, np, nonlin, s);
}
template void launch_pass<double>(const PassArgs<double> &, const PassArgs<double> *, int, bool, hipStream_t);
template void launch_pass<float>(const PassArgs<float> &, const PassArgs<float> *, int, bool, hipStream_t);

// Batched x march (no residual words, no plasma): tile height of the one-grid kernel of the
// same variant unless `rows` (3 or 7) is given
template <typename T>
int march_batch_rows(int rows, int np, bool nonlin, bool cp)
{
    return rows ? rows : march_auto_rows<T, false, false>(np, nonlin, cp);
}

template <typename T, int R>
static void launch_march_batch_r(const PassArgs<T> *table, const BatchRun *runs, int nruns, int total, int np,
                                 bool nonlin, bool cp, hipStream_t s)
{
    dim3 grd(total), blk(64, R + 1);
#define MPX_MB(NPV, NL, CP)                                                                                      \
    hipLaunchKernelGGL((k_march_batch<T, NPV, NL, CP, false, R, false>), grd, blk, 0, s, table, runs, nruns, \
                       total)
#define MPX_MB_NP(NL, CP)             \
    switch (np) {                     \
    case 0: MPX_MB(0, NL, CP); break; \
    case 1: MPX_MB(1, NL, CP); break; \
    case 2: MPX_MB(2, NL, CP); break; \
    default: MPX_MB(3, NL, CP); break; \
    }
    if (nonlin) {
        if (cp) { MPX_MB_NP(true, true) } else { MPX_MB_NP(true, false) }
    } else {
        if (cp) { MPX_MB_NP(false, true) } else { MPX_MB_NP(false, false) }
    }
#undef MPX_MB_NP
#undef MPX_MB
}

template <typename T>
void launch_march_batch(const PassArgs<T> *table, const BatchRun *runs, int nruns, int total, int rows, int np,
                        bool nonlin, bool cp, hipStream_t s)
{
    if (rows == 3)
        launch_march_batch_r<T, 3>(table, runs, nruns, total, np, nonlin, cp, s);
    else
        launch_march_batch_r<T, 7>(table, runs, nruns, total, np, nonlin, cp, s);
}
template int march_batch_rows<double>(int, int, bool, bool);
template int march_batch_rows<float>(int, int, bool, bool);
template void launch_march_batch<double>(const PassArgs<double> *, const BatchRun *, int, int, int, int, bool, bool,
                                         hipStream_t);
template void launch_march_batch<float>(const PassArgs<float> *, const BatchRun *, int, int, int, int, bool, bool,
                                        hipStream_t);

}  // namespace mpx
