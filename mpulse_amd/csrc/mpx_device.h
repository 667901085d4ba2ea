// Device helpers shared by the gfx950 kernel files: global-address loads and stores,
// nontemporal streams, DPP lane shuffles, the PLRC + Kerr E update core, the CPML face
// update, the float32 + residual storage form, tile maps.
#pragma once
#include <hip/hip_runtime.h>

#include "mpx_types.h"

namespace mpx {

// ------------------------------------------------------------------------------------
// PLRC + Kerr E update core, shared by both paths:
// FDTD_PLRCNonlinCore::plrcStepD src/fdtd_plrc.cpp:304-427,
// FDTD_PLRCLinCore::plrcStepD    src/fdtd_plrc.cpp:94-261.
// d* are the B differences in the reference's order, kd* the kappa*delta factors.
// ------------------------------------------------------------------------------------
// Every grid, table and layer pointer of the solver points into hipMalloc'd memory.  Kernels
// that read their argument block from device memory see those pointers as generic ones, and
// generic (flat) accesses also count against the scalar-load counter, so each wait for an
// argument load would wait for the outstanding grid loads too.  gld / gst access through the
// global address space instead.
#define MPX_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T gld(const T *p)
{
    return *(const MPX_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ void gst(T *p, T v)
{
    *(MPX_GLOBAL T *)p = v;
}

// Words a launch touches exactly once (the pole accumulators, a cell's own E in the E kernel
// and B in the B kernel) move with nontemporal loads and stores: they then do not evict from L2
// the neighbour planes that other threads of the launch re-read.
typedef double double2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 load_once(const double2 *p)
{
    const double2v v = __builtin_nontemporal_load((const MPX_GLOBAL double2v *)p);
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void store_once(double2 *p, double2 v)
{
    double2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (MPX_GLOBAL double2v *)p);
}
template <typename T>
__device__ __forceinline__ T load_once(const T *p)
{
    return __builtin_nontemporal_load((const MPX_GLOBAL T *)p);
}
template <typename T>
__device__ __forceinline__ void store_once(T *p, T v)
{
    __builtin_nontemporal_store(v, (MPX_GLOBAL T *)p);
}

// The value a neighbouring lane of the wavefront holds, by DPP: lane_prev gives lane l the
// value of lane l - 1 (wave_shr:1), lane_next that of lane l + 1 (wave_shl:1).  Lanes of a
// fused-kernel wave hold consecutive z cells of one row, so the z neighbour of a cell is in
// the next lane except at the ends of the wave and of the row, where it is loaded.
__device__ __forceinline__ float lane_prev(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_next(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_prev(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_next(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Storage forms of a field word: T (float64 or float32) on its own, or (RES, T = float) the
// float32 word plus a slot of the cell's residual word (res_encode / res_decode).  V is the
// type a loaded value is carried in: T itself, or the decoded float64 value.
template <typename T, bool RES>
struct Word {
    using V = T;
};
template <>
struct Word<float, true> {
    using V = double;
};
// the residual word of cell c (0 without residual storage: nothing is loaded)
template <bool RES>
__device__ __forceinline__ unsigned lo_ld(const unsigned *lo, size_t c)
{
    if constexpr (RES) return gld(lo + c);
    else return 0u;
}
template <bool RES>
__device__ __forceinline__ unsigned lo_once(const unsigned *lo, size_t c)
{
    if constexpr (RES) return load_once(lo + c);
    else return 0u;
}
// component q of cell c, read from its word array p and the residual word w
template <typename T, bool RES>
__device__ __forceinline__ typename Word<T, RES>::V rd(const T *p, unsigned w, int q, size_t c)
{
    if constexpr (RES) return res_decode(gld(p + c), w, q);
    else return gld(p + c);
}
template <typename T, bool RES>
__device__ __forceinline__ typename Word<T, RES>::V rd_once(const T *p, unsigned w, int q, size_t c)
{
    if constexpr (RES) return res_decode(load_once(p + c), w, q);
    else return load_once(p + c);
}
// the word stored for x (its residual OR-ed into slot q of w), and the value that word holds
template <typename T, bool RES>
__device__ __forceinline__ T put(double x, unsigned &w, int q)
{
    if constexpr (RES) return res_encode(x, w, q);
    else return (T)x;
}
template <typename T, bool RES>
__device__ __forceinline__ double got(T h, unsigned w, int q)
{
    if constexpr (RES) return res_decode(h, w, q);
    else return (double)h;
}

// quotient x / k, or x * k when k already holds the reciprocal (RECIP)
template <bool RECIP>
__device__ __forceinline__ double qt(double x, double k)
{
    return RECIP ? x * k : x / k;
}

// RECIP: kdx, kdy, kdz and denom hold reciprocals (1/(kappa delta), 1/denom) and the divisions
// become products; otherwise the reference's quotients are formed.
template <int NP, bool NONLIN, bool RECIP, typename PIN, typename POUT>
__device__ __forceinline__ void plrc_e_update_pp(double &ex, double &ey, double &ez, double dBz_y, double dBy_z,
                                                 double dBx_z, double dBz_x, double dBy_x, double dBx_y,
                                                 double kdx, double kdy, double kdz, double jx, double jy,
                                                 double jz, const PIN &psi_in, const POUT &psi, size_t c,
                                                 bool store, const PoleConsts &pc, double dt, double denom,
                                                 double numer, double chi, int ncap = kNewtonCap,
                                                 unsigned *flag = nullptr)
{
    double Psix = 0, Psiy = 0, Psiz = 0;
    const size_t pq = psi_at(c);
#pragma unroll
    for (int n = 0; n < NP; ++n) {
        double2 px = psi_in[n][0][pq], py = psi_in[n][1][pq], pz = psi_in[n][2][pq];
        const double2 D = pc.dd[n], X = pc.dxi0[n], C = pc.crec[n];
        double pxr = D.x * ex + px.x, pxi = D.y * ex + px.y;
        double pyr = D.x * ey + py.x, pyi = D.y * ey + py.y;
        double pzr = D.x * ez + pz.x, pzi = D.y * ez + pz.y;
        Psix += pxr;
        Psiy += pyr;
        Psiz += pzr;
        if (store) {
            store_once(&psi[n][0][pq], make_double2(X.x * ex + (C.x * pxr - C.y * pxi), X.y * ex + (C.x * pxi + C.y * pxr)));
            store_once(&psi[n][1][pq], make_double2(X.x * ey + (C.x * pyr - C.y * pyi), X.y * ey + (C.x * pyi + C.y * pyr)));
            store_once(&psi[n][2][pq], make_double2(X.x * ez + (C.x * pzr - C.y * pzi), X.y * ez + (C.x * pzi + C.y * pzr)));
        }
    }
    if (NONLIN) {
        double E2 = ex * ex + ey * ey + ez * ez;
        double f = numer + chi * E2;
        double cx = qt<RECIP>(f * ex + dt * ((qt<RECIP>(dBz_y, kdy) - qt<RECIP>(dBy_z, kdz)) + jx) + Psix, denom);
        double cy = qt<RECIP>(f * ey + dt * ((qt<RECIP>(dBx_z, kdz) - qt<RECIP>(dBz_x, kdx)) + jy) + Psiy, denom);
        double cz = qt<RECIP>(f * ez + dt * ((qt<RECIP>(dBy_x, kdx) - qt<RECIP>(dBx_y, kdy)) + jz) + Psiz, denom);
        double E = sqrt(E2);
        double A = qt<RECIP>(chi, denom);
        double Cm = sqrt(cx * cx + cy * cy + cz * cz);
        // a non-finite input makes Cm NaN: the reference then stores E = 0 (its C > 0 test
        // fails) and the accumulators keep the NaN; reproduced, and reported
        if (flag && !isfinite(Cm)) *flag |= 2u;
        if (Cm > 0) {
            // the reference's stopping rule |E - Eold| <= 1e-9; the cap only guards a cell that
            // could never meet it (non-finite data, or fields far from O(1))
            double Eold;
            int it = 0;
            do {
                Eold = E;
                double Q = E * E;
                E = (2 * A * Q * E + Cm) / (3 * A * Q + 1);
            } while (fabs(E - Eold) > 1e-9 && ++it < ncap);
            if (flag && it >= ncap) *flag |= 1u;  // stopped by the cap, not by the rule
            if (RECIP) {
                const double sc = E / Cm;
                ex = sc * cx;
                ey = sc * cy;
                ez = sc * cz;
            } else {
                ex = E * cx / Cm;
                ey = E * cy / Cm;
                ez = E * cz / Cm;
            }
        } else {
            ex = 0.0;
            ey = 0.0;
            ez = 0.0;
        }
    } else {
        double exn = qt<RECIP>(numer * ex + (dt * ((qt<RECIP>(dBz_y, kdy) - qt<RECIP>(dBy_z, kdz)) + jx) + Psix), denom);
        double eyn = qt<RECIP>(numer * ey + (dt * ((qt<RECIP>(dBx_z, kdz) - qt<RECIP>(dBz_x, kdx)) + jy) + Psiy), denom);
        double ezn = qt<RECIP>(numer * ez + (dt * ((qt<RECIP>(dBy_x, kdx) - qt<RECIP>(dBx_y, kdy)) + jz) + Psiz), denom);
        ex = exn;
        ey = eyn;
        ez = ezn;
    }
}

// Denominator and numerator of the E update at cell c: the uniform ones, or those of a
// per-cell permittivity formed with the pole sums in the order the uniform ones are formed.
// sigma = Sigma dt / 2 of the plasma resistivity (src/fdtd_plrc.cpp:356-359).
template <bool NONLIN, typename A>
__device__ __forceinline__ void cell_eps(const A &a, size_t c, double sigma, double &denom, double &numer)
{
    if (a.epsf || a.plasma) {
        const double e = a.epsf ? gld(a.epsf + c) : a.eps;
        denom = NONLIN ? e + a.sumChi0 - a.sumXi0 + sigma : e + (a.sumChi0 - a.sumXi0) + sigma;
        numer = e - a.sumXi0 - sigma;
    } else {
        denom = a.denom;
        numer = a.numer;
    }
}

// PlasmaDensity::stepScheme at one cell (src/plasmadensity.cpp:40-53): rho <- (rho (1 + n I)
// + m I^K) / (1 - n I) with I = |E|^2; returns I^(K-1), which Sigma receives times w.
__device__ __forceinline__ double plasma_density_cell(double ex, double ey, double ez, double &rho, double n,
                                                      double m, int Km)
{
    const double I = ex * ex + ey * ey + ez * ez;
    double IKm = 1.0;
    for (int q = 0; q < Km; ++q) IKm *= I;  // ipow, src/util.h:67-73
    const double IK = I * IKm;
    rho = (rho * (1 + n * I) + m * IK) / (1 - n * I);
    return IKm;
}

// PlasmaCurrent::stepScheme at one cell (src/plasmacurrent.cpp:63-71)
__device__ __forceinline__ double plasma_current_comp(double j, double e, double rho_t, double gdtn, double gdtd,
                                                      double emdt)
{
    return (j * gdtn - emdt * e * rho_t) / gdtd;
}

template <int NP, bool NONLIN, bool RECIP>
__device__ __forceinline__ void plrc_e_update(double &ex, double &ey, double &ez, double dBz_y, double dBy_z,
                                              double dBx_z, double dBz_x, double dBy_x, double dBx_y,
                                              double kdx, double kdy, double kdz, double jx, double jy, double jz,
                                              double2 *const psi[3][3], size_t c, const PoleConsts &pc,
                                              double dt, double denom, double numer, double chi, int ncap,
                                              unsigned *flag)
{
    plrc_e_update_pp<NP, NONLIN, RECIP>(ex, ey, ez, dBz_y, dBy_z, dBx_z, dBz_x, dBy_x, dBx_y, kdx, kdy, kdz, jx, jy, jz, psi,
                                 psi, c, true, pc, dt, denom, numer, chi, ncap, flag);
}

// Device status of a solver (MPX_ERR_NUMERIC): bit 0 a Kerr Newton solve stopped at the
// iteration cap, bit 1 a non-finite value in an E update (its inputs or its result).  flag holds this lane's bits; one lane of the wave ORs
// the wave's bits into *status (a vector atomic), so a wave costs at most one atomic.
__device__ __forceinline__ unsigned finite_flag(double x, double y, double z)
{
    return (isfinite(x) && isfinite(y) && isfinite(z)) ? 0u : 2u;
}
__device__ __forceinline__ void report(unsigned *status, unsigned flag)
{
    const unsigned long long b1 = __ballot((flag & 1u) != 0), b2 = __ballot((flag & 2u) != 0);
    const unsigned long long any = b1 | b2;
    if (any && status && (int)__lane_id() == __ffsll((unsigned long long)any) - 1)
        atomicOr(status, (b1 ? 1u : 0u) | (b2 ? 2u : 0u));
}


// ------------------------------------------------------------------------------------
// fused path
// ------------------------------------------------------------------------------------

// accumulators of the active poles held in registers, indexable like the psi pointer table
// (psi[n][component][cell]) so that the shared E update reads them in place
template <int NP>
struct PsiRegs {
    double2 v[NP > 0 ? NP : 1][3];
    struct Comp {
        double2 x;
        __device__ double2 operator[](size_t) const { return x; }
    };
    struct Pole {
        const double2 *p;
        __device__ Comp operator[](int q) const { return Comp{p[q]}; }
    };
    __device__ Pole operator[](int n) const { return Pole{v[n]}; }
};


// CPML auxiliaries of one face updated inside the fused kernels.  The face map is
// x: (psi0 -> Jy from dBz, psi1 -> Jz from dBy), y: (psi0 -> Jz from dBx, psi1 -> Jx from dBz),
// z: (psi0 -> Jx from dBy, psi1 -> Jy from dBx) (src/cpml_border.cpp:357-386, 456-485).
// E side: CPMLBorderECurrent::stepScheme, src/cpml_border.cpp:398-425 (psi0 gets -c d1/dx);
// H side: CPMLBorderHCurrent::stepScheme, src/cpml_border.cpp:499-527 (psi0 gets +c d1/dx).
template <typename T, bool HSIDE, bool RES = false>
__device__ __forceinline__ void cpml_face(const CpmlFace &f, int coord, int l, double d1, double d0, double &v0,
                                          double &v1)
{
    int q = coord - f.lo[f.dim];
    T *p0 = (T *)f.psi0, *p1 = (T *)f.psi1;
    double bq = gld(f.b + q), cq = gld(f.c + q);
    // ordinary loads: a face row covers only part of a cache line (z faces: 16 cells), and the
    // rest of the line is read by the neighbouring row
    const unsigned w = lo_ld<RES>(f.rw, l);
    const double o0 = (double)rd<T, RES>(p0, w, 0, l), o1 = (double)rd<T, RES>(p1, w, 1, l);
    double n0, n1;
    if (HSIDE) {
        n0 = bq * o0 + cq * d1 * f.idxn;
        n1 = bq * o1 - cq * d0 * f.idxn;
    } else {
        n0 = bq * o0 - cq * d1 * f.idxn;
        n1 = bq * o1 + cq * d0 * f.idxn;
    }
    // the current sums see the stored values
    unsigned wn = 0;
    const T h0 = put<T, RES>(n0, wn, 0), h1 = put<T, RES>(n1, wn, 1);
    gst(p0 + l, h0);
    gst(p1 + l, h1);
    if (RES) gst(f.rw + l, wn);
    v0 = got<T, RES>(h0, wn, 0);
    v1 = got<T, RES>(h1, wn, 1);
}

// 32-bit layer index of cell (i, j, k) in a face box
__device__ __forceinline__ int layer_index(const CpmlFace &f, int i, int j, int k)
{
    return ((i - f.lo[0]) * f.n1 + (j - f.lo[1])) * f.n2 + (k - f.lo[2]);
}

// add the current-sheet layers that contain cell (i, j, k), in list order
template <typename T>
__device__ __forceinline__ void add_sheets(const StepArgs<T> &a, int i, int j, int k, double &x, double &y,
                                           double &z)
{
    for (int s = 0; s < a.nsheets; ++s)
        if (k == a.sheet[s].kp) {
            int q = i * a.g.py + j;
            x += gld(a.sheet[s].J[0] + q);
            y += gld(a.sheet[s].J[1] + q);
            z += gld(a.sheet[s].J[2] + q);
        }
}

// Map the linear block id to a tile so that each XCD (blocks are dealt round-robin over
// the 8 XCDs) walks a contiguous slab of x planes and keeps the x-neighbour planes in its L2.
__device__ __forceinline__ void xcd_tile(int nbx, int nby, int nbz, int &bx, int &by, int &bz)
{
    unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    unsigned total = (unsigned)nbx * nby * nbz;
    unsigned per = total / 8u;
    unsigned xcd = lin & 7u, idx = lin >> 3;
    unsigned t = (lin < per * 8u) ? xcd * per + idx : lin;
    bx = t % nbx;
    t /= nbx;
    by = t % nby;
    bz = t / nby;
}

// Banded order: the tile rows are cut into bands of R rows; inside a band the z tiles run
// fastest, then the rows of the band, then the x chunks, so the x-neighbour plane of a tile
// was read a band-plane (nbx * R tiles) earlier and is still in the XCD's L2.  Each XCD
// takes a contiguous share of the order as in xcd_tile.  With rev the order runs from its end,
// which the B-side kernel uses to begin on the tiles the E-side kernel finished last.
__device__ __forceinline__ void band_tile(int nbx, int nby, int nbz, int R, int rev, int &bx, int &by, int &bz)
{
    unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    unsigned total = (unsigned)nbx * nby * nbz;
    unsigned per = total / 8u;
    unsigned xcd = lin & 7u, idx = lin >> 3;
    unsigned t = (lin < per * 8u) ? xcd * per + idx : lin;
    if (rev) t = total - 1u - t;
    const unsigned full = (unsigned)nbx * R * nbz;  // tiles of a full band; only the last is shorter
    const unsigned b = t / full;
    unsigned rem = t - b * full;
    const unsigned rows = min((unsigned)R, (unsigned)nby - b * R);
    bx = rem % nbx;
    rem /= nbx;
    by = b * R + rem % rows;
    bz = rem / rows;
}

__device__ __forceinline__ void fused_tile_of(int order, int band, int rev, int nbx, int nby, int nbz, int &bx,
                                              int &by, int &bz)
{
    if (order == 2)
        band_tile(nbx, nby, nbz, band, rev, bx, by, bz);
    else if (order == 1)
        xcd_tile(nbz, nbx, nby, bz, bx, by);
    else
        xcd_tile(nbx, nby, nbz, bx, by, bz);
}

// Whether coordinate coord of dimension dim lies in the layer of the high-side ([0]) and
// low-side ([1]) CPML face of that dimension.  The face index itself stays uniform across
// the wave (fbd[dim][s]), so the face descriptors are read with scalar loads.
__device__ __forceinline__ void in_faces(const int fbd[3][2], const CpmlFace *faces, int dim, int coord, bool out[2])
{
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        int f = fbd[dim][s];
        out[s] = f >= 0 && coord >= faces[f].lo[dim] && coord <= faces[f].hi[dim];
    }
}

// (j, k) of this thread in a tile of the fused kernels: 64 z cells x 4 y rows, or (LIN) 256 consecutive cells of
// the y-z plane taken row after row, which keeps every lane busy on grids whose z extent is
// not a multiple of 64.  lo = first index of the range in y and z.
template <bool LIN>
__device__ __forceinline__ void tile_jk(const GridDims &g, const int lo[3], int tbx, int tby, int &j, int &k)
{
    if (LIN) {
        const int nk = g.n[2] - lo[2] + 1;
        const int lin = tbx * 256 + threadIdx.y * 64 + threadIdx.x;
        const int q = lin / nk;
        j = q + lo[1];
        k = lin - q * nk + lo[2];
    } else {
        k = tbx * 64 + threadIdx.x + lo[2];
        j = tby * 4 + threadIdx.y + lo[1];
    }
}

// add the current-sheet layers of list `sh` that contain cell (i, j, k), in list order
__device__ __forceinline__ void add_sheet_planes(const Sheet *sh, int n, int py, int i, int j, int k, double &x,
                                                 double &y, double &z)
{
    for (int s = 0; s < n; ++s)
        if (k == sh[s].kp) {
            int q = i * py + j;
            x += gld(sh[s].J[0] + q);
            y += gld(sh[s].J[1] + q);
            z += gld(sh[s].J[2] + q);
        }
}

}  // namespace mpx
