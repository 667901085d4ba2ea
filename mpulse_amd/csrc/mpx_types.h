// Plain-old-data descriptors shared by the host solver and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "mpx_math.h"

namespace mpx {

constexpr int kMaxSheets = 8;   // incident-source current sheets per side
constexpr int kMaxFaces = 6;    // CPML face currents per side
constexpr int kNewtonCap = 200; // iteration bound of the device Kerr Newton loops

// Storage of one (nx+2) x (ny+2) x (nz+2) grid in HBM: z fastest, the z row padded to pz
// elements and shifted by z0 so that interior cell k = 1 starts an aligned segment.
struct GridDims {
    int n[3];      // interior cells per dimension
    int py;        // ny + 2
    int pz;        // padded z row
    int z0;        // element offset of k = 0 inside a row
    long long sx;  // py * pz
    __host__ __device__ __forceinline__ size_t at(int i, int j, int k) const
    {
        return (size_t)i * (size_t)sx + (size_t)j * (size_t)pz + (size_t)(k + z0);
    }
};

// Layout of the PLRC accumulators: the three components of one pole's accumulator are stored
// in blocks of 64 cells, [chunk][component][64] double2, so that the 64 z-consecutive cells a
// wavefront updates form one contiguous 3 KB run for all three components.  Component q's
// pointer is base + 64 q and cell c (a storage index of GridDims::at) sits at psi_at(c) from
// it.  The shift of 32 starts a chunk at the interior cell k = 1 + 64 t of a row (storage
// offset z0 + 1 = 32) when the row pitch is a multiple of 64.
constexpr size_t PSI_SHIFT = 32;
__host__ __device__ __forceinline__ size_t psi_at(size_t c)
{
    return ((c + PSI_SHIFT) >> 6) * 192 + ((c + PSI_SHIFT) & 63);
}
// double2 elements of one pole's block over `cells` storage cells
inline size_t psi_block_elems(size_t cells) { return (((cells - 1 + PSI_SHIFT) >> 6) + 1) * 192; }

// Float32 + residual storage (MPX_PREC_FP32_RES): a value x is held as the float32 word
// hi = fl32(x) plus, in a 10-bit slot of a residual word, the two's-complement
// r = round((bits(x) - bits(hi)) / 2^19): the difference of the float64 bit patterns of x and
// of hi widened to float64 (below 2^29 in magnitude, so the low words suffice), clamped to
// [-512, 511]; decoding adds r 2^19 to the bit pattern of hi.  In the binade of hi a unit of r
// is 2^-10 ulp(hi), so the stored value is within 2^-10 ulp(hi) of x, the clamp included
// (2^-33 relative); both sides work on the magnitude bits, so the sign and a hi rounded up
// across a power of two need no special case.  Zero, denormal, infinite and NaN words keep
// r = 0.  The residuals of the components of one cell (E or B: three, a CPML face: its two
// auxiliaries) share a 32-bit word, slot q in bits 10q .. 10q + 9.  Integer operations only;
// both helpers run on the host (get/set of fields) and on the device and give the same bits.
__host__ __device__ __forceinline__ float res_encode(double x, unsigned &w, int q)
{
    const float hi = (float)x;
    const unsigned e = (__builtin_bit_cast(unsigned, hi) >> 23) & 0xffu;
    const unsigned lx = (unsigned)__builtin_bit_cast(unsigned long long, x);
    const unsigned lh = (unsigned)__builtin_bit_cast(unsigned long long, (double)hi);
    int r = ((int)(lx - lh) + (1 << 18)) >> 19;
    r = r > 511 ? 511 : (r < -512 ? -512 : r);
    r = (e == 0u || e == 255u) ? 0 : r;
    w |= ((unsigned)r & 0x3ffu) << (10 * q);
    return hi;
}
__host__ __device__ __forceinline__ double res_decode(float hi, unsigned w, int q)
{
    const int r = (int)(w << (22 - 10 * q)) >> 22;  // sign-extended slot q
    return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, (double)hi) +
                                          ((unsigned long long)(long long)r << 19));
}

// One CPML border current (a face of CPMLBorder{E,H}Current): its layer box, its two live
// auxiliaries, and the b/c coefficient lines indexed by the normal coordinate.
struct CpmlFace {
    int dim;          // normal dimension
    int lo[3], hi[3]; // layer box (inclusive), as Storage::getBorderExtent builds it
    int n1, n2;       // box extents of dims 1 and 2
    void *psi0;       // component psi0 (see the face map of CPMLBorder*Current::initStorage)
    void *psi1;
    unsigned *rw;     // float32 + residual storage: residual words (slot 0 psi0, slot 1 psi1), else null
    int c0, c1;       // current components fed by psi0 / psi1
    int f0, f1;       // field components read for psi0 (f1 difference) / psi1 (f0 difference)
    const double *b;  // [normal - lo[dim]]
    const double *c;
    double dxn;       // grid spacing of the normal dimension
    double idxn;      // 1 / dxn
    __host__ __device__ __forceinline__ size_t at(int i, int j, int k) const
    {
        return ((size_t)(i - lo[0]) * (size_t)n1 + (size_t)(j - lo[1])) * (size_t)n2 + (size_t)(k - lo[2]);
    }
    __host__ __device__ __forceinline__ bool contains(int i, int j, int k) const
    {
        return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
    }
};

// One incident-source current sheet: a z plane kp of the down face, three float64
// component planes of (nx+2) x (ny+2), index i*(ny+2)+j.
struct Sheet {
    int kp;
    double *J[3];
};

// Source-function parameters for the sheet evaluation kernel.
struct PlaneParams {
    int gauss;
    double k[3], E[3], H[3], E_bg[3], H_bg[3];
    double ramp, dt, om, dx, dy, dz, width, offset;
};

// Focused pulse: the transverse incident fields of one carrier-frequency angular spectrum,
// precomputed on the two sampled planes ((nx+2) x (ny+2), index i*(ny+2)+j, (re, im)).
struct FocusedParams {
    const double2 *ex, *bx, *by;
    double zE, zB, T, om0, v, TShift, phase, amp, DT;
};

// GaussInject samples of the block in force (GaussInjectSourceFunc::getEField/getHField,
// src/gaussinject.cpp:203-224): F[(t - blockToff) (nx+2) + i) (ny+2) + j]
struct GaussSheetParams {
    const double *F1, *F2;
    int active, blockToff;
    double amp, sqeps;  // amp, sqrt(eps)
};

struct SheetEval {
    int kind;        // 1 short pulse, 2 plane wave / gauss, 3 focused pulse, 4 GaussInject
    int isH;         // 0: E current (samples H), 1: H current (samples E)
    int off;         // offset of the sampled cell along z (IncidentSource*Current::stepScheme)
    int t1, t2;      // transverse components
    double DX;       // normal spacing
    int time;
    int nx2, ny2;    // nx+2, ny+2
    Sheet sheet;
    PulseParams pulse;
    PlaneParams plane;
    FocusedParams focused;
    GaussSheetParams gauss;
};

// PLRC pole constants (PLRCData, src/fdtd_plrc.h:71-78) of the active poles, precomputed on the
// host with the formulas of src/fdtd_plrc.t:198-216.
struct PoleConsts {
    double2 dd[3];    // dchi0 - dxi0
    double2 dxi0[3];
    double2 crec[3];
};

template <typename T>
struct StepArgs {
    GridDims g;
    int periodic[3];
    T *E[3];
    T *B[3];
    unsigned *Elo, *Blo;  // float32 + residual storage: residual words of E and B, else null
    double2 *psi[3][3];  // [active pole][component]
    const double *kdE[3];  // kappaE * delta per dimension (lines 0..n+1)
    const double *kdH[3];
    const double *ikdE[3];  // their reciprocals (fused kernels)
    const double *ikdH[3];
    double dt;
    double eps, chi;
    double denom, numer;  // eps + sumChi0 - sumXi0 (+sigma), eps - sumXi0 (-sigma), sigma = 0
    double idenom;        // 1 / denom
    const double *epsf;   // per-cell permittivity (full storage) or nullptr for the uniform eps
    double sumChi0, sumXi0;
    // plasma: the PlasmaDensity OptField (plasma) and one PlasmaCurrent (pcur), float64 state
    // over the storage box; pcur_pos = number of current groups (CPML, sheets) listed before it
    int plasma, pcur, pcur_pos, pk;
    double pn, pm, pw;              // dt nu / 2, dt mpa, mpa Wion
    double gdtn, gdtd, emdt, rho_bg;
    double *rho, *jp[3], *sigma;
    PoleConsts pc;
    int nfaces;
    CpmlFace face[kMaxFaces];   // E-side faces for the E kernel, H-side faces for the B kernel
    int face_by_dim[3][2];      // per dimension: [0] high-side face (N/E/U, first in the list), [1] low side, or -1
    int nsheets;
    Sheet sheet[kMaxSheets];
    int sheets_first;           // a source factory precedes the CPML factory in the list
    // reference path only: current totals
    const T *Jt[3];
    int haveJ;
    int blo[3];                 // first updated B index per dimension (fused path)
    unsigned *status;           // device status word (report) and Newton iteration cap
    int ncap;
    int march;                  // fused path: x planes walked by one thread
    int order;                  // fused path: 0 = z tiles fastest in the block order, 1 = x chunks fastest,
                                // 2 = bands of `band` tile rows, walked plane by plane
    int band;
    int rev;                    // fused path: walk the block order backwards (B side)
};

// Single-pass sweep path: "o" = buffers of step n (read), "n" = buffers of step n+1 (written).
template <typename T>
struct SweepArgs {
    GridDims g;
    const T *Eo[3];
    T *En[3];
    const T *Bo[3];
    T *Bn[3];
    const double2 *psio[3][3];
    double2 *psin[3][3];
    const double *kdE[3];
    const double *kdH[3];
    const double *ikdE[3];
    const double *ikdH[3];
    double dt, chi, denom, numer, idenom;
    PoleConsts pc;
    int nefaces, nhfaces;
    CpmlFace eface[kMaxFaces];       // psi0/psi1 = new E auxiliaries
    const void *epsi_old[kMaxFaces][2];
    int efbd[3][2];
    CpmlFace hface[kMaxFaces];       // updated in place (owner cells only)
    int hfbd[3][2];
    int nsheets, nhsheets;
    Sheet sheet[kMaxSheets];
    Sheet hsheet[kMaxSheets];
    int sheets_first;
    int seg;       // x planes per block
    int variant;   // thread-block shape
    unsigned *status;
    int ncap;
};

// Single-pass leapfrog path: one kernel advances B of step n (B^n -> B^(n+1), from E^(n+1))
// and then E of step n + 1 (E^(n+1) -> E^(n+2)).  "o" buffers are read, "n" buffers written
// (E, B, their residual words, the CPML H auxiliaries); the pole accumulators and the CPML E
// auxiliaries, which no halo reads, are updated in place.
template <typename T>
struct PassArgs {
    GridDims g;
    const T *Eo[3];
    T *En[3];
    const T *Bo[3];
    T *Bn[3];
    const unsigned *Eolo, *Bolo;  // residual words (float32 + residual storage), else null
    unsigned *Enlo, *Bnlo;
    double2 *psi[3][3];
    const double *ikdE[3];
    const double *ikdH[3];
    double dt, eps, chi, denom, numer, idenom, sumChi0, sumXi0;
    const double *epsf;           // per-cell permittivity or null
    // plasma OptField and plasma current, as in StepArgs (updated in place by the cell's owner)
    int plasma, pcur, pcur_pos, pk;
    double pn, pm, pw, gdtn, gdtd, emdt, rho_bg;
    double *rho, *jp[3], *sigma;
    PoleConsts pc;
    int nefaces, nhfaces;
    CpmlFace eface[kMaxFaces];    // E auxiliaries, in place
    int efbd[3][2];
    CpmlFace hface[kMaxFaces];    // psi0 / psi1 / rw: the new H auxiliaries
    const void *hold[kMaxFaces][2];   // the old H auxiliaries
    const unsigned *hold_rw[kMaxFaces];
    int hfbd[3][2];
    int nsheets, nhsheets;
    Sheet sheet[kMaxSheets];      // E-side sheets (J of step n + 1)
    Sheet hsheet[kMaxSheets];     // H-side sheets (M of step n)
    int sheets_first;
    int band;                     // rows of a band in the banded tile order
    unsigned *status;             // device status word and Newton iteration cap
    int ncap;
    int kernel;                   // 0: one cell per thread (k_pass), 1: x march (k_march)
    int seg;                      // k_march: x planes per block
    int rows;                     // k_march: tile rows per block (3 or 7; 0: chosen per variant)
    int zper;                     // z periodic (single_periodic); 0: xy-periodic, z ghosts as stored
};

// Adjoint of one step of the fused float64 PLRC update without Lorentz poles.  The adjoints
// of the six curl differences of each kernel are kept per cell in the off-diagonal slots
// [dim][comp] (slot dim*2 + (comp > dim ? comp-1 : comp)): the difference along dim of field
// component comp.
struct AdjArgs {
    GridDims g;
    double dt, chi, eps;      // eps: uniform permittivity when epsf is null
    const double *epsf;
    const double *ikdE[3], *ikdH[3];
    double *lE[3], *lB[3];    // adjoints of E and B
    double *dEb[6];           // adjoints of the E differences of the B update
    double *dBb[6];           // adjoints of the B differences of the E update
    double *epsb, *chib;      // per-cell gradient accumulators
    const double *E0[3], *E1[3];  // forward E before and after the step
    int kobj;                 // objective plane (z index)
    double wobj;
    int nefaces, nhfaces;
    CpmlFace eface[kMaxFaces];
    CpmlFace hface[kMaxFaces];
    int efbd[3][2], hfbd[3][2];
    double *lpsiE[kMaxFaces][2], *lpsiH[kMaxFaces][2];  // adjoints of the CPML auxiliaries
    // Lorentz poles: constants of the active poles and the adjoints of their accumulators,
    // [pole][component] at cell c (storage index)
    int np;
    PoleConsts pc;
    double sumChi0, sumXi0;
    double2 *lpsi[3][3];
};

// one run of a batch: its tile grid and the number of its first tile in the batch's list
struct BatchRun {
    int first;
    int ntz, nty, nseg;
};

}  // namespace mpx
