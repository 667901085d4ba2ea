// Host-side launchers of the gfx950 kernels (defined in mpx_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mpx_types.h"

namespace mpx {

constexpr int kMaxSumViews = kMaxFaces + kMaxSheets;

// one current's layer as seen by sumCurrents: its box and up to three component arrays
struct SumView {
    int lo[3], hi[3];
    int n1, n2;
    int padded;  // comp[] are full-storage grids (GridDims layout) rather than compact boxes
    const double *comp[3];
};
struct SumList {
    int n;
    SumView v[kMaxSumViews];
};

// fdtd-plain (kind 0), fdtd-nl (kind 1), fdtd-disp (kind 2)
struct SimpleArgs {
    GridDims g;
    int kind;
    double *E[3];
    double *B[3];
    double *P[3][3];
    double *Pp[3][3];
    double dx, dy, dz, dt, eps, chi;
    double eps2[3], om2[3];
    unsigned *status;
    int ncap;
};

template <typename T>
struct ProbeArgs {
    int n;
    const T *f[6];
    const unsigned *lo[2];  // residual words of E and B (float32 + residual storage), else null
    const int *comp;
    const size_t *idx;
};

void launch_sheet_eval(const SheetEval &a, const FaddeevaTables &ft, hipStream_t s);
void launch_ref_cpml(const StepArgs<double> &a, int fi, bool isH, hipStream_t s);
void launch_ref_sum(const GridDims &g, const SumList &L, double *const T[3], hipStream_t s);
template <typename T>
void launch_wrap(const GridDims &g, T *const f[3], int dim, hipStream_t s);
void launch_ref_stepD(const StepArgs<double> &a, int np, bool nonlin, hipStream_t s);
void launch_ref_stepB(const StepArgs<double> &a, hipStream_t s);
void launch_ref_plasma_density(const StepArgs<double> &a, hipStream_t s);
void launch_ref_plasma_current(const StepArgs<double> &a, hipStream_t s);
void launch_init_accumulator(const StepArgs<double> &a, int np, hipStream_t s);
// `a` is the host copy (launch shape), `dev` the same block in device memory
template <typename T>
void launch_fused_E(const StepArgs<T> &a, const StepArgs<T> *dev, int np, bool nonlin, hipStream_t s);
template <typename T>
void launch_fused_B(const StepArgs<T> &a, const StepArgs<T> *dev, hipStream_t s);
template <typename T>
void launch_sweep(const SweepArgs<T> &a, int np, bool nonlin, hipStream_t s);
// single-pass leapfrog kernel (mpx_pass.hip): blocks of 64 z cells x 4 rows of one plane
void pass_tiles(const GridDims &g, int &nbx, int &nby, int &nbz);
template <typename T>
void launch_pass(const PassArgs<T> &a, const PassArgs<T> *dev, int np, bool nonlin, hipStream_t s);
void sweep_tiles(const GridDims &g, int variant, int seg, int &ntz, int &nty, int &nseg);
template <typename T>
void launch_sweep_batch(const SweepArgs<T> *args, const BatchRun *runs, int nruns, int total, int variant,
                        int blocks, int np, bool nonlin, bool cpml, hipStream_t s);
void fused_tiles(const GridDims &g, const int blo[3], bool hside, int L, int &ntz, int &nty, int &nseg);
template <typename T>
void launch_fused_E_batch(const StepArgs<T> *args, const BatchRun *runs, int nruns, int total, int blocks, int L,
                          int np, bool nonlin, bool cpml, hipStream_t s);
template <typename T>
void launch_fused_B_batch(const StepArgs<T> *args, const BatchRun *runs, int nruns, int total, int blocks, int L,
                          bool cpml, hipStream_t s);
// the sheets of the sides in `sides` (bit 0: E currents, bit 1: H currents); H currents at
// step *time + tstep - hshift, E currents at *time + tstep
void launch_sheet_eval_batch(const SheetEval *list, const int *first, int n, int blocks, const FaddeevaTables &ft,
                             const int *time, int tstep, hipStream_t s, int hshift = 0, int sides = 3);
// batched x march (mpx_pass.hip): run r's tiles are runs[r].first .. of a list of `total`
template <typename T>
int march_batch_rows(int rows, int np, bool nonlin, bool cp);
template <typename T>
void launch_march_batch(const PassArgs<T> *table, const BatchRun *runs, int nruns, int total, int rows, int np,
                        bool nonlin, bool cp, hipStream_t s);
void launch_advance_time(int *time, int k, hipStream_t s);
void launch_adjoint_step(const AdjArgs &a, hipStream_t s);
void launch_plane_objective(const GridDims &g, const double *const E[3], int kobj, double w, double *acc,
                            hipStream_t s);
void launch_simple_stepD(const SimpleArgs &a, hipStream_t s);
void launch_simple_stepB(const SimpleArgs &a, hipStream_t s);
template <typename T>
void launch_probe(const ProbeArgs<T> &a, double *row, hipStream_t s);
size_t energy_partials(const GridDims &g);
void launch_hbm_copy(const void *src, void *dst, size_t bytes, hipStream_t s);
// lo / q: the residual words and slot of a float32 + residual grid (null otherwise)
template <typename T>
void launch_pack_xfast(const GridDims &g, const T *src, double *dst, hipStream_t s, const unsigned *lo = nullptr,
                       int q = 0);
template <typename T>
void launch_sample_gather(const GridDims &g, const T *src, int dim, const int *c, int len, int wrapmask, double *dst,
                         hipStream_t s, const unsigned *lo = nullptr, int q = 0);
template <typename T>
void launch_energy(const GridDims &g, T *const e[3], T *const b[3], double *part, double *out, hipStream_t s,
                   const unsigned *elo = nullptr, const unsigned *blo = nullptr, double scale = 1.0);

}  // namespace mpx
