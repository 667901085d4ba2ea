/*
 * mpulse_hip.h -- C ABI of the MI355X (gfx950) time-stepping core of MPulse.
 *
 * The library owns the field state on one GPU and advances it with hand-written HIP
 * kernels.  Every entry point takes plain pointers and sizes; no C++ or torch types
 * cross this boundary.  All functions return MPX_OK (0) or a negative MPX_ERR_* code;
 * mpx_last_error() then returns a message for the calling thread.
 *
 * Reference interfaces each entry point replaces (paths relative to the MPulse tree):
 *   mpx_create        FieldSolver::initStorage              src/fieldsolver.h:12
 *                     + CurrentFactory::initCurrents        src/currentfactory.h:14
 *                     + Storage::resize/addGrid             src/storage.cpp:30-76
 *                     + Boundary::init                      src/boundary.h:44
 *   mpx_step          FDTD_PLRCSolver<Impl>::stepScheme     src/fdtd_plrc.t:127-168
 *                     (and FDTD_Plain/Nonlinear/Dispersion::stepScheme,
 *                      src/fdtd_plain.cpp:31-35, src/fdtd_nl.cpp:32-36, src/fdtd_disp.cpp:78-92)
 *                     driven by the time loop of Process::run  src/process.cpp:17-31
 *   mpx_step_init     FieldSolver::stepSchemeInit           src/fdtd_plrc.t:90-124
 *   mpx_get_field     Storage::getGrid / getBorderLayer     src/storage.cpp:48-90
 *   mpx_set_field     (same grids, written by initialisers, e.g. src/waveinit.cpp:9-60)
 *   mpx_energy        FieldEnergyDiag::calculate            src/fielddiag.cpp:255-286
 *   mpx_set_probes /  interior point histories of E and B, recorded inside every step
 *   mpx_get_probes    (no reference counterpart)
 *   mpx_sampler_*     the per-step DataGrid reads of FrequencyDiag::calculate and
 *                     FollowMax::calculate     src/freqdiag.cpp:45-70, src/followmax.cpp:73-113
 *   mpx_last_error    std::cerr + exit(-1) error reporting  (e.g. src/fdtd_plrc.t:20-24)
 *
 * Grid convention (src/globals.cpp:44-45): interior cells are 1..N in each dimension,
 * storage spans 0..N+1.  Host-side arrays exchanged through this ABI are float64,
 * shape (Nx+2, Ny+2, Nz+2), z fastest (C order).
 */
#ifndef MPULSE_HIP_H
#define MPULSE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPX_ABI_VERSION 3
#define MPX_MAX_FACTORIES 8
#define MPX_MAX_POLES 3

/* status codes */
#define MPX_OK               0
#define MPX_ERR_INVALID     -1   /* invalid descriptor or argument */
#define MPX_ERR_HIP         -2   /* a HIP runtime call failed */
#define MPX_ERR_NOMEM       -3   /* device allocation failed */
#define MPX_ERR_UNSUPPORTED -4   /* combination not provided by this build */
#define MPX_ERR_NAME        -5   /* unknown grid name */
#define MPX_ERR_NUMERIC     -6   /* non-finite value or Newton failure detected on device */

/* FieldSolver registry keys, src/fieldsim.cpp:177-195 */
#define MPX_SOLVER_PLRC_NONLIN 0  /* "fdtd-plrc-nonlin" */
#define MPX_SOLVER_PLRC_LIN    1  /* "fdtd-plrc-lin"    */
#define MPX_SOLVER_PLAIN       2  /* "fdtd-plain"       */
#define MPX_SOLVER_NL          3  /* "fdtd-nl"          */
#define MPX_SOLVER_DISP        4  /* "fdtd-disp"        */

/* Boundary registry keys, src/fieldsim.cpp:198-204 */
#define MPX_BOUNDARY_PERIODIC    0  /* "single_periodic"    */
#define MPX_BOUNDARY_XY_PERIODIC 1  /* "single_xy_periodic" */

/* storage precision of the bulk fields */
#define MPX_PREC_FP64      0  /* every word float64 */
#define MPX_PREC_FP32_BULK 1  /* E, B and CPML psi float32; PLRC accumulators and Kerr solve float64 */
#define MPX_PREC_FP32_RES  2  /* as FP32_BULK, each E, B and CPML psi word backed by a 10-bit residual
                                 (three of a cell's E or B components, or a face's two auxiliaries,
                                 share one 32-bit word): stored values within 2^-10 float32 ulp,
                                 about 1000x closer to float64 storage; PLRC accumulators and
                                 all arithmetic float64; fused and pass paths */

/* kernel structure */
#define MPX_PATH_FUSED     0  /* one fused E-side and one fused B-side kernel per step */
#define MPX_PATH_REFERENCE 1  /* one kernel per reference loop: currents, totals, update, wrap */
#define MPX_PATH_SWEEP     2  /* one kernel per step (E and B in a single x sweep; periodic boundary) */
#define MPX_PATH_PASS      3  /* one kernel per step: B of step n and E of step n+1 in one x sweep, so E,
                                 B and the accumulators cross HBM once per step; the first E half and
                                 the B half owed at a read run on the fused kernels; PLRC solvers,
                                 either boundary */

/* CurrentFactory registry keys, src/fdtd_plrc.t:431-469 */
#define MPX_FACTORY_CPML          1  /* "cpml_border"          src/cpml_border.cpp:14-140   */
#define MPX_FACTORY_CPML_1D       2  /* "cpml_border_1d"       src/cpml_border.cpp:158-255  */
#define MPX_FACTORY_SHORT_PULSE   3  /* "short_pulse_inject"   src/shortpulseinject.cpp      */
#define MPX_FACTORY_PLANE_WAVE    4  /* "wave_inject"          src/sources.cpp:8-173         */
#define MPX_FACTORY_PLANE_GAUSS   5  /* "plane_gauss_inject"   src/sources.cpp:179-331       */
#define MPX_FACTORY_PLASMA_CURRENT 7 /* "plasma_current"     src/plasmacurrent.cpp:6-74   */
#define MPX_FACTORY_GAUSS_INJECT  8  /* "gauss_inject"         src/gaussinject.cpp:33-49: a Gaussian
                                        pulse built in (kx, ky, omega) space and brought to the
                                        injection plane by three Fourier transforms (hipFFT) in
                                        blocks of time steps, injected on the down face */
#define MPX_FACTORY_FOCUSED_PULSE 6  /* "focused_pulse_inject" keys of src/focusedpulseinject.cpp:38-56;
                                        the field model is this library's own (the reference's does
                                        not compile): x-polarised Gaussian pupil of waist "width",
                                        non-paraxial angular spectrum at om0, Gaussian envelope of
                                        duration "length", injected on the down face */

typedef struct mpx_factory {
    int    kind;          /* MPX_FACTORY_* */
    int    d;             /* cpml: thickness "d" (default 8); sources: distance "d" (default 15) */
    /* cpml_border keys, src/cpml_border.cpp:146-149 */
    double kappa_max;     /* "kappaMax" (15)   */
    double a_max;         /* "aMax"     (0.25) */
    double sigma_max;     /* "sigmaMax" (3)    */
    /* short_pulse_inject keys, src/shortpulseinject.cpp:38-47 */
    double length;        /* "length" T            */
    double width;         /* "width"  w0           */
    double om0;           /* "om0"    (2 pi)       */
    double tshift;        /* "TShift"              */
    double zshift;        /* "ZShift"              */
    double phase;         /* "Phase" (units of 2 pi) */
    double amp;           /* "amp"                 */
    double eps;           /* "eps" (also wave_inject / plane_gauss_inject "eps") */
    /* wave_inject / plane_gauss_inject keys, src/sources.cpp:58-82, 225-242 */
    double k[3];          /* "kx" "ky" "kz"        */
    double h[3];          /* "Hx" "Hy" "Hz"        */
    double h_bg[3];       /* "Hx_bg" ...           */
    double e_bg[3];       /* "Ex_bg" ...           */
    double ramp;          /* "ramp" (0.5)          */
    double gwidth;        /* plane_gauss "width"  (10) */
    double goffset;       /* plane_gauss "offset" (40) */
    /* plasma_current keys, src/plasmacurrent.cpp:6-15 */
    double em;            /* "em"     (1)   e/m of the carriers */
    double gamma;         /* "gamma"  (0.1) friction           */
    double rho_bg;        /* "rho_bg" (0)   background density */
    /* gauss_inject keys (also "length" (1), "width" (1), "amp" (1), "eps" (1)),
       src/gaussinject.cpp:37-47 */
    int    gsize_t;       /* "SizeT"  (4096) time samples of the spectrum */
    int    toff;          /* "toff"   (0)    first step of the pulse      */
    double k0;            /* "k0"     (2 pi) centre wave number           */
    double gshift;        /* "Shift"  (0)    focus shift along z          */
    double gtime;         /* "Time"   (0)    time offset                  */
} mpx_factory;

typedef struct mpx_desc {
    int    abi_version;   /* MPX_ABI_VERSION */
    /* globals block, src/globals.cpp:21-36 */
    int    nx, ny, nz;    /* "grid-x" ... */
    double dx, dy, dz;    /* "grid-dx" ... */
    double dt;            /* "dt" */
    int    solver;        /* MPX_SOLVER_* */
    int    boundary;      /* MPX_BOUNDARY_* */
    int    precision;     /* MPX_PREC_* */
    int    path;          /* MPX_PATH_* */
    /* FDTD_PLRCSolver keys, src/fdtd_plrc.t:417-429, src/fdtd_plrc.cpp:463;
       FDTD_Nonlinear src/fdtd_nl.cpp:150-151 */
    double eps;           /* "eps" (1)   */
    double chi;           /* "chi" (0.1) */
    double l_eps[MPX_MAX_POLES];    /* "L1eps".. (0) */
    double l_delta[MPX_MAX_POLES];  /* "L1delta".. (0) */
    double l_om2[MPX_MAX_POLES];    /* "L1Om2".. (1) */
    /* FDTD_Dispersion keys, src/fdtd_disp.cpp:256-265 */
    double d_eps2[MPX_MAX_POLES];   /* "L1eps2".. */
    double d_om2[MPX_MAX_POLES];    /* "L1Om2"..  */
    int    n_factories;   /* current factories in setup-file order */
    mpx_factory factories[MPX_MAX_FACTORIES];
    /* "plasma" block of the PLRC solvers (PlasmaDensity, src/plasmadensity.cpp:62-73):
       multiphoton + avalanche ionisation feeding the density and the resistivity Sigma */
    int    plasma;        /* 1 when the block is present */
    int    plasma_k;      /* "K"    (5) multiphoton order */
    double plasma_nu;     /* "nu"   (1) avalanche rate    */
    double plasma_mpa;    /* "mpa"  (1) multiphoton rate  */
    double plasma_wion;   /* "Wion" (1) ionisation energy */
} mpx_desc;

typedef struct mpx_solver mpx_solver;

/* Kernel timing filled by mpx_time_steps: mean device time per launch of each kernel class. */
#define MPX_MAX_KERNEL_CLASSES 16
typedef struct mpx_timing {
    int    n_classes;
    char   name[MPX_MAX_KERNEL_CLASSES][32];
    double mean_ms[MPX_MAX_KERNEL_CLASSES];     /* mean duration of one launch */
    int    launches[MPX_MAX_KERNEL_CLASSES];    /* launches timed */
    double step_ms;                              /* mean wall time of one whole step */
} mpx_timing;

/* Fill a descriptor with the reference defaults (MakeParamMap defaults). */
void mpx_desc_defaults(mpx_desc *desc);
void mpx_factory_defaults(mpx_factory *f, int kind);

int  mpx_create(const mpx_desc *desc, mpx_solver **out);
int  mpx_destroy(mpx_solver *s);

/* Advance nsteps full steps.  The time index seen by the sources is the solver's own
   counter, starting at 0, exactly as Process::time (src/process.cpp:22). */
int  mpx_step(mpx_solver *s, int nsteps);
/* FieldSolver::stepSchemeInit: never called by the reference driver; provided for parity. */
int  mpx_step_init(mpx_solver *s);
int  mpx_get_time(mpx_solver *s, int *time);
int  mpx_set_time(mpx_solver *s, int time);

/* Full-storage grids by reference name: "Ex".."Bz", "PsiR1x".."PsiI3z" (PLRC),
   "P1x".."P3zp" (fdtd-disp), "JxTotal".."MzTotal" (reference path only), and "Eps": a
   per-cell permittivity of the PLRC solvers (no reference counterpart; reads give the uniform
   eps until one is set; setting it needs the fused or reference-structured path).
   Buffer: (Nx+2)*(Ny+2)*(Nz+2) float64, z fastest.  Ghosts: the reference-structured path and
   the plain / Kerr-only / ADE solvers read the ghost values of E and B as set until their first
   exchange, as the reference does; the fused and sweep paths of the PLRC solvers read a
   periodic dimension's ghost as the periodic image of the interior (what the exchange puts
   there), so a state whose ghosts differ from those images steps differently there. */
int  mpx_get_field(mpx_solver *s, const char *name, double *host, size_t count);
int  mpx_set_field(mpx_solver *s, const char *name, const double *host, size_t count);
/* Field output.  FieldDiag hands its HDF5 stream the grid as Storage holds it: the storage box
   0..N+1 in every dimension, x fastest (src/fielddiag.cpp:17-24, src/hdfstream.t:88-187).
   mpx_fetch_field queues a copy of grid `name` into `host` in that order, as float64, behind
   every step queued before it, and returns at once with a handle: a device kernel packs the
   grid (strips the padding, transposes to x fastest, widens float32) into a staging buffer,
   then the copy runs on the solver's stream, so steps queued afterwards run on while the copy
   lands and the caller writes its file.  mpx_fetch_wait -- callable from any host thread --
   returns when the data is in `host` and releases the handle; `host` stays untouched by the
   caller until then.  Page-locked memory from mpx_host_alloc lets the copy overlap the
   steps.  Grids assembled on the host (PLRC accumulator parts, plasma grids, an unset Eps)
   are read synchronously.  Replaces the synchronous grid read of FieldDiag::write; the
   asynchronous handle has no counterpart in the reference. */
typedef struct mpx_fetch mpx_fetch;
int  mpx_fetch_field(mpx_solver *s, const char *name, double *host, size_t count, mpx_fetch **handle);
int  mpx_fetch_wait(mpx_fetch *handle);
int  mpx_host_alloc(size_t bytes, void **ptr);
int  mpx_host_free(void *ptr);

/* CPML border layers by reference name and Direction (0 north,1 south,2 west,3 east,4 up,5 down;
   src/mpulse.h:29): "CPMLPsiEx".."CPMLPsiBz".  Buffer is the layer's inclusive box
   lo..hi (returned through lo/hi), z fastest. */
int  mpx_get_layer(mpx_solver *s, const char *name, int dir, double *host, size_t count,
                   int lo[3], int hi[3]);

/* Point histories: after every step the listed components are sampled.
   comp: 0..2 = Ex,Ey,Ez, 3..5 = Bx,By,Bz.  Capacity = max steps recorded. */
int  mpx_set_probes(mpx_solver *s, int nprobes, const int *ijk, const int *comp, int capacity);
int  mpx_get_probes(mpx_solver *s, double *host, int *nrecorded);

/* Samplers: the per-step reads of FrequencyDiag::calculate (one cell, src/freqdiag.cpp:45-70)
   and FollowMax::calculate (a storage line, src/followmax.cpp:73-113), which read
   DataGrid(x, y, z) every step.  A sampler of grid `name` records the value at storage cell
   pos (dim = -1) or the N[dim] + 2 values of the storage line through pos along dim (0..2),
   including ghosts, once per mpx_sampler_record, in step order, without waiting for the GPU;
   mpx_sampler_drain waits and copies the records made since the last drain (nrecords x
   values per record, float64) and empties the sampler.  Recording into a full sampler
   (capacity records pending) is an error.  pos must lie in the storage box 0..N+1.  E and B
   are gathered on the device; other grids are read through mpx_get_field.  Destroy a
   sampler before its solver. */
typedef struct mpx_sampler mpx_sampler;
int  mpx_sampler_create(mpx_solver *s, const char *name, int dim, const int pos[3], int capacity,
                        mpx_sampler **out);
int  mpx_sampler_record(mpx_sampler *p);
int  mpx_sampler_drain(mpx_sampler *p, double *host, size_t count, int *nrecords);
int  mpx_sampler_destroy(mpx_sampler *p);

/* sum over interior cells of (E^2 + B^2) * dx*dy*dz  (src/fielddiag.cpp:255-286) */
int  mpx_energy(mpx_solver *s, double *energy);
/* mpx_energy without allocating and without waiting: the same value, bit for bit, written to
   the caller's device memory d_energy (one double) by kernels queued on the solver's stream;
   the caller waits (mpx_sync) before reading it.  No counterpart in the reference: for callers
   that cannot allocate device memory -- code under HIP graph capture, tight serving loops.  The
   partial sums live in a file-scope device array of MPX_NOALLOC_PARTIALS doubles (4 MiB) that
   each GPU holds until the process ends; a grid needing more partials, ceil(nz/64) x
   ceil(ny/4) x nx > MPX_NOALLOC_PARTIALS (512^3 fits), is refused with MPX_ERR_UNSUPPORTED.
   Every call on one GPU shares that array: calls that overlap on one GPU (other streams, other
   solvers, other threads) race.  A pass-path solver owing a half step is refused (settle it
   with any read first). */
#define MPX_NOALLOC_PARTIALS (1 << 19)
int  mpx_energy_to(mpx_solver *s, double *d_energy);

/* Waits for the queued steps.  Returns MPX_ERR_NUMERIC (and clears the condition) when, since
   the last check, a Kerr Newton solve stopped at its iteration cap (tuning key "newton_cap",
   default 200; the reference iterates without bound, src/fdtd_plrc.cpp:406-425) or an E update
   met a non-finite value, in its inputs or its result (the reference goes on silently: with
   Kerr it stores E = 0 where C is NaN, src/fdtd_plrc.cpp:405-424, and the accumulators keep
   the NaN); mpx_last_error says which.  The state is left as computed. */
int  mpx_sync(mpx_solver *s);
/* Run nsteps with per-kernel HIP events on the solver's stream. */
int  mpx_time_steps(mpx_solver *s, int nsteps, mpx_timing *out);
/* Device-side byte size of the resident state. */
int  mpx_state_bytes(mpx_solver *s, size_t *bytes);

/* Launch-shape parameters: "march_e" / "march_b" = x planes walked by one thread in the
   E-side / B-side fused kernel (1, 2, 4); "sweep_seg" = x planes per block of the sweep kernel;
   "sweep_variant" = its thread-block shape (0: 64x8, 1: 64x4, 2: 64x16 threads; one
   tile row per thread, the last row and lane of a tile being its halo).  "tile_order" = block order
   of the fused kernels (0: z tiles fastest, 1: x chunks fastest, 2, the default: bands of
   "tile_band" tile rows walked x plane by x plane; "tile_band" 0, the default, chooses the
   height from the grid); "b_reverse" 1, the default, walks that order backwards in the B-side
   kernel, so it starts on the tiles the E-side kernel wrote last (0: forwards).  No
   counterpart in the reference: the hook of
   the tile autotuner.  "newton_cap" (>= 1, default 200) bounds the Kerr Newton iterations of a
   cell (see mpx_sync).  Pass path: "pass_kernel" 1, the default, marches each block over
   "march_seg" x planes (0, the default: 16, halved down to 2 on grids too small to fill the
   chip) of a 64-cell x "march_rows" tile (3 or 7; 0, the default, takes
   7 where the kernel variant fits four waves per SIMD, else 3);
   0 updates one cell per thread. */
int  mpx_set_tuning(mpx_solver *s, const char *key, int value);

/* Streaming copy of `bytes` between two fresh device buffers, `reps` times; mean ms per copy.
   No counterpart in the reference: a measurement aid that calibrates HBM byte counters and
   gives the practical copy bandwidth beside the stencil kernels. */
int  mpx_copy_probe(size_t bytes, int reps, double *ms_per_copy);

/* ---- batched independent runs -------------------------------------------------------
   No counterpart in the reference, which runs one simulation per process.  For uneven-size
   parameter sweeps (pulse energy, chi3, grid size): every run is an FDTD_PLRC solver on a
   periodic domain built exactly as mpx_create builds it (sweep path), and all runs advance in
   lockstep, one launch per kernel and step for all of them, replayed from a captured HIP
   graph.  The runs must share precision, solver (lin / nonlin), number of active poles and
   whether a CPML is present.  mpx_batch_run hands out a run's handle for reading and writing
   fields (mpx_get_field, mpx_energy, ...); stepping, timing or destroying such a handle
   directly is refused. */
#define MPX_MAX_BATCH 1024
typedef struct mpx_batch mpx_batch;
int  mpx_batch_create(const mpx_desc *descs, int nruns, mpx_batch **out);
int  mpx_batch_destroy(mpx_batch *b);
int  mpx_batch_size(mpx_batch *b, int *nruns);
int  mpx_batch_run(mpx_batch *b, int run, mpx_solver **out);
int  mpx_batch_step(mpx_batch *b, int nsteps);
int  mpx_batch_sync(mpx_batch *b);
int  mpx_batch_get_time(mpx_batch *b, int *time);
/* "structure" (0: fused E and fused B kernels over all runs, the default; 1: single-pass sweep;
   2: x-march pass kernel, one launch per step for B of one step and E of the next, with the
   fused E kernel opening and the fused B kernel closing every mpx_batch_step call, so that E
   and B belong to the same step between calls), "march_b" (fused B: x planes per thread,
   1 / 2 / 4), "fused_blocks_per_cu" (fused persistent grid = value x CUs), "variant" (sweep
   block shape 0..2, as sweep_variant), "seg" (sweep x planes per tile), "blocks_per_cu" (sweep
   persistent grid), "pass_seg" (x-march planes per block), "pass_rows" (x-march rows per
   tile: 3, 7, or 0 for the variant's default), "graph_steps" (steps per captured graph;
   0 = none) */
int  mpx_batch_set_tuning(mpx_batch *b, const char *key, int value);
/* Times the candidate launch shapes of the three structures for steps_per_trial steps each,
   keeps the fastest, then returns every run to its all-zero state at time 0.  result =
   {structure, then march_b, blocks per CU, 0 (fused) or variant, seg, blocks per CU (sweep) or
   pass_seg, pass_rows, 0 (x-march)}.  Only before the first step. */
int  mpx_batch_autotune(mpx_batch *b, int steps_per_trial, int result[4], double *best_ms_per_step);
/* Advance nsteps and report the mean device time per step (HIP events on the batch stream). */
int  mpx_batch_time_steps(mpx_batch *b, int nsteps, double *ms_per_step);

/* ---- gradient of a field objective by the discrete adjoint ---------------------------
   No counterpart in the reference.  Objective: J = weight * sum over the nsteps steps of
   sum_{i,j} |E(i, j, plane_k)|^2 after each step.  The solver is stepped nsteps from its
   current state; the adjoint of every step is then run backwards through the same discrete
   update (fused float64 kernels), recomputing the forward E of each segment of
   checkpoint_every steps from a stored checkpoint.  Gradients: dJ/dEps(x) for every interior
   cell (full-storage layout, zero in the ghosts) and dJ/dchi (0 for the linear core).  The
   solver ends at the final forward state.  Scope: FDTD_PLRC solvers (Kerr or linear core,
   with or without Lorentz poles), float64, fused path, periodic boundary (CPML and current
   sheets allowed; no plasma, no gauss_inject).  With dJ_deps and dJ_dchi both NULL only the
   forward objective is computed. */
typedef struct mpx_objective {
    int    plane_k;   /* z index of the observation plane, 1..nz */
    double weight;
} mpx_objective;
int  mpx_adjoint_grad(mpx_solver *s, const mpx_objective *obj, int nsteps, int checkpoint_every, double *J,
                      double *dJ_deps, double *dJ_dchi);

/* The value a field word holds once stored in `precision` (MPX_PREC_*): x itself (float64),
   its float32 rounding, or the float32 word plus residual; what mpx_set_field stores for x.
   No counterpart in the reference. */
double mpx_store_round(double x, int precision);

const char *mpx_last_error(void);
int  mpx_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MPULSE_HIP_H */
