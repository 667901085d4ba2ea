// Host side of the C ABI declared in include/mpulse_hip.h: builds the device descriptors
// from an mpx_desc the way the reference builds its currents and storage, owns the device
// state and enqueues the kernels of one time step on the solver's HIP stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <array>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/mpulse_hip.h"
#include "mpx_gauss.h"
#include "mpx_launch.h"

using mpx::CpmlFace;
using mpx::GridDims;
using mpx::StepArgs;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define MPX_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(MPX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

#define MPX_LAUNCHED(what)                                                                      \
    do {                                                                                        \
        hipError_t e_ = hipGetLastError();                                                      \
        if (e_ != hipSuccess) return fail(MPX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Direction enum of src/mpulse.h:28
enum { NORTH = 0, SOUTH = 1, WEST = 2, EAST = 3, UP = 4, DOWN = 5 };
const int kFaceOrder[6] = {NORTH, SOUTH, EAST, WEST, UP, DOWN};  // src/cpml_border.cpp:16-51

int dim_of(int dir) { return (dir == EAST || dir == WEST) ? 0 : ((dir == NORTH || dir == SOUTH) ? 1 : 2); }
bool is_high(int dir) { return dir == EAST || dir == NORTH || dir == UP; }

enum CurType { CUR_CPML_E = 1, CUR_CPML_H, CUR_INC_E, CUR_INC_H, CUR_PLASMA };

struct Layer {
    std::string name;
    int dir;
    int lo[3], hi[3];
    void *comp[3];  // device arrays (nullptr where the component is never written)
    unsigned *rw = nullptr;  // float32 + residual storage of a CPML auxiliary: the face's residual words
    int slot = 0;            // and this component's slot in them
    size_t count() const
    {
        return (size_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
    }
};

struct Current {
    int type;
    int dir;
    int factory;
    int layer = -1;   // index into layers (first of the three component layers)
    int face = -1;    // CPML: index into the side's face list
    // source evaluation
    mpx::SheetEval ev{};
    std::shared_ptr<mpx::GaussInjectDev> gauss;  // GaussInject: block schedule and samples
};

struct Timer {
    bool on = false;
    std::vector<int> cls;
    std::vector<hipEvent_t> ev0, ev1;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    hipEvent_t get()
    {
        if (used < pool.size()) return pool[used++];
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        pool.push_back(e);
        ++used;
        return e;
    }
};

const char *kClassNames[] = {"sheet_eval", "cpml_E", "sum_J", "stepD_ref", "wrap_E", "cpml_H", "sum_M",
                             "stepB_ref", "wrap_B", "fused_E", "fused_B", "simple_E", "simple_B", "probe", "sweep",
                             "pass"};
enum {
    K_SHEET = 0, K_CPML_E, K_SUM_J, K_STEPD_REF, K_WRAP_E, K_CPML_H, K_SUM_M, K_STEPB_REF, K_WRAP_B,
    K_FUSED_E, K_FUSED_B, K_SIMPLE_E, K_SIMPLE_B, K_PROBE, K_SWEEP, K_PASS, K_NCLASSES
};

}  // namespace

struct mpx_solver {
    mpx_desc d{};
    int solver = 0, prec = 0, path = 0;
    bool plrc = false, nonlin = false;
    size_t esz = 8;
    bool res = false;  // float32 + residual storage (MPX_PREC_FP32_RES)
    GridDims g{};
    size_t cells = 0;  // padded element count of one grid
    hipStream_t stream = nullptr;
    int device = 0;  // the GPU current when the solver was created
    int periodic[3] = {1, 1, 1};
    double delta[3] = {0, 0, 0};

    std::vector<void *> allocs;
    void *E[3] = {nullptr, nullptr, nullptr};
    void *B[3] = {nullptr, nullptr, nullptr};
    unsigned *Elo = nullptr, *Blo = nullptr;  // residual words of E and B (float32 + residual storage)
    double2 *psi[3][3] = {};
    int np = 0;
    int pole_ref[3] = {-1, -1, -1};  // active pole slot -> reference pole index
    std::vector<double> kE[3], kH[3];
    double *d_kdE[3] = {}, *d_kdH[3] = {};
    double *d_ikdE[3] = {}, *d_ikdH[3] = {};  // reciprocals, for the fused and sweep kernels
    double *d_eps = nullptr;                    // per-cell permittivity, once set (full storage)
    // plasma: PlasmaDensity OptField and PlasmaCurrent (float64 grids over the storage box)
    bool plasma = false, pcur = false;
    int pcur_pos = 0;
    double pcur_em = 0, pcur_gamma = 0, pcur_rhobg = 0;
    double *d_rho = nullptr, *d_jp[3] = {}, *d_sigma = nullptr;

    std::vector<Layer> layers;
    std::vector<Current> ecur, hcur;
    std::vector<CpmlFace> efaces, hfaces;
    int sheets_first = 0;

    double *Jt[3] = {}, *Mt[3] = {};
    bool haveJ = false, haveM = false;
    double *P[3][3] = {}, *Pp[3][3] = {};

    mpx::PoleConsts pc{};
    double sumChi0 = 0, sumXi0 = 0, denom = 1, numer = 1;
    mpx::FaddeevaTables ft{};

    int nprobe = 0, probe_cap = 0, probe_rec = 0;
    int *d_probe_comp = nullptr;
    size_t *d_probe_idx = nullptr;
    double *d_probe_hist = nullptr;

    double *d_part = nullptr, *d_energy = nullptr;
    unsigned *d_status = nullptr;  // device status (MPX_ERR_NUMERIC bits), read and cleared by mpx_sync
    int ncap = mpx::kNewtonCap;    // Kerr Newton iteration cap
    int time = 0;
    // x planes per thread in the fused kernels (1, 2 or 4); with the pole accumulators and the
    // cell's own field streamed nontemporally, E: 1 and B: 2 are fastest in both precisions
    int march_e = 1, march_b = 2;
    int tile_order = 2;            // fused kernels: block order, 0 = z tiles fastest, 1 = x chunks fastest,
                                   // 2 = bands of tile_band rows walked plane by plane (fastest)
    int tile_band = 0;             // 0: chosen from the grid by auto_band
    int b_reverse = 1;             // B kernel walks the banded order backwards
    // pass path: 1 = x march (k_march, `march_seg` planes x `march_rows` rows per block),
    // 0 = one cell per thread (k_pass)
    // seg 0: 16 planes, fewer where the grid has too few tiles to fill the chip
    int pass_kernel = 1, march_seg = 0, march_rows = 0;  // rows 0: chosen per kernel variant
    double *d_stage = nullptr;     // x-fastest float64 copy of one grid for mpx_fetch_field
    // sweep path: second set of buffers (swapped with the first after every step)
    void *E2[3] = {}, *B2[3] = {};
    double2 *psi2[3][3] = {};
    std::vector<void *> epsi2[2];      // per E face: alternate psi0 / psi1
    std::vector<int> eface_layer[2];   // per E face: layer holding psi0 / psi1
    int sweep_seg = 32, sweep_variant = 2;
    // pass path: between calls B lags E by the B half of the last step (pending) until a read
    // or write of the state settles it; E, B, their residual words and the CPML H
    // auxiliaries alternate between two sets (the second: E2, B2, E2lo, B2lo, hpsi2, hrw2)
    bool pending = false;
    int parity = 0;
    unsigned *E2lo = nullptr, *B2lo = nullptr;
    std::vector<void *> hpsi2[2];      // per H face: alternate psi0 / psi1
    std::vector<unsigned *> hrw2;      // per H face: alternate residual words
    std::vector<int> hface_layer[2];   // per H face: layer holding psi0 / psi1
    Timer timer;
    hipStream_t owner = nullptr;  // stream of the batch that steps this run (nullptr: none)
    // device copies of the fused kernels' argument blocks (E side, B side), refreshed when the
    // host-side block changes
    // (E side, B side) x buffer parity of the pass path; the pass kernel's own per parity
    void *d_fargs[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<unsigned char> h_fargs[4];
    void *d_pargs[2] = {nullptr, nullptr};
    std::vector<unsigned char> h_pargs[2];

    ~mpx_solver()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : allocs) (void)hipFree(p);
        for (hipEvent_t e : timer.pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename T>
    T *dalloc(size_t count, int &rc)
    {
        void *p = nullptr;
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            rc = fail(MPX_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
            return nullptr;
        }
        allocs.push_back(p);
        if (hipMemsetAsync(p, 0, count * sizeof(T), stream) != hipSuccess) {
            rc = fail(MPX_ERR_HIP, "hipMemsetAsync failed");
            return nullptr;
        }
        return (T *)p;
    }

    void tick(int cls)
    {
        if (!timer.on) return;
        timer.cls.push_back(cls);
        hipEvent_t e = timer.get();
        (void)hipEventRecord(e, stream);
        timer.ev0.push_back(e);
    }
    void tock()
    {
        if (!timer.on) return;
        hipEvent_t e = timer.get();
        (void)hipEventRecord(e, stream);
        timer.ev1.push_back(e);
    }
};

// a run owned by a batch is stepped only through the batch; state reads wait for its stream
#define MPX_NOT_IN_BATCH(s)                                                                          \
    do {                                                                                             \
        if ((s) && (s)->owner) return fail(MPX_ERR_INVALID, "run belongs to a batch: use mpx_batch_*"); \
    } while (0)
#define MPX_OWNER_SYNC(s)                                          \
    do {                                                           \
        if ((s) && (s)->owner) MPX_HIP(hipStreamSynchronize((s)->owner)); \
    } while (0)

// synchronous host-to-device copy ordered after the allocations' clears on the solver stream
static int upload(mpx_solver *s, void *dst, const void *src, size_t bytes)
{
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess) return 1;
    return hipStreamSynchronize(s->stream) != hipSuccess;
}

// ---------------------------------------------------------------------------------------
// set-up helpers mirroring the reference's Storage / current construction
// ---------------------------------------------------------------------------------------

// Storage::getBorderExtent, src/storage.cpp:208-318 (single domain, bordercells = 0)
static bool border_extent(const mpx_solver *s, int dir, int thickness, int distance, int blow[3], int bhigh[3])
{
    int low[3] = {0, 0, 0}, high[3], glow[3] = {0, 0, 0}, ghigh[3];
    for (int d = 0; d < 3; ++d) {
        high[d] = s->g.n[d] + 1;
        ghigh[d] = high[d];
        blow[d] = low[d];
        bhigh[d] = high[d];
    }
    switch (dir) {
    case WEST:
        if (low[0] < glow[0] + thickness + distance) {
            bhigh[0] = glow[0] + thickness - 1 + distance;
            blow[0] = glow[0] + distance;
            return true;
        }
        return false;
    case SOUTH:
        if (low[1] < glow[1] + thickness + distance) {
            bhigh[1] = glow[1] + thickness - 1 + distance;
            blow[1] = glow[1] + distance;
            return true;
        }
        return false;
    case DOWN:
        if ((low[2] < glow[2] + thickness + distance) && (high[2] >= glow[2] + thickness)) {
            bhigh[2] = glow[2] + thickness - 1 + distance;
            blow[2] = glow[2] + distance;
            return true;
        }
        return false;
    case EAST:
        if (high[0] > ghigh[0] - thickness - distance) {
            blow[0] = ghigh[0] - thickness + 1 - distance;
            bhigh[0] = ghigh[0] - distance;
            return true;
        }
        return false;
    case NORTH:
        if (high[1] > ghigh[1] - thickness - distance) {
            blow[1] = ghigh[1] - thickness + 1 - distance;
            bhigh[1] = ghigh[1] - distance;
            return true;
        }
        return false;
    case UP:
        if (high[2] > ghigh[2] - thickness - distance) {
            blow[2] = ghigh[2] - thickness + 1 - distance;
            bhigh[2] = ghigh[2] - distance;
            return true;
        }
        return false;
    }
    return false;
}

// Storage::addBorderLayer, src/storage.cpp:106-142: one grid per (name, direction); a second
// request for the same pair returns the first grid.  Returns the layer index or -1.
static int add_border_layer(mpx_solver *s, const std::string &name, int dir, int thickness, int distance)
{
    for (size_t q = 0; q < s->layers.size(); ++q)
        if (s->layers[q].dir == dir && s->layers[q].name == name) return (int)q;
    Layer L;
    L.name = name;
    L.dir = dir;
    L.comp[0] = L.comp[1] = L.comp[2] = nullptr;
    if (!border_extent(s, dir, thickness, distance, L.lo, L.hi)) return -1;
    s->layers.push_back(L);
    return (int)s->layers.size() - 1;
}

// CPMLBorder::initCoefficients, src/cpml_border.cpp:58-140
static void cpml_kappa(mpx_solver *s, int thickness, double kappaMax)
{
    for (int dim = 0; dim < 3; ++dim) {
        std::vector<double> &KE = s->kE[dim], &KH = s->kH[dim];
        int low = 0, high = s->g.n[dim] + 1, glow = 0, ghigh = high;
        if (low < glow + thickness) {
            for (double &v : KE) v = 1.0;
            for (double &v : KH) v = 1.0;
            for (int i = 0; i <= thickness; ++i) {
                double x = 1 - double(i) / double(thickness);
                double x3 = x * x * x;
                if (low + i <= high) KE[low + i] = 1 + (kappaMax - 1) * x3;
            }
            for (int i = 0; i < thickness; ++i) {
                double x = 1 - (double(i) + 0.5) / double(thickness);
                double x3 = x * x * x;
                if (low + i <= high) KH[low + i] = 1 + (kappaMax - 1) * x3;
            }
        }
        if (high > ghigh - thickness) {
            for (int i = 0; i <= thickness; ++i) {
                double x = 1 - double(i) / double(thickness);
                double x3 = x * x * x;
                if (high - i >= low) KH[high - i] = 1 + (kappaMax - 1) * x3;
            }
            for (int i = 0; i < thickness; ++i) {
                double x = 1 - (double(i) + 0.5) / double(thickness);
                double x3 = x * x * x;
                if (high - i >= low) KE[high - i] = 1 + (kappaMax - 1) * x3;
            }
        }
    }
}

// CPMLBorderOneD::initCoefficients, src/cpml_border.cpp:179-244
static void cpml1d_kappa(mpx_solver *s, int thickness, double kappaMax)
{
    for (int dim = 0; dim < 2; ++dim) {
        for (double &v : s->kE[dim]) v = 1.0;
        for (double &v : s->kH[dim]) v = 1.0;
    }
    std::vector<double> &KE = s->kE[2], &KH = s->kH[2];
    int low = 0, high = s->g.n[2] + 1, glow = 0, ghigh = high;
    if (low < glow + thickness) {
        for (double &v : KE) v = 1.0;
        for (double &v : KH) v = 1.0;
        for (int i = 0; i < thickness; ++i) {
            double x = 1 - double(i) / double(thickness);
            double x3 = x * x * x;
            if (low + i + 1 <= high) KE[low + i + 1] = 1 + (kappaMax - 1) * x3;
        }
        for (int i = 0; i < thickness; ++i) {
            double x = 1 - (double(i) + 0.5) / double(thickness);
            double x3 = x * x * x;
            if (low + i <= high) KH[low + i] = 1 + (kappaMax - 1) * x3;
        }
    }
    if (high > ghigh - thickness) {
        for (int i = 0; i < thickness; ++i) {
            double x = 1 - (double(i)) / double(thickness);
            double x3 = x * x * x;
            if (high - i - 1 >= low) KE[high - i - 1] = 1 + (kappaMax - 1) * x3;
        }
        for (int i = 0; i < thickness; ++i) {
            double x = 1 - double(i + 0.5) / double(thickness);
            double x3 = x * x * x;
            if (high - i - 1 >= low) KH[high - i - 1] = 1 + (kappaMax - 1) * x3;
        }
    }
}

// CPMLBorderCurrent::makeCoeff, src/cpml_border.cpp:287-339
static void cpml_coeff(const mpx_factory &f, int dir, bool isH, int lowk, int highk, double dt,
                       std::vector<double> &b, std::vector<double> &c)
{
    bool reverse = !is_high(dir);
    b.assign(highk - lowk + 1, 0.0);
    c.assign(highk - lowk + 1, 0.0);
    double offset = isH ? 0.5 : 0.0;
    double eps = 1;  // CPMLBorder passes eps = 1, src/cpml_border.cpp:17-52
    for (int k = 0; k < f.d; ++k) {
        double x = 1 - (double(k) - offset) / double(f.d);
        double x3 = x * x * x;
        int pos = reverse ? (lowk + k) : (highk - k);
        double sigma = x3 * f.sigma_max;
        double kappa = 1 + (f.kappa_max - 1) * x3;
        double a = f.a_max * (1 - x);
        double bb = exp(-(sigma / kappa + a) * dt / eps);
        double cc = sigma * (bb - 1) / (kappa * (sigma + kappa * a));
        if (pos - lowk >= 0 && pos - lowk < (int)b.size()) {
            b[pos - lowk] = bb;
            c[pos - lowk] = cc;
        }
    }
}

// pole constants, src/fdtd_plrc.t:181-216
static void pole_constants(const double LEps[3], const double LDelta[3], const double LOm2[3], double dt,
                           std::complex<double> Crec[3], std::complex<double> dchi0[3],
                           std::complex<double> dxi0[3], double &sumChi0, double &sumXi0)
{
    std::complex<double> I(0, 1), sumChi = 0.0, sumXi = 0.0;
    for (int n = 0; n < 3; ++n) {
        double beta = std::sqrt(LOm2[n] - LDelta[n] * LDelta[n]);
        double gamma = LEps[n] * LOm2[n] / beta;
        std::complex<double> phasor = LDelta[n] - I * beta;
        Crec[n] = std::exp(-phasor * dt);
        std::complex<double> chi0 = (I * gamma * (Crec[n] - 1.0) / phasor);
        std::complex<double> xi0 = I * gamma * (Crec[n] * (phasor * dt + 1.0) - 1.0) / (phasor * phasor * dt);
        dchi0[n] = chi0 * (1.0 - Crec[n]);
        dxi0[n] = xi0 * (1.0 - Crec[n]);
        sumChi += chi0;
        sumXi += xi0;
    }
    sumChi0 = std::real(sumChi);
    sumXi0 = std::real(sumXi);
}

namespace mpx {
double pulse_norm_device(const PulseParams &P, const FaddeevaTables &ft, hipStream_t s);
}

// Focused pulse model (the library's own; the reference's does not compile): an x-polarised
// Gaussian pupil A(k) = exp(-(kx^2+ky^2) w0^2/4) at the carrier om0 with vector plane waves
// E(k) = A (1, 0, -kx/kz), B(k) = k x E / om0, kz = sqrt(eps om0^2 - k_perp^2) (evanescent
// waves dropped), summed over the periodic transverse grid and normalised to |E_x| = 1 at
// the focus.  comp: 0 = E_x, 1 = B_x, 2 = B_y; (sx, sy) = Yee offsets of the component.
static void focused_plane(const mpx_factory &f, int nx, int ny, double dx, double dy, double z, int comp, double sx,
                          double sy, std::vector<double2> &out)
{
    typedef std::complex<double> cplx;
    double k = std::sqrt(f.eps) * f.om0, w0 = f.width;
    double cx = 0.5 * (nx + 1), cy = 0.5 * (ny + 1);
    int nx2 = nx + 2, ny2 = ny + 2;
    std::vector<double> kxs(nx), kys(ny);
    for (int m = 0; m < nx; ++m) kxs[m] = 2 * M_PI * (m - nx / 2) / (nx * dx);
    for (int m = 0; m < ny; ++m) kys[m] = 2 * M_PI * (m - ny / 2) / (ny * dy);
    // spectrum G(kx, ky) at plane z, and the normalisation sum of A
    std::vector<cplx> G((size_t)nx * ny);
    double norm = 0;
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) {
            double kp2 = kxs[a] * kxs[a] + kys[b] * kys[b];
            if (kp2 >= k * k) continue;
            double kz = std::sqrt(k * k - kp2), A = std::exp(-kp2 * w0 * w0 / 4);
            norm += A;
            double g = comp == 0 ? A : (comp == 1 ? -kxs[a] * kys[b] * A / (kz * f.om0) : A * (kz * kz + kxs[a] * kxs[a]) / (kz * f.om0));
            G[(size_t)a * ny + b] = g * std::exp(cplx(0, kz * z));
        }
    // separable synthesis with phasor tables: over ky, then over kx
    std::vector<cplx> ey((size_t)ny * ny2), ex((size_t)nx * nx2);
    for (int b = 0; b < ny; ++b)
        for (int j = 0; j < ny2; ++j) ey[(size_t)b * ny2 + j] = std::exp(cplx(0, kys[b] * (j + sy - cy) * dy));
    for (int a = 0; a < nx; ++a)
        for (int i = 0; i < nx2; ++i) ex[(size_t)a * nx2 + i] = std::exp(cplx(0, kxs[a] * (i + sx - cx) * dx));
    std::vector<cplx> tmp((size_t)nx * ny2);
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) {
            cplx g = G[(size_t)a * ny + b];
            if (g == cplx(0, 0)) continue;
            cplx *row = &tmp[(size_t)a * ny2];
            const cplx *ph = &ey[(size_t)b * ny2];
            for (int j = 0; j < ny2; ++j) row[j] += g * ph[j];
        }
    std::vector<cplx> res((size_t)nx2 * ny2);
    for (int a = 0; a < nx; ++a)
        for (int i = 0; i < nx2; ++i) {
            cplx ph = ex[(size_t)a * nx2 + i];
            const cplx *row = &tmp[(size_t)a * ny2];
            cplx *dst = &res[(size_t)i * ny2];
            for (int j = 0; j < ny2; ++j) dst[j] += ph * row[j];
        }
    out.resize((size_t)nx2 * ny2);
    for (size_t q = 0; q < out.size(); ++q) out[q] = make_double2(res[q].real() / norm, res[q].imag() / norm);
}

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------

extern "C" {

const char *mpx_last_error(void) { return g_err.c_str(); }

double mpx_store_round(double x, int precision)
{
    if (precision == MPX_PREC_FP32_BULK) return (double)(float)x;
    if (precision == MPX_PREC_FP32_RES) {
        unsigned w = 0;
        const float hi = mpx::res_encode(x, w, 0);
        return mpx::res_decode(hi, w, 0);
    }
    return x;
}
int mpx_abi_version(void) { return MPX_ABI_VERSION; }

void mpx_factory_defaults(mpx_factory *f, int kind)
{
    std::memset(f, 0, sizeof(*f));
    f->kind = kind;
    if (kind == MPX_FACTORY_CPML || kind == MPX_FACTORY_CPML_1D) {
        f->d = 8;  // src/cpml_border.cpp:146-149
        f->kappa_max = 15;
        f->a_max = 0.25;
        f->sigma_max = 3;
    } else {
        f->d = 15;  // src/incsource.cpp:84
        f->length = 1.;  // src/shortpulseinject.cpp:38-47
        f->width = 1.;
        f->om0 = 2 * M_PI;
        f->amp = 1;
        f->eps = 1;
        f->k[2] = 1;  // src/sources.cpp:62-80
        f->ramp = 0.5;
        f->gwidth = 10;  // src/sources.cpp:237-239
        f->goffset = 40;
        if (kind == MPX_FACTORY_PLANE_WAVE || kind == MPX_FACTORY_PLANE_GAUSS) f->amp = 0;
    }
    f->em = 1.0;  // src/plasmacurrent.cpp:10-12
    f->gamma = 0.1;
    f->rho_bg = 0.0;
    f->gsize_t = 4096;  // src/gaussinject.cpp:37-47
    f->toff = 0;
    f->k0 = 2 * M_PI;
    f->gshift = 0;
    f->gtime = 0;
}

void mpx_desc_defaults(mpx_desc *d)
{
    std::memset(d, 0, sizeof(*d));
    d->abi_version = MPX_ABI_VERSION;
    d->nx = d->ny = d->nz = 128;  // src/globals.cpp:23-33
    d->dx = d->dy = d->dz = 0.1;
    d->dt = 1e-2;
    d->solver = MPX_SOLVER_PLRC_NONLIN;
    d->boundary = MPX_BOUNDARY_PERIODIC;
    d->precision = MPX_PREC_FP64;
    d->path = MPX_PATH_FUSED;
    d->eps = 1.0;  // src/fdtd_plrc.t:417
    d->chi = 0.1;  // src/fdtd_plrc.cpp:463
    for (int n = 0; n < 3; ++n) d->l_om2[n] = 1.0;  // src/fdtd_plrc.t:427-429
    d->plasma_k = 5;  // src/plasmadensity.cpp:66-69
    d->plasma_nu = d->plasma_mpa = d->plasma_wion = 1.0;
}

static int build(mpx_solver *s, const mpx_desc *desc);

int mpx_create(const mpx_desc *desc, mpx_solver **out)
{
    if (!desc || !out) return fail(MPX_ERR_INVALID, "null descriptor or output pointer");
    *out = nullptr;
    if (desc->abi_version != MPX_ABI_VERSION) return fail(MPX_ERR_INVALID, "descriptor ABI version mismatch");
    if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1) return fail(MPX_ERR_INVALID, "grid sizes must be positive");
    if (!(desc->dt > 0) || !(desc->dx > 0) || !(desc->dy > 0) || !(desc->dz > 0))
        return fail(MPX_ERR_INVALID, "grid spacing and dt must be positive");
    if (desc->solver < MPX_SOLVER_PLRC_NONLIN || desc->solver > MPX_SOLVER_DISP)
        return fail(MPX_ERR_INVALID, "unknown solver");
    if (desc->boundary != MPX_BOUNDARY_PERIODIC && desc->boundary != MPX_BOUNDARY_XY_PERIODIC)
        return fail(MPX_ERR_INVALID, "unknown boundary");
    if (desc->plasma && desc->solver != MPX_SOLVER_PLRC_NONLIN && desc->solver != MPX_SOLVER_PLRC_LIN)
        return fail(MPX_ERR_UNSUPPORTED, "the plasma block belongs to the FDTD_PLRC solvers");
    if (desc->precision != MPX_PREC_FP64 && desc->precision != MPX_PREC_FP32_BULK &&
        desc->precision != MPX_PREC_FP32_RES)
        return fail(MPX_ERR_INVALID, "unknown precision");
    if (desc->path != MPX_PATH_FUSED && desc->path != MPX_PATH_REFERENCE && desc->path != MPX_PATH_SWEEP &&
        desc->path != MPX_PATH_PASS)
        return fail(MPX_ERR_INVALID, "unknown kernel path");
    if (desc->n_factories < 0 || desc->n_factories > MPX_MAX_FACTORIES)
        return fail(MPX_ERR_INVALID, "too many current factories");
    mpx_solver *s = new (std::nothrow) mpx_solver();
    if (!s) return fail(MPX_ERR_NOMEM, "host allocation failed");
    int rc = build(s, desc);
    if (rc != MPX_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return MPX_OK;
}

int mpx_destroy(mpx_solver *s)
{
    MPX_NOT_IN_BATCH(s);
    delete s;
    return MPX_OK;
}

}  // extern "C"

static int build(mpx_solver *s, const mpx_desc *desc)
{
    s->d = *desc;
    s->solver = desc->solver;
    s->prec = desc->precision;
    s->path = desc->path;
    s->plrc = (s->solver == MPX_SOLVER_PLRC_NONLIN || s->solver == MPX_SOLVER_PLRC_LIN);
    s->nonlin = (s->solver == MPX_SOLVER_PLRC_NONLIN);
    if (s->prec != MPX_PREC_FP64 && s->prec != MPX_PREC_FP32_BULK && s->prec != MPX_PREC_FP32_RES)
        return fail(MPX_ERR_INVALID, "unknown precision " + std::to_string(s->prec));
    s->res = s->prec == MPX_PREC_FP32_RES;
    s->esz = (s->prec == MPX_PREC_FP64) ? 8 : 4;
    // the residual B kernel keeps its planes' decoded values in float64 registers: one plane
    // per thread runs at 8 waves per SIMD against 5 for two (1.72 against 1.82 ms at 512^3)
    if (s->res) s->march_b = 1;
    s->delta[0] = desc->dx;
    s->delta[1] = desc->dy;
    s->delta[2] = desc->dz;
    s->periodic[2] = (desc->boundary == MPX_BOUNDARY_PERIODIC) ? 1 : 0;

    if (!s->plrc && desc->n_factories > 0)
        // FieldSolver::addCurrent default, src/fieldsolver.h:15-23
        return fail(MPX_ERR_INVALID, "Additional currents cannot be used when using this Field Solver");
    if (!s->plrc && s->prec != MPX_PREC_FP64)
        return fail(MPX_ERR_UNSUPPORTED, "fdtd-plain/nl/disp run in float64 only");
    if (s->plrc && (desc->l_om2[0] == 0 || desc->l_om2[1] == 0 || desc->l_om2[2] == 0))
        return fail(MPX_ERR_INVALID, "FDTD_PLRCSolver: Omega should not be zero!");  // src/fdtd_plrc.t:20-24
    if (s->plrc && s->path == MPX_PATH_REFERENCE && s->prec != MPX_PREC_FP64)
        return fail(MPX_ERR_UNSUPPORTED, "the reference-structured path runs in float64 only");
    if (s->res && s->path == MPX_PATH_SWEEP)
        return fail(MPX_ERR_UNSUPPORTED, "float32 + residual storage runs on the fused path");
    if (s->path == MPX_PATH_SWEEP && (!s->plrc || desc->boundary != MPX_BOUNDARY_PERIODIC))
        return fail(MPX_ERR_UNSUPPORTED, "the sweep path needs a PLRC solver and the single_periodic boundary");
    if (s->path == MPX_PATH_PASS && !s->plrc)
        return fail(MPX_ERR_UNSUPPORTED, "the pass path needs a PLRC solver");
    for (int n = 0; n < 3; ++n)
        if (s->plrc && !(desc->l_om2[n] > desc->l_delta[n] * desc->l_delta[n]))
            return fail(MPX_ERR_INVALID, "pole " + std::to_string(n + 1) + ": L*Om2 must exceed L*delta^2");

    // storage geometry: rows of (nz + 2) padded so that k = 1 starts a 32-element segment
    const int align = 32;
    GridDims &g = s->g;
    g.n[0] = desc->nx;
    g.n[1] = desc->ny;
    g.n[2] = desc->nz;
    g.py = desc->ny + 2;
    g.z0 = align - 1;
    g.pz = align + ((desc->nz + 1 + align - 1) / align) * align;
    g.sx = (long long)g.py * g.pz;
    s->cells = (size_t)(desc->nx + 2) * (size_t)g.sx;
    // the fused kernels index grids with 32-bit offsets
    if (s->plrc && s->path != MPX_PATH_REFERENCE && s->cells >= (size_t)0x7fffffff)
        return fail(MPX_ERR_UNSUPPORTED, "grid too large for the fused kernels' 32-bit indexing");

    if (hipGetDevice(&s->device) != hipSuccess) return fail(MPX_ERR_HIP, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(MPX_ERR_HIP, "hipStreamCreate failed");

    // Faddeeva node tables: n = (k h)^2, e = exp(-n)
    const double hs[2] = {0.5, 0.53};
    for (int t = 0; t < 2; ++t)
        for (int q = 0; q < 12; ++q) {
            double kh = hs[t] * (double)(q + 1);
            s->ft.n[t][q] = kh * kh;
            s->ft.e[t][q] = std::exp(-s->ft.n[t][q]);
        }

    int rc = MPX_OK;
    for (int c = 0; c < 3; ++c) {
        if (s->esz == 8) {
            s->E[c] = s->dalloc<double>(s->cells, rc);
            s->B[c] = s->dalloc<double>(s->cells, rc);
        } else {
            s->E[c] = s->dalloc<float>(s->cells, rc);
            s->B[c] = s->dalloc<float>(s->cells, rc);
        }
        if (rc) return rc;
    }
    if (s->res) {
        s->Elo = s->dalloc<unsigned>(s->cells, rc);
        s->Blo = s->dalloc<unsigned>(s->cells, rc);
        if (rc) return rc;
    }

    if (!s->plrc) {
        if (s->solver == MPX_SOLVER_DISP)
            for (int n = 0; n < 3; ++n)
                for (int c = 0; c < 3; ++c) {
                    s->P[n][c] = s->dalloc<double>(s->cells, rc);
                    s->Pp[n][c] = s->dalloc<double>(s->cells, rc);
                    if (rc) return rc;
                }
    } else {
        for (int d = 0; d < 3; ++d) {
            s->kE[d].assign(g.n[d] + 2, 1.0);
            s->kH[d].assign(g.n[d] + 2, 1.0);
        }
        // currentFactories->initCurrents in setup-file order, src/fdtd_plrc.t:39-45
        int ncpml = 0;
        for (int fi = 0; fi < desc->n_factories; ++fi) {
            const mpx_factory &f = desc->factories[fi];
            if (f.kind == MPX_FACTORY_CPML || f.kind == MPX_FACTORY_CPML_1D) {
                if (++ncpml > 1) return fail(MPX_ERR_UNSUPPORTED, "at most one CPML factory is supported");
                if (f.d < 1) return fail(MPX_ERR_INVALID, "CPML thickness d must be >= 1");
                bool one_d = f.kind == MPX_FACTORY_CPML_1D;
                const int faces1d[2] = {UP, DOWN};
                int nf = one_d ? 2 : 6;
                for (int h = 0; h < 2; ++h)
                    for (int q = 0; q < nf; ++q) {
                        Current c;
                        c.type = h ? CUR_CPML_H : CUR_CPML_E;
                        c.dir = one_d ? faces1d[q] : kFaceOrder[q];
                        c.factory = fi;
                        (h ? s->hcur : s->ecur).push_back(c);
                    }
                if (one_d)
                    cpml1d_kappa(s, f.d, f.kappa_max);
                else
                    cpml_kappa(s, f.d, f.kappa_max);
            } else if (f.kind == MPX_FACTORY_SHORT_PULSE || f.kind == MPX_FACTORY_PLANE_WAVE ||
                       f.kind == MPX_FACTORY_PLANE_GAUSS || f.kind == MPX_FACTORY_FOCUSED_PULSE ||
                       f.kind == MPX_FACTORY_GAUSS_INJECT) {
                if (ncpml == 0) s->sheets_first = 1;
                // needCurrent(down) only: src/shortpulseinject.cpp:28-31, src/sources.cpp:8-11
                for (int h = 0; h < 2; ++h) {
                    Current c;
                    c.type = h ? CUR_INC_H : CUR_INC_E;
                    c.dir = DOWN;
                    c.factory = fi;
                    (h ? s->hcur : s->ecur).push_back(c);
                }
            } else if (f.kind == MPX_FACTORY_PLASMA_CURRENT) {
                // PlasmaCurrentFactory::initCurrents, src/plasmacurrent.cpp:18-21
                if (s->pcur) return fail(MPX_ERR_UNSUPPORTED, "at most one plasma_current is supported");
                s->pcur = true;
                s->pcur_em = f.em;
                s->pcur_gamma = f.gamma;
                s->pcur_rhobg = f.rho_bg;
                bool seen_cpml = false, seen_sheet = false;
                for (const Current &c : s->ecur) {
                    seen_cpml |= c.type == CUR_CPML_E;
                    seen_sheet |= c.type == CUR_INC_E;
                }
                s->pcur_pos = (int)seen_cpml + (int)seen_sheet;
                Current c;
                c.type = CUR_PLASMA;
                c.dir = -1;
                c.factory = fi;
                s->ecur.push_back(c);
            } else {
                return fail(MPX_ERR_UNSUPPORTED, "unsupported current factory kind " + std::to_string(f.kind));
            }
        }

        // coreInitStorage, src/fdtd_plrc.cpp:13-86: layers, coefficients, invalid currents dropped
        auto init_side = [&](std::vector<Current> &list, bool isH) -> int {
            std::vector<Current> kept;
            for (Current c : list) {
                if (c.type == CUR_PLASMA) {  // a grid current over the storage box, no layer
                    kept.push_back(c);
                    continue;
                }
                const mpx_factory &f = desc->factories[c.factory];
                if (c.type == CUR_CPML_E || c.type == CUR_CPML_H) {
                    int distance = isH ? (is_high(c.dir) ? 1 : 0) : 1;
                    const char *nm[3] = {isH ? "CPMLPsiBx" : "CPMLPsiEx", isH ? "CPMLPsiBy" : "CPMLPsiEy",
                                         isH ? "CPMLPsiBz" : "CPMLPsiEz"};
                    int l0 = add_border_layer(s, nm[0], c.dir, f.d, distance);
                    int l1 = add_border_layer(s, nm[1], c.dir, f.d, distance);
                    int l2 = add_border_layer(s, nm[2], c.dir, f.d, distance);
                    if (l0 < 0 || l1 < 0 || l2 < 0) continue;  // CurrentInvalidPredicate
                    c.layer = l0;
                    int dim = dim_of(c.dir);
                    CpmlFace F{};
                    F.dim = dim;
                    for (int d = 0; d < 3; ++d) {
                        F.lo[d] = s->layers[l0].lo[d];
                        F.hi[d] = s->layers[l0].hi[d];
                    }
                    F.n1 = F.hi[1] - F.lo[1] + 1;
                    F.n2 = F.hi[2] - F.lo[2] + 1;
                    // face map, src/cpml_border.cpp:357-386 / 456-485
                    if (dim == 0) { F.c0 = 1; F.c1 = 2; F.f0 = 1; F.f1 = 2; }
                    else if (dim == 1) { F.c0 = 2; F.c1 = 0; F.f0 = 2; F.f1 = 0; }
                    else { F.c0 = 0; F.c1 = 1; F.f0 = 0; F.f1 = 1; }
                    F.dxn = s->delta[dim];
                    F.idxn = 1.0 / F.dxn;
                    size_t cnt = s->layers[l0].count();
                    int lc[3] = {l0, l1, l2};
                    for (int w = 0; w < 2; ++w) {
                        void *p = (s->esz == 8) ? (void *)s->dalloc<double>(cnt, rc) : (void *)s->dalloc<float>(cnt, rc);
                        if (rc) return rc;
                        int comp = w == 0 ? F.c0 : F.c1;
                        s->layers[lc[comp]].comp[0] = p;  // each component layer owns one array
                        if (w == 0) F.psi0 = p; else F.psi1 = p;
                    }
                    if (s->res) {
                        F.rw = s->dalloc<unsigned>(cnt, rc);
                        if (rc) return rc;
                        s->layers[lc[F.c0]].rw = s->layers[lc[F.c1]].rw = F.rw;
                        s->layers[lc[F.c0]].slot = 0;
                        s->layers[lc[F.c1]].slot = 1;
                    }
                    std::vector<double> bb, cc;
                    cpml_coeff(f, c.dir, isH, F.lo[dim], F.hi[dim], desc->dt, bb, cc);
                    double *db = s->dalloc<double>(bb.size(), rc);
                    double *dc = s->dalloc<double>(cc.size(), rc);
                    if (rc) return rc;
                    if (upload(s, db, bb.data(), bb.size() * 8) || upload(s, dc, cc.data(), cc.size() * 8))
                        return fail(MPX_ERR_HIP, "coefficient upload failed");
                    F.b = db;
                    F.c = dc;
                    std::vector<CpmlFace> &faces = isH ? s->hfaces : s->efaces;
                    c.face = (int)faces.size();
                    faces.push_back(F);
                } else {
                    // IncidentSource{E,H}Current::initStorage, src/incsource.t:15-31, 87-103
                    int distance = f.d;
                    if (isH && !is_high(c.dir)) distance -= 1;
                    const char *nm[3] = {isH ? "IncidentHJx" : "IncidentEJx", isH ? "IncidentHJy" : "IncidentEJy",
                                         isH ? "IncidentHJz" : "IncidentEJz"};
                    int l0 = add_border_layer(s, nm[0], c.dir, 1, distance);
                    int l1 = add_border_layer(s, nm[1], c.dir, 1, distance);
                    int l2 = add_border_layer(s, nm[2], c.dir, 1, distance);
                    if (l0 < 0 || l1 < 0 || l2 < 0) continue;
                    c.layer = l0;
                    int lc[3] = {l0, l1, l2};
                    for (int w = 0; w < 3; ++w)
                        if (!s->layers[lc[w]].comp[0]) {
                            s->layers[lc[w]].comp[0] = s->dalloc<double>(s->layers[lc[w]].count(), rc);
                            if (rc) return rc;
                        }
                    mpx::SheetEval &ev = c.ev;
                    ev.isH = isH ? 1 : 0;
                    // src/incsource.t:54 (E: offset -1) and :126 (H: +1 for east/north/up, else 0)
                    ev.off = isH ? (is_high(c.dir) ? 1 : 0) : -1;
                    ev.t1 = 0;  // down face: transverse1 = x, transverse2 = y, src/incsource.cpp:123-131
                    ev.t2 = 1;
                    ev.DX = desc->dz;
                    ev.nx2 = g.n[0] + 2;
                    ev.ny2 = g.n[1] + 2;
                    ev.sheet.kp = s->layers[l0].lo[2];
                    for (int w = 0; w < 3; ++w) ev.sheet.J[w] = (double *)s->layers[lc[w]].comp[0];
                    if (f.kind == MPX_FACTORY_GAUSS_INJECT) {
                        ev.kind = 4;
                        if (f.gsize_t < 2 || !(f.eps > 0))
                            return fail(MPX_ERR_INVALID, "gauss_inject: SizeT must be >= 2 and eps > 0");
                        c.gauss = std::make_shared<mpx::GaussInjectDev>();
                        int grc = c.gauss->setup(f, isH ? 1 : 0, g.n[0], g.n[1], desc->dx, desc->dy, desc->dt, 0,
                                                 g.n[2] + 1, s->stream);
                        if (grc == 1) return fail(MPX_ERR_NOMEM, "gauss_inject: device memory for the transforms");
                        if (grc) return fail(MPX_ERR_HIP, "gauss_inject: hipFFT plan creation failed");
                        c.gauss->params(ev.gauss);
                    } else if (f.kind == MPX_FACTORY_FOCUSED_PULSE) {
                        ev.kind = 3;
                        if (!(f.eps > 0) || !(f.width > 0) || !(f.length != 0) || !(f.om0 > 0))
                            return fail(MPX_ERR_INVALID, "focused_pulse_inject: eps, width, om0 must be > 0, length != 0");
                        mpx::FocusedParams &F = ev.focused;
                        double cz = 0.5 * (g.n[2] + 1);
                        int kk = f.d - 1;  // the plane both currents sample (down face)
                        F.zE = (kk - cz) * desc->dz - f.zshift;
                        F.zB = (kk - 0.5 - cz) * desc->dz - f.zshift;
                        F.T = f.length;
                        F.om0 = f.om0;
                        F.v = std::sqrt(1 / f.eps);
                        F.TShift = f.tshift;
                        F.phase = 2 * M_PI * f.phase;
                        F.amp = f.amp;
                        F.DT = desc->dt;
                        std::vector<double2> pl;
                        double2 *dp[3];
                        for (int w = 0; w < 3; ++w) {
                            focused_plane(f, g.n[0], g.n[1], desc->dx, desc->dy, w == 0 ? F.zE : F.zB, w,
                                          w == 1 ? 0.0 : 0.5, w == 1 ? 0.5 : 0.0, pl);
                            dp[w] = s->dalloc<double2>(pl.size(), rc);
                            if (rc) return rc;
                            if (upload(s, dp[w], pl.data(), pl.size() * sizeof(double2)))
                                return fail(MPX_ERR_HIP, "focused-pulse plane upload failed");
                        }
                        F.ex = dp[0];
                        F.bx = dp[1];
                        F.by = dp[2];
                    } else if (f.kind == MPX_FACTORY_SHORT_PULSE) {
                        ev.kind = 1;
                        // ShortPulseInjectSourceFunc::setParam, src/shortpulseinject.cpp:56-113
                        mpx::PulseParams &P = ev.pulse;
                        P.DX = desc->dx;
                        P.DY = desc->dy;
                        P.DZ = desc->dz;
                        P.DT = desc->dt;
                        P.centrex = 0.5 * double(g.n[0] + 1 + 0);
                        P.centrey = 0.5 * double(g.n[1] + 1 + 0);
                        P.centrez = 0.5 * double(g.n[2] + 1 + 0);
                        P.length = f.length;
                        P.om0 = f.om0;
                        if (!(f.eps > 0) || !(f.length != 0) || !(f.width > 0))
                            return fail(MPX_ERR_INVALID, "short_pulse_inject: eps, width must be > 0 and length != 0");
                        P.lightspeed = std::sqrt(1 / f.eps);
                        P.ZRl = 0.5 * f.om0 * f.width * f.width / P.lightspeed;
                        P.Phase = 0.0;
                        P.TShift = 0.0;
                        P.ZShift = 0.0;
                        P.amp = 1.0;
                        double Exmax = mpx::pulse_norm_device(P, s->ft, s->stream);
                        P.amp = f.amp / Exmax;
                        P.Phase = 2 * M_PI * f.phase;
                        P.TShift = f.tshift;
                        P.ZShift = f.zshift;
                    } else {
                        ev.kind = 2;
                        // PlaneWaveSource/PlaneGaussSource::make*Current, src/sources.cpp:14-56, 185-223
                        mpx::PlaneParams &W = ev.plane;
                        W.gauss = f.kind == MPX_FACTORY_PLANE_GAUSS;
                        double kx = f.k[0], ky = f.k[1], kz = f.k[2];
                        double Hx = f.h[0], Hy = f.h[1], Hz = f.h[2];
                        double Ev[3] = {ky * Hz - kz * Hy, kz * Hx - kx * Hz, kx * Hy - ky * Hx};
                        double bmag = std::sqrt(Hx * Hx + Hy * Hy + Hz * Hz);
                        double factor = -bmag / std::sqrt(Ev[0] * Ev[0] + Ev[1] * Ev[1] + Ev[2] * Ev[2]);
                        for (int w = 0; w < 3; ++w) {
                            Ev[w] *= factor / std::sqrt(f.eps);
                            W.k[w] = f.k[w];
                            W.H[w] = f.h[w];
                            W.E[w] = Ev[w];
                            W.H_bg[w] = W.gauss ? 0 : f.h_bg[w];
                            W.E_bg[w] = W.gauss ? 0 : f.e_bg[w];
                        }
                        W.ramp = f.ramp;
                        W.dt = desc->dt;
                        W.om = std::sqrt(W.k[0] * W.k[0] + W.k[1] * W.k[1] + W.k[2] * W.k[2]) / std::sqrt(f.eps);
                        W.dx = desc->dx;
                        W.dy = desc->dy;
                        W.dz = desc->dz;
                        W.width = f.gwidth;
                        W.offset = f.goffset;
                    }
                }
                kept.push_back(c);
            }
            list.swap(kept);
            return MPX_OK;
        };
        if ((rc = init_side(s->ecur, false)) != MPX_OK) return rc;
        if ((rc = init_side(s->hcur, true)) != MPX_OK) return rc;
        if ((int)s->efaces.size() > mpx::kMaxFaces || (int)s->hfaces.size() > mpx::kMaxFaces)
            return fail(MPX_ERR_UNSUPPORTED, "too many CPML faces");
        int nse = 0, nsh = 0;
        for (auto &c : s->ecur) nse += c.type == CUR_INC_E;
        for (auto &c : s->hcur) nsh += c.type == CUR_INC_H;
        if (nse > mpx::kMaxSheets || nsh > mpx::kMaxSheets) return fail(MPX_ERR_UNSUPPORTED, "too many source sheets");
        s->haveJ = !s->ecur.empty();
        s->haveM = !s->hcur.empty();

        // the plasma OptField (src/fdtd_plrc.t:431-433, PlasmaDensity) and the grids it shares
        // with the plasma current ("PlasmaDensity", "PlasmaJx".., "Sigma")
        s->plasma = desc->plasma != 0;
        if ((s->plasma || s->pcur) && s->path == MPX_PATH_SWEEP)
            return fail(MPX_ERR_UNSUPPORTED, "plasma needs the fused, pass or reference-structured path");
        if (s->plasma && desc->plasma_k < 1) return fail(MPX_ERR_INVALID, "plasma K must be >= 1");
        if (s->plasma || s->pcur) {
            s->d_rho = s->dalloc<double>(s->cells, rc);
            if (rc) return rc;
        }
        if (s->pcur)
            for (int c = 0; c < 3; ++c) {
                s->d_jp[c] = s->dalloc<double>(s->cells, rc);
                if (rc) return rc;
            }
        if (s->plasma) {
            s->d_sigma = s->dalloc<double>(s->cells, rc);
            if (rc) return rc;
        }

        // kappa*delta lines, as plrcStepD/B form them per cell (src/fdtd_plrc.cpp:313-315, 435-437)
        for (int d = 0; d < 3; ++d) {
            std::vector<double> kdE(g.n[d] + 2), kdH(g.n[d] + 2);
            for (int q = 0; q < g.n[d] + 2; ++q) {
                kdE[q] = s->kE[d][q] * s->delta[d];
                kdH[q] = s->kH[d][q] * s->delta[d];
            }
            std::vector<double> ikdE(kdE.size()), ikdH(kdH.size());
            for (size_t q = 0; q < kdE.size(); ++q) {
                ikdE[q] = 1.0 / kdE[q];
                ikdH[q] = 1.0 / kdH[q];
            }
            s->d_kdE[d] = s->dalloc<double>(kdE.size(), rc);
            s->d_kdH[d] = s->dalloc<double>(kdH.size(), rc);
            s->d_ikdE[d] = s->dalloc<double>(kdE.size(), rc);
            s->d_ikdH[d] = s->dalloc<double>(kdH.size(), rc);
            if (rc) return rc;
            if (upload(s, s->d_kdE[d], kdE.data(), kdE.size() * 8) || upload(s, s->d_kdH[d], kdH.data(), kdH.size() * 8) ||
                upload(s, s->d_ikdE[d], ikdE.data(), kdE.size() * 8) ||
                upload(s, s->d_ikdH[d], ikdH.data(), kdH.size() * 8))
                return fail(MPX_ERR_HIP, "kappa upload failed");
        }

        // active poles: Delta eps = 0 gives gamma = 0, so the recursion stays identically zero
        std::complex<double> Crec[3], dchi0[3], dxi0[3];
        pole_constants(desc->l_eps, desc->l_delta, desc->l_om2, desc->dt, Crec, dchi0, dxi0, s->sumChi0, s->sumXi0);
        s->np = 0;
        for (int n = 0; n < 3; ++n)
            if (desc->l_eps[n] != 0.0) {
                int a = s->np++;
                s->pole_ref[a] = n;
                std::complex<double> D = dchi0[n] - dxi0[n];
                s->pc.dd[a] = make_double2(D.real(), D.imag());
                s->pc.dxi0[a] = make_double2(dxi0[n].real(), dxi0[n].imag());
                s->pc.crec[a] = make_double2(Crec[n].real(), Crec[n].imag());
                double2 *base = s->dalloc<double2>(mpx::psi_block_elems(s->cells), rc);
                if (rc) return rc;
                for (int c = 0; c < 3; ++c) s->psi[a][c] = base + 64 * c;
            }
        // the 64x16 sweep block (at most 128 registers) spills with two or three poles; 64x8 does not
        if (s->np >= 2) s->sweep_variant = 0;
        double sigma = 0.5 * 0.0 * desc->dt;  // no OptField writes Sigma on this path
        if (s->nonlin)
            s->denom = desc->eps + s->sumChi0 - s->sumXi0 + sigma;
        else
            s->denom = desc->eps + (s->sumChi0 - s->sumXi0) + sigma;
        s->numer = desc->eps - s->sumXi0 - sigma;

        if (s->path == MPX_PATH_REFERENCE) {
            for (int c = 0; c < 3; ++c) {
                if (s->haveJ) s->Jt[c] = s->dalloc<double>(s->cells, rc);
                if (s->haveM) s->Mt[c] = s->dalloc<double>(s->cells, rc);
                if (rc) return rc;
            }
        }
        if (s->path == MPX_PATH_SWEEP) {
            for (int c = 0; c < 3; ++c) {
                if (s->esz == 8) {
                    s->E2[c] = s->dalloc<double>(s->cells, rc);
                    s->B2[c] = s->dalloc<double>(s->cells, rc);
                } else {
                    s->E2[c] = s->dalloc<float>(s->cells, rc);
                    s->B2[c] = s->dalloc<float>(s->cells, rc);
                }
                if (rc) return rc;
            }
            for (int a = 0; a < s->np; ++a) {
                double2 *base = s->dalloc<double2>(mpx::psi_block_elems(s->cells), rc);
                if (rc) return rc;
                for (int c = 0; c < 3; ++c) s->psi2[a][c] = base + 64 * c;
            }
            for (size_t f = 0; f < s->efaces.size(); ++f) {
                const CpmlFace &F = s->efaces[f];
                size_t cnt = (size_t)(F.hi[0] - F.lo[0] + 1) * F.n1 * F.n2;
                for (int w = 0; w < 2; ++w) {
                    void *p = (s->esz == 8) ? (void *)s->dalloc<double>(cnt, rc) : (void *)s->dalloc<float>(cnt, rc);
                    if (rc) return rc;
                    s->epsi2[w].push_back(p);
                    int li = -1;
                    for (size_t q = 0; q < s->layers.size(); ++q)
                        if (s->layers[q].comp[0] == (w == 0 ? F.psi0 : F.psi1)) li = (int)q;
                    s->eface_layer[w].push_back(li);
                }
            }
        }
    }
    // the pass path's second set; the sweep path (whose E2 and B2 are already there) takes the
    // H auxiliaries' second set too, so that a batch of sweep-path runs can run the pass
    if (s->path == MPX_PATH_PASS || (s->path == MPX_PATH_SWEEP && s->plrc)) {
        for (int c = 0; c < 3 && s->path == MPX_PATH_PASS; ++c) {
            if (s->esz == 8) {
                s->E2[c] = s->dalloc<double>(s->cells, rc);
                s->B2[c] = s->dalloc<double>(s->cells, rc);
            } else {
                s->E2[c] = s->dalloc<float>(s->cells, rc);
                s->B2[c] = s->dalloc<float>(s->cells, rc);
            }
            if (rc) return rc;
        }
        if (s->res) {
            s->E2lo = s->dalloc<unsigned>(s->cells, rc);
            s->B2lo = s->dalloc<unsigned>(s->cells, rc);
            if (rc) return rc;
        }
        for (size_t f = 0; f < s->hfaces.size(); ++f) {
            const CpmlFace &F = s->hfaces[f];
            const size_t cnt = (size_t)(F.hi[0] - F.lo[0] + 1) * F.n1 * F.n2;
            for (int w = 0; w < 2; ++w) {
                void *p = (s->esz == 8) ? (void *)s->dalloc<double>(cnt, rc) : (void *)s->dalloc<float>(cnt, rc);
                if (rc) return rc;
                s->hpsi2[w].push_back(p);
                int li = -1;
                for (size_t q = 0; q < s->layers.size(); ++q)
                    if (s->layers[q].comp[0] == (w == 0 ? F.psi0 : F.psi1)) li = (int)q;
                s->hface_layer[w].push_back(li);
            }
            unsigned *rw = nullptr;
            if (s->res && !(rw = s->dalloc<unsigned>(cnt, rc))) return rc;
            s->hrw2.push_back(rw);
        }
    }
    s->d_status = s->dalloc<unsigned>(1, rc);
    if (rc) return rc;
    s->d_part = s->dalloc<double>(mpx::energy_partials(g), rc);
    s->d_energy = s->dalloc<double>(1, rc);
    if (rc) return rc;
    MPX_HIP(hipStreamSynchronize(s->stream));
    s->time = 0;
    return MPX_OK;
}

// ---------------------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------------------

// Band height of the banded block order: the largest power of two that keeps a band-plane
// near 512 tiles (E side) or 256 tiles (B side) of 64 z cells; on 512^3 that is 64 and 32
// rows, the fastest of the heights measured (power-of-two bands keep the XCD shares of the
// block order aligned to whole bands).
static int auto_band(const mpx::GridDims &g, const int *blo, bool hside)
{
    const int lo1 = blo ? blo[1] : 1, lo2 = blo ? blo[2] : 1;
    const int nzt = std::max(1, (g.n[2] - lo2 + 1) / 64), nby = (g.n[1] - lo1 + 1 + 3) / 4;
    const int target = std::max(1, (hside ? 256 : 512) / nzt);
    int band = 1;
    while (band * 2 <= target && band * 2 <= nby) band *= 2;
    return band;
}

template <typename T>
static StepArgs<T> step_args(mpx_solver *s, bool hside)
{
    StepArgs<T> a{};
    a.g = s->g;
    for (int d = 0; d < 3; ++d) {
        a.periodic[d] = s->periodic[d];
        a.kdE[d] = s->d_kdE[d];
        a.kdH[d] = s->d_kdH[d];
        a.ikdE[d] = s->d_ikdE[d];
        a.ikdH[d] = s->d_ikdH[d];
        a.E[d] = (T *)s->E[d];
        a.B[d] = (T *)s->B[d];
        a.blo[d] = s->periodic[d] ? 1 : 0;
    }
    a.Elo = s->Elo;
    a.Blo = s->Blo;
    for (int n = 0; n < 3; ++n)
        for (int c = 0; c < 3; ++c) a.psi[n][c] = s->psi[n][c];
    a.dt = s->d.dt;
    a.eps = s->d.eps;
    a.chi = s->d.chi;
    a.denom = s->denom;
    a.idenom = 1.0 / s->denom;
    a.numer = s->numer;
    a.epsf = s->d_eps;
    a.sumChi0 = s->sumChi0;
    a.sumXi0 = s->sumXi0;
    a.pc = s->pc;
    if (!hside) {
        // src/plasmadensity.cpp:30-33, src/plasmacurrent.cpp:57-59
        const double dt = s->d.dt;
        a.plasma = s->plasma ? 1 : 0;
        a.pcur = s->pcur ? 1 : 0;
        a.pcur_pos = s->pcur_pos;
        a.pk = s->d.plasma_k;
        a.pn = 0.5 * dt * s->d.plasma_nu;
        a.pm = dt * s->d.plasma_mpa;
        a.pw = s->d.plasma_mpa * s->d.plasma_wion;
        a.gdtn = 1 - 0.5 * s->pcur_gamma * dt;
        a.gdtd = 1 + 0.5 * s->pcur_gamma * dt;
        a.emdt = s->pcur_em * dt;
        a.rho_bg = s->pcur_rhobg;
        a.rho = s->d_rho;
        a.sigma = s->d_sigma;
        for (int c = 0; c < 3; ++c) a.jp[c] = s->d_jp[c];
    }
    const std::vector<CpmlFace> &faces = hside ? s->hfaces : s->efaces;
    a.nfaces = (int)faces.size();
    for (int d = 0; d < 3; ++d) a.face_by_dim[d][0] = a.face_by_dim[d][1] = -1;
    const std::vector<Current> &list = hside ? s->hcur : s->ecur;
    for (size_t q = 0; q < faces.size(); ++q) a.face[q] = faces[q];
    for (const Current &c : list)
        if (c.face >= 0) {
            int dim = dim_of(c.dir);
            a.face_by_dim[dim][is_high(c.dir) ? 0 : 1] = c.face;
        }
    a.nsheets = 0;
    for (const Current &c : list)
        if (c.type == CUR_INC_E || c.type == CUR_INC_H) a.sheet[a.nsheets++] = c.ev.sheet;
    a.sheets_first = s->sheets_first;
    a.haveJ = 0;
    a.status = s->d_status;
    a.ncap = s->ncap;
    a.march = hside ? s->march_b : s->march_e;
    a.order = s->tile_order;
    a.band = s->tile_band > 0 ? s->tile_band : auto_band(s->g, hside ? a.blo : nullptr, hside);
    a.rev = hside ? s->b_reverse : 0;
    return a;
}

static mpx::SumList sum_list(mpx_solver *s, bool hside)
{
    mpx::SumList L{};
    const std::vector<Current> &list = hside ? s->hcur : s->ecur;
    const std::vector<CpmlFace> &faces = hside ? s->hfaces : s->efaces;
    for (const Current &c : list) {
        mpx::SumView &v = L.v[L.n++];
        if (c.type == CUR_PLASMA) {  // the plasma current's grid spans the storage box
            for (int d = 0; d < 3; ++d) {
                v.lo[d] = 0;
                v.hi[d] = s->g.n[d] + 1;
                v.comp[d] = s->d_jp[d];
            }
            v.n1 = v.n2 = 0;
            v.padded = 1;
            continue;
        }
        const Layer &Ly = s->layers[c.layer];
        for (int d = 0; d < 3; ++d) {
            v.lo[d] = Ly.lo[d];
            v.hi[d] = Ly.hi[d];
            v.comp[d] = nullptr;
        }
        v.n1 = v.hi[1] - v.lo[1] + 1;
        v.n2 = v.hi[2] - v.lo[2] + 1;
        if (c.face >= 0) {
            const CpmlFace &F = faces[c.face];
            v.comp[F.c0] = (const double *)F.psi0;
            v.comp[F.c1] = (const double *)F.psi1;
        } else {
            for (int d = 0; d < 3; ++d) v.comp[d] = c.ev.sheet.J[d];
        }
    }
    return L;
}

// the step index of a sheet evaluation; a GaussInject sheet first runs its block schedule
// (GaussInjectSourceFunc::setTime, called by IncidentSource*Current::stepScheme)
static int sheet_time(mpx_solver *s, Current &c, int t)
{
    c.ev.time = t;
    if (c.gauss) {
        int rc = c.gauss->setTime(t);
        if (rc) return fail(MPX_ERR_HIP, "gauss_inject: transform of a sample block failed");
        c.gauss->params(c.ev.gauss);
        if (c.ev.gauss.active && (t - c.ev.gauss.blockToff < 0 || t - c.ev.gauss.blockToff >= c.gauss->Nt))
            return fail(MPX_ERR_INVALID, "gauss_inject: step outside the sample block");
    }
    return MPX_OK;
}
static int sheet_time(mpx_solver *s, Current &c) { return sheet_time(s, c, s->time); }

// the sheets of one side at step t (the step counter of the step they belong to)
static int eval_sheets(mpx_solver *s, std::vector<Current> &list, int type, int t)
{
    for (Current &c : list)
        if (c.type == type) {
            int rc = sheet_time(s, c, t);
            if (rc) return rc;
            s->tick(K_SHEET);
            mpx::launch_sheet_eval(c.ev, s->ft, s->stream);
            s->tock();
            MPX_LAUNCHED("sheet evaluation");
        }
    return MPX_OK;
}

template <typename T>
static int wrap_group(mpx_solver *s, void *const f[3], int cls)
{
    T *ff[3] = {(T *)f[0], (T *)f[1], (T *)f[2]};
    for (int dim = 0; dim < 3; ++dim) {
        if (!s->periodic[dim]) continue;  // SingleXYPeriodicBoundary::exchangeZ is empty
        s->tick(cls);
        mpx::launch_wrap<T>(s->g, ff, dim, s->stream);
        s->tock();
        MPX_LAUNCHED("ghost wrap");
    }
    return MPX_OK;
}

static int record_probes(mpx_solver *s)
{
    if (s->nprobe == 0 || s->probe_rec >= s->probe_cap) return MPX_OK;
    double *row = s->d_probe_hist + (size_t)s->probe_rec * s->nprobe;
    s->tick(K_PROBE);
    if (s->esz == 8) {
        mpx::ProbeArgs<double> p{};
        p.n = s->nprobe;
        for (int c = 0; c < 3; ++c) {
            p.f[c] = (const double *)s->E[c];
            p.f[3 + c] = (const double *)s->B[c];
        }
        p.comp = s->d_probe_comp;
        p.idx = s->d_probe_idx;
        mpx::launch_probe<double>(p, row, s->stream);
    } else {
        mpx::ProbeArgs<float> p{};
        p.n = s->nprobe;
        for (int c = 0; c < 3; ++c) {
            p.f[c] = (const float *)s->E[c];
            p.f[3 + c] = (const float *)s->B[c];
        }
        p.lo[0] = s->Elo;
        p.lo[1] = s->Blo;
        p.comp = s->d_probe_comp;
        p.idx = s->d_probe_idx;
        mpx::launch_probe<float>(p, row, s->stream);
    }
    s->tock();
    MPX_LAUNCHED("probe");
    s->probe_rec++;
    return MPX_OK;
}

static mpx::SimpleArgs simple_args(mpx_solver *s)
{
    mpx::SimpleArgs a{};
    a.g = s->g;
    a.kind = s->solver == MPX_SOLVER_PLAIN ? 0 : (s->solver == MPX_SOLVER_NL ? 1 : 2);
    for (int c = 0; c < 3; ++c) {
        a.E[c] = (double *)s->E[c];
        a.B[c] = (double *)s->B[c];
        for (int n = 0; n < 3; ++n) {
            a.P[n][c] = s->P[n][c];
            a.Pp[n][c] = s->Pp[n][c];
        }
    }
    a.dx = s->d.dx;
    a.dy = s->d.dy;
    a.dz = s->d.dz;
    a.dt = s->d.dt;
    a.eps = s->d.eps;
    a.chi = s->d.chi;
    for (int n = 0; n < 3; ++n) {
        a.eps2[n] = s->d.d_eps2[n];
        a.om2[n] = s->d.d_om2[n];
    }
    a.status = s->d_status;
    a.ncap = s->ncap;
    return a;
}

template <typename T>
static mpx::SweepArgs<T> sweep_args(mpx_solver *s)
{
    mpx::SweepArgs<T> a{};
    a.g = s->g;
    for (int c = 0; c < 3; ++c) {
        a.Eo[c] = (const T *)s->E[c];
        a.En[c] = (T *)s->E2[c];
        a.Bo[c] = (const T *)s->B[c];
        a.Bn[c] = (T *)s->B2[c];
        a.kdE[c] = s->d_kdE[c];
        a.kdH[c] = s->d_kdH[c];
        a.ikdE[c] = s->d_ikdE[c];
        a.ikdH[c] = s->d_ikdH[c];
        for (int n = 0; n < 3; ++n) {
            a.psio[n][c] = s->psi[n][c];
            a.psin[n][c] = s->psi2[n][c];
        }
    }
    a.dt = s->d.dt;
    a.chi = s->d.chi;
    a.denom = s->denom;
    a.idenom = 1.0 / s->denom;
    a.numer = s->numer;
    a.pc = s->pc;
    a.nefaces = (int)s->efaces.size();
    a.nhfaces = (int)s->hfaces.size();
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < 2; ++q) a.efbd[d][q] = a.hfbd[d][q] = -1;
    for (size_t f = 0; f < s->efaces.size(); ++f) {
        a.eface[f] = s->efaces[f];
        a.eface[f].psi0 = s->epsi2[0][f];
        a.eface[f].psi1 = s->epsi2[1][f];
        a.epsi_old[f][0] = s->efaces[f].psi0;
        a.epsi_old[f][1] = s->efaces[f].psi1;
    }
    for (size_t f = 0; f < s->hfaces.size(); ++f) a.hface[f] = s->hfaces[f];
    for (const Current &c : s->ecur)
        if (c.face >= 0) a.efbd[dim_of(c.dir)][is_high(c.dir) ? 0 : 1] = c.face;
    for (const Current &c : s->hcur)
        if (c.face >= 0) a.hfbd[dim_of(c.dir)][is_high(c.dir) ? 0 : 1] = c.face;
    for (const Current &c : s->ecur)
        if (c.type == CUR_INC_E) a.sheet[a.nsheets++] = c.ev.sheet;
    for (const Current &c : s->hcur)
        if (c.type == CUR_INC_H) a.hsheet[a.nhsheets++] = c.ev.sheet;
    a.sheets_first = s->sheets_first;
    a.seg = s->sweep_seg;
    a.variant = s->sweep_variant;
    a.status = s->d_status;
    a.ncap = s->ncap;
    return a;
}

// after a sweep step the "new" buffers hold the state: exchange the two sets
static void swap_sweep_buffers(mpx_solver *s)
{
    for (int c = 0; c < 3; ++c) {
        std::swap(s->E[c], s->E2[c]);
        std::swap(s->B[c], s->B2[c]);
        for (int n = 0; n < 3; ++n) std::swap(s->psi[n][c], s->psi2[n][c]);
    }
    for (size_t f = 0; f < s->efaces.size(); ++f) {
        std::swap(s->efaces[f].psi0, s->epsi2[0][f]);
        std::swap(s->efaces[f].psi1, s->epsi2[1][f]);
        if (s->eface_layer[0][f] >= 0) s->layers[s->eface_layer[0][f]].comp[0] = s->efaces[f].psi0;
        if (s->eface_layer[1][f] >= 0) s->layers[s->eface_layer[1][f]].comp[0] = s->efaces[f].psi1;
    }
}

// device copy of a fused kernel's argument block; uploaded again only when it changed
template <typename T>
static const StepArgs<T> *dev_args(mpx_solver *s, const StepArgs<T> &a, int side)
{
    side += 2 * s->parity;
    const unsigned char *p = (const unsigned char *)&a;
    std::vector<unsigned char> &h = s->h_fargs[side];
    if (h.size() == sizeof(a) && std::memcmp(h.data(), p, sizeof(a)) == 0) return (const StepArgs<T> *)s->d_fargs[side];
    if (!s->d_fargs[side]) {
        if (hipMalloc(&s->d_fargs[side], sizeof(StepArgs<double>)) != hipSuccess) {
            s->d_fargs[side] = nullptr;
            return nullptr;
        }
        s->allocs.push_back(s->d_fargs[side]);
    }
    h.assign(p, p + sizeof(a));
    if (hipMemcpyAsync(s->d_fargs[side], h.data(), sizeof(a), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return nullptr;
    return (const StepArgs<T> *)s->d_fargs[side];
}

// The E half (E-side sheets at step t, fused E kernel) or the B half (H-side sheets at step t,
// fused B kernel) of step t on the fused path.
template <typename T>
static int fused_half_t(mpx_solver *s, bool hside)
{
    StepArgs<T> a = step_args<T>(s, hside);
    const StepArgs<T> *d = dev_args(s, a, hside ? 1 : 0);
    if (!d) return fail(MPX_ERR_HIP, "argument upload failed");
    s->tick(hside ? K_FUSED_B : K_FUSED_E);
    if (hside)
        mpx::launch_fused_B<T>(a, d, s->stream);
    else
        mpx::launch_fused_E<T>(a, d, s->np, s->nonlin, s->stream);
    s->tock();
    MPX_LAUNCHED(hside ? "fused B" : "fused E");
    return MPX_OK;
}

static int fused_half(mpx_solver *s, bool hside, int t)
{
    int rc = eval_sheets(s, hside ? s->hcur : s->ecur, hside ? CUR_INC_H : CUR_INC_E, t);
    if (rc) return rc;
    return s->esz == 8 ? fused_half_t<double>(s, hside) : fused_half_t<float>(s, hside);
}

template <typename T>
static mpx::PassArgs<T> pass_args(mpx_solver *s)
{
    mpx::PassArgs<T> a{};
    a.g = s->g;
    for (int c = 0; c < 3; ++c) {
        a.Eo[c] = (const T *)s->E[c];
        a.En[c] = (T *)s->E2[c];
        a.Bo[c] = (const T *)s->B[c];
        a.Bn[c] = (T *)s->B2[c];
        a.ikdE[c] = s->d_ikdE[c];
        a.ikdH[c] = s->d_ikdH[c];
        for (int n = 0; n < 3; ++n) a.psi[n][c] = s->psi[n][c];
    }
    a.Eolo = s->Elo;
    a.Bolo = s->Blo;
    a.Enlo = s->E2lo;
    a.Bnlo = s->B2lo;
    a.dt = s->d.dt;
    a.eps = s->d.eps;
    a.chi = s->d.chi;
    a.denom = s->denom;
    a.numer = s->numer;
    a.idenom = 1.0 / s->denom;
    a.sumChi0 = s->sumChi0;
    a.sumXi0 = s->sumXi0;
    a.epsf = s->d_eps;
    {  // src/plasmadensity.cpp:30-33, src/plasmacurrent.cpp:57-59 (the fused E kernel's values)
        const double dt = s->d.dt;
        a.plasma = s->plasma ? 1 : 0;
        a.pcur = s->pcur ? 1 : 0;
        a.pcur_pos = s->pcur_pos;
        a.pk = s->d.plasma_k;
        a.pn = 0.5 * dt * s->d.plasma_nu;
        a.pm = dt * s->d.plasma_mpa;
        a.pw = s->d.plasma_mpa * s->d.plasma_wion;
        a.gdtn = 1 - 0.5 * s->pcur_gamma * dt;
        a.gdtd = 1 + 0.5 * s->pcur_gamma * dt;
        a.emdt = s->pcur_em * dt;
        a.rho_bg = s->pcur_rhobg;
        a.rho = s->d_rho;
        a.sigma = s->d_sigma;
        for (int c = 0; c < 3; ++c) a.jp[c] = s->d_jp[c];
    }
    a.pc = s->pc;
    a.nefaces = (int)s->efaces.size();
    a.nhfaces = (int)s->hfaces.size();
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < 2; ++q) a.efbd[d][q] = a.hfbd[d][q] = -1;
    for (size_t f = 0; f < s->efaces.size(); ++f) a.eface[f] = s->efaces[f];
    for (size_t f = 0; f < s->hfaces.size(); ++f) {
        a.hface[f] = s->hfaces[f];
        a.hface[f].psi0 = s->hpsi2[0][f];
        a.hface[f].psi1 = s->hpsi2[1][f];
        a.hface[f].rw = s->hrw2[f];
        a.hold[f][0] = s->hfaces[f].psi0;
        a.hold[f][1] = s->hfaces[f].psi1;
        a.hold_rw[f] = s->hfaces[f].rw;
    }
    for (const Current &c : s->ecur)
        if (c.face >= 0) a.efbd[dim_of(c.dir)][is_high(c.dir) ? 0 : 1] = c.face;
    for (const Current &c : s->hcur)
        if (c.face >= 0) a.hfbd[dim_of(c.dir)][is_high(c.dir) ? 0 : 1] = c.face;
    for (const Current &c : s->ecur)
        if (c.type == CUR_INC_E) a.sheet[a.nsheets++] = c.ev.sheet;
    for (const Current &c : s->hcur)
        if (c.type == CUR_INC_H) a.hsheet[a.nhsheets++] = c.ev.sheet;
    a.sheets_first = s->sheets_first;
    a.band = s->tile_band > 0 ? s->tile_band : auto_band(s->g, nullptr, false);
    a.status = s->d_status;
    a.ncap = s->ncap;
    a.kernel = s->pass_kernel;
    int seg = s->march_seg;
    if (seg <= 0) {
        // 16 planes per block; shorter segments (down to 2) while fewer than 2048 blocks (eight
        // per CU) would cover the grid with the 3-row tiles of the smallest blocks
        const long long tiles = (long long)((s->g.n[2] + 63) / 64) * ((s->g.n[1] + 2) / 3);
        seg = 16;
        while (seg > 2 && tiles * ((s->g.n[0] + seg - 1) / seg) < 2048) seg /= 2;
    }
    a.seg = seg < s->g.n[0] ? seg : s->g.n[0];
    a.rows = s->march_rows;
    a.zper = s->periodic[2];
    return a;
}

// after a pass the new set holds the state: exchange the two sets
static void swap_pass_buffers(mpx_solver *s)
{
    for (int c = 0; c < 3; ++c) {
        std::swap(s->E[c], s->E2[c]);
        std::swap(s->B[c], s->B2[c]);
    }
    std::swap(s->Elo, s->E2lo);
    std::swap(s->Blo, s->B2lo);
    for (size_t f = 0; f < s->hfaces.size(); ++f) {
        std::swap(s->hfaces[f].psi0, s->hpsi2[0][f]);
        std::swap(s->hfaces[f].psi1, s->hpsi2[1][f]);
        std::swap(s->hfaces[f].rw, s->hrw2[f]);
        for (int w = 0; w < 2; ++w)
            if (s->hface_layer[w][f] >= 0) {
                Layer &L = s->layers[s->hface_layer[w][f]];
                L.comp[0] = w == 0 ? s->hfaces[f].psi0 : s->hfaces[f].psi1;
                L.rw = s->hfaces[f].rw;
            }
    }
    s->parity ^= 1;
}

template <typename T>
static int pass_launch_t(mpx_solver *s)
{
    const mpx::PassArgs<T> a = pass_args<T>(s);
    if (a.kernel == 0 && (!a.zper || a.plasma || a.pcur))
        return fail(MPX_ERR_UNSUPPORTED, "the one-cell pass kernel (pass_kernel 0) takes neither plasma nor the "
                                         "single_xy_periodic boundary");
    // device copy per buffer parity, uploaded again only when it changed
    const int side = s->parity;
    const unsigned char *p = (const unsigned char *)&a;
    std::vector<unsigned char> &h = s->h_pargs[side];
    if (!(h.size() == sizeof(a) && std::memcmp(h.data(), p, sizeof(a)) == 0)) {
        if (!s->d_pargs[side]) {
            if (hipMalloc(&s->d_pargs[side], sizeof(mpx::PassArgs<double>)) != hipSuccess) {
                s->d_pargs[side] = nullptr;
                return fail(MPX_ERR_NOMEM, "argument block allocation failed");
            }
            s->allocs.push_back(s->d_pargs[side]);
        }
        h.assign(p, p + sizeof(a));
        if (hipMemcpyAsync(s->d_pargs[side], h.data(), sizeof(a), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipStreamSynchronize(s->stream) != hipSuccess)
            return fail(MPX_ERR_HIP, "argument upload failed");
    }
    s->tick(K_PASS);
    mpx::launch_pass<T>(a, (const mpx::PassArgs<T> *)s->d_pargs[side], s->np, s->nonlin, s->stream);
    s->tock();
    MPX_LAUNCHED("pass");
    return MPX_OK;
}

// One step on the pass path.  Settled state at time t (E^t, B^t): the E half of step t.
// Pending state at time t (E^t, B^(t-1)): one pass, the B half of step t - 1 (H-side sheets
// at t - 1) and the E half of step t (E-side sheets at t).  Either way the state is pending at
// time t + 1 afterwards.
static int pass_step(mpx_solver *s)
{
    int rc;
    if (!s->pending) {
        if ((rc = fused_half(s, false, s->time))) return rc;
    } else {
        if ((rc = eval_sheets(s, s->hcur, CUR_INC_H, s->time - 1))) return rc;
        if ((rc = eval_sheets(s, s->ecur, CUR_INC_E, s->time))) return rc;
        rc = s->esz == 8 ? pass_launch_t<double>(s) : pass_launch_t<float>(s);
        if (rc) return rc;
        swap_pass_buffers(s);
    }
    s->pending = true;
    s->time++;
    return MPX_OK;
}

// Completes a pending B half (pass path) so that E and B belong to the same step.  Every
// entry point that reads or writes the state calls it first.
static int settle(mpx_solver *s)
{
    if (!s->pending) return MPX_OK;
    int rc = fused_half(s, true, s->time - 1);
    if (rc) return rc;
    s->pending = false;
    return MPX_OK;
}

// one full step: FDTD_PLRCSolver::stepScheme (src/fdtd_plrc.t:127-168) or the plain/nl/disp
// stepScheme (src/fdtd_plain.cpp:31-35)
static int step_once(mpx_solver *s)
{
    int rc;
    if (!s->plrc) {
        mpx::SimpleArgs a = simple_args(s);
        s->tick(K_SIMPLE_E);
        mpx::launch_simple_stepD(a, s->stream);
        s->tock();
        MPX_LAUNCHED("simple stepD");
        if ((rc = wrap_group<double>(s, s->E, K_WRAP_E))) return rc;
        s->tick(K_SIMPLE_B);
        mpx::launch_simple_stepB(a, s->stream);
        s->tock();
        MPX_LAUNCHED("simple stepB");
        if ((rc = wrap_group<double>(s, s->B, K_WRAP_B))) return rc;
    } else if (s->path == MPX_PATH_REFERENCE) {
        StepArgs<double> ae = step_args<double>(s, false);
        // Sigma := 0 and the E-side OptFields, src/fdtd_plrc.t:129-135
        if (s->plasma) {
            mpx::launch_ref_plasma_density(ae, s->stream);
            MPX_LAUNCHED("plasma density");
        }
        // E currents in list order
        for (Current &c : s->ecur) {
            if (c.type == CUR_CPML_E) {
                s->tick(K_CPML_E);
                mpx::launch_ref_cpml(ae, c.face, false, s->stream);
                s->tock();
                MPX_LAUNCHED("cpml E");
            } else if (c.type == CUR_PLASMA) {
                mpx::launch_ref_plasma_current(ae, s->stream);
                MPX_LAUNCHED("plasma current");
            } else {
                if ((rc = sheet_time(s, c))) return rc;
                s->tick(K_SHEET);
                mpx::launch_sheet_eval(c.ev, s->ft, s->stream);
                s->tock();
                MPX_LAUNCHED("sheet E");
            }
        }
        if (s->haveJ) {
            mpx::SumList L = sum_list(s, false);
            s->tick(K_SUM_J);
            mpx::launch_ref_sum(s->g, L, s->Jt, s->stream);
            s->tock();
            MPX_LAUNCHED("sum J");
            ae.haveJ = 1;
            for (int c = 0; c < 3; ++c) ae.Jt[c] = s->Jt[c];
        }
        s->tick(K_STEPD_REF);
        mpx::launch_ref_stepD(ae, s->np, s->nonlin, s->stream);
        s->tock();
        MPX_LAUNCHED("stepD");
        if ((rc = wrap_group<double>(s, s->E, K_WRAP_E))) return rc;
        StepArgs<double> ah = step_args<double>(s, true);
        for (Current &c : s->hcur) {
            if (c.type == CUR_CPML_H) {
                s->tick(K_CPML_H);
                mpx::launch_ref_cpml(ah, c.face, true, s->stream);
                s->tock();
                MPX_LAUNCHED("cpml H");
            } else {
                if ((rc = sheet_time(s, c))) return rc;
                s->tick(K_SHEET);
                mpx::launch_sheet_eval(c.ev, s->ft, s->stream);
                s->tock();
                MPX_LAUNCHED("sheet H");
            }
        }
        if (s->haveM) {
            mpx::SumList L = sum_list(s, true);
            s->tick(K_SUM_M);
            mpx::launch_ref_sum(s->g, L, s->Mt, s->stream);
            s->tock();
            MPX_LAUNCHED("sum M");
            ah.haveJ = 1;
            for (int c = 0; c < 3; ++c) ah.Jt[c] = s->Mt[c];
        }
        s->tick(K_STEPB_REF);
        mpx::launch_ref_stepB(ah, s->stream);
        s->tock();
        MPX_LAUNCHED("stepB");
        if ((rc = wrap_group<double>(s, s->B, K_WRAP_B))) return rc;
    } else if (s->path == MPX_PATH_SWEEP) {
        if ((rc = eval_sheets(s, s->ecur, CUR_INC_E, s->time))) return rc;
        if ((rc = eval_sheets(s, s->hcur, CUR_INC_H, s->time))) return rc;
        s->tick(K_SWEEP);
        if (s->esz == 8)
            mpx::launch_sweep<double>(sweep_args<double>(s), s->np, s->nonlin, s->stream);
        else
            mpx::launch_sweep<float>(sweep_args<float>(s), s->np, s->nonlin, s->stream);
        s->tock();
        MPX_LAUNCHED("sweep");
        swap_sweep_buffers(s);
    } else {
        if ((rc = fused_half(s, false, s->time))) return rc;
        if ((rc = fused_half(s, true, s->time))) return rc;
    }
    s->time++;
    return record_probes(s);
}

// ghost planes of E and B as the reference stores them (the fused path wraps by index)
static int materialise_ghosts(mpx_solver *s)
{
    if (s->plrc && s->path != MPX_PATH_REFERENCE) {
        int rc;
        if (s->esz == 8) {
            if ((rc = wrap_group<double>(s, s->E, K_WRAP_E))) return rc;
            if ((rc = wrap_group<double>(s, s->B, K_WRAP_B))) return rc;
        } else {
            if ((rc = wrap_group<float>(s, s->E, K_WRAP_E))) return rc;
            if ((rc = wrap_group<float>(s, s->B, K_WRAP_B))) return rc;
        }
        if (s->res) {
            void *lo[2][3] = {{s->Elo, nullptr, nullptr}, {s->Blo, nullptr, nullptr}};
            if ((rc = wrap_group<unsigned>(s, lo[0], K_WRAP_E))) return rc;
            if ((rc = wrap_group<unsigned>(s, lo[1], K_WRAP_B))) return rc;
        }
    }
    return MPX_OK;
}

extern "C" {

int mpx_step(mpx_solver *s, int nsteps)
{
    if (!s || nsteps < 0) return fail(MPX_ERR_INVALID, "invalid solver or step count");
    MPX_NOT_IN_BATCH(s);
    // probes sample E and B after every step: on the pass path they take the fused halves
    const bool pass = s->path == MPX_PATH_PASS && s->nprobe == 0;
    if (!pass) {
        int rc = settle(s);
        if (rc) return rc;
    }
    for (int n = 0; n < nsteps; ++n) {
        int rc = pass ? pass_step(s) : step_once(s);
        if (rc) return rc;
    }
    return MPX_OK;
}

// FieldSolver::stepSchemeInit: src/fdtd_plrc.t:90-124 (PLRC), src/fdtd_plain.cpp:26-29 (others)
int mpx_step_init(mpx_solver *s)
{
    if (!s) return fail(MPX_ERR_INVALID, "null solver");
    MPX_NOT_IN_BATCH(s);
    if (s->esz != 8) return fail(MPX_ERR_UNSUPPORTED, "stepSchemeInit is provided in float64 only");
    int rc;
    if ((rc = settle(s))) return rc;
    if (!s->plrc) {
        mpx::SimpleArgs a = simple_args(s);
        a.dt = 0.5 * s->d.dt;
        mpx::launch_simple_stepB(a, s->stream);
        MPX_LAUNCHED("simple stepB");
        return wrap_group<double>(s, s->B, K_WRAP_B);
    }
    if ((rc = materialise_ghosts(s))) return rc;
    StepArgs<double> ae = step_args<double>(s, false);
    mpx::launch_init_accumulator(ae, s->np, s->stream);
    MPX_LAUNCHED("initAccumulator");
    // stepB(0.5*dt) with the magnetic currents as they stand
    StepArgs<double> ah = step_args<double>(s, true);
    ah.dt = 0.5 * s->d.dt;
    if (s->haveM) {
        for (int c = 0; c < 3; ++c)
            if (!s->Mt[c]) {
                s->Mt[c] = s->dalloc<double>(s->cells, rc);
                if (rc) return rc;
            }
        mpx::SumList L = sum_list(s, true);
        mpx::launch_ref_sum(s->g, L, s->Mt, s->stream);
        MPX_LAUNCHED("sum M");
        ah.haveJ = 1;
        for (int c = 0; c < 3; ++c) ah.Jt[c] = s->Mt[c];
    }
    mpx::launch_ref_stepB(ah, s->stream);
    MPX_LAUNCHED("stepB");
    if ((rc = wrap_group<double>(s, s->B, K_WRAP_B))) return rc;
    // currents' stepSchemeInit: CPML E is empty, incident E/H and CPML H run one stepScheme
    if ((rc = eval_sheets(s, s->ecur, CUR_INC_E, s->time))) return rc;
    for (Current &c : s->hcur) {
        if (c.type == CUR_CPML_H) {
            mpx::launch_ref_cpml(ah, c.face, true, s->stream);
            MPX_LAUNCHED("cpml H");
        } else {
            if ((rc = sheet_time(s, c))) return rc;
            mpx::launch_sheet_eval(c.ev, s->ft, s->stream);
            MPX_LAUNCHED("sheet H");
        }
    }
    return MPX_OK;
}

int mpx_get_time(mpx_solver *s, int *t)
{
    if (!s || !t) return fail(MPX_ERR_INVALID, "null argument");
    *t = s->time;
    return MPX_OK;
}

int mpx_set_time(mpx_solver *s, int t)
{
    if (!s) return fail(MPX_ERR_INVALID, "null solver");
    MPX_NOT_IN_BATCH(s);
    if (t == s->time) return MPX_OK;  // nothing changes: a pending pass-path half step stays pending
    if (int rs = settle(s)) return rs;
    s->time = t;
    return MPX_OK;
}

// the device status since the last check (cleared): MPX_ERR_NUMERIC when a bit is set
static int check_status(mpx_solver *s)
{
    unsigned st = 0;
    MPX_HIP(hipMemcpyAsync(&st, s->d_status, sizeof st, hipMemcpyDeviceToHost, s->stream));
    MPX_HIP(hipStreamSynchronize(s->stream));
    if (!st) return MPX_OK;
    const unsigned zero = 0;
    MPX_HIP(hipMemcpyAsync(s->d_status, &zero, sizeof zero, hipMemcpyHostToDevice, s->stream));
    MPX_HIP(hipStreamSynchronize(s->stream));
    std::string what;
    if (st & 1u) what = "a Kerr Newton solve stopped at the iteration cap (" + std::to_string(s->ncap) + ")";
    if (st & 2u) what += std::string(what.empty() ? "" : "; ") + "a non-finite value in an E update";
    return fail(MPX_ERR_NUMERIC, what + " on the device before step " + std::to_string(s->time));
}

int mpx_sync(mpx_solver *s)
{
    if (!s) return fail(MPX_ERR_INVALID, "null solver");
    MPX_OWNER_SYNC(s);
    MPX_HIP(hipStreamSynchronize(s->stream));
    return check_status(s);
}

}  // extern "C"

// the residual words and slot of grid ptr when it is stored as float32 + residual (E and B)
static const unsigned *res_of(const mpx_solver *s, const void *ptr, int *slot)
{
    if (!s->res || !ptr) return nullptr;
    for (int q = 0; q < 3; ++q) {
        *slot = q;
        if (ptr == s->E[q]) return s->Elo;
        if (ptr == s->B[q]) return s->Blo;
    }
    return nullptr;
}

// host copy of a device array of `n` residual words
static int fetch_words(mpx_solver *s, const unsigned *d, size_t n, std::vector<unsigned> &out)
{
    out.resize(n);
    if (hipMemcpyAsync(out.data(), d, n * sizeof(unsigned), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return fail(MPX_ERR_HIP, "copy of residual words failed");
    return MPX_OK;
}

// resolve a full-storage grid name; *elem = element size; nullptr + MPX_OK means "all zero"
static int find_grid(mpx_solver *s, const char *name, void **ptr, size_t *elem, bool *isPsi, int *psiPart)
{
    *ptr = nullptr;
    *elem = s->esz;
    *isPsi = false;
    static const char *eb[6] = {"Ex", "Ey", "Ez", "Bx", "By", "Bz"};
    for (int q = 0; q < 6; ++q)
        if (!std::strcmp(name, eb[q])) {
            *ptr = q < 3 ? s->E[q] : s->B[q - 3];
            return MPX_OK;
        }
    const char comp[3] = {'x', 'y', 'z'};
    for (int n = 0; n < 3; ++n)
        for (int c = 0; c < 3; ++c) {
            char nr[16], ni[16], np_[16], npp[16];
            std::snprintf(nr, sizeof nr, "PsiR%d%c", n + 1, comp[c]);
            std::snprintf(ni, sizeof ni, "PsiI%d%c", n + 1, comp[c]);
            std::snprintf(np_, sizeof np_, "P%d%c", n + 1, comp[c]);
            std::snprintf(npp, sizeof npp, "P%d%cp", n + 1, comp[c]);
            bool r = !std::strcmp(name, nr), i = !std::strcmp(name, ni);
            if (r || i) {
                if (!s->plrc) return MPX_ERR_NAME;
                *isPsi = true;
                *psiPart = r ? 0 : 1;
                *elem = 8;
                for (int a = 0; a < s->np; ++a)
                    if (s->pole_ref[a] == n) *ptr = s->psi[a][c];
                return MPX_OK;
            }
            if (!std::strcmp(name, np_) || !std::strcmp(name, npp)) {
                if (s->solver != MPX_SOLVER_DISP) return MPX_ERR_NAME;
                *elem = 8;
                *ptr = !std::strcmp(name, np_) ? s->P[n][c] : s->Pp[n][c];
                return MPX_OK;
            }
        }
    static const char *jm[6] = {"JxTotal", "JyTotal", "JzTotal", "MxTotal", "MyTotal", "MzTotal"};
    for (int q = 0; q < 6; ++q)
        if (!std::strcmp(name, jm[q])) {
            double *p = q < 3 ? s->Jt[q] : s->Mt[q - 3];
            if (!p) return MPX_ERR_NAME;
            *elem = 8;
            *ptr = p;
            return MPX_OK;
        }
    if (!std::strcmp(name, "Sigma") && s->plrc) {
        *elem = 8;
        *ptr = s->d_sigma;  // stored with the plasma OptField; otherwise identically zero
        return MPX_OK;
    }
    if (!std::strcmp(name, "PlasmaDensity") && s->d_rho) {
        *elem = 8;
        *ptr = s->d_rho;
        return MPX_OK;
    }
    static const char *pj[3] = {"PlasmaJx", "PlasmaJy", "PlasmaJz"};
    for (int q = 0; q < 3; ++q)
        if (!std::strcmp(name, pj[q]) && s->d_jp[q]) {
            *elem = 8;
            *ptr = s->d_jp[q];
            return MPX_OK;
        }
    if (!std::strcmp(name, "Eps") && s->plrc) {
        *elem = 8;
        *ptr = s->d_eps;  // nullptr: the uniform eps of the descriptor
        return MPX_OK;
    }
    return MPX_ERR_NAME;
}

extern "C" {

}  // extern "C"

// Ghost values of the plasma grids on the fused path, which updates interior cells only.
// The reference updates the density and Sigma over the whole storage box from ghost E values
// that are periodic images, so their ghosts are the images too; its plasma current loop stops
// at N (src/plasmacurrent.cpp:61-63), so the high ghosts stay zero.  Along a non-periodic z
// the ghost E is zero: the density and current stay zero there and Sigma is w 0^(K-1).
static void plasma_ghosts(const mpx_solver *s, double *h, int kind /* 0 density, 1 Sigma, 2 current */)
{
    const int n[3] = {s->g.n[0], s->g.n[1], s->g.n[2]};
    const size_t e[3] = {(size_t)n[0] + 2, (size_t)n[1] + 2, (size_t)n[2] + 2};
    auto at = [&](size_t i, size_t j, size_t k) -> double & { return h[(i * e[1] + j) * e[2] + k]; };
    const double zero_sigma = s->d.plasma_k == 1 ? s->d.plasma_mpa * s->d.plasma_wion : 0.0;
    for (int d = 0; d < 3; ++d) {
        for (size_t a = 0; a < e[(d + 1) % 3]; ++a)
            for (size_t b = 0; b < e[(d + 2) % 3]; ++b) {
                size_t idx[3];
                auto cell = [&](size_t v) -> double & {
                    idx[d] = v;
                    idx[(d + 1) % 3] = a;
                    idx[(d + 2) % 3] = b;
                    return at(idx[0], idx[1], idx[2]);
                };
                if (s->periodic[d]) {
                    cell(0) = cell((size_t)n[d]);
                    cell((size_t)n[d] + 1) = kind == 2 ? 0.0 : cell(1);
                } else {
                    cell(0) = kind == 1 ? zero_sigma : 0.0;
                    cell((size_t)n[d] + 1) = kind == 1 ? zero_sigma : 0.0;
                }
            }
    }
}

extern "C" {

int mpx_get_field(mpx_solver *s, const char *name, double *host, size_t count)
{
    if (!s || !name || !host) return fail(MPX_ERR_INVALID, "null argument");
    MPX_OWNER_SYNC(s);
    if (int rs = settle(s)) return rs;
    const GridDims &g = s->g;
    size_t nx2 = g.n[0] + 2, ny2 = g.n[1] + 2, nz2 = g.n[2] + 2;
    if (count != nx2 * ny2 * nz2) return fail(MPX_ERR_INVALID, "buffer size must be (nx+2)(ny+2)(nz+2)");
    void *ptr;
    size_t elem;
    bool isPsi;
    int part = 0;
    int rc = find_grid(s, name, &ptr, &elem, &isPsi, &part);
    if (rc) return fail(rc, std::string("unknown grid ") + name);
    if (!ptr) {
        if (!std::strcmp(name, "Eps"))
            for (size_t q = 0; q < count; ++q) host[q] = s->d.eps;
        else
            std::memset(host, 0, count * sizeof(double));
        return MPX_OK;
    }
    if ((rc = materialise_ghosts(s))) return rc;
    // an accumulator component spans its pole's block from its own offset (psi_at)
    std::vector<unsigned char> buf(isPsi ? (mpx::psi_at(s->cells - 1) + 1) * 16 : s->cells * elem);
    MPX_HIP(hipMemcpyAsync(buf.data(), ptr, buf.size(), hipMemcpyDeviceToHost, s->stream));
    MPX_HIP(hipStreamSynchronize(s->stream));
    int slot = 0;
    const unsigned *dlo = res_of(s, ptr, &slot);
    std::vector<unsigned> lo;
    if (dlo && (rc = fetch_words(s, dlo, s->cells, lo))) return rc;
    for (size_t i = 0; i < nx2; ++i)
        for (size_t j = 0; j < ny2; ++j)
            for (size_t k = 0; k < nz2; ++k) {
                size_t src = g.at((int)i, (int)j, (int)k);
                size_t dst = (i * ny2 + j) * nz2 + k;
                if (isPsi)
                    host[dst] = ((const double *)buf.data())[2 * mpx::psi_at(src) + part];
                else if (elem == 8)
                    host[dst] = ((const double *)buf.data())[src];
                else if (dlo)
                    host[dst] = mpx::res_decode(((const float *)buf.data())[src], lo[src], slot);
                else
                    host[dst] = ((const float *)buf.data())[src];
            }
    if (s->path != MPX_PATH_REFERENCE) {
        if (ptr == s->d_rho)
            plasma_ghosts(s, host, 0);
        else if (ptr == s->d_sigma)
            plasma_ghosts(s, host, 1);
        else if (ptr == s->d_jp[0] || ptr == s->d_jp[1] || ptr == s->d_jp[2])
            plasma_ghosts(s, host, 2);
    }
    return MPX_OK;
}

}  // extern "C"

struct mpx_fetch {
    hipEvent_t done = nullptr;  // null: the data was already in place when the handle was made
};

extern "C" {

int mpx_fetch_field(mpx_solver *s, const char *name, double *host, size_t count, mpx_fetch **handle)
{
    if (!s || !name || !host || !handle) return fail(MPX_ERR_INVALID, "null argument");
    MPX_OWNER_SYNC(s);
    if (int rs = settle(s)) return rs;
    const GridDims &g = s->g;
    const size_t nx2 = g.n[0] + 2, ny2 = g.n[1] + 2, nz2 = g.n[2] + 2;
    if (count != nx2 * ny2 * nz2) return fail(MPX_ERR_INVALID, "buffer size must be (nx+2)(ny+2)(nz+2)");
    void *ptr;
    size_t elem;
    bool isPsi;
    int part = 0;
    int rc = find_grid(s, name, &ptr, &elem, &isPsi, &part);
    if (rc) return fail(rc, std::string("unknown grid ") + name);
    const bool plasma = ptr && (ptr == s->d_rho || ptr == s->d_sigma || ptr == s->d_jp[0] || ptr == s->d_jp[1] ||
                                ptr == s->d_jp[2]);
    if (!ptr || isPsi || plasma) {
        // grids whose host view is assembled on the host: read synchronously and reorder
        std::vector<double> tmp(count);
        if ((rc = mpx_get_field(s, name, tmp.data(), count))) return rc;
        for (size_t i = 0; i < nx2; ++i)
            for (size_t j = 0; j < ny2; ++j)
                for (size_t k = 0; k < nz2; ++k) host[(k * ny2 + j) * nx2 + i] = tmp[(i * ny2 + j) * nz2 + k];
        *handle = new mpx_fetch;
        return MPX_OK;
    }
    if (!s->d_stage) {
        if (!(s->d_stage = s->dalloc<double>(count, rc))) return rc;
    }
    if ((rc = materialise_ghosts(s))) return rc;
    int slot = 0;
    const unsigned *dlo = res_of(s, ptr, &slot);
    if (elem == 8)
        mpx::launch_pack_xfast<double>(g, (const double *)ptr, s->d_stage, s->stream);
    else
        mpx::launch_pack_xfast<float>(g, (const float *)ptr, s->d_stage, s->stream, dlo, slot);
    MPX_HIP(hipGetLastError());
    MPX_HIP(hipMemcpyAsync(host, s->d_stage, count * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    mpx_fetch *f = new mpx_fetch;
    if (hipEventCreateWithFlags(&f->done, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(f->done, s->stream) != hipSuccess) {
        if (f->done) (void)hipEventDestroy(f->done);
        delete f;
        return fail(MPX_ERR_HIP, "event for the field copy failed");
    }
    *handle = f;
    return MPX_OK;
}

int mpx_fetch_wait(mpx_fetch *f)
{
    if (!f) return fail(MPX_ERR_INVALID, "null handle");
    int rc = MPX_OK;
    if (f->done) {
        if (hipEventSynchronize(f->done) != hipSuccess) rc = fail(MPX_ERR_HIP, "waiting for the field copy failed");
        (void)hipEventDestroy(f->done);
    }
    delete f;
    return rc;
}

int mpx_host_alloc(size_t bytes, void **ptr)
{
    if (!ptr) return fail(MPX_ERR_INVALID, "null argument");
    if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *ptr = nullptr;
        return fail(MPX_ERR_NOMEM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return MPX_OK;
}

int mpx_host_free(void *ptr)
{
    if (ptr && hipHostFree(ptr) != hipSuccess) return fail(MPX_ERR_HIP, "hipHostFree failed");
    return MPX_OK;
}

int mpx_set_field(mpx_solver *s, const char *name, const double *host, size_t count)
{
    if (!s || !name || !host) return fail(MPX_ERR_INVALID, "null argument");
    MPX_OWNER_SYNC(s);
    if (int rs = settle(s)) return rs;
    const GridDims &g = s->g;
    size_t nx2 = g.n[0] + 2, ny2 = g.n[1] + 2, nz2 = g.n[2] + 2;
    if (count != nx2 * ny2 * nz2) return fail(MPX_ERR_INVALID, "buffer size must be (nx+2)(ny+2)(nz+2)");
    if (!std::strcmp(name, "Eps") && s->plrc && !s->d_eps) {
        // a per-cell permittivity: the fused and reference-structured kernels read it
        if (s->path == MPX_PATH_SWEEP) return fail(MPX_ERR_UNSUPPORTED, "a per-cell Eps needs the fused or reference path");
        if (s->owner) return fail(MPX_ERR_INVALID, "run belongs to a batch: use mpx_batch_*");
        int rc = 0;
        s->d_eps = s->dalloc<double>(s->cells, rc);
        if (rc) return rc;
    }
    void *ptr;
    size_t elem;
    bool isPsi;
    int part = 0;
    int rc = find_grid(s, name, &ptr, &elem, &isPsi, &part);
    if (rc) return fail(rc, std::string("unknown grid ") + name);
    if (!ptr) return fail(MPX_ERR_UNSUPPORTED, std::string("grid ") + name + " is not stored (inactive pole or Sigma)");
    // an accumulator component spans its pole's block from its own offset (psi_at)
    std::vector<unsigned char> buf(isPsi ? (mpx::psi_at(s->cells - 1) + 1) * 16 : s->cells * elem);
    MPX_HIP(hipMemcpyAsync(buf.data(), ptr, buf.size(), hipMemcpyDeviceToHost, s->stream));
    MPX_HIP(hipStreamSynchronize(s->stream));
    int slot = 0;
    unsigned *dlo = const_cast<unsigned *>(res_of(s, ptr, &slot));
    std::vector<unsigned> lo;
    if (dlo && (rc = fetch_words(s, dlo, s->cells, lo))) return rc;
    for (size_t i = 0; i < nx2; ++i)
        for (size_t j = 0; j < ny2; ++j)
            for (size_t k = 0; k < nz2; ++k) {
                size_t dst = g.at((int)i, (int)j, (int)k);
                size_t src = (i * ny2 + j) * nz2 + k;
                if (isPsi)
                    ((double *)buf.data())[2 * mpx::psi_at(dst) + part] = host[src];
                else if (elem == 8)
                    ((double *)buf.data())[dst] = host[src];
                else if (dlo) {
                    unsigned w = 0;
                    ((float *)buf.data())[dst] = mpx::res_encode(host[src], w, slot);
                    lo[dst] = (lo[dst] & ~(0x3ffu << (10 * slot))) | w;
                } else
                    ((float *)buf.data())[dst] = (float)host[src];
            }
    MPX_HIP(hipMemcpyAsync(ptr, buf.data(), buf.size(), hipMemcpyHostToDevice, s->stream));
    if (dlo) MPX_HIP(hipMemcpyAsync(dlo, lo.data(), lo.size() * sizeof(unsigned), hipMemcpyHostToDevice, s->stream));
    MPX_HIP(hipStreamSynchronize(s->stream));
    return MPX_OK;
}

int mpx_get_layer(mpx_solver *s, const char *name, int dir, double *host, size_t count, int lo[3], int hi[3])
{
    if (!s || !name) return fail(MPX_ERR_INVALID, "null argument");
    MPX_OWNER_SYNC(s);
    if (int rs = settle(s)) return rs;
    for (const Layer &L : s->layers)
        if (L.dir == dir && L.name == name) {
            if (lo && hi)
                for (int d = 0; d < 3; ++d) {
                    lo[d] = L.lo[d];
                    hi[d] = L.hi[d];
                }
            if (!host) return MPX_OK;
            if (count != L.count()) return fail(MPX_ERR_INVALID, "layer buffer size mismatch");
            if (!L.comp[0]) {
                std::memset(host, 0, count * sizeof(double));
                return MPX_OK;
            }
            bool sheet = L.name.rfind("Incident", 0) == 0;
            size_t elem = sheet ? 8 : s->esz;
            std::vector<unsigned char> buf(count * elem);
            MPX_HIP(hipMemcpyAsync(buf.data(), L.comp[0], buf.size(), hipMemcpyDeviceToHost, s->stream));
            MPX_HIP(hipStreamSynchronize(s->stream));
            std::vector<unsigned> lo;
            int rc;
            if (!sheet && L.rw && (rc = fetch_words(s, L.rw, count, lo))) return rc;
            for (size_t q = 0; q < count; ++q)
                host[q] = elem == 8 ? ((const double *)buf.data())[q]
                                    : (lo.empty() ? (double)((const float *)buf.data())[q]
                                                  : mpx::res_decode(((const float *)buf.data())[q], lo[q], L.slot));
            return MPX_OK;
        }
    return fail(MPX_ERR_NAME, std::string("no layer ") + name + " on direction " + std::to_string(dir));
}

int mpx_set_probes(mpx_solver *s, int nprobes, const int *ijk, const int *comp, int capacity)
{
    if (!s || nprobes < 0 || capacity < 0 || (nprobes > 0 && (!ijk || !comp)))
        return fail(MPX_ERR_INVALID, "invalid probe arguments");
    MPX_NOT_IN_BATCH(s);
    if (int rs = settle(s)) return rs;
    std::vector<size_t> idx(nprobes);
    for (int p = 0; p < nprobes; ++p) {
        for (int d = 0; d < 3; ++d)
            if (ijk[3 * p + d] < 1 || ijk[3 * p + d] > s->g.n[d])
                return fail(MPX_ERR_INVALID, "probe cell outside the interior 1..N");
        if (comp[p] < 0 || comp[p] > 5) return fail(MPX_ERR_INVALID, "probe component must be 0..5");
        idx[p] = s->g.at(ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2]);
    }
    int rc = MPX_OK;
    s->nprobe = 0;
    s->probe_rec = 0;
    s->probe_cap = capacity;
    if (nprobes == 0) return MPX_OK;
    s->d_probe_comp = s->dalloc<int>(nprobes, rc);
    s->d_probe_idx = s->dalloc<size_t>(nprobes, rc);
    s->d_probe_hist = s->dalloc<double>((size_t)nprobes * (capacity > 0 ? capacity : 1), rc);
    if (rc) return rc;
    if (upload(s, s->d_probe_comp, comp, nprobes * sizeof(int)) ||
        upload(s, s->d_probe_idx, idx.data(), nprobes * sizeof(size_t)))
        return fail(MPX_ERR_HIP, "probe upload failed");
    s->nprobe = nprobes;
    return MPX_OK;
}

int mpx_get_probes(mpx_solver *s, double *host, int *nrecorded)
{
    if (!s || !nrecorded) return fail(MPX_ERR_INVALID, "null argument");
    *nrecorded = s->probe_rec;
    if (host && s->nprobe && s->probe_rec) {
        MPX_HIP(hipMemcpyAsync(host, s->d_probe_hist, (size_t)s->probe_rec * s->nprobe * sizeof(double),
                               hipMemcpyDeviceToHost, s->stream));
        MPX_HIP(hipStreamSynchronize(s->stream));
    }
    return MPX_OK;
}

}  // extern "C"

// A sampler: values of one grid at a storage cell or along a storage line, one record per
// mpx_sampler_record call, kept in a device ring of `capacity` records until drained.  Records
// of E and B are gathered on the device with the ghost images the wrap would give; other grids
// go through mpx_get_field on the host (their ghosts are assembled there).
struct mpx_sampler {
    mpx_solver *s = nullptr;
    std::string name;
    int dim = -1, pos[3] = {0, 0, 0}, len = 1, capacity = 0, pending = 0;
    bool device = false;         // a grid of E or B, gathered on the device
    int wrapmask = 0;
    double *ring = nullptr;      // device, capacity * len
    std::vector<double> host;    // records of the host-read grids
};

extern "C" {

int mpx_sampler_create(mpx_solver *s, const char *name, int dim, const int pos[3], int capacity, mpx_sampler **out)
{
    if (!s || !name || !pos || !out) return fail(MPX_ERR_INVALID, "null argument");
    MPX_NOT_IN_BATCH(s);
    if (dim < -1 || dim > 2) return fail(MPX_ERR_INVALID, "sampler dim must be -1 (cell) or 0..2 (line)");
    for (int d = 0; d < 3; ++d)
        if (pos[d] < 0 || pos[d] > s->g.n[d] + 1) return fail(MPX_ERR_INVALID, "sampler cell outside the storage box 0..N+1");
    const int len = dim < 0 ? 1 : s->g.n[dim] + 2;
    if (capacity < 1 || (size_t)capacity * len > ((size_t)1 << 27))
        return fail(MPX_ERR_INVALID, "sampler capacity must be 1.. with capacity * values per record <= 2^27");
    void *ptr;
    size_t elem;
    bool isPsi;
    int part = 0;
    int rc = find_grid(s, name, &ptr, &elem, &isPsi, &part);
    if (rc) return fail(rc, std::string("unknown grid ") + name);
    std::unique_ptr<mpx_sampler> p(new mpx_sampler);
    p->s = s;
    p->name = name;
    p->dim = dim;
    for (int d = 0; d < 3; ++d) p->pos[d] = pos[d];
    p->len = len;
    p->capacity = capacity;
    for (int c = 0; c < 3 && ptr && !isPsi; ++c)
        if (ptr == s->E[c] || ptr == s->B[c]) p->device = true;
    if (p->device) {
        // the fused and sweep paths of the PLRC solvers leave the ghosts of E and B stale
        if (s->plrc && s->path != MPX_PATH_REFERENCE)
            for (int d = 0; d < 3; ++d) p->wrapmask |= s->periodic[d] ? 1 << d : 0;
        if (hipMalloc((void **)&p->ring, (size_t)capacity * len * sizeof(double)) != hipSuccess)
            return fail(MPX_ERR_NOMEM, "hipMalloc of the sampler ring failed");
    }
    *out = p.release();
    return MPX_OK;
}

int mpx_sampler_record(mpx_sampler *p)
{
    if (!p) return fail(MPX_ERR_INVALID, "null sampler");
    mpx_solver *s = p->s;
    MPX_NOT_IN_BATCH(s);
    if (p->pending >= p->capacity) return fail(MPX_ERR_INVALID, "sampler full: drain it first");
    if (int rs = settle(s)) return rs;
    if (p->device) {
        // looked up at every record: the sweep path swaps its two sets of E and B after each step
        void *grid;
        size_t elem;
        bool isPsi;
        int part = 0;
        if (find_grid(s, p->name.c_str(), &grid, &elem, &isPsi, &part) || !grid)
            return fail(MPX_ERR_INVALID, "sampler grid is gone");
        double *dst = p->ring + (size_t)p->pending * p->len;
        if (elem == 8)
            mpx::launch_sample_gather<double>(s->g, (const double *)grid, p->dim, p->pos, p->len, p->wrapmask, dst,
                                              s->stream);
        else {
            int slot = 0;
            const unsigned *dlo = res_of(s, grid, &slot);
            if (dlo && !(s->plrc && s->path != MPX_PATH_REFERENCE))
                return fail(MPX_ERR_UNSUPPORTED, "residual words without index wrap");
            mpx::launch_sample_gather<float>(s->g, (const float *)grid, p->dim, p->pos, p->len, p->wrapmask, dst,
                                             s->stream, dlo, slot);
        }
        MPX_HIP(hipGetLastError());
    } else {
        const GridDims &g = s->g;
        const size_t ny2 = g.n[1] + 2, nz2 = g.n[2] + 2;
        std::vector<double> f((size_t)(g.n[0] + 2) * ny2 * nz2);
        int rc = mpx_get_field(s, p->name.c_str(), f.data(), f.size());
        if (rc) return rc;
        int c[3] = {p->pos[0], p->pos[1], p->pos[2]};
        for (int q = 0; q < p->len; ++q) {
            if (p->dim >= 0) c[p->dim] = q;
            p->host.push_back(f[((size_t)c[0] * ny2 + c[1]) * nz2 + c[2]]);
        }
    }
    p->pending++;
    return MPX_OK;
}

int mpx_sampler_drain(mpx_sampler *p, double *host, size_t count, int *nrecords)
{
    if (!p || !nrecords || (!host && p->pending)) return fail(MPX_ERR_INVALID, "null argument");
    const size_t n = (size_t)p->pending * p->len;
    if (count < n) return fail(MPX_ERR_INVALID, "buffer smaller than the pending records");
    if (n && p->device) {
        MPX_HIP(hipMemcpyAsync(host, p->ring, n * sizeof(double), hipMemcpyDeviceToHost, p->s->stream));
        MPX_HIP(hipStreamSynchronize(p->s->stream));
    } else if (n) {
        std::memcpy(host, p->host.data(), n * sizeof(double));
    }
    p->host.clear();
    *nrecords = p->pending;
    p->pending = 0;
    return MPX_OK;
}

int mpx_sampler_destroy(mpx_sampler *p)
{
    if (!p) return MPX_OK;
    int rc = MPX_OK;
    if (p->ring) {
        // records still queued on the solver's stream write into the ring
        if (hipStreamSynchronize(p->s->stream) != hipSuccess || hipFree(p->ring) != hipSuccess)
            rc = fail(MPX_ERR_HIP, "freeing the sampler ring failed");
    }
    delete p;
    return rc;
}

int mpx_energy(mpx_solver *s, double *energy)
{
    if (!s || !energy) return fail(MPX_ERR_INVALID, "null argument");
    MPX_OWNER_SYNC(s);
    if (int rs = settle(s)) return rs;
    if (s->esz == 8) {
        double *e[3] = {(double *)s->E[0], (double *)s->E[1], (double *)s->E[2]};
        double *b[3] = {(double *)s->B[0], (double *)s->B[1], (double *)s->B[2]};
        mpx::launch_energy<double>(s->g, e, b, s->d_part, s->d_energy, s->stream);
    } else {
        float *e[3] = {(float *)s->E[0], (float *)s->E[1], (float *)s->E[2]};
        float *b[3] = {(float *)s->B[0], (float *)s->B[1], (float *)s->B[2]};
        mpx::launch_energy<float>(s->g, e, b, s->d_part, s->d_energy, s->stream, s->Elo, s->Blo);
    }
    MPX_LAUNCHED("energy");
    double v = 0;
    MPX_HIP(hipMemcpyAsync(&v, s->d_energy, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    MPX_HIP(hipStreamSynchronize(s->stream));
    *energy = v * (s->d.dx * s->d.dy * s->d.dz);
    return MPX_OK;
}

}  // extern "C"

// Scratch of mpx_energy_to: the partial sums of the energy reduction.  Each GPU the library
// runs on holds its own copy of the array until the process ends.  In host code the array's
// name is no device address: the address on each GPU comes from hipGetSymbolAddress, taken
// once per device (with that device current) at the entry point's first call there.
static __device__ __align__(256) double scratch[MPX_NOALLOC_PARTIALS];
static std::mutex g_scratch_mu;
static std::vector<double *> g_scratch;  // per device id; null until looked up

static int scratch_on(int device, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
    if (!g_scratch[device]) {
        int cur = 0;
        MPX_HIP(hipGetDevice(&cur));
        if (cur != device) MPX_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) MPX_HIP(hipSetDevice(cur));
        if (e != hipSuccess || !p) return fail(MPX_ERR_HIP, std::string("hipGetSymbolAddress of the scratch failed: ") +
                                                                 hipGetErrorString(e));
        g_scratch[device] = (double *)p;
    }
    *out = g_scratch[device];
    return MPX_OK;
}

extern "C" {

int mpx_energy_to(mpx_solver *s, double *d_energy)
{
    if (!s || !d_energy) return fail(MPX_ERR_INVALID, "null argument");
    MPX_NOT_IN_BATCH(s);
    if (s->pending)
        return fail(MPX_ERR_INVALID, "the pass path owes the B half of the last step: read the state or call "
                                     "mpx_energy once before mpx_energy_to");
    const size_t np = mpx::energy_partials(s->g);
    if (np > (size_t)MPX_NOALLOC_PARTIALS)
        return fail(MPX_ERR_UNSUPPORTED, "the grid needs " + std::to_string(np) + " partial sums (ceil(nz/64) ceil(ny/4) nx); "
                                         "the scratch holds " + std::to_string(MPX_NOALLOC_PARTIALS) + ": use mpx_energy");
    double *part = nullptr;
    if (int rc = scratch_on(s->device, &part)) return rc;
    const double dv = s->d.dx * s->d.dy * s->d.dz;
    if (s->esz == 8) {
        double *e[3] = {(double *)s->E[0], (double *)s->E[1], (double *)s->E[2]};
        double *b[3] = {(double *)s->B[0], (double *)s->B[1], (double *)s->B[2]};
        mpx::launch_energy<double>(s->g, e, b, part, d_energy, s->stream, nullptr, nullptr, dv);
    } else {
        float *e[3] = {(float *)s->E[0], (float *)s->E[1], (float *)s->E[2]};
        float *b[3] = {(float *)s->B[0], (float *)s->B[1], (float *)s->B[2]};
        mpx::launch_energy<float>(s->g, e, b, part, d_energy, s->stream, s->Elo, s->Blo, dv);
    }
    MPX_LAUNCHED("energy");
    return MPX_OK;
}

int mpx_time_steps(mpx_solver *s, int nsteps, mpx_timing *out)
{
    if (!s || !out || nsteps < 1) return fail(MPX_ERR_INVALID, "invalid arguments");
    MPX_NOT_IN_BATCH(s);
    std::memset(out, 0, sizeof(*out));
    hipEvent_t t0, t1;
    MPX_HIP(hipEventCreate(&t0));
    MPX_HIP(hipEventCreate(&t1));
    s->timer.on = true;
    s->timer.used = 0;
    s->timer.cls.clear();
    s->timer.ev0.clear();
    s->timer.ev1.clear();
    (void)hipEventRecord(t0, s->stream);
    int rc = mpx_step(s, nsteps);
    (void)hipEventRecord(t1, s->stream);
    s->timer.on = false;
    if (rc == MPX_OK) {
        if (hipEventSynchronize(t1) != hipSuccess) rc = fail(MPX_ERR_HIP, "event sync failed");
    }
    if (rc == MPX_OK) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t0, t1);
        out->step_ms = ms / nsteps;
        double sum[K_NCLASSES] = {0};
        int cnt[K_NCLASSES] = {0};
        for (size_t q = 0; q < s->timer.cls.size(); ++q) {
            float m = 0;
            (void)hipEventElapsedTime(&m, s->timer.ev0[q], s->timer.ev1[q]);
            sum[s->timer.cls[q]] += m;
            cnt[s->timer.cls[q]]++;
        }
        for (int c = 0; c < K_NCLASSES; ++c)
            if (cnt[c]) {
                int q = out->n_classes++;
                std::snprintf(out->name[q], sizeof out->name[q], "%s", kClassNames[c]);
                out->mean_ms[q] = sum[c] / cnt[c];
                out->launches[q] = cnt[c];
            }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return rc;
}

int mpx_state_bytes(mpx_solver *s, size_t *bytes)
{
    if (!s || !bytes) return fail(MPX_ERR_INVALID, "null argument");
    size_t b = 0;
    for (void *p : s->allocs) (void)p;
    b += 6 * s->cells * s->esz;
    if (s->res) b += 2 * s->cells * sizeof(unsigned);
    b += (size_t)s->np * mpx::psi_block_elems(s->cells) * 16;
    *bytes = b;
    return MPX_OK;
}

// Streaming-copy probe: copies `bytes` between two fresh device buffers `reps` times and
// reports the mean kernel time.  No reference counterpart: it calibrates HBM byte counters
// and measures the practical copy bandwidth next to the stencil kernels.
int mpx_copy_probe(size_t bytes, int reps, double *ms_per_copy)
{
    if (bytes < 16 || reps < 1 || !ms_per_copy) return fail(MPX_ERR_INVALID, "invalid copy probe arguments");
    void *a = nullptr, *b = nullptr;
    hipStream_t st;
    hipEvent_t e0, e1;
    MPX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        (void)hipStreamDestroy(st);
        return fail(MPX_ERR_NOMEM, "copy probe allocation failed");
    }
    MPX_HIP(hipMemsetAsync(a, 0, bytes, st));
    MPX_HIP(hipEventCreate(&e0));
    MPX_HIP(hipEventCreate(&e1));
    mpx::launch_hbm_copy(a, b, bytes, st);  // warm-up
    MPX_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) mpx::launch_hbm_copy(a, b, bytes, st);
    MPX_HIP(hipEventRecord(e1, st));
    MPX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MPX_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipStreamDestroy(st);
    return MPX_OK;
}

// Launch-shape parameters of the fused kernels (autotuner hook; no reference counterpart).
int mpx_set_tuning(mpx_solver *s, const char *key, int value)
{
    if (!s || !key) return fail(MPX_ERR_INVALID, "null argument");
    MPX_NOT_IN_BATCH(s);
    bool march = !std::strcmp(key, "march_e") || !std::strcmp(key, "march_b");
    if (march && value != 1 && value != 2 && value != 4) return fail(MPX_ERR_INVALID, "march must be 1, 2 or 4");
    if (!std::strcmp(key, "march_e"))
        s->march_e = value;
    else if (!std::strcmp(key, "march_b"))
        s->march_b = value;
    else if (!std::strcmp(key, "tile_order")) {
        if (value < 0 || value > 2) return fail(MPX_ERR_INVALID, "tile_order must be 0, 1 or 2");
        s->tile_order = value;
    } else if (!std::strcmp(key, "b_reverse")) {
        if (value != 0 && value != 1) return fail(MPX_ERR_INVALID, "b_reverse must be 0 or 1");
        s->b_reverse = value;
    } else if (!std::strcmp(key, "tile_band")) {
        if (value < 0) return fail(MPX_ERR_INVALID, "tile_band must be >= 0 (0: chosen from the grid)");
        s->tile_band = value;
    } else if (!std::strcmp(key, "sweep_seg")) {
        if (value < 1) return fail(MPX_ERR_INVALID, "sweep_seg must be >= 1");
        s->sweep_seg = value < s->g.n[0] ? value : s->g.n[0];
    } else if (!std::strcmp(key, "newton_cap")) {
        if (value < 1) return fail(MPX_ERR_INVALID, "newton_cap must be >= 1");
        s->ncap = value;
    } else if (!std::strcmp(key, "pass_kernel")) {
        if (value != 0 && value != 1) return fail(MPX_ERR_INVALID, "pass_kernel must be 0 or 1");
        s->pass_kernel = value;
    } else if (!std::strcmp(key, "march_seg")) {
        if (value < 0) return fail(MPX_ERR_INVALID, "march_seg must be >= 1 (0: chosen from the grid)");
        s->march_seg = value;
    } else if (!std::strcmp(key, "march_rows")) {
        if (value != 0 && value != 3 && value != 7)
            return fail(MPX_ERR_INVALID, "march_rows must be 3 or 7 (0: chosen from the kernel's registers)");
        s->march_rows = value;
    } else if (!std::strcmp(key, "sweep_variant")) {
        if (value < 0 || value > 2) return fail(MPX_ERR_INVALID, "sweep_variant must be 0, 1 or 2");
        s->sweep_variant = value;
    } else
        return fail(MPX_ERR_NAME, std::string("unknown tuning key ") + key);
    return MPX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// batched independent runs: one persistent sweep launch per step for all runs
// ---------------------------------------------------------------------------------------
//
// The runs are ordinary solvers built by `build` (sweep path, so that both kernel structures
// have their buffers), stepped in lockstep from one stream.  Per step the source sheets of
// every run are evaluated in one launch, then either the fused E and fused B kernels run over
// the tiles of every run (structure 0) or the single-pass sweep does (structure 1).  The step
// counter lives on the device, so a captured graph of K steps can be replayed at any later
// step.  Under the sweep the runs' buffer sets alternate every step, and two argument tables
// (one per parity) hold both arrangements.

struct mpx_batch {
    std::vector<mpx_solver *> runs;
    hipStream_t stream = nullptr;
    size_t esz = 8;
    int np = 0;
    bool nonlin = false, cpml = false;
    // 0: fused E + fused B kernels, 1: single-pass sweep, 2: x-march pass kernel (B of one step
    // and E of the next per launch; the fused E kernel opens a call, the fused B kernel closes it)
    int structure = 0;
    int variant = 2, seg = 32, blocks_per_cu = 2, graph_steps = 50;  // sweep tiling
    int march_b = 4, fused_bpc = 8;  // fused B planes per thread, fused blocks per CU
    int pass_seg = 16, pass_rows = 0;  // x-march: planes per block, rows per tile (0: per variant)
    int prows = 7;                     // x-march rows in use
    int cus = 256;
    int time = 0;
    int parity = 0;  // steps taken since creation, mod 2
    void *d_args[2] = {nullptr, nullptr};
    mpx::BatchRun *d_runs = nullptr;
    int total_tiles = 0, blocks = 0;
    // fused E / B argument tables [side][arrangement]: the x-march swaps the field sets every
    // step, and the fused kernels run on whichever set holds the state (the other one, as laid
    // out by swap_pass_buffers, only under structure 2)
    void *d_fargs[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    mpx::BatchRun *d_fruns[2] = {nullptr, nullptr};  // fused E / B tile tables
    int ftotal[2] = {0, 0}, fblocks[2] = {0, 0};
    void *d_pargs[2] = {nullptr, nullptr};  // x-march argument tables per arrangement
    mpx::BatchRun *d_pruns = nullptr;
    int ptotal = 0;
    mpx::SheetEval *d_sheets = nullptr;
    int *d_sheet_first = nullptr;
    int nsheets = 0, sheet_blocks = 0;
    int *d_time = nullptr;
    mpx::FaddeevaTables ft{};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    std::vector<void *> allocs;

    void drop_graphs()
    {
        for (int p = 0; p < 2; ++p)
            if (graph[p]) {
                (void)hipGraphExecDestroy(graph[p]);
                graph[p] = nullptr;
            }
    }
    ~mpx_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (mpx_solver *s : runs) {
            s->owner = nullptr;
            delete s;
        }
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int alloc(void **p, size_t bytes)
    {
        if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) return fail(MPX_ERR_NOMEM, "batch table allocation failed");
        allocs.push_back(*p);
        return MPX_OK;
    }
};

// argument tables, tile list and sheet list for the current tuning; drops captured graphs
template <typename T>
static int batch_prepare_t(mpx_batch *b)
{
    const int n = (int)b->runs.size();
    std::vector<mpx::SweepArgs<T>> a[2];
    std::vector<mpx::BatchRun> runs(n);
    int total = 0;
    for (int r = 0; r < n; ++r) {
        mpx_solver *s = b->runs[r];
        s->sweep_variant = b->variant;
        s->sweep_seg = b->seg < s->g.n[0] ? b->seg : s->g.n[0];
        a[b->parity].push_back(sweep_args<T>(s));
        swap_sweep_buffers(s);
        a[b->parity ^ 1].push_back(sweep_args<T>(s));
        swap_sweep_buffers(s);
        mpx::BatchRun &R = runs[r];
        R.first = total;
        mpx::sweep_tiles(s->g, b->variant, s->sweep_seg, R.ntz, R.nty, R.nseg);
        long long t = (long long)R.ntz * R.nty * R.nseg;
        if (total + t > (1LL << 30)) return fail(MPX_ERR_UNSUPPORTED, "batch has too many tiles");
        total += (int)t;
    }
    b->drop_graphs();
    MPX_HIP(hipStreamSynchronize(b->stream));
    size_t bytes = n * sizeof(mpx::SweepArgs<T>);
    for (int p = 0; p < 2; ++p) {
        if (!b->d_args[p] && b->alloc(&b->d_args[p], bytes)) return MPX_ERR_NOMEM;
        MPX_HIP(hipMemcpyAsync(b->d_args[p], a[p].data(), bytes, hipMemcpyHostToDevice, b->stream));
    }
    if (!b->d_runs && b->alloc((void **)&b->d_runs, n * sizeof(mpx::BatchRun))) return MPX_ERR_NOMEM;
    MPX_HIP(hipMemcpyAsync(b->d_runs, runs.data(), n * sizeof(mpx::BatchRun), hipMemcpyHostToDevice, b->stream));
    MPX_HIP(hipStreamSynchronize(b->stream));
    b->total_tiles = total;
    // persistent grid: blocks_per_cu blocks on every CU, a multiple of the 8 XCDs, no more
    // than the tiles need
    auto grid = [&](int bpc, int tiles) {
        int want = bpc * b->cus, need = (tiles + 7) / 8 * 8;
        return ((want < need ? want : need) + 7) / 8 * 8;
    };
    b->blocks = grid(b->blocks_per_cu, total);
    // fused structure: E side (one plane per thread) and B side (march_b planes per thread),
    // for the arrangement of the field sets as they are (index b->parity) and as the next
    // x-march pass leaves them (b->parity ^ 1)
    for (int side = 0; side < 2; ++side) {
        std::vector<StepArgs<T>> fa[2];
        std::vector<mpx::BatchRun> fr(n);
        int ft = 0;
        for (int r = 0; r < n; ++r) {
            mpx_solver *s = b->runs[r];
            for (int q = 0; q < 2; ++q) {
                if (q) swap_pass_buffers(s);
                StepArgs<T> sa = step_args<T>(s, side == 1);
                sa.march = side ? b->march_b : 1;
                fa[b->parity ^ q].push_back(sa);
                if (q) swap_pass_buffers(s);
            }
            fr[r].first = ft;
            mpx::fused_tiles(s->g, fa[0].back().blo, side == 1, fa[0].back().march, fr[r].ntz, fr[r].nty,
                             fr[r].nseg);
            long long t = (long long)fr[r].ntz * fr[r].nty * fr[r].nseg;
            if (ft + t > (1LL << 30)) return fail(MPX_ERR_UNSUPPORTED, "batch has too many tiles");
            ft += (int)t;
        }
        size_t fb = n * sizeof(StepArgs<T>);
        for (int q = 0; q < 2; ++q) {
            if (!b->d_fargs[side][q] && b->alloc(&b->d_fargs[side][q], fb)) return MPX_ERR_NOMEM;
            MPX_HIP(hipMemcpyAsync(b->d_fargs[side][q], fa[q].data(), fb, hipMemcpyHostToDevice, b->stream));
        }
        if (!b->d_fruns[side] && b->alloc((void **)&b->d_fruns[side], n * sizeof(mpx::BatchRun))) return MPX_ERR_NOMEM;
        MPX_HIP(hipMemcpyAsync(b->d_fruns[side], fr.data(), n * sizeof(mpx::BatchRun), hipMemcpyHostToDevice,
                               b->stream));
        MPX_HIP(hipStreamSynchronize(b->stream));
        b->ftotal[side] = ft;
        b->fblocks[side] = grid(b->fused_bpc, ft);
    }
    // x-march structure: one block per tile of 64 z cells x prows rows x pass_seg planes
    {
        b->prows = mpx::march_batch_rows<T>(b->pass_rows, b->np, b->nonlin, b->cpml);
        std::vector<mpx::PassArgs<T>> pa[2];
        std::vector<mpx::BatchRun> pr(n);
        int pt = 0;
        for (int r = 0; r < n; ++r) {
            mpx_solver *s = b->runs[r];
            for (int q = 0; q < 2; ++q) {
                if (q) swap_pass_buffers(s);
                mpx::PassArgs<T> a = pass_args<T>(s);
                a.kernel = 1;
                a.seg = b->pass_seg < s->g.n[0] ? b->pass_seg : s->g.n[0];
                a.rows = b->prows;
                pa[b->parity ^ q].push_back(a);
                if (q) swap_pass_buffers(s);
            }
            const mpx::PassArgs<T> &a = pa[0].back();
            pr[r].first = pt;
            pr[r].ntz = (s->g.n[2] + 63) / 64;
            pr[r].nty = (s->g.n[1] + b->prows - 1) / b->prows;
            pr[r].nseg = (s->g.n[0] + a.seg - 1) / a.seg;
            long long t = (long long)pr[r].ntz * pr[r].nty * pr[r].nseg;
            if (pt + t > (1LL << 30)) return fail(MPX_ERR_UNSUPPORTED, "batch has too many tiles");
            pt += (int)t;
        }
        size_t pb = n * sizeof(mpx::PassArgs<T>);
        for (int q = 0; q < 2; ++q) {
            if (!b->d_pargs[q] && b->alloc(&b->d_pargs[q], pb)) return MPX_ERR_NOMEM;
            MPX_HIP(hipMemcpyAsync(b->d_pargs[q], pa[q].data(), pb, hipMemcpyHostToDevice, b->stream));
        }
        if (!b->d_pruns && b->alloc((void **)&b->d_pruns, n * sizeof(mpx::BatchRun))) return MPX_ERR_NOMEM;
        MPX_HIP(hipMemcpyAsync(b->d_pruns, pr.data(), n * sizeof(mpx::BatchRun), hipMemcpyHostToDevice, b->stream));
        MPX_HIP(hipStreamSynchronize(b->stream));
        b->ptotal = pt;
    }
    return MPX_OK;
}

static int batch_prepare(mpx_batch *b)
{
    return b->esz == 8 ? batch_prepare_t<double>(b) : batch_prepare_t<float>(b);
}

// the fused E (side 0) or fused B (side 1) kernel over every run, field arrangement par
static void batch_fused(mpx_batch *b, int side, int par)
{
    const int n = (int)b->runs.size();
    if (b->esz == 8) {
        if (side == 0)
            mpx::launch_fused_E_batch<double>((const StepArgs<double> *)b->d_fargs[0][par], b->d_fruns[0], n,
                                              b->ftotal[0], b->fblocks[0], 1, b->np, b->nonlin, b->cpml, b->stream);
        else
            mpx::launch_fused_B_batch<double>((const StepArgs<double> *)b->d_fargs[1][par], b->d_fruns[1], n,
                                              b->ftotal[1], b->fblocks[1], b->march_b, b->cpml, b->stream);
    } else {
        if (side == 0)
            mpx::launch_fused_E_batch<float>((const StepArgs<float> *)b->d_fargs[0][par], b->d_fruns[0], n,
                                             b->ftotal[0], b->fblocks[0], 1, b->np, b->nonlin, b->cpml, b->stream);
        else
            mpx::launch_fused_B_batch<float>((const StepArgs<float> *)b->d_fargs[1][par], b->d_fruns[1], n,
                                             b->ftotal[1], b->fblocks[1], b->march_b, b->cpml, b->stream);
    }
}

// the sheets of the sides in `sides` (bit 0: E currents, bit 1: H currents) at the counter
// + j, the H currents hshift steps earlier
static void batch_sheets(mpx_batch *b, int j, int hshift, int sides)
{
    if (b->nsheets)
        mpx::launch_sheet_eval_batch(b->d_sheets, b->d_sheet_first, b->nsheets, b->sheet_blocks, b->ft, b->d_time, j,
                                     b->stream, hshift, sides);
}

// enqueue step j of a sequence that starts at the device counter, with argument parity par.
// Structure 2: the pass at counter + j (the B half of step counter + j - 1, the E half of
// step counter + j).
static void batch_enqueue(mpx_batch *b, int j, int par)
{
    const int n = (int)b->runs.size();
    if (b->structure == 2) {
        batch_sheets(b, j, 1, 3);
        if (b->esz == 8)
            mpx::launch_march_batch<double>((const mpx::PassArgs<double> *)b->d_pargs[par], b->d_pruns, n, b->ptotal,
                                            b->prows, b->np, b->nonlin, b->cpml, b->stream);
        else
            mpx::launch_march_batch<float>((const mpx::PassArgs<float> *)b->d_pargs[par], b->d_pruns, n, b->ptotal,
                                           b->prows, b->np, b->nonlin, b->cpml, b->stream);
        return;
    }
    batch_sheets(b, j, 0, 3);
    if (b->structure == 0) {
        batch_fused(b, 0, par);
        batch_fused(b, 1, par);
        return;
    }
    if (b->esz == 8)
        mpx::launch_sweep_batch<double>((const mpx::SweepArgs<double> *)b->d_args[par], b->d_runs, n,
                                        b->total_tiles, b->variant, b->blocks, b->np, b->nonlin, b->cpml, b->stream);
    else
        mpx::launch_sweep_batch<float>((const mpx::SweepArgs<float> *)b->d_args[par], b->d_runs, n,
                                       b->total_tiles, b->variant, b->blocks, b->np, b->nonlin, b->cpml, b->stream);
}

// a graph of graph_steps steps (structure 2: passes at counter + 1 .. counter + graph_steps)
static int batch_capture(mpx_batch *b, int par)
{
    hipGraph_t gr = nullptr;
    const int j0 = b->structure == 2 ? 1 : 0;
    MPX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < b->graph_steps; ++j) batch_enqueue(b, j0 + j, b->structure ? (par + j) & 1 : par);
    mpx::launch_advance_time(b->d_time, b->graph_steps, b->stream);
    hipError_t e = hipStreamEndCapture(b->stream, &gr);
    if (e != hipSuccess) return fail(MPX_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph[par], gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
        b->graph[par] = nullptr;
        return fail(MPX_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    return MPX_OK;
}

// nsteps in lockstep; graph replays for whole groups of graph_steps, direct launches after.
// Structure 2 takes the E half of the first step with the fused E kernel, nsteps - 1 passes
// and the B half of the last step with the fused B kernel, so that E and B belong to the same
// step between calls.
static int batch_step(mpx_batch *b, int nsteps, bool use_graph)
{
    if (nsteps == 0) return MPX_OK;
    const bool march = b->structure == 2;
    if (march) {
        batch_sheets(b, 0, 0, 1);
        batch_fused(b, 0, b->parity);
        MPX_LAUNCHED("batch E half");
    }
    const int work = march ? nsteps - 1 : nsteps;
    const int j0 = march ? 1 : 0;
    int done = 0;
    const int K = b->graph_steps;
    if (use_graph && K > 0)
        while (work - done >= K) {
            if (!b->graph[b->parity]) {
                int rc = batch_capture(b, b->parity);
                if (rc) return rc;
            }
            MPX_HIP(hipGraphLaunch(b->graph[b->parity], b->stream));
            if (b->structure) b->parity = (b->parity + K) & 1;
            done += K;
        }
    for (; done < work; ++done) {
        batch_enqueue(b, j0, b->parity);
        mpx::launch_advance_time(b->d_time, 1, b->stream);
        MPX_LAUNCHED("batch step");
        if (b->structure) b->parity ^= 1;
    }
    if (march) {
        batch_sheets(b, 0, 0, 2);
        batch_fused(b, 1, b->parity);
        mpx::launch_advance_time(b->d_time, 1, b->stream);
        MPX_LAUNCHED("batch B half");
    }
    for (mpx_solver *s : b->runs) {
        s->time += nsteps;
        if (b->structure == 1 && (nsteps & 1)) swap_sweep_buffers(s);
        if (march && (work & 1)) swap_pass_buffers(s);
    }
    b->time += nsteps;
    return MPX_OK;
}

// every mutable array of a run back to zero (the state a fresh solver starts from)
static int reset_state(mpx_solver *s, hipStream_t st)
{
    auto z = [&](void *p, size_t bytes) -> int {
        if (p && hipMemsetAsync(p, 0, bytes, st) != hipSuccess) return fail(MPX_ERR_HIP, "reset failed");
        return MPX_OK;
    };
    int rc = 0;
    for (int c = 0; c < 3; ++c) {
        rc |= z(s->E[c], s->cells * s->esz) | z(s->B[c], s->cells * s->esz);
        rc |= z(s->E2[c], s->cells * s->esz) | z(s->B2[c], s->cells * s->esz);
        rc |= z(s->Jt[c], s->cells * 8) | z(s->Mt[c], s->cells * 8);
        for (int n = 0; n < 3; ++n) {
            if (c == 0) rc |= z(s->psi[n][0], mpx::psi_block_elems(s->cells) * 16) | z(s->psi2[n][0], mpx::psi_block_elems(s->cells) * 16);
            rc |= z(s->P[n][c], s->cells * 8) | z(s->Pp[n][c], s->cells * 8);
        }
    }
    for (const Layer &L : s->layers) {
        size_t el = L.name.compare(0, 4, "CPML") == 0 ? s->esz : 8;
        for (int c = 0; c < 3; ++c) rc |= z(L.comp[c], L.count() * el);
    }
    for (int w = 0; w < 2; ++w)
        for (size_t f = 0; f < s->epsi2[w].size(); ++f) {
            const CpmlFace &F = s->efaces[f];
            rc |= z(s->epsi2[w][f], (size_t)(F.hi[0] - F.lo[0] + 1) * F.n1 * F.n2 * s->esz);
        }
    s->time = 0;
    s->probe_rec = 0;
    return rc ? MPX_ERR_HIP : MPX_OK;
}

extern "C" {

int mpx_batch_create(const mpx_desc *descs, int nruns, mpx_batch **out)
{
    if (!descs || !out || nruns < 1 || nruns > MPX_MAX_BATCH) return fail(MPX_ERR_INVALID, "invalid batch arguments");
    *out = nullptr;
    mpx_batch *b = new mpx_batch();
    auto bail = [&](int rc) {
        delete b;
        return rc;
    };
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return bail(fail(MPX_ERR_HIP, "batch stream or device query failed"));
    b->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    for (int r = 0; r < nruns; ++r) {
        mpx_desc d = descs[r];
        d.path = MPX_PATH_SWEEP;
        if (d.solver != MPX_SOLVER_PLRC_NONLIN && d.solver != MPX_SOLVER_PLRC_LIN)
            return bail(fail(MPX_ERR_UNSUPPORTED, "batched runs must use an FDTD_PLRC solver"));
        if (d.plasma) return bail(fail(MPX_ERR_UNSUPPORTED, "batched runs do not take the plasma block"));
        if (d.precision == MPX_PREC_FP32_RES)
            return bail(fail(MPX_ERR_UNSUPPORTED, "batched runs store float64 or float32 (no residual words)"));
        for (int q = 0; q < d.n_factories; ++q)
            if (d.factories[q].kind == MPX_FACTORY_PLASMA_CURRENT)
                return bail(fail(MPX_ERR_UNSUPPORTED, "batched runs do not take a plasma current"));
            else if (d.factories[q].kind == MPX_FACTORY_GAUSS_INJECT)
                return bail(fail(MPX_ERR_UNSUPPORTED, "batched runs do not take gauss_inject (its sample blocks "
                                                      "are scheduled from the host)"));
        mpx_solver *s = nullptr;
        int rc = mpx_create(&d, &s);
        if (rc) return bail(rc);
        if (hipStreamSynchronize(s->stream) != hipSuccess) {
            delete s;
            return bail(fail(MPX_ERR_HIP, "run set-up failed"));
        }
        s->owner = b->stream;
        b->runs.push_back(s);
        bool cp = !s->efaces.empty();
        if (r == 0) {
            b->esz = s->esz;
            b->np = s->np;
            b->nonlin = s->nonlin;
            b->cpml = cp;
            b->ft = s->ft;
        } else if (s->esz != b->esz || s->np != b->np || s->nonlin != b->nonlin || cp != b->cpml) {
            return bail(fail(MPX_ERR_UNSUPPORTED,
                             "batched runs must share precision, solver, active pole count and CPML presence"));
        }
        if (s->time != 0) return bail(fail(MPX_ERR_INVALID, "batched runs start at time 0"));
    }
    // source sheets of all runs
    std::vector<mpx::SheetEval> sheets;
    std::vector<int> first;
    for (mpx_solver *s : b->runs)
        for (int side = 0; side < 2; ++side)
            for (const Current &c : side ? s->hcur : s->ecur)
                if (c.type == (side ? CUR_INC_H : CUR_INC_E)) {
                    first.push_back(b->sheet_blocks);
                    sheets.push_back(c.ev);
                    b->sheet_blocks += c.ev.nx2 * ((c.ev.ny2 + 63) / 64);
                }
    b->nsheets = (int)sheets.size();
    if (b->alloc((void **)&b->d_time, sizeof(int))) return bail(MPX_ERR_NOMEM);
    if (b->nsheets) {
        if (b->alloc((void **)&b->d_sheets, sheets.size() * sizeof(mpx::SheetEval)) ||
            b->alloc((void **)&b->d_sheet_first, first.size() * sizeof(int)))
            return bail(MPX_ERR_NOMEM);
        if (hipMemcpyAsync(b->d_sheets, sheets.data(), sheets.size() * sizeof(mpx::SheetEval), hipMemcpyHostToDevice,
                           b->stream) != hipSuccess ||
            hipMemcpyAsync(b->d_sheet_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice,
                           b->stream) != hipSuccess)
            return bail(fail(MPX_ERR_HIP, "sheet table upload failed"));
    }
    if (hipMemsetAsync(b->d_time, 0, sizeof(int), b->stream) != hipSuccess)
        return bail(fail(MPX_ERR_HIP, "time counter clear failed"));
    int rc = batch_prepare(b);
    if (rc) return bail(rc);
    *out = b;
    return MPX_OK;
}

int mpx_batch_destroy(mpx_batch *b)
{
    delete b;
    return MPX_OK;
}

int mpx_batch_size(mpx_batch *b, int *nruns)
{
    if (!b || !nruns) return fail(MPX_ERR_INVALID, "null argument");
    *nruns = (int)b->runs.size();
    return MPX_OK;
}

int mpx_batch_run(mpx_batch *b, int run, mpx_solver **out)
{
    if (!b || !out || run < 0 || run >= (int)b->runs.size()) return fail(MPX_ERR_INVALID, "invalid run index");
    *out = b->runs[run];
    return MPX_OK;
}

int mpx_batch_step(mpx_batch *b, int nsteps)
{
    if (!b || nsteps < 0) return fail(MPX_ERR_INVALID, "invalid batch or step count");
    return batch_step(b, nsteps, true);
}

int mpx_batch_sync(mpx_batch *b)
{
    if (!b) return fail(MPX_ERR_INVALID, "null batch");
    MPX_HIP(hipStreamSynchronize(b->stream));
    for (size_t r = 0; r < b->runs.size(); ++r)
        if (int rc = check_status(b->runs[r])) return fail(rc, "run " + std::to_string(r) + ": " + g_err);
    return MPX_OK;
}

int mpx_batch_get_time(mpx_batch *b, int *time)
{
    if (!b || !time) return fail(MPX_ERR_INVALID, "null argument");
    *time = b->time;
    return MPX_OK;
}

int mpx_batch_set_tuning(mpx_batch *b, const char *key, int value)
{
    if (!b || !key) return fail(MPX_ERR_INVALID, "null argument");
    if (!std::strcmp(key, "structure")) {
        if (value < 0 || value > 2)
            return fail(MPX_ERR_INVALID, "structure must be 0 (fused), 1 (sweep) or 2 (x-march pass)");
        b->structure = value;
    } else if (!std::strcmp(key, "pass_seg")) {
        if (value < 1) return fail(MPX_ERR_INVALID, "pass_seg must be >= 1");
        b->pass_seg = value;
    } else if (!std::strcmp(key, "pass_rows")) {
        if (value != 0 && value != 3 && value != 7) return fail(MPX_ERR_INVALID, "pass_rows must be 0, 3 or 7");
        b->pass_rows = value;
    } else if (!std::strcmp(key, "march_b")) {
        if (value != 1 && value != 2 && value != 4) return fail(MPX_ERR_INVALID, "march_b must be 1, 2 or 4");
        b->march_b = value;
    } else if (!std::strcmp(key, "fused_blocks_per_cu")) {
        if (value < 1 || value > 32) return fail(MPX_ERR_INVALID, "fused_blocks_per_cu must be 1..32");
        b->fused_bpc = value;
    } else if (!std::strcmp(key, "variant")) {
        if (value < 0 || value > 2) return fail(MPX_ERR_INVALID, "variant must be 0, 1 or 2");
        b->variant = value;
    } else if (!std::strcmp(key, "seg")) {
        if (value < 1) return fail(MPX_ERR_INVALID, "seg must be >= 1");
        b->seg = value;
    } else if (!std::strcmp(key, "blocks_per_cu")) {
        if (value < 1 || value > 16) return fail(MPX_ERR_INVALID, "blocks_per_cu must be 1..16");
        b->blocks_per_cu = value;
    } else if (!std::strcmp(key, "graph_steps")) {
        if (value < 0 || value > 1000) return fail(MPX_ERR_INVALID, "graph_steps must be 0..1000");
        b->graph_steps = value;
        b->drop_graphs();
        return MPX_OK;
    } else
        return fail(MPX_ERR_NAME, std::string("unknown batch tuning key ") + key);
    return batch_prepare(b);
}

int mpx_batch_time_steps(mpx_batch *b, int nsteps, double *ms_per_step)
{
    if (!b || !ms_per_step || nsteps < 1) return fail(MPX_ERR_INVALID, "invalid arguments");
    hipEvent_t e0, e1;
    MPX_HIP(hipEventCreate(&e0));
    MPX_HIP(hipEventCreate(&e1));
    (void)hipEventRecord(e0, b->stream);
    int rc = batch_step(b, nsteps, true);
    (void)hipEventRecord(e1, b->stream);
    if (rc == MPX_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(MPX_ERR_HIP, "event sync failed");
    float ms = 0;
    if (rc == MPX_OK) (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_step = ms / nsteps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Autotuner: times every candidate launch shape -- the fused structure with 1, 2 or 4 B planes
// per thread and 4 or 8 blocks per CU, the sweep with block shapes 64x16 / 64x8 / 64x4, x
// segments of 8 / 16 / 32 planes and 1 or 2 blocks per CU, and the x-march pass with segments
// of 8 / 16 / 32 planes and tiles of 3 or 7 rows -- for `steps_per_trial` steps on the batch
// itself, keeps the fastest, then clears every run back to its initial all-zero state at time
// 0.  Allowed only before the first step.  result = {structure, a, b, c}: fused
// {0, march_b, blocks_per_cu, 0}, sweep {1, variant, seg, blocks_per_cu}, x-march
// {2, pass_seg, rows, 0}.
int mpx_batch_autotune(mpx_batch *b, int steps_per_trial, int result[4], double *best_ms)
{
    if (!b || steps_per_trial < 1) return fail(MPX_ERR_INVALID, "invalid arguments");
    if (b->time != 0) return fail(MPX_ERR_INVALID, "autotune runs before the first step only");
    struct Cand {
        int structure, p0, p1, p2;
    };
    std::vector<Cand> cands;
    for (int mb : {1, 2, 4})
        for (int bpc : {4, 8}) cands.push_back({0, mb, bpc, 0});
    for (int v : {2, 0, 1})
        for (int sg : {8, 16, 32})
            for (int bpc : {1, 2}) cands.push_back({1, v, sg, bpc});
    for (int sg : {8, 16, 32})
        for (int rows : {3, 7}) cands.push_back({2, sg, rows, 0});
    auto apply = [&](const Cand &c) {
        b->structure = c.structure;
        if (c.structure == 0) {
            b->march_b = c.p0;
            b->fused_bpc = c.p1;
        } else if (c.structure == 1) {
            b->variant = c.p0;
            b->seg = c.p1;
            b->blocks_per_cu = c.p2;
        } else {
            b->pass_seg = c.p0;
            b->pass_rows = c.p1;
        }
    };
    double best = 1e300;
    Cand bc = b->structure == 0   ? Cand{0, b->march_b, b->fused_bpc, 0}
              : b->structure == 1 ? Cand{1, b->variant, b->seg, b->blocks_per_cu}
                                  : Cand{2, b->pass_seg, b->pass_rows, 0};
    hipEvent_t e0, e1;
    MPX_HIP(hipEventCreate(&e0));
    MPX_HIP(hipEventCreate(&e1));
    int rc = MPX_OK;
    for (const Cand &c : cands) {
        apply(c);
        if ((rc = batch_prepare(b))) break;
        if ((rc = batch_step(b, 1, false))) break;  // warm-up
        (void)hipEventRecord(e0, b->stream);
        if ((rc = batch_step(b, steps_per_trial, false))) break;
        (void)hipEventRecord(e1, b->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(MPX_ERR_HIP, "event sync failed");
            break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            bc = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    for (mpx_solver *s : b->runs)
        if ((rc = reset_state(s, b->stream))) return rc;
    MPX_HIP(hipMemsetAsync(b->d_time, 0, sizeof(int), b->stream));
    MPX_HIP(hipStreamSynchronize(b->stream));
    b->time = 0;
    apply(bc);
    if ((rc = batch_prepare(b))) return rc;
    if (result) {
        result[0] = bc.structure;
        result[1] = bc.p0;
        result[2] = bc.p1;
        result[3] = bc.p2;
    }
    if (best_ms) *best_ms = best / steps_per_trial;
    return MPX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// gradient of a plane objective by the discrete adjoint (kernels in mpx_adjoint.hip)
// ---------------------------------------------------------------------------------------

namespace {

size_t face_count(const CpmlFace &F) { return (size_t)(F.hi[0] - F.lo[0] + 1) * F.n1 * F.n2; }

// device buffers owned by one adjoint call
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *q : p) (void)hipFree(q);
    }
    double *get(size_t count, hipStream_t st)
    {
        void *q = nullptr;
        if (hipMalloc(&q, (count ? count : 1) * sizeof(double)) != hipSuccess) return nullptr;
        p.push_back(q);
        if (hipMemsetAsync(q, 0, (count ? count : 1) * sizeof(double), st) != hipSuccess) return nullptr;
        return (double *)q;
    }
};

// the forward state of the fused float64 path: E, B, the pole accumulators and the CPML
// auxiliaries
struct StateRef {
    std::vector<std::pair<void *, size_t>> arr;  // (device pointer, doubles)
};

StateRef state_of(mpx_solver *s)
{
    StateRef r;
    for (int c = 0; c < 3; ++c) {
        r.arr.push_back({s->E[c], s->cells});
        r.arr.push_back({s->B[c], s->cells});
    }
    for (int p = 0; p < s->np; ++p) r.arr.push_back({s->psi[p][0], mpx::psi_block_elems(s->cells) * 2});
    for (const CpmlFace &F : s->efaces) {
        r.arr.push_back({F.psi0, face_count(F)});
        r.arr.push_back({F.psi1, face_count(F)});
    }
    for (const CpmlFace &F : s->hfaces) {
        r.arr.push_back({F.psi0, face_count(F)});
        r.arr.push_back({F.psi1, face_count(F)});
    }
    return r;
}

}  // namespace

extern "C" {

int mpx_adjoint_grad(mpx_solver *s, const mpx_objective *obj, int nsteps, int K, double *J, double *dJ_deps,
                     double *dJ_dchi)
{
    if (!s || !obj || !J || nsteps < 1 || K < 1) return fail(MPX_ERR_INVALID, "invalid adjoint arguments");
    MPX_NOT_IN_BATCH(s);
    if (!s->plrc || s->esz != 8 || s->path != MPX_PATH_FUSED || s->plasma || s->pcur)
        return fail(MPX_ERR_UNSUPPORTED, "the adjoint covers the float64 fused FDTD_PLRC update without plasma");
    if (!s->periodic[0] || !s->periodic[1] || !s->periodic[2])
        return fail(MPX_ERR_UNSUPPORTED, "the adjoint needs the periodic boundary");
    for (const std::vector<Current> *list : {&s->ecur, &s->hcur})
        for (const Current &c : *list)
            if (c.gauss)
                return fail(MPX_ERR_UNSUPPORTED, "the adjoint does not take gauss_inject (its sample blocks advance "
                                                 "with the forward steps only)");
    if (obj->plane_k < 1 || obj->plane_k > s->g.n[2]) return fail(MPX_ERR_INVALID, "objective plane out of range");
    const bool grad = dJ_deps || dJ_dchi;
    hipStream_t st = s->stream;
    DevBufs mem;
    double *dJ = mem.get(1, st);
    if (!dJ) return fail(MPX_ERR_NOMEM, "adjoint allocation failed");
    const int t0 = s->time;
    const int nseg = (nsteps + K - 1) / K;
    StateRef cur = state_of(s);
    auto snapshot = [&](std::vector<double *> &dst) -> int {
        if (dst.empty())
            for (auto &a : cur.arr) {
                double *q = mem.get(a.second, st);
                if (!q) return fail(MPX_ERR_NOMEM, "checkpoint allocation failed");
                dst.push_back(q);
            }
        for (size_t q = 0; q < cur.arr.size(); ++q)
            MPX_HIP(hipMemcpyAsync(dst[q], cur.arr[q].first, cur.arr[q].second * 8, hipMemcpyDeviceToDevice, st));
        return MPX_OK;
    };
    auto restore = [&](const std::vector<double *> &src) -> int {
        for (size_t q = 0; q < cur.arr.size(); ++q)
            MPX_HIP(hipMemcpyAsync(cur.arr[q].first, src[q], cur.arr[q].second * 8, hipMemcpyDeviceToDevice, st));
        return MPX_OK;
    };
    std::vector<std::vector<double *>> ckpt(grad ? nseg : 0);
    int rc;
    // forward pass: objective after every step, a checkpoint at the start of every segment
    for (int n = 0; n < nsteps; ++n) {
        if (grad && n % K == 0 && (rc = snapshot(ckpt[n / K]))) return rc;
        if ((rc = step_once(s))) return rc;
        double *E[3] = {(double *)s->E[0], (double *)s->E[1], (double *)s->E[2]};
        mpx::launch_plane_objective(s->g, E, obj->plane_k, obj->weight, dJ, st);
        MPX_LAUNCHED("plane objective");
    }
    MPX_HIP(hipMemcpyAsync(J, dJ, sizeof(double), hipMemcpyDeviceToHost, st));
    MPX_HIP(hipStreamSynchronize(st));
    if (!grad) return MPX_OK;
    std::vector<double *> fin;
    if ((rc = snapshot(fin))) return rc;

    // adjoint state, zero after the last step
    mpx::AdjArgs a{};
    a.g = s->g;
    a.dt = s->d.dt;
    a.chi = s->nonlin ? s->d.chi : 0.0;  // the linear core has no Kerr term
    a.eps = s->d.eps;
    a.epsf = s->d_eps;
    a.np = s->np;
    a.pc = s->pc;
    a.sumChi0 = s->sumChi0;
    a.sumXi0 = s->sumXi0;
    for (int p = 0; p < s->np; ++p)
        for (int q = 0; q < 3; ++q) {
            a.lpsi[p][q] = (double2 *)mem.get(2 * s->cells, st);
            if (!a.lpsi[p][q]) return fail(MPX_ERR_NOMEM, "adjoint allocation failed");
        }
    for (int d = 0; d < 3; ++d) {
        a.ikdE[d] = s->d_ikdE[d];
        a.ikdH[d] = s->d_ikdH[d];
        a.lE[d] = mem.get(s->cells, st);
        a.lB[d] = mem.get(s->cells, st);
    }
    for (int q = 0; q < 6; ++q) {
        a.dEb[q] = mem.get(s->cells, st);
        a.dBb[q] = mem.get(s->cells, st);
    }
    a.epsb = mem.get(s->cells, st);
    a.chib = mem.get(s->cells, st);
    a.kobj = obj->plane_k;
    a.wobj = obj->weight;
    a.nefaces = (int)s->efaces.size();
    a.nhfaces = (int)s->hfaces.size();
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < 2; ++q) a.efbd[d][q] = a.hfbd[d][q] = -1;
    for (size_t f = 0; f < s->efaces.size(); ++f) {
        a.eface[f] = s->efaces[f];
        for (int w = 0; w < 2; ++w) a.lpsiE[f][w] = mem.get(face_count(s->efaces[f]), st);
    }
    for (size_t f = 0; f < s->hfaces.size(); ++f) {
        a.hface[f] = s->hfaces[f];
        for (int w = 0; w < 2; ++w) a.lpsiH[f][w] = mem.get(face_count(s->hfaces[f]), st);
    }
    for (const Current &c : s->ecur)
        if (c.face >= 0) a.efbd[dim_of(c.dir)][is_high(c.dir) ? 0 : 1] = c.face;
    for (const Current &c : s->hcur)
        if (c.face >= 0) a.hfbd[dim_of(c.dir)][is_high(c.dir) ? 0 : 1] = c.face;
    // forward E of one segment: (K + 1) planes of E
    std::vector<std::array<double *, 3>> hist(K + 1);
    for (auto &h : hist)
        for (int c = 0; c < 3; ++c) h[c] = mem.get(s->cells, st);
    for (int d = 0; d < 3; ++d)
        if (!a.lE[d] || !a.lB[d]) return fail(MPX_ERR_NOMEM, "adjoint allocation failed");
    for (int q = 0; q < 6; ++q)
        if (!a.dEb[q] || !a.dBb[q]) return fail(MPX_ERR_NOMEM, "adjoint allocation failed");
    if (!a.epsb || !a.chib) return fail(MPX_ERR_NOMEM, "adjoint allocation failed");
    for (auto &h : hist)
        for (int c = 0; c < 3; ++c)
            if (!h[c]) return fail(MPX_ERR_NOMEM, "adjoint history allocation failed");

    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int len = std::min(K, nsteps - seg * K);
        if ((rc = restore(ckpt[seg]))) return rc;
        s->time = t0 + seg * K;
        auto keepE = [&](int m) -> int {
            for (int c = 0; c < 3; ++c)
                MPX_HIP(hipMemcpyAsync(hist[m][c], s->E[c], s->cells * 8, hipMemcpyDeviceToDevice, st));
            return MPX_OK;
        };
        if ((rc = keepE(0))) return rc;
        for (int m = 1; m <= len; ++m) {
            if ((rc = step_once(s))) return rc;
            if ((rc = keepE(m))) return rc;
        }
        for (int m = len - 1; m >= 0; --m) {
            for (int c = 0; c < 3; ++c) {
                a.E0[c] = hist[m][c];
                a.E1[c] = hist[m + 1][c];
            }
            mpx::launch_adjoint_step(a, st);
            MPX_LAUNCHED("adjoint step");
        }
    }
    if ((rc = restore(fin))) return rc;
    s->time = t0 + nsteps;
    // gradients to the host: dJ/dEps in full-storage layout, dJ/dchi summed in a fixed order
    const GridDims &g = s->g;
    const size_t nx2 = g.n[0] + 2, ny2 = g.n[1] + 2, nz2 = g.n[2] + 2;
    std::vector<double> hb(s->cells);
    if (dJ_deps) {
        MPX_HIP(hipMemcpyAsync(hb.data(), a.epsb, s->cells * 8, hipMemcpyDeviceToHost, st));
        MPX_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < nx2; ++i)
            for (size_t j = 0; j < ny2; ++j)
                for (size_t k = 0; k < nz2; ++k)
                    dJ_deps[(i * ny2 + j) * nz2 + k] = hb[g.at((int)i, (int)j, (int)k)];
    }
    if (dJ_dchi) {
        MPX_HIP(hipMemcpyAsync(hb.data(), a.chib, s->cells * 8, hipMemcpyDeviceToHost, st));
        MPX_HIP(hipStreamSynchronize(st));
        double sum = 0.0;
        for (int i = 1; i <= g.n[0]; ++i)
            for (int j = 1; j <= g.n[1]; ++j)
                for (int k = 1; k <= g.n[2]; ++k) sum += hb[g.at(i, j, k)];
        *dJ_dchi = s->nonlin ? sum : 0.0;
    }
    MPX_HIP(hipStreamSynchronize(st));
    return MPX_OK;
}

}  // extern "C"
