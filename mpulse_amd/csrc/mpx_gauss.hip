// GaussInject on the device: GaussInjectSourceFunc (src/gaussinject.cpp:56-482).
//
// Every Nt steps the sheet samples of the next block are rebuilt: the pulse is set up as a
// spectrum A(kx, ky, omega) (a Gaussian around k0 times a polarisation factor and a phase for
// the sample positions and the time offset), transformed backward over omega (only the Nt time
// samples of the block are kept), then forward over y and over x, with the index shifts and
// scale factors of transformT / transformY / transformX.  The transforms are batched hipFFT
// double-complex transforms with FFTW's sign conventions (backward +i, forward -i,
// unnormalised); the spectrum is formed in slabs of x so that its buffer stays below 1 GiB.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>

#include "mpx_gauss.h"

namespace mpx {

namespace {

enum { COMP_EX = 0, COMP_EY = 1, COMP_BX = 2 };

struct SpectrumArgs {
    int SX, SY, ST, SXH, SYH, STH;
    double BoxX, BoxY, BoxT, length, width, k0, Time;
    double sx, sy, sz;
    int comp;
};

// GaussFunc components, src/gaussinject.h:150-200
__device__ double gauss_comp(int comp, double x, double y, double z)
{
    double kp = sqrt(x * x + z * z), k = sqrt(x * x + y * y + z * z);
    if (comp == COMP_EX) {
        double D = k * kp;
        return D == 0 ? 0 : (-x * y / D);
    }
    if (comp == COMP_EY) return k == 0 ? 0 : (kp / k);
    return kp == 0 ? 0 : (-z / kp);
}

// S[(il SY + j) ST + k] = uo[k] = A(i0 + il, j, (k + SizeTH) % SizeT): the spectrum of fillEx /
// fillEy / fillBx in the input order of transformT
__global__ __launch_bounds__(256) void k_gauss_spectrum(SpectrumArgs a, int i0, int slab, double2 *S)
{
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t total = (size_t)slab * a.SY * a.ST;
    if (idx >= total) return;
    const int k = (int)(idx % a.ST);
    const int j = (int)((idx / a.ST) % a.SY);
    const int i = i0 + (int)(idx / ((size_t)a.ST * a.SY));
    const int m = (k + a.STH) % a.ST;
    const double kx = 2 * M_PI * (i - a.SXH) / (a.BoxX);
    const double ky = 2 * M_PI * (j - a.SYH) / (a.BoxY);
    const double om = 2 * M_PI * (m - a.STH) / (a.BoxT);
    double K = om * om - ky * ky - kx * kx;
    if (fabs(K) < 1e-9) K = 0;
    double2 v = make_double2(0.0, 0.0);
    if (!(K < 0)) {
        const double kz = sqrt(K) * (om > 0 ? 1 : (om < 0 ? -1 : 0));
        const double zoff = kz - a.k0;
        const double g = exp(-0.25 * zoff * zoff * a.length * a.length - 0.25 * (kx * kx + ky * ky) * a.width * a.width);
        const double amp = gauss_comp(a.comp, kx, ky, kz) * g;
        const double ph = -(kx * a.sx + ky * a.sy + kz * a.sz) + 0.5 * M_PI;
        const double th = ph + a.Time * om;  // amp exp(-i th)
        v = make_double2(amp * cos(th), -(amp * sin(th)));
    }
    S[idx] = v;
}

// transformT's kept samples, placed as transformY's input:
// P[(i Nt + kh) SY + (j + SizeYH) % SizeY] = ft S[(il SY + j) ST + (kh + blockmin + SizeTH) % SizeT]
__global__ __launch_bounds__(256) void k_gauss_keep(const double2 *S, double2 *P, int SY, int ST, int SYH, int STH,
                                                     int Nt, int blockmin, int i0, int slab, double ft)
{
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t total = (size_t)slab * SY * Nt;
    if (idx >= total) return;
    const int kh = (int)(idx % Nt);
    const int j = (int)((idx / Nt) % SY);
    const int il = (int)(idx / ((size_t)Nt * SY));
    const double2 s = S[((size_t)il * SY + j) * ST + (kh + blockmin + STH) % ST];
    P[((size_t)(i0 + il) * Nt + kh) * SY + (j + SYH) % SY] = make_double2(ft * s.x, ft * s.y);
}

// transformY's output placed as transformX's input:
// Q[(j Nt + kh) SX + (i + SizeXH) % SizeX] = fy P[(i Nt + kh) SY + (j + SizeYH) % SizeY]
__global__ __launch_bounds__(256) void k_gauss_turn(const double2 *P, double2 *Q, int SX, int SY, int SXH, int SYH,
                                                     int Nt, double fy)
{
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t total = (size_t)SX * SY * Nt;
    if (idx >= total) return;
    const int i = (int)(idx % SX);
    const int kh = (int)((idx / SX) % Nt);
    const int j = (int)(idx / ((size_t)SX * Nt));
    const double2 p = P[((size_t)i * Nt + kh) * SY + (j + SYH) % SY];
    Q[((size_t)j * Nt + kh) * SX + (i + SXH) % SX] = make_double2(fy * p.x, fy * p.y);
}

// transformX's real part: F[(kh SX + i) SY + j] = fx Re Q[(j Nt + kh) SX + (i + SizeXH) % SizeX]
__global__ __launch_bounds__(256) void k_gauss_final(const double2 *Q, double *F, int SX, int SY, int SXH, int Nt,
                                                      double fx)
{
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t total = (size_t)SX * SY * Nt;
    if (idx >= total) return;
    const int j = (int)(idx % SY);
    const int i = (int)((idx / SY) % SX);
    const int kh = (int)(idx / ((size_t)SX * SY));
    F[idx] = fx * Q[((size_t)j * Nt + kh) * SX + (i + SXH) % SX].x;
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

bool plan1d(hipfftHandle *p, int n, int batch, hipStream_t s)
{
    int len = n;
    if (hipfftPlanMany(p, 1, &len, nullptr, 1, n, nullptr, 1, n, HIPFFT_Z2Z, batch) != HIPFFT_SUCCESS) return false;
    return hipfftSetStream(*p, s) == HIPFFT_SUCCESS;
}

}  // namespace

GaussInjectDev::~GaussInjectDev()
{
    for (hipfftHandle *p : {&planT, &planTr, &planY, &planX})
        if (*p) hipfftDestroy(*p);
    for (void *q : {(void *)F1, (void *)F2, (void *)S, (void *)P, (void *)Q})
        if (q) (void)hipFree(q);
}

int GaussInjectDev::setup(const mpx_factory &f, int isH_, int nx, int ny, double dx, double dy, double dt, int lowz,
                          int highz, hipStream_t s)
{
    // GaussInjectSourceFunc::setParam, src/gaussinject.cpp:56-106; initSourceFunc :108-150
    isH = isH_;
    SizeT = f.gsize_t;
    length = f.length;
    width = f.width;
    k0 = f.k0;
    Shift = f.gshift;
    Time = f.gtime;
    amp = f.amp;
    eps = f.eps;
    toff = f.toff;
    const double lightspeed = std::sqrt(1 / eps);
    SizeX = nx + 2;
    SizeY = ny + 2;
    SizeTH = SizeT / 2;
    SizeXH = SizeX / 2;
    SizeYH = SizeY / 2;
    DX = dx;
    DY = dy;
    DT = dt * lightspeed;
    BoxX = SizeX * DX;
    BoxY = SizeY * DY;
    BoxT = SizeT * DT;
    Nt = 1 << (int)std::round(std::log((double)(highz - lowz)) / std::log(2.0) + 1e-6);
    blockToff = toff;
    stream = s;
    const size_t row = (size_t)SizeY * SizeT * sizeof(double2);
    slab = (int)std::max<size_t>(1, std::min<size_t>((size_t)SizeX, ((size_t)1 << 30) / row));
    const size_t nF = (size_t)Nt * SizeX * SizeY;
    if (hipMalloc(&F1, nF * sizeof(double)) != hipSuccess || hipMalloc(&F2, nF * sizeof(double)) != hipSuccess ||
        hipMalloc(&S, (size_t)slab * row) != hipSuccess || hipMalloc(&P, nF * sizeof(double2)) != hipSuccess ||
        hipMalloc(&Q, nF * sizeof(double2)) != hipSuccess)
        return 1;
    if (hipMemsetAsync(F1, 0, nF * sizeof(double), s) != hipSuccess ||
        hipMemsetAsync(F2, 0, nF * sizeof(double), s) != hipSuccess)
        return 1;
    if (!plan1d(&planT, SizeT, slab * SizeY, s)) return 2;
    if (SizeX % slab && !plan1d(&planTr, SizeT, (SizeX % slab) * SizeY, s)) return 2;
    if (!plan1d(&planY, SizeY, SizeX * Nt, s) || !plan1d(&planX, SizeX, SizeY * Nt, s)) return 2;
    return 0;
}

int GaussInjectDev::fill(int comp, double sx, double sy, double sz, double *F)
{
    SpectrumArgs a;
    a.SX = SizeX, a.SY = SizeY, a.ST = SizeT, a.SXH = SizeXH, a.SYH = SizeYH, a.STH = SizeTH;
    a.BoxX = BoxX, a.BoxY = BoxY, a.BoxT = BoxT, a.length = length, a.width = width, a.k0 = k0, a.Time = Time;
    a.sx = sx, a.sy = sy, a.sz = sz, a.comp = comp;
    const int blockmin = blockCount * Nt;
    const double ft = 2 * M_PI * length / BoxT, fy = width / BoxY, fx = width / BoxX;
    for (int i0 = 0; i0 < SizeX; i0 += slab) {
        const int n = std::min(slab, SizeX - i0);
        hipLaunchKernelGGL(k_gauss_spectrum, dim3(blocks_for((size_t)n * SizeY * SizeT)), dim3(256), 0, stream, a,
                           i0, n, (double2 *)S);
        if (hipGetLastError() != hipSuccess) return 1;
        if (hipfftExecZ2Z(n == slab ? planT : planTr, (hipfftDoubleComplex *)S, (hipfftDoubleComplex *)S,
                          HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
            return 2;
        hipLaunchKernelGGL(k_gauss_keep, dim3(blocks_for((size_t)n * SizeY * Nt)), dim3(256), 0, stream,
                           (const double2 *)S, (double2 *)P, SizeY, SizeT, SizeYH, SizeTH, Nt, blockmin, i0, n, ft);
        if (hipGetLastError() != hipSuccess) return 1;
    }
    if (hipfftExecZ2Z(planY, (hipfftDoubleComplex *)P, (hipfftDoubleComplex *)P, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        return 2;
    const size_t nF = (size_t)SizeX * SizeY * Nt;
    hipLaunchKernelGGL(k_gauss_turn, dim3(blocks_for(nF)), dim3(256), 0, stream, (const double2 *)P, (double2 *)Q,
                       SizeX, SizeY, SizeXH, SizeYH, Nt, fy);
    if (hipGetLastError() != hipSuccess) return 1;
    if (hipfftExecZ2Z(planX, (hipfftDoubleComplex *)Q, (hipfftDoubleComplex *)Q, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        return 2;
    hipLaunchKernelGGL(k_gauss_final, dim3(blocks_for(nF)), dim3(256), 0, stream, (const double2 *)Q, F, SizeX,
                       SizeY, SizeXH, Nt, fx);
    return hipGetLastError() != hipSuccess ? 1 : 0;
}

int GaussInjectDev::setTime(int time)
{
    // GaussInjectSourceFunc::setTime, src/gaussinject.cpp:152-182
    if (time < toff + blockCount * Nt) return 0;
    if (time >= toff + SizeT) {
        active = 0;
        return 0;
    }
    int rc;
    if (!isH) {
        if ((rc = fill(COMP_BX, 0, 0.5 * DX, Shift + 0.5 * DX, F1))) return rc;
    } else {
        Time -= 0.5 * DT;
        if ((rc = fill(COMP_EX, 0.5 * DX, 0, Shift, F1))) return rc;
        if ((rc = fill(COMP_EY, 0, 0.5 * DX, Shift, F2))) return rc;
        Time += 0.5 * DT;
    }
    blockToff = toff + blockCount * Nt;
    blockCount++;
    active = 1;
    return 0;
}

void GaussInjectDev::params(GaussSheetParams &p) const
{
    p.F1 = F1;
    p.F2 = F2;
    p.active = active;
    p.blockToff = blockToff;
    p.amp = amp;
    p.sqeps = std::sqrt(eps);
}

}  // namespace mpx
