// GaussInject sheet source on the device (GaussInjectSourceFunc, src/gaussinject.cpp:56-482):
// the block schedule of setTime on the host, the spectrum and the three transforms on the
// solver's stream.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include "../../include/mpulse_hip.h"
#include "mpx_types.h"

namespace mpx {

class GaussInjectDev {
  public:
    ~GaussInjectDev();
    // returns 0, 1 (device memory), 2 (FFT plan)
    int setup(const mpx_factory &f, int isH, int nx, int ny, double dx, double dy, double dt, int lowz, int highz,
              hipStream_t s);
    // GaussInjectSourceFunc::setTime: rebuilds the samples when `time` enters a new block
    int setTime(int time);
    void params(GaussSheetParams &p) const;
    int Nt = 0, active = 0, blockToff = 0, blockCount = 0;

  private:
    int fill(int comp, double sx, double sy, double sz, double *F);
    int isH = 0, SizeX = 0, SizeY = 0, SizeT = 0, SizeXH = 0, SizeYH = 0, SizeTH = 0, toff = 0, slab = 1;
    double BoxX = 0, BoxY = 0, BoxT = 0, DX = 0, DY = 0, DT = 0;
    double length = 0, width = 0, k0 = 0, Shift = 0, Time = 0, amp = 0, eps = 1;
    double *F1 = nullptr, *F2 = nullptr;   // [kh][i][j] samples of the block
    void *S = nullptr, *P = nullptr, *Q = nullptr;  // spectrum slab, transform buffers (double2)
    hipfftHandle planT = 0, planTr = 0, planY = 0, planX = 0;
    hipStream_t stream = nullptr;
};

}  // namespace mpx
