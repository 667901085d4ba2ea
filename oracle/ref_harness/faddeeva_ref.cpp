// C entry point around the reference's own Faddeeva_2 (src/specfunc.cpp:57-157), compiled
// from the reference tree in place.  Test infrastructure only: it produces golden vectors.
#include "specfunc.h"

extern "C" void ref_faddeeva(double re, double im, double *out)
{
    std::complex<double> w = Faddeeva_2(std::complex<double>(re, im));
    out[0] = w.real();
    out[1] = w.imag();
}
