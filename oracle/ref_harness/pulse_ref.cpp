// Driver around the reference's standalone short-pulse harness (test/testshortpulse.cpp),
// compiled in place: its main() is renamed and its ShortPulseInjectSourceFunc::Efunc/Bfunc
// are evaluated at the points read from stdin, printed with 17 significant digits.
// Input lines:  <E|X|Y> length width om0 phase eps x y z t
// Output lines: re im
// Test infrastructure only; run it in a scratch directory (the included file opens
// debug1.out..debug6.out in the working directory at start-up).
#define main testshortpulse_main
#include "testshortpulse.cpp"
#undef main

#include <cstdio>

int main()
{
    char kind[8];
    double length, width, om0, phase, eps, x, y, z, t;
    while (std::scanf("%7s %lf %lf %lf %lf %lf %lf %lf %lf %lf", kind, &length, &width, &om0,
                      &phase, &eps, &x, &y, &z, &t) == 10) {
        ShortPulseInjectSourceFunc src;
        src.setParam(length, width, om0, 0.0, phase, 1.0, eps);
        Complex v;
        if (kind[0] == 'E')
            v = src.Efunc(x, y, z, t);
        else
            v = src.Bfunc(x, y, z, t, kind[0] == 'X');
        std::printf("%.17g %.17g\n", v.real(), v.imag());
    }
    return 0;
}
