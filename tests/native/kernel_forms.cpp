// Host evaluation of the library's pair-term forms (common.hpp), for the CPU tests: for each q
// on stdin prints W and -q dW/dq - 3W from the reference form (kernWdW, FP64) and from the
// mixed form (kernVD in FP32 at q rounded to FP32, constants applied in FP64).
#include "../../cosmosphere_amd/csrc/common.hpp"

#include <cstdio>

int main()
{
    double q;
    while (std::scanf("%lf", &q) == 1)
    {
        double w, dw;
        csph::kernWdW<double>(q * q, q, w, dw);
        float V, qD;
        const float qf = (float)q;
        csph::kernVD<float>(qf, qf * qf, V, qD);
        const double wm = csph::kWw * (double)V;
        const double gm = csph::kGp * (double)qD - csph::kGw * (double)V;
        std::printf("%.17g %.17g %.17g %.17g\n", w, -q * dw - 3.0 * w, wm, gm);
    }
    return 0;
}
