// A C++ caller written against the reference's public headers (include/density.hpp,
// include/force.hpp), here include/cosmosphere.hpp, linked against libcosmoSPHere.a with the
// host compiler the way Phantom links it.  It solves a jittered lattice through solveDensH
// (defaults of density.hpp:85-91) and through the C extension cosmo_ext_solve_dens_h, then
// runs computeForces(gpuState(), ...) and force_gpu_c on the same tree, and prints whether the
// two routes agree bit for bit.
#include "cosmosphere.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

int main(int argc, char** argv)
{
    const int m = argc > 1 ? std::atoi(argv[1]) : 16;
    const int n = m * m * m;
    const double dx = 1.0 / m, pmass = 1.0 / n;
    std::vector<double> pos(3 * (size_t)n), h(n, 1.2 * dx), v(3 * (size_t)n), th(4 * (size_t)n);
    unsigned long long r = 12345;
    auto rnd = [&r]() { r = r * 6364136223846793005ull + 1442695040888963407ull; return (double)(r >> 11) / 9007199254740992.0; };
    for (int i = 0; i < n; ++i)
    {
        const int a = i % m, b = (i / m) % m, c = i / (m * m);
        pos[i] = (a + 0.5 + 0.2 * (rnd() - 0.5)) * dx;
        pos[n + i] = (b + 0.5 + 0.2 * (rnd() - 0.5)) * dx;
        pos[2 * (size_t)n + i] = (c + 0.5 + 0.2 * (rnd() - 0.5)) * dx;
        for (int k = 0; k < 3; ++k) v[(size_t)k * n + i] = rnd() - 0.5;
        th[i] = 1.0; th[n + i] = 1.0; th[2 * (size_t)n + i] = 1.0; th[3 * (size_t)n + i] = 1.0;
    }
    cosmo_arrays_init(n);
    auto send = [&]() {
        cosmo_upload(COSMO_POS, pos.data(), n);
        cosmo_upload(COSMO_HSML, h.data(), n);
        cosmo_upload(COSMO_VEL, v.data(), n);
        cosmo_upload(COSMO_ACCEL, std::vector<double>(3 * (size_t)n, 0.0).data(), n);
    };
    std::vector<double> h1(n), h2(n), d1(2 * (size_t)n), d2(2 * (size_t)n);
    send();
    const DensTimings t = solveDensH(n, pmass);
    cosmo_download(COSMO_HSML, h1.data(), n);
    cosmo_download(COSMO_DENS_OUT, d1.data(), n);
    send();
    const int iters = cosmo_ext_solve_dens_h(n, pmass, 0, 0, 0, nullptr, 1e-4, -1.0);
    cosmo_download(COSMO_HSML, h2.data(), n);
    cosmo_download(COSMO_DENS_OUT, d2.data(), n);
    const bool densSame = std::memcmp(h1.data(), h2.data(), n * sizeof(double)) == 0 &&
                          std::memcmp(d1.data(), d2.data(), 2 * (size_t)n * sizeof(double)) == 0;
    std::printf("density n=%d iters=%d/%d leaves=%d mode=%d same=%d\n", t.nParticles, t.itersRun, iters, t.nLeavesOut,
                (int)t.kernelMode, (int)densSame);

    std::vector<double> f1(6 * (size_t)n), f2(6 * (size_t)n);
    send();
    densityiterate_gpu_c(n, pmass, 0, nullptr, 1e-4, -1.0);
    cosmo_upload_sorted(COSMO_THERMO, th.data(), n);
    ForceTimings ft;
    computeForces(gpuState(), pmass, 2.0, 1.0, false, true, true, ft);
    cosmo_download(COSMO_FORCE_OUT, f1.data(), n);
    force_gpu_c(n, pmass, 2.0, 1.0, 0, 1, 1);
    cosmo_download(COSMO_FORCE_OUT, f2.data(), n);
    const bool forceSame = std::memcmp(f1.data(), f2.data(), 6 * (size_t)n * sizeof(double)) == 0;
    double fmax = 0.0;
    for (size_t i = 0; i < 3 * (size_t)n; ++i) fmax = std::fmax(fmax, std::fabs(f1[i]));
    std::printf("force kernel_s=%.3e same=%d fmax=%.6e\n", ft.kernel, (int)forceSame, fmax);
    return (densSame && forceSame && t.itersRun == iters && t.nParticles == n) ? 0 : 1;
}
