/*
 * cxxapi.hip -- the C++ interface of include/cosmosphere.hpp (the reference's solveDensH,
 * KernelMode, DensTimings, gpuState, computeForces and ForceTimings: include/density.hpp:20-91,
 * include/force.hpp:18-70) over the same device state as the C entry points.
 */
#include "../../include/cosmosphere.hpp"
#include "solver.hpp"

#include <cstdio>
#include <cstdlib>

struct GpuState
{
    csph::State* st;
};

GpuState& gpuState()
{
    static GpuState g{&csph::state()};
    return g;
}

DensTimings solveDensH(int n, double pmass, KernelMode mode, bool withGradients, const double* periodicBox,
                       double tolh, double hfact)
{
    csph::State& s = *gpuState().st;
    if (s.sizedFor != n)
    {
        std::fprintf(stderr,
                     "FATAL: solveDensH called with n=%d but the device arrays are sized for %d "
                     "(cosmo_arrays_init not called, or called with a different count)\n",
                     n, s.sizedFor);
        std::abort();
    }
    csph::ensureStream(s);
    s.n = n;
    s.nSets = 1;
    s.hSetFirst.assign(1, 0);
    s.hSetN.assign(1, n);
    const csph::Timings t =
        csph::solveDensH(s, pmass, mode == KernelMode::WARP_PER_LEAF ? 1 : 0, withGradients, periodicBox, tolh, hfact);
    DensTimings d{};
    d.upload = s.uploadSeconds;
    d.bboxAndSetup = t.bbox;
    d.keysAndSort = t.keys + t.sort + t.permute;
    d.treeBuild = t.tree;
    d.nodeCenters = t.nodes;
    d.jleafBuild = t.jbuild;
    d.densKernel = t.nrkern + t.compact;
    d.gradJleafBuild = t.gjbuild;
    d.gradKernel = t.gradkern;
    d.download = s.downloadSeconds;
    d.itersRun = t.iters;
    d.nParticles = n;
    d.nLeavesOut = t.nLeaves;
    d.kernelMode = mode;
    return d;
}

void computeForces(GpuState& g, double pmass, double beta, double alphau, bool discViscosity, bool pdvHeating,
                   bool shockHeating, ForceTimings& ft)
{
    csph::State& s = *g.st;
    if (!s.readyForForce(s.sizedFor))
    {
        std::fprintf(stderr,
                     "FATAL: computeForces called with no GPU density solve for this particle set "
                     "(n=%d state.ngas=%d token=%llu)\n",
                     s.sizedFor, s.n, (unsigned long long)s.token);
        std::abort();
    }
    csph::ensureStream(s);
    const csph::ForceTimings f = csph::computeForces(s, pmass, beta, alphau, discViscosity, pdvHeating, shockHeating);
    ft.hmaxUpsweep = f.upsweep;
    ft.jleafBuild = f.jbuild;
    ft.upload = s.uploadSeconds;
    ft.prep = f.prep;
    ft.kernel = f.kernel;
    ft.download = s.downloadSeconds;
}
