/*
 * cosmosphere.hpp -- the C++ interface of the library, for callers written against the
 * reference's public headers (include/density.hpp:20-91 and include/force.hpp:18-70): the
 * same names, argument meaning, defaults and mangled symbols.  Includable from translation
 * units that are not compiled for the GPU (no HIP header is pulled in).
 *
 * Everything runs on the library's process-wide device state: positions and h are put on
 * the device with cosmo_upload (cosmosphere.h) and results fetched with cosmo_download.
 */
#ifndef COSMOSPHERE_HPP
#define COSMOSPHERE_HPP

#include "cosmosphere.h"

/* The density kernel of the Newton-Raphson loop (reference include/density.hpp:20-34):
 * FLAT_PARTICLE evaluates only the particles still iterating; WARP_PER_LEAF re-evaluates every
 * particle of a leaf that still holds one. */
enum class KernelMode
{
    FLAT_PARTICLE,
    WARP_PER_LEAF
};

/* Per-call breakdown of solveDensH, in seconds (reference include/density.hpp:37-53).
 * upload / download: host-timed slot transfers since the last stats line (they sit outside
 * the solve; the figures cover the whole step). */
struct DensTimings
{
    double     upload;
    double     bboxAndSetup;
    double     keysAndSort;
    double     treeBuild;
    double     nodeCenters;
    double     jleafBuild;
    double     densKernel;
    double     gradJleafBuild;
    double     gradKernel;
    double     download;
    int        itersRun;
    int        nParticles;
    int        nLeavesOut;
    KernelMode kernelMode;
};

/* Solve h and rho for the n particles uploaded with cosmo_upload (POS, HSML and, with
 * withGradients, VEL and ACCEL); leaves h, DENS_OUT and GRAD_OUT on the device.  periodicBox:
 * nullptr (open) or {xmin, xmax, ymin, ymax, zmin, zmax}.  hfact < 0: the kernel's default.
 * Aborts on misuse or an unconverged solve, as densityiterate_gpu_c.
 * Replaces reference include/density.hpp:85-91 (src/density_base.cu:720-1015). */
DensTimings solveDensH(int n, double pmass, KernelMode mode = KernelMode::FLAT_PARTICLE, bool withGradients = false,
                       const double* periodicBox = nullptr, double tolh = 1.0e-4, double hfact = -1.0);

/* Per-phase cost of one force pass, in seconds (reference include/force.hpp:18-27).
 * upload / download: host-timed slot transfers since the last stats line. */
struct ForceTimings
{
    double hmaxUpsweep = 0.0;
    double jleafBuild  = 0.0;
    double upload      = 0.0;
    double prep        = 0.0;
    double kernel      = 0.0;
    double download    = 0.0;
};

/* The library's device state (opaque here; one per process, never destroyed).
 * Replaces reference include/gpu_state.hpp:126. */
struct GpuState;
GpuState& gpuState();

/* Force pass on the tree and h of the last solveDensH / densityiterate_gpu_c of this state;
 * THERMO (and VEL) uploaded with cosmo_upload_sorted; leaves FORCE_OUT on the device.  Aborts
 * when no density solve precedes it.  Replaces reference include/force.hpp:67-70
 * (src/force.cu:385-455). */
void computeForces(GpuState& s, double pmass, double beta, double alphau, bool discViscosity, bool pdvHeating,
                   bool shockHeating, ForceTimings& ft);

#endif /* COSMOSPHERE_HPP */
