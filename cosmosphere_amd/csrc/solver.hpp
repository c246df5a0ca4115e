/*
 * solver.hpp -- host drivers of the density solve and the force pass (solver.hip).
 */
#pragma once

#include "state.hpp"

namespace csph
{

// Size every particle-length device array for n particles (idempotent per n):
// src/gpu_state.cu:21-46.
void sizeParticleArrays(State& s, int n);

// Creates the library's stream on first device use.
void ensureStream(State& s);

// The density / smoothing-length solve over the sets described in s (s.n, s.nSets,
// s.hSetFirst, s.hSetN), with inputs already on the device in the caller's order.
// mode 0 = FLAT_PARTICLE, 1 = WARP_PER_LEAF (include/density.hpp:20-34).
Timings solveDensH(State& s, double pmass, int mode, bool withGradients, const double* periodicBox, double tolh,
                   double hfact);

// Force pass on the tree and h of the last solve: src/force.cu:385-455.
ForceTimings computeForces(State& s, double pmass, double beta, double alphau, bool disc, bool pdv, bool shock);

} // namespace csph
