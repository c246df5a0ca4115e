/*
 * loop.hpp -- the device-driven Newton-Raphson loop (loop.hip).
 */
#pragma once

#include "density.hpp"

namespace csph
{

// The loop's starting point, after the host-driven first compaction: the active leaves and
// their item counts, the buffers of the compaction, and the pass's density arguments.
struct DeviceLoopArgs
{
    int*      leaves;
    int*      leavesAlt;
    uint32_t* cnt;
    uint32_t* cntAlt;
    uint32_t* flag;
    uint32_t* pos;
    uint32_t* itemOff;
    int*      itemPos;
    int*      itemPart;
    int       nAL;
    unsigned long long nIterating;
    int       sparse;
    bool      periodic;
    unsigned long long* listCand;   // list candidates of the passes (nullptr: not counted)
    DensArgs  da;
};

// Runs the NR passes (FLAT_PARTICLE) until no leaf is active or s.maxIter NR passes have run:
// s.graphPasses passes per HIP graph replay (s.useGraph) or per host round trip.  Adds the
// pass records to t and s.passLog; returns the particles still iterating (0 on convergence).
int deviceLoop(State& s, const DeviceLoopArgs& in, Timings& t);

} // namespace csph
