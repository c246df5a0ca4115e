/*
 * tile.hpp -- staging of neighbour-leaf particles into LDS tiles, shared by the density,
 * gradient and force kernels, and the layout of the device work counters.
 */
#pragma once

#include "common.hpp"

namespace csph
{

constexpr int kCounterSlots  = 256;
constexpr int kCounterStride = 16;   // u64 per slot: one 128-byte line each

// Stage the j-particles of list entries [lo, hi) as 64-slot tiles, skipping the entries
// whose leaf keep(jl) rejects.  For every tile, each lane with a slot calls load(slot, j) and
// each lane without one pad(slot); then consume(nt) runs with the tile complete.
template<class Keep, class Load, class Pad, class Consume>
__device__ inline void forEachTileIf(int* __restrict__ segStart, int* __restrict__ segJ,
                                     const uint32_t* __restrict__ layout, const int* __restrict__ jList, uint32_t lo,
                                     uint32_t hi, Keep keep, Load load, Pad pad, Consume consume)
{
    const int lane = laneId();
    for (uint32_t e0 = lo; e0 < hi; e0 += kWave)
    {
        const uint32_t e = e0 + lane;
        int js = 0, jc = 0;
        if (e < hi)
        {
            const int jl = jList[e];
            if (keep(jl))
            {
                js = (int)layout[jl];
                jc = (int)layout[jl + 1] - js;
            }
        }
        const int incl  = waveInclusiveScan(jc);
        const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, kWave - 1, kWave));   // wave-uniform
        segStart[lane] = incl - jc;
        segJ[lane]     = js;
        const int nseg = (int)min((uint32_t)kWave, hi - e0);
        waveLdsSync();
        for (int s0 = 0; s0 < total; s0 += kWave)
        {
            const int slot = s0 + lane;
            if (slot < total)
            {
                int a = 0, b = nseg;   // last t with segStart[t] <= slot
                while (b - a > 1)
                {
                    const int m = (a + b) >> 1;
                    if (segStart[m] <= slot) a = m;
                    else                     b = m;
                }
                load(lane, segJ[a] + slot - segStart[a]);
            }
            else
            {
                pad(lane);
            }
            waveLdsSync();
            consume(__builtin_amdgcn_readfirstlane(min(kWave, total - s0)));
            waveLdsSync();
        }
    }
}

constexpr double kCullSlack = 1.0 + 1e-5;

// Squared gap between the point p and the tight box of leaf jl's particles (lb: centres x, y, z
// then half-sizes x, y, z, n doubles each; nearest periodic image of the centre per axis).
// Every particle of the leaf lies in the box (the half-sizes are rounded up), so none is
// nearer to p: a leaf whose gap exceeds R h (with kCullSlack) holds no pair inside that support.
// The octree cell of a leaf is up to twice its particles' extent per axis; the tight box takes
// a tenth of the candidates off the 8M-particle Plummer mixture's density passes.
template<bool PER>
__device__ inline double pointLeafGap2(double px, double py, double pz, const double* __restrict__ lb, int n, int jl,
                                       const DBox& box)
{
    const double p[3] = {px, py, pz};
    double d2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        double d = lb[(size_t)a * n + jl] - p[a];
        if constexpr (PER) d -= box.l[a] * rint(d * box.il[a]);
        const double v = fmax(fabs(d) - lb[(size_t)(3 + a) * n + jl], 0.0);
        d2 += v * v;
    }
    return d2;
}

template<class Load, class Consume>
__device__ inline void forEachTile(int* __restrict__ segStart, int* __restrict__ segJ,
                                   const uint32_t* __restrict__ layout, const int* __restrict__ jList, uint32_t lo,
                                   uint32_t hi, Load load, Consume consume)
{
    forEachTileIf(segStart, segJ, layout, jList, lo, hi, [](int) { return true; }, load, [](int) {}, consume);
}

} // namespace csph
