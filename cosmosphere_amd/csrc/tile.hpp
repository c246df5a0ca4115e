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

// The list entry owning slot s0 + lane of a tile.  Entry t (lane t) owns slots [start, start +
// count) of the batch (start: exclusive prefix sum of the counts).  Each entry overlapping the
// tile writes its index at its first tile position in `head` (distinct positions: two entries
// with slots cannot start at the same one), and an inclusive prefix maximum over the lanes
// carries it to the entry's other slots.  Lanes past the batch's total get -1 or a stale entry;
// the callers use the owner of slots below the total only.
__device__ inline int tileOwner(int* __restrict__ head, int s0, int start, int count)
{
    const int lane = laneId();
    head[lane] = -1;
    if (count > 0 && start < s0 + kWave && start + count > s0) head[max(start, s0) - s0] = lane;
    waveLdsSync();
    const int own = waveInclusiveMax(head[lane], -1);
    waveLdsSync();   // head is rewritten for the next tile
    return own;
}

// Stage the j-particles of list entries [lo, hi) as tiles of TD slots (64 or 128: the tile
// depth), skipping the entries whose leaf keep(jl) rejects.  For every tile, each lane stages
// TD / 64 slots: load(slot, j) for a slot with a particle, pad(slot) for one past the end;
// then consume(nt) runs with the tile complete.  A deeper tile halves the consume calls (their
// per-tile setup) against twice the LDS.
template<int TD = kWave, class Keep, class Load, class Pad, class Consume>
__device__ inline void forEachTileIf(int* __restrict__ segStart, int* __restrict__ segJ, int* __restrict__ head,
                                     const uint32_t* __restrict__ layout, const int* __restrict__ jList, uint32_t lo,
                                     uint32_t hi, Keep keep, Load load, Pad pad, Consume consume)
{
    static_assert(TD == kWave || TD == 2 * kWave, "tile depth 64 or 128");
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
        const int total = lastLane(incl);   // wave-uniform
        segStart[lane] = incl - jc;
        segJ[lane]     = js;
        waveLdsSync();
        for (int s0 = 0; s0 < total; s0 += TD)
        {
#pragma unroll
            for (int part = 0; part < TD / kWave; ++part)
            {
                const int base = s0 + part * kWave;   // wave-uniform
                if (base < total)
                {
                    const int slot = base + lane;
                    const int a = tileOwner(head, base, incl - jc, jc);
                    if (slot < total) load(part * kWave + lane, segJ[a] + slot - segStart[a]);
                    else              pad(part * kWave + lane);
                }
                else
                {
                    pad(part * kWave + lane);
                }
            }
            waveLdsSync();
            consume(__builtin_amdgcn_readfirstlane(min(TD, total - s0)));
            waveLdsSync();
        }
    }
}

// Stage, in list order, the j-particles of list entries [lo, hi) whose leaf keep(jl) accepts
// and which keepJ(j) accepts, as full 64-slot tiles: each round every lane tests one candidate,
// the accepted ones join a queue (LDS, 2 kWave entries), and each time 64 are queued (and once
// more for the rest at the end) load(slot, j) runs for every slot, then consume(nt).  A
// particle-level test that rejects most of a leaf's particles then costs no pair sweeps.
template<class Keep, class KeepJ, class Load, class Consume>
__device__ inline void forEachTileCompact(int* __restrict__ segStart, int* __restrict__ segJ, int* __restrict__ head,
                                          int* __restrict__ queue,
                                          const uint32_t* __restrict__ layout, const int* __restrict__ jList,
                                          uint32_t lo, uint32_t hi, Keep keep, KeepJ keepJ, Load load,
                                          Consume consume)
{
    const int lane = laneId();
    int fill = 0;   // queued entries, wave-uniform, < kWave between rounds
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
        const int total = lastLane(incl);   // wave-uniform
        segStart[lane] = incl - jc;
        segJ[lane]     = js;
        waveLdsSync();
        for (int s0 = 0; s0 < total; s0 += kWave)
        {
            const int slot = s0 + lane;
            const int a = tileOwner(head, s0, incl - jc, jc);
            int j = -1;
            if (slot < total) j = segJ[a] + slot - segStart[a];
            const bool ok = j >= 0 && keepJ(j);
            const uint64_t m = ballot64(ok);
            if (ok) queue[fill + __popcll(m & lanemaskLt())] = j;   // < 2 kWave
            fill += __popcll(m);
            if (fill >= kWave)
            {
                waveLdsSync();
                load(lane, queue[lane]);
                const int rest = queue[kWave + lane];   // entries kWave .. fill - 1 move to the front
                waveLdsSync();
                consume(kWave);
                waveLdsSync();
                fill -= kWave;
                if (lane < fill) queue[lane] = rest;
            }
        }
        waveLdsSync();   // segStart, segJ and the queue are rewritten next
    }
    if (fill > 0)
    {
        if (lane < fill) load(lane, queue[lane]);
        waveLdsSync();
        consume(fill);
        waveLdsSync();
    }
}

// Tiles of at most TD slots (64 or 128) filled in place: the j-particles of the list entries
// keep(jl) accepts, those probe(j, ok) accepts, in list order.  probe returns the slot's value
// (computed before the test, which it may use) and put(position, value) writes it at the next
// free position of the tile.  A tile is consumed once a round of 64 might not fit (more than
// TD - 64 slots: tiles of TD - 63 .. TD slots, and the last one), its positions past the end
// padded by pad(position).  No queue beyond the tile itself.
template<int TD, class Keep, class Probe, class Put, class Pad, class Consume>
__device__ inline void forEachTileInPlace(int* __restrict__ segStart, int* __restrict__ segJ, int* __restrict__ head,
                                          const uint32_t* __restrict__ layout, const int* __restrict__ jList,
                                          uint32_t lo, uint32_t hi, Keep keep, Probe probe, Put put, Pad pad,
                                          Consume consume)
{
    static_assert(TD == kWave || TD == 2 * kWave, "tile depth 64 or 128");
    const int lane = laneId();
    int fill = 0;              // slots in the tile (wave-uniform)
    int total = 0, s0 = 0;     // the staged chunk's slots, the next slot batch (wave-uniform)
    int start = 0, cnt = 0;    // this lane's entry of the chunk: first slot, slots
    uint32_t e0 = lo;
    bool done = false;
    while (!done)
    {
        while (fill <= TD - kWave)
        {
            if (s0 >= total)
            {
                if (e0 >= hi)
                {
                    done = true;
                    break;
                }
                // the next 64 list entries: their slot ranges
                waveLdsSync();   // segStart and segJ are rewritten
                const uint32_t e = e0 + lane;
                int js = 0;
                cnt = 0;
                if (e < hi)
                {
                    const int jl = jList[e];
                    if (keep(jl))
                    {
                        js = (int)layout[jl];
                        cnt = (int)layout[jl + 1] - js;
                    }
                }
                const int incl = waveInclusiveScan(cnt);
                total = lastLane(incl);   // wave-uniform
                start = incl - cnt;
                segStart[lane] = start;
                segJ[lane] = js;
                waveLdsSync();
                e0 += kWave;
                s0 = 0;
                continue;
            }
            const int slot = s0 + lane;
            const int a = tileOwner(head, s0, start, cnt);
            const bool in = slot < total;
            const int j = in ? segJ[a] + slot - segStart[a] : segJ[0];
            bool ok = false;
            const auto v = probe(j, ok);
            ok = ok && in;
            const uint64_t m = ballot64(ok);
            if (ok) put(fill + (int)__popcll(m & lanemaskLt()), v);   // < TD
            fill += __builtin_amdgcn_readfirstlane(__popcll(m));
            s0 += kWave;
        }
        if (fill == 0) continue;   // (done: the loop ends)
#pragma unroll
        for (int part = 0; part < TD / kWave; ++part)
        {
            const int k = part * kWave + lane;
            if (k >= fill) pad(k);
        }
        waveLdsSync();
        consume(fill);
        waveLdsSync();
        fill = 0;
    }
}

// forEachTileInPlace with the probe in two parts: fetch(j) loads a slot's raw data and
// finish(raw, ok) turns it into the slot's value and test.  Before a tile is consumed, the next
// 64 slots of the staged chunk are fetched, so their loads are in flight during the sweep.  The
// tiles, their slots and their order are forEachTileInPlace's.
template<int TD, class Keep, class Fetch, class Finish, class Put, class Pad, class Consume>
__device__ inline void forEachTileInPlaceAhead(int* __restrict__ segStart, int* __restrict__ segJ,
                                               int* __restrict__ head, const uint32_t* __restrict__ layout,
                                               const int* __restrict__ jList, uint32_t lo, uint32_t hi, Keep keep,
                                               Fetch fetch, Finish finish, Put put, Pad pad, Consume consume)
{
    static_assert(TD == kWave || TD == 2 * kWave, "tile depth 64 or 128");
    const int lane = laneId();
    int fill = 0;              // slots in the tile (wave-uniform)
    int total = 0, s0 = 0;     // the staged chunk's slots, the next slot batch (wave-uniform)
    int start = 0, cnt = 0;    // this lane's entry of the chunk: first slot, slots
    uint32_t e0 = lo;
    bool done = false;
    bool ahead = false;        // slots s0 .. s0 + 63 already fetched (wave-uniform)
    bool inA = false;
    decltype(fetch(0)) rawA{};
    // the slot's particle (a valid index for lanes past the chunk) and whether it is one
    auto slotJ = [&](bool& in) {
        const int slot = s0 + lane;
        const int a = tileOwner(head, s0, start, cnt);
        in = slot < total;
        return in ? segJ[a] + slot - segStart[a] : segJ[0];
    };
    while (!done)
    {
        while (fill <= TD - kWave)
        {
            if (s0 >= total)
            {
                if (e0 >= hi)
                {
                    done = true;
                    break;
                }
                // the next 64 list entries: their slot ranges
                waveLdsSync();   // segStart and segJ are rewritten
                const uint32_t e = e0 + lane;
                int js = 0;
                cnt = 0;
                if (e < hi)
                {
                    const int jl = jList[e];
                    if (keep(jl))
                    {
                        js = (int)layout[jl];
                        cnt = (int)layout[jl + 1] - js;
                    }
                }
                const int incl = waveInclusiveScan(cnt);
                total = lastLane(incl);   // wave-uniform
                start = incl - cnt;
                segStart[lane] = start;
                segJ[lane] = js;
                waveLdsSync();
                e0 += kWave;
                s0 = 0;
                continue;
            }
            bool in = inA;
            auto raw = rawA;
            if (!ahead) raw = fetch(slotJ(in));
            ahead = false;
            bool ok = false;
            const auto v = finish(raw, ok);
            ok = ok && in;
            const uint64_t m = ballot64(ok);
            if (ok) put(fill + (int)__popcll(m & lanemaskLt()), v);   // < TD
            fill += __builtin_amdgcn_readfirstlane(__popcll(m));
            s0 += kWave;
        }
        if (fill == 0) continue;   // (done: the loop ends)
        if (s0 < total)
        {
            rawA = fetch(slotJ(inA));
            ahead = true;
        }
#pragma unroll
        for (int part = 0; part < TD / kWave; ++part)
        {
            const int k = part * kWave + lane;
            if (k >= fill) pad(k);
        }
        waveLdsSync();
        consume(fill);
        waveLdsSync();
        fill = 0;
    }
}

constexpr double kCullSlack = 1.0 + 1e-5;

// Bounding box of the points of the lanes with own set (centre c, half-sizes s rounded up so
// that every point lies inside) and the largest r over them, in every lane.
__device__ inline void waveBox(bool own, double x, double y, double z, double r, double c[3], double s[3],
                               double& rmax)
{
    const double p[3] = {x, y, z};
    double lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        lo[a] = own ? p[a] : __builtin_inf();
        hi[a] = own ? p[a] : -__builtin_inf();
    }
    double rm = own ? r : 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        lo[a] = waveReduceMin(lo[a]);
        hi[a] = waveReduceMax(hi[a]);
    }
    rm = waveReduceMax(rm);
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        c[a] = 0.5 * (lo[a] + hi[a]);
        s[a] = 0.5 * (hi[a] - lo[a]) + 0x1p-50 * fmax(fabs(lo[a]), fabs(hi[a]));
    }
    rmax = rm;
}

// Squared gap between the point p and the box (centre c, half-sizes s; nearest periodic image of
// the centre per axis).
template<bool PER>
__device__ inline double pointBoxGap2(double px, double py, double pz, const double c[3], const double s[3],
                                      const DBox& box)
{
    const double p[3] = {px, py, pz};
    double d2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        double d = c[a] - p[a];
        if constexpr (PER) d -= box.l[a] * rint(d * box.il[a]);
        const double v = fmax(fabs(d) - s[a], 0.0);
        d2 += v * v;
    }
    return d2;
}

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
__device__ inline void forEachTile(int* __restrict__ segStart, int* __restrict__ segJ, int* __restrict__ head,
                                   const uint32_t* __restrict__ layout, const int* __restrict__ jList, uint32_t lo,
                                   uint32_t hi, Load load, Consume consume)
{
    forEachTileIf(segStart, segJ, head, layout, jList, lo, hi, [](int) { return true; }, load, [](int) {}, consume);
}

} // namespace csph
