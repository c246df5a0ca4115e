/*
 * nr.hip -- one Newton-Raphson pass of the smoothing-length solve: the density kernel with its
 * fused NR update and closure, the work items, and the compaction of the active set.
 *
 * Restates src/density_base.cu:237-477 (density sums and NR update) and :856-886 (compaction)
 * for gfx950.
 *
 * Work items.  Each active leaf becomes either
 *   - one whole-leaf item: one wavefront; the j-particles of the leaf's list are staged 64 at a
 *     time in an LDS tile (coalesced loads, once per wavefront) and the c particles still to
 *     evaluate (up to 64 at a time) each take g = floor(64 / c) lanes, lane s of a group
 *     sweeping tile slots s, s+g, ... with broadcast reads; a group reduction closes each
 *     particle.  A leaf of more than 32 runs lane = particle (g = 1, list order); a leaf with
 *     20 particles left uses 60 lanes and sweeps each tile in 22 steps instead of 64; or
 *   - in the last passes (few active leaves, so the chip would sit mostly idle), one item per
 *     particle of a leaf with at most `sparse` left: one wavefront per particle, the 64 lanes
 *     splitting its j-candidates, then a wavefront reduction.  The unconverged particles of
 *     the last passes are isolated ones whose h reaches into dense regions, with lists of up
 *     to millions of candidates: spreading them over wavefronts shortens the pass.
 *
 * Pruning.  The lists of a FLAT pass are the reference's lists trimmed to the reach of the
 * particles still to evaluate (jlist.hip, tight lists).  Items further skip the list entries
 * whose leaf's particles (their tight box) lie beyond R h_i of every particle evaluated (a
 * 1e-5 relative margin on the squared reach).  Skipped leaves hold no pair with q < R, so the
 * accepted pairs are unchanged.
 *
 * Per-particle state (u8): 1 = still iterating, 2 = converged, closure pending (rho and
 * d(rho)/d(h) are evaluated once more, at the final h, in the next pass), 0 = done.  With the
 * closure off (the reference's density-only behaviour) a converged particle goes straight to 0
 * and keeps the rho of its last update, taken at the h before it (src/density_base.cu:299-333).
 *
 * Precision forms:
 *   plain  FP64 throughout, the reference's arithmetic; whole-leaf items sum pair terms in list
 *          order, per-particle items by a wavefront tree;
 *   mixed  pair separations are formed exactly and rounded once to FP32 -- 32-bit fixed-point
 *          offsets from the i-leaf centre in whole-leaf items, an FP64 difference in
 *          per-particle items -- then r^2, q, the square root and the kernel brackets in FP32;
 *          whole-leaf items sum the W brackets in FP32 partials of at most 16 pairs added into
 *          FP64 (per-particle items add each one into FP64), the d(rho)/d(h) brackets in an FP32
 *          partial flushed into FP64 once per tile; the NR update, the convergence test and the
 *          closure's rho(h) formula are FP64.
 */
#include "density.hpp"
#include "prims.hpp"
#include "tile.hpp"

#include <algorithm>
#include <type_traits>

namespace csph
{

constexpr int    kNrWaves    = 4;
constexpr int    kNrThreads  = kNrWaves * kWave;
constexpr double kPruneSlack = 1.0 + 1e-5;
constexpr int    kFarFixed   = 1 << 30;           // padding j of the fixed-point tiles

// The reference's NR update, in FP64: src/density_base.cu:294-333.
__device__ inline void nrUpdate(double wsum, double gsum, double pmass, double tolh, double hfact, double& h,
                                double& rhoOut, double& gradOut, bool& conv)
{
#pragma clang fp contract(off)
    const double hi = h, hiOld = h;
    const double hi1 = 1.0 / hi, hi31 = hi1 * hi1 * hi1, hi41 = hi31 * hi1;
    const double rho_i  = wsum * kCnormk * pmass * hi31;
    const double grad_i = gsum * kCnormk * pmass * hi41;
    if (!(rho_i > 0.0))
    {
        rhoOut = 0.0; gradOut = 0.0; conv = false;
        return;
    }
    const double rhoh_i  = pmass * (hfact * hi1) * (hfact * hi1) * (hfact * hi1);
    const double funci   = rhoh_i - rho_i;
    const double dhdrhoi = -hi / (3.0 * rhoh_i);
    const double omegai  = 1.0 - dhdrhoi * grad_i;
    const double safe    = (omegai > 0.0) ? omegai : fabs(omegai + 1e-300);
    const double hraw    = hi - funci * dhdrhoi / safe;
    double hnew;
    if      (hraw > 1.2 * hi) hnew = 1.2 * hi;
    else if (hraw < 0.8 * hi) hnew = 0.8 * hi;
    else                      hnew = hraw;
    rhoOut  = rho_i;
    gradOut = grad_i;
    h       = hnew;
    conv    = fabs((hnew - hi) / hiOld) < tolh;
}

template<int MODE>
__device__ inline bool isClosureEval(uint8_t st) { return MODE == 2 || (MODE == 0 && st == 2); }

// The end of one particle's evaluation: NR update or closure, and its next state.
template<int MODE>
__device__ inline void finishParticle(const DensArgs& A, uint32_t i, int iLeaf, uint8_t st, double wsum, double gsum,
                                      double hcur)
{
    const double pmass = A.pmassSet ? A.pmassSet[A.leafSet[iLeaf]] : A.pmass;
    if (isClosureEval<MODE>(st))
    {
#pragma clang fp contract(off)
        const double hi1 = 1.0 / hcur, hi31 = hi1 * hi1 * hi1, hi41 = hi31 * hi1;
        A.rho[i]   = wsum * kCnormk * pmass * hi31;
        A.gradh[i] = gsum * kCnormk * pmass * hi41;
        if (MODE == 0) A.state[i] = 0;
        return;
    }
    double r, g;
    bool   conv;
    nrUpdate(wsum, gsum, pmass, A.tolh, A.hfact, hcur, r, g, conv);
    A.rho[i]   = r;
    A.gradh[i] = g;
    A.h[i]     = hcur;
    if (MODE == 0)            A.state[i] = conv ? (A.closure ? 2 : 0) : 1;
    else if (st != 0 && conv) A.state[i] = 0;   // WARP_PER_LEAF: converged particles are recomputed
}

// Block sum of four counters, one atomic each per block.
__device__ inline void blockCounters(unsigned long long* __restrict__ counters, const unsigned long long v[4])
{
    __shared__ unsigned long long red[4][kNrWaves];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const unsigned long long s = waveReduceSum(v[k]);
        if (laneId() == 0) red[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4)
    {
        const int idx[4] = {kCtrPairs, kCtrCand, kCtrClosPairs, kCtrClosCand};
        unsigned long long s = 0;
        for (int q = 0; q < kNrWaves; ++q) s += red[threadIdx.x][q];
        if (s) atomicAdd(counters + (blockIdx.x % kCounterSlots) * kCounterStride + idx[threadIdx.x], s);
    }
}

template<bool PER>
__device__ inline void imageFP64(double& dx, double& dy, double& dz, const DBox& box)
{
    if constexpr (PER)
    {
        if (fabs(dx) > 0.5 * box.l[0]) dx = dx - box.l[0] * copysign(1.0, dx);
        if (fabs(dy) > 0.5 * box.l[1]) dy = dy - box.l[1] * copysign(1.0, dy);
        if (fabs(dz) > 0.5 * box.l[2]) dz = dz - box.l[2] * copysign(1.0, dz);
    }
}

// Plain pair term (the reference's arithmetic): sums W and -q dW/dq - 3W when q < R.
template<bool PER>
__device__ inline void pairPlain(double dx, double dy, double dz, const DBox& box, double hs, double& sw, double& sg,
                                 int& acc)
{
    imageFP64<PER>(dx, dy, dz, box);
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double q2 = r2 * hs;
    if (q2 < kRad2)
    {
        const double q = sqrt(q2);
        double wk, dwk;
        kernWdW<double>(q2, q, wk, dwk);
        sw += wk;
        sg += -q * dwk - 3.0 * wk;
        ++acc;
    }
}

// Mixed pair term: FP32 separation (exact to one rounding), r^2, support test, q and kernel
// brackets V, q D (kernVD), when q < R, added to the caller's sums sv (V) and sp (q D).
// q^2 = r^2 hsr, hsr = 1/h^2 (in units of the separation's) rounded to FP32: its relative
// rounding, the same for all of a particle's pairs, is taken out of the W sum once per
// particle (particleSums).
template<class VSum>
__device__ inline void pairMixed(float dx, float dy, float dz, float hsr, VSum& sv, float& sp, int& acc)
{
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float q2 = r2 * hsr;
    if (q2 < float(kRad2))
    {
        const float q = __builtin_amdgcn_sqrtf(q2);
        float V, qD;
        kernVD<float>(q, q2, V, qD);
        sv += (VSum)V;
        sp += qD;
        ++acc;
    }
}

// V brackets of at most this many pairs are summed in FP32 before they enter the FP64 sum.
// The brackets are positive, so a partial's relative rounding error is below 15 u = 8.9e-7
// (u = 2^-24), and so is the FP64 total's; the error of h = hfact (m / rho)^(1/3) is a third of
// rho's.  Against a conversion and an FP64 add per accepted pair this takes 4% off the density
// kernel of the 8M-particle Plummer mixture, where mixed-against-FP64 errors stay below
// 1.7e-7 (h) and 3.8e-7 (rho) over all particles.
constexpr int kVChunk = 16;

// One particle's mixed pairs with the fixed-point tile slots ptr[0], ptr[stride], ... (steps
// of them), each slot read one step ahead: W brackets into sv (FP32 partials of <= kVChunk
// pairs), q D brackets into sp (the caller's FP32 partial).
__device__ inline void sweepFixed(const int4* ptr, int stride, int steps, int xq, int yq, int zq, float hsr,
                                  double& sv, float& sp, int& acc)
{
    int4 nx = *ptr;
    for (int t0 = 0; t0 < steps; t0 += kVChunk)
    {
        const int n = min(kVChunk, steps - t0);
        float v = 0.0f;
#pragma unroll 4
        for (int t = 0; t < n; ++t)
        {
            const int4 q = nx;
            ptr += stride;
            nx = *ptr;
            pairMixed((float)(xq - q.x), (float)(yq - q.y), (float)(zq - q.z), hsr, v, sp, acc);
        }
        sv += (double)v;
    }
}

// The mixed form's 1/h^2 as the FP32 hsr of the pair loop and the relative rounding eps of
// hsr (1 + eps) = hsd (|eps| <= 2^-24).
__device__ inline float mixedHsr(double hsd, double& eps)
{
    const float hsr = float(hsd);
    eps = (hsd - (double)hsr) / (double)hsr;
    return hsr;
}

// Work units of a whole-leaf batch of c particles.  With g = floor(64 / c) lanes per particle,
// 64 - g c lanes idle through every tile: a third of the wavefront for c = 22, almost half for
// c = 33.  Instead each particle's tile is cut into S units (slots s, s + S, ...: ceil(64 / S)
// steps each) and every lane takes U consecutive units of the c S; a tile then costs U unit
// sweeps plus a per-unit overhead instead of ceil(64 / g) steps.  S >= U, so a lane's units
// belong to at most two particles.  The shape is chosen per c by an estimated cost (41 per
// step, `unit` per unit, 8 per tile of the plain mapping), U in {2, 3, 4}, S <= 32 (so that a
// sweep reads no further than slot 2 S + 62 of the 128-slot tile); U = 1 keeps the plain
// mapping.  The per-unit cost that measured fastest on the 8M Plummer mixture is 15 for the
// mixed sweep and 5 for the FP64 one, whose steps cost more against the same overhead.
struct UnitShapes
{
    unsigned char U[kWave + 1];
    unsigned char S[kWave + 1];
};

constexpr UnitShapes makeUnitShapes(int unit)
{
    UnitShapes t{};
    constexpr int kStep = 41, kTile = 8;
    for (int c = 1; c <= kWave; ++c)
    {
        const int g = kWave / c;
        const int plain = ((kWave + g - 1) / g) * kStep + kTile;
        int bestU = 1, bestS = g, best = plain;
        for (int U = 2; U <= 4; ++U)
            for (int S = U; S <= 32; ++S)
            {
                if (c * S > kWave * U || c * S <= kWave * (U - 1)) continue;
                const int cost = U * (((kWave + S - 1) / S) * kStep + unit);
                if (cost < best && 100 * cost < 97 * plain)
                {
                    best = cost;
                    bestU = U;
                    bestS = S;
                }
            }
        t.U[c] = (unsigned char)bestU;
        t.S[c] = (unsigned char)bestS;
    }
    return t;
}

__constant__ constexpr UnitShapes kUnitShapesMixed = makeUnitShapes(15);
__constant__ constexpr UnitShapes kUnitShapesPlain = makeUnitShapes(5);

// The end of one sub-item of a split work item: particle r's sums (lanes with r >= 0) go to the
// item's slot; the wavefront whose sub-item completes the item (the count of finished sub-items
// reaches the item's count) adds every sub-item's partials in sub-item order and returns true
// with the totals in (s0, s1), after which it ends the particles.  The release fence before
// the count and the acquire fence after it order the partials across the chip's L2 caches.
__device__ inline bool splitCombine(const DensArgs& A, uint32_t subInfo, int r, double& s0, double& s1)
{
    const uint32_t sub = subInfo & 0xFFu, nSub = (subInfo >> 8) & 0xFFu, slot = subInfo >> 16;
    double* P = A.splitPart + (size_t)slot * kMaxSub * kWave * 2;
    if (r >= 0)
    {
        P[((size_t)sub * kWave + r) * 2] = s0;
        P[((size_t)sub * kWave + r) * 2 + 1] = s1;
    }
    __threadfence();
    uint32_t done = 0;
    if (laneId() == 0) done = atomicAdd(&A.splitDone[slot], 1u);
    done = __builtin_amdgcn_readfirstlane(done);
    if (done != nSub - 1u) return false;
    __threadfence();
    if (r >= 0)
    {
        s0 = 0.0;
        s1 = 0.0;
        for (uint32_t q = 0; q < nSub; ++q)
        {
            s0 += P[((size_t)q * kWave + r) * 2];
            s1 += P[((size_t)q * kWave + r) * 2 + 1];
        }
    }
    if (laneId() == 0) A.splitDone[slot] = 0u;   // the slot is free for the next pass
    return true;
}

// Particle sums from the accumulators: (sum W, sum -q dW/dq - 3W).  Mixed: the pair loop took
// q at 1/h^2 = hsr (1 + eps), so each q is (1 + eps/2) short; to first order the W brackets at
// the exact q sum to s0 + (eps/2) sum q V'(q) = s0 - (eps/2) (kGp / kWw) s1 (q V' = -(kGp / kWw)
// q D for both kernels); the remainder is O(eps^2) = 4e-15.
template<bool MIXED>
__device__ inline void particleSums(double s0, double s1, double eps, double& wsum, double& gsum)
{
#pragma clang fp contract(off)
    if constexpr (MIXED)
    {
        const double v = s0 - (0.5 * eps) * (kGp / kWw) * s1;
        wsum = kWw * v;
        gsum = kGp * s1 - kGw * v;
    }
    else
    {
        (void)eps;
        wsum = s0;
        gsum = s1;
    }
}

// PP: the launch holds per-particle items (a separate instance keeps the register budget of
// the whole-leaf path low when it does not).
// QUEUE: the device-driven pass's persistent grid (A.ctl), a separate instance so the
// register budget of the host-driven kernels is unchanged; held to 5 waves per SIMD (up to 96
// VGPRs: the instances need 87-91 and do not spill; at 6 waves, 80 VGPRs, they spilled 14-69
// registers and the 8M-particle Plummer mixture solved 0.5% (mixed) and 2% (FP64) slower).
// CULL: whole-leaf items stage only the list entries some particle's own support reaches (a
// separate instance: the culling table and test cost registers the plain instances keep).
#ifndef COSMO_QUEUE_WAVES
#define COSMO_QUEUE_WAVES 5
#endif
// TD: slots per LDS tile of the whole-leaf items (the tile depth, 64 or 128; a tuning option of
// the device-driven loop).
template<class Real, bool PER, int MODE, bool PP, bool QUEUE = false, bool CULL = false, int TD = kWave>
__global__ __launch_bounds__(kNrThreads) __attribute__((amdgpu_waves_per_eu(QUEUE ? COSMO_QUEUE_WAVES : 1)))
void densityKernel(DensArgs A)
{
    constexpr bool kMixed = sizeof(Real) == 4;
    __shared__ int segStart[kNrWaves][kWave];
    __shared__ int segJ[kNrWaves][kWave];
    __shared__ int    tileHead[kNrWaves][kWave];   // first tile position of each list entry (tileOwner)
    __shared__ int rankLane[kNrWaves][kWave];   // whole-leaf items: k-th particle to evaluate -> lane
    __shared__ double cpw[CULL ? kNrWaves : 1][4][CULL ? kWave : 1];   // ... its x, y, z, (R h)^2 slack
    // FP64 culling: the batch's box (centre, half-sizes) and largest reach (R hmax)^2 slack, and
    // the j-particles that pass the box test, queued (kCompactJ)
    constexpr bool kCompactJ = CULL && !kMixed;
    __shared__ double cbx[kCompactJ ? kNrWaves : 1][8];
    // whole-leaf tiles: mixed = fixed-point offsets (x, y, z, pad), one 16-byte LDS read per j;
    // plain = FP64 positions
    // (TD slots and the padding read beyond them: TD + 64)
    __shared__ int4   tq[kMixed ? kNrWaves : 1][kMixed ? TD + kWave : 1];
    // mixed whole-leaf items in work units: the batch's particles (fixed-point offsets, hsr bits)
    __shared__ int4   ptab[kMixed ? kNrWaves : 1][kMixed ? kWave : 1];
    // plain whole-leaf items in work units (culling instances): positions from cpw, 1/h^2 here
    __shared__ double ptabH[(!kMixed && CULL) ? kNrWaves : 1][(!kMixed && CULL) ? kWave : 1];
    __shared__ double tx[kMixed ? 1 : kNrWaves][kMixed ? 1 : TD + kWave];
    __shared__ double ty[kMixed ? 1 : kNrWaves][kMixed ? 1 : TD + kWave];
    __shared__ double tz[kMixed ? 1 : kNrWaves][kMixed ? 1 : TD + kWave];

    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = laneId();
    unsigned long long ctr[4] = {0, 0, 0, 0};   // NR pairs, NR candidates, closure pairs, closure candidates

    // subInfo: 0, or (sub-item, count, slot) of a split item: list range sub of count, partial sums
    auto run = [&](const int item, unsigned long long (&cc)[4], const uint32_t subInfo) {
        const int a = PP ? A.itemPos[item] : item;
        const int part = PP ? __builtin_amdgcn_readfirstlane(A.itemPart[item]) : -1;
        const int iLeaf = A.leaves ? A.leaves[a] : a;
        const uint32_t nSub = (subInfo >> 8) & 0xFFu;
        uint32_t lo = A.jOffset[a], hi = A.jOffset[a + 1];
        if (nSub > 1u)
        {
            const uint32_t len = hi - lo, chunk = (len + nSub - 1u) / nSub, sub = subInfo & 0xFFu;
            const uint32_t b = lo;
            lo = b + min(len, sub * chunk);
            hi = b + min(len, (sub + 1u) * chunk);
        }
        const DBox box = A.boxes[A.leafSet[iLeaf]];

        if (PP && part >= 0)
        {
            // ---- one particle, 64 lanes over its candidates ----
            const uint32_t i = (uint32_t)part;
            const uint8_t st = MODE == 2 ? 0 : A.state[i];
            const double xi = A.x[i], yi = A.y[i], zi = A.z[i], hcur = A.h[i];
            const double hsi = 1.0 / (hcur * hcur);
            double eps = 0.0;
            const float hsr = mixedHsr(hsi, eps);
            const double reach2 = (kRad * hcur) * (kRad * hcur) * kPruneSlack;
            double s0 = 0, s1 = 0;
            int acc = 0;
            uint32_t cand = 0;
            auto keep = [&](int jl) { return pointLeafGap2<PER>(xi, yi, zi, A.lbox, A.lboxN, jl, box) <= reach2; };
            auto pair = [&](int, int j) {
                double dx = xi - A.x[j], dy = yi - A.y[j], dz = zi - A.z[j];
                if constexpr (kMixed)
                {
                    imageFP64<PER>(dx, dy, dz, box);
                    float sp = 0.0f;   // one pair per lane per tile
                    pairMixed(float(dx), float(dy), float(dz), hsr, s0, sp, acc);
                    s1 += (double)sp;
                }
                else
                {
                    pairPlain<PER>(dx, dy, dz, box, hsi, s0, s1, acc);
                }
                ++cand;
            };
            forEachTileIf(segStart[w], segJ[w], tileHead[w], A.layout, A.jList, lo, hi, keep, pair, [](int) {}, [](int) {});
            s0 = waveReduceSum(s0);
            s1 = waveReduceSum(s1);
            const int slot = isClosureEval<MODE>(st) ? 2 : 0;
            cc[slot] += (unsigned long long)acc;
            cc[slot + 1] += cand;
            if (nSub > 1u && !splitCombine(A, subInfo, lane == 0 ? 0 : -1, s0, s1)) return;
            double wsum, gsum;
            particleSums<kMixed>(s0, s1, eps, wsum, gsum);
            if (lane == 0) finishParticle<MODE>(A, i, iLeaf, st, wsum, gsum, hcur);
        }
        else
        {
            // ---- whole leaf: its particles still to evaluate, up to 64 at a time, each with
            // g = floor(64 / count) lanes splitting every tile ----
            const uint32_t p0 = A.layout[iLeaf], p1 = A.layout[iLeaf + 1];
            const int node = A.leafNode[iLeaf];
            const double c0 = A.cx[node], c1 = A.cy[node], c2 = A.cz[node];
            const double s0n = A.sx[node], s1n = A.sy[node], s2n = A.sz[node];
            double hmList = A.hmaxLeaf[iLeaf];   // bounds the h of every particle evaluated here
            if (!isfinite(hmList) || hmList <= 0.0) hmList = 0.0;
            // mixed: offsets from the i-leaf centre in units of range / 2^29, range = leaf
            // half-size + R hmax (bounds every i offset); integer differences are exact, so a
            // pair separation is rounded once, to FP32, relative to itself.  j offsets are
            // clamped to +-2^30: a clamped j is beyond R h_i of every i, differences fit int32.
            double S = 0.0, invS2 = 0.0;
            if constexpr (kMixed)
            {
                const double range = fmax(fmax(s0n, s1n), s2n) + kRad * hmList;
                S = 536870912.0 / range;   // 2^29
                invS2 = 1.0 / (S * S);
            }
            auto toFixed = [&](double v) -> int {
                v = rint(v * S);
                v = fmin(fmax(v, -1073741824.0), 1073741824.0);   // +-2^30
                return (int)v;
            };
            for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
            {
                const uint32_t ic = q0 + lane;
                const uint8_t stc = (ic < p1 && MODE != 2) ? A.state[ic] : 0;
                const bool want = ic < p1 && (MODE != 0 || stc != 0);
                const uint64_t mask = ballot64(want);
                if (mask == 0) continue;
                const int c = __builtin_amdgcn_readfirstlane(__popcll(mask));
                const int g = __builtin_amdgcn_readfirstlane(kWave / c);   // lanes per particle
                if (want)
                {
                    const int r = __popcll(mask & lanemaskLt());
                    rankLane[w][r] = lane;
                    if constexpr (CULL)
                    {
                        const double hc = A.h[ic];
                        cpw[w][0][r] = A.x[ic]; cpw[w][1][r] = A.y[ic]; cpw[w][2][r] = A.z[ic];
                        cpw[w][3][r] = (kRad * hc) * (kRad * hc) * kCullSlack;
                    }
                }
                // mixed culling: the batch's box in fixed-point units and its largest reach (R hmax S
                // + 2 units: each fixed-point position is within half a unit of the particle's) squared,
                // with slack for the FP32 gap; a j-particle beyond it has no pair with q < R in the batch
                int bq0 = 0, bq1 = 0, bq2 = 0, bq3 = 0, bq4 = 0, bq5 = 0;
                float rq2 = 0.0f;
                if constexpr (kMixed && CULL)
                {
                    const uint32_t icc = want ? ic : p0;
                    const int xq = toFixed(A.x[icc] - c0), yq = toFixed(A.y[icc] - c1), zq = toFixed(A.z[icc] - c2);
                    bq0 = waveReduceMin(want ? xq : 0x7fffffff);
                    bq1 = waveReduceMin(want ? yq : 0x7fffffff);
                    bq2 = waveReduceMin(want ? zq : 0x7fffffff);
                    bq3 = waveReduceMax(want ? xq : -0x7fffffff);
                    bq4 = waveReduceMax(want ? yq : -0x7fffffff);
                    bq5 = waveReduceMax(want ? zq : -0x7fffffff);
                    const double reach = waveReduceMax(want ? kRad * A.h[icc] : 0.0) * S + 2.0;
                    rq2 = (float)(reach * reach * kCullSlack);
                }
                if constexpr (kCompactJ)
                {
                    // the batch's bounding box and largest reach: a j-particle farther from the box
                    // than that has no pair with q < R in the batch
                    const uint32_t icc = want ? ic : p0;
                    double c3[3], s3[3], rm;
                    waveBox(want, A.x[icc], A.y[icc], A.z[icc], want ? kRad * A.h[icc] : 0.0, c3, s3, rm);
                    if (lane == 0)
                    {
                        cbx[w][0] = c3[0]; cbx[w][1] = c3[1]; cbx[w][2] = c3[2];
                        cbx[w][3] = s3[0]; cbx[w][4] = s3[1]; cbx[w][5] = s3[2];
                        cbx[w][6] = rm * rm * kCullSlack;
                    }
                }
                waveLdsSync();
                // with culling, a list entry is staged only if its leaf's tight box lies within the support
                // of one of the c particles (a dropped leaf holds no pair with q < R for them)
                auto keep = [&](int jl) {
                    if constexpr (CULL)
                    {
                        for (int k = 0; k < c; ++k)
                            if (pointLeafGap2<PER>(cpw[w][0][k], cpw[w][1][k], cpw[w][2][k], A.lbox, A.lboxN, jl, box) <=
                                cpw[w][3][k])
                                return true;
                        return false;
                    }
                    else
                    {
                        (void)jl;
                        return true;
                    }
                };
                // FP64: of a kept entry, a j-particle only if it lies within the batch's largest reach
                // of its box.  It drops 15% of the 8M Plummer mixture's candidates and 7% of the
                // FP64 solve's time; the mixed form's pair terms cost too little for the test's
                // extra round trip per 64 slots to pay (its first passes ran 5% slower).
                auto keepJ = [&](const double3& pj) {
                    const double p3[3] = {pj.x, pj.y, pj.z};
                    double d2 = 0.0;
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                    {
                        double d = cbx[w][k] - p3[k];
                        if constexpr (PER) d -= box.l[k] * rint(d * box.il[k]);
                        const double v = fmax(fabs(d) - cbx[w][3 + k], 0.0);
                        d2 += v * v;
                    }
                    return d2 <= cbx[w][6];
                };
                const UnitShapes& shapes = kMixed ? kUnitShapesMixed : kUnitShapesPlain;
                const int U = ((kMixed || CULL) && A.units) ? (int)shapes.U[c] : 1;
                // FP64 tiles in sweep order: slot l at position (l mod Sl) Tl + floor(l / Sl), so
                // that the slots l0, l0 + Sl, l0 + 2 Sl, ... of one work unit (Sl = S) or of one
                // lane of a group (Sl = g) sit at consecutive positions (LDS reads with immediate
                // offsets: 1.6% off the FP64 density kernel of the 8M Plummer mixture).  Mixed
                // tiles keep slot order (position = slot), which measured 3% faster than sweep
                // order there.
                const int Sl = U > 1 ? (int)shapes.S[c] : g;
                const int Tl = (TD + Sl - 1) / Sl;
                // positions of the slots this lane stages: `lane` and, with 128-slot tiles, 64 + lane
                const int pos0 = kMixed ? lane : (lane % Sl) * Tl + lane / Sl;
                const int pos1 = kMixed ? kWave + lane : ((kWave + lane) % Sl) * Tl + (kWave + lane) / Sl;
                auto load = [&](int slot, int j) {
                    const int pos = slot < kWave ? pos0 : pos1;
                    double xj = A.x[j], yj = A.y[j], zj = A.z[j];
                    if constexpr (kMixed)
                    {
                        xj -= c0; yj -= c1; zj -= c2;
                        imageFP64<PER>(xj, yj, zj, box);
                        tq[w][pos] = make_int4(toFixed(xj), toFixed(yj), toFixed(zj), 0);
                    }
                    else
                    {
                        tx[w][pos] = xj; ty[w][pos] = yj; tz[w][pos] = zj;
                    }
                };
                // slots past the tile's end hold a j beyond every support: 2^30 fixed-point
                // units (>= 2 range) from the centre, or 1e200 in FP64 (r^2 overflows to inf)
                auto padAt = [&](int k) {
                    if constexpr (kMixed) tq[w][k] = make_int4(kFarFixed, kFarFixed, kFarFixed, 0);
                    else { tx[w][k] = 1e200; ty[w][k] = 1e200; tz[w][k] = 1e200; }
                };
                auto pad = [&](int slot) { padAt(slot < kWave ? pos0 : pos1); };
                // mixed culling: a slot's fixed-point offsets, and whether they lie within the
                // batch's reach of its box
                auto fetch = [&](int j) { return make_double3(A.x[j], A.y[j], A.z[j]); };
                auto finish = [&](const double3& r, bool& ok) {
                    double xj = r.x - c0, yj = r.y - c1, zj = r.z - c2;
                    imageFP64<PER>(xj, yj, zj, box);
                    const int xq = toFixed(xj), yq = toFixed(yj), zq = toFixed(zj);
                    const float dx = (float)max(max(bq0 - xq, xq - bq3), 0);
                    const float dy = (float)max(max(bq1 - yq, yq - bq4), 0);
                    const float dz = (float)max(max(bq2 - zq, zq - bq5), 0);
                    ok = dx * dx + dy * dy + dz * dz <= rq2;
                    return make_int4(xq, yq, zq, 0);
                };
                auto put = [&](int k, const int4& v) {
                    if constexpr (kMixed) tq[w][k] = v;
                };
                // FP64 culling: a slot's position, and whether it lies within the batch's reach of
                // its box; slot k goes to its sweep-order position
                const uint32_t mSl = (65536u + (uint32_t)Sl - 1u) / (uint32_t)Sl;   // k / Sl = (k mSl) >> 16, k < 1024
                auto posOf = [&](int k) {
                    const int q = (int)(((uint32_t)k * mSl) >> 16);
                    return (k - q * Sl) * Tl + q;
                };
                auto finish64 = [&](const double3& p, bool& ok) {
                    ok = keepJ(p);
                    return p;
                };
                auto put64 = [&](int k, const double3& p) {
                    if constexpr (!kMixed)
                    {
                        const int q = posOf(k);
                        tx[w][q] = p.x; ty[w][q] = p.y; tz[w][q] = p.z;
                    }
                };
                auto pad64 = [&](int k) { padAt(posOf(k)); };
                // the tiles stage slots 0 .. 63; the slots a sweep reads beyond them hold padding
                // through the batch: mixed, slots below 2 S + 63 <= 127 (the strided sweeps with
                // their read-ahead); FP64, slots below Sl Tl <= 95 (sweep order, the read-ahead
                // reaches position Sl Tl <= 127 at most)
                if constexpr (kMixed) padAt(TD + lane);
                else if (lane < Sl * Tl - TD) padAt(((TD + lane) % Sl) * Tl + (TD + lane) / Sl);
                if constexpr (kMixed || CULL)
                {
                    if (U > 1)
                    {
                        // ---- work units: each particle's tile in S units, U consecutive units
                        // per lane, a lane's units summed per particle (A: its first, B: the
                        // next), the particle's sums gathered from its lanes at the end ----
                        const int S = shapes.S[c];
                        const int nU = c * S;
                        const uint32_t mS = (65536u + (uint32_t)S - 1u) / (uint32_t)S;   // u / S = (u mS) >> 16, u < 1024
                        const bool own = lane < c;   // lane r holds particle r of the batch
                        const uint32_t i = own ? q0 + (uint32_t)rankLane[w][lane] : p0;
                        const uint8_t st = (own && MODE != 2) ? A.state[i] : 0;
                        const double hcur = own ? A.h[i] : 1.0;
                        double eps = 0.0;
                        if (own)
                        {
                            if constexpr (kMixed)
                            {
                                const float hsr = mixedHsr((1.0 / (hcur * hcur)) * invS2, eps);
                                ptab[w][lane] = make_int4(toFixed(A.x[i] - c0), toFixed(A.y[i] - c1),
                                                          toFixed(A.z[i] - c2), __float_as_int(hsr));
                            }
                            else
                            {
                                ptabH[w][lane] = 1.0 / (hcur * hcur);   // positions: cpw[w][0..2][lane]
                            }
                        }
                        waveLdsSync();
                        const int u0 = lane * U;
                        const int pA = (int)(((uint32_t)u0 * mS) >> 16);
                        double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
                        int accA = 0, accB = 0;
                        uint32_t candA = 0, candB = 0;
                        auto consumeUnits = [&](int nt) {
                            const int steps = __builtin_amdgcn_readfirstlane((nt + S - 1) / S);
                            for (int k = 0; k < U; ++k)
                            {
                                const int u = u0 + k;
                                if (u < nU)
                                {
                                    const int pu = (int)(((uint32_t)u * mS) >> 16);
                                    const int sg = u - pu * S;
                                    double sv = 0.0, sq = 0.0;
                                    int acc = 0;
                                    if constexpr (kMixed)
                                    {
                                        const int4 pp = ptab[w][pu];
                                        float sp = 0.0f;
                                        sweepFixed(&tq[w][sg], S, steps, pp.x, pp.y, pp.z, __int_as_float(pp.w), sv,
                                                   sp, acc);
                                        sq = (double)sp;
                                    }
                                    else
                                    {
                                        // slots sg + t S, t < steps <= Tl: positions sg Tl + t, below S Tl <= 95
                                        const double xu = cpw[w][0][pu], yu = cpw[w][1][pu], zu = cpw[w][2][pu];
                                        const double hsu = ptabH[w][pu];
#pragma unroll 4
                                        for (int t = 0; t < steps; ++t)
                                        {
                                            const int kk = sg * Tl + t;
                                            pairPlain<PER>(xu - tx[w][kk], yu - ty[w][kk], zu - tz[w][kk], box, hsu, sv, sq,
                                                           acc);
                                        }
                                    }
                                    const uint32_t cn = sg == 0 ? (uint32_t)nt : 0u;
                                    if (pu == pA) { a0 += sv; a1 += sq; accA += acc; candA += cn; }
                                    else          { b0 += sv; b1 += sq; accB += acc; candB += cn; }
                                }
                            }
                        };
                        if constexpr (kCompactJ)
                            forEachTileInPlaceAhead<TD>(segStart[w], segJ[w], tileHead[w], A.layout, A.jList, lo, hi,
                                                        keep, fetch, finish64, put64, pad64, consumeUnits);
                        else if constexpr (kMixed && CULL)
                            forEachTileInPlaceAhead<TD>(segStart[w], segJ[w], tileHead[w], A.layout, A.jList, lo, hi,
                                                        keep, fetch, finish, put, padAt, consumeUnits);
                        else
                            forEachTileIf<TD>(segStart[w], segJ[w], tileHead[w], A.layout, A.jList, lo, hi, keep, load,
                                              pad, consumeUnits);
                        // counters by the state of each lane's particles, read before any update
                        if (u0 < nU)
                        {
                            const int pB = (int)(((uint32_t)min(u0 + U - 1, nU - 1) * mS) >> 16);
                            const uint8_t sA = MODE != 2 ? A.state[q0 + (uint32_t)rankLane[w][pA]] : 0;
                            const uint8_t sB = MODE != 2 ? A.state[q0 + (uint32_t)rankLane[w][pB]] : 0;
                            const int slotA = isClosureEval<MODE>(sA) ? 2 : 0, slotB = isClosureEval<MODE>(sB) ? 2 : 0;
                            cc[slotA] += (unsigned long long)accA;
                            cc[slotA + 1] += candA;
                            cc[slotB] += (unsigned long long)accB;
                            cc[slotB + 1] += candB;
                        }
                        // each lane's partial sums through the tile area (free until the next
                        // batch); particle r's units lie on lanes floor(r S / U) .. floor((r S + S - 1) / U)
                        waveLdsSync();
                        double s0 = 0, s1 = 0;
                        if constexpr (kMixed)
                        {
                            // slots 2 l, 2 l + 1 of the 128: lane l's (a0, a1), (b0, b1) as bit patterns
                            auto pack = [](double x, double y) {
                                const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
                                return make_int4((int)bx, (int)(bx >> 32), (int)by, (int)(by >> 32));
                            };
                            auto lo64 = [](const int4& v) {
                                return __longlong_as_double(((long long)v.y << 32) | (long long)(uint32_t)v.x);
                            };
                            auto hi64 = [](const int4& v) {
                                return __longlong_as_double(((long long)v.w << 32) | (long long)(uint32_t)v.z);
                            };
                            tq[w][2 * lane] = pack(a0, a1);
                            tq[w][2 * lane + 1] = pack(b0, b1);
                            waveLdsSync();
                            if (own)
                            {
                                const int l0 = (lane * S) / U, l1 = (lane * S + S - 1) / U;
                                for (int l = l0; l <= l1; ++l)
                                {
                                    const int pl = (int)(((uint32_t)(l * U) * mS) >> 16);
                                    const int4 v = tq[w][2 * l + (pl == lane ? 0 : 1)];
                                    s0 += lo64(v);
                                    s1 += hi64(v);
                                }
                            }
                        }
                        else
                        {
                            // tx slots 2 l, 2 l + 1: lane l's (a0, a1); ty: its (b0, b1)
                            tx[w][2 * lane] = a0;
                            tx[w][2 * lane + 1] = a1;
                            ty[w][2 * lane] = b0;
                            ty[w][2 * lane + 1] = b1;
                            waveLdsSync();
                            if (own)
                            {
                                const int l0 = (lane * S) / U, l1 = (lane * S + S - 1) / U;
                                for (int l = l0; l <= l1; ++l)
                                {
                                    const int pl = (int)(((uint32_t)(l * U) * mS) >> 16);
                                    const double* src = pl == lane ? &tx[w][2 * l] : &ty[w][2 * l];
                                    s0 += src[0];
                                    s1 += src[1];
                                }
                            }
                        }
                        waveLdsSync();   // the tile area is the next batch's
                        if (nSub > 1u && !splitCombine(A, subInfo, own ? lane : -1, s0, s1)) continue;
                        double wsum, gsum;
                        particleSums<kMixed>(s0, s1, eps, wsum, gsum);
                        if (own) finishParticle<MODE>(A, i, iLeaf, st, wsum, gsum, hcur);
                        continue;
                    }
                }
                {
                    const int grp = lane / g, sub = lane - grp * g;
                    const bool has = grp < c;
                    const uint32_t i = has ? q0 + (uint32_t)rankLane[w][grp] : p0;
                    const uint8_t st = (has && MODE != 2) ? A.state[i] : 0;
                    const double xi = A.x[i], yi = A.y[i], zi = A.z[i];
                    const double hcur = has ? A.h[i] : 1.0;
                    const double hsi = 1.0 / (hcur * hcur);
                    int xq = 0, yq = 0, zq = 0;
                    if constexpr (kMixed) { xq = toFixed(xi - c0); yq = toFixed(yi - c1); zq = toFixed(zi - c2); }
                    double eps = 0.0;
                    const float hsr = mixedHsr(hsi * invS2, eps);
                    double s0 = 0, s1 = 0;
                    int acc = 0;
                    uint32_t cand = 0;
                    // lane sub of a group: pairs with slots sub, sub + g, ... (FP64: positions
                    // sub Tl, sub Tl + 1, ...), each read one step ahead.  Slots from nt on hold
                    // padding ([nt, 64) staged with the tile, the rest for the batch); the mixed
                    // read-ahead reaches slot g - 1 + ceil(nt / g) g < 128.
                    auto sweep = [&](int first, int steps) {
                        if constexpr (kMixed)
                        {
                            float sp = 0.0f;
                            sweepFixed(&tq[w][first], g, steps, xq, yq, zq, hsr, s0, sp, acc);
                            s1 += (double)sp;
                        }
                        else
                        {
    #pragma unroll 4
                            for (int t = 0; t < steps; ++t)
                            {
                                const int k = first * Tl + t;
                                pairPlain<PER>(xi - tx[w][k], yi - ty[w][k], zi - tz[w][k], box, hsi, s0, s1, acc);
                            }
                        }
                    };
                    auto consume = [&](int nt) {
                        if (g == 1)
                        {
                            sweep(0, nt);   // list order, as the reference
                        }
                        else
                        {
                            sweep(sub, __builtin_amdgcn_readfirstlane((nt + g - 1) / g));
                        }
                        cand += sub == 0 ? (uint32_t)nt : 0u;
                    };
                    if constexpr (kCompactJ)
                        forEachTileInPlaceAhead<TD>(segStart[w], segJ[w], tileHead[w], A.layout, A.jList, lo, hi,
                                                    keep, fetch, finish64, put64, pad64, consume);
                    else if constexpr (kMixed && CULL)
                        forEachTileInPlaceAhead<TD>(segStart[w], segJ[w], tileHead[w], A.layout, A.jList, lo, hi,
                                                    keep, fetch, finish, put, padAt, consume);
                    else
                        forEachTileIf<TD>(segStart[w], segJ[w], tileHead[w], A.layout, A.jList, lo, hi, keep, load, pad,
                                          consume);
                    if (has)
                    {
                        const int slot = isClosureEval<MODE>(st) ? 2 : 0;
                        cc[slot] += (unsigned long long)acc;
                        cc[slot + 1] += cand;
                    }
                    // group sums, a tree over lanes sub .. g-1 of each group (lane sub = 0 ends with it)
                    for (int off = 1; off < g; off <<= 1)
                    {
                        const double t0 = __shfl_down(s0, off, kWave), t1 = __shfl_down(s1, off, kWave);
                        if (sub + off < g) { s0 += t0; s1 += t1; }
                    }
                    if (nSub > 1u && !splitCombine(A, subInfo, (has && sub == 0) ? grp : -1, s0, s1)) continue;
                    double wsum, gsum;
                    particleSums<kMixed>(s0, s1, eps, wsum, gsum);
                    if (has && sub == 0) finishParticle<MODE>(A, i, iLeaf, st, wsum, gsum, hcur);
                }
            }
        }
    };

    if constexpr (QUEUE)
    {
        // device-driven pass: a persistent grid; each wavefront pulls the next item of the
        // pass's queue (cost-descending when qOrder is given) until the queue is empty
        PassCtl* const C = A.ctl;
        if (C->skip || C->perParticle != (PP ? 1u : 0u)) return;   // the whole block
        if (threadIdx.x == 0) atomicMin(&C->tStart, (unsigned long long)wall_clock64());
        const int nItems = (int)C->nItems;
        const int nHeavy = A.qOrder ? (int)C->qBucket[1] : 0;
        for (;;)
        {
            int k = 0;
            if (lane == 0) k = (int)atomicAdd(&C->queueHead, 1u);
            k = __builtin_amdgcn_readfirstlane(k);
            // one call site: an item's arithmetic is the same code wherever the queue puts it
            int item;
            uint32_t subInfo = 0u;
            if (k < nHeavy)
            {
                item = A.qOrder[k];
                if (A.qSub) subInfo = A.qSub[k];
            }
            else
            {
                k -= nHeavy;
                if (k >= nItems) break;
                if (nHeavy && A.qFlag[k]) continue;   // taken from the head already
                item = k;
            }
            run(__builtin_amdgcn_readfirstlane(item), ctr, __builtin_amdgcn_readfirstlane(subInfo));
        }
    }
    else
    {
        const int item = blockIdx.x * kNrWaves + w;
        if (item < A.nItems) run(item, ctr, 0u);
    }
    blockCounters(A.counters, ctr);
    if (QUEUE && threadIdx.x == 0) atomicMax(&A.ctl->tEnd, (unsigned long long)wall_clock64());
}

// ---------------------------------------------------------------------------------------
// Compaction and work items
// ---------------------------------------------------------------------------------------
__device__ inline uint32_t itemsOf(uint32_t remaining, int sparse)
{
    return remaining == 0 ? 0u : (remaining <= (uint32_t)sparse ? remaining : 1u);
}

// Wave per active leaf: flag[a] = leaf a has a particle in state != 0, count[a] = its work
// items in the next pass; counters: particles in state 1, in state != 0, items.
__global__ __launch_bounds__(256) void leafFlagKernel(const int* __restrict__ leaves, int nLeaves,
                                                      const uint32_t* __restrict__ layout,
                                                      const uint8_t* __restrict__ state, int sparse,
                                                      uint32_t* __restrict__ flag, uint32_t* __restrict__ count,
                                                      unsigned long long* __restrict__ counters,
                                                      const PassCtl* __restrict__ ctl)
{
    if (ctl && ctl->skip) return;
    const int n = ctl ? (int)ctl->nAL : nLeaves;
    unsigned long long* slot = counters + (blockIdx.x % kCounterSlots) * kCounterStride;
    for (int a = blockIdx.x * 4 + (threadIdx.x >> 6); a < n; a += gridDim.x * 4)
    {
        const int l = leaves[a];
        const uint32_t p0 = layout[l], p1 = layout[l + 1];
        uint32_t it = 0, rem = 0;
        for (uint32_t p = p0 + laneId(); p < p1; p += kWave)
        {
            const uint8_t s = state[p];
            it += (s == 1) ? 1u : 0u;
            rem += s ? 1u : 0u;
        }
        it = waveReduceSum(it);
        rem = waveReduceSum(rem);
        if (laneId() == 0)
        {
            const uint32_t items = itemsOf(rem, sparse);
            flag[a] = rem ? 1u : 0u;
            count[a] = items;
            if (it) atomicAdd(slot + kCtrIterating, (unsigned long long)it);
            if (rem) atomicAdd(slot + kCtrRemaining, (unsigned long long)rem);
            if (items) atomicAdd(slot + kCtrItems, (unsigned long long)items);
        }
    }
}

__global__ void leafScatterKernel(const int* __restrict__ leaves, const uint32_t* __restrict__ count, int nLeaves,
                                  const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                  int* __restrict__ out, uint32_t* __restrict__ outCount,
                                  const PassCtl* __restrict__ ctl)
{
    if (ctl && ctl->skip) return;
    const int n = ctl ? (int)ctl->nAL : nLeaves;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x)
    {
        if (flag[a])
        {
            out[pos[a]] = leaves[a];
            outCount[pos[a]] = count[a];
        }
    }
}

// Wave per active leaf: its items at off[a]: (a, -1) for a whole-leaf item, else (a, p) for
// every particle p in state != 0, in particle order.  The rule matches leafFlagKernel's count.
__global__ __launch_bounds__(256) void itemFillKernel(const int* __restrict__ leaves, int nLeaves,
                                                      const uint32_t* __restrict__ layout,
                                                      const uint8_t* __restrict__ state, int sparse,
                                                      const uint32_t* __restrict__ off, int* __restrict__ itemPos,
                                                      int* __restrict__ itemPart, const PassCtl* __restrict__ ctl)
{
    if (ctl && (ctl->skip || !ctl->perParticle)) return;
    const int n = ctl ? (int)ctl->nAL : nLeaves;
    for (int a = blockIdx.x * 4 + (threadIdx.x >> 6); a < n; a += gridDim.x * 4)
    {
        const int l = leaves[a];
        const uint32_t p0 = layout[l], p1 = layout[l + 1];
        uint32_t rem = 0;
        for (uint32_t p = p0 + laneId(); p < p1; p += kWave) rem += state[p] ? 1u : 0u;
        rem = waveReduceSum(rem);
        if (rem == 0) continue;
        uint32_t o = off[a];
        if (rem > (uint32_t)sparse)
        {
            if (laneId() == 0) { itemPos[o] = a; itemPart[o] = -1; }
            continue;
        }
        for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
        {
            const uint32_t p = q0 + laneId();
            const bool live = p < p1 && state[p] != 0;
            const uint64_t m = ballot64(live);
            if (live)
            {
                const uint32_t k = o + (uint32_t)__popcll(m & lanemaskLt());
                itemPos[k] = a;
                itemPart[k] = (int)p;
            }
            o += (uint32_t)__popcll(m);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Host wrappers
// ---------------------------------------------------------------------------------------
void launchDensity(const DensArgs& a, int precision, int mode, bool periodic, hipStream_t st)
{
    if (a.nItems <= 0) return;
    const int grid = iceil(a.nItems, kNrWaves);
    const bool pp = a.itemPart != nullptr;
#define CSPH_DENS(R, P, M, Q) densityKernel<R, P, M, Q><<<grid, kNrThreads, 0, st>>>(a)
#define CSPH_DENS_C(R, P, M, Q)                                                          \
    do {                                                                                 \
        if (a.cull) densityKernel<R, P, M, Q, false, true><<<grid, kNrThreads, 0, st>>>(a); \
        else        CSPH_DENS(R, P, M, Q);                                               \
    } while (0)
#define CSPH_DENS_PP(R, P, M)                     \
    do {                                          \
        if (pp) CSPH_DENS_C(R, P, M, true);       \
        else    CSPH_DENS_C(R, P, M, false);      \
    } while (0)
#define CSPH_DENS_MODE(R, P)                      \
    do {                                          \
        if (mode == 0)      CSPH_DENS_PP(R, P, 0); \
        else if (mode == 1) CSPH_DENS(R, P, 1, false); \
        else                CSPH_DENS(R, P, 2, false); \
    } while (0)
    if (precision == kMixed)
    {
        if (periodic) CSPH_DENS_MODE(float, true);
        else          CSPH_DENS_MODE(float, false);
    }
    else
    {
        if (periodic) CSPH_DENS_MODE(double, true);
        else          CSPH_DENS_MODE(double, false);
    }
#undef CSPH_DENS_MODE
#undef CSPH_DENS_PP
#undef CSPH_DENS_C
#undef CSPH_DENS
    LAUNCH_CHECK();
}

void launchDensityPersistent(const DensArgs& a, int precision, bool periodic, int blocks, hipStream_t st)
{
#define CSPH_DENS_T(R, P, C, T)                                                                   \
    do {                                                                                          \
        densityKernel<R, P, 0, false, true, C, T><<<blocks, kNrThreads, 0, st>>>(a);              \
        densityKernel<R, P, 0, true, true, C, T><<<blocks, kNrThreads, 0, st>>>(a);               \
    } while (0)
#define CSPH_DENS_P(R, P)                                                                         \
    do {                                                                                          \
        if (a.cull) { if (a.tileDepth == 2 * kWave) CSPH_DENS_T(R, P, true, 2 * kWave);           \
                      else CSPH_DENS_T(R, P, true, kWave); }                                      \
        else        { if (a.tileDepth == 2 * kWave) CSPH_DENS_T(R, P, false, 2 * kWave);          \
                      else CSPH_DENS_T(R, P, false, kWave); }                                     \
    } while (0)
    if (precision == kMixed)
    {
        if (periodic) CSPH_DENS_P(float, true);
        else          CSPH_DENS_P(float, false);
    }
    else
    {
        if (periodic) CSPH_DENS_P(double, true);
        else          CSPH_DENS_P(double, false);
    }
#undef CSPH_DENS_P
#undef CSPH_DENS_T
    LAUNCH_CHECK();
}

// Resident blocks of the whole-leaf instance over the chip (the persistent grid).
int densityPersistentBlocks(int precision, bool periodic)
{
    int dev = 0, cus = 0, per = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const void* f = precision == kMixed
                        ? (periodic ? (const void*)densityKernel<float, true, 0, false, true>
                                    : (const void*)densityKernel<float, false, 0, false, true>)
                        : (periodic ? (const void*)densityKernel<double, true, 0, false, true>
                                    : (const void*)densityKernel<double, false, 0, false, true>);
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, kNrThreads, 0));
    return std::max(cus, 1) * std::max(per, 1);
}

int counterSlots() { return kCounterSlots; }
int counterStride() { return kCounterStride; }

// With ctl the grids are capped (grid-stride loops over ctl->nAL), so a device-driven pass
// launches a fixed shape whatever the count.
static int strideGrid(int units, int perBlock, const PassCtl* ctl)
{
    const int g = iceil(units, perBlock);
    return ctl ? std::min(g, 4096) : g;
}

void launchLeafFlags(const int* leaves, int nLeaves, const uint32_t* layout, const uint8_t* state, int sparse,
                     uint32_t* flag, uint32_t* itemCount, unsigned long long* counters, hipStream_t st,
                     const PassCtl* ctl)
{
    if (nLeaves <= 0) return;
    leafFlagKernel<<<strideGrid(nLeaves, 4, ctl), 256, 0, st>>>(leaves, nLeaves, layout, state, sparse, flag,
                                                                itemCount, counters, ctl);
    LAUNCH_CHECK();
}

void launchLeafScatter(const int* leaves, const uint32_t* count, int nLeaves, const uint32_t* flag,
                       const uint32_t* pos, int* out, uint32_t* outCount, hipStream_t st, const PassCtl* ctl)
{
    if (nLeaves <= 0) return;
    leafScatterKernel<<<strideGrid(nLeaves, 256, ctl), 256, 0, st>>>(leaves, count, nLeaves, flag, pos, out,
                                                                     outCount, ctl);
    LAUNCH_CHECK();
}

void launchItemFill(const int* leaves, int nLeaves, const uint32_t* layout, const uint8_t* state, int sparse,
                    const uint32_t* off, int* itemPos, int* itemPart, hipStream_t st, const PassCtl* ctl)
{
    if (nLeaves <= 0) return;
    itemFillKernel<<<strideGrid(nLeaves, 4, ctl), 256, 0, st>>>(leaves, nLeaves, layout, state, sparse, off,
                                                                itemPos, itemPart, ctl);
    LAUNCH_CHECK();
}

} // namespace csph
