/*
 * resort.hip -- incremental re-sort of nearly sorted keys (resident Hilbert order across
 * steps; the reference names the idea and leaves it unbuilt, include/gpu_state.hpp:17-23).
 *
 * When the solve's inputs were uploaded in the previous solve's order (entry i is caller
 * particle order[i]), the keys of a drifted particle set are nearly sorted.  The result must
 * equal the full sort exactly: ascending key, ties broken by caller index (the stable sort of
 * src/tree.cu:88-91 over the caller's order).  So the sort key is the pair (key, caller
 * index), which is unique per particle.
 *
 *  1. A particle stays when its new key lies in the key range of the leaf it belonged to in
 *     the previous tree (and that leaf held at most 64 particles).  The previous leaves are
 *     disjoint key ranges in ascending order, and their particles are contiguous in the
 *     input, so the stayers, sorted within each previous leaf, form one sorted sequence A.
 *  2. One wavefront per previous leaf ranks its stayers by (key, caller index) against each
 *     lane's entry read as a scalar, and writes them to their place in A.
 *  3. The movers (new key outside their leaf's range, dead particles, leaves of more than 64
 *     particles) are compacted and radix-sorted: by caller index, then stably by key.
 *  4. A merge-path merge of A and the sorted movers writes the sorted keys, the input
 *     position of each sorted entry (the gather index of the permutation) and its caller
 *     index (the new `order`).
 *
 * Cost: three streaming passes over the keys plus a radix sort of the movers (a few percent
 * of the particles after a small drift), against the full sort's passes over all of them.
 */
#include "prims.hpp"
#include "tree.hpp"

#include <algorithm>

namespace csph
{

constexpr int kSegMax = 64;              // previous leaves up to one wavefront are sorted in place
constexpr int kMergeThreads = 256;
constexpr int kMergeItems = 8;
constexpr int kMergeTile = kMergeThreads * kMergeItems;

__device__ inline bool pairLess(uint64_t ka, uint32_t ca, uint64_t kb, uint32_t cb)
{
    return ka < kb || (ka == kb && ca < cb);
}

// 1 for a stayer, 0 for a mover.
__global__ __launch_bounds__(256) void stayFlagKernel(const uint64_t* __restrict__ keys, int n, int oldAlive,
                                                      const int* __restrict__ oldLeafOf,
                                                      const uint64_t* __restrict__ oldLeafStart,
                                                      const int* __restrict__ oldLeafLevel,
                                                      const uint32_t* __restrict__ oldLayout,
                                                      uint32_t* __restrict__ stay)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t st = 0;
    if (i < oldAlive)
    {
        const int l = oldLeafOf[i];
        const uint64_t k = keys[i];
        if (k != kDeadKey && oldLayout[l + 1] - oldLayout[l] <= (uint32_t)kSegMax)
        {
            const uint64_t lo = oldLeafStart[l];
            const uint64_t span = uint64_t(1) << (3 * (kMaxLevel - oldLeafLevel[l]));
            st = (k >= lo && k - lo < span) ? 1u : 0u;
        }
    }
    stay[i] = st;
}

// Movers, in input order, to B = [na, n) of the work arrays; their caller index is also the
// first radix key.
__global__ __launch_bounds__(256) void moverWriteKernel(const uint64_t* __restrict__ keys,
                                                        const int* __restrict__ order,
                                                        const uint32_t* __restrict__ stay,
                                                        const uint32_t* __restrict__ pos, int n, int na,
                                                        uint64_t* __restrict__ wKey, uint32_t* __restrict__ wCal,
                                                        uint32_t* __restrict__ wIdx, uint64_t* __restrict__ sKey,
                                                        uint32_t* __restrict__ sVal)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || stay[i]) return;
    const int j = i - (int)pos[i];
    wKey[na + j] = keys[i];
    wCal[na + j] = (uint32_t)order[i];
    wIdx[na + j] = (uint32_t)i;
    sKey[j] = (uint64_t)(uint32_t)order[i];
    sVal[j] = (uint32_t)j;
}

// Stayers of each previous leaf, ranked within the leaf, to A = [0, na).  One wavefront per leaf.
__global__ __launch_bounds__(256) void staySortKernel(const uint64_t* __restrict__ keys,
                                                      const int* __restrict__ order,
                                                      const uint32_t* __restrict__ stay,
                                                      const uint32_t* __restrict__ pos,
                                                      const uint32_t* __restrict__ oldLayout, int oldLeaves,
                                                      uint64_t* __restrict__ wKey, uint32_t* __restrict__ wCal,
                                                      uint32_t* __restrict__ wIdx)
{
    const int l = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
    if (l >= oldLeaves) return;
    const int lane = threadIdx.x % kWave;
    const int a = __builtin_amdgcn_readfirstlane((int)oldLayout[l]);
    const int c = __builtin_amdgcn_readfirstlane((int)oldLayout[l + 1] - a);
    if (c > kSegMax) return;   // flagged as movers
    const int i = a + lane;
    uint64_t k = kDeadKey;
    uint32_t cal = 0xFFFFFFFFu;
    int st = 0;
    if (lane < c)
    {
        st = (int)stay[i];
        k = keys[i];
        cal = (uint32_t)order[i];
    }
    int rank = 0;
    for (int t = 0; t < c; ++t)
    {
        // t is wave-uniform: lane t's entry as scalars
        const uint64_t kt = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(k >> 32), t) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, t);
        const uint32_t ct = (uint32_t)__builtin_amdgcn_readlane((int)cal, t);
        const int stt = __builtin_amdgcn_readlane(st, t);
        rank += (stt && pairLess(kt, ct, k, cal)) ? 1 : 0;
    }
    if (lane < c && st)
    {
        const int p = (int)pos[a] + rank;
        wKey[p] = k;
        wCal[p] = cal;
        wIdx[p] = (uint32_t)i;
    }
}

// After the radix sort by caller index: the key of each mover in that order, for the stable
// sort by key.
__global__ __launch_bounds__(256) void moverKeysKernel(const uint64_t* __restrict__ wKeyB,
                                                       const uint32_t* __restrict__ perm, int nb,
                                                       uint64_t* __restrict__ k2, uint32_t* __restrict__ v2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const uint32_t p = perm[j];
    k2[j] = wKeyB[p];
    v2[j] = p;
}

struct MergeIn
{
    const uint64_t* aKey;
    const uint32_t* aCal;
    const uint32_t* aIdx;
    const uint64_t* bKey;    // movers, input order; bPerm[j] is the j-th smallest
    const uint32_t* bCal;
    const uint32_t* bIdx;
    const uint32_t* bPerm;
    int na, nb;
};

__device__ inline void loadA(const MergeIn& m, int t, uint64_t& k, uint32_t& c)
{
    k = m.aKey[t];
    c = m.aCal[t];
}

__device__ inline void loadB(const MergeIn& m, int t, uint64_t& k, uint32_t& c)
{
    const uint32_t p = m.bPerm[t];
    k = m.bKey[p];
    c = m.bCal[p];
}

// Number of A entries among the d smallest of A and B.
__device__ int mergeSplitGlobal(const MergeIn& m, int d)
{
    int lo = max(0, d - m.nb), hi = min(d, m.na);
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        uint64_t ka, kb;
        uint32_t ca, cb;
        loadA(m, mid, ka, ca);
        loadB(m, d - mid - 1, kb, cb);
        if (pairLess(ka, ca, kb, cb)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kMergeThreads) void mergeKernel(MergeIn m, uint64_t* __restrict__ outKey,
                                                             uint32_t* __restrict__ outIdx,
                                                             int* __restrict__ outOrder)
{
    __shared__ uint64_t sKey[kMergeTile];
    __shared__ uint32_t sCal[kMergeTile];
    __shared__ uint32_t sIdx[kMergeTile];
    __shared__ int split[2];
    const int n = m.na + m.nb;
    const int d0 = blockIdx.x * kMergeTile;
    const int d1 = min(d0 + kMergeTile, n);
    if (threadIdx.x < 2) split[threadIdx.x] = mergeSplitGlobal(m, threadIdx.x ? d1 : d0);
    __syncthreads();
    const int a0 = split[0], a1 = split[1];
    const int b0 = d0 - a0, b1 = d1 - a1;
    const int la = a1 - a0, lb = b1 - b0;
    for (int t = threadIdx.x; t < la + lb; t += kMergeThreads)
    {
        if (t < la)
        {
            sKey[t] = m.aKey[a0 + t];
            sCal[t] = m.aCal[a0 + t];
            sIdx[t] = m.aIdx[a0 + t];
        }
        else
        {
            const uint32_t p = m.bPerm[b0 + t - la];
            sKey[t] = m.bKey[p];
            sCal[t] = m.bCal[p];
            sIdx[t] = m.bIdx[p];
        }
    }
    __syncthreads();
    const int e0 = threadIdx.x * kMergeItems;
    if (e0 >= la + lb) return;
    const int e1 = min(e0 + kMergeItems, la + lb);
    // this thread's split within the tile: A entries among the e0 smallest
    int lo = max(0, e0 - lb), hi = min(e0, la);
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (pairLess(sKey[mid], sCal[mid], sKey[la + e0 - mid - 1], sCal[la + e0 - mid - 1])) lo = mid + 1;
        else hi = mid;
    }
    int ia = lo, ib = e0 - lo;
    for (int e = e0; e < e1; ++e)
    {
        bool takeA;
        if (ia >= la) takeA = false;
        else if (ib >= lb) takeA = true;
        else takeA = pairLess(sKey[ia], sCal[ia], sKey[la + ib], sCal[la + ib]);
        const int s = takeA ? ia++ : la + ib++;
        outKey[d0 + e] = sKey[s];
        outIdx[d0 + e] = sIdx[s];
        outOrder[d0 + e] = (int)sCal[s];
    }
}

int resortIncremental(State& s, const uint64_t* keys, uint64_t* keysOut, uint32_t* idxOut, int* orderOut,
                      SortScratch& sc)
{
    hipStream_t st = s.stream;
    const int n = s.n;
    const int oldAlive = s.totalAlive, oldLeaves = s.nLeaves;
    uint32_t* stay = s.rsStay.get<uint32_t>((size_t)n + 1);
    uint32_t* pos = s.rsPos.get<uint32_t>((size_t)n + 2);
    uint64_t* wKey = s.rsKey.get<uint64_t>((size_t)n + 1);
    uint32_t* wCal = s.rsCal.get<uint32_t>((size_t)n + 1);
    uint32_t* wIdx = s.rsIdx.get<uint32_t>((size_t)n + 1);
    uint64_t* sKey = s.rsSortKey.get<uint64_t>(2 * ((size_t)n + 1));
    uint32_t* sVal = s.rsSortVal.get<uint32_t>(2 * ((size_t)n + 1));
    uint64_t* sKeyAlt = sKey + n + 1;
    uint32_t* sValAlt = sVal + n + 1;
    const int* order = s.order.as<int>();

    stayFlagKernel<<<iceil(n, 256), 256, 0, st>>>(keys, n, oldAlive, s.particleLeaf.as<int>(),
                                                  s.leafStart.as<uint64_t>(), s.leafLevel.as<int>(),
                                                  s.layout.as<uint32_t>(), stay);
    LAUNCH_CHECK();
    scanExclusive<uint32_t>(stay, pos, n, s.scanTmp, nullptr, st);
    uint32_t na = 0;
    HIP_CHECK(hipMemcpyAsync(&na, pos + n, sizeof(na), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    const int nb = n - (int)na;
    moverWriteKernel<<<iceil(n, 256), 256, 0, st>>>(keys, order, stay, pos, n, (int)na, wKey, wCal, wIdx, sKey, sVal);
    LAUNCH_CHECK();
    if (oldLeaves > 0)
    {
        staySortKernel<<<iceil(oldLeaves, 256 / kWave), 256, 0, st>>>(keys, order, stay, pos, s.layout.as<uint32_t>(),
                                                                      oldLeaves, wKey, wCal, wIdx);
        LAUNCH_CHECK();
    }
    // movers: by caller index, then stably by key
    const uint32_t* perm = sVal;
    if (nb > 1)
    {
        if (radixSortPairs(sKey, sKeyAlt, sVal, sValAlt, nb, 32, sc, st))
        {
            std::swap(sKey, sKeyAlt);
            std::swap(sVal, sValAlt);
        }
        moverKeysKernel<<<iceil(nb, 256), 256, 0, st>>>(wKey + na, sVal, nb, sKeyAlt, sValAlt);
        LAUNCH_CHECK();
        std::swap(sKey, sKeyAlt);
        std::swap(sVal, sValAlt);
        if (radixSortPairs(sKey, sKeyAlt, sVal, sValAlt, nb, 64, sc, st)) perm = sValAlt;
        else perm = sVal;
    }
    MergeIn m;
    m.aKey = wKey; m.aCal = wCal; m.aIdx = wIdx;
    m.bKey = wKey + na; m.bCal = wCal + na; m.bIdx = wIdx + na;
    m.bPerm = perm;
    m.na = (int)na;
    m.nb = nb;
    if (n > 0)
    {
        mergeKernel<<<iceil(n, kMergeTile), kMergeThreads, 0, st>>>(m, keysOut, idxOut, orderOut);
        LAUNCH_CHECK();
    }
    return nb;
}

} // namespace csph
