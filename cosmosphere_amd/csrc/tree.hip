/*
 * tree.hip -- root box, Hilbert keys, sort, permutation and the octree, for a batch of
 * independent particle sets in one pass of kernels.  Restates src/tree.cu:28-200 (and the
 * Cornerstone builders it calls) in a form built for gfx950.
 *
 * Leaf tree without iteration.  Cornerstone's updateOctreeGpu loop (tree.cu:121-136)
 * converges to the unique tree "split a node iff it holds more than 64 live keys and is
 * above level 21".  Such a node exists exactly where 65 consecutive sorted keys share its
 * prefix, so one pass over the sorted keys finds every internal node: particle p opens the
 * levels l with  commonDigits(k[p-1], k[p]) < l <= min(commonDigits(k[p], k[p+64]), 20).
 * The emitted list is in key (pre-)order; one radix sort puts it in (set, level, key)
 * order, the node numbering used here (children of internal node k are the 8 consecutive
 * nodes 1 + 8k of its set).  A leaf's index in key order follows in closed form from the
 * number A of internal nodes whose start key precedes it, the number a of its ancestors
 * that do, and its digit sum:  leaf = 7 (A - a) + sum of digits.  No host round trip per
 * level, no rebalance iterations.
 */
#include "hilbert.hpp"
#include "prims.hpp"
#include "tree.hpp"

#include <algorithm>
#include <vector>

namespace csph
{

// ---------------------------------------------------------------------------------------
// Root box: src/tree.cu:46-73 (open box over ALL particles of the set, dead included,
// padded x1.00001 on the largest span and centred).
// ---------------------------------------------------------------------------------------
constexpr int kBoxChunk = 16384;

__global__ __launch_bounds__(256) void bboxPartialKernel(const double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ z, const int* __restrict__ chunkSet,
                                                         const int* __restrict__ chunkBegin,
                                                         const int* __restrict__ chunkEnd, double* __restrict__ part)
{
    (void)chunkSet;
    __shared__ double red[6][4];
    const int c = blockIdx.x;
    const int b = chunkBegin[c], e = chunkEnd[c];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = b + threadIdx.x; i < e; i += blockDim.x)
    {
        const double v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            mn[a] = fmin(mn[a], v[a]);
            mx[a] = fmax(mx[a], v[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        mn[a] = waveReduceMin(mn[a]);
        mx[a] = waveReduceMax(mx[a]);
    }
    const int w = threadIdx.x / kWave;
    if (laneId() == 0)
        for (int a = 0; a < 3; ++a) { red[a][w] = mn[a]; red[3 + a][w] = mx[a]; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int a = 0; a < 3; ++a)
        {
            double m0 = red[a][0], m1 = red[3 + a][0];
            for (int ww = 1; ww < (int)(blockDim.x / kWave); ++ww)
            {
                m0 = fmin(m0, red[a][ww]);
                m1 = fmax(m1, red[3 + a][ww]);
            }
            part[6 * c + a]     = m0;
            part[6 * c + 3 + a] = m1;
        }
    }
}

__global__ void bboxFinalKernel(const double* __restrict__ part, const int* __restrict__ setChunkFirst,
                                const int* __restrict__ setChunkN, DBox* __restrict__ boxes, int nSets)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nSets) return;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = setChunkFirst[b]; c < setChunkFirst[b] + setChunkN[b]; ++c)
        for (int a = 0; a < 3; ++a)
        {
            mn[a] = fmin(mn[a], part[6 * c + a]);
            mx[a] = fmax(mx[a], part[6 * c + 3 + a]);
        }
    double span = fmax(fmax(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]) * 1.00001;
    const double half = 0.5 * span;
    double bb[6];
    for (int a = 0; a < 3; ++a)
    {
        const double c = 0.5 * (mn[a] + mx[a]);
        bb[2 * a]     = c - half;
        bb[2 * a + 1] = c + half;
    }
    boxes[b] = makeDBox(bb, 0);
}

// Set that owns index i, given the ascending first indices of the sets.
__device__ inline int setOf(const int* __restrict__ first, int nSets, int i)
{
    int lo = 0, hi = nSets;   // last b with first[b] <= i
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= i) lo = mid;
        else                 hi = mid;
    }
    return lo;
}

// Hilbert keys: include/tree.cuh:89-105 (h <= 0 -> ~0, parked after the live particles)
__global__ __launch_bounds__(256) void keysKernel(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ z, const double* __restrict__ h, int n,
                                                  const int* __restrict__ setFirst, int nSets,
                                                  const DBox* __restrict__ boxes, uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = nSets > 1 ? setOf(setFirst, nSets, i) : 0;
    const DBox& bx = boxes[b];
    uint64_t k = kDeadKey;
    if (h[i] > 0.0)
        k = hilbertEncode(quantise(x[i], bx.lo[0], bx.il[0]), quantise(y[i], bx.lo[1], bx.il[1]),
                          quantise(z[i], bx.lo[2], bx.il[2]));
    keys[i] = k;
    vals[i] = (uint32_t)i;
}

// Batch only: second sort key = set of the particle, or nSets for a dead one.
__global__ void setKeyKernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int n,
                             const int* __restrict__ setFirst, int nSets, uint64_t* __restrict__ setKey,
                             uint32_t* __restrict__ pos)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    setKey[i] = (keys[i] == kDeadKey) ? (uint64_t)nSets : (uint64_t)setOf(setFirst, nSets, (int)vals[i]);
    pos[i] = (uint32_t)i;
}

__global__ void gatherKeysVals(const uint64_t* __restrict__ k1, const uint32_t* __restrict__ v1,
                               const uint32_t* __restrict__ pos, int n, uint64_t* __restrict__ k2,
                               uint32_t* __restrict__ v2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = pos[i];
    k2[i] = k1[p];
    v2[i] = v1[p];
}

// Live ranges.  Single set: count of live keys.  Batch: first/end of each set's run in
// the (set, key) order, dead particles (key nSets) last.
__global__ void aliveSingleKernel(const uint64_t* __restrict__ keys, int n, int* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (keys[i] != kDeadKey && (i == n - 1 || keys[i + 1] == kDeadKey)) out[0] = i + 1;
}

__global__ void aliveBatchKernel(const uint64_t* __restrict__ setKey, int n, int nSets, int* __restrict__ first,
                                 int* __restrict__ end)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t sk = setKey[i];
    if (sk >= (uint64_t)nSets) return;
    if (i == 0 || setKey[i - 1] != sk) first[sk] = i;
    if (i == n - 1 || setKey[i + 1] != sk) end[sk] = i + 1;
}

// Gather of up to 10 FP64 arrays through the sort permutation: src/tree.cu:99-108
struct PermArgs
{
    const double* src[10];
    double*       dst[10];
    int           count;
};

__global__ __launch_bounds__(256) void permuteKernel(PermArgs a, const int* __restrict__ order, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = order[i];
#pragma unroll
    for (int k = 0; k < 10; ++k)
        if (k < a.count) a.dst[k][i] = a.src[k][o];
}

__global__ void valsToOrder(const uint32_t* __restrict__ vals, int* __restrict__ order, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) order[i] = (int)vals[i];
}

// ---------------------------------------------------------------------------------------
// Internal nodes from 65-key windows
// ---------------------------------------------------------------------------------------
struct WinInfo
{
    int e, cmax, b;
};

// bucket: a node holding more than `bucket` keys is split (64 is the reference's bucket and
// the size at which tree and lists are compared bit for bit; other sizes are speed options).
__device__ inline WinInfo windowOf(const uint64_t* __restrict__ keys, int i, const int* __restrict__ aliveFirst,
                                   const int* __restrict__ alive, int nSets, int bucket)
{
    WinInfo w;
    w.b = nSets > 1 ? setOf(aliveFirst, nSets, i) : 0;
    const int q = i - aliveFirst[w.b], nb = alive[w.b];
    const uint64_t k = keys[i];
    const int c = (q + bucket < nb) ? commonDigits(k, keys[i + bucket]) : -1;
    w.e    = (q > 0) ? commonDigits(keys[i - 1], k) : -1;
    w.cmax = c < kMaxLevel - 1 ? c : kMaxLevel - 1;
    return w;
}

__global__ __launch_bounds__(256) void windowCountKernel(const uint64_t* __restrict__ keys, int nAliveTot,
                                                         const int* __restrict__ aliveFirst,
                                                         const int* __restrict__ alive, int nSets, int bucket,
                                                         uint32_t* __restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nAliveTot) return;
    const WinInfo w = windowOf(keys, i, aliveFirst, alive, nSets, bucket);
    cnt[i] = (w.cmax > w.e) ? (uint32_t)(w.cmax - w.e) : 0u;
}

__global__ __launch_bounds__(256) void windowEmitKernel(const uint64_t* __restrict__ keys, int nAliveTot,
                                                        const int* __restrict__ aliveFirst,
                                                        const int* __restrict__ alive, int nSets, int bucket,
                                                        const uint32_t* __restrict__ off,
                                                        uint64_t* __restrict__ preStart, int* __restrict__ preLevel,
                                                        int* __restrict__ preSet, uint64_t* __restrict__ sortKey,
                                                        uint32_t* __restrict__ sortVal)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nAliveTot) return;
    const WinInfo w = windowOf(keys, i, aliveFirst, alive, nSets, bucket);
    const uint64_t k = keys[i];
    uint32_t o = off[i];
    for (int l = w.e + 1; l <= w.cmax; ++l, ++o)
    {
        preStart[o] = keyPrefix(k, l);
        preLevel[o] = l;
        preSet[o]   = w.b;
        sortKey[o]  = ((uint64_t)w.b << 40) | ((uint64_t)l << 35) | (uint64_t)o;
        sortVal[o]  = o;
    }
}

__global__ void gatherInternal(const uint32_t* __restrict__ idx, int nInt, const uint64_t* __restrict__ preStart,
                               const int* __restrict__ preLevel, const int* __restrict__ preSet,
                               uint64_t* __restrict__ intStart, int* __restrict__ intLevel, int* __restrict__ intSet)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nInt) return;
    const uint32_t p = idx[k];
    intStart[k] = preStart[p];
    intLevel[k] = preLevel[p];
    intSet[k]   = preSet[p];
}

__global__ void setInternalCounts(const uint32_t* __restrict__ off, const int* __restrict__ aliveFirst,
                                  const int* __restrict__ alive, int nSets, int* __restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nSets) return;
    out[b] = (int)(off[aliveFirst[b] + alive[b]] - off[aliveFirst[b]]);
}

// ---------------------------------------------------------------------------------------
// Linking: node ids, leaves, geometry
// ---------------------------------------------------------------------------------------
struct TreeOut
{
    int*      err;          // device counter of inconsistent leaf indices (must stay 0)
    int*      nodeChild;
    double*   cx; double* cy; double* cz; double* sx; double* sy; double* sz;
    uint64_t* leafStart;
    int*      leafLevel;
    int*      leafSet;
    int*      leafNode;
    int*      intNode;
};

__device__ inline void writeGeometry(const TreeOut& t, int node, uint64_t start, int level, const DBox& box)
{
    double c[3], s[3];
    nodeCentreSize(start, level, box, c, s);
    t.cx[node] = c[0]; t.cy[node] = c[1]; t.cz[node] = c[2];
    t.sx[node] = s[0]; t.sy[node] = s[1]; t.sz[node] = s[2];
}

__global__ void rootKernel(TreeOut t, const int* __restrict__ intFirst, const int* __restrict__ intN,
                           const int* __restrict__ leafFirst, const int* __restrict__ nodeFirst,
                           const DBox* __restrict__ boxes, int nSets)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nSets) return;
    const int root = nodeFirst[b];
    writeGeometry(t, root, 0, 0, boxes[b]);
    if (intN[b] > 0)
    {
        t.nodeChild[root]       = root + 1;
        t.intNode[intFirst[b]] = root;
    }
    else
    {
        const int leaf     = leafFirst[b];
        t.nodeChild[root]  = ~leaf;
        t.leafStart[leaf]  = 0;
        t.leafLevel[leaf]  = 0;
        t.leafSet[leaf]    = b;
        t.leafNode[leaf]   = root;
    }
}

// (level, start) lexicographic search of set b's internal nodes [lo, hi)
__device__ inline int findInternal(const uint64_t* __restrict__ intStart, const int* __restrict__ intLevel, int lo,
                                   int hi, int level, uint64_t start)
{
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        const int ml  = intLevel[mid];
        if (ml < level || (ml == level && intStart[mid] < start)) lo = mid + 1;
        else                                                       hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void linkKernel(TreeOut t, int nInt, const uint64_t* __restrict__ intStart,
                                                  const int* __restrict__ intLevel, const int* __restrict__ intSet,
                                                  const uint64_t* __restrict__ preStart,
                                                  const int* __restrict__ intFirst, const int* __restrict__ intN,
                                                  const int* __restrict__ leafFirst,
                                                  const int* __restrict__ nodeFirst, const DBox* __restrict__ boxes)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= 8 * nInt) return;
    const int k = tid >> 3, d = tid & 7;
    const int b = intSet[k], l = intLevel[k];
    const int i0 = intFirst[b], ni = intN[b];
    const int child = nodeFirst[b] + 1 + 8 * (k - i0) + d;
    const int cl = l + 1;
    const uint64_t cs = intStart[k] + (uint64_t)d * levelSpan(cl);
    int found = -1;
    if (cl <= kMaxLevel - 1)
    {
        const int p = findInternal(intStart, intLevel, i0, i0 + ni, cl, cs);
        if (p < i0 + ni && intLevel[p] == cl && intStart[p] == cs) found = p;
    }
    if (found >= 0)
    {
        t.nodeChild[child] = nodeFirst[b] + 1 + 8 * (found - i0);
        t.intNode[found]   = child;
    }
    else
    {
        // internal nodes of the set with start key < cs (pre-order list is sorted by start)
        int lo = i0, hi = i0 + ni;
        while (lo < hi)
        {
            const int mid = (lo + hi) >> 1;
            if (preStart[mid] < cs) lo = mid + 1;
            else                    hi = mid;
        }
        const int A = lo - i0;
        int a = 0, ds = 0;
        for (int tl = 0; tl < cl; ++tl)
        {
            const uint64_t low = (uint64_t(1) << (63 - 3 * tl)) - 1;
            a += (cs & low) != 0;
            ds += (int)((cs >> (63 - 3 * (tl + 1))) & 7u);
        }
        const int local      = 7 * (A - a) + ds;
        if (local < 0 || local > 7 * ni)
        {
            atomicAdd(t.err, 1);
            return;
        }
        const int leaf       = leafFirst[b] + local;
        t.nodeChild[child]   = ~leaf;
        t.leafStart[leaf]    = cs;
        t.leafLevel[leaf]    = cl;
        t.leafSet[leaf]      = b;
        t.leafNode[leaf]     = child;
    }
    writeGeometry(t, child, cs, cl, boxes[b]);
}

// Leaves under each internal node: a set's leaves are numbered in start-key order, so the
// leaves of the node's key range [start, start + span) are those between two lower bounds.
__global__ __launch_bounds__(256) void nodeLeavesKernel(int nInt, const uint64_t* __restrict__ intStart,
                                                        const int* __restrict__ intLevel,
                                                        const int* __restrict__ intSet,
                                                        const int* __restrict__ intNode,
                                                        const uint64_t* __restrict__ leafStart,
                                                        const int* __restrict__ leafFirst,
                                                        const int* __restrict__ intN, int* __restrict__ nodeLeaves)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nInt) return;
    const int b = intSet[k];
    const int l0 = leafFirst[b], nl = 7 * intN[b] + 1;
    const uint64_t start = intStart[k], end = start + levelSpan(intLevel[k]);
    auto lower = [&](uint64_t key) {
        int lo = l0, hi = l0 + nl;
        while (lo < hi)
        {
            const int mid = (lo + hi) >> 1;
            if (leafStart[mid] < key) lo = mid + 1;
            else                      hi = mid;
        }
        return lo;
    };
    nodeLeaves[intNode[k]] = lower(end) - lower(start);
}

// The full-walk order: the list the depth-first walk of a list would produce from node N if
// every node under N were accepted, out_full(N) = N's leaf children in octant order, then
// out_full(C) of its internal children C in reverse octant order (the LIFO pops).  Each
// node's out_full is a contiguous range of one permutation of the leaves: fullStart[N] is
// where it begins (its length is nodeLeaves[N]), fullOrder[position] the leaf.  Built top
// down, one launch per level; a set's root starts at the set's first leaf.
__global__ void fullRootKernel(const int* __restrict__ nodeFirst, const int* __restrict__ leafFirst,
                               const int* __restrict__ intN, const int* __restrict__ nodeChild, int nSets,
                               int* __restrict__ fullStart, int* __restrict__ fullOrder)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nSets) return;
    const int root = nodeFirst[b];
    fullStart[root] = leafFirst[b];
    if (intN[b] == 0) fullOrder[leafFirst[b]] = ~nodeChild[root];
}

__global__ void fullLevelKernel(const int* __restrict__ intLevel, const int* __restrict__ intNode, int nInt, int level,
                                const int* __restrict__ nodeChild, const int* __restrict__ nodeLeaves,
                                int* __restrict__ fullStart, int* __restrict__ fullOrder)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nInt || intLevel[k] != level) return;
    const int node = intNode[k];
    const int c0 = nodeChild[node];
    int ref[8];
    int nLeafChildren = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d)
    {
        ref[d] = nodeChild[c0 + d];
        nLeafChildren += ref[d] < 0 ? 1 : 0;
    }
    const int s0 = fullStart[node];
    int pl = s0, pi = s0 + nLeafChildren;
#pragma unroll
    for (int d = 0; d < 8; ++d)
        if (ref[d] < 0) fullOrder[pl++] = ~ref[d];
#pragma unroll
    for (int d = 7; d >= 0; --d)
        if (ref[d] >= 0)
        {
            fullStart[c0 + d] = pi;
            pi += nodeLeaves[c0 + d];
        }
}

// Tight box of each leaf's particles: centre and half-size per axis, the half-size rounded up
// by 2^-50 of the coordinates' magnitude so that the box contains every particle.  One
// wavefront per leaf.
__global__ __launch_bounds__(256) void leafBoxKernel(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const uint32_t* __restrict__ layout,
                                                     int nLeaves, double* __restrict__ lb)
{
    const int l = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (l >= nLeaves) return;
    const uint32_t p0 = layout[l], p1 = layout[l + 1];
    const double* c[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        double lo = INFINITY, hi = -INFINITY;
        for (uint32_t p = p0 + laneId(); p < p1; p += kWave)
        {
            const double v = c[a][p];
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
        lo = waveReduceMin(lo);
        hi = waveReduceMax(hi);
        if (laneId() == 0)
        {
            double cen = 0.0, half = 0.0;
            if (p1 > p0)
            {
                cen = 0.5 * (lo + hi);
                half = 0.5 * (hi - lo) + 0x1p-50 * fmax(fabs(lo), fabs(hi));
            }
            lb[(size_t)a * nLeaves + l] = cen;
            lb[(size_t)(3 + a) * nLeaves + l] = half;
        }
    }
}

__global__ __launch_bounds__(256) void layoutKernel(const uint64_t* __restrict__ keys,
                                                    const uint64_t* __restrict__ leafStart,
                                                    const int* __restrict__ leafSet, int nLeaves,
                                                    const int* __restrict__ aliveFirst, const int* __restrict__ alive,
                                                    uint32_t* __restrict__ layout, int nAliveTot)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l == nLeaves) layout[nLeaves] = (uint32_t)nAliveTot;
    if (l >= nLeaves) return;
    const int b = leafSet[l];
    int lo = aliveFirst[b], hi = lo + alive[b];
    const uint64_t s = leafStart[l];
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < s) lo = mid + 1;
        else               hi = mid;
    }
    layout[l] = (uint32_t)lo;
}

__global__ __launch_bounds__(256) void particleLeafKernel(const uint32_t* __restrict__ layout, int nLeaves,
                                                          int nAliveTot, int* __restrict__ particleLeaf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nAliveTot) return;
    int lo = 0, hi = nLeaves;   // last leaf with layout[leaf] <= i
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (layout[mid] <= (uint32_t)i) lo = mid;
        else                            hi = mid;
    }
    particleLeaf[i] = lo;
}

// ---------------------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------------------
// Host vectors here are temporaries: copy on the library stream and wait.
template<class T>
static void upload(DevBuf& buf, const std::vector<T>& v, hipStream_t st)
{
    T* p = buf.get<T>(v.size() + 1);
    if (v.empty()) return;
    HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

struct Events
{
    hipEvent_t e[8];
    Events() { for (auto& x : e) HIP_CHECK(hipEventCreate(&x)); }
    ~Events() { for (auto& x : e) (void)hipEventDestroy(x); }
};

static double elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms * 1e-3;
}

void buildTree(State& s, bool withVelAcc, const double* periodicBox, SortScratch& sc, bool incremental)
{
    hipStream_t st = s.stream;
    sc.err = s.ovf.as<int>() + 2;   // ovf[2]: sort / tree consistency errors
    const int n = s.n, B = s.nSets;
    Events ev;
    HIP_CHECK(hipEventRecord(ev.e[0], st));

    // ---- boxes ----
    upload(s.setFirst, s.hSetFirst, st);
    upload(s.setN, s.hSetN, st);
    s.hBox.assign(B, DBox{});
    if (periodicBox)
    {
        for (int b = 0; b < B; ++b) s.hBox[b] = makeDBox(periodicBox + 6 * b, 1);
        upload(s.box, s.hBox, st);
    }
    else
    {
        std::vector<int> cSet, cBeg, cEnd, sFirst(B), sN(B);
        for (int b = 0; b < B; ++b)
        {
            sFirst[b] = (int)cSet.size();
            for (int i = s.hSetFirst[b]; i < s.hSetFirst[b] + s.hSetN[b]; i += kBoxChunk)
            {
                cSet.push_back(b);
                cBeg.push_back(i);
                cEnd.push_back(std::min(i + kBoxChunk, s.hSetFirst[b] + s.hSetN[b]));
            }
            sN[b] = (int)cSet.size() - sFirst[b];
        }
        const size_t nc = cSet.size();
        int* dI = s.boxDesc.get<int>(3 * nc + 2 * B + 8);
        std::vector<int> packed;
        packed.insert(packed.end(), cSet.begin(), cSet.end());
        packed.insert(packed.end(), cBeg.begin(), cBeg.end());
        packed.insert(packed.end(), cEnd.begin(), cEnd.end());
        packed.insert(packed.end(), sFirst.begin(), sFirst.end());
        packed.insert(packed.end(), sN.begin(), sN.end());
        HIP_CHECK(hipMemcpyAsync(dI, packed.data(), packed.size() * sizeof(int), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
        double* part = s.partials.get<double>(6 * nc + 6);
        DBox* boxes = s.box.get<DBox>(B);
        if (nc > 0)
        {
            bboxPartialKernel<<<(int)nc, 256, 0, st>>>(s.x.as<double>(), s.y.as<double>(), s.z.as<double>(), dI, dI + nc,
                                                      dI + 2 * nc, part);
            LAUNCH_CHECK();
        }
        bboxFinalKernel<<<iceil(B, 64), 64, 0, st>>>(part, dI + 3 * nc, dI + 3 * nc + B, boxes, B);
        LAUNCH_CHECK();
        HIP_CHECK(hipMemcpyAsync(s.hBox.data(), boxes, B * sizeof(DBox), hipMemcpyDeviceToHost, st));
    }
    HIP_CHECK(hipEventRecord(ev.e[1], st));

    // ---- keys + sort ----
    uint64_t* keys = s.keys.get<uint64_t>(n + 1);
    uint64_t* keysAlt = s.keysAlt.get<uint64_t>(n + 1);
    uint32_t* vals = s.vals.get<uint32_t>(n + 1);
    uint32_t* valsAlt = s.valsAlt.get<uint32_t>(n + 1);
    if (n > 0)
    {
        keysKernel<<<iceil(n, 256), 256, 0, st>>>(s.x.as<double>(), s.y.as<double>(), s.z.as<double>(),
                                                  s.h.as<double>(), n, s.setFirst.as<int>(), B, s.box.as<DBox>(),
                                                  keys, vals);
        LAUNCH_CHECK();
    }
    HIP_CHECK(hipEventRecord(ev.e[2], st));
    // Incremental re-sort (inputs in the previous solve's order): sorted keys into keysAlt,
    // the input position of each sorted entry into valsAlt, the caller index into orderAlt.
    s.t.resortMovers = -1;
    if (incremental && n > 0)
    {
        int* orderNext = s.orderAlt.get<int>((size_t)n + 1);
        s.t.resortMovers = resortIncremental(s, keys, keysAlt, valsAlt, orderNext, sc);
        std::swap(s.keys, s.keysAlt);
        std::swap(s.vals, s.valsAlt);
        keys = s.keys.as<uint64_t>(); keysAlt = s.keysAlt.as<uint64_t>();
        vals = s.vals.as<uint32_t>(); valsAlt = s.valsAlt.as<uint32_t>();
    }
    else if (radixSortPairs(keys, keysAlt, vals, valsAlt, n, 64, sc, st))
    {
        std::swap(s.keys, s.keysAlt);
        std::swap(s.vals, s.valsAlt);
        keys = s.keys.as<uint64_t>(); keysAlt = s.keysAlt.as<uint64_t>();
        vals = s.vals.as<uint32_t>(); valsAlt = s.valsAlt.as<uint32_t>();
    }
    s.hAliveFirst.assign(B, 0);
    s.hAlive.assign(B, 0);
    int* dAlive = s.aliveFirst.get<int>(2 * B + 2);
    fillU32(reinterpret_cast<uint32_t*>(dAlive), 0u, 2 * B + 2, st);
    if (B > 1 && n > 0)
    {
        uint64_t* setKey = s.sortSetKey.get<uint64_t>(2 * (size_t)n + 2);
        uint64_t* setKeyAlt = setKey + n + 1;
        uint32_t* pos = s.tmpA.get<uint32_t>(2 * (size_t)n + 2);
        uint32_t* posAlt = pos + n + 1;
        setKeyKernel<<<iceil(n, 256), 256, 0, st>>>(keys, vals, n, s.setFirst.as<int>(), B, setKey, pos);
        LAUNCH_CHECK();
        const bool f = radixSortPairs(setKey, setKeyAlt, pos, posAlt, n, 32, sc, st);
        if (f) { std::swap(setKey, setKeyAlt); std::swap(pos, posAlt); }
        gatherKeysVals<<<iceil(n, 256), 256, 0, st>>>(keys, vals, pos, n, keysAlt, valsAlt);
        LAUNCH_CHECK();
        std::swap(s.keys, s.keysAlt);
        std::swap(s.vals, s.valsAlt);
        keys = s.keys.as<uint64_t>();
        vals = s.vals.as<uint32_t>();
        aliveBatchKernel<<<iceil(n, 256), 256, 0, st>>>(setKey, n, B, dAlive, dAlive + B);
        LAUNCH_CHECK();
        std::vector<int> fe(2 * B);
        HIP_CHECK(hipMemcpyAsync(fe.data(), dAlive, 2 * B * sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        int prevEnd = 0;
        for (int b = 0; b < B; ++b)
        {
            const int cnt = fe[B + b] - fe[b];
            s.hAlive[b] = cnt > 0 ? cnt : 0;
            s.hAliveFirst[b] = cnt > 0 ? fe[b] : prevEnd;
            prevEnd = s.hAliveFirst[b] + s.hAlive[b];
        }
    }
    else if (n > 0)
    {
        aliveSingleKernel<<<iceil(n, 256), 256, 0, st>>>(keys, n, dAlive);
        LAUNCH_CHECK();
        int na = 0;
        HIP_CHECK(hipMemcpyAsync(&na, dAlive, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        s.hAlive[0] = na;
    }
    s.totalAlive = 0;
    for (int b = 0; b < B; ++b) s.totalAlive += s.hAlive[b];
    upload(s.aliveFirst, s.hAliveFirst, st);
    upload(s.alive, s.hAlive, st);
    HIP_CHECK(hipEventRecord(ev.e[3], st));

    // ---- permute ----
    int* order = s.order.get<int>(n + 1);
    if (n > 0)
    {
        // gather index: the caller index (full sort of caller-order inputs) or the input
        // position (incremental re-sort of inputs in the previous order; the new order was
        // composed by the re-sort)
        const int* gidx = reinterpret_cast<const int*>(vals);
        if (incremental) std::swap(s.order, s.orderAlt);
        else
        {
            valsToOrder<<<iceil(n, 256), 256, 0, st>>>(vals, order, n);
            LAUNCH_CHECK();
            gidx = order;
        }
        std::vector<DevBuf*> fields = {&s.x, &s.y, &s.z, &s.h};
        if (withVelAcc) for (DevBuf* f : {&s.vx, &s.vy, &s.vz, &s.ax, &s.ay, &s.az}) fields.push_back(f);
        // gather into the persistent alternates, then swap them in (passes of at most 10)
        PermArgs a{};
        a.count = (int)fields.size();
        for (int k = 0; k < a.count; ++k)
        {
            a.src[k] = fields[k]->as<double>();
            a.dst[k] = s.permAlt[k].get<double>(std::max(n, s.sizedFor) + 1);
        }
        permuteKernel<<<iceil(n, 256), 256, 0, st>>>(a, gidx, n);
        LAUNCH_CHECK();
        for (int k = 0; k < a.count; ++k) std::swap(*fields[k], s.permAlt[k]);
    }
    s.haveOrder = 1;
    s.orderSets = B;
    HIP_CHECK(hipEventRecord(ev.e[4], st));

    // ---- internal nodes ----
    const int na = s.totalAlive;
    uint32_t* wcnt = s.winCount.get<uint32_t>(na + 1);
    uint32_t* woff = s.winOffset.get<uint32_t>(na + 2);
    if (na > 0)
    {
        windowCountKernel<<<iceil(na, 256), 256, 0, st>>>(keys, na, s.aliveFirst.as<int>(), s.alive.as<int>(), B,
                                                           s.bucket, wcnt);
        LAUNCH_CHECK();
    }
    scanExclusive<uint32_t>(wcnt, woff, na, s.scanTmp, nullptr, st);
    int nInt = 0;
    HIP_CHECK(hipMemcpyAsync(&nInt, woff + na, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    s.nInternal = nInt;

    uint64_t* preStart = s.preStart.get<uint64_t>(nInt + 1);
    int* preLevel = s.preLevel.get<int>(nInt + 1);
    int* preSet = s.preSet.get<int>(nInt + 1);
    uint64_t* bk = s.bfsKey.get<uint64_t>(nInt + 1);
    uint64_t* bkA = s.bfsKeyAlt.get<uint64_t>(nInt + 1);
    uint32_t* bi = s.bfsIdx.get<uint32_t>(nInt + 1);
    uint32_t* biA = s.bfsIdxAlt.get<uint32_t>(nInt + 1);
    if (na > 0)
    {
        windowEmitKernel<<<iceil(na, 256), 256, 0, st>>>(keys, na, s.aliveFirst.as<int>(), s.alive.as<int>(), B,
                                                         s.bucket, woff, preStart, preLevel, preSet, bk, bi);
        LAUNCH_CHECK();
    }
    if (radixSortPairs(bk, bkA, bi, biA, nInt, 64, sc, st)) { std::swap(bk, bkA); std::swap(bi, biA); }
    uint64_t* intStart = s.intStart.get<uint64_t>(nInt + 1);
    int* intLevel = s.intLevel.get<int>(nInt + 1);
    int* intSet = s.intSet.get<int>(nInt + 1);
    if (nInt > 0)
    {
        gatherInternal<<<iceil(nInt, 256), 256, 0, st>>>(bi, nInt, preStart, preLevel, preSet, intStart, intLevel, intSet);
        LAUNCH_CHECK();
    }

    // per-set counts -> leaf / node / internal bases
    s.hIntN.assign(B, 0);
    {
        int* dcnt = s.intFirst.get<int>(2 * B + 2);
        setInternalCounts<<<iceil(B, 64), 64, 0, st>>>(woff, s.aliveFirst.as<int>(), s.alive.as<int>(), B, dcnt);
        LAUNCH_CHECK();
        HIP_CHECK(hipMemcpyAsync(s.hIntN.data(), dcnt, B * sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
    s.hIntFirst.assign(B, 0);
    s.hLeafFirst.assign(B, 0);
    s.hLeafN.assign(B, 0);
    s.hNodeFirst.assign(B, 0);
    int accI = 0, accL = 0, accM = 0;
    for (int b = 0; b < B; ++b)
    {
        s.hIntFirst[b]  = accI;
        s.hLeafFirst[b] = accL;
        s.hNodeFirst[b] = accM;
        s.hLeafN[b]     = 7 * s.hIntN[b] + 1;
        accI += s.hIntN[b];
        accL += s.hLeafN[b];
        accM += 8 * s.hIntN[b] + 1;
    }
    s.nLeaves = accL;
    s.nNodes  = accM;
    std::vector<int> pack;
    pack.insert(pack.end(), s.hIntFirst.begin(), s.hIntFirst.end());
    pack.insert(pack.end(), s.hIntN.begin(), s.hIntN.end());
    pack.insert(pack.end(), s.hLeafFirst.begin(), s.hLeafFirst.end());
    pack.insert(pack.end(), s.hNodeFirst.begin(), s.hNodeFirst.end());
    upload(s.intFirst, pack, st);
    const int* dIntFirst = s.intFirst.as<int>();
    const int* dIntN = dIntFirst + B;
    const int* dLeafFirst = dIntFirst + 2 * B;
    const int* dNodeFirst = dIntFirst + 3 * B;
    HIP_CHECK(hipEventRecord(ev.e[5], st));

    // ---- link ----
    const int L = s.nLeaves, M = s.nNodes;
    TreeOut t;
    t.err = s.ovf.as<int>() + 2;
    t.nodeChild = s.nodeChild.get<int>(M + 1);
    t.cx = s.nodeCx.get<double>(M + 1); t.cy = s.nodeCy.get<double>(M + 1); t.cz = s.nodeCz.get<double>(M + 1);
    t.sx = s.nodeSx.get<double>(M + 1); t.sy = s.nodeSy.get<double>(M + 1); t.sz = s.nodeSz.get<double>(M + 1);
    t.leafStart = s.leafStart.get<uint64_t>(L + 1);
    t.leafLevel = s.leafLevel.get<int>(L + 1);
    t.leafSet = s.leafSet.get<int>(L + 1);
    t.leafNode = s.leafNode.get<int>(L + 1);
    t.intNode = s.intNode.get<int>(nInt + 1);
    rootKernel<<<iceil(B, 64), 64, 0, st>>>(t, dIntFirst, dIntN, dLeafFirst, dNodeFirst, s.box.as<DBox>(), B);
    LAUNCH_CHECK();
    if (nInt > 0)
    {
        linkKernel<<<iceil(8LL * nInt, 256), 256, 0, st>>>(t, nInt, intStart, intLevel, intSet, preStart, dIntFirst,
                                                            dIntN, dLeafFirst, dNodeFirst, s.box.as<DBox>());
        LAUNCH_CHECK();
    }
    int* nodeLeaves = s.nodeLeaves.get<int>(M + 1);
    if (nInt > 0)
    {
        nodeLeavesKernel<<<iceil(nInt, 256), 256, 0, st>>>(nInt, intStart, intLevel, intSet, t.intNode, t.leafStart,
                                                            dLeafFirst, dIntN, nodeLeaves);
        LAUNCH_CHECK();
    }
    {
        int* fullStart = s.fullStart.get<int>(M + 1);
        int* fullOrder = s.fullOrder.get<int>(L + 1);
        fullRootKernel<<<iceil(B, 64), 64, 0, st>>>(dNodeFirst, dLeafFirst, dIntN, t.nodeChild, B, fullStart, fullOrder);
        LAUNCH_CHECK();
        if (nInt > 0)
            for (int lvl = 0; lvl < kMaxLevel; ++lvl)
            {
                fullLevelKernel<<<iceil(nInt, 256), 256, 0, st>>>(s.intLevel.as<int>(), t.intNode, nInt, lvl, t.nodeChild,
                                                                 nodeLeaves, fullStart, fullOrder);
                LAUNCH_CHECK();
            }
    }
    uint32_t* layout = s.layout.get<uint32_t>(L + 2);
    layoutKernel<<<iceil(L + 1, 256), 256, 0, st>>>(keys, t.leafStart, t.leafSet, L, s.aliveFirst.as<int>(),
                                                    s.alive.as<int>(), layout, na);
    LAUNCH_CHECK();
    double* lbox = s.leafBox.get<double>((size_t)6 * L + 1);
    leafBoxKernel<<<iceil(L, 4), 256, 0, st>>>(s.x.as<double>(), s.y.as<double>(), s.z.as<double>(), layout, L, lbox);
    LAUNCH_CHECK();
    int* pleaf = s.particleLeaf.get<int>(std::max(n, 1));
    if (na > 0)
    {
        particleLeafKernel<<<iceil(na, 256), 256, 0, st>>>(layout, L, na, pleaf);
        LAUNCH_CHECK();
    }

    HIP_CHECK(hipEventRecord(ev.e[6], st));
    HIP_CHECK(hipEventSynchronize(ev.e[6]));
    int herr = 0;
    HIP_CHECK(hipMemcpyAsync(&herr, s.ovf.as<int>() + 2, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (herr != 0)
    {
        std::fprintf(stderr, "FATAL: buildTree: %d inconsistent sort / tree entries\n", herr);
        std::abort();
    }
    s.t.bbox    = elapsed(ev.e[0], ev.e[1]);
    s.t.keys    = elapsed(ev.e[1], ev.e[2]);
    s.t.sort    = elapsed(ev.e[2], ev.e[3]);
    s.t.permute = elapsed(ev.e[3], ev.e[4]);
    s.t.tree    = elapsed(ev.e[4], ev.e[5]);
    s.t.nodes   = elapsed(ev.e[5], ev.e[6]);
    s.t.nLeaves = L;
}

} // namespace csph
