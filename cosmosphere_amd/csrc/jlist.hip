/*
 * jlist.hip -- smoothing-length maxima and the neighbour-leaf (j-leaf) lists.
 *
 * Restates include/tree.cuh:176-389 and src/force.cu:19-88.  The list of an i-leaf is the
 * reference's depth-first walk: test the root; then for each popped node visit its 8
 * children in key order, append accepted leaves, push accepted internal nodes, pop LIFO.
 *
 * Mapping for gfx950: a group of 8 lanes walks one i-leaf, lane o testing child o of the
 * current node, so the 8 children are tested at once and read with unit stride from the
 * SoA node geometry; a ballot per group turns the accepted children into list slots (in
 * octant order) and stack pushes (in octant order, so pops are LIFO exactly as the
 * reference's).  Eight groups share a wavefront (64 lanes busy), and each group's stack
 * lives in LDS instead of a 768-byte private array.  A LIFO depth-first stack never holds
 * more than 7 siblings per level plus the 8 children just pushed: 7*20 + 8 = 148 entries.
 *
 * The accept tests are evaluated exactly as written (no contraction), which keeps the
 * lists bit-identical to the FP64 oracle's.
 *
 * Tight lists (Newton-Raphson passes).  The reference's test accepts a node that overlaps
 * the i-leaf box grown by R hmax along each axis: a cube around the support spheres, whose
 * corners hold no pair.  A tight list adds a second test: the Euclidean gap between the two
 * boxes (each widened by one finest cell: a particle lies within one finest cell of its
 * node's box, since the key quantisation floor(x m) - lo m truncates up to one cell low;
 * relative slack 1e-5) within R max(h) of the particles still to be evaluated.  A node it
 * drops holds no pair inside any such particle's support, and a dropped internal node's
 * leaves are dropped with it, so the tight list is the reference list minus leaves without
 * such pairs, in the same depth-first order: the accepted pairs and their order are
 * unchanged.
 */
#include "density.hpp"
#include "tile.hpp"
#include "tree.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace csph
{

// hmax_leaf = max finite h of the leaf's particles: include/tree.cuh:177-191.  With `state`,
// also hRem = max finite h of the particles in state != 0 (the tight-list reach).
// One wavefront per leaf, 4 per block.
__global__ __launch_bounds__(256) void hmaxLeafKernel(const double* __restrict__ h, const uint32_t* __restrict__ layout,
                                                      const int* __restrict__ leaves, int nLeaves,
                                                      double* __restrict__ hmaxLeaf, const uint8_t* __restrict__ state,
                                                      double* __restrict__ hRem, const PassCtl* __restrict__ ctl)
{
    if (ctl && ctl->skip) return;
    const int n = ctl ? (int)ctl->nAL : nLeaves;
    for (int a = blockIdx.x * 4 + (threadIdx.x >> 6); a < n; a += gridDim.x * 4)
    {
        const int l = leaves ? leaves[a] : a;
        const uint32_t p0 = layout[l], p1 = layout[l + 1];
        double m = 0.0, r = 0.0;
        for (uint32_t p = p0 + laneId(); p < p1; p += kWave)
        {
            const double v = h[p];
            if (isfinite(v))
            {
                m = fmax(m, v);
                if (state && state[p] != 0) r = fmax(r, v);
            }
        }
        m = waveReduceMax(m);
        if (laneId() == 0) hmaxLeaf[l] = m;
        if (state)
        {
            r = waveReduceMax(r);
            if (laneId() == 0) hRem[l] = r;
        }
    }
}

// Upsweep: include/tree.cuh:196-217 (a leaf node takes its hmax_leaf if finite and > 0,
// else 0; an internal node the max of its 8 children).
__global__ void nodeHmaxLeafKernel(const double* __restrict__ hmaxLeaf, const int* __restrict__ leafNode, int nLeaves,
                                   double* __restrict__ nodeHmax)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nLeaves) return;
    const double hm = hmaxLeaf[l];
    nodeHmax[leafNode[l]] = (isfinite(hm) && hm > 0.0) ? hm : 0.0;
}

__global__ void nodeHmaxInternalKernel(const int* __restrict__ intLevel, const int* __restrict__ intNode, int nInt,
                                       int level, const int* __restrict__ nodeChild, double* __restrict__ nodeHmax)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nInt || intLevel[k] != level) return;
    const int node = intNode[k];
    const int c = nodeChild[node];
    double m = 0.0;
#pragma unroll
    for (int o = 0; o < 8; ++o) m = fmax(m, nodeHmax[c + o]);
    nodeHmax[node] = m;
}

// Cornerstone minDistance (box-box), squared: |cb - ca| - sa - sb, clipped at 0 as
// (d + |d|) / 2; periodic: nearest image of cb - ca first.
template<bool PER>
__device__ inline double boxDist2(const double ac[3], const double as[3], double bcx, double bcy, double bcz, double bsx,
                                  double bsy, double bsz, const DBox& box)
{
#pragma clang fp contract(off)
    const double bc[3] = {bcx, bcy, bcz}, bs[3] = {bsx, bsy, bsz};
    double d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        double dx = bc[a] - ac[a];
        if constexpr (PER) dx -= box.l[a] * rint(dx * box.il[a]);
        double v = fabs(dx) - as[a] - bs[a];
        v += fabs(v);
        v *= 0.5;
        d[a] = v;
    }
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

constexpr int kListThreads = 256;
constexpr int kGroups      = kListThreads / 8;   // i-leaves per block
constexpr int kStack       = 152;                // >= 7*20 + 8

// Long walks.  A group walks one list sequentially, one node per step, each step a chain of
// dependent loads; a list of tens of thousands of entries (a leaf whose particles still to
// evaluate reach into a dense region) then takes milliseconds while the rest of the chip
// idles.  A count walk that passes kHeavySteps steps hands its list to a whole block: the
// list is out(root), where out(N) = N's accepted leaf children in octant order followed by
// out(C) of its accepted internal children C in reverse octant order (the LIFO order of the
// walk).  The block rewrites out(root) as a sequence of tokens (a leaf, or out(C) of a node
// C), expanding every out(C) of a node with more than kExpandLeaves leaves while the token
// buffer holds the result; the groups then walk the out(C) tokens at once, count them, scan
// the token sizes and fill each token at its offset.  The list is the sequential walk's,
// entry for entry.
constexpr int kHeavySteps   = 256;    // per 32768 lists of the launch (a launch of many lists
                                      // keeps the chip busy through long walks); 128, 512 and
                                      // 1024 were slower on the 8M Plummer mixture.  Lowered
                                      // to a multiple of the typical walk once one is known
                                      // (ListArgs::heavyFactor).
constexpr int kHeavyMinSteps = 64;
constexpr int kTokCap       = 2048;
constexpr int kExpandLeaves = 256;

struct ListArgs
{
    const int*      leaves;
    int             nLeaves;
    const int*      leafNode;
    const int*      leafSet;
    const int*      nodeFirst;
    const int*      nodeChild;
    const double*   cx; const double* cy; const double* cz;
    const double*   sx; const double* sy; const double* sz;
    const double*   hmaxLeaf;
    const double*   hRemLeaf;   // tight lists: reach of the particles still to evaluate (nullptr: off)
    const double*   nodeHmax;
    const DBox*     boxes;
    const uint32_t* jOffset;
    int*            jList;
    uint32_t*       jCount;
    int*            ovf;
    int             nNodes;   // walk cap: a tree walk visits each node at most once
    const PassCtl*  ctl;      // device-driven pass: ctl->nAL leaves, nothing when ctl->skip
    const int*      nodeLeaves;   // long walks: leaves under each internal node
    int*            heavyN;       // ... [0] lists handed to the block walk (count), nullptr: none;
                                  //     [1] tokens taken from tokPool
    int*            heavyList;    // ... their positions
    uint8_t*        heavyFlag;    // ... per position: 1 if the block walk owns the list
    int             heavySteps;   // ... hand-over threshold per 32768 lists (kHeavySteps)
    int             expandLeaves; // ... expansion threshold (kExpandLeaves)
    int2*           tokPool;      // ... the count pass's tokens and their offsets in the list, for the fill
    uint2*          tokAt;        // ... per hand-over position: (first token in tokPool or ~0u, tokens)
    unsigned long long listMean;  // entries per leaf of the last build of this kind (0: unknown; ctl: its own)
    int             heavyFactor;  // hand-over after heavyFactor x listMean / 4 steps (0: off)
    const int*      fullStart;    // the full-walk order (buildTree): per node, its start; nullptr: off
    const int*      fullOrder;    // ... per position, the leaf
    int*            slotTok;      // light lists: the count walk's tokens, kListSlotTok per position
                                  // (>= 0: a leaf; < 0: ~C, a node whose whole subtree is listed)
    uint32_t*       slotN;        // ... per position: their number, ~0u if not kept (the fill walks)
};

// The accept test of one i-leaf's list (include/tree.cuh:228-252 and the tight test).
template<bool SYM, bool PER>
struct Reach
{
    double ic[3], is[3], aSize[3], pad[3];
    double radHi, tight2;
    DBox   box;

    __device__ Reach(const ListArgs& A, int iLeaf)
    {
        const int in = A.leafNode[iLeaf];
        box = A.boxes[A.leafSet[iLeaf]];
        ic[0] = A.cx[in]; ic[1] = A.cy[in]; ic[2] = A.cz[in];
        is[0] = A.sx[in]; is[1] = A.sy[in]; is[2] = A.sz[in];
        double hmax = A.hmaxLeaf[iLeaf];
        if (!isfinite(hmax) || hmax <= 0.0) hmax = 0.0;
        radHi = kRad * hmax;
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            aSize[a] = SYM ? is[a] : is[a] + radHi;
            pad[a] = 2.0 * kFinestCell * box.l[a];
        }
        // tight-list reach, squared (negative: no second test)
        tight2 = -1.0;
        if (!SYM && A.hRemLeaf)
        {
            const double hr = A.hRemLeaf[iLeaf];
            tight2 = (kRad * hr) * (kRad * hr) * (1.0 + 1e-5);
        }
    }

    // Whether every node under `node` passes the test: the node's box lies within the reach, with
    // a margin (1e-10 of the operands' magnitudes) far above the rounding of the descendants' own
    // tests.  Gather lists: the box inside the i-cell grown by R hmax along each axis and, with
    // the tight test, its farthest point within the tight reach of the padded i-cell; symmetric
    // lists: its farthest point within R hmax_i of the i-cell (every node's reach is at least that).
    __device__ bool inside(const ListArgs& A, int node) const
    {
        const double bc[3] = {A.cx[node], A.cy[node], A.cz[node]};
        const double bs[3] = {A.sx[node], A.sy[node], A.sz[node]};
        double far2 = 0.0, scale2 = 0.0;
        bool in = true;
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            double dx = bc[a] - ic[a];
            if constexpr (PER) dx -= box.l[a] * rint(dx * box.il[a]);
            const double ad = fabs(dx);
            const double mag = ad + bs[a] + aSize[a] + pad[a];
            if constexpr (!SYM) in = in && (ad + bs[a] <= aSize[a] - 1e-10 * mag);
            const double e = fmax(ad + bs[a] - (SYM ? is[a] : is[a] + pad[a]), 0.0);
            far2 += e * e;
            scale2 += mag * mag;
        }
        if constexpr (SYM)
        {
            return far2 <= radHi * radHi - 1e-10 * (radHi * radHi + scale2);
        }
        else
        {
            if (!in) return false;
            if (tight2 < 0.0) return true;
            return far2 <= tight2 - 1e-10 * (tight2 + scale2);
        }
    }

    __device__ bool operator()(const ListArgs& A, int node) const
    {
#pragma clang fp contract(off)
        const double d2 = boxDist2<PER>(ic, aSize, A.cx[node], A.cy[node], A.cz[node], A.sx[node], A.sy[node],
                                        A.sz[node], box);
        if constexpr (SYM)
        {
            const double rs = fmax(radHi, kRad * A.nodeHmax[node]);
            return d2 <= rs * rs;
        }
        else
        {
            if (!(d2 <= 0.0)) return false;
            if (tight2 < 0.0) return true;
            const double isPad[3] = {is[0] + pad[0], is[1] + pad[1], is[2] + pad[2]};
            const double g2 = boxDist2<PER>(ic, isPad, A.cx[node], A.cy[node], A.cz[node], A.sx[node], A.sy[node],
                                            A.sz[node], box);
            return g2 <= tight2;
        }
    }
};

// out(root) of an accepted internal node by the 8 lanes of a group (lane o tests child o of
// the current node): the count, and without COUNT the entries at jList[base ...] below limit.
// Stops after stepCap steps with capped set.  The stack holds each pushed node's first child
// (read with the accept test), so a step waits for one round of loads, not two.
// STORE (with COUNT): the walk also keeps its tokens in the fill's order, the fill's own walk:
// each leaf as it is accepted, each whole subtree (~C) when it is popped; the first
// kListSlotTok go to tok[], their number (more once past the slot, and then not exact) to *nTok.
// Counts and steps are COUNT's.
template<bool COUNT, bool STORE = false, class R>
__device__ inline uint32_t groupWalk(const ListArgs& A, const R& inRange, int root, int* __restrict__ stk,
                                     uint32_t base, uint32_t limit, int stepCap, bool& capped,
                                     int* __restrict__ tok = nullptr, uint32_t* nTok = nullptr)
{
    static_assert(COUNT || !STORE, "tokens are kept by count walks");
    const int o = threadIdx.x & 7;
    const int shift = (laneId() >> 3) << 3;
    const uint32_t below = (1u << o) - 1u;
    uint32_t count = 0, nt = 0;
    int sp = 0, first = A.nodeChild[root], steps = 0;
    capped = false;
    while (true)
    {
        if (++steps > stepCap)
        {
            capped = true;
            break;
        }
        const int  c    = first + o;
        const bool acc  = inRange(A, c);
        const int  cref = A.nodeChild[c];
        const bool isLeaf = acc && cref < 0;
        const bool isInt  = acc && cref >= 0;
        // an internal child whose whole subtree passes: the count takes its leaves at once, the
        // fill copies them from the full-walk order when the walk would pop it
        const bool full = isInt && A.fullOrder != nullptr && inRange.inside(A, c);
        // (STORE: a list past the slot keeps no more tokens, so it counts them like COUNT)
        const bool walk = isInt && !(COUNT && full && (!STORE || nt > (uint32_t)kListSlotTok));
        const uint32_t mL = (uint32_t)(ballot64(isLeaf) >> shift) & 0xFFu;
        const uint32_t mI = (uint32_t)(ballot64(walk) >> shift) & 0xFFu;
        if (!COUNT && isLeaf)
        {
            const uint32_t at = base + count + __popc(mL & below);
            if (at < limit) A.jList[at] = ~cref;
            else            atomicAdd(&A.ovf[0], 1);
        }
        if constexpr (STORE)
        {
            const uint32_t t = nt + __popc(mL & below);
            if (isLeaf && t < kListSlotTok) tok[t] = ~cref;
            nt += __popc(mL);
        }
        count += __popc(mL);
        if constexpr (COUNT)
        {
            int nl = full ? A.nodeLeaves[c] : 0;   // summed over the group's 8 lanes
            nl += __shfl_xor(nl, 1);
            nl += __shfl_xor(nl, 2);
            nl += __shfl_xor(nl, 4);
            count += (uint32_t)nl;
            // the hand-over budget counts them as walked (about 4 entries a step): a list goes to
            // the block walk as it would without the shortcut, since its fill walks what is not full
            steps += nl >> 2;
        }
        const int push = __popc(mI);
        if (sp + push <= kStack)
        {
            if (walk) stk[sp + __popc(mI & below)] = full ? ~c : cref;   // < 0: a full node to copy
            sp += push;
        }
        else if (o == 0)
        {
            atomicAdd(&A.ovf[1], 1);   // subtree dropped; cannot happen for a depth-21 tree
        }
        waveLdsSync();
        bool more = false;
        while (sp > 0)
        {
            const int v = stk[--sp];
            if (v >= 0)
            {
                first = v;
                more = true;
                break;
            }
            if constexpr (STORE)
            {
                // a full node (counted when accepted): its token
                if (o == 0 && nt < kListSlotTok) tok[nt] = v;
                ++nt;
                continue;
            }
            // (fill) a full node: its subtree's leaves, in the walk's order, 64 per round trip
            const int node = ~v;
            const int n = A.nodeLeaves[node], f0 = A.fullStart[node];
            for (int t0 = 0; t0 < n; t0 += 64)
            {
                int e[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                {
                    const int t = t0 + 8 * u + o;
                    e[u] = t < n ? A.fullOrder[f0 + t] : 0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                {
                    const int t = t0 + 8 * u + o;
                    if (t < n)
                    {
                        const uint32_t at = base + count + (uint32_t)t;
                        if (at < limit) A.jList[at] = e[u];
                        else            atomicAdd(&A.ovf[0], 1);
                    }
                }
            }
            count += (uint32_t)n;
        }
        waveLdsSync();
        if (!more) break;
    }
    if constexpr (STORE) *nTok = nt;
    return count;
}

// (fill) A light list from the tokens its count walk kept: leaves as they are, whole subtrees
// copied from the full-walk order; the same entries in the same order as the fill walk.
__device__ inline void expandTokens(const ListArgs& A, int a, uint32_t nT)
{
    const int o = threadIdx.x & 7;
    const int shift = (laneId() >> 3) << 3;
    const int* tk = A.slotTok + (size_t)a * kListSlotTok;
    const uint32_t limit = A.jOffset[a + 1];
    uint32_t at = A.jOffset[a];
    for (uint32_t t0 = 0; t0 < nT; t0 += 8)
    {
        const uint32_t t = t0 + o;
        const bool has = t < nT;
        const int v = has ? tk[t] : 0;
        const uint32_t sz = has ? (v >= 0 ? 1u : (uint32_t)A.nodeLeaves[~v]) : 0u;
        uint32_t incl = sz;   // inclusive scan over the group's 8 lanes
        for (int d = 1; d < 8; d <<= 1)
        {
            const uint32_t y = __shfl_up(incl, d, 8);
            if (o >= d) incl += y;
        }
        const uint32_t off = at + incl - sz;
        if (has && v >= 0)
        {
            if (off < limit) A.jList[off] = v;
            else             atomicAdd(&A.ovf[0], 1);
        }
        uint32_t fm = (uint32_t)(ballot64(has && v < 0) >> shift) & 0xFFu;
        while (fm)
        {
            const int l = __builtin_ctz(fm);
            fm &= fm - 1;
            const int node = ~__shfl(v, l, 8);
            const uint32_t o0 = __shfl(off, l, 8);
            const int n = A.nodeLeaves[node], f0 = A.fullStart[node];
            for (int k = o; k < n; k += 8)
            {
                const uint32_t p = o0 + (uint32_t)k;
                if (p < limit) A.jList[p] = A.fullOrder[f0 + k];
                else           atomicAdd(&A.ovf[0], 1);
            }
        }
        at += __shfl(incl, 7, 8);
    }
}

template<bool SYM, bool COUNT, bool PER>
__global__ __launch_bounds__(kListThreads) __attribute__((amdgpu_waves_per_eu(8))) void jlistKernel(ListArgs A)
{
    __shared__ int stk[kGroups][kStack];
    const int g = threadIdx.x >> 3, o = threadIdx.x & 7;
    if (A.ctl && A.ctl->skip) return;
    const int nL = A.ctl ? (int)A.ctl->nAL : A.nLeaves;
    const bool heavy = A.heavyList != nullptr;
    // hand-over threshold: the launch-size rule, and once a typical list length is known (the
    // previous pass's, or the previous build's of this kind) heavyFactor times its walk (about
    // a step per 4 entries), at least kHeavyMinSteps.  Where most lists are short (carried
    // steps, force lists) the few long ones then go to the block walk early instead of holding
    // the launch for thousands of steps; where most are long (first passes) the rule stands.
    int heavyCap = A.heavySteps * max(1, (nL + 32767) / 32768);
    const unsigned long long mean = A.ctl ? A.ctl->listMean : A.listMean;
    if (mean > 0ull && A.heavyFactor > 0)
        heavyCap = min(heavyCap, (int)max((unsigned long long)kHeavyMinSteps,
                                          min(mean * (unsigned long long)A.heavyFactor / 4ull, 1ull << 30)));
    for (int a = blockIdx.x * kGroups + g; a < nL; a += gridDim.x * kGroups)
    {
        if (!COUNT && heavy && A.heavyFlag[a]) continue;   // filled by the block walk
        if constexpr (!COUNT)
        {
            const uint32_t nT = A.slotN[a];
            if (nT != ~0u)
            {
                expandTokens(A, a, nT);
                continue;
            }
        }
        if (COUNT && o == 0) A.slotN[a] = ~0u;   // kept below once the walk ends within the slot
        const int iLeaf = A.leaves ? A.leaves[a] : a;
        const Reach<SYM, PER> inRange(A, iLeaf);
        const uint32_t base = COUNT ? 0u : A.jOffset[a];
        const int root = A.nodeFirst[A.leafSet[iLeaf]];
        if (COUNT && heavy && o == 0) A.heavyFlag[a] = 0;
        if (!inRange(A, root))
        {
            if (COUNT && o == 0) A.jCount[a] = 0;
            continue;
        }
        const int rc = A.nodeChild[root];
        if (rc < 0)
        {
            if (o == 0)
            {
                if (COUNT) A.jCount[a] = 1;
                else       A.jList[base] = ~rc;
            }
            continue;
        }
        bool capped;
        uint32_t nt = 0;
        const uint32_t count =
            groupWalk<COUNT, COUNT>(A, inRange, root, stk[g], base, COUNT ? 0u : A.jOffset[a + 1],
                                    (COUNT && heavy) ? heavyCap : A.nNodes, capped,
                                    COUNT ? A.slotTok + (size_t)a * kListSlotTok : nullptr, &nt);
        if (capped)
        {
            if (COUNT && heavy)
            {
                if (o == 0)
                {
                    A.heavyList[atomicAdd(A.heavyN, 1)] = a;
                    A.heavyFlag[a] = 1;
                }
                continue;
            }
            if (o == 0) atomicAdd(&A.ovf[1], 1);
        }
        if (COUNT && o == 0)
        {
            A.jCount[a] = count;
            if (nt <= (uint32_t)kListSlotTok) A.slotN[a] = nt;
        }
    }
}

// Block-wide exclusive scan of v[0, n) in place (n <= kTokCap); returns the total.
__device__ inline uint32_t blockScan(uint32_t* __restrict__ v, int n, uint32_t* __restrict__ red)
{
    constexpr int kPer = kTokCap / kListThreads;
    const int t = threadIdx.x;
    uint32_t loc[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k)
    {
        const int i = t * kPer + k;
        loc[k] = i < n ? v[i] : 0u;
        sum += loc[k];
    }
    const uint32_t incl = waveInclusiveScan(sum);
    if (laneId() == kWave - 1) red[t >> 6] = incl;
    __syncthreads();
    uint32_t before = incl - sum;
    for (int w = 0; w < (t >> 6); ++w) before += red[w];
    const uint32_t total = red[0] + red[1] + red[2] + red[3];
#pragma unroll
    for (int k = 0; k < kPer; ++k)
    {
        const int i = t * kPer + k;
        if (i < n) v[i] = before;
        before += loc[k];
    }
    __syncthreads();
    return total;
}

// Fill pass of a block walk: the list's tokens at their offsets (a leaf, or out(C) of node C by a
// group walk), after the list's start jOffset[a].
template<bool PER, class R>
__device__ inline void fillTokens(const ListArgs& A, const R& inRange, int a, const int* __restrict__ tk,
                                  const uint32_t* __restrict__ toff, int nTok, int* __restrict__ stk)
{
    const int g = threadIdx.x >> 3, o = threadIdx.x & 7;
    const uint32_t base = A.jOffset[a], limit = A.jOffset[a + 1];
    for (int t = g; t < nTok; t += kGroups)
    {
        const int v = tk[t];
        const uint32_t at = base + toff[t];
        if (v >= 0)
        {
            if (o == 0)
            {
                if (at < limit) A.jList[at] = v;
                else            atomicAdd(&A.ovf[0], 1);
            }
        }
        else
        {
            bool capped;
            (void)groupWalk<false>(A, inRange, ~v, stk, at, limit, A.nNodes, capped);
        }
    }
}

// The lists that count walks handed over (heavyList[0, *heavyN)), one block per list.  The count
// pass keeps each list's tokens and their offsets (tokPool, while it has room), so the fill pass
// walks each token once more instead of expanding and counting the list again.
template<bool SYM, bool COUNT, bool PER>
__global__ __launch_bounds__(kListThreads) void jlistHeavyKernel(ListArgs A)
{
    static_assert(kListThreads == 4 * kWave, "blockScan assumes four wavefronts");
    __shared__ int stk[kGroups][kStack];
    __shared__ int tok[2][kTokCap];       // >= 0: a leaf; < 0: ~C, the list out(C) of node C
    __shared__ uint32_t tmask[kTokCap];   // expansion: bit 16 expanded, bits 0-7 leaf, 8-15 internal children
    __shared__ uint32_t toff[kTokCap];    // token sizes, then offsets
    __shared__ uint32_t red[kListThreads / kWave];
    __shared__ int nNew;
    if (A.ctl && A.ctl->skip) return;
    const int g = threadIdx.x >> 3, o = threadIdx.x & 7;
    const int shift = (laneId() >> 3) << 3;
    const uint32_t below = (1u << o) - 1u;
    __shared__ uint32_t keepAt;
    const int nH = *A.heavyN;
    for (int k = blockIdx.x; k < nH; k += gridDim.x)
    {
        const int a = A.heavyList[k];
        const int iLeaf = A.leaves ? A.leaves[a] : a;
        const Reach<SYM, PER> inRange(A, iLeaf);
        int cur = 0, nTok = 1, ne = 1;   // ne: out(C) tokens made by the last round (bounds the next)
        if constexpr (!COUNT)
        {
            // the count pass kept this list's tokens and offsets: only the fill walks remain
            const uint2 at = A.tokAt[k];
            if (at.x != ~0u)
            {
                nTok = (int)at.y;
                for (int t = threadIdx.x; t < nTok; t += kListThreads)
                {
                    const int2 v = A.tokPool[at.x + (uint32_t)t];
                    tok[0][t] = v.x;
                    toff[t] = (uint32_t)v.y;
                }
                __syncthreads();
                fillTokens<PER>(A, inRange, a, tok[0], toff, nTok, stk[g]);
                __syncthreads();   // the block's LDS is reused by the next list
                continue;
            }
        }
        const int root = A.nodeFirst[A.leafSet[iLeaf]];   // internal and accepted: its walk was long
        if (threadIdx.x == 0) tok[0][0] = ~root;
        __syncthreads();
        while (ne > 0 && nTok + 7 * ne <= kTokCap)
        {
            for (int t = g; t < nTok; t += kGroups)
            {
                const int v = tok[cur][t];
                uint32_t m = 0;
                if (v < 0 && A.nodeLeaves[~v] > A.expandLeaves)
                {
                    const int c = A.nodeChild[~v] + o;
                    const bool acc = inRange(A, c);
                    const int cref = A.nodeChild[c];
                    const uint32_t mL = (uint32_t)(ballot64(acc && cref < 0) >> shift) & 0xFFu;
                    const uint32_t mI = (uint32_t)(ballot64(acc && cref >= 0) >> shift) & 0xFFu;
                    m = 0x10000u | mL | (mI << 8);
                }
                if (o == 0) tmask[t] = m;
            }
            __syncthreads();
            for (int t = threadIdx.x; t < nTok; t += kListThreads)
            {
                const uint32_t m = tmask[t];
                toff[t] = (m & 0x10000u) ? (uint32_t)(__popc(m & 0xFFu) + __popc((m >> 8) & 0xFFu)) : 1u;
            }
            if (threadIdx.x == 0) nNew = 0;
            __syncthreads();
            const int n2 = (int)blockScan(toff, nTok, red);
            for (int t = g; t < nTok; t += kGroups)
            {
                const int v = tok[cur][t];
                const uint32_t m = tmask[t], at = toff[t];
                if (!(m & 0x10000u))
                {
                    if (o == 0) tok[cur ^ 1][at] = v;
                }
                else
                {
                    const uint32_t mL = m & 0xFFu, mI = (m >> 8) & 0xFFu;
                    const int c = A.nodeChild[~v] + o;
                    if ((mL >> o) & 1u) tok[cur ^ 1][at + __popc(mL & below)] = ~A.nodeChild[c];
                    if ((mI >> o) & 1u) tok[cur ^ 1][at + __popc(mL) + __popc(mI >> (o + 1))] = ~c;
                    if (o == 0 && mI) atomicAdd(&nNew, (int)__popc(mI));
                }
            }
            __syncthreads();
            cur ^= 1;
            nTok = n2;
            ne = nNew;
            __syncthreads();
        }
        // token sizes: a leaf 1, out(C) by a group walk
        for (int t = g; t < nTok; t += kGroups)
        {
            const int v = tok[cur][t];
            uint32_t n = 1;
            if (v < 0)
            {
                bool capped;
                n = groupWalk<true>(A, inRange, ~v, stk[g], 0u, 0u, A.nNodes, capped);
                if (capped && o == 0) atomicAdd(&A.ovf[1], 1);
            }
            if (o == 0) toff[t] = n;
        }
        __syncthreads();
        const uint32_t total = blockScan(toff, nTok, red);
        if constexpr (COUNT)
        {
            // keep the tokens and their offsets for the fill pass while the pool has room
            if (threadIdx.x == 0)
            {
                A.jCount[a] = total;
                uint32_t at = ~0u;
                if (A.tokPool)
                {
                    const uint32_t s0 = (uint32_t)atomicAdd(&A.heavyN[1], nTok);
                    if (s0 + (uint32_t)nTok <= (uint32_t)kListTokPool) at = s0;
                }
                keepAt = at;
                if (A.tokAt) A.tokAt[k] = make_uint2(at, (uint32_t)nTok);
            }
            __syncthreads();
            const uint32_t at = keepAt;
            if (at != ~0u)
                for (int t = threadIdx.x; t < nTok; t += kListThreads)
                    A.tokPool[at + (uint32_t)t] = make_int2(tok[cur][t], (int)toff[t]);
        }
        else
        {
            fillTokens<PER>(A, inRange, a, tok[cur], toff, nTok, stk[g]);
        }
        __syncthreads();   // the block's LDS is reused by the next list
    }
}

// ---------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------
void hmaxLeaves(State& s, const int* leaves, int nLeaves, const uint8_t* state, const PassCtl* ctl)
{
    if (nLeaves <= 0) return;
    double* hm = s.hmaxLeaf.get<double>(s.nLeaves + 1);
    double* hr = state ? s.hRemLeaf.get<double>(s.nLeaves + 1) : nullptr;
    const int grid = ctl ? std::min(iceil(nLeaves, 4), 4096) : iceil(nLeaves, 4);
    hmaxLeafKernel<<<grid, 256, 0, s.stream>>>(s.h.as<double>(), s.layout.as<uint32_t>(), leaves, nLeaves, hm, state,
                                               hr, ctl);
    LAUNCH_CHECK();
}

void hmaxUpsweep(State& s)
{
    double* nh = s.nodeHmax.get<double>(s.nNodes + 1);
    nodeHmaxLeafKernel<<<iceil(s.nLeaves, 256), 256, 0, s.stream>>>(s.hmaxLeaf.as<double>(), s.leafNode.as<int>(),
                                                                      s.nLeaves, nh);
    LAUNCH_CHECK();
    for (int lvl = kMaxLevel - 1; lvl >= 0; --lvl)
    {
        if (s.nInternal == 0) break;
        nodeHmaxInternalKernel<<<iceil(s.nInternal, 256), 256, 0, s.stream>>>(
            s.intLevel.as<int>(), s.intNode.as<int>(), s.nInternal, lvl, s.nodeChild.as<int>(), nh);
        LAUNCH_CHECK();
    }
}

constexpr int kHeavyBlocks = 4096;

// The walks of one form (count, or fill at jOffset), then the block walks of the long lists.
template<bool SYM, bool COUNT>
static void launchList(const ListArgs& a, bool periodic, hipStream_t st)
{
    const int grid = a.ctl ? std::min(iceil(a.nLeaves, kGroups), 8192) : iceil(a.nLeaves, kGroups);
    if (periodic) jlistKernel<SYM, COUNT, true><<<grid, kListThreads, 0, st>>>(a);
    else          jlistKernel<SYM, COUNT, false><<<grid, kListThreads, 0, st>>>(a);
    LAUNCH_CHECK();
    const int hgrid = std::min(a.nLeaves, kHeavyBlocks);
    if (periodic) jlistHeavyKernel<SYM, COUNT, true><<<hgrid, kListThreads, 0, st>>>(a);
    else          jlistHeavyKernel<SYM, COUNT, false><<<hgrid, kListThreads, 0, st>>>(a);
    LAUNCH_CHECK();
}

static ListArgs listArgs(State& s, const int* leaves, int nLeaves, bool symmetric, bool tight)
{
    ListArgs a;
    std::memset(&a, 0, sizeof(a));
    a.leaves = leaves;
    a.nLeaves = nLeaves;
    a.leafNode = s.leafNode.as<int>();
    a.leafSet = s.leafSet.as<int>();
    a.nodeFirst = s.intFirst.as<int>() + 3 * s.nSets;
    a.nodeChild = s.nodeChild.as<int>();
    a.cx = s.nodeCx.as<double>(); a.cy = s.nodeCy.as<double>(); a.cz = s.nodeCz.as<double>();
    a.sx = s.nodeSx.as<double>(); a.sy = s.nodeSy.as<double>(); a.sz = s.nodeSz.as<double>();
    a.hmaxLeaf = s.hmaxLeaf.as<double>();
    a.hRemLeaf = (tight && !symmetric) ? s.hRemLeaf.as<double>() : nullptr;
    a.nodeHmax = symmetric ? s.nodeHmax.as<double>() : nullptr;
    a.boxes = s.box.as<DBox>();
    a.ovf = s.ovf.get<int>(4);
    a.nNodes = s.nNodes;
    a.nodeLeaves = s.nodeLeaves.as<int>();
    a.heavyN = s.heavyN.get<int>(4);
    a.heavyList = s.heavyList.get<int>((size_t)s.nLeaves + 1);
    a.heavyFlag = s.heavyFlag.get<uint8_t>((size_t)s.nLeaves + 1);
    a.tokPool = s.tokPool.get<int2>((size_t)kListTokPool);
    a.fullStart = s.fullStart.as<int>();
    a.fullOrder = s.fullStart.p ? s.fullOrder.as<int>() : nullptr;
    a.tokAt = s.tokAt.get<uint2>((size_t)s.nLeaves + 1);
    a.slotTok = s.slotTok.get<int>(((size_t)s.nLeaves + 1) * kListSlotTok);
    a.slotN = s.slotN.get<uint32_t>((size_t)s.nLeaves + 1);
    a.heavySteps = s.listHeavySteps > 0 ? s.listHeavySteps : kHeavySteps;
    a.expandLeaves = s.listExpandLeaves >= 0 ? s.listExpandLeaves : kExpandLeaves;
    a.heavyFactor = s.listHeavyFactor;
    a.listMean = (unsigned long long)(symmetric ? s.listMeanSym : (leaves ? 0.0 : s.listMeanGather));
    return a;
}

void launchGatherListsDevice(State& s, const int* leaves, int nCap, uint32_t* jCount, const uint32_t* jOffset,
                             int* jList, const PassCtl* ctl)
{
    if (nCap <= 0) return;
    // tight lists, except in a solve that counts list candidates: the reference's lists there
    ListArgs a = listArgs(s, leaves, nCap, false, !s.countList);
    a.ctl = ctl;
    a.jCount = jCount;
    a.jOffset = jOffset;
    a.jList = jList;
    const bool per = s.hBox.size() && s.hBox[0].periodic;
    if (jList)
    {
        launchList<false, false>(a, per, s.stream);
    }
    else
    {
        fillU32(reinterpret_cast<uint32_t*>(a.heavyN), 0u, 2, s.stream);
        launchList<false, true>(a, per, s.stream);
    }
}

long long buildLists(State& s, const int* leaves, int nLeaves, bool symmetric, bool tight)
{
    hipStream_t st = s.stream;
    uint32_t* jc = s.jCount.get<uint32_t>(nLeaves + 1);
    uint32_t* jo = s.jOffset.get<uint32_t>(nLeaves + 2);
    ListArgs a = listArgs(s, leaves, nLeaves, symmetric, tight);
    a.jCount = jc;
    const bool per = s.hBox.size() && s.hBox[0].periodic;
    fillU32(reinterpret_cast<uint32_t*>(a.heavyN), 0u, 2, st);
    if (nLeaves > 0)
    {
        if (symmetric) launchList<true, true>(a, per, st);
        else           launchList<false, true>(a, per, st);
    }
    unsigned long long* tot =
        s.counters.get<unsigned long long>((size_t)kCounterSlots * kCounterStride + 64) + (size_t)kCounterSlots * kCounterStride + 16;
    fillU32(reinterpret_cast<uint32_t*>(tot), 0u, 2, st);
    scanExclusive<uint32_t>(jc, jo, nLeaves, s.scanTmp, tot, st);
    unsigned long long total = 0;
    HIP_CHECK(hipMemcpyAsync(&total, tot, sizeof(total), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (total > 0xFFFFFFF0ull)
    {
        std::fprintf(stderr, "FATAL: neighbour-leaf lists hold %llu entries, beyond 32-bit offsets\n", total);
        std::abort();
    }
    a.jOffset = jo;
    a.jList = s.jList.get<int>(total + 1);
    if (nLeaves > 0)
    {
        if (symmetric) launchList<true, false>(a, per, st);
        else           launchList<false, false>(a, per, st);
    }
    s.listTotal = (int)total;
    if (!leaves && nLeaves > 0)
        (symmetric ? s.listMeanSym : s.listMeanGather) = (double)total / (double)nLeaves;
    return (long long)total;
}

} // namespace csph
