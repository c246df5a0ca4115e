/*
 * loop.hip -- the Newton-Raphson loop driven from the device.
 *
 * The host-driven loop (solver.hip) reads the active-leaf count, the list total and the
 * counters back after every pass and sizes each launch from them.  Here every kernel of a
 * pass takes its counts from a small control block in device memory (PassCtl), so a pass is
 * a fixed sequence of launches: K passes are captured once in a HIP graph and the graph is
 * replayed until the loop is over, one host synchronisation per replay.  The density kernel
 * runs as a persistent grid whose wavefronts pull work items from the pass's queue: the
 * items with the longest lists first, so the pass does not end on a long item started last,
 * then the rest in the leaves' Hilbert order, so neighbouring items share j-particles in L2.
 *
 * One pass (the pass of src/density_base.cu:782-887):
 *   passBegin   halt (no active leaf, or the NR pass cap reached), NR or closure-only,
 *               item form; clears the queue, the heavy-item count and the kernel span
 *   counters    cleared (fill kernel: no memset node in the graph)
 *   hmax        hmax and hRem of the active leaves
 *   lists       count walk, scan, listCheck (capacity), fill walk: the tight gather lists
 *   items       scan of the per-leaf item counts, item fill, itemCount
 *   order       the heavy items (list length within 4x of the longest) to the queue's head
 *   density     both item forms as persistent grids; the one not selected exits at once
 *   compaction  leaf flags, scan, scatter, passEnd (log record, counts of the next pass),
 *               copy back
 *
 * A pass whose lists would overflow the list buffer stops at listCheck with nothing changed;
 * the host grows the buffer, captures again and replays.  The items of a pass, their lists
 * and every sum are those of the host-driven pass; only the order in which wavefronts take
 * items differs, and no item's result depends on it, so h, rho and d(rho)/d(h) are
 * bit-identical to the host-driven loop's.
 */
#include "density.hpp"
#include "loop.hpp"
#include "tile.hpp"
#include "tree.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace csph
{

constexpr int kCostBuckets = 32;

__device__ inline unsigned long long clockNow() { return (unsigned long long)wall_clock64(); }

// A pass after a halt or an overflow changes nothing in the control block: the host reads
// the overflowing pass's list need after the replay.
__global__ void passBeginKernel(PassCtl* __restrict__ C, int maxIter, int sparse, int tailLeaves, int flat)
{
    const int t = threadIdx.x;
    if (C->halt | C->overflow)
    {
        if (t == 0) C->skip = 1u;
        return;
    }
    if (t < kCostBuckets) C->qBucket[t] = 0u;
    if (t != 0) return;
    C->queueHead = 0u;
    C->nSplit = 0u;
    C->qExtra = 0u;
    C->tStart = ~0ull;
    C->tEnd = 0ull;
    C->tBegin = clockNow();
    C->listNeed = 0ull;
    uint32_t skip = 0u;
    const bool nr = C->nIterating > 0ull;
    if (C->nAL == 0u || (nr && (int)C->iters >= maxIter))
    {
        C->halt = 1u;
        skip = 1u;
    }
    C->isNR = nr ? 1u : 0u;
    C->perParticle = (sparse > 0 && (flat || (int)C->nAL <= tailLeaves)) ? 1u : 0u;
    C->skip = skip;
}

// Lists of the pass: entries needed against the buffer's capacity.
__global__ void listCheckKernel(PassCtl* __restrict__ C, const uint32_t* __restrict__ jOffset,
                                unsigned long long cap)
{
    if (C->skip) return;
    const unsigned long long tot = C->nAL ? (unsigned long long)jOffset[C->nAL] : 0ull;
    C->listNeed = tot;
    if (C->nAL) C->listMean = tot / C->nAL;
    if (tot > cap)
    {
        C->overflow = 1u;
        C->skip = 1u;
    }
}

__global__ void itemCountKernel(PassCtl* __restrict__ C, const uint32_t* __restrict__ itemOff)
{
    if (C->skip) return;
    C->nItems = C->perParticle ? (C->nAL ? itemOff[C->nAL] : 0u) : C->nAL;
}

// Cost bucket of an item: floor(log2) of its leaf's list length.
__device__ inline int costBucket(const PassCtl* C, const int* __restrict__ itemPos, const uint32_t* __restrict__ jCount,
                                 int k)
{
    const int a = C->perParticle ? itemPos[k] : k;
    const uint32_t len = jCount[a];
    return 31 - __clz((int)(len | 1u));
}

constexpr int kCostThreads = 256;
constexpr int kHeavySpan   = 2;   // heavy: a list at least 1/4 as long as the longest (in buckets)
constexpr int kHeavyMean   = 8;   // ... and at least 8 times the mean list length of the pass
constexpr int kFewItems    = 4096;   // a pass of at most this many items leaves most of the chip idle

// Largest cost bucket of the pass's items -> qBucket[0].
__global__ __launch_bounds__(kCostThreads) void costMaxKernel(PassCtl* __restrict__ C, const int* __restrict__ itemPos,
                                                               const uint32_t* __restrict__ jCount)
{
    if (C->skip) return;
    const int n = (int)C->nItems;
    int m = 0;
    for (int k = blockIdx.x * kCostThreads + threadIdx.x; k < n; k += gridDim.x * kCostThreads)
        m = max(m, costBucket(C, itemPos, jCount, k));
    m = waveReduceMax(m);
    if (laneId() == 0 && m > 0) atomicMax(&C->qBucket[0], (uint32_t)m);
}

// The work queue: the heavy items first (qOrder[0, qBucket[1]), in no particular order: there
// are few, and started last they would be the pass's tail), then every item in its own order
// (the leaves' Hilbert order, which keeps the j-particles that neighbouring items share in L2),
// skipping the heavy ones (qFlag).  A workload whose lists are all alike has no heavy item.
//
// A heavy item whose list exceeds splitLen entries (a whole leaf of at most 64 particles, or one
// particle) becomes up to kMaxSub sub-items at the head, while slots and head room last.
__global__ __launch_bounds__(kCostThreads) void costHeavyKernel(PassCtl* __restrict__ C,
                                                                 const int* __restrict__ itemPos,
                                                                 const uint32_t* __restrict__ jCount,
                                                                 int* __restrict__ qOrder, uint8_t* __restrict__ qFlag,
                                                                 uint32_t* __restrict__ qSub,
                                                                 const int* __restrict__ leaves,
                                                                 const uint32_t* __restrict__ layout,
                                                                 const int* __restrict__ itemPart, int splitLen)
{
    if (C->skip) return;
    const int n = (int)C->nItems;
    const int cut = (int)C->qBucket[0] - kHeavySpan;
    const unsigned long long mean = C->nAL ? C->listNeed / C->nAL : 0ull;   // entries per active leaf
    for (int k = blockIdx.x * kCostThreads + threadIdx.x; k < n; k += gridDim.x * kCostThreads)
    {
        const int a = C->perParticle ? itemPos[k] : k;
        // heavy: among the longest lists and, unless the pass has too few items to fill the chip,
        // well above the mean
        const bool heavy = costBucket(C, itemPos, jCount, k) >= cut &&
                           (n <= kFewItems || (unsigned long long)jCount[a] >= kHeavyMean * mean);
        qFlag[k] = heavy ? 1 : 0;
        if (!heavy) continue;
        uint32_t nSub = 1u, slot = 0u;
        const uint32_t len = jCount[a];
        if (qSub && splitLen > 0 && len > (uint32_t)splitLen)
        {
            const int l = leaves[a];
            const bool onePart = C->perParticle && itemPart[k] >= 0;   // a whole leaf: one batch only
            if (onePart || layout[l + 1] - layout[l] <= (uint32_t)kWave)
            {
                const uint32_t want = min((len + (uint32_t)splitLen - 1u) / (uint32_t)splitLen, (uint32_t)kMaxSub);
                const uint32_t sl = atomicAdd(&C->nSplit, 1u);
                const uint32_t ex = atomicAdd(&C->qExtra, want - 1u);
                if (sl < (uint32_t)kMaxSplit && ex + want - 1u <= (uint32_t)(kMaxSplit * kMaxSub))
                {
                    nSub = want;
                    slot = sl;
                }
            }
        }
        const uint32_t base = atomicAdd(&C->qBucket[1], nSub);
        for (uint32_t s = 0; s < nSub; ++s)
        {
            qOrder[base + s] = k;
            if (qSub) qSub[base + s] = s | (nSub << 8) | (slot << 16);
        }
    }
}

// End of a pass: one log record, the counts of the next pass.  One block of 256 threads,
// one counter slot each.
__global__ __launch_bounds__(256) void passEndKernel(PassCtl* __restrict__ C,
                                                     const unsigned long long* __restrict__ counters,
                                                     PassLogDev* __restrict__ log, int logCap,
                                                     const uint32_t* __restrict__ pos)
{
    __shared__ unsigned long long red[kNumCounters][4];
    if (C->skip) return;
    const int t = threadIdx.x, w = t >> 6;
    for (int c = 0; c < kNumCounters; ++c)
    {
        const unsigned long long v = waveReduceSum(counters[(size_t)t * kCounterStride + c]);
        if (laneId() == 0) red[c][w] = v;
    }
    __syncthreads();
    if (t != 0) return;
    unsigned long long sum[kNumCounters];
    for (int c = 0; c < kNumCounters; ++c) sum[c] = red[c][0] + red[c][1] + red[c][2] + red[c][3];
    const uint32_t nAL = C->nAL;
    const uint32_t next = nAL ? pos[nAL] : 0u;
    if ((int)C->passes < logCap)
    {
        PassLogDev& r = log[C->passes];
        r.nAL = nAL;
        r.nItems = C->nItems;
        r.nIterating = C->nIterating;
        r.cand = sum[kCtrCand];
        r.pairs = sum[kCtrPairs];
        r.closCand = sum[kCtrClosCand];
        r.closPairs = sum[kCtrClosPairs];
        r.tBegin = C->tBegin;
        r.tKStart = C->tStart;
        r.tKEnd = C->tEnd;
        r.tEnd = clockNow();
        r.perParticle = C->perParticle;
        r.listTotal = C->listNeed;
        r.isNR = C->isNR;
        r.maxListLog2 = C->qBucket[0];
        r.heavyItems = C->qBucket[1];
        r.splitItems = min(C->nSplit, (uint32_t)kMaxSplit);
    }
    C->passes += 1u;
    if (C->isNR) C->iters += 1u;
    C->nIterating = sum[kCtrIterating];
    C->nAL = next;
}

// The compacted leaves and their item counts back into the pass's arrays.
__global__ void copyBackKernel(const PassCtl* __restrict__ C, const int* __restrict__ leavesAlt,
                               const uint32_t* __restrict__ cntAlt, int* __restrict__ leaves,
                               uint32_t* __restrict__ cnt)
{
    if (C->skip) return;
    const int n = (int)C->nAL;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x)
    {
        leaves[a] = leavesAlt[a];
        cnt[a] = cntAlt[a];
    }
}

// ---------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------
namespace
{

struct Plan
{
    PassCtl*    ctl;
    PassLogDev* log;
    int         logCap;
    int*        leaves;
    int*        leavesAlt;
    uint32_t*   cnt;
    uint32_t*   cntAlt;
    uint32_t*   flag;
    uint32_t*   pos;
    uint32_t*   itemOff;
    int*        itemPos;
    int*        itemPart;
    int*        qOrder;
    uint8_t*    qFlag;
    uint32_t*   qSub;
    double*     splitPart;
    uint32_t*   splitDone;
    int         splitLen;
    uint32_t*   jc;
    uint32_t*   jo;
    int*        jList;
    unsigned long long listCap;
    unsigned long long* counters;
    unsigned long long* listCand;
    uint8_t*    state;
    int         L;
    int         maxIter, sparse, tailLeaves, flat;
    int         lpt;
    int         precision;
    bool        periodic;
    int         blocks;
    int         costBlocks;
    DensArgs    da;
};

void enqueuePass(State& s, const Plan& P)
{
    hipStream_t st = s.stream;
    const uint32_t* layout = s.layout.as<uint32_t>();
    passBeginKernel<<<1, 64, 0, st>>>(P.ctl, P.maxIter, P.sparse, P.tailLeaves, P.flat);
    LAUNCH_CHECK();
    fillU32(reinterpret_cast<uint32_t*>(P.counters), 0u, 2 * (size_t)kCounterSlots * kCounterStride, st);
    hmaxLeaves(s, P.leaves, P.L, P.state, P.ctl);
    launchGatherListsDevice(s, P.leaves, P.L, P.jc, nullptr, nullptr, P.ctl);
    scanExclusive<uint32_t>(P.jc, P.jo, P.L, s.scanTmp, nullptr, st, P.ctl);
    listCheckKernel<<<1, 1, 0, st>>>(P.ctl, P.jo, P.listCap);
    LAUNCH_CHECK();
    launchGatherListsDevice(s, P.leaves, P.L, P.jc, P.jo, P.jList, P.ctl);
    if (P.listCand) launchListCand(P.ctl, P.L, P.leaves, P.jo, P.jc, P.jList, layout, P.state, P.listCand, st);
    scanExclusive<uint32_t>(P.cnt, P.itemOff, P.L, s.scanTmp, nullptr, st, P.ctl);
    launchItemFill(P.leaves, P.L, layout, P.state, P.sparse, P.itemOff, P.itemPos, P.itemPart, st, P.ctl);
    itemCountKernel<<<1, 1, 0, st>>>(P.ctl, P.itemOff);
    LAUNCH_CHECK();
    if (P.lpt)
    {
        costMaxKernel<<<P.costBlocks, kCostThreads, 0, st>>>(P.ctl, P.itemPos, P.jc);
        costHeavyKernel<<<P.costBlocks, kCostThreads, 0, st>>>(P.ctl, P.itemPos, P.jc, P.qOrder, P.qFlag, P.qSub,
                                                               P.leaves, layout, P.itemPart, P.splitLen);
        LAUNCH_CHECK();
    }
    launchDensityPersistent(P.da, P.precision, P.periodic, P.blocks, st);
    launchLeafFlags(P.leaves, P.L, layout, P.state, P.sparse, P.flag, P.cnt, P.counters, st, P.ctl);
    scanExclusive<uint32_t>(P.flag, P.pos, P.L, s.scanTmp, nullptr, st, P.ctl);
    launchLeafScatter(P.leaves, P.cnt, P.L, P.flag, P.pos, P.leavesAlt, P.cntAlt, st, P.ctl);
    passEndKernel<<<1, 256, 0, st>>>(P.ctl, P.counters, P.log, P.logCap, P.pos);
    LAUNCH_CHECK();
    copyBackKernel<<<std::min(iceil(P.L, 256), 1024), 256, 0, st>>>(P.ctl, P.leavesAlt, P.cntAlt, P.leaves, P.cnt);
    LAUNCH_CHECK();
}

// Everything a captured launch depends on: a graph is replayed only while this is unchanged.
std::vector<unsigned long long> planKey(const State& s, const Plan& P, int K)
{
    std::vector<unsigned long long> k;
    auto add = [&](const void* p, size_t bytes) {
        const size_t n = (bytes + 7) / 8;
        const size_t at = k.size();
        k.resize(at + n, 0ull);
        std::memcpy(k.data() + at, p, bytes);
    };
    add(&P, sizeof(Plan));
    for (const DevBuf* b : {&s.layout, &s.leafNode, &s.leafSet, &s.intFirst, &s.nodeChild, &s.nodeCx, &s.nodeCy,
                            &s.nodeCz, &s.nodeSx, &s.nodeSy, &s.nodeSz, &s.hmaxLeaf, &s.hRemLeaf, &s.box, &s.h,
                            &s.ovf, &s.scanTmp, &s.nodeLeaves, &s.heavyN, &s.heavyList, &s.heavyFlag, &s.tokPool,
                            &s.tokAt, &s.slotTok, &s.slotN})
        k.push_back((unsigned long long)(uintptr_t)b->p);
    k.push_back((unsigned long long)s.nSets);
    k.push_back((unsigned long long)s.nNodes);
    k.push_back((unsigned long long)s.nLeaves);
    k.push_back((unsigned long long)K);
    return k;
}

} // namespace

struct LoopGraph
{
    hipGraph_t     graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned long long> key;
    unsigned long long lastUse = 0;

    void drop()
    {
        if (exec) HIP_CHECK(hipGraphExecDestroy(exec));
        if (graph) HIP_CHECK(hipGraphDestroy(graph));
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
};

// Captured graphs by key.  The particle arrays alternate between two buffers from one solve to
// the next (the permutation gathers into the alternates and swaps them in), so consecutive
// solves of one workload need two graphs; a few entries, least recently used replaced.
struct GraphCache
{
    static constexpr int kEntries = 4;
    LoopGraph e[kEntries];
    unsigned long long clock = 0;
    int captures = 0;

    LoopGraph* find(const std::vector<unsigned long long>& key)
    {
        for (LoopGraph& g : e)
            if (g.exec && g.key == key)
            {
                g.lastUse = ++clock;
                return &g;
            }
        return nullptr;
    }
    LoopGraph& victim()
    {
        LoopGraph* v = &e[0];
        for (LoopGraph& g : e)
        {
            if (!g.exec) return g;
            if (g.lastUse < v->lastUse) v = &g;
        }
        return *v;
    }
};

static GraphCache& graphCache()
{
    static GraphCache* g = new GraphCache();   // lives with the process, as the device state
    return *g;
}

static double wallClockHz()
{
    static double hz = 0.0;
    if (hz == 0.0)
    {
        int dev = 0, khz = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        hz = khz > 0 ? 1e3 * khz : 1e8;
    }
    return hz;
}

int deviceLoop(State& s, const DeviceLoopArgs& in, Timings& t)
{
    hipStream_t st = s.stream;
    const int L = s.nLeaves, na = s.totalAlive;
    Plan P;
    std::memset(&P, 0, sizeof(P));
    P.L = L;
    P.leaves = in.leaves;
    P.leavesAlt = in.leavesAlt;
    P.cnt = in.cnt;
    P.cntAlt = in.cntAlt;
    P.flag = in.flag;
    P.pos = in.pos;
    P.itemOff = in.itemOff;
    P.itemPos = in.itemPos;
    P.itemPart = in.itemPart;
    P.state = s.active.as<uint8_t>();
    P.listCand = in.listCand;
    P.maxIter = s.maxIter;
    P.sparse = in.sparse;
    P.tailLeaves = s.tailLeaves;
    P.flat = s.mapping == kMapFlat ? 1 : 0;
    P.lpt = s.lpt;
    P.precision = s.precision;
    P.periodic = in.periodic;
    P.logCap = s.maxIter + 4;
    P.ctl = s.passCtl.get<PassCtl>(1);
    P.log = s.passLogDev.get<PassLogDev>((size_t)P.logCap);
    // queue head: a heavy item per item at most, plus the sub-items of the split ones
    const size_t qCap = (size_t)std::max(na, L) + 1 + (size_t)kMaxSplit * kMaxSub;
    P.qOrder = s.qOrder.get<int>(qCap);
    P.qFlag = s.qFlag.get<uint8_t>(qCap);
    P.splitLen = s.splitLen >= 0 ? s.splitLen : kSplitLen;
    P.qSub = P.splitLen > 0 ? s.qSub.get<uint32_t>(qCap) : nullptr;
    P.splitPart = s.splitPart.get<double>((size_t)kMaxSplit * kMaxSub * kWave * 2);
    P.splitDone = s.splitDone.get<uint32_t>((size_t)kMaxSplit);
    fillU32(P.splitDone, 0u, (size_t)kMaxSplit, st);   // outside the graph: every pass leaves them at 0
    P.jc = s.jCount.get<uint32_t>((size_t)L + 1);
    P.jo = s.jOffset.get<uint32_t>((size_t)L + 2);
    P.counters = s.counters.get<unsigned long long>((size_t)kCounterSlots * kCounterStride + 64);
    s.scanTmp.get<char>(scanScratchBytes(L));
    s.ovf.get<int>(4);
    s.hmaxLeaf.get<double>((size_t)L + 1);
    s.hRemLeaf.get<double>((size_t)L + 1);
    s.heavyN.get<int>(4);   // the list walks take these in the captured passes: sized here
    s.heavyList.get<int>((size_t)L + 1);
    s.heavyFlag.get<uint8_t>((size_t)L + 1);
    s.tokPool.get<int2>((size_t)kListTokPool);
    s.tokAt.get<uint2>((size_t)L + 1);
    s.slotTok.get<int>(((size_t)L + 1) * kListSlotTok);
    s.slotN.get<uint32_t>((size_t)L + 1);
    if (s.loopBlocks <= 0 || s.loopBlocksPrecision != s.precision || s.loopBlocksPeriodic != (int)in.periodic)
    {
        s.loopBlocks = densityPersistentBlocks(s.precision, in.periodic);
        s.loopBlocksPrecision = s.precision;
        s.loopBlocksPeriodic = (int)in.periodic;
    }
    P.blocks = std::max(1, s.loopBlocks * std::max(s.loopBlockScale, 1) / 4);
    P.costBlocks = std::max(1, std::min(iceil(std::max(na, L), kCostThreads), 1024));

    // list capacity: what the buffer holds, at least 16 entries per leaf (or the verification
    // setting, to exercise the overflow path)
    unsigned long long cap = s.jList.cap / sizeof(int);
    cap = std::max<unsigned long long>(cap, 16ull * (unsigned long long)L + 1024ull);
    if (cap > 0xFFFFFFF0ull) cap = 0xFFFFFFF0ull;
    P.jList = s.jList.get<int>((size_t)cap);
    // the capacity the buffer really has (a grown buffer holds more than asked), so that the
    // next solve finds the same value and replays the same graph
    cap = std::min<unsigned long long>(s.jList.cap / sizeof(int), 0xFFFFFFF0ull);
    if (s.loopListCap > 0) cap = std::min<unsigned long long>(cap, (unsigned long long)s.loopListCap);
    P.listCap = cap;

    // byte copies: the graph key covers the Plan's bytes, padding included
    DensArgs da;
    std::memcpy(&da, &in.da, sizeof(da));
    // every buffer sized above is taken here, after its last resize (a pointer taken before a
    // resize would point at freed memory, or at nothing on a process's first solve)
    da.hmaxLeaf = s.hRemLeaf.as<double>();   // FLAT passes: bounds every h evaluated
    da.ctl = P.ctl;
    da.qOrder = P.lpt ? P.qOrder : nullptr;
    da.qFlag = P.lpt ? P.qFlag : nullptr;
    da.qSub = P.lpt ? P.qSub : nullptr;
    da.splitPart = P.splitPart;
    da.splitDone = P.splitDone;
    da.itemPos = P.itemPos;
    da.itemPart = P.itemPart;
    da.jOffset = P.jo;
    da.jList = P.jList;
    da.counters = P.counters;
    std::memcpy(&P.da, &da, sizeof(da));

    PassCtl h0;
    std::memset(&h0, 0, sizeof(h0));
    h0.nAL = (uint32_t)in.nAL;
    h0.nIterating = in.nIterating;
    // the first pass's typical list: the last all-leaf lists' (a carried step's h is near the
    // converged h they were built at; the symmetric lists of the gradient sweep, a superset,
    // when no gather lists were built); none for a first solve
    h0.listMean = (unsigned long long)(s.listMeanGather > 0.0 ? s.listMeanGather : s.listMeanSym);
    HIP_CHECK(hipMemcpyAsync(P.ctl, &h0, sizeof(h0), hipMemcpyHostToDevice, st));

    const int K = std::max(1, s.graphPasses);
    GraphCache& cache = graphCache();
    LoopGraph* G = nullptr;
    auto capture = [&]() {
        if (!G) G = &cache.victim();
        G->drop();
        HIP_CHECK(hipStreamSynchronize(st));
        HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        for (int k = 0; k < K; ++k) enqueuePass(s, P);
        HIP_CHECK(hipStreamEndCapture(st, &G->graph));
        HIP_CHECK(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
        G->key = planKey(s, P, K);
        G->lastUse = ++cache.clock;
        ++cache.captures;
    };
    if (s.useGraph)
    {
        G = cache.find(planKey(s, P, K));
        if (!G) capture();
    }

    PassCtl hc;
    int replays = 0, overflows = 0;
    // every replay without an overflow completes at least one pass, and a solve has at most
    // maxIter + 1 passes; every overflow grows the buffer past the need it reported: bound
    // both, so that a broken invariant aborts instead of spinning
    const int maxReplays = s.maxIter + 8, maxOverflows = 64;
    for (;;)
    {
        if (replays - overflows > maxReplays || overflows > maxOverflows)
        {
            std::fprintf(stderr, "FATAL: solveDensH: device-driven loop did not finish (%d replays, %d list overflows)\n",
                         replays, overflows);
            std::abort();
        }
        if (s.useGraph) HIP_CHECK(hipGraphLaunch(G->exec, st));
        else
            for (int k = 0; k < K; ++k) enqueuePass(s, P);
        ++replays;
        HIP_CHECK(hipMemcpyAsync(&hc, P.ctl, sizeof(hc), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (hc.overflow)
        {
            // grow the list buffer and redo the pass that stopped (it changed nothing)
            ++overflows;
            if (hc.listNeed <= P.listCap)
            {
                std::fprintf(stderr, "FATAL: solveDensH: list overflow reported for %llu entries within capacity %llu\n",
                             hc.listNeed, P.listCap);
                std::abort();
            }
            unsigned long long want = hc.listNeed + hc.listNeed / 4 + 1024ull;
            if (hc.listNeed > 0xFFFFFFF0ull)
            {
                std::fprintf(stderr, "FATAL: neighbour-leaf lists hold %llu entries, beyond 32-bit offsets\n",
                             hc.listNeed);
                std::abort();
            }
            want = std::min<unsigned long long>(want, 0xFFFFFFF0ull);
            HIP_CHECK(hipStreamSynchronize(st));
            s.jList.reserve((size_t)want * sizeof(int));
            P.jList = s.jList.as<int>();
            P.listCap = std::min<unsigned long long>(s.jList.cap / sizeof(int), 0xFFFFFFF0ull);
            P.da.jList = P.jList;
            hc.overflow = 0u;
            hc.skip = 0u;
            HIP_CHECK(hipMemcpyAsync(P.ctl, &hc, sizeof(hc), hipMemcpyHostToDevice, st));
            if (s.useGraph) capture();
            continue;
        }
        if (hc.halt) break;
    }

    // per-pass records -> timings, counters and the pass log
    const int nPass = std::min((int)hc.passes, P.logCap);
    std::vector<PassLogDev> rec((size_t)std::max(nPass, 1));
    if (nPass > 0)
        HIP_CHECK(hipMemcpyAsync(rec.data(), P.log, (size_t)nPass * sizeof(PassLogDev), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    const double hz = wallClockHz();
    for (int p = 0; p < nPass; ++p)
    {
        const PassLogDev& r = rec[p];
        const bool ran = r.tKStart != ~0ull && r.tKEnd >= r.tKStart && r.tKStart >= r.tBegin;
        const double tk = ran ? (double)(r.tKEnd - r.tKStart) / hz : 0.0;
        const double tj = ran ? (double)(r.tKStart - r.tBegin) / hz : 0.0;
        const double tc = ran ? (double)(r.tEnd - r.tKEnd) / hz : (double)(r.tEnd - r.tBegin) / hz;
        s.passLog.push_back(PassRecord{(double)r.nAL, (double)r.nItems, (double)r.nIterating, (double)r.cand,
                                       (double)r.pairs, (double)r.closCand, tk, (double)r.perParticle, tj, tc,
                                       (double)r.listTotal, (double)r.maxListLog2, (double)r.heavyItems,
                                       (double)r.splitItems});
        if (r.isNR)
        {
            t.jbuild += tj;
            t.nrkern += tk;
            t.compact += tc;
            t.iters += 1;
            t.kernLaunches += 1;
        }
        else
        {
            t.closure += tj + tk + tc;
        }
        t.pairsAccepted += (double)r.pairs;
        t.pairsCandidate += (double)r.cand;
        t.closurePairs += (double)r.closPairs;
        t.closureCand += (double)r.closCand;
        s.listTotal = (int)std::min<unsigned long long>(r.listTotal, 0x7FFFFFFFull);
    }
    t.graphReplays += replays;
    t.graphCaptures = cache.captures;
    return hc.nIterating > 0ull ? (int)std::min<unsigned long long>(hc.nIterating, 0x7FFFFFFFull) : 0;
}

} // namespace csph
