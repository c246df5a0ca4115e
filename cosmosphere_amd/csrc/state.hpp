/*
 * state.hpp -- device state that outlives a call (the MI355X counterpart of the
 * reference's GpuState, include/gpu_state.hpp:36-139).
 *
 * Layout in HBM: every particle field is its own FP64 array (SoA), in Hilbert order after
 * a solve, so a leaf's particles are one contiguous run of each array and a wavefront
 * reads them with unit stride.  Buffers grow and are never shrunk; nothing is freed
 * during a solve.  The state is created on first use and deliberately never destroyed
 * (the reference's leaked-global convention, src/gpu_state.cu:4-19).
 *
 * Particle sets: everything below is batched.  A single Phantom solve is a batch of one
 * set; the ensemble API packs many independent sets into the same arrays.  After a solve
 * the live particles of set b occupy [aliveFirst[b], aliveFirst[b] + nAlive[b]) of the
 * sorted arrays, sets in order, and every dead particle (h <= 0) sits after all of them.
 */
#pragma once

#include <vector>

#include "common.hpp"

namespace csph
{

// Grow-only device allocation.
struct DevBuf
{
    void*  p   = nullptr;
    size_t cap = 0;

    void reserve(size_t bytes)
    {
        if (bytes <= cap) return;
        if (p) HIP_CHECK(hipFree(p));
        size_t want = bytes + bytes / 8 + 256;
        HIP_CHECK(hipMalloc(&p, want));
        cap = want;
    }
    template<class T> T* as() const { return static_cast<T*>(p); }
    template<class T> T* get(size_t n) { reserve(n * sizeof(T)); return as<T>(); }
};

// Per-phase device time of one solve (seconds), and work counters.
struct Timings
{
    double bbox = 0, keys = 0, sort = 0, permute = 0, tree = 0, nodes = 0;
    double jbuild = 0, nrkern = 0, compact = 0, gjbuild = 0, gradkern = 0, closure = 0;
    double total = 0;
    int    iters = 0;
    int    nLeaves = 0;
    double pairsAccepted = 0, pairsCandidate = 0;   // summed over NR iterations
    double kernLaunches = 0;                         // NR pair-kernel launches
    double closurePairs = 0, closureCand = 0;
    // with count_list on: the j-particles of the list leaves of every evaluated particle (NR and
    // closure evaluations) over the reference's lists (that solve builds them untrimmed), the
    // candidate count the algorithm fixes whatever the lists are trimmed to or the kernel culls
    double listCand = 0, closureListCand = 0;
    double graphReplays = 0, graphCaptures = 0;      // device-driven loop
    double resortMovers = -1;                        // incremental re-sort: particles re-sorted from scratch (-1: full sort)
};

// One NR pass: active leaves, work items, particles iterating at its start, NR candidates and
// accepted pairs, closure candidates, kernel seconds, 1 if per-particle items were used.
struct PassRecord
{
    double leaves, items, iterating, candidates, pairs, closureCandidates, kernelSeconds, perParticle;
    double listSeconds;      // pass start to kernel start: hmax, lists, items
    double compactSeconds;   // kernel end to pass end: compaction
    double listEntries;      // entries of the pass's lists
    double maxListLog2;      // device-driven passes: floor(log2) of the longest list (with the cost order on)
    double heavyItems;       // device-driven passes: queue-head entries (heavy items and their sub-items)
    double splitItems;       // device-driven passes: heavy items split over several wavefronts
};

struct ForceTimings
{
    double upsweep = 0, jbuild = 0, prep = 0, kernel = 0, total = 0;
    double pairsAccepted = 0, pairsCandidate = 0;
    double listCand = 0;   // with count_list on: the j-particles of every particle's list leaves
};

enum Precision : int { kPlainFP64 = 0, kMixed = 1 };
enum InputTag : uint8_t { kTagCaller = 0, kTagAligned = 1 };
enum Mapping : int { kMapAuto = 0, kMapWavePerLeaf = 1, kMapFlat = 2 };

struct State
{
    hipStream_t stream = nullptr;

    // ---- particle arrays (length capacity >= n) ----
    int n = 0;          // particles of the last solve (sum over sets)
    int sizedFor = 0;   // cosmo_arrays_init count, 0 if never
    DevBuf x, y, z, h, vx, vy, vz, ax, ay, az;
    DevBuf rho, gradh, divv, xi, ddivvdt;
    DevBuf pro2, cs, alpha, u;
    DevBuf fx, fy, fz, f4, vsig, divvF;
    DevBuf order;       // sorted index -> caller index (int32)
    DevBuf orderAlt;    // the next order while the incremental re-sort composes it
    int    haveOrder = 0;
    int    orderSets = 0;   // sets of the solve that built `order`
    DevBuf stage;       // transfer staging (up to 6 components of n doubles)
    DevBuf stash;       // pristine copy of POS|HSML|VEL|ACCEL for repeated benchmark solves
    int    stashN = 0;
    // input fields x y z h vx vy vz ax ay az: kTagCaller (caller's order) or kTagAligned
    // (the order `order` describes: entry i is caller particle order[i])
    uint8_t inTag[10] = {};
    int    resident = 1;     // cosmo_upload of input slots lands in the resident order when it can
    int    xferThreads = 8;  // host threads of the page-locked transfer staging
    int    ringNext = 0;     // next staging chunk
    unsigned ringBusy = 0;   // staging chunks with a DMA possibly in flight
    // incremental re-sort work arrays
    DevBuf rsStay, rsPos, rsKey, rsCal, rsIdx, rsSortKey, rsSortVal;

    // ---- sets ----
    int nSets = 0;
    std::vector<int> hSetFirst, hSetN;        // caller-side ranges, input order
    std::vector<int> hAliveFirst, hAlive;     // sorted ranges of live particles
    std::vector<int> hLeafFirst, hLeafN, hIntFirst, hIntN, hNodeFirst;
    std::vector<DBox> hBox;
    DevBuf setFirst, setN, aliveFirst, alive, leafFirst, intFirst, nodeFirst, box;
    DevBuf pmassSet;          // ensemble: particle mass of each set
    int usePmassSet = 0;
    int totalAlive = 0;
    int sweepHeavyItems = 0;   // per-particle items of the last gradient or force sweep

    // ---- sort / tree scratch ----
    DevBuf keys, keysAlt, vals, valsAlt, hist, sortSetKey;
    DevBuf partials, boxDesc;
    DevBuf permAlt[10];                       // permute targets, swapped with the fields
    DevBuf tmpA, tmpB;                        // particle-sized scratch for the permute
    DevBuf winCount, winOffset;               // internal-node emission
    DevBuf preStart, preLevel, preSet;        // internal nodes in key (pre-)order
    DevBuf bfsKey, bfsKeyAlt, bfsIdx, bfsIdxAlt;
    DevBuf intStart, intLevel, intSet;        // internal nodes in (set, level, key) order
    DevBuf scanTmp;

    // ---- tree ----
    int nLeaves = 0, nInternal = 0, nNodes = 0;
    DevBuf leafStart;    // u64 start key of each leaf (the csTree without its end)
    DevBuf leafLevel;    // int
    DevBuf leafSet;      // int
    DevBuf layout;       // u32, nLeaves + 1: leaf l owns sorted particles [layout[l], layout[l+1])
    DevBuf leafNode;     // int: node id of each leaf
    DevBuf particleLeaf; // int
    DevBuf nodeChild;    // int: first child node (internal) or ~leaf (leaf)
    DevBuf nodeCx, nodeCy, nodeCz, nodeSx, nodeSy, nodeSz;   // FP64 geometry (SoA)
    DevBuf nodeHmax;     // force upsweep
    DevBuf intNode;      // node id of each internal node (BFS order)
    DevBuf nodeLeaves;   // int: leaves under each internal node
    DevBuf fullStart, fullOrder;   // the full-walk order: per node its start, per position the leaf
    DevBuf leafBox;      // FP64 tight boxes of the leaves' particles: 3 centres, 3 half-sizes, nLeaves each
    std::vector<int> hIntLevelFirst;          // level ranges of internal nodes (all sets)

    // ---- NR and lists ----
    DevBuf hmaxLeaf;
    DevBuf hRemLeaf;     // max h of the particles still to evaluate (tight NR lists)
    DevBuf activeLeaves, activeLeavesAlt, leafFlag;
    DevBuf active;       // u8 per particle: 1 iterating, 2 closure pending, 0 done
    DevBuf itemCount, itemCountAlt, itemOff, itemPos, itemPart;   // NR work items
    DevBuf jCount, jOffset, jList;
    DevBuf heavyList, heavyFlag, heavyN;   // lists walked by a whole block (long walks)
    DevBuf slotTok, slotN;                 // light lists: the count walk's tokens (kListSlotTok per position) and their number
    DevBuf tokPool, tokAt;                 // ... their tokens, kept by the count pass for the fill
    DevBuf qSub, splitPart, splitDone;     // device-driven loop: heavy items split over wavefronts
    DevBuf sweepFlag, sweepCount, sweepOff, sweepItems;   // gradient / force sweeps: per-particle items of long lists
    int    listLeaves = 0;    // number of leaves the CSR lists cover (all-leaf lists)
    int    listTotal  = 0;
    int    listSymmetric = 0; // 1 when the lists are the force (symmetric) lists
    DevBuf counters;          // u64 device counters (pairs, candidates, unconverged)
    DevBuf pinnedCounters;    // host-visible copy target
    DevBuf ovf;               // int[2] list truncation / stack drop (never expected)
    DevBuf listCandCtr;       // u64[2] list candidates of NR and closure evaluations (count_list on)
    DevBuf passCtl;           // device-driven loop: control block, pass records, work queue order
    DevBuf passLogDev;
    DevBuf qOrder, qFlag;

    // ---- force prep ----
    DevBuf hsqinv, hinv, rhoh, rho1, grkfac, pres, auterm, divfac;
    DevBuf forceF32;          // mixed form: FP32 copies of the factors the pair terms read

    // ---- bookkeeping ----
    double   hfact = 0.0;
    double   lastPmass = 0.0;
    uint64_t token = 0, jlistToken = 0;
    int      maxIter = 100;
    int      precision = kPlainFP64;
    int      mapping = kMapAuto;   // auto: whole-leaf items, per-particle below `sparse`
    int      sparse = 16;          // per-particle items when at most this many remain in a leaf
    int      tailLeaves = 2048;    // ... and at most this many leaves are active
    int      bucket = kBucket;     // leaf bucket size (64: the reference's; others are speed options)
    int      loopMode = 1;         // NR passes: 0 host-driven, 1 device-driven (FLAT_PARTICLE)
    int      useGraph = 1;         // device-driven passes replayed as a captured HIP graph
    int      graphPasses = 4;      // passes per graph replay (per host round trip)
    int      lpt = 1;              // work queue: the heaviest items first
    int      cull = 1;             // NR whole-leaf items: per-particle culling of list entries
    int      units = 1;            // mixed NR whole-leaf items: lanes over work units where that idles fewer lanes
    int      tileDepth = 128;      // device-driven NR passes: j-slots per LDS tile of whole-leaf items (64 or 128)
    int      loopBlockScale = 4;   // persistent density grid: resident blocks x scale / 4
    int      loopBlocks = 0, loopBlocksPrecision = -1, loopBlocksPeriodic = -1;
    long long loopListCap = 0;     // verification: starting list capacity of the device-driven loop (0: auto)
    int listHeavySteps = 0;        // verification: walk steps before a list goes to the block walk (0: default)
    int gradSymLists = 1;          // gradient sweep over the symmetric lists, which the force pass reuses
    int listHeavyFactor = 8;       // block walk for lists past this many times a typical walk (0: launch-size rule only)
    double listMeanGather = 0.0;   // entries per leaf of the last all-leaf gather / symmetric list build
    double listMeanSym = 0.0;
    int listExpandLeaves = -1;     // verification: block walk expands nodes above this many leaves (-1: default)
    int splitLen = -1;             // device-driven loop: heavy items with longer lists are split (-1: default, 0: off)
    int forceRounds = 1;           // mixed force pass: 1 shares each tile's accepted pairs over the lanes, 0 lanes keep their own slots
    int countList = 0;             // count the list candidates of each pass (a separate kernel; for reports), over the reference's untrimmed lists
    int sweepSplit = 512;          // gradient / force sweeps: leaves with more list entries go one particle per wavefront (0: off)
    std::vector<PassRecord> passLog;   // one record per NR pass of the last solve
    int      closureOn = -1;   // -1: on for mixed, off for plain
    int      treeOnly = 0;     // verification: stop after the tree build
    double   uploadSeconds = 0.0, downloadSeconds = 0.0;   // host-timed transfers since the last stats line
    double   uploadTotal = 0.0, downloadTotal = 0.0;       // ... and since the process started
    Timings      t;
    ForceTimings ft;

    bool readyForForce(int nn) const { return n == nn && token != 0; }
};

State& state();

} // namespace csph
