/*
 * density.hpp -- launch interface of the Newton-Raphson pass (nr.hip: density kernel,
 * closure, work items, compaction), of the gradient sweep (density.hip) and of the force
 * kernels (force.hip).
 */
#pragma once

#include "state.hpp"

namespace csph
{

// Device-resident state of the Newton-Raphson loop when its passes run without the host
// (loop.hip): every kernel of a pass reads its counts and its go/no-go from here instead of
// from its launch size.  Written only by the one-thread bookkeeping kernels of a pass
// (passBegin / listCheck / itemCount / passEnd) and by the work queue's atomics.
struct PassCtl
{
    uint32_t nAL;           // active leaves: positions [0, nAL) of the leaf list
    uint32_t nItems;        // work items of the current pass
    uint32_t iters;         // NR passes completed
    uint32_t passes;        // passes completed (NR and closure-only): next log record
    uint32_t halt;          // the loop is over: no active leaf, or the NR pass cap reached
    uint32_t overflow;      // the current pass needs more list entries than the list buffer holds
    uint32_t skip;          // the current pass does nothing (halt or overflow)
    uint32_t perParticle;   // the current pass uses per-particle items for sparse leaves
    uint32_t isNR;          // the current pass holds NR evaluations (else closures only)
    uint32_t queueHead;     // next position of the work queue
    uint32_t qBucket[32];   // work queue: [0] largest cost bucket, [1] entries at its head (heavy items, split)
    uint32_t nSplit;        // heavy items split over several wavefronts this pass (slots taken)
    uint32_t qExtra;        // queue-head entries beyond one per heavy item
    unsigned long long nIterating;      // particles in state 1 at the start of the pass
    unsigned long long listNeed;        // list entries of the current pass
    unsigned long long tBegin;          // wall clock at the start of the pass
    unsigned long long tStart, tEnd;    // density kernel span (wall clock) of the current pass
    unsigned long long listMean;        // list entries per active leaf of the last pass that built lists (0: none yet)
};

// One record per pass of a device-driven loop (u64 fields, wall-clock ticks for times).
struct PassLogDev
{
    unsigned long long nAL, nItems, nIterating, cand, pairs, closCand, closPairs;
    unsigned long long tBegin, tKStart, tKEnd, tEnd, perParticle, listTotal, isNR, maxListLog2, heavyItems, splitItems;
};

struct DensArgs
{
    const int*      leaves;     // active leaf ids (nullptr: leaf a is leaf a)
    int             nLeaves;
    const int*      itemPos;    // work item -> position in `leaves` (nullptr: item k is position k)
    const int*      itemPart;   // work item -> particle, or -1 for a whole-leaf item (nullptr: all -1)
    int             nItems;
    const uint32_t* jOffset;    // CSR of the lists, indexed by position in `leaves`
    const int*      jList;
    const uint32_t* layout;
    const int*      leafNode;
    const int*      leafSet;
    const double*   cx; const double* cy; const double* cz;
    const double*   sx; const double* sy; const double* sz;
    const double*   lbox;       // tight boxes of the leaves' particles (culling), lboxN leaves
    int             lboxN;
    const double*   hmaxLeaf;
    const double*   x; const double* y; const double* z;
    double*         h;
    double*         rho;
    double*         gradh;
    uint8_t*        state;      // 1 iterating, 2 converged with closure pending, 0 done
    const DBox*     boxes;
    double          pmass, tolh, hfact;
    const double*   pmassSet;   // ensemble: particle mass of each set (nullptr: pmass)
    int             closure;    // converged particles pass through state 2
    int             cull;       // whole-leaf items: stage only list entries some particle's support reaches
    int             units;      // mixed whole-leaf items: lanes over work units where that idles fewer lanes
    int             tileDepth;  // device-driven pass: j-slots per LDS tile of whole-leaf items (64 or 128)
    unsigned long long* counters;
    PassCtl*        ctl;        // device-driven pass: persistent waves pull items from a queue
    const int*      qOrder;     // ... heavy items qOrder[0, ctl->qBucket[1]) first, then every item
    const uint8_t*  qFlag;      //     not flagged heavy, in item order (nullptr: item order only)
    const uint32_t* qSub;       // ... per head entry: sub-item | count << 8 | slot << 16 (nullptr: none split)
    double*         splitPart;  // split items: per slot, sub-item and particle, the two partial sums
    uint32_t*       splitDone;  // ... per slot, sub-items finished (the last one resets it)
};

// Split work items: a heavy item with a list of more than kSplitLen entries is cut into at
// most kMaxSub sub-items (list ranges) run by different wavefronts; the last to finish adds
// the partial sums in sub-item order and ends the particles.  At most kMaxSplit per pass.
constexpr int kSplitLen = 2048;
constexpr int kMaxSub   = 32;
constexpr int kMaxSplit = 256;

// Device work counters (u64, summed over the counter slots).
enum Counter
{
    kCtrPairs = 0,        // accepted pairs of NR evaluations
    kCtrCand = 1,         // candidate pairs of NR evaluations
    kCtrIterating = 2,    // particles in state 1 after the pass
    kCtrRemaining = 3,    // particles in state != 0 after the pass
    kCtrItems = 4,        // work items of the next pass
    kCtrClosPairs = 5,    // accepted pairs of closure evaluations
    kCtrClosCand = 6,     // candidate pairs of closure evaluations
    kNumCounters = 7
};

// Work items of a gradient or force sweep, one wavefront each: items [0, nHeavy) are single
// particles (heavy[k] = (leaf, particle)) of the leaves whose lists are long (flag[leaf] != 0);
// item nHeavy + l is leaf l, skipped if flagged.  A whole-leaf wavefront runs the leaf's list
// once per chunk of up to 64 particles; a long list (a leaf whose largest h reaches far, most of
// it beyond the other particles' supports) then keeps one wavefront busy long after the rest of
// the grid has drained.  Its particles spread over wavefronts finish together with the others.
struct SweepItems
{
    int            nHeavy = 0;
    const int2*    heavy = nullptr;
    const uint8_t* flag = nullptr;

    // item k -> leaf a and its particle range [p0, p1); false for a skipped item
    __device__ bool resolve(int k, int nLeaves, const uint32_t* layout, int& a, uint32_t& p0, uint32_t& p1) const
    {
        if (k < nHeavy)
        {
            const int2 it = heavy[k];
            a = it.x;
            p0 = (uint32_t)it.y;
            p1 = p0 + 1u;
            return true;
        }
        a = k - nHeavy;
        if (a >= nLeaves || (flag && flag[a])) return false;
        p0 = layout[a];
        p1 = layout[a + 1];
        return true;
    }
    int count(int nLeaves) const { return nLeaves > 0 ? nHeavy + nLeaves : 0; }
};

struct GradArgs
{
    int             nLeaves;
    const uint32_t* jOffset;
    const int*      jList;
    const uint32_t* layout;
    const int*      leafSet;
    const double*   x; const double* y; const double* z;
    const double*   vx; const double* vy; const double* vz;
    const double*   ax; const double* ay; const double* az;
    const double*   h;
    double*         rho; double* gradh; double* divv; double* xi; double* ddivvdt;
    const DBox*     boxes;
    double          pmass;
    const int*      leafNode;                                 // per-particle culling of list entries
    const double*   lbox;                                     // ... against the leaves' tight boxes
    int             lboxN;
    const double*   cx; const double* cy; const double* cz;
    const double*   sx; const double* sy; const double* sz;
    SweepItems      items;                                    // one wavefront per item
};

struct ForceArgs
{
    int             nLeaves;
    const uint32_t* jOffset;
    const int*      jList;
    const uint32_t* layout;
    const int*      leafSet;
    const double*   x; const double* y; const double* z;
    const double*   vx; const double* vy; const double* vz;
    const double*   h;
    const double*   pro2; const double* cs; const double* alpha; const double* u;
    const double*   hsqinv; const double* hinv; const double* rhoh; const double* rho1;
    const double*   grkfac; const double* pres; const double* auterm; const double* divfac;
    double*         fx; double* fy; double* fz; double* f4; double* vsig; double* divv;
    const DBox*     boxes;
    double          pmass, beta;
    int             pdv, shock;
    unsigned long long* counters;
    const int*      leafNode;                                 // per-particle culling of list entries
    const double*   lbox;                                     // ... against the leaves' tight boxes
    int             lboxN;
    const double*   cx; const double* cy; const double* cz;
    const double*   sx; const double* sy; const double* sz;
    const double*   hmaxLeaf;                                 // reach of a j-leaf's own supports
    const float*    f32;      // mixed form: FP32 copies of hsqinv, hinv, grkfac, rhoh, rho1, pro2, cs,
    size_t          f32Stride; // alpha, pres, auterm (array k at f32 + k f32Stride)
    SweepItems      items;    // one wavefront per item
};

// mode 0: NR on the particles in state != 0 (FLAT_PARTICLE; state 2 = closure only);
// 1: NR on every particle of an active leaf (WARP_PER_LEAF); 2: closure of every particle,
// rho and d(rho)/d(h) at the current h, no update.  a.nItems work items.
void launchDensity(const DensArgs& a, int precision, int mode, bool periodic, hipStream_t st);
// Device-driven pass: both item forms are launched as persistent grids of `blocks` blocks; the
// one ctl->perParticle does not select exits at once.
void launchDensityPersistent(const DensArgs& a, int precision, bool periodic, int blocks, hipStream_t st);
int  densityPersistentBlocks(int precision, bool periodic);
void launchGradients(const GradArgs& a, bool periodic, bool mixed, hipStream_t st);
// Per-particle items of the leaves with more than `entries` list entries (0: none); reads the
// item count back (host-synchronising).
SweepItems buildSweepItems(State& s, int nLeaves, int entries, hipStream_t st);
void launchDeadGradients(int first, int n, double* divv, double* xi, double* ddivvdt, hipStream_t st);
// Per active leaf: flag (has a particle in state != 0) and its work-item count (1 whole-leaf
// item, or one item per remaining particle when at most `sparse` remain); counters 2-4.
// With ctl (device-driven pass), nLeaves is the capacity and ctl->nAL the count.
// List candidates of a pass: for each active leaf position a (n of them, or ctl->nAL), the
// particle counts of its list's leaves summed (entries [jOffset[a], jOffset[a] + jCount[a]), or
// to jOffset[a + 1] without jCount), times the leaf's particles in state 1 into out[0] and in
// state 2 into out[1]; without state, times all its particles into out[0].
void launchListCand(const PassCtl* ctl, int n, const int* leaves, const uint32_t* jOffset, const uint32_t* jCount,
                    const int* jList, const uint32_t* layout, const uint8_t* state, unsigned long long* out,
                    hipStream_t st);
void launchLeafFlags(const int* leaves, int nLeaves, const uint32_t* layout, const uint8_t* state, int sparse,
                     uint32_t* flag, uint32_t* itemCount, unsigned long long* counters, hipStream_t st,
                     const PassCtl* ctl = nullptr);
// Compaction: out[pos[a]] = leaves[a], outCount[pos[a]] = count[a] for flagged a.
void launchLeafScatter(const int* leaves, const uint32_t* count, int nLeaves, const uint32_t* flag,
                       const uint32_t* pos, int* out, uint32_t* outCount, hipStream_t st,
                       const PassCtl* ctl = nullptr);
// Work items of the active leaves from their counts' exclusive scan `off`.
void launchItemFill(const int* leaves, int nLeaves, const uint32_t* layout, const uint8_t* state, int sparse,
                    const uint32_t* off, int* itemPos, int* itemPart, hipStream_t st, const PassCtl* ctl = nullptr);
void launchInitActive(uint8_t* active, int nAlive, int n, double* rho, double* gradh, hipStream_t st);
void launchIota(int* v, int n, hipStream_t st);
void launchMaxH(const double* h, int n, unsigned long long* out, hipStream_t st);
int  counterSlots();
int  counterStride();

void launchForcePrep(const double* h, const double* gradh, const double* pro2, double* hsqinv, double* hinv,
                     double* rhoh, double* rho1, double* grkfac, double* pres, double* auterm, double* divfac, int n,
                     double pmass, double alphau, double hfact, hipStream_t st);
// rounds: the mixed form with each tile's accepted pairs shared evenly over the lanes
// (forceKernelMixedRounds, the default), else lanes that keep their own slots (forceKernelMixed)
void launchForce(const ForceArgs& a, bool periodic, bool disc, bool full, bool mixed, bool rounds, hipStream_t st);
// FP32 copies of the force factors for the mixed form (ForceArgs::f32 layout), n particles.
void launchForceF32(const ForceArgs& a, int n, float* dst, hipStream_t st);
void launchDeadForce(int first, int n, double* fx, double* fy, double* fz, double* f4, double* vsig, double* divv,
                     hipStream_t st);

} // namespace csph
