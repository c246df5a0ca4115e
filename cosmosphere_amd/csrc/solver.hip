/*
 * solver.hip -- host drivers: the density / smoothing-length solve and the force pass.
 *
 * solveDensH restates src/density_base.cu:720-1015: build the tree, iterate Newton-Raphson
 * on the shrinking active set (hmax of the active leaves -> their neighbour-leaf lists ->
 * one density kernel -> compaction), abort on an unconverged solve, check the periodic
 * support, optionally sweep the gradients at the converged h, and bump the token the force
 * pass checks.  computeForces restates src/force.cu:385-455.
 *
 * Differences of form, not of result: lists are built for the active leaves only (the
 * reference builds all leaves and reads only the active ones; tree.cuh:455-491), and the
 * active set is kept as a per-particle flag plus a compacted leaf list rather than
 * copy_if + sort + unique (density_base.cu:856-886).  The host reads the list total and
 * the counters (active leaves, particles still iterating, work items) once per pass.
 *
 * Closure.  With the closure on, a converged particle is evaluated once more at its final h
 * in the following pass (state 2), so the returned rho belongs to the returned h; passes that
 * hold only closures are not counted as iterations.
 */
#include "density.hpp"
#include "loop.hpp"
#include "solver.hpp"
#include "tree.hpp"
#include "xfer.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace csph
{

State& state()
{
    static State* s = new State();   // deliberately leaked, see state.hpp
    return *s;
}

// The stream is created on the first call that does device work, so that argument checks
// (which abort before any device call, as the reference's) need no device at all.
void ensureStream(State& s)
{
    if (!s.stream) HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
}

static SortScratch& sortScratch()
{
    static SortScratch* sc = new SortScratch();
    return *sc;
}

void sizeParticleArrays(State& s, int n)
{
    if (s.sizedFor == n) return;
    ensureStream(s);
    const size_t m = (size_t)std::max(n, 1) + 1;
    for (DevBuf* b : {&s.x, &s.y, &s.z, &s.h, &s.vx, &s.vy, &s.vz, &s.ax, &s.ay, &s.az, &s.rho, &s.gradh, &s.divv, &s.xi,
                      &s.ddivvdt, &s.pro2, &s.cs, &s.alpha, &s.u, &s.fx, &s.fy, &s.fz, &s.f4, &s.vsig, &s.divvF,
                      &s.hsqinv, &s.hinv, &s.rhoh, &s.rho1, &s.grkfac, &s.pres, &s.auterm, &s.divfac})
        b->get<double>(m);
    s.order.get<int>(m);
    s.particleLeaf.get<int>(m);
    s.active.get<uint8_t>(m);
    s.sizedFor = n;
    // the arrays may have moved: no order and no resident inputs survive a resize
    s.haveOrder = 0;
    resetInputTags(s);
}

struct Ev
{
    hipEvent_t e;
    Ev() { HIP_CHECK(hipEventCreate(&e)); }
    ~Ev() { (void)hipEventDestroy(e); }
    void rec(hipStream_t st) { HIP_CHECK(hipEventRecord(e, st)); }
};

static double secs(const Ev& a, const Ev& b)
{
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a.e, b.e));
    return ms * 1e-3;
}

static unsigned long long* zeroCounters(State& s)
{
    unsigned long long* c = s.counters.get<unsigned long long>((size_t)counterSlots() * counterStride() + 64);
    fillU32(reinterpret_cast<uint32_t*>(c), 0u, 2 * (size_t)counterSlots() * counterStride(), s.stream);
    return c;
}

// Sums every counter over the slots (after the stream is synchronised).
static void readCounters(State& s, unsigned long long out[kNumCounters])
{
    const int ns = counterSlots(), st = counterStride();
    std::vector<unsigned long long> h((size_t)ns * st);
    HIP_CHECK(hipMemcpyAsync(h.data(), s.counters.as<unsigned long long>(), h.size() * 8, hipMemcpyDeviceToHost,
                             s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
    for (int c = 0; c < kNumCounters; ++c) out[c] = 0;
    for (int k = 0; k < ns; ++k)
        for (int c = 0; c < kNumCounters; ++c) out[c] += h[(size_t)k * st + c];
}

static bool periodicOf(const State& s) { return !s.hBox.empty() && s.hBox[0].periodic; }

static DensArgs densArgs(State& s, const int* leaves, int nLeaves, double pmass, double tolh, double hfact,
                         unsigned long long* counters)
{
    DensArgs a;
    std::memset(&a, 0, sizeof(a));   // padding too: the device-driven loop keys its graph on the bytes
    a.leaves = leaves;
    a.nLeaves = nLeaves;
    a.jOffset = s.jOffset.as<uint32_t>();
    a.jList = s.jList.as<int>();
    a.layout = s.layout.as<uint32_t>();
    a.leafNode = s.leafNode.as<int>();
    a.leafSet = s.leafSet.as<int>();
    a.cx = s.nodeCx.as<double>(); a.cy = s.nodeCy.as<double>(); a.cz = s.nodeCz.as<double>();
    a.lbox = s.leafBox.as<double>(); a.lboxN = s.nLeaves;
    a.sx = s.nodeSx.as<double>(); a.sy = s.nodeSy.as<double>(); a.sz = s.nodeSz.as<double>();
    a.hmaxLeaf = s.hmaxLeaf.as<double>();
    a.x = s.x.as<double>(); a.y = s.y.as<double>(); a.z = s.z.as<double>();
    a.h = s.h.as<double>();
    a.rho = s.rho.as<double>();
    a.gradh = s.gradh.as<double>();
    a.state = s.active.as<uint8_t>();
    a.boxes = s.box.as<DBox>();
    a.pmass = pmass;
    a.tolh = tolh;
    a.hfact = hfact;
    a.pmassSet = s.usePmassSet ? s.pmassSet.as<double>() : nullptr;
    a.counters = counters;
    a.nItems = nLeaves;   // whole-leaf items unless the caller sets the item arrays
    a.cull = s.cull;
    a.units = s.units;
    a.tileDepth = s.tileDepth;
    return a;
}

Timings solveDensH(State& s, double pmass, int mode, bool withGradients, const double* periodicBox, double tolh,
                   double hfact)
{
    hipStream_t st = s.stream;
    Timings t{};
    if (!(hfact > 0.0)) hfact = kHfactDefault;
    s.hfact = hfact;
    s.lastPmass = pmass;
    const bool mixed = s.precision == kMixed;
    const bool closure = (s.closureOn < 0) ? mixed : (s.closureOn != 0);

    Ev e0, e1;
    e0.rec(st);
    int* ovf = s.ovf.get<int>(4);
    fillU32(reinterpret_cast<uint32_t*>(ovf), 0u, 4, st);
    s.t = Timings{};
    // inputs uploaded in the resident order (cosmo_upload after a solve of the same set) are
    // re-sorted incrementally; the rest are brought to one form first
    const bool incremental = prepareInputs(s, withGradients);
    buildTree(s, withGradients, periodicBox, sortScratch(), incremental);
    resetInputTags(s);
    t = s.t;
    if (s.treeOnly) return t;
    const int n = s.n, na = s.totalAlive, L = s.nLeaves;
    const bool per = periodicOf(s);

    uint8_t* state = s.active.as<uint8_t>();
    const uint32_t* layout = s.layout.as<uint32_t>();
    launchInitActive(state, na, n, s.rho.as<double>(), s.gradh.as<double>(), st);
    int* leaves = s.activeLeaves.get<int>(L + 1);
    int* leavesAlt = s.activeLeavesAlt.get<int>(L + 1);
    uint32_t* flag = s.leafFlag.get<uint32_t>(2 * (size_t)L + 4);
    uint32_t* pos = flag + L + 2;
    uint32_t* cnt = s.itemCount.get<uint32_t>(L + 1);
    uint32_t* cntAlt = s.itemCountAlt.get<uint32_t>(L + 1);
    uint32_t* itemOff = s.itemOff.get<uint32_t>(L + 2);
    int* itemPos = s.itemPos.get<int>((size_t)na + 1);
    int* itemPart = s.itemPart.get<int>((size_t)na + 1);
    // work-item rule: whole leaves only for WARP_PER_LEAF (every particle of an active leaf is
    // recomputed) and for the whole-leaf mapping; per-particle items for every remaining
    // particle with the flat mapping
    const int sparse = (mode == 1 || s.mapping == kMapWavePerLeaf) ? 0
                       : (s.mapping == kMapFlat ? (1 << 30) : std::max(s.sparse, 0));

    // Compaction of the active leaves after a pass: flags and item counts, scan, scatter; the
    // host reads the new leaf count and the counters (one synchronisation per pass).
    unsigned long long c[kNumCounters];
    int nAL = 0;
    auto compact = [&](int nIn, unsigned long long* ctr, Ev* done) {
        launchLeafFlags(leaves, nIn, layout, state, sparse, flag, cnt, ctr, st);
        scanExclusive<uint32_t>(flag, pos, nIn, s.scanTmp, nullptr, st);
        launchLeafScatter(leaves, cnt, nIn, flag, pos, leavesAlt, cntAlt, st);
        if (done) done->rec(st);
        unsigned int nNext = 0;
        if (nIn > 0) HIP_CHECK(hipMemcpyAsync(&nNext, pos + nIn, sizeof(nNext), hipMemcpyDeviceToHost, st));
        readCounters(s, c);   // synchronises
        std::swap(leaves, leavesAlt);   // the buffers keep their places: the same two each solve
        std::swap(cnt, cntAlt);
        nAL = (int)nNext;
    };

    // list candidates (count_list on, MODE 0 solves): accumulated on the device over the passes
    unsigned long long* listCand = nullptr;
    if (s.countList && mode == 0)
    {
        listCand = s.listCandCtr.get<unsigned long long>(2);
        fillU32(reinterpret_cast<uint32_t*>(listCand), 0u, 4, st);
    }

    s.passLog.clear();
    launchIota(leaves, L, st);
    compact(na > 0 ? L : 0, zeroCounters(s), nullptr);   // drops the leaves without live particles
    unsigned long long nIterating = c[kCtrIterating], nItems = c[kCtrItems];

    Ev a0, a1, a2, a3;
    bool devLoop = false;
    if (mode == 0 && s.loopMode == 1 && nAL > 0)
    {
        // device-driven passes (loop.hip): the same passes without a host round trip each
        DeviceLoopArgs dl;
        std::memset(&dl, 0, sizeof(dl));
        dl.leaves = leaves;
        dl.leavesAlt = leavesAlt;
        dl.cnt = cnt;
        dl.cntAlt = cntAlt;
        dl.flag = flag;
        dl.pos = pos;
        dl.itemOff = itemOff;
        dl.itemPos = itemPos;
        dl.itemPart = itemPart;
        dl.nAL = nAL;
        dl.nIterating = nIterating;
        dl.sparse = sparse;
        dl.periodic = per;
        s.hRemLeaf.get<double>((size_t)L + 1);   // sized before its pointer is taken
        dl.da = densArgs(s, leaves, L, pmass, tolh, hfact, nullptr);
        dl.da.hmaxLeaf = s.hRemLeaf.as<double>();
        dl.da.closure = closure && !withGradients;
        dl.listCand = listCand;
        nIterating = (unsigned long long)deviceLoop(s, dl, t);
        nAL = 0;
        devLoop = true;
    }
    while (nAL > 0)
    {
        const bool nrPass = nIterating > 0;   // else: closures only
        if (nrPass && t.iters >= s.maxIter) break;
        // per-particle items once few leaves remain (the chip is then underfilled and long
        // lists dominate); whole-leaf items with grouped lanes otherwise
        const bool perParticle = sparse > 0 && (s.mapping == kMapFlat || nAL <= s.tailLeaves);
        a0.rec(st);
        hmaxLeaves(s, leaves, nAL, mode == 0 ? state : nullptr);
        // tight: reach of the particles still to evaluate (the reference's lists when counting
        // list candidates)
        buildLists(s, leaves, nAL, false, mode == 0 && !s.countList);
        if (listCand)
            launchListCand(nullptr, nAL, leaves, s.jOffset.as<uint32_t>(), nullptr, s.jList.as<int>(), layout, state,
                           listCand, st);
        if (perParticle)
        {
            scanExclusive<uint32_t>(cnt, itemOff, nAL, s.scanTmp, nullptr, st);
            launchItemFill(leaves, nAL, layout, state, sparse, itemOff, itemPos, itemPart, st);
        }
        a1.rec(st);
        unsigned long long* ctr = zeroCounters(s);
        DensArgs da = densArgs(s, leaves, nAL, pmass, tolh, hfact, ctr);
        if (mode == 0) da.hmaxLeaf = s.hRemLeaf.as<double>();   // bounds every h evaluated in the pass
        da.itemPos = perParticle ? itemPos : nullptr;
        da.itemPart = perParticle ? itemPart : nullptr;
        da.nItems = perParticle ? (int)nItems : nAL;
        da.closure = closure && mode == 0 && !withGradients;   // the gradient sweep recomputes rho
        launchDensity(da, s.precision, mode, per, st);
        a2.rec(st);
        const int nALpass = nAL;
        compact(nAL, ctr, &a3);
        const double tj = secs(a0, a1), tk = secs(a1, a2), tc = secs(a2, a3);
        s.passLog.push_back(PassRecord{(double)nALpass, (double)da.nItems, (double)nIterating, (double)c[kCtrCand],
                                       (double)c[kCtrPairs], (double)c[kCtrClosCand], tk, (double)perParticle,
                                       tj, tc, (double)s.listTotal, 0.0, 0.0, 0.0});
        if (nrPass)
        {
            t.jbuild += tj;
            t.nrkern += tk;
            t.compact += tc;
            t.iters += 1;
        }
        else
        {
            t.closure += tj + tk + tc;
        }
        t.pairsAccepted += (double)c[kCtrPairs];
        t.pairsCandidate += (double)c[kCtrCand];
        t.closurePairs += (double)c[kCtrClosPairs];
        t.closureCand += (double)c[kCtrClosCand];
        if (nrPass) t.kernLaunches += 1;   // density launches timed in nrkern
        nIterating = c[kCtrIterating];
        nItems = c[kCtrItems];
    }
    (void)devLoop;
    const unsigned long long nActive = na > 0 ? nIterating : 0;
    if (listCand)
    {
        unsigned long long lc[2] = {0, 0};
        HIP_CHECK(hipMemcpyAsync(lc, listCand, sizeof(lc), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        t.listCand = (double)lc[0];
        t.closureListCand = (double)lc[1];
    }

    int hovf[2] = {0, 0};
    HIP_CHECK(hipMemcpyAsync(hovf, ovf, sizeof(hovf), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (nActive > 0 || hovf[0] > 0 || hovf[1] > 0)
    {
        std::fprintf(stderr, "FATAL: solveDensH: unconverged=%llu jlist_trunc=%d stack_drops=%d (iters=%d)\n",
                     nActive, hovf[0], hovf[1], t.iters);
        std::abort();
    }
    if (per && na > 0)
    {
        unsigned long long* m = s.counters.as<unsigned long long>() + (size_t)counterSlots() * counterStride() + 8;
        fillU32(reinterpret_cast<uint32_t*>(m), 0u, 2, st);
        launchMaxH(s.h.as<double>(), na, m, st);
        unsigned long long hm = 0;
        HIP_CHECK(hipMemcpyAsync(&hm, m, 8, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        double hmax;
        std::memcpy(&hmax, &hm, 8);
        double lmin = INFINITY;
        for (const DBox& b : s.hBox) lmin = std::min({lmin, b.l[0], b.l[1], b.l[2]});
        if (kRad * hmax >= 0.5 * lmin)
        {
            std::fprintf(stderr,
                         "FATAL: solveDensH: kernel support %gh = %g reaches half the periodic box "
                         "(smallest side %g); too few particles for this box\n",
                         kRad, kRad * hmax, lmin);
            std::abort();
        }
    }

    // WARP_PER_LEAF with the closure: one all-leaf pass at the final h (the FLAT solve folds
    // its closures into the iterations)
    if (closure && mode == 1 && !withGradients && L > 0 && na > 0)
    {
        Ev c0, c1;
        c0.rec(st);
        hmaxLeaves(s, nullptr, L);
        buildLists(s, nullptr, L, false);
        unsigned long long* ctr = zeroCounters(s);
        launchDensity(densArgs(s, nullptr, L, pmass, tolh, hfact, ctr), s.precision, 2, per, st);
        c1.rec(st);
        readCounters(s, c);
        t.closure += secs(c0, c1);
        t.closurePairs += (double)c[kCtrClosPairs];
        t.closureCand += (double)c[kCtrClosCand];
        s.listLeaves = L;
        s.listSymmetric = 0;
    }

    // the gradient sweep over the symmetric (force) lists: a superset of the gather lists in the
    // same order (a node within reach of the i-leaf's hmax is within reach of the larger of the
    // two hmax), so the sweep's pairs are the same; the force pass then reuses the lists
    const bool symLists = withGradients && s.gradSymLists;
    if (withGradients)
    {
        Ev g0, g1, g2;
        g0.rec(st);
        hmaxLeaves(s, nullptr, L);
        if (symLists) hmaxUpsweep(s);
        buildLists(s, nullptr, L, symLists);
        g1.rec(st);
        GradArgs g{};
        g.nLeaves = na > 0 ? L : 0;
        g.jOffset = s.jOffset.as<uint32_t>();
        g.jList = s.jList.as<int>();
        g.layout = s.layout.as<uint32_t>();
        g.leafSet = s.leafSet.as<int>();
        g.x = s.x.as<double>(); g.y = s.y.as<double>(); g.z = s.z.as<double>();
        g.vx = s.vx.as<double>(); g.vy = s.vy.as<double>(); g.vz = s.vz.as<double>();
        g.ax = s.ax.as<double>(); g.ay = s.ay.as<double>(); g.az = s.az.as<double>();
        g.h = s.h.as<double>();
        g.rho = s.rho.as<double>(); g.gradh = s.gradh.as<double>();
        g.divv = s.divv.as<double>(); g.xi = s.xi.as<double>(); g.ddivvdt = s.ddivvdt.as<double>();
        g.boxes = s.box.as<DBox>();
        g.pmass = pmass;
        g.leafNode = s.leafNode.as<int>();
        g.cx = s.nodeCx.as<double>(); g.cy = s.nodeCy.as<double>(); g.cz = s.nodeCz.as<double>();
        g.lbox = s.leafBox.as<double>(); g.lboxN = s.nLeaves;
        g.sx = s.nodeSx.as<double>(); g.sy = s.nodeSy.as<double>(); g.sz = s.nodeSz.as<double>();
        g.items = buildSweepItems(s, g.nLeaves, s.sweepSplit, st);
        launchGradients(g, per, s.precision == kMixed, st);
        launchDeadGradients(na, n, s.divv.as<double>(), s.xi.as<double>(), s.ddivvdt.as<double>(), st);
        g2.rec(st);
        HIP_CHECK(hipEventSynchronize(g2.e));
        t.gjbuild = secs(g0, g1);
        t.gradkern = secs(g1, g2);
        s.listLeaves = L;
        s.listSymmetric = symLists ? 1 : 0;
    }
    e1.rec(st);
    HIP_CHECK(hipEventSynchronize(e1.e));
    t.total = secs(e0, e1);
    s.t = t;
    ++s.token;
    if (symLists) s.jlistToken = s.token;   // the force pass of this solve reuses them
    return t;
}

ForceTimings computeForces(State& s, double pmass, double beta, double alphau, bool disc, bool pdv, bool shock)
{
    hipStream_t st = s.stream;
    ForceTimings ft{};
    const int n = s.n, L = s.nLeaves, na = s.totalAlive;
    const bool per = periodicOf(s);
    Ev e0, e1, e2, e3;
    e0.rec(st);
    if (s.jlistToken != s.token || !s.listSymmetric)
    {
        int* ovf = s.ovf.get<int>(4);
        fillU32(reinterpret_cast<uint32_t*>(ovf), 0u, 4, st);
        hmaxLeaves(s, nullptr, L);
        hmaxUpsweep(s);
        buildLists(s, nullptr, L, true);
        s.listLeaves = L;
        s.listSymmetric = 1;
        s.jlistToken = s.token;
        int hovf[2] = {0, 0};
        HIP_CHECK(hipMemcpyAsync(hovf, ovf, sizeof(hovf), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (hovf[0] > 0 || hovf[1] > 0)
            std::fprintf(stderr, "WARNING! buildForceJLeafList: jlist_trunc=%d stack_drops=%d\n", hovf[0], hovf[1]);
    }
    e1.rec(st);
    launchForcePrep(s.h.as<double>(), s.gradh.as<double>(), s.pro2.as<double>(), s.hsqinv.as<double>(),
                    s.hinv.as<double>(), s.rhoh.as<double>(), s.rho1.as<double>(), s.grkfac.as<double>(),
                    s.pres.as<double>(), s.auterm.as<double>(), s.divfac.as<double>(), n, pmass, alphau, s.hfact, st);
    e2.rec(st);
    unsigned long long* ctr = zeroCounters(s);
    ForceArgs a{};
    a.nLeaves = na > 0 ? L : 0;
    a.jOffset = s.jOffset.as<uint32_t>();
    a.jList = s.jList.as<int>();
    a.layout = s.layout.as<uint32_t>();
    a.leafSet = s.leafSet.as<int>();
    a.x = s.x.as<double>(); a.y = s.y.as<double>(); a.z = s.z.as<double>();
    a.vx = s.vx.as<double>(); a.vy = s.vy.as<double>(); a.vz = s.vz.as<double>();
    a.h = s.h.as<double>();
    a.pro2 = s.pro2.as<double>(); a.cs = s.cs.as<double>(); a.alpha = s.alpha.as<double>(); a.u = s.u.as<double>();
    a.hsqinv = s.hsqinv.as<double>(); a.hinv = s.hinv.as<double>(); a.rhoh = s.rhoh.as<double>();
    a.rho1 = s.rho1.as<double>(); a.grkfac = s.grkfac.as<double>(); a.pres = s.pres.as<double>();
    a.auterm = s.auterm.as<double>(); a.divfac = s.divfac.as<double>();
    a.fx = s.fx.as<double>(); a.fy = s.fy.as<double>(); a.fz = s.fz.as<double>();
    a.f4 = s.f4.as<double>(); a.vsig = s.vsig.as<double>(); a.divv = s.divvF.as<double>();
    a.boxes = s.box.as<DBox>();
    a.pmass = pmass;
    a.beta = beta;
    a.pdv = pdv ? 1 : 0;
    a.shock = shock ? 1 : 0;
    a.counters = ctr;
    a.leafNode = s.leafNode.as<int>();
    a.cx = s.nodeCx.as<double>(); a.cy = s.nodeCy.as<double>(); a.cz = s.nodeCz.as<double>();
    a.lbox = s.leafBox.as<double>(); a.lboxN = s.nLeaves;
    a.sx = s.nodeSx.as<double>(); a.sy = s.nodeSy.as<double>(); a.sz = s.nodeSz.as<double>();
    a.hmaxLeaf = s.hmaxLeaf.as<double>();
    const bool mixed = s.precision == kMixed;
    if (mixed)
    {
        a.f32Stride = (size_t)n + 1;
        float* f32 = s.forceF32.get<float>(10 * a.f32Stride);
        launchForceF32(a, n, f32, st);
        a.f32 = f32;
    }
    a.items = buildSweepItems(s, a.nLeaves, s.sweepSplit, st);
    launchForce(a, per, disc, pdv && shock, mixed, s.forceRounds != 0, st);
    launchDeadForce(na, n, s.fx.as<double>(), s.fy.as<double>(), s.fz.as<double>(), s.f4.as<double>(),
                    s.vsig.as<double>(), s.divvF.as<double>(), st);
    e3.rec(st);
    unsigned long long c[kNumCounters];
    readCounters(s, c);
    ft.upsweep = 0.0;
    ft.jbuild = secs(e0, e1);
    ft.prep = secs(e1, e2);
    ft.kernel = secs(e2, e3);
    ft.total = secs(e0, e3);
    ft.pairsAccepted = (double)c[0];
    ft.pairsCandidate = (double)c[1];
    if (s.countList && a.nLeaves > 0)
    {
        // after the timed span: the candidates the symmetric lists fix, every particle evaluated
        unsigned long long* lc = s.listCandCtr.get<unsigned long long>(2);
        fillU32(reinterpret_cast<uint32_t*>(lc), 0u, 4, st);
        launchListCand(nullptr, a.nLeaves, nullptr, a.jOffset, nullptr, a.jList, a.layout, nullptr, lc, st);
        unsigned long long h[2] = {0, 0};
        HIP_CHECK(hipMemcpyAsync(h, lc, sizeof(h), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        ft.listCand = (double)h[0];
    }
    s.ft = ft;
    return ft;
}

} // namespace csph
