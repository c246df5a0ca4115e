/*
 * abi.hip -- the C ABI: the ten entry points Phantom links against (declared in
 * include/cosmosphere.h, same names, arguments and abort-on-misuse behaviour as the
 * reference's src/dens_c_api.cu, src/force_c_api.cu, src/arrays_c_api.cu and
 * src/pin_c_api.cu), plus the extensions used by the ensemble solve, the tests and the
 * benchmark (cosmo_ext_*, cosmo_batch_*).
 */
#include "../../include/cosmosphere.h"
#include "density.hpp"
#include "solver.hpp"
#include "tree.hpp"
#include "tune.hpp"
#include "xfer.hpp"

#include <sys/resource.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>

using namespace csph;

namespace
{

[[noreturn]] void fatal(const char* msg)
{
    std::fprintf(stderr, "FATAL: %s\n", msg);
    std::abort();
}

void requireSized(State& s, int n, const char* what)
{
    if (s.sizedFor != n)
    {
        std::fprintf(stderr,
                     "FATAL: %s called with n=%d but the device arrays are sized for %d "
                     "(cosmo_arrays_init not called, or called with a different count)\n",
                     what, n, s.sizedFor);
        std::abort();
    }
}

// Transfers check their arguments before any device work and abort with the reference's
// messages (Phantom's logs are read against them): size, slot number, ordering.
[[noreturn]] void fatalf(const char* fmt, ...)
{
    std::fputs("FATAL: ", stderr);
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fputc('\n', stderr);
    std::abort();
}

const SlotDesc& slotFor(const State& s, int slot, int n, const char* what, bool needOrder)
{
    if (s.sizedFor != n) fatalf("%s(slot=%d) with n=%d but the device arrays are sized for %d", what, slot, n, s.sizedFor);
    const SlotDesc* d = slotDesc(slot);
    if (!d) fatalf("%s called for slot %d, which has no arrays", what, slot);
    if (needOrder && !(s.haveOrder && s.n == n))
        fatalf("%s(slot=%d) with no ordering: a density solve must have run first", what, slot);
    return *d;
}

using Clock = std::chrono::steady_clock;

double secondsSince(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// Input slots land in the resident order when the last solve was this set; everything else
// in the caller's order (it is then sorted by the solve).
constexpr int kSlotHsml = 2;

void upload(int slot, const double* host, int n, bool sortedCall)
{
    if (n <= 0) return;
    State& s = state();
    const SlotDesc& d = slotFor(s, slot, n, sortedCall ? "cosmo_upload_sorted" : "cosmo_upload", sortedCall);
    ensureStream(s);
    const auto t0 = Clock::now();
    const bool resident = !sortedCall && d.firstInput >= 0 && residentEligible(s, n);
    transferIn(s, d, host, n, sortedCall || resident);
    s.uploadTotal += secondsSince(t0);
    // a sorted upload is read as caller order by a later solve that finds it not uploaded
    // again (the reference's convention); a resident one as the current order
    for (int c = 0; d.firstInput >= 0 && c < d.ncomp; ++c) s.inTag[d.firstInput + c] = resident ? kTagAligned : kTagCaller;
    // new smoothing lengths: a force pass builds its lists again (their reach is hmax)
    if (slot == kSlotHsml) s.jlistToken = 0;
    s.uploadSeconds += secondsSince(t0);
}

// COSMO_MEM: device memory in use after the first solve (hipMemGetInfo) and the process's
// peak resident set (getrusage, KiB), once per process.
void memoryReport(int n)
{
    static bool reported = false;
    if (reported || !std::getenv("COSMO_MEM")) return;
    reported = true;
    size_t freeB = 0, totalB = 0;
    HIP_CHECK(hipMemGetInfo(&freeB, &totalB));
    rusage ru{};
    getrusage(RUSAGE_SELF, &ru);
    const double used = double(totalB - freeB);
    std::fprintf(stderr, "COSMO_MEM n=%d device_used=%.2f GB of %.2f GB (%.0f bytes/particle) host_peak=%.2f GB\n",
                 n, used / 1e9, totalB / 1e9, used / std::max(n, 1), double(ru.ru_maxrss) / 1e6);
}

void setSingle(State& s, int n)
{
    s.n = n;
    s.nSets = 1;
    s.hSetFirst.assign(1, 0);
    s.hSetN.assign(1, n);
}

bool statsEnabled()
{
    static const bool on = std::getenv("COSMO_DENS_STATS") != nullptr;
    return on;
}

void applyEnvPrecision(State& s)
{
    static int env = -2;
    if (env == -2)
    {
        const char* p = std::getenv("COSMO_PRECISION");
        env = (p && std::strcmp(p, "mixed") == 0) ? kMixed : (p && std::strcmp(p, "fp64") == 0) ? kPlainFP64 : -1;
        if (env >= 0) s.precision = env;
    }
}

} // namespace

// =========================================================================================
// The ten reference entry points
// =========================================================================================
extern "C" void densityiterate_gpu_c(int n, double pmass, int periodic, const double* box, double tolh, double hfact)
{
    const auto tEntry = Clock::now();
    State& s = state();
    requireSized(s, n, "densityiterate_gpu_c");
    ensureStream(s);
    applyEnvPrecision(s);
    setSingle(s, n);
    const double vecin = secondsSince(tEntry);   // host work before the solve
    const auto t0 = Clock::now();
    const Timings t = solveDensH(s, pmass, 0, true, periodic ? box : nullptr, tolh, hfact);
    const double solve = secondsSince(t0);
    memoryReport(n);
    if (statsEnabled())
    {
        // device phases from the solve's events; upload / download are the host-timed
        // transfers since the previous stats line (they sit outside the solve)
        const double gpu = t.bbox + t.keys + t.sort + t.permute + t.tree + t.nodes + t.jbuild + t.nrkern + t.compact +
                           t.gjbuild + t.gradkern;
        std::fprintf(stderr,
                     "COSMO_STATS n=%d leaves=%d iters=%d | vecin=%.2f upload=%.2f bbox=%.2f "
                     "keysort=%.2f tree=%.2f nodes=%.2f jbuild=%.2f nrkern=%.2f gjbuild=%.2f "
                     "gradkern=%.2f download=%.2f | gpusum=%.2f solve=%.2f unaccounted=%.2f total=%.2f\n",
                     n, t.nLeaves, t.iters, 1e3 * vecin, 1e3 * s.uploadSeconds, 1e3 * t.bbox,
                     1e3 * (t.keys + t.sort + t.permute), 1e3 * t.tree, 1e3 * t.nodes, 1e3 * t.jbuild,
                     1e3 * (t.nrkern + t.compact), 1e3 * t.gjbuild, 1e3 * t.gradkern, 1e3 * s.downloadSeconds,
                     1e3 * gpu, 1e3 * solve, 1e3 * (solve - gpu), 1e3 * secondsSince(tEntry));
        s.uploadSeconds = 0.0;
        s.downloadSeconds = 0.0;
    }
}

extern "C" double cosmo_kernel_radius(void) { return kRad; }

extern "C" void force_gpu_c(int n, double pmass, double beta, double alphau, int disc_viscosity, int pdv_heating,
                            int shock_heating)
{
    const auto t0 = Clock::now();
    State& s = state();
    requireSized(s, n, "force_gpu_c");
    if (!s.readyForForce(n))
    {
        std::fprintf(stderr,
                     "FATAL: force_gpu_c called with no GPU density solve for this particle set "
                     "(n=%d state.ngas=%d token=%llu)\n",
                     n, s.n, (unsigned long long)s.token);
        std::abort();
    }
    ensureStream(s);
    const ForceTimings ft =
        computeForces(s, pmass, beta, alphau, disc_viscosity != 0, pdv_heating > 0, shock_heating > 0);
    if (statsEnabled())
    {
        // upload / download: host-timed transfers since the previous stats line (the THERMO
        // upload before this pass, the downloads after the density solve)
        const double wall = secondsSince(t0);
        const double gpu = ft.upsweep + ft.jbuild + ft.prep + ft.kernel;
        std::fprintf(stderr,
                     "COSMO_FORCE n=%d leaves=%d | upsweep=%.2f jbuild=%.2f upload=%.2f prep=%.2f kernel=%.2f "
                     "download=%.2f | gpusum=%.2f wall=%.2f unaccounted=%.2f\n",
                     s.n, s.nLeaves, 1e3 * ft.upsweep, 1e3 * ft.jbuild, 1e3 * s.uploadSeconds, 1e3 * ft.prep,
                     1e3 * ft.kernel, 1e3 * s.downloadSeconds, 1e3 * gpu, 1e3 * wall, 1e3 * (wall - gpu));
        s.uploadSeconds = 0.0;
        s.downloadSeconds = 0.0;
    }
}

extern "C" void cosmo_arrays_init(int n)
{
    if (n <= 0) return;
    sizeParticleArrays(state(), n);
}

extern "C" int cosmo_slot_ncomp(int slot)
{
    const SlotDesc* d = slotDesc(slot);
    return d ? d->ncomp : 0;
}

extern "C" void cosmo_upload(int slot, const double* host, int n) { upload(slot, host, n, false); }

extern "C" void cosmo_upload_sorted(int slot, const double* host, int n) { upload(slot, host, n, true); }

extern "C" void cosmo_download(int slot, double* host, int n)
{
    if (n <= 0) return;
    State& s = state();
    const SlotDesc& d = slotFor(s, slot, n, "cosmo_download", true);
    ensureStream(s);
    const auto t0 = Clock::now();
    transferOut(s, d, host, n);
    s.downloadSeconds += secondsSince(t0);
    s.downloadTotal += secondsSince(t0);
}

// cosmo_download allocating nothing (no counterpart in the reference): for callers that cannot
// allocate.  Sizes beyond the static buffer are refused before any device work.
extern "C" int cosmo_download_noalloc(int slot, double* host, int n)
{
    if (n > kNoAllocMaxParticles)
    {
        std::fprintf(stderr,
                     "cosmo_download_noalloc: n=%d exceeds the static scatter buffer of %d particles; nothing "
                     "copied (cosmo_download has no such limit)\n",
                     n, kNoAllocMaxParticles);
        return -1;
    }
    if (n <= 0) return 0;
    State& s = state();
    const SlotDesc& d = slotFor(s, slot, n, "cosmo_download_noalloc", true);
    // an ordering exists, so a solve has run and created the stream
    const auto t0 = Clock::now();
    transferOutNoAlloc(s, d, host, n);
    s.downloadSeconds += secondsSince(t0);
    s.downloadTotal += secondsSince(t0);
    return 0;
}

// Page-locking is best effort in the reference (pin_c_api.cu:22-34: failure is cleared and
// ignored).  Registering memory the caller already owns is refused on this platform, and a
// refused registration leaves an error for the next runtime call to report; the copies go
// through the runtime's own staging instead, so these calls accept and ignore the range.
extern "C" void cosmo_pin_host(void* ptr, size_t nbytes)
{
    (void)ptr;
    (void)nbytes;
}

extern "C" void cosmo_unpin_host(void* ptr) { (void)ptr; }

// =========================================================================================
// Extensions
// =========================================================================================
extern "C" int cosmo_ext_solve_dens_h(int n, double pmass, int mode, int withGradients, int periodic,
                                      const double* box, double tolh, double hfact)
{
    State& s = state();
    requireSized(s, n, "cosmo_ext_solve_dens_h");
    ensureStream(s);
    setSingle(s, n);
    Timings t = solveDensH(s, pmass, mode, withGradients != 0, periodic ? box : nullptr, tolh, hfact);
    return t.iters;
}

extern "C" void cosmo_ext_set_option(const char* name, double value)
{
    State& s = state();
    const std::string nm(name);
    if (nm == "precision") s.precision = (int)value;
    else if (nm == "max_iter") s.maxIter = (int)value;
    else if (nm == "mapping") s.mapping = (int)value;
    else if (nm == "sparse") s.sparse = (int)value;
    else if (nm == "tail_leaves") s.tailLeaves = (int)value;
    else if (nm == "closure") s.closureOn = (int)value;
    else if (nm == "tree_only") s.treeOnly = (int)value;
    else if (nm == "bucket")
    {
        if (!(value >= 8 && value <= 1024)) fatal("cosmo_ext_set_option: bucket must lie in [8, 1024]");
        s.bucket = (int)value;
    }
    else if (nm == "loop") s.loopMode = (int)value;
    else if (nm == "graph") s.useGraph = (int)value;
    else if (nm == "graph_passes") s.graphPasses = std::max(1, (int)value);
    else if (nm == "lpt") s.lpt = (int)value;
    else if (nm == "cull") s.cull = (int)value;
    else if (nm == "units") s.units = (int)value;
    else if (nm == "loop_blocks_scale") s.loopBlockScale = std::max(1, (int)value);
    else if (nm == "loop_list_cap") s.loopListCap = (long long)value;
    else if (nm == "list_heavy_steps") s.listHeavySteps = (int)value;
    else if (nm == "list_heavy_factor") s.listHeavyFactor = std::max(0, (int)value);
    else if (nm == "grad_sym_lists") s.gradSymLists = (int)value;
    else if (nm == "split_len") s.splitLen = (int)value;
    else if (nm == "sweep_split") s.sweepSplit = std::max(0, (int)value);
    else if (nm == "count_list") s.countList = (int)value;
    else if (nm == "list_expand_leaves") s.listExpandLeaves = (int)value;
    else if (nm == "resident") s.resident = (int)value;
    else if (nm == "force_rounds") s.forceRounds = (int)value;
    else if (nm == "tile_depth")
    {
        if (value != 64 && value != 128) fatal("cosmo_ext_set_option: tile_depth must be 64 or 128");
        s.tileDepth = (int)value;
    }
    else if (nm == "xfer_threads") s.xferThreads = std::max(1, (int)value);
    else
    {
        std::fprintf(stderr, "FATAL: cosmo_ext_set_option: unknown option '%s'\n", name);
        std::abort();
    }
}

double tuneStat(const std::string& nm);

extern "C" double cosmo_ext_get_stat(const char* name)
{
    State& s = state();
    const std::string nm(name);
    const Timings& t = s.t;
    const ForceTimings& f = s.ft;
    if (nm == "bbox") return t.bbox;
    if (nm == "keys") return t.keys;
    if (nm == "sort") return t.sort;
    if (nm == "permute") return t.permute;
    if (nm == "tree") return t.tree;
    if (nm == "nodes") return t.nodes;
    if (nm == "jbuild") return t.jbuild;
    if (nm == "nrkern") return t.nrkern;
    if (nm == "compact") return t.compact;
    if (nm == "gjbuild") return t.gjbuild;
    if (nm == "gradkern") return t.gradkern;
    if (nm == "closure") return t.closure;
    if (nm == "total") return t.total;
    if (nm == "iters") return t.iters;
    if (nm == "pairs") return t.pairsAccepted;
    if (nm == "candidates") return t.pairsCandidate;
    if (nm == "launches") return t.kernLaunches;
    if (nm == "closure_pairs") return t.closurePairs;
    if (nm == "closure_candidates") return t.closureCand;
    if (nm == "list_candidates") return t.listCand;
    if (nm == "closure_list_candidates") return t.closureListCand;
    if (nm == "force_jbuild") return f.jbuild;
    if (nm == "force_prep") return f.prep;
    if (nm == "force_kernel") return f.kernel;
    if (nm == "force_total") return f.total;
    if (nm == "force_pairs") return f.pairsAccepted;
    if (nm == "force_candidates") return f.pairsCandidate;
    if (nm == "force_list_candidates") return f.listCand;
    if (nm == "n_alive") return s.totalAlive;
    if (nm == "n_leaves") return s.nLeaves;
    if (nm == "n_internal") return s.nInternal;
    if (nm == "n_nodes") return s.nNodes;
    if (nm == "list_total") return s.listTotal;
    if (nm == "list_leaves") return s.listLeaves;
    if (nm == "precision") return s.precision;
    if (nm == "max_iter") return s.maxIter;
    if (nm == "graph_replays") return t.graphReplays;
    if (nm == "graph_captures") return t.graphCaptures;
    if (nm == "loop") return s.loopMode;
    if (nm == "bucket") return s.bucket;
    if (nm.compare(0, 5, "tune_") == 0) return tuneStat(nm);
    if (nm == "graph") return s.useGraph;
    if (nm == "graph_passes") return s.graphPasses;
    if (nm == "lpt") return s.lpt;
    if (nm == "cull") return s.cull;
    if (nm == "units") return s.units;
    if (nm == "mapping") return s.mapping;
    if (nm == "sparse") return s.sparse;
    if (nm == "tail_leaves") return s.tailLeaves;
    if (nm == "loop_blocks_scale") return s.loopBlockScale;
    if (nm == "sweep_split") return s.sweepSplit;
    if (nm == "count_list") return s.countList;
    if (nm == "sweep_items") return s.sweepHeavyItems;
    if (nm == "resident") return s.resident;
    if (nm == "force_rounds") return s.forceRounds;
    if (nm == "list_heavy_factor") return s.listHeavyFactor;
    if (nm == "grad_sym_lists") return s.gradSymLists;
    if (nm == "tile_depth") return s.tileDepth;
    if (nm == "resort_movers") return t.resortMovers;
    if (nm == "upload_s") return s.uploadTotal;
    if (nm == "download_s") return s.downloadTotal;
    return -1.0;
}

// Copies a named device array (Hilbert order where it is per particle) to host memory.
// Returns the number of elements, or -1 for an unknown name; copies at most maxBytes.
extern "C" long long cosmo_ext_get_array(const char* name, void* dst, long long maxBytes)
{
    State& s = state();
    const std::string nm(name);
    const void* src = nullptr;
    long long count = 0, esz = 8;
    if (nm == "passes")   // host records: one PassRecord (all doubles) per NR pass of the last solve
    {
        const long long cnt = (long long)(s.passLog.size() * (sizeof(PassRecord) / sizeof(double)));
        const long long bytes = std::min(cnt * 8, maxBytes);
        if (bytes > 0) std::memcpy(dst, s.passLog.data(), (size_t)bytes);
        return cnt;
    }
    if (nm == "keys") { src = s.keys.p; count = s.n; }
    else if (nm == "order") { src = s.order.p; count = s.n; esz = 4; }
    else if (nm == "leafStart") { src = s.leafStart.p; count = s.nLeaves; }
    else if (nm == "leafLevel") { src = s.leafLevel.p; count = s.nLeaves; esz = 4; }
    else if (nm == "leafSet") { src = s.leafSet.p; count = s.nLeaves; esz = 4; }
    else if (nm == "layout") { src = s.layout.p; count = s.nLeaves + 1; esz = 4; }
    else if (nm == "particleLeaf") { src = s.particleLeaf.p; count = s.totalAlive; esz = 4; }
    else if (nm == "hmaxLeaf") { src = s.hmaxLeaf.p; count = s.nLeaves; }
    else if (nm == "jOffset") { src = s.jOffset.p; count = s.listLeaves + 1; esz = 4; }
    else if (nm == "jCount") { src = s.jCount.p; count = s.listLeaves; esz = 4; }
    else if (nm == "jList") { src = s.jList.p; count = s.listTotal; esz = 4; }
    else if (nm == "intStart") { src = s.intStart.p; count = s.nInternal; }
    else if (nm == "intLevel") { src = s.intLevel.p; count = s.nInternal; esz = 4; }
    else if (nm == "nodeChild") { src = s.nodeChild.p; count = s.nNodes; esz = 4; }
    else if (nm == "nodeCx") { src = s.nodeCx.p; count = s.nNodes; }
    else if (nm == "nodeCy") { src = s.nodeCy.p; count = s.nNodes; }
    else if (nm == "nodeCz") { src = s.nodeCz.p; count = s.nNodes; }
    else if (nm == "nodeSx") { src = s.nodeSx.p; count = s.nNodes; }
    else if (nm == "nodeSy") { src = s.nodeSy.p; count = s.nNodes; }
    else if (nm == "nodeSz") { src = s.nodeSz.p; count = s.nNodes; }
    else if (nm == "leafNode") { src = s.leafNode.p; count = s.nLeaves; esz = 4; }
    else if (nm == "x") { src = s.x.p; count = s.n; }
    else if (nm == "y") { src = s.y.p; count = s.n; }
    else if (nm == "z") { src = s.z.p; count = s.n; }
    else if (nm == "h") { src = s.h.p; count = s.n; }
    else if (nm == "rho") { src = s.rho.p; count = s.n; }
    else if (nm == "gradh") { src = s.gradh.p; count = s.n; }
    else if (nm == "box") { src = s.box.p; count = (long long)s.nSets * (long long)(sizeof(DBox) / 8); }
    else if (nm == "aliveFirst") { src = s.aliveFirst.p; count = s.nSets; esz = 4; }
    else if (nm == "alive") { src = s.alive.p; count = s.nSets; esz = 4; }
    else return -1;
    if (!s.stream) return 0;
    const long long bytes = std::min(count * esz, maxBytes);
    if (bytes > 0 && src)
    {
        HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, s.stream));
        HIP_CHECK(hipStreamSynchronize(s.stream));
    }
    return count;
}

// Benchmark support: keep a device copy of POS|HSML|VEL|ACCEL in the caller's order and put
// it back before each repeated solve, so every solve starts from the same inputs.
extern "C" void cosmo_ext_stash_inputs(void)
{
    State& s = state();
    ensureStream(s);
    inputsToCallerOrder(s);
    const int n = s.sizedFor;
    double* st = s.stash.get<double>((size_t)10 * n + 1);
    for (int slot = COSMO_POS; slot <= COSMO_ACCEL; ++slot)
    {
        const SlotDesc& d = *slotDesc(slot);
        for (int c = 0; c < d.ncomp; ++c)
            HIP_CHECK(hipMemcpyAsync(st + (size_t)(d.firstInput + c) * n, (s.*d.comp[c]).as<double>(), (size_t)n * 8,
                                     hipMemcpyDeviceToDevice, s.stream));
    }
    HIP_CHECK(hipStreamSynchronize(s.stream));
    s.stashN = n;
}

extern "C" void cosmo_ext_restore_inputs(void)
{
    State& s = state();
    const int n = s.stashN;
    if (n != s.sizedFor) fatal("cosmo_ext_restore_inputs: no stash for the current particle count");
    const double* st = s.stash.as<double>();
    for (int slot = COSMO_POS; slot <= COSMO_ACCEL; ++slot)
    {
        const SlotDesc& d = *slotDesc(slot);
        for (int c = 0; c < d.ncomp; ++c)
            HIP_CHECK(hipMemcpyAsync((s.*d.comp[c]).as<double>(), st + (size_t)(d.firstInput + c) * n, (size_t)n * 8,
                                     hipMemcpyDeviceToDevice, s.stream));
    }
    HIP_CHECK(hipStreamSynchronize(s.stream));
    resetInputTags(s);
}

// =========================================================================================
// Ensemble: many independent particle sets solved together.  The sets are stored one after
// another in every slot (set b owns [sum of earlier counts, + counts[b]) of the caller's
// order); each set has its own root box, tree, particle mass and convergence, and one
// sequence of kernel launches covers all of them.
// =========================================================================================
namespace
{
std::vector<int> gBatchCounts;
}

extern "C" void cosmo_batch_init(int nSets, const int* counts)
{
    if (nSets <= 0 || counts == nullptr) fatal("cosmo_batch_init: no sets");
    long long total = 0;
    for (int b = 0; b < nSets; ++b)
    {
        if (counts[b] < 0) fatal("cosmo_batch_init: negative set size");
        total += counts[b];
    }
    if (total <= 0 || total > 2000000000LL) fatal("cosmo_batch_init: total particle count out of range");
    gBatchCounts.assign(counts, counts + nSets);
    sizeParticleArrays(state(), (int)total);
}

static int batchSolve(const double* pmass, int periodic, const double* boxes, double tolh, double hfact);

extern "C" int cosmo_batch_solve(const double* pmass, int periodic, const double* boxes, double tolh, double hfact)
{
    return batchSolve(pmass, periodic, boxes, tolh, hfact);
}

static int batchSolve(const double* pmass, int periodic, const double* boxes, double tolh, double hfact)
{
    State& s = state();
    if (gBatchCounts.empty()) fatal("cosmo_batch_solve: cosmo_batch_init has not been called");
    long long total = 0;
    for (int c : gBatchCounts) total += c;
    requireSized(s, (int)total, "cosmo_batch_solve");
    ensureStream(s);
    const int B = (int)gBatchCounts.size();
    s.n = (int)total;
    s.nSets = B;
    s.hSetFirst.assign(B, 0);
    s.hSetN.assign(gBatchCounts.begin(), gBatchCounts.end());
    for (int b = 1; b < B; ++b) s.hSetFirst[b] = s.hSetFirst[b - 1] + s.hSetN[b - 1];
    double* pm = s.pmassSet.get<double>(B + 1);
    HIP_CHECK(hipMemcpyAsync(pm, pmass, B * sizeof(double), hipMemcpyHostToDevice, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
    s.usePmassSet = 1;
    Timings t = solveDensH(s, pmass[0], 0, false, periodic ? boxes : nullptr, tolh, hfact);
    s.usePmassSet = 0;
    return t.iters;
}

extern "C" void cosmo_ext_sync(void)
{
    if (state().stream) HIP_CHECK(hipStreamSynchronize(state().stream));
}

// =========================================================================================
// Autotuner (tune.hip): picks the speed options of the solve for the uploaded workload.
// =========================================================================================
namespace
{
TuneResult gTune;

std::string signatureOf(const char* kind, const State& s, int n, int periodic, double tolh, double hfact,
                        const std::vector<int>& counts)
{
    unsigned long long h = 1469598103934665603ull;   // FNV-1a of the set sizes
    for (int c : counts)
        for (int k = 0; k < 4; ++k)
        {
            h ^= (unsigned long long)((c >> (8 * k)) & 0xFF);
            h *= 1099511628211ull;
        }
    char b[256];
    std::snprintf(b, sizeof(b), "%s n=%d sets=%zu sizes=%016llx R=%g precision=%d periodic=%d tolh=%g hfact=%g "
                                "closure=%d max_iter=%d",
                  kind, n, counts.size(), h, kRad, s.precision, periodic, tolh, hfact, s.closureOn, s.maxIter);
    return b;
}
} // namespace

extern "C" int cosmo_ext_autotune(int n, double pmass, int periodic, const double* box, double tolh, double hfact,
                                  double budgetSeconds)
{
    State& s = state();
    requireSized(s, n, "cosmo_ext_autotune");
    ensureStream(s);
    const std::string sig = signatureOf("single", s, n, periodic, tolh, hfact, std::vector<int>{n});
    gTune = autotune(s, sig, [&]() {
        setSingle(s, n);
        solveDensH(s, pmass, 0, false, periodic ? box : nullptr, tolh, hfact);
    }, budgetSeconds);
    return gTune.trials;
}

extern "C" int cosmo_batch_autotune(const double* pmass, int periodic, const double* boxes, double tolh, double hfact,
                                    double budgetSeconds)
{
    State& s = state();
    if (gBatchCounts.empty()) fatal("cosmo_batch_autotune: cosmo_batch_init has not been called");
    long long total = 0;
    for (int c : gBatchCounts) total += c;
    requireSized(s, (int)total, "cosmo_batch_autotune");
    ensureStream(s);
    const std::string sig = signatureOf("batch", s, (int)total, periodic, tolh, hfact, gBatchCounts);
    gTune = autotune(s, sig, [&]() { batchSolve(pmass, periodic, boxes, tolh, hfact); }, budgetSeconds);
    return gTune.trials;
}

double tuneStat(const std::string& nm)
{
    if (nm == "tune_trials") return gTune.trials;
    if (nm == "tune_best_s") return gTune.bestSeconds;
    if (nm == "tune_default_s") return gTune.defaultSeconds;
    if (nm == "tune_seconds") return gTune.tuneSeconds;
    if (nm == "tune_cached") return gTune.cached ? 1.0 : 0.0;
    return -1.0;
}
