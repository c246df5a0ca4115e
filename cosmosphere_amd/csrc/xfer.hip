/*
 * xfer.hip -- the slot transfers behind cosmo_upload / cosmo_upload_sorted / cosmo_download
 * (reference src/arrays_c_api.cu:120-193), designed for this platform:
 *
 *   - One table maps a slot to its device arrays (member pointers into State).
 *   - Pageable caller memory cannot be registered here, so large blocks stream through a
 *     ring of kRing page-locked chunks.  A pool of host threads fills (or drains) chunk k
 *     while the DMA engine moves chunk k-1; the stream is synchronised once per call.
 *   - Gathers into and scatters out of the device order handle every component of a slot
 *     in one kernel, reading `order` once.
 *   - Resident order: input slots uploaded after a single-set solve of the same size are
 *     gathered into that solve's order (tagged kTagAligned), so the next solve sees nearly
 *     sorted keys.
 */
#include "xfer.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace csph
{

// ---------------------------------------------------------------------------------------
// Slot table
// ---------------------------------------------------------------------------------------
static const SlotDesc kSlots[9] = {
    {0, -1, {}},
    {3, 0, {&State::x, &State::y, &State::z}},                                          // POS
    {1, 3, {&State::h}},                                                                 // HSML
    {3, 4, {&State::vx, &State::vy, &State::vz}},                                        // VEL
    {3, 7, {&State::ax, &State::ay, &State::az}},                                        // ACCEL
    {2, -1, {&State::rho, &State::gradh}},                                               // DENS_OUT
    {3, -1, {&State::divv, &State::xi, &State::ddivvdt}},                                // GRAD_OUT
    {4, -1, {&State::pro2, &State::cs, &State::alpha, &State::u}},                       // THERMO
    {6, -1, {&State::fx, &State::fy, &State::fz, &State::f4, &State::vsig, &State::divvF}},   // FORCE_OUT
};

const SlotDesc* slotDesc(int slot) { return (slot >= 1 && slot <= 8) ? &kSlots[slot] : nullptr; }

static DevBuf* inputField(State& s, int k)
{
    static DevBuf State::* const f[kNumInputs] = {&State::x,  &State::y,  &State::z,  &State::h,  &State::vx,
                                                  &State::vy, &State::vz, &State::ax, &State::ay, &State::az};
    return &(s.*f[k]);
}

// ---------------------------------------------------------------------------------------
// Host copy pool: one memcpy split over worker threads (>= 1 MiB per part)
// ---------------------------------------------------------------------------------------
namespace
{

class CopyPool
{
public:
    explicit CopyPool(int threads)
    {
        for (int k = 1; k < threads; ++k) workers_.emplace_back([this, k] { work(k); });
    }

    void copy(void* dst, const void* src, size_t bytes)
    {
        const int parts = (int)std::min<size_t>(workers_.size() + 1, std::max<size_t>(1, bytes >> 20));
        if (parts <= 1)
        {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            parts_ = parts;
            left_ = parts - 1;
            ++gen_;
        }
        go_.notify_all();
        piece(dst_, src_, bytes, parts, 0);
        std::unique_lock<std::mutex> lk(m_);
        fin_.wait(lk, [this] { return left_ == 0; });
    }

private:
    static void piece(char* d, const char* s, size_t bytes, int parts, int k)
    {
        const size_t per = ((bytes / parts) + 4095) & ~size_t(4095);
        const size_t off = std::min(bytes, per * k);
        const size_t len = std::min(bytes - off, per);
        if (len) std::memcpy(d + off, s + off, len);
    }

    void work(int k)
    {
        unsigned long long seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;)
        {
            go_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            if (k >= parts_) continue;
            char* d = dst_;
            const char* s = src_;
            const size_t b = bytes_;
            const int p = parts_;
            lk.unlock();
            piece(d, s, b, p, k);
            lk.lock();
            if (--left_ == 0) fin_.notify_one();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable go_, fin_;
    unsigned long long gen_ = 0;
    int left_ = 0, parts_ = 1;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
};

// Page-locked staging: kRing chunks of kChunk bytes, each with the event of its last DMA.
constexpr int    kRing = 4;
constexpr size_t kChunk = size_t(16) << 20;
constexpr size_t kDirectBytes = size_t(2) << 20;   // smaller blocks go straight from pageable memory

struct Staging
{
    char*      buf[kRing] = {};
    hipEvent_t ev[kRing] = {};
    CopyPool*  pool = nullptr;
};

// Created on the first large transfer and kept for the life of the process, like the
// device state (the pool's threads wait on a condition variable between calls).
Staging& staging(int threads)
{
    static Staging* st = nullptr;
    if (!st)
    {
        st = new Staging();
        char* p = nullptr;
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), kChunk * kRing, hipHostMallocDefault));
        for (int r = 0; r < kRing; ++r)
        {
            st->buf[r] = p + kChunk * r;
            HIP_CHECK(hipEventCreateWithFlags(&st->ev[r], hipEventDisableTiming));
        }
        st->pool = new CopyPool(std::max(1, threads));
    }
    return *st;
}

// Caller memory -> device, through the ring.  Returns with the copies queued on the
// stream; the chunks are free again once the stream has passed their events.
void streamToDevice(State& s, const char* host, char* dev, size_t bytes)
{
    if (bytes <= kDirectBytes)
    {
        HIP_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s.stream));
        return;
    }
    Staging& g = staging(s.xferThreads);
    for (size_t off = 0; off < bytes; off += kChunk)
    {
        const int r = s.ringNext;
        s.ringNext = (r + 1) % kRing;
        if (s.ringBusy & (1u << r)) HIP_CHECK(hipEventSynchronize(g.ev[r]));
        const size_t len = std::min(kChunk, bytes - off);
        g.pool->copy(g.buf[r], host + off, len);
        HIP_CHECK(hipMemcpyAsync(dev + off, g.buf[r], len, hipMemcpyHostToDevice, s.stream));
        HIP_CHECK(hipEventRecord(g.ev[r], s.stream));
        s.ringBusy |= 1u << r;
    }
}

// Device -> caller memory, through the ring, up to kRing DMAs in flight ahead of the host
// copies.  Returns when every byte is in the caller's memory.
void streamToHost(State& s, const char* dev, char* host, size_t bytes)
{
    if (bytes <= kDirectBytes)
    {
        HIP_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s.stream));
        HIP_CHECK(hipStreamSynchronize(s.stream));
        return;
    }
    Staging& g = staging(s.xferThreads);
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    size_t issued = 0, drained = 0;
    while (drained < chunks)
    {
        while (issued < chunks && issued - drained < (size_t)kRing)
        {
            const int r = (int)(issued % kRing);
            const size_t off = issued * kChunk, len = std::min(kChunk, bytes - off);
            HIP_CHECK(hipMemcpyAsync(g.buf[r], dev + off, len, hipMemcpyDeviceToHost, s.stream));
            HIP_CHECK(hipEventRecord(g.ev[r], s.stream));
            ++issued;
        }
        const int r = (int)(drained % kRing);
        const size_t off = drained * kChunk, len = std::min(kChunk, bytes - off);
        HIP_CHECK(hipEventSynchronize(g.ev[r]));
        g.pool->copy(host + off, g.buf[r], len);
        ++drained;
    }
    s.ringBusy = 0;
}

// ---------------------------------------------------------------------------------------
// Gather / scatter of every component of a slot in one pass over `order`
// ---------------------------------------------------------------------------------------
struct CompPtrs
{
    double* p[6];
};

// dst[c][i] = src[c * n + order[i]]: caller order -> device order
__global__ __launch_bounds__(256) void gatherSlotKernel(const double* __restrict__ src, const int* __restrict__ order,
                                                        CompPtrs dst, int ncomp, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)order[i];
    for (int c = 0; c < ncomp; ++c) dst.p[c][i] = src[(size_t)c * n + o];
}

// dst[c * n + order[i]] = src[c][i]: device order -> caller order
__global__ __launch_bounds__(256) void scatterSlotKernel(CompPtrs src, const int* __restrict__ order,
                                                         double* __restrict__ dst, int ncomp, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)order[i];
    for (int c = 0; c < ncomp; ++c) dst[(size_t)c * n + o] = src.p[c][i];
}

// One field between the two orders, through a spare buffer that is then swapped in.
__global__ __launch_bounds__(256) void reorderKernel(const double* __restrict__ src, const int* __restrict__ order,
                                                     double* __restrict__ dst, int n, int toAligned)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = order[i];
    if (toAligned) dst[i] = src[o];
    else dst[o] = src[i];
}

void reorderField(State& s, int k, bool toAligned)
{
    const int n = s.n;
    DevBuf* f = inputField(s, k);
    DevBuf& spare = s.permAlt[k];
    double* dst = spare.get<double>((size_t)std::max(n, s.sizedFor) + 1);
    if (n > 0)
    {
        reorderKernel<<<iceil(n, 256), 256, 0, s.stream>>>(f->as<double>(), s.order.as<int>(), dst, n,
                                                             toAligned ? 1 : 0);
        LAUNCH_CHECK();
    }
    std::swap(*f, spare);
    s.inTag[k] = toAligned ? kTagAligned : kTagCaller;
}

} // namespace

// ---------------------------------------------------------------------------------------
// Transfers
// ---------------------------------------------------------------------------------------
void transferIn(State& s, const SlotDesc& d, const double* host, int n, bool aligned)
{
    const size_t nb = (size_t)n * sizeof(double);
    if (!aligned)
    {
        for (int c = 0; c < d.ncomp; ++c)
            streamToDevice(s, reinterpret_cast<const char*>(host + (size_t)c * n), (s.*d.comp[c]).as<char>(), nb);
    }
    else
    {
        double* stage = s.stage.get<double>((size_t)d.ncomp * n + 1);
        streamToDevice(s, reinterpret_cast<const char*>(host), reinterpret_cast<char*>(stage), nb * d.ncomp);
        CompPtrs dst{};
        for (int c = 0; c < d.ncomp; ++c) dst.p[c] = (s.*d.comp[c]).as<double>();
        gatherSlotKernel<<<iceil(n, 256), 256, 0, s.stream>>>(stage, s.order.as<int>(), dst, d.ncomp, n);
        LAUNCH_CHECK();
    }
    HIP_CHECK(hipStreamSynchronize(s.stream));
    s.ringBusy = 0;
    if (d.firstInput >= 0)
        for (int c = 0; c < d.ncomp; ++c) s.inTag[d.firstInput + c] = aligned ? kTagAligned : kTagCaller;
}

void transferOut(State& s, const SlotDesc& d, double* host, int n)
{
    double* stage = s.stage.get<double>((size_t)d.ncomp * n + 1);
    CompPtrs src{};
    for (int c = 0; c < d.ncomp; ++c) src.p[c] = (s.*d.comp[c]).as<double>();
    scatterSlotKernel<<<iceil(n, 256), 256, 0, s.stream>>>(src, s.order.as<int>(), stage, d.ncomp, n);
    LAUNCH_CHECK();
    streamToHost(s, reinterpret_cast<const char*>(stage), reinterpret_cast<char*>(host),
                 (size_t)d.ncomp * n * sizeof(double));
}

// ---------------------------------------------------------------------------------------
// Allocation-free download: each component is scattered into a static device buffer (the
// array `scratch`, one copy per GPU, held until the process ends) and copied from there
// straight into the caller's memory; no staging buffer, page-locked ring or thread pool.
// The scatter writes every entry of [0, n) before the copy reads it, so nothing needs
// clearing between calls.  Calls on one GPU share the buffer: overlapping calls race.
// ---------------------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kNoAllocMaxParticles];

namespace
{
constexpr int kMaxDevices = 64;
double* gScratch[kMaxDevices] = {};   // device address of `scratch` on each GPU

// The address of `scratch` on the current device, looked up on the first call there (the
// array's host-side name is not a device address and must never reach a kernel).
double* scratchOnDevice()
{
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
    {
        std::fprintf(stderr, "FATAL: cosmo_download_noalloc: device %d beyond the %d this library tracks\n", dev,
                     kMaxDevices);
        std::abort();
    }
    if (!gScratch[dev])
    {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        gScratch[dev] = static_cast<double*>(p);
    }
    return gScratch[dev];
}
} // namespace

void transferOutNoAlloc(State& s, const SlotDesc& d, double* host, int n)
{
    double* buf = scratchOnDevice();
    for (int c = 0; c < d.ncomp; ++c)
    {
        CompPtrs src{};
        src.p[0] = (s.*d.comp[c]).as<double>();
        scatterSlotKernel<<<iceil(n, 256), 256, 0, s.stream>>>(src, s.order.as<int>(), buf, 1, n);
        LAUNCH_CHECK();
        // stream order: the next component's scatter waits for this copy
        HIP_CHECK(hipMemcpyAsync(host + (size_t)c * n, buf, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                                 s.stream));
    }
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

// ---------------------------------------------------------------------------------------
// Resident order
// ---------------------------------------------------------------------------------------
bool residentEligible(const State& s, int n)
{
    return s.resident && s.haveOrder && s.n == n && s.orderSets == 1 && n > 0;
}

bool prepareInputs(State& s, bool withVelAcc)
{
    bool anyAligned = false, posAligned = true;
    for (int k = 0; k < kNumInputs; ++k) anyAligned |= s.inTag[k] == kTagAligned;
    for (int k = 0; k < 4; ++k) posAligned &= s.inTag[k] == kTagAligned;
    if (!anyAligned) return false;
    // aligned tags exist only while the order they refer to is the current one
    const bool incremental = posAligned && s.nSets == 1 && s.hSetN.size() == 1 && s.hSetN[0] == s.n;
    for (int k = 0; k < kNumInputs; ++k)
    {
        const bool want = incremental && (k < 4 || withVelAcc);
        if (want && s.inTag[k] != kTagAligned) reorderField(s, k, true);
        if (!want && s.inTag[k] == kTagAligned) reorderField(s, k, false);
    }
    return incremental;
}

void inputsToCallerOrder(State& s)
{
    for (int k = 0; k < kNumInputs; ++k)
        if (s.inTag[k] == kTagAligned) reorderField(s, k, false);
}

void resetInputTags(State& s)
{
    for (int k = 0; k < kNumInputs; ++k) s.inTag[k] = kTagCaller;
}

} // namespace csph
