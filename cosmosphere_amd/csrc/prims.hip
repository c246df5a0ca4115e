/*
 * prims.hip -- device-wide primitives written for gfx950: exclusive scan, a stable LSD
 * radix sort of (u64 key, u32 value) pairs, and a key-spread reduction that lets the
 * sort skip digits every key shares.  These replace the Thrust calls of the reference
 * (sequence + sort_by_key, src/tree.cu:90-91; the inclusive scans of tree.cu:151 and
 * tree.cuh:385).
 *
 * Sort: 8-bit digits.  Each 256-thread block owns a tile of 4096 keys; each wavefront a
 * contiguous 1024 of them, visited in 16 rounds of 64 coalesced loads.  Per pass:
 *   radixHist    per-block digit histogram (LDS atomics), written digit-major;
 *   scan         over the digit-major histogram -> global base of (digit, block);
 *   radixScatter wave-level multisplit: 8 ballots give each lane the mask of lanes with
 *                its digit, the rank is the popcount below it, and a per-wave LDS counter
 *                carries ranks across rounds -- so the order within a digit is the input
 *                order, which is what makes the sort stable.
 */
#include "density.hpp"
#include "prims.hpp"

#include <algorithm>

namespace csph
{

// ---------------------------------------------------------------------------------------
// Exclusive scan
// ---------------------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems   = 8;
constexpr int kScanTile    = kScanThreads * kScanItems;

// Count of a device-driven scan: entries from nEff on read as zero; tiles past nEff are
// not written (only out[0..nEff] is defined).
__device__ inline int scanCount(const PassCtl* ctl, int n) { return ctl ? min(n, (int)ctl->nAL) : n; }

template<class T>
__global__ __launch_bounds__(kScanThreads) void scanTileKernel(const T* __restrict__ in, T* __restrict__ out,
                                                               T* __restrict__ blockSums, int n,
                                                               unsigned long long* __restrict__ total64,
                                                               const PassCtl* __restrict__ ctl)
{
    __shared__ T waveTot[kScanThreads / kWave];
    if (ctl && ctl->skip) return;
    const int nEff = scanCount(ctl, n);
    if (ctl && (int)blockIdx.x * kScanTile > nEff) return;   // the whole block
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    T v[kScanItems];
    T sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        const int i = base + k;
        v[k] = (i < nEff) ? in[i] : T(0);
        sum += v[k];
    }
    // block-wide exclusive scan of the per-thread sums
    const T incl = waveInclusiveScan(sum);
    const int w = threadIdx.x / kWave;
    if (laneId() == kWave - 1) waveTot[w] = incl;
    __syncthreads();
    T waveOff = 0, blockTot = 0;
#pragma unroll
    for (int j = 0; j < kScanThreads / kWave; ++j)
    {
        if (j < w) waveOff += waveTot[j];
        blockTot += waveTot[j];
    }
    T run = waveOff + incl - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        const int i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (threadIdx.x == 0)
    {
        if (blockSums) blockSums[blockIdx.x] = blockTot;
        if (total64) atomicAdd(total64, (unsigned long long)blockTot);
    }
}

template<class T>
__global__ void scanAddKernel(T* __restrict__ out, const T* __restrict__ blockOffs, int n,
                              const PassCtl* __restrict__ ctl)
{
    if (ctl && ctl->skip) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (!ctl || i <= scanCount(ctl, n))) out[i] += blockOffs[i / kScanTile];
}

template<class T>
__global__ void scanFinishKernel(const T* __restrict__ in, T* __restrict__ out, int n, const PassCtl* __restrict__ ctl)
{
    // out[n] = out[n-1] + in[n-1] (device-driven: only when every entry counts)
    if (ctl && (ctl->skip || scanCount(ctl, n) < n)) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = (n > 0) ? out[n - 1] + in[n - 1] : T(0);
}

size_t scanScratchBytes(int n)
{
    size_t bytes = 0;
    while (n > kScanTile)
    {
        n = iceil(n, kScanTile);
        bytes += (size_t)(n + 1) * 8 * 2 + 256;
    }
    return bytes + 1024;
}

// The block sums of a device-driven scan are complete up to the tile holding nEff; the
// levels above scan them over their full length (the sums of tiles past nEff are stale and
// only reach offsets no kernel reads).
template<class T>
static void scanRec(const T* in, T* out, int n, char* scratch, unsigned long long* total64, hipStream_t st,
                    const PassCtl* ctl)
{
    const int nb = iceil(n, kScanTile);
    if (nb <= 1)
    {
        scanTileKernel<T><<<1, kScanThreads, 0, st>>>(in, out, nullptr, n, total64, ctl);
        LAUNCH_CHECK();
        return;
    }
    T* sums = reinterpret_cast<T*>(scratch);
    T* offs = sums + nb + 1;
    char* next = reinterpret_cast<char*>(offs + nb + 1);
    next = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(next) + 255) & ~uintptr_t(255));
    scanTileKernel<T><<<nb, kScanThreads, 0, st>>>(in, out, sums, n, total64, ctl);
    LAUNCH_CHECK();
    scanRec<T>(sums, offs, nb, next, nullptr, st, nullptr);
    scanAddKernel<T><<<iceil(n, 256), 256, 0, st>>>(out, offs, n, ctl);
    LAUNCH_CHECK();
}

template<class T>
void scanExclusive(const T* in, T* out, int n, DevBuf& scratch, unsigned long long* total64, hipStream_t st,
                   const PassCtl* ctl)
{
    char* sc = scratch.get<char>(scanScratchBytes(n));
    if (n > 0) scanRec<T>(in, out, n, sc, total64, st, ctl);
    scanFinishKernel<T><<<1, 64, 0, st>>>(in, out, n, ctl);
    LAUNCH_CHECK();
}

template void scanExclusive<uint32_t>(const uint32_t*, uint32_t*, int, DevBuf&, unsigned long long*, hipStream_t,
                                      const PassCtl*);
template void scanExclusive<int>(const int*, int*, int, DevBuf&, unsigned long long*, hipStream_t, const PassCtl*);

// ---------------------------------------------------------------------------------------
// Key spread: OR over i of (key_i XOR key_0), so a digit no key differs in is skipped.
// ---------------------------------------------------------------------------------------
__global__ void keySpreadKernel(const uint64_t* __restrict__ keys, int n, unsigned long long* __restrict__ out)
{
    const uint64_t k0 = keys[0];
    uint64_t acc = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc |= keys[i] ^ k0;
    acc = waveReduceOr(acc);
    if (laneId() == 0 && acc) atomicOr(out, (unsigned long long)acc);
}

// ---------------------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortWaves   = kSortThreads / kWave;
constexpr int kSortRounds  = 16;
constexpr int kSortTile    = kSortThreads * kSortRounds;   // 4096

__global__ __launch_bounds__(kSortThreads) void radixHistKernel(const uint64_t* __restrict__ keys, int n, int shift,
                                                                uint32_t* __restrict__ hist, int numBlocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = laneId();
    const int base = blockIdx.x * kSortTile + w * (kSortRounds * kWave);
#pragma unroll 4
    for (int r = 0; r < kSortRounds; ++r)
    {
        const int i = base + r * kWave + lane;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * numBlocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void radixScatterKernel(const uint64_t* __restrict__ keysIn,
                                                                   const uint32_t* __restrict__ valsIn,
                                                                   uint64_t* __restrict__ keysOut,
                                                                   uint32_t* __restrict__ valsOut, int n, int shift,
                                                                   const uint32_t* __restrict__ histOffs, int numBlocks,
                                                                   int* __restrict__ err)
{
    __shared__ uint32_t cnt[kSortWaves][256];
    for (int d = threadIdx.x; d < kSortWaves * 256; d += kSortThreads) (&cnt[0][0])[d] = 0;
    __syncthreads();

    const int w = threadIdx.x / kWave, lane = laneId();
    const int base = blockIdx.x * kSortTile + w * (kSortRounds * kWave);
    uint64_t k[kSortRounds];
    uint32_t v[kSortRounds];
    uint32_t rank[kSortRounds];
    uint32_t dig[kSortRounds];

#pragma unroll
    for (int r = 0; r < kSortRounds; ++r)
    {
        const int  i     = base + r * kWave + lane;
        const bool valid = i < n;
        k[r]   = valid ? keysIn[i] : 0;
        v[r]   = valid ? valsIn[i] : 0;
        const uint32_t d = (uint32_t)(k[r] >> shift) & 0xFF;
        dig[r] = d;
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            const bool     bit = (d >> b) & 1u;
            const uint64_t m   = ballot64(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lanemaskLt());
        uint32_t pre = 0;
        if (valid) pre = cnt[w][d];
        waveLdsSync();
        if (valid && below == 0) cnt[w][d] = pre + (uint32_t)__popcll(peers);
        waveLdsSync();
        rank[r] = pre + below;
    }
    __syncthreads();
    // per digit: offset of this wave's run = global base + counts of earlier waves
    {
        const int d = threadIdx.x;   // 256 threads, one digit each
        uint32_t run = histOffs[d * numBlocks + blockIdx.x];
#pragma unroll
        for (int ww = 0; ww < kSortWaves; ++ww)
        {
            const uint32_t c = cnt[ww][d];
            cnt[ww][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r)
    {
        const int i = base + r * kWave + lane;
        if (i < n)
        {
            const uint32_t dst = cnt[w][dig[r]] + rank[r];
            if (dst < (uint32_t)n)
            {
                keysOut[dst] = k[r];
                valsOut[dst] = v[r];
            }
            else
            {
                atomicAdd(err, 1);
            }
        }
    }
}

__global__ void fillU32Kernel(uint32_t* __restrict__ p, uint32_t v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

void fillU32(uint32_t* p, uint32_t v, size_t n, hipStream_t st)
{
    if (n == 0) return;
    fillU32Kernel<<<(int)std::min<size_t>((n + 255) / 256, 4096), 256, 0, st>>>(p, v, n);
    LAUNCH_CHECK();
}

__global__ void iotaKernel(uint32_t* __restrict__ v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

void iota(uint32_t* v, int n, hipStream_t st)
{
    if (n <= 0) return;
    iotaKernel<<<iceil(n, 256), 256, 0, st>>>(v, n);
    LAUNCH_CHECK();
}

bool radixSortPairs(uint64_t* keys, uint64_t* keysAlt, uint32_t* vals, uint32_t* valsAlt, int n, int bits,
                    SortScratch& sc, hipStream_t st)
{
    if (n <= 1) return false;
    // which digits differ at all
    unsigned long long* spread = sc.counter.get<unsigned long long>(1);
    fillU32(reinterpret_cast<uint32_t*>(spread), 0u, 2, st);
    keySpreadKernel<<<std::min(iceil(n, 256), 1024), 256, 0, st>>>(keys, n, spread);
    LAUNCH_CHECK();
    unsigned long long hspread = 0;
    HIP_CHECK(hipMemcpyAsync(&hspread, spread, sizeof(hspread), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));

    const int nb = iceil(n, kSortTile);
    uint32_t* hist = sc.hist.get<uint32_t>((size_t)256 * nb + 1);
    uint32_t* offs = sc.offs.get<uint32_t>((size_t)256 * nb + 1);
    bool flipped = false;
    uint64_t* kin = keys;  uint64_t* kout = keysAlt;
    uint32_t* vin = vals;  uint32_t* vout = valsAlt;
    for (int shift = 0; shift < bits; shift += 8)
    {
        if (((hspread >> shift) & 0xFF) == 0) continue;
        radixHistKernel<<<nb, kSortThreads, 0, st>>>(kin, n, shift, hist, nb);
        LAUNCH_CHECK();
        scanExclusive<uint32_t>(hist, offs, 256 * nb, sc.scan, nullptr, st);
        radixScatterKernel<<<nb, kSortThreads, 0, st>>>(kin, vin, kout, vout, n, shift, offs, nb, sc.err);
        LAUNCH_CHECK();
        std::swap(kin, kout);
        std::swap(vin, vout);
        flipped = !flipped;
    }
    return flipped;   // true: the result is in keysAlt / valsAlt
}

} // namespace csph
