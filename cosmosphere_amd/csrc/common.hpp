/*
 * common.hpp -- shared definitions of the gfx950 SPH library: error handling, the
 * smoothing kernel chosen at compile time, key constants and wavefront helpers.
 *
 * The smoothing kernel is fixed at build time as in the reference (include/kernel.hpp:
 * 80-100, KERNEL=cubic|quintic): the build emits libcosmoSPHere.so (M4 cubic) and
 * libcosmoSPHere_quintic.so (M6 quintic, -DCOSMO_KERNEL_QUINTIC) from the same sources.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <type_traits>

// Abort on any runtime error, naming the call site (reference convention,
// include/gpu_check.hpp:14-22).
#define HIP_CHECK(call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "HIP error %s:%d  %s\n", __FILE__, __LINE__,            \
                         hipGetErrorString(e_));                                         \
            std::abort();                                                                \
        }                                                                                \
    } while (0)

#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

namespace csph
{

constexpr int      kMaxLevel  = 21;                       // 63-bit Hilbert keys
constexpr uint64_t kRootRange = uint64_t(1) << 63;        // 8^21
constexpr uint64_t kDeadKey   = ~uint64_t(0);
constexpr int      kWave      = 64;                       // CDNA wavefront
constexpr int      kBucket    = 64;                       // leaf bucket size (tree parity size)
constexpr double   kPi        = 3.141592653589793;
constexpr double   kFinestCell = 1.0 / 2097152.0;         // level-21 cell per unit box length

#ifdef COSMO_KERNEL_QUINTIC
constexpr double kRad    = 3.0;
constexpr double kRad2   = 9.0;
constexpr double kCnormk = 1.0 / (120.0 * kPi);
constexpr double kHfactDefault = 1.0;
constexpr const char* kKernelName = "M_6 quintic";
#else
constexpr double kRad    = 2.0;
constexpr double kRad2   = 4.0;
constexpr double kCnormk = 1.0 / kPi;
constexpr double kHfactDefault = 1.2;
constexpr const char* kKernelName = "M_4 cubic";
#endif

// M4 cubic (include/kernel.hpp:35-47) and M6 quintic (:53-78): W and dW/dq up to the
// normalisation, in the arithmetic type T (double for the plain path, float for mixed).
template<class T>
__host__ __device__ inline void kernWdW(T q2, T q, T& w, T& dw)
{
#ifdef COSMO_KERNEL_QUINTIC
    if (q < T(1))
    {
        const T q4 = q2 * q2;
        w  = T(-10) * q4 * q + T(30) * q4 - T(60) * q2 + T(66);
        dw = q * (T(-50) * q2 * q + T(120) * q2 - T(120));
    }
    else if (q < T(2))
    {
        const T a = q - T(3), b = q - T(2), a2 = a * a, b2 = b * b;
        w  = -(a2 * a2 * a) + T(6) * (b2 * b2 * b);
        dw = T(-5) * (a2 * a2) + T(30) * (b2 * b2);
    }
    else if (q < T(3))
    {
        const T a = q - T(3), a2 = a * a;
        w  = -(a2 * a2 * a);
        dw = T(-5) * (a2 * a2);
    }
    else { w = T(0); dw = T(0); }
#else
    (void)q2;
    const T a = (T(2) - q > T(0)) ? (T(2) - q) : T(0);
    const T b = (T(1) - q > T(0)) ? (T(1) - q) : T(0);
    const T a2 = a * a, a3 = a2 * a, b2 = b * b, b3 = b2 * b;
    w  = T(0.25) * (a3 - T(4) * b3);
    dw = T(0.25) * (T(-3) * a2 + T(12) * b2);
#endif
}

// Branch-free pair terms of the mixed form (a wavefront whose q values straddle several
// pieces would otherwise execute every piece), FP32, for 0 <= q <= R, with q2 = q^2.
// With V the bracket of W and D the bracket of dW/dq,
//   quintic  W = V,       dW/dq = -5 D,    kWw = 1,    kGp = 5,    kGw = 3
//   cubic    W = V / 4,   dW/dq = -3/4 D,  kWw = 1/4,  kGp = 3/4,  kGw = 3/4
// so that W = kWw V and -q dW/dq - 3 W = kGp (q D) - kGw V: a pair loop accumulates V and q D
// and applies the constants once per particle.  Two forms, selected without a branch:
//   q >= 1   the truncated powers (the reference's outer pieces, include/kernel.hpp:35-78):
//            quintic V = a^5 - 6 b^5, D = a^4 - 6 b^4 (a = 3-q, b = (2-q)_+);
//            cubic   V = a^3,         D = a^2         (a = 2-q)
//   q < 1    the inner piece expanded in q, free of the cancellation a^5 - 6 b^5 + 15 c^5 =
//            243 - 192 + 15 that costs FP32 ~1.5e-6 of W(0) near q = 0:
//            quintic V = 66 - 60 q^2 + 30 q^4 - 10 q^5,  q D = 24 q^2 - 24 q^4 + 10 q^5;
//            cubic   V = 4 - 6 q^2 + 3 q^3,              q D = 4 q^2 - 3 q^3
// All coefficients are exact in FP32, so V(0) (the self pair of every particle) is exact.
#ifdef COSMO_KERNEL_QUINTIC
constexpr double kWw = 1.0, kGp = 5.0, kGw = 3.0;
#else
constexpr double kWw = 0.25, kGp = 0.75, kGw = 0.75;
#endif

template<class T>
__host__ __device__ inline void kernVD(T q, T q2, T& V, T& qD)
{
    const bool inner = q2 < T(1);   // q < 1 (at q = 1 both forms hold), known before the root
#ifdef COSMO_KERNEL_QUINTIC
    // b is used for q >= 1 only, where 2 - q <= 1: clamped to [0, 1] (one VALU op with the
    // clamp modifier) it equals the reference's (2 - q)_+
    const T a = T(3) - q;
    const T b = fmin(fmax(T(2) - q, T(0)), T(1));
    const T a2 = a * a, b2 = b * b;
    const T a4 = a2 * a2, b4 = b2 * b2;
    const T vOut = fma(T(-6), b4 * b, a4 * a);
    const T dOut = fma(T(-6), b4, a4);
    const T vIn  = fma(q2, fma(q2, fma(T(-10), q, T(30)), T(-60)), T(66));
    const T qdIn = q2 * fma(q2, fma(T(10), q, T(-24)), T(24));
    V  = inner ? vIn : vOut;
    qD = inner ? qdIn : q * dOut;
#else
    const T a = T(2) - q;
    const T a2 = a * a;
    const T vIn  = fma(q2, fma(T(3), q, T(-6)), T(4));
    const T qdIn = q2 * fma(T(-3), q, T(4));
    V  = inner ? vIn : a2 * a;
    qD = inner ? qdIn : q * a2;
#endif
}

// Root box of one particle set (Cornerstone Box restated: lx = xmax - xmin, ilx = 1/lx).
struct DBox
{
    double lo[3];
    double hi[3];
    double l[3];
    double il[3];
    int    periodic;
    int    pad;
};

__host__ __device__ inline DBox makeDBox(const double b[6], int periodic)
{
    DBox box;
    for (int a = 0; a < 3; ++a)
    {
        box.lo[a] = b[2 * a];
        box.hi[a] = b[2 * a + 1];
        box.l[a]  = box.hi[a] - box.lo[a];
        box.il[a] = 1.0 / box.l[a];
    }
    box.periodic = periodic;
    box.pad      = 0;
    return box;
}

// ---------------------------------------------------------------------------------------
// Wavefront helpers (wave64).
// ---------------------------------------------------------------------------------------
__device__ inline int laneId() { return __lane_id(); }

__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }

__device__ inline uint64_t lanemaskLt() { return (uint64_t(1) << laneId()) - 1; }

// Cross-lane moves by DPP (data-parallel primitives: the operand of a VALU instruction read
// from another lane of the same row, no LDS traffic).  GFX9 controls: row_shr:n shifts within
// rows of 16 lanes; row_bcast:15 / row_bcast:31 hand lane 15 of each row (lane 31) to the next
// row (rows 2 and 3).  Lanes without a source, or outside row_mask, receive `old`.
constexpr int kDppRowShr1    = 0x111;
constexpr int kDppRowShr2    = 0x112;
constexpr int kDppRowShr4    = 0x114;
constexpr int kDppRowShr8    = 0x118;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

template<int CTRL, int ROW_MASK, class T>
__device__ inline T dppMove(T old, T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4)
    {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v),
                                                                 CTRL, ROW_MASK, 0xF, false));
    }
    else
    {
        const long long o = __builtin_bit_cast(long long, old), x = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, ROW_MASK, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, ROW_MASK, 0xF, false);
        return __builtin_bit_cast(T, ((long long)hi << 32) | (long long)(uint32_t)lo);
    }
}

// Inclusive scan over the 64 lanes with an associative op and its identity: Hillis-Steele
// within each row (shifts 1, 2, 4, 8), then the rows joined by the two broadcasts.
template<class T, class Op>
__device__ inline T waveScanDpp(T v, T id, Op op)
{
    v = op(v, dppMove<kDppRowShr1, 0xF>(id, v));
    v = op(v, dppMove<kDppRowShr2, 0xF>(id, v));
    v = op(v, dppMove<kDppRowShr4, 0xF>(id, v));
    v = op(v, dppMove<kDppRowShr8, 0xF>(id, v));
    v = op(v, dppMove<kDppRowBcast15, 0xA>(id, v));
    v = op(v, dppMove<kDppRowBcast31, 0xC>(id, v));
    return v;
}

// The value of lane 63, in every lane (a scalar read).
template<class T>
__device__ inline T lastLane(T v)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), kWave - 1));
    else
    {
        const long long x = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_readlane((int)x, kWave - 1), hi = __builtin_amdgcn_readlane((int)(x >> 32), kWave - 1);
        return __builtin_bit_cast(T, ((long long)hi << 32) | (long long)(uint32_t)lo);
    }
}

// Sum over the wave, in every lane (the scan's total in lane 63).
template<class T>
__device__ inline T waveReduceSum(T v)
{
    return lastLane(waveScanDpp(v, T(0), [](T a, T b) { return a + b; }));
}

// Identities of max and min.
template<class T>
__device__ inline T lowestOf()
{
    if constexpr (std::is_floating_point<T>::value) return T(-__builtin_inf());
    else return std::numeric_limits<T>::lowest();
}

template<class T>
__device__ inline T highestOf()
{
    if constexpr (std::is_floating_point<T>::value) return T(__builtin_inf());
    else return std::numeric_limits<T>::max();
}

template<class T>
__device__ inline T waveReduceMax(T v)
{
    return lastLane(waveScanDpp(v, lowestOf<T>(), [](T a, T b) { return a > b ? a : b; }));
}

template<class T>
__device__ inline T waveReduceMin(T v)
{
    return lastLane(waveScanDpp(v, highestOf<T>(), [](T a, T b) { return a < b ? a : b; }));
}

template<class T>
__device__ inline T waveReduceOr(T v)
{
    return lastLane(waveScanDpp(v, T(0), [](T a, T b) { return a | b; }));
}

// Inclusive prefix sum over the wave.
template<class T>
__device__ inline T waveInclusiveScan(T v)
{
    return waveScanDpp(v, T(0), [](T a, T b) { return a + b; });
}

// Inclusive prefix maximum over the wave.
template<class T>
__device__ inline T waveInclusiveMax(T v, T id)
{
    return waveScanDpp(v, id, [](T a, T b) { return a > b ? a : b; });
}

// Orders memory operations of one wavefront through LDS: what a lane stored is visible
// to the other lanes of the same wave after this point.
__device__ inline void waveLdsSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

inline int iceil(long long a, long long b) { return (int)((a + b - 1) / b); }

} // namespace csph
