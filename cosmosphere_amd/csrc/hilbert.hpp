/*
 * hilbert.hpp -- 63-bit 3-D Hilbert keys (21 octal digits) and their decoding.
 *
 * The curve is the one Cornerstone's hilbert3D produces, restated from its published algorithm: levels
 * 20..0, digit mortonToHilbert[octant] with table {0,1,3,2,7,6,4,5}, then the
 * coordinate frame is complemented and rotated.  Here the running frame is held as an
 * axis permutation plus a complement bit per axis, packed into one small integer state
 * and advanced through a 48-state x 8-octant transition table built once at compile time.
 * The table form is what the GPU evaluates; it is derived from the frame rule, so the CPU
 * oracle (which applies the rule bit by bit) is an independent check of it.
 */
#pragma once

#include "common.hpp"

namespace csph
{

// Frame state: perm index p (0..5, permutation of the three axes) and complement mask f
// (3 bits): transformed axis a reads original axis kPerm[p][a], complemented if bit a of f.
struct HilbertTables
{
    uint8_t digit[48][8];   // [state][original-axis octant] -> Hilbert digit
    uint8_t next[48][8];    // [state][original-axis octant] -> next state
    uint8_t octOf[48][8];   // [state][digit] -> original-axis octant (decoder)
};

__host__ __device__ constexpr int permOf(int p, int a)
{
    // the six permutations of (0,1,2)
    return (p == 0) ? a : (p == 1) ? (a == 0 ? 1 : a == 1 ? 2 : 0)
         : (p == 2) ? (a == 0 ? 2 : a == 1 ? 0 : 1) : (p == 3) ? (a == 0 ? 2 : a == 1 ? 1 : 0)
         : (p == 4) ? (a == 0 ? 0 : a == 1 ? 2 : 1) : (a == 0 ? 1 : a == 1 ? 0 : 2);
}

constexpr int permIndex(int a0, int a1, int a2)
{
    for (int p = 0; p < 6; ++p)
        if (permOf(p, 0) == a0 && permOf(p, 1) == a1 && permOf(p, 2) == a2) return p;
    return -1;
}

constexpr HilbertTables makeHilbertTables()
{
    HilbertTables t{};
    const int m2h[8] = {0, 1, 3, 2, 7, 6, 4, 5};
    for (int st = 0; st < 48; ++st)
    {
        const int p = st >> 3, f = st & 7;
        for (int oct = 0; oct < 8; ++oct)   // oct = original-axis bits (x<<2 | y<<1 | z)
        {
            int ob[3] = {(oct >> 2) & 1, (oct >> 1) & 1, oct & 1};
            int tb[3] = {0, 0, 0};
            for (int a = 0; a < 3; ++a) tb[a] = ob[permOf(p, a)] ^ ((f >> (2 - a)) & 1);
            const int xi = tb[0], yi = tb[1], zi = tb[2];
            const int d = m2h[(xi << 2) | (yi << 1) | zi];
            // complement: flips the transformed axes
            int fl[3] = {(f >> 2) & 1, (f >> 1) & 1, f & 1};
            fl[0] ^= (xi & ((!yi) | zi));
            fl[1] ^= ((xi & (yi | zi)) | (yi & (!zi)));
            fl[2] ^= ((xi & (!yi) & (!zi)) | (yi & (!zi)));
            int pm[3] = {permOf(p, 0), permOf(p, 1), permOf(p, 2)};
            if (zi)
            {
                const int p0 = pm[0], f0 = fl[0];
                pm[0] = pm[1]; fl[0] = fl[1];
                pm[1] = pm[2]; fl[1] = fl[2];
                pm[2] = p0;    fl[2] = f0;
            }
            else if (!yi)
            {
                int tp = pm[0]; pm[0] = pm[2]; pm[2] = tp;
                int tf = fl[0]; fl[0] = fl[2]; fl[2] = tf;
            }
            const int np = permIndex(pm[0], pm[1], pm[2]);
            const int nf = (fl[0] << 2) | (fl[1] << 1) | fl[2];
            t.digit[st][oct] = (uint8_t)d;
            t.next[st][oct]  = (uint8_t)((np << 3) | nf);
            t.octOf[st][d]   = (uint8_t)oct;
        }
    }
    return t;
}

__constant__ constexpr HilbertTables kHilbert = makeHilbertTables();

// Integer coordinates (each < 2^21) -> key.
__device__ inline uint64_t hilbertEncode(unsigned ix, unsigned iy, unsigned iz)
{
    uint64_t key = 0;
    int      st  = 0;
#pragma unroll
    for (int level = kMaxLevel - 1; level >= 0; --level)
    {
        const int oct = (((ix >> level) & 1u) << 2) | (((iy >> level) & 1u) << 1) | ((iz >> level) & 1u);
        key = (key << 3) | kHilbert.digit[st][oct];
        st  = kHilbert.next[st][oct];
    }
    return key;
}

// Key -> integer coordinates of its level-21 cell.
__device__ inline void hilbertDecode(uint64_t key, unsigned& ix, unsigned& iy, unsigned& iz)
{
    ix = iy = iz = 0;
    int st = 0;
#pragma unroll
    for (int level = kMaxLevel - 1; level >= 0; --level)
    {
        const int d   = (int)((key >> (3 * level)) & 7u);
        const int oct = kHilbert.octOf[st][d];
        ix |= unsigned((oct >> 2) & 1) << level;
        iy |= unsigned((oct >> 1) & 1) << level;
        iz |= unsigned(oct & 1) << level;
        st = kHilbert.next[st][oct];
    }
}

// Cornerstone's hilbert3D quantisation: ix = int(floor(x*mx) - xmin*mx), mx = 2^21/lx,
// clamped to [0, 2^21-1].
__device__ inline unsigned quantise(double x, double lo, double il)
{
#pragma clang fp contract(off)
    const double m  = 2097152.0 * il;
    int          ix = (int)(floor(x * m) - lo * m);
    ix = ix > (1 << kMaxLevel) - 1 ? (1 << kMaxLevel) - 1 : ix;
    ix = ix < 0 ? 0 : ix;
    return (unsigned)ix;
}

// Node (start key, level) -> FP64 centre and half-size (hilbertIBox + centerAndSize,
// include/tree.cuh:110-125).
__device__ inline void nodeCentreSize(uint64_t start, int level, const DBox& b, double c[3], double s[3])
{
#pragma clang fp contract(off)
    unsigned ic[3];
    hilbertDecode(start, ic[0], ic[1], ic[2]);
    const unsigned cube = (1u << kMaxLevel) >> level;
    const unsigned mask = ~(cube - 1u);
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const unsigned imin = ic[a] & mask, imax = imin + cube;
        const double   half = 0.5 * (1.0 / 2097152.0) * b.l[a];
        c[a] = b.lo[a] + double(int(imax + imin)) * half;
        s[a] = double(int(imax - imin)) * half;
    }
}

// Number of leading octal digits two 63-bit keys share (0..21).
__device__ inline int commonDigits(uint64_t a, uint64_t b)
{
    const uint64_t x = a ^ b;
    if (x == 0) return kMaxLevel;
    return (__clzll((long long)x) - 1) / 3;
}

// The level-l prefix of a key: its top 3l bits, the rest zero.
__host__ __device__ inline uint64_t keyPrefix(uint64_t key, int level)
{
    return level == 0 ? 0 : (key & ~((uint64_t(1) << (63 - 3 * level)) - 1)) & (kRootRange - 1);
}

__host__ __device__ inline uint64_t levelSpan(int level)   // keys covered by a node at level
{
    return uint64_t(1) << (3 * (kMaxLevel - level));
}

} // namespace csph
