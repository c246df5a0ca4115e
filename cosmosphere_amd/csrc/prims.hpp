/*
 * prims.hpp -- device-wide scan and radix sort (prims.hip).
 */
#pragma once

#include "state.hpp"

namespace csph
{

struct PassCtl;

struct SortScratch
{
    DevBuf hist, offs, scan, counter;
    int*   err = nullptr;   // device counter of out-of-range scatters (must stay 0)
};

size_t scanScratchBytes(int n);

// out[0..n] = exclusive prefix sums of in[0..n), out[n] = total.  If total64 is given,
// the total is also added to it as a 64-bit sum (overflow check).
// With ctl (device-driven pass) n is the capacity: entries from ctl->nAL on count as zero,
// out[ctl->nAL] = total, and nothing runs when ctl->skip.
template<class T>
void scanExclusive(const T* in, T* out, int n, DevBuf& scratch, unsigned long long* total64, hipStream_t st,
                   const PassCtl* ctl = nullptr);

// Stable ascending sort of (keys, vals) over bits [0, bits).  Returns true when the
// sorted result sits in keysAlt / valsAlt rather than keys / vals.  Host-synchronising
// (one 8-byte read of the key spread).
bool radixSortPairs(uint64_t* keys, uint64_t* keysAlt, uint32_t* vals, uint32_t* valsAlt, int n, int bits,
                    SortScratch& sc, hipStream_t st);

void iota(uint32_t* v, int n, hipStream_t st);

// Fill with a kernel rather than hipMemsetAsync, so the call is safe inside graph capture.
void fillU32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);

} // namespace csph
