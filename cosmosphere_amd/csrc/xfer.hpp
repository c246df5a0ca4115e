/*
 * xfer.hpp -- host <-> device transfers of the slot interface (xfer.hip).
 *
 * A slot is a fixed group of per-particle FP64 fields (the numbering Phantom's ibun_* uses,
 * reference include/arrays.hpp:15-25).  The caller's block for a slot is ncomp runs of n
 * doubles; on the device each component is its own array in the solve's (Hilbert) order.
 *
 * Transfers stream through a small ring of page-locked chunks: host worker threads copy
 * the caller's pageable memory into one chunk while the DMA engine moves the previous one,
 * and every component of a slot is gathered into (or scattered out of) the device order by
 * one kernel.  A call returns when its data has arrived.
 *
 * Resident order.  The ten solve inputs (x, y, z, h, v, a) carry a tag: kTagCaller (the
 * device array holds the caller's order) or kTagAligned (it holds the order the last solve
 * left, so entry i belongs to caller index order[i]).  cosmo_upload of an input slot lands
 * aligned when the last solve was a single set of the same size, which lets the next solve
 * re-sort the nearly sorted keys incrementally instead of from scratch.
 */
#pragma once

#include "state.hpp"

namespace csph
{

// Input fields that the tree build permutes, in State::inTag order.
constexpr int kNumInputs = 10;

struct SlotDesc
{
    int ncomp;
    int firstInput;              // index of the slot's first field in State::inTag, -1 for outputs
    DevBuf State::* comp[6];     // the device arrays, in the caller's component order
};

// The slot's descriptor, or nullptr for a number no slot has.
const SlotDesc* slotDesc(int slot);

// Host block -> device.  aligned: gather through `order` into the device order (required for
// cosmo_upload_sorted; chosen for input slots by cosmo_upload when the order is resident).
void transferIn(State& s, const SlotDesc& d, const double* host, int n, bool aligned);

// Device -> host block, scattered through `order` into the caller's order.
void transferOut(State& s, const SlotDesc& d, double* host, int n);

// Device -> host block like transferOut, allocating nothing: the scatter goes through a static
// device buffer of kNoAllocMaxParticles doubles per GPU (n must not exceed it).
constexpr int kNoAllocMaxParticles = 4194304;
void transferOutNoAlloc(State& s, const SlotDesc& d, double* host, int n);

// True when cosmo_upload of an input slot should land aligned (resident order on, an order
// for exactly these n particles from a single-set solve).
bool residentEligible(const State& s, int n);

// Before a solve: makes the input tags consistent and returns true when the solve can
// re-sort incrementally (x, y, z, h aligned, one set).  The other inputs are brought to
// the form the permutation expects: aligned when they are permuted (withVelAcc), the
// caller's order when they are not.  Returns false after bringing every input back to the
// caller's order.
bool prepareInputs(State& s, bool withVelAcc);

// Every input back in the caller's order (the stash of repeated benchmark solves).
void inputsToCallerOrder(State& s);

// After a solve: what is left on the device counts as the caller's order for the next one
// (the reference's convention for fields not uploaded again).
void resetInputTags(State& s);

} // namespace csph
