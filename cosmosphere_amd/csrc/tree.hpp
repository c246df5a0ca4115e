/*
 * tree.hpp -- tree build, neighbour-leaf lists and the physics passes (host entry points
 * into the kernels of tree.hip, jlist.hip, density.hip and force.hip).
 */
#pragma once

#include "prims.hpp"

namespace csph
{

struct PassCtl;

// Block walks of long lists: tokens the count pass keeps for the fill pass (16M, 128 MiB).
constexpr int kListTokPool = 1 << 24;
constexpr int kListSlotTok = 256;  // light lists: tokens the count walk keeps per list for the fill

// Box, keys, sort, permutation (x, y, z, h and, if withVelAcc, v and a), leaf tree, node
// geometry, layout and particle -> leaf map for the sets described in s.
// periodicBox: nullptr (open, fitted boxes) or 6 doubles per set.
// incremental: the inputs are in the previous solve's order (x, y, z, h and, with withVelAcc,
// v and a), and the previous tree is still in s; the keys are then re-sorted incrementally
// (resort.hip) with the same result as the full sort.
void buildTree(State& s, bool withVelAcc, const double* periodicBox, SortScratch& sc, bool incremental = false);

// Incremental re-sort of keys given in the previous solve's order (s.order, the previous
// tree's leaves and particle -> leaf map): sorted keys, the input position of each sorted
// entry and its caller index (the new order).  Returns the number of particles that left
// their previous leaf's key range (sorted from scratch).
int resortIncremental(State& s, const uint64_t* keys, uint64_t* keysOut, uint32_t* idxOut, int* orderOut,
                      SortScratch& sc);

// hmax of the given leaves (leaves == nullptr: all of them); with `state`, also the max h of
// their particles in state != 0 into s.hRemLeaf (the reach of tight lists).
// With ctl (device-driven pass) nLeaves is the capacity and ctl->nAL the count.
void hmaxLeaves(State& s, const int* leaves, int nLeaves, const uint8_t* state = nullptr, const PassCtl* ctl = nullptr);

// hmax per node, leaves first then internal levels deepest first (force upsweep).
void hmaxUpsweep(State& s);

// CSR neighbour-leaf lists of the given leaves (leaves == nullptr: all leaves) into
// s.jCount / s.jOffset / s.jList, indexed by position in the leaf list.  Returns the
// total number of entries.  tight: gather lists trimmed to s.hRemLeaf (same order, same
// accepted pairs for the particles in state != 0).
long long buildLists(State& s, const int* leaves, int nLeaves, bool symmetric, bool tight = false);

// Device-driven pass: the count walk (jList == nullptr: counts into jCount) or the fill walk
// (into jList at jOffset) of the tight gather lists of ctl->nAL leaves; nothing when
// ctl->skip.  nCap bounds ctl->nAL.
void launchGatherListsDevice(State& s, const int* leaves, int nCap, uint32_t* jCount, const uint32_t* jOffset,
                             int* jList, const PassCtl* ctl);

} // namespace csph
