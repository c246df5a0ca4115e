/*
 * tree.hpp -- tree build, neighbour-leaf lists and the physics passes (host entry points
 * into the kernels of tree.hip, jlist.hip, density.hip and force.hip).
 */
#pragma once

#include "prims.hpp"

namespace csph
{

// Box, keys, sort, permutation (x, y, z, h and, if withVelAcc, v and a), leaf tree, node
// geometry, layout and particle -> leaf map for the sets described in s.
// periodicBox: nullptr (open, fitted boxes) or 6 doubles per set.
void buildTree(State& s, bool withVelAcc, const double* periodicBox, SortScratch& sc);

// hmax of the given leaves (leaves == nullptr: all of them); with `state`, also the max h of
// their particles in state != 0 into s.hRemLeaf (the reach of tight lists).
void hmaxLeaves(State& s, const int* leaves, int nLeaves, const uint8_t* state = nullptr);

// hmax per node, leaves first then internal levels deepest first (force upsweep).
void hmaxUpsweep(State& s);

// CSR neighbour-leaf lists of the given leaves (leaves == nullptr: all leaves) into
// s.jCount / s.jOffset / s.jList, indexed by position in the leaf list.  Returns the
// total number of entries.  tight: gather lists trimmed to s.hRemLeaf (same order, same
// accepted pairs for the particles in state != 0).
long long buildLists(State& s, const int* leaves, int nLeaves, bool symmetric, bool tight = false);

} // namespace csph
