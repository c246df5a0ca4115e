/*
 * tune.hpp -- the autotuner of the smoothing-length solve (tune.hip).
 */
#pragma once

#include <functional>
#include <string>

#include "state.hpp"

namespace csph
{

// The speed options a solve's results do not depend on (up to the summation order of the
// pair sums; the tree and lists are the reference's at bucket 64 only).
struct TuneParams
{
    int loop;          // NR passes: 0 host-driven, 1 device-driven
    int bucket;        // leaf bucket size
    int mapping;       // NR work items: 0 auto, 1 whole leaves, 2 per particle
    int sparse;        // auto mapping: per-particle items when at most this many remain in a leaf
    int tailLeaves;    // ... and at most this many leaves are active
    int blockScale;    // device loop: persistent density grid, resident blocks x scale / 4
    int graphPasses;   // device loop: passes per graph replay
    int lpt;           // device loop: heaviest items first in the work queue
    int cull;          // NR whole-leaf items: per-particle culling of list entries
    int units;         // mixed NR whole-leaf items: lanes over work units
    int tileDepth;     // device loop: j-slots per LDS tile of whole-leaf items (64 or 128)
};

TuneParams currentParams(const State& s);
void       applyParams(State& s, const TuneParams& p);
std::string describe(const TuneParams& p);

struct TuneResult
{
    TuneParams best;
    double     bestSeconds = 0.0;      // solve time with the chosen options
    double     defaultSeconds = 0.0;   // ... with the options in force when tuning started
    double     tuneSeconds = 0.0;      // wall time of the whole search
    int        trials = 0;
    bool       cached = false;         // taken from an earlier search of the same signature
};

// Coordinate search over TuneParams: each candidate is timed by complete solves (solve(),
// after the inputs are put back), the best of two; the search stops when no single-option
// change improves the time or the budget is spent.  The inputs as uploaded (caller order) are
// kept before the first trial and put back after the last, so the caller's next solve starts
// from them.  `signature` names the workload: a repeated signature reuses the earlier choice.
TuneResult autotune(State& s, const std::string& signature, const std::function<void()>& solve,
                    double budgetSeconds);

} // namespace csph
