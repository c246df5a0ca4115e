#!/usr/bin/env python3
"""Benchmark: particle h-solves/s and pair-interactions/s of the smoothing-length solve on
one MI355X, mixed precision form vs plain FP64, plus the FP64 CPU oracle on a bounded
sample of the same workload.

A step is one complete solve of the workload from the same device-resident inputs:
restore positions and h (device-to-device), then bounding box, Hilbert keys, sort,
permutation, leaf tree, node geometry, and the Newton-Raphson loop (hmax, neighbour-leaf
lists, density kernel, compaction) until every particle has converged -- the reference's
solveDensH (src/density_base.cu:720-1015) in FLAT_PARTICLE mode, density only.  The mixed
form also evaluates rho(h) at the converged h (its closure), so its outputs are consistent.

Before the timed steps the mixed form is autotuned (untimed, --tune-budget seconds, default
45): the library times complete solves of the same inputs over its speed options (loop driver,
leaf bucket, item mapping, persistent grid, graph depth, queue order), keeps the fastest and puts
the inputs back; the chosen options are in the JSON line.  The FP64 path keeps the library's
defaults unless --tune-forms names it.

Other workloads: --config C1 | C2 (single sets), C4 (ensemble of 256 sets solved as one batch,
against 256 sequential plain FP64 solves), C5 (density + force on a 16M-particle Plummer
mixture with velocities, a 10-step drift loop through the reference's call sequence).

Prints ONE JSON line on stdout; progress goes to stderr.

  python bench.py [--steps K] [--warmup W] [--config C3] [--no-cpu] [--tune-budget S]
                  [--opt name=value ...]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cosmosphere_amd import Library, workloads  # noqa: E402
from cosmosphere_amd.api import ACCEL, HSML, POS, VEL  # noqa: E402

METRIC = "particle h-solves/s + pair-interactions/s, mixed vs FP64, 1 MI355X"
FP32_PEAK = 157.3   # TFLOP/s, MI355X vector FP32
FP64_PEAK = 78.6    # TFLOP/s, MI355X vector FP64
# Flop convention of the density pair sum (fixed): 9 per candidate pair (3 sub, r^2, scale),
# 21 more per accepted pair (sqrt, kernel polynomial, accumulation).
FLOPS_CAND, FLOPS_ACC = 9.0, 21.0
# A candidate is a j-particle in a leaf of i's list, for each evaluated i: the lists are the
# reference's, so the count is the algorithm's, whatever the kernel culls before its pair test.
# The kernel's own candidates give the "executed" figure beside it.
CONVENTION_NOTE = ("9 per candidate + 21 per accepted pair; candidates = j-particles of the list "
                   "leaves of each evaluated particle (list_candidates); *_executed: the kernel's "
                   "own candidates after its culling")
# Flop convention of the force pair sum (fixed): 12 per candidate pair (separation, r^2, the two
# support tests), 120 more per accepted pair (unit vector, signal speeds, viscosity, both kernel
# gradients, the i- and j-side force, heating and div v terms)
FORCE_FLOPS_CAND, FORCE_FLOPS_ACC = 12.0, 120.0
# NR work-item options (the library's defaults)
DEFAULT_ITEMS = {"mapping": 0, "sparse": 16, "tail_leaves": 2048}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_config(name):
    if name == "C1":
        return workloads.config1()
    if name == "C2":
        return workloads.config2()
    if name == "C3":
        return workloads.config3()
    if name.startswith("C3:"):
        return workloads.config3(n=int(name[3:]))
    raise SystemExit("unknown config " + name)


def run_precision(lib, c, prec, steps, warmup, opts=(), count_lists=True):
    n = c["n"]
    lib.set_option("precision", prec)
    lib.set_option("max_iter", c.get("max_iter", 100))
    lib.set_option("closure", -1)
    lib.set_option("tree_only", 0)
    for k, v in DEFAULT_ITEMS.items():
        lib.set_option(k, v)
    for k, v in opts:
        lib.set_option(k, v)

    def one():
        lib.restore_inputs()
        return lib.solve_dens_h(n, c["pmass"], 0, False, None, 1e-4, c["hfact"])

    for w in range(warmup):
        t = time.perf_counter()
        it = one()
        log("  warmup %d: %d iterations, %.1f ms" % (w, it, 1e3 * (time.perf_counter() - t)))
    keys = ["nrkern", "launches", "pairs", "candidates", "closure", "closure_pairs", "closure_candidates",
            "jbuild", "compact", "total", "iters", "bbox", "keys", "sort", "permute", "tree", "nodes"]
    acc = {k: 0.0 for k in keys}
    lib.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        it = one()
        for key in keys:
            acc[key] += lib.stat(key)
        log("  step %d: %d iterations" % (k, it))
    lib.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    alive = lib.stat("n_alive")
    if count_lists:
        lc, lc_pairs = list_candidates(lib, one)
    else:   # --no-list-count: the kernel's own candidates stand in
        lc, lc_pairs = acc["candidates"] / steps, acc["pairs"] / steps
    # the convention's candidates are the lists' (every j-particle in the list leaves of each
    # evaluated particle); the kernel's own, after its culling, give the executed figure
    flops = FLOPS_CAND * lc * steps + FLOPS_ACC * acc["pairs"]
    flops_exec = FLOPS_CAND * acc["candidates"] + FLOPS_ACC * acc["pairs"]
    res = {
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "h_solves_per_s": alive * steps / wall,
        "pairs_per_s_kernel": acc["pairs"] / acc["nrkern"] if acc["nrkern"] > 0 else 0.0,
        "pairs_per_s_solve": acc["pairs"] / wall,
        "pairs_per_step": acc["pairs"] / steps,
        "candidates_per_step": acc["candidates"] / steps,
        "list_candidates_per_step": lc if count_lists else None,
        # the counting solve's pairs against the timed steps' (equal in FP64; the mixed form's
        # FP32 partials group differently over the longer lists, so its h may round apart)
        "list_count_pairs_rel_diff": (lc_pairs - acc["pairs"] / steps) / max(acc["pairs"] / steps, 1.0),
        "iterations": acc["iters"] / steps,
        "kernel_launches_per_step": acc["launches"] / steps,
        "kernel_avg_ms": 1e3 * acc["nrkern"] / max(acc["launches"], 1),
        "kernel_tflops": flops / acc["nrkern"] / 1e12 if acc["nrkern"] > 0 else 0.0,
        "kernel_tflops_executed": flops_exec / acc["nrkern"] / 1e12 if acc["nrkern"] > 0 else 0.0,
        "phase_ms": {k: 1e3 * acc[k] / steps for k in ("bbox", "keys", "sort", "permute", "tree", "nodes", "jbuild",
                                                          "nrkern", "compact", "closure", "total")},
        "n_leaves": lib.stat("n_leaves"),
        "stages": streaming_stages(lib, n, acc, steps),
    }
    res["h_sorted"] = lib.array("h")
    res["rho_sorted"] = lib.array("rho")
    return res


HBM_PEAK = 8000.0   # GB/s, MI355X HBM3E


def list_candidates(lib, solve):
    """One more solve, untimed, with the list count on: it builds the reference's (untrimmed) lists
    and counts their candidates for the NR evaluations; returns them and its accepted pairs.  The
    pairs equal the timed steps' in FP64; the mixed form's FP32 partials group differently over
    the longer lists, so its pairs may differ by a convergence flip or two.  The count is a
    separate kernel in each pass, kept out of the timed solves (--no-list-count skips it)."""
    lib.set_option("count_list", 1)
    try:
        solve()
    finally:
        lib.set_option("count_list", 0)
    return lib.stat("list_candidates"), lib.stat("pairs")


def streaming_stages(lib, n, acc, steps):
    """Achieved GB/s of the bandwidth stages of one solve against 8 TB/s, with the algorithmic bytes
    of the fixed conventions: open box 24 B/particle, keys 40, sort 20 (single-pass lower bound),
    permute 68, leaf tree 8 B/particle + 12 B/leaf, linked tree + geometry + maps 64 B/node +
    8 B/leaf + 4 B/particle; per NR pass, hmax 8 B per particle evaluated + 8 B per active leaf,
    j-lists 4 B per entry + 8 B per active leaf + 56 B per node, compaction 12 B per particle
    evaluated.  Times are the library's device-event phases averaged over the timed steps; the
    per-pass counts come from the last solve's pass records (the same inputs every step)."""
    L, nodes = lib.stat("n_leaves"), lib.stat("n_nodes")
    P = lib.passes()
    f = {k: i for i, k in enumerate(Library.PASS_FIELDS)}
    leaves, iterating, entries = P[:, f["leaves"]].sum(), P[:, f["iterating"]].sum(), P[:, f["list_entries"]].sum()
    nb = len(P)
    algo = {
        "bbox": 24.0 * n, "keys": 40.0 * n, "sort": 20.0 * n, "permute": 68.0 * n,
        "tree": 8.0 * n + 12.0 * L, "nodes": 64.0 * nodes + 8.0 * L + 4.0 * n,
        "jbuild": 8.0 * iterating + 8.0 * leaves + 4.0 * entries + 8.0 * leaves + 56.0 * nodes * nb,
        "compact": 12.0 * iterating,
    }
    out = {}
    for k, b in algo.items():
        ms = 1e3 * acc[k] / steps
        gbps = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[k] = {"ms": ms, "algorithmic_bytes": b, "gb_per_s": gbps, "frac": gbps / HBM_PEAK}
    out["jbuild"]["note"] = "hmax + count walk + scan + fill walk over %d passes (latency-bound DFS)" % nb
    return out


def measured_traffic(config, form):
    """HBM bytes per launch of the density kernel for this workload and form, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/density_<config>_<form>_hbm_traffic.md),
    or None when no such measurement exists."""
    import re
    path = os.path.join(ROOT, "profiles", "density_%s_%s_hbm_traffic.md" % (config.lower(), form))
    if not os.path.exists(path):
        return None, None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+-]+)", open(path).read())
    return (float(m.group(1)), os.path.relpath(path, ROOT)) if m else (None, None)


def rel_stats(a, b):
    e = np.abs(a - b) / np.abs(b)
    return {"max": float(e.max()), "p99": float(np.percentile(e, 99)), "p50": float(np.percentile(e, 50)),
            "n_gt_1e-6": int((e > 1e-6).sum()), "n_gt_1e-5": int((e > 1e-5).sum())}


def fp64_closure_reference(lib, c):
    """One untimed plain FP64 solve with the closure on: rho at the converged h, the quantity the
    mixed form returns (the timed FP64 solve returns rho at the h before the last update, as the
    reference does)."""
    lib.set_option("precision", 0)
    lib.set_option("closure", 1)
    lib.restore_inputs()
    lib.solve_dens_h(c["n"], c["pmass"], 0, False, None, 1e-4, c["hfact"])
    lib.set_option("closure", -1)
    return lib.array("h"), lib.array("rho")


def build_oracle(c):
    """The FP64 oracle's tree of the workload (keys, sort, leaf tree), timed."""
    from oracle import oracle as O
    o = O.Oracle(quintic=c["quintic"])
    o.set_particles(c["x"], c["y"], c["z"], c["h"])
    t = time.perf_counter()
    o.build()
    return o, time.perf_counter() - t


def cpu_baseline(c, o, t_tree, gpu_h_fp64, gpu_h_mixed, seconds):
    """FP64 oracle on a bounded sample: all particles of randomly chosen leaves, solved over
    the full tree (in FLAT mode a particle's h depends on positions only, so the sample's h
    equal the full solve's).  Rate = sampled particles / (sample NR time + the tree build's
    share for the sample)."""
    from oracle import oracle as O
    na = o.geti("nAlive")
    pl = o.get("particleLeaf")[:na]
    rng = np.random.default_rng(0)

    def sample(frac):
        pick = rng.choice(na, size=max(1, int(frac * na)), replace=False)
        return np.unique(pl[pick])

    frac = 2e-4
    leaves = sample(frac)
    t = time.perf_counter()
    o.density_solve_leaves(c["pmass"], leaves, 1e-4, c["hfact"], c.get("max_iter", 100))
    t_nr = time.perf_counter() - t
    if t_nr < 0.3 * seconds:
        frac = min(1.0, frac * seconds / max(t_nr, 1e-3))
        leaves = sample(frac)
        t = time.perf_counter()
        o.density_solve_leaves(c["pmass"], leaves, 1e-4, c["hfact"], c.get("max_iter", 100))
        t_nr = time.perf_counter() - t
    lay = o.get("layout")
    idx = np.concatenate([np.arange(lay[l], lay[l + 1]) for l in leaves])
    npart = idx.size
    rate = npart / (t_nr + t_tree * npart / na)
    h_ref = o.get("h")[idx]
    err = {"max_rel_err_h_fp64": float(np.max(np.abs(gpu_h_fp64[idx] - h_ref) / h_ref)) if gpu_h_fp64 is not None else None,
           "max_rel_err_h_mixed": float(np.max(np.abs(gpu_h_mixed[idx] - h_ref) / h_ref)) if gpu_h_mixed is not None else None,
           "p99_rel_err_h_mixed": float(np.percentile(np.abs(gpu_h_mixed[idx] - h_ref) / h_ref, 99))
           if gpu_h_mixed is not None else None,
           "particles": int(npart)}
    err["sample_idx"] = idx
    err["sample_h"] = h_ref
    err["sample_rho"] = o.get("rho")[idx]
    base = {"value": rate, "unit": "particle h-solves/s", "cores": O.num_threads(), "kind": "port",
            "sample": "FP64 oracle (oracle/sph_oracle.cpp, OpenMP) on all %d particles of %d randomly chosen leaves "
                      "(%.3f%% of the live particles), solved to convergence over the full %d-particle tree; tree "
                      "build %.2f s charged pro rata, NR %.2f s" % (npart, leaves.size, 100.0 * npart / na, na, t_tree,
                                                                     t_nr)}
    return base, err


def parity_full(lib, c, o, sample):
    """Full-size parity against the FP64 oracle (the GPU's plain FP64 path, the oracle's tree):
    keys, order, leaf tree and layout bit for bit; the all-leaf gather lists at the GPU's
    converged h (one FP64 solve with the gradient sweep over gather lists, grad_sym_lists=0)
    bit for bit; h and
    rho of the sampled leaves' particles (the cpu_baseline sample), with the particles whose
    Newton iteration stopped one step apart counted separately (|dh/h| above 1e-12: summation
    order alone moves h by rounding, about 1e-15; a |dh/h| < tolh decision that rounding tipped
    moves it by one Newton step, and rho, taken at the h before the last update, further: |drho/rho|
    above 1e-10 marks such a particle too)."""
    t0 = time.perf_counter()
    lib.set_option("precision", 0)
    lib.set_option("closure", -1)
    lib.restore_inputs()
    sym = lib.stat("grad_sym_lists")
    lib.set_option("grad_sym_lists", 0)
    try:
        lib.solve_dens_h(c["n"], c["pmass"], 0, True, None, 1e-4, c["hfact"])
    finally:
        lib.set_option("grad_sym_lists", sym)
    out = {}
    cmp = {"keys": "keys", "order": "order", "leafLevel": "leafLevel", "layout": "layout"}
    for g, r in cmp.items():
        a, b = lib.array(g), o.get(r)
        out[g] = bool(a.size == b.size and np.array_equal(a, b))
    ls, cs = lib.array("leafStart"), o.get("csTree")[:-1]
    out["leafStart"] = bool(ls.size == cs.size and np.array_equal(ls, cs))
    h_gpu = lib.array("h")
    o.set_h_sorted(h_gpu)
    o.hmax_all()
    o.build_lists(symmetric=False)
    jo, jl = lib.array("jOffset"), lib.array("jList")
    ro, rl = o.get("jOffset"), o.get("jList")
    out["gather_lists"] = bool(jo.size == ro.size and np.array_equal(jo.astype(np.int64), ro.astype(np.int64))
                               and np.array_equal(jl, rl))
    out["list_entries"] = int(rl.size)
    out["leaves"] = int(o.geti("nLeaves"))
    idx, h_ref, rho_ref = sample["sample_idx"], sample["sample_h"], sample["sample_rho"]
    h_fp64, rho_fp64 = sample["gpu_h_fp64_sorted"], sample["gpu_rho_fp64_sorted"]
    eh = np.abs(h_fp64[idx] - h_ref) / h_ref
    er = np.abs(rho_fp64[idx] - rho_ref) / rho_ref
    # a particle whose Newton loop stopped one step apart: its h differs by the last step, which
    # can be far below 1e-12 when the loop converges quadratically, but its rho (taken at the h
    # before the last update) differs by the step before, so either difference marks it
    flip = (eh > 1e-12) | (er > 1e-10)
    out["sample_particles"] = int(idx.size)
    out["sample_flips"] = int(flip.sum())
    out["sample_max_rel_err_h_unflipped"] = float(eh[~flip].max()) if (~flip).any() else 0.0
    out["sample_max_rel_err_rho_unflipped"] = float(er[~flip].max()) if (~flip).any() else 0.0
    out["sample_max_rel_err_h"] = float(eh.max())
    out["seconds"] = time.perf_counter() - t0
    out["bit_exact"] = all(out[k] for k in ("keys", "order", "leafLevel", "layout", "leafStart", "gather_lists"))
    return out


def bench_ensemble(args):
    """C4: 256 independent sets solved as one ensemble (mixed form) against the same sets
    solved one after another with the plain FP64 single-set path."""
    t = time.perf_counter()
    sets = workloads.config4()
    counts = [s["n"] for s in sets]
    log("ensemble: %d sets, %d particles (generated in %.1f s)" % (len(sets), sum(counts), time.perf_counter() - t))
    lib = Library("cubic")
    # sequential plain FP64 baseline, solve time only (inputs uploaded before each solve)
    lib.set_option("precision", 0)
    lib.set_option("closure", 0)
    seq = 0.0
    for s in sets:
        n = s["n"]
        lib.arrays_init(n)
        lib.upload(POS, np.stack([s["x"], s["y"], s["z"]]))
        lib.upload(HSML, s["h"])
        lib.sync()
        t0 = time.perf_counter()
        lib.solve_dens_h(n, s["pmass"], 0, False, None, 1e-4, 1.2)
        seq += time.perf_counter() - t0
    log("  sequential fp64: %.1f ms for %d solves" % (1e3 * seq, len(sets)))
    n = lib.batch_init(counts)
    lib.upload(POS, np.concatenate([np.stack([s["x"], s["y"], s["z"]]) for s in sets], axis=1))
    lib.upload(HSML, np.concatenate([s["h"] for s in sets]))
    lib.upload(VEL, np.zeros((3, n)))
    lib.upload(ACCEL, np.zeros((3, n)))
    lib.stash_inputs()
    pm = [s["pmass"] for s in sets]
    out = {}
    tuned = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        lib.set_option("precision", prec)
        lib.set_option("closure", -1)
        for k, v in args.opts:
            lib.set_option(k, v)
        if args.tune_budget > 0:
            # untimed: the autotuner times complete batched solves of these inputs and leaves the
            # fastest options in force (the inputs are put back afterwards)
            lib.restore_inputs()
            tuned[name] = lib.batch_autotune(pm, tolh=1e-4, hfact=1.2, budget_s=args.tune_budget)
            log("  autotune %s: %s" % (name, tuned[name]))
        for _ in range(args.warmup):
            lib.restore_inputs()
            lib.batch_solve(pm, 1e-4, 1.2)
        lib.sync()
        tot = {"pairs": 0.0, "candidates": 0.0, "nrkern": 0.0, "launches": 0.0, "iters": 0.0}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            lib.restore_inputs()
            lib.batch_solve(pm, 1e-4, 1.2)
            for k in tot:
                tot[k] += lib.stat(k)
        lib.sync()
        wall = time.perf_counter() - t0
        alive = lib.stat("n_alive")

        def again():
            lib.restore_inputs()
            lib.batch_solve(pm, 1e-4, 1.2)

        lc = list_candidates(lib, again)[0] if args.count_lists else tot["candidates"] / args.steps
        flops = FLOPS_CAND * lc * args.steps + FLOPS_ACC * tot["pairs"]
        flops_exec = FLOPS_CAND * tot["candidates"] + FLOPS_ACC * tot["pairs"]
        out[name] = {"ms_per_step": 1e3 * wall / args.steps, "h_solves_per_s": alive * args.steps / wall,
                     "pairs_per_s_kernel": tot["pairs"] / max(tot["nrkern"], 1e-12),
                     "kernel_tflops": flops / max(tot["nrkern"], 1e-12) / 1e12,
                     "kernel_tflops_executed": flops_exec / max(tot["nrkern"], 1e-12) / 1e12,
                     "list_candidates_per_step": lc, "candidates_per_step": tot["candidates"] / args.steps,
                     "kernel_avg_ms": 1e3 * tot["nrkern"] / max(tot["launches"], 1),
                     "iterations": tot["iters"] / args.steps}
        log("  batched %s: %.1f ms/step, %.3e h-solves/s" % (name, out[name]["ms_per_step"],
                                                              out[name]["h_solves_per_s"]))
    m = out["mixed"]
    res = {
        "metric": METRIC, "value": m["h_solves_per_s"], "unit": "particle h-solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 pair sums + fp64 Newton", "data": "synthetic",
        "config": {"workload": "C4 ensemble of %d sets, %d particles" % (len(sets), n), "kernel": "cubic",
                   "sets": len(sets), "n": n, "parallelism": "single GPU"},
        "pair_interactions_per_s": m["pairs_per_s_kernel"],
        "sequential_fp64": {"ms": 1e3 * seq, "h_solves_per_s": sum(counts) / seq},
        "batched_fp64": out["fp64"],
        "batched_mixed_vs_sequential_fp64": m["h_solves_per_s"] / (sum(counts) / seq),
        "autotune": tuned,
        "roofline": {"bound": "valu", "achieved": m["kernel_tflops"], "peak": FP32_PEAK, "unit": "TFLOP/s",
                     "frac": m["kernel_tflops"] / FP32_PEAK, "traffic": None, "kernel_avg_ms": m["kernel_avg_ms"],
                     "flops_convention": CONVENTION_NOTE,
                     "achieved_executed": m["kernel_tflops_executed"],
                     "frac_executed": m["kernel_tflops_executed"] / FP32_PEAK},
    }
    if not args.no_cpu:
        log("cpu baseline (oracle on a sample of the sets)")
        res["cpu_baseline"] = ensemble_cpu_baseline(sets, args.cpu_seconds)
    print(json.dumps(res), flush=True)


def ensemble_cpu_baseline(sets, seconds):
    """FP64 oracle on whole sets of the ensemble, drawn at random until about `seconds` of work:
    tree build and FLAT Newton-Raphson solve of each, one after another."""
    from oracle import oracle as O
    rng = np.random.default_rng(0)
    done, parts, t_all = [], 0, 0.0
    for b in rng.permutation(len(sets)):
        s = sets[b]
        o = O.Oracle(quintic=False)
        o.set_particles(s["x"], s["y"], s["z"], s["h"])
        t = time.perf_counter()
        o.build()
        o.density_solve(s["pmass"], 1e-4, 1.2, 100, 0, False)
        t_all += time.perf_counter() - t
        parts += s["n"]
        done.append(int(b))
        if t_all > seconds:
            break
    return {"value": parts / t_all, "unit": "particle h-solves/s", "cores": O.num_threads(), "kind": "port",
            "sample": "FP64 oracle (oracle/sph_oracle.cpp, OpenMP): %d whole sets (%d particles) solved one after "
                      "another in %.2f s" % (len(done), parts, t_all)}


def force_traffic(form):
    """HBM bytes per launch of the force kernel on C5, from the committed rocprofv3 summary, or None."""
    import re
    path = os.path.join(ROOT, "profiles", "force_c5_%s_counters.md" % form)
    if not os.path.exists(path):
        return None, None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+-]+)", open(path).read())
    return (float(m.group(1)), os.path.relpath(path, ROOT)) if m else (None, None)


def force_cpu_baseline(c, st, seconds):
    """FP64 oracle on a bounded sample of one carried C5 step: all particles of randomly chosen
    leaves.  The density solve of the sample starts from the step's carried h (as the GPU's);
    the force pass of the sample runs at the GPU FP64 path's converged h and d(rho)/d(h) (the
    symmetric sums read h_j of neighbours outside the sample).  Rate = sampled particles / (sample
    density time + sample force time + the tree build's share for the sample).  The same sample
    checks the GPU FP64 step against the oracle at full size: h, and the forces and du/dt."""
    from oracle import oracle as O
    n, pmass, hfact = c["n"], c["pmass"], c["hfact"]
    o = O.Oracle(quintic=True)
    o.set_particles(st["x"], st["y"], st["z"], st["h_in"], v=st["v"], a=st["a_in"])
    t = time.perf_counter()
    o.build()
    t_tree = time.perf_counter() - t
    na = o.geti("nAlive")
    pl = o.get("particleLeaf")[:na]
    order = o.get("order")
    rng = np.random.default_rng(0)

    def run(frac):
        leaves = np.unique(pl[rng.choice(na, size=max(1, int(frac * na)), replace=False)])
        t0 = time.perf_counter()
        o.density_solve_leaves(pmass, leaves, 1e-4, hfact, 100)
        t1 = time.perf_counter()
        return leaves, t1 - t0

    frac = 2e-4
    leaves, t_d = run(frac)
    if t_d < 0.1 * seconds:
        frac = min(1.0, frac * 0.3 * seconds / max(t_d, 1e-3))
        leaves, t_d = run(frac)
    lay = o.get("layout")
    idx = np.concatenate([np.arange(lay[l], lay[l + 1]) for l in leaves])
    h_cpu = o.get("h")[idx].copy()
    # the force pass of the sample at the GPU's converged state
    o.set_h_sorted(st["h_out"][order])
    o.set_gradh_sorted(st["gradh_out"][order])
    o.set_thermo(*st["thermo"])
    t0 = time.perf_counter()
    o.force_leaves(pmass, leaves, hfact, c["beta"], c["alphau"], 0, 1, 1)
    t_f = time.perf_counter() - t0
    npart = idx.size
    rate = npart / (t_d + t_f + t_tree * npart / na)
    cal = order[idx]
    eh = np.abs(st["h_out"][cal] - h_cpu) / h_cpu
    flip = eh > 1e-12
    f_gpu = st["f_out"]
    fscale = np.abs(f_gpu[:3]).max()
    err = {"particles": int(npart), "h_flips": int(flip.sum()),
           "max_rel_err_h_unflipped": float(eh[~flip].max()) if (~flip).any() else 0.0}
    for k, nm in enumerate(("fx", "fy", "fz", "du")):
        ref = o.get(nm)[idx]
        scale = fscale if k < 3 else np.abs(f_gpu[3]).max()
        err[nm + "_max_err_rel_to_max"] = float(np.max(np.abs(f_gpu[k][cal] - ref)) / scale)
    base = {"value": rate, "unit": "particle density+force updates/s", "cores": O.num_threads(), "kind": "port",
            "sample": "FP64 oracle (oracle/sph_oracle.cpp, OpenMP) on all %d particles of %d randomly chosen leaves "
                      "(%.3f%% of the live particles) of the first carried step: density solve from the carried h "
                      "%.2f s, force pass %.2f s, tree build %.2f s charged pro rata"
                      % (npart, leaves.size, 100.0 * npart / na, t_d, t_f, t_tree)}
    return base, err


def bench_force(args, n_override=None):
    """C5: density + force on the 16M-particle Plummer mixture with velocities, through the
    reference's call sequence (densityiterate_gpu_c with gradients -> THERMO -> force_gpu_c), and
    a loop of 10 drift steps x += v dt that carries h and the accelerations forward (11 solves).
    Mixed vs plain FP64; per-output errors of the mixed form against the plain FP64 GPU path at
    steps 1 and 10; the FP64 path's step 1 against the FP64 oracle on a sample (cpu_baseline).
    Inputs are uploaded every step into the resident order (incremental re-sort); the library's
    own transfer times (host-timed, PCIe included) are reported beside the resident step time."""
    from cosmosphere_amd.api import DENS_OUT, FORCE_OUT, THERMO
    t = time.perf_counter()
    c = workloads.config5() if n_override is None else workloads.config5(n=n_override)
    n, pmass, hfact, dt = c["n"], c["pmass"], c["hfact"], c["dt"]
    nsteps = c["nsteps"] if args.force_steps is None else args.force_steps
    log("workload %s: %d particles (generated in %.1f s)" % (c["name"], n, time.perf_counter() - t))
    v = np.stack([c["vx"], c["vy"], c["vz"]])
    gamma = c["gamma"]
    lib = Library(c["kernel"])
    lib.arrays_init(n)
    keep_steps = (1, nsteps)
    res, snaps, step1 = {}, {}, None
    pos = np.empty((3, n))
    th = np.empty((4, n))
    th[1], th[2], th[3] = c["cs"], c["alpha"], c["u"]
    for name, prec in (("fp64", 0), ("mixed", 1)):
        if args.only not in ("both", name):
            continue
        lib.set_option("precision", prec)
        lib.set_option("max_iter", 100)
        lib.set_option("closure", -1)
        lib.set_option("tree_only", 0)
        for k, val in args.opts:
            lib.set_option(k, val)
        pos[0], pos[1], pos[2] = c["x"], c["y"], c["z"]
        h = c["h"].copy()
        a = np.zeros((3, n))
        rec = []
        snaps[name] = {}
        for step in range(nsteps + 1):
            x0 = lib.stat("upload_s") + lib.stat("download_s")
            t0 = time.perf_counter()
            lib.upload(POS, pos)
            lib.upload(HSML, h)
            lib.upload(VEL, v)
            lib.upload(ACCEL, a)
            t1 = time.perf_counter()
            lib.densityiterate(n, pmass, tolh=1e-4, hfact=hfact)
            t2 = time.perf_counter()
            dens = {k: lib.stat(k) for k in ("total", "iters", "nrkern", "gradkern", "gjbuild", "jbuild", "pairs",
                                             "candidates", "resort_movers", "sort")}
            hnew = lib.download(HSML, n)[0]
            np.multiply(np.power(hfact / hnew, 3), pmass, out=th[0])
            np.divide((gamma - 1.0) * c["u"], th[0], out=th[0])
            lib.upload_sorted(THERMO, th)
            t3 = time.perf_counter()
            lib.force(n, pmass, beta=c["beta"], alphau=c["alphau"], disc_viscosity=0, pdv_heating=1, shock_heating=1)
            t4 = time.perf_counter()
            frc = {k: lib.stat(k) for k in ("force_total", "force_kernel", "force_jbuild", "force_pairs",
                                            "force_candidates")}
            f = lib.download(FORCE_OUT, n)
            t5 = time.perf_counter()
            xfer = lib.stat("upload_s") + lib.stat("download_s") - x0
            rec.append({"step": step, "density_s": t2 - t1, "force_s": t4 - t3, "step_wall_s": t5 - t0,
                        "transfer_s": xfer, **dens, **frc})
            log("  %s step %d: density %.1f ms (%d iters, re-sort movers %d), force %.1f ms (kernel %.1f ms), "
                "transfers %.1f ms, step with transfers %.1f ms"
                % (name, step, 1e3 * (t2 - t1), dens["iters"], dens["resort_movers"], 1e3 * (t4 - t3),
                   1e3 * frc["force_kernel"], 1e3 * xfer, 1e3 * (t5 - t0)))
            if step in keep_steps:
                d = lib.download(DENS_OUT, n)
                snaps[name][step] = {"h": hnew.copy(), "rho": d[0].copy(), "f": f.copy()}
                if name == "fp64" and step == 1 and not args.no_cpu:
                    step1 = {"x": pos[0].copy(), "y": pos[1].copy(), "z": pos[2].copy(), "h_in": h.copy(),
                             "v": v, "a_in": a.copy(), "h_out": hnew.copy(), "gradh_out": d[1].copy(),
                             "f_out": f.copy(), "thermo": [th[0].copy(), th[1].copy(), th[2].copy(), th[3].copy()]}
            # drift, carrying h and the accelerations into the next solve
            pos[0] += v[0] * dt
            pos[1] += v[1] * dt
            pos[2] += v[2] * dt
            h = hnew
            a = f[:3].copy()
        # the last step's force pass once more (its lists and inputs are still in place, so its
        # pairs are the same), outside the step totals, with the list count on
        if args.count_lists:
            lib.set_option("count_list", 1)
            try:
                lib.force(n, pmass, beta=c["beta"], alphau=c["alphau"], disc_viscosity=0, pdv_heating=1,
                          shock_heating=1)
            finally:
                lib.set_option("count_list", 0)
            again = {k: lib.stat(k) for k in ("force_kernel", "force_pairs", "force_candidates",
                                              "force_list_candidates")}
        else:   # --no-list-count: the last step's own figures, its candidates standing in
            again = {"force_kernel": rec[-1]["force_kernel"], "force_pairs": rec[-1]["force_pairs"],
                     "force_candidates": rec[-1]["force_candidates"],
                     "force_list_candidates": rec[-1]["force_candidates"]}
        tot_d = sum(r["density_s"] for r in rec)
        tot_f = sum(r["force_s"] for r in rec)
        tot_k = sum(r["force_kernel"] for r in rec)
        fl = FORCE_FLOPS_CAND * sum(r["force_candidates"] for r in rec) + FORCE_FLOPS_ACC * sum(r["force_pairs"] for r in rec)
        fl_list = FORCE_FLOPS_CAND * again["force_list_candidates"] + FORCE_FLOPS_ACC * again["force_pairs"]
        carried = rec[1:] if len(rec) > 1 else rec
        res[name] = {
            "particle_updates_per_s": n * len(rec) / (tot_d + tot_f),
            "ms_per_step": 1e3 * (tot_d + tot_f) / len(rec),
            "ms_per_step_with_transfers": 1e3 * sum(r["step_wall_s"] for r in rec) / len(rec),
            "transfer_ms_per_step": 1e3 * sum(r["transfer_s"] for r in rec) / len(rec),
            "ms_per_carried_step": 1e3 * sum(r["density_s"] + r["force_s"] for r in carried) / len(carried),
            "ms_per_carried_step_with_transfers": 1e3 * sum(r["step_wall_s"] for r in carried) / len(carried),
            "resort_movers_carried": [int(r["resort_movers"]) for r in rec[1:]],
            "sort_ms_carried": 1e3 * sum(r["sort"] for r in carried) / len(carried),
            "density_ms_step0": 1e3 * rec[0]["density_s"],
            "density_ms_carried": 1e3 * sum(r["density_s"] for r in rec[1:]) / max(len(rec) - 1, 1),
            "iters_step0": rec[0]["iters"],
            "iters_carried": [int(r["iters"]) for r in rec[1:]],
            "gradient_sweep_ms": 1e3 * sum(r["gradkern"] for r in rec) / len(rec),
            "gradient_lists_ms": 1e3 * sum(r["gjbuild"] for r in rec) / len(rec),
            "nr_kernel_ms_carried": 1e3 * sum(r["nrkern"] for r in rec[1:]) / max(len(rec) - 1, 1),
            "nr_lists_ms_carried": 1e3 * sum(r["jbuild"] for r in rec[1:]) / max(len(rec) - 1, 1),
            "force_candidates_per_pair": sum(r["force_candidates"] for r in rec) / max(sum(r["force_pairs"] for r in rec), 1),
            "force_ms": 1e3 * tot_f / len(rec),
            "force_kernel_ms": 1e3 * tot_k / len(rec),
            "force_pair_interactions_per_s": sum(r["force_pairs"] for r in rec) / max(tot_k, 1e-12),
            "force_kernel_tflops": fl_list / max(again["force_kernel"], 1e-12) / 1e12,
            "force_kernel_tflops_executed": fl / max(tot_k, 1e-12) / 1e12,
            "force_list_count_pass": {"kernel_ms": 1e3 * again["force_kernel"], "pairs": again["force_pairs"],
                                      "candidates": again["force_candidates"],
                                      "list_candidates": again["force_list_candidates"],
                                      "pairs_as_last_step": again["force_pairs"] == rec[-1]["force_pairs"]},
            "density_pair_interactions_per_s": sum(r["pairs"] for r in rec) / max(sum(r["nrkern"] for r in rec), 1e-12),
        }
    out = {"metric": METRIC + " (density + force)", "unit": "particle density+force updates/s", "n_gpus": 1,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic",
           "steps": nsteps + 1, "warmup": 0,
           "config": {"workload": c["name"], "n": n, "kernel": c["kernel"], "solves": nsteps + 1,
                      "dt": dt, "gamma": gamma, "alpha_AV": 1.0, "beta": c["beta"], "parallelism": "single GPU"}}
    head = "mixed" if "mixed" in res else "fp64"
    peak = FP32_PEAK if head == "mixed" else FP64_PEAK
    traffic, traffic_src = force_traffic(head)
    out.update({"value": res[head]["particle_updates_per_s"], "ms_per_step": res[head]["ms_per_step"],
                "dtype": "fp32 pair sums + fp64 Newton" if head == "mixed" else "fp64", head: res[head],
                "pair_interactions_per_s": res[head]["force_pair_interactions_per_s"],
                "roofline": {"bound": "valu", "achieved": res[head]["force_kernel_tflops"], "peak": peak,
                             "unit": "TFLOP/s", "frac": res[head]["force_kernel_tflops"] / peak, "traffic": traffic,
                             "traffic_source": traffic_src,
                             "kernel": ("forceKernelMixed%s (force pair sums)"
                                        % ("Rounds" if lib.stat("force_rounds") else "")) if head == "mixed"
                                       else "forceKernel (force pair sums)",
                             "kernel_avg_ms": res[head]["force_kernel_ms"],
                             "flops_convention": ("12 per candidate + 120 per accepted pair; candidates = "
                                                  "j-particles of the list leaves of each particle, over the "
                                                  "last step's force pass run again with the list count "
                                                  "(force_list_count_pass); *_executed: the kernel's own "
                                                  "candidates after its culling, over every step"),
                             "achieved_executed": res[head]["force_kernel_tflops_executed"],
                             "frac_executed": res[head]["force_kernel_tflops_executed"] / peak}})
    for k in res:
        out[k] = res[k]
    if "fp64" in res and "mixed" in res:
        out["mixed_vs_fp64"] = res["mixed"]["particle_updates_per_s"] / res["fp64"]["particle_updates_per_s"]
        err = {}
        for step in keep_steps:
            a_, b_ = snaps["mixed"][step], snaps["fp64"][step]
            fmax = np.abs(b_["f"][:3]).max()
            e = {"h": rel_stats(a_["h"], b_["h"]), "rho": rel_stats(a_["rho"], b_["rho"])}
            for k, nm in enumerate(("fx", "fy", "fz", "dudt")):
                scale = fmax if k < 3 else np.abs(b_["f"][3]).max()
                d = np.abs(a_["f"][k] - b_["f"][k]) / scale
                e[nm + "_rel_to_max"] = {"max": float(d.max()), "p99": float(np.percentile(d, 99))}
            err["step%d" % step] = e
        out["mixed_vs_fp64_error"] = err
    if step1 is not None:
        log("cpu baseline (oracle sample of the first carried step)")
        base, perr = force_cpu_baseline(c, step1, args.cpu_seconds)
        out["cpu_baseline"] = base
        out["parity_sample_fp64_step1"] = perr
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--only", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--force-steps", type=int, default=None, help="C5: drift steps after the first solve (default 10)")
    ap.add_argument("--tune-budget", type=float, default=None,
                    help="seconds of autotuning per precision form before the timed steps (default: 60 for C4, "
                         "45 for C1-C3, 0 for C5)")
    ap.add_argument("--tune-forms", default="mixed,fp64",
                    help="comma-separated precision forms of a single-set workload to autotune (default: both, so "
                         "mixed_vs_fp64 compares tuned paths)")
    ap.add_argument("--no-parity", action="store_true", help="skip the full-size parity leg of C1-C3")
    ap.add_argument("--no-list-count", dest="count_lists", action="store_false",
                    help="skip the untimed solve (C4: batch solve; C5: force pass) that counts the "
                         "reference lists' candidates; the roofline then counts the kernel's own")
    ap.add_argument("--opt", action="append", default=[], metavar="NAME=VALUE",
                    help="library option (cosmo_ext_set_option), e.g. loop=0 for the host-driven NR loop")
    args = ap.parse_args()
    args.opts = [(kv.split("=")[0], float(kv.split("=")[1])) for kv in args.opt]
    if args.tune_budget is None:
        args.tune_budget = 60.0 if args.config == "C4" else (0.0 if args.config.startswith("C5") else 45.0)
    if args.gpus != 1:
        print("bench.py measures one GPU: this solver runs a particle set (or an ensemble of sets) on one MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.config == "C4":
        bench_ensemble(args)
        return
    if args.config == "C5" or args.config.startswith("C5:"):
        bench_force(args, int(args.config[3:]) if args.config.startswith("C5:") else None)
        return

    t = time.perf_counter()
    c = load_config(args.config)
    log("workload %s: %d particles, %s kernel (generated in %.1f s)" % (c["name"], c["n"], c["kernel"],
                                                                        time.perf_counter() - t))
    lib = Library(c["kernel"])
    n = c["n"]
    lib.arrays_init(n)
    lib.upload(POS, np.stack([c["x"], c["y"], c["z"]]))
    lib.upload(HSML, c["h"])
    lib.upload(VEL, np.zeros((3, n)))
    lib.upload(ACCEL, np.zeros((3, n)))
    lib.stash_inputs()

    res = {}
    tuned = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        if args.only not in ("both", name):
            continue
        log("precision %s" % name)
        if args.tune_budget > 0 and name in args.tune_forms.split(","):
            # untimed: the autotuner times complete solves of these inputs and leaves the fastest
            # options in force (the inputs are put back afterwards)
            lib.set_option("precision", prec)
            lib.set_option("max_iter", c.get("max_iter", 100))
            lib.set_option("closure", -1)
            lib.set_option("tree_only", 0)
            for k, v in DEFAULT_ITEMS.items():
                lib.set_option(k, v)
            for k, v in args.opts:
                lib.set_option(k, v)
            lib.restore_inputs()
            tuned[name] = lib.autotune(n, c["pmass"], tolh=1e-4, hfact=c["hfact"], budget_s=args.tune_budget)
            log("  autotune %s: %s" % (name, tuned[name]))
            args_opts = [(k, tuned[name][k]) for k in Library.TUNE_OPTIONS]
        else:
            args_opts = args.opts
        res[name] = run_precision(lib, c, prec, args.steps, args.warmup, args_opts, args.count_lists)
        r = res[name]
        log("  %s: %.1f ms/step, %.3e h-solves/s, %.3e pairs/s (kernel), kernel %.2f TFLOP/s, iters %.1f"
            % (name, r["ms_per_step"], r["h_solves_per_s"], r["pairs_per_s_kernel"], r["kernel_tflops"],
               r["iterations"]))

    head = res.get("mixed") or res.get("fp64")
    head_name = "mixed" if "mixed" in res else "fp64"
    peak = FP32_PEAK if head_name == "mixed" else FP64_PEAK
    traffic, traffic_src = measured_traffic(args.config, head_name)
    out = {
        "metric": METRIC,
        "value": head["h_solves_per_s"],
        "unit": "particle h-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair sums + fp64 Newton" if head_name == "mixed" else "fp64",
        "data": "synthetic",
        "config": {"workload": c["name"], "n": n, "kernel": c["kernel"], "hfact": c["hfact"], "tolh": 1e-4,
                   "max_iter": c.get("max_iter", 100), "mode": "FLAT_PARTICLE density-only solve",
                   "parallelism": "single GPU"},
        "pair_interactions_per_s": head["pairs_per_s_kernel"],
        "pair_interactions_per_s_solve": head["pairs_per_s_solve"],
        "pairs_per_step": head["pairs_per_step"],
        "candidates_per_step": head["candidates_per_step"],
        "iterations": head["iterations"],
        "phase_ms": head["phase_ms"],
        "stages": head["stages"],
        "options": {k: lib.stat(k) for k in Library.TUNE_OPTIONS},
        "autotune": tuned,
        "roofline": {"bound": "valu", "achieved": head["kernel_tflops"], "peak": peak, "unit": "TFLOP/s",
                     "frac": head["kernel_tflops"] / peak, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "densityKernel (%s NR pair sums)" % head_name,
                     "kernel_avg_ms": head["kernel_avg_ms"],
                     "flops_convention": CONVENTION_NOTE,
                     "achieved_executed": head["kernel_tflops_executed"],
                     "frac_executed": head["kernel_tflops_executed"] / peak},
        "list_candidates_per_step": head["list_candidates_per_step"],
        "list_count_pairs_rel_diff": head["list_count_pairs_rel_diff"],
    }
    if "fp64" in res and "mixed" in res:
        f = res["fp64"]
        out["fp64"] = {"value": f["h_solves_per_s"], "ms_per_step": f["ms_per_step"],
                       "pair_interactions_per_s": f["pairs_per_s_kernel"], "kernel_tflops": f["kernel_tflops"],
                       "roofline_frac": f["kernel_tflops"] / FP64_PEAK,
                       "kernel_tflops_executed": f["kernel_tflops_executed"], "iterations": f["iterations"],
                       "phase_ms": f["phase_ms"]}
        out["mixed_vs_fp64"] = res["mixed"]["h_solves_per_s"] / f["h_solves_per_s"]
        na = int(lib.stat("n_alive"))
        hf, rf = fp64_closure_reference(lib, c)
        hm, rm = res["mixed"]["h_sorted"], res["mixed"]["rho_sorted"]
        # every live particle of the full workload: h and rho(h) of the mixed form against the
        # plain FP64 GPU solve (itself pinned to the oracle by the parity tests); n_gt_1e-5
        # counts convergence flips (a particle stopping one iteration apart)
        out["mixed_vs_fp64_error"] = {"h": rel_stats(hm[:na], hf[:na]), "rho": rel_stats(rm[:na], rf[:na]),
                                      "particles": na}
    if not args.no_cpu:
        log("cpu baseline (oracle sample)")
        o, t_tree = build_oracle(c)
        base, err = cpu_baseline(c, o, t_tree, res.get("fp64", {}).get("h_sorted"),
                                 res.get("mixed", {}).get("h_sorted"), args.cpu_seconds)
        out["cpu_baseline"] = base
        if "fp64" in res and not args.no_parity:
            log("full-size parity (keys, order, tree, gather lists; h and rho of the sample)")
            err["gpu_h_fp64_sorted"] = res["fp64"]["h_sorted"]
            err["gpu_rho_fp64_sorted"] = res["fp64"]["rho_sorted"]
            out["parity_full"] = parity_full(lib, c, o, err)
            log("  %s" % json.dumps(out["parity_full"]))
        out["parity_sample"] = {k: v for k, v in err.items() if not isinstance(v, np.ndarray)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
