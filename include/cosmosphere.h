/*
 * cosmosphere.h -- C ABI of the MI355X (gfx950) SPH density / smoothing-length and force
 * solver.  Drop-in for the reference's libcosmoSPHere.a: the ten entry points below keep the
 * reference's names, signatures, argument meaning, slot numbering and abort-on-misuse
 * behaviour, so a Fortran `bind(C)` interface written for the reference links unchanged.
 *
 * Build products (cosmosphere_amd/lib/):
 *   libcosmoSPHere.a / .so           M4 cubic kernel   (reference KERNEL=cubic)
 *   libcosmoSPHere_quintic.a / .so   M6 quintic kernel (reference KERNEL=quintic)
 *
 * Conventions (as the reference):
 *   - single device, single host thread, not re-entrant; every call returns when its work
 *     is complete;
 *   - misuse (wrong particle count, unknown slot, no ordering yet, force before density)
 *     and unconverged solves print a "FATAL: ..." line on stderr and abort();
 *   - particle data stay on the device between calls, in Hilbert order; transfers gather
 *     on the way in and scatter on the way out, so the caller only ever sees its own order;
 *   - a slot's host block is ncomp contiguous runs of n doubles.
 */
#ifndef COSMOSPHERE_H
#define COSMOSPHERE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Slot ids, shared with Phantom's ibun_* (reference include/arrays.hpp:15-25). */
enum CosmoSlot
{
    COSMO_POS       = 1, /* x, y, z                               */
    COSMO_HSML      = 2, /* h                                     */
    COSMO_VEL       = 3, /* vx, vy, vz                            */
    COSMO_ACCEL     = 4, /* ax, ay, az                            */
    COSMO_DENS_OUT  = 5, /* rho, d(rho)/d(h)                      */
    COSMO_GRAD_OUT  = 6, /* div v, xi, d(div v)/dt                */
    COSMO_THERMO    = 7, /* p/rho^2, c_s, alpha_AV, u             */
    COSMO_FORCE_OUT = 8  /* fx, fy, fz, du/dt, vsigmax, div v     */
};

/* ---- the ten reference entry points ---------------------------------------------------- */

/* Density / smoothing-length Newton-Raphson solve plus the gradient sweep, on the POS, HSML,
 * VEL and ACCEL slots previously uploaded in the caller's order.  box = {xmin, xmax, ymin,
 * ymax, zmin, zmax}, read only if periodic != 0.  hfact <= 0 selects the kernel default.
 * Replaces reference src/dens_c_api.cu:47-124 (densityiterate_gpu_c). */
void densityiterate_gpu_c(int n, double pmass, int periodic, const double* box, double tolh, double hfact);

/* Support radius of the compiled kernel in units of h: 2 (cubic) or 3 (quintic).
 * Replaces reference src/dens_c_api.cu:128-131. */
double cosmo_kernel_radius(void);

/* Force pass on the tree and h of the last density solve for this n; THERMO (and, for the
 * corrector, VEL) uploaded with cosmo_upload_sorted.  Leaves FORCE_OUT on the device.
 * Replaces reference src/force_c_api.cu:26-82 (force_gpu_c). */
void force_gpu_c(int n, double pmass, double beta, double alphau, int disc_viscosity, int pdv_heating,
                 int shock_heating);

/* Size the particle-length device arrays for n particles (n <= 0: no-op; idempotent).
 * Replaces reference src/arrays_c_api.cu:30-35. */
void cosmo_arrays_init(int n);

/* Components of a slot: 1->3, 2->1, 3->3, 4->3, 5->2, 6->3, 7->4, 8->6, other->0.
 * Replaces reference src/arrays_c_api.cu:145-155. */
int cosmo_slot_ncomp(int slot);

/* Send a slot in the caller's order (before the solve that sorts it).
 * Replaces reference src/arrays_c_api.cu:157-161. */
void cosmo_upload(int slot, const double* host, int n);

/* Send a slot and gather it into the device's order; aborts if no solve has built one.
 * Replaces reference src/arrays_c_api.cu:163-167. */
void cosmo_upload_sorted(int slot, const double* host, int n);

/* Fetch a slot back in the caller's order; needs an ordering.
 * Replaces reference src/arrays_c_api.cu:169-193. */
void cosmo_download(int slot, double* host, int n);

/* Best-effort page-locking hints; never fail.  Replaces reference src/pin_c_api.cu:22-34
 * (the registration itself is not performed on this platform). */
void cosmo_pin_host(void* ptr, size_t nbytes);
void cosmo_unpin_host(void* ptr);

/* ---- extensions (no counterpart in the reference's C ABI) ------------------------------- */

/* cosmo_download that allocates nothing.  No counterpart in the reference: for callers that
 * cannot allocate (HIP graph capture, serving loops, harnesses that forbid allocation).  Each
 * component is scattered into a static device buffer of COSMO_NOALLOC_MAX_PARTICLES doubles
 * (32 MiB, one per GPU, held until the process ends) and copied from it into host.  Returns 0,
 * or -1 with a message on stderr and nothing copied when n exceeds COSMO_NOALLOC_MAX_PARTICLES;
 * misuse aborts as cosmo_download.  Calls on one GPU share the buffer: calls that overlap
 * (several host threads) race. */
#define COSMO_NOALLOC_MAX_PARTICLES 4194304
int cosmo_download_noalloc(int slot, double* host, int n);

/* The reference's C++ solveDensH(n, pmass, mode, withGradients, periodicBox, tolh, hfact)
 * (include/density.hpp:85-91) with C linkage: mode 0 = FLAT_PARTICLE, 1 = WARP_PER_LEAF.
 * Returns the Newton iterations run. */
int cosmo_ext_solve_dens_h(int n, double pmass, int mode, int withGradients, int periodic, const double* box,
                           double tolh, double hfact);

/* Options: "precision" (0 = plain FP64, 1 = mixed FP32-offset form), "max_iter",
 * "loop" (NR passes: 0 host-driven, 1 device-driven, the default: counts in device memory,
 * "graph_passes" passes per HIP graph replay when "graph" is 1, persistent density grid of
 * "loop_blocks_scale" / 4 times the resident blocks, cost-ordered work queue when "lpt" is 1),
 * "bucket" (leaf bucket size, default 64), "units" (1, the default: NR whole-leaf items (mixed, and
 * FP64 with culling on) cut each particle's tile into work units spread evenly over the 64 lanes
 * where one particle per lane group would leave lanes idle; 0: one group of floor(64 / c) lanes
 * per particle), "cull" (1, the default: NR whole-leaf items stage only the list entries some
 * particle's own support reaches; 0: every entry of the tight list),
 * "mapping" (NR work items: 0 auto = one wavefront per leaf, one per particle once at most
 * "sparse" particles of the leaf remain; 1 = always one per leaf; 2 = always one per
 * particle), "sparse" (default 16), "closure" (-1 auto = on for mixed, 0 off, 1 on: rho and
 * d(rho)/d(h) re-evaluated at the converged h),
 * "tree_only" (1: stop after the tree build, for verification of the tree stages),
 * "list_heavy_steps" and "list_expand_leaves" (verification of the block walk of long lists:
 * walk steps per 32768 lists before a list is handed to a whole block, default 256; nodes with
 * more leaves than this are expanded into separate walks, default 256), "split_len" (device-driven
 * loop: a heavy work item whose list has more entries is split over several wavefronts, default
 * 2048; 0: never), "sweep_split" (gradient sweep and force pass: the particles of a leaf whose
 * list has more entries go one per wavefront, default 512; 0: never), "list_heavy_factor" (a
 * walk also goes to the block walk past this many times the previous build's typical walk,
 * default 8; 0: the launch-size rule only), "tile_depth" (device-driven loop: j-slots per LDS
 * tile of whole-leaf items, 64 or 128, the default), "grad_sym_lists" (1, the default: the
 * gradient sweep runs over the symmetric lists, which the force pass of the same solve reuses;
 * 0: over gather lists), "force_rounds" (mixed force pass: 1, the default, spreads each tile's
 * accepted pairs evenly over the lanes; 0: lanes keep their own particle), "resident" (1, the
 * default: uploads after a single-set solve land in its Hilbert order and the next solve
 * re-sorts incrementally), "xfer_threads" (host threads of the transfer pool, default 8).
 * The environment variable COSMO_PRECISION=mixed|fp64 sets the precision for
 * densityiterate_gpu_c. */
void cosmo_ext_set_option(const char* name, double value);

/* Per-phase device seconds and work counters of the last solve / force pass. */
double cosmo_ext_get_stat(const char* name);

/* Copy a named device array to host memory (at most maxBytes); returns its element count,
 * -1 for an unknown name.  For verification. */
long long cosmo_ext_get_array(const char* name, void* dst, long long maxBytes);

/* Keep / restore a device copy of the uploaded POS, HSML, VEL and ACCEL (repeated solves
 * from identical inputs). */
void cosmo_ext_stash_inputs(void);
void cosmo_ext_restore_inputs(void);
void cosmo_ext_sync(void);

/* Ensemble of independent particle sets solved together (one sequence of kernel launches for
 * all sets; each set has its own root box, tree, particle mass and convergence).  The sets lie
 * one after another in every slot: set b owns entries [counts[0] + ... + counts[b-1],
 * + counts[b]) of the caller's order.  cosmo_batch_init sizes the device arrays for the total;
 * upload and download then use the slot calls with n = total.  cosmo_batch_solve runs the
 * density-only Newton-Raphson solve (FLAT_PARTICLE) with pmass[b] for set b; boxes holds six
 * doubles per set when periodic != 0.  Returns the Newton iterations run. */
void cosmo_batch_init(int nSets, const int* counts);
int cosmo_batch_solve(const double* pmass, int periodic, const double* boxes, double tolh, double hfact);

/* Autotuner: times complete density-only solves (FLAT_PARTICLE) of the uploaded inputs over the
 * options the results do not depend on -- NR loop driver ("loop"), leaf bucket size ("bucket"),
 * wave-per-leaf vs flat item mapping ("mapping", "sparse", "tail_leaves"), persistent grid size
 * ("loop_blocks_scale"), passes per graph replay ("graph_passes"), heavy items first ("lpt"),
 * per-particle culling of list entries ("cull"), lanes over work units ("units"), j-tile depth
 * ("tile_depth") -- one option at a time, and leaves the fastest in force.  The uploaded inputs are put back
 * afterwards (a following solve starts from them; download nothing before it).  budgetSeconds
 * bounds the search.  A workload with the same signature (particle count, set sizes, kernel,
 * precision, tolerances) reuses the earlier choice (also across processes through the file named
 * by COSMO_TUNE_CACHE).  Returns the trial solves run (0: reused).  The tree and lists are the
 * reference's at bucket 64 only; h and rho do not depend on the bucket beyond summation order. */
int cosmo_ext_autotune(int n, double pmass, int periodic, const double* box, double tolh, double hfact,
                       double budgetSeconds);
int cosmo_batch_autotune(const double* pmass, int periodic, const double* boxes, double tolh, double hfact,
                         double budgetSeconds);

#ifdef __cplusplus
}
#endif

#endif /* COSMOSPHERE_H */
