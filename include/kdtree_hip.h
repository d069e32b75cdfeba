/*
 * kdtree_hip.h -- C ABI of the MI355X (gfx950) kD-tree Barnes-Hut step.
 *
 * Drop-in boundary for the reference's Rust crate rust_kdtree_nbody
 * (Parallel/RustVersion/src/lib.rs:1-3).  The reference exposes a Rust library
 * API and no FFI; every entry point below names the reference function whose
 * work it replaces.  Plain C types only: host pointers, int64 sizes, int status.
 *
 * Conventions
 *  - Every function returns KD_OK (0) or a negative KD_ERR_* code; the message
 *    for the last failure on the calling thread is kd_last_error().
 *  - A kd_ctx owns the device buffers for up to cap_n bodies and one HIP
 *    stream.  One context must not be used from two threads at once; distinct
 *    contexts are independent.
 *  - Bodies are passed as structure-of-arrays in the caller's (original)
 *    order; every result comes back in that order.
 *  - Host arrays are copied; device-pointer entry points (suffix _device) read
 *    device memory that the caller keeps alive for the duration of the call.
 */
#ifndef KDTREE_HIP_H
#define KDTREE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KD_OK 0
#define KD_ERR_ARG (-1)        /* bad argument (null pointer, n out of range, ...) */
#define KD_ERR_CAPACITY (-2)   /* n above the context capacity / output too small */
#define KD_ERR_HIP (-3)        /* a HIP runtime call failed (message has details) */
#define KD_ERR_NOMEM (-4)      /* device or host allocation failed */
#define KD_ERR_STATE (-5)      /* call out of order (e.g. step before upload) */
#define KD_ERR_NONFINITE (-6)  /* NaN or Inf in positions or masses */
#define KD_ERR_NODEVICE (-7)   /* no usable gfx950 device */

/* Arithmetic of the force traversal. */
#define KD_PREC_FP64 0         /* reference arithmetic: FP64, sqrt then divide, no FMA */
#define KD_PREC_MIXED 1        /* FP32 far-field and P2P, compensated FP32 sums,
                                  FP64 state, integrator and node moments */
#define KD_PREC_MIXED_D64 2    /* as MIXED but coordinate differences taken in FP64 */

/* Node-vector layouts of the reference. */
#define KD_LAYOUT_COMPACT 0    /* Sequential/RustVersion/src/array_kd_tree.rs:128-134 */
#define KD_LAYOUT_PADDED 1     /* Parallel/RustVersion/src/array_kd_tree.rs:565-581 */

typedef struct kd_params {
  double theta;      /* opening angle; THETA = 0.3 at array_kd_tree.rs:15 */
  double dt;         /* time step; dt = 1e-3 at Parallel/RustVersion/src/main.rs:23 */
  int32_t leaf_size; /* MAX_PARTS: 7 (Sequential :6) or 8 (Parallel :14); 2..32 */
  int32_t precision; /* KD_PREC_* */
  int32_t count_interactions; /* nonzero: kernels also count tests/M2P/P2P */
  int32_t profile;   /* nonzero: time build / traversal / kick-drift with HIP
                        events on the context's stream (see kd_timings) */
} kd_params;

typedef struct kd_ctx kd_ctx;

/* Flat view of a node vector in one of the reference layouts
 * (KDTree enum, Sequential/RustVersion/src/array_kd_tree.rs:10-26).  Leaves
 * list their members as perm[first .. first+num_parts); perm holds original
 * body indices in particle order (leaves contiguous, depth-first, members in
 * ascending original index).  Unused padded slots are leaves with
 * num_parts == 0, as allocate_node_vec fills them. */
typedef struct kd_tree_view {
  int64_t n_nodes;     /* in: capacity of every node array; out: nodes written */
  int32_t *is_leaf;
  int32_t *split_dim;
  int64_t *num_parts;
  int64_t *first;
  double *split_val;
  double *m;
  double *cmx, *cmy, *cmz;
  double *size;
  int64_t *left, *right;
  int64_t *perm;       /* n entries */
} kd_tree_view;

const char *kd_last_error(void);
/* Library version string and the compiled target ("gfx950"). */
const char *kd_version(void);

/* theta 0.3, dt 1e-3, leaf 7, FP64, no counting. */
int kd_default_params(kd_params *out);

/* Context with capacity for cap_n bodies on the current HIP device.
 * Replaces the per-call allocations of simple_sim (acc, tree, indices:
 * Parallel/RustVersion/src/array_kd_tree.rs:624-630).  KD_ERR_CAPACITY when
 * the tree of cap_n bodies at params->leaf_size would exceed 2^26 - 1 heap
 * slots (about leaf_size * 2^25 bodies); kd_set_params and kd_batch_create
 * (per system) check the same limit. */
int kd_create(int64_t cap_n, const kd_params *params, kd_ctx **out);
int kd_destroy(kd_ctx *ctx);
int kd_set_params(kd_ctx *ctx, const kd_params *params);
int kd_get_params(const kd_ctx *ctx, kd_params *out);

/* Copy n bodies in (the caller's Vec<Particle>, array_particle.rs:3-8).
 * -0.0 coordinates are stored as +0.0; NaN/Inf positions or masses and
 * non-positive masses are rejected with KD_ERR_NONFINITE / KD_ERR_ARG. */
int kd_upload(kd_ctx *ctx, int64_t n, const double *x, const double *y,
              const double *z, const double *vx, const double *vy,
              const double *vz, const double *m);
int kd_upload_device(kd_ctx *ctx, int64_t n, const double *x, const double *y,
                     const double *z, const double *vx, const double *vy,
                     const double *vz, const double *m);
/* Copy the current state out (any pointer may be NULL to skip it). */
int kd_download(kd_ctx *ctx, double *x, double *y, double *z, double *vx,
                double *vy, double *vz);
int kd_download_device(kd_ctx *ctx, double *x, double *y, double *z,
                       double *vx, double *vy, double *vz);
int64_t kd_size(const kd_ctx *ctx);

/* Build the tree of the current positions: build_tree
 * (Sequential/RustVersion/src/array_kd_tree.rs:49) / build_tree_par4
 * (Parallel/RustVersion/src/array_kd_tree.rs:515), including node moments. */
int kd_build(kd_ctx *ctx);

/* Export the last built tree in a reference layout (allocate_node_vec +
 * build_tree's node vector, array_kd_tree.rs:41-46 / :55-60).  If
 * view->n_nodes is smaller than needed, returns KD_ERR_CAPACITY and stores
 * the required count in view->n_nodes. */
int kd_export_tree(kd_ctx *ctx, int32_t layout, kd_tree_view *view);
/* Node-vector length the reference would allocate for n bodies. */
int64_t kd_nodes_needed(int64_t n, int32_t leaf_size, int32_t layout);

/* calc_accel(p, particles, nodes) over a caller-supplied node vector in one
 * of the reference layouts (array_kd_tree.rs:619 / Sequential :174): the
 * tree is not rebuilt, it is traversed as given (FP64 reference arithmetic).
 * Leaf members are tree->perm[first .. first+num_parts) (body indices).
 * Computes the k bodies idx[0..k) (idx NULL: all n bodies in order).
 * Stateless: host arrays in and out, on the current HIP device.  A tree
 * whose indices leave their arrays is rejected with KD_ERR_ARG. */
int kd_accel_tree(int64_t n, const double *x, const double *y, const double *z,
                  const double *m, const kd_tree_view *tree, double theta, int64_t k,
                  const int64_t *idx, double *ax, double *ay, double *az);

/* print_tree (Parallel/RustVersion/src/array_kd_tree.rs:666-692): write a node
 * vector to `path` in the reference's text format (read by
 * TreeVisualizer/src/main/scala/ViewTrees.scala:32-80): the node count, then
 * per node "I split_dim split_val left right" or "L k" and k lines "x y z" of
 * the members' positions (x, y, z indexed by body).  Numbers as Rust's `{}`
 * prints them (shortest round-trip decimal, no exponent).  Host only. */
int kd_print_tree(const kd_tree_view *tree, int64_t n, const double *x, const double *y,
                  const double *z, const char *path);

/* Build, then accelerations of every body: calc_accel(p, ..) for all p
 * (array_kd_tree.rs:619 / Sequential :174), in original order. */
int kd_accel(kd_ctx *ctx, double *ax, double *ay, double *az);
int kd_accel_device(kd_ctx *ctx, double *ax, double *ay, double *az);

/* steps x (build, accel, kick-drift v += dt*a; p += dt*v): simple_sim
 * (Parallel/RustVersion/src/array_kd_tree.rs:623-664). */
int kd_step(kd_ctx *ctx, int64_t steps);
/* Wait for the context's stream. */
int kd_sync(kd_ctx *ctx);

/* Counters of the last traversal (needs params.count_interactions):
 * out[0] node opening tests, out[1] accepted far-field interactions, out[2]
 * P2P pairs, each summed over bodies (the reference's per-body work); out[3]
 * node visits and out[4] leaf visits summed over wavefronts (the work the
 * 64-wide wavefronts actually issue). */
int kd_counters(kd_ctx *ctx, int64_t out[5]);

/* Launch-shape knobs of the kernels (no counterpart in the reference):
 * "leaf_predicated" 0/1 -- compute whole leaf batches with predication
 * instead of branching per body; "fp64_batched_walk" 0/1 -- run the FP64
 * path on the batched walk of the mixed paths; "lazy_dimension" 0/1 (default
 * 1) -- let the build skip the sort and the device-wide partitions of a
 * dimension that the previous full build found could not be split at those
 * levels (a thin system's third dimension); every build checks that this
 * holds and builds again in full when it does not, so trees are the same
 * either way; "fused_kick_drift" 0/1 (default 1) -- kd_step's walk applies
 * the kick-drift to each body as it finishes it, with no acceleration arrays
 * and no separate pass (steps with params.count_interactions set keep the
 * separate pass), bit-identical either way; "mid_levels" 0/1 (default 1) --
 * when a build starts from the previous tree order, its device-wide levels
 * whose segments hold at most 65,536 bodies (for a single system from level
 * 8 on) run as one launch of one workgroup per segment, with the same trees
 * (batches always do so from tree order).  Unknown keys
 * return KD_ERR_ARG. */
int kd_set_tuning(kd_ctx *ctx, const char *key, int64_t value);

/* Reads a knob of kd_set_tuning, or the build's state: "lazy_next" (the
 * dimension the next build skips, 3 for none), "lazy_builds" (builds that
 * skipped one), "lazy_redos" (of those, builds repeated in full) and
 * "mid_builds" (builds that ran workgroup levels).  No counterpart in the
 * reference. */
int kd_get_tuning(kd_ctx *ctx, const char *key, int64_t *value);

/* With params.profile set, kd_step accumulates per phase the HIP-event time
 * on the context's stream: out = {build ms, traversal ms, kick-drift ms,
 * steps timed} (with the fused kick-drift, the traversal time includes it
 * and the kick-drift time is ~0).  kd_reset_timings zeroes the sums.  No counterpart in the
 * reference (which times whole runs, Parallel/RustVersion/src/main.rs:25-31). */
int kd_timings(kd_ctx *ctx, double out[4]);
int kd_reset_timings(kd_ctx *ctx);

/* FP64 O(N) direct sum (calc_pp_accel form, array_particle.rs:67-76) for k
 * sample bodies given by original index, over all bodies; abs_* (nullable)
 * receive the per-component sum of |a_ij|. */
int kd_direct_sum(kd_ctx *ctx, int64_t k, const int64_t *idx, double *ax,
                  double *ay, double *az, double *absx, double *absy,
                  double *absz);

/* ---- standalone selection ----
 * quickstat_index(indices, goal, |i, j| vals[i] < vals[j])
 * (Parallel/RustVersion/src/quickstat.rs:9-34): reorder indices[0..len) so
 * that the element of rank goal under the order (vals[i], i) sits at goal,
 * every element before it compares less and every element after it greater
 * (the reference's postcondition, quickstat.rs:199-215).  Both sides keep
 * their input order.  Host arrays; indices must lie in [0, nvals). */
int kd_quickstat_index(int64_t len, int64_t *indices, int64_t goal, const double *vals,
                       int64_t nvals);
/* The same on device arrays, with a caller-owned device workspace of at
 * least kd_quickstat_workspace_bytes(len) bytes.  No counterpart in the
 * reference beyond quickstat_index. */
int64_t kd_quickstat_workspace_bytes(int64_t len);
int kd_quickstat_index_device(int64_t len, int64_t *indices, int64_t goal,
                              const double *vals, int64_t nvals, void *workspace,
                              int64_t workspace_bytes);

/* The same selection with no workspace argument and no allocation of any
 * kind: no counterpart in the reference.  It is for callers that cannot
 * allocate device memory (code under stream capture, serving loops, hosts
 * that forbid allocation).  The workspace is a fixed device array of
 * KD_NOALLOC_SCRATCH_BYTES inside the library, one per device, held until the
 * process ends; len above KD_NOALLOC_SELECT_MAX_LEN returns KD_ERR_CAPACITY.
 * Every call in a process on one device shares that array and leaves its
 * contents behind: calls that overlap on one device (from two host threads)
 * race and must be serialised by the caller.  Runs on the null stream and
 * returns after it completes. */
#define KD_NOALLOC_SCRATCH_BYTES (64ll << 20)
#define KD_NOALLOC_SELECT_MAX_LEN 4000000ll
int kd_quickstat_index_device_noalloc(int64_t len, int64_t *indices, int64_t goal,
                                      const double *vals, int64_t nvals);

/* ---- batches: many independent systems in one context ----
 * A batch holds nsys independent systems of n[i] bodies each and runs
 * simple_sim (Parallel/RustVersion/src/array_kd_tree.rs:623-664) on all of
 * them at once: one build of the whole forest (every tree with the
 * reference's split rule and leaf size), one walk over all systems, one
 * kick-drift.  Bodies are passed concatenated in system order; results come
 * back the same way.  kd_batch_step replays the step as a HIP graph (two
 * graphs, one per state buffer), captured on first use.  No counterpart in
 * the reference beyond the per-system calls it replaces. */
typedef struct kd_batch kd_batch;
int kd_batch_create(int64_t nsys, const int64_t *n, const kd_params *params,
                    kd_batch **out);
int kd_batch_destroy(kd_batch *b);
/* Bodies over all systems (sum of n[i]). */
int64_t kd_batch_size(const kd_batch *b);
int kd_batch_upload(kd_batch *b, const double *x, const double *y, const double *z,
                    const double *vx, const double *vy, const double *vz,
                    const double *m);
int kd_batch_download(kd_batch *b, double *x, double *y, double *z, double *vx,
                      double *vy, double *vz);
/* Build every tree, then calc_accel of every body (array_kd_tree.rs:619). */
int kd_batch_accel(kd_batch *b, double *ax, double *ay, double *az);
/* steps x (build, accel, kick-drift) on every system. */
int kd_batch_step(kd_batch *b, int64_t steps);
/* Tree of system sys in a reference layout, as kd_export_tree. */
int kd_batch_export_tree(kd_batch *b, int64_t sys, int32_t layout, kd_tree_view *view);
/* Launch-shape knobs: "graph" 0/1 (replay the step as a HIP graph),
 * "walk_threads" 64/128/256/512 (threads per block of the walk),
 * "walk_waves_per_cu" 0..64 (0: one wavefront per task; otherwise a
 * persistent grid of that many wavefronts per CU takes tasks from a queue),
 * "fused_kick_drift" 0/1 (default 1: the walk applies the kick-drift to its
 * bodies, as kd_set_tuning's key of that name; bit-identical either way),
 * "lazy_dimension" 0/1 (default 1: as kd_set_tuning's key, with the check
 * read after each step -- the host waits for every step -- and a step whose
 * skip was not exact taken again from the same state; trees and states are
 * the same either way). */
int kd_batch_set_tuning(kd_batch *b, const char *key, int64_t value);
/* Reads a knob of kd_batch_set_tuning, or the lazy-dimension state
 * ("lazy_next", "lazy_builds", "lazy_redos", as kd_get_tuning).  No
 * counterpart in the reference. */
int kd_batch_get_tuning(kd_batch *b, const char *key, int64_t *value);
/* Time the walk over the uploaded batch for every launch shape (best of reps
 * runs each) and keep the fastest: out = {threads, waves_per_cu, ms}.  The
 * walk is stackless, so its block size and persistence are the shape. */
int kd_batch_autotune(kd_batch *b, int64_t reps, double out[3]);
/* Counters as kd_counters, summed over the systems and over the traversals
 * since the last kd_batch_accel / kd_batch_step began. */
int kd_batch_counters(kd_batch *b, int64_t out[5]);

/* ---- differentiation ----
 * steps of simple_sim (Parallel/RustVersion/src/array_kd_tree.rs:623-664)
 * with the tree topology frozen within each step (the step's tree and every
 * body's accept/open decisions are fixed by the forward run), and the
 * reverse sweep of that function: gradients of a loss of the final state
 * with respect to the initial positions, initial velocities and masses.
 * The forward keeps a checkpoint of the state every `every` steps; the
 * reverse sweep recomputes each segment's trees from its checkpoint.
 * precision KD_PREC_FP64 or KD_PREC_MIXED (terms of the reverse sweep in the
 * forward's arithmetic, adjoint sums in FP64).  All arrays are DEVICE
 * pointers in original body order.  No counterpart in the reference. */
typedef struct kd_adj kd_adj;
int kd_adj_create(int64_t n, int64_t steps, int64_t every, const kd_params *params,
                  kd_adj **out);
int kd_adj_destroy(kd_adj *a);
/* Run the steps from (x0..vz0, m); the final state goes to x..vz (any may
 * be NULL). */
int kd_adj_forward(kd_adj *a, const double *x0, const double *y0, const double *z0,
                   const double *vx0, const double *vy0, const double *vz0,
                   const double *m, double *x, double *y, double *z, double *vx,
                   double *vy, double *vz);
/* Given dL/dp_final (gp*) and dL/dv_final (gv*; NULL = 0), the gradients
 * dL/dp0, dL/dv0 and dL/dm (any output may be NULL) of the LAST completed
 * kd_adj_forward: a context keeps one run's checkpoints.  A forward that
 * fails part way leaves none, and this call then returns KD_ERR_STATE. */
int kd_adj_backward(kd_adj *a, const double *gpx, const double *gpy, const double *gpz,
                    const double *gvx, const double *gvy, const double *gvz,
                    double *gx0, double *gy0, double *gz0, double *gvx0,
                    double *gvy0, double *gvz0, double *gm);

/* Synthetic initial conditions (host only, no device needed).
 * kd_ic_ring: body 0 mass 1 at rest at the origin; bodies 1..n-1 mass 1e-14,
 * r ~ U[1,2), phi ~ U[0,2pi), z ~ N(0, 1e-3), v = sqrt(1/r)(-sin phi, cos phi, 0).
 * kd_ic_circular_orbits: circular_orbits(n) of array_particle.rs:19-44 (n+1
 * bodies, d = 0.1 + 5i/n, angle U[0,6.28), planar) with a seeded stream.
 * kd_ic_two_bodies: two_bodies() of array_particle.rs:10-17 (2 bodies).
 * Random streams are splitmix64(seed); uniforms are (u >> 11) * 2^-53. */
int kd_ic_ring(uint64_t seed, int64_t n, double *x, double *y, double *z,
               double *vx, double *vy, double *vz, double *m);
int kd_ic_circular_orbits(uint64_t seed, int64_t n, double *x, double *y,
                          double *z, double *vx, double *vy, double *vz,
                          double *m);
int kd_ic_two_bodies(double *x, double *y, double *z, double *vx, double *vy,
                     double *vz, double *m);

#ifdef __cplusplus
}
#endif
#endif
