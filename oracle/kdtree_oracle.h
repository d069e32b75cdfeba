/*
 * kdtree_oracle.h -- CPU oracle for the kD-tree Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * It restates, in plain C17 compiled without contraction or fast-math, the
 * algorithm of the reference's Rust path:
 *   build_tree          Sequential/RustVersion/src/array_kd_tree.rs:49-138
 *   build_tree_par4     Parallel/RustVersion/src/array_kd_tree.rs:515-583
 *   nodes_needed / allocate_node_vec
 *                       Sequential/RustVersion/src/array_kd_tree.rs:37-46,
 *                       Parallel/RustVersion/src/array_kd_tree.rs:45-60
 *   quickstat_index     Parallel/RustVersion/src/quickstat.rs:9-34
 *   accel_recur         Sequential/RustVersion/src/array_kd_tree.rs:140-176
 *   calc_pp_accel       Sequential/RustVersion/src/array_particle.rs:67-76
 *   simple_sim          Sequential/RustVersion/src/array_kd_tree.rs:178-218
 *
 * Deliberate, documented deviations (all inside the reference's own freedom):
 *   - the quickselect pivot stream is a seeded splitmix64 instead of the
 *     unseeded fastrand thread-local RNG, so runs are reproducible;
 *   - keys compare under the total order (coordinate, original index); with
 *     distinct coordinates this is identical to the reference's strict '<',
 *     with ties it picks one of the outcomes the reference's random pivots
 *     can produce;
 *   - leaf members are stored in ascending original index (the canonical
 *     particle order), where the reference keeps the pivot-dependent order.
 *
 * Parity pins (each checked in tests/test_oracle.py): the reference C build
 * compiled from its own sources (oracle/_ref) on identical inputs, the only
 * reference known-answer test (quickstat.rs:191-197), closed-form node counts,
 * theta = 0 == O(N^2) direct sum, the two_bodies closed form, and brute-force
 * membership by full sort on small inputs.
 */
#ifndef KDTREE_ORACLE_H
#define KDTREE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LAYOUT_COMPACT = 0, ORC_LAYOUT_PADDED = 1 };

/* One node of the reference's node vector (KDTree enum, array_kd_tree.rs:10-26).
 * Leaves keep their members as a run [first, first + num_parts) of tree->perm. */
typedef struct {
  int32_t is_leaf;
  int32_t split_dim;
  int64_t num_parts;
  int64_t first;
  double split_val;
  double m;
  double cm[3];
  double size;
  int64_t left;
  int64_t right;
} orc_node;

typedef struct {
  int64_t n;         /* particles */
  int64_t n_nodes;   /* length of the node vector, as the reference sizes it */
  int64_t used;      /* last node index written + 1 */
  int32_t leaf_size;
  int32_t layout;
  orc_node *nodes;
  int64_t *perm;     /* particle order: leaves are contiguous runs, DFS order */
} orc_tree;

typedef struct {
  int64_t node_tests;  /* internal nodes whose opening test was evaluated */
  int64_t m2p;         /* accepted far-field (node) interactions */
  int64_t p2p;         /* particle-particle interactions (self excluded) */
} orc_counters;

/* Node-vector sizing.  Compact: 2*(n/(L/2)+1) grown on demand;
 * padded: 1 if n <= L else 2*2^ceil(log2(n/(L/2))) - 1. */
int64_t orc_nodes_needed(int64_t n, int leaf_size, int layout);
/* Nodes actually used by the compact DFS layout for n particles. */
int64_t orc_compact_count(int64_t n, int leaf_size);

/* Sequential quickselect on an index slice under (keys[i], i) order. */
void orc_quickstat_index(int64_t *indices, int64_t len, int64_t goal,
                         const double *keys, uint64_t pivot_seed);

/* Build the tree of n particles.  Returns 0, or -1 on allocation failure,
 * -2 if the padded layout overflows its allocation (the reference panics). */
int orc_build(int64_t n, const double *x, const double *y, const double *z,
              const double *m, int leaf_size, int layout, uint64_t pivot_seed,
              orc_tree *out);
void orc_tree_free(orc_tree *t);

/* Accelerations of all n particles (reference summation structure).  The
 * optional abs_* arrays receive sum_j |a_ij| per component (conditioning). */
int orc_accel(const orc_tree *t, int64_t n, const double *x, const double *y,
              const double *z, const double *m, double theta, double *ax,
              double *ay, double *az, double *absx, double *absy, double *absz,
              orc_counters *cnt);

/* Accelerations of k bodies idx[0..k) only (same arithmetic as orc_accel). */
int orc_accel_subset(const orc_tree *t, int64_t n, const double *x,
                     const double *y, const double *z, const double *m,
                     double theta, int64_t k, const int64_t *idx, double *ax,
                     double *ay, double *az, double *absx, double *absy,
                     double *absz, orc_counters *cnt);

/* One particle, as calc_accel(p, particles, nodes). */
void orc_calc_accel(const orc_tree *t, int64_t p, const double *x,
                    const double *y, const double *z, const double *m,
                    double theta, double out[3]);

/* simple_sim: per step identity indices, build, accel, kick-drift.  The
 * optional acc_hist receives steps*3*n doubles (per step: ax, ay, az). */
int orc_simple_sim(int64_t n, double *x, double *y, double *z, double *vx,
                   double *vy, double *vz, const double *m, double dt,
                   int64_t steps, double theta, int leaf_size, int layout,
                   uint64_t pivot_seed, double *acc_hist);

/* O(N) direct sum for k sample particles (FP64, calc_pp_accel form), plus the
 * per-component sum of |terms|. */
void orc_direct(int64_t n, const double *x, const double *y, const double *z,
                const double *m, int64_t k, const int64_t *idx, double *ax,
                double *ay, double *az, double *absx, double *absy,
                double *absz);

/* The same terms as orc_direct, summed with Neumaier's compensated
 * summation: the sums are exact to about one rounding, so they measure the
 * summation error of any other order of the same terms. */
void orc_direct_comp(int64_t n, const double *x, const double *y, const double *z,
                     const double *m, int64_t k, const int64_t *idx, double *ax,
                     double *ay, double *az, double *absx, double *absy,
                     double *absz);

/* Kick-drift as simple_sim does it (array_kd_tree.rs:203-216). */
void orc_kick_drift(int64_t n, double *x, double *y, double *z, double *vx,
                    double *vy, double *vz, const double *ax, const double *ay,
                    const double *az, double dt);

/* ---- the differentiated step (kdtree_adjoint.c) ----
 * simple_sim with the tree topology frozen per step: step k's tree and
 * every body's interaction list come from the positions at the start of the
 * step, node moments are recomputed from the members.  A tape records the
 * trees, lists and start positions of every step. */
typedef struct orc_tape orc_tape;
/* steps of the frozen-topology step, updating (x..vz) in place.  record:
 * receives a new tape (build + lists from the run's own positions).
 * replay: use that tape's trees and lists instead (the same function at
 * other inputs).  One of the two must be given. */
int orc_adj_forward(int64_t n, double *x, double *y, double *z, double *vx, double *vy,
                    double *vz, const double *m, double dt, int64_t steps, double theta,
                    int leaf_size, orc_tape **record, const orc_tape *replay);
/* Reverse sweep: given dL/dp_N and dL/dv_N, the gradients with respect to
 * the initial positions, initial velocities and masses. */
int orc_adj_backward(const orc_tape *t, const double *m, const double *lpx, const double *lpy,
                     const double *lpz, const double *lvx, const double *lvy, const double *lvz,
                     double *gx0, double *gy0, double *gz0, double *gvx0, double *gvy0,
                     double *gvz0, double *gm);
void orc_tape_free(orc_tape *t);
/* Test hook: plant a mistake in the reverse sweep (1: drop node-to-member
 * terms, 2: drop P2P partner terms, 3: drop mass terms; 0: none). */
void orc_adj_debug_drop(int which);

#ifdef __cplusplus
}
#endif
#endif
