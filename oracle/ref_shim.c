/*
 * ref_shim.c -- thin driver around the reference's own C implementation of
 * the kD-tree step, linked with the reference sources where they lie
 * (Sequential/CVersion/src/{kdtree.c,particle.c} and
 * Parallel/CVersion/src/{kdtree.c,particle.c}) into oracle/_ref/*.so.
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle (membership, split values,
 * accelerations on identical inputs) and as the CPU baseline in bench.py.
 * The reference's build_tree and calc_accel are called unmodified; this file
 * only converts SoA arrays to the reference's Particle records and exports the
 * node vector as flat arrays.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "kdtree.h"
#include "particle.h"

/* Declared in the reference's kdtree.c without a header prototype. */
void calc_accel(size_t p, const Particle_array_t *particles,
                const KDTree_array_t *nodes, double acc[3]);

static Particle_array_t make_particles(int64_t n, const double *x,
                                       const double *y, const double *z,
                                       const double *m) {
  Particle_array_t a = new_array_t((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    a.ptr[i].p[0] = x[i];
    a.ptr[i].p[1] = y[i];
    a.ptr[i].p[2] = z[i];
    a.ptr[i].v[0] = a.ptr[i].v[1] = a.ptr[i].v[2] = 0.0;
    a.ptr[i].r = 0.0;
    a.ptr[i].m = m[i];
  }
  return a;
}

int64_t ref_max_parts(void) { return (int64_t)MAX_PARTS; }

/* Build with the reference's build_tree (compact pre-order layout, rand()
 * pivots seeded by srand(seed)).  Node arrays must hold cap entries; returns
 * the node count written, or -1 if cap is too small. */
int64_t ref_build(int64_t n, const double *x, const double *y, const double *z,
                  const double *m, unsigned seed, int64_t cap,
                  int64_t *num_parts, int64_t *parts /* cap*MAX_PARTS */,
                  int64_t *split_dim, double *split_val, double *nm,
                  double *cmx, double *cmy, double *cmz, double *size,
                  int64_t *left, int64_t *right, int64_t *indices_out) {
  srand(seed);
  Particle_array_t ps = make_particles(n, x, y, z, m);
  KDTree_array_t nodes = allocate_node_vec((size_t)n);
  size_t_array_t ix = new_range(0, (size_t)n);
  size_t last = build_tree(&ix, 0, (size_t)n, &ps, 0, &nodes);
  int64_t used = (int64_t)last + 1;
  int64_t rc = used;
  if (used > cap) {
    rc = -1;
  } else {
    for (int64_t i = 0; i < used; ++i) {
      KDTree *t = nodes.ptr + i;
      num_parts[i] = (int64_t)t->num_parts;
      for (size_t k = 0; k < MAX_PARTS; ++k)
        parts[i * (int64_t)MAX_PARTS + (int64_t)k] =
            k < t->num_parts ? (int64_t)t->particles[k] : -1;
      split_dim[i] = (int64_t)t->split_dim;
      split_val[i] = t->split_val;
      nm[i] = t->m;
      cmx[i] = t->cm[0];
      cmy[i] = t->cm[1];
      cmz[i] = t->cm[2];
      size[i] = t->size;
      left[i] = (int64_t)t->left;
      right[i] = (int64_t)t->right;
    }
    for (int64_t i = 0; i < n; ++i) indices_out[i] = (int64_t)ix.ptr[i];
  }
  FREE_ARRAY(ps);
  FREE_ARRAY(nodes);
  FREE_ARRAY(ix);
  return rc;
}

/* One step's accelerations through the reference's build_tree + calc_accel. */
int ref_accel(int64_t n, const double *x, const double *y, const double *z,
              const double *m, unsigned seed, double *ax, double *ay,
              double *az) {
  srand(seed);
  Particle_array_t ps = make_particles(n, x, y, z, m);
  KDTree_array_t nodes = allocate_node_vec((size_t)n);
  size_t_array_t ix = new_range(0, (size_t)n);
  build_tree(&ix, 0, (size_t)n, &ps, 0, &nodes);
  for (int64_t i = 0; i < n; ++i) {
    double a[3] = {0.0, 0.0, 0.0};
    calc_accel((size_t)i, &ps, &nodes, a);
    ax[i] = a[0];
    ay[i] = a[1];
    az[i] = a[2];
  }
  FREE_ARRAY(ps);
  FREE_ARRAY(nodes);
  FREE_ARRAY(ix);
  return 0;
}

/* steps of build + accel + the reference's kick-drift (kdtree.c simple_sim
 * integrator: v += dt*a then p += dt*v).  acc_hist gets steps*3*n doubles. */
int ref_sim(int64_t n, double *x, double *y, double *z, double *vx, double *vy,
            double *vz, const double *m, double dt, int64_t steps,
            unsigned seed, double *acc_hist) {
  double *a = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  if (!a) return -1;
  for (int64_t s = 0; s < steps; ++s) {
    ref_accel(n, x, y, z, m, seed + (unsigned)s, a, a + n, a + 2 * n);
    if (acc_hist) memcpy(acc_hist + (size_t)s * 3 * (size_t)n, a, sizeof(double) * 3 * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      vx[i] += dt * a[i];
      vy[i] += dt * a[n + i];
      vz[i] += dt * a[2 * n + i];
      x[i] += dt * vx[i];
      y[i] += dt * vy[i];
      z[i] += dt * vz[i];
    }
  }
  free(a);
  return 0;
}

/* CPU baseline: the reference's build_tree (single thread, as its
 * simple_sim does) over all n bodies, then its calc_accel over a sample of k
 * bodies (OpenMP parallel-for, as Parallel/CVersion/src/kdtree.c:237-240),
 * then its kick-drift loop over all n (parallel-for, :244).  Times in
 * seconds go to t_out[0..2] = build, accel sample, integrate. */
int ref_cpu_sample(int64_t n, const double *x, const double *y,
                   const double *z, const double *vx, const double *vy,
                   const double *vz, const double *m, int64_t k,
                   const int64_t *sample, int threads, double dt,
                   double *ax_out, double *t_out) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  srand(12345u);
  Particle_array_t ps = make_particles(n, x, y, z, m);
  for (int64_t i = 0; i < n; ++i) {
    ps.ptr[i].v[0] = vx[i];
    ps.ptr[i].v[1] = vy[i];
    ps.ptr[i].v[2] = vz[i];
  }
  KDTree_array_t nodes = allocate_node_vec((size_t)n);
  size_t_array_t ix = new_range(0, (size_t)n);
  vect3_array_t acc = new_vect3_array_t((size_t)n);
#ifdef _OPENMP
  double t0 = omp_get_wtime();
#else
  double t0 = 0.0;
#endif
  build_tree(&ix, 0, (size_t)n, &ps, 0, &nodes);
#ifdef _OPENMP
  double t1 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 64)
#else
  double t1 = 0.0;
#endif
  for (int64_t s = 0; s < k; ++s) {
    double a[3] = {0.0, 0.0, 0.0};
    calc_accel((size_t)sample[s], &ps, &nodes, a);
    ax_out[3 * s + 0] = a[0];
    ax_out[3 * s + 1] = a[1];
    ax_out[3 * s + 2] = a[2];
  }
#ifdef _OPENMP
  double t2 = omp_get_wtime();
#pragma omp parallel for
#else
  double t2 = 0.0;
#endif
  /* the sample holds only k accelerations; the integrate loop is timed over
   * all n bodies with a zero acceleration array of the reference's shape */
  for (int64_t i = 0; i < n; ++i) {
    ps.ptr[i].v[0] += dt * acc.ptr[i].v[0];
    ps.ptr[i].v[1] += dt * acc.ptr[i].v[1];
    ps.ptr[i].v[2] += dt * acc.ptr[i].v[2];
    ps.ptr[i].p[0] += dt * ps.ptr[i].v[0];
    ps.ptr[i].p[1] += dt * ps.ptr[i].v[1];
    ps.ptr[i].p[2] += dt * ps.ptr[i].v[2];
  }
#ifdef _OPENMP
  double t3 = omp_get_wtime();
#else
  double t3 = 0.0;
#endif
  t_out[0] = t1 - t0;
  t_out[1] = t2 - t1;
  t_out[2] = t3 - t2;
  FREE_ARRAY(acc);
  FREE_ARRAY(ps);
  FREE_ARRAY(nodes);
  FREE_ARRAY(ix);
  return 0;
}

/* CPU baseline, whole steps: the reference's own simple_sim
 * (Parallel/CVersion/src/kdtree.c:221-260: per step identity indices,
 * build_tree on one thread, calc_accel of every body and the kick-drift as
 * OpenMP parallel-for loops) on n bodies for `steps` steps, `threads` OpenMP
 * threads.  The state is updated in place; t_out[0] = wall seconds of the
 * simple_sim call alone (no particle conversion). */
int ref_simple_sim_timed(int64_t n, double *x, double *y, double *z, double *vx,
                         double *vy, double *vz, const double *m, double dt,
                         int64_t steps, int threads, double *t_out) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  srand(12345u);
  Particle_array_t ps = make_particles(n, x, y, z, m);
  for (int64_t i = 0; i < n; ++i) {
    ps.ptr[i].v[0] = vx[i];
    ps.ptr[i].v[1] = vy[i];
    ps.ptr[i].v[2] = vz[i];
  }
#ifdef _OPENMP
  double t0 = omp_get_wtime();
#else
  double t0 = 0.0;
#endif
  simple_sim(&ps, dt, (int)steps);
#ifdef _OPENMP
  t_out[0] = omp_get_wtime() - t0;
#else
  t_out[0] = -1.0;
#endif
  for (int64_t i = 0; i < n; ++i) {
    x[i] = ps.ptr[i].p[0];
    y[i] = ps.ptr[i].p[1];
    z[i] = ps.ptr[i].p[2];
    vx[i] = ps.ptr[i].v[0];
    vy[i] = ps.ptr[i].v[1];
    vz[i] = ps.ptr[i].v[2];
  }
  FREE_ARRAY(ps);
  return 0;
}
