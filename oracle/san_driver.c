/*
 * san_driver.c -- exercises every oracle entry point on small inputs so that
 * the oracle can be run under AddressSanitizer and UndefinedBehaviorSanitizer
 * (oracle/Makefile.san, tests/test_oracle_sanitized.py).
 * TEST INFRASTRUCTURE ONLY, like the oracle itself.
 *
 * Inputs: seeded rings (a unit central mass and bodies at radius 1..2), a
 * lattice with equal coordinates (ties at the medians), a plane (z == 0) and
 * the tiny cases 1, 2, leaf, leaf + 1 bodies.  Exit status 0 when every call
 * returns success; the sanitizers abort on the first error they find.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "kdtree_oracle.h"

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static double unif(void) {
  rng_state += 0x9e3779b97f4a7c15ull;
  uint64_t z = rng_state;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

typedef struct {
  int64_t n;
  double *x, *y, *z, *vx, *vy, *vz, *m;
} bodies;

static bodies alloc_bodies(int64_t n) {
  bodies b;
  b.n = n;
  double **f[7] = {&b.x, &b.y, &b.z, &b.vx, &b.vy, &b.vz, &b.m};
  for (int i = 0; i < 7; ++i) {
    *f[i] = calloc((size_t)n, sizeof(double));
    if (!*f[i]) exit(3);
  }
  return b;
}

static void free_bodies(bodies *b) {
  free(b->x), free(b->y), free(b->z), free(b->vx), free(b->vy), free(b->vz), free(b->m);
}

static bodies ring(int64_t n, int planar) {
  bodies b = alloc_bodies(n);
  b.m[0] = 1.0;
  for (int64_t i = 1; i < n; ++i) {
    const double r = 1.0 + unif(), phi = 6.283185307179586 * unif();
    b.x[i] = r * cos(phi);
    b.y[i] = r * sin(phi);
    b.z[i] = planar ? 0.0 : 1e-3 * (unif() - 0.5);
    b.vx[i] = -sin(phi) / sqrt(r);
    b.vy[i] = cos(phi) / sqrt(r);
    b.m[i] = 1e-14;
  }
  return b;
}

static bodies lattice(int k) {
  bodies b = alloc_bodies((int64_t)k * k * k);
  int64_t i = 0;
  for (int a = 0; a < k; ++a)
    for (int c = 0; c < k; ++c)
      for (int d = 0; d < k; ++d, ++i) {
        b.x[i] = a, b.y[i] = c, b.z[i] = d;
        b.m[i] = 1.0 / (double)b.n;
      }
  return b;
}

static int check(int rc, const char *what, int64_t n) {
  if (rc != 0) {
    fprintf(stderr, "%s failed (rc %d) at n = %lld\n", what, rc, (long long)n);
    exit(1);
  }
  return rc;
}

static void exercise(bodies *b, int leaf) {
  const int64_t n = b->n;
  for (int layout = 0; layout < 2; ++layout) {
    orc_tree t;
    check(orc_build(n, b->x, b->y, b->z, b->m, leaf, layout, 7, &t), "orc_build", n);
    double *a = calloc((size_t)(6 * n), sizeof(double));
    if (!a) exit(3);
    orc_counters cnt;
    check(orc_accel(&t, n, b->x, b->y, b->z, b->m, 0.3, a, a + n, a + 2 * n, a + 3 * n,
                    a + 4 * n, a + 5 * n, &cnt),
          "orc_accel", n);
    int64_t idx[3] = {0, n / 2, n - 1};
    double o[18];  /* six arrays of the three samples */
    check(orc_accel_subset(&t, n, b->x, b->y, b->z, b->m, 0.5, 3, idx, o, o + 3, o + 6, NULL,
                           NULL, NULL, &cnt),
          "orc_accel_subset", n);
    double one[3];
    orc_calc_accel(&t, n - 1, b->x, b->y, b->z, b->m, 0.0, one);
    orc_direct(n, b->x, b->y, b->z, b->m, 3, idx, o, o + 3, o + 6, o + 9, o + 12, o + 15);
    free(a);
    orc_tree_free(&t);
  }
}

static void sim_and_adjoint(bodies *b, int leaf) {
  const int64_t n = b->n;
  bodies c = alloc_bodies(n);
  for (int64_t i = 0; i < n; ++i) {
    c.x[i] = b->x[i], c.y[i] = b->y[i], c.z[i] = b->z[i];
    c.vx[i] = b->vx[i], c.vy[i] = b->vy[i], c.vz[i] = b->vz[i], c.m[i] = b->m[i];
  }
  double *hist = calloc((size_t)(3 * 3 * n), sizeof(double));
  if (!hist) exit(3);
  check(orc_simple_sim(n, c.x, c.y, c.z, c.vx, c.vy, c.vz, c.m, 1e-3, 3, 0.3, leaf, 0, 7, hist),
        "orc_simple_sim", n);
  free(hist);
  orc_tape *tape = NULL;
  check(orc_adj_forward(n, c.x, c.y, c.z, c.vx, c.vy, c.vz, c.m, 1e-3, 4, 0.3, leaf, &tape,
                        NULL),
        "orc_adj_forward", n);
  bodies g = alloc_bodies(n);
  double *gm = calloc((size_t)n, sizeof(double));
  if (!gm) exit(3);
  check(orc_adj_backward(tape, c.m, c.x, c.y, c.z, c.vx, c.vy, c.vz, g.x, g.y, g.z, g.vx, g.vy,
                         g.vz, gm),
        "orc_adj_backward", n);
  check(orc_adj_forward(n, c.x, c.y, c.z, c.vx, c.vy, c.vz, c.m, 1e-3, 4, 0.3, leaf, NULL, tape),
        "orc_adj_forward replay", n);
  orc_tape_free(tape);
  free(gm);
  free_bodies(&g);
  free_bodies(&c);
}

int main(void) {
  const int leaves[3] = {2, 7, 32};
  const int64_t sizes[7] = {1, 2, 7, 8, 33, 1000, 5000};
  for (int li = 0; li < 3; ++li)
    for (int si = 0; si < 7; ++si) {
      bodies b = ring(sizes[si], 0);
      exercise(&b, leaves[li]);
      if (sizes[si] <= 1000) sim_and_adjoint(&b, leaves[li]);
      free_bodies(&b);
    }
  bodies p = ring(3000, 1);
  exercise(&p, 7);
  free_bodies(&p);
  bodies l = lattice(12);
  exercise(&l, 7);
  sim_and_adjoint(&l, 7);
  free_bodies(&l);
  int64_t ix[7] = {0, 1, 2, 3, 4, 5, 6};
  const double keys[7] = {2.3, 9.8, 3.1, 1.6, 6.7, 7.8, 8.6};
  orc_quickstat_index(ix, 7, 3, keys, 1);
  if (ix[3] != 4) {
    fprintf(stderr, "quickstat known answer: indices[3] = %lld\n", (long long)ix[3]);
    return 1;
  }
  printf("oracle sanitizer run ok\n");
  return 0;
}
