/*
 * kdtree_oracle.c -- CPU oracle (test infrastructure only, see the header).
 * Compiled with -O2 -ffp-contract=off -fno-fast-math: every product and sum
 * below rounds exactly once, in source order, as the reference's Rust does.
 */
#include "kdtree_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- pivot stream ------------------------------------------------------ */

static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* ---- sizing ------------------------------------------------------------ */

/* ceil(log2(x)) for x >= 1, exact (f64::log2(x).ceil() in
 * Parallel/RustVersion/src/array_kd_tree.rs:50 for the integers it sees). */
static int ceil_log2_u64(uint64_t x) {
  int p = 0;
  while (((uint64_t)1 << p) < x) ++p;
  return p;
}

int64_t orc_nodes_needed(int64_t n, int leaf_size, int layout) {
  int64_t half = leaf_size / 2;
  if (layout == ORC_LAYOUT_PADDED) {
    /* Parallel/RustVersion/src/array_kd_tree.rs:45-53 */
    if (n <= leaf_size) return 1;
    int64_t min_leaves = n / half;
    int64_t leaves = (int64_t)1 << ceil_log2_u64((uint64_t)min_leaves);
    return 2 * leaves - 1;
  }
  /* Sequential/RustVersion/src/array_kd_tree.rs:37-39 */
  return 2 * (n / half + 1);
}

int64_t orc_compact_count(int64_t n, int leaf_size) {
  if (n <= leaf_size) return 1;
  return 1 + orc_compact_count(n / 2, leaf_size) +
         orc_compact_count(n - n / 2, leaf_size);
}

/* ---- selection ---------------------------------------------------------- */

/* (key, index) total order; equal to the reference's strict '<' on keys
 * whenever keys are distinct. */
static int key_lt(const double *keys, int64_t a, int64_t b) {
  double ka = keys[a], kb = keys[b];
  return ka < kb || (ka == kb && a < b);
}

/* Parallel/RustVersion/src/quickstat.rs:9-34 (same loop structure: random
 * pivot swapped to s, strict-less two-pointer partition, pivot to high). */
static void quickstat(int64_t *ix, int64_t len, int64_t goal, const double *keys,
                      uint64_t *rng) {
  int64_t s = 0, e = len;
  while (s + 1 < e) {
    int64_t pivot = s + (int64_t)(splitmix64(rng) % (uint64_t)(e - s));
    int64_t t = ix[s]; ix[s] = ix[pivot]; ix[pivot] = t;
    int64_t low = s + 1, high = e - 1;
    while (low <= high) {
      if (key_lt(keys, ix[low], ix[s])) {
        low += 1;
      } else {
        t = ix[low]; ix[low] = ix[high]; ix[high] = t;
        high -= 1;
      }
    }
    t = ix[s]; ix[s] = ix[high]; ix[high] = t;
    if (high < goal) s = high + 1;
    else if (high > goal) e = high;
    else s = e;
  }
}

void orc_quickstat_index(int64_t *indices, int64_t len, int64_t goal,
                         const double *keys, uint64_t pivot_seed) {
  uint64_t rng = pivot_seed;
  quickstat(indices, len, goal, keys, &rng);
}

/* ---- build -------------------------------------------------------------- */

typedef struct {
  const double *c[3];
  const double *m;
  int leaf;
  int layout;
  uint64_t rng;
  orc_tree *t;
  int err;
} build_ctx;

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

static int ensure_nodes(orc_tree *t, int64_t need) {
  if (need <= t->n_nodes) return 0;
  orc_node *nn = (orc_node *)realloc(t->nodes, sizeof(orc_node) * (size_t)need);
  if (!nn) return -1;
  for (int64_t i = t->n_nodes; i < need; ++i) {
    memset(&nn[i], 0, sizeof(orc_node));
    nn[i].is_leaf = 1; /* KDTree::leaf(0, NEGS) */
  }
  t->nodes = nn;
  t->n_nodes = need;
  return 0;
}

static void make_leaf(build_ctx *b, int64_t start, int64_t end, int64_t cur) {
  orc_tree *t = b->t;
  /* canonical particle order inside a leaf */
  qsort(t->perm + start, (size_t)(end - start), sizeof(int64_t), cmp_i64);
  orc_node *nd = &t->nodes[cur];
  memset(nd, 0, sizeof(*nd));
  nd->is_leaf = 1;
  nd->num_parts = end - start;
  nd->first = start;
  if (cur + 1 > t->used) t->used = cur + 1;
}

/* One pass for m, sum(m*p) and the bounding box, in the segment's current
 * order (Sequential/RustVersion/src/array_kd_tree.rs:72-97), then the split
 * dimension with strict '>' (ties keep the lower dimension). */
static void node_stats(build_ctx *b, int64_t start, int64_t end, orc_node *nd) {
  const int64_t *ix = b->t->perm;
  double mn[3] = {1e100, 1e100, 1e100};
  double mx[3] = {-1e100, -1e100, -1e100};
  double m = 0.0;
  double cm[3] = {0.0, 0.0, 0.0};
  for (int64_t i = start; i < end; ++i) {
    int64_t p = ix[i];
    double mp = b->m[p];
    m += mp;
    for (int d = 0; d < 3; ++d) {
      double v = b->c[d][p];
      cm[d] += mp * v;
      mn[d] = fmin(mn[d], v);
      mx[d] = fmax(mx[d], v);
    }
  }
  for (int d = 0; d < 3; ++d) cm[d] /= m;
  int sd = 0;
  for (int d = 1; d < 3; ++d)
    if (mx[d] - mn[d] > mx[sd] - mn[sd]) sd = d;
  memset(nd, 0, sizeof(*nd));
  nd->is_leaf = 0;
  nd->split_dim = sd;
  nd->m = m;
  nd->cm[0] = cm[0]; nd->cm[1] = cm[1]; nd->cm[2] = cm[2];
  nd->size = mx[sd] - mn[sd];
}

/* Compact pre-order layout: Sequential/RustVersion/src/array_kd_tree.rs:49-138 */
static int64_t build_compact(build_ctx *b, int64_t start, int64_t end, int64_t cur) {
  orc_tree *t = b->t;
  int64_t np = end - start;
  if (ensure_nodes(t, cur + 1)) { b->err = -1; return cur; }
  if (np <= b->leaf) {
    make_leaf(b, start, end, cur);
    return cur;
  }
  orc_node nd;
  node_stats(b, start, end, &nd);
  int64_t mid = (start + end) / 2;
  quickstat(t->perm + start, np, mid - start, b->c[nd.split_dim], &b->rng);
  nd.split_val = b->c[nd.split_dim][t->perm[mid]];
  int64_t left = build_compact(b, start, mid, cur + 1);
  int64_t right = build_compact(b, mid, end, left + 1);
  if (b->err) return right;
  if (ensure_nodes(t, cur + 1)) { b->err = -1; return right; }
  nd.left = cur + 1;
  nd.right = left + 1;
  t->nodes[cur] = nd;
  if (cur + 1 > t->used) t->used = cur + 1;
  return right;
}

/* Padded layout: Parallel/RustVersion/src/array_kd_tree.rs:515-583.  The
 * node slice of this subtree is [cur, limit); the reference panics when a
 * slice is too short, here that is error -2. */
static void build_padded(build_ctx *b, int64_t start, int64_t end, int64_t cur,
                         int64_t limit) {
  orc_tree *t = b->t;
  int64_t np = end - start;
  if (cur >= limit) { b->err = -2; return; }
  if (np <= b->leaf) {
    make_leaf(b, start, end, cur);
    return;
  }
  orc_node nd;
  node_stats(b, start, end, &nd);
  int64_t mid = start + np / 2;
  quickstat(t->perm + start, np, mid - start, b->c[nd.split_dim], &b->rng);
  nd.split_val = b->c[nd.split_dim][t->perm[mid]];
  int64_t num_nodes = orc_nodes_needed(mid - start, b->leaf, ORC_LAYOUT_PADDED);
  if (num_nodes > limit - cur - 1) { b->err = -2; return; }
  build_padded(b, start, mid, cur + 1, cur + 1 + num_nodes);
  build_padded(b, mid, end, cur + 1 + num_nodes, limit);
  nd.left = cur + 1;
  nd.right = cur + 1 + num_nodes;
  t->nodes[cur] = nd;
  if (cur + 1 > t->used) t->used = cur + 1;
}

int orc_build(int64_t n, const double *x, const double *y, const double *z,
              const double *m, int leaf_size, int layout, uint64_t pivot_seed,
              orc_tree *out) {
  memset(out, 0, sizeof(*out));
  if (n < 1 || leaf_size < 2) return -3;
  out->n = n;
  out->leaf_size = leaf_size;
  out->layout = layout;
  out->perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  if (!out->perm) return -1;
  for (int64_t i = 0; i < n; ++i) out->perm[i] = i;
  if (ensure_nodes(out, orc_nodes_needed(n, leaf_size, layout))) return -1;
  build_ctx b = {{x, y, z}, m, leaf_size, layout, pivot_seed, out, 0};
  if (layout == ORC_LAYOUT_PADDED)
    build_padded(&b, 0, n, 0, out->n_nodes);
  else
    build_compact(&b, 0, n, 0);
  return b.err;
}

void orc_tree_free(orc_tree *t) {
  free(t->nodes);
  free(t->perm);
  memset(t, 0, sizeof(*t));
}

/* ---- traversal ----------------------------------------------------------- */

typedef struct {
  const orc_tree *t;
  const double *x, *y, *z, *m;
  double theta2;
  int want_abs;
  orc_counters cnt;
} acc_ctx;

/* Sequential/RustVersion/src/array_particle.rs:67-76 */
static void pp_accel(const acc_ctx *a, int64_t i, int64_t j, double out[3]) {
  double dx = a->x[i] - a->x[j];
  double dy = a->y[i] - a->y[j];
  double dz = a->z[i] - a->z[j];
  double dp2 = dx * dx + dy * dy + dz * dz;
  double dist = sqrt(dp2);
  double magi = -a->m[j] / (dist * dist * dist);
  out[0] = magi * dx;
  out[1] = magi * dy;
  out[2] = magi * dz;
}

/* Sequential/RustVersion/src/array_kd_tree.rs:140-172: a leaf sums its
 * members (self excluded by index); an internal node either accepts
 * (monopole) or returns left + right. */
static void accel_rec(acc_ctx *a, int64_t h, int64_t p, double out[3],
                      double ab[3]) {
  const orc_node *nd = &a->t->nodes[h];
  if (nd->is_leaf) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < nd->num_parts; ++i) {
      int64_t j = a->t->perm[nd->first + i];
      if (j != p) {
        double pp[3];
        pp_accel(a, p, j, pp);
        acc[0] += pp[0];
        acc[1] += pp[1];
        acc[2] += pp[2];
        a->cnt.p2p++;
        if (a->want_abs) {
          ab[0] += fabs(pp[0]); ab[1] += fabs(pp[1]); ab[2] += fabs(pp[2]);
        }
      }
    }
    out[0] = acc[0]; out[1] = acc[1]; out[2] = acc[2];
    return;
  }
  double dx = a->x[p] - nd->cm[0];
  double dy = a->y[p] - nd->cm[1];
  double dz = a->z[p] - nd->cm[2];
  double dist_sqr = dx * dx + dy * dy + dz * dz;
  a->cnt.node_tests++;
  if (nd->size * nd->size < a->theta2 * dist_sqr) {
    double dist = sqrt(dist_sqr);
    double magi = -nd->m / (dist_sqr * dist);
    out[0] = dx * magi;
    out[1] = dy * magi;
    out[2] = dz * magi;
    a->cnt.m2p++;
    if (a->want_abs) {
      ab[0] += fabs(out[0]); ab[1] += fabs(out[1]); ab[2] += fabs(out[2]);
    }
  } else {
    double l[3], r[3];
    accel_rec(a, nd->left, p, l, ab);
    accel_rec(a, nd->right, p, r, ab);
    out[0] = l[0] + r[0];
    out[1] = l[1] + r[1];
    out[2] = l[2] + r[2];
  }
}

int orc_accel(const orc_tree *t, int64_t n, const double *x, const double *y,
              const double *z, const double *m, double theta, double *ax,
              double *ay, double *az, double *absx, double *absy, double *absz,
              orc_counters *cnt) {
  acc_ctx a = {t, x, y, z, m, theta * theta, absx != NULL, {0, 0, 0}};
  for (int64_t i = 0; i < n; ++i) {
    double o[3], ab[3] = {0.0, 0.0, 0.0};
    accel_rec(&a, 0, i, o, ab);
    ax[i] = o[0]; ay[i] = o[1]; az[i] = o[2];
    if (absx) { absx[i] = ab[0]; absy[i] = ab[1]; absz[i] = ab[2]; }
  }
  if (cnt) *cnt = a.cnt;
  return 0;
}

int orc_accel_subset(const orc_tree *t, int64_t n, const double *x,
                     const double *y, const double *z, const double *m,
                     double theta, int64_t k, const int64_t *idx, double *ax,
                     double *ay, double *az, double *absx, double *absy,
                     double *absz, orc_counters *cnt) {
  (void)n;
  acc_ctx a = {t, x, y, z, m, theta * theta, absx != NULL, {0, 0, 0}};
  for (int64_t s = 0; s < k; ++s) {
    double o[3], ab[3] = {0.0, 0.0, 0.0};
    accel_rec(&a, 0, idx[s], o, ab);
    ax[s] = o[0]; ay[s] = o[1]; az[s] = o[2];
    if (absx) { absx[s] = ab[0]; absy[s] = ab[1]; absz[s] = ab[2]; }
  }
  if (cnt) *cnt = a.cnt;
  return 0;
}

void orc_calc_accel(const orc_tree *t, int64_t p, const double *x,
                    const double *y, const double *z, const double *m,
                    double theta, double out[3]) {
  acc_ctx a = {t, x, y, z, m, theta * theta, 0, {0, 0, 0}};
  double ab[3] = {0.0, 0.0, 0.0};
  accel_rec(&a, 0, p, out, ab);
}

/* ---- integrator / driver ------------------------------------------------- */

void orc_kick_drift(int64_t n, double *x, double *y, double *z, double *vx,
                    double *vy, double *vz, const double *ax, const double *ay,
                    const double *az, double dt) {
  for (int64_t i = 0; i < n; ++i) {
    vx[i] += dt * ax[i];
    vy[i] += dt * ay[i];
    vz[i] += dt * az[i];
    double dx = dt * vx[i];
    double dy = dt * vy[i];
    double dz = dt * vz[i];
    x[i] += dx;
    y[i] += dy;
    z[i] += dz;
  }
}

int orc_simple_sim(int64_t n, double *x, double *y, double *z, double *vx,
                   double *vy, double *vz, const double *m, double dt,
                   int64_t steps, double theta, int leaf_size, int layout,
                   uint64_t pivot_seed, double *acc_hist) {
  double *ax = (double *)malloc(sizeof(double) * (size_t)n * 3);
  if (!ax) return -1;
  double *ay = ax + n, *az = ax + 2 * n;
  uint64_t seed = pivot_seed;
  for (int64_t s = 0; s < steps; ++s) {
    orc_tree t;
    int rc = orc_build(n, x, y, z, m, leaf_size, layout, splitmix64(&seed), &t);
    if (rc) { orc_tree_free(&t); free(ax); return rc; }
    orc_accel(&t, n, x, y, z, m, theta, ax, ay, az, NULL, NULL, NULL, NULL);
    orc_tree_free(&t);
    if (acc_hist) memcpy(acc_hist + (size_t)s * 3 * (size_t)n, ax, sizeof(double) * 3 * (size_t)n);
    orc_kick_drift(n, x, y, z, vx, vy, vz, ax, ay, az, dt);
  }
  free(ax);
  return 0;
}

void orc_direct(int64_t n, const double *x, const double *y, const double *z,
                const double *m, int64_t k, const int64_t *idx, double *ax,
                double *ay, double *az, double *absx, double *absy,
                double *absz) {
  acc_ctx a = {NULL, x, y, z, m, 0.0, 0, {0, 0, 0}};
  for (int64_t s = 0; s < k; ++s) {
    int64_t i = idx[s];
    double acc[3] = {0.0, 0.0, 0.0}, ab[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      double pp[3];
      pp_accel(&a, i, j, pp);
      acc[0] += pp[0]; acc[1] += pp[1]; acc[2] += pp[2];
      ab[0] += fabs(pp[0]); ab[1] += fabs(pp[1]); ab[2] += fabs(pp[2]);
    }
    ax[s] = acc[0]; ay[s] = acc[1]; az[s] = acc[2];
    if (absx) { absx[s] = ab[0]; absy[s] = ab[1]; absz[s] = ab[2]; }
  }
}

/* Neumaier's compensated summation: s + c holds the running sum to about
 * one rounding of the result, whatever the order or the cancellation. */
static void neumaier_add(double *s, double *c, double v) {
  const double t = *s + v;
  if (fabs(*s) >= fabs(v))
    *c += (*s - t) + v;
  else
    *c += (v - t) + *s;
  *s = t;
}

void orc_direct_comp(int64_t n, const double *x, const double *y, const double *z,
                     const double *m, int64_t k, const int64_t *idx, double *ax,
                     double *ay, double *az, double *absx, double *absy, double *absz) {
  acc_ctx a = {NULL, x, y, z, m, 0.0, 0, {0, 0, 0}};
  for (int64_t s = 0; s < k; ++s) {
    int64_t i = idx[s];
    double acc[3] = {0.0, 0.0, 0.0}, cc[3] = {0.0, 0.0, 0.0};
    double ab[3] = {0.0, 0.0, 0.0}, cb[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      double pp[3];
      pp_accel(&a, i, j, pp);
      for (int c = 0; c < 3; ++c) {
        neumaier_add(&acc[c], &cc[c], pp[c]);
        neumaier_add(&ab[c], &cb[c], fabs(pp[c]));
      }
    }
    ax[s] = acc[0] + cc[0]; ay[s] = acc[1] + cc[1]; az[s] = acc[2] + cc[2];
    if (absx) { absx[s] = ab[0] + cb[0]; absy[s] = ab[1] + cb[1]; absz[s] = ab[2] + cb[2]; }
  }
}
