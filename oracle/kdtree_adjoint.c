/*
 * kdtree_adjoint.c -- CPU oracle of the differentiated step.
 *
 * TEST INFRASTRUCTURE ONLY (see kdtree_oracle.h): the FP64 reference gradient
 * that the GPU adjoint is checked against, and the frozen-topology forward
 * runs its own finite-difference pin uses.
 *
 * The function differentiated is simple_sim (Parallel/RustVersion/src/
 * array_kd_tree.rs:623-664) with the tree topology frozen within each step:
 * step k builds its tree on the positions p_k (orc_build, compact layout),
 * and the interaction list of every body (the nodes it accepts and the
 * bodies it meets in leaves, accel_recur at Sequential/RustVersion/src/
 * array_kd_tree.rs:140-172) is fixed by that build.  Within the step the
 * accelerations are then a smooth function of (p, m):
 *   a_i = sum_{j in P2P(i)} f(p_i - p_j, m_j) + sum_{n in M2P(i)} f(p_i - c_n, M_n),
 *   f(d, m) = -m d / |d|^3 (array_particle.rs:67-76),
 *   M_n = sum_{j in n} m_j,  c_n = sum_{j in n} m_j p_j / M_n,
 * followed by the kick-drift v += dt a; p += dt v (array_kd_tree.rs:649-662).
 * A replayed run uses a recorded tape's trees and lists with its own (p, m):
 * central differences of replayed runs are derivatives of this same function.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kdtree_oracle.h"

typedef struct {
  int64_t n_nodes;
  int64_t *lo, *hi;      /* member range [lo, hi) of every node in perm */
  int64_t *perm;         /* particle order of the step's tree */
  int32_t *is_leaf;
  int64_t *off;          /* n+1 offsets into ent */
  int64_t *ent;          /* >= 0: body j (P2P); < 0: node -(e+1) (M2P) */
  double *x, *y, *z;     /* positions at the start of the step */
} adj_step;

struct orc_tape {
  int64_t n, steps;
  double dt;
  adj_step *s;
};

/* ---- recording ----------------------------------------------------------- */

static int64_t ranges(const orc_tree *t, int64_t h, int64_t *lo, int64_t *hi) {
  const orc_node *nd = &t->nodes[h];
  if (nd->is_leaf) {
    lo[h] = nd->first;
    hi[h] = nd->first + nd->num_parts;
  } else {
    ranges(t, nd->left, lo, hi);
    ranges(t, nd->right, lo, hi);
    lo[h] = lo[nd->left];
    hi[h] = hi[nd->right];
  }
  return 0;
}

typedef struct {
  int64_t *v;
  int64_t len, cap;
} vec64;

static int push(vec64 *v, int64_t e) {
  if (v->len == v->cap) {
    int64_t c = v->cap ? 2 * v->cap : 1024;
    int64_t *q = (int64_t *)realloc(v->v, sizeof(int64_t) * (size_t)c);
    if (!q) return -1;
    v->v = q;
    v->cap = c;
  }
  v->v[v->len++] = e;
  return 0;
}

/* accel_recur's visit order, recording instead of summing; the opening
 * test uses the tree's own moments (as the forward traversal does). */
static int list_rec(const orc_tree *t, int64_t h, int64_t p, const double *x, const double *y,
                    const double *z, double theta2, vec64 *out) {
  const orc_node *nd = &t->nodes[h];
  if (nd->is_leaf) {
    for (int64_t i = 0; i < nd->num_parts; ++i) {
      int64_t j = t->perm[nd->first + i];
      if (j != p && push(out, j)) return -1;
    }
    return 0;
  }
  double dx = x[p] - nd->cm[0];
  double dy = y[p] - nd->cm[1];
  double dz = z[p] - nd->cm[2];
  double d2 = dx * dx + dy * dy + dz * dz;
  if (nd->size * nd->size < theta2 * d2) return push(out, -(h + 1));
  if (list_rec(t, nd->left, p, x, y, z, theta2, out)) return -1;
  return list_rec(t, nd->right, p, x, y, z, theta2, out);
}

static int record_step(adj_step *s, int64_t n, const double *x, const double *y, const double *z,
                       const double *m, double theta, int leaf) {
  orc_tree t;
  int rc = orc_build(n, x, y, z, m, leaf, ORC_LAYOUT_COMPACT, 1, &t);
  if (rc) {
    orc_tree_free(&t);
    return rc;
  }
  const int64_t nn = t.n_nodes;
  s->n_nodes = nn;
  s->lo = (int64_t *)calloc((size_t)nn, sizeof(int64_t));
  s->hi = (int64_t *)calloc((size_t)nn, sizeof(int64_t));
  s->is_leaf = (int32_t *)calloc((size_t)nn, sizeof(int32_t));
  s->perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  s->off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  s->x = (double *)malloc(sizeof(double) * (size_t)n);
  s->y = (double *)malloc(sizeof(double) * (size_t)n);
  s->z = (double *)malloc(sizeof(double) * (size_t)n);
  if (!s->lo || !s->hi || !s->is_leaf || !s->perm || !s->off || !s->x || !s->y || !s->z) {
    orc_tree_free(&t);
    return -1;
  }
  ranges(&t, 0, s->lo, s->hi);
  for (int64_t i = 0; i < nn; ++i) s->is_leaf[i] = t.nodes[i].is_leaf;
  memcpy(s->perm, t.perm, sizeof(int64_t) * (size_t)n);
  vec64 v = {NULL, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    s->off[i] = v.len;
    if (list_rec(&t, 0, i, x, y, z, theta * theta, &v)) {
      free(v.v);
      orc_tree_free(&t);
      return -1;
    }
  }
  s->off[n] = v.len;
  s->ent = v.v ? v.v : (int64_t *)malloc(sizeof(int64_t));
  orc_tree_free(&t);
  return 0;
}

/* ---- the frozen-topology force ------------------------------------------- */

/* Node mass and centre of mass over the member range, in particle order. */
static void moments(const adj_step *s, const double *x, const double *y, const double *z,
                    const double *m, double *M, double *cx, double *cy, double *cz) {
  for (int64_t h = 0; h < s->n_nodes; ++h) {
    if (s->is_leaf[h]) continue;
    double mm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (int64_t q = s->lo[h]; q < s->hi[h]; ++q) {
      int64_t j = s->perm[q];
      mm += m[j];
      sx += m[j] * x[j];
      sy += m[j] * y[j];
      sz += m[j] * z[j];
    }
    M[h] = mm;
    cx[h] = sx / mm;
    cy[h] = sy / mm;
    cz[h] = sz / mm;
  }
}

static void forces(const adj_step *s, int64_t n, const double *x, const double *y,
                   const double *z, const double *m, double *ax, double *ay, double *az) {
  const int64_t nn = s->n_nodes;
  double *M = (double *)calloc((size_t)nn, sizeof(double));
  double *cx = (double *)calloc((size_t)nn, sizeof(double));
  double *cy = (double *)calloc((size_t)nn, sizeof(double));
  double *cz = (double *)calloc((size_t)nn, sizeof(double));
  moments(s, x, y, z, m, M, cx, cy, cz);
  for (int64_t i = 0; i < n; ++i) {
    double a[3] = {0.0, 0.0, 0.0};
    for (int64_t e = s->off[i]; e < s->off[i + 1]; ++e) {
      int64_t q = s->ent[e];
      double dx, dy, dz, mq;
      if (q >= 0) {
        dx = x[i] - x[q];
        dy = y[i] - y[q];
        dz = z[i] - z[q];
        mq = m[q];
      } else {
        int64_t h = -(q + 1);
        dx = x[i] - cx[h];
        dy = y[i] - cy[h];
        dz = z[i] - cz[h];
        mq = M[h];
      }
      double d2 = dx * dx + dy * dy + dz * dz;
      double r = sqrt(d2);
      double g = -mq / (d2 * r);
      a[0] += g * dx;
      a[1] += g * dy;
      a[2] += g * dz;
    }
    ax[i] = a[0];
    ay[i] = a[1];
    az[i] = a[2];
  }
  free(M);
  free(cx);
  free(cy);
  free(cz);
}

void orc_tape_free(orc_tape *t) {
  if (!t) return;
  for (int64_t k = 0; k < t->steps; ++k) {
    adj_step *s = &t->s[k];
    free(s->lo);
    free(s->hi);
    free(s->perm);
    free(s->is_leaf);
    free(s->off);
    free(s->ent);
    free(s->x);
    free(s->y);
    free(s->z);
  }
  free(t->s);
  free(t);
}

int orc_adj_forward(int64_t n, double *x, double *y, double *z, double *vx, double *vy,
                    double *vz, const double *m, double dt, int64_t steps, double theta,
                    int leaf_size, orc_tape **record, const orc_tape *replay) {
  if (n < 1 || steps < 0) return -3;
  if (replay && (replay->n != n || replay->steps < steps)) return -3;
  orc_tape *tp = NULL;
  if (record) {
    tp = (orc_tape *)calloc(1, sizeof(orc_tape));
    if (!tp) return -1;
    tp->n = n;
    tp->dt = dt;
    tp->s = (adj_step *)calloc((size_t)(steps ? steps : 1), sizeof(adj_step));
    if (!tp->s) {
      free(tp);
      return -1;
    }
  }
  double *ax = (double *)malloc(sizeof(double) * (size_t)n);
  double *ay = (double *)malloc(sizeof(double) * (size_t)n);
  double *az = (double *)malloc(sizeof(double) * (size_t)n);
  int rc = (ax && ay && az) ? 0 : -1;
  for (int64_t k = 0; k < steps && rc == 0; ++k) {
    const adj_step *use;
    if (replay) {
      use = &replay->s[k];
    } else {
      adj_step *s = &tp->s[k];
      tp->steps = k + 1;
      rc = record_step(s, n, x, y, z, m, theta, leaf_size);
      if (rc) break;
      memcpy(s->x, x, sizeof(double) * (size_t)n);
      memcpy(s->y, y, sizeof(double) * (size_t)n);
      memcpy(s->z, z, sizeof(double) * (size_t)n);
      use = s;
    }
    forces(use, n, x, y, z, m, ax, ay, az);
    orc_kick_drift(n, x, y, z, vx, vy, vz, ax, ay, az, dt);
  }
  free(ax);
  free(ay);
  free(az);
  if (record) {
    if (rc) {
      orc_tape_free(tp);
      *record = NULL;
    } else {
      *record = tp;
    }
  }
  return rc;
}

/* ---- the adjoint ----------------------------------------------------------- */

/* Planted mistakes for the tests that show the finite-difference checks can
 * fail: 1 drops the node-to-member terms, 2 the P2P partner terms, 3 the
 * mass terms.  0 (the default) computes the adjoint. */
static int g_drop = 0;
void orc_adj_debug_drop(int which) { g_drop = which; }

/* u . f(d, mq) and its derivatives: gd = d/dd (u . f), gmq = d/dmq (u . f).
 * f = -mq d / r^3; d/dd = -mq (u / r^3 - 3 (d.u) d / r^5); d/dmq = -(u.d)/r^3. */
static void vjp_term(const double d[3], double mq, const double u[3], double gd[3],
                     double *gmq) {
  double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double r = sqrt(d2);
  double r3 = d2 * r;
  double du = d[0] * u[0] + d[1] * u[1] + d[2] * u[2];
  double k = 3.0 * du / (r3 * d2);
  for (int c = 0; c < 3; ++c) gd[c] = -mq * (u[c] / r3 - k * d[c]);
  *gmq = -du / r3;
}

/* g_p, g_m += (d a / d(p, m))^T u over one recorded step at positions (x,y,z). */
static void step_vjp(const adj_step *s, int64_t n, const double *x, const double *y,
                     const double *z, const double *m, const double *ux, const double *uy,
                     const double *uz, double *gx, double *gy, double *gz, double *gm) {
  const int64_t nn = s->n_nodes;
  double *M = (double *)calloc((size_t)nn, sizeof(double));
  double *cx = (double *)calloc((size_t)nn, sizeof(double));
  double *cy = (double *)calloc((size_t)nn, sizeof(double));
  double *cz = (double *)calloc((size_t)nn, sizeof(double));
  double *Gx = (double *)calloc((size_t)nn, sizeof(double));
  double *Gy = (double *)calloc((size_t)nn, sizeof(double));
  double *Gz = (double *)calloc((size_t)nn, sizeof(double));
  double *H = (double *)calloc((size_t)nn, sizeof(double));
  moments(s, x, y, z, m, M, cx, cy, cz);
  for (int64_t i = 0; i < n; ++i) {
    const double u[3] = {ux[i], uy[i], uz[i]};
    for (int64_t e = s->off[i]; e < s->off[i + 1]; ++e) {
      int64_t q = s->ent[e];
      double d[3], gd[3], gmq;
      if (q >= 0) {
        d[0] = x[i] - x[q];
        d[1] = y[i] - y[q];
        d[2] = z[i] - z[q];
        vjp_term(d, m[q], u, gd, &gmq);
        if (g_drop != 2) {
          gx[q] -= gd[0];
          gy[q] -= gd[1];
          gz[q] -= gd[2];
        }
        if (g_drop != 3) gm[q] += gmq;
      } else {
        int64_t h = -(q + 1);
        d[0] = x[i] - cx[h];
        d[1] = y[i] - cy[h];
        d[2] = z[i] - cz[h];
        vjp_term(d, M[h], u, gd, &gmq);
        Gx[h] -= gd[0];
        Gy[h] -= gd[1];
        Gz[h] -= gd[2];
        if (g_drop != 3) H[h] += gmq;
      }
      gx[i] += gd[0];
      gy[i] += gd[1];
      gz[i] += gd[2];
    }
  }
  /* node moments back to their members: dc/dp_j = m_j / M, dc/dm_j =
   * (p_j - c) / M, dM/dm_j = 1 */
  for (int64_t h = 0; h < nn && g_drop != 1; ++h) {
    if (s->is_leaf[h] || (Gx[h] == 0.0 && Gy[h] == 0.0 && Gz[h] == 0.0 && H[h] == 0.0))
      continue;
    const double ax = Gx[h] / M[h], ay = Gy[h] / M[h], az = Gz[h] / M[h];
    for (int64_t qq = s->lo[h]; qq < s->hi[h]; ++qq) {
      int64_t j = s->perm[qq];
      gx[j] += m[j] * ax;
      gy[j] += m[j] * ay;
      gz[j] += m[j] * az;
      gm[j] += H[h] + (x[j] - cx[h]) * ax + (y[j] - cy[h]) * ay + (z[j] - cz[h]) * az;
    }
  }
  free(M);
  free(cx);
  free(cy);
  free(cz);
  free(Gx);
  free(Gy);
  free(Gz);
  free(H);
}

int orc_adj_backward(const orc_tape *t, const double *m, const double *lpx, const double *lpy,
                     const double *lpz, const double *lvx, const double *lvy, const double *lvz,
                     double *gx0, double *gy0, double *gz0, double *gvx0, double *gvy0,
                     double *gvz0, double *gm) {
  const int64_t n = t->n;
  const double dt = t->dt;
  double *lp[3] = {gx0, gy0, gz0};
  double *lv[3] = {gvx0, gvy0, gvz0};
  const double *lp_in[3] = {lpx, lpy, lpz};
  const double *lv_in[3] = {lvx, lvy, lvz};
  for (int c = 0; c < 3; ++c) {
    memcpy(lp[c], lp_in[c], sizeof(double) * (size_t)n);
    memcpy(lv[c], lv_in[c], sizeof(double) * (size_t)n);
  }
  memset(gm, 0, sizeof(double) * (size_t)n);
  double *u[3];
  for (int c = 0; c < 3; ++c) u[c] = (double *)malloc(sizeof(double) * (size_t)n);
  if (!u[0] || !u[1] || !u[2]) {
    for (int c = 0; c < 3; ++c) free(u[c]);
    return -1;
  }
  for (int64_t k = t->steps - 1; k >= 0; --k) {
    const adj_step *s = &t->s[k];
    /* p' = p + dt v', v' = v + dt a(p): the adjoint of v' collects dt lp' */
    for (int c = 0; c < 3; ++c)
      for (int64_t i = 0; i < n; ++i) {
        lv[c][i] += dt * lp[c][i];
        u[c][i] = dt * lv[c][i];
      }
    step_vjp(s, n, s->x, s->y, s->z, m, u[0], u[1], u[2], lp[0], lp[1], lp[2], gm);
  }
  for (int c = 0; c < 3; ++c) free(u[c]);
  return 0;
}
