// kdtree.hpp -- C++ host interface of the gfx950 kD-tree step, mirroring the
// reference crate rust_kdtree_nbody (Parallel/RustVersion/src/lib.rs:1-3):
// Particle, KDTree, allocate_node_vec, build_tree, build_tree_par4,
// calc_accel, simple_sim, calc_pp_accel, two_bodies, circular_orbits.
// Same names, argument meaning and results; every call runs on the current
// HIP device through the C ABI of kdtree_hip.h.  Where the reference panics,
// these functions throw kdnbody::Error (message from kd_last_error()).
//
// Header-only; link with libkdtree_hip.so.
#ifndef KDTREE_HPP
#define KDTREE_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kdtree_hip.h"

namespace kdnbody {

// Largest leaf size a node record can hold (the reference fixes MAX_PARTS at
// 7, Sequential/RustVersion/src/array_kd_tree.rs:6, or 8,
// Parallel/RustVersion/src/array_kd_tree.rs:14).
#ifndef KDNBODY_MAX_PARTS
#define KDNBODY_MAX_PARTS 8
#endif
constexpr size_t MAX_PARTS = KDNBODY_MAX_PARTS;
static_assert(MAX_PARTS >= 2 && MAX_PARTS <= 32, "leaf size must be in 2..32");
// THETA of array_kd_tree.rs:15.
constexpr double THETA = 0.3;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc) {
  if (rc != KD_OK) throw Error(rc, kd_last_error());
}

// array_particle.rs:3-8 (r is carried but unused by the step).
struct Particle {
  std::array<double, 3> p;
  std::array<double, 3> v;
  double r;
  double m;
};

// KDTree enum of array_kd_tree.rs:19-35: a leaf lists up to MAX_PARTS body
// indices; an internal node carries its split, moments, extent and children.
struct KDTree {
  bool is_leaf = true;
  size_t num_parts = 0;
  std::array<size_t, MAX_PARTS> leaf_parts{};
  size_t split_dim = 0;
  double split_val = 0.0;
  double m = 0.0;
  std::array<double, 3> cm{};
  double size = 0.0;
  size_t left = 0, right = 0;

  static KDTree leaf(size_t num_parts, const std::array<size_t, MAX_PARTS> &parts) {
    KDTree t;
    t.num_parts = num_parts;
    t.leaf_parts = parts;
    return t;
  }
};

enum class Layout : int32_t { Compact = KD_LAYOUT_COMPACT, Padded = KD_LAYOUT_PADDED };

// allocate_node_vec (Sequential :41-46 compact; Parallel :55-60 padded):
// a vector of empty leaves of the length the reference allocates.
inline std::vector<KDTree> allocate_node_vec(size_t num_parts, Layout layout = Layout::Padded,
                                             int leaf_size = (int)MAX_PARTS) {
  const int64_t nn = kd_nodes_needed((int64_t)num_parts, leaf_size, (int32_t)layout);
  if (nn < 1) throw Error(KD_ERR_ARG, "allocate_node_vec: bad size");
  return std::vector<KDTree>((size_t)nn, KDTree{});
}

namespace detail {

// RAII context over kd_create / kd_destroy.
struct Ctx {
  kd_ctx *c = nullptr;
  Ctx(int64_t n, double theta, double dt, int leaf, int precision) {
    kd_params p;
    check(kd_default_params(&p));
    p.theta = theta;
    p.dt = dt;
    p.leaf_size = leaf;
    p.precision = precision;
    check(kd_create(n, &p, &c));
  }
  ~Ctx() {
    if (c) (void)kd_destroy(c);
  }
  Ctx(const Ctx &) = delete;
  Ctx &operator=(const Ctx &) = delete;
};

struct Soa {
  std::vector<double> x, y, z, vx, vy, vz, m;
  explicit Soa(size_t n) : x(n), y(n), z(n), vx(n), vy(n), vz(n), m(n) {}
  void set(size_t i, const Particle &q) {
    x[i] = q.p[0], y[i] = q.p[1], z[i] = q.p[2];
    vx[i] = q.v[0], vy[i] = q.v[1], vz[i] = q.v[2];
    m[i] = q.m;
  }
};

// Flat arrays of a node vector (kd_tree_view), both directions.
struct Flat {
  std::vector<int32_t> is_leaf, split_dim;
  std::vector<int64_t> num_parts, first, left, right, perm;
  std::vector<double> split_val, m, cmx, cmy, cmz, size;
  explicit Flat(size_t nn, size_t n)
      : is_leaf(nn), split_dim(nn), num_parts(nn), first(nn), left(nn), right(nn), perm(n),
        split_val(nn), m(nn), cmx(nn), cmy(nn), cmz(nn), size(nn) {}
  kd_tree_view view() {
    kd_tree_view v;
    v.n_nodes = (int64_t)is_leaf.size();
    v.is_leaf = is_leaf.data();
    v.split_dim = split_dim.data();
    v.num_parts = num_parts.data();
    v.first = first.data();
    v.split_val = split_val.data();
    v.m = m.data();
    v.cmx = cmx.data();
    v.cmy = cmy.data();
    v.cmz = cmz.data();
    v.size = size.data();
    v.left = left.data();
    v.right = right.data();
    v.perm = perm.data();
    return v;
  }
};

inline Flat flatten(const std::vector<KDTree> &nodes, size_t n_particles) {
  size_t members = 0;
  for (const KDTree &t : nodes)
    if (t.is_leaf) members += std::min(t.num_parts, MAX_PARTS);
  Flat f(nodes.size(), members);
  size_t at = 0;
  for (size_t i = 0; i < nodes.size(); ++i) {
    const KDTree &t = nodes[i];
    f.is_leaf[i] = t.is_leaf ? 1 : 0;
    if (t.is_leaf) {
      if (t.num_parts > MAX_PARTS) throw Error(KD_ERR_ARG, "leaf with more than MAX_PARTS bodies");
      f.num_parts[i] = (int64_t)t.num_parts;
      f.first[i] = (int64_t)at;
      for (size_t q = 0; q < t.num_parts; ++q) {
        if (t.leaf_parts[q] >= n_particles) throw Error(KD_ERR_ARG, "leaf member out of range");
        f.perm[at++] = (int64_t)t.leaf_parts[q];
      }
    } else {
      f.split_dim[i] = (int32_t)t.split_dim;
      f.split_val[i] = t.split_val;
      f.m[i] = t.m;
      f.cmx[i] = t.cm[0], f.cmy[i] = t.cm[1], f.cmz[i] = t.cm[2];
      f.size[i] = t.size;
      f.left[i] = (int64_t)t.left;
      f.right[i] = (int64_t)t.right;
    }
  }
  return f;
}

// Build over the bodies indices[start..end) and export in `layout`; node i of
// the export goes to nodes[cur_node + i] with child indices shifted alike,
// and indices[start..end) is rewritten in particle order (leaves contiguous,
// depth first).  Bodies are handed over in ascending particle index so that
// key ties fall to the lower particle index, as for a whole-vector build.
// Returns the number of nodes written.
inline size_t build_into(std::vector<size_t> &indices, size_t start, size_t end,
                         const std::vector<Particle> &particles, size_t cur_node,
                         std::vector<KDTree> &nodes, Layout layout, bool grow, int leaf) {
  if (start >= end || end > indices.size()) throw Error(KD_ERR_ARG, "build_tree: bad range");
  const size_t n = end - start;
  std::vector<size_t> ids(indices.begin() + (ptrdiff_t)start, indices.begin() + (ptrdiff_t)end);
  std::sort(ids.begin(), ids.end());
  Soa s(n);
  for (size_t i = 0; i < n; ++i) {
    if (ids[i] >= particles.size()) throw Error(KD_ERR_ARG, "build_tree: index out of range");
    if (i && ids[i] == ids[i - 1]) throw Error(KD_ERR_ARG, "build_tree: repeated index");
    s.set(i, particles[ids[i]]);
  }
  Ctx c((int64_t)n, THETA, 1e-3, leaf, KD_PREC_FP64);
  check(kd_upload(c.c, (int64_t)n, s.x.data(), s.y.data(), s.z.data(), s.vx.data(), s.vy.data(),
                  s.vz.data(), s.m.data()));
  check(kd_build(c.c));
  kd_tree_view probe{};  // n_nodes 0: asks for the length of the export
  const int rc = kd_export_tree(c.c, (int32_t)layout, &probe);
  if (rc != KD_ERR_CAPACITY) check(rc);
  Flat f((size_t)probe.n_nodes, n);
  kd_tree_view v = f.view();
  check(kd_export_tree(c.c, (int32_t)layout, &v));
  // compact vectors are allocated with slack: count the nodes the tree uses
  // (its largest reachable index + 1); padded slices are copied whole
  size_t used = (size_t)v.n_nodes;
  if (layout == Layout::Compact) {
    size_t top = 0;
    std::vector<size_t> todo{0};
    while (!todo.empty()) {
      const size_t h = todo.back();
      todo.pop_back();
      top = std::max(top, h);
      if (!f.is_leaf[h]) {
        todo.push_back((size_t)f.left[h]);
        todo.push_back((size_t)f.right[h]);
      }
    }
    used = top + 1;
  }
  if (cur_node + used > nodes.size()) {
    if (!grow) throw Error(KD_ERR_CAPACITY, "build_tree_par4: node slice too small");
    nodes.resize(cur_node + used, KDTree{});
  }
  for (size_t i = 0; i < used; ++i) {
    KDTree t;
    t.is_leaf = f.is_leaf[i] != 0;
    if (t.is_leaf) {
      t.num_parts = (size_t)f.num_parts[i];
      for (size_t q = 0; q < t.num_parts; ++q) t.leaf_parts[q] = ids[(size_t)f.perm[(size_t)f.first[i] + q]];
    } else {
      t.split_dim = (size_t)f.split_dim[i];
      t.split_val = f.split_val[i];
      t.m = f.m[i];
      t.cm = {f.cmx[i], f.cmy[i], f.cmz[i]};
      t.size = f.size[i];
      t.left = cur_node + (size_t)f.left[i];
      t.right = cur_node + (size_t)f.right[i];
    }
    nodes[cur_node + i] = t;
  }
  for (size_t i = 0; i < n; ++i) indices[start + i] = ids[(size_t)f.perm[i]];
  return used;
}

}  // namespace detail

// build_tree (Sequential/RustVersion/src/array_kd_tree.rs:49-138): compact
// pre-order layout from cur_node on, growing `nodes` when needed; permutes
// indices[start..end) into particle order; returns the index of the last
// node used.  Leaf size MAX_PARTS unless given.
inline size_t build_tree(std::vector<size_t> &indices, size_t start, size_t end,
                         const std::vector<Particle> &particles, size_t cur_node,
                         std::vector<KDTree> &nodes, int leaf_size = (int)MAX_PARTS) {
  const size_t used = detail::build_into(indices, start, end, particles, cur_node, nodes,
                                         Layout::Compact, true, leaf_size);
  return cur_node + used - 1;
}

// build_tree_par4 (Parallel/RustVersion/src/array_kd_tree.rs:515-583): the
// padded layout into nodes[cur_node..], which must be long enough (as
// allocate_node_vec sizes it).  thread_cnt is accepted for signature parity;
// the build runs on the GPU.
inline void build_tree_par4(std::vector<size_t> &indices, size_t cur_node,
                            const std::vector<Particle> &particles, std::vector<KDTree> &nodes,
                            size_t thread_cnt = 1, int leaf_size = (int)MAX_PARTS) {
  (void)thread_cnt;
  detail::build_into(indices, 0, indices.size(), particles, cur_node, nodes, Layout::Padded,
                     false, leaf_size);
}

// calc_accel for every body in order (the per-body loop of simple_sim,
// array_kd_tree.rs:647) over the node vector as given, FP64.
inline std::vector<std::array<double, 3>> calc_accel_all(const std::vector<Particle> &particles,
                                                         const std::vector<KDTree> &nodes,
                                                         double theta = THETA) {
  const size_t n = particles.size();
  detail::Soa s(n);
  for (size_t i = 0; i < n; ++i) s.set(i, particles[i]);
  detail::Flat f = detail::flatten(nodes, n);
  kd_tree_view v = f.view();
  std::vector<double> ax(n), ay(n), az(n);
  check(kd_accel_tree((int64_t)n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), &v, theta,
                      (int64_t)n, nullptr, ax.data(), ay.data(), az.data()));
  std::vector<std::array<double, 3>> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = {ax[i], ay[i], az[i]};
  return out;
}

// calc_accel (array_kd_tree.rs:619-621): acceleration of body p from the tree
// `nodes`.  One call is one device round trip; use calc_accel_all for many.
inline std::array<double, 3> calc_accel(size_t p, const std::vector<Particle> &particles,
                                        const std::vector<KDTree> &nodes, double theta = THETA) {
  const size_t n = particles.size();
  if (p >= n) throw Error(KD_ERR_ARG, "calc_accel: particle index out of range");
  detail::Soa s(n);
  for (size_t i = 0; i < n; ++i) s.set(i, particles[i]);
  detail::Flat f = detail::flatten(nodes, n);
  kd_tree_view v = f.view();
  const int64_t idx = (int64_t)p;
  double a[3];
  check(kd_accel_tree((int64_t)n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), &v, theta, 1,
                      &idx, &a[0], &a[1], &a[2]));
  return {a[0], a[1], a[2]};
}

// simple_sim (Parallel/RustVersion/src/array_kd_tree.rs:623-664): `steps`
// rounds of build, accelerations and kick-drift, updating bodies in place.
// precision KD_PREC_FP64 is the reference arithmetic; KD_PREC_MIXED the
// FP32-force step.
inline void simple_sim(std::vector<Particle> &bodies, double dt, int64_t steps,
                       int leaf_size = (int)MAX_PARTS, double theta = THETA,
                       int precision = KD_PREC_FP64) {
  const size_t n = bodies.size();
  if (n == 0 || steps < 0) throw Error(KD_ERR_ARG, "simple_sim: bad arguments");
  detail::Soa s(n);
  for (size_t i = 0; i < n; ++i) s.set(i, bodies[i]);
  detail::Ctx c((int64_t)n, theta, dt, leaf_size, precision);
  check(kd_upload(c.c, (int64_t)n, s.x.data(), s.y.data(), s.z.data(), s.vx.data(), s.vy.data(),
                  s.vz.data(), s.m.data()));
  check(kd_step(c.c, steps));
  check(kd_download(c.c, s.x.data(), s.y.data(), s.z.data(), s.vx.data(), s.vy.data(),
                    s.vz.data()));
  for (size_t i = 0; i < n; ++i) {
    bodies[i].p = {s.x[i], s.y[i], s.z[i]};
    bodies[i].v = {s.vx[i], s.vy[i], s.vz[i]};
  }
}

// calc_pp_accel (array_particle.rs:67-76): pairwise term, host FP64.
inline std::array<double, 3> calc_pp_accel(const Particle &pi, const Particle &pj) {
  const double dx = pi.p[0] - pj.p[0], dy = pi.p[1] - pj.p[1], dz = pi.p[2] - pj.p[2];
  const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
  const double magi = -pj.m / (dist * dist * dist);
  return {dx * magi, dy * magi, dz * magi};
}

// two_bodies (array_particle.rs:10-17).
inline std::vector<Particle> two_bodies() {
  std::vector<double> c(7 * 2);
  check(kd_ic_two_bodies(&c[0], &c[2], &c[4], &c[6], &c[8], &c[10], &c[12]));
  std::vector<Particle> out(2);
  for (size_t i = 0; i < 2; ++i)
    out[i] = Particle{{c[i], c[2 + i], c[4 + i]}, {c[6 + i], c[8 + i], c[10 + i]},
                      i == 0 ? 1.0 : 1e-4, c[12 + i]};
  return out;
}

namespace detail {
inline std::vector<Particle> from_ic(size_t n, const std::vector<double> &c, double r0, double r) {
  std::vector<Particle> out(n);
  for (size_t i = 0; i < n; ++i)
    out[i] = Particle{{c[i], c[n + i], c[2 * n + i]},
                      {c[3 * n + i], c[4 * n + i], c[5 * n + i]}, i == 0 ? r0 : r, c[6 * n + i]};
  return out;
}
}  // namespace detail

// circular_orbits (array_particle.rs:19-44): n+1 bodies, the angles drawn
// from a seeded stream instead of the unseeded thread-local generator.
inline std::vector<Particle> circular_orbits(size_t n, uint64_t seed = 0) {
  const size_t t = n + 1;
  std::vector<double> c(7 * t);
  check(kd_ic_circular_orbits(seed, (int64_t)n, &c[0], &c[t], &c[2 * t], &c[3 * t], &c[4 * t],
                              &c[5 * t], &c[6 * t]));
  return detail::from_ic(t, c, 0.00465047, 1e-7);
}

// The ring distribution of the benchmark: a unit mass at rest at the origin
// and n-1 bodies of mass 1e-14 on circular orbits, r in [1,2).
inline std::vector<Particle> ring(size_t n, uint64_t seed = 0) {
  std::vector<double> c(7 * n);
  check(kd_ic_ring(seed, (int64_t)n, &c[0], &c[n], &c[2 * n], &c[3 * n], &c[4 * n], &c[5 * n],
                   &c[6 * n]));
  return detail::from_ic(n, c, 0.00465047, 1e-7);
}

}  // namespace kdnbody

#endif
