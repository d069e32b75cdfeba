// Unit tests of the C++ interface (include/kdtree.hpp), the reference's own
// tests restated over it: single_node, two_leaves, big_solar,
// big_solar_with_steps and their par4 variants
// (Parallel/RustVersion/src/array_kd_tree.rs:698-877), plus calc_accel on
// two_bodies against the closed form.  Needs a GPU; prints one line per test
// and exits nonzero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

#include "kdtree.hpp"

using namespace kdnbody;

#define REQUIRE(c)                                                                  \
  do {                                                                              \
    if (!(c)) {                                                                     \
      std::fprintf(stderr, "%s:%d: requirement failed: %s\n", __FILE__, __LINE__, #c); \
      std::exit(1);                                                                 \
    }                                                                               \
  } while (0)

// Every leaf member lies in [min, max) of the box its ancestors' splits cut
// out (recur_test_tree_struct, array_kd_tree.rs:834-877; all three dims here,
// the reference loop stops at dim 1).
static void check_struct(size_t node, const std::vector<KDTree> &nodes,
                         const std::vector<Particle> &parts, std::array<double, 3> mn,
                         std::array<double, 3> mx, size_t &seen) {
  const KDTree &t = nodes[node];
  if (t.is_leaf) {
    for (size_t q = 0; q < t.num_parts; ++q) {
      const size_t i = t.leaf_parts[q];
      for (int d = 0; d < 3; ++d) {
        REQUIRE(parts[i].p[d] >= mn[d]);
        REQUIRE(parts[i].p[d] < mx[d]);
      }
      ++seen;
    }
    return;
  }
  const size_t d = t.split_dim;
  const double tmax = mx[d];
  mx[d] = t.split_val;
  check_struct(t.left, nodes, parts, mn, mx, seen);
  mx[d] = tmax;
  mn[d] = t.split_val;
  check_struct(t.right, nodes, parts, mn, mx, seen);
}

static void check_tree(const std::vector<KDTree> &nodes, const std::vector<Particle> &parts) {
  size_t seen = 0;
  check_struct(0, nodes, parts, {-1e100, -1e100, -1e100}, {1e100, 1e100, 1e100}, seen);
  REQUIRE(seen == parts.size());
}

static std::vector<size_t> iota_indices(size_t n) {
  std::vector<size_t> v(n);
  std::iota(v.begin(), v.end(), 0);
  return v;
}

static void single_node() {
  const auto parts = two_bodies();
  auto nodes = allocate_node_vec(parts.size());
  // the reference asserts 2 here, which nodes_needed_for_particles (one node
  // for n <= MAX_PARTS, array_kd_tree.rs:45-53) never returns
  REQUIRE(nodes.size() == 1);
  auto idx = iota_indices(parts.size());
  REQUIRE(build_tree(idx, 0, parts.size(), parts, 0, nodes) == 0);
  REQUIRE(nodes[0].is_leaf && nodes[0].num_parts == parts.size());
}

static void two_leaves(bool par4) {
  const auto parts = circular_orbits(11);
  auto nodes = allocate_node_vec(parts.size());
  auto idx = iota_indices(parts.size());
  if (par4)
    build_tree_par4(idx, 0, parts, nodes, 1);
  else
    build_tree(idx, 0, parts.size(), parts, 0, nodes);
  check_tree(nodes, parts);
  REQUIRE(!nodes[0].is_leaf);
  REQUIRE(nodes[1].is_leaf && nodes[2].is_leaf);
  REQUIRE(nodes[1].num_parts + nodes[2].num_parts == 12);
}

static void big_solar(bool par4) {
  const auto parts = circular_orbits(5000);
  auto nodes = allocate_node_vec(parts.size());
  auto idx = iota_indices(parts.size());
  if (par4)
    build_tree_par4(idx, 0, parts, nodes, 1);
  else
    build_tree(idx, 0, parts.size(), parts, 0, nodes);
  check_tree(nodes, parts);
  // indices come back as the particle order: a permutation, leaves contiguous
  std::vector<char> hit(parts.size(), 0);
  for (size_t i : idx) hit[i] = 1;
  REQUIRE(std::accumulate(hit.begin(), hit.end(), 0) == (int)parts.size());
}

static void big_solar_with_steps() {
  auto parts = circular_orbits(5000);
  simple_sim(parts, 1e-3, 10);
  auto nodes = allocate_node_vec(parts.size());
  auto idx = iota_indices(parts.size());
  build_tree(idx, 0, parts.size(), parts, 0, nodes);
  check_tree(nodes, parts);
}

static void accel_two_bodies() {
  const auto parts = two_bodies();
  auto nodes = allocate_node_vec(parts.size());
  auto idx = iota_indices(parts.size());
  build_tree(idx, 0, parts.size(), parts, 0, nodes);
  const auto a1 = calc_accel(1, parts, nodes);
  const auto a0 = calc_accel(0, parts, nodes);
  REQUIRE(a1[0] == -1.0 && a1[1] == 0.0 && a1[2] == 0.0);
  REQUIRE(a0[0] == 1e-20 && a0[1] == 0.0 && a0[2] == 0.0);
  const auto pp = calc_pp_accel(parts[1], parts[0]);
  REQUIRE(pp[0] == a1[0] && pp[1] == a1[1] && pp[2] == a1[2]);
}

// theta = 0 opens every internal node: the tree sum is a direct sum.
static void accel_theta_zero_is_direct() {
  const auto parts = circular_orbits(2000, 7);
  auto nodes = allocate_node_vec(parts.size());
  auto idx = iota_indices(parts.size());
  build_tree(idx, 0, parts.size(), parts, 0, nodes);
  const auto a = calc_accel_all(parts, nodes, 0.0);
  for (size_t i = 0; i < parts.size(); i += 37) {
    double s[3] = {0, 0, 0}, sa[3] = {0, 0, 0};
    for (size_t j = 0; j < parts.size(); ++j) {
      if (j == i) continue;
      const auto t = calc_pp_accel(parts[i], parts[j]);
      for (int d = 0; d < 3; ++d) s[d] += t[d], sa[d] += std::fabs(t[d]);
    }
    for (int d = 0; d < 3; ++d) REQUIRE(std::fabs(a[i][d] - s[d]) <= 1e-12 * (sa[d] + 1e-300));
  }
}

// A subrange build (start > 0, cur_node > 0) writes its nodes at cur_node on.
static void subrange_build() {
  const auto parts = circular_orbits(300, 3);
  auto idx = iota_indices(parts.size());
  std::vector<KDTree> nodes(5);
  const size_t last = build_tree(idx, 100, 250, parts, 5, nodes);
  REQUIRE(last + 1 == nodes.size());
  for (size_t i = 0; i < 100; ++i) REQUIRE(idx[i] == i);
  for (size_t i = 250; i < parts.size(); ++i) REQUIRE(idx[i] == i);
  size_t seen = 0;
  check_struct(5, nodes, parts, {-1e100, -1e100, -1e100}, {1e100, 1e100, 1e100}, seen);
  REQUIRE(seen == 150);
}

int main() {
  const std::vector<std::pair<const char *, std::function<void()>>> tests = {
      {"single_node", single_node},
      {"two_leaves", [] { two_leaves(false); }},
      {"two_leaves_par4", [] { two_leaves(true); }},
      {"big_solar", [] { big_solar(false); }},
      {"big_solar_par4", [] { big_solar(true); }},
      {"big_solar_with_steps", big_solar_with_steps},
      {"accel_two_bodies", accel_two_bodies},
      {"accel_theta_zero_is_direct", accel_theta_zero_is_direct},
      {"subrange_build", subrange_build},
  };
  for (const auto &t : tests) {
    try {
      t.second();
    } catch (const Error &e) {
      std::fprintf(stderr, "%s: %s\n", t.first, e.what());
      return 1;
    }
    std::printf("ok %s\n", t.first);
    std::fflush(stdout);
  }
  return 0;
}
