// kd_ptrwalk.hip -- force traversal of a caller-supplied node vector in the
// reference layouts (explicit left/right children, leaf member lists), as
// calc_accel(p, particles, nodes) does (Sequential/RustVersion/src/
// array_kd_tree.rs:140-176).  Used by the C++ and Rust host mirrors, whose
// node vectors need not come from this library's build.
//
// One wavefront serves 64 bodies; the current node is wave-uniform and the
// right children still to visit live on a per-wave stack in LDS, with the
// depth of each entry.  Lanes take part in a node iff its depth <= D (the
// depth at which the lane last accepted), as in kd_traverse.hip.  FP64
// reference arithmetic, unfused; self excluded by body index.
#include "kd_internal.hpp"

namespace kd {

constexpr int PW_WAVES = 4;   // wavefronts per block

__global__ __launch_bounds__(PW_WAVES * 64) void k_walk_ptr(PtrTree t, const double *__restrict__ x,
                                                            const double *__restrict__ y,
                                                            const double *__restrict__ z,
                                                            const double *__restrict__ mass,
                                                            int64_t n, const int64_t *__restrict__ idx,
                                                            int64_t k, double theta2,
                                                            double *__restrict__ ax,
                                                            double *__restrict__ ay,
                                                            double *__restrict__ az, int *err) {
#pragma clang fp contract(off)
  __shared__ int64_t snode[PW_WAVES][PW_STACK];
  __shared__ int32_t sdepth[PW_WAVES][PW_STACK];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t s = (int64_t)blockIdx.x * (PW_WAVES * 64) + threadIdx.x;
  const int64_t wbase = s - lane;
  if (wbase >= k) return;
  const bool valid = s < k;
  const int64_t me = valid ? (idx ? idx[s] : s) : -1;
  const double px = valid ? x[me] : 0.0, py = valid ? y[me] : 0.0, pz = valid ? z[me] : 0.0;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  int D = valid ? 0x7fffffff : -1;
  int sp = 0;
  int64_t h = 0;
  int depth = 0;
  for (;;) {
    const bool active = depth <= D;
    bool open = false;
    if (t.is_leaf[h]) {
      if (active) {
        const int64_t f = t.first[h], c = t.num_parts[h];
        for (int64_t q = 0; q < c; ++q) {
          const int64_t j = t.parts[f + q];
          if (j != me) {
            double dx = px - x[j];
            double dy = py - y[j];
            double dz = pz - z[j];
            double dp2 = dx * dx + dy * dy + dz * dz;
            double dist = sqrt(dp2);
            double magi = -mass[j] / (dist * dist * dist);
            sx += magi * dx;
            sy += magi * dy;
            sz += magi * dz;
          }
        }
      }
    } else if (active) {
      double dx = px - t.cmx[h];
      double dy = py - t.cmy[h];
      double dz = pz - t.cmz[h];
      double d2 = dx * dx + dy * dy + dz * dz;
      double sz2 = t.size[h];
      if (sz2 * sz2 < theta2 * d2) {
        double dist = sqrt(d2);
        double magi = -t.m[h] / (d2 * dist);
        sx += dx * magi;
        sy += dy * magi;
        sz += dz * magi;
        D = depth;
      } else {
        D = 0x7fffffff;
        open = true;
      }
    }
    if (__any(open)) {
      if (sp == PW_STACK) {  // deeper than any tree the reference builds
        if (lane == 0) atomicOr(err, 1);
        break;
      }
      // every lane writes the same (uniform) entry
      snode[w][sp] = t.right[h];
      sdepth[w][sp] = depth + 1;
      ++sp;
      h = t.left[h];
      ++depth;
    } else {
      if (sp == 0) break;
      --sp;
      h = snode[w][sp];
      depth = sdepth[w][sp];
    }
  }
  if (valid) {
    ax[s] = sx;
    ay[s] = sy;
    az[s] = sz;
  }
}

hipError_t launch_walk_ptr(const PtrTree &t, const double *x, const double *y, const double *z,
                           const double *m, int64_t n, const int64_t *idx, int64_t k,
                           double theta, double *ax, double *ay, double *az, int *err,
                           hipStream_t s) {
  const int64_t blocks = (k + PW_WAVES * 64 - 1) / (PW_WAVES * 64);
  k_walk_ptr<<<dim3((unsigned)blocks), dim3(PW_WAVES * 64), 0, s>>>(t, x, y, z, m, n, idx, k,
                                                                   theta * theta, ax, ay, az, err);
  return hipGetLastError();
}

}  // namespace kd
