// kd_traverse.hip -- theta-opened force traversal on gfx950.
//
// One wavefront serves 64 consecutive bodies in tree order.  It walks the
// heap-ordered tree depth first without a stack: the current node h
// (1-based) is wave-uniform, so node records and leaf bodies arrive through
// scalar loads; descending goes to 2h, finishing a subtree goes to
// (h >> trailing_ones(h)) + 1.  Each lane keeps the depth D at which it last
// accepted a node (or INT_MAX while it keeps opening): a lane takes part in
// node h iff depth(h) <= D, so every lane sees exactly the nodes and leaves
// its own recursive traversal (Sequential/RustVersion/src/array_kd_tree.rs:
// 140-172) would see.  The wave descends when any active lane opens.
//
// FP64 kernel: the reference arithmetic, unfused (accept iff
// size*size < theta^2*d2, M2P -m/(d2*sqrt(d2)), P2P -m/(dist*dist*dist)).
// Mixed kernels: FP32 arithmetic with v_rsq_f32 and compensated (Kahan) FP32
// accumulation; FP64 results handed to the FP64 integrator.
#include "kd_internal.hpp"

namespace kd {

struct Counters {
  unsigned long long *c;  // [0] node tests, [1] M2P, [2] P2P
};

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <bool COUNT>
__device__ __forceinline__ void flush_counters(Counters cnt, unsigned long long t,
                                               unsigned long long a,
                                               unsigned long long p) {
  if (!COUNT) return;
  t = wave_sum_u64(t);
  a = wave_sum_u64(a);
  p = wave_sum_u64(p);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&cnt.c[0], t);
    atomicAdd(&cnt.c[1], a);
    atomicAdd(&cnt.c[2], p);
  }
}

__device__ __forceinline__ uint32_t uni(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

// Advance past the subtree of h; returns 0 when the root is finished.
__device__ __forceinline__ uint32_t heap_next(uint32_t h, int &depth) {
  int t = __builtin_ctz(~h);
  h >>= t;
  depth -= t;
  return h == 0 ? 0 : h + 1;
}

template <bool COUNT>
__global__ __launch_bounds__(256) void k_traverse_f64(const NodeRec64 *__restrict__ nodes,
                                                      const double4 *__restrict__ pos,
                                                      uint32_t n, double theta2,
                                                      uint32_t leaf, double *__restrict__ ax,
                                                      double *__restrict__ ay,
                                                      double *__restrict__ az,
                                                      Counters cnt) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
  if (wbase >= n) return;
  const uint32_t me = wbase + lane;
  const bool valid = me < n;
  const double4 p = valid ? pos[me] : make_double4(0.0, 0.0, 0.0, 0.0);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  int D = valid ? 0x7fffffff : -1;
  unsigned long long nt = 0, na = 0, np = 0;
  uint32_t h = 1;
  int depth = 0;
  while (h != 0) {
    const NodeRec64 r = nodes[uni(h) - 1];
    const bool active = depth <= D;
    if (r.count <= leaf) {
      if (active) {
        for (uint32_t jj = 0; jj < r.count; ++jj) {
          const uint32_t j = r.start + jj;
          const double4 q = pos[j];
          if (j != me) {
            double dx = p.x - q.x;
            double dy = p.y - q.y;
            double dz = p.z - q.z;
            double dp2 = dx * dx + dy * dy + dz * dz;
            double dist = sqrt(dp2);
            double magi = -q.w / (dist * dist * dist);
            sx += magi * dx;
            sy += magi * dy;
            sz += magi * dz;
            if (COUNT) ++np;
          }
        }
      }
      h = heap_next(h, depth);
      continue;
    }
    bool open = false;
    if (active) {
      double dx = p.x - r.cx;
      double dy = p.y - r.cy;
      double dz = p.z - r.cz;
      double d2 = dx * dx + dy * dy + dz * dz;
      if (COUNT) ++nt;
      if (r.size2 < theta2 * d2) {
        double dist = sqrt(d2);
        double magi = -r.m / (d2 * dist);
        sx += dx * magi;
        sy += dy * magi;
        sz += dz * magi;
        D = depth;
        if (COUNT) ++na;
      } else {
        D = 0x7fffffff;
        open = true;
      }
    }
    if (__any(open)) {
      h = 2 * h;
      ++depth;
    } else {
      h = heap_next(h, depth);
    }
  }
  if (valid) {
    ax[me] = sx;
    ay[me] = sy;
    az[me] = sz;
  }
  flush_counters<COUNT>(cnt, nt, na, np);
}

// Compensated FP32 accumulator (Kahan): sum s, running compensation c.
struct Kahan3 {
  float s[3] = {0.f, 0.f, 0.f};
  float c[3] = {0.f, 0.f, 0.f};
  __device__ __forceinline__ void add(float tx, float ty, float tz) {
    float t[3] = {tx, ty, tz};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      float y = t[i] - c[i];
      float u = s[i] + y;
      c[i] = (u - s[i]) - y;
      s[i] = u;
    }
  }
  __device__ __forceinline__ double get(int i) const { return (double)s[i] - (double)c[i]; }
};

template <bool COUNT>
__global__ __launch_bounds__(256) void k_traverse_mixed(const NodeRec32 *__restrict__ nodes,
                                                        const float4 *__restrict__ pos,
                                                        uint32_t n, float theta2,
                                                        uint32_t leaf,
                                                        double *__restrict__ ax,
                                                        double *__restrict__ ay,
                                                        double *__restrict__ az,
                                                        Counters cnt) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
  if (wbase >= n) return;
  const uint32_t me = wbase + lane;
  const bool valid = me < n;
  const float4 p = valid ? pos[me] : make_float4(0.f, 0.f, 0.f, 0.f);
  Kahan3 acc;
  int D = valid ? 0x7fffffff : -1;
  unsigned long long nt = 0, na = 0, np = 0;
  uint32_t h = 1;
  int depth = 0;
  while (h != 0) {
    const NodeRec32 r = nodes[uni(h) - 1];
    const bool active = depth <= D;
    if (r.count <= leaf) {
      if (active) {
        for (uint32_t jj = 0; jj < r.count; ++jj) {
          const uint32_t j = r.start + jj;
          const float4 q = pos[j];
          if (j != me) {
            float dx = p.x - q.x;
            float dy = p.y - q.y;
            float dz = p.z - q.z;
            float d2 = dx * dx + dy * dy + dz * dz;
            float ri = __builtin_amdgcn_rsqf(d2);
            float f = -q.w * (ri * ri * ri);
            acc.add(f * dx, f * dy, f * dz);
            if (COUNT) ++np;
          }
        }
      }
      h = heap_next(h, depth);
      continue;
    }
    bool open = false;
    if (active) {
      float dx = p.x - r.cx;
      float dy = p.y - r.cy;
      float dz = p.z - r.cz;
      float d2 = dx * dx + dy * dy + dz * dz;
      if (COUNT) ++nt;
      if (r.size2 < theta2 * d2) {
        float ri = __builtin_amdgcn_rsqf(d2);
        float f = -r.m * (ri * ri * ri);
        acc.add(f * dx, f * dy, f * dz);
        D = depth;
        if (COUNT) ++na;
      } else {
        D = 0x7fffffff;
        open = true;
      }
    }
    if (__any(open)) {
      h = 2 * h;
      ++depth;
    } else {
      h = heap_next(h, depth);
    }
  }
  if (valid) {
    ax[me] = acc.get(0);
    ay[me] = acc.get(1);
    az[me] = acc.get(2);
  }
  flush_counters<COUNT>(cnt, nt, na, np);
}

// Mixed precision with the coordinate difference taken in FP64.
template <bool COUNT>
__global__ __launch_bounds__(256) void k_traverse_mixed_d64(const NodeRecD *__restrict__ nodes,
                                                            const double4 *__restrict__ pos,
                                                            uint32_t n, float theta2,
                                                            uint32_t leaf,
                                                            double *__restrict__ ax,
                                                            double *__restrict__ ay,
                                                            double *__restrict__ az,
                                                            Counters cnt) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
  if (wbase >= n) return;
  const uint32_t me = wbase + lane;
  const bool valid = me < n;
  const double4 p = valid ? pos[me] : make_double4(0.0, 0.0, 0.0, 0.0);
  Kahan3 acc;
  int D = valid ? 0x7fffffff : -1;
  unsigned long long nt = 0, na = 0, np = 0;
  uint32_t h = 1;
  int depth = 0;
  while (h != 0) {
    const NodeRecD r = nodes[uni(h) - 1];
    const bool active = depth <= D;
    if (r.count <= leaf) {
      if (active) {
        for (uint32_t jj = 0; jj < r.count; ++jj) {
          const uint32_t j = r.start + jj;
          const double4 q = pos[j];
          if (j != me) {
            float dx = (float)(p.x - q.x);
            float dy = (float)(p.y - q.y);
            float dz = (float)(p.z - q.z);
            float d2 = dx * dx + dy * dy + dz * dz;
            float ri = __builtin_amdgcn_rsqf(d2);
            float f = -(float)q.w * (ri * ri * ri);
            acc.add(f * dx, f * dy, f * dz);
            if (COUNT) ++np;
          }
        }
      }
      h = heap_next(h, depth);
      continue;
    }
    bool open = false;
    if (active) {
      float dx = (float)(p.x - r.cx);
      float dy = (float)(p.y - r.cy);
      float dz = (float)(p.z - r.cz);
      float d2 = dx * dx + dy * dy + dz * dz;
      if (COUNT) ++nt;
      if (r.size2 < theta2 * d2) {
        float ri = __builtin_amdgcn_rsqf(d2);
        float f = -r.m * (ri * ri * ri);
        acc.add(f * dx, f * dy, f * dz);
        D = depth;
        if (COUNT) ++na;
      } else {
        D = 0x7fffffff;
        open = true;
      }
    }
    if (__any(open)) {
      h = 2 * h;
      ++depth;
    } else {
      h = heap_next(h, depth);
    }
  }
  if (valid) {
    ax[me] = acc.get(0);
    ay[me] = acc.get(1);
    az[me] = acc.get(2);
  }
  flush_counters<COUNT>(cnt, nt, na, np);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

hipError_t launch_traverse(int precision, bool count, const TreeDev &t,
                           const double4 *pos64, const float4 *pos32, uint32_t n,
                           double theta, int leaf, double *ax, double *ay, double *az,
                           unsigned long long *counters, hipStream_t s) {
  const unsigned blocks = (n + 255) / 256;
  Counters c{counters};
  const double theta2 = theta * theta;
  const float theta2f = (float)theta2;
  if (precision == 0) {
    if (count)
      k_traverse_f64<true><<<blocks, 256, 0, s>>>(t.rec64, pos64, n, theta2, leaf, ax, ay, az, c);
    else
      k_traverse_f64<false><<<blocks, 256, 0, s>>>(t.rec64, pos64, n, theta2, leaf, ax, ay, az, c);
  } else if (precision == 1) {
    if (count)
      k_traverse_mixed<true><<<blocks, 256, 0, s>>>(t.rec32, pos32, n, theta2f, leaf, ax, ay, az, c);
    else
      k_traverse_mixed<false><<<blocks, 256, 0, s>>>(t.rec32, pos32, n, theta2f, leaf, ax, ay, az, c);
  } else {
    if (count)
      k_traverse_mixed_d64<true><<<blocks, 256, 0, s>>>(t.recd, pos64, n, theta2f, leaf, ax, ay, az, c);
    else
      k_traverse_mixed_d64<false><<<blocks, 256, 0, s>>>(t.recd, pos64, n, theta2f, leaf, ax, ay, az, c);
  }
  return hipGetLastError();
}

}  // namespace kd
