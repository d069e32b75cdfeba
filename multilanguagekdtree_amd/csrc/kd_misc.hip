// kd_misc.hip -- integrator, reordering and checking kernels.
#include <algorithm>
#include <vector>

#include "kd_internal.hpp"

namespace kd {

// Kick-drift, Parallel/RustVersion/src/array_kd_tree.rs:649-662 (semi-implicit
// Euler, unfused FP64).  Flags a non-finite result in *bad.  With range set,
// also reduces the new positions' coordinate range for the next build (the
// grid is then capped at RANGE_BLOCKS blocks and strides over the bodies).
__global__ __launch_bounds__(256) void k_kick_drift(StateDev s, const double *__restrict__ ax,
                                                    const double *__restrict__ ay,
                                                    const double *__restrict__ az, uint32_t n,
                                                    double dt, int *bad,
                                                    unsigned long long *range) {
#pragma clang fp contract(off)
  __shared__ uint64_t sh[6][4];
  uint64_t lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double vx = s.vx[i] + dt * ax[i];
    double vy = s.vy[i] + dt * ay[i];
    double vz = s.vz[i] + dt * az[i];
    double x = s.x[i] + dt * vx;
    double y = s.y[i] + dt * vy;
    double z = s.z[i] + dt * vz;
    s.vx[i] = vx;
    s.vy[i] = vy;
    s.vz[i] = vz;
    s.x[i] = x;
    s.y[i] = y;
    s.z[i] = z;
    if (!(isfinite(x) && isfinite(y) && isfinite(z))) *bad = 1;
    if (range) {
      const uint64_t k[3] = {ordkey(x), ordkey(y), ordkey(z)};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        lo[d] = k[d] < lo[d] ? k[d] : lo[d];
        hi[d] = k[d] > hi[d] ? k[d] : hi[d];
      }
    }
  }
  if (range) range_commit(lo, hi, sh, range);  // range is the same for the whole grid
}

// out[oid[t]] = a[t] for three components.
__global__ void k_to_original(const uint32_t *__restrict__ oid, const double *__restrict__ a0,
                              const double *__restrict__ a1, const double *__restrict__ a2,
                              double *__restrict__ o0, double *__restrict__ o1,
                              double *__restrict__ o2, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t o = oid[i];
  if (o0) o0[o] = a0[i];
  if (o1) o1[o] = a1[i];
  if (o2) o2[o] = a2[i];
}

__global__ void k_inverse(const uint32_t *__restrict__ oid, uint32_t *__restrict__ inv,
                          uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[oid[i]] = i;
}

__global__ void k_iota(uint32_t *__restrict__ v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

// Device-side upload check: -0.0 -> +0.0 for positions, flags non-finite.
__global__ void k_sanitize(double *x, double *y, double *z, const double *vx,
                           const double *vy, const double *vz, const double *m,
                           uint32_t n, int *bad) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = x[i], b = y[i], c = z[i];
  if (a == 0.0) x[i] = 0.0;
  if (b == 0.0) y[i] = 0.0;
  if (c == 0.0) z[i] = 0.0;
  if (!(isfinite(a) && isfinite(b) && isfinite(c) && isfinite(vx[i]) && isfinite(vy[i]) &&
        isfinite(vz[i]) && isfinite(m[i])))
    *bad = 1;
}

// FP64 direct sum (calc_pp_accel form, Sequential/RustVersion/src/
// array_particle.rs:67-76) for sample bodies at working positions wpos[s].
// One block per sample; per-thread partial sums combined in a fixed tree.
__global__ __launch_bounds__(256) void k_direct(const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                const double *__restrict__ z,
                                                const double *__restrict__ m, uint32_t n,
                                                const uint32_t *__restrict__ wpos,
                                                double *__restrict__ out /* k*6 */) {
#pragma clang fp contract(off)
  __shared__ double sh[6][256];
  const uint32_t i = wpos[blockIdx.x];
  const double px = x[i], py = y[i], pz = z[i];
  double a[3] = {0.0, 0.0, 0.0}, ab[3] = {0.0, 0.0, 0.0};
  for (uint32_t j = threadIdx.x; j < n; j += 256) {
    if (j == i) continue;
    double dx = px - x[j];
    double dy = py - y[j];
    double dz = pz - z[j];
    double dp2 = dx * dx + dy * dy + dz * dz;
    double dist = sqrt(dp2);
    double magi = -m[j] / (dist * dist * dist);
    double t0 = magi * dx, t1 = magi * dy, t2 = magi * dz;
    a[0] += t0;
    a[1] += t1;
    a[2] += t2;
    ab[0] += fabs(t0);
    ab[1] += fabs(t1);
    ab[2] += fabs(t2);
  }
  for (int c = 0; c < 3; ++c) {
    sh[c][threadIdx.x] = a[c];
    sh[3 + c][threadIdx.x] = ab[c];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 6; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 6) out[(size_t)blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

hipError_t launch_kick_drift(const StateDev &s, const double *ax, const double *ay,
                             const double *az, uint32_t n, double dt, int *bad,
                             hipStream_t st, unsigned long long *range) {
  uint32_t blocks = (n + 255) / 256;
  if (range) {
    hipError_t e = launch_range_reset(range, st);
    if (e != hipSuccess) return e;
    blocks = std::min<uint32_t>(blocks, RANGE_BLOCKS);
  }
  k_kick_drift<<<blocks, 256, 0, st>>>(s, ax, ay, az, n, dt, bad, range);
  return hipGetLastError();
}

hipError_t launch_to_original(const uint32_t *oid, const double *a0, const double *a1,
                              const double *a2, double *o0, double *o1, double *o2,
                              uint32_t n, hipStream_t st) {
  k_to_original<<<(n + 255) / 256, 256, 0, st>>>(oid, a0, a1, a2, o0, o1, o2, n);
  return hipGetLastError();
}

hipError_t launch_inverse(const uint32_t *oid, uint32_t *inv, uint32_t n, hipStream_t st) {
  k_inverse<<<(n + 255) / 256, 256, 0, st>>>(oid, inv, n);
  return hipGetLastError();
}

hipError_t launch_iota(uint32_t *v, uint32_t n, hipStream_t st) {
  k_iota<<<(n + 255) / 256, 256, 0, st>>>(v, n);
  return hipGetLastError();
}

hipError_t launch_sanitize(const StateDev &s, uint32_t n, int *bad, hipStream_t st) {
  k_sanitize<<<(n + 255) / 256, 256, 0, st>>>(s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m, n, bad);
  return hipGetLastError();
}

hipError_t launch_direct(const StateDev &s, uint32_t n, const uint32_t *wpos, uint32_t k,
                         double *out, hipStream_t st) {
  k_direct<<<k, 256, 0, st>>>(s.x, s.y, s.z, s.m, n, wpos, out);
  return hipGetLastError();
}

}  // namespace kd

namespace kd {

// Fill of a 2D region (the work of a memset node): width elements of
// elem_size bytes per row, height rows pitch bytes apart.
__global__ void k_fill(char *dst, unsigned elem_size, size_t width, size_t height, size_t pitch,
                       unsigned value) {
  const size_t total = width * height;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    char *p = dst + (i / width) * pitch + (i % width) * elem_size;
    if (elem_size == 1)
      *reinterpret_cast<uint8_t *>(p) = (uint8_t)value;
    else if (elem_size == 2)
      *reinterpret_cast<uint16_t *>(p) = (uint16_t)value;
    else
      *reinterpret_cast<uint32_t *>(p) = value;
  }
}

hipError_t graph_replace_memsets(hipGraph_t g, int *replaced) {
  *replaced = 0;
  size_t nn = 0;
  hipError_t e = hipGraphGetNodes(g, nullptr, &nn);
  if (e != hipSuccess) return e;
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn && (e = hipGraphGetNodes(g, nodes.data(), &nn)) != hipSuccess) return e;
  for (hipGraphNode_t m : nodes) {
    hipGraphNodeType ty;
    if ((e = hipGraphNodeGetType(m, &ty)) != hipSuccess) return e;
    if (ty != hipGraphNodeTypeMemset) continue;
    hipMemsetParams mp;
    if ((e = hipGraphMemsetNodeGetParams(m, &mp)) != hipSuccess) return e;
    if (mp.elementSize != 1 && mp.elementSize != 2 && mp.elementSize != 4)
      return hipErrorInvalidValue;
    size_t nd = 0, nt = 0;
    if ((e = hipGraphNodeGetDependencies(m, nullptr, &nd)) != hipSuccess) return e;
    if ((e = hipGraphNodeGetDependentNodes(m, nullptr, &nt)) != hipSuccess) return e;
    std::vector<hipGraphNode_t> deps(nd), dents(nt);
    if (nd && (e = hipGraphNodeGetDependencies(m, deps.data(), &nd)) != hipSuccess) return e;
    if (nt && (e = hipGraphNodeGetDependentNodes(m, dents.data(), &nt)) != hipSuccess) return e;
    char *dst = static_cast<char *>(mp.dst);
    unsigned elem = mp.elementSize;
    size_t width = mp.width, height = mp.height ? mp.height : 1, pitch = mp.pitch;
    unsigned value = mp.value;
    void *args[] = {&dst, &elem, &width, &height, &pitch, &value};
    const size_t total = width * height;
    hipKernelNodeParams kp{};
    kp.func = reinterpret_cast<void *>(&k_fill);
    kp.blockDim = dim3(256);
    kp.gridDim = dim3((unsigned)std::min<size_t>((total + 255) / 256 + 1, 4096));
    kp.kernelParams = args;
    kp.extra = nullptr;
    kp.sharedMemBytes = 0;
    hipGraphNode_t k = nullptr;
    if ((e = hipGraphAddKernelNode(&k, g, deps.data(), nd, &kp)) != hipSuccess) return e;
    for (hipGraphNode_t d : dents)
      if ((e = hipGraphAddDependencies(g, &k, &d, 1)) != hipSuccess) return e;
    if ((e = hipGraphDestroyNode(m)) != hipSuccess) return e;
    ++*replaced;
  }
  return hipSuccess;
}

}  // namespace kd
