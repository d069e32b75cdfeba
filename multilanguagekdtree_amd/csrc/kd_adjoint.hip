// kd_adjoint.hip -- reverse sweep of the kD-tree step on gfx950.
//
// The step (simple_sim, Parallel/RustVersion/src/array_kd_tree.rs:623-664)
// with its tree topology frozen: the tree of step k and every body's
// accept/open decisions are those of the forward run, so within the step the
// accelerations are a smooth function of positions p and masses m:
//   a_i = sum_{j in P2P(i)} f(p_i - p_j, m_j) + sum_{n in M2P(i)} f(p_i - c_n, M_n),
//   f(d, m) = -m d / |d|^3,  M_n = sum m_j,  c_n = sum m_j p_j / M_n.
// Backward through v' = v + dt a(p), p' = p + dt v' (array_kd_tree.rs:649-662):
//   lv += dt lp;  u = dt lv;  lp += (da/dp)^T u;  lm += (da/dm)^T u.
// k_adj_walk replays the forward walk (same decisions) and adds, per term,
//   d/dp_i (u_i . f) = -m (u_i / r^3 - 3 (d.u_i) d / r^5)  to body i,
// its negative to the partner body (P2P) or to the node's centre of mass
// (M2P), and -(u_i . d) / r^3 to the partner's mass or the node's mass.
// k_adj_nodes carries node terms to the members: dc/dp_j = m_j / M,
// dc/dm_j = (p_j - c) / M, dM/dm_j = 1.  All adjoint sums are FP64 (global
// FP64 atomics); the mixed path evaluates its terms in FP32 like its forward
// walk.  Adjoint arrays are in original body order.
#include <algorithm>

#include "kd_internal.hpp"

namespace kd {

__device__ __forceinline__ uint32_t uni_a(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

__device__ __forceinline__ uint32_t heap_next_a(uint32_t h, int &depth) {
  int t = __builtin_ctz(~h);
  h >>= t;
  depth -= t;
  return h == 0 ? 0 : h + 1;
}

// gd = d/dd (u . f(d, mq)), gm = d/dmq (u . f(d, mq)), FP64.
__device__ __forceinline__ void vjp64(double dx, double dy, double dz, double mq, double ux,
                                      double uy, double uz, double &gx, double &gy, double &gz,
                                      double &gm) {
#pragma clang fp contract(off)
  const double d2 = dx * dx + dy * dy + dz * dz;
  const double r = sqrt(d2);
  const double r3 = d2 * r;
  const double du = dx * ux + dy * uy + dz * uz;
  const double k = 3.0 * du / (r3 * d2);
  gx = -mq * (ux / r3 - k * dx);
  gy = -mq * (uy / r3 - k * dy);
  gz = -mq * (uz / r3 - k * dz);
  gm = -du / r3;
}

// The same in FP32 with a hardware reciprocal square root (mixed path).
__device__ __forceinline__ void vjp32(float dx, float dy, float dz, float mq, float ux, float uy,
                                      float uz, float &gx, float &gy, float &gz, float &gm) {
  const float d2 = dx * dx + dy * dy + dz * dz;
  const float ri = __builtin_amdgcn_rsqf(d2);
  const float ri2 = ri * ri;
  const float ri3 = ri2 * ri;
  const float du = dx * ux + dy * uy + dz * uz;
  const float k = 3.f * du * ri3 * ri2;
  gx = -mq * (ux * ri3 - k * dx);
  gy = -mq * (uy * ri3 - k * dy);
  gz = -mq * (uz * ri3 - k * dz);
  gm = -du * ri3;
}

struct AdjWalk {
  const NodeRec64 *rec64;  // one step's tree (heap order)
  const NodeRec32 *rec32;
  const double4 *pos64;    // tree order (x, y, z, m)
  const float4 *pos32;
  const uint32_t *oid;     // tree position -> original index
  const double *ux, *uy, *uz;  // seed, original order
  double *lpx, *lpy, *lpz, *lm;  // adjoint accumulators, original order
  double *gx, *gy, *gz, *gh;     // per node: d/dc_n and d/dM_n
  double *body[4];               // lpx, lpy, lpz, lm
  double *node[4];               // gx, gy, gz, gh
  uint32_t n;
  uint32_t leaf;
  double theta2;
};

// Reaction terms (to a partner body or to a node) are summed over the
// wavefront before their FP64 atomics.  Instead of four 64-lane reductions
// per term, a wavefront buffers up to EV terms in LDS, one row of 4 EV values
// per lane, then lane l sums column l over the 64 rows (in row order) and
// issues the one atomic for value l: the EV terms' 4 EV atomics go out as
// one instruction.  The buffer is private to the wavefront and every access
// happens in wave-uniform control flow, so no workgroup barrier is needed:
// the flush waits for the wavefront's own LDS writes (lgkmcnt 0) before the
// column reads, and the asm barriers keep the compiler from moving them.
constexpr int EV = 8;
constexpr int EV_ROW = 4 * EV + 2;  // doubles per lane row; rows stay 16-byte aligned
constexpr int ADJ_THREADS = 256;
constexpr uint32_t EV_NODE = 0x80000000u;  // tag bit: node arrays, else body arrays

struct EvBuf {
  double (*v)[EV_ROW];  // this wavefront's 64 rows
  uint32_t *tag;        // per buffered term: target index | EV_NODE
  uint32_t n;           // buffered terms (wave-uniform)
};

__device__ __forceinline__ void ev_flush(EvBuf &b, const AdjWalk &a) {
  const uint32_t lane = threadIdx.x & 63;
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt left at their maximum
  if (lane < 4 * b.n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll 4
    for (int k = 0; k < 64; k += 4) {
      s0 += b.v[k][lane];
      s1 += b.v[k + 1][lane];
      s2 += b.v[k + 2][lane];
      s3 += b.v[k + 3][lane];
    }
    const double sum = (s0 + s1) + (s2 + s3);
    const uint32_t tag = b.tag[lane >> 2], c = lane & 3u;
    double *const *arr = (tag & EV_NODE) ? a.node : a.body;
    atomicAdd(arr[c] + (tag & ~EV_NODE), sum);
  }
  asm volatile("" ::: "memory");
  b.n = 0;
}

// Add one term (this lane's four values; lanes without a term pass zeros).
// Called only where some lane has a term (wave-uniform).
__device__ __forceinline__ void ev_add(EvBuf &b, const AdjWalk &a, double x, double y, double z,
                                       double w, uint32_t tag) {
  const uint32_t lane = threadIdx.x & 63;
  double *row = b.v[lane] + 4 * b.n;
  row[0] = x;
  row[1] = y;
  row[2] = z;
  row[3] = w;
  if (lane == 0) b.tag[b.n] = tag;
  if (++b.n == (uint32_t)EV) ev_flush(b, a);
}

template <bool MIXED>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_walk(AdjWalk a) {
#pragma clang fp contract(off)
  __shared__ __align__(16) double ev_v[ADJ_THREADS / 64][64][EV_ROW];
  __shared__ uint32_t ev_t[ADJ_THREADS / 64][EV];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = (xcd_block() * blockDim.x + threadIdx.x) & ~63u;
  if (wbase >= a.n) return;
  const uint32_t me = wbase + lane;
  const bool valid = me < a.n;
  const uint32_t ms = valid ? me : wbase;
  const double4 p = a.pos64[ms];
  const float4 pf = a.pos32[ms];
  const uint32_t o = a.oid[ms];
  const double ux = valid ? a.ux[o] : 0.0, uy = valid ? a.uy[o] : 0.0,
               uz = valid ? a.uz[o] : 0.0;
  const float uxf = (float)ux, uyf = (float)uy, uzf = (float)uz;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  int D = valid ? 0x7fffffff : -1;
  EvBuf eb;
  eb.v = ev_v[threadIdx.x >> 6];
  eb.tag = ev_t[threadIdx.x >> 6];
  eb.n = 0;
  uint32_t h = 1;
  int depth = 0;
  while (h != 0) {
    const NodeRec64 r = a.rec64[uni_a(h) - 1];
    const bool active = depth <= D;
    if (r.count <= a.leaf) {
      for (uint32_t jj = 0; jj < r.count; ++jj) {
        const uint32_t j = r.start + jj;
        const bool use = active && j != me;
        double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0;
        if (use) {
          if (MIXED) {
            const float4 q = a.pos32[j];
            float fx, fy, fz, fm;
            vjp32(pf.x - q.x, pf.y - q.y, pf.z - q.z, q.w, uxf, uyf, uzf, fx, fy, fz, fm);
            gx = fx, gy = fy, gz = fz, gm = fm;
          } else {
            const double4 q = a.pos64[j];
            vjp64(p.x - q.x, p.y - q.y, p.z - q.z, q.w, ux, uy, uz, gx, gy, gz, gm);
          }
          sx += gx;
          sy += gy;
          sz += gz;
        }
        if (__any(use)) ev_add(eb, a, -gx, -gy, -gz, gm, uni_a(a.oid[j]));
      }
      h = heap_next_a(h, depth);
      continue;
    }
    bool ok;
    double dx, dy, dz;
    float fdx = 0.f, fdy = 0.f, fdz = 0.f;
    if (MIXED) {
      const NodeRec32 q = a.rec32[h - 1];
      fdx = pf.x - q.cx;
      fdy = pf.y - q.cy;
      fdz = pf.z - q.cz;
      const float d2 = fdx * fdx + fdy * fdy + fdz * fdz;
      ok = q.size2 < d2;  // size^2 / theta^2 < d2, as the forward walk
      dx = dy = dz = 0.0;
    } else {
      dx = p.x - r.cx;
      dy = p.y - r.cy;
      dz = p.z - r.cz;
      const double d2 = dx * dx + dy * dy + dz * dz;
      ok = r.size2 < a.theta2 * d2;
    }
    const bool take = active && ok;
    const bool open = active && !ok;
    double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0;
    if (take) {
      if (MIXED) {
        float fx, fy, fz, fm;
        vjp32(fdx, fdy, fdz, (float)r.m, uxf, uyf, uzf, fx, fy, fz, fm);
        gx = fx, gy = fy, gz = fz, gm = fm;
      } else {
        vjp64(dx, dy, dz, r.m, ux, uy, uz, gx, gy, gz, gm);
      }
      sx += gx;
      sy += gy;
      sz += gz;
    }
    if (__any(take)) ev_add(eb, a, -gx, -gy, -gz, gm, (h - 1) | EV_NODE);
    D = take ? depth : (open ? 0x7fffffff : D);
    if (__any(open)) {
      h = 2 * h;
      ++depth;
    } else {
      h = heap_next_a(h, depth);
    }
  }
  if (eb.n) ev_flush(eb, a);
  if (valid) {
    atomicAdd(a.lpx + o, sx);
    atomicAdd(a.lpy + o, sy);
    atomicAdd(a.lpz + o, sz);
  }
}

// Node terms to the members.  One thread per heap slot that is a leaf: the
// sums over its internal ancestors of A = G/M and B = H - c.A give
//   lp_j += m_j A,  lm_j += p_j . A + B  for every member j.
__global__ void k_adj_nodes(const uint32_t *__restrict__ seg_count,
                            const NodeRec64 *__restrict__ rec64,
                            const double4 *__restrict__ pos64, const uint32_t *__restrict__ oid,
                            const double *__restrict__ gx, const double *__restrict__ gy,
                            const double *__restrict__ gz, const double *__restrict__ gh,
                            uint32_t slots, uint32_t leaf, double *lpx, double *lpy, double *lpz,
                            double *lm) {
#pragma clang fp contract(off)
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= slots) return;
  const uint32_t c = seg_count[h];
  if (c == 0 || c > leaf) return;
  if (h > 0 && seg_count[(h - 1) / 2] <= leaf) return;  // below a leaf: unused slot
  double ax = 0.0, ay = 0.0, az = 0.0, b = 0.0;
  for (uint32_t q = h; q > 0;) {
    q = (q - 1) / 2;
    if (gx[q] == 0.0 && gy[q] == 0.0 && gz[q] == 0.0 && gh[q] == 0.0) continue;
    const NodeRec64 r = rec64[q];
    const double Ax = gx[q] / r.m, Ay = gy[q] / r.m, Az = gz[q] / r.m;
    ax += Ax;
    ay += Ay;
    az += Az;
    b += gh[q] - (r.cx * Ax + r.cy * Ay + r.cz * Az);
  }
  if (ax == 0.0 && ay == 0.0 && az == 0.0 && b == 0.0) return;
  const NodeRec64 lr = rec64[h];
  for (uint32_t t = lr.start; t < lr.start + lr.count; ++t) {
    const double4 p = pos64[t];
    const uint32_t o = oid[t];
    lpx[o] += p.w * ax;
    lpy[o] += p.w * ay;
    lpz[o] += p.w * az;
    lm[o] += p.x * ax + p.y * ay + p.z * az + b;
  }
}

// lv += dt lp;  u = dt lv  (original order).
__global__ void k_adj_integrate(double *lpx, double *lpy, double *lpz, double *lvx, double *lvy,
                                double *lvz, double *ux, double *uy, double *uz, uint32_t n,
                                double dt) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = lvx[i] + dt * lpx[i], b = lvy[i] + dt * lpy[i], c = lvz[i] + dt * lpz[i];
  lvx[i] = a;
  lvy[i] = b;
  lvz[i] = c;
  ux[i] = dt * a;
  uy[i] = dt * b;
  uz[i] = dt * c;
}

__global__ void k_zero(double *p, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

hipError_t launch_adj_step(const AdjStep &st, hipStream_t s) {
  const uint32_t n = st.n;
  const unsigned nb = (n + 255) / 256;
  k_adj_integrate<<<nb, 256, 0, s>>>(st.lp[0], st.lp[1], st.lp[2], st.lv[0], st.lv[1], st.lv[2],
                                     st.u[0], st.u[1], st.u[2], n, st.dt);
  const size_t zc = 4 * (size_t)st.slots;
  k_zero<<<(unsigned)std::min<size_t>((zc + 255) / 256, 4096), 256, 0, s>>>(st.gnode, zc);
  AdjWalk a;
  a.rec64 = st.rec64;
  a.rec32 = st.rec32;
  a.pos64 = st.pos64;
  a.pos32 = st.pos32;
  a.oid = st.oid;
  a.ux = st.u[0];
  a.uy = st.u[1];
  a.uz = st.u[2];
  a.lpx = st.lp[0];
  a.lpy = st.lp[1];
  a.lpz = st.lp[2];
  a.lm = st.lm;
  a.gx = st.gnode;
  a.gy = st.gnode + st.slots;
  a.gz = st.gnode + 2 * (size_t)st.slots;
  a.gh = st.gnode + 3 * (size_t)st.slots;
  a.body[0] = a.lpx;
  a.body[1] = a.lpy;
  a.body[2] = a.lpz;
  a.body[3] = a.lm;
  a.node[0] = a.gx;
  a.node[1] = a.gy;
  a.node[2] = a.gz;
  a.node[3] = a.gh;
  a.n = n;
  a.leaf = st.leaf;
  a.theta2 = st.theta * st.theta;
  const unsigned wb = (n + ADJ_THREADS - 1) / ADJ_THREADS;
  if (st.mixed)
    k_adj_walk<true><<<wb, ADJ_THREADS, 0, s>>>(a);
  else
    k_adj_walk<false><<<wb, ADJ_THREADS, 0, s>>>(a);
  k_adj_nodes<<<(st.slots + 255) / 256, 256, 0, s>>>(st.seg_count, st.rec64, st.pos64, st.oid,
                                                     a.gx, a.gy, a.gz, a.gh, st.slots, st.leaf,
                                                     st.lp[0], st.lp[1], st.lp[2], st.lm);
  return hipGetLastError();
}

}  // namespace kd
