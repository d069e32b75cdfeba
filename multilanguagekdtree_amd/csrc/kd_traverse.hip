// kd_traverse.hip -- theta-opened force traversal on gfx950.
//
// One wavefront serves 64 consecutive bodies in tree order.  It walks the
// heap-ordered tree depth first without a stack: the current node h
// (1-based) is wave-uniform, so node records and leaf bodies arrive through
// scalar loads; descending goes to 2h, finishing a subtree goes to
// (h >> trailing_ones(h)) + 1.  Each lane keeps the level D (root 1) at which
// it last accepted a node (~0u while it keeps opening): a lane takes part in
// node h iff level(h) <= D, so every lane sees exactly the nodes and leaves its
// own recursive traversal (Sequential/RustVersion/src/array_kd_tree.rs:
// 140-172) sees.  The wave descends when any active lane opens.
//
// Arithmetic policies:
//   F64   the reference arithmetic, unfused: accept iff size*size <
//         (theta*theta)*d2, M2P -m/(d2*sqrt(d2)), P2P -m/(dist*dist*dist),
//         FP64 sums (array_kd_tree.rs:600-613, array_particle.rs:67-76);
//   F32   FP32 coordinates, v_rsq_f32, opening threshold size^2/theta^2
//         baked into the record, blocked compensated (Kahan) FP32 sums:
//         terms accumulate in a plain FP32 partial that is folded into the
//         compensated sum after every leaf (the partial then carries the
//         negated compensation into the next block); the self pair adds 0
//         through a per-body floor on d2 instead of an index compare;
//   D64   as F32 but the coordinate difference is taken in FP64 (self pair
//         dropped by index, as F64).
#include <algorithm>

#include "kd_internal.hpp"

namespace kd {

struct Counters {
  unsigned long long *c;  // node tests, M2P, P2P, wave node visits, wave leaf visits
};

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ uint32_t uni(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

// One node record as a single wide value (a vector of 32-bit words), so that
// the scalar unit fetches it with one load instead of per-field loads.
// nodes_m1 points one record before the first, so the 1-based heap index h
// is the byte offset h * sizeof(Rec).
template <class Rec>
__device__ __forceinline__ Rec load_rec_at(const Rec *__restrict__ nodes_m1, uint32_t off) {
  typedef uint32_t words __attribute__((ext_vector_type(sizeof(Rec) / 4)));
  const char *base = reinterpret_cast<const char *>(nodes_m1);
  const words w = *reinterpret_cast<const words *>(base + off);
  Rec r;
  __builtin_memcpy(&r, &w, sizeof(Rec));
  return r;
}

// Advance past the subtree of h; returns 0 when the root is finished.
__device__ __forceinline__ uint32_t heap_next(uint32_t h, int &depth) {
  int t = __builtin_ctz(~h);
  h >>= t;
  depth -= t;
  return h == 0 ? 0 : h + 1;
}

// ---------------------------------------------------------------------------
// arithmetic policies
// ---------------------------------------------------------------------------

struct F64 {
  static constexpr int CHUNK = 4;  // leaf bodies per batch of scalar loads
  static constexpr bool SELF_BY_INDEX = true;  // self pair dropped by a select
  static constexpr bool COMPACT = false;  // leaves are never split across lanes
  using Rec = NodeRec64;
  using Body = double4;
  using Self = double4;
  struct Acc {
    double s[3] = {0.0, 0.0, 0.0};
  };
  struct Diff {
    double dx, dy, dz, d2;
  };
  __device__ static Self self(const Body &b) { return b; }
  __device__ static bool test(const Self &p, const Rec &r, double theta2, Diff &f) {
#pragma clang fp contract(off)
    f.dx = p.x - r.cx;
    f.dy = p.y - r.cy;
    f.dz = p.z - r.cz;
    f.d2 = f.dx * f.dx + f.dy * f.dy + f.dz * f.dz;
    return r.size2 < theta2 * f.d2;
  }
  __device__ static void m2p(Acc &a, const Rec &r, const Diff &f) {
#pragma clang fp contract(off)
    double dist = sqrt(f.d2);
    double magi = -r.m / (f.d2 * dist);
    a.s[0] += f.dx * magi;
    a.s[1] += f.dy * magi;
    a.s[2] += f.dz * magi;
  }
  __device__ static void p2p(Acc &a, const Self &p, const Body &q, bool self_pair) {
#pragma clang fp contract(off)
    double dx = p.x - q.x;
    double dy = p.y - q.y;
    double dz = p.z - q.z;
    double dp2 = dx * dx + dy * dy + dz * dz;
    double dist = sqrt(dp2);
    double magi = -q.w / (dist * dist * dist);
    magi = self_pair ? 0.0 : magi;  // self is skipped: adds exactly +0
    a.s[0] += magi * dx;
    a.s[1] += magi * dy;
    a.s[2] += magi * dz;
  }
  __device__ static void m2p_sel(Acc &a, const Rec &r, const Diff &f, bool take) {
    if (take) m2p(a, r, f);
  }
  __device__ static void fold(Acc &) {}
  __device__ static double result(const Acc &a, int i) { return a.s[i]; }
};

// Blocked Kahan summation with the compensation carried in the partial:
// after a fold the partial holds -c (the rounding error of the last sum,
// negated), and the next block of terms is added on top of it.  A fold is
// t = s + p; p -= t - s; s = t: three operations per component and no reset.
struct KahanAcc {
  float p[3] = {0.f, 0.f, 0.f};  // the current block of terms minus the compensation
  float s[3] = {0.f, 0.f, 0.f};  // compensated sum
};

__device__ __forceinline__ void kahan_fold(KahanAcc &a) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float t = a.s[i] + a.p[i];
    a.p[i] -= t - a.s[i];
    a.s[i] = t;
  }
}

// F32 pair terms take d2 + e2 with a per-body floor e2 (added inside the
// first FMA, so it costs nothing).  The self pair (d = 0 exactly) then gives
// finite g times 0 = 0 with no compare or select: e2 >= max(m, 1) 2^-84 keeps
// m r^-3 below 2^126.  e2 >= (|p|_max 2^-30)^2 places the floor 64 times below
// the FP32 resolution of the body's own coordinates, so distinct pairs move
// by less than their coordinate rounding already moves them.
struct F32 {
  static constexpr int CHUNK = 8;
  static constexpr bool SELF_BY_INDEX = false;
  static constexpr bool COMPACT = true;
  using Rec = NodeRec32;
  using Body = float4;
  using Self = float4;  // x, y, z, e2
  using Acc = KahanAcc;
  struct Diff {
    float dx, dy, dz, d2;
  };
  __device__ static Self self(const Body &b) {
    const float a = fmaxf(fabsf(b.x), fmaxf(fabsf(b.y), fabsf(b.z))) * 0x1p-30f;
    return make_float4(b.x, b.y, b.z, fmaxf(a * a, fmaxf(b.w, 1.f) * 0x1p-84f));
  }
  __device__ static bool test(const Self &p, const Rec &r, double, Diff &f) {
    f.dx = p.x - r.cx;
    f.dy = p.y - r.cy;
    f.dz = p.z - r.cz;
    f.d2 = f.dx * f.dx + f.dy * f.dy + f.dz * f.dz;
    return r.size2 < f.d2;  // size2 holds size^2 / theta^2
  }
  // Every lane evaluates the term; lanes that do not accept add +0.  This
  // keeps the walk free of divergent branches (no exec-mask save/restore).
  __device__ static void m2p_sel(Acc &a, const Rec &r, const Diff &f, bool take) {
    float ri = __builtin_amdgcn_rsqf(f.d2);
    float g = -r.m * (ri * ri * ri);
    g = take ? g : 0.f;
    a.p[0] += g * f.dx;
    a.p[1] += g * f.dy;
    a.p[2] += g * f.dz;
  }
  __device__ static void p2p(Acc &a, const Self &p, const Body &q, bool skip) {
    float dx = p.x - q.x;
    float dy = p.y - q.y;
    float dz = p.z - q.z;
    float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, p.w)));
    float ri = __builtin_amdgcn_rsqf(d2);
    float g = -q.w * (ri * ri * ri);
    g = skip ? 0.f : g;  // padding of a predicated batch; constant false otherwise
    a.p[0] += g * dx;
    a.p[1] += g * dy;
    a.p[2] += g * dz;
  }
  __device__ static void fold(Acc &a) { kahan_fold(a); }
  __device__ static double result(const Acc &a, int i) {
    return (double)a.s[i] + (double)a.p[i];
  }
};

struct D64 {
  static constexpr int CHUNK = 4;
  static constexpr bool SELF_BY_INDEX = true;
  static constexpr bool COMPACT = false;
  using Rec = NodeRecD;
  using Body = double4;
  using Self = double4;
  using Acc = KahanAcc;
  struct Diff {
    float dx, dy, dz, d2;
  };
  __device__ static Self self(const Body &b) { return b; }
  __device__ static bool test(const Self &p, const Rec &r, double, Diff &f) {
    f.dx = (float)(p.x - r.cx);
    f.dy = (float)(p.y - r.cy);
    f.dz = (float)(p.z - r.cz);
    f.d2 = f.dx * f.dx + f.dy * f.dy + f.dz * f.dz;
    return r.size2 < f.d2;
  }
  // Every lane evaluates the term; lanes that do not accept add +0.  This
  // keeps the walk free of divergent branches (no exec-mask save/restore).
  __device__ static void m2p_sel(Acc &a, const Rec &r, const Diff &f, bool take) {
    float ri = __builtin_amdgcn_rsqf(f.d2);
    float g = -r.m * (ri * ri * ri);
    g = take ? g : 0.f;
    a.p[0] += g * f.dx;
    a.p[1] += g * f.dy;
    a.p[2] += g * f.dz;
  }
  __device__ static void p2p(Acc &a, const Self &p, const Body &q, bool self_pair) {
    float dx = (float)(p.x - q.x);
    float dy = (float)(p.y - q.y);
    float dz = (float)(p.z - q.z);
    float d2 = dx * dx + dy * dy + dz * dz;
    float ri = __builtin_amdgcn_rsqf(d2);
    float g = -(float)q.w * (ri * ri * ri);
    g = self_pair ? 0.f : g;
    a.p[0] += g * dx;
    a.p[1] += g * dy;
    a.p[2] += g * dz;
  }
  __device__ static void fold(Acc &a) { kahan_fold(a); }
  __device__ static double result(const Acc &a, int i) {
    return (double)a.s[i] + (double)a.p[i];
  }
};

template <class P>
struct WalkArgs {
  const typename P::Rec *__restrict__ nodes;
  const typename P::Body *__restrict__ bodies;  // tree order, padded by MAX_LEAF
  uint32_t n;
  uint32_t leaf;
  double theta2;
  double *__restrict__ ax;
  double *__restrict__ ay;
  double *__restrict__ az;
  Counters cnt;
  KickDrift kd{};  // used by the walks that apply the kick-drift themselves
};

// A body's state, loaded when its wavefront starts so that the loads are
// done by the time the walk ends.  The state is read and written with
// non-temporal accesses, which keep its 96 bytes a body from displacing the
// tree's records and bodies in L2: at 10 M ring bodies the fused mixed step
// takes 13.38 ms against 13.59 ms with ordinary accesses (the walk 10.29 ms
// against 10.50) and 13.66 ms with the separate kick-drift pass.
struct BodyState {
  double x, y, z, vx, vy, vz;
};
__device__ __forceinline__ BodyState body_state(const KickDrift &k, uint32_t i) {
  return BodyState{__builtin_nontemporal_load(k.s.x + i),  __builtin_nontemporal_load(k.s.y + i),
                   __builtin_nontemporal_load(k.s.z + i),  __builtin_nontemporal_load(k.s.vx + i),
                   __builtin_nontemporal_load(k.s.vy + i), __builtin_nontemporal_load(k.s.vz + i)};
}

// v += dt a; p += dt v for body i, unfused FP64 exactly as k_kick_drift
// (array_kd_tree.rs:649-662); the updated position's ordered keys go into
// the lane's range.
__device__ __forceinline__ void kick_drift_one(const KickDrift &k, uint32_t i, const BodyState &b,
                                               double ax, double ay, double az, uint64_t lo[3],
                                               uint64_t hi[3]) {
#pragma clang fp contract(off)
  const double dt = k.dt;
  const double vx = b.vx + dt * ax;
  const double vy = b.vy + dt * ay;
  const double vz = b.vz + dt * az;
  const double x = b.x + dt * vx;
  const double y = b.y + dt * vy;
  const double z = b.z + dt * vz;
  __builtin_nontemporal_store(vx, k.s.vx + i);
  __builtin_nontemporal_store(vy, k.s.vy + i);
  __builtin_nontemporal_store(vz, k.s.vz + i);
  __builtin_nontemporal_store(x, k.s.x + i);
  __builtin_nontemporal_store(y, k.s.y + i);
  __builtin_nontemporal_store(z, k.s.z + i);
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) *k.bad = 1;
  const uint64_t kk[3] = {ordkey(x), ordkey(y), ordkey(z)};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = kk[d];
    hi[d] = kk[d];
  }
}

// The wavefront's range (lanes without a body hold the empty range) into
// its 6 words.
__device__ __forceinline__ void wave_range_store(unsigned long long *w, uint64_t lo[3],
                                                 uint64_t hi[3], uint32_t lane) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t a = __shfl_xor(lo[d], o, 64), b = __shfl_xor(hi[d], o, 64);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  }
  if (lane < 6) {
    const uint64_t v = lane == 0   ? lo[0]
                       : lane == 1 ? lo[1]
                       : lane == 2 ? lo[2]
                       : lane == 3 ? hi[0]
                       : lane == 4 ? hi[1]
                                   : hi[2];
    w[lane] = v;
  }
}

__global__ __launch_bounds__(256) void k_range_waves(const unsigned long long *__restrict__ w,
                                                     uint32_t nw, unsigned long long *range) {
  __shared__ uint64_t sh[6][4];
  uint64_t lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const uint64_t a = w[6 * (size_t)i + d], b = w[6 * (size_t)i + 3 + d];
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  }
  range_commit(lo, hi, sh, range);
}

hipError_t launch_range_from_waves(const unsigned long long *wave_range, uint32_t nw,
                                   unsigned long long *range, hipStream_t s) {
  hipError_t e = launch_range_reset(range, s);
  if (e != hipSuccess) return e;
  const uint32_t blocks = std::min<uint32_t>((nw + 255) / 256, RANGE_BLOCKS);
  k_range_waves<<<std::max(blocks, 1u), 256, 0, s>>>(wave_range, nw, range);
  return hipGetLastError();
}

// P2P of the lane's body with n (1..B) consecutive leaf bodies starting at
// tree position s0.  The batch loads are unconditional (the body arrays
// carry BODY_PAD spare entries); terms are added in leaf order.
template <class P, int B, bool COUNT, bool PRED>
__device__ __forceinline__ void leaf_batch(typename P::Acc &acc, const typename P::Self &p,
                                           const typename P::Body *__restrict__ base,
                                           uint32_t n, uint32_t s0, uint32_t me, bool active,
                                           unsigned long long &np) {
  typename P::Body q[B];
#pragma unroll
  for (int j = 0; j < B; ++j) q[j] = base[j];
  if (!active) return;
  if (PRED) {
    // every body of the batch is computed; those past n contribute +0
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const bool use = (uint32_t)j < n;
      P::p2p(acc, p, q[j], !use || (P::SELF_BY_INDEX && s0 + j == me));
      if (COUNT) np += (use && s0 + j != me) ? 1 : 0;
    }
    return;
  }
  // one wave-uniform branch per body, terms in leaf order; each term updates
  // the accumulators in place, so the joins need no register copies
#pragma unroll
  for (int j = 0; j < B; ++j) {
    if ((uint32_t)j < n) {
      P::p2p(acc, p, q[j], P::SELF_BY_INDEX && s0 + j == me);
      if (COUNT) np += (s0 + j != me) ? 1 : 0;
    }
  }
}

// Lane i's value of v from the lane that the DPP control CTRL names (0 where
// that lane is outside the row).
template <int CTRL>
__device__ __forceinline__ float dpp_from(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}

// Leaf work split across the wavefront when at most 8 of its lanes take
// part in a leaf of c bodies.  The active lanes are compacted through the
// wavefront's 64 LDS slots: the active lane of rank o (its place among the
// active lanes) puts its position in slot o, and lane l then computes, for
// owner o = l / 8, the terms of leaf bodies l mod 8, l mod 8 + 8, ...:
// ceil(c / 8) terms a lane instead of c.  The 8 lanes of an owner sum their
// partials across lanes (DPP within a row of 16), the last of them puts the
// sum in the owner's slot, and the owner adds it to its accumulator.  A
// leaf's terms are thus summed per body slot and then across slots, not in
// leaf order: this moves only the FP32 rounding of the leaf's sum.  Lanes
// whose owner slot is not an active lane's compute terms from stale slot
// values that no owner reads.  Body loads past the leaf (at most 7 bodies)
// stay inside the body arrays, which carry BODY_PAD spare entries past the
// last body, and their terms are selected out.  The terms are F32::p2p's,
// with the self pair adding 0 through the floor e2.
constexpr uint32_t SPLIT_MAX = 8;  // owners of a split leaf; also its bodies a step

template <class P, bool COUNT>
__device__ __forceinline__ void leaf_split(typename P::Acc &acc, const float4 &p,
                                           const float4 *__restrict__ base, uint32_t c,
                                           uint32_t s0, uint32_t me, uint64_t actm, bool active,
                                           float4 *__restrict__ slot, unsigned long long &np) {
  constexpr uint32_t B = SPLIT_MAX;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t b = lane & (B - 1u);  // the lane's body slot
  const float4 q0 = base[b];           // loads while the positions are exchanged
  const uint32_t rank =
      __builtin_amdgcn_mbcnt_hi((uint32_t)(actm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)actm, 0u));
  wave_sync();  // the previous leaf's reads of the slots come first
  if (active) slot[rank] = p;
  wave_sync();
  float *mine = (float *)(slot + lane / B);
  const float4 own = *(const float4 *)mine;
  float dx, dy, dz;
  auto term = [&](const float4 &q, uint32_t j) {
    dx = own.x - q.x;
    dy = own.y - q.y;
    dz = own.z - q.z;
    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, own.w)));
    const float ri = __builtin_amdgcn_rsqf(d2);
    const float g = -q.w * (ri * ri * ri);
    return j < c ? g : 0.f;
  };
  float g = term(q0, b);
  float sx = g * dx, sy = g * dy, sz = g * dz;
  for (uint32_t j0 = B; j0 < c; j0 += B) {
    g = term(base[j0 + b], j0 + b);
    sx += g * dx;
    sy += g * dy;
    sz += g * dz;
  }
  // the owner's 8 partials: quads by two exchanges, then the group's last
  // lane adds the other quad (a shift of 4)
  sx += dpp_from<0xB1>(sx);
  sy += dpp_from<0xB1>(sy);
  sz += dpp_from<0xB1>(sz);
  sx += dpp_from<0x4E>(sx);
  sy += dpp_from<0x4E>(sy);
  sz += dpp_from<0x4E>(sz);
  sx += dpp_from<0x114>(sx);
  sy += dpp_from<0x114>(sy);
  sz += dpp_from<0x114>(sz);
  wave_sync();
  if (b == B - 1u) {
    mine[0] = sx;
    mine[1] = sy;
    mine[2] = sz;
  }
  wave_sync();
  if (active) {
    const float *t = (const float *)(slot + rank);
    acc.p[0] += t[0];
    acc.p[1] += t[1];
    acc.p[2] += t[2];
    if (COUNT) np += c - (me - s0 < c ? 1u : 0u);
  }
}

// Lanes set in a wave-uniform mask, counted on the scalar unit.
__device__ __forceinline__ uint32_t mask_count(uint64_t m) {
  uint32_t k;
  asm("s_bcnt1_i32_b64 %0, %1" : "=s"(k) : "s"(m) : "scc");
  return k;
}

// The wavefront's 64 LDS slots for split leaves (none for policies that do
// not split them).
template <class P, int NT>
__device__ __forceinline__ float4 *walk_slots() {
  if constexpr (P::COMPACT) {
    __shared__ float4 slots[NT];
    return slots + (threadIdx.x & ~63u);
  } else {
    return nullptr;
  }
}

// The walk of one wavefront: bodies [wbase, wbase + 64) of a tree whose
// records start one before nodes_m1[1] and whose bodies start at bodies[0].
template <class P, bool COUNT, bool PRED, bool KD = false>
__device__ __forceinline__ void walk_wave(const typename P::Rec *__restrict__ nodes_m1,
                                          const typename P::Body *__restrict__ bodies,
                                          const WalkArgs<P> &a, uint32_t wbase_in,
                                          float4 *__restrict__ slot) {
  // wave-uniform by construction; readfirstlane lets the compiler know it
  const uint32_t wbase = uni(wbase_in);
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t me = wbase + lane;
  const bool valid = me < a.n;
  const typename P::Self p = P::self(bodies[valid ? me : wbase]);
  BodyState bs{};
  if (KD) bs = body_state(a.kd, valid ? me : wbase);
  typename P::Acc acc;
  // The position is the byte offset of the current record, off = h RS (the
  // records' skip pointers are stored the same way), so no multiply is
  // needed per visit.  The lane's acceptance level D (root 1; unbounded
  // while it keeps opening) is kept as F = 32 - D - log2(RS) and compared
  // with clz(off) = 32 - level(h) - log2(RS), which saves forming the level:
  // F = 32 - log2(RS) keeps a lane out of every node, F = -log2(RS) lets it
  // into every node.
  constexpr uint32_t RS = sizeof(typename P::Rec);
  static_assert(RS == 32 || RS == 64, "records of 32 or 64 bytes");
  constexpr int RSL = RS == 32 ? 5 : 6;
  int E = valid ? -RSL : 32 - RSL;
  unsigned long long nt = 0, na = 0, np = 0, wv = 0, wl = 0;
  uint32_t h = RS;  // the root, heap index 1
  do {
    const typename P::Rec r = load_rec_at(nodes_m1, uni(h));
    const int lz = __builtin_clz(h);  // 32 - level(h) - log2(RS)
    // the lanes taking part, as a scalar mask (kept on the scalar unit) and
    // as the lane's own condition
    const uint64_t actm = __builtin_amdgcn_ballot_w64(lz >= E);
    const bool active = __builtin_amdgcn_inverse_ballot_w64(actm);
    if (COUNT) ++wv;
    if (r.count <= a.leaf) {
      if (COUNT) ++wl;
      bool split = false;
      if constexpr (P::COMPACT) {
        const uint32_t k = mask_count(actm);
        if (k <= SPLIT_MAX) {
          leaf_split<P, COUNT>(acc, p, bodies + r.start, r.count, r.start, me, actm, active, slot,
                               np);
          split = true;
        }
      }
      if (!split)
        for (uint32_t b = 0; b < r.count; b += P::CHUNK)
          leaf_batch<P, P::CHUNK, COUNT, PRED>(acc, p, bodies + (r.start + b),
                                               min((uint32_t)P::CHUNK, r.count - b),
                                               r.start + b, me, active, np);
      P::fold(acc);  // on every lane (a fold is valid at any point): no exec-mask change
      h = r.skip;
    } else {
      // the test runs on every lane (no divergence); only active lanes use it
      typename P::Diff f;
      const bool ok = P::test(p, r, a.theta2, f);
      // lane masks combined on the scalar unit, then used as lane conditions
      const uint64_t okm = __builtin_amdgcn_ballot_w64(ok);
      // the lanes that open the node, and from the same SCC the next node:
      // the first child if any lane opens, else the skip (the compiler
      // would test om against 0 a second time)
      uint64_t om;
      uint32_t hn;
      asm("s_andn2_b64 %0, %2, %3\n\ts_cselect_b32 %1, %4, %5"
          : "=&s"(om), "=s"(hn)
          : "s"(actm), "s"(okm), "s"(2 * h), "s"(r.skip)
          : "scc");
      const bool take = __builtin_amdgcn_inverse_ballot_w64(actm & okm);
      const bool open = __builtin_amdgcn_inverse_ballot_w64(om);
      // wave-uniform: the far-field term is evaluated only when a lane takes it
      if (actm & okm) P::m2p_sel(acc, r, f, take);
      // D = open ? unbounded : min(D, level): a taking lane had D >= level,
      // a lane outside D < level
      E = open ? -RSL : max(E, lz);
      if (COUNT) {
        nt += active ? 1 : 0;
        na += take ? 1 : 0;
      }
      h = hn;
    }
  } while (h != 0);
  P::fold(acc);
  if (KD) {
    uint64_t lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
    if (valid)
      kick_drift_one(a.kd, me, bs, P::result(acc, 0), P::result(acc, 1), P::result(acc, 2), lo,
                     hi);
    if (a.kd.wave_range) wave_range_store(a.kd.wave_range + 6 * (size_t)(wbase >> 6), lo, hi, lane);
  } else if (valid) {
    a.ax[me] = P::result(acc, 0);
    a.ay[me] = P::result(acc, 1);
    a.az[me] = P::result(acc, 2);
  }
  if (COUNT) {
    nt = wave_sum_u64(nt);
    na = wave_sum_u64(na);
    np = wave_sum_u64(np);
    if (lane == 0) {
      atomicAdd(&a.cnt.c[0], nt);
      atomicAdd(&a.cnt.c[1], na);
      atomicAdd(&a.cnt.c[2], np);
      atomicAdd(&a.cnt.c[3], wv);
      atomicAdd(&a.cnt.c[4], wl);
    }
  }
}

// Eight wavefronts a SIMD (at most 64 VGPRs): a first form of the split
// leaves needed 65 registers, which left seven wavefronts a SIMD and took
// 11.36 ms against 10.71 ms held to 64.  The walk now needs 48 (the fused
// F32 form), so the bound only guards against such growth.
#ifndef KD_WALK_WAVES
#define KD_WALK_WAVES 8
#endif
#ifndef KD_WALK_XCD_CHUNK
#define KD_WALK_XCD_CHUNK 256
#endif
// The walk's block order: chunks of C consecutive blocks (C = 256: 65,536
// bodies) dealt to the 8 XCDs round robin, the last partial round split into
// one contiguous range an XCD as xcd_block does (C = 0: xcd_block itself).
// Each XCD walks parts of the whole tree, which evens out the work of
// regions of different density: with one range an XCD the 10 M ring's XCDs
// were 5 % apart in work and the walk took 0.1 ms longer (C = 32: 0.2 ms
// longer, the XCDs' L2s sharing less).
__device__ __forceinline__ uint32_t walk_block() {
  constexpr uint32_t C = KD_WALK_XCD_CHUNK;
  if (C == 0) return xcd_block();
  const uint32_t nb = gridDim.x, full = nb / (8u * C) * (8u * C);
  const uint32_t x = blockIdx.x % 8u, i = blockIdx.x / 8u;
  if (blockIdx.x < full) return (i / C) * (8u * C) + x * C + i % C;
  const uint32_t t = nb - full, q = t / 8u, r = t % 8u, it = (blockIdx.x - full) / 8u;
  return full + x * q + (x < r ? x : r) + it;
}
template <class P, bool COUNT, bool PRED, bool KD = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KD_WALK_WAVES))) void k_walk(
    const typename P::Rec *__restrict__ nodes_m1, const typename P::Body *__restrict__ bodies,
    WalkArgs<P> a) {
  const uint32_t wbase = (walk_block() * blockDim.x + threadIdx.x) & ~63u;
  if (wbase >= a.n) return;
  walk_wave<P, COUNT, PRED, KD>(nodes_m1, bodies, a, wbase, walk_slots<P, 256>());
}

// Forest walk: a task is one wavefront of one system (x: system, y: first
// body).  PERSIST: a grid sized to the machine (a multiple of 8 blocks) takes
// tasks from one atomic counter per XCD (reset by k_queue_reset before the
// launch) until its XCD's tasks are done: the tasks are dealt to the 8 XCDs
// round robin in chunks of 2^cl2 consecutive wavefronts, so neighbouring
// bodies share an L2 while the largest-first order stays balanced across
// XCDs; otherwise wavefront w of the grid runs task w.

template <class P, bool COUNT, bool PERSIST, int NT, bool KD = false>
// (eight wavefronts per SIMD: the persistent grid is sized for 32 a CU; the
// kick-drift's pointers would otherwise take the SGPRs of an eighth)
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_walk_forest(
    const typename P::Rec *__restrict__ nodes, const typename P::Body *__restrict__ bodies,
    WalkArgs<P> a, const SysDesc *__restrict__ sys, const uint2 *__restrict__ tasks,
    uint32_t ntasks, uint32_t *queue, uint32_t cl2) {
  // the wavefront index, uniform within a wavefront (NT is a multiple of 64)
  uint32_t t = uni(((PERSIST ? blockIdx.x : xcd_block()) * NT + threadIdx.x) >> 6);
  for (;;) {
    if (PERSIST) {
      const uint32_t x = blockIdx.x % 8u;  // the XCD the dispatcher placed this block on
      uint32_t got = 0;
      if ((threadIdx.x & 63) == 0) got = atomicAdd(queue + x, 1u);
      got = uni(__shfl(got, 0, 64));
      // the XCD's got-th task; increasing in got, so past ntasks it is done
      t = ((((got >> cl2) << 3) + x) << cl2) | (got & ((1u << cl2) - 1u));
    }
    if (t >= ntasks) return;
    const uint2 task = tasks[t];
    const SysDesc S = sys[task.x];
    WalkArgs<P> b = a;
    b.n = S.n;
    b.ax += S.body_off;
    b.ay += S.body_off;
    b.az += S.body_off;
    if (KD) {
      // the system's state; its wavefront range goes to slot t (walk_wave
      // indexes the slots by the wavefront's first body within the system)
      StateDev &st = b.kd.s;
      st.x += S.body_off;
      st.y += S.body_off;
      st.z += S.body_off;
      st.vx += S.body_off;
      st.vy += S.body_off;
      st.vz += S.body_off;
      if (b.kd.wave_range) b.kd.wave_range += 6 * ((size_t)t - (task.y >> 6));
    }
    walk_wave<P, COUNT, false, KD>(nodes + S.node_off - 1, bodies + S.body_off, b, task.y,
                                   walk_slots<P, NT>());
    if (!PERSIST) return;
  }
}

__global__ void k_queue_reset(uint32_t *queue) {
  if (threadIdx.x < 8) queue[threadIdx.x] = 0u;
}

// The plain FP64 path keeps a simpler loop: it is VALU-bound (FP64 sqrt and
// divide dominate), and this form issues fewer vector instructions per visit
// than the batched walk above: leaf bodies are read one at a time under the
// active-lane mask and the self pair is skipped by branch.
template <bool COUNT, bool KD = false>
__global__ __launch_bounds__(256) void k_walk_f64(const NodeRec64 *__restrict__ nodes,
                                                  const double4 *__restrict__ pos,
                                                  WalkArgs<F64> a) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = (xcd_block() * blockDim.x + threadIdx.x) & ~63u;
  if (wbase >= a.n) return;
  const uint32_t me = wbase + lane;
  const bool valid = me < a.n;
  const double4 p = valid ? pos[me] : make_double4(0.0, 0.0, 0.0, 0.0);
  BodyState bs{};
  if (KD) bs = body_state(a.kd, valid ? me : wbase);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  int D = valid ? 0x7fffffff : -1;
  unsigned long long nt = 0, na = 0, np = 0, wv = 0, wl = 0;
  uint32_t h = 1;
  int depth = 0;
  while (h != 0) {
    const NodeRec64 r = nodes[uni(h) - 1];
    const bool active = depth <= D;
    if (COUNT) ++wv;
    if (r.count <= a.leaf) {
      if (COUNT) ++wl;
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
      if (r.size2 < a.theta2 * d2) {
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
  if (KD) {
    uint64_t lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
    if (valid) kick_drift_one(a.kd, me, bs, sx, sy, sz, lo, hi);
    if (a.kd.wave_range) wave_range_store(a.kd.wave_range + 6 * (size_t)(wbase >> 6), lo, hi, lane);
  } else if (valid) {
    a.ax[me] = sx;
    a.ay[me] = sy;
    a.az[me] = sz;
  }
  if (COUNT) {
    nt = wave_sum_u64(nt);
    na = wave_sum_u64(na);
    np = wave_sum_u64(np);
    if (lane == 0) {
      atomicAdd(&a.cnt.c[0], nt);
      atomicAdd(&a.cnt.c[1], na);
      atomicAdd(&a.cnt.c[2], np);
      atomicAdd(&a.cnt.c[3], wv);
      atomicAdd(&a.cnt.c[4], wl);
    }
  }
}

template <class P>
static void launch_policy(const WalkArgs<P> &a, bool count, int variant, hipStream_t s,
                          bool kick = false) {
  const unsigned blocks = (a.n + 255) / 256;
  const bool pred = (variant & 2) != 0;
  if (kick) {
    if (pred)
      k_walk<P, false, true, true><<<blocks, 256, 0, s>>>(a.nodes - 1, a.bodies, a);
    else
      k_walk<P, false, false, true><<<blocks, 256, 0, s>>>(a.nodes - 1, a.bodies, a);
  } else if (count) {
    if (pred)
      k_walk<P, true, true><<<blocks, 256, 0, s>>>(a.nodes - 1, a.bodies, a);
    else
      k_walk<P, true, false><<<blocks, 256, 0, s>>>(a.nodes - 1, a.bodies, a);
  } else {
    if (pred)
      k_walk<P, false, true><<<blocks, 256, 0, s>>>(a.nodes - 1, a.bodies, a);
    else
      k_walk<P, false, false><<<blocks, 256, 0, s>>>(a.nodes - 1, a.bodies, a);
  }
}

hipError_t launch_traverse(int precision, bool count, int variant, const TreeDev &t,
                           const double4 *pos64, const float4 *pos32, uint32_t n,
                           double theta, int leaf, double *ax, double *ay, double *az,
                           unsigned long long *counters, hipStream_t s) {
  Counters c{counters};
  const double theta2 = theta * theta;
  if (precision == 0) {
    WalkArgs<F64> a{t.rec64, pos64, n, (uint32_t)leaf, theta2, ax, ay, az, c};
    const unsigned blocks = (n + 255) / 256;
    if (variant & 4) {
      launch_policy(a, count, variant, s);
    } else if (count) {
      k_walk_f64<true><<<blocks, 256, 0, s>>>(a.nodes, a.bodies, a);
    } else {
      k_walk_f64<false><<<blocks, 256, 0, s>>>(a.nodes, a.bodies, a);
    }
  } else if (precision == 1) {
    WalkArgs<F32> a{t.rec32, pos32, n, (uint32_t)leaf, theta2, ax, ay, az, c};
    launch_policy(a, count, variant, s);
  } else {
    WalkArgs<D64> a{t.recd, pos64, n, (uint32_t)leaf, theta2, ax, ay, az, c};
    launch_policy(a, count, variant, s);
  }
  return hipGetLastError();
}

hipError_t launch_traverse_kick(int precision, int variant, const TreeDev &t, const double4 *pos64,
                                const float4 *pos32, uint32_t n, double theta, int leaf,
                                const KickDrift &kd, hipStream_t s) {
  Counters c{nullptr};
  const double theta2 = theta * theta;
  if (precision == 0) {
    WalkArgs<F64> a{t.rec64, pos64, n, (uint32_t)leaf, theta2, nullptr, nullptr, nullptr, c, kd};
    if (variant & 4)
      launch_policy(a, false, variant, s, true);
    else
      k_walk_f64<false, true><<<(n + 255) / 256, 256, 0, s>>>(a.nodes, a.bodies, a);
  } else if (precision == 1) {
    WalkArgs<F32> a{t.rec32, pos32, n, (uint32_t)leaf, theta2, nullptr, nullptr, nullptr, c, kd};
    launch_policy(a, false, variant, s, true);
  } else {
    WalkArgs<D64> a{t.recd, pos64, n, (uint32_t)leaf, theta2, nullptr, nullptr, nullptr, c, kd};
    launch_policy(a, false, variant, s, true);
  }
  return hipGetLastError();
}

template <class P, bool COUNT, bool PERSIST, bool KD>
static void launch_forest_nt(const WalkArgs<P> &a, const WalkShape &ws, const SysDesc *sys,
                             const uint2 *tasks, uint32_t ntasks, uint32_t *queue, int cus,
                             hipStream_t s) {
  const int nt = ws.threads;
  const unsigned wpb = (unsigned)(nt / 64);
  // chunks of 2^cl2 tasks: as long as 1024 (65,536 bodies), while every XCD
  // still gets at least 8 chunks (at 512 systems and 24.7 M bodies the
  // persistent walk takes 28.5 ms with 1024, 28.9 with 64, 30.2 with 4)
  uint32_t cl2 = 0;
  while (cl2 < 10 && (ntasks >> (cl2 + 1)) >= 64u) ++cl2;
  unsigned blocks;
  if (PERSIST)  // every XCD's queue needs blocks: a multiple of 8, at least 8
    blocks = ((unsigned)std::max(1, cus * ws.waves_per_cu / (int)wpb) + 7u) / 8u * 8u;
  else
    blocks = (ntasks + wpb - 1) / wpb;
#define KD_FOREST(NT)                                                                    \
  case NT:                                                                               \
    k_walk_forest<P, COUNT, PERSIST, NT, KD><<<blocks, NT, 0, s>>>(a.nodes, a.bodies, a, sys, \
                                                              tasks, ntasks, queue, cl2); \
    break;
  switch (nt) {
    KD_FOREST(64)
    KD_FOREST(128)
    KD_FOREST(512)
    default:
    KD_FOREST(256)
  }
#undef KD_FOREST
}

template <class P>
static void launch_forest(const WalkArgs<P> &a, bool count, bool kick, const WalkShape &ws,
                          const SysDesc *sys, const uint2 *tasks, uint32_t ntasks,
                          uint32_t *queue, int cus, hipStream_t s) {
  const bool persist = ws.waves_per_cu > 0;
  if (persist) k_queue_reset<<<1, 64, 0, s>>>(queue);
  if (kick) {
    if (persist)
      launch_forest_nt<P, false, true, true>(a, ws, sys, tasks, ntasks, queue, cus, s);
    else
      launch_forest_nt<P, false, false, true>(a, ws, sys, tasks, ntasks, queue, cus, s);
  } else if (count) {
    if (persist)
      launch_forest_nt<P, true, true, false>(a, ws, sys, tasks, ntasks, queue, cus, s);
    else
      launch_forest_nt<P, true, false, false>(a, ws, sys, tasks, ntasks, queue, cus, s);
  } else {
    if (persist)
      launch_forest_nt<P, false, true, false>(a, ws, sys, tasks, ntasks, queue, cus, s);
    else
      launch_forest_nt<P, false, false, false>(a, ws, sys, tasks, ntasks, queue, cus, s);
  }
}

hipError_t launch_traverse_forest(int precision, bool count, const WalkShape &ws,
                                  const TreeDev &t, const double4 *pos64, const float4 *pos32,
                                  const Forest &f, const uint2 *tasks, uint32_t ntasks,
                                  uint32_t *queue, int cus, double theta, int leaf, double *ax,
                                  double *ay, double *az, unsigned long long *counters,
                                  hipStream_t s, const KickDrift *kick) {
  if (ws.threads != 64 && ws.threads != 128 && ws.threads != 256 && ws.threads != 512)
    return hipErrorInvalidValue;
  Counters c{counters};
  const double theta2 = theta * theta;
  const KickDrift kd = kick ? *kick : KickDrift{};
  const bool k = kick != nullptr;
  if (precision == 0) {
    WalkArgs<F64> a{t.rec64, pos64, 0, (uint32_t)leaf, theta2, ax, ay, az, c, kd};
    launch_forest(a, count && !k, k, ws, f.sys, tasks, ntasks, queue, cus, s);
  } else if (precision == 1) {
    WalkArgs<F32> a{t.rec32, pos32, 0, (uint32_t)leaf, theta2, ax, ay, az, c, kd};
    launch_forest(a, count && !k, k, ws, f.sys, tasks, ntasks, queue, cus, s);
  } else {
    WalkArgs<D64> a{t.recd, pos64, 0, (uint32_t)leaf, theta2, ax, ay, az, c, kd};
    launch_forest(a, count && !k, k, ws, f.sys, tasks, ntasks, queue, cus, s);
  }
  return hipGetLastError();
}

}  // namespace kd
