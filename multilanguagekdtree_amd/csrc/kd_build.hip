// kd_build.hip -- gfx950 tree build: level-synchronous median partition.
//
// The tree shape is a pure function of N (every split is at floor(len/2),
// Sequential/RustVersion/src/array_kd_tree.rs:100), so segment bounds are
// known up front and the tree is kept in heap order.  Each step:
//   1. three lists of body ids sorted by (x), (y), (z): a device radix sort
//      on 32-bit fixed-point keys over each coordinate's range, then runs of
//      shared keys put in exact (coordinate, id) order;
//   2. device-wide levels while segments exceed SUB_MAX bodies: per segment,
//      bbox from the ends of the three sorted lists, split dimension with the
//      reference's strict '>' (:91-96), split value = key at mid (:125); the
//      split-dim list is already partitioned at mid, the other two are
//      stably partitioned by a left/right bit per id (per-tile counts, an
//      exclusive scan over tiles, scatter);
//   3. one workgroup per remaining segment finishes its subtree in LDS,
//      emits leaves in canonical order (ascending original index), node
//      moments bottom-up, and writes the body state in tree order;
//   4. one workgroup reduces the moments of the device-wide levels.
// A forest (many independent systems) runs the same kernels with the system
// index in blockIdx.y; see SysDesc.
// Ties on a split coordinate are resolved by original index, lazily: only a
// run of equal keys that straddles the median is reordered.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>

#include "kd_internal.hpp"

namespace kd {

// ---------------------------------------------------------------------------
// block scans
// ---------------------------------------------------------------------------

// Exclusive sum over the block; total receives the block sum.  sh needs
// NT/64 entries.
template <int NT, typename T>
__device__ __forceinline__ T block_excl_sum(T v, T *sh, T &total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  if (w == 0) {
    T t = lane < NT / 64 ? sh[lane] : T(0);
#pragma unroll
    for (int o = 1; o < NT / 64; o <<= 1) {
      T y = __shfl_up(t, o, 64);
      if (lane >= o) t += y;
    }
    if (lane < NT / 64) sh[lane] = t;
  }
  __syncthreads();
  T before = w ? sh[w - 1] : T(0);
  total = sh[NT / 64 - 1];
  __syncthreads();
  return before + x - v;
}

// Segmented inclusive-scan element: value and "segment starts here" flag.
struct SegV {
  uint64_t v;
  uint32_t h;
};
__device__ __forceinline__ SegV seg_combine(SegV a, SegV b) {
  SegV r;
  r.v = b.h ? b.v : a.v + b.v;
  r.h = a.h | b.h;
  return r;
}

// Exclusive segmented carry for each thread: combination of all earlier
// threads' aggregates.  shv/shh need NT/64 entries.  One barrier: every
// wavefront combines the per-wavefront totals itself, so the caller must
// pass a barrier between two calls before shv/shh are written again.
template <int NT>
__device__ __forceinline__ SegV block_seg_carry(SegV agg, uint64_t *shv,
                                                uint32_t *shh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  SegV x = agg;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    SegV y;
    y.v = __shfl_up(x.v, o, 64);
    y.h = __shfl_up(x.h, o, 64);
    if (lane >= o) x = seg_combine(y, x);
  }
  if (lane == 63) {
    shv[w] = x.v;
    shh[w] = x.h;
  }
  __syncthreads();
  SegV t;
  t.v = lane < NT / 64 ? shv[lane] : 0;
  t.h = lane < NT / 64 ? shh[lane] : 0;
#pragma unroll
  for (int o = 1; o < NT / 64; o <<= 1) {
    SegV y;
    y.v = __shfl_up(t.v, o, 64);
    y.h = __shfl_up(t.h, o, 64);
    if (lane >= o) t = seg_combine(y, t);
  }
  // exclusive within the wave
  SegV prev;
  prev.v = __shfl_up(x.v, 1, 64);
  prev.h = __shfl_up(x.h, 1, 64);
  if (lane == 0) {
    prev.v = 0;
    prev.h = 0;
  }
  // the combination of the totals of wavefronts 0 .. w-1
  SegV carry;
  carry.v = __shfl(t.v, w > 0 ? w - 1 : 0, 64);
  carry.h = __shfl(t.h, w > 0 ? w - 1 : 0, 64);
  if (w == 0) {
    carry.v = 0;
    carry.h = 0;
  }
  return lane == 0 ? carry : seg_combine(carry, prev);
}

// ---------------------------------------------------------------------------
// shape
// ---------------------------------------------------------------------------

// Segment of every heap slot: descend from the root splitting at floor(c/2);
// slots below a leaf get count 0.
__global__ void k_shape(uint32_t *seg_start, uint32_t *seg_count, uint8_t *lmask,
                        const SysDesc *sys, int leaf) {
  const SysDesc S = sys[blockIdx.y];
  int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= S.slots) return;
  seg_start += S.node_off;
  seg_count += S.node_off;
  lmask[S.node_off + h] = 0;  // every list of a root starts in the first buffer
  const uint32_t n = S.n;
  uint32_t hp = (uint32_t)h + 1;
  int lev = 31 - __clz(hp);
  uint32_t k = hp - (1u << lev);
  uint32_t s = 0, c = n;
  bool ok = true;
  for (int b = lev - 1; b >= 0; --b) {
    if (c <= (uint32_t)leaf) {
      ok = false;
      break;
    }
    uint32_t half = c >> 1;
    if ((k >> b) & 1u) {
      s += half;
      c -= half;
    } else {
      c = half;
    }
  }
  seg_start[h] = ok ? s : 0;
  seg_count[h] = ok ? c : 0;
}

// ---------------------------------------------------------------------------
// sort keys
// ---------------------------------------------------------------------------

// The id lists are sorted by a 32-bit fixed-point image of each coordinate
// over its current range, key = floor((x - min) (2^32 - 2) / (max - min))
// (computed on halves, so no range overflows).  Every rounding step of the
// map is monotone, so the key order never contradicts the coordinate order;
// bodies whose distinct coordinates share a key are put in coordinate order
// afterwards (k_fix_runs, k_fix_long).  Equal coordinates keep the sort's
// stable (id) order.

__device__ __forceinline__ double ordkey_inv(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

// range[2d], range[2d+1]: ordered-key minimum and maximum of dimension d.
__global__ void k_range_init(unsigned long long *range) {
  const int t = threadIdx.x;
  if (t < 3) {
    range[2 * t] = ~0ull;
    range[2 * t + 1] = 0ull;
  }
}

hipError_t launch_range_reset(unsigned long long *range, hipStream_t s) {
  k_range_init<<<1, 64, 0, s>>>(range);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_range(const double *__restrict__ x,
                                               const double *__restrict__ y,
                                               const double *__restrict__ z, uint32_t n,
                                               unsigned long long *range) {
  __shared__ uint64_t sh[6][4];
  uint64_t lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  // two positions per iteration, so more loads are in flight
  for (; i + stride < n; i += 2 * stride) {
    const double p[6] = {x[i], y[i], z[i], x[i + stride], y[i + stride], z[i + stride]};
#pragma unroll
    for (int e = 0; e < 6; ++e) {
      const uint64_t k = ordkey(p[e]);
      const int d = e % 3;
      lo[d] = k < lo[d] ? k : lo[d];
      hi[d] = k > hi[d] ? k : hi[d];
    }
  }
  if (i < n) {
    const uint64_t k[3] = {ordkey(x[i]), ordkey(y[i]), ordkey(z[i])};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = k[d] < lo[d] ? k[d] : lo[d];
      hi[d] = k[d] > hi[d] ? k[d] : hi[d];
    }
  }
  range_commit(lo, hi, sh, range);
}

hipError_t launch_range(const StateDev &st, uint32_t n, unsigned long long *range, hipStream_t s) {
  k_range_init<<<1, 64, 0, s>>>(range);
  k_range<<<std::max(1u, std::min<uint32_t>((n + 255) / 256, RANGE_BLOCKS)), 256, 0, s>>>(
      st.x, st.y, st.z, n, range);
  return hipGetLastError();
}

// Body keys span [0, PAD_KEY - 1]; in a forest, the padding positions that
// round each system up to whole wavefronts take PAD_KEY, so they sort last
// and never share a key with a body.
constexpr uint32_t PAD_KEY = 0xFFFFFFFFu;

// The sort key of every body position j (the values the sort carries are
// the positions themselves, read from a counting iterator).  In a forest
// (sob set) the key is the system index in the top sysbits bits over the
// top 32 - sysbits bits of the coordinate key, so one sort orders every
// system's bodies by coordinate and places the systems one after another
// (a body's key still never orders it against its system's coordinate
// order: the coordinate part is a monotone map); with sysbits =
// bits_for(nsys) no body key reaches PAD_KEY.
__global__ void k_make_keys32(const double *__restrict__ p, const unsigned long long *range,
                              uint32_t *__restrict__ keys, uint32_t n, uint32_t *runs,
                              const uint32_t *__restrict__ sob,
                              const SysDesc *__restrict__ sys, int sysbits) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) runs[0] = 0;
  if (j >= n) return;
  const double lo = 0.5 * ordkey_inv(range[0]), hi = 0.5 * ordkey_inv(range[1]);
  const double scale = hi > lo ? 4294967294.0 / (hi - lo) : 0.0;
  const double v = 0.5 * p[j];
  uint32_t k = (uint32_t)fmin(fmax((v - lo) * scale, 0.0), 4294967294.0);
  if (sob != nullptr) {
    const uint32_t q = sob[j >> 6];
    const SysDesc S = sys[q];
    k = j - S.body_off >= S.n ? PAD_KEY : (q << (32 - sysbits)) | (k >> sysbits);
  }
  keys[j] = k;
}

// Runs of equal keys: a run of at most FIX_SHORT ids is put in (coordinate,
// id) order by its first thread; the start of a longer run is queued for
// k_fix_long, which finds its length and checks its coordinates in parallel.

__device__ __forceinline__ bool coord_lt(const double *p, uint32_t a, uint32_t b) {
  const double pa = p[a], pb = p[b];
  return pa < pb || (pa == pb && a < b);
}

// The run starting at i (keys[i] == keys[i + 1], and i is 0 or keys[i - 1]
// differs), fixed by one thread.
__device__ void fix_run(const uint32_t *__restrict__ keys, uint32_t *__restrict__ ids,
                        const double *__restrict__ p, uint32_t n, uint32_t *runs, uint32_t i) {
  const uint32_t k = keys[i];
  if (k == PAD_KEY) return;  // forest padding only
  const uint32_t lim = min(n, i + FIX_SHORT + 1);
  uint32_t j = i + 1;
  while (j < lim && keys[j] == k) ++j;
  if (j == i + FIX_SHORT + 1) {  // longer than FIX_SHORT: one workgroup takes it
    const uint32_t slot = atomicAdd(&runs[0], 1u);
    runs[1 + 2 * slot] = i;
    runs[2 + 2 * slot] = 0;
    return;
  }
  bool distinct = false;
  const double p0 = p[ids[i]];
  for (uint32_t a = i + 1; a < j; ++a) distinct |= p[ids[a]] != p0;
  if (!distinct) return;  // equal coordinates: the stable order stands
  for (uint32_t a = i + 1; a < j; ++a) {
    const uint32_t v = ids[a];
    uint32_t b = a;
    while (b > i && coord_lt(p, v, ids[b - 1])) {
      ids[b] = ids[b - 1];
      --b;
    }
    ids[b] = v;
  }
}

// Each thread scans FIX_PER consecutive sorted keys (two 16-byte loads plus
// one key either side) for run starts; the few it finds are fixed serially.
// Runs are disjoint, so threads never touch the same ids.
constexpr uint32_t FIX_PER = 8;

__global__ __launch_bounds__(256) void k_fix_runs(const uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ ids,
                                                  const double *__restrict__ p, uint32_t n,
                                                  const unsigned long long *range,
                                                  uint32_t *runs) {
  const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * FIX_PER;
  if (i0 >= n) return;
  // every coordinate of this dimension equal (a planar system's z): the
  // stable order of the sort is already the exact one
  if (range[0] == range[1]) return;
  uint32_t k[FIX_PER + 2];  // k[0] = keys[i0 - 1], k[1 + j] = keys[i0 + j]
  const uint32_t m = min(FIX_PER, n - i0);
  if (m == FIX_PER) {
    const uint4 a = *reinterpret_cast<const uint4 *>(keys + i0);
    const uint4 b = *reinterpret_cast<const uint4 *>(keys + i0 + 4);
    k[1] = a.x, k[2] = a.y, k[3] = a.z, k[4] = a.w;
    k[5] = b.x, k[6] = b.y, k[7] = b.z, k[8] = b.w;
  } else {
#pragma unroll
    for (uint32_t j = 0; j < FIX_PER; ++j) k[1 + j] = j < m ? keys[i0 + j] : 0u;
  }
  k[0] = i0 > 0 ? keys[i0 - 1] : 0u;
  k[FIX_PER + 1] = i0 + FIX_PER < n ? keys[i0 + FIX_PER] : 0u;
  uint32_t starts = 0;
#pragma unroll
  for (uint32_t j = 0; j < FIX_PER; ++j) {
    const bool next_eq = i0 + j + 1 < n && k[2 + j] == k[1 + j];
    const bool prev_eq = i0 + j > 0 && k[j] == k[1 + j];
    starts |= (next_eq && !prev_eq) ? (1u << j) : 0u;
  }
  while (starts) {
    const uint32_t j = __builtin_ctz(starts);
    starts &= starts - 1;
    fix_run(keys, ids, p, n, runs, i0 + j);
  }
}

// Long runs, one workgroup each: chunks of LONG_LDS ids are bitonic-sorted in
// LDS, then merged pairwise in global memory (merge path: each thread finds
// the split of its output slice by binary search), ping-ponging between the
// id list and a scratch list at the same offsets.  Runs are disjoint, so
// workgroups never share memory.
constexpr uint32_t LONG_LDS = 4096;

__device__ void lds_bitonic(double *sk, uint32_t *si, uint32_t m) {
  for (uint32_t size = 2; size <= m; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
        const uint32_t u = t ^ stride;
        if (u > t) {
          const bool up = (t & size) == 0;
          const bool gt = sk[t] > sk[u] || (sk[t] == sk[u] && si[t] > si[u]);
          if (gt == up) {
            const double tk = sk[t];
            sk[t] = sk[u];
            sk[u] = tk;
            const uint32_t ti = si[t];
            si[t] = si[u];
            si[u] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Sort ids[0..len) of one chunk (len <= LONG_LDS) in LDS.
__device__ void sort_chunk(uint32_t *ids, uint32_t len, const double *__restrict__ p, double *sk,
                           uint32_t *si) {
  uint32_t m = 1;
  while (m < len) m <<= 1;
  for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
    if (t < len) {
      si[t] = ids[t];
      sk[t] = p[si[t]];
    } else {
      si[t] = 0xffffffffu;
      sk[t] = __longlong_as_double(0x7ff0000000000000ll);  // +inf: padding sorts last
    }
  }
  __syncthreads();
  lds_bitonic(sk, si, m);
  for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) ids[t] = si[t];
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_fix_long(const uint32_t *runs,
                                                  const uint32_t *__restrict__ keys, uint32_t n,
                                                  uint32_t *ids, uint32_t *scratch,
                                                  const double *__restrict__ p) {
  __shared__ double sk[LONG_LDS];
  __shared__ uint32_t si[LONG_LDS];
  __shared__ uint32_t s_end;
  const uint32_t nruns = runs[0];
  for (uint32_t r = blockIdx.x; r < nruns; r += gridDim.x) {
    const uint32_t st = runs[1 + 2 * r];
    // the run's end: the first position after st whose key differs
    const uint32_t k = keys[st];
    if (k == PAD_KEY) continue;  // forest padding only (uniform branch)
    if (threadIdx.x == 0) s_end = n;
    __syncthreads();
    uint32_t end = n;
    for (uint32_t base = st + 1; base < end; base += blockDim.x) {
      const uint32_t j = base + threadIdx.x;
      if (j < n && keys[j] != k) atomicMin(&s_end, j);
      __syncthreads();
      end = s_end;
      __syncthreads();
      if (end < base + blockDim.x) break;
    }
    const uint32_t len = end - st;
    // equal coordinates throughout: the stable order stands
    const double p0 = p[ids[st]];
    int distinct = 0;
    for (uint32_t j = st + 1 + threadIdx.x; j < end; j += blockDim.x)
      distinct |= p[ids[j]] != p0 ? 1 : 0;
    if (!__syncthreads_or(distinct)) continue;
    uint32_t *a = ids + st, *b = scratch + st;
    for (uint32_t c = 0; c < len; c += LONG_LDS)
      sort_chunk(a + c, min(LONG_LDS, len - c), p, sk, si);
    for (uint32_t w = LONG_LDS; w < len; w <<= 1) {
      // merge [lo, lo+w) and [lo+w, lo+2w) of a into b
      for (uint32_t lo = 0; lo < len; lo += 2 * w) {
        const uint32_t mid = min(lo + w, len), hi = min(lo + 2 * w, len);
        const uint32_t na = mid - lo, nb = hi - mid, tot = na + nb;
        const uint32_t per = (tot + blockDim.x - 1) / blockDim.x;
        const uint32_t o0 = min(threadIdx.x * per, tot), o1 = min(o0 + per, tot);
        // merge path: the first o0 outputs are a[lo, lo+i) and a[mid, mid+o0-i)
        uint32_t l = o0 > nb ? o0 - nb : 0, h = min(o0, na);
        while (l < h) {
          const uint32_t i = (l + h) / 2, j = o0 - i - 1;
          if (coord_lt(p, a[mid + j], a[lo + i])) h = i; else l = i + 1;
        }
        uint32_t i = l, j = o0 - l;
        for (uint32_t o = o0; o < o1; ++o) {
          const bool take_b = j < nb && (i >= na || coord_lt(p, a[mid + j], a[lo + i]));
          b[lo + o] = take_b ? a[mid + j++] : a[lo + i++];
        }
      }
      __threadfence_block();
      __syncthreads();
      uint32_t *t = a;
      a = b;
      b = t;
    }
    if (a != ids + st)
      for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) ids[st + t] = a[t];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// device-wide levels
// ---------------------------------------------------------------------------

struct Lists {
  uint32_t *l[3];
};
struct Coords {
  const double *c[3];
};

// The two buffers of the three sorted lists.  A node's lists live in A or B
// per its mask (bit d: list d in B); splitting a node moves the two lists
// that are partitioned to the other buffer and leaves the split-dimension
// list where it is (it is already partitioned at the median).
struct ListBufs {
  Lists A, B;
  __device__ uint32_t *in(uint32_t m, int d) const { return ((m >> d) & 1u) ? B.l[d] : A.l[d]; }
  __device__ uint32_t *out(uint32_t m, int d) const { return ((m >> d) & 1u) ? A.l[d] : B.l[d]; }
};

// A run of equal keys straddles the median of segment h (list ls of its
// split dimension, [s, s + c)): the run's members that go left are those
// with the smallest original indices.  The block radix-selects the threshold
// index (the original index of the first body that goes right)
// and stably partitions the run (scratch = the same list's output buffer,
// not live for this node).
template <int NT>
__device__ void tie_fix(uint32_t *ls, uint32_t *tmp, const double *key,
                        const uint32_t *__restrict__ oid, uint32_t s, uint32_t c) {
  const uint32_t mid = s + c / 2;
  const double kv = key[ls[mid]];
  __shared__ uint32_t r0s, r1s, need_s, prefix_s, mask_s;
  __shared__ uint32_t hist[256];
  __shared__ uint32_t scan_sh[NT / 64];
  if (threadIdx.x == 0) {
    // keys are sorted non-decreasing inside [s, s+c)
    uint32_t lo = s, hi = mid;  // first position with key == kv
    while (lo < hi) {
      uint32_t m = lo + (hi - lo) / 2;
      if (key[ls[m]] < kv) lo = m + 1; else hi = m;
    }
    uint32_t r0 = lo;
    lo = mid + 1;
    hi = s + c;  // first position with key > kv
    while (lo < hi) {
      uint32_t m = lo + (hi - lo) / 2;
      if (key[ls[m]] > kv) hi = m; else lo = m + 1;
    }
    r0s = r0;
    r1s = lo;
    need_s = mid - r0;
    prefix_s = 0;
    mask_s = 0;
  }
  __syncthreads();
  const uint32_t r0 = r0s, r1 = r1s;
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int i = threadIdx.x; i < 256; i += NT) hist[i] = 0;
    __syncthreads();
    const uint32_t pre = prefix_s, msk = mask_s;
    for (uint32_t p = r0 + threadIdx.x; p < r1; p += NT) {
      uint32_t o = oid[ls[p]];
      if ((o & msk) == pre) atomicAdd(&hist[(o >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t rem = need_s, cum = 0, b = 0;
      for (; b < 255; ++b) {
        if (cum + hist[b] > rem) break;
        cum += hist[b];
      }
      need_s = rem - cum;
      prefix_s = pre | (b << shift);
      mask_s = msk | (255u << shift);
    }
    __syncthreads();
  }
  const uint32_t thr_oid = prefix_s;  // original index of the first body that goes right
  uint32_t nleft = mid - r0;
  uint32_t off_l = 0, off_r = nleft;
  for (uint32_t base = r0; base < r1; base += NT) {
    uint32_t p = base + threadIdx.x;
    uint32_t id = p < r1 ? ls[p] : 0;
    uint32_t f = (p < r1 && oid[id] < thr_oid) ? 1u : 0u;
    uint32_t tot;
    uint32_t ex = block_excl_sum<NT>(f, scan_sh, tot);
    if (p < r1) {
      uint32_t idx = p - base;
      uint32_t dst = f ? off_l + ex : off_r + (idx - ex);
      tmp[r0 + dst] = id;
    }
    uint32_t cnt = min((uint32_t)NT, r1 - base);
    off_l += tot;
    off_r += cnt - tot;
  }
  __syncthreads();
  for (uint32_t p = r0 + threadIdx.x; p < r1; p += NT) ls[p] = tmp[p];
}

// Split dimension of extents e: the largest, ties to the lower dimension
// (array_kd_tree.rs:91-96, strict '>').
__device__ __forceinline__ int split_dim_of(const double e[3]) {
  int sd = 0;
  if (e[1] > e[sd]) sd = 1;
  if (e[2] > e[sd]) sd = 2;
  return sd;
}

// The split of every segment of a device-wide level, one block per segment:
// bbox from the ends of the three sorted lists, split dimension with the
// reference's strict '>' (array_kd_tree.rs:91-96), split value = key at mid
// (:125) and the children's list-buffer masks; if equal keys straddle the
// median, the block orders them (tie_fix).
//
// Lazy dimension (lazy < 3): its list is not kept per segment, so its
// extent is known only to be at most its global extent g (from range; a
// segment's extent hi - lo cannot exceed hi_all - lo_all, and rounding keeps
// that order).  The split dimension only grows toward `lazy` as its extent
// grows, so if the dimension the rule picks with the lazy extent set to g is
// not `lazy`, the rule picks the same dimension for every extent in [0, g];
// otherwise the segment sets bit `lazy` of *viol and the build is not exact.
// With viol set and lazy == 3, the block reports every dimension d that the
// rule would pick with its extent raised to its global extent.
constexpr int SPLIT_THREADS = 256;
// The split of segment h ([s, s + c) of its lists, list-buffer mask m): the
// node's dimension, value and size, the children's masks, and whether equal
// keys straddle the median (the caller then orders that run).
__device__ __forceinline__ int split_decide(const ListBufs &LB, const Coords &P, const TreeDev &t,
                                            uint32_t h, uint32_t s, uint32_t c, uint32_t m,
                                            const unsigned long long *range, int lazy,
                                            uint32_t *viol, bool &tie) {
  double ext[3], glob[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    glob[d] = ordkey_inv(range[2 * d + 1]) - ordkey_inv(range[2 * d]);
    if (d == lazy) {
      ext[d] = glob[d];
      continue;
    }
    const uint32_t *ld = LB.in(m, d);
    double lo = P.c[d][ld[s]];
    double hi = P.c[d][ld[s + c - 1]];
    ext[d] = hi - lo;
  }
  int sd;
  if (lazy < 3) {
    if (split_dim_of(ext) == lazy) atomicOr(viol, 1u << lazy);
    ext[lazy] = -1.0;
    sd = split_dim_of(ext);
  } else {
    sd = split_dim_of(ext);
    if (viol != nullptr) {
      uint32_t unsafe = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double e[3] = {ext[0], ext[1], ext[2]};
        e[d] = glob[d];
        unsafe |= split_dim_of(e) == d ? 1u << d : 0u;
      }
      if (unsafe) atomicOr(viol, unsafe);
    }
  }
  const uint32_t mid = s + c / 2;
  const double kv = P.c[sd][LB.in(m, sd)[mid]];
  const double kp = P.c[sd][LB.in(m, sd)[mid - 1]];
  const uint8_t cm = (uint8_t)(m ^ (7u ^ (1u << sd)));
  t.lmask[2 * h + 1] = cm;
  t.lmask[2 * h + 2] = cm;
  t.split_dim[h] = (uint8_t)sd;
  t.split_val[h] = kv;
  t.size[h] = ext[sd];
  tie = kp == kv;
  return sd;
}

__global__ __launch_bounds__(SPLIT_THREADS) void k_level_split(ListBufs LB, Coords P,
                                                               TreeDev t0, const SysDesc *sys,
                                                               const uint32_t *oid,
                                                               uint32_t level,
                                                               const unsigned long long *range,
                                                               int lazy, uint32_t *viol) {
  const SysDesc S = sys[blockIdx.y];
  if (level >= S.sub_level) return;
  const uint32_t k = blockIdx.x;
  const TreeDev t = tree_at(t0, S.node_off);
  const uint32_t h = (1u << level) - 1 + k;
  const uint32_t s = t.seg_start[h] + S.list_off, c = t.seg_count[h];
  const uint32_t m = t.lmask[h];
  __shared__ int sd_s, tie_s;
  if (threadIdx.x == 0) {
    bool tie;
    sd_s = split_decide(LB, P, t, h, s, c, m, range, lazy, viol, tie);
    tie_s = tie;
  }
  __syncthreads();
  if (tie_s)
    tie_fix<SPLIT_THREADS>(LB.in(m, sd_s), LB.out(m, sd_s), P.c[sd_s], oid, s, c);
}

__global__ __launch_bounds__(LVL_THREADS) void k_level_flags(ListBufs LB, TreeDev t0,
                                                             const SysDesc *sys,
                                                             uint32_t level,
                                                             uint32_t T,
                                                             uint8_t *flags) {
  const SysDesc S = sys[blockIdx.y];
  if (level >= S.sub_level) return;
  const TreeDev t = tree_at(t0, S.node_off);
  uint32_t k = blockIdx.x / T, q = blockIdx.x % T;
  uint32_t h = (1u << level) - 1 + k;
  uint32_t s = t.seg_start[h] + S.list_off, c = t.seg_count[h];
  int sd = t.split_dim[h];
  uint32_t mid = s + c / 2, e = s + c;
  if (c == 0) return;
  const uint32_t *ls = LB.in(t.lmask[h], sd);
  uint32_t base = s + q * LVL_TILE;
  // every load first (indices clamped into the segment, so all are in
  // bounds and independent), then the stores of the items that exist
  uint32_t id[LVL_ITEMS];
#pragma unroll
  for (int i = 0; i < LVL_ITEMS; ++i) id[i] = ls[min(base + i * LVL_THREADS + threadIdx.x, e - 1)];
#pragma unroll
  for (int i = 0; i < LVL_ITEMS; ++i) {
    uint32_t j = base + i * LVL_THREADS + threadIdx.x;
    if (j < e) flags[id[i]] = j < mid ? 1 : 0;
  }
}

// The left/right flags packed to one bit per id: 1/8 of the bytes, so the
// random reads of the count and scatter passes stay in L2.
__global__ void k_level_pack(const uint8_t *__restrict__ flags, uint32_t *__restrict__ bits,
                             uint32_t words) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words) return;
  const uint4 *src = reinterpret_cast<const uint4 *>(flags) + 2 * (size_t)w;
  const uint4 a = src[0], b = src[1];
  const uint32_t q[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    // bytes are 0 or 1: gather bit 0 of each byte
    const uint32_t x = q[i];
    const uint32_t nib = (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
    v |= nib << (4 * i);
  }
  bits[w] = v;
}

__device__ __forceinline__ uint32_t flag_bit(const uint32_t *__restrict__ bits, uint32_t id) {
  return (bits[id >> 5] >> (id & 31u)) & 1u;
}

__global__ __launch_bounds__(LVL_THREADS) void k_level_count(ListBufs LB, TreeDev t0,
                                                             const SysDesc *sys,
                                                             uint32_t level,
                                                             uint32_t T,
                                                             const uint32_t *__restrict__ bits,
                                                             uint32_t *cnt, int lazy) {
  __shared__ uint32_t sh[LVL_THREADS / 64];
  const SysDesc S = sys[blockIdx.y];
  if (level >= S.sub_level) return;
  const TreeDev t = tree_at(t0, S.node_off);
  uint32_t nseg = 1u << level;
  cnt += (size_t)blockIdx.y * 3 * nseg * T;
  uint32_t k = blockIdx.x / T, q = blockIdx.x % T;
  uint32_t h = nseg - 1 + k;
  uint32_t s = t.seg_start[h] + S.list_off, c = t.seg_count[h];
  int sd = t.split_dim[h];
  uint32_t e = s + c;
  uint32_t base = s + q * LVL_TILE;
  if (c == 0) return;
  for (int d = 0; d < 3; ++d) {
    // the split list is not partitioned, the lazy one not kept (uniform)
    if (d == sd || d == lazy) continue;
    // list entries, then their flag words, each as a batch of independent
    // loads (indices clamped into the segment); items past e count 0
    const uint32_t *ld = LB.in(t.lmask[h], d);
    uint32_t id[LVL_ITEMS], w[LVL_ITEMS];
#pragma unroll
    for (int i = 0; i < LVL_ITEMS; ++i)
      id[i] = ld[min(base + i * LVL_THREADS + threadIdx.x, e - 1)];
#pragma unroll
    for (int i = 0; i < LVL_ITEMS; ++i) w[i] = bits[id[i] >> 5];
    uint32_t local = 0;
#pragma unroll
    for (int i = 0; i < LVL_ITEMS; ++i) {
      uint32_t j = base + i * LVL_THREADS + threadIdx.x;
      local += j < e ? (w[i] >> (id[i] & 31u)) & 1u : 0u;
    }
    // block total: wavefront sums, then the first thread adds them
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
#pragma unroll
      for (int w2 = 0; w2 < LVL_THREADS / 64; ++w2) tot += sh[w2];
      cnt[((size_t)d * nseg + k) * T + q] = tot;
    }
    __syncthreads();
  }
}

// Exclusive scan over the T tile counts of one (list, segment): thread i
// sums a run of ceil(T / NT) consecutive counts, one block scan joins the
// runs, and each thread writes its run's prefixes (so a level-0 row of
// ~10^4 tiles takes one pass instead of one per 256 counts).  Rows of at
// most SCAN_SMALL_T tiles (the deeper levels) take one wavefront each.
constexpr uint32_t SCAN_SMALL_T = 1024;
template <int NT>
__global__ __launch_bounds__(NT) void k_level_scan(uint32_t *cnt, uint32_t T, const SysDesc *sys,
                                                   uint32_t level) {
  __shared__ uint32_t sh[NT / 64];
  if (level >= sys[blockIdx.y].sub_level) return;
  const uint32_t rows = 3u << level;
  uint32_t *row = cnt + ((size_t)blockIdx.y * rows + blockIdx.x) * T;
  const uint32_t per = (T + NT - 1) / NT;
  const uint32_t b = min(threadIdx.x * per, T), e = min(b + per, T);
  uint32_t sum = 0;
  for (uint32_t i = b; i < e; ++i) sum += row[i];
  uint32_t tot;
  uint32_t run = block_excl_sum<NT>(sum, sh, tot);
  for (uint32_t i = b; i < e; ++i) {
    const uint32_t v = row[i];
    row[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(LVL_THREADS) void k_level_scatter(ListBufs LB,
                                                               TreeDev t0,
                                                               const SysDesc *sys,
                                                               uint32_t level,
                                                               uint32_t T,
                                                               const uint32_t *__restrict__ bits,
                                                               const uint32_t *cnt, int lazy) {
  __shared__ uint32_t sh[LVL_THREADS / 64];
  const SysDesc S = sys[blockIdx.y];
  if (level >= S.sub_level) return;
  const TreeDev t = tree_at(t0, S.node_off);
  uint32_t nseg = 1u << level;
  cnt += (size_t)blockIdx.y * 3 * nseg * T;
  uint32_t k = blockIdx.x / T, q = blockIdx.x % T;
  uint32_t h = nseg - 1 + k;
  uint32_t s = t.seg_start[h] + S.list_off, c = t.seg_count[h];
  int sd = t.split_dim[h];
  if (c == 0) return;
  uint32_t mid = s + c / 2, e = s + c;
  uint32_t base = s + q * LVL_TILE + threadIdx.x * LVL_ITEMS;
  const uint32_t m = t.lmask[h];
  for (int d = 0; d < 3; ++d) {
    // the split list is already partitioned at the median and stays in
    // place; the lazy list is not kept
    if (d == sd || d == lazy) continue;
    const uint32_t *ld = LB.in(m, d);
    uint32_t *od = LB.out(m, d);
    uint32_t id[LVL_ITEMS];
    uint32_t f[LVL_ITEMS];
    uint32_t local = 0;
    // batches of independent loads, as in k_level_count
#pragma unroll
    for (int i = 0; i < LVL_ITEMS; ++i) id[i] = ld[min(base + i, e - 1)];
#pragma unroll
    for (int i = 0; i < LVL_ITEMS; ++i) f[i] = bits[id[i] >> 5];
#pragma unroll
    for (int i = 0; i < LVL_ITEMS; ++i) {
      f[i] = base + i < e ? (f[i] >> (id[i] & 31u)) & 1u : 0u;
      local += f[i];
    }
    uint32_t tot;
    uint32_t ex = block_excl_sum<LVL_THREADS>(local, sh, tot);
    uint32_t before = cnt[((size_t)d * nseg + k) * T + q] + ex;  // lefts before j in segment
#pragma unroll
    for (int i = 0; i < LVL_ITEMS; ++i) {
      uint32_t j = base + i;
      if (j < e) {
        uint32_t dst = f[i] ? s + before : mid + (j - s - before);
        od[dst] = id[i];
      }
      before += f[i];
    }
  }
}

// Device-wide levels [mid_level, sub_level) of each system, one workgroup per
// segment of level mid_level (at most MID_MAX bodies; a task lists the
// system and the segment): each level takes the same splits and
// the same stable partitions as k_level_split, k_level_flags, k_level_count
// and k_level_scatter, with the workgroup's barriers between the steps in
// place of kernel boundaries (no tile counts, no scan pass, no bit packing:
// the partition gathers the flag bytes of its own bodies).
constexpr int MID_THREADS = 1024, MID_ITEMS = 4;
constexpr uint32_t MID_TILE = MID_THREADS * MID_ITEMS;
#ifndef KD_MID_MAX
#define KD_MID_MAX 65536
#endif
constexpr uint32_t MID_MAX = KD_MID_MAX;
// and only from a level with at least this many segments (fewer workgroups
// leave most CUs idle, where the device-wide kernels spread every level)
#ifndef KD_MID_MIN_LEVEL
#define KD_MID_MIN_LEVEL 8
#endif
constexpr int MID_MIN_LEVEL = KD_MID_MIN_LEVEL;
constexpr uint32_t MID_SUBS = 64;  // segments of one workgroup at its deepest level
static_assert(MID_MAX / SUB_MAX <= MID_SUBS, "sub-segment table too small");

__global__ __launch_bounds__(MID_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void
k_level_mid(ListBufs LB, Coords P, TreeDev t0, const SysDesc *sys, const uint2 *tasks,
            const uint32_t *oid, const unsigned long long *range, int lazy, uint32_t *viol,
            uint8_t *flags) {
  __shared__ uint32_t s_start[MID_SUBS], s_cnt[MID_SUBS];
  __shared__ uint8_t s_sd[MID_SUBS], s_mask[MID_SUBS], s_tie[MID_SUBS];
  __shared__ uint32_t scan_sh[MID_THREADS / 64];
  // this workgroup's system and its segment at the system's mid_level
  const uint2 task = tasks[blockIdx.x];
  const SysDesc S = sys[task.x];
  const TreeDev t = tree_at(t0, S.node_off);
  const uint32_t k = task.y, L0 = S.mid_level, L1 = S.sub_level;
  const uint32_t tid = threadIdx.x;
  for (uint32_t lev = L0; lev < L1; ++lev) {
    const uint32_t nsub = 1u << (lev - L0);
    const uint32_t hb = (1u << lev) - 1 + (k << (lev - L0));
    // the splits, one thread per segment
    for (uint32_t j = tid; j < nsub; j += MID_THREADS) {
      const uint32_t h = hb + j;
      const uint32_t s = t.seg_start[h] + S.list_off, c = t.seg_count[h];
      const uint32_t m = t.lmask[h];
      bool tie;
      s_sd[j] = (uint8_t)split_decide(LB, P, t, h, s, c, m, range, lazy, viol, tie);
      s_start[j] = s;
      s_cnt[j] = c;
      s_mask[j] = (uint8_t)m;
      s_tie[j] = tie ? 1 : 0;
    }
    __syncthreads();
    for (uint32_t j = 0; j < nsub; ++j) {
      if (!s_tie[j]) continue;
      const int sd = s_sd[j];
      tie_fix<MID_THREADS>(LB.in(s_mask[j], sd), LB.out(s_mask[j], sd), P.c[sd], oid, s_start[j],
                           s_cnt[j]);
      __threadfence_block();
      __syncthreads();
    }
    // left/right flag of every body, from its segment's split list (four
    // positions a thread at a time: the list loads, then the flag stores)
    for (uint32_t j = 0; j < nsub; ++j) {
      const uint32_t s = s_start[j], e = s + s_cnt[j], mid = s + s_cnt[j] / 2;
      const uint32_t *ls = LB.in(s_mask[j], s_sd[j]);
      for (uint32_t p0 = s + tid; p0 < e; p0 += 4 * MID_THREADS) {
        uint32_t id[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t p = p0 + q * MID_THREADS;
          id[q] = ls[min(p, e - 1)];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t p = p0 + q * MID_THREADS;
          if (p < e) flags[id[q]] = p < mid ? 1 : 0;
        }
      }
    }
    __threadfence_block();
    __syncthreads();
    // stable partition of each list that moves (not the split list, which is
    // partitioned already, nor the lazy one), in tiles of MID_TILE positions
    for (int d = 0; d < 3; ++d) {
      if (d == lazy) continue;
      for (uint32_t j = 0; j < nsub; ++j) {
        if (d == s_sd[j]) continue;
        const uint32_t s = s_start[j], e = s + s_cnt[j], mid = s + s_cnt[j] / 2;
        const uint32_t *ld = LB.in(s_mask[j], d);
        uint32_t *od = LB.out(s_mask[j], d);
        uint32_t left = s, right = mid;  // next destinations of the tile
        // software pipelined: while tile t is scanned and scattered, the flag
        // gathers of tile t + 1 and the list loads of tile t + 2 are in flight
        auto load_ids = [&](uint32_t base, uint32_t (&id)[MID_ITEMS]) {
          const uint32_t b0 = base + tid * MID_ITEMS;
#pragma unroll
          for (int i = 0; i < MID_ITEMS; ++i) id[i] = ld[min(b0 + i, e - 1)];
        };
        auto gather = [&](uint32_t base, const uint32_t (&id)[MID_ITEMS],
                          uint32_t (&f)[MID_ITEMS]) {
          const uint32_t b0 = base + tid * MID_ITEMS;
#pragma unroll
          for (int i = 0; i < MID_ITEMS; ++i) f[i] = b0 + i < e ? flags[id[i]] : 0u;
        };
        const uint32_t ntile = (e - s + MID_TILE - 1) / MID_TILE;
        uint32_t idc[MID_ITEMS], fc[MID_ITEMS], idn[MID_ITEMS] = {};
        load_ids(s, idc);
        gather(s, idc, fc);
        if (ntile > 1) load_ids(s + MID_TILE, idn);
        for (uint32_t t = 0; t < ntile; ++t) {
          const uint32_t base = s + t * MID_TILE;
          uint32_t fn[MID_ITEMS] = {}, idnn[MID_ITEMS] = {};
          if (t + 1 < ntile) gather(base + MID_TILE, idn, fn);
          if (t + 2 < ntile) load_ids(base + 2 * MID_TILE, idnn);
          const uint32_t b0 = base + tid * MID_ITEMS;
          uint32_t local = 0;
#pragma unroll
          for (int i = 0; i < MID_ITEMS; ++i) local += fc[i];
          uint32_t tot;
          const uint32_t ex = block_excl_sum<MID_THREADS>(local, scan_sh, tot);
          uint32_t dl = left + ex, dr = right + (b0 - base) - ex;
#pragma unroll
          for (int i = 0; i < MID_ITEMS; ++i) {
            if (b0 + i >= e) break;
            if (fc[i]) od[dl++] = idc[i];
            else od[dr++] = idc[i];
          }
          left += tot;
          right += min(MID_TILE, e - base) - tot;
#pragma unroll
          for (int i = 0; i < MID_ITEMS; ++i) {
            idc[i] = idn[i];
            fc[i] = fn[i];
            idn[i] = idnn[i];
          }
        }
      }
    }
    __threadfence_block();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// subtree builder: one workgroup per segment of <= SUB_MAX bodies
// ---------------------------------------------------------------------------

struct SubArgs {
  ListBufs LB;      // sorted lists; a node's buffers per its mask
  const SysDesc *sys;
  Coords P;         // positions, working order
  const double *m;  // masses, working order
  StateDev in;      // state, working order
  StateDev out;     // state, tree order
  TreeDev t;
  uint32_t *loc;    // scratch: working id -> local index
  uint32_t *perm;   // tree position -> working id
  double4 *pos64;   // tree order (x, y, z, m)
  float4 *pos32;
  int leaf;
  double theta2;
  int lazy;  // a dimension whose list the device-wide levels did not keep (3: none)
  // (system, segment) of every workgroup in a forest (nullptr: a single
  // system, workgroup k takes segment k)
  const uint2 *tasks;
};

// Heap index (1-based) of the node after the subtree of heap slot h (0-based).
// Records store it times their own size: the byte offset at which the walk
// loads the next record (below 2^32 by the 2^26 heap-slot limit).
__device__ __forceinline__ uint32_t skip_of(uint32_t h) {
  uint32_t x = h + 1;
  x >>= __builtin_ctz(~x);
  return x == 0 ? 0u : x + 1;
}

__device__ void write_leaf_rec(const TreeDev &t, uint32_t h, uint32_t start,
                               uint32_t count) {
  const uint32_t skip = skip_of(h);
  NodeRec64 r;
  r.cx = r.cy = r.cz = r.m = r.size2 = 0.0;
  r.start = start;
  r.count = count;
  r.skip = skip * (uint32_t)sizeof(NodeRec64);
  r.pad0 = 0;
  r.pad1 = 0.0;
  if (t.recs & REC_64) t.rec64[h] = r;
  NodeRec32 q;
  q.cx = q.cy = q.cz = q.m = q.size2 = 0.f;
  q.start = start;
  q.count = count;
  q.skip = skip * (uint32_t)sizeof(NodeRec32);
  if (t.recs & REC_32) t.rec32[h] = q;
  NodeRecD dd;
  dd.cx = dd.cy = dd.cz = 0.0;
  dd.m = dd.size2 = 0.f;
  dd.start = start;
  dd.count = count;
  dd.skip = skip * (uint32_t)sizeof(NodeRecD);
  dd.pad0 = 0;
  dd.pad1 = 0.0;
  if (t.recs & REC_D) t.recd[h] = dd;
}

__device__ void write_internal_rec(const TreeDev &t, uint32_t h, uint32_t start,
                                   uint32_t count, double theta2) {
#pragma clang fp contract(off)
  double m = t.mm[h];
  double cx = t.mx[h] / m, cy = t.my[h] / m, cz = t.mz[h] / m;
  double sz = t.size[h];
  double s2 = sz * sz;
  float open2 = (float)(s2 / theta2);  // theta = 0: +inf (or NaN): never accepted
  const uint32_t skip = skip_of(h);
  NodeRec64 r;
  r.cx = cx;
  r.cy = cy;
  r.cz = cz;
  r.m = m;
  r.size2 = s2;
  r.start = start;
  r.count = count;
  r.skip = skip * (uint32_t)sizeof(NodeRec64);
  r.pad0 = 0;
  r.pad1 = 0.0;
  if (t.recs & REC_64) t.rec64[h] = r;
  NodeRec32 q;
  q.cx = (float)cx;
  q.cy = (float)cy;
  q.cz = (float)cz;
  q.m = (float)m;
  q.size2 = open2;
  q.start = start;
  q.count = count;
  q.skip = skip * (uint32_t)sizeof(NodeRec32);
  if (t.recs & REC_32) t.rec32[h] = q;
  NodeRecD dd;
  dd.cx = cx;
  dd.cy = cy;
  dd.cz = cz;
  dd.m = (float)m;
  dd.size2 = open2;
  dd.start = start;
  dd.count = count;
  dd.skip = skip * (uint32_t)sizeof(NodeRecD);
  dd.pad0 = 0;
  dd.pad1 = 0.0;
  if (t.recs & REC_D) t.recd[h] = dd;
}

// Body range [s, s + cnt) of node c at depth lam below a segment of n0
// bodies (every split at floor(count / 2)); false if an ancestor is a leaf,
// so that the node does not exist.
__device__ __forceinline__ bool sub_node(uint32_t c, int lam, uint32_t n0, int leaf,
                                         uint32_t &s, uint32_t &cnt) {
  s = 0;
  cnt = n0;
  for (int b = lam - 1; b >= 0; --b) {
    if (cnt <= (uint32_t)leaf) return false;
    const uint32_t half = cnt >> 1;
    if ((c >> b) & 1u) {
      s += half;
      cnt -= half;
    } else {
      cnt = half;
    }
  }
  return true;
}

// The lazy dimension's order inside a segment: the order of the presorted
// lists, (coordinate, working id).  Sorted by a bitonic network in the form
// whose comparators all sort upward: the merge of two sorted blocks into one
// of `size` positions first compares mirror positions x and x ^ (size - 1),
// then x and x + stride for strides size/4 .. 1.  Positions at or past n0
// act as +inf: they are never read or written, and every comparator whose
// upper position is one of them is skipped.  A pass runs up to LZ_G
// consecutive stages of one merge in registers, 2^g positions a thread, so
// the values cross LDS once per LZ_G stages, not once per stage.  At 10 M
// ring bodies (segments of ~1220) a build takes 3.07 ms with 2 stages a pass
// on 32-bit values (3.09 ms with 3 stages, half the threads idle; 3.17 ms on
// 64-bit keys; 3.45 ms for one LDS round trip per stage over every position
// of the padded block).
#ifndef KD_LZ_G
#define KD_LZ_G 2
#endif
constexpr int LZ_G = KD_LZ_G;

// The values the network sorts, in two forms.  Packed: a 21-bit fixed-point
// key of the coordinate over the segment's own range and the 11-bit local
// index in one word (4 bytes a value, one compare a comparator); bodies whose
// keys collide are put in exact order afterwards.  Wide: the 64-bit ordered
// key of the coordinate and the local index, equal keys by working id (exact
// by itself; the form a segment falls back to when keys collide in runs
// longer than LZ_RUN).
struct LzPacked {
  uint32_t *k;
  using V = uint32_t;
  __device__ V get(uint32_t p) const { return k[p]; }
  __device__ void put(uint32_t p, V v) const { k[p] = v; }
  __device__ void order(V &a, V &b) const {
    const V lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
  }
};
struct LzWide {
  uint64_t *k;
  uint16_t *l;
  const uint32_t *ids;
  struct V {
    uint64_t k;
    uint16_t l;
  };
  __device__ V get(uint32_t p) const { return V{k[p], l[p]}; }
  __device__ void put(uint32_t p, const V &v) const {
    k[p] = v.k;
    l[p] = v.l;
  }
  __device__ void order(V &a, V &b) const {
    if (a.k > b.k || (a.k == b.k && ids[a.l] > ids[b.l])) {
      const V t = a;
      a = b;
      b = t;
    }
  }
};
constexpr uint32_t LZ_KEY_BITS = 21, LZ_KEY_MAX = (1u << LZ_KEY_BITS) - 1, LZ_RUN = 16;
static_assert(SUB_MAX <= (1 << (32 - LZ_KEY_BITS)), "local indices fill the packed low bits");

// Merges of sizes 2 .. 2^G over runs of 2^G consecutive positions.
template <int NT, int G, class Pol>
__device__ __forceinline__ void lz_first(const Pol &pol, uint32_t n0, uint32_t M, int tid) {
  constexpr int E = 1 << G;
  for (uint32_t t = tid; t < (M >> G); t += NT) {
    const uint32_t a = t << G;
    if (a >= n0) break;
    typename Pol::V v[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (a + e < n0) v[e] = pol.get(a + e);
#pragma unroll
    for (int q = 1; q <= G; ++q) {
#pragma unroll
      for (int r = q - 1; r >= 0; --r) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          if (e & (1 << r)) continue;
          const int f = r == q - 1 ? e ^ ((1 << q) - 1) : e + (1 << r);
          if (a + f < n0) pol.order(v[e], v[f]);
        }
      }
    }
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (a + e < n0) pol.put(a + e, v[e]);
  }
}

// G stages of the merge into blocks of `size`, with strides 2^ls ..
// 2^(ls-G+1), less the first S of them; MIRROR: the first is the mirror
// stage (2^ls = size / 2).  Set t
// holds the positions a + the sums of subsets of those strides, a = t with
// zero bits inserted at the strides' bit positions; with MIRROR the upper
// half of the set is the mirror image of the lower half, listed so that
// positions still increase with the set index.
template <int NT, int G, bool MIRROR, int S, class Pol>
__device__ __forceinline__ void lz_pass(const Pol &pol, uint32_t n0, uint32_t M, uint32_t size,
                                        int ls, int tid) {
  constexpr int E = 1 << G;
  const int lo = ls - G + 1;
  for (uint32_t t = tid; t < (M >> G); t += NT) {
    const uint32_t a = ((t >> lo) << (ls + 1)) | (t & ((1u << lo) - 1));
    if (a >= n0) break;  // a grows with t, and is the set's lowest position
    uint32_t pos[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      uint32_t off = 0, full = 0;
#pragma unroll
      for (int b = 0; b < G; ++b) {
        if (MIRROR && b == G - 1) break;
        if ((e >> b) & 1) off += 1u << (lo + b);
        full += 1u << (lo + b);
      }
      pos[e] = (MIRROR && (e >> (G - 1))) ? (a + full - off) ^ (size - 1) : a + off;
    }
    typename Pol::V v[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (pos[e] < n0) v[e] = pol.get(pos[e]);
#pragma unroll
    for (int j = S; j < G; ++j) {
      const int bit = 1 << (G - 1 - j);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (e & bit) continue;
        const int f = (MIRROR && j == 0) ? e ^ (E - 1) : e | bit;
        if (pos[f] < n0) pol.order(v[e], v[f]);
      }
    }
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (pos[e] < n0) pol.put(pos[e], v[e]);
  }
}

// The network over the n0 <= SUB_MAX values already in LDS; ends with a
// barrier.  Every pass gives thread t one set of 2^G positions (a merge's
// last stages that do not fill a pass run as the last pass shape with its
// first stages skipped), so a pass whose largest stride is below 2^(6+G)
// keeps wavefront w within positions [64 w 2^G, 64 (w+1) 2^G): between two
// such passes the wavefront only waits for its own LDS writes.
template <int NT, int G, class Pol>
__device__ void lz_sort(const Pol &pol, uint32_t n0, int tid) {
  static_assert((SUB_MAX >> G) <= NT, "one set of positions a thread");
  static_assert(G >= 2 && G <= 3, "passes of 2 or 3 stages");
  constexpr int WL = 6 + G;
  int lm = G;
  while ((1u << lm) < n0) ++lm;
  const uint32_t M = 1u << lm;
  lz_first<NT, G>(pol, n0, M, tid);
  int prev = G - 1;  // the largest stride of the last pass, as a power of 2
  auto sync = [&](int next) {
    if (prev + 1 <= WL && next + 1 <= WL) wave_sync();
    else __syncthreads();
    prev = next;
  };
  for (int m = G + 1; m <= lm; ++m) {
    const uint32_t size = 1u << m;
    // m > G stages: the mirror stage and G - 1 strides, then the rest
    sync(m - 1);
    lz_pass<NT, G, true, 0>(pol, n0, M, size, m - 1, tid);
    for (int ls = m - 1 - G; ls >= 0; ls -= G) {
      if (ls + 1 >= G) {
        sync(ls);
        lz_pass<NT, G, false, 0>(pol, n0, M, size, ls, tid);
      } else {
        // the last ls + 1 < G stages: strides 2^ls .. 1
        sync(G - 1);
        if (G - 1 - ls == 1) lz_pass<NT, G, false, 1>(pol, n0, M, size, G - 1, tid);
        else lz_pass<NT, G, false, (G >= 3 ? 2 : 1)>(pol, n0, M, size, G - 1, tid);
      }
    }
  }
  __syncthreads();
}

// The lazy dimension's order of a segment: local indices into iz.  buf: 8
// SUB_MAX bytes of LDS; shd: 2 NT / 64 doubles; flagp: one word.  The packed
// form first; a run of more than LZ_RUN colliding keys (a segment whose
// coordinates are all equal, or nearly) sends the segment to the wide form.
template <int NT>
__device__ void lz_order(unsigned char *buf, uint16_t *iz, const uint32_t *ids, const double *zc,
                         uint32_t n0, int tid, double *shd, uint32_t *flagp) {
  constexpr int IT = SUB_MAX / NT;
  constexpr int NW = NT / 64;
  double z[IT];
  double lo = INFINITY, hi = -INFINITY;
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const uint32_t i = tid + q * NT;
    z[q] = i < n0 ? zc[ids[i]] : 0.0;
    if (i < n0) {
      lo = fmin(lo, z[q]);
      hi = fmax(hi, z[q]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  if ((tid & 63) == 0) {
    shd[tid >> 6] = lo;
    shd[NW + (tid >> 6)] = hi;
  }
  if (tid == 0) *flagp = 0;
  __syncthreads();
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    lo = fmin(lo, shd[w]);
    hi = fmax(hi, shd[NW + w]);
  }
  // all coordinates equal (a planar system) leaves the order to the working
  // ids, which the wide form gives directly
  if (hi > lo) {
    // a monotone map of the coordinate, so distinct keys are in coordinate
    // order and only equal keys need the exact comparison
    const double scale = (double)LZ_KEY_MAX / (hi - lo);
    uint32_t *kp = reinterpret_cast<uint32_t *>(buf);
#pragma unroll
    for (int q = 0; q < IT; ++q) {
      const uint32_t i = tid + q * NT;
      if (i < n0) {
        const uint32_t key = (uint32_t)fmin(fmax((z[q] - lo) * scale, 0.0), (double)LZ_KEY_MAX);
        kp[i] = (key << (32 - LZ_KEY_BITS)) | i;
      }
    }
    __syncthreads();
    lz_sort<NT, LZ_G>(LzPacked{kp}, n0, tid);
    constexpr uint32_t IMASK = (1u << (32 - LZ_KEY_BITS)) - 1;
#pragma unroll
    for (int q = 0; q < IT; ++q) {
      const uint32_t i = tid + q * NT;
      if (i < n0) iz[i] = (uint16_t)(kp[i] & IMASK);
    }
    __syncthreads();
    // runs of equal keys, each ordered by the thread at its start by
    // (coordinate, working id)
#pragma unroll
    for (int q = 0; q < IT; ++q) {
      const uint32_t i = tid + q * NT;
      if (i + 1 >= n0) continue;
      const uint32_t key = kp[i] >> (32 - LZ_KEY_BITS);
      if ((kp[i + 1] >> (32 - LZ_KEY_BITS)) != key) continue;
      if (i > 0 && (kp[i - 1] >> (32 - LZ_KEY_BITS)) == key) continue;
      uint32_t j = i + 1;
      while (j < n0 && j - i <= LZ_RUN && (kp[j] >> (32 - LZ_KEY_BITS)) == key) ++j;
      if (j - i > LZ_RUN) {
        *flagp = 1;
        continue;
      }
      for (uint32_t x = i + 1; x < j; ++x) {
        const uint16_t v = iz[x];
        const uint32_t idv = ids[v];
        const double zv = zc[idv];
        uint32_t y = x;
        while (y > i) {
          const uint16_t u = iz[y - 1];
          const uint32_t idu = ids[u];
          const double zu = zc[idu];
          if (zu < zv || (zu == zv && idu < idv)) break;
          iz[y] = u;
          --y;
        }
        iz[y] = v;
      }
    }
  } else if (tid == 0) {
    *flagp = 1;
  }
  __syncthreads();
  if (*flagp) {
    uint64_t *kz = reinterpret_cast<uint64_t *>(buf);
#pragma unroll
    for (int q = 0; q < IT; ++q) {
      const uint32_t i = tid + q * NT;
      if (i < n0) {
        kz[i] = ordkey(z[q]);
        iz[i] = (uint16_t)i;
      }
    }
    __syncthreads();
    lz_sort<NT, 2>(LzWide{kz, iz, ids}, n0, tid);
  }
}

// At most 64 VGPRs, so that four workgroups (the LDS limit) fit on a CU
// rather than three: the build takes 3.27 ms against 3.54 ms at 10 M bodies.
#ifndef KD_SUB_WAVES
#define KD_SUB_WAVES 8
#endif
__global__ __launch_bounds__(SUB_THREADS) __attribute__((amdgpu_waves_per_eu(KD_SUB_WAVES))) void
k_subtree(SubArgs a) {
#pragma clang fp contract(off)
  constexpr int NT = SUB_THREADS, IT = SUB_ITEMS;
  // one LDS buffer: ids (working id of every local index), lst (the three
  // sorted lists of local indices, double-buffered), flag, sdarr; once the
  // bodies are placed, the node moments reuse it as staging and node sums
  __shared__ __align__(16) unsigned char smem[18 * SUB_MAX];
  uint32_t *ids = reinterpret_cast<uint32_t *>(smem);
  uint16_t(*lst)[3][SUB_MAX] = reinterpret_cast<uint16_t(*)[3][SUB_MAX]>(smem + 4 * SUB_MAX);
  uint8_t *flag = smem + 16 * SUB_MAX;
  uint8_t *sdarr = smem + 17 * SUB_MAX;
  __shared__ uint64_t shv[NT / 64];

  __shared__ uint32_t shh[NT / 64];
  // per level (by parity): some segment of the level splits, so bodies are
  // still alive below it
  __shared__ uint32_t splits[2];
  // the lazy dimension's sort: block range of the coordinate, fallback flag
  __shared__ double shd[2 * NT / 64];
  __shared__ uint32_t lzflag;

  const uint2 task = a.tasks ? a.tasks[blockIdx.x] : make_uint2(0u, blockIdx.x);
  const SysDesc S = a.sys[task.x];
  const uint32_t k = task.y;
  const uint32_t level = S.sub_level;
  if (k >= (1u << level)) return;
  const TreeDev t = tree_at(a.t, S.node_off);
  const uint32_t m0 = t.lmask[(1u << level) - 1 + k];
  // the segment's ids come from list g0; g1 (and 2 when no dimension is
  // lazy) are read as local indices, the lazy one is sorted here
  const int lz = a.lazy;
  const int g0 = lz == 0 ? 1 : 0, g1 = lz == 1 || lz == 0 ? 2 : 1;
  const uint32_t *L0 = a.LB.in(m0, g0), *L1 = a.LB.in(m0, g1), *L2 = a.LB.in(m0, 2);
  const uint32_t h0 = (1u << level) - 1 + k;
  // s0l: first body of the segment within its system; s0: in the body
  // arrays (tree order); sl: in the sorted lists
  const uint32_t s0l = t.seg_start[h0], n0 = t.seg_count[h0];
  const uint32_t s0 = s0l + S.body_off, sl = s0l + S.list_off;
  if (n0 == 0) return;
  const int leaf = a.leaf;
  const int tid = threadIdx.x;
  if (tid < 2) splits[tid] = 0;

  for (uint32_t i = tid; i < n0; i += NT) {
    uint32_t id = L0[sl + i];
    ids[i] = id;
    lst[0][g0][i] = (uint16_t)i;
    a.loc[id] = i;
  }
  __threadfence_block();
  __syncthreads();
  for (uint32_t i = tid; i < n0; i += NT) {
    lst[0][g1][i] = (uint16_t)a.loc[L1[sl + i]];
    if (lz == 3) lst[0][2][i] = (uint16_t)a.loc[L2[sl + i]];
  }
  __syncthreads();
  if (lz < 3) {
    // the lazy dimension's order; the keys use the LDS of the second list
    // buffer, flags and sdarr (not live yet), the local indices end in
    // lst[0][lz]
    static_assert(18 * SUB_MAX - 10 * SUB_MAX >= 8 * SUB_MAX, "lazy-sort keys need 8 SUB_MAX bytes");
    lz_order<NT>(smem + 10 * SUB_MAX, lst[0][lz], ids, a.P.c[lz], n0, tid, shd, &lzflag);
  }

  uint32_t st[IT], ct[IT], cc[IT], hl[IT];
  bool alive[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    uint32_t j = tid * IT + it;
    alive[it] = j < n0;
    st[it] = 0;
    ct[it] = n0;
    cc[it] = 0;
    hl[it] = h0;
  }
  int cur = 0;
  int lam_end = 0;
  for (int lam = 0;; ++lam) {
    const uint32_t hbase = (1u << (level + lam)) - 1 + (k << lam);
    // one thread per segment of this level: split decision or leaf mark
    for (uint32_t c = tid; c < (1u << lam); c += NT) {
      uint32_t sgs, sgc;
      if (!sub_node(c, lam, n0, leaf, sgs, sgc)) continue;  // an ancestor is a leaf
      const uint32_t h = hbase + c;
      if (sgc <= (uint32_t)leaf) {
        sdarr[sgs] = 0xFF;  // leaf: its bodies are finished at this level
      } else {
        // the ends and the two middle keys of every dimension as one batch of
        // independent loads (one round trip; the dimension index is a
        // constant in each load, so no pointer is fetched at run time)
        const uint32_t mid = sgs + sgc / 2;
        const uint16_t *l0 = lst[cur][0], *l1 = lst[cur][1], *l2 = lst[cur][2];
        const double *c0 = a.P.c[0], *c1 = a.P.c[1], *c2 = a.P.c[2];
        const double e0 = c0[ids[l0[sgs + sgc - 1]]] - c0[ids[l0[sgs]]];
        const double e1 = c1[ids[l1[sgs + sgc - 1]]] - c1[ids[l1[sgs]]];
        const double e2 = c2[ids[l2[sgs + sgc - 1]]] - c2[ids[l2[sgs]]];
        const double m0 = c0[ids[l0[mid]]], p0 = c0[ids[l0[mid - 1]]];
        const double m1 = c1[ids[l1[mid]]], p1 = c1[ids[l1[mid - 1]]];
        const double m2 = c2[ids[l2[mid]]], p2 = c2[ids[l2[mid - 1]]];
        // argmax with ties to the lower dimension, by selects
        const bool s1 = e1 > e0;
        const double b01 = s1 ? e1 : e0;
        const bool s2 = e2 > b01;
        const int sd = s2 ? 2 : (s1 ? 1 : 0);
        const double best = s2 ? e2 : b01;
        const double kv = s2 ? m2 : (s1 ? m1 : m0);
        const double kprev = s2 ? p2 : (s1 ? p1 : p0);
        uint16_t *ls = lst[cur][sd];
        const double *key = a.P.c[sd];
        if (kprev == kv) {
          // equal keys straddle the median: order the run by original index
          uint32_t r0 = mid, r1 = mid + 1;
          while (r0 > sgs && key[ids[ls[r0 - 1]]] == kv) --r0;
          while (r1 < sgs + sgc && key[ids[ls[r1]]] == kv) ++r1;
          for (uint32_t x = r0 + 1; x < r1; ++x) {
            uint16_t v = ls[x];
            uint32_t ov = a.in.oid[ids[v]];
            int y = (int)x - 1;
            while (y >= (int)r0 && a.in.oid[ids[ls[y]]] > ov) {
              ls[y + 1] = ls[y];
              --y;
            }
            ls[y + 1] = v;
          }
        }
        t.split_dim[h] = (uint8_t)sd;
        t.split_val[h] = kv;
        t.size[h] = best;
        sdarr[sgs] = (uint8_t)sd;
        splits[lam & 1] = 1;
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      if (alive[it] && sdarr[st[it]] == 0xFF) {
        alive[it] = false;
        hl[it] = hbase + cc[it];  // heap slot of the leaf this body ends in
      }
    }
    lam_end = lam;
    // a body is still alive iff its segment splits at this level
    if (!splits[lam & 1]) break;
    // the other parity's flag was last read at the previous level, before
    // the barrier above; the next level sets it after later barriers
    if (tid == 0) splits[(lam + 1) & 1] = 0;

    // flags from the split-dim list
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      if (!alive[it]) continue;
      uint32_t j = tid * IT + it;
      int sd = sdarr[st[it]];
      uint32_t mid = st[it] + ct[it] / 2;
      flag[lst[cur][sd][j]] = j < mid ? 1 : 0;
    }
    __syncthreads();

    // segmented exclusive scan of the left flags, three lists packed
    uint64_t v[IT];
    uint32_t hd[IT];
    SegV agg;
    agg.v = 0;
    agg.h = 0;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      uint32_t j = tid * IT + it;
      v[it] = 0;
      hd[it] = 1;
      if (alive[it]) {
        int sd = sdarr[st[it]];
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if (d != sd) v[it] += (uint64_t)flag[lst[cur][d][j]] << (16 * d);
        hd[it] = (j == st[it]) ? 1u : 0u;
      }
      SegV e;
      e.v = v[it];
      e.h = hd[it];
      agg = seg_combine(agg, e);
    }
    SegV carry = block_seg_carry<NT>(agg, shv, shh);
    uint64_t run = carry.v;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      uint32_t j = tid * IT + it;
      if (hd[it]) run = 0;
      uint64_t ex = run;  // lefts before j in its segment, per list
      run += v[it];
      if (j >= n0) continue;
      int sd = alive[it] ? sdarr[st[it]] : -1;
      uint32_t mid = st[it] + ct[it] / 2;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        uint16_t x = lst[cur][d][j];
        uint32_t dst = j;
        if (alive[it] && d != sd) {
          uint32_t e = (uint32_t)((ex >> (16 * d)) & 0xFFFFu);
          dst = flag[x] ? st[it] + e : mid + (j - st[it] - e);
        }
        lst[cur ^ 1][d][dst] = x;
      }
    }
    __syncthreads();
    cur ^= 1;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      if (!alive[it]) continue;
      uint32_t j = tid * IT + it;
      uint32_t mid = st[it] + ct[it] / 2;
      if (j < mid) {
        ct[it] = ct[it] / 2;
        cc[it] = 2 * cc[it];
      } else {
        ct[it] = ct[it] - ct[it] / 2;
        st[it] = mid;
        cc[it] = 2 * cc[it] + 1;
      }
    }
  }

  // Leaves in canonical order.  List 0 holds every leaf's members (in list
  // order); each body ranks itself among its leaf by original index, using
  // the spare list buffer as staging for those indices, and its state goes
  // straight to its slot in tree order.
  uint32_t *oidl = reinterpret_cast<uint32_t *>(&lst[cur ^ 1][0][0]);  // two lists' room
  uint16_t *permloc = &lst[cur ^ 1][2][0];  // tree position -> local index
  uint32_t ow[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    uint32_t j = tid * IT + it;
    ow[it] = 0;
    if (j < n0) {
      ow[it] = a.in.oid[ids[lst[cur][0][j]]];
      oidl[j] = ow[it];
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    uint32_t j = tid * IT + it;
    if (j >= n0) continue;
    uint32_t rank = 0;
    for (uint32_t q = st[it]; q < st[it] + ct[it]; ++q) rank += oidl[q] < ow[it] ? 1u : 0u;
    permloc[st[it] + rank] = lst[cur][0][j];
  }
  __syncthreads();
  for (uint32_t i = tid; i < n0; i += NT) {
    const uint32_t w = ids[permloc[i]];
    const uint32_t tpos = s0 + i;
    a.perm[tpos] = w;
    double x = a.in.x[w], y = a.in.y[w], z = a.in.z[w], m = a.in.m[w];
    a.out.x[tpos] = x;
    a.out.y[tpos] = y;
    a.out.z[tpos] = z;
    a.out.vx[tpos] = a.in.vx[w];
    a.out.vy[tpos] = a.in.vy[w];
    a.out.vz[tpos] = a.in.vz[w];
    a.out.m[tpos] = m;
    a.out.oid[tpos] = a.in.oid[w];
    // the walk's body records, for the arithmetic that reads them
    if (t.recs & (REC_64 | REC_D)) a.pos64[tpos] = make_double4(x, y, z, m);
    if (t.recs & REC_32) a.pos32[tpos] = make_float4((float)x, (float)y, (float)z, (float)m);
  }
  __threadfence_block();
  __syncthreads();
  // Node moments (m, m x, m y, m z).  The members' terms of one quantity at a
  // time are staged in LDS in tree order; each leaf's first thread adds its
  // members in canonical order, and the internal nodes are summed bottom-up
  // (left + right) in LDS, one level per barrier, so no level waits on
  // global memory.  Every sum also goes to the node arrays.  When the node
  // sums of all four quantities fit beside the staged terms (nslot heap
  // slots of the subtree each: the ring's segments of ~1220 bodies need 511),
  // the four bottom-up passes share their barriers; otherwise each quantity
  // takes its own pass with SUB_MAX slots (the deepest leaf of at most
  // SUB_MAX bodies lies above depth log2(SUB_MAX)).
  double *stg = reinterpret_cast<double *>(smem);
  static_assert(sizeof(smem) >= 2 * SUB_MAX * sizeof(double), "moments need 16 SUB_MAX bytes");
  double *const dsts[4] = {t.mm, t.mx, t.my, t.mz};
  const uint32_t nslot = (2u << lam_end) - 1;
  const bool joint = (n0 + 4 * nslot) * sizeof(double) <= sizeof(smem);
  // node sums: quantity q's at nsum + q * qstride
  double *nsum = joint ? stg + n0 : stg + SUB_MAX;
  const uint32_t qstride = joint ? nslot : 0;
  for (int q = 0; q < 4; ++q) {
    const double *pc = q == 1 ? a.out.x : (q == 2 ? a.out.y : a.out.z);
    for (uint32_t i = tid; i < n0; i += NT) {
      const double bm = a.out.m[s0 + i];
      stg[i] = q == 0 ? bm : bm * pc[s0 + i];
    }
    __syncthreads();
    double *dst = dsts[q];
    double *ns = nsum + q * qstride;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const uint32_t j = tid * IT + it;
      if (j >= n0 || j != st[it]) continue;
      double sum = 0.0;
      for (uint32_t r = 0; r < ct[it]; ++r) sum += stg[st[it] + r];
      dst[hl[it]] = sum;
      // the leaf's slot in the subtree's own heap: hl + 1 = 2^(level + lam)
      // + (k << lam) + index within depth lam
      const uint32_t lam = (31 - __clz(hl[it] + 1)) - level;
      ns[hl[it] + (1u << lam) - (1u << (level + lam)) - (k << lam)] = sum;
    }
    __syncthreads();
    if (joint && q < 3) continue;  // the bottom-up sums of all four follow the last
    for (int lam = lam_end - 1; lam >= 0; --lam) {
      const uint32_t hbase = (1u << (level + lam)) - 1 + (k << lam);
      for (uint32_t c = tid; c < (1u << lam); c += NT) {
        uint32_t s, cnt;
        if (!sub_node(c, lam, n0, leaf, s, cnt) || cnt <= (uint32_t)leaf) continue;
        const uint32_t rel = (1u << lam) - 1 + c;
        for (int r = joint ? 0 : q; r <= q; ++r) {
          double *nr = nsum + r * qstride;
          const double v = nr[2 * rel + 1] + nr[2 * rel + 2];
          nr[rel] = v;
          dsts[r][hbase + c] = v;
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const uint32_t j = tid * IT + it;
    if (j < n0 && j == st[it]) write_leaf_rec(t, hl[it], s0l + st[it], ct[it]);
  }
  // internal records, over the subtree's heap slots (at most two a thread):
  // their sums were written by the passes above
  for (uint32_t rel = tid; rel + 1 < (1u << lam_end); rel += NT) {
    const int lam = 31 - __clz(rel + 1);
    const uint32_t c = rel + 1 - (1u << lam);
    uint32_t s, cnt;
    if (!sub_node(c, lam, n0, leaf, s, cnt) || cnt <= (uint32_t)leaf) continue;
    write_internal_rec(t, (1u << (level + lam)) - 1 + (k << lam) + c, s0l + s, cnt, a.theta2);
  }
}

// Moments and records of the device-wide levels [0, level), bottom-up.
__global__ __launch_bounds__(1024) void k_moments_top(TreeDev t0, const SysDesc *sys,
                                                      double theta2) {
#pragma clang fp contract(off)
  const SysDesc S = sys[blockIdx.x];
  const TreeDev t = tree_at(t0, S.node_off);
  const uint32_t level = S.sub_level;
  for (int lev = (int)level - 1; lev >= 0; --lev) {
    uint32_t nseg = 1u << lev;
    for (uint32_t k = threadIdx.x; k < nseg; k += blockDim.x) {
      uint32_t h = nseg - 1 + k;
      uint32_t l = 2 * h + 1, r = 2 * h + 2;
      t.mm[h] = t.mm[l] + t.mm[r];
      t.mx[h] = t.mx[l] + t.mx[r];
      t.my[h] = t.my[l] + t.my[r];
      t.mz[h] = t.mz[l] + t.mz[r];
      write_internal_rec(t, h, t.seg_start[h], t.seg_count[h], theta2);
    }
    __threadfence_block();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

Shape make_shape(int64_t n, int leaf) {
  Shape s;
  s.n = n;
  s.leaf = leaf;
  int l = 0;
  int64_t mx = n;
  while (mx > leaf) {
    mx = (mx + 1) / 2;
    ++l;
  }
  s.levels = l + 1;
  int ls = 0;
  mx = n;
  while (mx > SUB_MAX) {
    mx = (mx + 1) / 2;
    ++ls;
  }
  s.sub_level = ls;
  s.slots = ((int64_t)1 << s.levels) - 1;
  return s;
}


SysDesc make_sys(const Shape &sh, uint32_t body_off, uint32_t node_off, uint32_t list_off) {
  SysDesc d{};
  d.body_off = body_off;
  d.list_off = list_off;
  d.n = (uint32_t)sh.n;
  d.node_off = node_off;
  d.slots = (uint32_t)sh.slots;
  d.sub_level = (uint32_t)sh.sub_level;
  d.levels = (uint32_t)sh.levels;
  return d;
}

// Radix sort of (32-bit key, 32-bit id) pairs: 8-bit digits, 1024 threads x
// 10 items per tile.  On gfx950 this sorts 10 M pairs in ~320 us, against
// ~400 us for rocprim's default tile of 1024 x 16 (wider digits, 10 or 11
// bits in three passes, measured slower).
#ifndef KD_SORT_BITS
#define KD_SORT_BITS 8
#endif
#ifndef KD_SORT_ITEMS
#define KD_SORT_ITEMS 10
#endif
using SortCfg = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, KD_SORT_ITEMS>,
                                        rocprim::kernel_config<1024, KD_SORT_ITEMS>, KD_SORT_BITS,
                                        rocprim::block_radix_rank_algorithm::match>>;

static int bits_for(uint32_t v) {
  int b = 1;
  while (b < 32 && (v >> b) != 0) ++b;
  return b;
}

size_t sort_temp_bytes(uint32_t n, int /*nsys*/) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs<SortCfg>(nullptr, bytes, (uint32_t *)nullptr,
                                           (uint32_t *)nullptr,
                                           rocprim::counting_iterator<uint32_t>(0u),
                                           (uint32_t *)nullptr, n, 0, 32);
  return bytes;
}

// Tiles per segment at a device-wide level.
static uint32_t level_tiles(const Forest &f, int lev) {
  const uint32_t nseg = 1u << lev;
  const uint32_t maxseg = (f.max_n + nseg - 1) / nseg;
  return (maxseg + LVL_TILE - 1) / LVL_TILE;
}

size_t level_cnt_entries(const Forest &f) {
  size_t m = 1;
  for (int lev = 0; lev < f.max_sub_level; ++lev)
    m = std::max(m, (size_t)f.nsys * 3 * ((size_t)1 << lev) * level_tiles(f, lev));
  return m;
}


hipError_t launch_shape(const TreeDev &t, const Forest &f, int leaf, hipStream_t s) {
  const unsigned blocks = (f.max_slots + 255) / 256;
  k_shape<<<dim3(blocks, (unsigned)f.nsys), dim3(256), 0, s>>>(t.seg_start, t.seg_count,
                                                               t.lmask, f.sys, leaf);
  return hipGetLastError();
}

size_t mid_tasks_max(int64_t n, int nsys) {
  // segments at a mid_level above MID_MIN_LEVEL hold more than MID_MAX / 2
  // bodies; at MID_MIN_LEVEL there are 2^MID_MIN_LEVEL of them
  return (size_t)(2 * n / MID_MAX) + (size_t)nsys + ((size_t)1 << MID_MIN_LEVEL) + 1;
}

void set_sub_tasks(const std::vector<SysDesc> &sys, std::vector<uint2> &tasks) {
  tasks.clear();
  for (size_t q = 0; q < sys.size(); ++q)
    for (uint32_t k = 0; k < (1u << sys[q].sub_level); ++k)
      tasks.push_back(make_uint2((uint32_t)q, k));
}

size_t sub_tasks_max(int64_t n, int nsys) {
  // a segment at sub_level holds more than SUB_MAX / 2 bodies unless its
  // system has none split device-wide
  return (size_t)(2 * n / SUB_MAX) + (size_t)nsys + 1;
}

void set_mid_levels(std::vector<SysDesc> &sys, std::vector<uint2> &tasks) {
  tasks.clear();
  for (size_t q = 0; q < sys.size(); ++q) {
    SysDesc &d = sys[q];
    uint32_t mx = d.n;
    int l = 0;
    while (mx > MID_MAX) {
      mx = (mx + 1) / 2;
      ++l;
    }
    if (sys.size() == 1) l = std::max(l, MID_MIN_LEVEL);
    d.mid_level = std::min((uint32_t)l, d.sub_level);
    if (d.mid_level < d.sub_level)
      for (uint32_t k = 0; k < (1u << d.mid_level); ++k) tasks.push_back(make_uint2((uint32_t)q, k));
  }
  // largest segments first, so that the longest workgroups start first
  std::stable_sort(tasks.begin(), tasks.end(), [&](const uint2 &a, const uint2 &b) {
    return (sys[a.x].n >> sys[a.x].mid_level) > (sys[b.x].n >> sys[b.x].mid_level);
  });
}

hipError_t launch_build(const Forest &f, int leaf, const StateDev &in, const StateDev &out,
                        const TreeDev &t, BuildBufs &b, double4 *pos64, float4 *pos32,
                        double theta, hipStream_t s, bool range_ready, int lazy, bool report,
                        bool mid) {
  const double *coord[3] = {in.x, in.y, in.z};
  const unsigned ny = (unsigned)f.nsys;
  hipError_t e;
  if (report) {
    e = hipMemsetAsync(b.viol, 0, sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
  }
  // a single system sorts its n bodies; a forest sorts every body position
  // by (system, coordinate), which leaves each system's bodies in coordinate
  // order at [list_off, list_off + n) (padding last)
  const uint32_t nsort = f.nsys == 1 ? f.max_n : f.total;
  if (sort_temp_bytes(nsort, f.nsys) > b.sort_tmp_bytes) return hipErrorInvalidValue;
  if (!range_ready) {
    k_range_init<<<1, 64, 0, s>>>(b.range);
    k_range<<<std::min<uint32_t>((nsort + 255) / 256, RANGE_BLOCKS), 256, 0, s>>>(
        in.x, in.y, in.z, nsort, b.range);
  }
  // the three coordinate sorts are independent: dimension 0 runs on s,
  // dimensions 1 and 2 on the side streams, forked after the range and
  // joined before the levels (a graph capture of s follows the fork)
  e = hipEventRecord(b.fork, s);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipStreamWaitEvent(b.side[i], b.fork, 0);
  if (e != hipSuccess) return e;
  for (int d = 0; d < 3; ++d) {
    const hipStream_t sd = d == 0 ? s : b.side[d - 1];
    if (d == lazy) continue;  // sorted per segment by the subtree builder
    uint32_t *key32 = b.keys_in[d];
    uint32_t *key32_out = b.keys_out[d];
    k_make_keys32<<<(nsort + 255) / 256, 256, 0, sd>>>(
        coord[d], b.range + 2 * d, key32, nsort, b.runs[d],
        f.nsys > 1 ? f.sys_of_block : nullptr, f.sys, bits_for((uint32_t)f.nsys));
    size_t bytes = b.sort_tmp_bytes;
    e = rocprim::radix_sort_pairs<SortCfg>(b.sort_tmp[d], bytes, key32, key32_out,
                                           rocprim::counting_iterator<uint32_t>(0u), b.L[d],
                                           nsort, 0, 32, sd);
    if (e != hipSuccess) return e;
    k_fix_runs<<<(nsort + 256 * FIX_PER - 1) / (256 * FIX_PER), 256, 0, sd>>>(
        key32_out, b.L[d], coord[d], nsort, b.range + 2 * d, b.runs[d]);
    k_fix_long<<<256, 256, 0, sd>>>(b.runs[d], key32_out, nsort, b.L[d], b.ids_in[d], coord[d]);
  }
  for (int i = 0; i < 2 && e == hipSuccess; ++i) {
    e = hipEventRecord(b.join[i], b.side[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, b.join[i], 0);
  }
  if (e != hipSuccess) return e;
  ListBufs LB;
  LB.A = Lists{{b.L[0], b.L[1], b.L[2]}};
  LB.B = Lists{{b.Lb[0], b.Lb[1], b.Lb[2]}};
  Coords P{{in.x, in.y, in.z}};
  // every system's deeper device-wide levels (segments of at most MID_MAX
  // bodies) run in one launch, one workgroup per segment; the kernel-by-
  // kernel levels before them read the table whose sub_level is mid_level
  const bool use_mid = mid && f.n_mid_tasks > 0;
  const SysDesc *lsys = use_mid ? f.sys_mid : f.sys;
  const int lev_end = use_mid ? f.max_mid_level : f.max_sub_level;
  for (int lev = 0; lev < lev_end; ++lev) {
    const uint32_t nseg = 1u << lev;
    const uint32_t T = level_tiles(f, lev);
    k_level_split<<<dim3(nseg, ny), SPLIT_THREADS, 0, s>>>(LB, P, t, lsys, in.oid, lev,
                                                           b.range, lazy,
                                                           report ? b.viol : nullptr);
    k_level_flags<<<dim3(nseg * T, ny), LVL_THREADS, 0, s>>>(LB, t, lsys, lev, T, b.flags);
    k_level_pack<<<(b.flag_words + 255) / 256, 256, 0, s>>>(b.flags, b.bits, b.flag_words);
    k_level_count<<<dim3(nseg * T, ny), LVL_THREADS, 0, s>>>(LB, t, lsys, lev, T, b.bits,
                                                             b.cnt, lazy);
    if (T > SCAN_SMALL_T)
      k_level_scan<1024><<<dim3(3 * nseg, ny), 1024, 0, s>>>(b.cnt, T, lsys, lev);
    else
      k_level_scan<64><<<dim3(3 * nseg, ny), 64, 0, s>>>(b.cnt, T, lsys, lev);
    k_level_scatter<<<dim3(nseg * T, ny), LVL_THREADS, 0, s>>>(LB, t, lsys, lev, T, b.bits,
                                                               b.cnt, lazy);
  }
  if (use_mid)
    k_level_mid<<<f.n_mid_tasks, MID_THREADS, 0, s>>>(LB, P, t, f.sys, f.mid_tasks, in.oid,
                                                      b.range, lazy, report ? b.viol : nullptr,
                                                      b.flags);
  if (report) {
    e = hipMemcpyAsync(b.viol_host, b.viol, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    // (a captured build is waited for by the stream, not the event)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (e == hipSuccess) e = hipStreamIsCapturing(s, &cs);
    if (e == hipSuccess && cs == hipStreamCaptureStatusNone) e = hipEventRecord(b.ev_viol, s);
    if (e != hipSuccess) return e;
  }
  SubArgs a;
  a.LB = LB;
  a.sys = f.sys;
  a.P = P;
  a.m = in.m;
  a.in = in;
  a.out = out;
  a.t = t;
  a.loc = b.loc;
  a.perm = b.perm;
  a.pos64 = pos64;
  a.pos32 = pos32;
  a.leaf = leaf;
  a.theta2 = theta * theta;
  a.lazy = lazy;
  a.tasks = f.nsys > 1 ? f.sub_tasks : nullptr;
  k_subtree<<<f.nsys > 1 ? f.n_sub_tasks : 1u << f.max_sub_level, SUB_THREADS, 0, s>>>(a);
  if (f.max_sub_level > 0) k_moments_top<<<ny, 1024, 0, s>>>(t, f.sys, theta * theta);
  return hipGetLastError();
}

}  // namespace kd
