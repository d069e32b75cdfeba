// kd_select.hip -- standalone selection on gfx950: quickstat_index
// (Parallel/RustVersion/src/quickstat.rs:9-34) for an index slice keyed by
// an FP64 value array.  The element of rank `goal` under the total order
// (value, index) is found by radix selection over the 64-bit ordered value
// keys (then over the indices, when values tie), and the slice is stably
// partitioned around it: [0, goal) before, the pivot at goal, the rest after
// -- the reference's postcondition (quickstat.rs:199-215).
//
// Passes per 8-bit digit: a histogram of the digit over the elements whose
// higher digits match the prefix chosen so far (LDS histograms, one global
// add per bin and block), then a one-block kernel that picks the digit
// holding the wanted rank.  All of it stays on the device.
#include <algorithm>

#include "kd_internal.hpp"

namespace kd {

__device__ __forceinline__ uint64_t ordkey_sel(double v) {
  uint64_t b = (uint64_t)__double_as_longlong(v);
  if ((b << 1) == 0) b = 0;  // -0.0 and +0.0 compare equal
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Selection state on the device: prefix and mask of the pivot key (and of
// its index once keys tie), and the rank still wanted inside the prefix.
struct SelState {
  uint64_t key_prefix, key_mask;
  uint64_t idx_prefix, idx_mask;
  uint64_t rank;
  uint64_t hist[256];
};

constexpr int SEL_THREADS = 256;
constexpr int SEL_ITEMS = 16;

__global__ __launch_bounds__(SEL_THREADS) void k_sel_keys(const int64_t *__restrict__ idx,
                                                          const double *__restrict__ vals,
                                                          uint64_t *__restrict__ keys,
                                                          int64_t n, SelState *st,
                                                          uint64_t goal) {
  const int64_t i = (int64_t)blockIdx.x * SEL_THREADS + threadIdx.x;
  if (i == 0) {
    st->key_prefix = st->key_mask = st->idx_prefix = st->idx_mask = 0;
    st->rank = goal;
  }
  if (i < 256) st->hist[i] = 0;
  if (i >= n) return;
  keys[i] = ordkey_sel(vals[idx[i]]);
}

// Histogram of digit `shift` of the key (on_index: of the index) over the
// elements matching the current prefix.
template <bool ON_INDEX>
__global__ __launch_bounds__(SEL_THREADS) void k_sel_hist(const uint64_t *__restrict__ keys,
                                                          const int64_t *__restrict__ idx,
                                                          int64_t n, SelState *st, int shift) {
  __shared__ unsigned int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t kp = st->key_prefix, km = st->key_mask;
  const uint64_t ip = st->idx_prefix, im = st->idx_mask;
  const int64_t base = (int64_t)blockIdx.x * SEL_THREADS * SEL_ITEMS + threadIdx.x;
#pragma unroll 4
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int64_t i = base + (int64_t)it * SEL_THREADS;
    if (i >= n) break;
    const uint64_t k = keys[i];
    if ((k & km) != kp) continue;
    if (ON_INDEX) {
      const uint64_t x = (uint64_t)idx[i];
      if (k != kp || (x & im) != ip) continue;
      atomicAdd(&h[(x >> shift) & 255u], 1u);
    } else {
      atomicAdd(&h[(k >> shift) & 255u], 1u);
    }
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd((unsigned long long *)&st->hist[threadIdx.x],
                                (unsigned long long)h[threadIdx.x]);
}

// Pick the digit that holds the wanted rank; clear the histogram.
template <bool ON_INDEX>
__global__ void k_sel_pick(SelState *st, int shift) {
  if (threadIdx.x == 0) {
    uint64_t rank = st->rank, cum = 0;
    int b = 0;
    for (; b < 255; ++b) {
      if (cum + st->hist[b] > rank) break;
      cum += st->hist[b];
    }
    st->rank = rank - cum;
    if (ON_INDEX) {
      st->idx_prefix |= (uint64_t)b << shift;
      st->idx_mask |= (uint64_t)255 << shift;
    } else {
      st->key_prefix |= (uint64_t)b << shift;
      st->key_mask |= (uint64_t)255 << shift;
    }
  }
  __syncthreads();
  st->hist[threadIdx.x] = 0;
}

// Counts of elements equal to the pivot key (decides whether index passes
// are needed) -- reuses hist[0].
__global__ __launch_bounds__(SEL_THREADS) void k_sel_count_eq(const uint64_t *__restrict__ keys,
                                                              int64_t n, SelState *st) {
  __shared__ unsigned int c;
  if (threadIdx.x == 0) c = 0;
  __syncthreads();
  const uint64_t kp = st->key_prefix;
  const int64_t base = (int64_t)blockIdx.x * SEL_THREADS * SEL_ITEMS + threadIdx.x;
  unsigned int local = 0;
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int64_t i = base + (int64_t)it * SEL_THREADS;
    if (i < n && keys[i] == kp) ++local;
  }
  if (local) atomicAdd(&c, local);
  __syncthreads();
  if (threadIdx.x == 0 && c) atomicAdd((unsigned long long *)&st->hist[0], (unsigned long long)c);
}

// 0: before the pivot, 1: the pivot, 2: after, under (key, index).
__device__ __forceinline__ int sel_class(uint64_t k, int64_t x, uint64_t pk, int64_t px) {
  if (k != pk) return k < pk ? 0 : 2;
  return x < px ? 0 : (x == px ? 1 : 2);
}

// Per-block counts of the "before" and "after" classes.
__global__ __launch_bounds__(SEL_THREADS) void k_sel_count(const uint64_t *__restrict__ keys,
                                                           const int64_t *__restrict__ idx,
                                                           int64_t n, const SelState *st,
                                                           unsigned long long *cnt_b,
                                                           unsigned long long *cnt_a) {
  __shared__ unsigned int cb, ca;
  if (threadIdx.x == 0) cb = ca = 0;
  __syncthreads();
  const uint64_t pk = st->key_prefix;
  const int64_t px = (int64_t)st->idx_prefix;
  const bool exact = st->idx_mask != 0;
  const int64_t base = (int64_t)blockIdx.x * SEL_THREADS * SEL_ITEMS;
  unsigned int lb = 0, la = 0;
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int64_t i = base + (int64_t)it * SEL_THREADS + threadIdx.x;
    if (i >= n) break;
    const uint64_t k = keys[i];
    const int cl = exact ? sel_class(k, idx[i], pk, px) : (k < pk ? 0 : (k == pk ? 1 : 2));
    lb += cl == 0 ? 1u : 0u;
    la += cl == 2 ? 1u : 0u;
  }
  if (lb) atomicAdd(&cb, lb);
  if (la) atomicAdd(&ca, la);
  __syncthreads();
  if (threadIdx.x == 0) {
    cnt_b[blockIdx.x] = cb;
    cnt_a[blockIdx.x] = ca;
  }
}

// Exclusive scan of nb per-block counts (one block).
__global__ __launch_bounds__(1024) void k_sel_scan(unsigned long long *cnt, int64_t nb) {
  __shared__ unsigned long long sh[1024];
  unsigned long long carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += 1024) {
    const int64_t i = b0 + threadIdx.x;
    const unsigned long long v = i < nb ? cnt[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const unsigned long long t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) cnt[i] = carry + sh[threadIdx.x] - v;
    carry += sh[1023];
    __syncthreads();
  }
}

// Stable partition: "before" elements to [0, goal) in order, the pivot to
// goal, "after" elements to (goal, n) in order.
__global__ __launch_bounds__(SEL_THREADS) void k_sel_scatter(const uint64_t *__restrict__ keys,
                                                             const int64_t *__restrict__ idx,
                                                             int64_t n, const SelState *st,
                                                             const unsigned long long *cnt_b,
                                                             const unsigned long long *cnt_a,
                                                             uint64_t goal,
                                                             int64_t *__restrict__ out) {
  __shared__ unsigned int wsum[SEL_THREADS / 64][2];
  const uint64_t pk = st->key_prefix;
  const int64_t px = (int64_t)st->idx_prefix;
  const bool exact = st->idx_mask != 0;
  const int64_t base = (int64_t)blockIdx.x * SEL_THREADS * SEL_ITEMS;
  unsigned long long off_b = cnt_b[blockIdx.x], off_a = goal + 1 + cnt_a[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int64_t i = base + (int64_t)it * SEL_THREADS + threadIdx.x;
    int cl = 3;
    int64_t x = 0;
    if (i < n) {
      x = idx[i];
      const uint64_t k = keys[i];
      cl = exact ? sel_class(k, x, pk, px) : (k < pk ? 0 : (k == pk ? 1 : 2));
    }
    const uint64_t mb = __builtin_amdgcn_ballot_w64(cl == 0);
    const uint64_t ma = __builtin_amdgcn_ballot_w64(cl == 2);
    if (lane == 0) {
      wsum[w][0] = (unsigned)__builtin_popcountll(mb);
      wsum[w][1] = (unsigned)__builtin_popcountll(ma);
    }
    __syncthreads();
    unsigned int pb = 0, pa = 0, tb = 0, ta = 0;
    for (int q = 0; q < SEL_THREADS / 64; ++q) {
      if (q < w) {
        pb += wsum[q][0];
        pa += wsum[q][1];
      }
      tb += wsum[q][0];
      ta += wsum[q][1];
    }
    if (cl == 0)
      out[off_b + pb + __builtin_popcountll(mb & below)] = x;
    else if (cl == 1)
      out[goal] = x;
    else if (cl == 2)
      out[off_a + pa + __builtin_popcountll(ma & below)] = x;
    off_b += tb;
    off_a += ta;
    __syncthreads();
  }
}

struct SelBufs {
  uint64_t *keys;
  int64_t *out;
  unsigned long long *cnt;  // 2 * blocks
  SelState *st;
};

size_t select_blocks(int64_t n) {
  return (size_t)((n + SEL_THREADS * SEL_ITEMS - 1) / (SEL_THREADS * SEL_ITEMS));
}

// Any index outside [0, nvals) sets *err.
__global__ void k_sel_check(const int64_t *__restrict__ idx, int64_t n, int64_t nvals,
                            int *err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (idx[i] < 0 || idx[i] >= nvals)) *err = 1;
}

// Workspace: keys (n), partitioned copy (n), two count arrays, state, flag.
size_t select_workspace_bytes(int64_t n) {
  const size_t nb = select_blocks(n);
  return 8 * (size_t)n + 8 * (size_t)n + 16 * nb + sizeof(SelState) + 256;
}

// Select rank `goal` of idx[0..n) keyed by vals[idx[i]] and partition idx in
// place.  Returns hipErrorInvalidValue if an index is outside [0, nvals).
hipError_t launch_select(int64_t *idx, const double *vals, int64_t n, int64_t nvals,
                         int64_t goal, void *ws, hipStream_t s) {
  const unsigned nbk = (unsigned)((n + SEL_THREADS - 1) / SEL_THREADS);
  const unsigned nb = (unsigned)select_blocks(n);
  char *w = static_cast<char *>(ws);
  uint64_t *keys = reinterpret_cast<uint64_t *>(w);
  int64_t *out = reinterpret_cast<int64_t *>(w + 8 * (size_t)n);
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(w + 16 * (size_t)n);
  SelState *st = reinterpret_cast<SelState *>(w + 16 * (size_t)n + 16 * (size_t)nb);
  int *err = reinterpret_cast<int *>(reinterpret_cast<char *>(st) + sizeof(SelState));
  hipError_t e;
  int herr = 0;
  if ((e = hipMemsetAsync(err, 0, sizeof(int), s)) != hipSuccess) return e;
  k_sel_check<<<nbk, SEL_THREADS, 0, s>>>(idx, n, nvals, err);
  if ((e = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (herr) return hipErrorInvalidValue;
  k_sel_keys<<<std::max(nbk, 1u), SEL_THREADS, 0, s>>>(idx, vals, keys, n, st, (uint64_t)goal);
  for (int shift = 56; shift >= 0; shift -= 8) {
    k_sel_hist<false><<<nb, SEL_THREADS, 0, s>>>(keys, idx, n, st, shift);
    k_sel_pick<false><<<1, 256, 0, s>>>(st, shift);
  }
  k_sel_count_eq<<<nb, SEL_THREADS, 0, s>>>(keys, n, st);
  unsigned long long eq = 0;
  if ((e = hipMemcpyAsync(&eq, &st->hist[0], 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(&st->hist[0], 0, 8, s)) != hipSuccess) return e;
  if (eq > 1) {
    // equal values: the pivot is the rank-th of them by index
    for (int shift = 56; shift >= 0; shift -= 8) {
      k_sel_hist<true><<<nb, SEL_THREADS, 0, s>>>(keys, idx, n, st, shift);
      k_sel_pick<true><<<1, 256, 0, s>>>(st, shift);
    }
  }
  k_sel_count<<<nb, SEL_THREADS, 0, s>>>(keys, idx, n, st, cnt, cnt + nb);
  k_sel_scan<<<1, 1024, 0, s>>>(cnt, nb);
  k_sel_scan<<<1, 1024, 0, s>>>(cnt + nb, nb);
  k_sel_scatter<<<nb, SEL_THREADS, 0, s>>>(keys, idx, n, st, cnt, cnt + nb, (uint64_t)goal, out);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return hipMemcpyAsync(idx, out, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
}

size_t select_state_bytes() { return sizeof(SelState); }

}  // namespace kd
