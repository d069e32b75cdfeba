// kd_internal.hpp -- shared layouts and host-side launch declarations of the
// gfx950 kD-tree Barnes-Hut step.  Device code lives in kd_build.hip,
// kd_traverse.hip and kd_misc.hip; the C ABI in kd_api.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace kd {

// Largest segment the one-workgroup subtree builder holds in LDS, and its
// launch shape (SUB_THREADS threads x SUB_ITEMS consecutive positions).
// 2048 (37 KB of LDS a workgroup): against 4096 (74 KB), one more device-wide
// level but more, smaller workgroups; at 10 M bodies a build takes 4.36 ms
// against 4.40, and a 512-system batch step 29.9 ms against 32.0.  Other
// shapes, at 10 M bodies: 4.06 ms for 512 x 4, against 4.17 for 256 x 4 and
// 4.46 for 512 x 2 (segments of 1024, one more device-wide level), 4.19 for
// 256 x 8 and 5.13 for 1024 x 2.
#ifndef KD_SUB_THREADS
#define KD_SUB_THREADS 512
#endif
#ifndef KD_SUB_ITEMS
#define KD_SUB_ITEMS 4
#endif
constexpr int SUB_THREADS = KD_SUB_THREADS;
constexpr int SUB_ITEMS = KD_SUB_ITEMS;
constexpr int SUB_MAX = SUB_THREADS * SUB_ITEMS;  // 2048

// Positions per block in the device-wide level kernels.  4 per thread: at
// 10 M bodies a build takes 4.45 ms, against 4.70 (8 per thread), 4.60 (2),
// 5.27 (1) and 5.13 (16); 128 or 512 threads of 4 were no faster.
constexpr int LVL_THREADS = 256;
constexpr int LVL_ITEMS = 4;
constexpr int LVL_TILE = LVL_THREADS * LVL_ITEMS;  // 1024

// Runs of equal sort keys up to this length are ordered by one thread; longer
// runs are queued (at most n / (FIX_SHORT + 1) of them) for a workgroup each.
constexpr uint32_t FIX_SHORT = 16;

constexpr int MAX_LEAF = 32;

// The block index a kernel should treat as its own so that each of the 8
// XCDs runs one contiguous range of the grid: the dispatcher hands blocks to
// the XCDs round robin (block b to XCD b mod 8), and neighbouring bodies
// share tree nodes and leaf bodies that each XCD caches in its own L2.  A
// bijection of [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, q = nb / 8u, r = nb % 8u;
  const uint32_t x = blockIdx.x % 8u, i = blockIdx.x / 8u;
  return x * q + (x < r ? x : r) + i;
}

// LDS writes of a wavefront made visible to its own lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Order-preserving 64-bit key of a double; -0.0 and +0.0 share one key.
__device__ __forceinline__ uint64_t ordkey(double v) {
  uint64_t b = (uint64_t)__double_as_longlong(v);
  if ((b << 1) == 0) b = 0;
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Coordinate range for the build's sort keys: range[2d], range[2d+1] hold the
// ordered-key minimum and maximum of dimension d.  A kernel that produces it
// reduces each block's share (wavefront shuffles, then LDS over the block's
// wavefronts, sh holding 6 x (blockDim.x / 64) entries) and only threads 0-5
// of a block update the six words; grids are capped at RANGE_BLOCKS blocks,
// so the atomics stay few.  Every thread of the block must call this.
constexpr uint32_t RANGE_BLOCKS = 1024;
__device__ __forceinline__ void range_commit(uint64_t lo[3], uint64_t hi[3], uint64_t (*sh)[4],
                                             unsigned long long *range) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t a = __shfl_xor(lo[d], o, 64), b = __shfl_xor(hi[d], o, 64);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      sh[2 * d][w] = lo[d];
      sh[2 * d + 1][w] = hi[d];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int v = threadIdx.x;
    uint64_t r = sh[v][0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
      const uint64_t c = sh[v][q];
      r = (v & 1) ? (c > r ? c : r) : (c < r ? c : r);
    }
    if (v & 1)
      atomicMax(&range[v], (unsigned long long)r);
    else
      atomicMin(&range[v], (unsigned long long)r);
  }
}

// Tree-ordered body arrays carry this many padding entries so that batched
// leaf loads (up to 8 bodies) never leave the allocation.
constexpr int BODY_PAD = 32;

// Node record read by the FP64 traversal: 64 bytes, one scalar load.
// Internal nodes: centre of mass, mass, size*size, and the particle range of
// the node.  Leaves: the particle range only (count <= leaf size).  skip is
// the 1-based heap index of the node that follows this subtree in depth-first
// order (0 after the last one).
struct alignas(64) NodeRec64 {
  double cx, cy, cz, m;
  double size2;
  uint32_t start, count;
  uint32_t skip, pad0;
  double pad1;
};

// Node record read by the mixed-precision traversal: 32 bytes.  size2 holds
// the opening threshold size^2 / theta^2 (accept iff size2 < d2).
struct alignas(32) NodeRec32 {
  float cx, cy, cz, m;
  float size2;
  uint32_t start, count;
  uint32_t skip;
};

// Node record for the mixed traversal with FP64 coordinate differences.
struct alignas(64) NodeRecD {
  double cx, cy, cz;
  float m, size2;
  uint32_t start, count;
  uint32_t skip, pad0;
  double pad1;
};

// Device-resident tree of one system in heap order: node h has children
// 2h+1 and 2h+2; level l holds heap slots [2^l - 1, 2^(l+1) - 1).
struct TreeDev {
  uint32_t *seg_start;  // first tree position of the node's bodies
  uint32_t *seg_count;  // number of bodies (0 for unused heap slots)
  uint8_t *split_dim;
  double *split_val;
  double *size;
  double *mm, *mx, *my, *mz;  // mass and mass-weighted position sums
  NodeRec64 *rec64;
  NodeRec32 *rec32;
  NodeRecD *recd;
  // build: bit d set iff the node's sorted list d sits in the second buffer
  uint8_t *lmask;
  // which walk records the build writes: REC_64 | REC_32 | REC_D
  uint32_t recs = 7;
};
constexpr uint32_t REC_64 = 1, REC_32 = 2, REC_D = 4;

// The same tree with every per-node array advanced by o slots.
__host__ __device__ inline TreeDev tree_at(const TreeDev &t, uint32_t o) {
  TreeDev r = t;
  r.seg_start += o;
  r.seg_count += o;
  r.split_dim += o;
  r.split_val += o;
  r.size += o;
  r.mm += o;
  r.mx += o;
  r.my += o;
  r.mz += o;
  r.rec64 += o;
  r.rec32 += o;
  r.recd += o;
  r.lmask += o;
  return r;
}

// One system of a forest: independent trees built and stepped together.
// Body arrays hold every system's bodies at [body_off, body_off + n) (body_off
// a multiple of 64, so no wavefront spans two systems); the sorted id lists
// hold them at [list_off, list_off + n) (systems back to back); node arrays
// hold its heap at [node_off, node_off + slots).  Node records store body
// positions relative to body_off.  A single system is a forest of one with
// every offset 0.
struct SysDesc {
  uint32_t body_off;
  uint32_t n;
  uint32_t node_off;
  uint32_t slots;
  uint32_t sub_level;  // levels split device-wide; the subtree builder takes the next
  uint32_t levels;
  uint32_t list_off;
  uint32_t mid_level;  // from this level on, the device-wide levels run in k_level_mid
};

struct Forest {
  int nsys = 0;
  const SysDesc *sys = nullptr;  // device copy, nsys entries
  // the same table with sub_level replaced by mid_level (what the kernel-
  // by-kernel levels read when the workgroup levels take the rest), the
  // workgroup levels' (system, segment) table and the deepest mid_level
  const SysDesc *sys_mid = nullptr;
  const uint2 *mid_tasks = nullptr;
  uint32_t n_mid_tasks = 0;
  int max_mid_level = 0;
  // (system, segment) of every subtree-builder workgroup of a forest
  const uint2 *sub_tasks = nullptr;
  uint32_t n_sub_tasks = 0;
  int max_sub_level = 0;
  uint32_t max_n = 0;
  uint32_t max_slots = 0;
  uint32_t total = 0;            // body positions, padding included
  const uint32_t *sys_of_block = nullptr;  // device: system of body positions [64i, 64i+64)
};

// Body state in "working order" (the previous step's tree order).
struct StateDev {
  double *x, *y, *z, *vx, *vy, *vz, *m;
  uint32_t *oid;  // original index of the body held at this position
};

struct Shape {
  int64_t n = 0;
  int leaf = 7;
  int levels = 0;       // deepest level + 1
  int sub_level = 0;    // level whose segments the subtree builder takes
  int64_t slots = 0;    // heap slots (2^levels - 1)
};

struct BuildBufs {
  uint32_t *L[3], *Lb[3];  // sorted id lists and their partition buffers
  // per dimension, so that the three coordinate sorts run concurrently:
  // dimension 0 on the build stream, 1 and 2 on the side streams
  uint32_t *keys_in[3], *keys_out[3];
  uint32_t *ids_in[3];  // scratch of the long-run fix-up (and a forest's system keys)
  uint32_t *runs[3];  // long runs of equal sort keys: count, then (start, length)
  void *sort_tmp[3];
  size_t sort_tmp_bytes;  // of each sort_tmp[d]
  hipStream_t side[2];    // default priority, like the build stream
  hipEvent_t fork, join[2];
  unsigned long long *range;  // ordered-key minimum and maximum of x, y, z
  uint8_t *flags;    // left/right flag per working id (capacity rounded up to 32)
  uint32_t *bits;    // the flags packed, one bit per working id
  uint32_t flag_words;  // 32-bit words of bits
  uint32_t *loc;     // working id -> local index inside a subtree block
  uint32_t *cnt;     // per-tile left counts of the device-wide levels
  uint32_t *perm;    // tree position -> working id
  // lazy dimension (single systems): bit d of *viol is set when dimension d
  // could be the split dimension of a device-wide segment, judged by its
  // global extent; viol_host (pinned) receives it after the levels, ev_viol
  // marks that copy
  uint32_t *viol = nullptr;
  uint32_t *viol_host = nullptr;
  hipEvent_t ev_viol = nullptr;
};

Shape make_shape(int64_t n, int leaf);
SysDesc make_sys(const Shape &sh, uint32_t body_off, uint32_t node_off, uint32_t list_off);
size_t sort_temp_bytes(uint32_t n, int nsys);
// per-tile count entries the device-wide levels of a forest need
size_t level_cnt_entries(const Forest &f);
hipError_t launch_shape(const TreeDev &t, const Forest &f, int leaf, hipStream_t s);
// range_ready: b.range already holds the range of `in` (computed by the
// kick-drift that produced it), so the build skips its own range pass.
// lazy (single systems, 0-2; 3 for none): that dimension is neither sorted
// nor partitioned at the device-wide levels, whose splits take its global
// extent as the bound of every segment's, and the subtree builder sorts it
// per segment; the tree is exact unless bit `lazy` of *b.viol is set.
// mid: a single system's device-wide levels from mid_level(f) on run as one
// launch of one workgroup per segment (mid_level(f) == f.max_sub_level: none).
// report: clear *b.viol before the levels, then copy it to b.viol_host and
// record b.ev_viol after them (a full build reports every dimension that a
// lazy build could not have skipped).
// Each system's mid_level (a single system's from level 8 on, so that its
// segments fill the GPU; a forest's from the first level whose segments hold
// at most MID_MAX bodies) and the (system, segment) table of the workgroup
// levels.
void set_mid_levels(std::vector<SysDesc> &sys, std::vector<uint2> &tasks);
size_t mid_tasks_max(int64_t n, int nsys);  // bound on the table's entries
// The (system, segment) table of a forest's subtree builder, and its bound.
void set_sub_tasks(const std::vector<SysDesc> &sys, std::vector<uint2> &tasks);
size_t sub_tasks_max(int64_t n, int nsys);
hipError_t launch_build(const Forest &f, int leaf, const StateDev &in, const StateDev &out,
                        const TreeDev &t, BuildBufs &b, double4 *pos64, float4 *pos32,
                        double theta, hipStream_t s, bool range_ready = false, int lazy = 3,
                        bool report = false, bool mid = true);
hipError_t launch_traverse(int precision, bool count, int variant, const TreeDev &t,
                           const double4 *pos64, const float4 *pos32, uint32_t n,
                           double theta, int leaf, double *ax, double *ay, double *az,
                           unsigned long long *counters, hipStream_t s);
// The kick-drift a walk applies to its own bodies as it finishes them (the
// step of kd_step): the tree-ordered state advanced in place (the walk reads
// only its own body records, never this state), dt, the non-finite flag, and
// when the next build wants the coordinate range, 6 words per wavefront that
// receive the ordered-key range of its updated positions (lo x, lo y, lo z,
// hi x, hi y, hi z).  The accelerations themselves are not stored.
struct KickDrift {
  StateDev s;
  double dt;
  int *bad;
  unsigned long long *wave_range;
};
hipError_t launch_traverse_kick(int precision, int variant, const TreeDev &t, const double4 *pos64,
                                const float4 *pos32, uint32_t n, double theta, int leaf,
                                const KickDrift &kd, hipStream_t s);
// The coordinate range (reset first) of the first n positions of a state.
hipError_t launch_range(const StateDev &s, uint32_t n, unsigned long long *range, hipStream_t st);
// The coordinate range (reset first) from nw wavefronts' ranges.
hipError_t launch_range_from_waves(const unsigned long long *wave_range, uint32_t nw,
                                   unsigned long long *range, hipStream_t s);
// Launch shape of the forest walk: threads per block (64..512) and resident
// wavefronts per CU for the persistent form (0: one wavefront per task).
struct WalkShape {
  int threads = 256;
  int waves_per_cu = 0;
};
hipError_t launch_traverse_forest(int precision, bool count, const WalkShape &ws,
                                  const TreeDev &t, const double4 *pos64, const float4 *pos32,
                                  const Forest &f, const uint2 *tasks, uint32_t ntasks,
                                  uint32_t *queue, int cus, double theta, int leaf, double *ax,
                                  double *ay, double *az, unsigned long long *counters,
                                  hipStream_t s, const KickDrift *kick = nullptr);
// One step of the reverse sweep: the step's recorded tree and tree-order
// bodies, and the adjoint arrays (original body order).
struct AdjStep {
  const NodeRec64 *rec64;
  const NodeRec32 *rec32;
  const double4 *pos64;
  const float4 *pos32;
  const uint32_t *oid;
  const uint32_t *seg_count;
  double *lp[3], *lv[3], *u[3];
  double *lm;
  double *gnode;  // 4 * slots: d/dc (x, y, z) and d/dM per node
  uint32_t n, slots, leaf;
  double theta, dt;
  bool mixed;
};
hipError_t launch_adj_step(const AdjStep &st, hipStream_t s);

// Standalone selection (quickstat_index): workspace size and launcher.
size_t select_workspace_bytes(int64_t n);
hipError_t launch_select(int64_t *idx, const double *vals, int64_t n, int64_t nvals,
                         int64_t goal, void *ws, hipStream_t s);

// Replace every memset node of a captured graph by a fill kernel node with
// the same dependencies (memset nodes are not replayed reliably here).
hipError_t graph_replace_memsets(hipGraph_t g, int *replaced);
// range (may be null): also computes the next build's coordinate range of
// the updated positions (reset first), so that build can skip its own pass.
hipError_t launch_kick_drift(const StateDev &s, const double *ax, const double *ay,
                             const double *az, uint32_t n, double dt, int *bad,
                             hipStream_t st, unsigned long long *range = nullptr);
hipError_t launch_range_reset(unsigned long long *range, hipStream_t s);
hipError_t launch_to_original(const uint32_t *oid, const double *a0, const double *a1,
                              const double *a2, double *o0, double *o1, double *o2,
                              uint32_t n, hipStream_t st);
hipError_t launch_inverse(const uint32_t *oid, uint32_t *inv, uint32_t n, hipStream_t st);
hipError_t launch_iota(uint32_t *v, uint32_t n, hipStream_t st);
hipError_t launch_sanitize(const StateDev &s, uint32_t n, int *bad, hipStream_t st);
hipError_t launch_direct(const StateDev &s, uint32_t n, const uint32_t *wpos, uint32_t k,
                         double *out, hipStream_t st);

constexpr int PW_STACK = 96;  // traversal stack entries per wavefront (bounds tree depth)
struct PtrTree {
  const int32_t *is_leaf;
  const int64_t *num_parts, *first;
  const double *m, *cmx, *cmy, *cmz, *size;
  const int64_t *left, *right;
  const int64_t *parts;  // leaf members: parts[first .. first+num_parts)
  int64_t n_nodes;
};
hipError_t launch_walk_ptr(const PtrTree &t, const double *x, const double *y, const double *z,
                           const double *m, int64_t n, const int64_t *idx, int64_t k,
                           double theta, double *ax, double *ay, double *az, int *err,
                           hipStream_t s);

}  // namespace kd
