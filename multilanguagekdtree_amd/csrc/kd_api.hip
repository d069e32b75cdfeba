// kd_api.hip -- C ABI of the gfx950 kD-tree Barnes-Hut step (kdtree_hip.h).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kdtree_hip.h"
#include "kd_internal.hpp"


using namespace kd;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                             \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(e_ == hipErrorOutOfMemory ? KD_ERR_NOMEM : KD_ERR_HIP,          \
                  "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                      \
  } while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

struct kd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  kd_params params{};
  int64_t cap = 0;
  int64_t n = 0;
  bool has_state = false;
  bool has_tree = false;
  Shape shape;
  int64_t slots_alloc = 0;

  StateDev st[2]{};
  int cur = 0;
  BuildBufs b{};
  TreeDev tree{};
  double4 *pos64 = nullptr;
  float4 *pos32 = nullptr;
  double *acc[3] = {nullptr, nullptr, nullptr};
  double *acc_o[3] = {nullptr, nullptr, nullptr};
  uint32_t *inv = nullptr;
  unsigned long long *counters = nullptr;
  int *bad = nullptr;
  std::vector<void *> allocs;
  std::vector<void *> node_allocs;
  // phase timing (params.profile): 4 events per step, drained every PROF_CHUNK steps
  static constexpr int PROF_CHUNK = 32;
  std::vector<hipEvent_t> ev;
  int ev_used = 0;  // steps recorded in the current chunk
  double t_ms[3] = {0.0, 0.0, 0.0};
  int64_t t_steps = 0;
  int traverse_variant = 0;  // bit 1: predicated leaf batches; bit 2: FP64 on the batched walk
  // systems: one for a plain context, many for a batch
  Forest forest;
  std::vector<SysDesc> hsys;
  SysDesc *dsys = nullptr;
  SysDesc *dsys_mid = nullptr;  // dsys with sub_level = mid_level
  uint2 *dmid = nullptr;        // (system, segment) of the workgroup levels
  size_t dmid_cap = 0;
  uint2 *dsub = nullptr;        // (system, segment) of the subtree builder (forests)
  size_t dsub_cap = 0;
  uint32_t *sob = nullptr;  // system of every 64 body positions
  uint32_t extra_recs = 0;  // walk records built beyond the precision's own
  // lazy dimension (one system, never captured): the build skips the sort and
  // the device-wide partitions of a dimension that a full build reported
  // could not be split there; a build that finds otherwise is repeated in
  // full, and lazy builds pause for LAZY_PAUSE builds
  static constexpr int LAZY_PAUSE = 16;
  bool lazy_allowed = false;  // single-system context
  bool lazy_on = true;        // kd_set_tuning("lazy_dimension", 0) turns it off
  // kd_step's walk applies the kick-drift to its own bodies (no acceleration
  // arrays, no kick-drift pass); kd_set_tuning("fused_kick_drift", 0) keeps
  // the separate pass.  wave_range: 6 words per wavefront of the walk.
  bool fused_kick = true;
  // kd_set_tuning("mid_levels", 0): every device-wide level of the build as
  // its own kernels (for comparison; the trees are the same)
  bool mid_levels = true;
  // the state is in the previous build's tree order (not an upload's order):
  // only then are a workgroup's bodies close in memory, which its levels
  // need (from an upload's order they take 2.1 ms at 10 M bodies, against
  // 0.27 ms from tree order)
  bool tree_order = false;
  int64_t mid_builds = 0;  // builds that ran workgroup levels
  unsigned long long *wave_range = nullptr;
  int lazy_dim = 3;           // dimension the next build skips (3: none)
  int lazy_pause = 0;
  int64_t lazy_builds = 0, lazy_redos = 0;
};

template <typename T>
static int dalloc(kd_ctx *c, std::vector<void *> &list, T **p, size_t count) {
  void *q = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess)
    return fail(KD_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  list.push_back(q);
  *p = (T *)q;
  return KD_OK;
}

#define ALLOC(list, ptr, count)                     \
  do {                                              \
    int rc_ = dalloc(c, list, &(ptr), (size_t)(count)); \
    if (rc_) return rc_;                            \
  } while (0)

static int check_params(const kd_params *p) {
  if (!p) return fail(KD_ERR_ARG, "params is NULL");
  if (!(p->theta >= 0.0) || !std::isfinite(p->theta))
    return fail(KD_ERR_ARG, "theta must be finite and >= 0 (got %g)", p->theta);
  if (!std::isfinite(p->dt)) return fail(KD_ERR_ARG, "dt must be finite");
  if (p->leaf_size < 2 || p->leaf_size > MAX_LEAF)
    return fail(KD_ERR_ARG, "leaf_size must be in [2, %d] (got %d)", MAX_LEAF, p->leaf_size);
  if (p->precision < KD_PREC_FP64 || p->precision > KD_PREC_MIXED_D64)
    return fail(KD_ERR_ARG, "unknown precision %d", p->precision);
  return KD_OK;
}

// The walks address a node record at the 32-bit byte offset h * (record
// size), one scalar instruction per visit cheaper than a 64-bit offset; with
// records of up to 64 bytes one tree may hold at most 2^26 - 1 heap slots
// (about leaf_size * 2^25 bodies: 235 M at leaf size 7, 1.07 G at 32).
constexpr int64_t MAX_TREE_SLOTS = ((int64_t)1 << 26) - 1;

static int check_tree_size(int64_t n, int leaf) {
  const Shape sh = make_shape(n, leaf);
  if (sh.slots > MAX_TREE_SLOTS)
    return fail(KD_ERR_CAPACITY,
                "a tree of %lld bodies at leaf_size %d has %lld heap slots, above the walk's "
                "limit of 2^26 - 1 (raise leaf_size, or split the system)",
                (long long)n, leaf, (long long)sh.slots);
  return KD_OK;
}

static int alloc_nodes(kd_ctx *c, int64_t s) {
  if (s > c->slots_alloc) {
    for (void *p : c->node_allocs) (void)hipFree(p);
    c->node_allocs.clear();
    c->slots_alloc = 0;
    TreeDev &t = c->tree;
    ALLOC(c->node_allocs, t.seg_start, s);
    ALLOC(c->node_allocs, t.seg_count, s);
    ALLOC(c->node_allocs, t.split_dim, s);
    ALLOC(c->node_allocs, t.split_val, s);
    ALLOC(c->node_allocs, t.size, s);
    ALLOC(c->node_allocs, t.mm, s);
    ALLOC(c->node_allocs, t.mx, s);
    ALLOC(c->node_allocs, t.my, s);
    ALLOC(c->node_allocs, t.mz, s);
    ALLOC(c->node_allocs, t.rec64, s);
    ALLOC(c->node_allocs, t.rec32, s);
    ALLOC(c->node_allocs, t.recd, s);
    ALLOC(c->node_allocs, t.lmask, s);
    c->slots_alloc = s;
  }
  return KD_OK;
}

// Copy the system table to the device and compute every tree's segments.
static int install_forest(kd_ctx *c) {
  const int nsys = (int)c->hsys.size();
  Forest &f = c->forest;
  f = Forest();
  f.nsys = nsys;
  f.sys = c->dsys;
  f.sys_mid = c->dsys_mid;
  f.mid_tasks = c->dmid;
  f.sys_of_block = c->sob;
  std::vector<uint2> tasks;
  set_mid_levels(c->hsys, tasks);
  if (tasks.size() > c->dmid_cap) return fail(KD_ERR_CAPACITY, "workgroup-level table too small");
  f.n_mid_tasks = (uint32_t)tasks.size();
  std::vector<uint2> sub;
  if (nsys > 1) {
    set_sub_tasks(c->hsys, sub);
    if (sub.size() > c->dsub_cap) return fail(KD_ERR_CAPACITY, "subtree task table too small");
  }
  f.sub_tasks = c->dsub;
  f.n_sub_tasks = (uint32_t)sub.size();
  std::vector<SysDesc> mid_sys = c->hsys;
  std::vector<uint32_t> sob;
  for (int i = 0; i < nsys; ++i) {
    const SysDesc &d = c->hsys[i];
    f.max_sub_level = std::max(f.max_sub_level, (int)d.sub_level);
    f.max_mid_level = std::max(f.max_mid_level, (int)d.mid_level);
    f.max_n = std::max(f.max_n, d.n);
    f.max_slots = std::max(f.max_slots, d.slots);
    f.total = std::max(f.total, d.body_off + d.n);
    sob.resize((d.body_off + d.n + 63) / 64, (uint32_t)i);
    mid_sys[i].sub_level = d.mid_level;
  }
  HIP_TRY(hipMemcpyAsync(c->dsys, c->hsys.data(), sizeof(SysDesc) * nsys, hipMemcpyHostToDevice,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(c->dsys_mid, mid_sys.data(), sizeof(SysDesc) * nsys,
                         hipMemcpyHostToDevice, c->stream));
  if (!tasks.empty())
    HIP_TRY(hipMemcpyAsync(c->dmid, tasks.data(), sizeof(uint2) * tasks.size(),
                           hipMemcpyHostToDevice, c->stream));
  if (!sub.empty())
    HIP_TRY(hipMemcpyAsync(c->dsub, sub.data(), sizeof(uint2) * sub.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->sob, sob.data(), 4 * sob.size(), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(launch_shape(c->tree, f, c->params.leaf_size, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));  // the host tables above are temporaries
  return KD_OK;
}

static int ensure_nodes(kd_ctx *c) {
  Shape sh = make_shape(c->n, c->params.leaf_size);
  c->shape = sh;
  int rc = alloc_nodes(c, sh.slots);
  if (rc) return rc;
  c->hsys.assign(1, make_sys(sh, 0, 0, 0));
  return install_forest(c);
}

extern "C" {

const char *kd_last_error(void) { return g_err.c_str(); }

const char *kd_version(void) { return "kdtree_hip 0.1 gfx950"; }

int kd_default_params(kd_params *out) {
  if (!out) return fail(KD_ERR_ARG, "out is NULL");
  memset(out, 0, sizeof(*out));
  out->theta = 0.3;
  out->dt = 1e-3;
  out->leaf_size = 7;
  out->precision = KD_PREC_FP64;
  out->count_interactions = 0;
  return KD_OK;
}

int64_t kd_nodes_needed(int64_t n, int32_t leaf, int32_t layout) {
  if (n < 0 || leaf < 2) return -1;
  int64_t half = leaf / 2;
  if (layout == KD_LAYOUT_PADDED) {
    if (n <= leaf) return 1;
    int64_t min_leaves = n / half;
    int p = 0;
    while (((int64_t)1 << p) < min_leaves) ++p;
    return 2 * ((int64_t)1 << p) - 1;
  }
  return 2 * (n / half + 1);
}

int kd_destroy(kd_ctx *c) {
  if (!c) return KD_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void *p : c->allocs) (void)hipFree(p);
  for (void *p : c->node_allocs) (void)hipFree(p);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  for (int i = 0; i < 2; ++i) {
    if (c->b.side[i]) (void)hipStreamDestroy(c->b.side[i]);
    if (c->b.join[i]) (void)hipEventDestroy(c->b.join[i]);
  }
  if (c->b.fork) (void)hipEventDestroy(c->b.fork);
  if (c->b.ev_viol) (void)hipEventDestroy(c->b.ev_viol);
  if (c->b.viol_host) (void)hipHostFree(c->b.viol_host);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return KD_OK;
}

// Buffer sizes beyond the body capacity: sort scratch, per-tile level
// counts, and the number of systems.
struct Sizing {
  size_t sort_bytes;
  size_t cnt_entries;
  int nsys;
};

static int create_impl(kd_ctx *c, int64_t cap, const Sizing &z) {
  HIP_TRY(hipGetDevice(&c->device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, c->device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(KD_ERR_NODEVICE, "device %d is %s; this library is built for gfx950",
                c->device, prop.gcnArchName);
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  size_t N = (size_t)cap;
  for (int i = 0; i < 2; ++i) {
    StateDev &s = c->st[i];
    ALLOC(c->allocs, s.x, N);
    ALLOC(c->allocs, s.y, N);
    ALLOC(c->allocs, s.z, N);
    ALLOC(c->allocs, s.vx, N);
    ALLOC(c->allocs, s.vy, N);
    ALLOC(c->allocs, s.vz, N);
    ALLOC(c->allocs, s.m, N);
    ALLOC(c->allocs, s.oid, N);
  }
  for (int d = 0; d < 3; ++d) {
    ALLOC(c->allocs, c->b.L[d], N);
    ALLOC(c->allocs, c->b.Lb[d], N);
    ALLOC(c->allocs, c->acc[d], N);
    ALLOC(c->allocs, c->acc_o[d], N);
  }
  ALLOC(c->allocs, c->wave_range, 6 * ((N + 63) / 64));
  c->b.sort_tmp_bytes = z.sort_bytes;
  for (int d = 0; d < 3; ++d) {
    ALLOC(c->allocs, c->b.keys_in[d], N);
    ALLOC(c->allocs, c->b.keys_out[d], N);
    ALLOC(c->allocs, c->b.ids_in[d], N);
    ALLOC(c->allocs, c->b.runs[d], 2 * (N / FIX_SHORT + 1) + 2);
    ALLOC(c->allocs, *(uint8_t **)&c->b.sort_tmp[d], c->b.sort_tmp_bytes);
  }
  for (int i = 0; i < 2; ++i) {
    HIP_TRY(hipStreamCreateWithFlags(&c->b.side[i], hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->b.join[i], hipEventDisableTiming));
  }
  HIP_TRY(hipEventCreateWithFlags(&c->b.fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&c->b.ev_viol, hipEventDisableTiming));
  ALLOC(c->allocs, c->b.viol, 1);
  HIP_TRY(hipHostMalloc((void **)&c->b.viol_host, sizeof(uint32_t), hipHostMallocDefault));
  *c->b.viol_host = 0;
  ALLOC(c->allocs, c->b.range, 8);
  c->b.flag_words = (uint32_t)((N + 31) / 32);
  ALLOC(c->allocs, c->b.flags, 32 * (size_t)c->b.flag_words);
  ALLOC(c->allocs, c->b.bits, c->b.flag_words);
  HIP_TRY(hipMemsetAsync(c->b.flags, 0, 32 * (size_t)c->b.flag_words, c->stream));
  ALLOC(c->allocs, c->b.loc, N);
  ALLOC(c->allocs, c->b.perm, N);
  ALLOC(c->allocs, c->inv, N);
  ALLOC(c->allocs, c->pos64, N + BODY_PAD);
  ALLOC(c->allocs, c->pos32, N + BODY_PAD);
  HIP_TRY(hipMemsetAsync(c->pos64, 0, sizeof(double4) * (N + BODY_PAD), c->stream));
  HIP_TRY(hipMemsetAsync(c->pos32, 0, sizeof(float4) * (N + BODY_PAD), c->stream));
  ALLOC(c->allocs, c->b.cnt, z.cnt_entries);
  ALLOC(c->allocs, c->dsys, z.nsys);
  ALLOC(c->allocs, c->dsys_mid, z.nsys);
  c->dmid_cap = mid_tasks_max(cap, z.nsys);
  ALLOC(c->allocs, c->dmid, c->dmid_cap);
  c->dsub_cap = z.nsys > 1 ? sub_tasks_max(cap, z.nsys) : 1;
  ALLOC(c->allocs, c->dsub, c->dsub_cap);
  ALLOC(c->allocs, c->sob, N / 64 + 1);
  ALLOC(c->allocs, c->counters, 8);
  ALLOC(c->allocs, c->bad, 4);
  return KD_OK;
}

int kd_create(int64_t cap_n, const kd_params *params, kd_ctx **out) {
  if (!out) return fail(KD_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (cap_n < 1 || cap_n > 0x7fffffffLL)
    return fail(KD_ERR_ARG, "capacity must be in [1, 2^31-1] (got %lld)", (long long)cap_n);
  kd_params p;
  if (params) {
    p = *params;
  } else {
    kd_default_params(&p);
  }
  int rc = check_params(&p);
  if (rc) return rc;
  if ((rc = check_tree_size(cap_n, p.leaf_size))) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(KD_ERR_NODEVICE, "no HIP device visible");
  kd_ctx *c = new kd_ctx();
  c->params = p;
  c->cap = cap_n;
  // a single system of up to cap_n bodies: at most 2*cap/SUB_MAX segments
  // reach the subtree builder, each device-wide level has at most
  // cap/LVL_TILE + 2^level tiles
  const size_t max_seg = 2 * ((size_t)cap_n / SUB_MAX) + 64;
  Sizing z{sort_temp_bytes((uint32_t)cap_n, 1),
           3 * ((size_t)cap_n / LVL_TILE + max_seg + 64), 1};
  rc = create_impl(c, cap_n, z);
  if (rc) {
    std::string msg = g_err;
    kd_destroy(c);
    g_err = msg;
    return rc;
  }
  c->lazy_allowed = true;
  *out = c;
  return KD_OK;
}

int kd_set_params(kd_ctx *c, const kd_params *p) {
  if (!c) return fail(KD_ERR_ARG, "ctx is NULL");
  int rc = check_params(p);
  if (rc) return rc;
  if ((rc = check_tree_size(c->cap, p->leaf_size))) return rc;
  bool leaf_changed = p->leaf_size != c->params.leaf_size;
  c->params = *p;
  if (leaf_changed) c->has_tree = false;
  if (leaf_changed && c->has_state) {
    HIP_TRY(hipSetDevice(c->device));
    rc = ensure_nodes(c);
    if (rc) return rc;
  }
  return KD_OK;
}

int kd_get_params(const kd_ctx *c, kd_params *out) {
  if (!c || !out) return fail(KD_ERR_ARG, "NULL argument");
  *out = c->params;
  return KD_OK;
}

int64_t kd_size(const kd_ctx *c) { return c ? c->n : -1; }

static int upload_common(kd_ctx *c, int64_t n) {
  StateDev &s = c->st[c->cur];
  HIP_TRY(launch_iota(s.oid, (uint32_t)n, c->stream));
  HIP_TRY(hipMemsetAsync(c->bad, 0, sizeof(int), c->stream));
  HIP_TRY(launch_sanitize(s, (uint32_t)n, c->bad, c->stream));
  int bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, c->bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (bad) {
    c->has_state = false;
    return fail(KD_ERR_NONFINITE, "non-finite position, velocity or mass in upload");
  }
  c->n = n;
  c->has_state = true;
  c->has_tree = false;
  c->tree_order = false;
  return ensure_nodes(c);
}

static int upload_impl(kd_ctx *c, int64_t n, const double *const src[7], hipMemcpyKind kind) {
  if (!c) return fail(KD_ERR_ARG, "ctx is NULL");
  for (int i = 0; i < 7; ++i)
    if (!src[i]) return fail(KD_ERR_ARG, "input array %d is NULL", i);
  if (n < 1) return fail(KD_ERR_ARG, "n must be >= 1 (got %lld)", (long long)n);
  if (n > c->cap)
    return fail(KD_ERR_CAPACITY, "n = %lld exceeds the context capacity %lld", (long long)n,
                (long long)c->cap);
  HIP_TRY(hipSetDevice(c->device));
  StateDev &s = c->st[c->cur];
  double *dst[7] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m};
  for (int i = 0; i < 7; ++i)
    HIP_TRY(hipMemcpyAsync(dst[i], src[i], sizeof(double) * (size_t)n, kind, c->stream));
  return upload_common(c, n);
}

int kd_upload(kd_ctx *c, int64_t n, const double *x, const double *y, const double *z,
              const double *vx, const double *vy, const double *vz, const double *m) {
  const double *src[7] = {x, y, z, vx, vy, vz, m};
  return upload_impl(c, n, src, hipMemcpyHostToDevice);
}

int kd_upload_device(kd_ctx *c, int64_t n, const double *x, const double *y, const double *z,
                     const double *vx, const double *vy, const double *vz, const double *m) {
  const double *src[7] = {x, y, z, vx, vy, vz, m};
  return upload_impl(c, n, src, hipMemcpyDeviceToDevice);
}

static int download_impl(kd_ctx *c, double *const out[6], hipMemcpyKind kind) {
  if (!c) return fail(KD_ERR_ARG, "ctx is NULL");
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  HIP_TRY(hipSetDevice(c->device));
  StateDev &s = c->st[c->cur];
  const double *src[6] = {s.x, s.y, s.z, s.vx, s.vy, s.vz};
  uint32_t n = (uint32_t)c->n;
  for (int i = 0; i < 6; i += 3) {
    HIP_TRY(launch_to_original(s.oid, src[i], src[i + 1], src[i + 2], c->acc_o[0], c->acc_o[1],
                               c->acc_o[2], n, c->stream));
    for (int k = 0; k < 3; ++k)
      if (out[i + k])
        HIP_TRY(hipMemcpyAsync(out[i + k], c->acc_o[k], sizeof(double) * n, kind, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return KD_OK;
}

int kd_download(kd_ctx *c, double *x, double *y, double *z, double *vx, double *vy, double *vz) {
  double *out[6] = {x, y, z, vx, vy, vz};
  return download_impl(c, out, hipMemcpyDeviceToHost);
}

int kd_download_device(kd_ctx *c, double *x, double *y, double *z, double *vx, double *vy,
                       double *vz) {
  double *out[6] = {x, y, z, vx, vy, vz};
  return download_impl(c, out, hipMemcpyDeviceToDevice);
}

// The walk records a build must write: the precision's own, plus those a
// reverse sweep reads.
static uint32_t records_for(const kd_ctx *c) {
  const int p = c->params.precision;
  const uint32_t own = p == KD_PREC_FP64 ? REC_64 : (p == KD_PREC_MIXED ? REC_32 : REC_D);
  return own | c->extra_recs;
}

// range_ready: the last kick-drift left the range of the current state in
// c->b.range (only within one kd_step call, where nothing else changes it).
static int build_impl(kd_ctx *c, bool range_ready = false) {
  int nxt = c->cur ^ 1;
  c->tree.recs = records_for(c);
  // a single-system context learns after each build (from the flags the
  // device-wide levels report) whether the next may skip a dimension; the
  // host waits for those flags while the GPU runs the subtree builder
  const bool report = c->lazy_allowed && c->lazy_on;
  const int lazy = report && c->lazy_pause == 0 ? c->lazy_dim : 3;
  const bool mid = c->mid_levels && c->tree_order;
  if (mid && c->forest.n_mid_tasks > 0) ++c->mid_builds;
  HIP_TRY(launch_build(c->forest, c->params.leaf_size, c->st[c->cur], c->st[nxt], c->tree, c->b,
                       c->pos64, c->pos32, c->params.theta, c->stream, range_ready, lazy,
                       report, mid));
  if (report) {
    HIP_TRY(hipEventSynchronize(c->b.ev_viol));
    uint32_t unsafe = *c->b.viol_host;
    bool full = lazy == 3;  // a full build reports every dimension
    if (lazy < 3) {
      ++c->lazy_builds;
      if (unsafe & (1u << lazy)) {
        // the skipped dimension could have been a device-wide split: the
        // tree is not exact, so build again from the same input with every
        // list (the range of `in` is in b.range by now)
        ++c->lazy_redos;
        HIP_TRY(launch_build(c->forest, c->params.leaf_size, c->st[c->cur], c->st[nxt], c->tree,
                             c->b, c->pos64, c->pos32, c->params.theta, c->stream, true, 3,
                             true, mid));
        HIP_TRY(hipEventSynchronize(c->b.ev_viol));
        unsafe = *c->b.viol_host;
        full = true;
        c->lazy_pause = kd_ctx::LAZY_PAUSE;
      }
    } else if (c->lazy_pause > 0) {
      --c->lazy_pause;
    }
    if (full) {
      // the next lazy dimension: the first of z, x, y this build found safe
      c->lazy_dim = 3;
      for (int d : {2, 0, 1})
        if (!(unsafe & (1u << d))) {
          c->lazy_dim = d;
          break;
        }
    }
  }
  c->cur = nxt;
  c->has_tree = true;
  c->tree_order = true;
  return KD_OK;
}

static int traverse_impl(kd_ctx *c) {
  bool count = c->params.count_interactions != 0;
  if (count) HIP_TRY(hipMemsetAsync(c->counters, 0, 8 * sizeof(unsigned long long), c->stream));
  HIP_TRY(launch_traverse(c->params.precision, count, c->traverse_variant, c->tree, c->pos64,
                          c->pos32,
                          (uint32_t)c->n, c->params.theta, c->params.leaf_size, c->acc[0],
                          c->acc[1], c->acc[2], c->counters, c->stream));
  return KD_OK;
}

int kd_build(kd_ctx *c) {
  if (!c) return fail(KD_ERR_ARG, "ctx is NULL");
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  HIP_TRY(hipSetDevice(c->device));
  int rc = build_impl(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return KD_OK;
}

static int accel_impl(kd_ctx *c, double *ax, double *ay, double *az, hipMemcpyKind kind) {
  if (!c) return fail(KD_ERR_ARG, "ctx is NULL");
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  HIP_TRY(hipSetDevice(c->device));
  int rc = build_impl(c);
  if (rc) return rc;
  rc = traverse_impl(c);
  if (rc) return rc;
  uint32_t n = (uint32_t)c->n;
  HIP_TRY(launch_to_original(c->st[c->cur].oid, c->acc[0], c->acc[1], c->acc[2], c->acc_o[0],
                             c->acc_o[1], c->acc_o[2], n, c->stream));
  double *out[3] = {ax, ay, az};
  for (int k = 0; k < 3; ++k)
    if (out[k])
      HIP_TRY(hipMemcpyAsync(out[k], c->acc_o[k], sizeof(double) * n, kind, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return KD_OK;
}

int kd_accel(kd_ctx *c, double *ax, double *ay, double *az) {
  return accel_impl(c, ax, ay, az, hipMemcpyDeviceToHost);
}

int kd_accel_device(kd_ctx *c, double *ax, double *ay, double *az) {
  return accel_impl(c, ax, ay, az, hipMemcpyDeviceToDevice);
}

// Fold the recorded event intervals of the current chunk into the sums.
static int drain_timings(kd_ctx *c) {
  if (c->ev_used == 0) return KD_OK;
  HIP_TRY(hipEventSynchronize(c->ev[4 * (c->ev_used - 1) + 3]));
  for (int s = 0; s < c->ev_used; ++s) {
    for (int p = 0; p < 3; ++p) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[4 * s + p], c->ev[4 * s + p + 1]));
      c->t_ms[p] += ms;
    }
  }
  c->t_steps += c->ev_used;
  c->ev_used = 0;
  return KD_OK;
}

static int record(kd_ctx *c, int phase) {
  if (!c->params.profile) return KD_OK;
  if (c->ev.empty()) {
    c->ev.resize(4 * kd_ctx::PROF_CHUNK);
    for (auto &e : c->ev) HIP_TRY(hipEventCreate(&e));
  }
  if (phase == 0 && c->ev_used == kd_ctx::PROF_CHUNK) {
    int rc = drain_timings(c);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(c->ev[4 * c->ev_used + phase], c->stream));
  if (phase == 3) ++c->ev_used;
  return KD_OK;
}

int kd_step(kd_ctx *c, int64_t steps) {
  if (!c) return fail(KD_ERR_ARG, "ctx is NULL");
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  if (steps < 0) return fail(KD_ERR_ARG, "steps must be >= 0");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemsetAsync(c->bad, 0, sizeof(int), c->stream));
  int rc;
  for (int64_t s = 0; s < steps; ++s) {
    if ((rc = record(c, 0))) return rc;
    // after the first step the previous kick-drift computed the range
    if ((rc = build_impl(c, s > 0))) return rc;
    if ((rc = record(c, 1))) return rc;
    const bool more = s + 1 < steps;
    if (c->fused_kick && !c->params.count_interactions) {
      // the walk advances each body as it finishes it; the next build's
      // range comes from the wavefronts' ranges
      const uint32_t n = (uint32_t)c->n;
      const KickDrift kd{c->st[c->cur], c->params.dt, c->bad, more ? c->wave_range : nullptr};
      HIP_TRY(launch_traverse_kick(c->params.precision, c->traverse_variant, c->tree, c->pos64,
                                   c->pos32, n, c->params.theta, c->params.leaf_size, kd,
                                   c->stream));
      if (more) HIP_TRY(launch_range_from_waves(c->wave_range, (n + 63) / 64, c->b.range, c->stream));
      if ((rc = record(c, 2))) return rc;
    } else {
      if ((rc = traverse_impl(c))) return rc;
      if ((rc = record(c, 2))) return rc;
      HIP_TRY(launch_kick_drift(c->st[c->cur], c->acc[0], c->acc[1], c->acc[2], (uint32_t)c->n,
                                c->params.dt, c->bad, c->stream,
                                more ? c->b.range : nullptr));
    }
    if ((rc = record(c, 3))) return rc;
  }
  if (c->params.profile && (rc = drain_timings(c))) return rc;
  int bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, c->bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (bad) return fail(KD_ERR_NONFINITE, "a body reached a non-finite position");
  return KD_OK;
}

int kd_sync(kd_ctx *c) {
  if (!c) return fail(KD_ERR_ARG, "ctx is NULL");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return KD_OK;
}

int kd_counters(kd_ctx *c, int64_t out[5]) {
  if (!c || !out) return fail(KD_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(c->device));
  unsigned long long h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 5; ++i) out[i] = (int64_t)h[i];
  return KD_OK;
}

int kd_set_tuning(kd_ctx *c, const char *key, int64_t value) {
  if (!c || !key) return fail(KD_ERR_ARG, "NULL argument");
  if (strcmp(key, "leaf_predicated") == 0) {
    c->traverse_variant = value ? (c->traverse_variant | 2) : (c->traverse_variant & ~2);
    return KD_OK;
  }
  if (strcmp(key, "fp64_batched_walk") == 0) {
    c->traverse_variant = value ? (c->traverse_variant | 4) : (c->traverse_variant & ~4);
    return KD_OK;
  }
  if (strcmp(key, "lazy_dimension") == 0) {
    c->lazy_on = value != 0;
    c->lazy_dim = 3;  // learnt again from the next full build
    c->lazy_pause = 0;
    return KD_OK;
  }
  if (strcmp(key, "fused_kick_drift") == 0) {
    c->fused_kick = value != 0;
    return KD_OK;
  }
  if (strcmp(key, "mid_levels") == 0) {
    c->mid_levels = value != 0;
    return KD_OK;
  }
  return fail(KD_ERR_ARG, "unknown tuning key '%s'", key);
}

int kd_get_tuning(kd_ctx *c, const char *key, int64_t *value) {
  if (!c || !key || !value) return fail(KD_ERR_ARG, "NULL argument");
  if (strcmp(key, "leaf_predicated") == 0) {
    *value = (c->traverse_variant & 2) ? 1 : 0;
  } else if (strcmp(key, "fp64_batched_walk") == 0) {
    *value = (c->traverse_variant & 4) ? 1 : 0;
  } else if (strcmp(key, "lazy_dimension") == 0) {
    *value = c->lazy_allowed && c->lazy_on ? 1 : 0;
  } else if (strcmp(key, "lazy_next") == 0) {
    *value = c->lazy_allowed && c->lazy_on && c->lazy_pause == 0 ? c->lazy_dim : 3;
  } else if (strcmp(key, "lazy_builds") == 0) {
    *value = c->lazy_builds;
  } else if (strcmp(key, "lazy_redos") == 0) {
    *value = c->lazy_redos;
  } else if (strcmp(key, "fused_kick_drift") == 0) {
    *value = c->fused_kick ? 1 : 0;
  } else if (strcmp(key, "mid_levels") == 0) {
    *value = c->mid_levels ? 1 : 0;
  } else if (strcmp(key, "mid_builds") == 0) {
    *value = c->mid_builds;
  } else {
    return fail(KD_ERR_ARG, "unknown tuning key '%s'", key);
  }
  return KD_OK;
}

int kd_timings(kd_ctx *c, double out[4]) {
  if (!c || !out) return fail(KD_ERR_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(c->device));
  int rc = drain_timings(c);
  if (rc) return rc;
  for (int p = 0; p < 3; ++p) out[p] = c->t_ms[p];
  out[3] = (double)c->t_steps;
  return KD_OK;
}

int kd_reset_timings(kd_ctx *c) {
  if (!c) return fail(KD_ERR_ARG, "ctx is NULL");
  HIP_TRY(hipSetDevice(c->device));
  int rc = drain_timings(c);
  if (rc) return rc;
  c->t_ms[0] = c->t_ms[1] = c->t_ms[2] = 0.0;
  c->t_steps = 0;
  return KD_OK;
}

int kd_direct_sum(kd_ctx *c, int64_t k, const int64_t *idx, double *ax, double *ay,
                  double *az, double *absx, double *absy, double *absz) {
  if (!c || !idx || !ax || !ay || !az) return fail(KD_ERR_ARG, "NULL argument");
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  if (k < 1 || k > c->n) return fail(KD_ERR_ARG, "k must be in [1, n]");
  for (int64_t i = 0; i < k; ++i)
    if (idx[i] < 0 || idx[i] >= c->n)
      return fail(KD_ERR_ARG, "sample index %lld out of range", (long long)idx[i]);
  HIP_TRY(hipSetDevice(c->device));
  uint32_t n = (uint32_t)c->n;
  StateDev &s = c->st[c->cur];
  HIP_TRY(launch_inverse(s.oid, c->inv, n, c->stream));
  std::vector<uint32_t> inv(n);
  HIP_TRY(hipMemcpyAsync(inv.data(), c->inv, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<uint32_t> wpos(k);
  for (int64_t i = 0; i < k; ++i) wpos[i] = inv[idx[i]];
  uint32_t *dw = nullptr;
  double *dout = nullptr;
  HIP_TRY(hipMalloc(&dw, sizeof(uint32_t) * k));
  hipError_t e = hipMalloc(&dout, sizeof(double) * 6 * k);
  if (e != hipSuccess) {
    (void)hipFree(dw);
    return fail(KD_ERR_NOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
  }
  std::vector<double> out(6 * k);
  e = hipMemcpyAsync(dw, wpos.data(), sizeof(uint32_t) * k, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = launch_direct(s, n, dw, (uint32_t)k, dout, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out.data(), dout, sizeof(double) * 6 * k, hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dw);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(KD_ERR_HIP, "direct sum failed: %s", hipGetErrorString(e));
  for (int64_t i = 0; i < k; ++i) {
    ax[i] = out[6 * i + 0];
    ay[i] = out[6 * i + 1];
    az[i] = out[6 * i + 2];
    if (absx) absx[i] = out[6 * i + 3];
    if (absy) absy[i] = out[6 * i + 4];
    if (absz) absz[i] = out[6 * i + 5];
  }
  return KD_OK;
}

// ---------------------------------------------------------------------------
// export to the reference node-vector layouts
// ---------------------------------------------------------------------------

struct HostTree {
  std::vector<uint32_t> start, count;
  std::vector<uint8_t> sd;
  std::vector<double> sv, size;
  std::vector<double> mm, mx, my, mz;  // mass and mass-weighted position sums
  int leaf;
};

// Compact pre-order layout: left = cur+1, right = last node of left + 1
// (Sequential/RustVersion/src/array_kd_tree.rs:128-134).  Returns the last
// node index used.
static int64_t emit_compact(const HostTree &t, uint32_t h, int64_t cur, kd_tree_view *v) {
  uint32_t cnt = t.count[h];
  if (cnt <= (uint32_t)t.leaf) {
    v->is_leaf[cur] = 1;
    v->num_parts[cur] = cnt;
    v->first[cur] = t.start[h];
    return cur;
  }
  int64_t left = emit_compact(t, 2 * h + 1, cur + 1, v);
  int64_t right = emit_compact(t, 2 * h + 2, left + 1, v);
  v->is_leaf[cur] = 0;
  v->split_dim[cur] = t.sd[h];
  v->split_val[cur] = t.sv[h];
  v->size[cur] = t.size[h];
  v->m[cur] = t.mm[h];
  v->cmx[cur] = t.mx[h] / t.mm[h];
  v->cmy[cur] = t.my[h] / t.mm[h];
  v->cmz[cur] = t.mz[h] / t.mm[h];
  v->left[cur] = cur + 1;
  v->right[cur] = left + 1;
  return right;
}

// Padded layout: right = cur + 1 + nodes_needed(floor(n/2))
// (Parallel/RustVersion/src/array_kd_tree.rs:565-581).  Returns false when a
// subtree overflows its slice (the reference panics there).
static bool emit_padded(const HostTree &t, uint32_t h, int64_t cur, int64_t limit,
                        kd_tree_view *v) {
  if (cur >= limit) return false;
  uint32_t cnt = t.count[h];
  if (cnt <= (uint32_t)t.leaf) {
    v->is_leaf[cur] = 1;
    v->num_parts[cur] = cnt;
    v->first[cur] = t.start[h];
    return true;
  }
  int64_t nn = kd_nodes_needed(cnt / 2, t.leaf, KD_LAYOUT_PADDED);
  if (nn > limit - cur - 1) return false;
  if (!emit_padded(t, 2 * h + 1, cur + 1, cur + 1 + nn, v)) return false;
  if (!emit_padded(t, 2 * h + 2, cur + 1 + nn, limit, v)) return false;
  v->is_leaf[cur] = 0;
  v->split_dim[cur] = t.sd[h];
  v->split_val[cur] = t.sv[h];
  v->size[cur] = t.size[h];
  v->m[cur] = t.mm[h];
  v->cmx[cur] = t.mx[h] / t.mm[h];
  v->cmy[cur] = t.my[h] / t.mm[h];
  v->cmz[cur] = t.mz[h] / t.mm[h];
  v->left[cur] = cur + 1;
  v->right[cur] = cur + 1 + nn;
  return true;
}

static int64_t compact_used(int64_t n, int leaf) {
  if (n <= leaf) return 1;
  return 1 + compact_used(n / 2, leaf) + compact_used(n - n / 2, leaf);
}

// Export the tree of system d (heap at d.node_off, bodies at d.body_off).
static int export_impl(kd_ctx *c, const SysDesc &d, int32_t layout, kd_tree_view *v) {
  if (!v) return fail(KD_ERR_ARG, "NULL argument");
  if (!c->has_tree) return fail(KD_ERR_STATE, "no tree built");
  if (layout != KD_LAYOUT_COMPACT && layout != KD_LAYOUT_PADDED)
    return fail(KD_ERR_ARG, "unknown layout %d", layout);
  const int leaf = c->params.leaf_size;
  const int64_t n = d.n;
  int64_t need;
  if (layout == KD_LAYOUT_PADDED) {
    need = kd_nodes_needed(n, leaf, layout);
  } else {
    need = std::max(kd_nodes_needed(n, leaf, layout), compact_used(n, leaf));
  }
  if (v->n_nodes < need) {
    v->n_nodes = need;
    return fail(KD_ERR_CAPACITY, "node arrays hold fewer than %lld nodes", (long long)need);
  }
  if (!v->is_leaf || !v->split_dim || !v->num_parts || !v->first || !v->split_val || !v->m ||
      !v->cmx || !v->cmy || !v->cmz || !v->size || !v->left || !v->right || !v->perm)
    return fail(KD_ERR_ARG, "a node array is NULL");
  HIP_TRY(hipSetDevice(c->device));
  const int64_t slots = d.slots;
  const TreeDev tr = tree_at(c->tree, d.node_off);
  HostTree t;
  t.leaf = leaf;
  t.start.resize(slots);
  t.count.resize(slots);
  t.sd.resize(slots);
  t.sv.resize(slots);
  t.size.resize(slots);
  t.mm.resize(slots);
  t.mx.resize(slots);
  t.my.resize(slots);
  t.mz.resize(slots);
  std::vector<uint32_t> oid(n);
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(t.start.data(), tr.seg_start, 4 * slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(t.count.data(), tr.seg_count, 4 * slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(t.sd.data(), tr.split_dim, slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(t.sv.data(), tr.split_val, 8 * slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(t.size.data(), tr.size, 8 * slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(t.mm.data(), tr.mm, 8 * slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(t.mx.data(), tr.mx, 8 * slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(t.my.data(), tr.my, 8 * slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(t.mz.data(), tr.mz, 8 * slots, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(oid.data(), c->st[c->cur].oid + d.body_off, 4 * n,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int64_t i = 0; i < need; ++i) {
    v->is_leaf[i] = 1;
    v->split_dim[i] = 0;
    v->num_parts[i] = 0;
    v->first[i] = 0;
    v->split_val[i] = v->m[i] = v->cmx[i] = v->cmy[i] = v->cmz[i] = v->size[i] = 0.0;
    v->left[i] = v->right[i] = 0;
  }
  for (int64_t i = 0; i < n; ++i) v->perm[i] = (int64_t)oid[i] - (int64_t)d.body_off;
  if (layout == KD_LAYOUT_PADDED) {
    if (!emit_padded(t, 0, 0, need, v))
      return fail(KD_ERR_CAPACITY, "padded layout overflows its %lld-node allocation",
                  (long long)need);
  } else {
    emit_compact(t, 0, 0, v);
  }
  v->n_nodes = need;
  return KD_OK;
}

int kd_export_tree(kd_ctx *c, int32_t layout, kd_tree_view *v) {
  if (!c) return fail(KD_ERR_ARG, "NULL argument");
  if (!c->has_tree) return fail(KD_ERR_STATE, "no tree built");
  return export_impl(c, c->hsys[0], layout, v);
}

// ---------------------------------------------------------------------------
// initial conditions (host)
// ---------------------------------------------------------------------------

static uint64_t sm64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unif(uint64_t &s) { return (double)(sm64(s) >> 11) * 0x1.0p-53; }

int kd_ic_ring(uint64_t seed, int64_t n, double *x, double *y, double *z, double *vx,
               double *vy, double *vz, double *m) {
  if (n < 1 || !x || !y || !z || !vx || !vy || !vz || !m)
    return fail(KD_ERR_ARG, "bad argument to kd_ic_ring");
  uint64_t s = seed;
  const double two_pi = 6.283185307179586;
  x[0] = y[0] = z[0] = vx[0] = vy[0] = vz[0] = 0.0;
  m[0] = 1.0;
  for (int64_t i = 1; i < n; ++i) {
    double r = 1.0 + unif(s);
    double phi = two_pi * unif(s);
    double u1 = unif(s), u2 = unif(s);
    double g = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(two_pi * u2);
    double v = std::sqrt(1.0 / r);
    double c = std::cos(phi), sn = std::sin(phi);
    x[i] = r * c;
    y[i] = r * sn;
    z[i] = 1e-3 * g;
    vx[i] = -v * sn;
    vy[i] = v * c;
    vz[i] = 0.0;
    m[i] = 1e-14;
  }
  return KD_OK;
}

int kd_ic_circular_orbits(uint64_t seed, int64_t n, double *x, double *y, double *z,
                          double *vx, double *vy, double *vz, double *m) {
  if (n < 1 || !x || !y || !z || !vx || !vy || !vz || !m)
    return fail(KD_ERR_ARG, "bad argument to kd_ic_circular_orbits");
  uint64_t s = seed;
  x[0] = y[0] = z[0] = vx[0] = vy[0] = vz[0] = 0.0;
  m[0] = 1.0;
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.1 + ((double)i * 5.0 / (double)n);
    double v = std::sqrt(1.0 / d);
    double theta = unif(s) * 6.28;
    x[i + 1] = d * std::cos(theta);
    y[i + 1] = d * std::sin(theta);
    z[i + 1] = 0.0;
    vx[i + 1] = -v * std::sin(theta);
    vy[i + 1] = v * std::cos(theta);
    vz[i + 1] = 0.0;
    m[i + 1] = 1e-14;
  }
  return KD_OK;
}

int kd_ic_two_bodies(double *x, double *y, double *z, double *vx, double *vy, double *vz,
                     double *m) {
  if (!x || !y || !z || !vx || !vy || !vz || !m)
    return fail(KD_ERR_ARG, "bad argument to kd_ic_two_bodies");
  x[0] = y[0] = z[0] = vx[0] = vy[0] = vz[0] = 0.0;
  m[0] = 1.0;
  x[1] = 1.0;
  y[1] = z[1] = 0.0;
  vx[1] = 0.0;
  vy[1] = 1.0;
  vz[1] = 0.0;
  m[1] = 1e-20;
  return KD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// traversal of a caller-supplied node vector
// ---------------------------------------------------------------------------

extern "C" int kd_accel_tree(int64_t n, const double *x, const double *y, const double *z,
                             const double *m, const kd_tree_view *tv, double theta, int64_t k,
                             const int64_t *idx, double *ax, double *ay, double *az) {
  if (n < 1 || !x || !y || !z || !m || !tv || !ax || !ay || !az)
    return fail(KD_ERR_ARG, "bad argument to kd_accel_tree");
  if (!idx) k = n;
  if (k < 1) return fail(KD_ERR_ARG, "k must be >= 1");
  if (!(theta >= 0.0) || !std::isfinite(theta)) return fail(KD_ERR_ARG, "bad theta");
  const int64_t nn = tv->n_nodes;
  if (nn < 1 || !tv->is_leaf || !tv->num_parts || !tv->first || !tv->m || !tv->cmx ||
      !tv->cmy || !tv->cmz || !tv->size || !tv->left || !tv->right || !tv->perm)
    return fail(KD_ERR_ARG, "incomplete tree view");
  // every index the kernel can follow must stay inside its array
  for (int64_t i = 0; i < n; ++i)
    if (tv->perm[i] < 0 || tv->perm[i] >= n)
      return fail(KD_ERR_ARG, "perm[%lld] out of range", (long long)i);
  for (int64_t i = 0; i < nn; ++i) {
    if (tv->is_leaf[i]) {
      if (tv->num_parts[i] < 0 || tv->first[i] < 0 || tv->first[i] + tv->num_parts[i] > n)
        return fail(KD_ERR_ARG, "leaf %lld members out of range", (long long)i);
    } else if (tv->left[i] < 0 || tv->left[i] >= nn || tv->right[i] < 0 || tv->right[i] >= nn) {
      return fail(KD_ERR_ARG, "node %lld children out of range", (long long)i);
    }
  }
  {
    // the walk visits every node it reaches at most once per wavefront only if
    // the node vector is a tree: reject shared subtrees and cycles, and trees
    // deeper than the per-wavefront stack
    std::vector<uint8_t> seen((size_t)nn, 0);
    std::vector<std::pair<int64_t, int>> todo{{0, 0}};
    while (!todo.empty()) {
      const auto [h, d] = todo.back();
      todo.pop_back();
      if (seen[h]) return fail(KD_ERR_ARG, "node %lld reached twice: not a tree", (long long)h);
      seen[h] = 1;
      if (d >= PW_STACK) return fail(KD_ERR_ARG, "tree deeper than %d levels", PW_STACK);
      if (!tv->is_leaf[h]) {
        todo.push_back({tv->right[h], d + 1});
        todo.push_back({tv->left[h], d + 1});
      }
    }
  }
  if (idx)
    for (int64_t i = 0; i < k; ++i)
      if (idx[i] < 0 || idx[i] >= n) return fail(KD_ERR_ARG, "idx[%lld] out of range", (long long)i);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(KD_ERR_NODEVICE, "no HIP device visible");

  std::vector<void *> bufs;
  auto release = [&]() {
    for (void *p : bufs) (void)hipFree(p);
  };
  auto dcopy = [&](const void *src, size_t bytes, void **dst) -> hipError_t {
    hipError_t e = hipMalloc(dst, bytes ? bytes : 8);
    if (e != hipSuccess) return e;
    bufs.push_back(*dst);
    return src ? hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
  };
  hipError_t e = hipSuccess;
  void *dx, *dy, *dz, *dm, *dil, *dnp, *dfi, *dnm, *dcx, *dcy, *dcz, *dsz, *dl, *dr, *dp, *didx = nullptr;
  void *oa, *ob, *oc, *derr;
  const size_t N8 = sizeof(double) * n, M8 = sizeof(double) * nn, MI = sizeof(int64_t) * nn;
  if ((e = dcopy(x, N8, &dx)) || (e = dcopy(y, N8, &dy)) || (e = dcopy(z, N8, &dz)) ||
      (e = dcopy(m, N8, &dm)) || (e = dcopy(tv->is_leaf, sizeof(int32_t) * nn, &dil)) ||
      (e = dcopy(tv->num_parts, MI, &dnp)) || (e = dcopy(tv->first, MI, &dfi)) ||
      (e = dcopy(tv->m, M8, &dnm)) || (e = dcopy(tv->cmx, M8, &dcx)) ||
      (e = dcopy(tv->cmy, M8, &dcy)) || (e = dcopy(tv->cmz, M8, &dcz)) ||
      (e = dcopy(tv->size, M8, &dsz)) || (e = dcopy(tv->left, MI, &dl)) ||
      (e = dcopy(tv->right, MI, &dr)) || (e = dcopy(tv->perm, sizeof(int64_t) * n, &dp)) ||
      (idx && (e = dcopy(idx, sizeof(int64_t) * k, &didx))) ||
      (e = dcopy(nullptr, sizeof(double) * k, &oa)) || (e = dcopy(nullptr, sizeof(double) * k, &ob)) ||
      (e = dcopy(nullptr, sizeof(double) * k, &oc)) || (e = dcopy(nullptr, sizeof(int), &derr))) {
    release();
    return fail(e == hipErrorOutOfMemory ? KD_ERR_NOMEM : KD_ERR_HIP, "kd_accel_tree: %s",
                hipGetErrorString(e));
  }
  int herr = 0;
  PtrTree t{(const int32_t *)dil, (const int64_t *)dnp, (const int64_t *)dfi,
            (const double *)dnm, (const double *)dcx, (const double *)dcy,
            (const double *)dcz, (const double *)dsz, (const int64_t *)dl,
            (const int64_t *)dr, (const int64_t *)dp, nn};
  e = hipMemset(derr, 0, sizeof(int));
  if (e == hipSuccess)
    e = launch_walk_ptr(t, (const double *)dx, (const double *)dy, (const double *)dz,
                        (const double *)dm, n, (const int64_t *)didx, k, theta, (double *)oa,
                        (double *)ob, (double *)oc, (int *)derr, nullptr);
  if (e == hipSuccess) e = hipMemcpy(ax, oa, sizeof(double) * k, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ay, ob, sizeof(double) * k, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(az, oc, sizeof(double) * k, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&herr, derr, sizeof(int), hipMemcpyDeviceToHost);
  release();
  if (e != hipSuccess) return fail(KD_ERR_HIP, "kd_accel_tree: %s", hipGetErrorString(e));
  if (herr) return fail(KD_ERR_ARG, "tree deeper than the traversal stack allows");
  return KD_OK;
}

// ---------------------------------------------------------------------------
// batches: many independent systems built and stepped together
// ---------------------------------------------------------------------------

struct kd_batch {
  kd_ctx *c = nullptr;              // buffers for every system (padded layout)
  std::vector<int64_t> n;           // bodies per system
  int64_t n_total = 0;              // sum of n (the caller's concatenated layout)
  uint2 *tasks = nullptr;           // (system, first body) per wavefront, largest systems first
  uint32_t ntasks = 0;
  uint32_t *queue = nullptr;        // task counters of the persistent walk, one per XCD
  int cus = 0;
  WalkShape ws;
  bool use_graph = true;
  // the walk applies the kick-drift to its bodies, and the next build takes
  // the coordinate range its wavefronts reduced (every step and upload
  // leaves a valid range in c->b.range)
  bool fused = true;
  // The lazy dimension, checked after the step: every step's build reports
  // which dimensions a build could have skipped (a full build) or whether
  // the skip was exact (a lazy one); a step whose skip was not exact is
  // taken again from the same state with a full build.
  bool lazy_on = true;
  int lazy_dim = 3;
  int lazy_pause = 0;
  int64_t lazy_builds = 0, lazy_redos = 0;
  // one step from state buffer 0 / 1, with lazy dimension 0, 1, 2 or none
  hipGraph_t graph[2][4] = {};
  hipGraphExec_t exec[2][4] = {};
};

static void drop_graphs(kd_batch *b) {
  for (int i = 0; i < 2; ++i) {
    for (int l = 0; l < 4; ++l) {
      if (b->exec[i][l]) (void)hipGraphExecDestroy(b->exec[i][l]);
      if (b->graph[i][l]) (void)hipGraphDestroy(b->graph[i][l]);
      b->exec[i][l] = nullptr;
      b->graph[i][l] = nullptr;
    }
  }
}

// Padded body position of caller index i of system s.
static void batch_to_padded(const kd_batch *b, const double *src, std::vector<double> &dst) {
  dst.assign(b->c->forest.total, 0.0);
  int64_t at = 0;
  for (size_t s = 0; s < b->n.size(); ++s) {
    const uint32_t off = b->c->hsys[s].body_off;
    std::memcpy(&dst[off], src + at, sizeof(double) * (size_t)b->n[s]);
    at += b->n[s];
  }
}

static void batch_from_padded(const kd_batch *b, const std::vector<double> &src, double *dst) {
  int64_t at = 0;
  for (size_t s = 0; s < b->n.size(); ++s) {
    const uint32_t off = b->c->hsys[s].body_off;
    std::memcpy(dst + at, &src[off], sizeof(double) * (size_t)b->n[s]);
    at += b->n[s];
  }
}

static int batch_traverse(kd_batch *b, const KickDrift *kick = nullptr) {
  kd_ctx *c = b->c;
  return launch_traverse_forest(c->params.precision, c->params.count_interactions != 0, b->ws,
                                c->tree, c->pos64, c->pos32, c->forest, b->tasks, b->ntasks,
                                b->queue, b->cus, c->params.theta, c->params.leaf_size,
                                c->acc[0], c->acc[1], c->acc[2], c->counters, c->stream,
                                kick) == hipSuccess
             ? KD_OK
             : fail(KD_ERR_HIP, "forest walk launch failed: %s",
                    hipGetErrorString(hipGetLastError()));
}

// One step (build, walk, kick-drift) enqueued on the context's stream.  The
// build takes the range the upload or the previous step left; the step
// leaves the range of its new positions.
static int batch_step_enqueue(kd_batch *b, int lazy) {
  kd_ctx *c = b->c;
  const int nxt = c->cur ^ 1;
  c->tree.recs = records_for(c);
  // the workgroup levels only from tree order (a step after an upload runs
  // eagerly with the kernel-by-kernel levels)
  HIP_TRY(launch_build(c->forest, c->params.leaf_size, c->st[c->cur], c->st[nxt], c->tree, c->b,
                       c->pos64, c->pos32, c->params.theta, c->stream, true, lazy, b->lazy_on,
                       c->mid_levels && c->tree_order));
  c->cur = nxt;
  c->has_tree = true;
  c->tree_order = true;
  int rc;
  if (b->fused && !c->params.count_interactions) {
    const KickDrift kd{c->st[c->cur], c->params.dt, c->bad, c->wave_range};
    if ((rc = batch_traverse(b, &kd))) return rc;
    HIP_TRY(launch_range_from_waves(c->wave_range, b->ntasks, c->b.range, c->stream));
    return KD_OK;
  }
  if ((rc = batch_traverse(b))) return rc;
  HIP_TRY(launch_kick_drift(c->st[c->cur], c->acc[0], c->acc[1], c->acc[2], c->forest.total,
                            c->params.dt, c->bad, c->stream, c->b.range));
  return KD_OK;
}

// Capture the step that starts from state buffer `from` into a graph.  The
// sort library records memset nodes; each becomes a fill kernel node, and a
// graph that still holds one is refused.
static int batch_capture(kd_batch *b, int from, int lazy) {
  kd_ctx *c = b->c;
  const int saved = c->cur;
  c->cur = from;
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  int rc = batch_step_enqueue(b, lazy);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->cur = saved;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(KD_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  int replaced = 0;
  e = graph_replace_memsets(g, &replaced);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(KD_ERR_HIP, "replacing memset nodes: %s", hipGetErrorString(e));
  }
  size_t nn = 0;
  e = hipGraphGetNodes(g, nullptr, &nn);
  std::vector<hipGraphNode_t> nodes(nn);
  if (e == hipSuccess && nn) e = hipGraphGetNodes(g, nodes.data(), &nn);
  for (size_t i = 0; e == hipSuccess && i < nn; ++i) {
    hipGraphNodeType ty;
    e = hipGraphNodeGetType(nodes[i], &ty);
    if (e == hipSuccess && ty == hipGraphNodeTypeMemset) {
      (void)hipGraphDestroy(g);
      return fail(KD_ERR_HIP, "the captured step holds a memset node");
    }
  }
  hipGraphExec_t x = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(KD_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  b->graph[from][lazy] = g;
  b->exec[from][lazy] = x;
  return KD_OK;
}

// One step from the current state buffer, as a graph replay (the graphs are
// captured from tree order) or eagerly.
static int batch_run_step(kd_batch *b, int lazy) {
  kd_ctx *c = b->c;
  if (!b->use_graph || !c->tree_order) return batch_step_enqueue(b, lazy);
  const int from = c->cur;
  int rc;
  if (!b->exec[from][lazy] && (rc = batch_capture(b, from, lazy))) return rc;
  HIP_TRY(hipGraphLaunch(b->exec[from][lazy], c->stream));
  c->cur ^= 1;
  c->has_tree = true;
  return KD_OK;
}

extern "C" int kd_batch_destroy(kd_batch *b) {
  if (!b) return KD_OK;
  if (b->c) {
    (void)hipSetDevice(b->c->device);
    if (b->c->stream) (void)hipStreamSynchronize(b->c->stream);
  }
  drop_graphs(b);
  if (b->tasks) (void)hipFree(b->tasks);
  if (b->queue) (void)hipFree(b->queue);
  if (b->c) kd_destroy(b->c);
  delete b;
  return KD_OK;
}

extern "C" int kd_batch_create(int64_t nsys, const int64_t *n, const kd_params *params,
                               kd_batch **out) {
  if (!out || !n) return fail(KD_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (nsys < 1 || nsys > 65535) return fail(KD_ERR_ARG, "nsys must be in [1, 65535]");
  kd_params p;
  if (params) p = *params;
  else kd_default_params(&p);
  int rc = check_params(&p);
  if (rc) return rc;
  // layout: systems at 64-aligned body offsets, heaps back to back
  std::vector<SysDesc> hs((size_t)nsys);
  uint64_t body = 0, slots = 0;
  int64_t total_n = 0;
  for (int64_t i = 0; i < nsys; ++i) {
    if (n[i] < 1 || n[i] > 0x7fffffffLL) return fail(KD_ERR_ARG, "n[%lld] out of range", (long long)i);
    if ((rc = check_tree_size(n[i], p.leaf_size))) return rc;
    const Shape sh = make_shape(n[i], p.leaf_size);
    hs[i] = make_sys(sh, (uint32_t)body, (uint32_t)slots, (uint32_t)total_n);
    body += ((uint64_t)n[i] + 63) & ~63ull;
    slots += (uint64_t)sh.slots;
    total_n += n[i];
    if (body > 0x7fff0000ull || slots > 0xffff0000ull)
      return fail(KD_ERR_CAPACITY, "batch too large for 32-bit body/node indices");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(KD_ERR_NODEVICE, "no HIP device visible");
  kd_batch *b = new kd_batch();
  b->n.assign(n, n + nsys);
  b->n_total = total_n;
  kd_ctx *c = new kd_ctx();
  b->c = c;
  c->params = p;
  c->cap = (int64_t)body;
  Forest f;
  f.nsys = (int)nsys;
  for (const SysDesc &d : hs) {
    f.max_sub_level = std::max(f.max_sub_level, (int)d.sub_level);
    f.max_n = std::max(f.max_n, d.n);
  }
  Sizing z{sort_temp_bytes((uint32_t)body, (int)nsys), level_cnt_entries(f), (int)nsys};
  auto bail = [&](int code) {
    std::string msg = g_err;
    kd_batch_destroy(b);
    g_err = msg;
    return code;
  };
  if ((rc = create_impl(c, (int64_t)body, z))) return bail(rc);
  if ((rc = alloc_nodes(c, (int64_t)slots))) return bail(rc);
  c->hsys = hs;
  if ((rc = install_forest(c))) return bail(rc);
  c->n = (int64_t)body;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
    fail(KD_ERR_HIP, "hipGetDeviceProperties failed");
    return bail(KD_ERR_HIP);
  }
  b->cus = prop.multiProcessorCount;
  // wavefront tasks, largest systems first (their walks are the longest)
  std::vector<int64_t> order((size_t)nsys);
  for (int64_t i = 0; i < nsys; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t q) { return n[a] > n[q]; });
  std::vector<uint2> tasks;
  for (int64_t i : order)
    for (int64_t w = 0; w < n[i]; w += 64) tasks.push_back(make_uint2((uint32_t)i, (uint32_t)w));
  b->ntasks = (uint32_t)tasks.size();
  if (hipMalloc(&b->tasks, sizeof(uint2) * tasks.size()) != hipSuccess ||
      hipMalloc(&b->queue, 8 * sizeof(uint32_t)) != hipSuccess) {
    fail(KD_ERR_NOMEM, "task table allocation failed");
    return bail(KD_ERR_NOMEM);
  }
  if (hipMemcpy(b->tasks, tasks.data(), sizeof(uint2) * tasks.size(), hipMemcpyHostToDevice) !=
      hipSuccess) {
    fail(KD_ERR_HIP, "task table upload failed");
    return bail(KD_ERR_HIP);
  }
  *out = b;
  return KD_OK;
}

extern "C" int64_t kd_batch_size(const kd_batch *b) { return b ? b->n_total : -1; }

extern "C" int kd_batch_upload(kd_batch *b, const double *x, const double *y, const double *z,
                               const double *vx, const double *vy, const double *vz,
                               const double *m) {
  if (!b) return fail(KD_ERR_ARG, "batch is NULL");
  const double *src[7] = {x, y, z, vx, vy, vz, m};
  for (int i = 0; i < 7; ++i)
    if (!src[i]) return fail(KD_ERR_ARG, "input array %d is NULL", i);
  kd_ctx *c = b->c;
  HIP_TRY(hipSetDevice(c->device));
  const uint32_t total = c->forest.total;
  // both state buffers: zero bodies in the padding, positions as original
  // indices (padding bodies never enter a tree and stay where they are)
  for (int k = 0; k < 2; ++k) {
    StateDev &s = c->st[k];
    double *dst[7] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m};
    for (int i = 0; i < 7; ++i)
      HIP_TRY(hipMemsetAsync(dst[i], 0, sizeof(double) * total, c->stream));
    HIP_TRY(launch_iota(s.oid, total, c->stream));
  }
  for (int d = 0; d < 3; ++d) HIP_TRY(hipMemsetAsync(c->acc[d], 0, sizeof(double) * total, c->stream));
  StateDev &s = c->st[c->cur];
  double *dst[7] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m};
  std::vector<double> stage;
  for (int i = 0; i < 7; ++i) {
    batch_to_padded(b, src[i], stage);
    HIP_TRY(hipMemcpyAsync(dst[i], stage.data(), sizeof(double) * total, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));  // stage is reused
  }
  HIP_TRY(hipMemsetAsync(c->bad, 0, sizeof(int), c->stream));
  HIP_TRY(launch_sanitize(s, total, c->bad, c->stream));
  int bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, c->bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (bad) {
    c->has_state = false;
    return fail(KD_ERR_NONFINITE, "non-finite position, velocity or mass in upload");
  }
  // the first step's build takes this range (every step leaves the next one)
  HIP_TRY(launch_range(s, total, c->b.range, c->stream));
  c->has_state = true;
  c->has_tree = false;
  c->tree_order = false;
  return KD_OK;
}

static int batch_fetch(kd_batch *b, const double *const src[3], double *const out[3]) {
  kd_ctx *c = b->c;
  const uint32_t total = c->forest.total;
  HIP_TRY(launch_to_original(c->st[c->cur].oid, src[0], src[1], src[2], c->acc_o[0], c->acc_o[1],
                             c->acc_o[2], total, c->stream));
  std::vector<double> stage(total);
  for (int k = 0; k < 3; ++k) {
    if (!out[k]) continue;
    HIP_TRY(hipMemcpyAsync(stage.data(), c->acc_o[k], sizeof(double) * total,
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    batch_from_padded(b, stage, out[k]);
  }
  return KD_OK;
}

extern "C" int kd_batch_download(kd_batch *b, double *x, double *y, double *z, double *vx,
                                 double *vy, double *vz) {
  if (!b) return fail(KD_ERR_ARG, "batch is NULL");
  kd_ctx *c = b->c;
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  HIP_TRY(hipSetDevice(c->device));
  StateDev &s = c->st[c->cur];
  const double *p[3] = {s.x, s.y, s.z}, *v[3] = {s.vx, s.vy, s.vz};
  double *po[3] = {x, y, z}, *vo[3] = {vx, vy, vz};
  int rc = batch_fetch(b, p, po);
  if (rc) return rc;
  return batch_fetch(b, v, vo);
}

extern "C" int kd_batch_accel(kd_batch *b, double *ax, double *ay, double *az) {
  if (!b) return fail(KD_ERR_ARG, "batch is NULL");
  kd_ctx *c = b->c;
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  HIP_TRY(hipSetDevice(c->device));
  int rc = build_impl(c);
  if (rc) return rc;
  if (c->params.count_interactions)
    HIP_TRY(hipMemsetAsync(c->counters, 0, 8 * sizeof(unsigned long long), c->stream));
  if ((rc = batch_traverse(b))) return rc;
  const double *a[3] = {c->acc[0], c->acc[1], c->acc[2]};
  double *o[3] = {ax, ay, az};
  return batch_fetch(b, a, o);
}

extern "C" int kd_batch_step(kd_batch *b, int64_t steps) {
  if (!b) return fail(KD_ERR_ARG, "batch is NULL");
  kd_ctx *c = b->c;
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  if (steps < 0) return fail(KD_ERR_ARG, "steps must be >= 0");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemsetAsync(c->bad, 0, sizeof(int), c->stream));
  if (c->params.count_interactions)
    HIP_TRY(hipMemsetAsync(c->counters, 0, 8 * sizeof(unsigned long long), c->stream));
  int rc;
  for (int64_t s = 0; s < steps; ++s) {
    const int from = c->cur;
    const int lazy = b->lazy_on && b->lazy_pause == 0 ? b->lazy_dim : 3;
    if ((rc = batch_run_step(b, lazy))) return rc;
    if (!b->lazy_on) continue;
    // the step's build report (its copy to the host is part of the step)
    HIP_TRY(hipStreamSynchronize(c->stream));
    uint32_t unsafe = *c->b.viol_host;
    bool full = lazy == 3;
    if (lazy < 3) {
      ++b->lazy_builds;
      if (unsafe & (1u << lazy)) {
        // the skipped dimension could have been split at a device-wide
        // level: take the step again from the same state (the step left the
        // next state's range, so the range of this one is reduced again)
        ++b->lazy_redos;
        c->cur = from;
        HIP_TRY(launch_range(c->st[from], c->forest.total, c->b.range, c->stream));
        if ((rc = batch_run_step(b, 3))) return rc;
        HIP_TRY(hipStreamSynchronize(c->stream));
        unsafe = *c->b.viol_host;
        full = true;
        b->lazy_pause = kd_ctx::LAZY_PAUSE;
      }
    } else if (b->lazy_pause > 0) {
      --b->lazy_pause;
    }
    if (full) {
      b->lazy_dim = 3;
      for (int d : {2, 0, 1})
        if (!(unsafe & (1u << d))) {
          b->lazy_dim = d;
          break;
        }
    }
  }
  int bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, c->bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (bad) return fail(KD_ERR_NONFINITE, "a body reached a non-finite position");
  return KD_OK;
}

extern "C" int kd_batch_export_tree(kd_batch *b, int64_t sys, int32_t layout, kd_tree_view *v) {
  if (!b) return fail(KD_ERR_ARG, "batch is NULL");
  if (sys < 0 || sys >= (int64_t)b->n.size()) return fail(KD_ERR_ARG, "system index out of range");
  if (!b->c->has_tree) return fail(KD_ERR_STATE, "no tree built");
  return export_impl(b->c, b->c->hsys[(size_t)sys], layout, v);
}

extern "C" int kd_batch_set_tuning(kd_batch *b, const char *key, int64_t value) {
  if (!b || !key) return fail(KD_ERR_ARG, "NULL argument");
  const std::string k = key;
  if (k == "graph") {
    b->use_graph = value != 0;
  } else if (k == "walk_threads") {
    if (value != 64 && value != 128 && value != 256 && value != 512)
      return fail(KD_ERR_ARG, "walk_threads must be 64, 128, 256 or 512");
    b->ws.threads = (int)value;
  } else if (k == "walk_waves_per_cu") {
    if (value < 0 || value > 64) return fail(KD_ERR_ARG, "walk_waves_per_cu must be in [0, 64]");
    b->ws.waves_per_cu = (int)value;
  } else if (k == "fused_kick_drift") {
    b->fused = value != 0;
  } else if (k == "lazy_dimension") {
    b->lazy_on = value != 0;
    b->lazy_dim = 3;  // learnt again from the next full build
    b->lazy_pause = 0;
  } else {
    return fail(KD_ERR_ARG, "unknown tuning key '%s'", key);
  }
  drop_graphs(b);  // the captured steps embed the launch shape
  return KD_OK;
}

extern "C" int kd_batch_get_tuning(kd_batch *b, const char *key, int64_t *value) {
  if (!b || !key || !value) return fail(KD_ERR_ARG, "NULL argument");
  const std::string k = key;
  if (k == "graph") *value = b->use_graph ? 1 : 0;
  else if (k == "walk_threads") *value = b->ws.threads;
  else if (k == "walk_waves_per_cu") *value = b->ws.waves_per_cu;
  else if (k == "fused_kick_drift") *value = b->fused ? 1 : 0;
  else if (k == "lazy_dimension") *value = b->lazy_on ? 1 : 0;
  else if (k == "lazy_next") *value = b->lazy_on && b->lazy_pause == 0 ? b->lazy_dim : 3;
  else if (k == "lazy_builds") *value = b->lazy_builds;
  else if (k == "lazy_redos") *value = b->lazy_redos;
  else return fail(KD_ERR_ARG, "unknown tuning key '%s'", key);
  return KD_OK;
}

extern "C" int kd_batch_autotune(kd_batch *b, int64_t reps, double out[3]) {
  if (!b) return fail(KD_ERR_ARG, "batch is NULL");
  kd_ctx *c = b->c;
  if (!c->has_state) return fail(KD_ERR_STATE, "no bodies uploaded");
  if (reps < 1) reps = 1;
  HIP_TRY(hipSetDevice(c->device));
  int rc = build_impl(c);  // the walk needs a tree; positions are unchanged
  if (rc) return rc;
  const int saved_count = c->params.count_interactions;
  c->params.count_interactions = 0;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const int threads[4] = {64, 128, 256, 512};
  const int wpc[5] = {0, 8, 16, 24, 32};
  WalkShape best = b->ws;
  float best_ms = 1e30f;
  for (int ti = 0; ti < 4 && rc == KD_OK; ++ti) {
    for (int wi = 0; wi < 5 && rc == KD_OK; ++wi) {
      b->ws.threads = threads[ti];
      b->ws.waves_per_cu = wpc[wi];
      if ((rc = batch_traverse(b))) break;  // warm-up
      float ms_min = 1e30f;
      for (int64_t r = 0; r < reps && rc == KD_OK; ++r) {
        (void)hipEventRecord(e0, c->stream);
        rc = batch_traverse(b);
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) rc = fail(KD_ERR_HIP, "event sync failed");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms_min = std::min(ms_min, ms);
      }
      if (rc == KD_OK && ms_min < best_ms) {
        best_ms = ms_min;
        best = b->ws;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->params.count_interactions = saved_count;
  b->ws = best;
  drop_graphs(b);
  if (rc) return rc;
  if (out) {
    out[0] = best.threads;
    out[1] = best.waves_per_cu;
    out[2] = best_ms;
  }
  return KD_OK;
}

extern "C" int kd_batch_counters(kd_batch *b, int64_t out[5]) {
  if (!b) return fail(KD_ERR_ARG, "batch is NULL");
  return kd_counters(b->c, out);
}

// ---------------------------------------------------------------------------
// differentiation: forward with checkpoints, reverse sweep with recompute
// ---------------------------------------------------------------------------

struct kd_adj {
  kd_ctx *c = nullptr;
  int64_t n = 0, steps = 0, every = 0, nck = 0;
  double *ck = nullptr;        // nck checkpoints x 7 arrays x n (working order)
  uint32_t *ck_oid = nullptr;  // nck x n
  // tape of one segment (every steps)
  double4 *t_pos64 = nullptr;
  float4 *t_pos32 = nullptr;
  uint32_t *t_oid = nullptr;
  NodeRec64 *t_rec64 = nullptr;
  NodeRec32 *t_rec32 = nullptr;
  int64_t slots = 0;
  double *lp[3] = {nullptr, nullptr, nullptr}, *lv[3] = {nullptr, nullptr, nullptr},
         *u[3] = {nullptr, nullptr, nullptr};
  double *lm = nullptr, *gnode = nullptr;
  std::vector<void *> allocs;
  bool forward_done = false;
};

extern "C" int kd_adj_destroy(kd_adj *a) {
  if (!a) return KD_OK;
  if (a->c) {
    (void)hipSetDevice(a->c->device);
    (void)hipStreamSynchronize(a->c->stream);
  }
  for (void *p : a->allocs) (void)hipFree(p);
  if (a->c) kd_destroy(a->c);
  delete a;
  return KD_OK;
}

template <typename T>
static int adj_alloc(kd_adj *a, T **p, size_t count) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, count ? count * sizeof(T) : 256);
  if (e != hipSuccess) return fail(KD_ERR_NOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
  a->allocs.push_back(q);
  *p = (T *)q;
  return KD_OK;
}

extern "C" int kd_adj_create(int64_t n, int64_t steps, int64_t every, const kd_params *params,
                             kd_adj **out) {
  if (!out) return fail(KD_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n < 1 || n > 0x7fffffffLL) return fail(KD_ERR_ARG, "n out of range");
  if (steps < 1 || every < 1) return fail(KD_ERR_ARG, "steps and checkpoint interval must be >= 1");
  if (params && params->precision == KD_PREC_MIXED_D64)
    return fail(KD_ERR_ARG, "the reverse sweep supports KD_PREC_FP64 and KD_PREC_MIXED");
  kd_adj *a = new kd_adj();
  int rc = kd_create(n, params, &a->c);
  if (rc) {
    std::string msg = g_err;
    delete a;
    g_err = msg;
    return rc;
  }
  a->c->extra_recs = REC_64 | REC_32;  // the reverse sweep reads both
  a->n = n;
  a->steps = steps;
  a->every = std::min(every, steps);
  a->nck = (steps + a->every - 1) / a->every;
  a->slots = make_shape(n, a->c->params.leaf_size).slots;
  const size_t N = (size_t)n, E = (size_t)a->every, S = (size_t)a->slots;
  auto bail = [&](int code) {
    std::string msg = g_err;
    kd_adj_destroy(a);
    g_err = msg;
    return code;
  };
  if ((rc = adj_alloc(a, &a->ck, (size_t)a->nck * 7 * N)) ||
      (rc = adj_alloc(a, &a->ck_oid, (size_t)a->nck * N)) ||
      (rc = adj_alloc(a, &a->t_pos64, E * N)) || (rc = adj_alloc(a, &a->t_pos32, E * N)) ||
      (rc = adj_alloc(a, &a->t_oid, E * N)) || (rc = adj_alloc(a, &a->t_rec64, E * S)) ||
      (rc = adj_alloc(a, &a->t_rec32, E * S)) || (rc = adj_alloc(a, &a->lm, N)) ||
      (rc = adj_alloc(a, &a->gnode, 4 * S)))
    return bail(rc);
  for (int d = 0; d < 3; ++d)
    if ((rc = adj_alloc(a, &a->lp[d], N)) || (rc = adj_alloc(a, &a->lv[d], N)) ||
        (rc = adj_alloc(a, &a->u[d], N)))
      return bail(rc);
  *out = a;
  return KD_OK;
}

static int adj_save(kd_adj *a, int64_t slot) {
  kd_ctx *c = a->c;
  const size_t N = (size_t)a->n;
  StateDev &s = c->st[c->cur];
  const double *src[7] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m};
  double *base = a->ck + (size_t)slot * 7 * N;
  for (int i = 0; i < 7; ++i)
    HIP_TRY(hipMemcpyAsync(base + i * N, src[i], 8 * N, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(a->ck_oid + (size_t)slot * N, s.oid, 4 * N, hipMemcpyDeviceToDevice,
                         c->stream));
  return KD_OK;
}

static int adj_restore(kd_adj *a, int64_t slot) {
  kd_ctx *c = a->c;
  const size_t N = (size_t)a->n;
  StateDev &s = c->st[c->cur];
  double *dst[7] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m};
  const double *base = a->ck + (size_t)slot * 7 * N;
  for (int i = 0; i < 7; ++i)
    HIP_TRY(hipMemcpyAsync(dst[i], base + i * N, 8 * N, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(s.oid, a->ck_oid + (size_t)slot * N, 4 * N, hipMemcpyDeviceToDevice,
                         c->stream));
  return KD_OK;
}

static int adj_step_forward(kd_adj *a) {
  kd_ctx *c = a->c;
  int rc = build_impl(c);
  if (rc) return rc;
  if ((rc = traverse_impl(c))) return rc;
  HIP_TRY(launch_kick_drift(c->st[c->cur], c->acc[0], c->acc[1], c->acc[2], (uint32_t)c->n,
                            c->params.dt, c->bad, c->stream));
  return KD_OK;
}

extern "C" int kd_adj_forward(kd_adj *a, const double *x0, const double *y0, const double *z0,
                              const double *vx0, const double *vy0, const double *vz0,
                              const double *m, double *x, double *y, double *z, double *vx,
                              double *vy, double *vz) {
  if (!a) return fail(KD_ERR_ARG, "adjoint context is NULL");
  // the checkpoints are overwritten from here on: until this run completes,
  // no reverse sweep may read them
  a->forward_done = false;
  kd_ctx *c = a->c;
  int rc = kd_upload_device(c, a->n, x0, y0, z0, vx0, vy0, vz0, m);
  if (rc) return rc;
  c->params.count_interactions = 0;
  HIP_TRY(hipMemsetAsync(c->bad, 0, sizeof(int), c->stream));
  for (int64_t k = 0; k < a->steps; ++k) {
    if (k % a->every == 0 && (rc = adj_save(a, k / a->every))) return rc;
    if ((rc = adj_step_forward(a))) return rc;
  }
  int bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, c->bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (bad) return fail(KD_ERR_NONFINITE, "a body reached a non-finite position");
  a->forward_done = true;
  return kd_download_device(c, x, y, z, vx, vy, vz);
}

extern "C" int kd_adj_backward(kd_adj *a, const double *gpx, const double *gpy, const double *gpz,
                               const double *gvx, const double *gvy, const double *gvz,
                               double *gx0, double *gy0, double *gz0, double *gvx0,
                               double *gvy0, double *gvz0, double *gm) {
  if (!a) return fail(KD_ERR_ARG, "adjoint context is NULL");
  if (!a->forward_done) return fail(KD_ERR_STATE, "kd_adj_forward has not run");
  kd_ctx *c = a->c;
  HIP_TRY(hipSetDevice(c->device));
  const size_t N = (size_t)a->n, S = (size_t)a->slots;
  const double *gin[6] = {gpx, gpy, gpz, gvx, gvy, gvz};
  double *acc[6] = {a->lp[0], a->lp[1], a->lp[2], a->lv[0], a->lv[1], a->lv[2]};
  for (int i = 0; i < 6; ++i) {
    if (gin[i])
      HIP_TRY(hipMemcpyAsync(acc[i], gin[i], 8 * N, hipMemcpyDeviceToDevice, c->stream));
    else
      HIP_TRY(hipMemsetAsync(acc[i], 0, 8 * N, c->stream));
  }
  HIP_TRY(hipMemsetAsync(a->lm, 0, 8 * N, c->stream));
  const bool mixed = c->params.precision == KD_PREC_MIXED;
  int rc;
  for (int64_t seg = a->nck - 1; seg >= 0; --seg) {
    const int64_t k0 = seg * a->every;
    const int64_t len = std::min(a->every, a->steps - k0);
    if ((rc = adj_restore(a, seg))) return rc;
    for (int64_t j = 0; j < len; ++j) {
      if ((rc = build_impl(c))) return rc;
      HIP_TRY(hipMemcpyAsync(a->t_pos64 + j * N, c->pos64, 32 * N, hipMemcpyDeviceToDevice,
                             c->stream));
      HIP_TRY(hipMemcpyAsync(a->t_pos32 + j * N, c->pos32, 16 * N, hipMemcpyDeviceToDevice,
                             c->stream));
      HIP_TRY(hipMemcpyAsync(a->t_oid + j * N, c->st[c->cur].oid, 4 * N,
                             hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(a->t_rec64 + j * S, c->tree.rec64, sizeof(NodeRec64) * S,
                             hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(a->t_rec32 + j * S, c->tree.rec32, sizeof(NodeRec32) * S,
                             hipMemcpyDeviceToDevice, c->stream));
      if (j + 1 < len) {
        if ((rc = traverse_impl(c))) return rc;
        HIP_TRY(launch_kick_drift(c->st[c->cur], c->acc[0], c->acc[1], c->acc[2],
                                  (uint32_t)c->n, c->params.dt, c->bad, c->stream));
      }
    }
    for (int64_t j = len - 1; j >= 0; --j) {
      AdjStep st;
      st.rec64 = a->t_rec64 + j * S;
      st.rec32 = a->t_rec32 + j * S;
      st.pos64 = a->t_pos64 + j * N;
      st.pos32 = a->t_pos32 + j * N;
      st.oid = a->t_oid + j * N;
      st.seg_count = c->tree.seg_count;
      for (int d = 0; d < 3; ++d) {
        st.lp[d] = a->lp[d];
        st.lv[d] = a->lv[d];
        st.u[d] = a->u[d];
      }
      st.lm = a->lm;
      st.gnode = a->gnode;
      st.n = (uint32_t)N;
      st.slots = (uint32_t)S;
      st.leaf = (uint32_t)c->params.leaf_size;
      st.theta = c->params.theta;
      st.dt = c->params.dt;
      st.mixed = mixed;
      HIP_TRY(launch_adj_step(st, c->stream));
    }
  }
  double *outs[7] = {gx0, gy0, gz0, gvx0, gvy0, gvz0, gm};
  const double *srcs[7] = {a->lp[0], a->lp[1], a->lp[2], a->lv[0], a->lv[1], a->lv[2], a->lm};
  for (int i = 0; i < 7; ++i)
    if (outs[i])
      HIP_TRY(hipMemcpyAsync(outs[i], srcs[i], 8 * N, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return KD_OK;
}

// ---------------------------------------------------------------------------
// tree dump (text interchange format)
// ---------------------------------------------------------------------------

// The shortest decimal that reads back as v, written positionally (no
// exponent), as Rust's `{}` prints an f64: 0.1 -> "0.1", 1.0 -> "1",
// 1e-7 -> "0.0000001", -0.0 -> "-0".
static std::string fmt_f64(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
  if (v == 0.0) return std::signbit(v) ? "-0" : "0";
  char buf[64];
  int p = 1;
  for (; p <= 17; ++p) {
    snprintf(buf, sizeof buf, "%.*e", p - 1, v);
    if (strtod(buf, nullptr) == v) break;
  }
  // buf = [-]d[.ddd]e[+-]xx
  std::string s = buf;
  const bool neg = s[0] == '-';
  if (neg) s.erase(0, 1);
  const size_t e = s.find('e');
  const int exp10 = atoi(s.c_str() + e + 1);
  std::string digits;
  for (size_t i = 0; i < e; ++i)
    if (s[i] != '.') digits += s[i];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out;
  if (exp10 >= 0) {
    const size_t intlen = (size_t)exp10 + 1;
    if (digits.size() <= intlen) {
      out = digits + std::string(intlen - digits.size(), '0');
    } else {
      out = digits.substr(0, intlen) + "." + digits.substr(intlen);
    }
  } else {
    out = "0." + std::string((size_t)(-exp10 - 1), '0') + digits;
  }
  return neg ? "-" + out : out;
}

// print_tree (Parallel/RustVersion/src/array_kd_tree.rs:666-692): the node
// count, then per node in vector order "I split_dim split_val left right" or
// "L k" followed by the k member positions "x y z".
extern "C" int kd_print_tree(const kd_tree_view *t, int64_t n, const double *x, const double *y,
                             const double *z, const char *path) {
  if (!t || !x || !y || !z || !path || n < 0) return fail(KD_ERR_ARG, "bad argument to kd_print_tree");
  if (!t->is_leaf || !t->num_parts || !t->first || !t->perm || !t->split_dim || !t->split_val ||
      !t->left || !t->right)
    return fail(KD_ERR_ARG, "incomplete tree view");
  FILE *f = fopen(path, "w");
  if (!f) return fail(KD_ERR_ARG, "cannot open %s for writing", path);
  std::string line;
  bool ok = fprintf(f, "%lld\n", (long long)t->n_nodes) > 0;
  for (int64_t i = 0; ok && i < t->n_nodes; ++i) {
    if (t->is_leaf[i]) {
      const int64_t k = t->num_parts[i], f0 = t->first[i];
      if (k < 0 || f0 < 0 || f0 + k > n) {
        fclose(f);
        return fail(KD_ERR_ARG, "leaf %lld members out of range", (long long)i);
      }
      ok = fprintf(f, "L %lld\n", (long long)k) > 0;
      for (int64_t q = 0; ok && q < k; ++q) {
        const int64_t p = t->perm[f0 + q];
        if (p < 0 || p >= n) {
          fclose(f);
          return fail(KD_ERR_ARG, "leaf member index out of range");
        }
        line = fmt_f64(x[p]) + " " + fmt_f64(y[p]) + " " + fmt_f64(z[p]) + "\n";
        ok = fputs(line.c_str(), f) >= 0;
      }
    } else {
      line = "I " + std::to_string(t->split_dim[i]) + " " + fmt_f64(t->split_val[i]) + " " +
             std::to_string(t->left[i]) + " " + std::to_string(t->right[i]) + "\n";
      ok = fputs(line.c_str(), f) >= 0;
    }
  }
  if (fclose(f) != 0) ok = false;
  if (!ok) return fail(KD_ERR_ARG, "write to %s failed", path);
  return KD_OK;
}

// ---------------------------------------------------------------------------
// standalone selection
// ---------------------------------------------------------------------------

extern "C" int64_t kd_quickstat_workspace_bytes(int64_t len) {
  return len < 1 ? -1 : (int64_t)select_workspace_bytes(len);
}

extern "C" int kd_quickstat_index_device(int64_t len, int64_t *indices, int64_t goal,
                                         const double *vals, int64_t nvals, void *workspace,
                                         int64_t workspace_bytes) {
  if (len < 1 || !indices || !vals || nvals < 1 || !workspace)
    return fail(KD_ERR_ARG, "bad argument to kd_quickstat_index_device");
  if (goal < 0 || goal >= len) return fail(KD_ERR_ARG, "goal must be in [0, len)");
  if (workspace_bytes < (int64_t)select_workspace_bytes(len))
    return fail(KD_ERR_CAPACITY, "workspace smaller than kd_quickstat_workspace_bytes(len)");
  hipError_t e = launch_select(indices, vals, len, nvals, goal, workspace, nullptr);
  if (e == hipErrorInvalidValue) return fail(KD_ERR_ARG, "an index is outside [0, nvals)");
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return fail(KD_ERR_HIP, "selection: %s", hipGetErrorString(e));
  return KD_OK;
}

// ---- selection without allocation ----
// kd_quickstat_index_device_noalloc keeps the selection's workspace (sort
// keys, the partitioned copy, per-block counts, the selection state) in this
// file-scope device array instead of a caller workspace.  Each device holds
// its own copy for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns on that device (the
// array's name is a host-side stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[KD_NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The address of `scratch` on the current device, looked up once per device.
static int scratch_on_device(void **out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
    return fail(KD_ERR_NODEVICE, "device %d is beyond the %d devices served", dev,
                SCRATCH_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(KD_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return KD_OK;
}

extern "C" int kd_quickstat_index_device_noalloc(int64_t len, int64_t *indices, int64_t goal,
                                                 const double *vals, int64_t nvals) {
  if (len < 1 || !indices || !vals || nvals < 1)
    return fail(KD_ERR_ARG, "bad argument to kd_quickstat_index_device_noalloc");
  if (goal < 0 || goal >= len) return fail(KD_ERR_ARG, "goal must be in [0, len)");
  if (len > KD_NOALLOC_SELECT_MAX_LEN)
    return fail(KD_ERR_CAPACITY,
                "len %lld is above KD_NOALLOC_SELECT_MAX_LEN (%lld), the most the fixed "
                "device workspace holds; use kd_quickstat_index_device with a workspace",
                (long long)len, (long long)KD_NOALLOC_SELECT_MAX_LEN);
  if (select_workspace_bytes(len) > KD_NOALLOC_SCRATCH_BYTES)
    return fail(KD_ERR_CAPACITY, "workspace of len %lld exceeds the fixed device workspace",
                (long long)len);
  void *ws = nullptr;
  int rc = scratch_on_device(&ws);
  if (rc) return rc;
  hipError_t e = launch_select(indices, vals, len, nvals, goal, ws, nullptr);
  if (e == hipErrorInvalidValue) return fail(KD_ERR_ARG, "an index is outside [0, nvals)");
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) return fail(KD_ERR_HIP, "selection: %s", hipGetErrorString(e));
  return KD_OK;
}

extern "C" int kd_quickstat_index(int64_t len, int64_t *indices, int64_t goal,
                                  const double *vals, int64_t nvals) {
  if (len < 1 || !indices || !vals || nvals < 1)
    return fail(KD_ERR_ARG, "bad argument to kd_quickstat_index");
  if (goal < 0 || goal >= len) return fail(KD_ERR_ARG, "goal must be in [0, len)");
  for (int64_t i = 0; i < len; ++i)
    if (indices[i] < 0 || indices[i] >= nvals)
      return fail(KD_ERR_ARG, "indices[%lld] is outside [0, nvals)", (long long)i);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(KD_ERR_NODEVICE, "no HIP device visible");
  const size_t wsb = select_workspace_bytes(len);
  void *di = nullptr, *dv = nullptr, *ws = nullptr;
  hipError_t e = hipMalloc(&di, 8 * (size_t)len);
  if (e == hipSuccess) e = hipMalloc(&dv, 8 * (size_t)nvals);
  if (e == hipSuccess) e = hipMalloc(&ws, wsb);
  if (e == hipSuccess) e = hipMemcpy(di, indices, 8 * (size_t)len, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dv, vals, 8 * (size_t)nvals, hipMemcpyHostToDevice);
  int rc = KD_OK;
  if (e == hipSuccess) {
    rc = kd_quickstat_index_device(len, (int64_t *)di, goal, (const double *)dv, nvals, ws,
                                   (int64_t)wsb);
    if (rc == KD_OK) e = hipMemcpy(indices, di, 8 * (size_t)len, hipMemcpyDeviceToHost);
  }
  if (di) (void)hipFree(di);
  if (dv) (void)hipFree(dv);
  if (ws) (void)hipFree(ws);
  if (rc) return rc;
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? KD_ERR_NOMEM : KD_ERR_HIP, "selection: %s",
                hipGetErrorString(e));
  return KD_OK;
}
