// Host runtime of the device BFS engine: graph residency in HBM, launch
// configuration, cooperative launch and timing.  C ABI in include/dfq_bfs.h.
//
// Reference host path replaced (design/shared-q/atomic-sa/our-sep/upper-bound/
// bfs.cu): managed allocations and prefetch (:186,194-197,273-277), occupancy
// query and grid = SMs x blocks (:255-271), cudaEvent timing around the
// cooperative launch (:279-288), cudaFree (:312-316).  Unlike the reference
// every HIP status is checked and returned.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "dfq_bfs.h"
#include "dfq_kernels.hpp"
#include "dfq_msbfs.hpp"
#include "host_common.hpp"

using namespace dfq;

namespace {

// Streaming copy used to calibrate traffic counters (16-byte vectors).
__global__ __launch_bounds__(256) void dfq_stream_copy_kernel(uint4 *__restrict__ dst,
                                                              const uint4 *__restrict__ src, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256)
    dst[i] = src[i];
}

hipError_t &last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

}  // namespace

#define HIP_TRY(expr)                 \
  do {                                \
    hipError_t e__ = (expr);          \
    if (e__ != hipSuccess) {          \
      last_hip_error() = e__;         \
      return DFQ_ERR_HIP;             \
    }                                 \
  } while (0)

struct dfq_graph {
  int device = -1;
  uint32_t launches = 0;  // number of the last plain launch (never 0; parity alternates)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t n_words = 0;
  bool symmetric = false;
  uint64_t max_deg = 0;  // largest out-degree
  int64_t *d_off = nullptr;
  uint32_t *d_col = nullptr;
  uint32_t *d_level = nullptr;   // outputs for the host-array entry point
  uint32_t *d_parent = nullptr;
  uint32_t *d_vis = nullptr;
  uint32_t *d_vis_snap = nullptr;
  uint32_t *d_fb = nullptr;
  uint32_t *d_iso = nullptr;     // isolated vertices (symmetric graphs only)
  uint32_t *d_col_bu = nullptr;  // rows by neighbour degree (symmetric graphs only)
  uint4 *d_head = nullptr;       // first four entries of every col_bu row (symmetric graphs only)
  uint2 *d_items = nullptr;
  uint32_t item_cap = 0;
  uint32_t *d_qseg[2] = {nullptr, nullptr};  // kSegments x n entries each
  uint64_t *d_prof = nullptr;
  uint32_t prof_levels = 0;
  // multi-source searches (allocated at the first one): seen, two frontier
  // mask arrays and the level planes; three queue buffers; per-search
  // counters, depths and sources
  uint64_t *d_ms = nullptr;
  uint32_t *d_msq = nullptr;
  uint64_t *d_mscnt = nullptr;
  uint64_t *d_msprof = nullptr;
  uint32_t msprof_levels = 0;
  uint2 *d_msitems = nullptr;  // work items of multi-source searches, kSegments x msitem_cap
  uint64_t msitem_cap = 0;
  int grid_ms = 0;
  // degree-ranked copy of the graph for multi-source searches (built at the
  // first one): offsets, columns as ranks (rows sorted), vertex -> rank
  int64_t *d_off_r = nullptr;
  uint32_t *d_col_r = nullptr;
  uint32_t *d_o2r = nullptr;
  uint32_t nz = 0;  // vertices with entries (ranks [0, nz))
  Control *d_ctl = nullptr;
  // back-to-back searches: events between launches, per-search results,
  // the sources of a one-launch sequence
  std::vector<hipEvent_t> evs;
  uint64_t *d_res = nullptr;
  uint32_t *d_src = nullptr;
  uint32_t res_cap = 0;
  dfq_config cfg{};
  int grid = 0;
  int grid_seq = 0;  // the sequence kernel's grid (its own occupancy)
  int grid_prof = 0; // the profiling kernel's grid (its own occupancy)
  int block = 0;
};

namespace {

int validate_config(const dfq_config &c) {
  if (c.block_size != 256 && c.block_size != 512 && c.block_size != 1024) return DFQ_ERR_INVALID_ARG;
  if (c.small_max > 64) return DFQ_ERR_INVALID_ARG;
  if (c.medium_max < c.small_max || c.large_max < c.medium_max) return DFQ_ERR_INVALID_ARG;
  if (c.direction_opt > 1) return DFQ_ERR_INVALID_ARG;
  if (c.direction_opt && (c.bu_enter_div == 0 || c.bu_exit_div == 0)) return DFQ_ERR_INVALID_ARG;
  if (c.launch_mode > DFQ_LAUNCH_PLAIN_ABANDON) return DFQ_ERR_INVALID_ARG;
  return DFQ_OK;
}

static_assert(DFQ_PROFILE_FIELDS == kProfFields, "header and kernel disagree on the profile row");
static_assert(DFQ_MULTI_MAX_SOURCES == kMsMaxSources, "header and kernel disagree on the multi-source width");
static_assert(DFQ_MULTI_PROFILE_FIELDS == kMsProfFields, "header and kernel disagree on the profile row");

// prof: the instantiation with the per-level counters compiled in (used by
// dfq_bfs_profile only; the search kernels carry none of that code).
// td: the top-down-only instantiation, for searches that cannot go
// bottom-up (directed graph or direction_opt 0).  low (with td, 1024-thread
// blocks): the graph's largest out-degree fits the small bin.
enum class Variant { kAny, kTd, kTdLow };

template <int B>
const void *kernel_at(bool prof, Variant v) {
  if constexpr (B == 1024)
    if (v == Variant::kTdLow)
      return prof ? (const void *)&dfq_bfs_kernel<B, true, true, true>
                  : (const void *)&dfq_bfs_kernel<B, false, true, true>;
  if (v != Variant::kAny)
    return prof ? (const void *)&dfq_bfs_kernel<B, true, true> : (const void *)&dfq_bfs_kernel<B, false, true>;
  return prof ? (const void *)&dfq_bfs_kernel<B, true> : (const void *)&dfq_bfs_kernel<B>;
}

template <int B>
const void *seq_kernel_at(Variant v) {
  if constexpr (B == 1024)
    if (v == Variant::kTdLow) return (const void *)&dfq_bfs_seq_kernel<B, true, true>;
  if (v != Variant::kAny) return (const void *)&dfq_bfs_seq_kernel<B, true>;
  return (const void *)&dfq_bfs_seq_kernel<B>;
}

const void *kernel_for(uint32_t block, bool prof = false, Variant v = Variant::kAny) {
  switch (block) {
    case 1024: return kernel_at<1024>(prof, v);
    case 512: return kernel_at<512>(prof, v);
    default: return kernel_at<256>(prof, v);
  }
}

const void *seq_kernel_for(uint32_t block, Variant v = Variant::kAny) {
  switch (block) {
    case 1024: return seq_kernel_at<1024>(v);
    case 512: return seq_kernel_at<512>(v);
    default: return seq_kernel_at<256>(v);
  }
}

// The instantiation a search with these arguments runs.
Variant variant_for(const dfq_graph *g, const BfsArgs &a) {
  if (a.bottom_up) return Variant::kAny;
  return g->max_deg <= a.small_max ? Variant::kTdLow : Variant::kTd;
}

// Co-resident workgroups per CU of a kernel, the smallest over its
// instantiations (the grid is shared by all).
template <class F>
hipError_t occupancy_all(int *per_cu, F kernel, int block) {
  int best = 1 << 30;
  for (Variant v : {Variant::kAny, Variant::kTd, Variant::kTdLow}) {
    int k = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, kernel(v), block, 0);
    if (e != hipSuccess) return e;
    best = std::min(best, k);
  }
  *per_cu = best;
  return hipSuccess;
}

// Grid = CUs x co-resident workgroups per CU (a cooperative launch needs
// every workgroup resident).
int configure(dfq_graph *g, const dfq_config &cfg) {
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, g->device));
  if (!prop.cooperativeLaunch) return DFQ_ERR_COOPERATIVE;
  int per_cu = 0;
  const uint32_t bs = cfg.block_size;
  HIP_TRY(occupancy_all(&per_cu, [bs](Variant v) { return kernel_for(bs, false, v); }, (int)bs));
  if (per_cu <= 0) return DFQ_ERR_COOPERATIVE;
  if (cfg.blocks_per_cu) per_cu = std::min(per_cu, (int)cfg.blocks_per_cu);
  int per_cu_seq = 0;
  HIP_TRY(occupancy_all(&per_cu_seq, [bs](Variant v) { return seq_kernel_for(bs, v); }, (int)bs));
  if (per_cu_seq <= 0) return DFQ_ERR_COOPERATIVE;
  per_cu_seq = std::min(per_cu_seq, per_cu);
  int per_cu_prof = 0;
  HIP_TRY(occupancy_all(&per_cu_prof, [bs](Variant v) { return kernel_for(bs, true, v); }, (int)bs));
  if (per_cu_prof <= 0) return DFQ_ERR_COOPERATIVE;
  per_cu_prof = std::min(per_cu_prof, per_cu);
  g->cfg = cfg;
  g->grid = prop.multiProcessorCount * per_cu;
  g->grid_seq = prop.multiProcessorCount * per_cu_seq;
  g->grid_prof = prop.multiProcessorCount * per_cu_prof;
  g->block = (int)cfg.block_size;
  return DFQ_OK;
}

void free_device(dfq_graph *g) {
  (void)hipFree(g->d_qseg[0]);
  (void)hipFree(g->d_qseg[1]);
  (void)hipFree(g->d_off);
  (void)hipFree(g->d_col);
  (void)hipFree(g->d_level);
  (void)hipFree(g->d_parent);
  (void)hipFree(g->d_vis);
  (void)hipFree(g->d_vis_snap);
  (void)hipFree(g->d_fb);
  (void)hipFree(g->d_iso);
  (void)hipFree(g->d_col_bu);
  (void)hipFree(g->d_head);
  (void)hipFree(g->d_items);
  (void)hipFree(g->d_prof);
  (void)hipFree(g->d_ms);
  (void)hipFree(g->d_msq);
  (void)hipFree(g->d_mscnt);
  (void)hipFree(g->d_msprof);
  (void)hipFree(g->d_msitems);
  (void)hipFree(g->d_off_r);
  (void)hipFree(g->d_col_r);
  (void)hipFree(g->d_o2r);
  (void)hipFree(g->d_ctl);
  (void)hipFree(g->d_res);
  (void)hipFree(g->d_src);
  for (hipEvent_t e : g->evs) (void)hipEventDestroy(e);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  if (g->stream) (void)hipStreamDestroy(g->stream);
}

BfsArgs make_args(dfq_graph *g, uint32_t source, uint32_t *d_level, uint32_t *d_parent, bool prof) {
  BfsArgs a;
  a.off = g->d_off;
  a.col = g->d_col;
  a.level = d_level;
  a.parent = d_parent;
  a.vis = g->d_vis;
  a.vis_snap = g->d_vis_snap;
  a.fb = g->d_fb;
  a.iso = g->d_iso;
  a.col_bu = g->d_col_bu;
  a.head = g->d_head;
  a.result = nullptr;
  a.qseg[0] = g->d_qseg[0];
  a.qseg[1] = g->d_qseg[1];
  a.items = g->d_items;
  a.item_cap = g->item_cap;
  a.ctl = g->d_ctl;
  a.prof = prof ? g->d_prof : nullptr;
  a.prof_levels = prof ? g->prof_levels : 0;
  a.n = g->n;
  a.n_words = g->n_words;
  a.source = source;
  a.small_max = g->cfg.small_max;
  a.deg_max = (uint32_t)std::min<uint64_t>(g->max_deg, 0xffffffffu);
  a.fuse_bits = fuse_bits_for(g->n, g->m);
  a.medium_max = g->cfg.medium_max;
  a.large_max = g->cfg.large_max;
  a.bottom_up = (g->cfg.direction_opt && g->symmetric) ? 1u : 0u;
  a.alpha = g->cfg.bu_enter_div;
  a.bu_exit = g->cfg.bu_exit_div ? g->n / g->cfg.bu_exit_div : 0;
  a.m = g->m;
  return a;
}

// Room for `count` searches' results and sources.
hipError_t reserve_results(dfq_graph *g, uint32_t count) {
  if (count <= g->res_cap) return hipSuccess;
  (void)hipFree(g->d_res);
  (void)hipFree(g->d_src);
  g->d_res = nullptr;
  g->d_src = nullptr;
  g->res_cap = 0;
  hipError_t e = hipMalloc(&g->d_res, sizeof(uint64_t) * kResultWords * count);
  if (e == hipSuccess) e = hipMalloc(&g->d_src, sizeof(uint32_t) * count);
  if (e != hipSuccess) return e;
  g->res_cap = count;
  return hipSuccess;
}

// Stats of one search from its kResultWords result words.
void fill_stats(const dfq_graph *g, const uint64_t *r, double ms, int grid, dfq_stats &st) {
  st.kernel_ms = ms;
  st.depth = (uint32_t)r[0];
  st.reached = r[1];
  st.edges = r[3];
  st.bottom_up_levels = (uint32_t)r[2];
  st.grid_blocks = (uint32_t)grid;
  st.block_size = (uint32_t)g->block;
}

// How long a workgroup of a plain launch waits for the rest of the grid
// before the launch is given up (a resident grid assembles in ~2 us).
constexpr uint64_t kPlainWaitNs = 1000000;

// Number of the next plain launch: never 0, and consecutive launches
// alternate in parity (the kernel's census words go by parity).
uint32_t next_launch(dfq_graph *g) {
  g->launches = g->launches == 0xffffffffu ? 2u : g->launches + 1u;
  return g->launches;
}

// Sets up `a` for a plain launch (numbered) or a cooperative one.
void launch_kind(dfq_graph *g, BfsArgs &a, bool plain) {
  a.plain = plain ? next_launch(g) : 0u;
  a.plain_wait_ns = g->cfg.launch_mode == DFQ_LAUNCH_PLAIN_ABANDON ? 0u : kPlainWaitNs;
}

// One launch of the single-search kernel on the graph's stream.  A plain
// launch costs ~1.5 us of dispatch between kernels where a cooperative
// one costs ~19 us on this stack; the kernel makes a plain launch safe by
// assembling its grid itself (and giving it up when it cannot).
hipError_t launch_search(dfq_graph *g, BfsArgs &a) {
  void *args[] = {&a};
  const bool prof = a.prof != nullptr;
  const void *k = kernel_for(g->cfg.block_size, prof, variant_for(g, a));
  const int grid = prof ? g->grid_prof : g->grid;
  if (a.plain) return hipLaunchKernel(k, dim3(grid), dim3(g->block), args, 0, g->stream);
  return hipLaunchCooperativeKernel(k, dim3(grid), dim3(g->block), args, 0, g->stream);
}

int run_bfs(dfq_graph *g, uint32_t source, uint32_t *d_level, uint32_t *d_parent, dfq_stats *stats,
            bool prof = false) {
  if (source >= g->n) return DFQ_ERR_RANGE;
  HIP_TRY(hipSetDevice(g->device));
  BfsArgs a = make_args(g, source, d_level, d_parent, prof);
  launch_kind(g, a, g->cfg.launch_mode != DFQ_LAUNCH_COOPERATIVE);
  HIP_TRY(hipEventRecord(g->ev0, g->stream));
  HIP_TRY(launch_search(g, a));
  HIP_TRY(hipEventRecord(g->ev1, g->stream));
  HIP_TRY(hipEventSynchronize(g->ev1));
  Control c;
  HIP_TRY(hipMemcpy(&c, g->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
  if (a.plain && c.abandoned == a.plain) {
    // the grid did not assemble (other work held CUs): run cooperatively
    if (prof) HIP_TRY(hipMemset(g->d_prof, 0, sizeof(uint64_t) * kProfFields * g->prof_levels));
    launch_kind(g, a, false);
    HIP_TRY(hipEventRecord(g->ev0, g->stream));
    HIP_TRY(launch_search(g, a));
    HIP_TRY(hipEventRecord(g->ev1, g->stream));
    HIP_TRY(hipEventSynchronize(g->ev1));
    HIP_TRY(hipMemcpy(&c, g->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
  }
  if (c.err) return DFQ_ERR_CAPACITY;
  if (stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    stats->kernel_ms = ms;
    stats->depth = c.depth;
    stats->reached = c.reached;
    stats->edges = c.edges;
    stats->grid_blocks = (uint32_t)(prof ? g->grid_prof : g->grid);
    stats->block_size = (uint32_t)g->block;
    stats->bottom_up_levels = c.bu_levels;
  }
  return DFQ_OK;
}

}  // namespace

extern "C" {

int dfq_config_default(dfq_config *cfg) {
  if (!cfg) return DFQ_ERR_INVALID_ARG;
  memset(cfg, 0, sizeof(*cfg));
  // measured on MI355X (RMAT-22, 64 sources): 1024-thread workgroups, one
  // per CU; rows above 256 entries cut into grid-wide work items
  cfg->block_size = 1024;
  cfg->blocks_per_cu = 0;
  cfg->small_max = 32;
  cfg->medium_max = 256;
  cfg->large_max = 256;
  cfg->direction_opt = 0;
  cfg->bu_enter_div = 40;
  // back to top-down below n / 96 frontier vertices: staying bottom-up for
  // the shrinking levels measured 1.5 us per search better than n / 24
  cfg->bu_exit_div = 96;
  return DFQ_OK;
}

int dfq_get_device_info(dfq_device_info *info) {
  if (!info) return DFQ_ERR_INVALID_ARG;
  memset(info, 0, sizeof(*info));
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DFQ_ERR_NO_DEVICE;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  info->device = dev;
  info->cu_count = prop.multiProcessorCount;
  info->cooperative = prop.cooperativeLaunch;
  info->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
  info->hbm_bytes = prop.totalGlobalMem;
  info->l2_bytes = (uint64_t)prop.l2CacheSize;
  int xcc = 0;
  if (hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess) xcc = 0;
  info->xcc_count = xcc;
  strncpy(info->arch, prop.gcnArchName, sizeof(info->arch) - 1);
  return DFQ_OK;
}

int dfq_graph_create(const int64_t *row_offsets, const uint32_t *col_idx, uint32_t n, uint64_t m,
                     dfq_graph **out) {
  if (!out || !row_offsets || (m && !col_idx) || n == 0 || n == 0xffffffffu) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  // Input checks on the host (the reference trusted its file blindly).
  if (m > kMaxEntries) return DFQ_ERR_RANGE;
  if (row_offsets[0] != 0 || (uint64_t)row_offsets[n] != m) return DFQ_ERR_RANGE;
  std::atomic<bool> bad{false};
  std::atomic<uint64_t> max_deg{0};
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    uint64_t mx = 0;
    for (uint64_t v = b; v < e; ++v) {
      if (row_offsets[v + 1] < row_offsets[v]) {
        bad = true;
        return;
      }
      mx = std::max(mx, (uint64_t)(row_offsets[v + 1] - row_offsets[v]));
    }
    uint64_t cur = max_deg.load();
    while (mx > cur && !max_deg.compare_exchange_weak(cur, mx)) {
    }
  });
  if (bad) return DFQ_ERR_RANGE;
  parallel_for(m, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t i = b; i < e; ++i)
      if (col_idx[i] >= n) {
        bad = true;
        return;
      }
  });
  if (bad) return DFQ_ERR_RANGE;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DFQ_ERR_NO_DEVICE;

  dfq_graph *g = new (std::nothrow) dfq_graph();
  if (!g) return DFQ_ERR_NOMEM;
  auto fail = [&](int code) {
    free_device(g);
    delete g;
    return code;
  };
#define HIP_OR_FAIL(expr)              \
  do {                                 \
    hipError_t e__ = (expr);           \
    if (e__ != hipSuccess) {           \
      last_hip_error() = e__;          \
      return fail(DFQ_ERR_HIP);        \
    }                                  \
  } while (0)
  HIP_OR_FAIL(hipGetDevice(&g->device));
  g->n = n;
  g->m = m;
  g->n_words = (n + 31) / 32;
  g->symmetric = is_symmetric(row_offsets, col_idx, n);
  g->max_deg = max_deg.load();
  HIP_OR_FAIL(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  HIP_OR_FAIL(hipEventCreate(&g->ev0));
  HIP_OR_FAIL(hipEventCreate(&g->ev1));
  HIP_OR_FAIL(hipMalloc(&g->d_off, sizeof(int64_t) * ((size_t)n + 1)));
  // +4 entries: bottom-up levels read columns as aligned 16-byte blocks
  HIP_OR_FAIL(hipMalloc(&g->d_col, sizeof(uint32_t) * (m + 4)));
  HIP_OR_FAIL(hipMalloc(&g->d_level, sizeof(uint32_t) * n));
  HIP_OR_FAIL(hipMalloc(&g->d_parent, sizeof(uint32_t) * n));
  HIP_OR_FAIL(hipMalloc(&g->d_vis, sizeof(uint32_t) * g->n_words));
  HIP_OR_FAIL(hipMalloc(&g->d_vis_snap, sizeof(uint32_t) * g->n_words));
  HIP_OR_FAIL(hipMalloc(&g->d_fb, sizeof(uint32_t) * 2 * ((size_t)g->n_words + 1)));
  // one level's items: sum over frontier vertices of ceil(deg / kItem)
  // <= m / kItem + n; any segment may receive all of them
  g->item_cap = (uint32_t)std::min<uint64_t>(m / kItem + n, 0xffffffffull);
  HIP_OR_FAIL(hipMalloc(&g->d_items, sizeof(uint2) * (size_t)kSegments * g->item_cap));
  for (int i = 0; i < 2; ++i)
    HIP_OR_FAIL(hipMalloc(&g->d_qseg[i], sizeof(uint32_t) * (size_t)kSegments * n));
  HIP_OR_FAIL(hipMalloc(&g->d_ctl, sizeof(Control)));
  HIP_OR_FAIL(hipMemcpy(g->d_off, row_offsets, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyHostToDevice));
  if (m) HIP_OR_FAIL(hipMemcpy(g->d_col, col_idx, sizeof(uint32_t) * m, hipMemcpyHostToDevice));
  if (g->symmetric) {
    // bottom-up levels read a second copy of the rows ordered by neighbour
    // degree (top-down levels keep the input order: degree order puts the
    // same hubs at the front of every row, and the probes of all waves then
    // land on the same bitmap lines at once); isolated vertices start visited
    std::vector<uint32_t> ordered;
    std::vector<uint32_t> iso(g->n_words);
    try {
      ordered.resize(m);
    } catch (...) {
      return fail(DFQ_ERR_NOMEM);
    }
    order_rows_by_degree(row_offsets, col_idx, n, ordered.data());
    isolated_bitmap(row_offsets, n, iso.data());
    HIP_OR_FAIL(hipMalloc(&g->d_col_bu, sizeof(uint32_t) * (m + 4)));
    if (m) HIP_OR_FAIL(hipMemcpy(g->d_col_bu, ordered.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice));
    {
      std::vector<uint32_t> heads;
      try {
        heads.resize(4 * (size_t)n);
      } catch (...) {
        return fail(DFQ_ERR_NOMEM);
      }
      row_heads(row_offsets, ordered.data(), n, heads.data());
      HIP_OR_FAIL(hipMalloc(&g->d_head, sizeof(uint4) * (size_t)n));
      HIP_OR_FAIL(hipMemcpy(g->d_head, heads.data(), sizeof(uint4) * (size_t)n, hipMemcpyHostToDevice));
    }
    HIP_OR_FAIL(hipMalloc(&g->d_iso, sizeof(uint32_t) * g->n_words));
    HIP_OR_FAIL(hipMemcpy(g->d_iso, iso.data(), sizeof(uint32_t) * g->n_words, hipMemcpyHostToDevice));
  }
  HIP_OR_FAIL(hipMemset(g->d_ctl, 0, sizeof(Control)));
#undef HIP_OR_FAIL
  dfq_config cfg;
  dfq_config_default(&cfg);
  int rc = configure(g, cfg);
  if (rc != DFQ_OK) return fail(rc);
  *out = g;
  return DFQ_OK;
}

void dfq_graph_destroy(dfq_graph *g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  free_device(g);
  delete g;
}

int dfq_graph_set_config(dfq_graph *g, const dfq_config *cfg) {
  if (!g || !cfg) return DFQ_ERR_INVALID_ARG;
  int rc = validate_config(*cfg);
  if (rc != DFQ_OK) return rc;
  HIP_TRY(hipSetDevice(g->device));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return configure(g, *cfg);
}

int dfq_graph_get_config(const dfq_graph *g, dfq_config *cfg) {
  if (!g || !cfg) return DFQ_ERR_INVALID_ARG;
  *cfg = g->cfg;
  return DFQ_OK;
}

int dfq_graph_is_symmetric(const dfq_graph *g) { return g && g->symmetric ? 1 : 0; }

int dfq_graph_device_arrays(const dfq_graph *g, const int64_t **d_row_offsets, const uint32_t **d_col_idx) {
  if (!g) return DFQ_ERR_INVALID_ARG;
  if (d_row_offsets) *d_row_offsets = g->d_off;
  if (d_col_idx) *d_col_idx = g->d_col;
  return DFQ_OK;
}

int dfq_bfs(dfq_graph *g, uint32_t source, uint32_t *level_out, uint32_t *parent_out, dfq_stats *stats) {
  if (!g || !level_out) return DFQ_ERR_INVALID_ARG;
  int rc = run_bfs(g, source, g->d_level, parent_out ? g->d_parent : nullptr, stats);
  if (rc != DFQ_OK) return rc;
  HIP_TRY(hipMemcpy(level_out, g->d_level, sizeof(uint32_t) * g->n, hipMemcpyDeviceToHost));
  if (parent_out) HIP_TRY(hipMemcpy(parent_out, g->d_parent, sizeof(uint32_t) * g->n, hipMemcpyDeviceToHost));
  return DFQ_OK;
}

int dfq_bfs_device(dfq_graph *g, uint32_t source, uint32_t *d_level, uint32_t *d_parent, dfq_stats *stats) {
  if (!g || !d_level) return DFQ_ERR_INVALID_ARG;
  return run_bfs(g, source, d_level, d_parent, stats);
}

int dfq_bfs_device_many(dfq_graph *g, uint32_t count, const uint32_t *sources, uint32_t *d_level,
                        uint32_t *d_parent, dfq_stats *stats) {
  if (!g || !d_level || (count && (!sources || !stats))) return DFQ_ERR_INVALID_ARG;
  for (uint32_t i = 0; i < count; ++i)
    if (sources[i] >= g->n) return DFQ_ERR_RANGE;
  if (count == 0) return DFQ_OK;
  HIP_TRY(hipSetDevice(g->device));
  while (g->evs.size() < (size_t)count + 1) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    g->evs.push_back(e);
  }
  HIP_TRY(reserve_results(g, count));
  std::vector<uint64_t> res((size_t)kResultWords * count);
  std::vector<uint32_t> num(count);
  bool plain = g->cfg.launch_mode != DFQ_LAUNCH_COOPERATIVE;
  for (;;) {
    // all launches are queued before the host waits, so each event pair
    // brackets one kernel and not the host's submission of it
    HIP_TRY(hipEventRecord(g->evs[0], g->stream));
    for (uint32_t i = 0; i < count; ++i) {
      BfsArgs a = make_args(g, sources[i], d_level, d_parent, false);
      a.result = g->d_res + (size_t)kResultWords * i;
      launch_kind(g, a, plain);
      num[i] = a.plain;
      HIP_TRY(launch_search(g, a));
      HIP_TRY(hipEventRecord(g->evs[i + 1], g->stream));
    }
    HIP_TRY(hipEventSynchronize(g->evs[count]));
    uint32_t err = 0;
    HIP_TRY(hipMemcpy(&err, &g->d_ctl->err, sizeof(err), hipMemcpyDeviceToHost));
    if (err) return DFQ_ERR_CAPACITY;
    HIP_TRY(hipMemcpy(res.data(), g->d_res, sizeof(uint64_t) * res.size(), hipMemcpyDeviceToHost));
    bool abandoned = false;
    for (uint32_t i = 0; i < count && plain; ++i) abandoned |= res[(size_t)kResultWords * i + 6] == num[i];
    if (!abandoned) break;
    // a grid did not assemble: the whole sequence again, cooperatively, so
    // the output still ends with the last source's search
    plain = false;
  }
  for (uint32_t i = 0; i < count; ++i) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g->evs[i], g->evs[i + 1]));
    fill_stats(g, &res[(size_t)kResultWords * i], ms, g->grid, stats[i]);
  }
  return DFQ_OK;
}

int dfq_bfs_device_seq(dfq_graph *g, uint32_t count, const uint32_t *sources, uint32_t *d_level,
                       uint64_t level_stride, uint32_t *d_parent, uint64_t parent_stride, dfq_stats *stats,
                       double *launch_ms) {
  if (!g || !d_level || (count && (!sources || !stats))) return DFQ_ERR_INVALID_ARG;
  if ((level_stride && level_stride < g->n) || (d_parent && parent_stride && parent_stride < g->n))
    return DFQ_ERR_INVALID_ARG;
  // element offsets of the last search must fit the kernel's 64-bit index
  if (count && (level_stride > UINT64_MAX / count || parent_stride > UINT64_MAX / count))
    return DFQ_ERR_INVALID_ARG;
  for (uint32_t i = 0; i < count; ++i)
    if (sources[i] >= g->n) return DFQ_ERR_RANGE;
  if (launch_ms) *launch_ms = 0.0;
  if (count == 0) return DFQ_OK;
  HIP_TRY(hipSetDevice(g->device));
  HIP_TRY(reserve_results(g, count));
  HIP_TRY(hipMemcpyAsync(g->d_src, sources, sizeof(uint32_t) * count, hipMemcpyHostToDevice, g->stream));
  BfsArgs a = make_args(g, sources[0], d_level, d_parent, false);
  a.result = g->d_res;
  const uint32_t *d_src = g->d_src;
  void *args[] = {&a, &d_src, &count, &level_stride, &parent_stride};
  HIP_TRY(hipEventRecord(g->ev0, g->stream));
  HIP_TRY(hipLaunchCooperativeKernel(seq_kernel_for(g->cfg.block_size, variant_for(g, a)), dim3(g->grid_seq),
                                     dim3(g->block), args, 0, g->stream));
  HIP_TRY(hipEventRecord(g->ev1, g->stream));
  HIP_TRY(hipEventSynchronize(g->ev1));
  uint32_t err = 0;
  HIP_TRY(hipMemcpy(&err, &g->d_ctl->err, sizeof(err), hipMemcpyDeviceToHost));
  if (err) return DFQ_ERR_CAPACITY;
  std::vector<uint64_t> res((size_t)kResultWords * count);
  HIP_TRY(hipMemcpy(res.data(), g->d_res, sizeof(uint64_t) * res.size(), hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < count; ++i) {
    const uint64_t *r = &res[(size_t)kResultWords * i];
    fill_stats(g, r, (double)(r[5] - r[4]) * 1e-6, g->grid_seq, stats[i]);
  }
  if (launch_ms) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    *launch_ms = ms;
  }
  return DFQ_OK;
}

namespace {
int run_multi(dfq_graph *g, uint32_t count, const uint32_t *sources, uint32_t *d_level, uint64_t level_stride,
              dfq_stats *stats, double *launch_ms, uint64_t *prof_rows, uint32_t prof_levels);
}

int dfq_bfs_device_multi(dfq_graph *g, uint32_t count, const uint32_t *sources, uint32_t *d_level,
                         uint64_t level_stride, dfq_stats *stats, double *launch_ms) {
  return run_multi(g, count, sources, d_level, level_stride, stats, launch_ms, nullptr, 0);
}

int dfq_bfs_multi_profile(dfq_graph *g, uint32_t count, const uint32_t *sources, uint32_t *d_level,
                          uint64_t level_stride, uint64_t *rows, uint32_t max_levels, dfq_stats *stats) {
  if (!rows || max_levels == 0) return DFQ_ERR_INVALID_ARG;
  return run_multi(g, count, sources, d_level, level_stride, stats, nullptr, rows, max_levels);
}

}  // extern "C"

namespace {
int run_multi(dfq_graph *g, uint32_t count, const uint32_t *sources, uint32_t *d_level, uint64_t level_stride,
              dfq_stats *stats, double *launch_ms, uint64_t *prof_rows, uint32_t prof_levels) {
  if (!g || !d_level || !sources || !stats || count == 0 || count > kMsMaxSources) return DFQ_ERR_INVALID_ARG;
  if (level_stride < g->n || level_stride > UINT64_MAX / count) return DFQ_ERR_INVALID_ARG;
  for (uint32_t i = 0; i < count; ++i)
    if (sources[i] >= g->n) return DFQ_ERR_RANGE;
  HIP_TRY(hipSetDevice(g->device));
  const size_t n = g->n;
  if (!g->d_ms) {
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)&dfq_msbfs_kernel<1024>, 1024, 0));
    if (per_cu <= 0) return DFQ_ERR_COOPERATIVE;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, g->device));
    hipError_t e = hipMalloc(&g->d_ms, sizeof(uint64_t) * n * (3 + kMsPlanes));
    if (e == hipSuccess) e = hipMalloc(&g->d_msq, sizeof(uint32_t) * n * 3 * kSegments);
    if (e == hipSuccess) e = hipMalloc(&g->d_mscnt, sizeof(uint64_t) * (2 * kMsMaxSources) + 2 * sizeof(uint32_t) * kMsMaxSources);
    if (e != hipSuccess) {
      (void)hipFree(g->d_ms);
      (void)hipFree(g->d_msq);
      (void)hipFree(g->d_mscnt);
      g->d_ms = nullptr;
      g->d_msq = nullptr;
      g->d_mscnt = nullptr;
      last_hip_error() = e;
      return DFQ_ERR_HIP;
    }
    g->grid_ms = prop.multiProcessorCount * per_cu;
  }
  if (!g->d_off_r) {
    // the degree-ranked copy, built on the host from the resident CSR
    std::vector<int64_t> off, off_r;
    std::vector<uint32_t> col, col_r, o2r, r2o;
    try {
      off.resize(n + 1);
      off_r.resize(n + 1);
      col.resize(g->m ? g->m : 1);
      col_r.resize(g->m ? g->m : 1);
      o2r.resize(n);
      r2o.resize(n);
    } catch (...) {
      return DFQ_ERR_NOMEM;
    }
    HIP_TRY(hipMemcpy(off.data(), g->d_off, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    if (g->m) HIP_TRY(hipMemcpy(col.data(), g->d_col, sizeof(uint32_t) * g->m, hipMemcpyDeviceToHost));
    rank_relabel(off.data(), col.data(), g->n, o2r.data(), r2o.data(), off_r.data(), col_r.data());
    uint32_t nz = g->n;
    while (nz > 0 && off_r[nz - 1] == off_r[nz]) --nz;
    g->nz = nz;
    hipError_t e = hipMalloc(&g->d_off_r, sizeof(int64_t) * (n + 1));
    // +4 entries: bottom-up levels read columns as aligned 16-byte blocks
    if (e == hipSuccess) e = hipMalloc(&g->d_col_r, sizeof(uint32_t) * (g->m + 4));
    if (e == hipSuccess) e = hipMalloc(&g->d_o2r, sizeof(uint32_t) * n);
    if (e == hipSuccess) e = hipMemcpy(g->d_off_r, off_r.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && g->m) e = hipMemcpy(g->d_col_r, col_r.data(), sizeof(uint32_t) * g->m, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g->d_o2r, o2r.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(g->d_off_r);
      (void)hipFree(g->d_col_r);
      (void)hipFree(g->d_o2r);
      g->d_off_r = nullptr;
      g->d_col_r = nullptr;
      g->d_o2r = nullptr;
      last_hip_error() = e;
      return DFQ_ERR_HIP;
    }
  }
  const uint64_t icap = ms_item_cap(g->n, g->m, g->cfg.large_max);
  if (icap > 0xffffffffull) return DFQ_ERR_CAPACITY;
  if (icap > g->msitem_cap) {
    (void)hipFree(g->d_msitems);
    g->d_msitems = nullptr;
    g->msitem_cap = 0;
    HIP_TRY(hipMalloc(&g->d_msitems, sizeof(uint2) * kSegments * icap));
    g->msitem_cap = icap;
  }
  uint32_t *d_depth = (uint32_t *)(g->d_mscnt + 2 * kMsMaxSources);
  uint32_t *d_src = d_depth + kMsMaxSources;
  HIP_TRY(hipMemcpyAsync(d_src, sources, sizeof(uint32_t) * count, hipMemcpyHostToDevice, g->stream));
  MsArgs a;
  memset(&a, 0, sizeof(a));
  a.off = g->d_off_r;
  a.col = g->d_col_r;
  a.col_bu = nullptr;  // the ranked rows are already highest degree first
  a.o2r = g->d_o2r;
  a.nz = g->nz;
  a.n = g->n;
  a.count = count;
  a.m = g->m;
  a.sources = d_src;
  a.level = d_level;
  a.level_stride = level_stride;
  a.seen = g->d_ms;
  a.f = g->d_ms + n;
  a.planes = g->d_ms + 3 * n;
  a.qb = g->d_msq;
  a.segs = kSegments;
  a.items = g->d_msitems;
  a.item_cap = (uint32_t)g->msitem_cap;
  a.ctl = g->d_ctl;
  a.counters = g->d_mscnt;
  a.depth = d_depth;
  a.small_max = g->cfg.small_max;
  a.medium_max = g->cfg.medium_max;
  a.large_max = g->cfg.large_max;
  a.bottom_up = (g->symmetric && g->cfg.direction_opt) ? 1u : 0u;
  a.bu_div = g->cfg.multi_bu_div ? g->cfg.multi_bu_div : kMsBuDiv;
  if (prof_rows) {
    if (prof_levels > g->msprof_levels) {
      (void)hipFree(g->d_msprof);
      g->d_msprof = nullptr;
      g->msprof_levels = 0;
      HIP_TRY(hipMalloc(&g->d_msprof, sizeof(uint64_t) * kMsProfFields * prof_levels));
      g->msprof_levels = prof_levels;
    }
    HIP_TRY(hipMemsetAsync(g->d_msprof, 0, sizeof(uint64_t) * kMsProfFields * prof_levels, g->stream));
    a.prof = g->d_msprof;
    a.prof_levels = prof_levels;
  }
  void *args[] = {&a};
  HIP_TRY(hipEventRecord(g->ev0, g->stream));
  HIP_TRY(hipLaunchCooperativeKernel((const void *)&dfq_msbfs_kernel<1024>, dim3(g->grid_ms), dim3(1024), args, 0,
                                     g->stream));
  HIP_TRY(hipEventRecord(g->ev1, g->stream));
  HIP_TRY(hipEventSynchronize(g->ev1));
  Control c;
  HIP_TRY(hipMemcpy(&c, g->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
  if (c.err == kMsErrDepth) return DFQ_ERR_CAPACITY;
  if (c.err) return DFQ_ERR_HIP;
  uint64_t cnt[2 * kMsMaxSources];
  uint32_t dep[kMsMaxSources];
  HIP_TRY(hipMemcpy(cnt, g->d_mscnt, sizeof(cnt), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(dep, d_depth, sizeof(dep), hipMemcpyDeviceToHost));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  for (uint32_t i = 0; i < count; ++i) {
    stats[i].kernel_ms = ms;
    stats[i].depth = dep[i];
    stats[i].reached = cnt[i];
    stats[i].edges = cnt[kMsMaxSources + i];
    stats[i].grid_blocks = (uint32_t)g->grid_ms;
    stats[i].block_size = 1024;
    stats[i].bottom_up_levels = c.bu_levels;
  }
  if (launch_ms) *launch_ms = ms;
  if (prof_rows)
    HIP_TRY(hipMemcpy(prof_rows, g->d_msprof, sizeof(uint64_t) * kMsProfFields * prof_levels, hipMemcpyDeviceToHost));
  return DFQ_OK;
}
}  // namespace

extern "C" {

int dfq_bfs_profile(dfq_graph *g, uint32_t source, uint64_t *rows, uint32_t max_levels, dfq_stats *stats) {
  if (!g || !rows || max_levels == 0) return DFQ_ERR_INVALID_ARG;
  HIP_TRY(hipSetDevice(g->device));
  if (max_levels > g->prof_levels) {
    (void)hipFree(g->d_prof);
    g->d_prof = nullptr;
    g->prof_levels = 0;
    HIP_TRY(hipMalloc(&g->d_prof, sizeof(uint64_t) * kProfFields * max_levels));
    g->prof_levels = max_levels;
  }
  HIP_TRY(hipMemset(g->d_prof, 0, sizeof(uint64_t) * kProfFields * g->prof_levels));
  int rc = run_bfs(g, source, g->d_level, nullptr, stats, true);
  if (rc != DFQ_OK) return rc;
  HIP_TRY(hipMemcpy(rows, g->d_prof, sizeof(uint64_t) * kProfFields * max_levels, hipMemcpyDeviceToHost));
  return DFQ_OK;
}

int dfq_device_alloc(size_t bytes, void **d_ptr) {
  if (!d_ptr || bytes == 0) return DFQ_ERR_INVALID_ARG;
  *d_ptr = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DFQ_ERR_NO_DEVICE;
  HIP_TRY(hipMalloc(d_ptr, bytes));
  return DFQ_OK;
}

int dfq_device_free(void *d_ptr) {
  if (!d_ptr) return DFQ_OK;
  HIP_TRY(hipFree(d_ptr));
  return DFQ_OK;
}

int dfq_copy_to_host(void *h_dst, const void *d_src, size_t bytes) {
  if (!h_dst || !d_src) return DFQ_ERR_INVALID_ARG;
  HIP_TRY(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  return DFQ_OK;
}

int dfq_copy_to_device(void *d_dst, const void *h_src, size_t bytes) {
  if (!d_dst || !h_src) return DFQ_ERR_INVALID_ARG;
  HIP_TRY(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  return DFQ_OK;
}

int dfq_stream_copy(void *d_dst, const void *d_src, size_t bytes, double *ms_out) {
  if (!d_dst || !d_src || bytes == 0 || bytes % 16) return DFQ_ERR_INVALID_ARG;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int dev = 0;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  const size_t n16 = bytes / 16;
  const size_t want = (n16 + 255) / 256;
  const unsigned grid = (unsigned)std::min<size_t>(want, (size_t)prop.multiProcessorCount * 8);
  HIP_TRY(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(dfq_stream_copy_kernel, dim3(grid), dim3(256), 0, nullptr, (uint4 *)d_dst,
                     (const uint4 *)d_src, n16);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(e1, nullptr));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_out) *ms_out = ms;
  return DFQ_OK;
}

int dfq_device_synchronize(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DFQ_ERR_NO_DEVICE;
  HIP_TRY(hipDeviceSynchronize());
  return DFQ_OK;
}

}  // extern "C"
