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

#include "dfq_bfs.h"
#include "dfq_kernels.hpp"
#include "host_common.hpp"

using namespace dfq;

#define HIP_TRY(expr)                 \
  do {                                \
    hipError_t e__ = (expr);          \
    if (e__ != hipSuccess) {          \
      last_hip_error() = e__;         \
      return DFQ_ERR_HIP;             \
    }                                 \
  } while (0)

namespace {

hipError_t &last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

struct LaunchShape {
  int grid = 0;
  int block = 0;
};

}  // namespace

struct dfq_graph {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t n_words = 0;
  int64_t *d_off = nullptr;
  uint32_t *d_col = nullptr;
  uint32_t *d_level = nullptr;   // outputs for the host-array entry point
  uint32_t *d_parent = nullptr;
  uint32_t *d_vis = nullptr;
  uint32_t *d_q0 = nullptr, *d_q1 = nullptr, *d_hq = nullptr;
  Control *d_ctl = nullptr;
  dfq_config cfg{};
  LaunchShape shape{};
};

namespace {

int validate_config(const dfq_config &c) {
  if (c.block_size != 256 && c.block_size != 512) return DFQ_ERR_INVALID_ARG;
  if (c.small_max > 64) return DFQ_ERR_INVALID_ARG;
  if (c.medium_max < c.small_max || c.large_max < c.medium_max) return DFQ_ERR_INVALID_ARG;
  if (c.reserved[0] || c.reserved[1] || c.reserved[2]) return DFQ_ERR_INVALID_ARG;
  return DFQ_OK;
}

const void *kernel_for(uint32_t block) {
  return block == 512 ? (const void *)&dfq_bfs_kernel<512> : (const void *)&dfq_bfs_kernel<256>;
}

// Grid = CUs x co-resident workgroups per CU (cooperative launch requires
// every workgroup to be resident).
int compute_shape(dfq_graph *g) {
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, g->device));
  if (!prop.cooperativeLaunch) return DFQ_ERR_COOPERATIVE;
  int per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_for(g->cfg.block_size),
                                                        (int)g->cfg.block_size, 0));
  if (per_cu <= 0) return DFQ_ERR_COOPERATIVE;
  if (g->cfg.blocks_per_cu) per_cu = std::min(per_cu, (int)g->cfg.blocks_per_cu);
  g->shape.grid = prop.multiProcessorCount * per_cu;
  g->shape.block = (int)g->cfg.block_size;
  return DFQ_OK;
}

void free_device(dfq_graph *g) {
  (void)hipFree(g->d_off);
  (void)hipFree(g->d_col);
  (void)hipFree(g->d_level);
  (void)hipFree(g->d_parent);
  (void)hipFree(g->d_vis);
  (void)hipFree(g->d_q0);
  (void)hipFree(g->d_q1);
  (void)hipFree(g->d_hq);
  (void)hipFree(g->d_ctl);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  if (g->stream) (void)hipStreamDestroy(g->stream);
}

int run_bfs(dfq_graph *g, uint32_t source, uint32_t *d_level, uint32_t *d_parent, dfq_stats *stats) {
  if (source >= g->n) return DFQ_ERR_RANGE;
  HIP_TRY(hipSetDevice(g->device));
  BfsArgs a;
  a.off = g->d_off;
  a.col = g->d_col;
  a.level = d_level;
  a.parent = d_parent;
  a.vis = g->d_vis;
  a.q[0] = g->d_q0;
  a.q[1] = g->d_q1;
  a.hq = g->d_hq;
  a.ctl = g->d_ctl;
  a.n = g->n;
  a.n_words = g->n_words;
  a.source = source;
  a.small_max = g->cfg.small_max;
  a.medium_max = g->cfg.medium_max;
  a.large_max = g->cfg.large_max;
  void *args[] = {&a};
  HIP_TRY(hipEventRecord(g->ev0, g->stream));
  HIP_TRY(hipLaunchCooperativeKernel(kernel_for(g->cfg.block_size), dim3(g->shape.grid),
                                     dim3(g->shape.block), args, 0, g->stream));
  HIP_TRY(hipEventRecord(g->ev1, g->stream));
  HIP_TRY(hipEventSynchronize(g->ev1));
  if (stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    Control c;
    HIP_TRY(hipMemcpy(&c, g->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
    stats->kernel_ms = ms;
    stats->depth = c.depth;
    stats->reached = c.reached;
    stats->grid_blocks = (uint32_t)g->shape.grid;
    stats->block_size = (uint32_t)g->shape.block;
    stats->reserved0 = 0;
  }
  return DFQ_OK;
}

}  // namespace

extern "C" {

int dfq_config_default(dfq_config *cfg) {
  if (!cfg) return DFQ_ERR_INVALID_ARG;
  memset(cfg, 0, sizeof(*cfg));
  cfg->block_size = 256;
  cfg->blocks_per_cu = 0;
  cfg->small_max = 32;
  cfg->medium_max = 2048;
  cfg->large_max = 32768;
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
  (void)xcc;
  info->xcc_count = 0;
  strncpy(info->arch, prop.gcnArchName, sizeof(info->arch) - 1);
  return DFQ_OK;
}

int dfq_graph_create(const int64_t *row_offsets, const uint32_t *col_idx, uint32_t n, uint64_t m,
                     dfq_graph **out) {
  if (!out || !row_offsets || (m && !col_idx) || n == 0 || n == 0xffffffffu) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  // Input checks on the host (the reference trusted its file blindly).
  if (row_offsets[0] != 0 || (uint64_t)row_offsets[n] != m) return DFQ_ERR_RANGE;
  std::atomic<bool> bad{false};
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t v = b; v < e; ++v)
      if (row_offsets[v + 1] < row_offsets[v]) {
        bad = true;
        return;
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
  int rc = DFQ_OK;
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
  dfq_config_default(&g->cfg);
  HIP_OR_FAIL(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  HIP_OR_FAIL(hipEventCreate(&g->ev0));
  HIP_OR_FAIL(hipEventCreate(&g->ev1));
  HIP_OR_FAIL(hipMalloc(&g->d_off, sizeof(int64_t) * ((size_t)n + 1)));
  HIP_OR_FAIL(hipMalloc(&g->d_col, sizeof(uint32_t) * (m ? m : 1)));
  HIP_OR_FAIL(hipMalloc(&g->d_level, sizeof(uint32_t) * n));
  HIP_OR_FAIL(hipMalloc(&g->d_parent, sizeof(uint32_t) * n));
  HIP_OR_FAIL(hipMalloc(&g->d_vis, sizeof(uint32_t) * g->n_words));
  HIP_OR_FAIL(hipMalloc(&g->d_q0, sizeof(uint32_t) * n));
  HIP_OR_FAIL(hipMalloc(&g->d_q1, sizeof(uint32_t) * n));
  HIP_OR_FAIL(hipMalloc(&g->d_hq, sizeof(uint32_t) * n));
  HIP_OR_FAIL(hipMalloc(&g->d_ctl, sizeof(Control)));
  HIP_OR_FAIL(hipMemcpy(g->d_off, row_offsets, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyHostToDevice));
  if (m) HIP_OR_FAIL(hipMemcpy(g->d_col, col_idx, sizeof(uint32_t) * m, hipMemcpyHostToDevice));
  HIP_OR_FAIL(hipMemset(g->d_ctl, 0, sizeof(Control)));
#undef HIP_OR_FAIL
  rc = compute_shape(g);
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
  dfq_config old = g->cfg;
  LaunchShape old_shape = g->shape;
  g->cfg = *cfg;
  rc = compute_shape(g);
  if (rc != DFQ_OK) {
    g->cfg = old;
    g->shape = old_shape;
  }
  return rc;
}

int dfq_graph_get_config(const dfq_graph *g, dfq_config *cfg) {
  if (!g || !cfg) return DFQ_ERR_INVALID_ARG;
  *cfg = g->cfg;
  return DFQ_OK;
}

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

}  // extern "C"
