// BFS entry point that allocates nothing: dfq_bfs_device_noalloc
// (include/dfq_bfs.h).
//
// The same persistent search kernel as dfq_bfs_device (dfq_kernels.hpp,
// top-down levels through the degree bins, default cut-offs), over a CSR the
// caller already holds in device memory, with every scratch buffer of the
// search - control block with the grid barrier, visited bitmap and its
// snapshot, frontier bitmaps, the two sets of per-XCD frontier segments and
// the work items of huge rows - carved from one file-scope device array.
// The library makes no hipMalloc / hipFree (or any other allocating call)
// on this path: for callers that cannot allocate (serving loops, code run
// under a harness that forbids device allocation).
//
// Per device, at the first call on it (with that device current): the
// array's device address is looked up once with hipGetSymbolAddress and
// kept; the launch grid (CUs x co-resident workgroups) and two timing
// events are set up alongside.  After that only the kept address is used
// (kernel arguments, copies): in host code the array's name is a host-side
// stand-in, not a device address.
//
// Every call on a device shares that device's array: calls that overlap on
// one GPU race and give wrong results.  The control block keeps the state
// of the last call (barrier generations, the launch-parity words of the
// self-assembling grid), which is what the next call expects; everything a
// search reads before writing it (levels, bitmaps, the level-0 frontier,
// the per-level tails, the error word) is set by the kernel at search start.

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "dfq_bfs.h"
#include "dfq_kernels.hpp"

using namespace dfq;

namespace {

constexpr uint64_t kNaMaxN = DFQ_NOALLOC_MAX_VERTICES;
constexpr uint64_t kNaMaxM = DFQ_NOALLOC_MAX_ENTRIES;
constexpr uint32_t kNaBlock = 1024;
// Bins of the default configuration.  Rows above kNaLargeMax >= kItem
// entries are cut into ceil(d / kItem) <= d / 128 work items, so one
// level's items never exceed m / 128: any segment may receive all of them.
constexpr uint32_t kNaSmallMax = 32, kNaMediumMax = 256, kNaLargeMax = 256;
static_assert(kNaLargeMax >= kItem, "work-item bound needs rows longer than one item");
constexpr uint64_t kNaItemCap = kNaMaxM / 128 + 1;

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }
constexpr uint64_t kNaWords = (kNaMaxN + 31) / 32;
// Layout of the scratch array (byte offsets, each 256-aligned).
constexpr uint64_t kOffCtl = 0;
constexpr uint64_t kOffErr = align256(kOffCtl + sizeof(Control));        // CSR check word
constexpr uint64_t kOffVis = align256(kOffErr + sizeof(uint32_t));
constexpr uint64_t kOffSnap = align256(kOffVis + 4 * kNaWords);
constexpr uint64_t kOffFb = align256(kOffSnap + 4 * kNaWords);
constexpr uint64_t kOffQ0 = align256(kOffFb + 4 * 2 * (kNaWords + 1));
constexpr uint64_t kOffQ1 = align256(kOffQ0 + 4 * (uint64_t)kSegments * kNaMaxN);
constexpr uint64_t kOffItems = align256(kOffQ1 + 4 * (uint64_t)kSegments * kNaMaxN);
constexpr uint64_t kScratchBytes = align256(kOffItems + 8 * (uint64_t)kSegments * kNaItemCap);
static_assert(kScratchBytes == DFQ_NOALLOC_SCRATCH_BYTES, "header states a different scratch size");

// The scratch array: one copy per GPU the library runs on, held until the
// process ends.  Zero at load, which is the control block's initial state.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// CSR check (the checks dfq_graph_create makes on the host): offsets start
// at 0, end at m and never decrease; every column index is below n.  Sets
// bits of *err.
__global__ __launch_bounds__(256) void noalloc_check_csr(const int64_t *__restrict__ off,
                                                         const uint32_t *__restrict__ col, uint32_t n, uint64_t m,
                                                         uint32_t *err) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  uint32_t bad = 0;
  const uint64_t i0 = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i0 == 0 && (off[0] != 0 || (uint64_t)off[n] != m)) bad |= 1u;
  for (uint64_t v = i0; v < n; v += stride)
    if (off[v + 1] < off[v]) bad |= 2u;
  for (uint64_t i = i0; i < m; i += stride)
    if (col[i] >= n) bad |= 4u;
  if (bad) atomicOr(err, bad);
}

__global__ void noalloc_clear_word(uint32_t *w) { *w = 0; }

struct DeviceState {
  bool ready = false;
  unsigned char *ptr = nullptr;  // device address of `scratch` on this device
  int grid = 0;
  uint32_t launches = 0;         // number of the last plain launch (never 0)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

constexpr int kMaxDevices = 64;
DeviceState g_dev[kMaxDevices];
std::mutex g_mu;

// First call on the current device: the array's address, the grid and the
// events.  Returns a DFQ status.
int setup(int dev, DeviceState *&out) {
  if (dev < 0 || dev >= kMaxDevices) return DFQ_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceState &s = g_dev[dev];
  if (!s.ready) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DFQ_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return DFQ_ERR_HIP;
    if (!prop.cooperativeLaunch) return DFQ_ERR_COOPERATIVE;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)&dfq_bfs_kernel<kNaBlock>,
                                                      (int)kNaBlock, 0) != hipSuccess)
      return DFQ_ERR_HIP;
    if (per_cu <= 0) return DFQ_ERR_COOPERATIVE;
    if (hipEventCreate(&s.ev0) != hipSuccess || hipEventCreate(&s.ev1) != hipSuccess) return DFQ_ERR_HIP;
    s.ptr = (unsigned char *)p;
    s.grid = prop.multiProcessorCount * per_cu;
    s.ready = true;
  }
  out = &s;
  return DFQ_OK;
}

}  // namespace

extern "C" int dfq_bfs_device_noalloc(const int64_t *d_row_offsets, const uint32_t *d_col_idx, uint32_t n,
                                      uint64_t m, uint32_t source, uint32_t *d_level, uint32_t *d_parent,
                                      dfq_stats *stats) {
  if (!d_row_offsets || (m && !d_col_idx) || !d_level || n == 0) return DFQ_ERR_INVALID_ARG;
  if (n > kNaMaxN || m > kNaMaxM) return DFQ_ERR_CAPACITY;
  if (source >= n) return DFQ_ERR_RANGE;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DFQ_ERR_NO_DEVICE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return DFQ_ERR_HIP;
  DeviceState *s = nullptr;
  int rc = setup(dev, s);
  if (rc != DFQ_OK) return rc;
  unsigned char *base = s->ptr;
  Control *ctl = (Control *)(base + kOffCtl);
  uint32_t *err = (uint32_t *)(base + kOffErr);

  // the caller's CSR is checked on the device before the search reads it
  hipLaunchKernelGGL(noalloc_clear_word, dim3(1), dim3(1), 0, nullptr, err);
  const uint64_t work = (uint64_t)n + 1 > m ? (uint64_t)n + 1 : m;
  const unsigned cgrid = (unsigned)((work + 255) / 256 < 1024 ? (work + 255) / 256 : 1024);
  hipLaunchKernelGGL(noalloc_check_csr, dim3(cgrid), dim3(256), 0, nullptr, d_row_offsets, d_col_idx, n, m,
                     err);
  if (hipGetLastError() != hipSuccess) return DFQ_ERR_HIP;
  uint32_t bad = 0;
  if (hipMemcpy(&bad, err, sizeof(bad), hipMemcpyDeviceToHost) != hipSuccess) return DFQ_ERR_HIP;
  if (bad) return DFQ_ERR_RANGE;

  BfsArgs a;
  memset(&a, 0, sizeof(a));
  a.off = d_row_offsets;
  a.col = d_col_idx;
  a.col_bu = nullptr;
  a.head = nullptr;
  a.level = d_level;
  a.parent = d_parent;
  a.iso = nullptr;
  a.vis = (uint32_t *)(base + kOffVis);
  a.vis_snap = (uint32_t *)(base + kOffSnap);
  a.fb = (uint32_t *)(base + kOffFb);
  a.qseg[0] = (uint32_t *)(base + kOffQ0);
  a.qseg[1] = (uint32_t *)(base + kOffQ1);
  a.items = (uint2 *)(base + kOffItems);
  a.item_cap = (uint32_t)kNaItemCap;
  a.ctl = ctl;
  a.prof = nullptr;
  a.prof_levels = 0;
  a.result = nullptr;
  a.n = n;
  a.n_words = (n + 31) / 32;
  a.source = source;
  a.small_max = kNaSmallMax;
  a.medium_max = kNaMediumMax;
  a.large_max = kNaLargeMax;
  a.bottom_up = 0;  // top-down levels only: no degree-ordered copy of the rows exists here
  a.alpha = 0;
  a.bu_exit = 0;
  a.m = m;
  a.fuse_bits = fuse_bits_for(n, m);
  // one plain launch whose grid assembles itself; cooperative when it does
  // not (other work held CUs), as dfq_bfs_device does
  Control c;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (attempt == 0) {
      s->launches = s->launches == 0xffffffffu ? 2u : s->launches + 1u;
      a.plain = s->launches;
      a.plain_wait_ns = 1000000;
    } else {
      a.plain = 0;
      a.plain_wait_ns = 0;
    }
    void *args[] = {&a};
    if (hipEventRecord(s->ev0, nullptr) != hipSuccess) return DFQ_ERR_HIP;
    hipError_t e = a.plain ? hipLaunchKernel((const void *)&dfq_bfs_kernel<kNaBlock>, dim3(s->grid), dim3(kNaBlock),
                                             args, 0, nullptr)
                           : hipLaunchCooperativeKernel((const void *)&dfq_bfs_kernel<kNaBlock>, dim3(s->grid),
                                                        dim3(kNaBlock), args, 0, nullptr);
    if (e != hipSuccess) return DFQ_ERR_HIP;
    if (hipEventRecord(s->ev1, nullptr) != hipSuccess || hipEventSynchronize(s->ev1) != hipSuccess)
      return DFQ_ERR_HIP;
    if (hipMemcpy(&c, ctl, sizeof(Control), hipMemcpyDeviceToHost) != hipSuccess) return DFQ_ERR_HIP;
    if (!(a.plain && c.abandoned == a.plain)) break;
  }
  if (c.err) return DFQ_ERR_CAPACITY;
  if (stats) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, s->ev0, s->ev1) != hipSuccess) return DFQ_ERR_HIP;
    stats->kernel_ms = ms;
    stats->depth = c.depth;
    stats->reached = c.reached;
    stats->edges = c.edges;
    stats->grid_blocks = (uint32_t)s->grid;
    stats->block_size = kNaBlock;
    stats->bottom_up_levels = c.bu_levels;
  }
  return DFQ_OK;
}
