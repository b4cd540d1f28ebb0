/*
 * dfq_bfs.h — device BFS engine of the MI355X DFQ-BFS library (plain C ABI).
 *
 * Replaces the reference's device path of
 * design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:
 *
 *   reference                                            replaced by
 *   ---------------------------------------------------  ---------------------------
 *   :186,194-197 cudaMallocManaged node/edge/cost        dfq_graph_create (explicit
 *   :273-277     cudaMemPrefetchAsync                     device copies, no managed
 *                                                         memory)
 *   :98-179      __global__ CUDA_BFS_KERNEL_CG_NEW(Node*, dfq_bfs / dfq_bfs_device:
 *                int*, short*, bool*, int*)              one persistent launch per
 *   :255-271     occupancy query, grid = SMs x blocks     BFS on gfx950 (plain by
 *   :279-288     cudaEvent timing around                  default: the grid assembles
 *                cudaLaunchCooperativeKernel + sync       itself, cooperative fallback;
 *                                                         see launch_mode), hipEvent
 *                                                         timing in dfq_stats.kernel_ms
 *   :312-316     cudaFree                                dfq_graph_destroy
 *
 * Output convention: level[v] is the hop distance from the source (uint32,
 * DFQ_UNREACHED when unreachable; the reference used short cost[] with -1).
 * parent[v] (optional) is a BFS-tree parent: parent[source] == source,
 * DFQ_UNREACHED for unreached vertices.
 *
 * All calls return DFQ_OK (0) or a negative DFQ_ERR_* code (dfq_graph.h).
 * A dfq_graph may be used by one host thread at a time.
 */
#ifndef DFQ_BFS_H
#define DFQ_BFS_H

#include <stddef.h>
#include <stdint.h>

#include "dfq_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tuning knobs of the persistent kernel (the reference fixed these with
 * #defines, bfs.cu:16-36 upper-bound, lower-bound/bfs.cu:16-41).
 * Degree bins of a top-down level:
 *              deg <= small_max       -> lane-owned, wave-gathered edges
 *              deg <= medium_max      -> one wavefront (64 lanes) per vertex
 *              deg <= large_max       -> one workgroup per vertex
 *              deg >  large_max       -> 256-entry work items over the grid
 * direction_opt = 1 lets a level run bottom-up (every unvisited vertex looks
 * for a parent in the frontier) when the frontier is growing and its
 * out-edges exceed the unexplored edges / bu_enter_div (Beamer's alpha;
 * default 40, measured on MI355X, where a top-down level costs relatively
 * more than on the CPUs alpha = 14 was tuned for), and returns to top-down
 * when a shrinking frontier holds fewer
 * than n / bu_exit_div vertices (beta, default 96, measured on MI355X:
 * bottom-up levels over few remaining vertices are cheap there).  It only takes effect on
 * symmetric graphs (checked when the graph is created); levels are the same
 * either way.  direction_opt = 0 runs every level top-down. */
typedef struct dfq_config {
  uint32_t block_size;     /* threads per workgroup: 256, 512 or 1024     */
  uint32_t blocks_per_cu;  /* 0 = as many as occupancy allows             */
  uint32_t small_max;      /* <= 64                                        */
  uint32_t medium_max;     /* >= small_max                                 */
  uint32_t large_max;      /* >= medium_max                                */
  uint32_t direction_opt;  /* 0 or 1                                       */
  uint32_t bu_enter_div;   /* alpha, > 0 when direction_opt                */
  uint32_t bu_exit_div;    /* beta, > 0 when direction_opt                 */
  uint32_t launch_mode;    /* DFQ_LAUNCH_*, default DFQ_LAUNCH_PLAIN        */
  uint32_t multi_bu_div;   /* dfq_bfs_device_multi: a level expands bottom-up
                              when its frontier's out-edges exceed m /
                              multi_bu_div (0 = default, 2)                  */
} dfq_config;

/* How dfq_bfs / dfq_bfs_device / dfq_bfs_device_many / dfq_bfs_profile
 * launch the search kernel.  The reference launched cooperatively
 * (bfs.cu:283); on this stack a cooperative launch costs about 19 us
 * between kernels against 1.5 us for a plain one, so by default the kernel
 * is launched plainly and assembles its grid itself: its workgroups count
 * themselves in with a bounded wait, and if the whole grid is not resident
 * within 1 ms (other work holding CUs) every workgroup gives up alike and
 * the search is run again with a cooperative launch.  Results are the same
 * either way.  dfq_bfs_device_seq always launches cooperatively (once per
 * sequence). */
#define DFQ_LAUNCH_PLAIN 0         /* plain launch, cooperative when it gives up */
#define DFQ_LAUNCH_COOPERATIVE 1   /* always cooperative                         */
#define DFQ_LAUNCH_PLAIN_ABANDON 2 /* plain launch that gives up at once (tests the fallback) */

typedef struct dfq_stats {
  double   kernel_ms;        /* hipEvent time of the search's launch      */
  uint32_t depth;            /* number of BFS levels (max level + 1)      */
  uint32_t grid_blocks;      /* workgroups launched                       */
  uint32_t block_size;       /* threads per workgroup                     */
  uint32_t bottom_up_levels; /* levels that ran bottom-up                 */
  uint64_t reached;          /* vertices with a level                     */
  uint64_t edges;            /* CSR entries of reached vertices (E_r)     */
} dfq_stats;

typedef struct dfq_device_info {
  int32_t  device;              /* HIP ordinal                            */
  int32_t  cu_count;            /* compute units                          */
  int32_t  xcc_count;           /* accelerator complex dies (0 = unknown) */
  int32_t  cooperative;         /* cooperative launch supported           */
  uint64_t lds_per_cu;          /* bytes of LDS per CU                    */
  uint64_t hbm_bytes;           /* global memory                          */
  uint64_t l2_bytes;            /* L2 per XCD as reported                 */
  char     arch[64];            /* gcnArchName                            */
} dfq_device_info;

typedef struct dfq_graph dfq_graph;   /* opaque, device-resident CSR */

/* Defaults used when no config is set (tuned for gfx950). */
int dfq_config_default(dfq_config *cfg);

/* Information about the current HIP device; DFQ_ERR_NO_DEVICE without one. */
int dfq_get_device_info(dfq_device_info *info);

/* Upload a host CSR (row_offsets[n+1] int64, col_idx[m] uint32) to HBM and
 * allocate the BFS working set (bitmap, queues, control block).  Row offsets
 * must be non-decreasing with row_offsets[0] == 0 and row_offsets[n] == m,
 * every column index < n, and m < 2^40 (the level loop's grid-wide sums are
 * 40-bit fields); violations give DFQ_ERR_RANGE. */
int dfq_graph_create(const int64_t *row_offsets, const uint32_t *col_idx,
                     uint32_t n, uint64_t m, dfq_graph **out);
void dfq_graph_destroy(dfq_graph *g);

int dfq_graph_set_config(dfq_graph *g, const dfq_config *cfg);
int dfq_graph_get_config(const dfq_graph *g, dfq_config *cfg);

/* Device pointers of the resident CSR (for callers that keep other device
 * data next to it).  Valid until dfq_graph_destroy.  The columns are the
 * input's, in the input's order. */
int dfq_graph_device_arrays(const dfq_graph *g, const int64_t **d_row_offsets,
                            const uint32_t **d_col_idx);

/* Single-source BFS.  level_out[n] is required, parent_out[n] optional
 * (NULL skips parent output); both are host arrays.  stats optional.
 * Blocks until the result is in the host arrays. */
int dfq_bfs(dfq_graph *g, uint32_t source, uint32_t *level_out,
            uint32_t *parent_out, dfq_stats *stats);

/* 1 when the graph is symmetric (every u->v has v->u, rows sorted), which
 * is what bottom-up levels need; 0 otherwise. */
int dfq_graph_is_symmetric(const dfq_graph *g);

/* Same BFS with outputs left in device memory (d_level[n] required,
 * d_parent[n] optional).  Blocks until the kernel has finished; the timed
 * region of stats->kernel_ms is the cooperative launch only. */
int dfq_bfs_device(dfq_graph *g, uint32_t source, uint32_t *d_level,
                   uint32_t *d_parent, dfq_stats *stats);

/* ---- BFS for callers that cannot allocate ---------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops that must not call hipMalloc, harnesses that forbid
 * it): the same BFS as dfq_bfs_device under the default configuration
 * (top-down levels through the degree bins) over a CSR the caller already
 * holds in device memory, d_row_offsets[n + 1] (int64) and d_col_idx[m].
 * The library allocates nothing on this path: every scratch buffer of the
 * search comes from one static device array of DFQ_NOALLOC_SCRATCH_BYTES
 * (~20 MiB) per GPU, held until the process ends.  Limits: n <=
 * DFQ_NOALLOC_MAX_VERTICES, m <= DFQ_NOALLOC_MAX_ENTRIES, else
 * DFQ_ERR_CAPACITY.  The CSR is checked on the device first (offsets from 0
 * to m, never decreasing, columns < n; DFQ_ERR_RANGE otherwise).  Outputs:
 * d_level[n] required, d_parent[n] optional, as for dfq_bfs_device; runs on
 * the current device's null stream and blocks until done.  Calls share the
 * device's array: calls that overlap on one GPU race and must not be made. */
#define DFQ_NOALLOC_MAX_VERTICES (1u << 18)
#define DFQ_NOALLOC_MAX_ENTRIES (1ull << 23)
#define DFQ_NOALLOC_SCRATCH_BYTES 21115904ull
int dfq_bfs_device_noalloc(const int64_t *d_row_offsets, const uint32_t *d_col_idx,
                           uint32_t n, uint64_t m, uint32_t source,
                           uint32_t *d_level, uint32_t *d_parent, dfq_stats *stats);

/* count searches back to back on the graph's stream, outputs in device
 * memory (each search overwrites the previous one's), one wait at the end.
 * stats[i] belongs to sources[i]; its kernel_ms is bracketed by hipEvents
 * recorded on the stream right before and after that launch, with every
 * launch already queued, so it does not include the host's submission
 * latency that a single dfq_bfs_device call does. */
int dfq_bfs_device_many(dfq_graph *g, uint32_t count, const uint32_t *sources,
                        uint32_t *d_level, uint32_t *d_parent, dfq_stats *stats);

/* count searches in ONE cooperative launch: the grid runs them in turn,
 * each one's initialisation barrier also closing the one before, so the
 * launch cost is paid once (the reference's interactive loop,
 * bfs.cu:202-214, relaunched the kernel per source).  Search i writes
 * d_level + i * level_stride and, when d_parent is given, d_parent +
 * i * parent_stride; a stride of 0 makes every search overwrite the same
 * array (else it must be >= n).  stats[i].kernel_ms is search i's
 * device-clock time inside the launch (its initialisation to its last
 * level); *launch_ms (optional) is the hipEvent time of the whole launch.
 * Same results as count dfq_bfs_device calls. */
int dfq_bfs_device_seq(dfq_graph *g, uint32_t count, const uint32_t *sources,
                       uint32_t *d_level, uint64_t level_stride,
                       uint32_t *d_parent, uint64_t parent_stride,
                       dfq_stats *stats, double *launch_ms);

/* Up to DFQ_MULTI_MAX_SOURCES searches of one graph in one traversal
 * (one cooperative launch): search i starts from sources[i] and writes its
 * levels to d_level + i * level_stride (level_stride >= n).  Each vertex
 * carries one bit per search (bit-parallel multi-source BFS), so a pass over
 * an edge serves every search whose frontier holds its tail.  Levels are
 * those of dfq_bfs for every search (sources may repeat).  No parents.
 * stats[i]: depth and reached of search i; kernel_ms the launch's event
 * time (also in *launch_ms when given), bottom_up_levels the traversal's.
 * Levels deeper than 255 are not covered: DFQ_ERR_CAPACITY (run such
 * sources with dfq_bfs_device). */
#define DFQ_MULTI_MAX_SOURCES 64u
int dfq_bfs_device_multi(dfq_graph *g, uint32_t count, const uint32_t *sources,
                         uint32_t *d_level, uint64_t level_stride, dfq_stats *stats,
                         double *launch_ms);

/* The same with per-level counters, rows[max_levels][DFQ_MULTI_PROFILE_FIELDS]:
 *   0 distinct vertices in the level's frontier (over all searches),
 *   1 their out-degree sum, 2 1 if the level expanded bottom-up,
 *   3 searches with a frontier at this level, 4 work items (rows cut into
 *   256-entry pieces), 5 ns of the commit phase, 6 ns of the expansion
 *   (both barriers included, timed by the first workgroup), 7 vertices put
 *   in the next level's frontier.  Levels past max_levels are not recorded. */
#define DFQ_MULTI_PROFILE_FIELDS 8
int dfq_bfs_multi_profile(dfq_graph *g, uint32_t count, const uint32_t *sources,
                          uint32_t *d_level, uint64_t level_stride, uint64_t *rows,
                          uint32_t max_levels, dfq_stats *stats);

/* BFS with per-level counters (the reference's frontier profiling build,
 * design/shared-q/atomic-sa/our-sep/upper-bound/bfs_ftr_prof.cu:98,134-140,
 * 200-209, and its per-thread workload counts,
 * design/shared-q/atomic-sa/sup-sep/upper-bound/profile/workload.cuh:8-45).
 * rows[max_levels][DFQ_PROFILE_FIELDS], one row per level:
 *   0 frontier vertices, 1 CSR entries read, 2 bitmap tests,
 *   3 atomic OR attempts, 4 vertices claimed, 5 work items of the grid
 *   phase (256 CSR entries each), 6 1 if the level ran bottom-up,
 *   7 wall ns of the level, 8 slowest workgroup's expansion ns,
 *   9 sum of the workgroups' ns,
 *   10-13 frontier vertices per degree bin (lane-owned, one wave each,
 *         one workgroup each, work items) on levels that go through the
 *         bins (0 on bottom-up levels and edge-balanced small frontiers),
 *   14-17 idle lane slots of those bins' last steps (64-lane steps;
 *         workgroup-size steps for the workgroup bin),
 *   18 CSR entries read by the busiest workgroup,
 *   19 unvisited vertices scanned by a bottom-up level.
 * Levels beyond max_levels are not recorded.  The counters slow the search;
 * time it with dfq_bfs_device instead. */
#define DFQ_PROFILE_FIELDS 20
int dfq_bfs_profile(dfq_graph *g, uint32_t source, uint64_t *rows,
                    uint32_t max_levels, dfq_stats *stats);

/* ---- device memory helpers ----------------------------------------------
 * No counterpart in the reference (it used cudaMallocManaged, bfs.cu:186,
 * 194-197).  They let callers that hold no HIP runtime of their own (the
 * ctypes tests and benchmark) own device buffers for dfq_bfs_device. */
int dfq_device_alloc(size_t bytes, void **d_ptr);
int dfq_device_free(void *d_ptr);
int dfq_copy_to_host(void *h_dst, const void *d_src, size_t bytes);
int dfq_copy_to_device(void *d_dst, const void *h_src, size_t bytes);
int dfq_device_synchronize(void);

/* Plain device-to-device copy of `bytes` (multiple of 16) through a
 * streaming HIP kernel, timed with hipEvents.  No counterpart in the
 * reference: it is the known-size workload against which HBM traffic
 * counters (FETCH_SIZE / WRITE_SIZE) of the BFS kernel are calibrated, and
 * it gives the attainable copy bandwidth on the running device. */
int dfq_stream_copy(void *d_dst, const void *d_src, size_t bytes, double *ms_out);

#ifdef __cplusplus
}
#endif
#endif /* DFQ_BFS_H */
