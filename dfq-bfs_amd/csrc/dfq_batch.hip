// Host side of the batched BFS (include/dfq_batch.h): resident graph set,
// query tables, tier plans per size bucket, hipGraph capture / replay of
// the one persistent launch, and the autotuner.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "dfq_batch.h"
#include "dfq_batch.hpp"
#include "host_common.hpp"

using namespace dfq;

#define BHIP(expr)                 \
  do {                             \
    if ((expr) != hipSuccess) return DFQ_ERR_HIP; \
  } while (0)

static_assert(DFQ_WG_MAX_VERTICES == kWgMaxN, "header and kernel disagree on the workgroup tier limit");

struct dfq_batch {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // graphs
  std::vector<BatchGraph> graphs;
  int64_t *d_off = nullptr;
  uint32_t *d_col = nullptr;
  BatchGraph *d_graphs = nullptr;
  uint32_t n_max = 0;
  uint64_t m_max = 0;
  // queries
  std::vector<uint32_t> q_graph, q_source;
  std::vector<uint64_t> lbase;
  uint64_t level_total = 0;
  uint32_t *d_level = nullptr;
  BatchQuery *d_queries = nullptr;
  BatchResult *d_results = nullptr;
  std::vector<uint32_t> dev_pos;  // query -> row of the device tables
  uint32_t table_cap = 0;
  // plans
  dfq_query_plan plan[32];
  uint32_t block = 1024;
  int grid = 0;
  // grid-tier scratch (largest graph), one copy per team, and
  // workgroup-tier queues
  uint32_t team_xcds = 0;   // XCDs per grid-tier team (0: the whole grid)
  uint32_t teams_cap = 0;   // scratch copies allocated
  uint32_t *d_vis = nullptr, *d_vis_snap = nullptr, *d_fb = nullptr, *d_qseg[2] = {nullptr, nullptr};
  uint2 *d_items = nullptr;
  uint32_t item_cap = 0;
  Control *d_ctl = nullptr;  // one per team, then the census's
  uint32_t *d_work = nullptr;
  uint32_t *d_wg_queue = nullptr;
  int wg_queue_grid = 0;
  uint32_t *d_iso = nullptr;  // isolated-vertex bitmaps of the symmetric graphs
  uint32_t *d_col_bu = nullptr;  // rows by neighbour degree (symmetric graphs' ranges)
  uint4 *d_head = nullptr;       // first four col_bu entries per vertex, indexed like d_off
  // launch state of the current tables
  BatchArgs args{};
  bool tables_ok = false;
  hipGraph_t hgraph = nullptr;
  hipGraphExec_t hexec = nullptr;
  bool graph_ok = false;
  bool graph_failed = false;
};

namespace {

const void *batch_kernel_for(uint32_t block) {
  return block == 512 ? (const void *)&dfq_batch_kernel<512> : (const void *)&dfq_batch_kernel<1024>;
}

uint32_t bucket_of(uint32_t n) {
  uint32_t b = 0;
  while ((2ull << b) <= n) ++b;
  return b;
}

void drop_graph_exec(dfq_batch *b) {
  if (b->hexec) (void)hipGraphExecDestroy(b->hexec);
  if (b->hgraph) (void)hipGraphDestroy(b->hgraph);
  b->hexec = nullptr;
  b->hgraph = nullptr;
  b->graph_ok = false;
}

void free_all(dfq_batch *b) {
  drop_graph_exec(b);
  void *ptrs[] = {b->d_off, b->d_col, b->d_graphs, b->d_level, b->d_queries, b->d_results,
                  b->d_vis, b->d_vis_snap, b->d_fb, b->d_qseg[0], b->d_qseg[1], b->d_items,
                  b->d_ctl, b->d_work, b->d_wg_queue, b->d_iso, b->d_col_bu, b->d_head};
  for (void *p : ptrs) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

// Grid-tier scratch for `copies` teams (each sized for the largest graph,
// laid out as dfq_batch_kernel strides it) and copies + 1 control blocks.
int alloc_team_scratch(dfq_batch *b, uint32_t copies) {
  void *old[] = {b->d_vis, b->d_vis_snap, b->d_fb, b->d_qseg[0], b->d_qseg[1], b->d_items, b->d_ctl};
  for (void *p : old) (void)hipFree(p);
  b->d_vis = b->d_vis_snap = b->d_fb = b->d_qseg[0] = b->d_qseg[1] = nullptr;
  b->d_items = nullptr;
  b->d_ctl = nullptr;
  b->teams_cap = 0;
  const size_t nw = ((size_t)b->n_max + 31) / 32, c = copies;
  BHIP(hipMalloc(&b->d_vis, sizeof(uint32_t) * nw * c));
  BHIP(hipMalloc(&b->d_vis_snap, sizeof(uint32_t) * nw * c));
  BHIP(hipMalloc(&b->d_fb, sizeof(uint32_t) * 2 * (nw + 1) * c));
  for (int i = 0; i < 2; ++i) BHIP(hipMalloc(&b->d_qseg[i], sizeof(uint32_t) * (size_t)kSegments * b->n_max * c));
  BHIP(hipMalloc(&b->d_items, sizeof(uint2) * (size_t)kSegments * b->item_cap * c));
  BHIP(hipMalloc(&b->d_ctl, sizeof(Control) * (c + 1)));
  BHIP(hipMemset(b->d_ctl, 0, sizeof(Control) * (c + 1)));
  b->teams_cap = copies;
  return DFQ_OK;
}

dfq_query_plan default_plan(uint32_t bucket) {
  dfq_query_plan p;
  memset(&p, 0, sizeof(p));
  p.tier = ((1ull << bucket) <= kWgMaxN / 2) ? DFQ_TIER_WORKGROUP : DFQ_TIER_GRID;
  p.small_max = 32;
  p.medium_max = 256;
  p.large_max = 256;
  p.direction_opt = 1;
  p.alpha = 40;
  p.beta = 24;
  return p;
}

int validate_plan(const dfq_query_plan &p) {
  if (p.tier > 1 || p.small_max > 64 || p.medium_max < p.small_max || p.large_max < p.medium_max ||
      p.direction_opt > 1 || (p.direction_opt && (p.alpha == 0 || p.beta == 0)) || p.reserved)
    return DFQ_ERR_INVALID_ARG;
  return DFQ_OK;
}

int ensure_shape(dfq_batch *b) {
  hipDeviceProp_t prop;
  BHIP(hipGetDeviceProperties(&prop, b->device));
  if (!prop.cooperativeLaunch) return DFQ_ERR_COOPERATIVE;
  int per_cu = 0;
  BHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_for(b->block), (int)b->block, 0));
  if (per_cu <= 0) return DFQ_ERR_COOPERATIVE;
  const int grid = prop.multiProcessorCount * per_cu;
  if (grid != b->wg_queue_grid) {
    (void)hipFree(b->d_wg_queue);
    b->d_wg_queue = nullptr;
    BHIP(hipMalloc(&b->d_wg_queue, sizeof(uint32_t) * (size_t)grid * 2 * kWgMaxN));
    b->wg_queue_grid = grid;
  }
  b->grid = grid;
  return DFQ_OK;
}

// Builds the device tables for the queries in `subset` (all when empty):
// grid-tier queries first, workgroup-tier ones by decreasing edge count.
int build_tables(dfq_batch *b, const std::vector<uint32_t> &subset) {
  const uint32_t nq = (uint32_t)b->q_graph.size();
  std::vector<uint32_t> idx;
  if (subset.empty()) {
    idx.resize(nq);
    for (uint32_t i = 0; i < nq; ++i) idx[i] = i;
  } else {
    idx = subset;
  }
  std::vector<uint32_t> grid_q, wg_q;
  for (uint32_t q : idx) {
    const BatchGraph &g = b->graphs[b->q_graph[q]];
    const dfq_query_plan &p = b->plan[bucket_of(g.n)];
    if (p.tier == DFQ_TIER_GRID || g.n > kWgMaxN) grid_q.push_back(q);
    else wg_q.push_back(q);
  }
  std::stable_sort(wg_q.begin(), wg_q.end(), [&](uint32_t x, uint32_t y) {
    return b->graphs[b->q_graph[x]].m > b->graphs[b->q_graph[y]].m;
  });
  std::vector<BatchQuery> table;
  table.reserve(idx.size());
  b->dev_pos.assign(nq, 0xffffffffu);
  for (int pass = 0; pass < 2; ++pass)
    for (uint32_t q : pass == 0 ? grid_q : wg_q) {
      const BatchGraph &g = b->graphs[b->q_graph[q]];
      const dfq_query_plan &p = b->plan[bucket_of(g.n)];
      BatchQuery t;
      memset(&t, 0, sizeof(t));
      t.lbase = b->lbase[q];
      t.graph = b->q_graph[q];
      t.source = b->q_source[q];
      t.tier = pass == 0 ? 1u : 0u;
      t.small_max = p.small_max;
      t.medium_max = p.medium_max;
      t.large_max = p.large_max;
      t.direction_opt = p.direction_opt;
      t.alpha = p.alpha;
      t.beta = p.beta;
      b->dev_pos[q] = (uint32_t)table.size();
      table.push_back(t);
    }
  if (!table.empty())
    BHIP(hipMemcpy(b->d_queries, table.data(), sizeof(BatchQuery) * table.size(), hipMemcpyHostToDevice));

  BatchArgs &a = b->args;
  memset(&a, 0, sizeof(a));
  a.off = b->d_off;
  a.col = b->d_col;
  a.graphs = b->d_graphs;
  a.queries = b->d_queries;
  a.results = b->d_results;
  a.level = b->d_level;
  a.n_queries = (uint32_t)table.size();
  a.n_grid = (uint32_t)grid_q.size();
  a.work = b->d_work;
  a.wg_queue = b->d_wg_queue;
  a.iso = b->d_iso;
  a.col_bu = b->d_col_bu;
  a.head = b->d_head;
  BfsArgs &s = a.grid;
  s.vis = b->d_vis;
  s.vis_snap = b->d_vis_snap;
  s.fb = b->d_fb;
  s.qseg[0] = b->d_qseg[0];
  s.qseg[1] = b->d_qseg[1];
  s.items = b->d_items;
  s.item_cap = b->item_cap;
  s.ctl = b->d_ctl;
  a.xpt = b->team_xcds;
  a.teams_cap = b->teams_cap;
  a.n_max = b->n_max;
  s.prof = nullptr;
  s.prof_levels = 0;
  s.result = nullptr;
  return DFQ_OK;
}

int launch(dfq_batch *b, bool use_graph, dfq_batch_stats *st) {
  BHIP(hipSetDevice(b->device));
  void *kargs[] = {&b->args};
  bool replayed = false;
  if (use_graph && !b->graph_failed) {
    if (!b->graph_ok) {
      drop_graph_exec(b);
      // capture the cooperative launch; if this runtime cannot, fall back to
      // direct launches for the life of the batch
      bool ok = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
      if (ok) {
        const hipError_t le = hipLaunchCooperativeKernel(batch_kernel_for(b->block), dim3(b->grid),
                                                         dim3(b->block), kargs, 0, b->stream);
        const hipError_t ce = hipStreamEndCapture(b->stream, &b->hgraph);
        ok = le == hipSuccess && ce == hipSuccess && b->hgraph;
        if (ok) ok = hipGraphInstantiate(&b->hexec, b->hgraph, nullptr, nullptr, 0) == hipSuccess;
      }
      if (!ok) {
        (void)hipGetLastError();
        drop_graph_exec(b);
        b->graph_failed = true;
      } else {
        b->graph_ok = true;
      }
    }
    if (b->graph_ok) {
      BHIP(hipEventRecord(b->ev0, b->stream));
      BHIP(hipGraphLaunch(b->hexec, b->stream));
      BHIP(hipEventRecord(b->ev1, b->stream));
      replayed = true;
    }
  }
  if (!replayed) {
    BHIP(hipEventRecord(b->ev0, b->stream));
    BHIP(hipLaunchCooperativeKernel(batch_kernel_for(b->block), dim3(b->grid), dim3(b->block), kargs, 0,
                                    b->stream));
    BHIP(hipEventRecord(b->ev1, b->stream));
  }
  BHIP(hipEventSynchronize(b->ev1));
  uint32_t err = 0;
  BHIP(hipMemcpy(&err, &b->d_ctl->err, sizeof(err), hipMemcpyDeviceToHost));
  if (err) return DFQ_ERR_CAPACITY;
  if (st) {
    float ms = 0.f;
    BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    st->kernel_ms = ms;
    st->queries = b->args.n_queries;
    st->grid_queries = b->args.n_grid;
    st->wg_queries = b->args.n_queries - b->args.n_grid;
    st->graph_replay = replayed ? 1u : 0u;
    st->block_size = b->block;
    st->grid_blocks = (uint32_t)b->grid;
    st->team_xcds = b->team_xcds;
    st->reserved = 0;
  }
  return DFQ_OK;
}

int run_subset(dfq_batch *b, const std::vector<uint32_t> &subset, dfq_batch_stats *st) {
  drop_graph_exec(b);
  b->tables_ok = false;
  int rc = build_tables(b, subset);
  if (rc != DFQ_OK) return rc;
  return launch(b, false, st);
}

}  // namespace

extern "C" {

int dfq_batch_create(uint32_t num_graphs, const uint32_t *n, const uint64_t *m, const int64_t *row_offsets,
                     const uint32_t *col_idx, dfq_batch **out) {
  if (!out || num_graphs == 0 || !n || !m || !row_offsets) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  uint64_t vtot = 0, etot = 0;
  std::vector<BatchGraph> graphs(num_graphs);
  for (uint32_t g = 0; g < num_graphs; ++g) {
    if (n[g] == 0 || n[g] == 0xffffffffu) return DFQ_ERR_INVALID_ARG;
    graphs[g].vbase = vtot;
    graphs[g].ebase = etot;
    graphs[g].n = n[g];
    graphs[g].m = m[g];
    vtot += (uint64_t)n[g] + 1;
    etot += m[g];
  }
  if (etot && !col_idx) return DFQ_ERR_INVALID_ARG;
  // same input checks as dfq_graph_create, per graph
  for (uint32_t g = 0; g < num_graphs; ++g) {
    const int64_t *off = row_offsets + graphs[g].vbase;
    const uint32_t *col = col_idx + graphs[g].ebase;
    if (m[g] > kMaxEntries) return DFQ_ERR_RANGE;
    if (off[0] != 0 || (uint64_t)off[n[g]] != m[g]) return DFQ_ERR_RANGE;
    std::atomic<bool> bad{false};
    parallel_for(n[g], [&](uint64_t bb, uint64_t e, unsigned) {
      for (uint64_t v = bb; v < e; ++v)
        if (off[v + 1] < off[v]) {
          bad = true;
          return;
        }
    });
    parallel_for(m[g], [&](uint64_t bb, uint64_t e, unsigned) {
      for (uint64_t i = bb; i < e; ++i)
        if (col[i] >= n[g]) {
          bad = true;
          return;
        }
    });
    if (bad) return DFQ_ERR_RANGE;
    graphs[g].symmetric = is_symmetric(off, col, n[g]) ? 1u : 0u;
  }
  uint64_t iso_words = 0;
  for (auto &gr : graphs) {
    gr.iso_base = gr.symmetric ? iso_words : kNoIso;
    if (gr.symmetric) iso_words += ((uint64_t)gr.n + 31) / 32;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DFQ_ERR_NO_DEVICE;

  dfq_batch *b = new (std::nothrow) dfq_batch();
  if (!b) return DFQ_ERR_NOMEM;
  auto fail = [&](int rc) {
    free_all(b);
    delete b;
    return rc;
  };
#define BFAIL(expr)                                  \
  do {                                               \
    if ((expr) != hipSuccess) return fail(DFQ_ERR_HIP); \
  } while (0)
  BFAIL(hipGetDevice(&b->device));
  b->graphs = graphs;
  for (auto &g : graphs) {
    b->n_max = std::max(b->n_max, g.n);
    b->m_max = std::max(b->m_max, g.m);
  }
  for (uint32_t k = 0; k < 32; ++k) b->plan[k] = default_plan(k);
  BFAIL(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BFAIL(hipEventCreate(&b->ev0));
  BFAIL(hipEventCreate(&b->ev1));
  BFAIL(hipMalloc(&b->d_off, sizeof(int64_t) * vtot));
  // +4 entries: bottom-up levels read columns as aligned 16-byte blocks
  BFAIL(hipMalloc(&b->d_col, sizeof(uint32_t) * (etot + 4)));
  BFAIL(hipMalloc(&b->d_col_bu, sizeof(uint32_t) * (etot + 4)));
  BFAIL(hipMalloc(&b->d_graphs, sizeof(BatchGraph) * num_graphs));
  BFAIL(hipMemcpy(b->d_off, row_offsets, sizeof(int64_t) * vtot, hipMemcpyHostToDevice));
  {
    // symmetric graphs: a second copy of the rows by neighbour degree for
    // bottom-up levels, and isolated vertices marked visited from the start
    std::vector<uint32_t> iso(iso_words ? iso_words : 1), ordered, heads;
    BFAIL(hipMalloc(&b->d_head, sizeof(uint4) * vtot));
    for (uint32_t g = 0; g < num_graphs; ++g) {
      const BatchGraph &gr = graphs[g];
      const int64_t *off = row_offsets + gr.vbase;
      const uint32_t *col = col_idx + gr.ebase;
      if (gr.m) BFAIL(hipMemcpy(b->d_col + gr.ebase, col, sizeof(uint32_t) * gr.m, hipMemcpyHostToDevice));
      if (!gr.symmetric) continue;
      ordered.resize(gr.m);
      order_rows_by_degree(off, col, gr.n, ordered.data());
      isolated_bitmap(off, gr.n, iso.data() + gr.iso_base);
      if (gr.m)
        BFAIL(hipMemcpy(b->d_col_bu + gr.ebase, ordered.data(), sizeof(uint32_t) * gr.m, hipMemcpyHostToDevice));
      heads.resize(4 * (size_t)gr.n);
      row_heads(off, ordered.data(), gr.n, heads.data());
      if (gr.n) BFAIL(hipMemcpy(b->d_head + gr.vbase, heads.data(), sizeof(uint4) * gr.n, hipMemcpyHostToDevice));
    }
    BFAIL(hipMalloc(&b->d_iso, sizeof(uint32_t) * iso.size()));
    BFAIL(hipMemcpy(b->d_iso, iso.data(), sizeof(uint32_t) * iso.size(), hipMemcpyHostToDevice));
  }
  BFAIL(hipMemcpy(b->d_graphs, graphs.data(), sizeof(BatchGraph) * num_graphs, hipMemcpyHostToDevice));
  // grid-tier scratch sized for the largest graph (one team: the grid)
  b->item_cap = (uint32_t)std::min<uint64_t>(b->m_max / kItem + b->n_max, 0xffffffffull);
  if (alloc_team_scratch(b, 1) != DFQ_OK) return fail(DFQ_ERR_HIP);
  // the workgroup tier's instance counter (any start value: each launch
  // counts from the value it finds)
  BFAIL(hipMalloc(&b->d_work, sizeof(uint32_t)));
#undef BFAIL
  int rc = ensure_shape(b);
  if (rc != DFQ_OK) return fail(rc);
  *out = b;
  return DFQ_OK;
}

void dfq_batch_destroy(dfq_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_all(b);
  delete b;
}

int dfq_batch_set_queries(dfq_batch *b, uint32_t nq, const uint32_t *graph, const uint32_t *source) {
  if (!b || (nq && (!graph || !source))) return DFQ_ERR_INVALID_ARG;
  for (uint32_t q = 0; q < nq; ++q) {
    if (graph[q] >= b->graphs.size()) return DFQ_ERR_RANGE;
    if (source[q] >= b->graphs[graph[q]].n) return DFQ_ERR_RANGE;
  }
  BHIP(hipSetDevice(b->device));
  drop_graph_exec(b);
  b->tables_ok = false;
  b->q_graph.assign(graph, graph + nq);
  b->q_source.assign(source, source + nq);
  b->lbase.resize(nq);
  uint64_t tot = 0;
  for (uint32_t q = 0; q < nq; ++q) {
    b->lbase[q] = tot;
    tot += b->graphs[graph[q]].n;
  }
  if (tot > b->level_total) {
    (void)hipFree(b->d_level);
    b->d_level = nullptr;
    b->level_total = 0;
    BHIP(hipMalloc(&b->d_level, sizeof(uint32_t) * (tot ? tot : 1)));
    b->level_total = tot;
  }
  if (nq > b->table_cap) {
    (void)hipFree(b->d_queries);
    (void)hipFree(b->d_results);
    b->d_queries = nullptr;
    b->d_results = nullptr;
    b->table_cap = 0;
    BHIP(hipMalloc(&b->d_queries, sizeof(BatchQuery) * nq));
    BHIP(hipMalloc(&b->d_results, sizeof(BatchResult) * nq));
    b->table_cap = nq;
  }
  return DFQ_OK;
}

int dfq_batch_set_plan(dfq_batch *b, uint32_t bucket, const dfq_query_plan *plan) {
  if (!b || !plan || bucket >= 32) return DFQ_ERR_INVALID_ARG;
  int rc = validate_plan(*plan);
  if (rc != DFQ_OK) return rc;
  b->plan[bucket] = *plan;
  drop_graph_exec(b);
  b->tables_ok = false;
  return DFQ_OK;
}

int dfq_batch_get_plan(const dfq_batch *b, uint32_t bucket, dfq_query_plan *plan) {
  if (!b || !plan || bucket >= 32) return DFQ_ERR_INVALID_ARG;
  *plan = b->plan[bucket];
  return DFQ_OK;
}

int dfq_batch_set_team_xcds(dfq_batch *b, uint32_t xcds_per_team) {
  if (!b || (xcds_per_team != 0 && xcds_per_team != 1 && xcds_per_team != 2 && xcds_per_team != 4 &&
             xcds_per_team != 8))
    return DFQ_ERR_INVALID_ARG;
  BHIP(hipSetDevice(b->device));
  // teams the launch can form: the device's XCDs (the census's bound when
  // the runtime does not say) over the team width
  int xcc = 0;
  if (hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, b->device) != hipSuccess || xcc <= 0 ||
      xcc > kMaxXcd) {
    (void)hipGetLastError();
    xcc = kMaxXcd;
  }
  const uint32_t copies = xcds_per_team ? ((uint32_t)xcc + xcds_per_team - 1) / xcds_per_team : 1u;
  if (copies != b->teams_cap) {
    const int rc = alloc_team_scratch(b, copies);
    if (rc != DFQ_OK) {
      // keep a usable batch: the whole-grid layout
      b->team_xcds = 0;
      if (alloc_team_scratch(b, 1) != DFQ_OK) return DFQ_ERR_HIP;
      drop_graph_exec(b);
      b->tables_ok = false;
      return DFQ_ERR_NOMEM;
    }
  }
  b->team_xcds = xcds_per_team;
  drop_graph_exec(b);
  b->tables_ok = false;
  return DFQ_OK;
}

int dfq_batch_set_block_size(dfq_batch *b, uint32_t block_size) {
  if (!b || (block_size != 512 && block_size != 1024)) return DFQ_ERR_INVALID_ARG;
  BHIP(hipSetDevice(b->device));
  BHIP(hipStreamSynchronize(b->stream));
  const uint32_t old = b->block;
  b->block = block_size;
  int rc = ensure_shape(b);
  if (rc != DFQ_OK) {
    b->block = old;
    return rc;
  }
  drop_graph_exec(b);
  b->tables_ok = false;
  return DFQ_OK;
}

int dfq_batch_run(dfq_batch *b, dfq_batch_stats *stats) {
  if (!b) return DFQ_ERR_INVALID_ARG;
  if (b->q_graph.empty()) return DFQ_ERR_STATE;
  BHIP(hipSetDevice(b->device));
  if (!b->tables_ok) {
    int rc = build_tables(b, {});
    if (rc != DFQ_OK) return rc;
    b->tables_ok = true;
  }
  return launch(b, true, stats);
}

int dfq_batch_levels(const dfq_batch *b, uint32_t query, uint32_t *level_out) {
  if (!b || !level_out) return DFQ_ERR_INVALID_ARG;
  if (query >= b->q_graph.size()) return DFQ_ERR_RANGE;
  BHIP(hipSetDevice(b->device));
  BHIP(hipMemcpy(level_out, b->d_level + b->lbase[query],
                 sizeof(uint32_t) * b->graphs[b->q_graph[query]].n, hipMemcpyDeviceToHost));
  return DFQ_OK;
}

int dfq_batch_result(const dfq_batch *b, uint32_t query, uint32_t *depth, uint64_t *reached) {
  if (!b) return DFQ_ERR_INVALID_ARG;
  if (query >= b->q_graph.size() || b->dev_pos.size() != b->q_graph.size() ||
      b->dev_pos[query] == 0xffffffffu)
    return DFQ_ERR_RANGE;
  BHIP(hipSetDevice(b->device));
  BatchResult r;
  BHIP(hipMemcpy(&r, b->d_results + b->dev_pos[query], sizeof(r), hipMemcpyDeviceToHost));
  if (depth) *depth = r.depth;
  if (reached) *reached = r.reached;
  return DFQ_OK;
}

int dfq_batch_autotune(dfq_batch *b, uint32_t repeats, dfq_tune_report *report) {
  if (!b || b->q_graph.empty()) return DFQ_ERR_STATE;
  if (repeats == 0) repeats = 1;
  BHIP(hipSetDevice(b->device));
  dfq_tune_report rep;
  memset(&rep, 0, sizeof(rep));
  // queries per bucket
  std::vector<std::vector<uint32_t>> per(32);
  for (uint32_t q = 0; q < b->q_graph.size(); ++q)
    per[bucket_of(b->graphs[b->q_graph[q]].n)].push_back(q);
  std::vector<dfq_query_plan> cands;
  auto time_plan = [&](uint32_t k, const dfq_query_plan &p, double &ms) -> int {
    b->plan[k] = p;
    double best = 1e30;
    for (uint32_t r = 0; r < repeats + 1; ++r) {  // first run warms up
      dfq_batch_stats st;
      int rc = run_subset(b, per[k], &st);
      if (rc != DFQ_OK) return rc;
      if (r > 0 || repeats == 0) best = std::min(best, st.kernel_ms);
    }
    ms = best;
    return DFQ_OK;
  };
  const double G = (double)b->grid;
  double total_cost = 0;
  std::vector<double> wg_single(32, 0.0);
  for (uint32_t k = 0; k < 32; ++k) {
    if (per[k].empty()) continue;
    const double nq = (double)per[k].size();
    cands.clear();
    for (uint32_t tier : {DFQ_TIER_WORKGROUP, DFQ_TIER_GRID}) {
      if (tier == DFQ_TIER_WORKGROUP && (1ull << k) > kWgMaxN) continue;
      for (uint32_t med : {64u, 256u, 1024u})
        for (uint32_t dir : {0u, 1u}) {
          dfq_query_plan p = default_plan(k);
          p.tier = tier;
          p.medium_max = med;
          p.large_max = tier == DFQ_TIER_GRID ? med : 32768u;
          p.direction_opt = dir;
          cands.push_back(p);
        }
    }
    double best_cost = 1e30, best_ms = 0;
    dfq_query_plan best = b->plan[k];
    for (const auto &p : cands) {
      double ms = 0;
      int rc = time_plan(k, p, ms);
      if (rc != DFQ_OK) return rc;
      ++rep.candidates;
      // workgroup-milliseconds per query: the workgroup tier occupies
      // min(queries, grid) workgroups for the sub-batch, the grid tier all
      const double busy = p.tier == DFQ_TIER_GRID ? G : std::min(nq, G);
      const double cost = ms * busy / nq;
      if (cost < best_cost) {
        best_cost = cost;
        best_ms = ms;
        best = p;
      }
    }
    b->plan[k] = best;
    total_cost += best_cost * nq;
    // one query of this bucket on one workgroup takes about this long
    wg_single[k] = best.tier == DFQ_TIER_WORKGROUP ? best_ms / std::ceil(nq / G) : 0.0;
    dfq_tune_entry &e = rep.entry[rep.entries++];
    e.bucket = k;
    e.queries = (uint32_t)per[k].size();
    e.plan = best;
    e.cost_us = best_cost * 1000.0;
    e.sub_batch_ms = best_ms;
  }
  // a workgroup-tier query longer than a quarter of the whole batch would
  // be the tail of the launch: run that bucket on the grid tier instead
  const double makespan = total_cost / G;
  for (uint32_t i = 0; i < rep.entries; ++i) {
    dfq_tune_entry &e = rep.entry[i];
    if (e.plan.tier == DFQ_TIER_WORKGROUP && wg_single[e.bucket] > 0.25 * makespan) {
      dfq_query_plan p = e.plan;
      p.tier = DFQ_TIER_GRID;
      p.large_max = p.medium_max;
      double ms = 0;
      int rc = time_plan(e.bucket, p, ms);
      if (rc != DFQ_OK) return rc;
      ++rep.candidates;
      b->plan[e.bucket] = p;
      e.plan = p;
      e.sub_batch_ms = ms;
      e.cost_us = ms * G / e.queries * 1000.0;
    }
  }
  // workgroup size on the whole batch
  double best_ms = 1e30;
  uint32_t best_block = b->block;
  for (uint32_t blk : {1024u, 512u}) {
    int rc = dfq_batch_set_block_size(b, blk);
    if (rc != DFQ_OK) continue;
    double ms = 1e30;
    for (uint32_t r = 0; r < repeats + 1; ++r) {
      dfq_batch_stats st;
      rc = run_subset(b, {}, &st);
      if (rc != DFQ_OK) return rc;
      if (r > 0) ms = std::min(ms, st.kernel_ms);
    }
    ++rep.candidates;
    if (ms < best_ms) {
      best_ms = ms;
      best_block = blk;
    }
  }
  int rc = dfq_batch_set_block_size(b, best_block);
  if (rc != DFQ_OK) return rc;
  rep.block_size = best_block;
  // team width of the grid tier on the whole batch (several grid-tier
  // queries at once on narrower teams)
  best_ms = 1e30;
  uint32_t best_x = b->team_xcds;
  for (uint32_t x : {0u, 4u, 2u, 1u}) {
    rc = dfq_batch_set_team_xcds(b, x);
    if (rc != DFQ_OK) continue;
    double ms = 1e30;
    for (uint32_t r = 0; r < repeats + 1; ++r) {
      dfq_batch_stats st;
      rc = run_subset(b, {}, &st);
      if (rc != DFQ_OK) return rc;
      if (r > 0) ms = std::min(ms, st.kernel_ms);
    }
    ++rep.candidates;
    if (ms < best_ms) {
      best_ms = ms;
      best_x = x;
    }
  }
  rc = dfq_batch_set_team_xcds(b, best_x);
  if (rc != DFQ_OK) return rc;
  rep.team_xcds = best_x;
  drop_graph_exec(b);
  b->tables_ok = false;
  if (report) *report = rep;
  return DFQ_OK;
}

}  // extern "C"
