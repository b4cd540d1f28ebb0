// Host side of the BFS server (include/dfq_serve.h): resident graphs, the
// working-set arena and its admission rule, the request ring in pinned
// host-mapped memory, the persistent launch, and the serving loop that
// replays an arrival trace.
//
// Ring protocol.  Ticket T goes to ring slot T mod R.  The host fills the
// request, clears the slot's result word, then stores seq = t + 1 (t = T
// minus the launch's first ticket) with release order.  On the device the
// next idle team takes t from a counter, waits for seq, runs the query and
// stores done = t + 1 after a system-scope write-back of the outputs.  A
// slot is reused only after its query has been released, and the host ends
// the launch only after every published query is done, by setting `stop`.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <iterator>
#include <map>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "dfq_serve.h"
#include "dfq_serve.hpp"
#include "host_common.hpp"

using namespace dfq;

static_assert(DFQ_SERVE_PARENTS == kServeParents, "flag mismatch");
static_assert(sizeof(ServeRequest) == 64 && sizeof(ServeResult) == 64, "ring entries are one cache line");

#define SHIP(expr)                                 \
  do {                                             \
    if ((expr) != hipSuccess) return DFQ_ERR_HIP;  \
  } while (0)

namespace {

using Clock = std::chrono::steady_clock;

double us_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
}

// First-fit allocator over [0, size) with coalescing of neighbours.
class Arena {
 public:
  void reset(uint64_t size) {
    free_.clear();
    if (size) free_[0] = size;
    size_ = size;
    used_ = 0;
  }
  bool alloc(uint64_t bytes, uint64_t *off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const uint64_t rest = it->second - bytes, start = it->first + bytes;
      free_.erase(it);
      if (rest) free_[start] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    used_ -= bytes;
    uint64_t start = off, len = bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && next->first == off + bytes) {
      len += next->second;
      free_.erase(next);
    }
    free_[start] = len;
  }
  uint64_t size() const { return size_; }
  uint64_t used() const { return used_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t size_ = 0, used_ = 0;
};

// A ring slot is busy from publication until the host has read the device's
// result; admitted queries are tracked apart, so a finished query the caller
// has not released holds arena room but no slot.
struct RingSlot {
  uint64_t ticket = 0;  // ring ticket of the task in this slot
  bool busy = false;
  bool copy = false;    // an internal copy task
  uint64_t qid0 = 0;    // its queries: qid0 .. qid0 + count - 1
  uint32_t count = 0;
};

// One admitted query.  Tickets handed to callers are query ids; a query
// runs in the ring task `ring_ticket`, alone or as member `member` of a
// group task whose arena range it shares (`group`).
struct Query {
  uint64_t work = 0, bytes = 0;  // arena range (the group's for members)
  uint64_t ring_ticket = 0;
  uint64_t group = 0;            // group id, 0 for a single query
  uint32_t member = 0;
  uint32_t graph = 0, source = 0, flags = 0;
  uint32_t item_cap = 0;
  uint32_t count = 1;            // searches of its task
  bool done = false;
  int status = DFQ_OK;           // DFQ_ERR_CAPACITY: the group went past its level limit
  dfq_serve_result result{};
};

struct GroupRange {
  uint64_t work = 0, bytes = 0;
  uint32_t held = 0;  // members not yet released
};

constexpr uint32_t kDrainMs = 120000;  // stop: wait for published queries
constexpr uint32_t kEndMs = 30000;     // stop: wait for the launch to end
constexpr uint32_t kCopyMs = 30000;    // level copies
constexpr size_t kStageBytes = 32u << 20;

}  // namespace

struct dfq_server {
  int device = -1;
  hipStream_t stream = nullptr;       // the persistent launch
  std::vector<ServeGraph> graphs;
  std::vector<dfq_query_plan> plans;
  int64_t *d_off = nullptr;
  uint32_t *d_col = nullptr;
  ServeGraph *d_graphs = nullptr;
  uint32_t *d_iso = nullptr;
  uint32_t *d_col_bu = nullptr;  // rows by neighbour degree (symmetric graphs' ranges)
  uint4 *d_head = nullptr;       // first four col_bu entries per vertex, indexed like d_off
  ServeCtl *d_ctl = nullptr;
  uint8_t *d_arena = nullptr;
  // pinned, host-mapped
  void *h_block = nullptr;
  ServeRequest *h_req = nullptr;
  ServeResult *h_res = nullptr;
  uint32_t *h_stop = nullptr;
  ServeStatus *h_status = nullptr;
  ServeArgs args{};
  uint8_t *h_stage = nullptr;  // pinned, host-mapped staging of copy tasks
  uint8_t *d_stage = nullptr;  // its device address
  uint32_t *h_gsrc = nullptr;  // pinned, host-mapped [ring][64] group sources
  ServeGroupResult *h_gres = nullptr;  // pinned, host-mapped [ring][64] group results
  int64_t *d_off_r = nullptr;  // degree-ranked copies (group mode), indexed like d_off
  uint32_t *d_col_r = nullptr;
  uint32_t *d_o2r = nullptr;
  uint64_t next_qid = 1;       // query ids (tickets); 0 is never one
  uint64_t next_group = 1;
  std::unordered_map<uint64_t, GroupRange> groups_held;
  dfq_serve_config cfg{};
  int grid = 0;
  bool running = false;
  bool wedged = false;  // a launch that did not end: memory is never freed
  uint64_t next_ticket = 0;
  uint64_t base = 0;    // first ticket of the current launch
  std::vector<RingSlot> ring;
  std::unordered_map<uint64_t, Query> queries;  // admitted, not yet released
  Arena arena;
};

namespace {

// Work-item slots per segment a group task's levels may use.
uint32_t group_item_cap(const ServeGraph &g, const dfq_query_plan &p) {
  return (uint32_t)std::min<uint64_t>(ms_item_cap(g.n, g.m, p.large_max), 0xffffffffull);
}

uint64_t group_footprint(const ServeGraph &g, const dfq_query_plan &p, uint32_t xpt, uint32_t count) {
  return serve_group_layout(g.n, count, group_item_cap(g, p), serve_segments(xpt)).bytes;
}

const void *serve_kernel_for(uint32_t block, bool groups) {
  if (groups)
    return block == 512 ? (const void *)&dfq_serve_kernel<512, true> : (const void *)&dfq_serve_kernel<1024, true>;
  return block == 512 ? (const void *)&dfq_serve_kernel<512, false> : (const void *)&dfq_serve_kernel<1024, false>;
}

dfq_query_plan default_serve_plan() {
  dfq_query_plan p;
  memset(&p, 0, sizeof(p));
  p.tier = DFQ_TIER_GRID;
  p.small_max = 32;
  p.medium_max = 256;
  p.large_max = 256;
  p.direction_opt = 1;
  p.alpha = 40;
  p.beta = 24;
  return p;
}

bool plan_ok(const dfq_query_plan &p) {
  return p.small_max <= 64 && p.medium_max >= p.small_max && p.large_max >= p.medium_max &&
         p.direction_opt <= 1 && (!p.direction_opt || (p.alpha && p.beta)) && !p.reserved;
}

bool shape_ok(uint32_t xpt, uint32_t block) {
  return (xpt == 0 || xpt == 1 || xpt == 2 || xpt == 4 || xpt == 8) && (block == 512 || block == 1024);
}

uint64_t footprint_of(const ServeGraph &g, const dfq_query_plan &p, uint32_t xpt, uint32_t flags,
                      uint32_t *item_cap) {
  const uint32_t cap = serve_item_cap(g.n, g.m, p.large_max);
  if (item_cap) *item_cap = cap;
  return serve_layout(g.n, cap, serve_segments(xpt), (flags & kServeParents) != 0).bytes;
}

bool kernel_ended(dfq_server *s) { return hipStreamQuery(s->stream) != hipErrorNotReady; }

void free_all(dfq_server *s) {
  if (s->wedged) return;  // a launch may still read this memory
  void *dev[] = {s->d_off,  s->d_col, s->d_col_bu, s->d_head,  s->d_graphs, s->d_iso,
                 s->d_ctl,  s->d_arena, s->d_off_r, s->d_col_r, s->d_o2r};
  for (void *p : dev) (void)hipFree(p);
  if (s->h_block) (void)hipHostFree(s->h_block);
  if (s->h_stage) (void)hipHostFree(s->h_stage);
  if (s->h_gres) (void)hipHostFree(s->h_gres);
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

// Reads the device result of ring slot idx if it is there; frees the slot.
// Returns true when the slot is not (or no longer) busy.
bool collect(dfq_server *s, uint32_t idx) {
  RingSlot &rs = s->ring[idx];
  if (!rs.busy) return true;
  const ServeResult *r = &s->h_res[idx];
  if (__atomic_load_n(&r->done, __ATOMIC_ACQUIRE) != rs.ticket - s->base + 1) return false;
  if (!rs.copy) {
    for (uint32_t i = 0; i < rs.count; ++i) {
      auto it = s->queries.find(rs.qid0 + i);
      if (it == s->queries.end()) continue;
      Query &q = it->second;
      q.result.graph = q.graph;
      q.result.source = q.source;
      q.result.bottom_up_levels = r->bu_levels;
      q.result.team = r->team;
      q.result.service_us = (double)(r->t_end_ns - r->t_start_ns) / 1000.0;
      if (q.group) {
        const ServeGroupResult &g = s->h_gres[(size_t)idx * kMsMaxSources + q.member];
        q.result.depth = __atomic_load_n(&g.depth, __ATOMIC_RELAXED);
        q.result.reached = __atomic_load_n(&g.reached, __ATOMIC_RELAXED);
        q.result.edges = __atomic_load_n(&g.edges, __ATOMIC_RELAXED);
        if (r->err) q.status = DFQ_ERR_CAPACITY;
      } else {
        q.result.depth = r->depth;
        q.result.reached = r->reached;
        q.result.edges = r->edges;
      }
      q.done = true;
    }
  }
  rs.busy = false;
  rs.copy = false;
  return true;
}

// The admitted query of a ticket, its result collected if available.
Query *find(dfq_server *s, uint64_t ticket) {
  auto it = s->queries.find(ticket);
  if (it == s->queries.end()) return nullptr;
  if (!it->second.done) {
    const uint64_t rt = it->second.ring_ticket;
    const uint32_t idx = (uint32_t)(rt % s->cfg.ring_slots);
    if (s->ring[idx].busy && s->ring[idx].ticket == rt) collect(s, idx);
  }
  return &it->second;
}

// Gives a finished query's arena share back.
void release_query(dfq_server *s, std::unordered_map<uint64_t, Query>::iterator it) {
  Query &q = it->second;
  if (q.group) {
    auto g = s->groups_held.find(q.group);
    if (g != s->groups_held.end() && --g->second.held == 0) {
      s->arena.release(g->second.work, g->second.bytes);
      s->groups_held.erase(g);
    }
  } else {
    s->arena.release(q.work, q.bytes);
  }
  s->queries.erase(it);
}

bool any_running(dfq_server *s) {
  for (uint32_t i = 0; i < s->cfg.ring_slots; ++i)
    if (!collect(s, i)) return true;
  return false;
}

// Fills and publishes the next ticket's slot (which must not be busy);
// its queries are qid0 .. qid0 + count - 1.
uint64_t publish(dfq_server *s, const ServeRequest &req, bool copy, uint64_t qid0 = 0, uint32_t count = 0) {
  const uint64_t T = s->next_ticket++;
  const uint32_t idx = (uint32_t)(T % s->cfg.ring_slots);
  ServeRequest *r = &s->h_req[idx];
  r->work = req.work;
  r->dst = req.dst;
  r->graph = req.graph;
  r->source = req.source;
  r->flags = req.flags;
  r->item_cap = req.item_cap;
  r->small_max = req.small_max;
  r->medium_max = req.medium_max;
  r->large_max = req.large_max;
  r->alpha = req.alpha;
  r->beta = req.beta;
  __atomic_store_n(&s->h_res[idx].done, 0ull, __ATOMIC_RELAXED);
  __atomic_store_n(&r->seq, T - s->base + 1, __ATOMIC_RELEASE);
  s->ring[idx].ticket = T;
  s->ring[idx].busy = true;
  s->ring[idx].copy = copy;
  s->ring[idx].qid0 = qid0;
  s->ring[idx].count = count;
  return T;
}

// Runs one copy task on the next ticket and waits for it.  The slot may
// still hold a query in flight: it finishes on its own, so wait for it.
int run_copy_task(dfq_server *s, uint64_t src_off, uint64_t words) {
  const uint32_t idx = (uint32_t)(s->next_ticket % s->cfg.ring_slots);
  const auto t0 = Clock::now();
  while (!collect(s, idx)) {
    if (kernel_ended(s)) return DFQ_ERR_STATE;
    if (us_since(t0) > kCopyMs * 1000.0) return DFQ_ERR_TIMEOUT;
    std::this_thread::yield();
  }
  ServeRequest req;
  memset(&req, 0, sizeof(req));
  req.work = src_off;
  req.dst = (uint64_t)(uintptr_t)s->d_stage;
  req.source = (uint32_t)words;
  req.flags = kServeCopy;
  publish(s, req, true);
  const auto t1 = Clock::now();
  for (uint64_t k = 0; !collect(s, idx); ++k) {
    if ((k & 255) == 255) {
      // a task left busy is collected by a later call or by the stop drain
      if (kernel_ended(s)) return DFQ_ERR_STATE;
      if (us_since(t1) > kCopyMs * 1000.0) return DFQ_ERR_TIMEOUT;
    }
  }
  return DFQ_OK;
}

// Copies `bytes` (a multiple of 4) from arena offset src_off to host memory:
// through copy tasks while the launch runs, by DMA otherwise.
int copy_out(dfq_server *s, uint64_t src_off, uint32_t *dst, uint64_t bytes) {
  SHIP(hipSetDevice(s->device));
  if (!s->running) {
    SHIP(hipMemcpy(dst, s->d_arena + src_off, bytes, hipMemcpyDeviceToHost));
    return DFQ_OK;
  }
  for (uint64_t done = 0; done < bytes;) {
    const uint64_t chunk = std::min<uint64_t>(bytes - done, kStageBytes);
    const int rc = run_copy_task(s, src_off + done, chunk / 4);
    if (rc != DFQ_OK) return rc;
    memcpy((uint8_t *)dst + done, s->h_stage, chunk);
    done += chunk;
  }
  return DFQ_OK;
}

}  // namespace

extern "C" {

int dfq_serve_config_default(dfq_serve_config *cfg) {
  if (!cfg) return DFQ_ERR_INVALID_ARG;
  memset(cfg, 0, sizeof(*cfg));
  cfg->block_size = 1024;
  cfg->xcds_per_team = 1;
  cfg->ring_slots = 1024;
  cfg->groups = 0;
  cfg->arena_bytes = 0;
  return DFQ_OK;
}

uint64_t dfq_serve_footprint(uint32_t n, uint64_t m, uint32_t large_max, uint32_t xcds_per_team,
                             uint32_t flags) {
  ServeGraph g{};
  g.n = n;
  g.m = m;
  dfq_query_plan p = default_serve_plan();
  p.large_max = large_max;
  return footprint_of(g, p, xcds_per_team, flags, nullptr);
}

uint64_t dfq_serve_group_footprint(uint32_t n, uint64_t m, uint32_t count, uint32_t xcds_per_team) {
  if (count == 0 || count > kMsMaxSources) return 0;
  ServeGraph g{};
  g.n = n;
  g.m = m;
  return group_footprint(g, default_serve_plan(), xcds_per_team, count);
}

int dfq_server_create(uint32_t num_graphs, const uint32_t *n, const uint64_t *m, const int64_t *row_offsets,
                      const uint32_t *col_idx, const dfq_serve_config *cfg_in, dfq_server **out) {
  if (!out || num_graphs == 0 || !n || !m || !row_offsets) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  dfq_serve_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else dfq_serve_config_default(&cfg);
  if (!shape_ok(cfg.xcds_per_team, cfg.block_size) || cfg.groups > 1 || cfg.ring_slots < 2 ||
      cfg.ring_slots > 65536 || (cfg.ring_slots & (cfg.ring_slots - 1)))
    return DFQ_ERR_INVALID_ARG;
  std::vector<ServeGraph> graphs(num_graphs);
  uint64_t vtot = 0, etot = 0;
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
  for (uint32_t g = 0; g < num_graphs; ++g) {
    const int64_t *off = row_offsets + graphs[g].vbase;
    const uint32_t *col = col_idx + graphs[g].ebase;
    if (m[g] > kMaxEntries || !csr_valid(off, col, n[g], m[g])) return DFQ_ERR_RANGE;
    graphs[g].symmetric = is_symmetric(off, col, n[g]) ? 1u : 0u;
  }
  uint64_t iso_words = 0, obase = 0;
  for (auto &gr : graphs) {
    gr.iso_base = gr.symmetric ? iso_words : ~0ull;
    if (gr.symmetric) iso_words += ((uint64_t)gr.n + 31) / 32;
    gr.obase = obase;
    obase += gr.n;
    gr.nz = gr.n;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DFQ_ERR_NO_DEVICE;

  dfq_server *s = new (std::nothrow) dfq_server();
  if (!s) return DFQ_ERR_NOMEM;
  auto fail = [&](int rc) {
    free_all(s);
    delete s;
    return rc;
  };
#define SFAIL(expr)                                     \
  do {                                                  \
    if ((expr) != hipSuccess) return fail(DFQ_ERR_HIP); \
  } while (0)
  SFAIL(hipGetDevice(&s->device));
  hipDeviceProp_t prop;
  SFAIL(hipGetDeviceProperties(&prop, s->device));
  if (!prop.cooperativeLaunch) return fail(DFQ_ERR_COOPERATIVE);
  s->cfg = cfg;
  s->graphs = graphs;
  s->plans.assign(num_graphs, default_serve_plan());
  SFAIL(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  SFAIL(hipMalloc(&s->d_off, sizeof(int64_t) * vtot));
  // +4 entries: bottom-up levels read columns as aligned 16-byte blocks
  SFAIL(hipMalloc(&s->d_col, sizeof(uint32_t) * (etot + 4)));
  SFAIL(hipMalloc(&s->d_col_bu, sizeof(uint32_t) * (etot + 4)));
  SFAIL(hipMalloc(&s->d_graphs, sizeof(ServeGraph) * num_graphs));
  SFAIL(hipMemcpy(s->d_off, row_offsets, sizeof(int64_t) * vtot, hipMemcpyHostToDevice));
  {
    // symmetric graphs: a second copy of the rows by neighbour degree for
    // bottom-up levels, and isolated vertices marked visited from the start
    std::vector<uint32_t> iso(iso_words ? iso_words : 1), ordered, heads;
    SFAIL(hipMalloc(&s->d_head, sizeof(uint4) * vtot));
    for (uint32_t g = 0; g < num_graphs; ++g) {
      const ServeGraph &gr = graphs[g];
      const int64_t *off = row_offsets + gr.vbase;
      const uint32_t *col = col_idx + gr.ebase;
      if (gr.m) SFAIL(hipMemcpy(s->d_col + gr.ebase, col, sizeof(uint32_t) * gr.m, hipMemcpyHostToDevice));
      if (!gr.symmetric) continue;
      ordered.resize(gr.m);
      order_rows_by_degree(off, col, gr.n, ordered.data());
      isolated_bitmap(off, gr.n, iso.data() + gr.iso_base);
      if (gr.m)
        SFAIL(hipMemcpy(s->d_col_bu + gr.ebase, ordered.data(), sizeof(uint32_t) * gr.m, hipMemcpyHostToDevice));
      heads.resize(4 * (size_t)gr.n);
      row_heads(off, ordered.data(), gr.n, heads.data());
      if (gr.n) SFAIL(hipMemcpy(s->d_head + gr.vbase, heads.data(), sizeof(uint4) * gr.n, hipMemcpyHostToDevice));
    }
    SFAIL(hipMalloc(&s->d_iso, sizeof(uint32_t) * iso.size()));
    SFAIL(hipMemcpy(s->d_iso, iso.data(), sizeof(uint32_t) * iso.size(), hipMemcpyHostToDevice));
  }
  if (cfg.groups) {
    // group tasks search degree-ranked copies of the graphs
    SFAIL(hipMalloc(&s->d_off_r, sizeof(int64_t) * vtot));
    SFAIL(hipMalloc(&s->d_col_r, sizeof(uint32_t) * (etot + 4)));
    SFAIL(hipMalloc(&s->d_o2r, sizeof(uint32_t) * (obase ? obase : 1)));
    std::vector<int64_t> off_r;
    std::vector<uint32_t> col_r, o2r, r2o;
    for (uint32_t g = 0; g < num_graphs; ++g) {
      ServeGraph &gr = graphs[g];
      try {
        off_r.resize((size_t)gr.n + 1);
        col_r.resize(gr.m ? gr.m : 1);
        o2r.resize(gr.n);
        r2o.resize(gr.n);
      } catch (...) {
        return fail(DFQ_ERR_NOMEM);
      }
      rank_relabel(row_offsets + gr.vbase, col_idx + gr.ebase, gr.n, o2r.data(), r2o.data(), off_r.data(),
                   col_r.data());
      uint32_t nz = gr.n;
      while (nz > 0 && off_r[nz - 1] == off_r[nz]) --nz;
      gr.nz = nz;
      SFAIL(hipMemcpy(s->d_off_r + gr.vbase, off_r.data(), sizeof(int64_t) * ((size_t)gr.n + 1),
                      hipMemcpyHostToDevice));
      if (gr.m) SFAIL(hipMemcpy(s->d_col_r + gr.ebase, col_r.data(), sizeof(uint32_t) * gr.m, hipMemcpyHostToDevice));
      SFAIL(hipMemcpy(s->d_o2r + gr.obase, o2r.data(), sizeof(uint32_t) * gr.n, hipMemcpyHostToDevice));
    }
    s->graphs = graphs;
  }
  SFAIL(hipMemcpy(s->d_graphs, graphs.data(), sizeof(ServeGraph) * num_graphs, hipMemcpyHostToDevice));
  SFAIL(hipMalloc(&s->d_ctl, sizeof(ServeCtl)));
  SFAIL(hipMemset(s->d_ctl, 0, sizeof(ServeCtl)));
  // arena: by default four working sets of the largest graph per team (two
  // full groups per team in group mode)
  uint64_t arena = cfg.arena_bytes;
  if (arena == 0) {
    uint64_t fmax = 0;
    for (uint32_t g = 0; g < num_graphs; ++g)
      fmax = std::max(fmax, cfg.groups ? group_footprint(graphs[g], s->plans[g], cfg.xcds_per_team, kMsMaxSources)
                                       : footprint_of(graphs[g], s->plans[g], cfg.xcds_per_team, kServeParents,
                                                      nullptr));
    const uint64_t teams = cfg.xcds_per_team ? (8 + cfg.xcds_per_team - 1) / cfg.xcds_per_team : 1;
    arena = (cfg.groups ? 2 : 4) * teams * fmax;
    size_t free_b = 0, total_b = 0;
    SFAIL(hipMemGetInfo(&free_b, &total_b));
    arena = std::min<uint64_t>(arena, free_b / 2);
  }
  arena = serve_align(arena);
  SFAIL(hipMalloc(&s->d_arena, arena));
  s->cfg.arena_bytes = arena;
  s->arena.reset(arena);
  // ring, results, stop word and status in one pinned, coherent block
  const uint32_t R = cfg.ring_slots;
  const size_t block_bytes = sizeof(ServeRequest) * R + sizeof(ServeResult) * R + 256 + sizeof(ServeStatus);
  // group blocks: sources and per-search results of every ring slot
  const size_t gblock_bytes = (sizeof(uint32_t) + sizeof(ServeGroupResult)) * kMsMaxSources * R;
  SFAIL(hipHostMalloc(&s->h_block, block_bytes, hipHostMallocMapped | hipHostMallocCoherent));
  memset(s->h_block, 0, block_bytes);
  uint8_t *hb = (uint8_t *)s->h_block;
  s->h_req = (ServeRequest *)hb;
  s->h_res = (ServeResult *)(hb + sizeof(ServeRequest) * R);
  s->h_stop = (uint32_t *)(hb + sizeof(ServeRequest) * R + sizeof(ServeResult) * R);
  s->h_status = (ServeStatus *)((uint8_t *)s->h_stop + 256);
  void *dblock = nullptr;
  SFAIL(hipHostGetDevicePointer(&dblock, s->h_block, 0));
  uint8_t *db = (uint8_t *)dblock;
  SFAIL(hipHostMalloc((void **)&s->h_stage, kStageBytes, hipHostMallocMapped | hipHostMallocCoherent));
  SFAIL(hipHostGetDevicePointer((void **)&s->d_stage, s->h_stage, 0));
  void *gblock = nullptr, *d_gblock = nullptr;
  SFAIL(hipHostMalloc(&gblock, gblock_bytes, hipHostMallocMapped | hipHostMallocCoherent));
  memset(gblock, 0, gblock_bytes);
  s->h_gres = (ServeGroupResult *)gblock;
  s->h_gsrc = (uint32_t *)((uint8_t *)gblock + sizeof(ServeGroupResult) * kMsMaxSources * R);
  SFAIL(hipHostGetDevicePointer(&d_gblock, gblock, 0));
  s->ring.assign(R, RingSlot());

  ServeArgs &a = s->args;
  a.off = s->d_off;
  a.col = s->d_col;
  a.graphs = s->d_graphs;
  a.iso = s->d_iso;
  a.col_bu = s->d_col_bu;
  a.head = s->d_head;
  a.req = (ServeRequest *)db;
  a.res = (ServeResult *)(db + sizeof(ServeRequest) * R);
  a.stop = (const uint32_t *)(db + sizeof(ServeRequest) * R + sizeof(ServeResult) * R);
  a.status = (ServeStatus *)((uint8_t *)a.stop + 256);
  a.ctl = s->d_ctl;
  a.arena = s->d_arena;
  a.ring = R;
  a.off_r = s->d_off_r;
  a.col_r = s->d_col_r;
  a.o2r = s->d_o2r;
  a.gres = (ServeGroupResult *)d_gblock;
  a.gsrc = (const uint32_t *)((uint8_t *)d_gblock + sizeof(ServeGroupResult) * kMsMaxSources * R);
  a.bu_div = kMsBuDiv;
#undef SFAIL
  *out = s;
  return DFQ_OK;
}

void dfq_server_destroy(dfq_server *s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->running) (void)dfq_server_stop(s);
  if (s->running) s->wedged = true;  // never free memory a live launch may touch
  free_all(s);
  if (!s->wedged) delete s;
}

int dfq_server_set_plan(dfq_server *s, uint32_t graph, const dfq_query_plan *plan) {
  if (!s || !plan || graph >= s->graphs.size()) return DFQ_ERR_INVALID_ARG;
  if (!plan_ok(*plan)) return DFQ_ERR_INVALID_ARG;
  s->plans[graph] = *plan;
  return DFQ_OK;
}

int dfq_server_get_plan(const dfq_server *s, uint32_t graph, dfq_query_plan *plan) {
  if (!s || !plan || graph >= s->graphs.size()) return DFQ_ERR_INVALID_ARG;
  *plan = s->plans[graph];
  return DFQ_OK;
}

int dfq_server_set_shape(dfq_server *s, uint32_t xcds_per_team, uint32_t block_size) {
  if (!s || !shape_ok(xcds_per_team, block_size)) return DFQ_ERR_INVALID_ARG;
  if (s->running || !s->queries.empty()) return DFQ_ERR_STATE;  // layouts of held ranges depend on it
  s->cfg.xcds_per_team = xcds_per_team;
  s->cfg.block_size = block_size;
  return DFQ_OK;
}

int dfq_server_get_info(const dfq_server *s, dfq_server_info *info) {
  if (!s || !info) return DFQ_ERR_INVALID_ARG;
  memset(info, 0, sizeof(*info));
  info->running = s->running ? 1u : 0u;
  const uint32_t teams = __atomic_load_n(&s->h_status->teams, __ATOMIC_ACQUIRE);
  info->teams = (s->running && teams != 0xffffffffu) ? teams : 0u;
  for (uint32_t t = 0; t < info->teams && t < 16; ++t)
    info->team_blocks[t] = __atomic_load_n(&s->h_status->team_blocks[t], __ATOMIC_RELAXED);
  info->grid_blocks = (uint32_t)s->grid;
  info->block_size = s->cfg.block_size;
  info->xcds_per_team = s->cfg.xcds_per_team;
  info->ring_slots = s->cfg.ring_slots;
  info->in_flight = (uint32_t)s->queries.size();
  info->arena_bytes = s->arena.size();
  info->arena_used = s->arena.used();
  return DFQ_OK;
}

int dfq_server_start(dfq_server *s) {
  if (!s) return DFQ_ERR_INVALID_ARG;
  if (s->running || s->wedged) return DFQ_ERR_STATE;
  SHIP(hipSetDevice(s->device));
  int per_cu = 0;
  SHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_kernel_for(s->cfg.block_size, s->cfg.groups != 0),
                                                    (int)s->cfg.block_size, 0));
  hipDeviceProp_t prop;
  SHIP(hipGetDeviceProperties(&prop, s->device));
  if (per_cu <= 0) return DFQ_ERR_COOPERATIVE;
  // one workgroup per CU: teams are whole XCDs either way, and the level
  // loop is tuned for it
  s->grid = prop.multiProcessorCount;
  // fresh ring (results of earlier launches live in the query table)
  const uint32_t R = s->cfg.ring_slots;
  for (uint32_t i = 0; i < R; ++i) {
    if (s->ring[i].busy) return DFQ_ERR_STATE;
    __atomic_store_n(&s->h_req[i].seq, 0ull, __ATOMIC_RELAXED);
    __atomic_store_n(&s->h_res[i].done, 0ull, __ATOMIC_RELAXED);
  }
  memset(s->h_status, 0, sizeof(ServeStatus));
  __atomic_store_n(s->h_stop, 0u, __ATOMIC_RELEASE);
  s->base = (s->next_ticket + R - 1) / R * R;
  s->next_ticket = s->base;
  SHIP(hipMemsetAsync(&s->d_ctl->head, 0, sizeof(BarrierWord), s->stream));
  s->args.xpt = s->cfg.xcds_per_team;
  void *kargs[] = {&s->args};
  SHIP(hipLaunchCooperativeKernel(serve_kernel_for(s->cfg.block_size, s->cfg.groups != 0), dim3(s->grid), dim3(s->cfg.block_size),
                                  kargs, 0, s->stream));
  s->running = true;
  // the teams report in once the launch is resident
  const auto t0 = Clock::now();
  for (;;) {
    const uint32_t teams = __atomic_load_n(&s->h_status->teams, __ATOMIC_ACQUIRE);
    if (teams == 0xffffffffu || (teams == 0 && kernel_ended(s))) {
      (void)hipStreamSynchronize(s->stream);
      s->running = false;
      return DFQ_ERR_COOPERATIVE;
    }
    if (teams) break;
    if (us_since(t0) > kEndMs * 1000.0) {
      // not resident in time: ask it to leave and report
      __atomic_store_n(s->h_stop, 1u, __ATOMIC_RELEASE);
      return DFQ_ERR_TIMEOUT;
    }
    std::this_thread::yield();
  }
  return DFQ_OK;
}

int dfq_server_stop(dfq_server *s) {
  if (!s) return DFQ_ERR_INVALID_ARG;
  if (!s->running) return DFQ_OK;
  SHIP(hipSetDevice(s->device));
  int rc = DFQ_OK;
  const auto t0 = Clock::now();
  for (;;) {
    bool pending = false;
    for (uint32_t i = 0; i < s->cfg.ring_slots; ++i)
      if (!collect(s, i)) pending = true;
    if (!pending) break;
    if (kernel_ended(s)) break;
    if (us_since(t0) > kDrainMs * 1000.0) {
      rc = DFQ_ERR_TIMEOUT;
      break;
    }
    std::this_thread::yield();
  }
  __atomic_store_n(s->h_stop, 1u, __ATOMIC_RELEASE);
  const auto t1 = Clock::now();
  while (!kernel_ended(s)) {
    if (us_since(t1) > kEndMs * 1000.0) return DFQ_ERR_TIMEOUT;
    std::this_thread::yield();
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess) rc = DFQ_ERR_HIP;
  s->running = false;
  // queries published but never run (a launch that ended early) give their
  // ranges back; their tickets become unknown
  for (uint32_t i = 0; i < s->cfg.ring_slots; ++i) {
    RingSlot &rs = s->ring[i];
    if (!rs.busy) continue;
    if (!rs.copy)
      for (uint32_t i = 0; i < rs.count; ++i) {
        auto it = s->queries.find(rs.qid0 + i);
        if (it != s->queries.end()) release_query(s, it);
      }
    rs.busy = false;
    rs.copy = false;
    if (rc == DFQ_OK) rc = DFQ_ERR_STATE;
  }
  return rc;
}

int dfq_server_submit_group(dfq_server *s, uint32_t graph, uint32_t count, const uint32_t *sources, uint32_t flags,
                            uint64_t *tickets);

int dfq_server_submit(dfq_server *s, uint32_t graph, uint32_t source, uint32_t flags, uint64_t *ticket) {
  if (!s || !ticket || graph >= s->graphs.size() || (flags & ~kServeParents)) return DFQ_ERR_INVALID_ARG;
  if (s->cfg.groups) return dfq_server_submit_group(s, graph, 1, &source, flags, ticket);
  if (source >= s->graphs[graph].n) return DFQ_ERR_RANGE;
  if (!s->running) return DFQ_ERR_STATE;
  const ServeGraph &g = s->graphs[graph];
  const dfq_query_plan &p = s->plans[graph];
  uint32_t item_cap = 0;
  const uint64_t bytes = footprint_of(g, p, s->cfg.xcds_per_team, flags, &item_cap);
  if (bytes > s->arena.size()) return DFQ_ERR_CAPACITY;
  if (!collect(s, (uint32_t)(s->next_ticket % s->cfg.ring_slots))) return DFQ_ERR_BUSY;
  uint64_t work = 0;
  if (!s->arena.alloc(bytes, &work)) return DFQ_ERR_BUSY;
  ServeRequest req;
  memset(&req, 0, sizeof(req));
  req.work = work;
  req.graph = graph;
  req.source = source;
  req.flags = (flags & kServeParents) | (p.direction_opt ? kServeDirOpt : 0u);
  req.item_cap = item_cap;
  req.small_max = p.small_max;
  req.medium_max = p.medium_max;
  req.large_max = p.large_max;
  req.alpha = p.alpha;
  req.beta = p.beta;
  Query q;
  q.work = work;
  q.bytes = bytes;
  q.graph = graph;
  q.source = source;
  q.flags = flags;
  q.item_cap = item_cap;
  q.ring_ticket = s->next_ticket;
  const uint64_t qid = s->next_qid++;
  s->queries[qid] = q;
  publish(s, req, false, qid, 1);
  *ticket = qid;
  return DFQ_OK;
}

int dfq_server_submit_group(dfq_server *s, uint32_t graph, uint32_t count, const uint32_t *sources, uint32_t flags,
                            uint64_t *tickets) {
  if (!s || !tickets || !sources || graph >= s->graphs.size() || flags) return DFQ_ERR_INVALID_ARG;
  if (!s->cfg.groups || count == 0 || count > kMsMaxSources) return DFQ_ERR_INVALID_ARG;
  for (uint32_t i = 0; i < count; ++i)
    if (sources[i] >= s->graphs[graph].n) return DFQ_ERR_RANGE;
  if (!s->running) return DFQ_ERR_STATE;
  const ServeGraph &g = s->graphs[graph];
  const dfq_query_plan &p = s->plans[graph];
  const uint32_t item_cap = group_item_cap(g, p);
  const uint64_t bytes = group_footprint(g, p, s->cfg.xcds_per_team, count);
  if (bytes > s->arena.size()) return DFQ_ERR_CAPACITY;
  const uint32_t idx = (uint32_t)(s->next_ticket % s->cfg.ring_slots);
  if (!collect(s, idx)) return DFQ_ERR_BUSY;
  uint64_t work = 0;
  if (!s->arena.alloc(bytes, &work)) return DFQ_ERR_BUSY;
  ServeRequest req;
  memset(&req, 0, sizeof(req));
  req.work = work;
  req.graph = graph;
  req.source = count;
  req.flags = kServeGroup | (p.direction_opt ? kServeDirOpt : 0u);
  req.item_cap = item_cap;
  req.small_max = p.small_max;
  req.medium_max = p.medium_max;
  req.large_max = p.large_max;
  req.alpha = p.alpha;
  req.beta = p.beta;
  // the sources travel in the slot's host-mapped group block
  memcpy(s->h_gsrc + (size_t)idx * kMsMaxSources, sources, sizeof(uint32_t) * count);
  const uint64_t gid = s->next_group++;
  GroupRange gr;
  gr.work = work;
  gr.bytes = bytes;
  gr.held = count;
  s->groups_held[gid] = gr;
  const uint64_t qid0 = s->next_qid;
  s->next_qid += count;
  for (uint32_t i = 0; i < count; ++i) {
    Query q;
    q.work = work;
    q.bytes = bytes;
    q.group = gid;
    q.member = i;
    q.graph = graph;
    q.source = sources[i];
    q.flags = 0;
    q.item_cap = item_cap;
    q.count = count;
    q.ring_ticket = s->next_ticket;
    s->queries[qid0 + i] = q;
    tickets[i] = qid0 + i;
  }
  // the host's stores to the group block precede the publication's release
  publish(s, req, false, qid0, count);
  return DFQ_OK;
}

int dfq_server_poll(dfq_server *s, uint64_t ticket, dfq_serve_result *result) {
  if (!s) return DFQ_ERR_INVALID_ARG;
  Query *q = find(s, ticket);
  if (!q) return DFQ_ERR_INVALID_ARG;
  if (!q->done) return DFQ_ERR_PENDING;
  if (result) *result = q->result;
  return q->status;
}

int dfq_server_wait(dfq_server *s, uint64_t ticket, uint32_t timeout_ms, dfq_serve_result *result) {
  const auto t0 = Clock::now();
  for (uint64_t k = 0;; ++k) {
    const int rc = dfq_server_poll(s, ticket, result);
    if (rc != DFQ_ERR_PENDING) return rc;
    if ((k & 1023) == 1023) {
      if (!s->running || kernel_ended(s)) return DFQ_ERR_STATE;
      if (us_since(t0) > timeout_ms * 1000.0) return DFQ_ERR_TIMEOUT;
    }
  }
}

int dfq_server_levels(dfq_server *s, uint64_t ticket, uint32_t *level_out, uint32_t *parent_out) {
  if (!s || !level_out) return DFQ_ERR_INVALID_ARG;
  Query *q = find(s, ticket);
  if (!q) return DFQ_ERR_INVALID_ARG;
  if (!q->done) return DFQ_ERR_PENDING;
  if (q->status != DFQ_OK) return q->status;
  const ServeGraph &g = s->graphs[q->graph];
  if (q->group) {
    if (parent_out) return DFQ_ERR_INVALID_ARG;
    const ServeGroupLayout l = serve_group_layout(g.n, q->count, q->item_cap, serve_segments(s->cfg.xcds_per_team));
    return copy_out(s, q->work + l.level + 4ull * l.n_pad * q->member, level_out, 4ull * g.n);
  }
  const bool parents = (q->flags & kServeParents) != 0;
  if (parent_out && !parents) return DFQ_ERR_INVALID_ARG;
  // the layout the query was admitted with (the shape cannot change while
  // ranges are held)
  const ServeLayout l = serve_layout(g.n, q->item_cap, serve_segments(s->cfg.xcds_per_team), parents);
  const uint64_t work = q->work;
  int rc = copy_out(s, work + l.level, level_out, 4ull * g.n);
  if (rc == DFQ_OK && parent_out) rc = copy_out(s, work + l.parent, parent_out, 4ull * g.n);
  return rc;
}

int dfq_server_release(dfq_server *s, uint64_t ticket) {
  if (!s) return DFQ_ERR_INVALID_ARG;
  Query *q = find(s, ticket);
  if (!q) return DFQ_ERR_INVALID_ARG;
  if (!q->done) return DFQ_ERR_PENDING;
  release_query(s, s->queries.find(ticket));
  return DFQ_OK;
}

}  // extern "C"

namespace {

// The serving loop of a group-mode server: arrivals wait per graph; while
// fewer than `teams + 2` group tasks are in flight, the graph of the oldest
// waiting query sends up to 64 of its waiting queries (oldest first) as one
// group task.  Queries that arrive while the teams are busy so wait for the
// next free team and share its traversal.
int replay_groups(dfq_server *s, uint32_t nq, const uint32_t *graph, const uint32_t *source, const double *arrival_us,
                  uint32_t *const *level_out, dfq_serve_record *records) {
  const uint32_t ngraphs = (uint32_t)s->graphs.size();
  std::vector<std::vector<uint32_t>> waiting(ngraphs);  // per graph, arrival order
  std::vector<size_t> head(ngraphs, 0);                 // first waiting entry not yet sent
  struct Live {
    std::vector<uint32_t> q;
    std::vector<uint64_t> t;
  };
  std::vector<Live> live;
  const uint32_t teams = std::max<uint32_t>(1u, __atomic_load_n(&s->h_status->teams, __ATOMIC_ACQUIRE));
  const size_t max_live = teams + 2;
  const auto t0 = Clock::now();
  uint32_t arrived = 0, oldest = 0, finished = 0;
  std::vector<uint8_t> sent(nq, 0);
  std::vector<uint32_t> members;
  std::vector<uint32_t> srcs;
  std::vector<uint64_t> tks(kMsMaxSources);
  uint64_t idle = 0;
  while (finished < nq) {
    bool progress = false;
    const double now = us_since(t0);
    while (arrived < nq && arrival_us[arrived] <= now) {
      waiting[graph[arrived]].push_back(arrived);
      ++arrived;
    }
    while (live.size() < max_live) {
      while (oldest < arrived && sent[oldest]) ++oldest;
      if (oldest >= arrived) break;
      const uint32_t g = graph[oldest];
      // the largest group whose working set the arena can hold
      uint32_t cap = kMsMaxSources;
      while (cap > 1 && group_footprint(s->graphs[g], s->plans[g], s->cfg.xcds_per_team, cap) > s->arena.size())
        cap /= 2;
      members.clear();
      srcs.clear();
      for (size_t k = head[g]; k < waiting[g].size() && members.size() < cap; ++k) {
        members.push_back(waiting[g][k]);
        srcs.push_back(source[waiting[g][k]]);
      }
      const int rc = dfq_server_submit_group(s, g, (uint32_t)members.size(), srcs.data(), 0, tks.data());
      if (rc == DFQ_ERR_BUSY && live.empty() && !any_running(s)) return DFQ_ERR_BUSY;
      if (rc == DFQ_ERR_BUSY) break;
      if (rc != DFQ_OK) return rc;
      const double admit = us_since(t0);
      Live l;
      for (size_t k = 0; k < members.size(); ++k) {
        const uint32_t q = members[k];
        sent[q] = 1;
        records[q].arrival_us = arrival_us[q];
        records[q].admit_us = admit;
        l.q.push_back(q);
        l.t.push_back(tks[k]);
      }
      head[g] += members.size();
      live.push_back(std::move(l));
      progress = true;
    }
    for (size_t i = 0; i < live.size();) {
      dfq_serve_result r;
      const int rc = dfq_server_poll(s, live[i].t[0], &r);
      if (rc == DFQ_ERR_PENDING) {
        ++i;
        continue;
      }
      const double done = us_since(t0);
      for (size_t k = 0; k < live[i].q.size(); ++k) {
        const uint32_t q = live[i].q[k];
        const uint64_t tk = live[i].t[k];
        const int prc = dfq_server_poll(s, tk, &r);
        if (prc != DFQ_OK) return prc;
        dfq_serve_record &rec = records[q];
        rec.done_us = done;
        rec.service_us = r.service_us;
        rec.reached = r.reached;
        rec.edges = r.edges;
        rec.depth = r.depth;
        rec.team = r.team;
        if (level_out && level_out[q]) {
          const int crc = dfq_server_levels(s, tk, level_out[q], nullptr);
          if (crc != DFQ_OK) return crc;
        }
        const int rrc = dfq_server_release(s, tk);
        if (rrc != DFQ_OK) return rrc;
        ++finished;
      }
      live[i] = std::move(live.back());
      live.pop_back();
      progress = true;
    }
    if (progress) {
      idle = 0;
    } else if ((++idle & 4095) == 0) {
      if (kernel_ended(s)) return DFQ_ERR_STATE;
    }
  }
  return DFQ_OK;
}

}  // namespace

extern "C" {

int dfq_server_replay(dfq_server *s, uint32_t nq, const uint32_t *graph, const uint32_t *source,
                      const double *arrival_us, uint32_t flags, uint32_t *const *level_out,
                      dfq_serve_record *records) {
  if (!s || (nq && (!graph || !source || !arrival_us || !records))) return DFQ_ERR_INVALID_ARG;
  if (flags & ~kServeParents) return DFQ_ERR_INVALID_ARG;
  for (uint32_t q = 0; q < nq; ++q) {
    if (graph[q] >= s->graphs.size()) return DFQ_ERR_INVALID_ARG;
    if (source[q] >= s->graphs[graph[q]].n) return DFQ_ERR_RANGE;
    if (q && arrival_us[q] < arrival_us[q - 1]) return DFQ_ERR_INVALID_ARG;
    const uint64_t fp = s->cfg.groups
                            ? group_footprint(s->graphs[graph[q]], s->plans[graph[q]], s->cfg.xcds_per_team, 1)
                            : footprint_of(s->graphs[graph[q]], s->plans[graph[q]], s->cfg.xcds_per_team, flags, nullptr);
    if (fp > s->arena.size()) return DFQ_ERR_CAPACITY;
  }
  if (!s->running) return DFQ_ERR_STATE;
  memset(records, 0, sizeof(dfq_serve_record) * nq);
  if (s->cfg.groups) {
    if (flags) return DFQ_ERR_INVALID_ARG;  // group tasks write no parents
    return replay_groups(s, nq, graph, source, arrival_us, level_out, records);
  }
  std::vector<std::pair<uint32_t, uint64_t>> live;
  live.reserve(s->cfg.ring_slots);
  const auto t0 = Clock::now();
  uint32_t next = 0, finished = 0;
  uint64_t idle = 0;
  while (finished < nq) {
    bool progress = false;
    // admission in arrival order while a slot and arena room are free
    while (next < nq && arrival_us[next] <= us_since(t0)) {
      uint64_t tk = 0;
      const int rc = dfq_server_submit(s, graph[next], source[next], flags, &tk);
      // refused while nothing runs on the device: the arena is held by
      // finished queries the caller has not released, so waiting cannot help
      if (rc == DFQ_ERR_BUSY && live.empty() && !any_running(s)) return DFQ_ERR_BUSY;
      if (rc == DFQ_ERR_BUSY) break;
      if (rc != DFQ_OK) return rc;
      records[next].arrival_us = arrival_us[next];
      records[next].admit_us = us_since(t0);
      live.emplace_back(next, tk);
      ++next;
      progress = true;
    }
    for (size_t i = 0; i < live.size();) {
      const uint64_t tk = live[i].second;
      dfq_serve_result r;
      const int rc = dfq_server_poll(s, tk, &r);
      if (rc == DFQ_ERR_PENDING) {
        ++i;
        continue;
      }
      if (rc != DFQ_OK) return rc;
      const uint32_t q = live[i].first;
      dfq_serve_record &rec = records[q];
      rec.done_us = us_since(t0);
      rec.service_us = r.service_us;
      rec.reached = r.reached;
      rec.edges = r.edges;
      rec.depth = r.depth;
      rec.team = r.team;
      if (level_out && level_out[q]) {
        const int crc = dfq_server_levels(s, tk, level_out[q], nullptr);
        if (crc != DFQ_OK) return crc;
      }
      const int rrc = dfq_server_release(s, tk);
      if (rrc != DFQ_OK) return rrc;
      live[i] = live.back();
      live.pop_back();
      ++finished;
      progress = true;
    }
    if (progress) {
      idle = 0;
    } else if ((++idle & 4095) == 0) {
      if (kernel_ended(s)) return DFQ_ERR_STATE;
    }
  }
  return DFQ_OK;
}

}  // extern "C"
