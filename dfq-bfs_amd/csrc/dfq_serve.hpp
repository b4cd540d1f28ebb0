// Serving kernel: one persistent cooperative launch split into teams of
// XCDs.  Each team repeatedly takes the next ticket of a request ring that
// the host fills in pinned, host-mapped memory, runs that query's BFS with
// the plain kernel's level loop (grid_bfs) on the team's own workgroups and
// barrier, and publishes the result in the ring.  No counterpart in the
// reference, whose interactive loop (design/shared-q/atomic-sa/our-sep/
// upper-bound/bfs.cu:202-214,302-308) relaunched the kernel per source.
//
// Per-query device memory (levels, bitmaps, frontier segments, work items)
// is a range of one arena that the host hands out at admission; the layout
// inside a range is serve_layout(), shared by host and device.
#pragma once

#include "dfq_kernels.hpp"
#include "dfq_msbfs.hpp"

namespace dfq {

constexpr uint32_t kServeParents = 1u;   // public flag DFQ_SERVE_PARENTS
constexpr uint32_t kServeDirOpt = 2u;    // internal: direction optimisation allowed
constexpr uint32_t kServeCopy = 0x100u;  // internal: a copy task, not a search
constexpr uint32_t kServeGroup = 0x200u; // internal: up to 64 searches of one graph in one traversal
constexpr int kMaxTeams = kMaxXcd;

struct ServeGraph {
  uint64_t vbase;      // first row offset of this graph in the concatenated offsets
  uint64_t ebase;      // first column index in the concatenated columns
  uint64_t m;
  uint32_t n;
  uint32_t symmetric;
  uint64_t iso_base;   // first word of its isolated-vertex bitmap, ~0 if none
  uint64_t obase;      // first entry of its vertex -> rank map (degree-ranked copy)
  uint32_t nz;         // vertices with entries (ranks [0, nz))
  uint32_t pad;
};

// A group task (kServeGroup) holds `source` searches (1..64) of `graph`,
// their sources in the host-mapped group block of the ring slot; each
// search's reached count, entries and depth go back there too.
struct ServeGroupResult {
  uint64_t reached;
  uint64_t edges;
  uint32_t depth;
  uint32_t pad;
};

// Written by the host, read by the team leader.  seq = ticket + 1 publishes.
// A copy task (kServeCopy) moves `source` 32-bit words from arena offset
// `work` to the host-mapped address `dst`: copy engines do not run beside
// the persistent launch, so results leave the device this way while it runs.
struct alignas(64) ServeRequest {
  uint64_t seq;
  uint64_t work;       // arena byte offset of the working set (copy: of the source)
  uint64_t dst;        // copy: device address of the pinned destination
  uint32_t graph, source;  // copy: source = words to move
  uint32_t flags;      // kServeParents | kServeDirOpt | kServeCopy
  uint32_t item_cap;   // work items per segment
  uint32_t small_max, medium_max, large_max;
  uint32_t alpha, beta;
  uint32_t pad;
};

// Written by the team leader, read by the host.  done = ticket + 1.
struct alignas(64) ServeResult {
  uint64_t done;
  uint64_t reached;
  uint64_t t_start_ns, t_end_ns;  // device clock
  uint32_t depth, bu_levels;
  uint32_t team, err;
  uint64_t edges;  // CSR entries of the reached vertices
  uint32_t pad[2];
};

// Host-mapped status the kernel reports once it runs.
struct ServeStatus {
  uint32_t teams;
  uint32_t team_blocks[kMaxTeams];
  uint32_t pad[15];
};

struct ServeCtl {
  GridBarrier census;            // kernel-start census only
  BarrierWord head;              // next ticket to take
  Control team[kMaxTeams];       // per-team barrier and level slots
  ServeRequest desc[kMaxTeams];  // the team's current request (device copy)
  uint64_t ticket[kMaxTeams];
};

// Offsets of one query's arrays inside its arena range (256-byte aligned).
struct ServeLayout {
  uint64_t level, parent, vis, snap, fb, qseg0, qseg1, items, bytes;
};

// Offsets of a group task's arrays inside its arena range: `count` level
// arrays of n_pad entries, then the multi-source search's working set.
struct ServeGroupLayout {
  uint64_t level, n_pad, seen, f, planes, qb, items, counters, depth, sources, bytes;
};

__host__ __device__ inline uint64_t serve_align(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

// Work items of one level fit in item_cap per segment: a row of d > large_max
// entries makes ceil(d / kItem) <= d / kItem + 1 items, and at most
// min(n, m / (large_max + 1)) rows are that long.
__host__ __device__ inline uint32_t serve_item_cap(uint32_t n, uint64_t m, uint32_t large_max) {
  uint64_t rows = m / ((uint64_t)large_max + 1);
  if (rows > n) rows = n;
  const uint64_t cap = m / kItem + rows + 1;
  return cap > 0xffffffffull ? 0xffffffffu : (uint32_t)cap;
}

// Frontier segments a query's working set holds: one per XCD of its team.
__host__ __device__ inline uint32_t serve_segments(uint32_t xpt) {
  return (xpt == 0 || xpt > (uint32_t)kSegments) ? (uint32_t)kSegments : xpt;
}

__host__ __device__ inline ServeLayout serve_layout(uint32_t n, uint32_t item_cap, uint32_t segs,
                                                    bool parents) {
  ServeLayout l;
  const uint64_t words = ((uint64_t)n + 31) / 32;
  uint64_t o = 0;
  l.level = o;  o += serve_align(4ull * n);
  l.parent = o; if (parents) o += serve_align(4ull * n);
  l.vis = o;    o += serve_align(4 * words);
  l.snap = o;   o += serve_align(4 * words);
  l.fb = o;     o += serve_align(8 * (words + 1));
  l.qseg0 = o;  o += serve_align(4ull * segs * n);
  l.qseg1 = o;  o += serve_align(4ull * segs * n);
  l.items = o;  o += serve_align(8ull * segs * item_cap);
  l.bytes = o;
  return l;
}

__host__ __device__ inline ServeGroupLayout serve_group_layout(uint32_t n, uint32_t count, uint32_t item_cap,
                                                              uint32_t segs) {
  ServeGroupLayout l;
  l.n_pad = ((uint64_t)n + 63) & ~(uint64_t)63;
  uint64_t o = 0;
  l.level = o;    o += serve_align(4ull * l.n_pad * count);
  l.seen = o;     o += serve_align(8ull * n);
  l.f = o;        o += serve_align(16ull * n);
  l.planes = o;   o += serve_align(8ull * kMsPlanes * n);
  l.qb = o;       o += serve_align(4ull * 3 * segs * n);
  l.items = o;    o += serve_align(8ull * segs * item_cap);
  l.counters = o; o += serve_align(8ull * 2 * kMsMaxSources);
  l.depth = o;    o += serve_align(4ull * kMsMaxSources);
  l.sources = o;  o += serve_align(4ull * kMsMaxSources);
  l.bytes = o;
  return l;
}

struct ServeArgs {
  const int64_t *off;           // concatenated row offsets
  const uint32_t *col;          // concatenated columns
  const uint32_t *col_bu;       // the same, rows by neighbour degree (symmetric graphs)
  const uint4 *head;            // first four col_bu entries per vertex, indexed like off
  const ServeGraph *graphs;
  const uint32_t *iso;          // concatenated isolated-vertex bitmaps (symmetric graphs)
  ServeRequest *req;            // host-mapped ring [ring]
  ServeResult *res;             // host-mapped ring [ring]
  const uint32_t *stop;         // host-mapped: set after the host has drained
  ServeStatus *status;          // host-mapped
  ServeCtl *ctl;
  uint8_t *arena;
  uint32_t ring;                // power of two
  uint32_t xpt;                 // XCDs per team (0 = all)
  const int64_t *off_r;         // degree-ranked copies of the graphs (group tasks), indexed like off
  const uint32_t *col_r;        // and like col
  const uint32_t *o2r;          // vertex -> rank, graph g's from graphs[g].obase
  const uint32_t *gsrc;         // host-mapped [ring][64] sources of group tasks
  ServeGroupResult *gres;       // host-mapped [ring][64] results of group tasks
  uint32_t bu_div;              // group tasks: bottom-up when frontier out-edges > m / bu_div
  uint32_t pad;
};

__device__ __forceinline__ uint64_t ld_system(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t ld_system(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Request and graph descriptors are read from memory by every thread alike;
// moved to scalar registers, everything derived from them stays scalar.
__device__ __forceinline__ ServeRequest uniform_request(const ServeRequest &r) {
  ServeRequest d;
  d.seq = uniform(r.seq);
  d.work = uniform(r.work);
  d.dst = uniform(r.dst);
  d.graph = uniform(r.graph);
  d.source = uniform(r.source);
  d.flags = uniform(r.flags);
  d.item_cap = uniform(r.item_cap);
  d.small_max = uniform(r.small_max);
  d.medium_max = uniform(r.medium_max);
  d.large_max = uniform(r.large_max);
  d.alpha = uniform(r.alpha);
  d.beta = uniform(r.beta);
  d.pad = 0;
  return d;
}

__device__ __forceinline__ ServeGraph uniform_graph(const ServeGraph &x) {
  ServeGraph g;
  g.vbase = uniform(x.vbase);
  g.ebase = uniform(x.ebase);
  g.m = uniform(x.m);
  g.n = uniform(x.n);
  g.symmetric = uniform(x.symmetric);
  g.iso_base = uniform(x.iso_base);
  g.obase = uniform(x.obase);
  g.nz = uniform(x.nz);
  g.pad = 0;
  return g;
}

// A group task on the team: the multi-source search of up to 64 queries of
// one graph, then their results into the host-mapped group block and the
// ring.
template <int BLOCK>
__device__ __forceinline__ void serve_group_task(const ServeArgs &s, const ServeRequest &d, SharedAll<BLOCK> &sh,
                                              Team &team, uint32_t id, uint64_t t0) {
  ServeCtl *c = s.ctl;
  GridBarrier *bar = &c->team[id].bar;
  const uint32_t tid = threadIdx.x;
    const ServeGraph g = uniform_graph(s.graphs[d.graph]);
    const uint32_t segs = serve_segments(s.xpt);
    const ServeGroupLayout gl = serve_group_layout(g.n, d.source, d.item_cap, segs);
    uint8_t *w = s.arena + d.work;
    MsArgs ms;
    ms.off = s.off_r + g.vbase;
    ms.col = s.col_r + g.ebase;
    ms.col_bu = nullptr;
    ms.o2r = s.o2r + g.obase;
    ms.nz = g.nz;
    ms.n = g.n;
    ms.count = d.source;
    ms.m = g.m;
    ms.sources = (const uint32_t *)(w + gl.sources);
    ms.level = (uint32_t *)(w + gl.level);
    ms.level_stride = gl.n_pad;
    ms.seen = (uint64_t *)(w + gl.seen);
    ms.f = (uint64_t *)(w + gl.f);
    ms.planes = (uint64_t *)(w + gl.planes);
    ms.qb = (uint32_t *)(w + gl.qb);
    ms.segs = segs;
    ms.items = (uint2 *)(w + gl.items);
    ms.item_cap = d.item_cap;
    ms.ctl = &c->team[id];
    ms.counters = (uint64_t *)(w + gl.counters);
    ms.depth = (uint32_t *)(w + gl.depth);
    ms.small_max = d.small_max;
    ms.medium_max = d.medium_max;
    ms.large_max = d.large_max;
    ms.bottom_up = ((d.flags & kServeDirOpt) && g.symmetric) ? 1u : 0u;
    ms.bu_div = s.bu_div;
    ms.prof_levels = 0;
    ms.prof = nullptr;
    ms_grid_search<BLOCK>(ms, sh, team);
    // outputs reach memory the host's copies read before the result appears
    grid_barrier<true>(bar, team.bst, 0, sh.bcast);
    if (sh.team_rank == 0 && tid == 0) {
      const uint64_t tk = c->ticket[id];
      const uint32_t slot = (uint32_t)(tk & (s.ring - 1));
      ServeGroupResult *gr = s.gres + (size_t)slot * kMsMaxSources;
      uint64_t reached = 0, edges = 0;
      for (uint32_t i = 0; i < d.source; ++i) {
        const uint64_t ri = ms.counters[i], ei = ms.counters[kMsMaxSources + i];
        __hip_atomic_store(&gr[i].reached, ri, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&gr[i].edges, ei, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&gr[i].depth, ms.depth[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        reached += ri;
        edges += ei;
      }
      ServeResult *o = &s.res[slot];
      o->reached = reached;
      o->depth = ms.ctl->depth;
      o->bu_levels = ms.ctl->bu_levels;
      o->edges = edges;
      o->team = id;
      o->err = ms.ctl->err;
      o->t_start_ns = t0;
      o->t_end_ns = now_ns();
      __hip_atomic_store(&o->done, tk + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// kGroups: the launch serves group tasks (multi-source searches) and copy
// tasks; otherwise single searches and copy tasks.  A launch carries only
// the search path it runs (both inlined together would not fit the
// registers of 1024-thread workgroups).
template <int BLOCK, bool kGroups>
__global__ __launch_bounds__(BLOCK) void dfq_serve_kernel(ServeArgs s) {
  __shared__ SharedAll<BLOCK> sh;
  Team team{};
  ServeCtl *c = s.ctl;
  if (!kernel_prologue<BLOCK>(&c->census, s.xpt, true, [c](uint32_t id) { return &c->team[id].bar; }, sh,
                              team)) {
    // a team would span more XCDs than there are frontier segments: report
    // no team, so the host sees the launch end without serving
    if (blockIdx.x == 0 && threadIdx.x == 0)
      __hip_atomic_store(&s.status->teams, 0xffffffffu, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  const uint32_t tid = threadIdx.x;
  const uint32_t id = sh.team_id;
  GridBarrier *bar = &c->team[id].bar;
  // every member has its ticket rank once this barrier opens
  grid_barrier(bar, team.bst, 0, sh.bcast);
  if (sh.team_rank == 0 && tid == 0) {
    team_ticket_reset(bar);
    __hip_atomic_store(&s.status->team_blocks[id], sh.team_size, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (id == 0) __hip_atomic_store(&s.status->teams, team.teams, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }

  for (;;) {
    uint64_t stop = 0, t = 0, t0 = 0;
    if (sh.team_rank == 0 && tid == 0) {
      t = __hip_atomic_fetch_add(&c->head.word, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const ServeRequest *r = &s.req[t & (s.ring - 1)];
      for (;;) {
        if (ld_system(&r->seq) == t + 1) break;
        if (ld_system(s.stop)) {
          stop = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if (!stop) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // system scope
        t0 = now_ns();
        ServeRequest d = *r;
        c->desc[id] = d;
        c->ticket[id] = t;
        if (d.flags & kServeGroup) {
          // the group's sources move from the host-mapped slot into the
          // working set (the team barrier below publishes them)
          const ServeGraph gg = s.graphs[d.graph];
          const ServeGroupLayout gl = serve_group_layout(gg.n, d.source, d.item_cap, serve_segments(s.xpt));
          uint32_t *dst = (uint32_t *)(s.arena + d.work + gl.sources);
          const uint32_t *src = s.gsrc + (size_t)(t & (s.ring - 1)) * kMsMaxSources;
          for (uint32_t i = 0; i < d.source; ++i) dst[i] = ld_system(&src[i]);
        }
      }
    }
    if (grid_barrier(bar, team.bst, stop, sh.bcast)) break;

    const ServeRequest d = uniform_request(c->desc[id]);
    if (d.flags & kServeCopy) {
      // all workgroups of the team stream the words out (arena ranges and
      // the staging buffer are 256-byte aligned)
      const uint32_t *src = (const uint32_t *)(s.arena + d.work);
      uint32_t *dst = (uint32_t *)d.dst;
      const uint64_t words = d.source, vec = words / 4;
      const uint64_t gtid = (uint64_t)sh.team_rank * BLOCK + tid, gstride = (uint64_t)sh.team_size * BLOCK;
      for (uint64_t i = gtid; i < vec; i += gstride) ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
      for (uint64_t i = vec * 4 + gtid; i < words; i += gstride) dst[i] = src[i];
      grid_barrier<true>(bar, team.bst, 0, sh.bcast);
      if (sh.team_rank == 0 && tid == 0) {
        const uint64_t tk = c->ticket[id];
        ServeResult *o = &s.res[tk & (s.ring - 1)];
        o->team = id;
        o->err = 0;
        o->t_start_ns = t0;
        o->t_end_ns = now_ns();
        __hip_atomic_store(&o->done, tk + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      continue;
    }
    if constexpr (kGroups) {
      serve_group_task<BLOCK>(s, d, sh, team, id, t0);
      continue;
    }
    const ServeGraph g = uniform_graph(s.graphs[d.graph]);
    const bool parents = (d.flags & kServeParents) != 0;
    const ServeLayout l = serve_layout(g.n, d.item_cap, serve_segments(s.xpt), parents);
    uint8_t *w = s.arena + d.work;
    BfsArgs a;
    a.off = s.off + g.vbase;
    a.col = s.col + g.ebase;
    a.col_bu = g.symmetric ? s.col_bu + g.ebase : nullptr;
    a.head = g.symmetric ? s.head + g.vbase : nullptr;
    a.level = (uint32_t *)(w + l.level);
    a.parent = parents ? (uint32_t *)(w + l.parent) : nullptr;
    a.iso = g.iso_base != ~0ull ? s.iso + g.iso_base : nullptr;
    a.vis = (uint32_t *)(w + l.vis);
    a.vis_snap = (uint32_t *)(w + l.snap);
    a.fb = (uint32_t *)(w + l.fb);
    a.qseg[0] = (uint32_t *)(w + l.qseg0);
    a.qseg[1] = (uint32_t *)(w + l.qseg1);
    a.items = (uint2 *)(w + l.items);
    a.item_cap = d.item_cap;
    a.ctl = &c->team[id];
    a.prof = nullptr;
    a.prof_levels = 0;
    a.result = nullptr;
    a.n = g.n;
    a.n_words = (g.n + 31) / 32;
    a.source = d.source;
    a.small_max = d.small_max;
    a.medium_max = d.medium_max;
    a.large_max = d.large_max;
    a.bottom_up = ((d.flags & kServeDirOpt) && g.symmetric) ? 1u : 0u;
    a.alpha = d.alpha;
    a.bu_exit = d.beta ? g.n / d.beta : 0;
    a.m = g.m;
    a.deg_max = 0;
    a.fuse_bits = fuse_bits_for(g.n, g.m);
    // the system-scope barrier right after writes every L2 back
    grid_bfs<BLOCK>(a, SearchIo{a.level, a.parent, a.result, a.source}, sh, team, GridBfsMode{false, true, 0});
    // outputs reach memory the host's copies read before the result appears
    grid_barrier<true>(bar, team.bst, 0, sh.bcast);
    if (sh.team_rank == 0 && tid == 0) {
      const uint64_t tk = c->ticket[id];
      ServeResult *o = &s.res[tk & (s.ring - 1)];
      o->reached = a.ctl->reached;
      o->depth = a.ctl->depth;
      o->bu_levels = a.ctl->bu_levels;
      o->edges = a.ctl->edges;
      o->team = id;
      o->err = a.ctl->err;
      o->t_start_ns = t0;
      o->t_end_ns = now_ns();
      __hip_atomic_store(&o->done, tk + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

}  // namespace dfq
