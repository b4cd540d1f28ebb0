// Batched DFQ-BFS: many BFS instances over many resident graphs in one
// persistent cooperative launch (no counterpart in the reference, which ran
// one search per process; the per-level machinery follows the plain kernel
// design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:98-179 as rebuilt in
// dfq_kernels.hpp).
//
// Two tiers, chosen per query by the host plan:
//   grid tier       a team of XCDs runs one instance at a time (grid_bfs,
//                   XCD-aware barrier per level) - for large graphs; with
//                   teams narrower than the grid (BatchArgs::xpt) several
//                   instances run at once, each team taking every
//                   teams-th one;
//   workgroup tier  one workgroup runs one instance: visited set and the
//                   current / next frontier bitmaps live in LDS, claims are
//                   LDS atomics, levels are separated by __syncthreads and the
//                   frontier queue is the workgroup's own partition of an HBM
//                   array.  Workgroups take instances from a shared counter
//                   (largest first) until none are left.
#pragma once

#include "dfq_kernels.hpp"

namespace dfq {

// Vertices a workgroup-tier instance may have (three LDS bitmaps of
// kWgMaxN / 8 bytes each).
constexpr uint32_t kWgMaxN = 1u << 17;
constexpr uint32_t kWgWords = kWgMaxN / 32;

struct BatchGraph {
  uint64_t vbase;      // first row offset of this graph in the concatenated offsets
  uint64_t ebase;      // first column index of this graph in the concatenated columns
  uint64_t m;          // CSR entries
  uint32_t n;          // vertices
  uint32_t symmetric;  // bottom-up levels allowed
  uint64_t iso_base;   // first word of its isolated-vertex bitmap, kNoIso if none
};

constexpr uint64_t kNoIso = ~0ull;

struct BatchQuery {
  uint64_t lbase;      // first entry of this query's level array in the output
  uint32_t graph;
  uint32_t source;
  uint32_t tier;       // 0 workgroup, 1 grid
  uint32_t small_max, medium_max, large_max;
  uint32_t direction_opt, alpha, beta;
  uint32_t pad;
};

struct BatchResult {
  uint32_t depth;
  uint32_t bu_levels;
  uint64_t reached;
};

struct BatchArgs {
  const int64_t *off;         // concatenated row offsets
  const uint32_t *col;        // concatenated columns
  const uint32_t *col_bu;     // the same, rows by neighbour degree (symmetric graphs)
  const uint4 *head;          // first four col_bu entries per vertex, indexed like off
  const BatchGraph *graphs;
  const BatchQuery *queries;  // grid-tier queries first, then workgroup tier
  BatchResult *results;
  uint32_t *level;            // concatenated level arrays
  uint32_t n_queries;
  uint32_t n_grid;            // queries [0, n_grid) are grid tier
  uint32_t *work;             // workgroup-tier instance counter
  uint32_t *wg_queue;         // per-workgroup frontier partitions [grid][2][kWgMaxN]
  const uint32_t *iso;        // concatenated isolated-vertex bitmaps (symmetric graphs)
  // scratch of the grid tier (sized for its largest graph), team 0's; team
  // t's arrays follow at t times the strides below, and its control block
  // is grid.ctl[t]; with teams, grid.ctl[teams_cap] is the census's
  BfsArgs grid;
  uint32_t xpt;               // XCDs per grid-tier team (0: the whole grid, one team)
  uint32_t teams_cap;         // scratch copies (teams the launch may form)
  uint32_t n_max;             // vertices of the largest graph (scratch strides)
  uint32_t pad;
};

// ---------------------------------------------------------------------------
// Workgroup tier
// ---------------------------------------------------------------------------
template <int BLOCK>
struct SharedWg {
  SharedA<BLOCK> a;                 // bin queues and gather scratch
  uint32_t vis[kWgWords];           // visited
  uint32_t fb[2][kWgWords];         // frontier bitmaps (current, next) by level parity
  uint32_t wave_sum[BLOCK / kWave];
  uint64_t mf;                      // out-degree sum of the next frontier
  uint32_t out_n;                   // next frontier size
  uint32_t med_n, big_n;
  uint32_t query;
};

template <int BLOCK>
union SharedBatch {
  SharedAll<BLOCK> grid;
  SharedWg<BLOCK> wg;
};

// Edge visitor of the workgroup tier (LDS visited set, LDS queue tail).
struct WgVisitor {
  const uint32_t *__restrict__ col;
  const int64_t *off;
  uint32_t *level;
  uint32_t *vis;       // LDS
  uint32_t *fb_next;   // LDS
  uint32_t *q_out;     // HBM partition
  uint32_t *out_n;     // LDS
  uint32_t next_level;
  bool track_mf;
  bool deg32;  // every degree < 2^32: the low words of two offsets give it
  uint64_t mf;

  __device__ __forceinline__ void append(bool mine, uint64_t mask, uint32_t w) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(out_n, (uint32_t)__popcll(mask));
    base = bcast_lane(base, 0);
    if (mine) q_out[base + mbcnt(mask)] = w;
  }

  // Claims in LDS; with deg32 the claimed vertices' row offsets (low
  // words) are summed after the whole batch, so their loads are waited for
  // once per batch instead of once per claim.
  template <int U>
  __device__ __forceinline__ void visit_batch(const bool (&act)[U], const uint32_t (&w)[U]) {
    const bool late = track_mf && deg32;
    uint32_t lo0[U], lo1[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
      bool claimed = false;
      lo0[i] = lo1[i] = 0;
      if (act[i]) {
        const uint32_t bit = 1u << (w[i] & 31u);
        if (!(vis[w[i] >> 5] & bit)) claimed = !(atomicOr(&vis[w[i] >> 5], bit) & bit);
        if (claimed) {
          level[w[i]] = next_level;
          atomicOr(&fb_next[w[i] >> 5], bit);
          if (late) {
            const uint32_t *o = reinterpret_cast<const uint32_t *>(off + w[i]);
            lo0[i] = o[0];
            lo1[i] = o[2];
          } else if (track_mf) {
            mf += (uint64_t)(off[w[i] + 1] - off[w[i]]);
          }
        }
      }
      const uint64_t mask = __ballot(claimed);
      if (mask) append(claimed, mask, w[i]);
    }
    if (late) {
#pragma unroll
      for (int i = 0; i < U; ++i) mf += (uint64_t)(lo1[i] - lo0[i]);  // 0 where nothing was claimed
    }
  }
};

template <int BLOCK>
__device__ __forceinline__ void wg_gather_small(WgVisitor &vis, uint32_t v, int64_t start, uint32_t deg,
                                                uint32_t *g_excl, int64_t *g_s) {
  const uint32_t lane = lane_id();
  const uint32_t incl = wave_inclusive_scan(deg);
  const uint32_t total = bcast_lane(incl, 63);
  if (total == 0) return;
  g_excl[lane] = incl - deg;
  g_s[lane] = start;
  (void)v;
  wave_sync();
  for (uint32_t b = 0; b < total; b += kWave * kUnroll) {
    bool act[kUnroll];
    uint32_t w[kUnroll];
#pragma unroll
    for (int i = 0; i < kUnroll; ++i) {
      const uint32_t e = b + i * kWave + lane;
      act[i] = e < total;
      uint32_t j = 0;
#pragma unroll
      for (uint32_t step = 32; step >= 1; step >>= 1)
        if (g_excl[j + step] <= e) j += step;
      w[i] = act[i] ? vis.col[g_s[j] + (int64_t)(e - g_excl[j])] : 0u;
    }
    vis.visit_batch<kUnroll>(act, w);
  }
  wave_sync();
}

// Expands `cnt` rows from LDS queue arrays, `width` lanes per row
// (64 = one wave per vertex, BLOCK = the whole workgroup per vertex).
template <int BLOCK, int WIDTH>
__device__ __forceinline__ void wg_expand_rows(WgVisitor &vis, const uint32_t *qd, const int64_t *qs,
                                               uint32_t cnt) {
  const uint32_t lane = WIDTH == BLOCK ? threadIdx.x : lane_id();
  const uint32_t first = WIDTH == BLOCK ? 0 : threadIdx.x / kWave;
  const uint32_t stride = WIDTH == BLOCK ? 1 : BLOCK / kWave;
  for (uint32_t j = first; j < cnt; j += stride) {
    const uint32_t d = qd[j];
    const int64_t s = qs[j];
    for (uint32_t k0 = 0; k0 < d; k0 += WIDTH * kUnroll) {
      bool act[kUnroll];
      uint32_t w[kUnroll];
#pragma unroll
      for (int i = 0; i < kUnroll; ++i) {
        const uint32_t k = k0 + i * WIDTH + lane;
        act[i] = k < d;
        w[i] = act[i] ? vis.col[s + k] : 0u;
      }
      vis.visit_batch<kUnroll>(act, w);
    }
  }
}

// One BFS by one workgroup.  off / col point at the instance's graph.
template <int BLOCK>
__device__ __forceinline__ void wg_bfs(SharedWg<BLOCK> &sh, const int64_t *off, const uint32_t *col, const uint32_t *col_bu,
                       uint32_t n, bool symmetric, const uint32_t *iso, const BatchQuery &q, uint32_t *level,
                       uint32_t *qbuf, BatchResult *res) {
  using S = KernelShape<BLOCK>;
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / kWave;
  const uint32_t words = (n + 31) / 32;
  const uint32_t src = q.source;
  for (uint32_t i = tid; i < n; i += BLOCK) level[i] = (i == src) ? 0u : kInf;
  for (uint32_t w = tid; w < words; w += BLOCK) {
    const uint32_t init = (w == (src >> 5)) ? (1u << (src & 31u)) : 0u;
    sh.vis[w] = init | (iso ? iso[w] : 0u);
    sh.fb[0][w] = init;
    sh.fb[1][w] = 0;
  }
  if (tid == 0) {
    qbuf[0] = src;
    sh.out_n = 0;
    sh.mf = 0;
  }
  __syncthreads();

  WgVisitor vis;
  vis.col = col;
  vis.off = off;
  vis.level = level;
  vis.vis = sh.vis;
  vis.out_n = &sh.out_n;
  vis.track_mf = q.direction_opt && symmetric;
  vis.deg32 = (uint64_t)(off[n] - off[0]) <= 0xffffffffull;  // no row can have 2^32 entries

  uint32_t n_cur = 1, n_prev = 0, bu_levels = 0;
  uint64_t reached = 1;
  uint64_t m_f = (uint64_t)(off[src + 1] - off[src]), m_seen = m_f;
  bool bottom_up = false;
  const uint32_t bu_exit = q.beta ? n / q.beta : 0;
  uint32_t L = 0;
  for (;; ++L) {
    const uint32_t *q_in = qbuf + (size_t)(L & 1) * kWgMaxN;
    uint32_t *fb_cur = sh.fb[L & 1];
    uint32_t *fb_next = sh.fb[(L + 1) & 1];
    vis.q_out = qbuf + (size_t)((L + 1) & 1) * kWgMaxN;
    vis.fb_next = fb_next;
    vis.next_level = L + 1;
    vis.mf = 0;
    if (vis.track_mf) {
      const uint64_t m_u = m_seen < (uint64_t)off[n] ? (uint64_t)off[n] - m_seen : 0;
      if (!bottom_up && n_cur > n_prev && m_f * q.alpha > m_u) bottom_up = true;
      else if (bottom_up && n_cur < bu_exit && n_cur < n_prev) bottom_up = false;
    }
    if (!bottom_up) {
      if (tid == 0) {
        sh.med_n = 0;
        sh.big_n = 0;
      }
      __syncthreads();
      for (uint32_t base = 0; base < n_cur; base += BLOCK) {
        const uint32_t p = base + tid;
        uint32_t v = 0, deg = 0;
        int64_t start = 0;
        if (p < n_cur) {
          v = q_in[p];
          start = off[v];
          deg = (uint32_t)(off[v + 1] - start);
        }
        const bool is_small = deg != 0 && deg <= q.small_max;
        const bool is_med = deg > q.small_max && deg <= q.medium_max;
        const bool is_big = deg > q.medium_max;
        const uint64_t mm = __ballot(is_med);
        if (mm) {
          uint32_t b0 = 0;
          if (lane_id() == 0) b0 = atomicAdd(&sh.med_n, (uint32_t)__popcll(mm));
          b0 = bcast_lane(b0, 0);
          if (is_med) {
            sh.a.med_d[b0 + mbcnt(mm)] = deg;
            sh.a.med_s[b0 + mbcnt(mm)] = start;
          }
        }
        const uint64_t bm = __ballot(is_big);
        if (bm) {
          uint32_t b0 = 0;
          if (lane_id() == 0) b0 = atomicAdd(&sh.big_n, (uint32_t)__popcll(bm));
          b0 = bcast_lane(b0, 0);
          if (is_big) {
            sh.a.big_d[b0 + mbcnt(bm)] = deg;
            sh.a.big_s[b0 + mbcnt(bm)] = start;
          }
        }
        wg_gather_small<BLOCK>(vis, v, is_small ? start : 0, is_small ? deg : 0u, sh.a.g_excl[wave],
                               sh.a.g_s[wave]);
        __syncthreads();
        const uint32_t mn = sh.med_n, bn = sh.big_n;
        __syncthreads();
        if (mn > (uint32_t)(S::kMedCap - BLOCK)) {
          wg_expand_rows<BLOCK, kWave>(vis, sh.a.med_d, sh.a.med_s, mn);
          __syncthreads();
          if (tid == 0) sh.med_n = 0;
        }
        if (bn > (uint32_t)(S::kBigCap - BLOCK)) {
          wg_expand_rows<BLOCK, BLOCK>(vis, sh.a.big_d, sh.a.big_s, bn);
          __syncthreads();
          if (tid == 0) sh.big_n = 0;
        }
        __syncthreads();
      }
      const uint32_t mn = sh.med_n, bn = sh.big_n;
      wg_expand_rows<BLOCK, kWave>(vis, sh.a.med_d, sh.a.med_s, mn);
      wg_expand_rows<BLOCK, BLOCK>(vis, sh.a.big_d, sh.a.big_s, bn);
    } else {
      ++bu_levels;
      // every unvisited vertex looks for a neighbour in the current
      // frontier; 64-vertex groups per wave own their bitmap words
      const uint32_t groups = (n + 63) / 64;
      for (uint32_t g = wave; g < groups; g += BLOCK / kWave) {
        const uint32_t v = g * 64 + lane_id();
        const uint32_t w0 = sh.vis[2 * g];
        const uint32_t w1 = 2 * g + 1 < words ? sh.vis[2 * g + 1] : 0xffffffffu;
        const uint64_t seen = ((uint64_t)w1 << 32) | w0;
        bool claimed = false;
        if (v < n && !((seen >> lane_id()) & 1ull)) {
          const int64_t s = off[v], e = off[v + 1];
          for (int64_t k = s; k < e && !claimed; k += kUnroll) {
            uint32_t w[kUnroll];
#pragma unroll
            for (int i = 0; i < kUnroll; ++i) w[i] = k + i < e ? col_bu[k + i] : 0xffffffffu;
#pragma unroll
            for (int i = 0; i < kUnroll; ++i)
              if (w[i] != 0xffffffffu && ((fb_cur[w[i] >> 5] >> (w[i] & 31u)) & 1u)) claimed = true;
          }
          if (claimed) {
            level[v] = L + 1;
            vis.mf += (uint64_t)(e - s);
          }
        }
        const uint64_t mask = __ballot(claimed);
        if (mask) {
          if (lane_id() == 0) {
            sh.vis[2 * g] = w0 | (uint32_t)mask;
            fb_next[2 * g] |= (uint32_t)mask;
          }
          if (lane_id() == 1 && 2 * g + 1 < words) {
            sh.vis[2 * g + 1] = w1 | (uint32_t)(mask >> 32);
            fb_next[2 * g + 1] |= (uint32_t)(mask >> 32);
          }
          vis.append(claimed, mask, v);
        }
      }
    }
    // next frontier size and out-degree sum
    uint64_t mfl = vis.mf;
    for (int d = 32; d >= 1; d >>= 1) mfl += __shfl_xor(mfl, d);
    if (lane_id() == 0 && mfl) atomicAdd((unsigned long long *)&sh.mf, (unsigned long long)mfl);
    __syncthreads();
    const uint32_t n_next = sh.out_n;
    const uint64_t mf_next = sh.mf;
    __syncthreads();
    // clear the bitmap that becomes the next level's "next"
    for (uint32_t w = tid; w < words; w += BLOCK) fb_cur[w] = 0;
    if (tid == 0) {
      sh.out_n = 0;
      sh.mf = 0;
    }
    __syncthreads();
    reached += n_next;
    if (n_next == 0) break;
    m_f = mf_next;
    m_seen += mf_next;
    n_prev = n_cur;
    n_cur = n_next;
  }
  if (tid == 0) {
    res->depth = L + 1;
    res->reached = reached;
    res->bu_levels = bu_levels;
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void dfq_batch_kernel(BatchArgs b) {
  __shared__ SharedBatch<BLOCK> sh;
  Team team{};
  Control *ctl = b.grid.ctl;
  const bool teams = b.xpt != 0;
  GridBarrier *census = teams ? &ctl[b.teams_cap].bar : &ctl[0].bar;
  // the workgroup tier's counter only grows: each launch counts from the
  // value read here, before this workgroup is counted by the census, which
  // no workgroup passes (and so takes an instance) before every one is
  // counted
  uint32_t work_base = 0;
  if (threadIdx.x == 0) work_base = ld_relaxed(b.work);
  if (!kernel_prologue<BLOCK>(census, b.xpt, teams, [ctl](uint32_t id) { return &ctl[id].bar; }, sh.grid,
                              team)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[0].err = 2;
    return;
  }
  const uint32_t id = uniform(sh.grid.team_id);
  if (threadIdx.x == 0) sh.grid.bcast[0] = team.teams;
  __syncthreads();
  const uint32_t nteams = uniform((uint32_t)sh.grid.bcast[0]);
  __syncthreads();
  GridBarrier *bar = &ctl[id].bar;
  // every member has its rank once this barrier opens (teams rank their
  // workgroups by ticket)
  grid_barrier(bar, team.bst, 0, sh.grid.bcast);
  if (teams && sh.grid.team_rank == 0 && threadIdx.x == 0) team_ticket_reset(bar);

  // grid tier: this team's instances, one at a time over the team
  const size_t nw = ((size_t)b.n_max + 31) / 32;
  for (uint32_t qi = id; qi < b.n_grid; qi += nteams) {
    const BatchQuery q = b.queries[qi];
    const BatchGraph gr = b.graphs[q.graph];
    BfsArgs a = b.grid;
    a.vis += id * nw;
    a.vis_snap += id * nw;
    a.fb += id * 2 * (nw + 1);
    a.qseg[0] += id * (size_t)kSegments * b.n_max;
    a.qseg[1] += id * (size_t)kSegments * b.n_max;
    a.items += id * (size_t)kSegments * a.item_cap;
    a.ctl = ctl + id;
    a.off = b.off + gr.vbase;
    a.col = b.col + gr.ebase;
    a.col_bu = gr.symmetric ? b.col_bu + gr.ebase : nullptr;
    a.head = gr.symmetric ? b.head + gr.vbase : nullptr;
    a.level = b.level + q.lbase;
    a.parent = nullptr;
    a.n = gr.n;
    a.n_words = (gr.n + 31) / 32;
    a.m = gr.m;
    a.fuse_bits = fuse_bits_for(gr.n, gr.m);
    a.iso = gr.iso_base != kNoIso ? b.iso + gr.iso_base : nullptr;
    a.source = q.source;
    a.small_max = q.small_max;
    a.medium_max = q.medium_max;
    a.large_max = q.large_max;
    a.bottom_up = (q.direction_opt && gr.symmetric) ? 1u : 0u;
    a.alpha = q.alpha;
    a.bu_exit = q.beta ? gr.n / q.beta : 0;
    grid_bfs<BLOCK>(a, SearchIo{a.level, a.parent, a.result, a.source}, sh.grid, team);
    if (sh.grid.team_rank == 0 && threadIdx.x == 0) {
      b.results[qi].depth = a.ctl->depth;
      b.results[qi].reached = a.ctl->reached;
      b.results[qi].bu_levels = a.ctl->bu_levels;
    }
  }
  // the grid tier's level barriers leave caches as they are (grid_bfs
  // invalidates before its own shared reads only)
  if (b.n_grid && threadIdx.x == 0) l2_invalidate();

  // workgroup tier: instances from the shared counter, one per workgroup
  uint32_t *qbuf = b.wg_queue + (size_t)blockIdx.x * 2 * kWgMaxN;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0)
      sh.wg.query = b.n_grid + (__hip_atomic_fetch_add(b.work, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                                work_base);
    __syncthreads();
    const uint32_t qi = sh.wg.query;
    if (qi >= b.n_queries) break;
    const BatchQuery q = b.queries[qi];
    const BatchGraph gr = b.graphs[q.graph];
    wg_bfs<BLOCK>(sh.wg, b.off + gr.vbase, b.col + gr.ebase,
                  (gr.symmetric ? b.col_bu : b.col) + gr.ebase, gr.n, gr.symmetric != 0,
                  gr.iso_base != kNoIso ? b.iso + gr.iso_base : nullptr, q,
                  b.level + q.lbase, qbuf, &b.results[qi]);
  }
}

}  // namespace dfq
