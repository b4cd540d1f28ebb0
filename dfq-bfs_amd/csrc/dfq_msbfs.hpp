// Multi-source search: up to 64 BFS of one graph share one traversal.
//
// No counterpart in the reference (one search per launch, design/shared-q/
// atomic-sa/our-sep/upper-bound/bfs.cu:98-179; its interactive loop :202-214,
// 302-308 re-runs the search per source).  The level-synchronous scheme is
// the bit-parallel multi-source BFS of Then, Kaufmann, Chirigati, Hoang-Vu,
// Pham, Kemper, Neumann and Vo, "The More the Merrier: Efficient Multi-Source
// Graph Traversal", PVLDB 8(4), 2014: search q owns bit q of a 64-bit mask
// per vertex, so one pass over an edge serves every search that has the
// edge's tail in its frontier.  Each search's levels are its own BFS
// distances, identical to a single-source search's.
//
// Per vertex (HBM):
//   seen[v]        searches that have reached v (levels <= L during level L)
//   f[L & 1][v]    searches with v in their level-L frontier; nonzero exactly
//                  for the vertices of the level-L queue, zero elsewhere
//   planes[b][v]   bit b of the level at which each search reached v (the
//                  levels stay packed until a final pass writes each search's
//                  uint32 level array with full-line stores)
// plus three rotating frontier queues (per-XCD segments, as in the
// single-source kernel), since a level's masks are cleared one level later
// through the queue that lists them.
//
// Per level L:
//   commit   over the level-L queue: seen |= f_L, the level's bits into the
//            planes; the previous queue's masks cleared; the level's active
//            searches (OR of f_L) and frontier out-degree sum reduced
//   expand   top-down through the single-source kernel's degree bins and
//            work items with a visitor that moves masks (new = f[u] & ~seen[w],
//            one 64-bit atomic OR, the 0 -> nonzero transition appends w to
//            the next queue), or, on symmetric graphs with a large frontier,
//            bottom-up: every vertex missing some active search ORs its
//            neighbours' frontier masks until it has all of them (no atomics)
#pragma once

#include "dfq_kernels.hpp"

namespace dfq {

constexpr uint32_t kMsMaxSources = 64;
// Level bit-planes: a multi-source search covers levels 0..255; a search
// that would go deeper stops with ctl->err = kMsErrDepth (callers run such
// sources one by one).
constexpr uint32_t kMsPlanes = 8;
constexpr uint32_t kMsErrDepth = 3;
// Bottom-up levels (symmetric graphs) when the frontier's out-edges exceed
// m / kMsBuDiv: a bottom-up level of many searches reads most rows to their
// end (a row stops early only once every missing search is found), so it
// pays only for frontiers that reach a large part of the edges.
constexpr uint32_t kMsBuDiv = 2;

struct MsArgs {
  const int64_t *off;      // row offsets [n + 1]
  const uint32_t *col;     // columns [m] (top-down levels)
  const uint32_t *col_bu;  // rows by neighbour degree (bottom-up levels), or null
  const uint32_t *o2r;     // the graph above is the degree-ranked copy: vertex v is o2r[v]
                           // (sources and output stay in the caller's ids); null: identity
  uint32_t nz;             // bottom-up levels look at vertices [0, nz) only: in the ranked
                           // copy the vertices without entries come last (n otherwise)
  uint32_t n;
  uint32_t count;          // searches, 1..64
  uint64_t m;
  const uint32_t *sources; // [count]
  uint32_t *level;         // search q writes level + q * level_stride [n]
  uint64_t level_stride;
  uint64_t *seen;          // [n]
  uint64_t *f;             // [2][n]
  uint64_t *planes;        // [kMsPlanes][n]
  uint32_t *qb;            // [3][kSegments or team segments][n]
  uint32_t segs;           // segments per queue buffer
  uint2 *items;            // work items of huge rows, segs x item_cap
  uint32_t item_cap;
  Control *ctl;
  uint64_t *counters;      // [2][64]: vertices reached and CSR entries of them, per search
  uint32_t *depth;         // [64] levels per search (max level + 1)
  uint32_t small_max, medium_max, large_max;
  uint32_t bottom_up;      // bottom-up levels allowed (symmetric graph)
  uint32_t bu_div;         // bottom-up when the frontier's out-edges exceed m / bu_div
  uint32_t prof_levels;    // rows of prof
  uint64_t *prof;          // [prof_levels][kMsProfFields] per-level counters, or null
};

// Per-level profile row of a multi-source search.
enum MsProfField : int {
  kMsProfFrontier = 0,  // vertices in the level's queue (distinct, over all searches)
  kMsProfEdges,         // their out-degree sum
  kMsProfBottomUp,      // 1 when the level expanded bottom-up
  kMsProfActive,        // searches with a frontier at this level
  kMsProfItems,         // work items (top-down huge rows or bottom-up long rows)
  kMsProfCommitNs,      // commit phase, barrier included (team's first workgroup)
  kMsProfExpandNs,      // expansion, barriers included
  kMsProfNext,          // vertices the expansion put in the next queue
  kMsProfFields = 8
};

// Top-down edge visitor of the multi-source search (the single-source
// kernel's phases call visit_batch / append / flush through it).
struct MsVisitor {
  const uint32_t *__restrict__ col;
  const uint64_t *seen;
  const uint64_t *fcur;   // frontier masks of the level being expanded
  uint64_t *fnxt;         // masks of the next level (zero at level start)
  uint32_t *seg;          // output segment of this workgroup's XCD
  uint32_t *seg_tail;
  uint32_t *stage;        // this wave's LDS staging buffer [kStage]
  uint32_t out_n;
  uint32_t claims;
  uint64_t mf;            // unused here (the commit pass sums degrees)

  __device__ __forceinline__ void flush_wave() {
    const uint32_t cnt = out_n;
    uint32_t base = 0;
    if (lane_id() == 0) base = __hip_atomic_fetch_add(seg_tail, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    base = bcast_lane(base, 0);
    wave_sync();
    for (uint32_t k = lane_id(); k < cnt; k += kWave) seg[base + k] = stage[k];
    wave_sync();
    out_n = 0;
  }

  __device__ __forceinline__ void append(bool mine, uint64_t mask, uint32_t w) {
    if (mine) stage[out_n + mbcnt(mask)] = w;
    const uint32_t cnt = (uint32_t)__popcll(mask);
    out_n += cnt;
    claims += cnt;
    if (out_n > kStage - kWave) flush_wave();
  }

  // U edges u[i] -> w[i] per lane: the searches of u's frontier that have
  // not reached w move to w's next-level mask.  A plain load of that mask
  // first skips the atomic when it already holds the bits (it only grows
  // during the level, so a stale copy can only hold fewer).  The atomic that
  // turns the mask from zero to nonzero appends w to the next queue.
  // kLive: no meaning here (the single-search visitor's live-bitmap probe).
  template <int U, bool kEarlyDeg, bool kLive = false>
  __device__ __forceinline__ void visit_batch(const bool (&act)[U], const uint32_t (&u)[U],
                                              const uint32_t (&w)[U]) {
    uint64_t nw[U], cur[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const uint64_t fu = act[i] ? fcur[u[i]] : 0ull;
      const uint64_t sw = act[i] ? seen[w[i]] : ~0ull;
      cur[i] = act[i] ? fnxt[w[i]] : 0ull;
      nw[i] = fu & ~sw;
    }
    uint64_t old[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
      old[i] = ~0ull;
      if (nw[i] & ~cur[i])
        old[i] = __hip_atomic_fetch_or(&fnxt[w[i]], nw[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const bool claimed = old[i] == 0ull;
      const uint64_t mask = __ballot(claimed);
      if (mask) append(claimed, mask, w[i]);
    }
  }
};

// Long rows are not scanned by one lane in bottom-up levels: a row must be
// read to its end whenever one of the missing searches reaches the vertex
// later, and one lane on a hub's row would hold up the level.  In the first
// bottom-up level of a traversal rows rarely finish early, so rows longer
// than kItem are cut into kItem-entry work items outright.  In later ones
// most rows find every missing search among their first (highest-degree)
// neighbours, so a row longer than kMsBuRow gets a probe - one wave reads
// its first kMsBuRow entries at once - and only a row still missing
// searches after it is cut into items for the rest.  (Probes of kItem
// entries in the first level measured slower: RMAT-22 1.71 -> 2.76 ms.)
constexpr uint32_t kMsBuRow = kWave;
// Items a wave pulls per trip (their loads in flight together); probes of
// kMsBuRow entries go kItem / kMsBuRow times as many per trip.
constexpr int kMsItemBatch = 4;

// Work-item slots per segment a multi-source search needs: top-down levels
// cut rows longer than large_max into kItem-entry items; a bottom-up level
// puts one probe per row longer than kMsBuRow at the front of the segment
// and the rest-of-row items at its back.
__host__ __device__ inline uint64_t ms_item_cap(uint32_t n, uint64_t m, uint32_t large_max) {
  uint64_t rows_td = m / ((uint64_t)large_max + 1);
  if (rows_td > n) rows_td = n;
  uint64_t rows_bu = m / ((uint64_t)kMsBuRow + 1);
  if (rows_bu > n) rows_bu = n;
  const uint64_t td = m / kItem + rows_td + 1;
  const uint64_t bu = rows_bu + (m / kItem + rows_bu) + 2;
  return td > bu ? td : bu;
}

// Bottom-up level of the multi-source search.  Each wave takes chunks of
// kBuChunk vertices, lists those missing some active search (alive &
// ~seen) in LDS, then every lane streams through the list as a three-stage
// pipeline - row offsets with the vertex's seen mask, an aligned block of
// four column indices, the four neighbours' frontier masks - until the
// vertex has collected every missing search or its row ends.  Each trip
// issues the loads of every stage from straight-line code (lanes with
// nothing to load for a kind read one shared address), so one trip costs
// one memory round trip.  The searches found become the vertex's next
// mask, written by its owner lane, and the vertex joins the next queue.
// Long rows go elsewhere: with `probe`, rows longer than kMsBuRow get a
// probe (front of `items`, count in item_tail), run by ms_bottomup_probes
// after a barrier; otherwise rows longer than kItem are cut into items for
// the whole row (back of `items`, count in item2_tail), pulled by
// ms_bottomup_items.
template <int BLOCK>
__device__ __forceinline__ void ms_bottomup_level(const MsArgs &a, SharedAll<BLOCK> &sh, MsVisitor &vis,
                                                  uint64_t alive, uint32_t wave, uint32_t *chunk_next,
                                                  uint2 *items, uint32_t *item_tail, uint32_t *item2_tail,
                                                  bool probe) {
  const int64_t long_len = probe ? (int64_t)kMsBuRow : (int64_t)kItem;
  const uint32_t lane = lane_id();
  const uint32_t *colb = a.col_bu ? a.col_bu : a.col;
  const uint64_t *fcur = vis.fcur;
  uint64_t *fnxt = vis.fnxt;
  uint32_t *cand = sh.u.cand[wave];
  // chunks of [0, nz) claimed one at a time from a counter of the level: in
  // the ranked copy the chunks of high-degree vertices come first and the
  // cheap ones last, so the waves finish together
  for (;;) {
    uint32_t c = 0;
    if (lane == 0) c = __hip_atomic_fetch_add(chunk_next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t base = bcast_lane(c, 0) * kBuChunk;
    if (base >= a.nz) break;
    const uint32_t r1 = min(base + kBuChunk, a.nz);
    uint32_t total = 0;
#pragma unroll
    for (uint32_t g0 = 0; g0 < kBuChunk / kWave; g0 += 4) {
      uint64_t sv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t w = base + (g0 + i) * kWave + lane;
        sv[i] = w < r1 ? a.seen[w] : ~0ull;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool want = (alive & ~sv[i]) != 0ull;
        const uint64_t mk = __ballot(want);
        if (want) cand[total + mbcnt(mk)] = base + (g0 + i) * kWave + lane;
        total += (uint32_t)__popcll(mk);
      }
    }
    wave_sync();
    uint32_t cursor = 0, stage = 0, v = 0;
    int64_t k = 0, e = 0, kb = 0;
    uint64_t miss = 0, got = 0;
    uint32_t cw[4] = {0u, 0u, 0u, 0u};
    for (;;) {
      const uint64_t need = __ballot(stage == 0);
      const uint32_t ci = cursor + mbcnt(need);
      cursor += (uint32_t)__popcll(need);
      if (stage == 0) {
        if (ci < total) v = cand[ci];
        else stage = 3;
      }
      if (!__ballot(stage != 3)) break;
      const bool st0 = stage == 0, st1 = stage == 1, st2 = stage == 2;
      const int64_t *po = a.off + (st0 ? v : 0u);
      const uint64_t *ps = a.seen + (st0 ? v : 0u);
      const uint32_t *pk = colb + (st1 ? k : 0);
      const uint32_t back = (uint32_t)(((uintptr_t)pk & 15) >> 2);
      const uint4 *pq = (const uint4 *)(pk - back);
      uint32_t fi[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) fi[i] = (st2 && kb + i >= k && kb + i < e) ? cw[i] : 0u;
      const int64_t o0 = po[0], o1 = po[1];
      const uint64_t sv = *ps;
      const uint4 q = *pq;
      uint64_t fw[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) fw[i] = fcur[fi[i]];
      // next state by selects only (one memory wait per trip)
      uint64_t acc = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc |= ((kb + i >= k) & (kb + i < e)) ? fw[i] : 0ull;
      const uint64_t got2 = got | (acc & miss);
      const int64_t kn = kb + 4;
      const bool fin2 = st2 && (got2 == miss || kn >= e);
      const bool claimed = fin2 && got2 != 0ull;
      const uint64_t miss0 = alive & ~sv;
      // a long row gets a probe or its items; this lane takes the next candidate
      const bool long_row = st0 && miss0 != 0ull && o1 - o0 > long_len;
      if (long_row) {
        if (probe) {
          const uint32_t b0 = __hip_atomic_fetch_add(item_tail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          items[b0] = make_uint2(v, 0u);
        } else {
          const uint32_t nit = (uint32_t)((o1 - o0 + kItem - 1) / kItem);
          const uint32_t b0 = __hip_atomic_fetch_add(item2_tail, nit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          for (uint32_t j = 0; j < nit; ++j) items[a.item_cap - 1 - (b0 + j)] = make_uint2(v, j);
        }
      }
      const uint32_t next0 = (o1 > o0 && miss0 != 0ull && !long_row) ? 1u : 0u;
      stage = st0 ? next0 : (st1 ? 2u : (st2 ? (fin2 ? 0u : 1u) : stage));
      cw[0] = st1 ? q.x : cw[0];
      cw[1] = st1 ? q.y : cw[1];
      cw[2] = st1 ? q.z : cw[2];
      cw[3] = st1 ? q.w : cw[3];
      kb = st1 ? k - (int64_t)back : kb;  // CSR index of cw[0]
      k = st0 ? o0 : (st2 ? kn : k);
      e = st0 ? o1 : e;
      miss = st0 ? miss0 : miss;
      got = st0 ? 0ull : (st2 ? got2 : got);
      if (claimed) fnxt[v] = got2;
      const uint64_t mask = __ballot(claimed);
      if (mask) vis.append(claimed, mask, v);
    }
    wave_sync();
  }
}

// The probes of a bottom-up level's long rows: `total` probes located
// through pre (per-segment counts, from the front of each segment),
// round-robin over every wave of the team, several per trip.  A wave ORs
// the frontier masks of the row's first PL x 64 neighbours, keeps the vertex's missing searches and writes them as its next
// mask (no other wave touches the vertex in this phase), appending the
// vertex to the next queue.  A row still missing searches after its probe
// puts kItem-entry items for the rest at the back of this workgroup's
// segment (count in item2_tail).
template <int BLOCK, int PL>
__device__ __forceinline__ void ms_bottomup_probes(const MsArgs &a, MsVisitor &vis, uint64_t alive,
                                                   const uint32_t *pre, uint32_t total, uint32_t gwave,
                                                   uint32_t nwaves, uint2 *seg_items, uint32_t *item2_tail) {
  // PL loads per lane: probes of PL x 64 entries, 2 x kUnroll loads per lane per trip
  constexpr int B = 2 * kUnroll / PL;
  constexpr uint32_t P = (uint32_t)PL * kWave;
  const uint32_t lane = lane_id();
  const uint32_t *colb = a.col_bu ? a.col_bu : a.col;
  for (uint32_t i0 = gwave; i0 < total; i0 += B * nwaves) {
    uint32_t v[B], len[B], deg[B];
    int64_t s0[B];
    uint64_t miss[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const uint32_t i = i0 + (uint32_t)b * nwaves;
      v[b] = 0;
      len[b] = 0;
      deg[b] = 0;
      s0[b] = 0;
      miss[b] = 0;
      if (i < total) {  // wave-uniform
        uint32_t j = 0;
#pragma unroll
        for (int k = 1; k < kSegments; ++k) j += pre[k] <= i;
        const uint32_t x = a.items[(size_t)j * a.item_cap + (i - pre[j])].x;
        v[b] = x;
        s0[b] = a.off[x];
        deg[b] = (uint32_t)(a.off[x + 1] - s0[b]);
        len[b] = min(deg[b], P);
        miss[b] = alive & ~a.seen[x];
      }
    }
    uint32_t c[B][PL];
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
      for (int k = 0; k < PL; ++k) {
        const uint32_t e = (uint32_t)k * kWave + lane;
        c[b][k] = e < len[b] ? colb[s0[b] + e] : 0u;
      }
    uint64_t acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      acc[b] = 0;
#pragma unroll
      for (int k = 0; k < PL; ++k) {
        const uint32_t e = (uint32_t)k * kWave + lane;
        const uint64_t fw = vis.fcur[c[b][k]];
        acc[b] |= e < len[b] ? fw : 0ull;
      }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      uint64_t x = acc[b];
      for (int d = 32; d >= 1; d >>= 1) x |= __shfl_xor(x, d);
      const uint64_t nb = x & miss[b];
      if (lane == 0 && nb) vis.fnxt[v[b]] = nb;
      const bool claimed = lane == 0 && nb != 0ull;
      const uint64_t mask = __ballot(claimed);
      if (mask) vis.append(claimed, mask, v[b]);
      if (nb != miss[b] && deg[b] > P) {  // wave-uniform
        const uint32_t nit = (deg[b] - P + kItem - 1) / kItem;
        uint32_t base = 0;
        if (lane == 0) base = __hip_atomic_fetch_add(item2_tail, nit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        base = bcast_lane(base, 0);
        for (uint32_t k = lane; k < nit; k += kWave) seg_items[a.item_cap - 1 - (base + k)] = make_uint2(v[b], k);
      }
    }
  }
}

// Work items of long rows: `total` items located through pre (per-segment
// counts, from the back of each segment), round-robin over every wave of
// the team.  Item (v, k) covers entries first + k * kItem .. of v's row
// (first = the level's probe length).  A wave ORs the
// frontier masks of the item's neighbours, keeps the vertex's missing
// searches, and merges them into the vertex's next mask with one atomic;
// the transition from zero appends the vertex to the next queue.  Items of
// a vertex that already holds every missing search skip their loads (the
// mask only grows, so a stale copy can only hold fewer).
template <int BLOCK>
__device__ __forceinline__ void ms_bottomup_items(const MsArgs &a, MsVisitor &vis, uint64_t alive,
                                                  const uint32_t *pre, uint32_t total, uint32_t gwave,
                                                  uint32_t nwaves, uint32_t first) {
  constexpr int B = kMsItemBatch;
  const uint32_t lane = lane_id();
  const uint32_t *colb = a.col_bu ? a.col_bu : a.col;
  for (uint32_t i0 = gwave; i0 < total; i0 += B * nwaves) {
    // B items per trip (items i0, i0 + nwaves, ...): their column and mask
    // loads are all in flight together
    uint32_t v[B], len[B];
    int64_t s0[B];
    uint64_t miss[B], have[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const uint32_t i = i0 + (uint32_t)b * nwaves;
      v[b] = 0;
      len[b] = 0;
      s0[b] = 0;
      miss[b] = 0;
      have[b] = 0;
      if (i < total) {  // wave-uniform
        uint32_t j = 0;
#pragma unroll
        for (int k = 1; k < kSegments; ++k) j += pre[k] <= i;
        const uint2 it = a.items[(size_t)j * a.item_cap + a.item_cap - 1 - (i - pre[j])];
        v[b] = it.x;
        s0[b] = a.off[it.x] + first + (int64_t)it.y * kItem;
        len[b] = (uint32_t)min((int64_t)kItem, a.off[it.x + 1] - s0[b]);
        miss[b] = alive & ~a.seen[it.x];
        have[b] = vis.fnxt[it.x];
        // a vertex that already holds every missing search skips the item
        if ((have[b] & miss[b]) == miss[b]) len[b] = 0;
      }
    }
    uint32_t c[B][kUnroll];
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) {
        const uint32_t e = (uint32_t)k * kWave + lane;
        c[b][k] = e < len[b] ? colb[s0[b] + e] : 0u;
      }
    uint64_t acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      acc[b] = 0;
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) {
        const uint32_t e = (uint32_t)k * kWave + lane;
        const uint64_t fw = vis.fcur[c[b][k]];
        acc[b] |= e < len[b] ? fw : 0ull;
      }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      uint64_t x = acc[b];
      for (int d = 32; d >= 1; d >>= 1) x |= __shfl_xor(x, d);
      const uint64_t nb = x & miss[b] & ~have[b];
      bool claimed = false;
      if (lane == 0 && nb)
        claimed = __hip_atomic_fetch_or(&vis.fnxt[v[b]], nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull;
      const uint64_t mask = __ballot(claimed);
      if (mask) vis.append(claimed, mask, v[b]);
    }
  }
}

// Entry i of a queue buffer whose segments (stride n) hold pre[k+1]-pre[k]
// entries each.
__device__ __forceinline__ uint32_t ms_queue_at(const uint32_t *qbuf, const uint32_t *pre, uint32_t n, uint32_t i) {
  uint32_t j = 0;
#pragma unroll
  for (int k = 1; k < kSegments; ++k) j += pre[k] <= i;
  return qbuf[(size_t)j * n + (i - pre[j])];
}

// One multi-source search over a team (every workgroup of the team calls
// it with the same arguments).  Results: the level arrays, a.counters,
// a.depth, ctl->depth (levels of the whole traversal), ctl->bu_levels,
// ctl->err.  The caller's next grid barrier publishes the outputs.
template <int BLOCK>
__device__ __forceinline__ void ms_grid_search(const MsArgs &a, SharedAll<BLOCK> &sh, Team &team) {
  using S = KernelShape<BLOCK>;
  __shared__ uint64_t ms_red[2 * kMsMaxSources];  // per-search reached / entries of this workgroup
  __shared__ uint64_t ms_or;
  __shared__ uint32_t ms_pre[kSegments + 1];
  const uint32_t tid = threadIdx.x, lane = lane_id();
  const uint32_t wave = uniform(tid / kWave);
  const uint32_t trank = uniform(sh.team_rank), tsize = uniform(sh.team_size);
  const uint32_t gtid = trank * BLOCK + tid, gstride = tsize * BLOCK;
  const uint32_t gwave = trank * S::kWaves + wave, nwaves = tsize * S::kWaves;
  BarrierState &bst = team.bst;
  Control *ctl = a.ctl;
  const uint32_t n = a.n;
  const uint64_t qstride = (uint64_t)a.segs * n;  // one queue buffer

  // single-source phase arguments for the top-down expansion
  BfsArgs ba;
  ba.off = a.off;
  ba.col = a.col;
  ba.col_bu = nullptr;
  ba.head = nullptr;
  ba.level = nullptr;
  ba.parent = nullptr;
  ba.iso = nullptr;
  ba.vis = nullptr;
  ba.vis_snap = nullptr;
  ba.fb = nullptr;
  ba.qseg[0] = ba.qseg[1] = nullptr;
  ba.items = a.items;
  ba.item_cap = a.item_cap;
  ba.ctl = ctl;
  ba.prof = nullptr;
  ba.result = nullptr;
  ba.plain = 0;
  ba.plain_wait_ns = 0;
  ba.prof_levels = 0;
  ba.n = n;
  ba.n_words = (n + 31) / 32;
  ba.source = 0;
  ba.small_max = a.small_max;
  ba.medium_max = a.medium_max;
  ba.large_max = a.large_max;
  ba.bottom_up = 0;
  ba.alpha = 0;
  ba.bu_exit = 0;
  ba.m = a.m;

  // initial state: masks, counters, level slots; then the sources
  for (uint32_t v = gtid; v < n; v += gstride) {
    a.seen[v] = 0ull;
    a.f[v] = 0ull;
    a.f[(size_t)n + v] = 0ull;
  }
  if (gtid < 2 * kMsMaxSources) a.counters[gtid] = 0ull;
  if (gtid < 4) {
    LevelSlot &sl = ctl->slot[gtid];
    for (int k = 0; k < kSegments; ++k) {
      sl.tail[k] = 0;
      sl.item_tail[k] = 0;
      sl.item2_tail[k] = 0;
    }
    sl.active = 0ull;
    sl.chunk_next = 0;
  }
  if (gtid == 0) ctl->err = 0;
  grid_barrier(&ctl->bar, bst, 0, sh.bcast);
  // level 0: search q starts at sources[q]; each distinct source joins the
  // level-0 queue once (segment 0 of buffer 0, tails in slot 3)
  if (gtid < a.count) {
    const uint32_t s = a.o2r ? a.o2r[a.sources[gtid]] : a.sources[gtid];
    __hip_atomic_fetch_or(&a.f[s], 1ull << gtid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool first = true;
    for (uint32_t j = 0; j < gtid; ++j) first &= a.sources[j] != a.sources[gtid];
    if (first) {
      const uint32_t pos = __hip_atomic_fetch_add(&ctl->slot[3].tail[0], 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
      a.qb[pos] = s;
    }
  }
  grid_barrier(&ctl->bar, bst, 0, sh.bcast);

  MsVisitor vis;
  vis.col = a.col;
  vis.seen = a.seen;
  vis.stage = sh.stage[wave];
  vis.out_n = 0;
  vis.claims = 0;
  vis.mf = 0;

  uint32_t my_depth = 0;  // thread q of the team's first workgroup: levels of search q
  uint32_t bu_levels = 0;
  uint32_t L = 0;
  for (;; ++L) {
    const uint32_t *in_q = a.qb + (size_t)(L % 3) * qstride;
    const uint32_t *prev_q = a.qb + (size_t)((L + 2) % 3) * qstride;
    uint32_t *next_q = a.qb + (size_t)((L + 1) % 3) * qstride;
    LevelSlot &in_slot = ctl->slot[(L + 3) & 3];
    LevelSlot &prev_slot = ctl->slot[(L + 2) & 3];
    LevelSlot &cur = ctl->slot[L & 3];
    uint64_t *fcur = a.f + (size_t)(L & 1) * n;
    uint64_t *fnxt = a.f + (size_t)((L + 1) & 1) * n;

    // ---- commit of level L
    uint64_t t_c = 0;
    if (a.prof && tid == 0) t_c = now_ns();
    if (tid < kWave) tails_prefix<kSegments>(in_slot.tail, sh.seg_pre);
    if (tid >= kWave && tid < 2 * kWave) tails_prefix<kSegments>(prev_slot.tail, ms_pre);
    if (tid == 0) ms_or = 0ull;
    __syncthreads();
    const uint32_t F = uniform(sh.seg_pre[kSegments]);
    const uint32_t P = L ? uniform(ms_pre[kSegments]) : 0u;
    if (F == 0) break;  // no search reached a vertex at level L: every search has ended
    if (L >= (1u << kMsPlanes)) {
      if (gtid == 0) ctl->err = kMsErrDepth;
      break;
    }
    uint64_t act = 0, mf = 0;
    for (uint32_t i = gtid; i < F; i += gstride) {
      const uint32_t u = ms_queue_at(in_q, sh.seg_pre, n, i);
      const uint64_t nb = fcur[u];
      a.seen[u] |= nb;
#pragma unroll
      for (uint32_t b = 0; b < kMsPlanes; ++b)
        if ((L >> b) & 1u) a.planes[(size_t)b * n + u] |= nb;
      act |= nb;
      mf += (uint64_t)(a.off[u + 1] - a.off[u]);
    }
    // the previous level's masks leave the buffer the next level fills
    for (uint32_t i = gtid; i < P; i += gstride) fnxt[ms_queue_at(prev_q, ms_pre, n, i)] = 0ull;
    if (gtid < 3 * kSegments) {
      LevelSlot &z = ctl->slot[(L + 1) & 3];
      if (gtid < kSegments) z.tail[gtid] = 0;
      else if (gtid < 2 * kSegments) z.item_tail[gtid - kSegments] = 0;
      else z.item2_tail[gtid - 2 * kSegments] = 0;
      if (gtid == 0) {
        z.active = 0ull;
        z.chunk_next = 0;
      }
    }
    for (int d = 32; d >= 1; d >>= 1) act |= __shfl_xor(act, d);
    if (lane == 0 && act) atomicOr((unsigned long long *)&ms_or, (unsigned long long)act);
    __syncthreads();
    if (tid == 0 && ms_or)
      __hip_atomic_fetch_or(&in_slot.active, ms_or, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    vis.mf = mf;
    // the out-degree sum is this barrier's only sum: it rides as the
    // payload, with no second chain of arrivals
    const uint64_t m_f = grid_barrier(&ctl->bar, bst, block_sum_mf<BLOCK>(sh, vis), sh.bcast);
    const uint64_t alive = uniform(ld_relaxed(&in_slot.active));
    if (trank == 0 && tid < a.count && ((alive >> tid) & 1ull)) my_depth = L + 1;
    uint64_t t_e = 0;
    if (a.prof && tid == 0) t_e = now_ns();

    // ---- expansion of level L
    // the plane level L + 1 needs first (L + 1 a power of two) is cleared
    // now, before the commit that writes it
    if (((L + 1) & L) == 0 && L + 1 < (1u << kMsPlanes)) {
      uint64_t *pl = a.planes + (size_t)(31 - __builtin_clz(L + 1)) * n;
      for (uint32_t v = gtid; v < n; v += gstride) pl[v] = 0ull;
    }
    vis.fcur = fcur;
    vis.fnxt = fnxt;
    vis.seg = next_q + (size_t)sh.seg * n;
    vis.seg_tail = &cur.tail[sh.seg];
    vis.claims = 0;
    uint64_t edges = 0;
    const bool bu = a.bottom_up && m_f * (uint64_t)a.bu_div > a.m;
    uint32_t items = 0, next = 0;
    if (!bu) {
      if (F <= (uint32_t)BLOCK && m_f >= kSmallFrontierEdges && m_f <= 0xffffffffull)
        topdown_small<BLOCK>(ba, sh, vis, in_q, in_slot, wave, gwave, nwaves, edges);
      else
        topdown_phase_a<BLOCK, false>(ba, sh, vis, in_q, in_slot, cur, wave, trank, tsize, edges);
      const uint32_t wg = flush_block<BLOCK>(vis, sh, wave, vis.claims);
      next = (uint32_t)grid_barrier(&ctl->bar, bst, wg, sh.bcast);
      if (tid < kWave) tails_prefix<kSegments>(cur.item_tail, sh.item_pre);
      __syncthreads();
      const uint32_t hc = uniform(sh.item_pre[kSegments]);
      items = hc;
      if (hc) {
        const uint32_t before = vis.claims;
        topdown_phase_h<BLOCK>(ba, sh, vis, hc, gwave, nwaves, edges);
        const uint32_t wg2 = flush_block<BLOCK>(vis, sh, wave, vis.claims - before);
        next += (uint32_t)grid_barrier(&ctl->bar, bst, wg2, sh.bcast);
      }
    } else {
      ++bu_levels;
      uint2 *seg_items = a.items + (size_t)sh.seg * a.item_cap;
      const bool probe = bu_levels > 1;
      ms_bottomup_level<BLOCK>(a, sh, vis, alive, wave, &cur.chunk_next, seg_items, &cur.item_tail[sh.seg],
                               &cur.item2_tail[sh.seg], probe);
      const uint32_t wg = flush_block<BLOCK>(vis, sh, wave, vis.claims);
      next = (uint32_t)grid_barrier(&ctl->bar, bst, wg, sh.bcast);
      if (tid < kWave) tails_prefix<kSegments>(probe ? cur.item_tail : cur.item2_tail, sh.item_pre);
      __syncthreads();
      const uint32_t hp = uniform(sh.item_pre[kSegments]);
      items = hp;
      if (hp && !probe) {
        const uint32_t before = vis.claims;
        ms_bottomup_items<BLOCK>(a, vis, alive, sh.item_pre, hp, gwave, nwaves, 0u);
        const uint32_t wg3 = flush_block<BLOCK>(vis, sh, wave, vis.claims - before);
        next += (uint32_t)grid_barrier(&ctl->bar, bst, wg3, sh.bcast);
      } else if (hp) {
        // probes of the long rows, then the rest of the rows they left open
        uint32_t before = vis.claims;
        ms_bottomup_probes<BLOCK, 1>(a, vis, alive, sh.item_pre, hp, gwave, nwaves, seg_items,
                                     &cur.item2_tail[sh.seg]);
        const uint32_t wg2 = flush_block<BLOCK>(vis, sh, wave, vis.claims - before);
        next += (uint32_t)grid_barrier(&ctl->bar, bst, wg2, sh.bcast);
        if (tid < kWave) tails_prefix<kSegments>(cur.item2_tail, sh.item_pre);
        __syncthreads();
        const uint32_t hc = uniform(sh.item_pre[kSegments]);
        items += hc;
        if (hc) {
          before = vis.claims;
          ms_bottomup_items<BLOCK>(a, vis, alive, sh.item_pre, hc, gwave, nwaves, kMsBuRow);
          const uint32_t wg3 = flush_block<BLOCK>(vis, sh, wave, vis.claims - before);
          next += (uint32_t)grid_barrier(&ctl->bar, bst, wg3, sh.bcast);
        }
      }
    }
    if (a.prof && trank == 0 && tid == 0 && L < a.prof_levels) {
      uint64_t *row = a.prof + (size_t)L * kMsProfFields;
      row[kMsProfFrontier] = F;
      row[kMsProfEdges] = m_f;
      row[kMsProfBottomUp] = bu ? 1 : 0;
      row[kMsProfActive] = (uint64_t)__popcll(alive);
      row[kMsProfItems] = items;
      row[kMsProfCommitNs] = t_e - t_c;
      row[kMsProfExpandNs] = now_ns() - t_e;
      row[kMsProfNext] = next;
    }
  }

  // ---- final pass: every search's level array (full 64-vertex stores per
  // search), and per-search reached vertices and their CSR entries.  The
  // levels are at most L - 1, so the planes below bit_length(L - 1) hold them.
  const uint32_t np = L > 1 ? 32u - __builtin_clz(L - 1) : 0u;
  const uint32_t K = a.count;
  if (tid < 2 * kMsMaxSources) ms_red[tid] = 0ull;
  __syncthreads();
  uint64_t r_cnt = 0, r_edges = 0;  // lane q: search q's share of this wave's vertices
  const uint32_t groups = (n + kWave - 1) / kWave;
  for (uint32_t g = gwave; g < groups; g += nwaves) {
    // w in the caller's ids (coalesced output), r its vertex in the graph searched
    const uint32_t w = g * kWave + lane;
    const bool valid = w < n;
    const uint32_t r = valid ? (a.o2r ? a.o2r[w] : w) : 0u;
    const uint64_t s = valid ? a.seen[r] : 0ull;
    const uint32_t deg = valid ? (uint32_t)(a.off[r + 1] - a.off[r]) : 0u;
    uint64_t pl[kMsPlanes];
#pragma unroll
    for (uint32_t b = 0; b < kMsPlanes; ++b) pl[b] = (b < np && valid) ? a.planes[(size_t)b * n + r] : 0ull;
    for (uint32_t q = 0; q < K; ++q) {
      uint32_t lv = 0;
#pragma unroll
      for (uint32_t b = 0; b < kMsPlanes; ++b) lv |= (uint32_t)((pl[b] >> q) & 1ull) << b;
      if (valid) a.level[q * a.level_stride + w] = ((s >> q) & 1ull) ? lv : kInf;
    }
    // lane q sums search q over the 64 vertices (broadcast from each lane)
    const uint32_t slo = (uint32_t)s, shi = (uint32_t)(s >> 32);
    for (int j = 0; j < kWave; ++j) {
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)slo, j);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)shi, j);
      const uint32_t dj = (uint32_t)__builtin_amdgcn_readlane((int)deg, j);
      const uint32_t bit = ((lane < 32 ? lo >> lane : hi >> (lane - 32)) & 1u);
      r_cnt += bit;
      r_edges += bit ? (uint64_t)dj : 0ull;
    }
  }
  if (lane < K) {
    if (r_cnt) atomicAdd((unsigned long long *)&ms_red[lane], (unsigned long long)r_cnt);
    if (r_edges) atomicAdd((unsigned long long *)&ms_red[kMsMaxSources + lane], (unsigned long long)r_edges);
  }
  __syncthreads();
  if (tid < 2 * kMsMaxSources && ms_red[tid])
    atomicAdd((unsigned long long *)&a.counters[tid], (unsigned long long)ms_red[tid]);
  if (trank == 0 && tid < K) a.depth[tid] = my_depth;
  if (trank == 0 && tid == 0) {
    ctl->depth = L;
    ctl->bu_levels = bu_levels;
  }
}

// Up to 64 searches of one graph in one cooperative launch over the whole
// grid (one team).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void dfq_msbfs_kernel(const MsArgs a) {
  __shared__ SharedAll<BLOCK> sh;
  Team team{};
  GridBarrier *bar = &a.ctl->bar;
  if (!kernel_prologue<BLOCK>(bar, 0, false, [bar](uint32_t) { return bar; }, sh, team)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl->err = 2;
    return;
  }
  ms_grid_search<BLOCK>(a, sh, team);
}

}  // namespace dfq
