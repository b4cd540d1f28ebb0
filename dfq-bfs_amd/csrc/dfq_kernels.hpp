// The persistent cooperative DFQ-BFS kernel for gfx950.
//
// One launch runs a whole BFS (reference: CUDA_BFS_KERNEL_CG_NEW,
// design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:98-179, launched
// cooperatively at :283).  Per level:
//
//   phase A  workgroups take chunks of the frontier queue (HBM, written by
//            the previous level) with an atomic cursor and bin each vertex by
//            out-degree:
//              deg <= small_max   lane-owned; the wave gathers the edges of
//                                 its 64 lanes' vertices through a DPP prefix
//                                 sum so every lane reads one edge per step
//              deg <= medium_max  LDS queue, one wavefront per vertex
//              deg <= large_max   LDS queue, one workgroup per vertex
//              deg >  large_max   HBM queue, expanded by the whole grid
//            LDS queues are expanded by the workgroup that filled them when
//            they come within one chunk of full, and drained at the end of
//            the level (reference :126-169, flush on near-full + drain).
//   barrier  grid.sync
//   phase H  only when the level has huge-degree vertices: all workgroups
//            split their concatenated edge ranges into workgroup-sized
//            chunks (reference E/F bins :171-176, which serialised F)
//            followed by a second grid.sync.
//
// Claimed vertices go to a per-wave LDS staging buffer and from there to the
// next-frontier queue with one atomic per flush.  The loop ends when the
// next frontier is empty; every workgroup reads the same counter after the
// barrier, so all leave on the same level.
#pragma once

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dfq_device.hpp"

namespace dfq {

namespace cg = cooperative_groups;

// Per-level control slot; three rotate so that the slot being zeroed for
// level L+2 is never touched during level L (+ padding against false sharing
// of the hot counters).
struct alignas(128) LevelSlot {
  uint32_t q_count;   // frontier entries of this level (written at level-1)
  uint32_t q_cursor;  // consumption cursor
  uint32_t h_count;   // huge-degree vertices of this level
  uint32_t pad[29];
};

struct Control {
  LevelSlot slot[3];
  uint32_t depth;     // levels executed (max level + 1)
  uint32_t pad0;
  uint64_t reached;   // vertices placed in any frontier
  uint32_t pad1[28];
};

struct BfsArgs {
  const int64_t *off;    // row offsets [n+1]
  const uint32_t *col;   // column indices [m]
  uint32_t *level;       // [n]
  uint32_t *parent;      // [n] or null
  uint32_t *vis;         // visited bitmap [n_words]
  uint32_t *q[2];        // frontier queues [n] each
  uint32_t *hq;          // huge-degree queue [n]
  Control *ctl;
  uint32_t n;
  uint32_t n_words;
  uint32_t source;
  uint32_t small_max;    // <= 64
  uint32_t medium_max;
  uint32_t large_max;
};

template <int BLOCK>
struct KernelShape {
  static constexpr int kWaves = BLOCK / kWave;
  static constexpr int kMedCap = 2 * BLOCK;  // medium LDS queue
  static constexpr int kBigCap = BLOCK;      // large LDS queue
  static constexpr int kStage = 256;         // per-wave next-frontier staging
  static constexpr int kHTile = 1024;        // huge vertices per phase-H tile
};

template <int BLOCK>
struct SharedA {
  uint32_t med_v[KernelShape<BLOCK>::kMedCap];
  uint32_t med_d[KernelShape<BLOCK>::kMedCap];
  int64_t med_s[KernelShape<BLOCK>::kMedCap];
  uint32_t big_v[KernelShape<BLOCK>::kBigCap];
  uint32_t big_d[KernelShape<BLOCK>::kBigCap];
  int64_t big_s[KernelShape<BLOCK>::kBigCap];
  uint32_t g_excl[KernelShape<BLOCK>::kWaves][kWave];
  uint32_t g_v[KernelShape<BLOCK>::kWaves][kWave];
  int64_t g_s[KernelShape<BLOCK>::kWaves][kWave];
};

template <int BLOCK>
struct SharedH {
  uint32_t hv[KernelShape<BLOCK>::kHTile];
  int64_t hs[KernelShape<BLOCK>::kHTile];
  uint64_t hx[KernelShape<BLOCK>::kHTile];  // exclusive prefix of degrees
  uint64_t wave_tot[KernelShape<BLOCK>::kWaves];
};

template <int BLOCK>
struct SharedAll {
  union {
    SharedA<BLOCK> a;
    SharedH<BLOCK> h;
  } u;
  uint32_t stage[KernelShape<BLOCK>::kWaves][KernelShape<BLOCK>::kStage];
  uint32_t stage_n[KernelShape<BLOCK>::kWaves];
  uint32_t stage_base[KernelShape<BLOCK>::kWaves];
  uint32_t med_n, big_n, chunk_base;
};

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-uniform gather of the edges of up to 64 lane-owned vertices.
template <int BLOCK>
__device__ __forceinline__ void gather_small(Visitor &vis, uint32_t v, int64_t start, uint32_t deg,
                                             uint32_t *g_excl, uint32_t *g_v, int64_t *g_s) {
  const uint32_t lane = lane_id();
  const uint32_t incl = wave_inclusive_scan(deg);
  const uint32_t total = bcast_lane(incl, 63);
  if (total == 0) return;
  g_excl[lane] = incl - deg;
  g_v[lane] = v;
  g_s[lane] = start;
  wave_sync();
  for (uint32_t b = 0; b < total; b += kWave) {
    const uint32_t e = b + lane;
    const bool act = e < total;
    // owner = last lane whose exclusive prefix is <= e; lanes with degree 0
    // share their prefix with the next lane and are never selected.
    uint32_t j = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1)
      if (g_excl[j + step] <= e) j += step;
    uint32_t w = 0, u = 0;
    if (act) {
      u = g_v[j];
      w = vis.col[g_s[j] + (int64_t)(e - g_excl[j])];
    }
    vis.visit(act, u, w);
  }
  wave_sync();
}

// Expands the medium LDS queue (one wavefront per vertex).
template <int BLOCK>
__device__ __forceinline__ void expand_medium(Visitor &vis, SharedAll<BLOCK> &sh, uint32_t cnt,
                                              uint32_t wave) {
  constexpr int W = KernelShape<BLOCK>::kWaves;
  const uint32_t lane = lane_id();
  for (uint32_t j = wave; j < cnt; j += W) {
    const uint32_t u = sh.u.a.med_v[j];
    const uint32_t d = sh.u.a.med_d[j];
    const int64_t s = sh.u.a.med_s[j];
    for (uint32_t k0 = 0; k0 < d; k0 += kWave) {
      const uint32_t k = k0 + lane;
      const bool act = k < d;
      const uint32_t w = act ? vis.col[s + k] : 0u;
      vis.visit(act, u, w);
    }
  }
}

// Expands the large LDS queue (the whole workgroup per vertex).
template <int BLOCK>
__device__ __forceinline__ void expand_large(Visitor &vis, SharedAll<BLOCK> &sh, uint32_t cnt) {
  for (uint32_t j = 0; j < cnt; ++j) {
    const uint32_t u = sh.u.a.big_v[j];
    const uint32_t d = sh.u.a.big_d[j];
    const int64_t s = sh.u.a.big_s[j];
    for (uint32_t k0 = 0; k0 < d; k0 += BLOCK) {
      const uint32_t k = k0 + threadIdx.x;
      const bool act = k < d;
      const uint32_t w = act ? vis.col[s + k] : 0u;
      vis.visit(act, u, w);
    }
  }
}

// Moves every wave's staged entries to the HBM queue with one atomic per
// workgroup.  Must be reached by all threads of the workgroup.
template <int BLOCK>
__device__ __forceinline__ void flush_block(Visitor &vis, SharedAll<BLOCK> &sh, uint32_t wave) {
  constexpr int W = KernelShape<BLOCK>::kWaves;
  if (lane_id() == 0) sh.stage_n[wave] = vis.out_n;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int i = 0; i < W; ++i) {
      sh.stage_base[i] = tot;
      tot += sh.stage_n[i];
    }
    uint32_t base = 0;
    if (tot) base = __hip_atomic_fetch_add(vis.q_count, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int i = 0; i < W; ++i) sh.stage_base[i] += base;
  }
  __syncthreads();
  const uint32_t cnt = vis.out_n, base = sh.stage_base[wave];
  for (uint32_t k = lane_id(); k < cnt; k += kWave) vis.qout[base + k] = vis.stage[k];
  vis.out_n = 0;
  __syncthreads();
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void dfq_bfs_kernel(BfsArgs a) {
  using S = KernelShape<BLOCK>;
  __shared__ SharedAll<BLOCK> sh;
  cg::grid_group grid = cg::this_grid();

  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / kWave;
  const uint32_t gtid = blockIdx.x * BLOCK + tid;
  const uint32_t gstride = gridDim.x * BLOCK;

  // ---- initialisation (reference :216 cost init, done on the device) ----
  for (uint32_t i = gtid; i < a.n; i += gstride) {
    a.level[i] = (i == a.source) ? 0u : kInf;
    if (a.parent) a.parent[i] = (i == a.source) ? a.source : kInf;
  }
  for (uint32_t w = gtid; w < a.n_words; w += gstride)
    a.vis[w] = (w == (a.source >> 5)) ? (1u << (a.source & 31u)) : 0u;
  if (gtid == 0) {
    for (int s = 0; s < 3; ++s) {
      a.ctl->slot[s].q_count = 0;
      a.ctl->slot[s].q_cursor = 0;
      a.ctl->slot[s].h_count = 0;
    }
    a.ctl->slot[0].q_count = 1;
    a.q[0][0] = a.source;
  }
  grid.sync();

  Visitor vis;
  vis.col = a.col;
  vis.level = a.level;
  vis.parent = a.parent;
  vis.vis = a.vis;
  vis.stage = sh.stage[wave];
  vis.stage_cap = S::kStage;
  vis.out_n = 0;

  uint64_t reached = 0;
  for (uint32_t L = 0;; ++L) {
    LevelSlot &cur = a.ctl->slot[L % 3];
    LevelSlot &nxt = a.ctl->slot[(L + 1) % 3];
    if (gtid == 0) {
      LevelSlot &z = a.ctl->slot[(L + 2) % 3];
      z.q_count = 0;
      z.q_cursor = 0;
      z.h_count = 0;
    }
    const uint32_t n_in = ld_agent(&cur.q_count);
    reached += n_in;
    const uint32_t *qin = a.q[L & 1];
    vis.qout = a.q[(L + 1) & 1];
    vis.q_count = &nxt.q_count;
    vis.next_level = L + 1;

    // ---- phase A: frontier chunks -> bins -> expansion ----
    if (tid == 0) {
      sh.med_n = 0;
      sh.big_n = 0;
    }
    __syncthreads();
    for (;;) {
      if (tid == 0)
        sh.chunk_base = __hip_atomic_fetch_add(&cur.q_cursor, (uint32_t)BLOCK, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      const uint32_t base = sh.chunk_base;
      if (base >= n_in) break;
      const uint32_t i = base + tid;
      uint32_t v = 0, deg = 0;
      int64_t start = 0;
      if (i < n_in) {
        v = qin[i];
        start = a.off[v];
        deg = (uint32_t)(a.off[v + 1] - start);
      }
      const bool is_small = deg != 0 && deg <= a.small_max;
      const bool is_med = deg > a.small_max && deg <= a.medium_max;
      const bool is_big = deg > a.medium_max && deg <= a.large_max;
      const bool is_huge = deg > a.large_max;
      // wave-aggregated LDS slot reservation for the medium / large bins
      {
        const uint64_t mm = __ballot(is_med);
        if (mm) {
          uint32_t b0 = 0;
          if (lane_id() == 0) b0 = atomicAdd(&sh.med_n, (uint32_t)__popcll(mm));
          b0 = bcast_lane(b0, 0);
          if (is_med) {
            const uint32_t p = b0 + mbcnt(mm);
            sh.u.a.med_v[p] = v;
            sh.u.a.med_d[p] = deg;
            sh.u.a.med_s[p] = start;
          }
        }
        const uint64_t bm = __ballot(is_big);
        if (bm) {
          uint32_t b0 = 0;
          if (lane_id() == 0) b0 = atomicAdd(&sh.big_n, (uint32_t)__popcll(bm));
          b0 = bcast_lane(b0, 0);
          if (is_big) {
            const uint32_t p = b0 + mbcnt(bm);
            sh.u.a.big_v[p] = v;
            sh.u.a.big_d[p] = deg;
            sh.u.a.big_s[p] = start;
          }
        }
        const uint64_t hm = __ballot(is_huge);
        if (hm) {
          uint32_t b0 = 0;
          if (lane_id() == 0)
            b0 = __hip_atomic_fetch_add(&cur.h_count, (uint32_t)__popcll(hm), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
          b0 = bcast_lane(b0, 0);
          if (is_huge) a.hq[b0 + mbcnt(hm)] = v;
        }
      }
      gather_small<BLOCK>(vis, is_small ? v : 0u, is_small ? start : 0, is_small ? deg : 0u,
                          sh.u.a.g_excl[wave], sh.u.a.g_v[wave], sh.u.a.g_s[wave]);
      __syncthreads();
      const uint32_t mn = sh.med_n, bn = sh.big_n;
      if (mn > (uint32_t)(S::kMedCap - BLOCK)) {
        expand_medium<BLOCK>(vis, sh, mn, wave);
        __syncthreads();
        if (tid == 0) sh.med_n = 0;
      }
      if (bn > (uint32_t)(S::kBigCap - BLOCK)) {
        expand_large<BLOCK>(vis, sh, bn);
        __syncthreads();
        if (tid == 0) sh.big_n = 0;
      }
      __syncthreads();
    }
    // drain (reference :161-169)
    expand_medium<BLOCK>(vis, sh, sh.med_n, wave);
    expand_large<BLOCK>(vis, sh, sh.big_n);
    flush_block<BLOCK>(vis, sh, wave);
    grid.sync();

    // ---- phase H: huge-degree vertices over the whole grid ----
    const uint32_t hc = ld_agent(&cur.h_count);
    if (hc) {
      constexpr int HT = S::kHTile;
      for (uint32_t t0 = 0; t0 < hc; t0 += HT) {
        const uint32_t tn = min((uint32_t)HT, hc - t0);
        // degrees of the tile and their exclusive prefix (per-thread runs of
        // HT/BLOCK entries, wave scan of run totals, cross-wave add)
        constexpr int RUN = HT / BLOCK;
        uint64_t run = 0;
        for (int r = 0; r < RUN; ++r) {
          const uint32_t idx = tid * RUN + r;
          uint32_t hvtx = 0, d = 0;
          int64_t s = 0;
          if (idx < tn) {
            hvtx = a.hq[t0 + idx];
            s = a.off[hvtx];
            d = (uint32_t)(a.off[hvtx + 1] - s);
          }
          sh.u.h.hv[idx] = hvtx;
          sh.u.h.hs[idx] = s;
          sh.u.h.hx[idx] = run;  // exclusive within the run for now
          run += d;
        }
        // 64-bit inclusive scan of the run totals across the wave, then
        // across waves through LDS.
        uint64_t incl = run;
        for (int d = 1; d < kWave; d <<= 1) {
          const uint64_t t = __shfl_up(incl, d);
          if (lane_id() >= (uint32_t)d) incl += t;
        }
        if (lane_id() == kWave - 1) sh.u.h.wave_tot[wave] = incl;
        __syncthreads();
        uint64_t wave_off = 0, tot = 0;
        for (int w = 0; w < S::kWaves; ++w) {
          if ((uint32_t)w < wave) wave_off += sh.u.h.wave_tot[w];
          tot += sh.u.h.wave_tot[w];
        }
        const uint64_t excl = wave_off + (incl - run);
        for (int r = 0; r < RUN; ++r) sh.u.h.hx[tid * RUN + r] += excl;
        __syncthreads();
        // Entries beyond tn have degree 0 and an exclusive prefix == tot, so
        // the owner search below never selects them.
        for (uint64_t ch = blockIdx.x; ch * BLOCK < tot; ch += gridDim.x) {
          const uint64_t e = ch * BLOCK + tid;
          const bool act = e < tot;
          uint32_t j = 0;
#pragma unroll
          for (uint32_t step = HT / 2; step >= 1; step >>= 1)
            if (sh.u.h.hx[j + step] <= e) j += step;
          uint32_t w = 0, u = 0;
          if (act) {
            u = sh.u.h.hv[j];
            w = a.col[sh.u.h.hs[j] + (int64_t)(e - sh.u.h.hx[j])];
          }
          vis.visit(act, u, w);
        }
        __syncthreads();
      }
      flush_block<BLOCK>(vis, sh, wave);
      grid.sync();
    }

    const uint32_t n_next = ld_agent(&nxt.q_count);
    if (n_next == 0) {
      if (gtid == 0) {
        a.ctl->depth = L + 1;
        a.ctl->reached = reached;
      }
      break;
    }
  }
}

}  // namespace dfq
