// Device-side building blocks of the DFQ-BFS kernels for gfx950 (CDNA4,
// wave64): wavefront primitives, the visited-bitmap claim, and the per-wave
// LDS staging of next-frontier entries.
//
// Reference counterparts (design/shared-q/atomic-sa/our-sep/upper-bound/
// bfs.cu): VISIT_tiled / VISIT_blk_tiled / VISIT_grid_tiled /
// VISIT_all_tiled (:44-94) test `cost[nid] < 0` and write `cost[ftr] + 1`
// without atomics, and find the next frontier by rescanning all N levels
// (:116-124).  Here a vertex is claimed exactly once through an atomic OR on
// a visited bitmap; the claimer alone writes its level (so levels are
// deterministic) and parent, and appends it to the next-frontier queue, so no
// level ever rescans the vertex set.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dfq {

constexpr uint32_t kInf = 0xffffffffu;
constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Orders LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of set bits of mask below this lane (v_mbcnt_lo/hi).
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t bcast_lane(uint32_t x, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

// Inclusive prefix sum over the 64 lanes of a wavefront.  row_shr DPP for
// the steps inside a 16-lane row, row_bcast:15 / row_bcast:31 across rows.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
  // update_dpp(old, src, dpp_ctrl, row_mask, bank_mask, bound_ctrl):
  // lanes whose source is out of the row read `old` (0 here).
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

// Visited-bitmap claim: a plain load filters vertices already seen (a stale
// copy can only say "unvisited", never the reverse, because bits are only
// set during a BFS); the atomic OR decides ownership.
__device__ __forceinline__ bool claim_vertex(uint32_t *vis, uint32_t w) {
  const uint32_t word = w >> 5, bit = 1u << (w & 31u);
  if (vis[word] & bit) return false;
  const uint32_t old = __hip_atomic_fetch_or(&vis[word], bit, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
  return (old & bit) == 0;
}

// Everything the edge visitor needs; one per thread, with `out_n` (entries
// staged by this wave) kept wave-uniform.
struct Visitor {
  const uint32_t *__restrict__ col;
  uint32_t *level;
  uint32_t *parent;   // may be null
  uint32_t *vis;
  uint32_t *qout;     // next-frontier queue (HBM)
  uint32_t *q_count;  // its tail counter
  uint32_t *stage;    // this wave's LDS staging buffer
  uint32_t stage_cap; // entries; a multiple of 64, >= 128
  uint32_t out_n;     // staged entries (wave-uniform)
  uint32_t next_level;

  // Reserve a contiguous block of the HBM queue for `cnt` entries and copy
  // the wave's staged entries there (one atomic per flush).
  __device__ __forceinline__ void flush_wave() {
    const uint32_t cnt = out_n;
    if (cnt == 0) return;
    uint32_t base = 0;
    if (lane_id() == 0)
      base = __hip_atomic_fetch_add(q_count, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    base = bcast_lane(base, 0);
    wave_sync();
    for (uint32_t k = lane_id(); k < cnt; k += kWave) qout[base + k] = stage[k];
    wave_sync();
    out_n = 0;
  }

  // Called by every lane of the wave (wave-uniform control flow); `act`
  // marks lanes that carry an edge u -> w.
  __device__ __forceinline__ void visit(bool act, uint32_t u, uint32_t w) {
    bool claimed = false;
    if (act) {
      claimed = claim_vertex(vis, w);
      if (claimed) {
        level[w] = next_level;
        if (parent) parent[w] = u;
      }
    }
    const uint64_t mask = __ballot(claimed);
    if (mask == 0) return;
    if (claimed) stage[out_n + mbcnt(mask)] = w;
    out_n += (uint32_t)__popcll(mask);
    if (out_n > stage_cap - kWave) flush_wave();
  }
};

}  // namespace dfq
