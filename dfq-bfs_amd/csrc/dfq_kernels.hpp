// The persistent cooperative DFQ-BFS kernel for gfx950.
//
// One launch runs a whole BFS (reference: CUDA_BFS_KERNEL_CG_NEW,
// design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:98-179, launched
// cooperatively at :283).  Per level (top-down):
//
//   phase A  every workgroup takes an equal share of the frontier (the
//            previous level's per-XCD segments, located from their tails; no
//            shared cursor) and bins each vertex by out-degree:
//              deg <= small_max   lane-owned; the wave gathers the edges of
//                                 its 64 lanes' vertices through a DPP prefix
//                                 sum so every lane reads one edge per step
//              deg <= medium_max  LDS queue, one wavefront per vertex
//              deg <= large_max   LDS queue, one workgroup per vertex
//              deg >  large_max   cut into 256-edge work items (HBM)
//            LDS queues are expanded by the workgroup that filled them when
//            they come within one chunk of full, and drained at the end of
//            the level (reference :126-169, flush on near-full + drain).
//   barrier  tree barrier; returns the number of vertices claimed
//   phase H  only when the level has work items: every wavefront of the grid
//            takes items round-robin (balanced to one item per wave; the
//            reference's E/F bins :171-176 serialised F), then a second
//            barrier.
//
// Optionally (symmetric graphs, direction_opt) a level whose frontier is
// large runs bottom-up instead: every unvisited vertex looks for a neighbour
// in the frontier bitmap and stops at the first one (Beamer, Asanovic,
// Patterson, "Direction-optimizing breadth-first search", SC 2012).  Each
// vertex has one owner lane, so that level needs no atomics.
//
// The loop ends when a level claims no vertex; every workgroup receives the
// same count from the barrier, so all leave on the same level.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dfq_device.hpp"

namespace dfq {

// Per-level control slot; four rotate: level L appends into slot L%4's
// segment tails and huge list, the consumer of level L+1 reads them, and the
// slot is zeroed during level L+2 for reuse at level L+4.
struct alignas(256) LevelSlot {
  uint32_t tail[kSegments];       // entries in each segment of level L's output
  uint32_t item_tail[kSegments];  // work items of level L's grid phase, per segment
  uint64_t active;                // multi-source searches: the searches with a frontier at level L
  uint32_t chunk_next;            // multi-source bottom-up levels: next vertex chunk to claim
  uint32_t item2_tail[kSegments]; // multi-source bottom-up levels: rest-of-row items, per segment
  uint32_t pad[64 - 3 * kSegments - 3];
};

// Huge-degree frontier vertices are cut into items of kItem CSR entries:
// (vertex, chunk index).  One item is one wavefront step of kUnroll x 64.

struct Control {
  GridBarrier bar;
  LevelSlot slot[4];
  uint32_t depth;     // levels executed (max level + 1)
  uint32_t err;       // nonzero: a queue capacity was exceeded
  uint64_t reached;   // vertices with a level
  uint32_t bu_levels; // levels run bottom-up
  uint32_t abandoned; // number of the last plain launch whose grid did not assemble
  uint64_t edges;     // CSR entries of reached vertices (single searches)
  uint32_t pad[56];
  BarrierWord solo_rel;        // tag of the last hand-back of solo levels
  uint64_t solo_state[32];     // search state handed back (first 8 words used)
};

// Profiling counters, one row per level (optional).
enum ProfField : int {
  kProfFrontier = 0,  // vertices in the level's frontier
  kProfEdges,         // CSR entries read by the expansion
  kProfProbes,        // visited / frontier-bitmap tests
  kProfAtomics,       // atomic OR attempts on the visited bitmap
  kProfClaims,        // vertices given the next level
  kProfHuge,          // work items of the grid phase (kItem entries each)
  kProfBottomUp,      // 1 when the level ran bottom-up
  kProfLevelNs,       // wall time of the level (workgroup 0, barrier to barrier)
  kProfBusyMaxNs,     // slowest workgroup's expansion time (before its barrier arrival)
  kProfBusySumNs,     // sum over workgroups of the same
  kProfBinSmall,      // frontier vertices per degree bin (levels through the bins):
  kProfBinMedium,     //   lane-owned, one wave each, one workgroup each,
  kProfBinLarge,      //   kItem-entry work items over the grid
  kProfBinHuge,
  kProfIdleSmall,     // idle lane slots of each bin's last step: 64-lane steps for
  kProfIdleMedium,    //   the lane-owned, wave and work-item bins, BLOCK-lane steps
  kProfIdleLarge,     //   for the workgroup bin (the reference's idle-lane counts)
  kProfIdleHuge,
  kProfEdgesMaxWg,    // CSR entries read by the busiest workgroup (per-workgroup work)
  kProfBuCandidates,  // unvisited vertices a bottom-up level scanned
  kProfFields = 20
};


struct BfsArgs {
  const int64_t *off;    // row offsets [n+1]
  const uint32_t *col;   // column indices [m] (top-down levels read these)
  const uint32_t *col_bu; // the same rows ordered by neighbour degree (bottom-up levels), or null
  const uint4 *head;     // first four entries of every col_bu row [n] (bottom-up levels), or null
  uint32_t *level;       // [n]
  uint32_t *parent;      // [n] or null
  const uint32_t *iso;   // isolated vertices [n_words] (symmetric graphs) or null
  uint32_t *vis;         // visited bitmap [n_words]
  uint32_t *vis_snap;    // its copy taken at the start of top-down levels with many out-edges [n_words]
  uint32_t *fb;          // two frontier bitmaps [2][n_words + 1] (bottom-up levels)
  uint32_t *qseg[2];     // frontier queues: kSegments segments of n entries each
  uint2 *items;          // kSegments segments of item_cap work items
  uint32_t item_cap;     // items per segment (m / kItem + n bounds one level)
  Control *ctl;
  uint64_t *prof;        // [prof_levels][kProfFields] or null
  uint64_t *result;      // [kResultWords] or null: depth, reached, bottom-up levels, edges,
                         // device-clock ns at the search's start and end, the launch
                         // number when a plain launch was abandoned
  uint32_t plain;        // 0: cooperative launch; else the host's number of this plain launch
  uint64_t plain_wait_ns; // plain launches: how long a workgroup waits for the grid to assemble
  uint32_t prof_levels;
  uint32_t n;
  uint32_t n_words;
  uint32_t source;
  uint32_t small_max;    // <= 64
  uint32_t medium_max;
  uint32_t large_max;
  uint32_t bottom_up;    // direction optimisation allowed (symmetric graph)
  uint32_t alpha;        // bottom-up when frontier edges > unexplored edges / alpha
  uint32_t bu_exit;      // back to top-down when a shrinking frontier has fewer vertices
  uint64_t m;            // CSR entries
  uint32_t deg_max = 0;  // largest out-degree (read by the low-degree instantiation only)
  uint32_t fuse_bits = 0; // nonzero: level barriers carry the claim count in the payload's
                          // low fuse_bits bits and the out-degree sum above them
};

// What changes from one search to the next on the same resident graph:
// kept apart from BfsArgs, so a launch that runs many searches leaves its
// argument block untouched (constant, re-read from the kernel arguments
// instead of held in registers).
struct SearchIo {
  uint32_t *level;   // [n]
  uint32_t *parent;  // [n] or null
  uint64_t *result;  // [kResultWords] or null
  uint32_t source;
};

// Independent 64-edge steps a wave keeps in flight in the expansions.
constexpr int kUnroll = 4;
constexpr uint32_t kItem = kWave * kUnroll;  // CSR entries per work item

template <int BLOCK>
struct KernelShape {
  static constexpr int kWaves = BLOCK / kWave;
  static constexpr int kMedCap = 2 * BLOCK;  // medium LDS queue
  static constexpr int kBigCap = BLOCK;      // large LDS queue
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

// Bottom-up levels work on chunks of kBuWords visited words (kBuChunk
// vertices) per wave.
constexpr uint32_t kBuWords = 32;
constexpr uint32_t kBuChunk = kBuWords * 32;

// Small top-down frontier (at most one entry per thread): vertices, row
// starts and the exclusive prefix of their degrees.
template <int BLOCK>
struct SharedSmall {
  uint32_t v[BLOCK];
  int64_t start[BLOCK];
  uint32_t pre[BLOCK];
  uint32_t wave_tot[KernelShape<BLOCK>::kWaves];
};

template <int BLOCK>
struct SharedAll {
  union {
    SharedA<BLOCK> a;                                       // top-down bins
    uint32_t cand[KernelShape<BLOCK>::kWaves][kBuChunk];   // bottom-up candidates
    SharedSmall<BLOCK> f;                                   // small frontiers
  } u;
  uint32_t bu_claim[KernelShape<BLOCK>::kWaves][kBuWords];  // bottom-up claims of the chunk
  uint32_t stage[KernelShape<BLOCK>::kWaves][kStage];
  uint32_t stage_n[KernelShape<BLOCK>::kWaves];
  uint32_t stage_base[KernelShape<BLOCK>::kWaves];
  uint32_t seg_pre[kSegments + 1];
  uint32_t item_pre[kSegments + 1];
  uint32_t seg;  // output segment of this workgroup (its XCD's rank in the team)
  uint32_t team_rank, team_size, team_id;
  uint32_t wave_sum[KernelShape<BLOCK>::kWaves];
  uint64_t wave_mf[KernelShape<BLOCK>::kWaves];
  uint64_t prof[kProfFields];
  uint64_t bcast[2];
  uint32_t med_n, big_n;
  uint32_t census[2 * kMaxXcd];  // kernel start: workgroups per XCD rank, team generations
};

// Wave-uniform gather of the edges of up to 64 lane-owned vertices.
template <int BLOCK, class V>
__device__ __forceinline__ void gather_small(V &vis, uint32_t v, int64_t start, uint32_t deg,
                                             uint32_t *g_excl, uint32_t *g_v, int64_t *g_s) {
  const uint32_t lane = lane_id();
  const uint32_t incl = wave_inclusive_scan(deg);
  const uint32_t total = bcast_lane(incl, 63);
  if (total == 0) return;
  g_excl[lane] = incl - deg;
  g_v[lane] = v;
  g_s[lane] = start;
  wave_sync();
  for (uint32_t b = 0; b < total; b += kWave * kUnroll) {
    bool act[kUnroll];
    uint32_t u[kUnroll], w[kUnroll];
#pragma unroll
    for (int i = 0; i < kUnroll; ++i) {
      const uint32_t e = b + i * kWave + lane;
      act[i] = e < total;
      // owner = last lane whose exclusive prefix is <= e; lanes with degree
      // 0 share their prefix with the next lane and are never selected.
      uint32_t j = 0;
#pragma unroll
      for (uint32_t step = 32; step >= 1; step >>= 1)
        if (g_excl[j + step] <= e) j += step;
      u[i] = g_v[j];
      w[i] = act[i] ? vis.col[g_s[j] + (int64_t)(e - g_excl[j])] : 0u;
    }
    vis.template visit_batch<kUnroll, true>(act, u, w);
  }
  wave_sync();
}

// Expands the medium LDS queue (one wavefront per vertex).
template <int BLOCK, class V>
__device__ __forceinline__ void expand_medium(V &vis, SharedAll<BLOCK> &sh, uint32_t cnt,
                                              uint32_t wave) {
  constexpr int W = KernelShape<BLOCK>::kWaves;
  const uint32_t lane = lane_id();
  for (uint32_t j = wave; j < cnt; j += W) {
    const uint32_t u = sh.u.a.med_v[j];
    const uint32_t d = sh.u.a.med_d[j];
    const int64_t s = sh.u.a.med_s[j];
    for (uint32_t k0 = 0; k0 < d; k0 += kWave * kUnroll) {
      bool act[kUnroll];
      uint32_t uu[kUnroll], w[kUnroll];
#pragma unroll
      for (int i = 0; i < kUnroll; ++i) {
        const uint32_t k = k0 + i * kWave + lane;
        act[i] = k < d;
        uu[i] = u;
        w[i] = act[i] ? vis.col[s + k] : 0u;
      }
      vis.template visit_batch<kUnroll, true>(act, uu, w);
    }
  }
}

// Expands the large LDS queue (the whole workgroup per vertex).
template <int BLOCK, class V>
__device__ __forceinline__ void expand_large(V &vis, SharedAll<BLOCK> &sh, uint32_t cnt) {
  for (uint32_t j = 0; j < cnt; ++j) {
    const uint32_t u = sh.u.a.big_v[j];
    const uint32_t d = sh.u.a.big_d[j];
    const int64_t s = sh.u.a.big_s[j];
    for (uint32_t k0 = 0; k0 < d; k0 += BLOCK * kUnroll) {
      bool act[kUnroll];
      uint32_t uu[kUnroll], w[kUnroll];
#pragma unroll
      for (int i = 0; i < kUnroll; ++i) {
        const uint32_t k = k0 + i * BLOCK + threadIdx.x;
        act[i] = k < d;
        uu[i] = u;
        w[i] = act[i] ? vis.col[s + k] : 0u;
      }
      vis.template visit_batch<kUnroll, true>(act, uu, w);
    }
  }
}

// Moves every wave's staged entries to the segment with one atomic per
// workgroup, and sums the wave-uniform `v` over the workgroup (thread 0
// gets the total in the return value).  With `mf` it also sums the lanes'
// out-degree sums of claimed vertices into *mf (every thread) and resets
// them, in the same pass.  Must be reached by all threads of the workgroup.
template <int BLOCK, class V>
__device__ __forceinline__ uint32_t flush_block(V &vis, SharedAll<BLOCK> &sh, uint32_t wave, uint32_t v = 0,
                                                uint64_t *mf = nullptr) {
  constexpr int W = KernelShape<BLOCK>::kWaves;
  if (lane_id() == 0) {
    sh.stage_n[wave] = vis.out_n;
    sh.wave_sum[wave] = v;
  }
  __syncthreads();
  uint32_t sum = 0;
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int i = 0; i < W; ++i) {
      sh.stage_base[i] = tot;
      tot += sh.stage_n[i];
      sum += sh.wave_sum[i];
    }
    uint32_t base = 0;
    if (tot) base = __hip_atomic_fetch_add(vis.seg_tail, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int i = 0; i < W; ++i) sh.stage_base[i] += base;
  }
  __syncthreads();
  const uint32_t cnt = vis.out_n, base = sh.stage_base[wave];
  for (uint32_t k = lane_id(); k < cnt; k += kWave) vis.seg[base + k] = vis.stage[k];
  vis.out_n = 0;
  if (mf) {
    const uint64_t w = wave_sum_u40(vis.mf);
    vis.mf = 0;
    if (lane_id() == 0) sh.wave_mf[wave] = w;
  }
  __syncthreads();
  if (mf) {
    // the slots are written again only after two more __syncthreads
    uint64_t t = 0;
    for (int i = 0; i < W; ++i) t += sh.wave_mf[i];
    *mf = uniform(t);
  }
  return sum;
}

// Workgroup sum of a per-lane value (the workgroup's total below 2^40), in
// every thread.  Per wave a DPP sum (no LDS permutes), then the wave sums
// through LDS.  Ends with a __syncthreads.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_sum_u40(SharedAll<BLOCK> &sh, uint64_t v) {
  constexpr int W = KernelShape<BLOCK>::kWaves;
  const uint64_t w = wave_sum_u40(v);
  if (lane_id() == 0) sh.wave_mf[threadIdx.x / kWave] = w;
  __syncthreads();
  uint64_t t = 0;
  for (int i = 0; i < W; ++i) t += sh.wave_mf[i];
  __syncthreads();  // the slots are free again
  return uniform(t);
}

// Workgroup sum of the per-lane out-degree sums of claimed vertices, in
// every thread; resets the lanes' counters.
template <int BLOCK, class V>
__device__ __forceinline__ uint64_t block_sum_mf(SharedAll<BLOCK> &sh, V &vis) {
  const uint64_t t = block_sum_u40<BLOCK>(sh, vis.mf);
  vis.mf = 0;
  return t;
}

// Bits of the level barrier's payload that carry the claim count when the
// out-degree sum shares the word (claims < 2^bits(N), sums <= M), or 0
// when the two do not fit its kCountShift bits together.
__host__ __device__ inline uint32_t fuse_bits_for(uint64_t n, uint64_t m) {
  uint32_t bn = 0, bm = 0;
  while (bn < 64 && (n >> bn)) ++bn;
  while (bm < 64 && (m >> bm)) ++bm;
  return bn + bm <= (uint32_t)kCountShift ? bn : 0u;
}

// Level barrier: returns the team's claim count and sets *mf_total to the
// out-degree sum of the claimed vertices.  With a.fuse_bits both travel in
// the one payload word (claims below 2^fuse_bits, the sum above them; the
// host sets it when the bits of N and of M fit the 54-bit payload), which
// spares the side sum's second chain of arrivals (~0.8 us per level).
__device__ __forceinline__ uint64_t level_barrier(const BfsArgs &a, BarrierState &bst, uint64_t claims,
                                                  uint64_t mf, uint64_t *s_word, uint64_t *mf_total) {
  const uint32_t fb = a.fuse_bits;
  if (fb) {
    const uint64_t t = grid_barrier<false, false>(&a.ctl->bar, bst, claims | (mf << fb), s_word);
    *mf_total = t >> fb;
    return t & ((1ull << fb) - 1);
  }
  return grid_barrier<false, false>(&a.ctl->bar, bst, claims, s_word, mf, mf_total);
}

// Sum over all lanes of the workgroup of a per-lane value into sh.prof[f]
// (profiling only).
template <int BLOCK>
__device__ __forceinline__ void prof_add(SharedAll<BLOCK> &sh, int f, uint64_t v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  if (lane_id() == 0) atomicAdd((unsigned long long *)&sh.prof[f], (unsigned long long)v);
}

// Bin counts and idle lane slots of one 64-vertex block of a phase A
// (profiling only; every lane of the wave calls it).
template <int BLOCK>
__device__ __forceinline__ void prof_bins(SharedAll<BLOCK> &sh, bool is_small, bool is_med, bool is_big,
                                          bool is_huge, uint32_t deg) {
  const uint32_t ns = (uint32_t)__popcll(__ballot(is_small)), nm = (uint32_t)__popcll(__ballot(is_med));
  const uint32_t nb = (uint32_t)__popcll(__ballot(is_big)), nh = (uint32_t)__popcll(__ballot(is_huge));
  uint32_t sdeg = is_small ? deg : 0u;
  uint32_t im = is_med ? (kWave - deg % kWave) % kWave : 0u;
  uint32_t ib = is_big ? ((uint32_t)BLOCK - deg % (uint32_t)BLOCK) % (uint32_t)BLOCK : 0u;
  uint32_t ih = is_huge ? (kWave - deg % kWave) % kWave : 0u;
  for (int d = 32; d >= 1; d >>= 1) {
    sdeg += __shfl_xor(sdeg, d);
    im += __shfl_xor(im, d);
    ib += __shfl_xor(ib, d);
    ih += __shfl_xor(ih, d);
  }
  if (lane_id() == 0) {
    // the wave gathers the block's lane-owned edges 64 per step
    const uint32_t is = (kWave - sdeg % kWave) % kWave;
    const uint32_t vals[8] = {ns, nm, nb, nh, is, im, ib, ih};
#pragma unroll
    for (int f = 0; f < 8; ++f)
      if (vals[f]) atomicAdd((unsigned long long *)&sh.prof[kProfBinSmall + f], (unsigned long long)vals[f]);
  }
}

template <int BLOCK>
__device__ __forceinline__ void prof_flush(const BfsArgs &a, SharedAll<BLOCK> &sh, uint32_t L) {
  __syncthreads();
  if (threadIdx.x < kProfFields) {
    // the workgroup's own entry count goes into the grid-wide maximum
    const uint64_t v = threadIdx.x == kProfEdgesMaxWg ? sh.prof[kProfEdges] : sh.prof[threadIdx.x];
    if (v && L < a.prof_levels) {
      unsigned long long *dst = (unsigned long long *)&a.prof[(size_t)L * kProfFields + threadIdx.x];
      if (threadIdx.x == kProfBusyMaxNs || threadIdx.x == kProfEdgesMaxWg) atomicMax(dst, (unsigned long long)v);
      else atomicAdd(dst, (unsigned long long)v);
    }
    sh.prof[threadIdx.x] = 0;
  }
  __syncthreads();
}

// Top-down phase A over this workgroup's equal share of the frontier.
// kProf: the per-level counters are compiled in (the profiling kernel only).
// kLow: every frontier vertex is known to be in the small bin (the graph's
// largest out-degree is at most a.small_max), so the other bins, their LDS
// queues and the workgroup barriers around them are compiled out.
// Shared data is first read after the workgroup barrier that follows the
// segment tails, so an invalidate by thread 0 before the call is ordered
// before those reads.
template <int BLOCK, bool kProf, bool kLow = false, class V>
__device__ __forceinline__ void topdown_phase_a(const BfsArgs &a, SharedAll<BLOCK> &sh, V &vis,
                                const uint32_t *in_seg, const LevelSlot &in_slot, LevelSlot &cur,
                                uint32_t wave, uint32_t trank, uint32_t tsize, uint64_t &edges) {

  using S = KernelShape<BLOCK>;
  const uint32_t tid = threadIdx.x;
  if (tid < kWave) tails_prefix<kSegments>(in_slot.tail, sh.seg_pre);
  if (tid == 0) {
    sh.med_n = 0;
    sh.big_n = 0;
  }
  __syncthreads();
  // 64-entry blocks of the frontier dealt round-robin over the workgroups:
  // coalesced reads, and a mix of early claims (often high degree) and late
  // ones in every workgroup
  const uint32_t total = sh.seg_pre[kSegments];
  const uint32_t nblk = (total + kWave - 1) / kWave;
  const uint32_t G = tsize, W = (uint32_t)S::kWaves;
  const uint32_t mine = nblk > trank ? (nblk - trank - 1) / G + 1 : 0;
  {
    for (uint32_t c0 = 0; c0 < mine; c0 += W) {
      const uint32_t c = c0 + wave;  // this wave's block ordinal within the workgroup
      const uint32_t p = (trank + c * G) * kWave + lane_id();
      uint32_t v = 0, deg = 0;
      int64_t start = 0;
      if (c < mine && p < total) {
        uint32_t j = 0;
#pragma unroll
        for (int k = 1; k < kSegments; ++k) j += sh.seg_pre[k] <= p;
        v = in_seg[(size_t)j * a.n + (p - sh.seg_pre[j])];
        start = a.off[v];
        deg = (uint32_t)(a.off[v + 1] - start);
        if (kLow || deg <= a.large_max) edges += deg;
      }
      const bool is_small = deg != 0 && (kLow || deg <= a.small_max);
      const bool is_med = !kLow && deg > a.small_max && deg <= a.medium_max;
      const bool is_big = !kLow && deg > a.medium_max && deg <= a.large_max;
      const bool is_huge = !kLow && deg > a.large_max;
      if (kProf && a.prof) prof_bins<BLOCK>(sh, is_small, is_med, is_big, is_huge, deg);
      // wave-aggregated slot reservation (ballot + mbcnt) for the bins
      const uint64_t mm = __ballot(is_med);
      if (mm) {
        uint32_t b0 = 0;
        if (lane_id() == 0) b0 = atomicAdd(&sh.med_n, (uint32_t)__popcll(mm));
        b0 = bcast_lane(b0, 0);
        if (is_med) {
          const uint32_t q = b0 + mbcnt(mm);
          sh.u.a.med_v[q] = v;
          sh.u.a.med_d[q] = deg;
          sh.u.a.med_s[q] = start;
        }
      }
      const uint64_t bm = __ballot(is_big);
      if (bm) {
        uint32_t b0 = 0;
        if (lane_id() == 0) b0 = atomicAdd(&sh.big_n, (uint32_t)__popcll(bm));
        b0 = bcast_lane(b0, 0);
        if (is_big) {
          const uint32_t q = b0 + mbcnt(bm);
          sh.u.a.big_v[q] = v;
          sh.u.a.big_d[q] = deg;
          sh.u.a.big_s[q] = start;
        }
      }
      // huge: kItem-entry work items, slots reserved by one atomic per wave
      const uint64_t hm = __ballot(is_huge);
      if (hm) {
        const uint32_t nit = is_huge ? (deg + kItem - 1) / kItem : 0u;
        const uint32_t incl = wave_inclusive_scan(nit);
        const uint32_t tot = bcast_lane(incl, 63);
        uint32_t b0 = 0;
        if (lane_id() == 0)
          b0 = __hip_atomic_fetch_add(&cur.item_tail[sh.seg], tot, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        b0 = bcast_lane(b0, 0);
        uint2 *dst = a.items + (size_t)sh.seg * a.item_cap + b0 + (incl - nit);
        for (uint32_t k = 0; k < nit; ++k) dst[k] = make_uint2(v, k);
      }
      gather_small<BLOCK>(vis, is_small ? v : 0u, is_small ? start : 0, is_small ? deg : 0u,
                          sh.u.a.g_excl[wave], sh.u.a.g_v[wave], sh.u.a.g_s[wave]);
      if (kLow) continue;  // gather_small is wave-local: no LDS queue to drain
      __syncthreads();
      const uint32_t mn = sh.med_n, bn = sh.big_n;
      __syncthreads();
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
  }
  __syncthreads();
  if (kLow) return;
  // drain (reference :161-169)
  const uint32_t mn = sh.med_n, bn = sh.big_n;
  expand_medium<BLOCK>(vis, sh, mn, wave);
  expand_large<BLOCK>(vis, sh, bn);
}

// Frontier edges from which a frontier of at most one entry per thread is
// expanded edge-balanced (below, degree bins cost less).
constexpr uint64_t kSmallFrontierEdges = 8192;

// Top-down level of a frontier with at most BLOCK entries: every workgroup
// reads the whole frontier and its degrees, and the team's waves split the
// concatenated edge list evenly (edge-balanced, no bins, no work items, so
// the level needs one barrier).  An edge's vertex is found by a binary
// search of the degree prefix in LDS.
template <int BLOCK, class V>
__device__ __forceinline__ void topdown_small(const BfsArgs &a, SharedAll<BLOCK> &sh, V &vis, const uint32_t *in_seg,
                              const LevelSlot &in_slot, uint32_t wave, uint32_t gwave, uint32_t nwaves,
                              uint64_t &edges) {
  constexpr int W = KernelShape<BLOCK>::kWaves;
  const uint32_t tid = threadIdx.x, lane = lane_id();
  if (tid < kWave) tails_prefix<kSegments>(in_slot.tail, sh.seg_pre);
  __syncthreads();
  const uint32_t total = sh.seg_pre[kSegments];
  uint32_t v = 0, deg = 0;
  int64_t start = 0;
  if (tid < total) {
    uint32_t j = 0;
#pragma unroll
    for (int k = 1; k < kSegments; ++k) j += sh.seg_pre[k] <= tid;
    v = in_seg[(size_t)j * a.n + (tid - sh.seg_pre[j])];
    start = a.off[v];
    deg = (uint32_t)(a.off[v + 1] - start);
  }
  // block-wide exclusive prefix of the degrees (the caller guarantees that
  // the frontier's out-degree sum fits 32 bits)
  const uint32_t incl = wave_inclusive_scan(deg);
  if (lane == kWave - 1) sh.u.f.wave_tot[wave] = incl;
  __syncthreads();
  uint32_t before = 0, T = 0;
  for (int i = 0; i < W; ++i) {
    const uint32_t t = (uint32_t)sh.u.f.wave_tot[i];
    if ((uint32_t)i < wave) before += t;
    T += t;
  }
  sh.u.f.v[tid] = v;
  sh.u.f.start[tid] = start;
  sh.u.f.pre[tid] = before + incl - deg;
  __syncthreads();
  const uint32_t lo = (uint32_t)((uint64_t)T * gwave / nwaves), hi = (uint32_t)((uint64_t)T * (gwave + 1) / nwaves);
  const uint32_t cnt = total ? total : 1;
  for (uint32_t e0 = lo; e0 < hi; e0 += kWave * kUnroll) {
    bool act[kUnroll];
    uint32_t u[kUnroll], w[kUnroll];
#pragma unroll
    for (int i = 0; i < kUnroll; ++i) {
      const uint32_t e = e0 + (uint32_t)i * kWave + lane;
      act[i] = e < hi;
      // last entry whose prefix is <= e: entries of degree 0 share their
      // prefix with the next one and are never chosen
      uint32_t j = 0;
      for (uint32_t step = 1u << 10; step >= 1; step >>= 1)
        if (j + step < cnt && sh.u.f.pre[j + step] <= e) j += step;
      u[i] = sh.u.f.v[j];
      w[i] = act[i] ? a.col[sh.u.f.start[j] + (int64_t)(e - sh.u.f.pre[j])] : 0u;
      if (act[i]) ++edges;
    }
    vis.template visit_batch<kUnroll, false>(act, u, w);
  }
  __syncthreads();
}

// Phase H: the level's work items, round-robin over every wavefront of the
// grid.  `total` items located through sh.item_pre.  A wave's items form a
// software pipeline: while item i's columns, probes and claims are in
// flight, item i+1's row offsets and item i+2's descriptor are loaded, so
// an item costs the trips of its own edges only.
template <int BLOCK, bool kLive = false, class V>
__device__ __forceinline__ void topdown_phase_h(const BfsArgs &a, SharedAll<BLOCK> &sh, V &vis, uint32_t total,
                                uint32_t gwave, uint32_t nwaves, uint64_t &edges) {

  const uint32_t lane = lane_id();
  auto item_at = [&](uint32_t i) {
    uint32_t j = 0;
#pragma unroll
    for (int k = 1; k < kSegments; ++k) j += sh.item_pre[k] <= i;
    return a.items[(size_t)j * a.item_cap + (i - sh.item_pre[j])];
  };
  uint32_t i = gwave;
  if (i >= total) return;
  uint2 it = item_at(i);
  int64_t s0 = a.off[it.x], e0 = a.off[it.x + 1];
  uint2 it1 = i + nwaves < total ? item_at(i + nwaves) : make_uint2(0u, 0u);
  for (; i < total; i += nwaves) {
    const uint32_t u = it.x;
    const int64_t s = s0 + (int64_t)it.y * kItem;
    const uint32_t len = (uint32_t)min((int64_t)kItem, e0 - s);
    if (lane == 0) edges += len;
    bool act[kUnroll];
    uint32_t uu[kUnroll], w[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      const uint32_t e = (uint32_t)k * kWave + lane;
      act[k] = e < len;
      uu[k] = u;
      w[k] = act[k] ? a.col[s + e] : 0u;
    }
    // wave-uniform: total < 2^32 - 2 * nwaves (8 segments of item_cap)
    int64_t s1 = 0, e1 = 0;
    if (i + nwaves < total) {
      s1 = a.off[it1.x];
      e1 = a.off[it1.x + 1];
    }
    const uint2 it2 = i + 2 * nwaves < total ? item_at(i + 2 * nwaves) : make_uint2(0u, 0u);
    vis.template visit_batch<kUnroll, false, kLive>(act, uu, w);
    it = it1;
    s0 = s1;
    e0 = e1;
    it1 = it2;
  }
}

// Frontier bitmap of a bottom-up level that follows a top-down one:
// fb = { v : level[v] == L }, one 64-vertex group per wave step, eight
// groups' level loads in flight at once.
__device__ __forceinline__ void build_frontier_bitmap(const BfsArgs &a, const uint32_t *level, uint32_t L, uint32_t *fb,
                                                      uint32_t gwave, uint32_t nwaves) {

  constexpr int B = 8;
  const uint32_t groups = (a.n + 63) / 64;
  const uint32_t lane = lane_id();
  for (uint32_t g0 = gwave; g0 < groups; g0 += B * nwaves) {
    uint32_t lv[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const uint32_t v = (g0 + i * nwaves) * 64 + lane;
      lv[i] = (g0 + i * nwaves < groups && v < a.n) ? level[v] : kInf;
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const uint32_t g = g0 + i * nwaves;
      if (g >= groups) break;  // wave-uniform
      const uint64_t m = __ballot(lv[i] == L);
      if (lane == 0) fb[2 * g] = (uint32_t)m;
      if (lane == 1 && 2 * g + 1 < a.n_words + 1) fb[2 * g + 1] = (uint32_t)(m >> 32);
    }
  }
}

// One bottom-up level.  Each wave takes chunks of kBuWords visited words:
// it lists the chunk's unvisited vertices in LDS, then every lane streams
// through its share of that list as a three-stage pipeline — row offsets
// (with the row's first four entries from the head array when there is
// one), four column indices, their frontier-bitmap words — so that each
// loop trip costs one memory round trip for the whole wave whatever stage
// each lane is in, and a lane that finishes a vertex starts the next at once.
// The first frontier neighbour in row order becomes the parent.  Claims are
// collected in LDS; the wave then writes the chunk's visited words and the
// next level's frontier-bitmap words itself (it owns them: no atomics).
template <int BLOCK, bool kProf>
__device__ __forceinline__ void bottomup_level(const BfsArgs &a, const SearchIo &io, SharedAll<BLOCK> &sh, Visitor &vis, uint32_t L,
                               const uint32_t *fb_cur, uint32_t *fb_nxt, uint32_t wave,
                               uint32_t gwave, uint32_t nwaves, uint64_t &edges) {

  const uint32_t lane = lane_id();
  const uint32_t *colb = a.col_bu ? a.col_bu : a.col;
  const uint4 *head = a.head;
  const bool hd = head != nullptr;
  uint32_t *cand = sh.u.cand[wave];
  uint32_t *cm = sh.bu_claim[wave];
  const uint32_t chunks = (a.n_words + kBuWords - 1) / kBuWords;
  for (uint32_t c = gwave; c < chunks; c += nwaves) {
    const uint32_t w0 = c * kBuWords;
    const uint32_t wi = w0 + lane;
    const bool own = lane < kBuWords && wi < a.n_words;
    const uint32_t word = own ? a.vis[wi] : 0xffffffffu;
    uint32_t unv = ~word;
    if (own && wi == a.n_words - 1 && (a.n & 31u)) unv &= (1u << (a.n & 31u)) - 1u;  // no vertex past n
    const uint32_t cnt = (uint32_t)__popc(unv);
    const uint32_t incl = wave_inclusive_scan(cnt);
    const uint32_t total = bcast_lane(incl, 63);
    if (kProf && a.prof && lane == 0 && total)
      atomicAdd((unsigned long long *)&sh.prof[kProfBuCandidates], (unsigned long long)total);
    if (lane < kBuWords) cm[lane] = 0;
    for (uint32_t pos = incl - cnt; unv; unv &= unv - 1u) cand[pos++] = (wi << 5) + (uint32_t)(__ffs(unv) - 1);
    wave_sync();

    // Lanes that need a vertex take the next ones of the list in order, so
    // the row-offset loads of one trip fall on neighbouring vertices.
    // Column indices come as one aligned 16-byte load per step (entries of
    // the block outside [k, e) are ignored), so a step is one request.
    // Every lane issues every load of a trip from straight-line code; lanes
    // with nothing to load for a kind read one shared address (the wave
    // merges them into one request), so the trip waits on memory once.
    uint32_t cursor = 0, stage = 0, v = 0;
    int64_t s0 = 0, k = 0, e = 0, kb = 0;
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
      const uint32_t *pk = colb + (st1 ? k : 0);
      const uint32_t back = (uint32_t)(((uintptr_t)pk & 15) >> 2);
      // with the head array, stage 0 fetches the row's first block with its
      // offsets: neighbouring candidates' heads are neighbouring 16-byte
      // words, so the trip's head loads are a few whole lines
      const uint4 *pq = (st0 && hd) ? head + v : (const uint4 *)(pk - back);
      uint32_t fi[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) fi[i] = (st2 && kb + i >= k && kb + i < e) ? (cw[i] >> 5) : 0u;
      const int64_t o0 = po[0], o1 = po[1];
      const uint4 q = *pq;  // derived from col or head: a global load
      uint32_t fw[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) fw[i] = fb_cur[fi[i]];
      // next state by selects only: a branch here would let the compiler
      // sink each load into the branch that uses it, one wait per branch
      bool hit = false;
      uint32_t par = 0;
#pragma unroll
      for (int i = 3; i >= 0; --i) {
        const bool h = (kb + i >= k) & (kb + i < e) & (((fw[i] >> (cw[i] & 31u)) & 1u) != 0u);
        hit = hit || h;
        par = h ? cw[i] : par;
      }
      const int64_t kn = kb + 4;
      const bool claimed = st2 && hit;
      const uint32_t scanned = st2 ? (uint32_t)(min(kn, e) - k) : 0u;
      vis.probes += scanned;
      edges += scanned;
      const uint32_t next2 = (!hit && kn < e) ? 1u : 0u;
      const uint32_t next0 = o1 > o0 ? (hd ? 2u : 1u) : 0u;
      stage = st0 ? next0 : (st1 ? 2u : (st2 ? next2 : stage));
      const bool take = st1 || (st0 && hd);
      cw[0] = take ? q.x : cw[0];
      cw[1] = take ? q.y : cw[1];
      cw[2] = take ? q.z : cw[2];
      cw[3] = take ? q.w : cw[3];
      kb = st1 ? k - (int64_t)back : ((st0 && hd) ? o0 : kb);  // CSR index of cw[0]
      const int64_t deg = e - s0;
      k = st0 ? o0 : (st2 ? kn : k);
      s0 = st0 ? o0 : s0;
      e = st0 ? o1 : e;
      if (claimed) {
        if (io.parent) io.parent[v] = par;
        vis.mf += (uint64_t)deg;
        atomicOr(&cm[(v >> 5) - w0], 1u << (v & 31u));
      }
      const uint64_t mask = __ballot(claimed);
      if (mask) vis.append(claimed, mask, v);
    }
    wave_sync();
    if (own) {
      const uint32_t got = cm[lane];
      if (got) a.vis[wi] = word | got;
      fb_nxt[wi] = got;
    }
    // levels of the chunk's claims, 64 consecutive vertices per store
    // instruction (a few requests instead of one per claim)
    for (uint32_t p = 0; p < kBuWords / 2; ++p) {
      const uint32_t bits = cm[2 * p + (lane >> 5)];
      if ((bits >> (lane & 31u)) & 1u) io.level[(w0 + 2 * p) * 32 + lane] = L + 1;
    }
    wave_sync();
  }
}

// Lazy invalidation after grid_barrier<.., false>: a workgroup about to
// read (with plain loads) data that other workgroups wrote since the
// barrier invalidates first, once per barrier; the flag is bit kFresh of a
// state word.  Every thread of the workgroup calls it with the same `need`.
constexpr uint32_t kFresh = 4;
__device__ __forceinline__ void fresh_bit(uint32_t &word, bool need) {
  if (need && !(word & kFresh)) {
    if (threadIdx.x == 0) l2_invalidate();
    __syncthreads();
    word |= kFresh;
  }
}

// A top-down level takes a visited-bitmap snapshot when its frontier has at
// least N / kSnapshotDiv out-edges and at least N / kSnapshotMinVisited
// vertices are visited.
constexpr uint64_t kSnapshotDiv = 16;
constexpr uint64_t kSnapshotMinVisited = 64;

// A snapshot level also probes the live bitmap (for the neighbours its
// snapshot shows unvisited) when its frontier has more than
// kLiveProbeRatio out-edges per visited vertex: most of its entries then
// name vertices claimed during the level, and the probe saves their
// atomics.  On a well-filled snapshot the probe is one more round trip
// per batch that saves few atomics.
constexpr uint64_t kLiveProbeRatio = 256;

// A top-down level without a snapshot whose frontier has at most this many
// out-edges claims by the atomic alone (Visitor::direct): the probe before
// it is one more wait on a level that is all latency, and the extra atomics
// are few (4096^2 grid 11.7 -> 11.3 us per level, RMAT-22 hybrid 159.8 ->
// 157.9 us per search, measured).
constexpr uint64_t kDirectMaxEdges = 65536;

// Solo levels: at most this many frontier vertices and out-edges.  Once
// team levels stopped invalidating every workgroup's caches, one
// workgroup's chain grew costlier than a team level's beyond ~512 vertices
// (4096^2 grid: 1,024 / 8,192 -> 512 / 4,096 took 10.9 -> 10.4 us per
// level; 256 / 2,048 the same, 64 / 512 10.5; RMAT-22 156.0 -> 154.9 us).
constexpr uint64_t kSoloMaxFrontier = 512;
constexpr uint64_t kSoloMaxEdges = 4096;

// End of a solo level: the workgroup's claims and out-degree sum, and its
// stores made visible to all its waves (same CU, so no cache maintenance).
template <int BLOCK>
__device__ __forceinline__ uint64_t solo_sync(SharedAll<BLOCK> &sh, uint32_t claims, uint64_t mf,
                                              uint64_t *mf_out) {
  wait_own_stores();
  if (threadIdx.x == 0) {
    sh.bcast[0] = claims;
    sh.bcast[1] = mf;
  }
  __syncthreads();
  *mf_out = uniform(sh.bcast[1]);
  const uint64_t n = uniform(sh.bcast[0]);
  __syncthreads();
  return n;
}

// Set in the hand-back tag when the search ended during the solo levels.
constexpr uint64_t kSoloDone = 1ull << 63;

// The solo workgroup hands the search back (or reports its end): its XCD's
// L2 is written back, the state goes to the control block, then the tag.
// quick_end: the search has ended and the caller ends the kernel or passes
// a grid barrier next, either of which writes the L2 back, so the tag goes
// out at once (the others only leave).
__device__ __forceinline__ void solo_publish(Control *ctl, uint64_t tag, uint32_t L, uint64_t n_cur,
                                             uint64_t n_prev, uint64_t m_f, uint64_t m_seen, uint64_t reached,
                                             uint32_t bu_levels, bool done, bool quick_end) {
  __syncthreads();
  if (threadIdx.x == 0) {
    if (!(done && quick_end)) {
      l2_writeback();  // waits for this thread's stores; the others waited before the barrier
      const uint64_t st[7] = {L, n_cur, n_prev, m_f, m_seen, reached, bu_levels};
      for (int i = 0; i < 7; ++i)
        __hip_atomic_store(&ctl->solo_state[i], st[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      wait_own_stores();
    }
    __hip_atomic_store(&ctl->solo_rel.word, tag | (done ? kSoloDone : 0ull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

// Other workgroups: wait for the solo workgroup's tag, then take the state.
// Returns true when the search ended during the solo levels.
template <int BLOCK>
__device__ __forceinline__ bool solo_wait(Control *ctl, SharedAll<BLOCK> &sh, uint64_t tag, uint32_t &L,
                                          uint64_t &n_cur, uint64_t &n_prev, uint64_t &m_f, uint64_t &m_seen,
                                          uint64_t &reached, uint32_t &bu_levels) {
  __shared__ uint64_t st[8];
  if (threadIdx.x < kWave) {
    const uint32_t lane = lane_id();
    if (lane == 0) {
      uint64_t w;
      while (((w = ld_relaxed(&ctl->solo_rel.word)) & ~kSoloDone) != tag) __builtin_amdgcn_s_sleep(2);
      st[7] = w & kSoloDone;
    }
    wave_sync();
    if (!st[7]) {
      if (lane == 0) l2_invalidate();
      // the seven state words in one request
      if (lane < 7) st[lane] = ld_relaxed(&ctl->solo_state[lane]);
    }
  }
  __syncthreads();
  L = uniform((uint32_t)st[0]);
  n_cur = uniform(st[1]);
  n_prev = uniform(st[2]);
  m_f = uniform(st[3]);
  m_seen = uniform(st[4]);
  reached = uniform(st[5]);
  bu_levels = uniform((uint32_t)st[6]);
  const bool done = uniform(st[7]) != 0;
  __syncthreads();
  return done;
}

// Initial state of a search (reference :216 cost init, done on the
// device), `gtid` of `gstride` threads: levels, visited bitmap and its
// snapshot (isolated vertices premarked), frontier bitmaps, and the
// level-0 frontier (the source alone, in segment 0 of buffer 0, whose
// tails live in the slot of "level -1", slot 3).  kThrough: every store is
// an agent-scope one, written through this XCD's L2, so once this thread
// has waited for its stores they are visible to every XCD without an L2
// write-back.
template <bool kThrough = false>
__device__ __forceinline__ void bfs_init(const BfsArgs &a, const SearchIo &io, uint32_t gtid, uint32_t gstride) {
  auto put = [](uint32_t *p, uint32_t x) {
    if (kThrough) __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = x;
  };
  for (uint32_t i = gtid; i < a.n; i += gstride) {
    put(&io.level[i], (i == io.source) ? 0u : kInf);
    if (io.parent) put(&io.parent[i], (i == io.source) ? io.source : kInf);
  }
  for (uint32_t w = gtid; w < a.n_words; w += gstride) {
    const uint32_t init = ((w == (io.source >> 5)) ? (1u << (io.source & 31u)) : 0u) | (a.iso ? a.iso[w] : 0u);
    put(&a.vis[w], init);
    put(&a.vis_snap[w], init);
  }
  const size_t fbs = (size_t)a.n_words + 1;  // words per frontier bitmap
  if (a.bottom_up)
    for (size_t w = gtid; w < 2 * fbs; w += gstride) put(&a.fb[w], 0u);
  // the four level slots' tails and item tails, one word per thread; the
  // level-0 frontier is the source alone in segment 0 of slot 3
  if (gtid < 4 * 2 * kSegments) {
    const uint32_t s = gtid / (2 * kSegments), k = gtid % (2 * kSegments);
    uint32_t *w = k < (uint32_t)kSegments ? &a.ctl->slot[s].tail[k] : &a.ctl->slot[s].item_tail[k - kSegments];
    put(w, (s == 3 && k == 0) ? 1u : 0u);
  }
  if (gtid == 0) {
    put(&a.qseg[0][0], io.source);
    put(&a.ctl->err, 0u);
  }
}

// Words of BfsArgs::result per search.
constexpr uint32_t kResultWords = 8;

// How a caller runs grid_bfs.  pre_init: bfs_init has already run over the
// grid and been released (the first grid barrier is skipped), and
// solo_seen is ctl->solo_rel as read before any workgroup could pass a
// barrier of this search.  quick_end: see solo_publish.
struct GridBfsMode {
  bool pre_init;
  bool quick_end;
  uint64_t solo_seen;
};

// One BFS over the whole grid (every workgroup must call it with the same
// arguments): initialisation, then the level loop until a level claims no
// vertex.  Results: the output arrays plus ctl->depth / reached / bu_levels.
// kProf: per-level counters into a.prof (compiled into the profiling kernel
// only, so the search kernels carry none of their registers).
// kLow (with kTd): the graph's largest out-degree is at most a.small_max:
// no medium, large or huge bins, no work items, no edge-balanced short
// frontiers (see topdown_phase_a).  Nor is the next frontier's out-degree
// sum counted per level (no degree loads for claimed vertices, no side sum
// in the level barrier): the level rules take the bound
// n_next * a.deg_max instead, and the CSR entries of reached vertices are
// summed once, at the end of the search, from the degrees each thread
// expanded (into result[3], or ctl->edges for a search without result
// words, which is then the launch's only search).
// kTd: every level top-down (a.bottom_up is ignored), compiled without the
// bottom-up code; the work items of a snapshot level with more than
// kLiveProbeRatio out-edges per visited vertex also probe the live bitmap.
// That probe stays out of the kernels that may go bottom-up: their large
// levels run bottom-up, so they seldom take a snapshot, and the probe's
// code alone cost the hybrid RMAT-22 search ~1 us.
template <int BLOCK, bool kProf = false, bool kTd = false, bool kLow = false>
__device__ __forceinline__ void grid_bfs(const BfsArgs &a, const SearchIo &io, SharedAll<BLOCK> &sh, Team &team,
                                         GridBfsMode mode = GridBfsMode{false, false, 0}) {
  using S = KernelShape<BLOCK>;
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = uniform(tid / kWave);  // wave-uniform: a scalar register
  const uint32_t trank = uniform(sh.team_rank), tsize = uniform(sh.team_size);
  const uint32_t gtid = trank * BLOCK + tid;
  const uint32_t gstride = tsize * BLOCK;
  const uint32_t gwave = trank * S::kWaves + wave;
  const uint32_t nwaves = tsize * S::kWaves;
  BarrierState &bst = team.bst;
  Control *ctl = a.ctl;
  if (kProf && tid < kProfFields) sh.prof[tid] = 0;

  if (io.result && gtid == 0) io.result[4] = now_ns();
  // kLow: summed by atomics at the end of the search; zeroed by an
  // agent-scope store, which leaves no cached copy of the word behind
  uint64_t *const e_sum = io.result ? &io.result[3] : &ctl->edges;
  if (kLow && gtid == 0) __hip_atomic_store(e_sum, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t solo_next = 0;  // tag of the next hand-back of solo levels (thread 0)
  if (mode.pre_init) {
    if (tid == 0) solo_next = (mode.solo_seen & ~kSoloDone) + 1;
  } else {
    if (tid == 0) solo_next = (ld_relaxed(&ctl->solo_rel.word) & ~kSoloDone) + 1;
    bfs_init(a, io, gtid, gstride);
    grid_barrier<false, false>(&ctl->bar, bst, 0, sh.bcast);
  }
  const size_t fbs = (size_t)a.n_words + 1;  // words per frontier bitmap

  Visitor vis;
  vis.col = a.col;
  vis.level = io.level;
  vis.parent = io.parent;
  vis.vis = a.vis;
  vis.snap = a.vis;
  vis.probe_live = false;
  vis.direct = false;
  vis.stage = sh.stage[wave];
  vis.out_n = 0;
  vis.probes = 0;
  vis.atomics = 0;

  vis.off = a.off;
  vis.fb_mark = nullptr;
  // out-degree sums of the claimed vertices: the direction heuristic, the
  // choice of the small-frontier expansion, and ctl->edges
  vis.track_mf = !kLow;
  vis.deg32 = a.m <= 0xffffffffull;  // no row can have 2^32 entries
  vis.mf = 0;

  uint64_t n_cur = 1, n_prev = 0, reached = 1;
  // m_f: out-degree sum of the current frontier; m_seen: of every vertex
  // placed in a frontier so far (m - m_seen = edges left to explore)
  uint64_t m_f = (uint64_t)(a.off[io.source + 1] - a.off[io.source]);
  uint64_t m_seen = m_f;
  uint64_t e_total = 0;  // kLow: CSR entries this thread expanded, all levels
  uint32_t bu_levels = 0;
  bool bottom_up = false;
  // Frontier bitmaps alternate by level parity: fb[L & 1] is level L's.
  // fb_zero bit b: fb[b] is known to be all zero; fb_valid: fb[L & 1] holds
  // exactly level L's frontier.  A top-down level zeroes fb[L & 1] (unused
  // there) and, when fb[(L + 1) & 1] is known zero, ORs its claims into it,
  // so a bottom-up level right after it needs no level-array scan; a
  // bottom-up level writes the next bitmap whole.
  // bit kFresh of the same word: this workgroup has invalidated its caches
  // since the last level barrier (those leave caches as they are; see
  // fresh_bit), kept here rather than in a register of its own
  uint32_t fb_zero = 3u | (mode.pre_init ? kFresh : 0u);
  bool fb_valid = false;
  // Solo levels: a small top-down level (and the small ones after it) is
  // run by the team's first workgroup alone, separated by workgroup
  // barriers only, while the others wait for it to publish the search
  // state (or the end of the search).
  bool solo = false;
  uint64_t solo_tag = 0;
  for (uint32_t L = 0;; ++L) {
    // selects, not a[L & 1]: an array member indexed at run time would put
    // the whole argument block in scratch memory
    const uint32_t *in_seg = (L & 1) ? a.qseg[1] : a.qseg[0];
    const LevelSlot &in_slot = ctl->slot[(L + 3) & 3];
    LevelSlot &cur = ctl->slot[L & 3];
    if (gtid == 0)
      for (int k = 0; k < kSegments; ++k) {
        ctl->slot[(L + 2) & 3].tail[k] = 0;
        ctl->slot[(L + 2) & 3].item_tail[k] = 0;
      }
    // direction choice (Beamer et al.): identical inputs in every workgroup
    if (!kTd && a.bottom_up) {
      const uint64_t m_u = a.m > m_seen ? a.m - m_seen : 0;
      if (!bottom_up && n_cur > n_prev && m_f * a.alpha > m_u) bottom_up = true;
      else if (bottom_up && n_cur < a.bu_exit && n_cur < n_prev) bottom_up = false;
    }
    const bool small = !bottom_up && tsize > 1 && n_cur <= kSoloMaxFrontier && m_f <= kSoloMaxEdges;
    if (solo && !small) {
      solo_publish(ctl, solo_tag, L, n_cur, n_prev, m_f, m_seen, reached, bu_levels, false, false);
      solo = false;
    } else if (!solo && small) {
      // tags count hand-backs: every workgroup read the last one before
      // the search's first barrier
      if (tid == 0) solo_tag = solo_next++;
      if (trank != 0) {
        // wait until the first workgroup is done with the small levels
        const bool done = solo_wait(ctl, sh, solo_tag, L, n_cur, n_prev, m_f, m_seen, reached, bu_levels);
        if (done) break;
        // solo levels leave the frontier bitmaps alone (fb_zero holds) but
        // do not build the next one
        fb_valid = false;
        fb_zero |= kFresh;  // solo_wait invalidated
        --L;  // the loop's ++L gives the level to resume at
        continue;
      }
      solo = true;
    }
    // this workgroup's share: the whole team's, or alone in solo levels
    const uint32_t r_gwave = solo ? wave : gwave, r_nwaves = solo ? (uint32_t)S::kWaves : nwaves;
    vis.seg = ((L & 1) ? a.qseg[0] : a.qseg[1]) + (size_t)sh.seg * a.n;
    vis.seg_tail = &cur.tail[sh.seg];
    vis.next_level = L + 1;
    vis.claims = 0;
    uint64_t edges = 0;
    uint64_t n_next;
    uint32_t hc = 0;
    uint64_t t_start = 0, busy = 0;
    if (kProf && a.prof && tid == 0) t_start = now_ns();
    if (!bottom_up) {
      bool fb_track = false;
      vis.fb_mark = nullptr;
      // A level with many out-edges filters through a snapshot of the
      // visited bitmap taken at its start (read-only while the atomics hit
      // the live words); words copied while claims go on are still subsets
      // of the visited set, which is all the filter needs.  Smaller levels
      // (and solo ones) probe the live bitmap alone, and so does a large
      // level while fewer than N / kSnapshotMinVisited vertices are
      // visited: its snapshot would filter nothing.
      const bool snapshot = !solo && m_f * kSnapshotDiv >= (uint64_t)a.n &&
                            reached * kSnapshotMinVisited >= (uint64_t)a.n;
      vis.snap = snapshot ? a.vis_snap : a.vis;
      vis.probe_live = kTd && snapshot && m_f > reached * kLiveProbeRatio;
      vis.direct = !snapshot && m_f <= kDirectMaxEdges;
      if (snapshot) {
        fresh_bit(fb_zero, true);
        for (uint32_t w = gtid; w < a.n_words; w += gstride) a.vis_snap[w] = a.vis[w];
      }
      // A growing frontier may turn the next level bottom-up: zero this
      // level's bitmap parity for later and OR the claims into the next
      // one when it is known to be zero.  Solo levels and shrinking
      // frontiers leave the bitmaps alone (a bottom-up level after them
      // builds its bitmap from the level array).
      if (!kTd && a.bottom_up && !solo && n_cur > n_prev) {
        uint32_t *fb_here = a.fb + (L & 1) * fbs;
        for (size_t w = gtid; w < fbs; w += gstride) fb_here[w] = 0;
        fb_track = (fb_zero >> ((L + 1) & 1)) & 1u;
        vis.fb_mark = fb_track ? a.fb + ((L + 1) & 1) * fbs : nullptr;
        fb_zero = (fb_zero & kFresh) | (1u << (L & 1));
      }
      fb_valid = fb_track;
      // a short frontier with many edges (hubs): edge-balanced over the
      // team in one pass; otherwise degree bins
      if (!kLow && !solo && n_cur <= (uint64_t)BLOCK && m_f >= kSmallFrontierEdges && m_f <= 0xffffffffull) {
        fresh_bit(fb_zero, true);  // every workgroup reads the whole frontier
        topdown_small<BLOCK>(a, sh, vis, in_seg, in_slot, wave, gwave, nwaves, edges);
      } else {
        // block b of the frontier goes to rank b mod the team size: only
        // ranks below the block count read shared data
        const uint32_t ra = solo ? 0u : trank;
        if (!(fb_zero & kFresh) && (uint64_t)ra * kWave < n_cur) {
          if (tid == 0) l2_invalidate();
          fb_zero |= kFresh;
        }
        topdown_phase_a<BLOCK, kProf, kLow>(a, sh, vis, in_seg, in_slot, cur, wave, ra, solo ? 1u : tsize, edges);
      }
      uint64_t wg_mf = 0;
      const uint32_t wg_claims = flush_block<BLOCK>(vis, sh, wave, vis.claims, kLow ? nullptr : &wg_mf);
      if (kProf && a.prof && tid == 0) busy += now_ns() - t_start;
      uint64_t mf1 = 0;
      if (solo) {
        n_next = solo_sync<BLOCK>(sh, wg_claims, wg_mf, &mf1);
      } else {
        n_next = kLow ? grid_barrier<false, false>(&ctl->bar, bst, wg_claims, sh.bcast)
                      : level_barrier(a, bst, wg_claims, wg_mf, sh.bcast, &mf1);
        fb_zero &= ~kFresh;
      }
      m_f = kLow ? n_next * a.deg_max : mf1;
      if (!kLow) {
        if (tid < kWave) tails_prefix<kSegments>(cur.item_tail, sh.item_pre);
        __syncthreads();
        hc = uniform(sh.item_pre[kSegments]);
      }
      if (hc) {
        const uint32_t before = vis.claims;
        uint64_t th = 0;
        if (kProf && a.prof && tid == 0) th = now_ns();
        // a workgroup reads work items only when its first wave has one
        fresh_bit(fb_zero, (solo ? 0u : trank * (uint32_t)S::kWaves) < hc);
        topdown_phase_h<BLOCK, kTd>(a, sh, vis, hc, r_gwave, r_nwaves, edges);
        uint64_t wg_mf2 = 0;
        const uint32_t wg2 = flush_block<BLOCK>(vis, sh, wave, vis.claims - before, &wg_mf2);
        if (kProf && a.prof && tid == 0) busy += now_ns() - th;
        uint64_t mf2 = 0;
        if (solo) {
          n_next += solo_sync<BLOCK>(sh, wg2, wg_mf2, &mf2);
        } else {
          n_next += level_barrier(a, bst, wg2, wg_mf2, sh.bcast, &mf2);
          fb_zero &= ~kFresh;
        }
        m_f += mf2;
      }
    } else {
      ++bu_levels;
      uint32_t *fb_cur = a.fb + (L & 1) * fbs;
      uint32_t *fb_nxt = a.fb + ((L + 1) & 1) * fbs;
      if (!fb_valid) {
        fresh_bit(fb_zero, true);
        build_frontier_bitmap(a, io.level, L, fb_cur, gwave, nwaves);
        grid_barrier<false, false>(&ctl->bar, bst, 0, sh.bcast);
        fb_zero &= ~kFresh;
      }
      fresh_bit(fb_zero, true);
      vis.fb_mark = nullptr;
      fb_zero &= kFresh;
      fb_valid = true;
      uint64_t tb = 0;
      if (kProf && a.prof && tid == 0) tb = now_ns();
      bottomup_level<BLOCK, kProf>(a, io, sh, vis, L, fb_cur, fb_nxt, wave, gwave, nwaves, edges);
      uint64_t wg_mf = 0;
      const uint32_t wg_claims = flush_block<BLOCK>(vis, sh, wave, vis.claims, &wg_mf);
      if (kProf && a.prof && tid == 0) busy += now_ns() - tb;
      uint64_t mf1 = 0;
      n_next = level_barrier(a, bst, wg_claims, wg_mf, sh.bcast, &mf1);
      fb_zero &= ~kFresh;
      m_f = mf1;
    }
    m_seen += m_f;
    if (kProf && a.prof) {
      prof_add<BLOCK>(sh, kProfEdges, edges);
      prof_add<BLOCK>(sh, kProfProbes, vis.probes);
      prof_add<BLOCK>(sh, kProfAtomics, vis.atomics);
      if (tid == 0) {
        sh.prof[kProfBusyMaxNs] = busy;
        sh.prof[kProfBusySumNs] = busy;
      }
      if (tid == 0 && trank == 0) {
        sh.prof[kProfFrontier] = n_cur;
        sh.prof[kProfClaims] = n_next;
        sh.prof[kProfHuge] = hc;
        sh.prof[kProfBottomUp] = bottom_up ? 1 : 0;
        sh.prof[kProfLevelNs] = now_ns() - t_start;
      }
      prof_flush<BLOCK>(a, sh, L);
      vis.probes = 0;
      vis.atomics = 0;
    }
    if (kLow) e_total += edges;
    reached += n_next;
    if (n_next == 0) {
      if (gtid == 0) {
        ctl->depth = L + 1;
        ctl->reached = reached;
        ctl->bu_levels = bu_levels;
        if (!kLow) ctl->edges = m_seen;
        if (io.result) {
          io.result[0] = L + 1;
          io.result[1] = reached;
          io.result[2] = bu_levels;
          if (!kLow) io.result[3] = m_seen;
          io.result[5] = now_ns();
        }
      }
      if (solo) solo_publish(ctl, solo_tag, L, n_cur, n_prev, m_f, m_seen, reached, bu_levels, true, mode.quick_end);
      break;
    }
    n_prev = n_cur;
    n_cur = n_next;
  }
  if (kLow) {
    const uint64_t wg_e = block_sum_u40<BLOCK>(sh, e_total);
    if (tid == 0 && wg_e) __hip_atomic_fetch_add(e_sum, wg_e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Kernel start: census of workgroups per XCD and formation of teams of
// `xpt` XCDs each, whose barriers are team_bar(id).  Returns false in every
// workgroup alike when a team would span more XCDs than frontier segments.
// plain: 0 for a cooperative launch, else the host's number of this plain
// launch, whose grid assembles itself (plain_census); when it does not,
// every workgroup returns false with *abandoned set.
template <int BLOCK, class TeamBar>
__device__ __forceinline__ bool kernel_prologue(GridBarrier *census_bar, uint32_t xpt, bool by_ticket, TeamBar team_bar,
                                SharedAll<BLOCK> &sh, Team &team, uint32_t plain = 0, uint64_t plain_wait_ns = 0,
                                bool *abandoned = nullptr) {
  CensusInfo ci;
  if (plain) {
    if (threadIdx.x < kWave) {
      const bool go = plain_census(census_bar, plain, plain_wait_ns, team_bar, ci, sh.census, sh.census + kMaxXcd);
      if (threadIdx.x == 0) sh.bcast[1] = go ? 1u : 0u;
    }
    __syncthreads();
    if (!sh.bcast[1]) {
      *abandoned = true;
      return false;
    }
  } else if (threadIdx.x < kWave) {
    barrier_census(census_bar, team_bar, ci, sh.census, sh.census + kMaxXcd);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (xpt == 0 || xpt > ci.nx) xpt = ci.nx;
    team_form(ci, sh.census, sh.census + kMaxXcd, xpt, team_bar(ci.xrank / xpt), by_ticket, team);
    sh.seg = team.seg;
    sh.team_rank = team.rank;
    sh.team_size = team.size;
    sh.team_id = team.id;
    sh.bcast[0] = xpt;
  }
  __syncthreads();
  return sh.bcast[0] <= (uint64_t)kSegments;
}

template <int BLOCK, bool kProf = false, bool kTd = false, bool kLow = false>
__global__ __launch_bounds__(BLOCK) void dfq_bfs_kernel(BfsArgs a) {
  __shared__ SharedAll<BLOCK> sh;
  Team team{};
  GridBarrier *bar = &a.ctl->bar;
  // One search per launch: each workgroup writes its blockIdx share of the
  // initial state with stores that go through to memory and waits for
  // them before it is counted, so the census, which every workgroup
  // passes only after all were counted, releases the state (no grid
  // barrier of its own).  solo_rel is read before this workgroup is
  // counted, so before any hand-back of this search.
  uint64_t solo_seen = 0;
  if (threadIdx.x == 0) solo_seen = ld_relaxed(&a.ctl->solo_rel.word);
  const SearchIo io{a.level, a.parent, a.result, a.source};
  bfs_init<true>(a, io, blockIdx.x * BLOCK + threadIdx.x, gridDim.x * BLOCK);
  wait_own_stores();
  __syncthreads();
  bool abandoned = false;
  if (!kernel_prologue<BLOCK>(bar, 0, false, [bar](uint32_t) { return bar; }, sh, team, a.plain, a.plain_wait_ns,
                              &abandoned)) {
    if (abandoned) {
      // a plain launch whose grid did not assemble: the host reruns it
      if (threadIdx.x == 0) {
        a.ctl->abandoned = a.plain;
        if (a.result) a.result[6] = a.plain;
      }
    } else if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.ctl->err = 2;
    }
    return;
  }
  // acquire: drop lines this CU or XCD may hold from before the release
  if (threadIdx.x == 0) l2_invalidate();
  __syncthreads();
  // the kernel ends after the search, which writes every L2 back
  grid_bfs<BLOCK, kProf, kTd, kLow>(a, io, sh, team, GridBfsMode{true, true, solo_seen});
}

// Many searches of one graph in one launch, in turn: search i starts from
// sources[i] and writes level + i * level_stride (and parent + i *
// parent_stride when parents are asked for) and result + i * kResultWords.
// Each search's first grid barrier also ends the previous one.
template <int BLOCK, bool kTd = false, bool kLow = false>
__global__ __launch_bounds__(BLOCK) void dfq_bfs_seq_kernel(const BfsArgs a, const uint32_t *sources, uint32_t count,
                                                            uint64_t level_stride, uint64_t parent_stride) {
  __shared__ SharedAll<BLOCK> sh;
  Team team{};
  GridBarrier *bar = &a.ctl->bar;
  if (!kernel_prologue<BLOCK>(bar, 0, false, [bar](uint32_t) { return bar; }, sh, team)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl->err = 2;
    return;
  }
  for (uint32_t i = 0; i < count; ++i) {
    const SearchIo io{a.level + i * level_stride, a.parent ? a.parent + i * parent_stride : nullptr,
                      a.result + (size_t)i * kResultWords, uniform(sources[i])};
    // the last search ends with the kernel; the others are followed by the
    // next one's initialisation, which needs their stores written back
    grid_bfs<BLOCK, false, kTd, kLow>(a, io, sh, team, GridBfsMode{false, i + 1 == count, 0});
  }
}

}  // namespace dfq
