// Device-side building blocks of the DFQ-BFS kernels for gfx950 (CDNA4,
// wave64): wavefront primitives, the visited-bitmap claim, the grid barrier
// and the per-wave output blocks of the decentralised frontier queue.
//
// Reference counterparts (design/shared-q/atomic-sa/our-sep/upper-bound/
// bfs.cu): VISIT_tiled / VISIT_blk_tiled / VISIT_grid_tiled /
// VISIT_all_tiled (:44-94) test `cost[nid] < 0` and write `cost[ftr] + 1`
// without atomics, find the next frontier by rescanning all N levels
// (:116-124) and separate levels with cooperative-groups grid.sync()
// (:114,170,177).  Here a vertex is claimed exactly once through an atomic OR
// on a visited bitmap; the claimer alone writes its level (so levels are
// deterministic) and parent and appends it to its wavefront's block of the
// next-frontier queue; levels are separated by a tree barrier that also sums
// the next frontier's size.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dfq {

constexpr uint32_t kInf = 0xffffffffu;
constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// 100 MHz constant-rate counter -> ns
__device__ __forceinline__ uint64_t now_ns() { return __builtin_amdgcn_s_memrealtime() * 10ull; }

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

// A value every lane of the wave holds alike (read from one LDS word, say),
// moved to a scalar register: branches on it are then scalar branches and
// it takes no vector register.
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ uint64_t uniform(uint64_t x) {
  const uint32_t lo = uniform((uint32_t)x), hi = uniform((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t bcast_lane(uint32_t x, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

// Inclusive prefix sum over the 64 lanes of a wavefront.  row_shr DPP for
// the steps inside a 16-lane row, row_bcast:15 / row_bcast:31 across rows.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
  // update_dpp(old, src, dpp_ctrl, row_mask, bank_mask, bound_ctrl): lanes
  // whose source is outside the row, or whose row is masked off, get `old`.
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

// Sum over the 64 lanes of a value below 2^40 (claim counts and out-degree
// sums are below 2^40), in a scalar register: DPP scans of its two 20-bit
// halves (neither half's sum can overflow 32 bits), no LDS permutes.
// Every lane of the wave must be active.
__device__ __forceinline__ uint64_t wave_sum_u40(uint64_t v) {
  const uint32_t a = wave_inclusive_scan((uint32_t)v & 0xfffffu);
  const uint32_t b = wave_inclusive_scan((uint32_t)(v >> 20) & 0xfffffu);
  return (uint64_t)bcast_lane(a, 63) + ((uint64_t)bcast_lane(b, 63) << 20);
}

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of the kSegments tails into pre[0..kSegments], the total
// in pre[kSegments].  Wave 0 only: lanes 0..kSegments-1 load one tail each,
// so the tails cost one memory round trip instead of a chain of them; the
// prefix is row_shr DPP steps inside the first 16-lane row.
template <int kN>
__device__ __forceinline__ void tails_prefix(const uint32_t *tails, uint32_t *pre) {
  static_assert(kN <= 16 && (kN & (kN - 1)) == 0, "tail count");
  const uint32_t lane = lane_id();
  const uint32_t t = lane < (uint32_t)kN ? ld_agent(&tails[lane]) : 0u;
  uint32_t x = t;
  // lanes with no source in the row get 0; lanes >= kN hold t = 0, so the
  // sums of lanes 0..kN-1 are their inclusive prefixes
  if (kN > 1) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  if (kN > 2) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  if (kN > 4) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  if (kN > 8) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  if (lane < (uint32_t)kN) pre[lane] = x - t;
  if (lane == (uint32_t)kN - 1) pre[kN] = x;
}

// ---------------------------------------------------------------------------
// Grid barrier, XCD-aware.
//
// MI355X has 8 XCDs, each with its own L2; a store is visible to another XCD
// only after its L2 line is written back, and a load on another XCD may hit a
// stale line of that XCD's L2 until it is invalidated.  The language-level
// agent fence does both per workgroup (buffer_wbl2 / buffer_inv sc1), which
// at 1024 workgroups costs ~16 us per barrier and, issued by workgroups that
// arrive early, keeps flushing the L2 under the ones still working.  Here:
//   * every wave waits for its own stores (s_waitcnt vmcnt(0)): they are then
//     in its XCD's L2;
//   * workgroups arrive at their XCD's counter; the last one of each XCD
//     writes that L2 back once (buffer_wbl2 sc1) and arrives at the root;
//   * the root's last arrival publishes generation + payload sum in one
//     64-bit word, which every workgroup polls with relaxed loads;
//   * on release every workgroup invalidates (buffer_inv sc1: device scope
//     is the only one that drops its CU's L1 lines, and it also drops the
//     XCD's L2 lines) before reading shared data.
// Workgroups find their XCD from the XCC_ID hardware register; the number
// of workgroups per XCD is counted at kernel start (barrier_census), because
// the dispatcher's block -> XCD rotation changes from launch to launch.
// Arrival words pack count (bits 54..63) and payload (bits 0..53); the
// release words generation (bits 54..63, modulo 1024) and payload sum.
// ---------------------------------------------------------------------------
constexpr int kMaxXcd = 16;
constexpr int kCountShift = 54;
constexpr uint64_t kArrive = 1ull << kCountShift;
constexpr uint64_t kPayloadMask = kArrive - 1;
constexpr uint32_t kGenMask = (1u << (64 - kCountShift)) - 1;
// Largest CSR entry count a search accepts: claim counts and out-degree
// sums travel as sums below 2^40 (wave and workgroup sums, barrier
// payloads).
constexpr uint64_t kMaxEntries = (1ull << 40) - 1;

struct alignas(256) BarrierWord {
  uint64_t word;
  uint64_t pad[31];
};

struct GridBarrier {
  BarrierWord xleaf[kMaxXcd];   // per-XCD arrivals
  BarrierWord xside[kMaxXcd];   // per-XCD side arrivals: count (bits 54..63) | side sum
  BarrierWord root;             // XCD-leader arrivals
  BarrierWord root_side;        // XCD side arrivals: count | side sum
  BarrierWord side_out;         // generation (bits 54..63) | side total of that round
  BarrierWord gensum;           // generation (bits 54..63) | payload sum (bits 0..53)
  BarrierWord census_epoch;     // launches counted so far (low bit picks the census set)
  BarrierWord ticket;           // workgroup ranks handed out within a team
  uint32_t census[2][kMaxXcd];  // workgroups per XCD, one set per launch parity
  // plain (non-cooperative) launches, by the parity of the host's launch
  // number: arrivals (low bits) | kPlainAbandon, and workgroups per XCD
  BarrierWord plain_word[2];
  uint32_t plain_census[2][kMaxXcd];
};

// Barrier state of one workgroup within its team (thread 0's copy counts).
struct BarrierState {
  uint32_t gen;     // generation modulo 2^24
  uint32_t xcc;     // this workgroup's XCD
  uint32_t xcc_n;   // workgroups of the team on this XCD
  uint32_t nx;      // XCDs in the team
};

// What the kernel-start census tells every workgroup.
struct CensusInfo {
  uint32_t xcc;                // this workgroup's XCD
  uint32_t xrank;              // its rank among the XCDs that have workgroups
  uint32_t nx;                 // XCDs with workgroups
};

// A team: the workgroups of `xpt` consecutive XCD ranks, with its own
// barrier words.  The whole grid is one team of all XCDs.
struct Team {
  uint32_t id;      // team index
  uint32_t rank;    // workgroup index within the team
  uint32_t size;    // workgroups in the team
  uint32_t seg;     // output segment: this workgroup's XCD rank within the team
  uint32_t teams;   // teams in the grid
  BarrierState bst; // thread 0 only
};

__device__ __forceinline__ uint32_t xcc_id() {
  uint32_t v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(v));
  return v & (kMaxXcd - 1);
}

__device__ __forceinline__ void wait_own_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void l2_writeback() {
  asm volatile("buffer_wbl2 sc1\n\ts_waitcnt vmcnt(0)" ::: "memory");
}
// Write-back at system scope: the lines become visible to the host and to
// copy engines, not only to the other XCDs.
__device__ __forceinline__ void l2_writeback_system() {
  asm volatile("buffer_wbl2 sc0 sc1\n\ts_waitcnt vmcnt(0)" ::: "memory");
}
__device__ __forceinline__ void l2_invalidate() {
  asm volatile("buffer_inv sc1\n\ts_waitcnt vmcnt(0)" ::: "memory");
}

template <class T>
__device__ __forceinline__ T ld_relaxed(T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Kernel start (wave 0 of every workgroup, all workgroups resident):
// counts workgroups per XCD.  Alternate launches count in alternate sets
// (census_epoch's low bit), and once a launch's count is complete its
// workgroup 0 clears the other set for the next launch and advances the
// epoch, so no workgroup waits for the last reader.  Lanes 0..15 poll the
// sixteen counters with one load; lanes 16..31 read the barrier generation
// of teams 0..15 alongside (stable until the team's first barrier, which
// needs this workgroup).  Results: `count` holds the non-zero counters in
// XCD order and `gen` the generations, both in LDS.
template <class TeamBar>
__device__ __forceinline__ void barrier_census(GridBarrier *b, TeamBar team_bar, CensusInfo &ci,
                                               uint32_t *count, uint32_t *gen) {
  const uint32_t lane = lane_id();
  ci.xcc = xcc_id();
  const uint64_t epoch = ld_relaxed(&b->census_epoch.word);
  uint32_t *cnt = b->census[epoch & 1];
  if (lane == 0) __hip_atomic_fetch_add(&cnt[ci.xcc], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t c = 0, g = 0;
  for (;;) {
    if (lane < (uint32_t)kMaxXcd) c = ld_relaxed(&cnt[lane]);
    else if (lane < 2u * kMaxXcd) g = (uint32_t)(ld_relaxed(&team_bar(lane - kMaxXcd)->gensum.word) >> kCountShift);
    uint32_t sum = c;
    sum += __shfl_xor(sum, 1);
    sum += __shfl_xor(sum, 2);
    sum += __shfl_xor(sum, 4);
    sum += __shfl_xor(sum, 8);
    if ((uint32_t)__shfl(sum, 0) >= gridDim.x) break;
    __builtin_amdgcn_s_sleep(1);
  }
  const uint32_t nz = (uint32_t)__ballot(lane < (uint32_t)kMaxXcd && c != 0);
  const uint32_t below = nz & ((1u << (lane & 31u)) - 1u);
  if (lane < (uint32_t)kMaxXcd && c != 0) count[__popc(below)] = c;
  if (lane >= (uint32_t)kMaxXcd && lane < 2u * kMaxXcd) gen[lane - kMaxXcd] = g;
  ci.nx = __popc(nz);
  ci.xrank = __popc(nz & ((1u << ci.xcc) - 1u));
  if (blockIdx.x == 0 && lane == 0) {
    uint32_t *other = b->census[(epoch & 1) ^ 1];
    for (int i = 0; i < kMaxXcd; ++i) __hip_atomic_store(&other[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&b->census_epoch.word, epoch + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Kernel start of a plain (non-cooperative) launch, wave 0 of every
// workgroup.  Nothing guarantees that every workgroup of such a launch is
// resident at once (other work may hold CUs), so the grid assembles itself
// with a bounded wait and one decision word: every workgroup adds itself to
// its XCD's counter and then to the arrival count; it proceeds once the
// count reaches the grid with the abandon bit clear, and gives up when the
// bit is set.  A workgroup that has waited `wait_ns` sets the bit with a
// compare-and-swap that only succeeds while the count is short, so a grid
// that completed can no longer be abandoned and an abandoned one never
// proceeds: all workgroups decide alike, late ones included (they see the
// bit in their own arrival's result).  The host then relaunches
// cooperatively.  `epoch` is the host's number of the launch: alternate
// launches use alternate words, and the workgroup that decides (workgroup
// 0 on success, the one whose swap succeeded otherwise) clears the other
// parity's words for the next launch, whose predecessor has ended.
// Returns false in every workgroup when the launch is abandoned; otherwise
// fills ci, count and gen as barrier_census does.
constexpr uint64_t kPlainAbandon = 1ull << 63;
template <class TeamBar>
__device__ __forceinline__ bool plain_census(GridBarrier *b, uint32_t epoch, uint64_t wait_ns, TeamBar team_bar,
                                             CensusInfo &ci, uint32_t *count, uint32_t *gen) {
  const uint32_t lane = lane_id();
  const uint32_t p = epoch & 1u;
  ci.xcc = xcc_id();
  uint32_t *cnt = b->plain_census[p];
  uint64_t *word = &b->plain_word[p].word;
  uint32_t go = 0;
  if (lane == 0) {
    __hip_atomic_fetch_add(&cnt[ci.xcc], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    wait_own_stores();  // the per-XCD count is in before the arrival counts
    const uint64_t old = __hip_atomic_fetch_add(word, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!(old & kPlainAbandon)) {
      const uint64_t t0 = now_ns();
      uint64_t w = ld_relaxed(word);
      for (;;) {
        if (w & kPlainAbandon) break;
        if ((w & ~kPlainAbandon) >= gridDim.x) {
          go = 1;
          break;
        }
        if (now_ns() - t0 >= wait_ns) {
          if (__hip_atomic_compare_exchange_strong(word, &w, w | kPlainAbandon, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            // this workgroup abandoned the launch: ready the next one's words
            uint32_t *oc = b->plain_census[p ^ 1u];
            for (int i = 0; i < kMaxXcd; ++i)
              __hip_atomic_store(&oc[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&b->plain_word[p ^ 1u].word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          continue;  // w now holds the word's current value
        }
        __builtin_amdgcn_s_sleep(1);
        w = ld_relaxed(word);
      }
    }
  }
  go = (uint32_t)__shfl((int)go, 0);
  if (!go) return false;
  // the grid is complete, so the per-XCD counts are final
  uint32_t c = 0, g = 0;
  if (lane < (uint32_t)kMaxXcd) c = ld_relaxed(&cnt[lane]);
  else if (lane < 2u * kMaxXcd) g = (uint32_t)(ld_relaxed(&team_bar(lane - kMaxXcd)->gensum.word) >> kCountShift);
  const uint32_t nz = (uint32_t)__ballot(lane < (uint32_t)kMaxXcd && c != 0);
  const uint32_t below = nz & ((1u << (lane & 31u)) - 1u);
  if (lane < (uint32_t)kMaxXcd && c != 0) count[__popc(below)] = c;
  if (lane >= (uint32_t)kMaxXcd && lane < 2u * kMaxXcd) gen[lane - kMaxXcd] = g;
  ci.nx = __popc(nz);
  ci.xrank = __popc(nz & ((1u << ci.xcc) - 1u));
  if (blockIdx.x == 0 && lane == 0) {
    uint32_t *oc = b->plain_census[p ^ 1u];
    for (int i = 0; i < kMaxXcd; ++i) __hip_atomic_store(&oc[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&b->plain_word[p ^ 1u].word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

// Team of XCD ranks [id * xpt, id * xpt + xpt) (thread 0).  `tb` is the
// team's barrier; `count` and `gen` are the census results.  by_ticket:
// ranks from the team's ticket counter (reset by team_ticket_reset after
// the team's first barrier); otherwise the grid is one team and the rank
// is blockIdx.x.
__device__ __forceinline__ void team_form(const CensusInfo &ci, const uint32_t *count, const uint32_t *gen,
                                          uint32_t xpt, GridBarrier *tb, bool by_ticket, Team &t) {
  t.id = ci.xrank / xpt;
  const uint32_t first = t.id * xpt;
  const uint32_t last = min(first + xpt, ci.nx);
  uint32_t size = 0;
  for (uint32_t r = first; r < last; ++r) size += count[r];
  t.size = size;
  t.seg = ci.xrank - first;
  t.teams = (ci.nx + xpt - 1) / xpt;
  t.rank = by_ticket ? (uint32_t)__hip_atomic_fetch_add(&tb->ticket.word, 1ull, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                     : blockIdx.x;
  t.bst.gen = gen[t.id];
  t.bst.xcc = ci.xcc;
  t.bst.xcc_n = count[ci.xrank];
  t.bst.nx = last - first;
}

__device__ __forceinline__ void team_ticket_reset(GridBarrier *tb) {
  __hip_atomic_store(&tb->ticket.word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// All threads of every workgroup must call it.  `payload` and `side` (each
// summing to < 2^54 over the grid) are read from thread 0; returns the
// grid-wide payload sum in every thread and, when side_total is given (by
// every workgroup alike), the side sum there.  `st` is only meaningful in
// thread 0.  `s_word` points at two 64-bit LDS slots.
// kSystem: the XCD leaders write their L2 back at system scope, so every
// store made before the barrier can be read by the host or a DMA copy.
// kInv = false: the workgroup does not invalidate its caches on release;
// the caller must invalidate (l2_invalidate) before it reads, with plain
// loads, data other workgroups wrote.  An invalidate acts on the XCD's L2
// as well as the CU's L1 (a device-scope buffer_inv is the only one that
// drops L1 lines on gfx950), and the workgroups of one XCD invalidating
// together wait for each other: 32 of them take ~2 us.  Levels where few
// workgroups have work skip most of that.
//
// The payload and the side sum travel in two arrival words each carrying
// its own count, added together by every workgroup (one round trip).  Per
// XCD the last arrival on the payload word writes the L2 back and arrives
// at the root's payload word; the last arrival on the side word forwards
// its sum to the root's side word (the two may be different workgroups).
// The last root arrivals publish generation | total in gensum and
// side_out; every workgroup polls both words in one round trip per poll.
template <bool kSystem = false, bool kInv = true>
__device__ __forceinline__ uint64_t grid_barrier(GridBarrier *b, BarrierState &st, uint64_t payload,
                                                 uint64_t *s_word, uint64_t side = 0,
                                                 uint64_t *side_total = nullptr) {
  const bool with_side = side_total != nullptr;
  wait_own_stores();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t next = (st.gen + 1) & kGenMask;
    const uint64_t tag = (uint64_t)next << kCountShift;
    // both arrivals in flight together
    uint64_t os = 0;
    if (with_side)
      os = __hip_atomic_fetch_add(&b->xside[st.xcc].word, kArrive | side, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t old = __hip_atomic_fetch_add(&b->xleaf[st.xcc].word, kArrive | payload,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // last arrivals of this XCD: on the side word it forwards the XCD's side
    // sum; on the payload word it writes the L2 back (every workgroup of the
    // XCD has its stores there) and forwards the payload sum.  One
    // workgroup may be both; its two root arrivals then go out together.
    const bool lead_s = with_side && (os >> kCountShift) + 1 == st.xcc_n;
    const bool lead_p = (old >> kCountShift) + 1 == st.xcc_n;
    const uint64_t xs = (os & kPayloadMask) + side;
    const uint64_t xsum = (old & kPayloadMask) + payload;
    if (lead_s) __hip_atomic_store(&b->xside[st.xcc].word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lead_p) __hip_atomic_store(&b->xleaf[st.xcc].word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the resets land before this XCD can count as arrived at the root
    if (lead_p) {
      if (kSystem) l2_writeback_system();
      else l2_writeback();
    } else if (lead_s) {
      wait_own_stores();
    }
    uint64_t rs = 0, rp = 0;
    if (lead_s)
      rs = __hip_atomic_fetch_add(&b->root_side.word, kArrive | xs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lead_p)
      rp = __hip_atomic_fetch_add(&b->root.word, kArrive | xsum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool fin_s = lead_s && (rs >> kCountShift) + 1 == st.nx;
    const bool fin_p = lead_p && (rp >> kCountShift) + 1 == st.nx;
    if (fin_s) __hip_atomic_store(&b->root_side.word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (fin_p) __hip_atomic_store(&b->root.word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (fin_s || fin_p) wait_own_stores();
    if (fin_s)
      __hip_atomic_store(&b->side_out.word, tag | ((rs & kPayloadMask) + xs), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (fin_p) {
      __hip_atomic_store(&b->gensum.word, tag | ((rp & kPayloadMask) + xsum), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      // generations repeat every 1024 barriers: keep the side word's current
      // too, so that a later barrier with a side sum cannot take a word
      // left over from 1024 barriers without one for its own
      if (!with_side)
        __hip_atomic_store(&b->side_out.word, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // every workgroup waits on the root's words themselves (no relay)
    uint64_t w, ws = tag;
    for (;;) {
      w = ld_relaxed(&b->gensum.word);
      if (with_side) ws = ld_relaxed(&b->side_out.word);
      if ((w >> kCountShift) == next && (ws >> kCountShift) == next) break;
      __builtin_amdgcn_s_sleep(1);
    }
    // this CU's L1 may hold stale lines; on gfx950 only the device-scope
    // invalidate reaches it (workgroup scope leaves L1 alone outside
    // thread-group-split mode), and it also drops this XCD's L2 lines
    if (kInv) l2_invalidate();
    st.gen = next;
    s_word[0] = w & kPayloadMask;
    s_word[1] = ws & kPayloadMask;
  }
  __syncthreads();
  if (with_side) *side_total = uniform(s_word[1]);
  return uniform(s_word[0]);
}

// ---------------------------------------------------------------------------
// Decentralised frontier queue.  The next frontier is split into one segment
// per XCD (capacity n each, so it cannot overflow): a wavefront stages its
// claims in LDS and moves them to its XCD's segment with one atomic on that
// segment's tail per flush; at the end of a phase the workgroup's waves flush
// together with one atomic.  No tail is shared by the whole grid, and the
// consumer of the next level reads just the segment tails to split the
// frontier evenly over all workgroups.
// ---------------------------------------------------------------------------
constexpr int kSegments = 8;
constexpr uint32_t kStage = 1024;  // per-wave LDS staging entries

// Per-thread state of the edge visitor; out_n / claims are wave-uniform.
struct Visitor {
  const uint32_t *__restrict__ col;
  uint32_t *level;
  uint32_t *parent;     // may be null
  uint32_t *vis;
  const uint32_t *snap; // visited bitmap as of the level's start (subset of vis), or vis itself
  bool probe_live;      // snap is a snapshot: probe the live bitmap after it
  bool direct;          // claim by the atomic alone, no bitmap probe before it
  uint32_t *seg;        // output segment of this workgroup's XCD
  uint32_t *seg_tail;   // its tail
  uint32_t *stage;      // this wave's LDS staging buffer [kStage]
  uint32_t out_n;       // staged entries
  uint32_t claims;      // entries appended in the current phase
  uint32_t next_level;
  uint32_t probes, atomics;  // profiling counters (per lane)
  const int64_t *off;        // row offsets, read for claimed vertices when track_mf
  bool track_mf;
  bool deg32;                // every degree < 2^32: the low words of two offsets give it
  uint64_t mf;               // out-degree sum of vertices claimed by this lane
  uint32_t *fb_mark;         // when set, claims also set their bit here (next frontier bitmap)

  // Moves this wave's staged entries to the segment (one atomic).
  __device__ __forceinline__ void flush_wave() {
    const uint32_t cnt = out_n;
    uint32_t base = 0;
    if (lane_id() == 0)
      base = __hip_atomic_fetch_add(seg_tail, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    base = bcast_lane(base, 0);
    wave_sync();
    for (uint32_t k = lane_id(); k < cnt; k += kWave) seg[base + k] = stage[k];
    wave_sync();
    out_n = 0;
  }

  // Appends the lanes flagged in `mask` (all lanes of the wave call it).
  __device__ __forceinline__ void append(bool mine, uint64_t mask, uint32_t w) {
    if (mine) stage[out_n + mbcnt(mask)] = w;
    const uint32_t cnt = (uint32_t)__popcll(mask);
    out_n += cnt;
    claims += cnt;
    if (out_n > kStage - kWave) flush_wave();
  }

  // Batched form: U edges per lane whose column indices are already loaded.
  // All U bitmap probes are issued before any result is used, then the
  // claims (atomic OR for the few that look unvisited) and appends follow,
  // so one wave keeps U probes in flight instead of one.
  //
  // Filter: a probe of `snap`, then (probe_live) a probe of the live
  // bitmap for the neighbours unvisited in `snap`, then the atomic for the
  // neighbours still unvisited.  On large levels `snap` is the level-start
  // snapshot of the visited bitmap (never written during the level, so its
  // lines stay in L1/L2), on other levels the live bitmap, probed once.
  // The loads can only under-report visited vertices, never over-report.
  //
  // kEarlyDeg (with deg32): the low words of the row offsets of the
  // vertices tried (for the out-degree sum) are loaded with the claim
  // attempts, so one wait covers both; otherwise the offsets are loaded
  // for the claimed vertices only, after the attempts return.
  template <int U, bool kEarlyDeg, bool kLive = false>
  __device__ __forceinline__ void visit_batch(const bool (&act)[U], const uint32_t (&u)[U],
                                              const uint32_t (&w)[U]) {
    uint32_t word[U];
    if (direct) {
#pragma unroll
      for (int i = 0; i < U; ++i) word[i] = act[i] ? 0u : 0xffffffffu;
    } else {
#pragma unroll
      for (int i = 0; i < U; ++i) word[i] = act[i] ? snap[w[i] >> 5] : 0xffffffffu;
    }
    if (kLive && probe_live) {
#pragma unroll
      for (int i = 0; i < U; ++i)
        if (!((word[i] >> (w[i] & 31u)) & 1u)) word[i] = vis[w[i] >> 5];
    }
    const bool early = kEarlyDeg && track_mf && deg32;
    uint32_t old[U], lo0[U], lo1[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const uint32_t bit = 1u << (w[i] & 31u);
      old[i] = bit;
      lo0[i] = lo1[i] = 0;
      if (!(word[i] & bit)) {
        ++atomics;
        old[i] = __hip_atomic_fetch_or(&vis[w[i] >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (early) {
          const uint32_t *o = reinterpret_cast<const uint32_t *>(off + w[i]);
          lo0[i] = o[0];
          lo1[i] = o[2];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < U; ++i) {
      bool claimed = false;
      if (act[i]) {
        ++probes;
        claimed = !(old[i] & (1u << (w[i] & 31u)));
        if (claimed) {
          level[w[i]] = next_level;
          if (parent) parent[w[i]] = u[i];
          if (early) mf += (uint64_t)(lo1[i] - lo0[i]);
          else if (track_mf) mf += (uint64_t)(off[w[i] + 1] - off[w[i]]);
          if (fb_mark)
            __hip_atomic_fetch_or(&fb_mark[w[i] >> 5], 1u << (w[i] & 31u), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      const uint64_t mask = __ballot(claimed);
      if (mask) append(claimed, mask, w[i]);
    }
  }
};

}  // namespace dfq
