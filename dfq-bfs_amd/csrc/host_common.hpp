// Host-side helpers shared by the graph utilities: a counter-based RNG whose
// streams do not depend on the thread count, and a small parallel-for over
// std::thread.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace dfq {

// SplitMix64 (Steele, Lea, Flood 2014): used both as a seeding hash and as
// the sequential generator of one stream.
inline uint64_t splitmix64_next(uint64_t &state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t hash64(uint64_t x) {
  uint64_t s = x;
  return splitmix64_next(s);
}

// One independent stream per (seed, stream id): edge i of a generator uses
// stream i, so the output is the same for any partition over threads.
struct Rng {
  uint64_t state;
  Rng(uint64_t seed, uint64_t stream) {
    state = hash64(seed ^ hash64(stream + 0x632BE59BD9B4E019ull));
  }
  uint64_t next() { return splitmix64_next(state); }
  // Uniform integer in [0, bound) by Lemire's multiply-shift with rejection.
  uint64_t below(uint64_t bound) {
    if (bound <= 1) return 0;
    for (;;) {
      uint64_t x = next();
      __uint128_t p = (__uint128_t)x * bound;
      uint64_t lo = (uint64_t)p;
      if (lo >= bound) return (uint64_t)(p >> 64);
      uint64_t thresh = (0 - bound) % bound;
      if (lo >= thresh) return (uint64_t)(p >> 64);
    }
  }
  // Uniform double in [0, 1) with 53 random bits.
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline unsigned host_threads() {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 1;
  return std::min(t, 64u);
}

// Runs fn(begin, end, tid) over [0, n) split into contiguous chunks.
template <class F>
void parallel_for(uint64_t n, F fn, unsigned threads = 0) {
  if (threads == 0) threads = host_threads();
  if (n < 4096 || threads == 1) {
    fn((uint64_t)0, n, 0u);
    return;
  }
  threads = (unsigned)std::min<uint64_t>(threads, n);
  std::vector<std::thread> pool;
  pool.reserve(threads);
  uint64_t chunk = (n + threads - 1) / threads;
  for (unsigned t = 0; t < threads; ++t) {
    uint64_t b = (uint64_t)t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([=, &fn] { fn(b, e, t); });
  }
  for (auto &th : pool) th.join();
}

// CSR input check shared by the device front ends: offsets start at 0, end
// at m and never decrease; every column index is a vertex.
inline bool csr_valid(const int64_t *off, const uint32_t *col, uint32_t n, uint64_t m) {
  if (off[0] != 0 || (uint64_t)off[n] != m) return false;
  std::atomic<bool> bad{false};
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t v = b; v < e; ++v)
      if (off[v + 1] < off[v]) {
        bad = true;
        return;
      }
  });
  if (bad) return false;
  parallel_for(m, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t i = b; i < e; ++i)
      if (col[i] >= n) {
        bad = true;
        return;
      }
  });
  return !bad;
}

// Bottom-up levels scan a vertex's out-edges as its in-edges, which is only
// valid when every u->v has a v->u.  Requires sorted rows (binary search);
// a graph with unsorted rows is treated as directed.
inline bool is_symmetric(const int64_t *off, const uint32_t *col, uint32_t n) {
  std::atomic<bool> ok{true};
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t u = b; u < e && ok.load(std::memory_order_relaxed); ++u)
      for (int64_t k = off[u] + 1; k < off[u + 1]; ++k)
        if (col[k - 1] > col[k]) {
          ok = false;
          return;
        }
  });
  if (!ok) return false;
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t u = b; u < e && ok.load(std::memory_order_relaxed); ++u)
      for (int64_t k = off[u]; k < off[u + 1]; ++k) {
        const uint32_t v = col[k];
        if (!std::binary_search(col + off[v], col + off[v + 1], (uint32_t)u)) {
          ok = false;
          return;
        }
      }
  });
  return ok;
}

// Bottom-up levels stop at a vertex's first neighbour in the frontier, and
// high-degree vertices join the frontier first, so rows are stored with
// their neighbours by decreasing degree (ties by id): the same neighbour
// set, found sooner (degree-ordered adjacency lists for bottom-up search as
// in Yasui, Fujisawa and Goto, IEEE BigData 2013).  out has m entries.
inline void order_rows_by_degree(const int64_t *off, const uint32_t *col, uint32_t n, uint32_t *out) {
  std::vector<uint32_t> deg(n);
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t v = b; v < e; ++v)
      deg[v] = (uint32_t)std::min<int64_t>(off[v + 1] - off[v], 0xffffffffll);
  });
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    std::vector<uint64_t> key;
    for (uint64_t u = b; u < e; ++u) {
      const int64_t s = off[u], t = off[u + 1];
      if (t - s < 2) {
        if (t > s) out[s] = col[s];
        continue;
      }
      key.resize((size_t)(t - s));
      for (int64_t k = s; k < t; ++k) {
        const uint32_t w = col[k];
        key[(size_t)(k - s)] = ((uint64_t)(0xffffffffu - deg[w]) << 32) | w;
      }
      std::sort(key.begin(), key.end());
      for (int64_t k = s; k < t; ++k) out[k] = (uint32_t)key[(size_t)(k - s)];
    }
  });
}

// Degree-ranked copy of a CSR for multi-source searches: vertex v gets rank
// o2r[v] (decreasing degree, ties by id; r2o is the inverse), row r of
// (off_r, col_r) is row r2o[r] with its neighbours given as ranks, sorted
// ascending, i.e. highest degree first.  The frontier masks of the
// highest-degree vertices - the tails of most edges in a skewed graph - then
// sit together at the front of every per-vertex array, where they stay in
// cache; and a row read in order meets its hubs first.
inline void rank_relabel(const int64_t *off, const uint32_t *col, uint32_t n, uint32_t *o2r, uint32_t *r2o,
                         int64_t *off_r, uint32_t *col_r) {
  std::vector<uint64_t> key(n);
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t v = b; v < e; ++v) {
      const uint64_t d = (uint64_t)std::min<int64_t>(off[v + 1] - off[v], 0xffffffffll);
      key[v] = ((0xffffffffull - d) << 32) | v;
    }
  });
  std::sort(key.begin(), key.end());
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t r = b; r < e; ++r) {
      const uint32_t v = (uint32_t)key[r];
      r2o[r] = v;
      o2r[v] = (uint32_t)r;
    }
  });
  off_r[0] = 0;
  for (uint32_t r = 0; r < n; ++r) off_r[r + 1] = off_r[r] + (off[r2o[r] + 1] - off[r2o[r]]);
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t r = b; r < e; ++r) {
      const uint32_t v = r2o[r];
      uint32_t *dst = col_r + off_r[r];
      const int64_t len = off[v + 1] - off[v];
      for (int64_t k = 0; k < len; ++k) dst[k] = o2r[col[off[v] + k]];
      std::sort(dst, dst + len);
    }
  });
}

// First four entries of every row of `ordered` (rows by neighbour degree),
// four words per vertex; words past a row's end are 0 (never tested).
inline void row_heads(const int64_t *off, const uint32_t *ordered, uint32_t n, uint32_t *out) {
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t v = b; v < e; ++v) {
      const int64_t s = off[v], t = off[v + 1];
      for (int i = 0; i < 4; ++i) out[4 * v + i] = s + i < t ? ordered[s + i] : 0u;
    }
  });
}

// Bit v set for every vertex without out-entries.  In a symmetric graph such
// a vertex has no in-entries either, so no search reaches it: searches may
// start with these bits already in the visited bitmap, and bottom-up levels
// then never look at them.
inline void isolated_bitmap(const int64_t *off, uint32_t n, uint32_t *words) {
  const uint32_t nw = (n + 31) / 32;
  parallel_for(nw, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t w = b; w < e; ++w) {
      uint32_t bits = 0;
      for (uint32_t i = 0; i < 32; ++i) {
        const uint64_t v = w * 32 + i;
        if (v < n && off[v + 1] == off[v]) bits |= 1u << i;
      }
      words[w] = bits;
    }
  });
}

}  // namespace dfq
