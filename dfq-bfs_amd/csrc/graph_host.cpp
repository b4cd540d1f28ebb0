// Host graph utilities: CSR construction, file formats, seeded generators,
// source selection and the BFS certificate check.  C ABI in dfq_graph.h.
//
// Reference behaviour kept:
//   * SNAP header contract "# Nodes: N Edges: M" / "# FromNodeId ToNodeId"
//     (README.md:19-32).
//   * ".dat" CSR text "N M SOURCE", N x "start length", M x "dst weight"
//     with the weight ignored (design/shared-q/atomic-sa/our-sep/upper-bound/
//     bfs.cu:188,198-199).
//   * report fields Dist / Touch / Error (bfs.cu:290-300) via
//     dfq_compare_levels and dfq_validate_bfs.

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "dfq_graph.h"
#include "host_common.hpp"

using namespace dfq;

namespace {

// Largest CSR entry count the library accepts anywhere (the device engine's
// barrier sums are 40-bit fields); also keeps every byte count below 2^64.
constexpr uint64_t kHostMaxEntries = (1ull << 40) - 1;

dfq_csr *csr_alloc(uint32_t n, uint64_t m) {
  if (m > kHostMaxEntries || n == 0xffffffffu) return nullptr;
  dfq_csr *g = (dfq_csr *)calloc(1, sizeof(dfq_csr));
  if (!g) return nullptr;
  g->n = n;
  g->m = m;
  g->row_offsets = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
  g->col_idx = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  if (!g->row_offsets || !g->col_idx) {
    free(g->row_offsets);
    free(g->col_idx);
    free(g);
    return nullptr;
  }
  return g;
}

// Seeded Fisher-Yates permutation of [0, n).
std::vector<uint32_t> make_permutation(uint32_t n, uint64_t seed) {
  std::vector<uint32_t> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = i;
  Rng r(seed, 0x5045524D55544532ull);
  for (uint32_t i = n; i > 1; --i) {
    uint32_t j = (uint32_t)r.below(i);
    std::swap(p[i - 1], p[j]);
  }
  return p;
}

// Core CSR builder from materialised src/dst arrays.
int build_csr(const uint32_t *src, const uint32_t *dst, uint64_t ne, uint32_t n,
              uint32_t flags, uint64_t perm_seed, dfq_csr **out) {
  const bool sym = flags & DFQ_CSR_SYMMETRIZE;
  const bool dedup = flags & DFQ_CSR_DEDUP;
  const bool drop_self = flags & DFQ_CSR_DROP_SELF;
  const bool permute = flags & DFQ_CSR_PERMUTE;

  std::atomic<bool> bad{false};
  parallel_for(ne, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t i = b; i < e; ++i)
      if (src[i] >= n || dst[i] >= n) { bad = true; return; }
  });
  if (bad) return DFQ_ERR_RANGE;

  std::vector<uint32_t> perm;
  if (permute) perm = make_permutation(n, perm_seed);
  auto map = [&](uint32_t x) -> uint32_t { return permute ? perm[x] : x; };

  std::vector<uint32_t> deg((size_t)n + 1, 0);
  parallel_for(ne, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t i = b; i < e; ++i) {
      uint32_t u = src[i], v = dst[i];
      if (drop_self && u == v) continue;
      __atomic_fetch_add(&deg[map(u)], 1u, __ATOMIC_RELAXED);
      if (sym) __atomic_fetch_add(&deg[map(v)], 1u, __ATOMIC_RELAXED);
    }
  });
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) off[v + 1] = off[v] + deg[v];
  const uint64_t m0 = (uint64_t)off[n];
  std::vector<uint32_t> cols;
  try {
    cols.resize(m0 ? m0 : 1);
  } catch (const std::bad_alloc &) {
    return DFQ_ERR_NOMEM;
  }
  std::vector<int64_t> cur(off.begin(), off.end() - 1);
  parallel_for(ne, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t i = b; i < e; ++i) {
      uint32_t u = src[i], v = dst[i];
      if (drop_self && u == v) continue;
      uint32_t pu = map(u), pv = map(v);
      int64_t s = __atomic_fetch_add(&cur[pu], 1, __ATOMIC_RELAXED);
      cols[s] = pv;
      if (sym) {
        int64_t t = __atomic_fetch_add(&cur[pv], 1, __ATOMIC_RELAXED);
        cols[t] = pu;
      }
    }
  });
  // Sort each row (makes the result independent of fill order) and, with
  // DEDUP, count the distinct entries.
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t v = b; v < e; ++v) {
      uint32_t *rb = cols.data() + off[v], *re = cols.data() + off[v + 1];
      std::sort(rb, re);
      if (dedup) deg[v] = (uint32_t)(std::unique(rb, re) - rb);
    }
  });
  uint64_t m = m0;
  if (dedup) {
    m = 0;
    for (uint32_t v = 0; v < n; ++v) m += deg[v];
  }
  dfq_csr *g = csr_alloc(n, m);
  if (!g) return DFQ_ERR_NOMEM;
  if (!dedup) {
    memcpy(g->row_offsets, off.data(), sizeof(int64_t) * ((size_t)n + 1));
    if (m) memcpy(g->col_idx, cols.data(), sizeof(uint32_t) * m);
  } else {
    g->row_offsets[0] = 0;
    for (uint32_t v = 0; v < n; ++v) g->row_offsets[v + 1] = g->row_offsets[v] + deg[v];
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t v = b; v < e; ++v)
        memcpy(g->col_idx + g->row_offsets[v], cols.data() + off[v], sizeof(uint32_t) * deg[v]);
    });
  }
  *out = g;
  return DFQ_OK;
}

// ---- text parsing helpers -------------------------------------------------

struct MappedFile {
  const char *data = nullptr;
  size_t size = 0;
  int fd = -1;
  bool open(const char *path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    size = (size_t)st.st_size;
    if (size == 0) {
      data = "";
      return true;
    }
    void *p = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) return false;
    data = (const char *)p;
    return true;
  }
  ~MappedFile() {
    if (data && size) munmap((void *)data, size);
    if (fd >= 0) ::close(fd);
  }
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

// Parses an unsigned decimal at p (after optional blanks on the same line).
// Returns false when no digit is found before the line end / buffer end.
// Accumulates the digits at p into v; false when the number does not fit
// in 64 bits (the digits are consumed either way).
inline bool read_digits(const char *&p, const char *end, uint64_t &v) {
  bool fits = true;
  v = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    const uint64_t d = (uint64_t)(*p - '0');
    if (v > (UINT64_MAX - d) / 10) fits = false;
    else v = v * 10 + d;
    ++p;
  }
  return fits;
}

inline bool parse_u64(const char *&p, const char *end, uint64_t &val) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  if (p >= end || *p < '0' || *p > '9') return false;
  uint64_t v = 0;
  if (!read_digits(p, end, v)) return false;
  val = v;
  return true;
}

// Same, skipping any whitespace including newlines (token stream formats).
inline bool next_u64(const char *&p, const char *end, uint64_t &val, bool &negative) {
  while (p < end && is_space(*p)) ++p;
  negative = false;
  if (p < end && *p == '-') {
    negative = true;
    ++p;
  }
  if (p >= end || *p < '0' || *p > '9') return false;
  uint64_t v = 0;
  if (!read_digits(p, end, v)) return false;
  if (p < end && !is_space(*p)) return false;
  val = v;
  return true;
}

}  // namespace

extern "C" {

const char *dfq_strerror(int status) {
  switch (status) {
    case DFQ_OK: return "ok";
    case DFQ_ERR_INVALID_ARG: return "invalid argument";
    case DFQ_ERR_IO: return "file I/O error";
    case DFQ_ERR_FORMAT: return "malformed input";
    case DFQ_ERR_RANGE: return "value out of range";
    case DFQ_ERR_NOMEM: return "out of host memory";
    case DFQ_ERR_HIP: return "HIP runtime error";
    case DFQ_ERR_NO_DEVICE: return "no usable GPU";
    case DFQ_ERR_COOPERATIVE: return "cooperative launch not possible";
    case DFQ_ERR_CAPACITY: return "capacity exceeded";
    case DFQ_ERR_BUSY: return "busy: admission refused";
    case DFQ_ERR_STATE: return "invalid state";
    case DFQ_ERR_PENDING: return "query still running";
    case DFQ_ERR_TIMEOUT: return "timed out";
    default: return "unknown error";
  }
}

void dfq_csr_free(dfq_csr *g) {
  if (!g) return;
  free(g->row_offsets);
  free(g->col_idx);
  free(g);
}

int dfq_csr_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t ne, uint32_t n,
                       uint32_t flags, uint64_t perm_seed, dfq_csr **out) {
  if (!out || (ne && (!src || !dst))) return DFQ_ERR_INVALID_ARG;
  if (flags & ~0xFu) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  try {
    return build_csr(src, dst, ne, n, flags, perm_seed, out);
  } catch (const std::bad_alloc &) {
    return DFQ_ERR_NOMEM;
  }
}

int dfq_csr_from_arrays(const int64_t *row_offsets, const uint32_t *col_idx, uint32_t n,
                        uint64_t m, dfq_csr **out) {
  if (!out || !row_offsets || (m && !col_idx)) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  if (m > kHostMaxEntries || n == 0xffffffffu) return DFQ_ERR_RANGE;
  if (row_offsets[0] != 0 || (uint64_t)row_offsets[n] != m) return DFQ_ERR_RANGE;
  for (uint32_t v = 0; v < n; ++v)
    if (row_offsets[v + 1] < row_offsets[v]) return DFQ_ERR_RANGE;
  for (uint64_t i = 0; i < m; ++i)
    if (col_idx[i] >= n) return DFQ_ERR_RANGE;
  dfq_csr *g = csr_alloc(n, m);
  if (!g) return DFQ_ERR_NOMEM;
  memcpy(g->row_offsets, row_offsets, sizeof(int64_t) * ((size_t)n + 1));
  if (m) memcpy(g->col_idx, col_idx, sizeof(uint32_t) * m);
  *out = g;
  return DFQ_OK;
}

// ---- SNAP edge list -----------------------------------------------------------

namespace {
int load_snap(const char *path, uint32_t flags, dfq_csr **out);
}

int dfq_load_snap(const char *path, uint32_t flags, dfq_csr **out) {
  if (!path || !out) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  try {
    return load_snap(path, flags, out);
  } catch (const std::bad_alloc &) {
    return DFQ_ERR_NOMEM;
  }
}

}  // extern "C"

namespace {
int load_snap(const char *path, uint32_t flags, dfq_csr **out) {
  MappedFile f;
  if (!f.open(path)) return DFQ_ERR_IO;
  const char *p = f.data, *end = f.data + f.size;
  int64_t n_header = -1;
  std::vector<uint32_t> src, dst;
  uint64_t max_id = 0;
  bool any = false;
  while (p < end) {
    const char *le = (const char *)memchr(p, '\n', (size_t)(end - p));
    if (!le) le = end;
    const char *q = p;
    while (q < le && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
    if (q < le && *q == '#') {
      // "# Nodes: N Edges: M" (README.md:19-32); any other comment is skipped.
      std::string line(q, le);
      const char *s = strstr(line.c_str(), "Nodes:");
      if (s) {
        s += 6;
        const char *se = line.c_str() + line.size();
        uint64_t nv;
        if (!parse_u64(s, se, nv) || nv > 0xFFFFFFFEull) return DFQ_ERR_FORMAT;
        n_header = (int64_t)nv;
      }
    } else if (q < le) {
      uint64_t u, v;
      const char *r = q;
      if (!parse_u64(r, le, u) || !parse_u64(r, le, v)) return DFQ_ERR_FORMAT;
      while (r < le && (*r == ' ' || *r == '\t' || *r == '\r')) ++r;
      // A third column (weight) is tolerated and ignored.
      if (r < le) {
        uint64_t w;
        bool neg = false;
        if (*r == '-') { neg = true; ++r; }
        (void)neg;
        if (!parse_u64(r, le, w)) return DFQ_ERR_FORMAT;
      }
      if (u > 0xFFFFFFFEull || v > 0xFFFFFFFEull) return DFQ_ERR_RANGE;
      src.push_back((uint32_t)u);
      dst.push_back((uint32_t)v);
      max_id = std::max(max_id, std::max(u, v));
      any = true;
    }
    p = le + 1;
  }
  uint32_t n;
  if (n_header >= 0) {
    n = (uint32_t)n_header;
    if (any && max_id >= (uint64_t)n) return DFQ_ERR_RANGE;
  } else {
    n = any ? (uint32_t)(max_id + 1) : 0;
  }
  return dfq_csr_from_edges(src.data(), dst.data(), src.size(), n, flags, 0, out);
}
}  // namespace

extern "C" {

int dfq_save_snap(const char *path, const dfq_csr *g) {
  if (!path || !g) return DFQ_ERR_INVALID_ARG;
  FILE *fp = fopen(path, "w");
  if (!fp) return DFQ_ERR_IO;
  std::vector<char> buf(1 << 20);
  setvbuf(fp, buf.data(), _IOFBF, buf.size());
  fprintf(fp, "# Nodes: %u Edges: %llu\n", g->n, (unsigned long long)g->m);
  fprintf(fp, "# FromNodeId\tToNodeId\n");
  for (uint32_t u = 0; u < g->n; ++u)
    for (int64_t k = g->row_offsets[u]; k < g->row_offsets[u + 1]; ++k)
      fprintf(fp, "%u\t%u\n", u, g->col_idx[k]);
  int rc = ferror(fp) ? DFQ_ERR_IO : DFQ_OK;
  if (fclose(fp) != 0) rc = DFQ_ERR_IO;
  return rc;
}

// ---- reference .dat CSR text ---------------------------------------------------

int dfq_load_dat(const char *path, dfq_csr **out, uint32_t *source_out) {
  if (!path || !out) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  MappedFile f;
  if (!f.open(path)) return DFQ_ERR_IO;
  const char *p = f.data, *end = f.data + f.size;
  uint64_t n, m, s;
  bool neg = false;
  if (!next_u64(p, end, n, neg) || neg) return DFQ_ERR_FORMAT;
  if (!next_u64(p, end, m, neg) || neg) return DFQ_ERR_FORMAT;
  if (!next_u64(p, end, s, neg)) return DFQ_ERR_FORMAT;
  if (n > 0xFFFFFFFEull || m > kHostMaxEntries) return DFQ_ERR_RANGE;
  if (neg || (n > 0 && s >= n) || (n == 0 && s != 0)) return DFQ_ERR_RANGE;
  // 2n + 2m more tokens follow, each at least one digit and one separator:
  // a header that promises more than the file can hold is refused before
  // anything is allocated for it
  const uint64_t room = (uint64_t)(end - p) / 2 + 1;
  if (n > room || m > room || n + m > room) return DFQ_ERR_FORMAT;
  try {
    std::vector<uint64_t> start(n), len(n);
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) {
      if (!next_u64(p, end, start[i], neg) || neg) return DFQ_ERR_FORMAT;
      if (!next_u64(p, end, len[i], neg) || neg) return DFQ_ERR_FORMAT;
      // row [start, start + len) inside edge[0, m), written so it cannot wrap
      if (len[i] > m || start[i] > m - len[i]) return DFQ_ERR_RANGE;
      // each step adds at most m < 2^40 to a total below 2^40: no wrap
      total += len[i];
      if (total > kHostMaxEntries) return DFQ_ERR_RANGE;
    }
    std::vector<uint32_t> edge(m);
    for (uint64_t i = 0; i < m; ++i) {
      uint64_t d, w;
      if (!next_u64(p, end, d, neg) || neg) return DFQ_ERR_FORMAT;
      if (!next_u64(p, end, w, neg)) return DFQ_ERR_FORMAT;
      if (d >= n) return DFQ_ERR_RANGE;
      edge[i] = (uint32_t)d;
    }
    // Rows may overlap or leave gaps in the reference layout; copy them out
    // into a contiguous CSR in vertex order.
    dfq_csr *g = csr_alloc((uint32_t)n, total);
    if (!g) return DFQ_ERR_NOMEM;
    g->row_offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) {
      g->row_offsets[i + 1] = g->row_offsets[i] + (int64_t)len[i];
      if (len[i]) memcpy(g->col_idx + g->row_offsets[i], edge.data() + start[i], sizeof(uint32_t) * len[i]);
    }
    *out = g;
    if (source_out) *source_out = (uint32_t)s;
    return DFQ_OK;
  } catch (const std::bad_alloc &) {
    return DFQ_ERR_NOMEM;
  }
}

int dfq_save_dat(const char *path, const dfq_csr *g, uint32_t source) {
  if (!path || !g) return DFQ_ERR_INVALID_ARG;
  if (g->n > 0 && source >= g->n) return DFQ_ERR_RANGE;
  FILE *fp = fopen(path, "w");
  if (!fp) return DFQ_ERR_IO;
  std::vector<char> buf(1 << 20);
  setvbuf(fp, buf.data(), _IOFBF, buf.size());
  fprintf(fp, "%u %llu %u\n", g->n, (unsigned long long)g->m, source);
  for (uint32_t v = 0; v < g->n; ++v)
    fprintf(fp, "%lld %lld\n", (long long)g->row_offsets[v],
            (long long)(g->row_offsets[v + 1] - g->row_offsets[v]));
  for (uint64_t i = 0; i < g->m; ++i) fprintf(fp, "%u 1\n", g->col_idx[i]);
  int rc = ferror(fp) ? DFQ_ERR_IO : DFQ_OK;
  if (fclose(fp) != 0) rc = DFQ_ERR_IO;
  return rc;
}

// ---- binary CSR cache -------------------------------------------------------------

static const char kCsrMagic[8] = {'D', 'F', 'Q', 'C', 'S', 'R', '0', '1'};

int dfq_save_csr_bin(const char *path, const dfq_csr *g) {
  if (!path || !g) return DFQ_ERR_INVALID_ARG;
  FILE *fp = fopen(path, "wb");
  if (!fp) return DFQ_ERR_IO;
  bool ok = fwrite(kCsrMagic, 1, 8, fp) == 8 && fwrite(&g->n, sizeof(uint32_t), 1, fp) == 1 &&
            fwrite(&g->m, sizeof(uint64_t), 1, fp) == 1 &&
            fwrite(g->row_offsets, sizeof(int64_t), (size_t)g->n + 1, fp) == (size_t)g->n + 1 &&
            (g->m == 0 || fwrite(g->col_idx, sizeof(uint32_t), g->m, fp) == g->m);
  if (fclose(fp) != 0) ok = false;
  return ok ? DFQ_OK : DFQ_ERR_IO;
}

int dfq_load_csr_bin(const char *path, dfq_csr **out) {
  if (!path || !out) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  FILE *fp = fopen(path, "rb");
  if (!fp) return DFQ_ERR_IO;
  char magic[8];
  uint32_t n = 0;
  uint64_t m = 0;
  if (fread(magic, 1, 8, fp) != 8 || memcmp(magic, kCsrMagic, 8) != 0 ||
      fread(&n, sizeof(uint32_t), 1, fp) != 1 || fread(&m, sizeof(uint64_t), 1, fp) != 1) {
    fclose(fp);
    return DFQ_ERR_FORMAT;
  }
  // the header's sizes must match the file's before anything is allocated
  struct stat st;
  const uint64_t head = 8 + sizeof(uint32_t) + sizeof(uint64_t);
  if (fstat(fileno(fp), &st) != 0) {
    fclose(fp);
    return DFQ_ERR_IO;
  }
  if (n == 0xffffffffu || m > kHostMaxEntries ||
      (uint64_t)st.st_size != head + sizeof(int64_t) * ((uint64_t)n + 1) + sizeof(uint32_t) * m) {
    fclose(fp);
    return DFQ_ERR_FORMAT;
  }
  dfq_csr *g = csr_alloc(n, m);
  if (!g) {
    fclose(fp);
    return DFQ_ERR_NOMEM;
  }
  bool ok = fread(g->row_offsets, sizeof(int64_t), (size_t)n + 1, fp) == (size_t)n + 1 &&
            (m == 0 || fread(g->col_idx, sizeof(uint32_t), m, fp) == m);
  fclose(fp);
  if (!ok) {
    dfq_csr_free(g);
    return DFQ_ERR_FORMAT;
  }
  if (g->row_offsets[0] != 0 || (uint64_t)g->row_offsets[n] != m) {
    dfq_csr_free(g);
    return DFQ_ERR_FORMAT;
  }
  for (uint32_t v = 0; v < n; ++v)
    if (g->row_offsets[v + 1] < g->row_offsets[v]) {
      dfq_csr_free(g);
      return DFQ_ERR_FORMAT;
    }
  for (uint64_t i = 0; i < m; ++i)
    if (g->col_idx[i] >= n) {
      dfq_csr_free(g);
      return DFQ_ERR_FORMAT;
    }
  *out = g;
  return DFQ_OK;
}

// ---- generators -----------------------------------------------------------------

int dfq_gen_amazon_like(uint32_t n, uint64_t seed, dfq_csr **out) {
  if (!out || n < 7) return DFQ_ERR_INVALID_ARG;  // 5 distinct non-self targets need n >= 6
  *out = nullptr;
  try {
    std::vector<uint32_t> src, dst;
    src.reserve((size_t)n * 3);
    dst.reserve((size_t)n * 3);
    Rng r(seed, 0x414D415A4F4E4C4Bull);
    const int64_t W = 2000;
    uint32_t picked[5];
    for (uint32_t v = 0; v < n; ++v) {
      uint32_t d = 1 + (uint32_t)r.below(5);
      uint32_t k = 0;
      while (k < d) {
        uint32_t t;
        if (r.uniform() < 0.7) {
          int64_t lo = std::max<int64_t>(0, (int64_t)v - W);
          int64_t hi = std::min<int64_t>((int64_t)n - 1, (int64_t)v + W);
          t = (uint32_t)(lo + (int64_t)r.below((uint64_t)(hi - lo + 1)));
        } else {
          t = (uint32_t)r.below(n);
        }
        if (t == v) continue;
        bool dup = false;
        for (uint32_t j = 0; j < k; ++j) dup |= picked[j] == t;
        if (dup) continue;
        picked[k++] = t;
      }
      for (uint32_t j = 0; j < d; ++j) {
        src.push_back(v);
        dst.push_back(picked[j]);
      }
    }
    return build_csr(src.data(), dst.data(), src.size(), n, 0, 0, out);
  } catch (const std::bad_alloc &) {
    return DFQ_ERR_NOMEM;
  }
}

int dfq_gen_rmat(uint32_t scale, uint32_t edgefactor, double a, double b, double c, uint64_t seed,
                 uint32_t flags, dfq_csr **out) {
  if (!out || scale == 0 || scale > 31 || edgefactor == 0) return DFQ_ERR_INVALID_ARG;
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return DFQ_ERR_INVALID_ARG;
  if (flags & ~0xFu) return DFQ_ERR_INVALID_ARG;
  *out = nullptr;
  const uint32_t n = (uint32_t)(1ull << scale) - (scale == 32 ? 1 : 0);
  const uint64_t ne = (uint64_t)edgefactor << scale;
  try {
    std::vector<uint32_t> src(ne), dst(ne);
    // Quadrant thresholds on 32-bit draws.
    const uint64_t ta = (uint64_t)std::llround(a * 4294967296.0);
    const uint64_t tb = (uint64_t)std::llround((a + b) * 4294967296.0);
    const uint64_t tc = (uint64_t)std::llround((a + b + c) * 4294967296.0);
    parallel_for(ne, [&](uint64_t lo, uint64_t hi, unsigned) {
      for (uint64_t i = lo; i < hi; ++i) {
        Rng r(seed, i);
        uint32_t u = 0, v = 0;
        uint64_t bits = 0;
        for (uint32_t l = 0; l < scale; ++l) {
          if ((l & 1) == 0) bits = r.next();
          uint64_t x = (l & 1) ? (bits >> 32) : (bits & 0xFFFFFFFFull);
          uint32_t bu = 0, bv = 0;
          if (x < ta) {
          } else if (x < tb) {
            bv = 1;
          } else if (x < tc) {
            bu = 1;
          } else {
            bu = 1;
            bv = 1;
          }
          u = (u << 1) | bu;
          v = (v << 1) | bv;
        }
        src[i] = u;
        dst[i] = v;
      }
    });
    return build_csr(src.data(), dst.data(), ne, n, flags, seed ^ 0x9E3779B97F4A7C15ull, out);
  } catch (const std::bad_alloc &) {
    return DFQ_ERR_NOMEM;
  }
}

int dfq_gen_grid2d(uint32_t width, uint32_t height, double drop_p, uint64_t seed, dfq_csr **out) {
  if (!out || width == 0 || height == 0 || !(drop_p >= 0.0 && drop_p <= 1.0)) return DFQ_ERR_INVALID_ARG;
  if ((uint64_t)width * height > 0xFFFFFFFEull) return DFQ_ERR_RANGE;
  *out = nullptr;
  const uint32_t n = width * height;
  const uint64_t nh = (uint64_t)(width - 1) * height;  // horizontal edges
  const uint64_t thr = drop_p < 1.0 ? (uint64_t)std::ldexp(drop_p, 64) : ~0ull;
  const bool drop_all = drop_p >= 1.0;
  // Edge id: horizontal (x,y)-(x+1,y) -> y*(width-1)+x; vertical (x,y)-(x,y+1)
  // -> nh + y*width + x.  Kept iff hash(seed, id) >= drop threshold.
  auto kept = [&](uint64_t id) -> bool {
    if (drop_all) return false;
    if (drop_p <= 0.0) return true;
    return hash64(seed * 0xD1B54A32D192ED03ull + id + 1) >= thr;
  };
  try {
    dfq_csr *g = nullptr;
    std::vector<uint8_t> deg(n, 0);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t v = b; v < e; ++v) {
        uint32_t x = (uint32_t)(v % width), y = (uint32_t)(v / width);
        uint8_t d = 0;
        if (y > 0 && kept(nh + (uint64_t)(y - 1) * width + x)) ++d;
        if (x > 0 && kept((uint64_t)y * (width - 1) + x - 1)) ++d;
        if (x + 1 < width && kept((uint64_t)y * (width - 1) + x)) ++d;
        if (y + 1 < height && kept(nh + (uint64_t)y * width + x)) ++d;
        deg[v] = d;
      }
    });
    uint64_t m = 0;
    for (uint32_t v = 0; v < n; ++v) m += deg[v];
    g = csr_alloc(n, m);
    if (!g) return DFQ_ERR_NOMEM;
    g->row_offsets[0] = 0;
    for (uint32_t v = 0; v < n; ++v) g->row_offsets[v + 1] = g->row_offsets[v] + deg[v];
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t v = b; v < e; ++v) {
        uint32_t x = (uint32_t)(v % width), y = (uint32_t)(v / width);
        uint32_t *c = g->col_idx + g->row_offsets[v];
        // ascending neighbour order: up, left, right, down
        if (y > 0 && kept(nh + (uint64_t)(y - 1) * width + x)) *c++ = (uint32_t)(v - width);
        if (x > 0 && kept((uint64_t)y * (width - 1) + x - 1)) *c++ = (uint32_t)(v - 1);
        if (x + 1 < width && kept((uint64_t)y * (width - 1) + x)) *c++ = (uint32_t)(v + 1);
        if (y + 1 < height && kept(nh + (uint64_t)y * width + x)) *c++ = (uint32_t)(v + width);
      }
    });
    *out = g;
    return DFQ_OK;
  } catch (const std::bad_alloc &) {
    return DFQ_ERR_NOMEM;
  }
}

// ---- sources, work and validation ------------------------------------------------

int dfq_pick_sources(const dfq_csr *g, uint32_t count, uint64_t seed, uint32_t *out) {
  if (!g || (count && !out)) return DFQ_ERR_INVALID_ARG;
  uint64_t nonzero = 0;
  for (uint32_t v = 0; v < g->n; ++v) nonzero += g->row_offsets[v + 1] > g->row_offsets[v];
  if (count && nonzero == 0) return DFQ_ERR_RANGE;
  const bool distinct = count <= nonzero;
  Rng r(seed, 0x534F555243455321ull);
  for (uint32_t i = 0; i < count; ++i) {
    for (;;) {
      uint32_t v = (uint32_t)r.below(g->n);
      if (g->row_offsets[v + 1] == g->row_offsets[v]) continue;
      bool dup = false;
      if (distinct)
        for (uint32_t j = 0; j < i; ++j) dup |= out[j] == v;
      if (dup) continue;
      out[i] = v;
      break;
    }
  }
  return DFQ_OK;
}

int dfq_reached_work(const dfq_csr *g, const uint32_t *level, uint64_t *edges_out,
                     uint64_t *vertices_out, uint32_t *depth_out) {
  if (!g || !level) return DFQ_ERR_INVALID_ARG;
  std::vector<uint64_t> pe(host_threads(), 0), pv(host_threads(), 0);
  std::vector<uint32_t> pd(host_threads(), 0);
  parallel_for(g->n, [&](uint64_t b, uint64_t e, unsigned t) {
    uint64_t se = 0, sv = 0;
    uint32_t d = 0;
    for (uint64_t v = b; v < e; ++v)
      if (level[v] != DFQ_UNREACHED) {
        se += (uint64_t)(g->row_offsets[v + 1] - g->row_offsets[v]);
        ++sv;
        d = std::max(d, level[v] + 1);
      }
    pe[t] = se;
    pv[t] = sv;
    pd[t] = d;
  });
  uint64_t se = 0, sv = 0;
  uint32_t d = 0;
  for (size_t t = 0; t < pe.size(); ++t) {
    se += pe[t];
    sv += pv[t];
    d = std::max(d, pd[t]);
  }
  if (edges_out) *edges_out = se;
  if (vertices_out) *vertices_out = sv;
  if (depth_out) *depth_out = d;
  return DFQ_OK;
}

int dfq_validate_bfs(const dfq_csr *g, uint32_t source, const uint32_t *level, const uint32_t *parent,
                     dfq_validate_report *report) {
  if (!g || !level || !report) return DFQ_ERR_INVALID_ARG;
  if (source >= g->n) return DFQ_ERR_RANGE;
  memset(report, 0, sizeof(*report));
  const uint32_t n = g->n;
  const int64_t *off = g->row_offsets;
  const uint32_t *col = g->col_idx;
  unsigned T = host_threads();
  std::vector<uint64_t> span(T, 0), missed(T, 0), reached(T, 0), badp(T, 0);
  std::vector<uint32_t> depth(T, 0);

  if (level[source] != 0) report->bad_source++;
  if (parent && parent[source] != source) report->bad_source++;

  // Edge rules.
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned t) {
    uint64_t s = 0, mi = 0, r = 0;
    uint32_t d = 0;
    for (uint64_t u = b; u < e; ++u) {
      uint32_t lu = level[u];
      if (lu == DFQ_UNREACHED) continue;
      ++r;
      d = std::max(d, lu + 1);
      for (int64_t k = off[u]; k < off[u + 1]; ++k) {
        uint32_t lv = level[col[k]];
        if (lv == DFQ_UNREACHED) ++mi;
        else if (lv > lu + 1) ++s;
      }
    }
    span[t] = s;
    missed[t] = mi;
    reached[t] = r;
    depth[t] = d;
  });

  // Parent rule (or in-edge witness when no parent array is given).
  if (parent) {
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned t) {
      uint64_t bp = 0;
      for (uint64_t v = b; v < e; ++v) {
        if (v == source) continue;
        uint32_t lv = level[v], p = parent[v];
        if (lv == DFQ_UNREACHED) {
          if (p != DFQ_UNREACHED) ++bp;
          continue;
        }
        if (p >= n || level[p] == DFQ_UNREACHED || level[p] + 1 != lv) {
          ++bp;
          continue;
        }
        if (!std::binary_search(col + off[p], col + off[p + 1], (uint32_t)v)) {
          // rows from foreign CSRs may be unsorted: fall back to a scan
          bool found = false;
          for (int64_t k = off[p]; k < off[p + 1] && !found; ++k) found = col[k] == v;
          if (!found) ++bp;
        }
      }
      badp[t] = bp;
    });
  } else {
    std::vector<uint8_t> witnessed(n, 0);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t u = b; u < e; ++u) {
        uint32_t lu = level[u];
        if (lu == DFQ_UNREACHED) continue;
        for (int64_t k = off[u]; k < off[u + 1]; ++k) {
          uint32_t v = col[k];
          if (level[v] != DFQ_UNREACHED && level[v] == lu + 1) __atomic_store_n(&witnessed[v], 1, __ATOMIC_RELAXED);
        }
      }
    });
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned t) {
      uint64_t bp = 0;
      for (uint64_t v = b; v < e; ++v)
        if (v != source && level[v] != DFQ_UNREACHED && !witnessed[v]) ++bp;
      badp[t] = bp;
    });
  }
  for (unsigned t = 0; t < T; ++t) {
    report->bad_edge_span += span[t];
    report->missed_vertex += missed[t];
    report->bad_parent += badp[t];
    report->reached += reached[t];
    report->depth = std::max(report->depth, depth[t]);
  }
  report->errors = report->bad_source + report->bad_edge_span + report->missed_vertex + report->bad_parent;
  return DFQ_OK;
}

int dfq_compare_levels(uint32_t n, const uint32_t *level, const uint32_t *expect, uint64_t *mismatches_out) {
  if (!level || !expect || !mismatches_out) return DFQ_ERR_INVALID_ARG;
  uint64_t bad = 0;
  for (uint32_t v = 0; v < n; ++v) bad += level[v] != expect[v];
  *mismatches_out = bad;
  return DFQ_OK;
}

}  // extern "C"
