// Sanitizer driver for the host graph library (dfq-bfs_amd/csrc/graph_host.cpp).
//
// Built by `make -C dfq-bfs_amd asan` with AddressSanitizer and
// UndefinedBehaviorSanitizer (host code only, no device code) and run once:
//   * crafted files that once read or wrote out of bounds: a ".dat" row whose
//     start + length wraps past 2^64, a binary CSR header whose entry count
//     makes the allocation size wrap, numbers longer than 20 digits, a header
//     promising more rows than the file holds;
//   * seeded mutations of small valid ".dat", SNAP and binary CSR files (byte
//     flips, truncations, inserted digits): every load must return a status
//     code, and a successful load must give a well-formed CSR;
//   * the generators, CSR construction, source selection, reached-work and the
//     certificate check on small graphs.
// Any sanitizer report aborts the run with a non-zero status; a wrong status
// code is reported and fails the run too.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

#include "dfq_graph.h"

static int failures = 0;

#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                          \
    }                                                                      \
  } while (0)

static std::string tmp_path(const char *name) {
  const char *d = getenv("TMPDIR");
  return std::string(d && *d ? d : "/tmp") + "/dfq_asan_" + std::to_string((long)getpid()) + "_" + name;
}

static void write_file(const std::string &p, const void *data, size_t size) {
  FILE *f = fopen(p.c_str(), "wb");
  if (!f) {
    perror(p.c_str());
    exit(2);
  }
  if (size) fwrite(data, 1, size, f);
  fclose(f);
}

static void write_text(const std::string &p, const std::string &s) { write_file(p, s.data(), s.size()); }

static std::vector<unsigned char> read_file(const std::string &p) {
  std::vector<unsigned char> v;
  FILE *f = fopen(p.c_str(), "rb");
  if (!f) return v;
  int c;
  while ((c = fgetc(f)) != EOF) v.push_back((unsigned char)c);
  fclose(f);
  return v;
}

// A loaded CSR must be well formed whatever the file held.
static void check_csr(const dfq_csr *g) {
  CHECK(g->row_offsets[0] == 0);
  CHECK((uint64_t)g->row_offsets[g->n] == g->m);
  for (uint32_t v = 0; v < g->n; ++v) CHECK(g->row_offsets[v + 1] >= g->row_offsets[v]);
  for (uint64_t i = 0; i < g->m; ++i) CHECK(g->col_idx[i] < g->n);
}

static int load_any(int kind, const std::string &p) {
  dfq_csr *g = nullptr;
  uint32_t s = 0;
  int rc = kind == 0 ? dfq_load_dat(p.c_str(), &g, &s)
                     : kind == 1 ? dfq_load_snap(p.c_str(), 0, &g) : dfq_load_csr_bin(p.c_str(), &g);
  if (rc == DFQ_OK) {
    check_csr(g);
    if (kind == 0 && g->n) CHECK(s < g->n);
    dfq_csr_free(g);
  } else {
    CHECK(g == nullptr);
  }
  return rc;
}

static void crafted() {
  const std::string p = tmp_path("crafted.dat");
  // start + length wraps past 2^64 (was a heap read out of bounds)
  write_text(p, "1 1 0\n18446744073709551615 2\n0 1\n");
  CHECK(load_any(0, p) == DFQ_ERR_RANGE);
  // start inside, length past the end
  write_text(p, "2 2 0\n0 1\n1 2\n0 1\n1 1\n");
  CHECK(load_any(0, p) == DFQ_ERR_RANGE);
  // numbers longer than 64 bits
  write_text(p, "1 1 0\n0 123456789012345678901234567890\n0 1\n");
  CHECK(load_any(0, p) == DFQ_ERR_FORMAT);
  write_text(p, "99999999999999999999999 1 0\n");
  CHECK(load_any(0, p) == DFQ_ERR_FORMAT);
  // a header promising far more rows and entries than the file holds (was an
  // uncaught std::bad_alloc out of the C ABI)
  write_text(p, "2 1000000000000000000 0\n0 0\n0 0\n");
  CHECK(load_any(0, p) != DFQ_OK);
  write_text(p, "4000000000 1 0\n0 1\n");
  CHECK(load_any(0, p) == DFQ_ERR_FORMAT);
  // the valid reference layout with overlapping rows still loads
  write_text(p, "3 4 0\n2 2\n0 1\n1 1\n2 5\n1 5\n0 5\n1 5\n");
  CHECK(load_any(0, p) == DFQ_OK);

  const std::string q = tmp_path("crafted.csr");
  // binary header with m = 2^62: the allocation size used to wrap
  {
    std::vector<unsigned char> b(8 + 4 + 8 + 16, 0);
    memcpy(b.data(), "DFQCSR01", 8);
    const uint32_t n = 1;
    const uint64_t m = 1ull << 62;
    memcpy(b.data() + 8, &n, 4);
    memcpy(b.data() + 12, &m, 8);
    write_file(q, b.data(), b.size());
    CHECK(load_any(2, q) == DFQ_ERR_FORMAT);
  }
  // header sizes that disagree with the file's length
  {
    std::vector<unsigned char> b(8 + 4 + 8 + 8 * 3 + 4 * 3, 0);
    memcpy(b.data(), "DFQCSR01", 8);
    const uint32_t n = 2;
    const uint64_t m = 7;
    memcpy(b.data() + 8, &n, 4);
    memcpy(b.data() + 12, &m, 8);
    write_file(q, b.data(), b.size());
    CHECK(load_any(2, q) == DFQ_ERR_FORMAT);
  }
  const std::string r = tmp_path("crafted.txt");
  write_text(r, "# Nodes: 99999999999999999999999 Edges: 1\n0 1\n");
  CHECK(load_any(1, r) == DFQ_ERR_FORMAT);
  write_text(r, "0 123456789012345678901234567890\n");
  CHECK(load_any(1, r) == DFQ_ERR_FORMAT);
  remove(p.c_str());
  remove(q.c_str());
  remove(r.c_str());
}

// Seeded mutations of valid files of each format.
static void mutations() {
  dfq_csr *g = nullptr;
  CHECK(dfq_gen_rmat(6, 4, 0.57, 0.19, 0.19, 3, 0xF, &g) == DFQ_OK);
  if (!g) return;
  const std::string paths[3] = {tmp_path("m.dat"), tmp_path("m.txt"), tmp_path("m.csr")};
  CHECK(dfq_save_dat(paths[0].c_str(), g, 1) == DFQ_OK);
  CHECK(dfq_save_snap(paths[1].c_str(), g) == DFQ_OK);
  CHECK(dfq_save_csr_bin(paths[2].c_str(), g) == DFQ_OK);
  dfq_csr_free(g);
  uint64_t state = 0x9E3779B97F4A7C15ull;
  auto rnd = [&state](uint64_t bound) {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return bound ? state % bound : 0;
  };
  const std::string mp = tmp_path("mut");
  for (int kind = 0; kind < 3; ++kind) {
    const std::vector<unsigned char> base = read_file(paths[kind]);
    CHECK(!base.empty());
    CHECK(load_any(kind, paths[kind]) == DFQ_OK);
    for (int trial = 0; trial < 400; ++trial) {
      std::vector<unsigned char> b = base;
      const int edits = 1 + (int)rnd(4);
      for (int e = 0; e < edits && !b.empty(); ++e) {
        const size_t at = (size_t)rnd(b.size());
        switch (rnd(5)) {
          case 0: b[at] ^= (unsigned char)(1u << rnd(8)); break;              // bit flip
          case 1: b.resize(at); break;                                         // truncate
          case 2: b.insert(b.begin() + (long)at, (unsigned char)('0' + rnd(10))); break;  // digit
          case 3: b[at] = "0123456789 \n-#"[rnd(15)]; break;                   // token char
          default: b.insert(b.begin() + (long)at, 20, (unsigned char)'9'); break;  // long number
        }
      }
      write_file(mp, b.data(), b.size());
      (void)load_any(kind, mp);
    }
    remove(paths[kind].c_str());
  }
  remove(mp.c_str());
}

static void library_paths() {
  dfq_csr *g = nullptr;
  CHECK(dfq_gen_amazon_like(3000, 5, &g) == DFQ_OK);
  if (g) {
    check_csr(g);
    dfq_csr_free(g);
  }
  g = nullptr;
  CHECK(dfq_gen_grid2d(70, 50, 0.02, 9, &g) == DFQ_OK);
  if (g) {
    check_csr(g);
    dfq_csr_free(g);
  }
  g = nullptr;
  CHECK(dfq_gen_rmat(12, 8, 0.57, 0.19, 0.19, 12, 0xF, &g) == DFQ_OK);
  if (!g) return;
  check_csr(g);
  uint32_t src[8];
  CHECK(dfq_pick_sources(g, 8, 2022, src) == DFQ_OK);
  // a BFS written out here (plain level-synchronous queue) gives the
  // certificate something true to accept and something planted to reject
  std::vector<uint32_t> level(g->n, DFQ_UNREACHED), parent(g->n, DFQ_UNREACHED), q;
  level[src[0]] = 0;
  parent[src[0]] = src[0];
  q.push_back(src[0]);
  for (size_t h = 0; h < q.size(); ++h) {
    const uint32_t u = q[h];
    for (int64_t k = g->row_offsets[u]; k < g->row_offsets[u + 1]; ++k) {
      const uint32_t v = g->col_idx[k];
      if (level[v] == DFQ_UNREACHED) {
        level[v] = level[u] + 1;
        parent[v] = u;
        q.push_back(v);
      }
    }
  }
  dfq_validate_report rep;
  CHECK(dfq_validate_bfs(g, src[0], level.data(), parent.data(), &rep) == DFQ_OK && rep.errors == 0);
  CHECK(dfq_validate_bfs(g, src[0], level.data(), nullptr, &rep) == DFQ_OK && rep.errors == 0);
  level[q.back()] += 1;
  CHECK(dfq_validate_bfs(g, src[0], level.data(), parent.data(), &rep) == DFQ_OK && rep.errors > 0);
  uint64_t edges = 0, verts = 0;
  uint32_t depth = 0;
  CHECK(dfq_reached_work(g, level.data(), &edges, &verts, &depth) == DFQ_OK && verts == q.size());
  dfq_csr *h = nullptr;
  CHECK(dfq_csr_from_arrays(g->row_offsets, g->col_idx, g->n, g->m, &h) == DFQ_OK);
  if (h) dfq_csr_free(h);
  dfq_csr_free(g);
}

int main() {
  crafted();
  mutations();
  library_paths();
  if (failures) {
    fprintf(stderr, "graph_host_check: %d failed checks\n", failures);
    return 1;
  }
  printf("graph_host_check: all checks passed\n");
  return 0;
}
