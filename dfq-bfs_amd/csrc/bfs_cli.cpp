// bfs — command-line driver with the reference's interface
// (design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:181-317):
//
//   bfs <graph> [x]        run one BFS from the file's source and exit
//   bfs <graph> $          interactive: after each run ask "Exit? [y/n]",
//                          on "n" read a new source from stdin
//   bfs <graph> --source S run from S (SNAP files carry no source; default 0)
//   bfs <graph> --json     one JSON line instead of the text report
//
// <graph> is a reference ".dat" CSR text (carries N M SOURCE), a SNAP edge
// list (.txt/.snap/.el; "# Nodes: N Edges: M" header, README.md:19-32) or a
// binary CSR cache (.csr).  The report keeps the reference's lines
// "# Nodes :", "# Edges :", "Source  :", "- Dist  =", "- Touch =",
// "- Error =", "- Time  = <ms>ms" (:206-208, :297-300) and the trailing
// attribute reset (:311), so harnesses that parse the time token
// (plot.py:32) keep working.  "Error" counts violations of the BFS
// certificate (dfq_validate_bfs: level and parent-tree rules over every
// vertex and edge) instead of re-running a CPU BFS.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "dfq_bfs.h"
#include "dfq_graph.h"

static bool ends_with(const std::string &s, const char *suf) {
  size_t n = strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

static int die(const char *what, int rc) {
  fprintf(stderr, "bfs: %s: %s\n", what, dfq_strerror(rc));
  return 1;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: bfs <graph.dat|.txt|.csr> [x|$] [--source S] [--json]\n");
    return 2;
  }
  std::string path = argv[1];
  bool interactive = argv[argc - 1][0] == '$';
  bool json = false;
  long long src_arg = -1;
  for (int i = 2; i < argc; ++i) {
    if (!strcmp(argv[i], "--json")) json = true;
    else if (!strcmp(argv[i], "--source") && i + 1 < argc) src_arg = atoll(argv[++i]);
  }

  dfq_csr *csr = nullptr;
  uint32_t source = 0;
  int rc;
  if (ends_with(path, ".dat")) rc = dfq_load_dat(path.c_str(), &csr, &source);
  else if (ends_with(path, ".csr")) rc = dfq_load_csr_bin(path.c_str(), &csr);
  else rc = dfq_load_snap(path.c_str(), 0, &csr);
  if (rc != DFQ_OK) return die(path.c_str(), rc);
  if (src_arg >= 0) source = (uint32_t)src_arg;

  dfq_graph *g = nullptr;
  rc = dfq_graph_create(csr->row_offsets, csr->col_idx, csr->n, csr->m, &g);
  if (rc != DFQ_OK) {
    dfq_csr_free(csr);
    return die("dfq_graph_create", rc);
  }
  std::vector<uint32_t> level(csr->n), parent(csr->n);
  // the graph's name in bold yellow, as the reference's first line (:183)
  if (!json) printf("\033[0;1;33m%s\033[0;1m\n", path.c_str());
  bool first = true;
  for (std::string answer; answer != "y";) {
    // Source prompt; unlike the reference an invalid source in the file is an
    // error instead of an endless loop (:204-214).
    if (!first) {
      long long s = -1;
      printf("Source  : ");
      fflush(stdout);
      if (!(std::cin >> s)) break;
      if (s < 0 || s >= (long long)csr->n) {
        printf("source out of range [0, %u)\n", csr->n);
        continue;
      }
      source = (uint32_t)s;
    } else if (!json) {
      printf("# Nodes : %u\n# Edges : %llu\nSource  : %u\n", csr->n, (unsigned long long)csr->m, source);
    }
    first = false;
    if (source >= csr->n) {
      dfq_graph_destroy(g);
      dfq_csr_free(csr);
      return die("source", DFQ_ERR_RANGE);
    }
    dfq_stats st;
    rc = dfq_bfs(g, source, level.data(), parent.data(), &st);
    if (rc != DFQ_OK) {
      dfq_graph_destroy(g);
      dfq_csr_free(csr);
      return die("dfq_bfs", rc);
    }
    dfq_validate_report rep;
    rc = dfq_validate_bfs(csr, source, level.data(), parent.data(), &rep);
    if (rc != DFQ_OK) {
      dfq_graph_destroy(g);
      dfq_csr_free(csr);
      return die("dfq_validate_bfs", rc);
    }
    uint64_t edges = 0;
    dfq_reached_work(csr, level.data(), &edges, nullptr, nullptr);
    const uint32_t dist = rep.depth ? rep.depth - 1 : 0;
    if (json) {
      printf("{\"nodes\": %u, \"edges\": %llu, \"source\": %u, \"dist\": %u, \"reached\": %llu, "
             "\"errors\": %llu, \"time_ms\": %.6f, \"edges_traversed\": %llu}\n",
             csr->n, (unsigned long long)csr->m, source, dist, (unsigned long long)rep.reached,
             (unsigned long long)rep.errors, st.kernel_ms, (unsigned long long)edges);
    } else {
      printf("- Dist  = %u\n", dist);
      printf("- Touch = %g %% (%llu / %u)\n", (double)rep.reached / (double)csr->n * 100.0,
             (unsigned long long)rep.reached, csr->n);
      if (rep.errors) printf("- Error = \033[31m%llu\033[0;1m\n", (unsigned long long)rep.errors);
      else printf("- Error = 0\n");
      printf("- Time  = %gms\n", st.kernel_ms);
    }
    if (!interactive) break;
    printf("Exit? [y/n] ");
    fflush(stdout);
    if (!(std::cin >> answer)) break;
  }
  // The text report ends with an attribute reset after the last line, as the
  // reference's does (:311): harnesses take the time from the next-to-last
  // whitespace token (plot.py:32, `stdout.split()[-2][:-2]`).
  if (!json) printf("\033[0m");
  fflush(stdout);
  dfq_graph_destroy(g);
  dfq_csr_free(csr);
  return 0;
}
