/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's serial BFS, used as the checker for the
 * GPU engine.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library; the product (libdfqbfs.so, the bfs
 * CLI) never links or calls it.
 *
 * Source followed: NTUDDSNLab/DFQ-BFS
 *   design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu
 *     :216       cost[] init (0 at the source, -1 elsewhere)
 *     :218-223   F = {SOURCE}; X[i] = (i == SOURCE); C[i] = 0
 *     :227-253   level-synchronous loop: the current frontier F is walked
 *                from its back (i = F.size()-1 .. 0); for each edge
 *                id -> nid with X[nid] false, set X[nid], C[nid] = C[id]+1
 *                and append nid to F_next; then F = F_next
 *     :290-300   compare: over vertices with X[i] set, count cost[i] != C[i],
 *                Dist = max level among matching vertices, Touch = |X|
 *
 * Differences in representation only: levels are uint32 with
 * ORACLE_UNREACHED (0xffffffff) for vertices the BFS does not reach (the
 * reference leaves C[i] = 0 there and masks them with X[]); row offsets are
 * int64 instead of Node{int start; int length}.
 *
 * How this oracle is pinned (tests/test_oracle.py): golden level vectors
 * computed by scipy.sparse.csgraph (a third-party BFS) for seeded graphs
 * committed under tests/golden/, closed forms (path, star, complete binary
 * tree, intact grid = x + y, disconnected pieces), and Floyd-Warshall brute
 * force on tiny random graphs over all sources.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_UNREACHED 0xffffffffu

/* Returns 0 on success, -1 on bad arguments, -5 on allocation failure.
 * parent_out (optional) receives the discovering vertex of each reached
 * vertex (source -> itself), ORACLE_UNREACHED elsewhere. */
int oracle_serial_bfs(uint32_t n, const int64_t *row_offsets, const uint32_t *col_idx,
                      uint32_t source, uint32_t *level_out, uint32_t *parent_out) {
  if (!row_offsets || !level_out || source >= n) return -1;
  unsigned char *X = (unsigned char *)calloc(n, 1);
  uint32_t *F = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
  uint32_t *F_next = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
  if (!X || !F || !F_next) {
    free(X);
    free(F);
    free(F_next);
    return -5;
  }
  for (uint32_t i = 0; i < n; ++i) level_out[i] = ORACLE_UNREACHED;
  if (parent_out)
    for (uint32_t i = 0; i < n; ++i) parent_out[i] = ORACLE_UNREACHED;
  /* bfs.cu:216-223 */
  X[source] = 1;
  level_out[source] = 0;
  if (parent_out) parent_out[source] = source;
  size_t fsz = 0;
  F[fsz++] = source;
  /* bfs.cu:227-253: walk the frontier from its back, append discoveries. */
  while (fsz != 0) {
    size_t nsz = 0;
    for (size_t i = fsz; i-- > 0;) {
      uint32_t id = F[i];
      int64_t start = row_offsets[id], end = row_offsets[id + 1];
      for (int64_t j = start; j < end; ++j) {
        uint32_t nid = col_idx[j];
        if (!X[nid]) {
          X[nid] = 1;
          level_out[nid] = level_out[id] + 1;
          if (parent_out) parent_out[nid] = id;
          F_next[nsz++] = nid;
        }
      }
    }
    uint32_t *t = F;
    F = F_next;
    F_next = t;
    fsz = nsz;
  }
  free(X);
  free(F);
  free(F_next);
  return 0;
}

/* bfs.cu:290-300 compare semantics over the CPU-reached set. */
void oracle_compare(uint32_t n, const uint32_t *cpu_level, const uint32_t *gpu_level,
                    uint32_t *dist_out, uint64_t *touch_out, uint64_t *errors_out) {
  uint32_t dist = 0;
  uint64_t touch = 0, errors = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (cpu_level[i] == ORACLE_UNREACHED) continue;
    ++touch;
    if (gpu_level[i] != cpu_level[i]) ++errors;
    else if (cpu_level[i] > dist) dist = cpu_level[i];
  }
  if (dist_out) *dist_out = dist;
  if (touch_out) *touch_out = touch;
  if (errors_out) *errors_out = errors;
}
