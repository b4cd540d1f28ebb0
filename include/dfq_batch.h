/*
 * dfq_batch.h — batched BFS over many resident graphs in one persistent
 * launch (plain C ABI).
 *
 * No counterpart in the reference: NTUDDSNLab/DFQ-BFS runs one search per
 * process (design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:181-317; the
 * interactive loop :202-214,302-308 re-runs a search on the same resident
 * graph, one launch each).  Here a batch of (graph, source) queries runs in
 * one cooperative launch, replayed from a captured hipGraph: large graphs are
 * searched by the whole grid (or by teams of XCDs, several at once) one
 * after another, small ones by one workgroup each with LDS-resident visited
 * and frontier bitmaps.  Levels are the same
 * as dfq_bfs's for every query.
 *
 * Graphs are given concatenated: graph g has n[g] vertices with local ids
 * 0..n[g]-1, its n[g]+1 row offsets (local, starting at 0) follow graph
 * g-1's in row_offsets, its m[g] column indices follow graph g-1's in
 * col_idx.  Status codes as in dfq_graph.h.
 */
#ifndef DFQ_BATCH_H
#define DFQ_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "dfq_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DFQ_TIER_WORKGROUP 0u  /* one workgroup per query, LDS bitmaps */
#define DFQ_TIER_GRID      1u  /* the whole grid per query, one after another */

/* Largest graph (vertices) the workgroup tier accepts. */
#define DFQ_WG_MAX_VERTICES 131072u

/* How queries on graphs of one size class run.  Bins as in dfq_config:
 * small_max <= 64 lane-owned, medium_max one wavefront per vertex, above
 * large_max grid work items (grid tier) / whole workgroup (workgroup tier). */
typedef struct dfq_query_plan {
  uint32_t tier;
  uint32_t small_max;
  uint32_t medium_max;
  uint32_t large_max;
  uint32_t direction_opt;  /* bottom-up levels on symmetric graphs */
  uint32_t alpha;          /* as dfq_config.bu_enter_div */
  uint32_t beta;           /* as dfq_config.bu_exit_div  */
  uint32_t reserved;       /* must be zero */
} dfq_query_plan;

typedef struct dfq_batch_stats {
  double   kernel_ms;      /* hipEvent time of the launch (graph replay)   */
  uint32_t queries;
  uint32_t grid_queries;
  uint32_t wg_queries;
  uint32_t graph_replay;   /* 1 when launched from a captured hipGraph     */
  uint32_t block_size;
  uint32_t grid_blocks;
  uint32_t team_xcds;      /* XCDs per grid-tier team (0: the whole grid)  */
  uint32_t reserved;
} dfq_batch_stats;

typedef struct dfq_tune_entry {
  uint32_t bucket;         /* floor(log2(vertices)) of the graphs          */
  uint32_t queries;        /* queries in this bucket                       */
  dfq_query_plan plan;     /* chosen plan                                  */
  double   cost_us;        /* estimated workgroup-microseconds per query / 1000 CUs */
  double   sub_batch_ms;   /* time of the bucket's queries alone, chosen plan */
} dfq_tune_entry;

typedef struct dfq_tune_report {
  uint32_t entries;        /* filled entries of `entry`                    */
  uint32_t block_size;     /* chosen workgroup size                         */
  uint32_t candidates;     /* plans timed                                   */
  uint32_t team_xcds;      /* chosen grid-tier team width (0: the whole grid) */
  dfq_tune_entry entry[32];
} dfq_tune_report;

typedef struct dfq_batch dfq_batch;

/* Uploads the graphs (checked as in dfq_graph_create) and allocates the
 * working set of both tiers. */
int dfq_batch_create(uint32_t num_graphs, const uint32_t *n, const uint64_t *m,
                     const int64_t *row_offsets, const uint32_t *col_idx,
                     dfq_batch **out);
void dfq_batch_destroy(dfq_batch *b);

/* Sets the queries (graph index, source vertex in that graph's ids).
 * Output levels of query q are kept until the next set_queries. */
int dfq_batch_set_queries(dfq_batch *b, uint32_t num_queries, const uint32_t *graph,
                          const uint32_t *source);

/* Plan for graphs with floor(log2(n)) == bucket (0..31).  Defaults: the
 * workgroup tier up to DFQ_WG_MAX_VERTICES, the grid tier above. */
int dfq_batch_set_plan(dfq_batch *b, uint32_t bucket, const dfq_query_plan *plan);
int dfq_batch_get_plan(const dfq_batch *b, uint32_t bucket, dfq_query_plan *plan);
/* 512 or 1024 threads per workgroup for the batched launch. */
int dfq_batch_set_block_size(dfq_batch *b, uint32_t block_size);
/* Grid-tier queries run on teams of this many XCDs (1, 2, 4 or 8), several
 * at once, each team on its own working set (one copy per team); 0, the
 * default, runs them on the whole grid one after another.  No counterpart
 * in the reference (one search per process there). */
int dfq_batch_set_team_xcds(dfq_batch *b, uint32_t xcds_per_team);

/* Runs every query in one launch; blocks until done. */
int dfq_batch_run(dfq_batch *b, dfq_batch_stats *stats);

/* Results of query q of the last run. */
int dfq_batch_levels(const dfq_batch *b, uint32_t query, uint32_t *level_out);
int dfq_batch_result(const dfq_batch *b, uint32_t query, uint32_t *depth,
                     uint64_t *reached);

/* Autotuner: for each size bucket of the current queries, times the
 * bucket's queries alone under candidate plans (tier x bin cut-offs x
 * direction optimisation), keeps the cheapest in workgroup-time per query,
 * moves buckets whose single query would outlast a quarter of the batch to
 * the grid tier, then picks the workgroup size and the grid tier's team
 * width on the whole batch.  The choices are set on the batch. */
int dfq_batch_autotune(dfq_batch *b, uint32_t repeats, dfq_tune_report *report);

#ifdef __cplusplus
}
#endif
#endif /* DFQ_BATCH_H */
