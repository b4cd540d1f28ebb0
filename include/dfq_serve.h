/*
 * dfq_serve.h — serving BFS queries on resident graphs: one persistent
 * launch, continuous batching, admission by working-set footprint (plain C
 * ABI).
 *
 * No counterpart in the reference.  Its closest analogue is the interactive
 * loop of design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:202-214,
 * 302-308, which re-runs a search on a resident graph for each source typed
 * in, one launch per search.  Here the graphs stay resident and one
 * cooperative launch stays on the GPU between dfq_server_start and
 * dfq_server_stop.  The launch is split into teams of `xcds_per_team` XCDs;
 * each team takes the next published query, runs it with the same level
 * loop as dfq_bfs on its own XCDs, publishes the result and takes the next.
 * Queries that arrive while others run are admitted as soon as a request
 * slot and room in the working-set arena are free, so several searches are
 * in flight at once.  Levels (and parents) are those of dfq_bfs for every
 * query, whatever else runs beside it.
 *
 * Graphs are given concatenated as in dfq_batch_create.  Status codes as in
 * dfq_graph.h.  One host thread should drive a server.
 */
#ifndef DFQ_SERVE_H
#define DFQ_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "dfq_batch.h"
#include "dfq_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Query flags */
#define DFQ_SERVE_PARENTS 0x1u  /* also write a BFS parent array */

typedef struct dfq_serve_config {
  uint32_t block_size;      /* 512 or 1024 threads per workgroup                 */
  uint32_t xcds_per_team;   /* XCDs per team: 1, 2, 4 or 8; 0 = one team of all  */
  uint32_t ring_slots;      /* requests published at once (power of two, 2..65536) */
  uint32_t groups;          /* 0: each query is one search on a team; 1: queries
                               of one graph share traversals in group tasks of up
                               to DFQ_SERVE_GROUP_MAX (bit-parallel multi-source
                               search, as dfq_bfs_device_multi); no parents, and
                               levels deeper than 255 are refused               */
  uint64_t arena_bytes;     /* device memory for in-flight working sets; 0 = room
                               for four queries (two full groups) on the largest
                               graph per team                                     */
} dfq_serve_config;

#define DFQ_SERVE_GROUP_MAX 64u

typedef struct dfq_server_info {
  uint32_t running;         /* 1 between start and stop                          */
  uint32_t teams;           /* teams of the running launch (0 before it reports) */
  uint32_t grid_blocks;
  uint32_t block_size;
  uint32_t xcds_per_team;
  uint32_t ring_slots;
  uint32_t in_flight;       /* admitted, not yet released                        */
  uint32_t reserved;
  uint64_t arena_bytes;
  uint64_t arena_used;      /* bytes held by admitted queries                    */
  uint32_t team_blocks[16]; /* workgroups per team                               */
} dfq_server_info;

typedef struct dfq_serve_result {
  uint32_t graph;
  uint32_t source;
  uint32_t depth;           /* levels (max level + 1)                            */
  uint32_t bottom_up_levels;
  uint64_t reached;
  uint64_t edges;           /* CSR entries of the reached vertices (sum of their
                               out-degrees): the edges the search traversed      */
  uint32_t team;            /* team that ran it                                  */
  uint32_t reserved;
  double   service_us;      /* device time from the team taking it to its result */
} dfq_serve_result;

/* One query of dfq_server_replay; times in microseconds from the start of
 * the replay. */
typedef struct dfq_serve_record {
  double   arrival_us;      /* from the trace                                    */
  double   admit_us;        /* published to the device                           */
  double   done_us;         /* completion seen by the host                       */
  double   service_us;      /* as in dfq_serve_result                            */
  uint64_t reached;
  uint64_t edges;           /* as in dfq_serve_result                            */
  uint32_t depth;
  uint32_t team;
} dfq_serve_record;

typedef struct dfq_server dfq_server;

int dfq_serve_config_default(dfq_serve_config *cfg);

/* Device bytes one query needs on a graph of n vertices and m CSR entries,
 * with the given row cut-off (plan.large_max), team width and flags: level
 * (and parent) output, visited / snapshot / frontier bitmaps, frontier
 * segments and work items. */
uint64_t dfq_serve_footprint(uint32_t n, uint64_t m, uint32_t large_max, uint32_t xcds_per_team,
                             uint32_t flags);

/* Device bytes of one group task of `count` searches (1..DFQ_SERVE_GROUP_MAX)
 * on a graph of n vertices and m CSR entries with the default plan: the
 * count level arrays and the multi-source search's working set. */
uint64_t dfq_serve_group_footprint(uint32_t n, uint64_t m, uint32_t count, uint32_t xcds_per_team);

/* Uploads the graphs (checked as in dfq_graph_create) and allocates the
 * arena and the request ring.  cfg may be NULL (defaults). */
int dfq_server_create(uint32_t num_graphs, const uint32_t *n, const uint64_t *m,
                      const int64_t *row_offsets, const uint32_t *col_idx,
                      const dfq_serve_config *cfg, dfq_server **out);
/* Stops the launch if it runs (as dfq_server_stop) and frees everything. */
void dfq_server_destroy(dfq_server *s);

/* Per-graph plan (tier is ignored); applies to queries submitted later.
 * Default: bins 32/256/256, direction optimisation on symmetric graphs. */
int dfq_server_set_plan(dfq_server *s, uint32_t graph, const dfq_query_plan *plan);
int dfq_server_get_plan(const dfq_server *s, uint32_t graph, dfq_query_plan *plan);
/* Team width and workgroup size; only while stopped. */
int dfq_server_set_shape(dfq_server *s, uint32_t xcds_per_team, uint32_t block_size);
int dfq_server_get_info(const dfq_server *s, dfq_server_info *info);

/* Starts the persistent launch. */
int dfq_server_start(dfq_server *s);
/* Waits (bounded) for every admitted query to finish, then ends the launch.
 * Results of finished, unreleased queries stay readable. */
int dfq_server_stop(dfq_server *s);

/* Admits one query: DFQ_ERR_BUSY when no request slot or not enough arena
 * is free right now (retry after a release), DFQ_ERR_CAPACITY when its
 * footprint exceeds the whole arena. */
int dfq_server_submit(dfq_server *s, uint32_t graph, uint32_t source, uint32_t flags,
                      uint64_t *ticket);
/* Group-mode servers: admits `count` queries (1..DFQ_SERVE_GROUP_MAX) of one
 * graph as one group task that searches them in one traversal; tickets[i]
 * is query i's ticket (poll, wait, levels and release work per ticket; the
 * group's arena range is freed with its last member).  flags must be 0.
 * DFQ_ERR_BUSY / DFQ_ERR_CAPACITY as dfq_server_submit.  In group mode
 * dfq_server_submit admits a group of one. */
int dfq_server_submit_group(dfq_server *s, uint32_t graph, uint32_t count, const uint32_t *sources,
                            uint32_t flags, uint64_t *tickets);
/* DFQ_OK with the result when finished, DFQ_ERR_PENDING otherwise;
 * DFQ_ERR_CAPACITY (result still filled) for a group query whose traversal
 * went past 255 levels (its levels are not available). */
int dfq_server_poll(dfq_server *s, uint64_t ticket, dfq_serve_result *result);
/* Polls until finished or timeout_ms passes (DFQ_ERR_TIMEOUT). */
int dfq_server_wait(dfq_server *s, uint64_t ticket, uint32_t timeout_ms, dfq_serve_result *result);
/* Copies a finished query's levels (and parents, when it was submitted with
 * DFQ_SERVE_PARENTS; parent_out may be NULL) to host arrays of n entries. */
int dfq_server_levels(dfq_server *s, uint64_t ticket, uint32_t *level_out, uint32_t *parent_out);
/* Frees a finished query's arena range and request slot. */
int dfq_server_release(dfq_server *s, uint64_t ticket);

/* The serving loop on a trace: query q arrives at arrival_us[q]
 * (non-decreasing); arrivals are admitted in order as slots and arena room
 * free up; finished queries are released at once (after their levels are
 * copied to level_out[q] when level_out and level_out[q] are not NULL).
 * Group mode: arrivals wait per graph, and while fewer than teams + 2 group
 * tasks are in flight, the graph of the oldest waiting query sends up to
 * DFQ_SERVE_GROUP_MAX of its waiting queries as one group (flags must be
 * 0), so queries that arrive while the teams are busy share the next
 * traversal of their graph.
 * The server must be running.  records[q] receives the timings.  Returns
 * DFQ_ERR_BUSY if an arrival cannot be admitted while none of the replay's
 * own queries is in flight (resources held by unreleased tickets). */
int dfq_server_replay(dfq_server *s, uint32_t num_queries, const uint32_t *graph,
                      const uint32_t *source, const double *arrival_us, uint32_t flags,
                      uint32_t *const *level_out, dfq_serve_record *records);

#ifdef __cplusplus
}
#endif
#endif /* DFQ_SERVE_H */
