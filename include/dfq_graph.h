/*
 * dfq_graph.h — host-side graph interface of the MI355X DFQ-BFS library.
 *
 * Plain C ABI (extern "C", plain pointers and sizes).  Everything here runs
 * on the host CPU and is what the reference's host driver does before and
 * after its cooperative launch:
 *
 *   reference                                          replaced by
 *   --------------------------------------------------  ------------------------------
 *   README.md:19-32  SNAP edge list with the           dfq_load_snap / dfq_save_snap
 *                    "# Nodes: N Edges: M" and
 *                    "# FromNodeId ToNodeId" header
 *   design/shared-q/atomic-sa/our-sep/upper-bound/      dfq_load_dat / dfq_save_dat
 *     bfs.cu:184-200  ".dat" reader (N M SOURCE,
 *                    N x "start length", M x "dst w")
 *   bfs.cu:38-42,194-199  Node{start,length} + edge[]   dfq_csr (int64 offsets, uint32 cols)
 *   bfs.cu:188,204-214  source from the file / stdin   dfq_pick_sources (seeded, degree > 0)
 *   bfs.cu:290-300  compare against the CPU BFS and    dfq_validate_bfs (Graph500-style
 *                   print Dist / Touch / Error            certificate: levels + parent tree)
 *
 * Status codes are shared with dfq_bfs.h (DFQ_OK == 0, negative on error).
 * No C++ exception crosses this boundary.
 */
#ifndef DFQ_GRAPH_H
#define DFQ_GRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (shared by every entry point of the library) ---------- */
#define DFQ_OK                 0
#define DFQ_ERR_INVALID_ARG   -1  /* null pointer, bad size, bad flag          */
#define DFQ_ERR_IO            -2  /* file cannot be opened / read / written     */
#define DFQ_ERR_FORMAT        -3  /* malformed header or body                   */
#define DFQ_ERR_RANGE         -4  /* vertex id >= N, source >= N, edge count    */
#define DFQ_ERR_NOMEM         -5  /* host allocation failed                     */
#define DFQ_ERR_HIP           -6  /* a HIP runtime call failed                  */
#define DFQ_ERR_NO_DEVICE     -7  /* no usable GPU                              */
#define DFQ_ERR_COOPERATIVE   -8  /* cooperative launch not possible            */
#define DFQ_ERR_CAPACITY      -9  /* a queue / slot / memory budget exceeded    */
#define DFQ_ERR_BUSY         -10  /* serving loop: admission refused for now    */
#define DFQ_ERR_STATE        -11  /* call not valid in the object's state       */
#define DFQ_ERR_PENDING      -12  /* serving loop: query not finished yet       */
#define DFQ_ERR_TIMEOUT      -13  /* a bounded wait ran out                     */

/* Unreached vertex marker in level and parent arrays. */
#define DFQ_UNREACHED 0xffffffffu

/* CSR graph owned by the library (free with dfq_csr_free).
 * row_offsets has n+1 entries, col_idx has m entries.  Each row's column
 * indices are sorted ascending when the graph was built by this library. */
typedef struct dfq_csr {
  uint32_t  n;            /* vertices                         */
  uint64_t  m;            /* directed CSR entries             */
  int64_t  *row_offsets;  /* [n + 1]                          */
  uint32_t *col_idx;      /* [m]                              */
} dfq_csr;

/* Flags for building a CSR from an edge list. */
#define DFQ_CSR_SYMMETRIZE  0x1u  /* add (v,u) for every (u,v)            */
#define DFQ_CSR_DEDUP       0x2u  /* drop repeated (u,v)                  */
#define DFQ_CSR_DROP_SELF   0x4u  /* drop (u,u)                           */
#define DFQ_CSR_PERMUTE     0x8u  /* seeded random relabelling of ids     */

const char *dfq_strerror(int status);

void dfq_csr_free(dfq_csr *g);

/* Build a CSR from n vertices and ne directed pairs (src[i] -> dst[i]).
 * Rows come out sorted; duplicates kept unless DFQ_CSR_DEDUP.
 * DFQ_CSR_PERMUTE uses perm_seed.  Returns DFQ_ERR_RANGE for ids >= n. */
int dfq_csr_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t ne,
                       uint32_t n, uint32_t flags, uint64_t perm_seed,
                       dfq_csr **out);

/* Deep copy of caller-owned CSR arrays into a library-owned dfq_csr. */
int dfq_csr_from_arrays(const int64_t *row_offsets, const uint32_t *col_idx,
                        uint32_t n, uint64_t m, dfq_csr **out);

/* ---- file formats -------------------------------------------------------- */

/* SNAP edge list.  "# Nodes: N Edges: M" gives N (vertex ids must be < N);
 * without it N = max id + 1.  Other '#' lines are skipped.  Each data line is
 * "from to" (whitespace separated).  flags as for dfq_csr_from_edges. */
int dfq_load_snap(const char *path, uint32_t flags, dfq_csr **out);
int dfq_save_snap(const char *path, const dfq_csr *g);

/* Reference ".dat" CSR text: "N M SOURCE", then N lines "start length",
 * then M lines "dst weight" (weight ignored on load, written as 1). */
int dfq_load_dat(const char *path, dfq_csr **out, uint32_t *source_out);
int dfq_save_dat(const char *path, const dfq_csr *g, uint32_t source);

/* Binary CSR cache: magic "DFQCSR01", n (u32), m (u64), offsets, cols. */
int dfq_load_csr_bin(const char *path, dfq_csr **out);
int dfq_save_csr_bin(const char *path, const dfq_csr *g);

/* ---- seeded synthetic generators (output independent of thread count) --- */

/* Directed graph, n vertices: out-degree uniform on 1..5; each target drawn
 * with p = 0.7 from [v-2000, v+2000] and otherwise uniform on [0, n);
 * no self loops, no duplicate edges (a repeated draw is redrawn). */
int dfq_gen_amazon_like(uint32_t n, uint64_t seed, dfq_csr **out);

/* Graph500-style R-MAT: 2^scale vertices, edgefactor * 2^scale generated
 * edges with quadrant probabilities a, b, c (d = 1-a-b-c), no noise.
 * flags: DFQ_CSR_* applied to the generated list (the canonical undirected
 * form is SYMMETRIZE|DEDUP|DROP_SELF|PERMUTE). */
int dfq_gen_rmat(uint32_t scale, uint32_t edgefactor, double a, double b,
                 double c, uint64_t seed, uint32_t flags, dfq_csr **out);

/* width x height 4-neighbour grid, v = y*width + x; each undirected edge is
 * dropped independently with probability drop_p; symmetric CSR. */
int dfq_gen_grid2d(uint32_t width, uint32_t height, double drop_p,
                   uint64_t seed, dfq_csr **out);

/* ---- sources and validation --------------------------------------------- */

/* count seeded sources drawn uniformly from vertices with out-degree > 0
 * (distinct while count <= number of such vertices). */
int dfq_pick_sources(const dfq_csr *g, uint32_t count, uint64_t seed,
                     uint32_t *out);

/* Sum of out-degrees over vertices with level != DFQ_UNREACHED (the directed
 * CSR entries a top-down BFS must read), and number of such vertices. */
int dfq_reached_work(const dfq_csr *g, const uint32_t *level,
                     uint64_t *edges_out, uint64_t *vertices_out,
                     uint32_t *depth_out);

typedef struct dfq_validate_report {
  uint64_t errors;            /* total violations (0 == valid BFS)          */
  uint64_t bad_source;        /* level[src] != 0 or parent[src] != src      */
  uint64_t bad_edge_span;     /* edge u->v, both reached, level[v] > level[u]+1 */
  uint64_t missed_vertex;     /* edge u->v, u reached, v unreached          */
  uint64_t bad_parent;        /* parent missing / not an edge / wrong level */
  uint64_t reached;           /* vertices with a level                       */
  uint32_t depth;             /* max level + 1 over reached vertices         */
} dfq_validate_report;

/* Certificate check of a BFS result without a second BFS.  For a directed
 * graph the level array equals the exact hop distances iff:
 * level[src]==0; for every edge u->v with u reached, v is reached and
 * level[v] <= level[u]+1; every reached v != src has an in-edge from a
 * vertex at level[v]-1.  With parent != NULL the last rule is checked on
 * parent[v] (edge parent[v]->v must exist, level[parent[v]] == level[v]-1,
 * unreached vertices carry DFQ_UNREACHED).  Without it, in-edges are found
 * by a pass over all edges. */
int dfq_validate_bfs(const dfq_csr *g, uint32_t source, const uint32_t *level,
                     const uint32_t *parent, dfq_validate_report *report);

/* Reference-style report line fields (bfs.cu:290-300): max level over
 * reached vertices, reached count, and count of level mismatches against a
 * second level array over vertices that array reaches. */
int dfq_compare_levels(uint32_t n, const uint32_t *level,
                       const uint32_t *expect, uint64_t *mismatches_out);

#ifdef __cplusplus
}
#endif
#endif /* DFQ_GRAPH_H */
