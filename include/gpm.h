/*
 * gpm.h -- C ABI of the MI355X (gfx950, wave64) graph-pattern-mining engine.
 *
 * Every entry point takes plain pointers and sizes and returns an int status
 * (GPM_OK == 0).  On failure the thread-local message behind gpm_last_error()
 * says what went wrong; the C++ solver shims in gpm-artifact_amd/include turn a
 * non-zero status into the reference's abort-with-log behaviour.
 *
 * Conventions
 *   - vertex ids and CSR offsets are uint32 (the reference uses int VID for
 *     both, src/graph/graph.h:422-423); graphs with >= 2^32 adjacency entries
 *     are rejected with GPM_EINVAL.
 *   - adjacency lists are sorted ascending and hold no duplicates or self loops
 *     unless a function says otherwise.
 *   - counts are uint64 (AccType = unsigned long long, src/utils/utils.h:25).
 *   - host arrays passed in are copied; arrays handed out (uint32_t** outs) are
 *     malloc'ed by the library and released with gpm_free().
 *   - device-side calls are synchronous (like the reference solvers, which
 *     end with a device synchronise + D2H of the counters) and run on the
 *     graph's own HIP stream; one graph handle must not be used by two host
 *     threads at once.
 */
#ifndef GPM_H_
#define GPM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum gpm_status {
  GPM_OK = 0,
  GPM_EINVAL = 1,      /* bad argument (null pointer, size, unknown id)      */
  GPM_EHIP = 2,        /* a HIP runtime call failed                          */
  GPM_ENOMEM = 3,      /* host or device allocation failed                   */
  GPM_EUNSUPPORTED = 4,/* valid request the engine does not implement        */
  GPM_EIO = 5          /* file could not be opened / parsed                  */
};

/* Message for the last failing call on this host thread ("" if none). */
const char* gpm_last_error(void);
/* ABI version, bumped on any signature change. */
int gpm_abi_version(void);
/* Release an array the library handed out. */
void gpm_free(void* p);

/* ------------------------------------------------ host graph preparation */

/* Flags for gpm_csr_from_edges. */
#define GPM_CSR_SYMMETRIZE 0x1u /* add (v,u) for every (u,v)                 */
#define GPM_CSR_DEDUP 0x2u      /* drop repeated (u,v)                         */
#define GPM_CSR_DROP_SELF 0x4u  /* drop (u,u)                                  */
#define GPM_CSR_SORT 0x8u       /* sort every adjacency list ascending         */

/*
 * Edge list -> CSR.  With flags == 0 this is Graph::Init's CSR build: lists
 * keep input order (src/graph/graph.h:271-287).  The configurations' graphs
 * use SYMMETRIZE|DEDUP|DROP_SELF|SORT.
 */
int gpm_csr_from_edges(uint32_t n, const uint32_t* src, const uint32_t* dst,
                       uint64_t n_edges, uint32_t flags, uint32_t** offsets,
                       uint32_t** adj, uint64_t* m);

/*
 * Degree orientation, replacing Graph::orientation (src/graph/graph.h:116-204):
 * keep u->v iff deg(v) > deg(u) or (deg(v) == deg(u) and v > u), where deg is
 * the list length in the input CSR (graph.h:128-129).  List order is kept.
 */
int gpm_csr_orient(uint32_t n, const uint32_t* offsets, const uint32_t* adj,
                   uint32_t** out_offsets, uint32_t** out_adj, uint64_t* out_m);

/*
 * Degree relabel, replacing Graph::SortCSRGraph (src/graph/graph.h:206-252):
 * stable sort of vertices by list length, descending; new_id[old] is written
 * to *perm (n entries) and every list is renamed and sorted ascending.
 */
int gpm_csr_sort_by_degree(uint32_t n, const uint32_t* offsets,
                           const uint32_t* adj, uint32_t** out_offsets,
                           uint32_t** out_adj, uint32_t** perm);

/*
 * Matrix-market edge reader with the semantics of Loader::_load_mtx
 * (src/graph/io.h:84-139): comment lines are skipped, the "n n m" header is
 * read and ignored, self loops are dropped, ids are rebased by the minimum id
 * seen and n = max - min + 1.  No symmetrisation, dedup or sort.
 */
int gpm_load_mtx(const char* path, uint32_t* n, uint32_t** src, uint32_t** dst,
                 uint64_t* n_edges);

/*
 * Labeled graph reader for the sub_ceci text format
 * (sub_ceci/graph/graph.cpp:86-166): "t |V| |E|", then "v id label degree"
 * lines in id order, then "e u v" lines with each undirected edge once.
 * Returns the symmetric CSR with sorted lists and the labels; *n_labels =
 * max label + 1.  Unlike the reference, inconsistent files (ids out of order,
 * declared degrees not met, counts not matching the header) are rejected
 * with GPM_EIO instead of being read past.  Query graphs use the same format.
 */
int gpm_load_labeled(const char* path, uint32_t* n, uint32_t** offsets, uint32_t** adj,
                     uint64_t* m, uint32_t** labels, uint32_t* n_labels);

/*
 * sub_ceci compressed graph (sub_ceci/graph/graph.cpp:184-304): a degree
 * file (int32 4, uint32 |V|, uint32 |E|, uint32 degree[|V|]), an adjacency
 * file (the |V| neighbour lists back to back, 2|E| uint32 ids) and, when
 * label_path is not NULL, a label file (uint32 label[|V|]).  Lists are sorted
 * on load, as in the reference; *n_labels = max label + 1.  Files whose sizes
 * disagree with the header, or ids >= |V|, are rejected with GPM_EIO.
 */
int gpm_load_ceci_bin(const char* degree_path, const char* adj_path, const char* label_path,
                      uint32_t* n, uint32_t** offsets, uint32_t** adj, uint64_t* m,
                      uint32_t** labels, uint32_t* n_labels);
/* Writes the same three files (Graph::storeComparessedGraph,
 * sub_ceci/graph/graph.cpp:306-366); label_path may be NULL. */
int gpm_store_ceci_bin(const char* degree_path, const char* adj_path, const char* label_path,
                       uint32_t n, const uint32_t* offsets, const uint32_t* adj,
                       const uint32_t* labels);

/*
 * GraphMiner binary directory (the output of the reference's converter,
 * scripts/gbin_reader/greader.h:129-272): graph.vertex.bin (int64 row
 * pointer, |V|+1), graph.edge.bin (int32 ids, nnz), graph.meta.txt (|V|,
 * nnz, 4, max degree).  Lists are returned in file order.  Inconsistent
 * files are rejected with GPM_EIO.
 */
int gpm_load_gbin(const char* dir, uint32_t* n, uint32_t** offsets, uint32_t** adj,
                  uint64_t* m, uint32_t* max_degree);
/* Writes a CSR as a GraphMiner binary directory (the directory must exist). */
int gpm_store_gbin(const char* dir, uint32_t n, const uint32_t* offsets, const uint32_t* adj);

/*
 * Seeded Erdos-Renyi G(n, M): M distinct undirected edges, no self loops,
 * emitted once each as (min, max).  No counterpart in the reference (which
 * only reads files); used for the triangle-counting workload.
 */
int gpm_gen_er(uint32_t n, uint64_t m_undirected, uint64_t seed,
               uint32_t** src, uint32_t** dst, uint64_t* n_edges);

/*
 * Seeded R-MAT: edge_factor * 2^scale samples with quadrant probabilities
 * (a, b, c, 1-a-b-c), no noise and no permutation.  Raw samples (self loops
 * and duplicates included); build the CSR with gpm_csr_from_edges.
 */
int gpm_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                 uint64_t seed, uint32_t** src, uint32_t** dst,
                 uint64_t* n_edges);

/* Seeded uniform labels in [0, n_labels).  Replaces the "randomK" branch of
 * Loader::LoadVLabel (src/graph/io.h:43-49), which used unseeded rand(). */
int gpm_gen_labels(uint32_t n, uint32_t n_labels, uint64_t seed,
                   uint32_t** labels);

/* --------------------------------------------------------- device graph */

typedef struct gpm_graph gpm_graph;

#define GPM_GRAPH_ORIENTED 0x1u /* CSR is a DAG (TC, k-clique)               */
#define GPM_GRAPH_SYMMETRIC 0x2u /* CSR is undirected (patterns, matching)   */

/*
 * Upload a CSR (and optional labels) to the current HIP device; the handle
 * owns the device copy until gpm_graph_destroy.  Replaces Graph::Init's H2D
 * copies and Graph::DeviceObject (src/graph/graph.h:311-366).  labels may be
 * NULL; when given, a label-partitioned adjacency index is also built for
 * the matcher.
 */
int gpm_graph_create(const uint32_t* offsets, uint32_t n, const uint32_t* adj,
                     uint64_t m, const uint32_t* labels, uint32_t n_labels,
                     uint32_t flags, gpm_graph** out);
int gpm_graph_destroy(gpm_graph* g);
/* n, m, max list length. */
int gpm_graph_info(const gpm_graph* g, uint32_t* n, uint64_t* m,
                   uint32_t* max_deg);
/*
 * Device-side orientation of a SYMMETRIC graph into a new ORIENTED handle,
 * same rule as gpm_csr_orient (src/graph/graph.h:116-204), computed on the GPU.
 */
int gpm_graph_orient(gpm_graph* sym, gpm_graph** dag);

/*
 * Builds a device graph from an edge list on the GPU, replacing the host
 * passes of Graph::Init (src/graph/graph.h:254-340): csr_flags as for
 * gpm_csr_from_edges (SYMMETRIZE, DEDUP, DROP_SELF, SORT; DEDUP implies
 * SORT; without SORT every list keeps edge-list order) -- the same CSR
 * gpm_csr_from_edges returns, built with device radix sorts.  labels /
 * n_labels / graph_flags as for gpm_graph_create.
 */
int gpm_graph_build(uint32_t n, const uint32_t* src, const uint32_t* dst, uint64_t n_edges,
                    uint32_t csr_flags, const uint32_t* labels, uint32_t n_labels,
                    uint32_t graph_flags, gpm_graph** out);
/*
 * Degree relabel on the GPU, replacing Graph::SortCSRGraph
 * (src/graph/graph.h:206-252) and matching gpm_csr_sort_by_degree: vertices
 * stable-sorted by list length, descending; *out is a new handle over the
 * renamed graph (lists sorted, labels carried over) and *perm (n entries,
 * release with gpm_free) holds new_id[old].
 */
int gpm_graph_relabel_by_degree(gpm_graph* g, gpm_graph** out, uint32_t** perm);
/* Copies the device CSR to caller buffers of n + 1 and m entries. */
int gpm_graph_copy_csr(const gpm_graph* g, uint32_t* offsets, uint32_t* adj);

/* Event-timed duration (ms) of the kernels of the last mining call on g. */
int gpm_last_kernel_ms(const gpm_graph* g, double* ms);
/*
 * Per-kernel split of the last mining call: ms[i] is the event-timed span
 * ending with the i-th main kernel (names[i] names it); *count receives the
 * number of kernels, of which min(capacity, *count) are written.
 */
int gpm_last_kernel_breakdown(const gpm_graph* g, double* ms, const char** names,
                              int capacity, int* count);

/* ------------------------------------------------------------ scheduling */

enum gpm_sched_policy {
  GPM_SCHED_DEGREE = 0,  /* plain degree-based scheduler (reference style)  */
  GPM_SCHED_LEARNED = 1, /* cost-model order + chunking from per-root costs */
  GPM_SCHED_DEGREE_SPLIT = 2, /* k-clique only: LEARNED's ordering, block
                                 promotion and heavy-root splitting driven by
                                 the degree estimate d+(v)^2 instead of a
                                 model (root_cost unused)                   */
};

typedef struct gpm_sched {
  int policy;              /* gpm_sched_policy                               */
  uint32_t chunk;          /* k-clique wave tier: roots per chunk of its
                              static interleave (wave w takes chunks w,
                              w + waves, ...), 0 = automatic (1)             */
  uint32_t root_begin;     /* root range [root_begin, root_end) (vertex ids); */
  uint32_t root_end;       /*   root_end == 0 means "all vertices"            */
  const float* root_cost;  /* LEARNED: predicted cost per vertex (n floats)   */
  uint32_t wave_cap;       /* local-set size served by one wave, 0 = 64       */
  uint32_t n_root_list;    /* > 0: only the roots in root_list (host array),  */
  const uint32_t* root_list; /*   each vertex at most once, ids < n           */
} gpm_sched;

/* ------------------------------------------------------------- mining */

enum gpm_tc_variant {
  GPM_TC_EDGE = 0,   /* e_centric: edge-parallel binary-search intersection */
  GPM_TC_VERTEX = 1, /* v_centric: vertex-parallel LDS hash table           */
};

/*
 * Triangle count of an ORIENTED graph, replacing TCSolver
 * (src/TCSolver.cuh:10-121; e_centric :27-59, v_centric :64-113).
 */
int gpm_tc(gpm_graph* g, int variant, uint64_t* out);

/*
 * gpm_tc repeated reps times back to back on the device (no host round trip
 * between repetitions), for throughput measurement of launches too short to
 * time one by one; *out is the count (every repetition recounts it) and
 * gpm_last_kernel_ms reports the mean time per repetition.  No counterpart
 * in the reference.
 */
int gpm_tc_repeat(gpm_graph* g, int variant, int reps, uint64_t* out);
/*
 * The same measurement for the edge-parallel count with up to 32 counts per
 * launch (one kernel, grid y = count index, each count into its own counter
 * slot), so independent counts overlap on the device without a graph of
 * launches.  No counterpart in the reference.
 */
int gpm_tc_repeat_fused(gpm_graph* g, int reps, uint64_t* out);

/*
 * Edge-parallel triangle count (the GPM_TC_EDGE kernel of gpm_tc) that
 * allocates nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory -- code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc.
 *
 * Inputs are device pointers to an oriented, sorted CSR (d_offsets: n + 1
 * entries with d_offsets[n] == m; d_adj: m ids < n); the count is written to
 * the device uint64 *d_count.  The call only enqueues kernels on `stream` (a
 * hipStream_t; NULL = the null stream) and returns: *d_count holds the count
 * once the stream has reached that point.  Its staging (the COO source of
 * every edge and the kernel's per-block partial sums) lives in a fixed
 * device array of 8 * 8,192 + 4 * 16,777,216 bytes that each GPU keeps until
 * the process ends; m > 16,777,216 is rejected with GPM_EINVAL.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (two host
 * threads, or two streams running at once) race and must be serialised by
 * the caller.
 */
int gpm_tc_edge_noalloc(const uint32_t* d_offsets, const uint32_t* d_adj, uint32_t n,
                        uint64_t m, uint64_t* d_count, void* stream);

/*
 * k-clique count (3 <= k <= 5) of an ORIENTED graph, replacing CFSolver
 * (src/CFSolver.cuh:99-157, which supports k = 4, 5).  sched may be NULL.
 */
int gpm_clique(gpm_graph* g, int k, const gpm_sched* sched, uint64_t* out);

/*
 * 3-, 4- and 5-clique counts of an ORIENTED graph in one pass over the roots
 * (each root's local DAG is built once and counted for every k):
 * out[0], out[1], out[2] = counts for k = 3, 4, 5.  No counterpart in the
 * reference, whose CFSolver counts one k per call (CFSolver.cuh:99-157).
 */
int gpm_clique_all(gpm_graph* g, const gpm_sched* sched, uint64_t out[3]);

/*
 * Out-degree limits of the k-clique kernels on this device: roots with
 * out-degree <= *wave_cap run on one wave (rows in registers), <= *lds_cap
 * on one block with LDS rows, <= *wide_cap on one block with wide LDS rows,
 * above that on one block with rows in HBM.
 */
int gpm_clique_tier_caps(uint32_t* wave_cap, uint32_t* lds_cap, uint32_t* wide_cap);

/*
 * Per-root work log for the cost model: work[v] = number of adjacency
 * entries the plain k-clique plan scans for root vertex v (n entries).
 */
int gpm_clique_root_work(gpm_graph* g, int k, uint64_t* work);

/*
 * Algorithmic bytes of the canonical edge-rooted k-clique plan
 * (CFSolver.cuh:15-97: materialise W1 = N+(v0) & N+(v1), ...):
 * 4(n+1) + 8|E+| + sum over set operations of 4(|A|+|B|) (+ 4|out| when the
 * result is materialised).  k == 3 gives the triangle-counting figure.
 */
int gpm_clique_plan_bytes(gpm_graph* g, int k, uint64_t* bytes);

enum gpm_pattern_id {
  GPM_P1 = 1, GPM_P2, GPM_P3, GPM_P4, GPM_P5, GPM_P6, GPM_P7, GPM_P8,
  GPM_MOTIF4 = 9
};
enum gpm_formulation {
  GPM_FORM_G2MINER = 0, /* direct enumeration (G2MinerSolver)              */
  GPM_FORM_PATTERN = 1, /* frequency x count - correction (PatternSolver)   */
  GPM_FORM_BLOCK_PREFILTER = 2, /* PatternSolver with the BLOCK_OPT block-
                                   mask prefilter for P4, P5, P6
                                   (PatternSolver.cuh:109-151)              */
};

/*
 * Fixed-pattern count on a SYMMETRIC graph, replacing G2MinerSolver
 * (src/G2MinerSolver.cuh:18-175) and PatternSolver (src/PatternSolver.cuh:
 * 17-264).  out[0..5]: for P1..P8 out[0] is the pattern count; for MOTIF4
 * out = induced {3-star, 4-path, tailed-triangle, 4-cycle, diamond,
 * 4-clique}.  Under the counting formulations, P2, P3, P7 and P8 also report
 * the frequency-weighted sum in out[1] and the correction in out[2]
 * (out[0] = out[1] - out[2]), and P4 (prefilter only), P5 and P6 the pair
 * count in out[1] (= out[0]).  *scalar receives what the matching reference
 * solver returns (G2Miner: counters[0]; PatternSolver: h[1] or h[1]-h[2],
 * modulo 2^64).
 */
int gpm_pattern(gpm_graph* g, int pattern_id, int formulation,
                uint64_t out[6], uint64_t* scalar);

/* ------------------------------------------------------ set primitives */

enum gpm_setop_kind {
  GPM_SETOP_ISECT_COUNT = 0, /* |{x in A & B : x < ub}|  (intersect_num)     */
  GPM_SETOP_ISECT = 1,       /* that list, sorted         (intersect)         */
  GPM_SETOP_DIFF_COUNT = 2,  /* |{x in A \ B : x < ub}| (difference_num)      */
  GPM_SETOP_DIFF = 3,        /* that list, sorted         (difference_set)    */
  GPM_SETOP_FUSE = 4,        /* A \ B, A & B and B \ A in one pass (fuse_set) */
};
enum gpm_set_variant {
  GPM_SETV_BINARY = 0,    /* per-lane binary search of B (search.cuh:20-35)   */
  GPM_SETV_TWO_PHASE = 1, /* 64 LDS pivots, then a 1/64 slice of B
                             (binary_search_2phase, search.cuh:54-87)        */
  GPM_SETV_MERGE = 2,     /* merge path: 64 diagonal segments, one per lane  */
  GPM_SETV_HASH = 3,      /* B in an LDS hash table (|B| <= 512, else binary) */
  GPM_SETV_BITMAP = 4,    /* B in an LDS bitmap over [B[0], B[|B|-1]]
                             (span <= 32768, else binary)                    */
};

/*
 * Batched sorted-set operations, one wave per pair, replacing the primitives
 * of src/graph/set_intersection.cuh:6-304 and set_difference.cuh:6-325.
 * Pair i is A_i = a[a_off[i] .. a_off[i+1]) and B_i = b[b_off[i] ..
 * b_off[i+1]), each strictly increasing.  A is the iterated list (A \ B is
 * never swapped); ub (NULL = none; ignored by FUSE) keeps only x < ub[i].
 * counts[i] receives the size (FUSE: counts[3i .. 3i+2] = |A\B|, |A&B|,
 * |B\A|).  ISECT, DIFF and FUSE write list i at out_a + a_off[i] (FUSE:
 * A \ B there, A & B at out_ab + a_off[i], B \ A at out_b + b_off[i]).
 * Host arrays; synchronous.  No single counterpart in the reference: its
 * primitives are device functions called inside the mining kernels.
 */
int gpm_setops(int op, int variant, uint32_t n_pairs, const uint64_t* a_off,
               const uint32_t* a, const uint64_t* b_off, const uint32_t* b,
               const uint32_t* ub, uint64_t* counts, uint32_t* out_a,
               uint32_t* out_ab, uint32_t* out_b);

/* Query graph for the labeled matcher (<= 16 vertices, connected). */
typedef struct gpm_query {
  uint32_t n;               /* vertices                                     */
  uint32_t n_edges;         /* undirected edges                              */
  const uint32_t* edges;    /* 2*n_edges endpoints (u0,v0,u1,v1,...)         */
  const uint32_t* labels;   /* n labels                                      */
} gpm_query;

/*
 * How the matcher tests a streamed window of 64 candidates for membership in
 * the other backward neighbours' label slices (the role of the reference's
 * binary_search / intersect variants, src/graph/search.cuh:6-87 and
 * src/gpu_kernels/tuning_schedules.cuh:169-202, chosen there at compile time).
 */
enum gpm_match_setv {
  GPM_MATCH_SETV_AUTO = 0,    /* autotune: MERGE when the plan streams slices
                                 of expected length >= 256 entries, else
                                 BINARY                                      */
  GPM_MATCH_SETV_BINARY = 1,  /* every lane binary-searches its candidate     */
  GPM_MATCH_SETV_MERGE = 2,   /* the wave walks the probed slice in 64-entry
                                 coalesced chunks from a per-stream cursor and
                                 each lane finds its candidate among the
                                 chunk's values; probes of short streams (< 4
                                 windows) or of slices much longer than the
                                 stream are binary-searched instead          */
};

typedef struct gpm_tuning {
  int autotune;             /* 1: project autotuner picks per-query chunks;
                               with policy LEARNED and no root_model also the
                               plan, by one trial launch over sampled tasks */
  int policy;               /* gpm_sched_policy for the matching order      */
  uint32_t chunk;           /* candidate roots per task, 0 = automatic       */
  int use_graph;            /* 1: replay the batch as one hipGraph; needs
                               autotune = 0 (else GPM_EUNSUPPORTED)          */
  const float* root_model;  /* LEARNED: packed root-cost MLP (mean[4] std[4]
                               W1[64x4] b1[64] W2[64x64] b2[64] W3[64] b3[1],
                               4553 floats) predicting log1p of a root's
                               weighted DFS steps from (degree, depth-1
                               slice length, 2-hop degree sum, query size);
                               it picks the matching order among the
                               candidate orders and sizes chunks and
                               heavy-root parts.  NULL: label statistics and
                               the plan trial instead                        */
  int set_variant;          /* gpm_match_setv: membership probes of the
                               matcher; AUTO lets the autotuner pick per
                               query (BINARY without autotune)               */
  uint32_t pass_budget;     /* autotune = 1: budgeted passes.  A running
                               task checks every pass_budget search steps
                               (windows + descents) whether the launch's task
                               queue is drained; then it stops and is counted
                               again, cut one level finer, in a later launch
                               of the same call.  0 = automatic: long
                               batches run in one launch; a short batch runs
                               passes of 512 steps on its first call, one
                               launch on its second (an identical call
                               replays the prepared batch), and the faster of
                               the two afterwards.  0xFFFFFFFF = always one
                               launch.  Counts are the same for every value  */
} gpm_tuning;

/*
 * Labeled subgraph matching: number of injective, label-preserving,
 * edge-preserving maps of each query into a SYMMETRIC labeled data graph
 * (non-induced; automorphisms counted) -- the count the reference's CPU CECI
 * matcher prints as "#Embeddings" (sub_ceci/myceci.cpp:540-668, 1115).
 * All queries run in one persistent launch; per_query[i] is query i's count.
 * tuning may be NULL: autotune = 1, policy DEGREE, chunk 0, use_graph = 0.
 */
int gpm_match_batch(gpm_graph* g, const gpm_query* queries, int n_queries,
                    const gpm_tuning* tuning, uint64_t* per_query);

/*
 * Training data of the learned matching model (no counterpart in the
 * reference, whose QueryingOrder.h is empty).  For the plan `policy` picks
 * for one query (no plan trial): work[v] = weighted DFS steps (windows and
 * descents weighed by their dependent loads) the exact miner spends below
 * root candidate v, 0 for other vertices (n entries); feat[4v .. 4v+3] = the
 * model's features of root candidate v (degree, depth-1 label-slice length,
 * degrees summed over that slice, query size), zeros elsewhere (4n).
 * Both currently return GPM_EUNSUPPORTED: the root log stopped responding
 * once on the device and stays switched off until the cause is found.
 */
int gpm_match_root_work(gpm_graph* g, const gpm_query* query, int policy, uint64_t* work);
int gpm_match_root_features(gpm_graph* g, const gpm_query* query, int policy, float* feat);

/* ------------------------------------------------------- cost model (MLP) */

/*
 * Per-vertex features for the work estimator, computed on the device:
 * feat[4*v + {0,1,2,3}] = {deg(v), out-degree d+(v) in dag, 2-hop size
 * bound, label (0 without labels)}.  The 2-hop bound is sum_{u in N+(v)} d+(u)
 * when dag is given (the entries a k-clique root streams) and
 * sum_{u in N(v)} deg(u) when dag is NULL (then d+(v) reads 0).
 */
int gpm_vertex_features(gpm_graph* sym, gpm_graph* dag, float* feat);

/*
 * 3-layer MLP (4 -> 64 -> 64 -> 1, ReLU) on BF16 matrix cores with FP32
 * accumulation: out[v] = W3 relu(W2 relu(W1 x_v + b1) + b2) + b3, x_v being
 * the 4 features of v after log1p and the given affine normalisation.
 * Weights are FP32 row-major (W1 64x4, W2 64x64, W3 1x64) and rounded to BF16.
 */
int gpm_mlp_predict(const float* feat, uint32_t n, const float* norm_mean,
                    const float* norm_std, const float* w1, const float* b1,
                    const float* w2, const float* b2, const float* w3,
                    const float* b3, float* out);

#ifdef __cplusplus
}
#endif
#endif /* GPM_H_ */
