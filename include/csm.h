/* C ABI of the MI355X continuous-subgraph-matching library (libcsm.so).
 *
 * The reference exposes this path only as the C++ class API of graph/graph.{h,cpp},
 * graph/plan.{h,cpp} and graph/match_gpu.{h,cu}, driven by TechniquePermutationProcess
 * (technique_permutation.h:29-225) and the csmg command line (main.cpp:162-189).  Each entry point
 * below names the reference interface it replaces.  Conventions: opaque session handle, plain
 * pointers and sizes, status codes instead of exit(), one GPU per session, calls on one session
 * are serialised by the caller.  csm_last_error() returns the message of the last failure.
 */
#ifndef CSM_H_
#define CSM_H_

/* Records at a Cartesian-product depth at which the semi-BFS levels stop expanding and the
 * Cartesian-product flattening takes over, when csm_options.semi_bfs_threshold is 0: the
 * reference's compiled-in 65536 (config.h:47 MIN_NUM_RESULTS_TO_GPU).  Results do not depend on
 * it; the autotuner (gcsm-bu_amd/autotune.py) may pick a larger value per workload, which moves
 * work from flattening to further levels. */
#define CSM_DEFAULT_SEMI_BFS_THRESHOLD 65536u

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csm_session csm_session;

typedef enum {
    CSM_OK = 0,
    CSM_ERR_ARGUMENT = 1,
    CSM_ERR_IO = 2,          /* missing or unreadable file (reference: exit, graph.cpp:198-202) */
    CSM_ERR_FORMAT = 3,      /* malformed file, vertex line in an update file (graph.cpp:502-506) */
    CSM_ERR_LIMIT = 4,       /* query over 16 vertices / 42 edges (graph.cpp:275-280), key width */
    CSM_ERR_DEVICE = 5,      /* HIP runtime failure (reference: cudaErrorCheck -> exit) */
    CSM_ERR_MEMORY = 6,      /* device pool budget or adjacency arena exhausted (match_gpu.cu:772-775) */
    CSM_ERR_STATE = 7,       /* call out of order (e.g. plan before build) */
    CSM_ERR_PRECONDITION = 8,/* batch inserts an existing edge or deletes a missing one */
    CSM_ERR_MISMATCH = 9     /* csm_tune: two launch settings gave different results (a library fault) */
} csm_status;

typedef struct {
    int32_t device;          /* main.cpp:168 --device */
    int32_t directed;        /* extension: arcs src -> dst only (the reference is undirected) */
    uint64_t pool_bytes;     /* HBM budget for the device pool, 0 = no cap */
    uint64_t ring_bytes;     /* partial-result ring of each lane (config.h:39 RES_SPACE), 0 = 1 GiB */
    uint64_t cp_list_bytes;  /* Cartesian-product leaf lists of each lane (config.h:43 SIZE_SPACE), 0 = 512 MiB */
    uint32_t max_batch_edges;/* largest batch the scratch is sized for, 0 = the stream's batch */
    uint32_t semi_bfs_threshold; /* config.h:47 MIN_NUM_RESULTS_TO_GPU, 0 = CSM_DEFAULT_SEMI_BFS_THRESHOLD */
    uint32_t lanes;          /* extension: concurrent step lanes (each with its own ring and scratch; grouped
                                in teams, csm_tunables.team), 0 = 32, at most 64 */
} csm_options;

/* TechniqueOption (utils/technique_option.h:37-56); values as the csmg flags. */
typedef struct {
    int32_t matching_order;          /* 0 indexing, 1 Gamma, 2 GCSM (default), 3 RI */
    int32_t search_strategy;         /* 0 DFS w/ stealing, 1 DFS w/o, 2 dynamic semi-BFS, 3 BFS */
    int32_t local_index;             /* 1 / 0 */
    int32_t cartesian_product;
    int32_t heavy_vertex_avoidance;
    int32_t relation_switch;
} csm_technique;

/* Per query per batch: the reference prints only the positive count, summed over batches
 * (technique_permutation.h:224); deletions and hashes are extensions. */
typedef struct {
    uint64_t positive;
    uint64_t negative;
    /* Embedding hash H(f) = fmix64(sum over query vertices q of fmix64(((f(q) << 4) | q) +
     * 0x9E3779B97F4A7C15)), fmix64 = MurmurHash3's 64-bit finaliser; the two hashes are sums of
     * H(f) over the new / destroyed embeddings, all mod 2^64. */
    uint64_t positive_hash;
    uint64_t negative_hash;
} csm_query_result;

typedef struct {
    uint64_t edges, inserts, deletes;
    double ms_total, ms_index, ms_match;
} csm_batch_info;

/* MatchGPU constructor + cudaSetDevice (match_gpu.cu:84-128, main.cpp:179). */
csm_status csm_session_create(const csm_options* options, csm_session** out);
void csm_session_destroy(csm_session* s);
const char* csm_last_error(const csm_session* s);

/* GraphFileIoManager::LoadInitial (graph.cpp:359-468). */
csm_status csm_load_graph_file(csm_session* s, const char* path);
/* GraphFileIoManager(query_path) + SetQueryMeta (graph.cpp:191-357).  Several queries may stand. */
csm_status csm_load_query_file(csm_session* s, const char* path, uint32_t* query_id);
/* GraphFileIoManager::LoadUpdate (graph.cpp:482-563); batch_size as --batch_size (main.cpp:172). */
csm_status csm_open_update_file(csm_session* s, const char* path, uint32_t batch_size, uint64_t* num_batches);
/* BuildTries(initial) + AllocRelations + UpdateGlobalIndex offline + GetSummary
 * (technique_permutation.h:62-97, match_gpu.cu:336-377, 589-834). */
csm_status csm_build_index(csm_session* s);
/* PlanManager::GenerateIndexingOrders + GenerateMatchingOrders (plan.cpp:185-193, 2117-2146). */
csm_status csm_plan(csm_session* s, uint32_t query_id, const csm_technique* technique);
/* PlanManager::PrintOrders text (plan.cpp:1200-1291).  *needed receives the size incl. NUL. */
csm_status csm_get_plan_text(csm_session* s, uint32_t query_id, char* buf, size_t cap, size_t* needed);
/* MatchGPU::GetSummary output: cardinalities and average degrees per relation (match_gpu.cu:589-609). */
csm_status csm_get_summary(csm_session* s, uint32_t query_id, uint32_t* cardinalities, float* avg_degrees, uint32_t cap,
                           uint32_t* num_relations);
/* One batch of the update file: the per-batch body of technique_permutation.h:127-215
 * (BuildTries, UpdateGlobalIndex, BuildLocalIndex_v2 / GammaBuildLocalIndex, Match per query edge). */
csm_status csm_apply_batch(csm_session* s, uint64_t batch_index, csm_query_result* results, uint32_t num_queries,
                           csm_batch_info* info);
/* Extension: the update-file edges [first, first + n) as one batch, so batches of any size can
 * run over the resident stream (the reference splits the file into fixed batches only,
 * graph.cpp:511-516).  Same semantics and results as csm_apply_batch. */
csm_status csm_apply_range(csm_session* s, uint64_t first, uint64_t n, csm_query_result* results, uint32_t num_queries,
                           csm_batch_info* info);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): csm_apply_range with every
 * lane's partial-result ring taken from the library's static device ring (256 MiB per GPU, split
 * equally over the session's lanes) instead of the session's pool; the call allocates nothing.
 * Same results as csm_apply_range.  Limits: at most 2^23 / lanes edges per call (262,144 with
 * the default 32 lanes), refused with CSM_ERR_LIMIT before any device work; a step whose first two
 * matching-order vertices are not its anchor edge needs its initial partial results to fit one
 * slice, else CSM_ERR_MEMORY.  All sessions on one GPU share the static ring: calls that overlap
 * on one GPU race. */
csm_status csm_apply_range_static(csm_session* s, uint64_t first, uint64_t n, csm_query_result* results, uint32_t num_queries,
                                  csm_batch_info* info);
/* Extension (serving; no counterpart in the reference, whose batches come only from the update
 * file, graph.cpp:482-563): one batch of n edges from host arrays; op[k] = 0 inserts, 1 deletes;
 * labels are raw edge labels already known to the session.  n <= csm_options.max_batch_edges (set
 * it when opening a session for in-memory batches; it also reserves arena room for 16 such
 * batches).  Admission: a batch whose grown lists do not fit the adjacency arena is rejected with
 * CSM_ERR_MEMORY before any device work, leaving the session as it was. */
csm_status csm_apply_edges(csm_session* s, const uint32_t* src, const uint32_t* dst, const uint32_t* elabel, const uint8_t* op,
                           uint64_t n, csm_query_result* results, uint32_t num_queries, csm_batch_info* info);
/* Extension (serving; no counterpart in the reference, whose batches come only from the update
 * file and run one after another, technique_permutation.h:127-215): continuous batching of
 * in-memory batches.  csm_serve_start starts the session's request loop (the session is built and
 * every query planned; set csm_options.max_batch_edges).  csm_serve_submit copies one batch (same
 * arrays as csm_apply_edges; they may be reused on return) and returns its ticket.  Requests apply
 * in submission order; while one runs on the GPU, the next is validated, keyed and copied to the
 * device on a copy stream into the other of two staging slots, and a request enters a slot as soon
 * as one is free; the adjacency arena then admits it (reusing the blocks earlier relocations
 * freed) or refuses it with CSM_ERR_MEMORY, changing nothing.  csm_serve_wait blocks until the
 * request is done and returns its results (equal to csm_apply_edges on the same sequence of
 * batches) and its latency from submission to completion.  While serving, the other batch calls
 * and snapshots return CSM_ERR_STATE.  csm_serve_stop finishes every submitted request and stops
 * the loop (session destruction stops it too). */
csm_status csm_serve_start(csm_session* s);
csm_status csm_serve_submit(csm_session* s, const uint32_t* src, const uint32_t* dst, const uint32_t* elabel, const uint8_t* op,
                            uint64_t n, uint64_t* ticket);
csm_status csm_serve_wait(csm_session* s, uint64_t ticket, csm_query_result* results, uint32_t num_queries, csm_batch_info* info,
                          double* latency_ms);
csm_status csm_serve_stop(csm_session* s);
/* Extension: adjacency arena words in use (the blocks of every current list; blocks that relocated
 * lists left behind are reused by later batches and do not count) and capacity. */
csm_status csm_arena_usage(csm_session* s, uint64_t* used_words, uint64_t* capacity_words);
/* Per-relation edge counts of the reference's grouping (GraphFileIoManager::PrintNumDataEdges,
 * graph.cpp:470-480): batch_index = UINT64_MAX for the initial graph. */
csm_status csm_get_relation_edge_counts(csm_session* s, uint32_t query_id, uint64_t batch_index, uint64_t* counts, uint32_t cap,
                                        uint32_t* num_relations);
/* Number of data vertices (DV_COUNT, graph.cpp:464) and of standing queries. */
csm_status csm_get_counts(csm_session* s, uint32_t* num_vertices, uint32_t* num_queries);
/* Device pool usage (HBM budget accounting). */
csm_status csm_pool_usage(csm_session* s, uint64_t* used, uint64_t* capacity);
/* Read-back for parity tests: candidate bitmap of query vertex u (global index), ceil(dv/32) words. */
csm_status csm_read_candidates(csm_session* s, uint32_t query_id, uint32_t u, uint32_t* words, uint32_t cap);
/* Read-back for parity tests: current neighbour ids of v (sorted by label run, then id). */
csm_status csm_read_adjacency(csm_session* s, uint32_t v, uint32_t direction, uint32_t* ids, uint32_t cap, uint32_t* degree);
csm_status csm_synchronize(csm_session* s);
/* Extension (no counterpart in the reference, whose graph only moves forward): a device copy of the
 * graph state between batches (adjacency lists, label-slot directory, candidate bitmaps), and its
 * restoration, after which the session behaves as if the batches applied since the snapshot had
 * never run.  Lets a benchmark or a serving replica replay an update stream without reloading the
 * data graph.  The copy is one device allocation outside the pool (about the adjacency arena's
 * size), reused by later snapshots.  restore also clears a failed session's refusal state. */
csm_status csm_snapshot(csm_session* s);
csm_status csm_restore(csm_session* s);

/* Launch settings that change scheduling only, never results.  The reference fixes them at compile
 * time: the semi-BFS switch (config.h:47 MIN_NUM_RESULTS_TO_GPU, match_gpu.cu:1961) and the grid
 * shapes of its enumeration launches (match_gpu.cu:1721-2294).  In csm_set_tunables a field of 0
 * keeps the current value; settings change between batches (not while serving). */
typedef struct {
    uint32_t semi_bfs_threshold; /* records at a Cartesian-product depth at which the semi-BFS levels stop
                                    expanding (csm_options.semi_bfs_threshold) */
    uint32_t waves_per_anchor;   /* depth-first counting: most waves one record's first-depth candidates are
                                    split over, in whole 64-candidate rounds (1 = one wave per record;
                                    1..65535, default 1024) */
    uint32_t stack_depth;        /* depth-first counting: most LDS stack rows (depths below the first) one
                                    launch descends; deeper searches run breadth-first levels first
                                    (1..16, default 16) */
    uint32_t dfs_blocks;         /* blocks of the persistent depth-first grid (64..16384, default 2048) */
    uint32_t local_blocks;       /* blocks of the local-index kernels (16..2048; default 512 for batches of
                                    2^18 edges or more, else 256) */
    uint32_t team;               /* lanes per team (1..64, default a quarter of the lanes, so four teams, one per
                                    hardware queue of the runtime's default four): a team runs that many steps at a time on
                                    one stream from one host thread, each stage for all of them before one
                                    wait, and their semi-BFS level chains as one chain of multi-step launches */
    uint32_t level_blocks;       /* blocks per step of a team's level kernels (64..4096, default 384) */
} csm_tunables;
csm_status csm_get_tunables(csm_session* s, csm_tunables* out);
csm_status csm_set_tunables(csm_session* s, const csm_tunables* t);
/* Autotuner (no counterpart in the reference).  Times launch settings on the resident update
 * stream's edges [first, first + (warmup + batches) * batch_edges), in batches of batch_edges (0:
 * the update file's batch size; the last batch may end early at the end of the stream), replayed
 * from a snapshot taken on entry (it replaces any earlier snapshot); the batches after the warm-up
 * ones are timed.  With candidates (fields of 0 keep the session's setting) every one is
 * evaluated; with num_candidates = 0 a coordinate search runs from the current settings (each
 * setting in turn over built-in values, the others held at the best so far).  Every evaluation's
 * per-batch results must equal the first one's, else CSM_ERR_MISMATCH.  On return the graph is as
 * it was on entry and the session holds the fastest settings; evaluated[k] / ms_per_batch[k] (for
 * k < cap, either may be NULL) list the evaluations in order, *num_evaluated their number and *best
 * the index of the fastest. */
csm_status csm_tune(csm_session* s, uint64_t first, uint64_t batch_edges, uint32_t warmup, uint32_t batches,
                    const csm_tunables* candidates, uint32_t num_candidates, csm_tunables* evaluated, double* ms_per_batch,
                    uint32_t cap, uint32_t* num_evaluated, uint32_t* best);

/* Measurement (no counterpart in the reference, whose timing is the TIME_* macros of
 * utils/cuda_helpers.h:78-102).  Kernel families: 0 depth-first enumeration (count), 1 depth-first
 * enumeration (write partial results), 2 Cartesian product, 3 local index, 4 batch build, 5 apply
 * batch, 6 Cartesian-product preparation, 7 anchors, 8 one-level extension (count), 9 one-level
 * extension (write a level), 10 one-level preparation (run lengths and their scan), 11 local-index
 * semi-join check, 12 merge of a batch's insertions into the adjacency lists (family 3 is the
 * local-index expansion, family 5 the whole graph update).  The bytes of
 * families 8 and 9 include the runs their preparation sizes.  Bit f of mask times family f with HIP events on the session's stream and,
 * for the enumeration kernels, counts algorithmic bytes (every view run read once plus candidates
 * written). */
csm_status csm_profile_enable(csm_session* s, uint32_t mask);
csm_status csm_profile_read(csm_session* s, uint32_t family, uint64_t* launches, double* ms, uint64_t* bytes);
/* Times a plain device copy of `words` 32-bit words (calibration of HBM traffic counters). */
csm_status csm_calibration_copy(csm_session* s, uint64_t words, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* CSM_H_ */
