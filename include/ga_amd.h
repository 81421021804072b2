/* ga_amd.h - C ABI of the MI355X-native triangle-count / triangle-density peeling / (k,triangle)-core library.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns an int status (GA_OK = 0). Host
 * buffers are caller-owned with explicit lengths; device memory is owned by the handles and freed by *_destroy.
 * A context owns one HIP stream on one device; a handle is used by one host thread at a time.
 *
 * Each declaration names the reference interface it replaces (file:line in graph_analyzer's src/). The reference binds
 * these as C++ functions and backend methods; the C++ headers in graph_analyzer_amd/host/ restate that API on top of
 * this ABI. */
#ifndef GA_AMD_H
#define GA_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GA_OK = 0,
    GA_ERR_ARG = 1,    /* bad argument (sizes, ids >= n, duplicate pairs where unique pairs are required, ...) */
    GA_ERR_HIP = 2,    /* HIP runtime error; the message names the call */
    GA_ERR_OOM = 3,    /* device allocation failed */
    GA_ERR_ARITH = 4,  /* an exact-arithmetic bound does not hold (peel_arith_ok, T >= 2^32 for u32 counters) */
    GA_ERR_STATE = 5,  /* device-side sanity check failed (removed 0 / more than alive, level decreased, T mismatch) */
};

typedef struct ga_ctx ga_ctx;       /* device + stream + scratch */
typedef struct ga_graph ga_graph;   /* oriented CSR resident in HBM (+ optional undirected CSR, count plan) */
typedef struct ga_peel ga_peel;     /* batch-peeling state of one graph */
typedef struct ga_core ga_core;     /* (k,triangle)-core state of one graph */
typedef struct ga_batch ga_batch;   /* many graphs resident together, counted / peeled in shared launches */
typedef struct ga_truss ga_truss;   /* k-truss decomposition state of one graph */

/* ---- context ------------------------------------------------------------------------------------------------ */
/* Replaces the driver set-up of gpu_tc.cu:83-85 (device properties) and the implicit default stream. */
int ga_ctx_create(int device, ga_ctx** out);
void ga_ctx_destroy(ga_ctx* ctx);
const char* ga_ctx_last_error(const ga_ctx* ctx);
int ga_ctx_sync(ga_ctx* ctx);
/* name (NUL-terminated, truncated to len), compute units, warp (wave) size, total HBM bytes */
int ga_ctx_device_info(ga_ctx* ctx, char* name, int len, int* cus, int* wave, uint64_t* hbm_bytes);
/* Copies `bytes` (rounded down to 16) between two fresh device buffers with one plain copy kernel (k_copy_probe):
 * the known-size reference that scales hardware traffic counters (rocprofv3 FETCH_SIZE / WRITE_SIZE).
 * No counterpart in the reference (measurement support). */
int ga_ctx_copy_probe(ga_ctx* ctx, uint64_t bytes, float* ms);
/* Random-segment gather (k_gather_probe): *segments reads of seg_bytes (32, 64, 128 or 256), each from a random
 * aligned offset in a fresh buffer of `bytes`, 8 independent reads in flight per lane group; ms = kernel time. The
 * measured ceiling of the count's access pattern (short probe lists at random places). No counterpart in the
 * reference (measurement support). */
int ga_ctx_gather_probe(ga_ctx* ctx, uint64_t bytes, uint32_t seg_bytes, uint64_t* segments, float* ms);
/* What a device-wide barrier costs inside a cooperative kernel of `workgroups` resident workgroups of 256 threads
 * (two-level arrival counters, agent-scope release/acquire, bounded wait), and what one launch of an equally shaped
 * trivial kernel costs back to back on the stream: us_per_barrier, us_per_launch, each averaged over `iters`. The
 * measurement behind running the core's rounds as separate launches. No counterpart in the reference (measurement
 * support). */
int ga_ctx_barrier_probe(ga_ctx* ctx, uint32_t workgroups, uint32_t iters, float* us_per_barrier, float* us_per_launch);
/* Test hook: every count, participation, support, peel and core launch of this context takes at most max_workgroups
 * workgroups (0: the library's own grids), so that grid-stride and item loops run many passes. Replaces the tiny
 * grids (1, 3 and 7 blocks) the reference's CPU emulator runs every kernel on (emu_main.cpp:40-57). Results must not
 * depend on it. */
int ga_ctx_set_grid_cap(ga_ctx* ctx, uint32_t max_workgroups);
/* Free and total device memory (hipMemGetInfo): the serving loop's admission budget. */
int ga_ctx_mem_info(ga_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* ---- graphs (graph.h:19-96, densest.h:162-171) -------------------------------------------------------------- */
/* n < 0: load_graph semantics on a raw pair list (graph.h:71-96 after the text parse): self loops dropped, pairs
 *        canonicalised to (min,max) and de-duplicated, n = max_id + 1 (1 for an empty list).
 * n >= 0: build_graph_from_edges semantics (graph.h:38-67): pairs must be unique, loop-free, ids < n; violations
 *        return GA_ERR_ARG.
 * Both: (degree, id) orientation, rows sorted ascending, src[] and max_out_degree, all on the GPU.
 * build_ms (may be NULL): device time of the build, excluding the host-to-device copy of the pairs. */
int ga_graph_from_pairs(ga_ctx* ctx, const uint32_t* pairs, uint64_t npairs, int64_t n, ga_graph** out, float* build_ms);
/* Upload of an oriented CSR built elsewhere (gpu_tc.cu:42-62 upload of rowptr/col/src); src is derived on device. */
int ga_graph_from_csr(ga_ctx* ctx, uint32_t n, uint32_t m, const uint32_t* rowptr, const uint32_t* col, ga_graph** out);
void ga_graph_destroy(ga_graph* g);
int ga_graph_info(const ga_graph* g, uint32_t* n, uint32_t* m, uint32_t* max_out_degree);
/* Download of the Graph fields (graph.h:19-26); any pointer may be NULL. Sizes n+1, m, m. */
int ga_graph_download(ga_ctx* ctx, const ga_graph* g, uint32_t* rowptr, uint32_t* col, uint32_t* src);
/* Both-direction sorted adjacency on the device (build_undirected_csr, densest.h:163-171); optional download of
 * urow (n+1) and ucol (2m). */
int ga_graph_undirected(ga_ctx* ctx, ga_graph* g, uint32_t* urow, uint32_t* ucol);
/* Work order of edges by floor(log2(min(d+u,d+v)+1)), heaviest bucket first, stable (build_edge_order,
 * graph.h:156-171), computed on the device; perm has m entries. */
int ga_graph_edge_order(ga_ctx* ctx, ga_graph* g, uint32_t* perm);

/* build_tasks (graph.h:138-150) on the device: every oriented edge's work s = min(d+(u), d+(v)) cut into chunks of
 * `chunk` elements, task i of edge e being (t_edge = e, t_start = chunk * i), edges in order. *ntasks gets the number of
 * tasks; with t_edge and t_start NULL the call only counts, otherwise both need `capacity` >= *ntasks entries
 * (GA_ERR_ARG if not). chunk = 0 is GA_ERR_ARG. */
int ga_graph_tasks(ga_ctx* ctx, ga_graph* g, uint32_t chunk, uint32_t* t_edge, uint32_t* t_start, uint64_t capacity, uint64_t* ntasks);

/* load_graph (graph.h:71-96) with the text parse on the GPU as well: `bytes` holds a SNAP edge list (len bytes, host
 * memory). Lines are taken as the reference's fgets(line, 512) loop sees them (physical lines cut into chunks of at most
 * 511 bytes), chunks starting with '#', '%' or '\n' or without two base-10 numbers (strtoul rules: leading white space,
 * optional sign, saturation at ULONG_MAX) are skipped, a == b is dropped, ids are cast to uint32; then the n < 0 build of
 * ga_graph_from_pairs. Equal to the host parse followed by ga_graph_from_pairs on every input. ms (may be NULL): device
 * time from the first parse kernel to the finished CSR (the upload of the text excluded). */
int ga_graph_from_text(ga_ctx* ctx, const char* bytes, uint64_t len, ga_graph** out, float* ms);
/* ga_graph_from_text on the whole content of the file at `path`, streamed to the device through two pinned 64 MiB
 * buffers (the host never holds the whole file). */
int ga_graph_load_text(ga_ctx* ctx, const char* path, ga_graph** out, float* ms);
/* Binary CSR file of this library (no reference counterpart; the fast admission path next to the text loader): a 64-byte
 * header {magic "GACSR001", n, m, max_out_degree as u64, 32 reserved bytes}, rowptr u32[n+1], col u32[m]. */
int ga_graph_save(ga_ctx* ctx, const ga_graph* g, const char* path);
/* Loads a ga_graph_save file. The device checks that the arrays are exactly what the builder makes for their edge set
 * (rowptr from 0 to m and never decreasing, col < n, rows strictly ascending, every edge oriented by the (degree, id)
 * rule of graph.h:43-45, max_out_degree as stored); anything else returns GA_ERR_ARG. The file is streamed to the
 * device through two pinned buffers. ms (may be NULL): device time of the checks and the src[] derivation (the upload,
 * overlapped with the file read, is not included). */
int ga_graph_load(ga_ctx* ctx, const char* path, ga_graph** out, float* ms);

/* induced_subgraph (dsg.h:103-114) on the GPU: the subgraph on {v : keep[v] != 0} with new ids in old-id order, rebuilt
 * with build_graph_from_edges semantics (orientation by the degrees inside the subgraph, rows sorted, src[],
 * max_out_degree). keep: n bytes. old_id (may be NULL): capacity n, gets old_id[new] for the n' = ga_graph_info(*out)
 * kept vertices. ms (may be NULL): device time from the first kernel to the finished CSR. */
int ga_graph_induced(ga_ctx* ctx, const ga_graph* g, const uint8_t* keep, ga_graph** out, uint32_t* old_id, float* ms);
/* prune_to_core (dsg.h:636-645) on the GPU: ga_graph_induced on {v : core[v] >= k}; core (host, n entries) are the
 * graph's core numbers; core_out (may be NULL, capacity n) gets core[old_id[i]], which are the pruned graph's own core
 * numbers (the containment argument of dsg.h:629-635). */
int ga_graph_prune_to_core(ga_ctx* ctx, const ga_graph* g, const uint32_t* core, uint64_t k, ga_graph** out, uint32_t* old_id,
                           uint32_t* core_out, float* ms);

/* ---- triangle counting (kernels.cuh:52-358, gpu_tc.cu:128-228) --------------------------------------------- */
enum {
    GA_TC_AUTO = 0,        /* the tuned default (currently GA_TC_HASH) */
    GA_TC_MERGE = 1,       /* one lane per oriented edge, two-pointer merge (role of v1) */
    GA_TC_GROUP4 = 2,      /* 4 lanes per edge, strided binary search of the shorter list (role of v5) */
    GA_TC_GROUP8 = 3,
    GA_TC_GROUP16 = 4,
    GA_TC_GROUP64 = 5,     /* a full wave per edge (role of v2) */
    GA_TC_HASH = 6,        /* longer list of each edge in an LDS hash set, shorter list streamed by lane groups
                              sized per degree bucket (roles of v3/v8 generalised, v4/v9 work ordering on device) */
    GA_TC_SORTED = 7,      /* one lane per edge in the device-built work order of ga_graph_edge_order (role of v9) */
    GA_TC_NOSRC = 8,       /* 4 lanes per edge without src[]: each edge's source recovered from rowptr, one binary search
                              per wavefront range and then a 64-entry rowptr window in the lanes (role of v7,
                              tc_edge_find_src, kernels.cuh:219-248) */
    GA_TC_VERTEX = 9,      /* vertex-centric without src[]: one lane merges the out-edges of a vertex of out-degree <= 16,
                              a whole wavefront those of a larger one (role of v6, tc_vertex_per_thread,
                              kernels.cuh:195-217) */
    GA_TC_NVARIANTS = 10
};
const char* ga_tc_variant_name(int variant);
/* T = sum over oriented edges (u,v) of |N+(u) ∩ N+(v)| (graph.h:110-121). ms (may be NULL): device time of the
 * counting launch(es), measured with HIP events on the context stream. Preparation of the count plan (GA_TC_HASH,
 * GA_TC_SORTED) happens on first use and is excluded; call ga_tc_prepare to time it. */
int ga_tc_count(ga_ctx* ctx, ga_graph* g, int variant, uint64_t* T, float* ms);
/* Counts `count` resident graphs on ctx's stream with one readback at the end (T[i] for graphs[i]); plans are built
 * first where missing. GA_TC_HASH / GA_TC_AUTO with two or more graphs: one launch per bucket shape covers every
 * graph (a device table of per-graph arguments and a workgroup-to-graph map; each graph runs what its own launch
 * would); graphs whose tuning uses the flat kernel, and other variants, get their own launches. The serving loop's
 * batched count. No counterpart in the reference. */
int ga_tc_count_many(ga_ctx* ctx, ga_graph* const* graphs, uint32_t count, int variant, uint64_t* T, float* ms);
/* ga_tc_count for GA_TC_NOSRC or GA_TC_VERTEX that allocates nothing. No counterpart in the reference: it is for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid allocation). The
 * count is enqueued on ctx's stream and the call returns without waiting; *d_T, a device word the caller owns, holds T
 * once the stream has reached that point (the caller synchronises). The per-workgroup partial sums and a completion
 * ticket live in a static device buffer of the library (GA_NOALLOC_MAX_BLOCKS + 1 words per device, held until the
 * process ends), shared by every call on that device: calls on one device that overlap in time race, so enqueue them
 * on one stream or serialise them. blocks = workgroups (0 = the library's choice); more than GA_NOALLOC_MAX_BLOCKS is
 * GA_ERR_ARG. Other variants (they need src[] or a count plan) are GA_ERR_ARG. */
#define GA_NOALLOC_MAX_BLOCKS 4096u
int ga_tc_count_noalloc(ga_ctx* ctx, const ga_graph* g, int variant, uint32_t blocks, uint64_t* d_T);
/* Builds the device-side count plan (hash-side work lists, degree buckets, work order); ms = its device time. */
int ga_tc_prepare(ga_ctx* ctx, ga_graph* g, float* ms);
/* Lane-group width / hash capacity / block shape per degree bucket (the autotuner's output, see ga_tc_tune). */
typedef struct ga_tc_tuning {
    int32_t lanes[5];      /* lanes per work item for the hashed buckets d+(h) <= 8, 32, 128, 512, hash_cap
                              (each one of 4, 8, 16, 32, 64; bucket 0 also 1: one lane per item with N+(h) in
                              registers, for plain and per-graph counts) */
    int32_t block;         /* threads per workgroup (64..1024, multiple of 64) */
    int32_t work_chunk;    /* probe elements per work item before a row is cut */
    int32_t batches;       /* lane-group batches (lanes probe lists each) per work item before a row is cut */
    int32_t load_shift;    /* hash table slots = next_pow2(d+(h)) << load_shift (1..3) */
    int32_t hash_cap;      /* largest d+(h) kept in an LDS table (<= 4096); larger rows use binary search in HBM */
    int32_t flat_chunk;    /* plain counts: rows with d+(h) <= 128 are counted by the flat kernel in workgroup chunks of
                              this many units (table slots, one per work entry, probe elements); 0 = lane groups only.
                              A multiple of flat_block in [flat_block, 8192]; needs hash_cap >= 128 */
    int32_t flat_block;    /* threads per workgroup of the flat kernel (64..1024, multiple of 64) */
    int32_t item_order;    /* 1: work items of a bucket ordered by shape (entries, longest probe list, d+(h)), so the
                              lane groups of one wavefront run loops of equal length (in a batch union: within each
                              graph, graphs in order); 0: in row order; 2: by shape within windows of 4,096
                              consecutive h ids, windows in id order; 3: as 2, each XCD (workgroups b with the same
                              b mod 8) taking one contiguous eighth of a bucket's items; 4, 5: as 2, 3 with windows of
                              65,536 ids */
    int32_t relabel;       /* 1: plain and per-graph counts run on a private copy of the oriented CSR whose vertices
                              are renumbered so that vertices joined by edges in two or more triangles sit together
                              (each vertex is keyed by the smallest id among itself and such neighbours); the
                              orientation, and so every count, is unchanged. 2: as 1, with the copy's rows placed so
                              that no row of at most 32 ids crosses a 128-byte line (about a fifth more column
                              memory). 0: on the CSR itself. Participation, support and the per-vertex states always
                              use the CSR's own ids */
    int32_t probe;         /* plain and per-graph counts, lane-group buckets: 1 = every probe list of a batch gets
                              ceil(length / 8) lanes and all of them are read in the same round; 0 = the lists are read
                              by length class (lists of length in (S/2, S] together, one round per class). Participation
                              and support always read by length class */
    int32_t waves;         /* plain and per-graph counts: 0 = as many waves per SIMD as registers and tables allow;
                              4..8 = at most this many, held by the workgroups' LDS allocation (fewer items in flight
                              keep more of their probe lists in the L2) */
    int32_t streams;       /* plain and per-graph counts: 1 = the bucket launches of one count run on four streams
                              (the caller's and three of the library's, joined back before the call returns), so
                              that launches overlap; 0 = one after another on the caller's stream. Per-launch
                              profiles and captured streams always use one stream */
} ga_tc_tuning;
int ga_tc_get_tuning(const ga_graph* g, ga_tc_tuning* out);
/* The library's default tuning (what a graph uses until ga_tc_set_tuning or ga_tc_tune). */
int ga_tc_default_tuning(ga_tc_tuning* out);
int ga_tc_set_tuning(ga_ctx* ctx, ga_graph* g, const ga_tc_tuning* t);
/* Count plan statistics: work entries (edges with min(d+) > 0), items per bucket (5 hashed + 1 binary-search). */
int ga_tc_plan_info(ga_ctx* ctx, ga_graph* g, uint32_t* nwork, uint32_t* nitems6);
/* Per-launch profile of the plain count as tuned (no reference counterpart: measurement): for each of the 5 hashed
 * buckets, the binary-search bucket and the flat kernel (slot 6, light rows, when it is on), the work entries, probe
 * elements (sum of min(d+u, d+v) over the entries) and the mean device time of the launch over reps counts. */
int ga_tc_bucket_profile(ga_ctx* ctx, ga_graph* g, int reps, uint64_t* entries7, uint64_t* elems7, float* ms7);
/* Times candidate tunings on this graph (each verified against the first count) and keeps the fastest. */
int ga_tc_tune(ga_ctx* ctx, ga_graph* g, ga_tc_tuning* best, float* best_ms);
/* variant: GA_TC_MERGE or GA_TC_HASH (GA_TC_AUTO = GA_TC_HASH) for the two calls below.
 * Per-vertex triangle participation in the subgraph induced by alive (n bytes; NULL = all alive):
 * cpu_triangle_participation (densest.h:81-101) / tc_per_vertex_alive (kernels.cuh:372-425). tri has n entries. */
int ga_tc_vertex(ga_ctx* ctx, ga_graph* g, int variant, const uint8_t* alive, uint32_t* tri, uint64_t* T);
/* Per-oriented-edge triangle support: sup[e] = triangles containing edge e (indices of for_each_triangle,
 * dsg.h:62-79). sup has m entries. */
int ga_tc_edge_support(ga_ctx* ctx, ga_graph* g, int variant, uint32_t* sup, uint64_t* T);

/* ---- batch peeling: the backend concept of batch_peel_loop (densest.h:265-303, gpu_densest.cu:70-126) ------- */
int ga_peel_create(ga_ctx* ctx, ga_graph* g, ga_peel** out);
/* Moves an idle peel state (no call in flight; creation and every run return with their work finished) to another
 * context: a serving worker runs it on its own stream. */
int ga_peel_set_ctx(ga_peel* p, ga_ctx* ctx);
void ga_peel_destroy(ga_peel* p);
int ga_peel_reset_state(ga_peel* p);                 /* alive = 1, removed_round = UINT32_MAX */
int ga_peel_begin_round(ga_peel* p);                 /* tri = 0, counters = 0 (fill kernels, no memset) */
int ga_peel_run_tc(ga_peel* p);                      /* participation of the alive subgraph */
int ga_peel_run_mark(ga_peel* p, uint64_t nalive, uint64_t num, uint64_t den, uint32_t round);
int ga_peel_read_stats(ga_peel* p, uint64_t* T, uint64_t* removed);   /* the one host sync per round */
int ga_peel_download_removed_round(ga_peel* p, uint32_t* removed_round);
/* Whole peel with the round loop on the device (rounds enqueued in blocks, one host sync per block).
 * stats = {best_num, best_den, best_round, rounds}; T_full = the caller's full count for the round-0 check
 * (UINT64_MAX = skip the check). Errors as batch_peel_loop: GA_ERR_ARITH, GA_ERR_STATE. */
int ga_peel_run(ga_peel* p, uint64_t T_full, uint64_t num, uint64_t den, uint64_t* stats, float* ms);
/* Fused peels of several resident graphs in shared launches, every round's begin, participation count, mark and end
 * covering all unfinished states (each state exactly as ga_peel_run runs it alone; the hash variant). peels: distinct
 * states on ctx's device, moved to ctx. T_full: per-state counts or NULL (UINT64_MAX entries: unknown). stats: 4 per
 * state {best_num, best_den, best_round, rounds}; removed rounds through ga_peel_download_removed_round. The serving
 * loop's batched peel. No counterpart in the reference: it peels one graph per process. */
int ga_peel_run_batch(ga_ctx* ctx, ga_peel* const* peels, uint32_t count, const uint64_t* T_full, uint64_t num, uint64_t den,
                      uint64_t* stats, float* ms);

/* ---- (k,triangle)-core: backends of core_decomp_loop / core_decomp_loop_inc (dsg.h:665-751, gpu_dsg.cu:88-199) - */
int ga_core_create(ga_ctx* ctx, ga_graph* g, int incremental, ga_core** out);
/* Moves an idle core state to another context (its stream); same rule as ga_peel_set_ctx. */
int ga_core_set_ctx(ga_core* c, ga_ctx* ctx);
/* Incremental decompositions of `count` graphs in one launch, one workgroup per graph running all its rounds
 * (core_decomp_loop_inc rules, dsg.h:719-751); for small graphs, where per-round launches would dominate. Each state
 * must be incremental; T_full (count entries, or NULL; UINT64_MAX = unchecked) is checked against the device count;
 * stats gets rounds and kmax per graph; core numbers stay in each state (ga_core_download). No counterpart in the
 * reference (it runs one graph per process). */
int ga_core_run_many(ga_ctx* ctx, ga_core* const* cores, uint32_t count, const uint64_t* T_full, uint32_t* stats, float* ms);
/* Fused incremental decompositions of `count` states in shared launches: each round's find / update / tail launch
 * covers every unfinished state, each on its own range of workgroups, so a state follows exactly the rounds of
 * ga_core_run and gets the same results; the round latency is paid once for the whole batch. For serving several
 * requests of any size at once. Every state must be incremental, belong to ctx (ga_core_set_ctx moves one) and appear
 * once; T_full, stats and errors as ga_core_run_many (an error names the first failing state). No counterpart in the
 * reference (it runs one graph per process). */
int ga_core_run_batch(ga_ctx* ctx, ga_core* const* cores, uint32_t count, const uint64_t* T_full, uint32_t* stats, float* ms);
void ga_core_destroy(ga_core* c);
int ga_core_reset_state(ga_core* c);                 /* alive/state = 1, core = UINT32_MAX */
int ga_core_begin_round(ga_core* c);
int ga_core_run_tc(ga_core* c);                      /* recount only */
int ga_core_run_min(ga_core* c);
int ga_core_run_peel(ga_core* c, uint64_t k);        /* recount only */
int ga_core_read_stats(ga_core* c, uint64_t* T, uint64_t* removed, uint64_t* k);
int ga_core_init_tri(ga_core* c, uint64_t* T);       /* incremental only: the one full count */
int ga_core_run_find(ga_core* c, uint64_t k);        /* incremental only */
int ga_core_run_update(ga_core* c);                  /* incremental only */
int ga_core_run_retire(ga_core* c);                  /* incremental only */
int ga_core_download(ga_core* c, uint32_t* core);
/* prune_to_core (dsg.h:636-645) on the core numbers a finished decomposition holds on the device (no host round trip of
 * core[]; gpu_dsg.cu:318-321 downloads them and prunes on the host). Arguments as ga_graph_prune_to_core; the new graph
 * belongs to the state's context. */
int ga_core_prune(ga_core* c, uint64_t k, ga_graph** out, uint32_t* old_id, uint32_t* core_out, float* ms);
/* Whole decomposition with the level loop on the device; stats = {rounds, kmax}. T_full as in ga_peel_run. */
int ga_core_run(ga_core* c, uint64_t T_full, uint32_t* stats, float* ms);

/* ---- k-truss decomposition (no reference function: it extends the per-edge supports of dsg.h:62-94 as the
 *      (k,triangle)-core extends per-vertex participation, with the same level rule as dsg.h:694-751) --------------- */
/* Allocates the state and builds the undirected adjacency with oriented edge ids on the device. */
int ga_truss_create(ga_ctx* ctx, ga_graph* g, ga_truss** out);
void ga_truss_destroy(ga_truss* t);
/* Whole decomposition on the device: supports by the hash count, then rounds of level = max(level, least present
 * support), removal of every present edge with support <= level (truss = level + 2) and support decrements for the
 * triangles they close; rounds enqueued in blocks with one readback per block. stats = {rounds, kmax} (kmax 0 when
 * m = 0); ms: device time including the support count. */
int ga_truss_run(ga_truss* t, uint32_t* stats, float* ms);
/* truss[e] for every oriented edge e (index into col), m entries; valid after ga_truss_run. */
int ga_truss_download(ga_truss* t, uint32_t* truss);

/* ---- Greedy++ for the triangle-density densest subgraph (Flowless Alg 2, triangles as hyperedges) ---------------- */
/* greedy_pp(g, passes) (dsg.h:189-248) and greedy(g) = passes 1 (dsg.h:250), as gpu_dsg runs Greedy++ on the graph
 * prune_to_core leaves (gpu_dsg.cu:344-356) and on the whole graph (gpu_dsg.cu:318-330). Each deletion takes the alive
 * vertex with the smallest (load + tri_H, id);
 * H is evaluated before every deletion and replaces the best only when strictly denser (exact cross products).
 * Outputs: *tri / *verts = the best density found (0 / 0 when the graph has no triangle, as DsgResult); per_pass
 * (2 * passes words, may be NULL) = the best (tri, verts) after each pass (DsgResult::per_pass); set (capacity n, may be
 * NULL) = the best subgraph's vertices ascending, *set_size of them (DsgResult::set). Up to ga_greedy_max_vertices(ctx)
 * vertices one workgroup keeps the per-vertex state in LDS; above it the state lives in HBM and the minimum comes from
 * two levels of block minima (then T must be below 2^32, else GA_ERR_ARITH, and passes at most 2^19). passes = 0 is
 * GA_ERR_ARG; GA_ERR_STATE when the triangle bookkeeping self-check of dsg.h:230 fails. ms: device time including the
 * per-vertex triangle count. */
int ga_greedy_pp(ga_ctx* ctx, ga_graph* g, uint32_t passes, uint64_t* tri, uint64_t* verts, uint64_t* per_pass, uint32_t* set,
                 uint32_t* set_size, float* ms);
/* The largest vertex count ga_greedy_pp keeps in LDS on this device, and the limit of ga_greedy_pp_many (no counterpart
 * in the reference, whose CPU code has none). */
uint32_t ga_greedy_max_vertices(ga_ctx* ctx);
/* core_exact (dsg.h:454-627, DSD paper Alg 4): the densest subgraph for the triangle density, exactly, by max-flow on
 * the connected components of the pruned cores. core: the (k,triangle)-core numbers of g (host, n entries; from
 * ga_core_download or core_decomposition). use_pruning12 = 0 is the reference's test hook (start from one triangle).
 * seed_tri_verts = {tri, verts} and seed_set (verts ids) of another algorithm's result on g (Greedy++), or NULL: it
 * is checked (ids, no duplicates, triangle count) and then bounds the search from below, as gpu_dsg.cu:348-361 uses it.
 * out7 = {tri, verts, set size, ok, flow_calls, components_searched, shrinks} (DsgResult); set: capacity n, the best
 * subgraph's vertices ascending. The triangle passes and induced subgraphs run on the device, the components and the
 * max-flow (Dinic) on the host. A failed self-check of the reference (inconsistent core numbers, an invalid seed, flow
 * capacities of 2^59 or more, a min cut not denser than alpha) is GA_ERR_STATE with out7[3] = 0 and the reference's
 * message in ga_ctx_last_error. */
int ga_core_exact(ga_ctx* ctx, ga_graph* g, const uint32_t* core, int use_pruning12, const uint64_t* seed_tri_verts,
                  const uint32_t* seed_set, uint64_t* out7, uint32_t* set);
/* exact_densest (dsg.h:402-428, Alg 1): the same search on one network of the whole graph. out7 and set as above. */
int ga_exact_densest(ga_ctx* ctx, ga_graph* g, uint64_t* out7, uint32_t* set);
/* Greedy++ with `passes` passes on each of `count` graphs of this context, graph by graph equal to ga_greedy_pp: one
 * workgroup per graph in one launch for graphs of up to 1,120 vertices (the LDS-bitmap form on MI355X), one launch
 * per graph for larger ones (each at most ga_greedy_max_vertices). stats: 2 per graph, {tri, verts} of the best density
 * (0 / 0 without triangles); in_set (may be NULL): n_0 + ... + n_{count-1} bytes, graph i's vertices at offset
 * n_0 + ... + n_{i-1}, 1 for the vertices of its best set. ms: device time of the whole call. No counterpart in the
 * reference, which runs one graph per process. */
int ga_greedy_pp_many(ga_ctx* ctx, ga_graph* const* graphs, uint32_t count, uint32_t passes, uint64_t* stats, uint8_t* in_set,
                      float* ms);

/* ---- batching: many graphs resident together (no reference counterpart: the reference runs one graph per process;
 *      per-graph semantics are exactly load_graph + the calls above, graph by graph) ----------------------------- */
/* Graph i's raw pairs are pairs[2*pair_offsets[i] .. 2*pair_offsets[i+1]) with load_graph semantics per graph
 * (n_i = max_id + 1, loops dropped, de-duplicated). All graphs are built in one GPU pass; build_ms = host wall time. */
int ga_batch_create(ga_ctx* ctx, uint32_t count, const uint64_t* pair_offsets, const uint32_t* pairs, ga_batch** out, float* build_ms);
void ga_batch_destroy(ga_batch* b);
int ga_batch_info(const ga_batch* b, uint32_t* count, uint64_t* total_n, uint64_t* total_m);
int ga_batch_graph_info(const ga_batch* b, uint32_t i, uint32_t* n, uint32_t* m, uint32_t* max_out_degree);
/* Graph i's oriented CSR with local ids (equals ga_graph_download of the graph built alone). */
int ga_batch_download_graph(ga_ctx* ctx, ga_batch* b, uint32_t i, uint32_t* rowptr, uint32_t* col, uint32_t* src);
/* Per-graph T (count entries) from one shared launch set; ms = host wall time from launch to result. */
int ga_batch_count(ga_ctx* ctx, ga_batch* b, uint64_t* T, float* ms);
/* The batch's algorithmic work and count time, for its roofline: *summin (may be NULL) = sum over every graph's oriented
 * edges of min(d+(u), d+(v)), computed on the device, so the batch's algorithmic bytes are 4 (total_n + count) +
 * 4 total_m + 4 *summin; *last_count_device_ms (may be NULL) = device time of the last ga_batch_count's launches (HIP
 * events on ctx's stream; 0 before the first). No counterpart in the reference (measurement support). */
int ga_batch_work(ga_ctx* ctx, ga_batch* b, uint64_t* summin, float* last_count_device_ms);
/* The same counts with one ga_tc_count per graph (the per-graph-launch baseline). */
int ga_batch_count_per_graph(ga_ctx* ctx, ga_batch* b, uint64_t* T, float* ms);
/* Autotunes the union's count (ga_tc_tune on the resident union); the per-graph baseline inherits the result. */
int ga_batch_tune(ga_ctx* ctx, ga_batch* b, ga_tc_tuning* best, float* best_ms);
/* Sets the union's count tuning; the per-graph baseline uses the same record (ga_tc_set_tuning on each graph). */
int ga_batch_set_tuning(ga_ctx* ctx, ga_batch* b, const ga_tc_tuning* t);
int ga_batch_get_tuning(const ga_batch* b, ga_tc_tuning* out);
/* Batch peeling of every graph with shared launches per round. stats: 4 per graph {best_num, best_den, best_round,
 * rounds}; removed_round: total_n entries, graph i at its vertex offset (or NULL); T_full: per-graph counts for the
 * round-0 check (or NULL). */
int ga_batch_peel(ga_ctx* ctx, ga_batch* b, uint64_t num, uint64_t den, const uint64_t* T_full, uint64_t* stats,
                  uint32_t* removed_round, float* ms);
/* Peel rounds per captured HIP graph in ga_batch_peel (default 8): the first call's block of rounds is captured once
 * and replayed until every graph is done, with one readback per 1, 2, 4, 8, 8, ... replays. 0 enqueues every round
 * directly (one readback per 1, 2, 4, ..., 64 rounds). Results are identical either way. No counterpart in the
 * reference, which syncs once per round (densest.h:279-303). */
int ga_batch_set_peel_graph(ga_batch* b, int rounds);
/* The same with one ga_peel_run per graph (the per-graph-launch baseline). */
int ga_batch_peel_per_graph(ga_ctx* ctx, ga_batch* b, uint64_t num, uint64_t den, const uint64_t* T_full, uint64_t* stats,
                            uint32_t* removed_round, float* ms);

#ifdef __cplusplus
}
#endif
#endif
