// ga_internal.hpp - shared host structures, error plumbing and wave64 device helpers of libga_amd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ga_amd.h"

// ---------------------------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------------------------
struct ga_ctx {
    int device = 0;
    int cus = 0;
    size_t lds_max = 65536;             // dynamic LDS bytes one workgroup may use
    size_t lds_cu = 163840;             // LDS bytes of one CU (shared by its resident workgroups)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t side[3] = {nullptr, nullptr, nullptr};   // tune.streams: the count's bucket launches besides `stream`
    hipEvent_t fork_ev = nullptr, join_ev[3] = {nullptr, nullptr, nullptr};   //   (default priority, created on first use)
    uint64_t* hstat = nullptr;          // pinned host staging for per-round statistics (512 bytes)
    uint64_t* dsmall = nullptr;         // 64 device words for counters of one call (never reallocated)
    void* scratch = nullptr;            // grow-only device scratch for library temporaries (sort / scan storage)
    size_t scratch_bytes = 0;
    uint64_t* dmany = nullptr;          // grow-only per-graph counts of ga_tc_count_many
    size_t dmany_n = 0;
    void* dcore = nullptr;              // grow-only descriptors and results of ga_core_run_many
    size_t dcore_bytes = 0;
    void* dtcm = nullptr;               // grow-only job tables and workgroup maps of ga_tc_count_many's shared launches
    size_t dtcm_bytes = 0;
    std::vector<uint8_t> htcm;          // their host copy: left untouched until the next call (each call ends with a sync)
    std::vector<uint8_t> hpeel;         // host copy of ga_peel_run_batch's job tables and maps (same rule)
    uint32_t grid_cap = 0;              // test hook: at most this many workgroups per count / peel / core launch (0: none)
    std::string err;
};

// Device-resident count plan of GA_TC_HASH: work items grouped by the endpoint whose out-list is hashed.
struct TcPlan {
    bool ready = false;
    uint32_t nitems[5] = {0, 0, 0, 0, 0};   // work items per hashed bucket
    uint32_t nhuge = 0;                  // work items whose d+(h) is too large for an LDS table (binary search)
    uint32_t* items = nullptr;           // 4 words per item {h, begin, rowptr[h], entries | d+(h) << 17}, bucket-major
    uint32_t* wcol = nullptr;            // probe-side vertex per work entry (rows of the hashed vertex, longest first)
    uint32_t* wedge = nullptr;           // oriented edge id of each work entry
    uint2* wpl = nullptr;                // {rowptr[p], d+(p)} of each work entry: start and length of its probe list, one
                                         // 8-byte load (rowptr: the row starts of the column array the counts read)
    const uint32_t* st_used = nullptr;   // the row starts that items[].z and wpl[].x hold
    uint32_t nwork = 0;                  // work entries (edges with min(d+) > 0)
    bool entries = false;                // wcol/wedge/wpl/wh hold the sorted work entries (they depend on the CSR only)
    uint32_t* wh = nullptr;              // hashed vertex h of each work entry
    uint32_t nlight = 0;                 // entries [0, nlight) have d+(h) <= 128 (the flat kernel's class)
    unsigned long long* wu = nullptr;    // flat unit offset of each light entry (nlight + 1, exclusive prefix)
    uint32_t* cstart = nullptr;          // light entry holding the first unit of each flat chunk (nch + 1)
    uint32_t* frec = nullptr;            // 4 words per light entry: the flat kernel's entry record
    uint32_t nch = 0, chunk_used = 0;    // flat chunks and the chunk size they were cut with
    unsigned long long utot = 0;         // flat units of all light entries
    ga_tc_tuning tune = {};             // block == 0: defaults not applied yet
    ga_tc_tuning built = {};            // the tuning the current arrays were built with (plan parameters only matter)
    const uint32_t* group = nullptr;     // locality group of each vertex (not owned; a disjoint union's graph ids): items
    uint32_t group_bits = 0;             //   are shape-ordered within each group, groups ascending
    uint32_t* perm = nullptr;            // GA_TC_SORTED work order (ga_graph_edge_order), lazily
};

struct ga_graph {
    ga_ctx* ctx = nullptr;
    uint32_t n = 0, m = 0, max_out = 0;
    uint32_t *rowptr = nullptr, *col = nullptr, *src = nullptr;
    uint32_t *urow = nullptr, *ucol = nullptr;   // undirected CSR, built on demand
    TcPlan plan;                                  // count plan on the CSR's own ids (every flagged launch; counts when
                                                  //   tune.relabel is 0)
    // tune.relabel = 1: plain and per-graph counts use a renumbered copy of the oriented CSR (same orientation, new
    // ids grouping vertices that share triangles) and its own plan; built on demand
    uint32_t *crow = nullptr, *ccol = nullptr, *csrc = nullptr;
    // tune.relabel = 2: the same rows placed so that no row of at most 32 ids crosses a 128-byte line (longer rows
    // start on a line when that saves one): row v at pcol[pst[v] .. pst[v] + d+(v)); built with the copy when its
    // padded length fits in 32 bits and in memory (else relabel 2 reads the unpadded copy)
    uint32_t *pst = nullptr, *pcol = nullptr;
    TcPlan cplan;
};

int ga_fail(ga_ctx* ctx, int code, const char* fmt, ...);

#define GA_HIP(ctx, call)                                                                                 \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess) return ga_fail((ctx), GA_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                                             hipGetErrorString(e_), __FILE__, __LINE__);                  \
    } while (0)

#define GA_TRY(call)                 \
    do {                             \
        int rc_ = (call);            \
        if (rc_ != GA_OK) return rc_; \
    } while (0)

int ga_malloc(ga_ctx* ctx, void** p, size_t bytes);
void ga_free(void* p);
int ga_scratch(ga_ctx* ctx, size_t bytes, void** p);   // returns a pointer to at least `bytes` of scratch
int ga_launch_check(ga_ctx* ctx, const char* what);

// device-side primitives implemented in ga_prims.hip
int ga_sort_u64(ga_ctx* ctx, uint64_t* keys, uint64_t* tmp, size_t count, int end_bit, uint64_t** result);
int ga_sort_pairs_u64(ga_ctx* ctx, unsigned long long* keys, unsigned long long* keys_tmp, uint32_t* vals, uint32_t* vals_tmp,
                      size_t count, int end_bit, unsigned long long** keys_out, uint32_t** vals_out);
int ga_sort_pairs_u32(ga_ctx* ctx, uint32_t* keys, uint32_t* keys_tmp, uint32_t* vals, uint32_t* vals_tmp, size_t count,
                      int end_bit, uint32_t** keys_out, uint32_t** vals_out);
int ga_exclusive_scan_u32(ga_ctx* ctx, const uint32_t* in, uint32_t* out, size_t count);
int ga_exclusive_scan_u64(ga_ctx* ctx, const uint64_t* in, uint64_t* out, size_t count);
int ga_fill_u32(ga_ctx* ctx, uint32_t* p, size_t count, uint32_t v);
int ga_fill_u8(ga_ctx* ctx, uint8_t* p, size_t count, uint8_t v);
int ga_fill_u64(ga_ctx* ctx, uint64_t* p, size_t count, uint64_t v);
int ga_max_u32(ga_ctx* ctx, const uint32_t* p, size_t count, uint32_t* out_host);

// graph helpers shared between files
int ga_build_from_device_pairs(ga_ctx* ctx, const uint32_t* dpairs, uint64_t k, int64_t n_in, bool strict, ga_graph** out,
                               float* build_ms);
int ga_graph_ensure_undirected(ga_ctx* ctx, ga_graph* g);
// prune_to_core on device-resident core numbers (ga_sub.hip): keep core_dev[v] >= k.
int ga_prune_core_dev(ga_ctx* ctx, const ga_graph* g, const uint32_t* core_dev, uint64_t k, ga_graph** out, uint32_t* old_id,
                      uint32_t* core_out, float* ms);
int ga_graph_ensure_perm(ga_ctx* ctx, ga_graph* g);
// Builds the count plan a launch uses: counts_only = the plan of plain and per-graph counts (on the renumbered copy
// when tune.relabel is set), else the plan on the CSR's own ids (participation, support, alive masks).
int ga_tc_ensure_plan(ga_ctx* ctx, ga_graph* g, bool counts_only = false);
// ga_tc_tune over the count form a caller runs: plain (gcnt NULL) or per graph of a disjoint union (gof, gcnt of
// ngraphs + 1 device words), as ga_batch_count launches it.
// Host wall clock in milliseconds (steady clock).
double wall_ms();
// Participation counts of several graphs in shared launches (one per bucket shape): graph i with alive[i], tri[i],
// its triangle total in *total[i] and its device flag done[i]. Built and uploaded once (into the context's buffer of
// ga_tc_count_many: one such call at a time on a context), then launched any number of times on ctx's stream.
struct ga_tc_many;
int ga_tc_participation_many(ga_ctx* ctx, ga_graph* const* graphs, const uint8_t* const* alive, uint32_t* const* tri,
                             unsigned long long* const* total, const uint32_t* const* done, uint32_t count, ga_tc_many** out);
int ga_tc_many_launch(ga_ctx* ctx, const ga_tc_many* M);
void ga_tc_many_free(ga_tc_many* M);
int ga_tc_tune_counts(ga_ctx* ctx, ga_graph* g, const uint32_t* gof, unsigned long long* gcnt, uint32_t ngraphs, ga_tc_tuning* best,
                      float* best_ms);
// One count of `variant` into *d_total (zeroed by the caller). Optional: alive mask, per-vertex tri, per-edge sup, a device
// `done` flag that turns the launch into a no-op, and per-graph counts gcnt[gof[h]] for a disjoint union of graphs.
int ga_tc_launch(ga_ctx* ctx, ga_graph* g, int variant, unsigned long long* d_total, const uint8_t* d_alive, uint32_t* d_tri,
                 uint32_t* d_sup, const uint32_t* d_done, const uint32_t* gof = nullptr, unsigned long long* gcnt = nullptr,
                 uint32_t gstride = 0);
// gcnt[g] = the sum of gcnt[g + n * x] over the eight XCD copies x of a per-graph count launched with gstride n.
int ga_tc_fold_xcd(ga_ctx* ctx, unsigned long long* gcnt, uint32_t n);

// GA_TC_NOSRC / GA_TC_VERTEX count into *total (zeroed by the caller): ga_tc_simple.hip.
int ga_tc_launch_simple(ga_ctx* ctx, ga_graph* g, int variant, unsigned long long* total);
// Looks up the device address of ga_noalloc.hip's static scratch on ctx's device (once per device).
int ga_noalloc_init(ga_ctx* ctx);

inline unsigned grid_for(uint64_t threads, unsigned block, unsigned cap = 1u << 20) {
    uint64_t b = (threads + block - 1) / block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// The grid of a count, peel or core launch under the context's cap (ga_ctx_set_grid_cap): with a cap of 1, 3 or 7
// workgroups every grid-stride and item loop runs many passes, the way the reference's emulator runs its kernels on
// 1-, 3- and 7-block grids (emu_main.cpp:40-57). Every kernel these launches use covers its work for any grid size.
inline unsigned ctx_grid(const ga_ctx* ctx, unsigned grid) { return ctx->grid_cap && grid > ctx->grid_cap ? ctx->grid_cap : grid; }

// ---------------------------------------------------------------------------------------------------------------
// device side: wave64 helpers
// ---------------------------------------------------------------------------------------------------------------
#define GA_WAVE 64

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t t = __shfl_xor(v, o, 64);
        v = t < v ? t : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t t = __shfl_xor(v, o, 64);
        v = t > v ? t : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of active lanes below this one in a ballot mask.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) { return (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull)); }

// First index in sorted a[lo, hi) with a[i] >= x.
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* __restrict__ a, uint32_t lo, uint32_t hi, uint32_t x) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// First index in sorted a[lo, hi) with a[i] > x.
__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t* __restrict__ a, uint32_t lo, uint32_t hi, uint32_t x) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
