// ga_core.hip - (k,triangle)-core decomposition, level-synchronous (dsg.h:665-751; kernels.cuh:461-626).
//
// Level rule: k = max(k, min over alive of tri); every alive v with tri[v] <= k is removed from the same snapshot
// with core[v] = k. Recount design: tri[] recounted on the alive subgraph every round. Incremental design: one full
// count, then each round subtracts exactly the triangles that touch the round's frontier, each triangle handled by
// its smallest-id frontier member (kernels.cuh:536-541), so tri[] after a round equals a recount.
//
// MI355X shape: wave64 min reduction with one atomic per wave; the frontier is appended with one atomic per wave
// (ballot + prefix popcount); the update kernel gives each frontier vertex u one wave whose lanes split u's sorted
// neighbour list, and each lane intersects N(u) after v with N(v) after v by walking the shorter list and binary-
// searching the longer one (cost min(|A|,|B|) log, so hubs on either side stay cheap).
// ga_core_run keeps k / nalive / rounds on the device (CoreDev) and enqueues rounds in blocks, one readback per block.
#include "ga_internal.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

__device__ __forceinline__ void group_sync_wave() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

#define GA_EMPTY 0xFFFFFFFFu

struct CoreDev {
    unsigned long long T;          // stats[0]: count of the recount / init pass
    unsigned long long removed;    // stats[1]: removed this round (= frontier size, incremental)
    unsigned long long k;          // stats[2]: level after this round
    unsigned long long nalive, k_prev, T_full;
    uint32_t min_tri, done, error, rounds;
};

enum { CORE_ERR_T = 1, CORE_ERR_BIG = 2, CORE_ERR_STATS = 3 };

struct CoreFusedMem;
static void cf_release(CoreFusedMem* m);

struct ga_core {
    CoreFusedMem* fm = nullptr;    // device buffers of the fused loop (first ga_core_run; kept for later runs)
    uint32_t* blk = nullptr;       // two alive lists of ga_core_run_many (first use; kept)
    ga_ctx* ctx = nullptr;
    ga_graph* g = nullptr;
    int incremental = 1;
    uint8_t* state = nullptr;      // 0 removed, 1 alive, 2 frontier of this round (recount uses 0/1)
    uint32_t* core = nullptr;
    uint32_t* tri = nullptr;
    uint32_t* frontier = nullptr;
    CoreDev* st = nullptr;
    int variant = GA_TC_HASH;
};

template <bool FUSED>
__global__ void k_core_begin(CoreDev* st, uint32_t* __restrict__ tri, uint32_t n, int reset_tri) {
    if (FUSED && st->done) return;
    if (reset_tri)
        for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) tri[v] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (reset_tri) st->T = 0;
        st->removed = 0;
        st->k = 0;
        st->min_tri = 0xFFFFFFFFu;
    }
}

// min over alive (state == 1) of tri (kernels.cuh:480-491).
template <bool FUSED>
__global__ void k_core_min(const uint32_t* __restrict__ tri, const uint8_t* __restrict__ state, uint32_t n, CoreDev* st) {
    if (FUSED && st->done) return;
    uint32_t mn = 0xFFFFFFFFu;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        if (state[v] == 1 && tri[v] < mn) mn = tri[v];
    mn = wave_min_u32(mn);
    if (lane_id() == 0 && mn != 0xFFFFFFFFu) atomicMin(&st->min_tri, mn);
}

// Recount level step (kernels.cuh:493-517).
__global__ void k_core_peel_level(const uint32_t* __restrict__ tri, uint8_t* __restrict__ alive, uint32_t* __restrict__ core,
                                  uint32_t n, unsigned long long k_in, CoreDev* st) {
    const unsigned long long mt = st->min_tri;
    const unsigned long long k = mt > k_in ? mt : k_in;
    uint64_t local = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        if (alive[v] == 1 && (unsigned long long)tri[v] <= k) {
            alive[v] = 0;
            core[v] = (uint32_t)k;
            local++;
        }
    }
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(&st->removed, (unsigned long long)local);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->k = k;
}

// Frontier: alive v with tri <= k (kernels.cuh:544-562); appended with one atomic per wave.
template <bool FUSED>
__global__ void k_core_find(const uint32_t* __restrict__ tri, uint8_t* __restrict__ state, uint32_t* __restrict__ core, uint32_t n,
                            unsigned long long k_in, uint32_t* __restrict__ frontier, CoreDev* st) {
    if (FUSED && st->done) return;
    const unsigned long long kp = FUSED ? st->k_prev : k_in;
    const unsigned long long mt = st->min_tri;
    const unsigned long long k = mt > kp ? mt : kp;
    const uint32_t lane = lane_id();
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t base = wave0 * 64; base < n; base += nwaves * 64) {   // wave-uniform loop
        const uint64_t v = base + lane;
        bool join = false;
        if (v < n && state[v] == 1 && (unsigned long long)tri[v] <= k) {
            join = true;
            state[v] = 2;
            core[v] = (uint32_t)k;
        }
        const uint64_t mask = __ballot(join);
        if (mask) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
            unsigned long long pos = 0;
            if (lane == leader) pos = atomicAdd(&st->removed, (unsigned long long)__popcll(mask));
            pos = __shfl(pos, leader, 64);
            if (join) frontier[pos + mask_rank(mask)] = (uint32_t)v;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->k = k;
}

// Triangles {u, v, w} through frontier vertex u, v = ucol[p] in N(u), w > v: subtract once, by the smallest-id
// frontier member (kernels.cuh:574-618). Walks the shorter of N(u)-after-v and N(v)-after-v and binary-searches the
// other, so hubs on either side stay cheap.
__device__ __forceinline__ void core_pair(const uint32_t* __restrict__ urow, const uint32_t* __restrict__ ucol, const uint8_t* __restrict__ state,
                                          uint32_t* __restrict__ tri, uint32_t u, uint32_t p, uint32_t ue) {
    const uint32_t v = ucol[p];
    const uint8_t sv = state[v];
    if (sv == 0) return;
    const uint32_t a = p + 1;                                                // N(u) after v: all > v
    const uint32_t vb = urow[v], ve = urow[v + 1];
    const uint32_t b = upper_bound_u32(ucol, vb, ve, v);                     // N(v) after v
    const bool a_short = (ue - a) <= (ve - b);
    uint32_t s = a_short ? a : b;
    const uint32_t se = a_short ? ue : ve;
    uint32_t l = a_short ? b : a;
    const uint32_t le = a_short ? ve : ue;
    for (; s < se && l < le; s++) {
        const uint32_t w = ucol[s];
        l = lower_bound_u32(ucol, l, le, w);
        if (l >= le) break;
        if (ucol[l] != w) continue;
        const uint8_t sw = state[w];
        if (sw == 0) continue;
        const bool smaller_frontier = (sv == 2 && v < u) || (sw == 2 && w < u);
        if (smaller_frontier) continue;
        if (sv == 1) atomicSub(&tri[v], 1u);
        if (sw == 1) atomicSub(&tri[w], 1u);
    }
}

// One wave per frontier vertex, lanes over N(u).
template <bool FUSED>
__global__ void k_core_update(const uint32_t* __restrict__ urow, const uint32_t* __restrict__ ucol, const uint8_t* __restrict__ state,
                              const uint32_t* __restrict__ frontier, CoreDev* st, uint32_t* __restrict__ tri) {
    if (FUSED && st->done) return;
    const unsigned long long F = st->removed;
    const uint32_t lane = lane_id();
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t i = wave0; i < F; i += nwaves) {
        const uint32_t u = frontier[i];
        const uint32_t ub = urow[u], ue = urow[u + 1];
        for (uint32_t p = ub + lane; p < ue; p += 64) core_pair(urow, ucol, state, tri, u, p, ue);
    }
}

template <bool FUSED>
__global__ void k_core_retire(uint8_t* __restrict__ state, const uint32_t* __restrict__ frontier, CoreDev* st) {
    if (FUSED && st->done) return;
    const unsigned long long F = st->removed;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < F; i += (unsigned long long)gridDim.x * blockDim.x)
        state[frontier[i]] = 0;
}

__global__ void k_core_init(CoreDev* st, unsigned long long n, unsigned long long T_full) {
    st->T = 0;
    st->removed = 0;
    st->k = 0;
    st->nalive = n;
    st->k_prev = 0;
    st->T_full = T_full;
    st->min_tri = 0xFFFFFFFFu;
    st->done = n == 0;
    st->error = 0;
    st->rounds = 0;
}

__global__ void k_core_check_T(CoreDev* st) {
    if (st->done) return;
    if (st->T_full != ~0ull && st->T != st->T_full) { st->error = CORE_ERR_T; st->done = 1; return; }
    if (st->T >= (1ull << 32)) { st->error = CORE_ERR_BIG; st->done = 1; }
}

// End of a fused round: the rules of core_decomp_loop_inc (dsg.h:744-747), then the next round's resets.
__global__ void k_core_end(CoreDev* st) {
    if (st->done) return;
    const unsigned long long removed = st->removed, k = st->k, nalive = st->nalive;
    if (removed == 0 || removed > nalive || k < st->k_prev) { st->error = CORE_ERR_STATS; st->done = 1; return; }
    st->nalive = nalive - removed;
    st->k_prev = k;
    st->rounds++;
    st->removed = 0;
    st->min_tri = 0xFFFFFFFFu;
    if (st->nalive == 0) st->done = 1;
}

static unsigned cgrid(ga_ctx* ctx, uint64_t n) { return ctx_grid(ctx, grid_for(n, 256, (unsigned)std::max(ctx->cus, 1) * 8u)); }
static unsigned ugrid(ga_ctx* ctx) { return ctx_grid(ctx, (unsigned)std::max(ctx->cus, 1) * 8u); }

extern "C" int ga_core_create(ga_ctx* ctx, ga_graph* g, int incremental, ga_core** out) {
    *out = nullptr;
    if (!ctx || !g) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    ga_core* c = new ga_core;
    c->ctx = ctx;
    c->g = g;
    c->incremental = incremental ? 1 : 0;
    const size_t n = g->n;
    if (ga_malloc(ctx, (void**)&c->state, n + 16) || ga_malloc(ctx, (void**)&c->core, (n + 1) * 4) ||
        ga_malloc(ctx, (void**)&c->tri, (n + 4) * 4) || ga_malloc(ctx, (void**)&c->frontier, (n + 1) * 4) ||
        ga_malloc(ctx, (void**)&c->st, sizeof(CoreDev))) {
        ga_core_destroy(c);
        return GA_ERR_OOM;
    }
    int rc = ga_tc_ensure_plan(ctx, g);
    if (!rc && c->incremental) rc = ga_graph_ensure_undirected(ctx, g);
    if (!rc) {
        k_core_init<<<1, 1, 0, ctx->stream>>>(c->st, g->n, ~0ull);
        rc = ga_launch_check(ctx, "k_core_init");
    }
    // finished before returning: the state may next run on another context's stream (ga_core_set_ctx)
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ga_fail(ctx, GA_ERR_HIP, "core state initialisation");
    if (rc) { ga_core_destroy(c); return rc; }
    *out = c;
    return GA_OK;
}

extern "C" int ga_core_set_ctx(ga_core* c, ga_ctx* ctx) {
    if (!c || !ctx || ctx->device != c->ctx->device) return GA_ERR_ARG;
    c->ctx = ctx;   // every call on the state ends with a stream sync, so an idle state has nothing in flight
    return GA_OK;
}

extern "C" void ga_core_destroy(ga_core* c) {
    if (!c) return;
    if (c->ctx) (void)hipStreamSynchronize(c->ctx->stream);
    ga_free(c->state);
    ga_free(c->core);
    ga_free(c->tri);
    ga_free(c->frontier);
    ga_free(c->st);
    ga_free(c->blk);
    cf_release(c->fm);
    delete c;
}

extern "C" int ga_core_reset_state(ga_core* c) {
    GA_TRY(ga_fill_u8(c->ctx, c->state, c->g->n, 1));
    GA_TRY(ga_fill_u32(c->ctx, c->core, c->g->n, UINT32_MAX));
    return GA_OK;
}

extern "C" int ga_core_begin_round(ga_core* c) {
    ga_ctx* ctx = c->ctx;
    k_core_begin<false><<<cgrid(ctx, c->g->n), 256, 0, ctx->stream>>>(c->st, c->tri, c->g->n, c->incremental ? 0 : 1);
    return ga_launch_check(ctx, "k_core_begin");
}

extern "C" int ga_core_run_tc(ga_core* c) {
    if (c->incremental) return ga_fail(c->ctx, GA_ERR_ARG, "run_tc belongs to the recount design");
    return ga_tc_launch(c->ctx, c->g, c->variant, &c->st->T, c->state, c->tri, nullptr, nullptr);
}

extern "C" int ga_core_run_min(ga_core* c) {
    ga_ctx* ctx = c->ctx;
    k_core_min<false><<<cgrid(ctx, c->g->n), 256, 0, ctx->stream>>>(c->tri, c->state, c->g->n, c->st);
    return ga_launch_check(ctx, "k_core_min");
}

extern "C" int ga_core_run_peel(ga_core* c, uint64_t k) {
    if (c->incremental) return ga_fail(c->ctx, GA_ERR_ARG, "run_peel belongs to the recount design");
    ga_ctx* ctx = c->ctx;
    k_core_peel_level<<<cgrid(ctx, c->g->n), 256, 0, ctx->stream>>>(c->tri, c->state, c->core, c->g->n, k, c->st);
    return ga_launch_check(ctx, "k_core_peel_level");
}

extern "C" int ga_core_read_stats(ga_core* c, uint64_t* T, uint64_t* removed, uint64_t* k) {
    ga_ctx* ctx = c->ctx;
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, c->st, 24, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (T) *T = ctx->hstat[0];
    if (removed) *removed = ctx->hstat[1];
    if (k) *k = ctx->hstat[2];
    return GA_OK;
}

static int init_tri_async(ga_core* c) {
    ga_ctx* ctx = c->ctx;
    k_core_begin<false><<<cgrid(ctx, c->g->n), 256, 0, ctx->stream>>>(c->st, c->tri, c->g->n, 1);
    GA_TRY(ga_launch_check(ctx, "k_core_begin"));
    return ga_tc_launch(ctx, c->g, c->variant, &c->st->T, c->state, c->tri, nullptr, nullptr);
}

extern "C" int ga_core_init_tri(ga_core* c, uint64_t* T) {
    if (!c->incremental) return ga_fail(c->ctx, GA_ERR_ARG, "init_tri belongs to the incremental design");
    GA_TRY(init_tri_async(c));
    return ga_core_read_stats(c, T, nullptr, nullptr);
}

extern "C" int ga_core_run_find(ga_core* c, uint64_t k) {
    if (!c->incremental) return ga_fail(c->ctx, GA_ERR_ARG, "run_find belongs to the incremental design");
    ga_ctx* ctx = c->ctx;
    k_core_find<false><<<cgrid(ctx, c->g->n), 256, 0, ctx->stream>>>(c->tri, c->state, c->core, c->g->n, k, c->frontier, c->st);
    return ga_launch_check(ctx, "k_core_find");
}

extern "C" int ga_core_run_update(ga_core* c) {
    if (!c->incremental) return ga_fail(c->ctx, GA_ERR_ARG, "run_update belongs to the incremental design");
    ga_ctx* ctx = c->ctx;
    k_core_update<false><<<ugrid(ctx), 256, 0, ctx->stream>>>(c->g->urow, c->g->ucol, c->state, c->frontier, c->st, c->tri);
    return ga_launch_check(ctx, "k_core_update");
}

extern "C" int ga_core_run_retire(ga_core* c) {
    if (!c->incremental) return ga_fail(c->ctx, GA_ERR_ARG, "run_retire belongs to the incremental design");
    ga_ctx* ctx = c->ctx;
    k_core_retire<false><<<ugrid(ctx), 256, 0, ctx->stream>>>(c->state, c->frontier, c->st);
    return ga_launch_check(ctx, "k_core_retire");
}

extern "C" int ga_core_prune(ga_core* c, uint64_t k, ga_graph** out, uint32_t* old_id, uint32_t* core_out, float* ms) {
    if (!c || !out) return GA_ERR_ARG;
    return ga_prune_core_dev(c->ctx, c->g, c->core, k, out, old_id, core_out, ms);
}

extern "C" int ga_core_download(ga_core* c, uint32_t* core) {
    ga_ctx* ctx = c->ctx;
    if (c->g->n) GA_HIP(ctx, hipMemcpyAsync(core, c->core, (size_t)c->g->n * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// fused incremental decomposition
// ---------------------------------------------------------------------------------------------------------------
// The level loop runs on the device with three structures that shrink with the alive set:
//   alist        the ids still alive at the last compaction (min / find scan it instead of all n vertices)
//   cbeg / cend  per-vertex ranges into a compacted copy of the undirected adjacency holding only neighbours that
//                were alive at the last compaction (sorted order kept), so late-round hubs carry short lists
//   entries      each frontier vertex u is split into (u, chunk) entries of at most CORE_CHUNK neighbours, one wave
//                per entry, so a hub's triangles are enumerated by many waves at once
// Compaction runs on the device (no host size readback) whenever at most half of alist is still alive: alist is
// filtered, and each surviving vertex copies its surviving neighbours into the other adjacency buffer, reserving
// space with one atomic per vertex. Every kernel returns at once once `done` is set.
#define CORE_CHUNK 8u

struct CoreFusedDev {
    uint32_t na;          // alist length
    uint32_t alist_cur;   // which alist buffer is current (0/1)
    uint32_t adj_cur;     // which adjacency buffer is current (0 = the graph's own ucol, 1/2 = copies)
    uint32_t compact;     // 1: compact after this round
    uint32_t na_new;      // compaction cursor for the new alist
    uint32_t nfv;         // frontier vertices this round
    uint32_t nent;        // frontier entries this round
    uint32_t min_next;    // min tri over the alive vertices after this round (the next round's level candidate)
    uint32_t ticket;      // workgroups of k_cf_tail that have finished this round
    uint32_t xcur;        // cross[xcur] holds this round's crossers (read by the find), cross[xcur ^ 1] the next round's
    uint32_t xcount[2];   // entries of each cross list
    unsigned long long adj_cursor;   // compaction cursor in the new adjacency buffer
};

struct CoreBufs {
    const uint32_t* bitmaps;  // bm_words words per hub
    size_t bm_words;
    uint32_t* alist[2];
    const uint32_t* adj[3];   // adj[0] = ucol; adj[1], adj[2] writable copies
    uint32_t* adjw[3];
    uint32_t* cbeg;
    uint32_t* cend;
    uint32_t* cgt;            // per vertex: position in its list of the first neighbour with a larger id (cN(v) after v)
    uint32_t* fvert;
    uint32_t* cross[2];       // vertices whose tri fell to the level or below in an update (the next frontier)
    uint32_t* ent;            // 2 words per entry: u, chunk
    const uint2* mref;        // per vertex: {bitmap index, 1 << 31} for a hub, else {first bucket, log2 of buckets}
    const uint4* mtab;        // membership tables of the original neighbour lists (4-slot buckets, load <= 1/2)
};

__global__ void k_cf_init(CoreFusedDev* f, uint32_t* __restrict__ alist0, uint32_t* __restrict__ cbeg, uint32_t* __restrict__ cend,
                          uint32_t* __restrict__ cgt, const uint32_t* __restrict__ urow, const uint32_t* __restrict__ ucol, uint32_t n) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        alist0[v] = v;
        cbeg[v] = urow[v];
        cend[v] = urow[v + 1];
        cgt[v] = upper_bound_u32(ucol, urow[v], urow[v + 1], v);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        f->na = n;
        f->alist_cur = 0;
        f->adj_cur = 0;
        f->compact = 0;
        f->na_new = 0;
        f->nfv = 0;
        f->nent = 0;
        f->min_next = 0xFFFFFFFFu;
        f->ticket = 0;
        f->xcur = 0;
        f->xcount[0] = f->xcount[1] = 0;
        f->adj_cursor = 0;
    }
}

__global__ void k_cf_min(CoreBufs B, const CoreFusedDev* __restrict__ f, const uint32_t* __restrict__ tri, const uint8_t* __restrict__ state,
                         CoreDev* st) {
    if (st->done) return;
    const uint32_t* __restrict__ al = B.alist[f->alist_cur];
    const uint32_t na = f->na;
    uint32_t mn = 0xFFFFFFFFu;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
        const uint32_t v = al[i];
        if (state[v] == 1 && tri[v] < mn) mn = tri[v];
    }
    mn = wave_min_u32(mn);
    if (lane_id() == 0 && mn != 0xFFFFFFFFu) atomicMin(&st->min_tri, mn);
}

// Frontier = alive v with tri <= k; appends v to fvert and, if it is in a triangle, its chunk entries (one atomic per
// wave for each). After a round whose update pushed some alive tri from above the level to the level or below, those
// crossers are exactly the next frontier and the level stays (min <= k_prev): the find then reads the cross list
// instead of the alive list. Otherwise k = the min over the alive vertices, taken by the previous tail.
__device__ __forceinline__ void cf_find_body(const CoreBufs& B, CoreFusedDev* f, const uint32_t* __restrict__ tri,
                                             uint8_t* __restrict__ state, uint32_t* __restrict__ core, CoreDev* st, uint32_t bid,
                                             uint32_t nb) {
    if (st->done) return;
    const uint32_t xc = f->xcur, nx = f->xcount[xc];
    const unsigned long long mt = st->min_tri, kp = st->k_prev;
    const unsigned long long k = nx ? kp : (mt > kp ? mt : kp);
    const uint32_t* __restrict__ al = nx ? B.cross[xc] : B.alist[f->alist_cur];
    const uint32_t na = nx ? nx : f->na;
    const uint32_t lane = lane_id();
    const uint64_t wave0 = ((uint64_t)bid * blockDim.x + threadIdx.x) / 64, nwaves = ((uint64_t)nb * blockDim.x) / 64;
    for (uint64_t base = wave0 * 64; base < na; base += nwaves * 64) {   // wave-uniform loop
        const uint64_t i = base + lane;
        bool join = false;
        uint32_t v = 0, nc = 0;
        if (i < na) {
            v = al[i];
            if (state[v] == 1 && (unsigned long long)tri[v] <= k) {
                join = true;
                state[v] = 2;
                core[v] = (uint32_t)k;
                // a vertex in no alive triangle has nothing to subtract: no entries
                const uint32_t d = B.cend[v] - B.cbeg[v];
                if (tri[v] && d) nc = (d + CORE_CHUNK - 1) / CORE_CHUNK;
            }
        }
        const uint64_t mask = __ballot(join);
        if (!mask) continue;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
        uint32_t inc = nc;   // inclusive scan over the wave
#pragma unroll
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o, 64);
            if (lane >= o) inc += t;
        }
        const uint32_t tot = __shfl(inc, 63, 64);
        uint32_t vpos = 0, epos = 0;
        if (lane == leader) {
            vpos = atomicAdd(&f->nfv, (uint32_t)__popcll(mask));
            epos = tot ? atomicAdd(&f->nent, tot) : 0u;
        }
        vpos = __shfl(vpos, leader, 64);
        epos = __shfl(epos, leader, 64);
        if (join) {
            B.fvert[vpos + mask_rank(mask)] = v;
            const uint32_t e0 = epos + inc - nc;
            for (uint32_t c = 0; c < nc; c++) {
                B.ent[2 * (e0 + c)] = v;
                B.ent[2 * (e0 + c) + 1] = c;
            }
        }
    }
    if (bid == 0 && threadIdx.x == 0) st->k = k;
}

__global__ void k_cf_find(CoreBufs B, CoreFusedDev* f, const uint32_t* __restrict__ tri, uint8_t* __restrict__ state,
                          uint32_t* __restrict__ core, CoreDev* st) {
    cf_find_body(B, f, tri, state, core, st, blockIdx.x, gridDim.x);
}

// Triangles {u, v, w} (w > v) through entry (u, chunk): subtract once, by the smallest-id frontier member.
// One wave per entry, 64 neighbours v at a time. For each pair the SHORTER of cN(u) after v and cN(v) after v is
// streamed and its elements are tested for membership in the other endpoint's ORIGINAL neighbour list: the hub
// bitmap (original degree > CORE_HCAP) or the static 4-slot-bucket table built once per state (cf_members). A member
// that has since been removed reads state 0 and is skipped, so the static sets answer like sets of alive neighbours.
// Streaming follows the count kernel: lists compacted in LDS longest first, read by sub-groups sized to their length,
// 4 loads in flight per lane. A hit w is a triangle {u, v, w}; its alive corners lose one triangle each.
#define CORE_HCAP 2048u

__device__ __forceinline__ bool bit_test(const uint32_t* __restrict__ bm, uint32_t w) { return (bm[w >> 5] >> (w & 31)) & 1u; }

// w in the original neighbour list of the vertex whose membership reference is (rx, ry)?
__device__ __forceinline__ bool cf_member(const CoreBufs& B, uint32_t rx, uint32_t ry, uint32_t w) {
    if (ry >> 31) return bit_test(B.bitmaps + (size_t)rx * B.bm_words, w);
    const uint32_t lb = ry & 31u, nb = 1u << lb;
    const uint4* __restrict__ t = B.mtab + rx;
    uint32_t b = lb ? (w * 2654435769u) >> (32u - lb) : 0u;
    for (uint32_t i = 0; i < nb; i++) {   // buckets fill front to back: a bucket with a free last slot ends the probe
        const uint4 q = t[b];
        if (q.x == w || q.y == w || q.z == w || q.w == w) return true;
        if (q.w == GA_EMPTY) return false;
        b = (b + 1) & (nb - 1);
    }
    return false;
}

// stg: 6 * 64 words of LDS (ps | len | v | sv + (mode << 8) | ref.x | ref.y); one 64-lane workgroup
__device__ __forceinline__ void cf_update_body(const CoreBufs& B, CoreFusedDev* f, const uint8_t* __restrict__ state,
                                               uint32_t* __restrict__ tri, const CoreDev* __restrict__ st, uint32_t bid, uint32_t nb,
                                               uint32_t* __restrict__ stg) {
    if (st->done) return;
    const uint32_t* __restrict__ adj = B.adj[f->adj_cur];
    // an alive vertex whose tri falls from above the level to the level or below joins the next round's frontier
    const unsigned long long lvl = st->k;
    const uint32_t xw = f->xcur ^ 1u;
    uint32_t* __restrict__ xlist = B.cross[xw];
    uint32_t* xcnt = &f->xcount[xw];
    const uint32_t nent = f->nent;
    const uint32_t lane = threadIdx.x;
    for (uint32_t ei = bid; ei < nent; ei += nb) {
        const uint32_t u = B.ent[2 * ei], chunk = B.ent[2 * ei + 1];
        const uint32_t ub = B.cbeg[u], ue = B.cend[u];
        const uint32_t pb = ub + chunk * CORE_CHUNK;
        const uint32_t pe = min(ue, pb + CORE_CHUNK);
        const uint2 uref = B.mref[u];
        for (uint32_t base = pb; base < pe; base += 64) {
            const uint32_t p = base + lane;
            uint32_t v = 0, sv = 0, ps = 0, len = 0, mode = 0;
            uint2 vref = make_uint2(0u, 0u);
            if (p < pe) {
                v = adj[p];
                sv = state[v];
                if (sv) {
                    const uint32_t lo = B.cgt[v], hi = B.cend[v];   // cN(v) after v
                    vref = B.mref[v];
                    const uint32_t a = ue - (p + 1), bl = hi - lo;
                    if (bl <= a) { ps = lo; len = bl; mode = 0; }   // stream cN(v) after v, test in N(u)
                    else { ps = p + 1; len = a; mode = 1; }         // stream cN(u) after v, test in N(v)
                }
            }
            const uint64_t gm = __ballot(len > 0);
            const uint32_t cnt = (uint32_t)__popcll(gm);
            if (len > 0) {
                const uint32_t r = mask_rank(gm);
                stg[r] = ps;
                stg[64 + r] = len;
                stg[128 + r] = v;
                stg[192 + r] = sv | (mode << 8);
                stg[256 + r] = vref.x;
                stg[320 + r] = vref.y;
            }
            group_sync_wave();
            const uint32_t lk = lane < cnt ? stg[64 + lane] : 0u;
            uint32_t start = 0;
#pragma unroll
            for (uint32_t S = 64; S >= 1; S >>= 1) {
                constexpr uint32_t U = 4;
                const uint32_t W = S >= U ? S / U : 1u;
                const uint32_t end = (uint32_t)__popcll(__ballot(lk > S / 2));
                for (uint32_t k0 = start; k0 < end; k0 += 64 / W) {
                    const uint32_t k = k0 + lane / W, r = lane % W;
                    uint32_t c = 0, kv = 0, ksv = 0;
                    if (k < end) {
                        const uint32_t kps = stg[k], klen = stg[64 + k];
                        kv = stg[128 + k];
                        const uint32_t sm = stg[192 + k];
                        ksv = sm & 0xFF;
                        // the other endpoint's membership reference: u's when streaming cN(v), v's when streaming cN(u)
                        const bool test_u = (sm >> 8) == 0;
                        const uint32_t rx = test_u ? uref.x : stg[256 + k], ry = test_u ? uref.y : stg[320 + k];
                        for (uint32_t off = r; off < klen; off += U * W) {
                            uint32_t x[U];
#pragma unroll
                            for (uint32_t q = 0; q < U; q++) x[q] = off + q * W < klen ? adj[kps + off + q * W] : 0u;
                            bool dm[U];
#pragma unroll
                            for (uint32_t q = 0; q < U; q++) {
                                dm[q] = false;
                                if (off + q * W >= klen) continue;
                                const uint32_t w = x[q];
                                if (!cf_member(B, rx, ry, w)) continue;
                                const uint8_t sw = state[w];
                                if (sw == 0) continue;
                                if ((ksv == 2 && kv < u) || (sw == 2 && w < u)) continue;   // a smaller frontier member owns it
                                dm[q] = sw == 1;
                                c++;
                            }
                            uint32_t old[U];
#pragma unroll
                            for (uint32_t q = 0; q < U; q++)   // all decrements in flight before the first crossing test
                                if (dm[q]) old[q] = atomicSub(&tri[x[q]], 1u);
#pragma unroll
                            for (uint32_t q = 0; q < U; q++)
                                if (dm[q] && (unsigned long long)old[q] == lvl + 1) xlist[atomicAdd(xcnt, 1u)] = x[q];
                        }
                    }
#pragma unroll
                    for (uint32_t o = W / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
                    if (r == 0 && k < end && c && ksv == 1) {
                        const uint32_t o = atomicSub(&tri[kv], c);
                        if ((unsigned long long)o > lvl && (unsigned long long)o <= lvl + c) xlist[atomicAdd(xcnt, 1u)] = kv;
                    }
                }
                start = end;
                if (S == 1) break;
            }
            group_sync_wave();
        }
    }
}

__global__ void __launch_bounds__(64) k_cf_update(CoreBufs B, CoreFusedDev* f, const uint8_t* __restrict__ state,
                                                  uint32_t* __restrict__ tri, const CoreDev* __restrict__ st) {
    __shared__ uint32_t stg[6 * 64];
    cf_update_body(B, f, state, tri, st, blockIdx.x, gridDim.x, stg);
}

// Membership bitmaps of the vertices whose ORIGINAL degree exceeds CORE_HCAP: hubid[v] = index or GA_EMPTY.
__global__ void k_cf_hub_ids(const uint32_t* __restrict__ urow, uint32_t n, uint32_t* __restrict__ hubid, uint32_t* __restrict__ nhub) {
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t base = wave0 * 64; base < n; base += nwaves * 64) {
        const uint64_t v = base + lane_id();
        const bool hub = v < n && urow[v + 1] - urow[v] > CORE_HCAP;
        const uint64_t mask = __ballot(hub);
        uint32_t pos = 0;
        if (mask) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
            if (lane_id() == leader) pos = atomicAdd(nhub, (uint32_t)__popcll(mask));
            pos = __shfl(pos, leader, 64);
        }
        if (v < n) hubid[v] = hub ? pos + mask_rank(mask) : GA_EMPTY;
    }
}

__global__ void k_cf_hub_bits(const uint32_t* __restrict__ urow, const uint32_t* __restrict__ ucol, const uint32_t* __restrict__ hubid, uint32_t n,
                              uint32_t* __restrict__ bitmaps, size_t words) {
    for (uint32_t v = blockIdx.x; v < n; v += gridDim.x) {
        const uint32_t h = hubid[v];
        if (h == GA_EMPTY) continue;
        uint32_t* bm = bitmaps + (size_t)h * words;
        for (uint32_t p = urow[v] + threadIdx.x; p < urow[v + 1]; p += blockDim.x) {
            const uint32_t w = ucol[p];
            atomicOr(&bm[w >> 5], 1u << (w & 31));
        }
    }
}

// Membership tables of the other vertices (1 <= original degree <= CORE_HCAP): 2^lb buckets of 4 slots with
// lb = max(ceil(log2 d) - 1, 0), so at most half of the slots are used. mref[v] = {first bucket, lb}; hubs get
// {bitmap index, 1 << 31}. nbk[v] = buckets of v (exclusive-scanned into the first buckets by the caller).
__device__ __forceinline__ uint32_t cf_table_log2(uint32_t d) {
    const uint32_t c = d > 1 ? 32u - (uint32_t)__clz(d - 1) : 0u;   // ceil(log2 d)
    return c ? c - 1u : 0u;
}

__global__ void k_cf_table_sizes(const uint32_t* __restrict__ urow, const uint32_t* __restrict__ hubid, uint32_t n, uint32_t* __restrict__ nbk) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += gridDim.x * blockDim.x) {
        const uint32_t d = v < n ? urow[v + 1] - urow[v] : 0u;
        nbk[v] = (d == 0 || hubid[v] != GA_EMPTY) ? 0u : 1u << cf_table_log2(d);
    }
}

__global__ void k_cf_table_refs(const uint32_t* __restrict__ urow, const uint32_t* __restrict__ hubid, const uint32_t* __restrict__ boff,
                                uint32_t n, uint2* __restrict__ mref) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t h = hubid[v];
        mref[v] = h != GA_EMPTY ? make_uint2(h, 1u << 31) : make_uint2(boff[v], cf_table_log2(urow[v + 1] - urow[v]));
    }
}

// One wave per vertex inserts its original neighbours (no duplicates in the undirected CSR) into its table: slots of a
// bucket are claimed in order, so a bucket fills front to back; a full bucket passes on to the next one.
__global__ void k_cf_table_fill(const uint32_t* __restrict__ urow, const uint32_t* __restrict__ ucol, const uint2* __restrict__ mref,
                                uint32_t n, uint32_t* __restrict__ mtab) {
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t v = wave0; v < n; v += nwaves) {
        const uint2 ref = mref[v];
        if (ref.y >> 31) continue;
        const uint32_t lb = ref.y, nb = 1u << lb;
        uint32_t* __restrict__ t = mtab + 4ull * ref.x;
        for (uint32_t p = urow[v] + lane_id(); p < urow[v + 1]; p += 64) {
            const uint32_t w = ucol[p];
            uint32_t b = lb ? (w * 2654435769u) >> (32u - lb) : 0u;
            bool placed = false;
            for (uint32_t i = 0; i < nb && !placed; i++) {
                for (uint32_t q = 0; q < 4 && !placed; q++) placed = atomicCAS(&t[4 * b + q], GA_EMPTY, w) == GA_EMPTY;
                b = (b + 1) & (nb - 1);
            }
        }
    }
}

// Round tail, one launch: retire the frontier, take the next round's min over
// the alive list and, when the alive list halves, compact the list and the adjacency into the other buffers (alive =
// state 1; frontier members read as 2 or 0, both excluded). The last workgroup to finish (atomic ticket, no waiting)
// applies the round's end rules (dsg.h:744-747) and switches to the compacted buffers.
// sh: 2 words of LDS (last, bmin)
__device__ __forceinline__ void cf_tail_body(const CoreBufs& B, CoreFusedDev* f, uint8_t* __restrict__ state, const uint32_t* __restrict__ tri,
                                             CoreDev* st, uint32_t bid, uint32_t nb, uint32_t* sh) {
    uint32_t& last = sh[0];
    uint32_t& bmin = sh[1];
    if (st->done) return;
    if (threadIdx.x == 0) bmin = 0xFFFFFFFFu;
    const uint32_t nfv = f->nfv, na = f->na;
    const unsigned long long nalive = st->nalive;
    const bool compact = nfv < nalive && 2ull * (nalive - nfv) <= na;   // the same decision in every workgroup
    // with crossers the next round keeps the level and reads them: no min is needed, so no scan unless compacting
    const uint32_t sna = (compact || f->xcount[f->xcur ^ 1u] == 0) ? na : 0u;
    const uint32_t lane = lane_id();
    const uint64_t tid = (uint64_t)bid * blockDim.x + threadIdx.x, nthr = (uint64_t)nb * blockDim.x;
    for (uint64_t i = tid; i < nfv; i += nthr) state[B.fvert[i]] = 0;
    const uint32_t* __restrict__ al = B.alist[f->alist_cur];
    uint32_t* __restrict__ out = B.alist[f->alist_cur ^ 1u];
    const uint64_t wave0 = tid / 64, nwaves = nthr / 64;
    uint32_t mn = 0xFFFFFFFFu;
    for (uint64_t base = wave0 * 64; base < sna; base += nwaves * 64) {   // wave-uniform loop
        const uint64_t i = base + lane;
        uint32_t v = 0;
        const bool keep = i < na && state[(v = al[i])] == 1;
        if (keep) mn = min(mn, tri[v]);
        if (compact) {
            const uint64_t mask = __ballot(keep);
            if (!mask) continue;
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
            uint32_t pos = 0;
            if (lane == leader) pos = atomicAdd(&f->na_new, (uint32_t)__popcll(mask));
            pos = __shfl(pos, leader, 64);
            if (keep) out[pos + mask_rank(mask)] = v;
        }
    }
    mn = wave_min_u32(mn);
    __syncthreads();   // bmin initialised
    if (lane == 0 && mn != 0xFFFFFFFFu) atomicMin(&bmin, mn);
    if (compact) {   // one wave per surviving vertex copies its alive neighbours, in order, into the next buffer
        const uint32_t cur = f->adj_cur, nxt = cur == 1 ? 2u : 1u;
        const uint32_t* __restrict__ src = B.adj[cur];
        uint32_t* __restrict__ dst = B.adjw[nxt];
        for (uint64_t i = wave0; i < na; i += nwaves) {
            const uint32_t v = al[i];
            if (state[v] != 1) continue;   // wave-uniform
            const uint32_t b = B.cbeg[v], e = B.cend[v];
            uint32_t cnt = 0;
            for (uint32_t p = b + lane; p < e; p += 64) cnt += state[src[p]] == 1;
            cnt = (uint32_t)wave_sum_u64(cnt);
            unsigned long long at = 0;
            if (lane == 0 && cnt) at = atomicAdd(&f->adj_cursor, (unsigned long long)cnt);
            at = __shfl(at, 0, 64);
            uint32_t w = (uint32_t)at, below = 0;   // below: kept neighbours with a smaller id than v
            for (uint32_t p0 = b; p0 < e; p0 += 64) {   // wave-uniform, keeps the sorted order
                const uint32_t p = p0 + lane;
                const uint32_t x = p < e ? src[p] : 0u;
                const bool k2 = p < e && state[x] == 1;
                const uint64_t mask = __ballot(k2);
                if (k2) dst[w + mask_rank(mask)] = x;
                w += (uint32_t)__popcll(mask);
                below += (uint32_t)__popcll(__ballot(k2 && x < v));
            }
            if (lane == 0) {
                B.cbeg[v] = (uint32_t)at;
                B.cend[v] = (uint32_t)at + cnt;
                B.cgt[v] = (uint32_t)at + below;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (bmin != 0xFFFFFFFFu) atomicMin(&f->min_next, bmin);   // one global atomic per workgroup
        __threadfence();
        last = atomicAdd(&f->ticket, 1u) == nb - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __threadfence();
    const unsigned long long removed = nfv, k = st->k;
    st->removed = removed;
    f->ticket = 0;
    if (removed == 0 || removed > nalive || k < st->k_prev) { st->error = CORE_ERR_STATS; st->done = 1; return; }
    st->nalive = nalive - removed;
    st->k_prev = k;
    st->rounds++;
    st->min_tri = atomicAdd(&f->min_next, 0u);
    f->min_next = 0xFFFFFFFFu;
    f->nfv = 0;
    f->nent = 0;
    f->xcount[f->xcur] = 0;   // this round's crossers are consumed; the next round reads the ones its update found
    f->xcur ^= 1u;
    if (compact) {
        f->na = atomicAdd(&f->na_new, 0u);
        f->alist_cur ^= 1u;
        f->adj_cur = f->adj_cur == 1 ? 2u : 1u;
    }
    f->na_new = 0;
    f->adj_cursor = 0;
    if (st->nalive == 0) st->done = 1;
}

__global__ void k_cf_tail(CoreBufs B, CoreFusedDev* f, uint8_t* __restrict__ state, const uint32_t* __restrict__ tri, CoreDev* st) {
    __shared__ uint32_t sh[2];
    cf_tail_body(B, f, state, tri, st, blockIdx.x, gridDim.x, sh);
}

struct CoreFusedMem {
    ga_ctx* ctx = nullptr;
    void* block = nullptr;
    void* hubs = nullptr;
    void* block_bm = nullptr;
    void* block_tab = nullptr;
    CoreBufs B{};
    CoreFusedDev* f = nullptr;
    ~CoreFusedMem() {
        ga_free(block);
        ga_free(hubs);
        ga_free(block_bm);
        ga_free(block_tab);
    }
};

static void cf_release(CoreFusedMem* m) { delete m; }

// The membership structures of the original undirected lists, built once per state: bitmaps for hubs, 4-slot-bucket
// tables for the other vertices (about 2-4 slots per listed neighbour).
static int cf_members(ga_ctx* ctx, ga_graph* g, CoreFusedMem& M) {
    const size_t n = g->n;
    void* idb = nullptr;
    // hubid (n+1) | counter (16) | bucket counts (n+1) | first buckets (n+1) | pad | mref (2n, 8-byte aligned)
    GA_TRY(ga_malloc(ctx, &idb, (5 * n + 24) * 4 + 64));
    M.hubs = idb;
    uint32_t* hubid = (uint32_t*)idb;
    uint32_t* cnt = hubid + n + 1;
    uint32_t* nbk = cnt + 16;
    uint32_t* boff = nbk + n + 1;
    uint2* mref = (uint2*)(((uintptr_t)(boff + n + 1) + 7) & ~(uintptr_t)7);
    GA_TRY(ga_fill_u32(ctx, cnt, 1, 0));
    k_cf_hub_ids<<<cgrid(ctx, n), 256, 0, ctx->stream>>>(g->urow, g->n, hubid, cnt);
    GA_TRY(ga_launch_check(ctx, "k_cf_hub_ids"));
    k_cf_table_sizes<<<cgrid(ctx, n + 1), 256, 0, ctx->stream>>>(g->urow, hubid, g->n, nbk);
    GA_TRY(ga_launch_check(ctx, "k_cf_table_sizes"));
    GA_TRY(ga_exclusive_scan_u32(ctx, nbk, boff, n + 1));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, cnt, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat + 1, boff + n, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t nh = ctx->hstat[0] & 0xFFFFFFFFu, nbuckets = ctx->hstat[1] & 0xFFFFFFFFu;
    if (nbuckets > 2ull * g->m) return ga_fail(ctx, GA_ERR_STATE, "membership table sizes exceed the edge count");
    const size_t words = (n + 31) / 32;
    M.B.bm_words = words;
    M.B.bitmaps = nullptr;
    if (nh) {
        GA_TRY(ga_malloc(ctx, &M.block_bm, nh * words * 4));
        GA_TRY(ga_fill_u32(ctx, (uint32_t*)M.block_bm, nh * words, 0));
        k_cf_hub_bits<<<std::min<uint32_t>(g->n, 65535), 256, 0, ctx->stream>>>(g->urow, g->ucol, hubid, g->n, (uint32_t*)M.block_bm, words);
        GA_TRY(ga_launch_check(ctx, "k_cf_hub_bits"));
        M.B.bitmaps = (const uint32_t*)M.block_bm;
    }
    GA_TRY(ga_malloc(ctx, &M.block_tab, std::max<size_t>(nbuckets, 1) * 16));
    GA_TRY(ga_fill_u32(ctx, (uint32_t*)M.block_tab, 4 * std::max<size_t>(nbuckets, 1), GA_EMPTY));
    k_cf_table_refs<<<cgrid(ctx, n), 256, 0, ctx->stream>>>(g->urow, hubid, boff, g->n, mref);
    GA_TRY(ga_launch_check(ctx, "k_cf_table_refs"));
    if (n) {
        k_cf_table_fill<<<grid_for(64ull * n, 256, 65535), 256, 0, ctx->stream>>>(g->urow, g->ucol, mref, g->n, (uint32_t*)M.block_tab);
        GA_TRY(ga_launch_check(ctx, "k_cf_table_fill"));
    }
    M.B.mref = mref;
    M.B.mtab = (const uint4*)M.block_tab;
    return GA_OK;
}

static int cf_alloc(ga_ctx* ctx, ga_graph* g, CoreFusedMem& M) {
    const size_t n = g->n, m2 = 2ull * g->m;
    const size_t nent = n + m2 / CORE_CHUNK + 16;
    const size_t words = 2 * n + 2 * std::max<size_t>(m2, 1) + 3 * n + n + 2 * n + 2 * nent + 64 + 16;
    if (m2 >= 0xFFFFFFFFull || nent >= 0x7FFFFFFFull)
        return ga_fail(ctx, GA_ERR_ARG, "graph too large for the fused core decomposition (32-bit offsets)");
    GA_TRY(ga_malloc(ctx, &M.block, words * 4));
    uint32_t* p = (uint32_t*)M.block;
    M.B.alist[0] = p; p += n;
    M.B.alist[1] = p; p += n;
    M.B.adj[0] = g->ucol;
    M.B.adjw[0] = nullptr;
    M.B.adjw[1] = p; M.B.adj[1] = p; p += std::max<size_t>(m2, 1);
    M.B.adjw[2] = p; M.B.adj[2] = p; p += std::max<size_t>(m2, 1);
    M.B.cbeg = p; p += n;
    M.B.cend = p; p += n;
    M.B.cgt = p; p += n;
    M.B.fvert = p; p += n;
    M.B.cross[0] = p; p += n;
    M.B.cross[1] = p; p += n;
    M.B.ent = p; p += 2 * nent;
    p = (uint32_t*)(((uintptr_t)p + 63) & ~(uintptr_t)63);
    M.f = (CoreFusedDev*)p;
    return GA_OK;
}

// First use of a state in the fused loop: its buffers and membership structures (kept for later runs).
static int cf_ensure(ga_core* c) {
    if (c->fm) return GA_OK;
    CoreFusedMem* fm = new CoreFusedMem;
    int rc = cf_alloc(c->ctx, c->g, *fm);
    if (!rc) rc = cf_members(c->ctx, c->g, *fm);
    if (rc) { delete fm; return rc; }
    c->fm = fm;
    return GA_OK;
}

// Enqueues the start of a fused decomposition: state reset, the one full participation count, the T check and round
// 0's level candidate.
static int cf_begin(ga_core* c, uint64_t T_full) {
    ga_ctx* ctx = c->ctx;
    ga_graph* g = c->g;
    CoreFusedMem& M = *c->fm;
    GA_TRY(ga_core_reset_state(c));
    k_core_init<<<1, 1, 0, ctx->stream>>>(c->st, g->n, T_full);
    GA_TRY(ga_launch_check(ctx, "k_core_init"));
    k_cf_init<<<cgrid(ctx, g->n), 256, 0, ctx->stream>>>(M.f, M.B.alist[0], M.B.cbeg, M.B.cend, M.B.cgt, g->urow, g->ucol, g->n);
    GA_TRY(ga_launch_check(ctx, "k_cf_init"));
    if (g->n) {
        GA_TRY(init_tri_async(c));
        k_core_check_T<<<1, 1, 0, ctx->stream>>>(c->st);
        GA_TRY(ga_launch_check(ctx, "k_core_check_T"));
    }
    k_cf_min<<<cgrid(ctx, g->n), 256, 0, ctx->stream>>>(M.B, M.f, c->tri, c->state, c->st);
    return ga_launch_check(ctx, "k_cf_min");
}

// Launch shapes of one decomposition's round for `alive` list entries: find / tail workgroups of 256 threads, update
// workgroups of 64. Small graphs and late rounds launch few workgroups, which leaves the device to other work.
struct CfGrid {
    unsigned v, u, t;
};

static CfGrid cf_grid(ga_ctx* ctx, uint64_t alive) {
    CfGrid r;
    r.v = cgrid(ctx, std::max<uint64_t>(alive, 1));
    r.u = ctx_grid(ctx, 4u * std::min<unsigned>(ugrid(ctx), std::max<unsigned>(4u, (unsigned)std::min<uint64_t>((8 * alive + 255) / 256, 1u << 20))));
    r.t = std::min<unsigned>(r.v, (unsigned)std::max(ctx->cus, 1) * 2u);   // few workgroups: few tickets
    return r;
}

static int cf_result(ga_ctx* ctx, const CoreDev* h, uint32_t* stats, const char* who) {
    switch (h->error) {
        case 0: break;
        case CORE_ERR_T: return ga_fail(ctx, GA_ERR_STATE, "%sdevice triangle count differs from the given count", who);
        case CORE_ERR_BIG: return ga_fail(ctx, GA_ERR_ARITH, "%striangle count >= 2^32: uint32 tri[] counters could overflow", who);
        default: return ga_fail(ctx, GA_ERR_STATE, "%sbad stats from device (removed 0, removed > alive, or level decreased)", who);
    }
    if (stats) {
        stats[0] = h->rounds;
        stats[1] = (uint32_t)h->k_prev;
    }
    return GA_OK;
}

// Whole decomposition with the round loop on the device (incremental design with compaction; the recount design is
// available through the per-round calls).
extern "C" int ga_core_run(ga_core* c, uint64_t T_full, uint32_t* stats, float* ms) {
    ga_ctx* ctx = c->ctx;
    ga_graph* g = c->g;
    if (!c->incremental) return ga_fail(ctx, GA_ERR_ARG, "ga_core_run runs the incremental design");
    if (T_full != UINT64_MAX && T_full >= (1ull << 32))
        return ga_fail(ctx, GA_ERR_ARITH, "triangle count >= 2^32: uint32 tri[] counters could overflow");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(cf_ensure(c));   // allocated once per core object: later runs allocate and free nothing
    CoreFusedMem& M = *c->fm;
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(cf_begin(c, T_full));
    int block_rounds = 8;
    CfGrid gr = cf_grid(ctx, g->n);
    for (uint64_t guard = 0;; guard++) {
        for (int r = 0; r < block_rounds; r++) {   // 3 launches per round
            k_cf_find<<<gr.v, 256, 0, ctx->stream>>>(M.B, M.f, c->tri, c->state, c->core, c->st);
            k_cf_update<<<gr.u, 64, 0, ctx->stream>>>(M.B, M.f, c->state, c->tri, c->st);
            k_cf_tail<<<gr.t, 256, 0, ctx->stream>>>(M.B, M.f, c->state, c->tri, c->st);
        }
        GA_TRY(ga_launch_check(ctx, "core round"));
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, c->st, sizeof(CoreDev), hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const CoreDev* h = (const CoreDev*)ctx->hstat;
        if (h->done) break;
        if (guard > (uint64_t)g->n + 16) return ga_fail(ctx, GA_ERR_STATE, "core decomposition did not finish");
        block_rounds = std::min(block_rounds * 2, 256);
        gr = cf_grid(ctx, 2 * h->nalive);   // the alive list is at most about twice the alive count (compaction at half)
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return cf_result(ctx, (const CoreDev*)ctx->hstat, stats, "");
}

// ---------------------------------------------------------------------------------------------------------------
// many decompositions in shared launches: each round's find / update / tail launch covers every unfinished state
// ---------------------------------------------------------------------------------------------------------------
// A state owns a contiguous range of each launch's workgroups (base, nblk); map[] gives the state of every
// workgroup. The kernels are the single-graph ones run on the state's own block range, so every state follows exactly
// the rounds it follows alone. Ranges follow each state's alive count and are rebuilt after every block of rounds;
// finished states get none.
struct CoreJob {
    CoreBufs B;
    CoreFusedDev* f;
    uint8_t* state;
    uint32_t* tri;
    uint32_t* core;
    CoreDev* st;
    uint32_t base[3], nblk[3];   // find, update, tail
};

__global__ void k_cfb_find(const CoreJob* __restrict__ jobs, const uint32_t* __restrict__ map) {
    const CoreJob* J = jobs + map[blockIdx.x];
    cf_find_body(J->B, J->f, J->tri, J->state, J->core, J->st, blockIdx.x - J->base[0], J->nblk[0]);
}

__global__ void __launch_bounds__(64) k_cfb_update(const CoreJob* __restrict__ jobs, const uint32_t* __restrict__ map) {
    __shared__ uint32_t stg[6 * 64];
    const CoreJob* J = jobs + map[blockIdx.x];
    cf_update_body(J->B, J->f, J->state, J->tri, J->st, blockIdx.x - J->base[1], J->nblk[1], stg);
}

__global__ void k_cfb_tail(const CoreJob* __restrict__ jobs, const uint32_t* __restrict__ map) {
    __shared__ uint32_t sh[2];
    const CoreJob* J = jobs + map[blockIdx.x];
    cf_tail_body(J->B, J->f, J->state, J->tri, J->st, blockIdx.x - J->base[2], J->nblk[2], sh);
}

__global__ void k_cfb_stats(const CoreJob* __restrict__ jobs, uint32_t count, CoreDev* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) out[i] = *jobs[i].st;
}

extern "C" int ga_core_run_batch(ga_ctx* ctx, ga_core* const* cores, uint32_t count, const uint64_t* T_full, uint32_t* stats, float* ms) {
    if (!ctx || (count && (!cores || !stats))) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (!count) return GA_OK;
    if (count > (1u << 20)) return ga_fail(ctx, GA_ERR_ARG, "ga_core_run_batch: at most 2^20 states per call");
    for (uint32_t i = 0; i < count; i++) {
        ga_core* c = cores[i];
        if (!c || !c->incremental || c->ctx != ctx)
            return ga_fail(ctx, GA_ERR_ARG, "core %u: not an incremental state of this context (ga_core_set_ctx moves one)", i);
        if (T_full && T_full[i] != UINT64_MAX && T_full[i] >= (1ull << 32))
            return ga_fail(ctx, GA_ERR_ARITH, "core %u: triangle count >= 2^32: uint32 tri[] counters could overflow", i);
    }
    {
        std::vector<ga_core*> sorted(cores, cores + count);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return ga_fail(ctx, GA_ERR_ARG, "ga_core_run_batch: a state appears twice");
    }
    for (uint32_t i = 0; i < count; i++) GA_TRY(cf_ensure(cores[i]));
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (uint32_t i = 0; i < count; i++) GA_TRY(cf_begin(cores[i], T_full ? T_full[i] : UINT64_MAX));

    // per state: the workgroups it would launch alone (they follow its alive count); when a launch's total exceeds
    // its budget, every state's share is scaled down by the same factor, so large alive lists keep proportionally more
    std::vector<uint64_t> alive(count);
    std::vector<uint8_t> done(count, 0);
    for (uint32_t i = 0; i < count; i++) alive[i] = cores[i]->g->n;
    const unsigned cus = (unsigned)std::max(ctx->cus, 1);
    std::vector<CoreJob> jobs(count);
    for (uint32_t i = 0; i < count; i++) {
        ga_core* c = cores[i];
        jobs[i] = CoreJob{c->fm->B, c->fm->f, c->state, c->tri, c->core, c->st, {0, 0, 0}, {0, 0, 0}};
    }
    std::vector<uint32_t> map[3];
    std::vector<uint8_t> host;
    size_t jobs_bytes = ((size_t)count * sizeof(CoreJob) + 255) & ~(size_t)255;
    const size_t stats_bytes = ((size_t)count * sizeof(CoreDev) + 255) & ~(size_t)255;
    int block_rounds = 8;
    std::vector<CoreDev> hst(count);
    const uint64_t budget[3] = {64ull * cus, 256ull * cus, 16ull * cus};   // find / update / tail workgroups per launch
    const unsigned floor_nb[3] = {1u, 4u, 1u};
    std::vector<CfGrid> grids(count);
    for (uint64_t guard = 0;; guard++) {
        uint32_t active = 0;
        for (uint32_t i = 0; i < count; i++) active += !done[i];
        if (!active) break;
        uint64_t total[3] = {0, 0, 0};
        for (uint32_t i = 0; i < count; i++) {
            if (done[i]) continue;
            grids[i] = cf_grid(ctx, guard ? 2 * alive[i] : alive[i]);
            total[0] += grids[i].v;
            total[1] += grids[i].u;
            total[2] += grids[i].t;
        }
        for (auto& m : map) m.clear();
        for (uint32_t i = 0; i < count; i++) {
            CoreJob& J = jobs[i];
            for (int q = 0; q < 3; q++) J.base[q] = J.nblk[q] = 0;
            if (done[i]) continue;
            unsigned nb[3] = {grids[i].v, grids[i].u, grids[i].t};
            for (int q = 0; q < 3; q++) {
                if (total[q] > budget[q]) nb[q] = std::max(floor_nb[q], (unsigned)((uint64_t)nb[q] * budget[q] / total[q]));
                J.base[q] = (uint32_t)map[q].size();
                J.nblk[q] = nb[q];
                map[q].insert(map[q].end(), nb[q], i);
            }
        }
        const size_t map_words = map[0].size() + map[1].size() + map[2].size();
        const size_t up_bytes = jobs_bytes + map_words * 4;
        const size_t stats_off = (up_bytes + 255) & ~(size_t)255;
        const size_t need = stats_off + stats_bytes;
        if (need > ctx->dcore_bytes) {   // grow-only: shared with ga_core_run_many, used by one call at a time
            ga_free(ctx->dcore);
            ctx->dcore = nullptr;
            ctx->dcore_bytes = 0;
            GA_TRY(ga_malloc(ctx, &ctx->dcore, need * 2));
            ctx->dcore_bytes = need * 2;
        }
        host.resize(up_bytes);
        memcpy(host.data(), jobs.data(), (size_t)count * sizeof(CoreJob));
        size_t at = jobs_bytes;
        for (auto& m : map) {
            memcpy(host.data() + at, m.data(), m.size() * 4);
            at += m.size() * 4;
        }
        char* dbase = (char*)ctx->dcore;
        const CoreJob* djobs = (const CoreJob*)dbase;
        const uint32_t* dmap[3];
        dmap[0] = (const uint32_t*)(dbase + jobs_bytes);
        dmap[1] = dmap[0] + map[0].size();
        dmap[2] = dmap[1] + map[1].size();
        CoreDev* dstats = (CoreDev*)(dbase + stats_off);
        // the host copy stays untouched until the stream sync below
        GA_HIP(ctx, hipMemcpyAsync(dbase, host.data(), up_bytes, hipMemcpyHostToDevice, ctx->stream));
        for (int r = 0; r < block_rounds; r++) {   // 3 launches per round for every unfinished state
            k_cfb_find<<<(unsigned)map[0].size(), 256, 0, ctx->stream>>>(djobs, dmap[0]);
            k_cfb_update<<<(unsigned)map[1].size(), 64, 0, ctx->stream>>>(djobs, dmap[1]);
            k_cfb_tail<<<(unsigned)map[2].size(), 256, 0, ctx->stream>>>(djobs, dmap[2]);
        }
        GA_TRY(ga_launch_check(ctx, "batched core round"));
        k_cfb_stats<<<grid_for(count, 256), 256, 0, ctx->stream>>>(djobs, count, dstats);
        GA_TRY(ga_launch_check(ctx, "k_cfb_stats"));
        GA_HIP(ctx, hipMemcpyAsync(hst.data(), dstats, (size_t)count * sizeof(CoreDev), hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        uint64_t nmax = 0;
        for (uint32_t i = 0; i < count; i++) {
            nmax = std::max<uint64_t>(nmax, cores[i]->g->n);
            if (hst[i].done) done[i] = 1;
            else alive[i] = hst[i].nalive;
        }
        if (guard > nmax + 16) return ga_fail(ctx, GA_ERR_STATE, "batched core decomposition did not finish");
        block_rounds = std::min(block_rounds * 2, 256);
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    for (uint32_t i = 0; i < count; i++) {
        char who[32];
        snprintf(who, sizeof who, "core %u: ", i);
        GA_TRY(cf_result(ctx, &hst[i], stats + 2 * i, who));
    }
    return GA_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// many small graphs in one launch: one workgroup runs a whole incremental decomposition
// ---------------------------------------------------------------------------------------------------------------
// The rounds are the same level-synchronous rounds as above (k = max(k, min alive tri); frontier = alive with
// tri <= k; subtract the frontier's triangles with the owner rule; retire), separated by workgroup barriers instead
// of kernel boundaries, so a graph costs no launches per round and a launch carries as many graphs as it has
// workgroups. Meant for graphs whose rounds are small (a workgroup walks the alive list every round).
struct CoreBlockDesc {
    const uint32_t* urow;
    const uint32_t* ucol;
    uint8_t* state;           // all 1 on entry
    uint32_t* core;
    uint32_t* tri;            // participation counts on entry
    uint32_t* front;          // n words
    uint32_t* alist;          // 2n words
    const CoreDev* st;        // T of the initial count
    uint32_t* out;            // 4 words: rounds, kmax, error, T (low 32 bits are enough: T < 2^32 is checked)
    uint32_t n;
};

__global__ void __launch_bounds__(1024) k_core_block(const CoreBlockDesc* __restrict__ descs) {
    const CoreBlockDesc d = descs[blockIdx.x];
    __shared__ uint32_t s_nf;
    __shared__ uint32_t wred[16];
    const uint32_t tid = threadIdx.x, T = blockDim.x, lane = lane_id(), wv = tid / 64, nw = T / 64;
    const uint32_t* __restrict__ urow = d.urow;
    const uint32_t* __restrict__ ucol = d.ucol;
    uint8_t* __restrict__ state = d.state;
    uint32_t* __restrict__ tri = d.tri;
    uint32_t* al = d.alist;
    uint32_t* al2 = d.alist + d.n;
    uint32_t na = d.n, nalive = d.n, rounds = 0, k_prev = 0, err = 0;
    for (uint32_t i = tid; i < na; i += T) al[i] = i;
    __syncthreads();
    while (nalive > 0) {   // uniform: every thread holds the same na / nalive
        uint32_t mn = 0xFFFFFFFFu;
        for (uint32_t i = tid; i < na; i += T) {
            const uint32_t v = al[i];
            if (state[v] == 1) mn = min(mn, tri[v]);
        }
        mn = wave_min_u32(mn);
        if (lane == 0) wred[wv] = mn;
        if (tid == 0) s_nf = 0;
        __syncthreads();
        uint32_t bm = 0xFFFFFFFFu;
        for (uint32_t w = 0; w < nw; w++) bm = min(bm, wred[w]);
        const uint32_t k = max(k_prev, bm);
        for (uint32_t base = wv * 64; base < na; base += T) {   // wave-uniform loop
            const uint32_t i = base + lane;
            uint32_t v = 0;
            const bool join = i < na && state[(v = al[i])] == 1 && tri[v] <= k;
            if (join) {
                state[v] = 2;
                d.core[v] = k;
            }
            const uint64_t mask = __ballot(join);
            if (!mask) continue;
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
            uint32_t pos = 0;
            if (lane == leader) pos = atomicAdd(&s_nf, (uint32_t)__popcll(mask));
            pos = __shfl(pos, leader, 64);
            if (join) d.front[pos + mask_rank(mask)] = v;
        }
        __syncthreads();
        const uint32_t nf = s_nf;
        if (nf == 0 || nf > nalive) { err = 1; break; }   // uniform; cannot happen: a vertex holds the minimum
        for (uint32_t fi = wv; fi < nf; fi += nw) {
            const uint32_t u = d.front[fi];
            if (tri[u] == 0) continue;   // in no alive triangle: nothing to subtract
            const uint32_t ub = urow[u], ue = urow[u + 1];
            for (uint32_t p = ub + lane; p < ue; p += 64) core_pair(urow, ucol, state, tri, u, p, ue);
        }
        __syncthreads();
        for (uint32_t i = tid; i < nf; i += T) state[d.front[i]] = 0;
        nalive -= nf;
        k_prev = k;
        rounds++;
        __syncthreads();
        if (nalive > 0 && 2 * nalive <= na) {   // the alive list has halved: compact it, in order
            uint32_t total = 0;
            for (uint32_t t0 = 0; t0 < na; t0 += T) {   // uniform loop
                const uint32_t i = t0 + tid;
                uint32_t v = 0;
                const bool keep = i < na && state[(v = al[i])] == 1;
                const uint64_t mask = __ballot(keep);
                if (lane == 0) wred[wv] = (uint32_t)__popcll(mask);
                __syncthreads();
                uint32_t off = total, add = 0;
                for (uint32_t w = 0; w < nw; w++) {
                    if (w < wv) off += wred[w];
                    add += wred[w];
                }
                if (keep) al2[off + mask_rank(mask)] = v;
                total += add;
                __syncthreads();
            }
            uint32_t* t = al;
            al = al2;
            al2 = t;
            na = total;
        }
    }
    if (tid == 0) {
        d.out[0] = rounds;
        d.out[1] = k_prev;
        d.out[2] = err;
        d.out[3] = (uint32_t)d.st->T;
    }
}

extern "C" int ga_core_run_many(ga_ctx* ctx, ga_core* const* cores, uint32_t count, const uint64_t* T_full, uint32_t* stats, float* ms) {
    if (!ctx || (count && (!cores || !stats))) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (!count) return GA_OK;
    for (uint32_t i = 0; i < count; i++) {
        ga_core* c = cores[i];
        if (!c || !c->incremental || c->ctx->device != ctx->device) return ga_fail(ctx, GA_ERR_ARG, "core %u: not an incremental state of this device", i);
        if (T_full && T_full[i] != UINT64_MAX && T_full[i] >= (1ull << 32))
            return ga_fail(ctx, GA_ERR_ARITH, "core %u: triangle count >= 2^32: uint32 tri[] counters could overflow", i);
        if (!c->blk) GA_TRY(ga_malloc(ctx, (void**)&c->blk, (2ull * c->g->n + 4) * 4));
        GA_TRY(ga_tc_ensure_plan(ctx, c->g));
    }
    const size_t desc_bytes = ((size_t)count * sizeof(CoreBlockDesc) + 255) & ~(size_t)255;
    const size_t need = desc_bytes + (size_t)count * 16;
    if (need > ctx->dcore_bytes) {
        ga_free(ctx->dcore);
        ctx->dcore = nullptr;
        ctx->dcore_bytes = 0;
        GA_TRY(ga_malloc(ctx, &ctx->dcore, need * 2));
        ctx->dcore_bytes = need * 2;
    }
    std::vector<CoreBlockDesc> h(count);
    uint32_t* dout = (uint32_t*)((char*)ctx->dcore + desc_bytes);
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (uint32_t i = 0; i < count; i++) {
        ga_core* c = cores[i];
        const uint32_t n = c->g->n;
        GA_TRY(ga_fill_u8(ctx, c->state, n, 1));
        GA_TRY(ga_fill_u32(ctx, c->core, n, UINT32_MAX));
        k_core_begin<false><<<cgrid(ctx, n), 256, 0, ctx->stream>>>(c->st, c->tri, n, 1);
        GA_TRY(ga_launch_check(ctx, "k_core_begin"));
        GA_TRY(ga_tc_launch(ctx, c->g, c->variant, &c->st->T, c->state, c->tri, nullptr, nullptr));
        h[i] = CoreBlockDesc{c->g->urow, c->g->ucol, c->state, c->core, c->tri, c->frontier, c->blk, c->st, dout + 4 * i, n};
    }
    GA_HIP(ctx, hipMemcpyAsync(ctx->dcore, h.data(), (size_t)count * sizeof(CoreBlockDesc), hipMemcpyHostToDevice, ctx->stream));
    k_core_block<<<count, 1024, 0, ctx->stream>>>((const CoreBlockDesc*)ctx->dcore);
    GA_TRY(ga_launch_check(ctx, "k_core_block"));
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<uint32_t> out((size_t)count * 4);
    GA_HIP(ctx, hipMemcpyAsync(out.data(), dout, (size_t)count * 16, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    for (uint32_t i = 0; i < count; i++) {
        if (out[4 * i + 2]) return ga_fail(ctx, GA_ERR_STATE, "core %u: bad round on the device (empty frontier)", i);
        if (T_full && T_full[i] != UINT64_MAX && (uint64_t)out[4 * i + 3] != T_full[i])
            return ga_fail(ctx, GA_ERR_STATE, "core %u: device triangle count differs from the given count", i);
        stats[2 * i] = out[4 * i];
        stats[2 * i + 1] = out[4 * i + 1];
    }
    return GA_OK;
}
