// ga_batch.hip - many independent graphs resident together and processed by shared launches.
//
// Layout: the batch is stored as ONE oriented CSR of the disjoint union of its graphs. Graph i owns the vertex range
// [voff[i], voff[i+1]) and, because rows are contiguous, the edge range [eoff[i], eoff[i+1]); gof[v] names v's graph.
// Orientation by (degree, id) and row order are unchanged by the id shift, so the union restricted to graph i is graph
// i's own CSR shifted by voff[i] (ga_batch_download_graph returns it with local ids).
// Counting: one hash-count launch set over the union, each work item adding its triangles to gcnt[gof[h]].
// Peeling: every graph runs the rounds of batch_peel_loop (densest.h:279-303) independently, but all active graphs
// share the same kernels per round: per-graph T / removed / nalive / best / round live in device arrays, a one-thread-
// per-graph epilogue applies the loop's rules, and a graph that has finished simply has no alive vertex left.
// The per-graph baselines run the same library once per graph (one count / one fused peel per graph).
#include <chrono>

#include "ga_internal.hpp"

enum { BP_ERR_ARITH = 2, BP_ERR_REMOVED = 3, BP_ERR_T = 1 };

struct BatchPeelDev {   // per-graph arrays of `count` entries, device
    unsigned long long *T, *removed, *nalive, *best_num, *best_den;
    uint32_t *round, *best_round, *rounds, *done, *error;
    uint32_t* active;      // graphs not finished yet (1 word)
    uint32_t* all_done;    // 1 when active == 0 (read by every kernel to become a no-op)
};

struct ga_batch {
    ga_ctx* ctx = nullptr;
    uint32_t count = 0;
    uint64_t total_n = 0, total_m = 0;
    ga_graph* u = nullptr;            // disjoint union
    uint32_t* gof = nullptr;          // graph of each union vertex
    std::vector<uint64_t> voff, eoff; // count + 1
    std::vector<uint32_t> gmax;       // max out-degree per graph
    std::vector<ga_graph*> slices;    // per-graph copies for the one-launch-per-graph baseline (lazily)
    // batched peel state
    uint8_t* alive = nullptr;
    uint32_t* tri = nullptr;
    uint32_t* rr = nullptr;
    void* pstate = nullptr;
    BatchPeelDev pd{};
    int graph_rounds = 8;             // peel rounds captured into one HIP graph (0: enqueue every round directly)
    float count_dev_ms = 0;           // device time of the last ga_batch_count's launches
    unsigned long long* gcnt = nullptr;  // per-graph counters of the batched count: 8 XCD copies + a total word (lazily)
};

namespace {

// graph id of pair i by binary search over the pair offsets
__device__ __forceinline__ uint32_t graph_of_pair(const unsigned long long* __restrict__ off, uint32_t count, unsigned long long i) {
    uint32_t lo = 0, hi = count;   // first g with off[g+1] > i
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid + 1] <= i) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Per-graph max id over non-loop pairs and "has an edge" flags. A graph's pairs are contiguous, so a wave almost
// always covers one graph: then one atomic per wave, else one per lane.
__global__ void k_batch_maxid(const uint32_t* __restrict__ pairs, unsigned long long k, const unsigned long long* __restrict__ off,
                              uint32_t count, uint32_t* __restrict__ mx, uint32_t* __restrict__ any) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long base = (unsigned long long)blockIdx.x * blockDim.x; base < k; base += stride) {   // wave-uniform loop
        const unsigned long long i = base + threadIdx.x;
        uint32_t g = 0xFFFFFFFFu, v = 0;
        bool has = false;
        if (i < k) {
            const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
            g = graph_of_pair(off, count, i);
            has = a != b;
            v = has ? max(a, b) : 0u;
        }
        const uint32_t g0 = __shfl(g, 0, 64);
        if (__all(g == g0 || g == 0xFFFFFFFFu)) {
            const uint32_t wm = wave_max_u32(v);
            const bool wa = __any(has);
            if (lane_id() == 0 && wa) {
                atomicMax(&mx[g0], wm);
                atomicOr(&any[g0], 1u);
            }
        } else if (has) {
            atomicMax(&mx[g], v);
            atomicOr(&any[g], 1u);
        }
    }
}

__global__ void k_batch_shift(uint32_t* __restrict__ pairs, unsigned long long k, const unsigned long long* __restrict__ off, uint32_t count,
                              const unsigned long long* __restrict__ voff) {
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (unsigned long long)gridDim.x * blockDim.x) {
        const uint32_t g = graph_of_pair(off, count, i);
        const uint32_t s = (uint32_t)voff[g];
        pairs[2 * i] += s;
        pairs[2 * i + 1] += s;
    }
}

__global__ void k_batch_gof(const unsigned long long* __restrict__ voff, uint32_t count, uint32_t* __restrict__ gof) {
    for (uint32_t g = blockIdx.x; g < count; g += gridDim.x)
        for (unsigned long long v = voff[g] + threadIdx.x; v < voff[g + 1]; v += blockDim.x) gof[v] = g;
}

__global__ void k_batch_edges(const uint32_t* __restrict__ rowptr, const unsigned long long* __restrict__ voff, uint32_t count,
                              unsigned long long* __restrict__ eoff) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= count; g += gridDim.x * blockDim.x) eoff[g] = rowptr[voff[g]];
}

__global__ void k_batch_maxout(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ gof, uint32_t n, uint32_t* __restrict__ gmax) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t d = rowptr[v + 1] - rowptr[v];
        if (d) atomicMax(&gmax[gof[v]], d);
    }
}

// Copy of graph slice [v0, v1) / [e0, e1) with local ids.
__global__ void k_batch_slice(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col, const uint32_t* __restrict__ src,
                              uint32_t v0, uint32_t n, uint32_t e0, uint32_t m, uint32_t* __restrict__ r_out, uint32_t* __restrict__ c_out,
                              uint32_t* __restrict__ s_out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
    for (uint32_t u = t; u <= n; u += st) r_out[u] = rowptr[v0 + u] - e0;
    for (uint32_t e = t; e < m; e += st) {
        c_out[e] = col[e0 + e] - v0;
        s_out[e] = src[e0 + e] - v0;
    }
}

// ---- batched peel ----
__global__ void k_bpeel_init(BatchPeelDev d, const unsigned long long* __restrict__ voff, uint32_t count) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x) {
        d.T[g] = 0;
        d.removed[g] = 0;
        d.nalive[g] = voff[g + 1] - voff[g];
        d.best_num[g] = 0;
        d.best_den[g] = 0;
        d.round[g] = 0;
        d.best_round[g] = 0;
        d.rounds[g] = 0;
        d.done[g] = d.nalive[g] == 0;
        d.error[g] = 0;
        if (d.done[g]) atomicSub(d.active, 1u);
    }
}

__global__ void k_bpeel_begin(BatchPeelDev d, uint32_t* __restrict__ tri, uint64_t n, uint32_t count) {
    if (*d.all_done) return;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) tri[v] = 0;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x) {
        d.T[g] = 0;
        d.removed[g] = 0;
    }
}

// Removal test of densest.h:151-154 per vertex with its own graph's T, nalive and round.
__global__ void k_bpeel_mark(BatchPeelDev d, const uint32_t* __restrict__ tri, uint8_t* __restrict__ alive, uint32_t* __restrict__ rr,
                             const uint32_t* __restrict__ gof, uint64_t n, unsigned long long num, unsigned long long den) {
    if (*d.all_done) return;
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t base = wave0 * 64; base < n; base += nwaves * 64) {   // wave-uniform loop
        const uint64_t v = base + lane_id();
        uint32_t g = 0;
        bool rem = false;
        if (v < n) {
            g = gof[v];
            if (alive[v]) {
                const unsigned long long rhs = 3ull * (den + num) * d.T[g];
                if ((unsigned long long)tri[v] * d.nalive[g] * den <= rhs) {
                    rem = true;
                    alive[v] = 0;
                    rr[v] = d.round[g];
                }
            }
        }
        const uint32_t g0 = __shfl(g, 0, 64);
        const bool uniform = __ballot(v < n && g != g0) == 0;
        if (uniform) {
            const uint64_t c = __popcll(__ballot(rem));
            if (lane_id() == 0 && c) atomicAdd(&d.removed[g0], (unsigned long long)c);
        } else if (rem) {
            atomicAdd(&d.removed[g], 1ull);
        }
    }
}

__device__ bool arith_ok(unsigned long long n, unsigned long long T0, unsigned long long num, unsigned long long den) {
    if (den < 1 || den > (1ull << 20) || num > (1ull << 20)) return false;
    if (T0 >= (1ull << 32)) return false;
    if (n > 0 && T0 > ~0ull / (n * den)) return false;
    if (T0 > ~0ull / (3ull * (den + num))) return false;
    return true;
}

// The per-round rules of batch_peel_loop, one thread per graph.
__global__ void k_bpeel_end(BatchPeelDev d, const unsigned long long* __restrict__ voff, const unsigned long long* __restrict__ T_full,
                            uint32_t count, unsigned long long num, unsigned long long den) {
    if (*d.all_done) return;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x) {
        if (d.done[g]) continue;
        const unsigned long long T = d.T[g], removed = d.removed[g], nalive = d.nalive[g];
        const uint32_t round = d.round[g];
        bool fail = false;
        if (round == 0) {
            if (T_full && T != T_full[g]) { d.error[g] = BP_ERR_T; fail = true; }
            else if (!arith_ok(voff[g + 1] - voff[g], T, num, den)) { d.error[g] = BP_ERR_ARITH; fail = true; }
        }
        if (!fail) {
            if (d.best_den[g] == 0 || T * d.best_den[g] > d.best_num[g] * nalive) {
                d.best_num[g] = T;
                d.best_den[g] = nalive;
                d.best_round[g] = round;
            }
            if (removed == 0 || removed > nalive) { d.error[g] = BP_ERR_REMOVED; fail = true; }
        }
        if (fail) {
            d.done[g] = 1;
        } else {
            d.nalive[g] = nalive - removed;
            d.rounds[g] = round + 1;
            d.round[g] = round + 1;
            if (nalive == removed) d.done[g] = 1;
        }
        if (d.done[g]) {
            if (atomicSub(d.active, 1u) == 1u) *d.all_done = 1;
        }
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

struct GraphExec {   // scoped captured graph and its executable
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphExec() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

int ensure_slices(ga_ctx* ctx, ga_batch* b) {
    if (!b->slices.empty()) return GA_OK;
    b->slices.assign(b->count, nullptr);
    for (uint32_t i = 0; i < b->count; i++) {
        const uint32_t n = (uint32_t)(b->voff[i + 1] - b->voff[i]), m = (uint32_t)(b->eoff[i + 1] - b->eoff[i]);
        ga_graph* g = new ga_graph;
        g->ctx = ctx;
        g->n = n;
        g->m = m;
        g->max_out = b->gmax[i];
        b->slices[i] = g;
        if (ga_malloc(ctx, (void**)&g->rowptr, ((size_t)n + 1) * 4) || ga_malloc(ctx, (void**)&g->col, std::max<size_t>(m, 1) * 4) ||
            ga_malloc(ctx, (void**)&g->src, std::max<size_t>(m, 1) * 4))
            return GA_ERR_OOM;
        k_batch_slice<<<grid_for(std::max(n + 1, m), 256, 2048), 256, 0, ctx->stream>>>(b->u->rowptr, b->u->col, b->u->src, (uint32_t)b->voff[i], n,
                                                                                         (uint32_t)b->eoff[i], m, g->rowptr, g->col, g->src);
        GA_TRY(ga_launch_check(ctx, "k_batch_slice"));
        g->plan.tune = b->u->plan.tune;
    }
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

int ensure_peel_state(ga_ctx* ctx, ga_batch* b) {
    if (b->pstate) return GA_OK;
    const size_t c = b->count;
    const size_t bytes = c * 8 * 5 + c * 4 * 5 + 64;
    GA_TRY(ga_malloc(ctx, &b->pstate, bytes));
    char* p = (char*)b->pstate;
    BatchPeelDev& d = b->pd;
    d.T = (unsigned long long*)p; p += c * 8;
    d.removed = (unsigned long long*)p; p += c * 8;
    d.nalive = (unsigned long long*)p; p += c * 8;
    d.best_num = (unsigned long long*)p; p += c * 8;
    d.best_den = (unsigned long long*)p; p += c * 8;
    d.round = (uint32_t*)p; p += c * 4;
    d.best_round = (uint32_t*)p; p += c * 4;
    d.rounds = (uint32_t*)p; p += c * 4;
    d.done = (uint32_t*)p; p += c * 4;
    d.error = (uint32_t*)p; p += c * 4;
    d.active = (uint32_t*)p;
    d.all_done = d.active + 1;
    const size_t n = b->total_n;
    GA_TRY(ga_malloc(ctx, (void**)&b->alive, n + 16));
    GA_TRY(ga_malloc(ctx, (void**)&b->tri, (n + 4) * 4));
    GA_TRY(ga_malloc(ctx, (void**)&b->rr, (n + 1) * 4));
    return GA_OK;
}

}  // namespace

extern "C" int ga_batch_create(ga_ctx* ctx, uint32_t count, const uint64_t* pair_offsets, const uint32_t* pairs, ga_batch** out,
                               float* build_ms) {
    *out = nullptr;
    if (!ctx || !count || !pair_offsets) return GA_ERR_ARG;
    const uint64_t k = pair_offsets[count] - pair_offsets[0];
    if (pair_offsets[0] != 0) return ga_fail(ctx, GA_ERR_ARG, "pair_offsets[0] must be 0");
    for (uint32_t i = 0; i < count; i++)
        if (pair_offsets[i + 1] < pair_offsets[i]) return ga_fail(ctx, GA_ERR_ARG, "pair_offsets decrease at %u", i);
    if (k && !pairs) return ga_fail(ctx, GA_ERR_ARG, "pairs is NULL");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    const double t0 = wall_ms();
    DevBuf dp, doff, dmx, dvoff;
    GA_TRY(ga_malloc(ctx, &dp.p, std::max<uint64_t>(k, 1) * 8));
    GA_TRY(ga_malloc(ctx, &doff.p, ((size_t)count + 1) * 8));
    GA_TRY(ga_malloc(ctx, &dmx.p, (size_t)count * 8));
    GA_TRY(ga_malloc(ctx, &dvoff.p, ((size_t)count + 1) * 8));
    if (k) GA_HIP(ctx, hipMemcpyAsync(dp.p, pairs, k * 8, hipMemcpyHostToDevice, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(doff.p, pair_offsets, ((size_t)count + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    uint32_t* mx = (uint32_t*)dmx.p;
    uint32_t* any = mx + count;
    GA_TRY(ga_fill_u32(ctx, mx, 2 * (size_t)count, 0));
    const unsigned long long* off = (const unsigned long long*)doff.p;
    if (k) {
        k_batch_maxid<<<grid_for(k, 256, 8192), 256, 0, ctx->stream>>>((const uint32_t*)dp.p, k, off, count, mx, any);
        GA_TRY(ga_launch_check(ctx, "k_batch_maxid"));
    }
    std::vector<uint32_t> hmx(2 * (size_t)count);
    GA_HIP(ctx, hipMemcpyAsync(hmx.data(), mx, 8 * (size_t)count, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ga_batch* b = new ga_batch;
    b->ctx = ctx;
    b->count = count;
    b->voff.assign((size_t)count + 1, 0);
    for (uint32_t i = 0; i < count; i++) {
        const uint64_t n = hmx[count + i] ? (uint64_t)hmx[i] + 1 : 1;   // load_graph: n = max_id + 1, empty -> 1
        b->voff[i + 1] = b->voff[i] + n;
    }
    b->total_n = b->voff[count];
    auto fail_b = [&](int rc) { ga_batch_destroy(b); return rc; };
    if (b->total_n > UINT32_MAX - 1) return fail_b(ga_fail(ctx, GA_ERR_ARG, "batch has %llu vertices; at most 2^32-2 fit", (unsigned long long)b->total_n));
    if (hipMemcpyAsync(dvoff.p, b->voff.data(), ((size_t)count + 1) * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return fail_b(ga_fail(ctx, GA_ERR_HIP, "voff upload failed"));
    int rc;
    if (k) {
        k_batch_shift<<<grid_for(k, 256, 8192), 256, 0, ctx->stream>>>((uint32_t*)dp.p, k, off, count, (const unsigned long long*)dvoff.p);
        if ((rc = ga_launch_check(ctx, "k_batch_shift"))) return fail_b(rc);
    }
    float ms = 0;
    if ((rc = ga_build_from_device_pairs(ctx, (const uint32_t*)dp.p, k, (int64_t)b->total_n, false, &b->u, &ms))) return fail_b(rc);
    b->total_m = b->u->m;
    ga_free(dp.p);   // the raw pairs are no longer needed: free them before the count plan is built
    dp.p = nullptr;
    if (ga_malloc(ctx, (void**)&b->gof, (b->total_n + 1) * 4)) return fail_b(GA_ERR_OOM);
    k_batch_gof<<<std::min<uint32_t>(count, 65535), 256, 0, ctx->stream>>>((const unsigned long long*)dvoff.p, count, b->gof);
    if ((rc = ga_launch_check(ctx, "k_batch_gof"))) return fail_b(rc);
    // edge offsets and per-graph max out-degree
    DevBuf deoff, dgmax;
    if (ga_malloc(ctx, &deoff.p, ((size_t)count + 1) * 8) || ga_malloc(ctx, &dgmax.p, (size_t)count * 4)) return fail_b(GA_ERR_OOM);
    k_batch_edges<<<grid_for(count + 1, 256), 256, 0, ctx->stream>>>(b->u->rowptr, (const unsigned long long*)dvoff.p, count,
                                                                     (unsigned long long*)deoff.p);
    if ((rc = ga_fill_u32(ctx, (uint32_t*)dgmax.p, count, 0))) return fail_b(rc);
    k_batch_maxout<<<grid_for(b->total_n, 256, 8192), 256, 0, ctx->stream>>>(b->u->rowptr, b->gof, (uint32_t)b->total_n, (uint32_t*)dgmax.p);
    if ((rc = ga_launch_check(ctx, "k_batch_maxout"))) return fail_b(rc);
    b->eoff.assign((size_t)count + 1, 0);
    b->gmax.assign(count, 0);
    if (hipMemcpyAsync(b->eoff.data(), deoff.p, ((size_t)count + 1) * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(b->gmax.data(), dgmax.p, (size_t)count * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail_b(ga_fail(ctx, GA_ERR_HIP, "batch offsets download failed"));
    b->u->plan.group = b->gof;   // shape-ordered items stay graph-major (cache locality per graph)
    b->u->plan.group_bits = 1;
    while (b->u->plan.group_bits < 31 && (1u << b->u->plan.group_bits) < count) b->u->plan.group_bits++;
    if ((rc = ga_tc_ensure_plan(ctx, b->u, true))) return fail_b(rc);
    if (build_ms) *build_ms = (float)(wall_ms() - t0);
    *out = b;
    return GA_OK;
}

extern "C" void ga_batch_destroy(ga_batch* b) {
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    ga_graph_destroy(b->u);
    ga_free(b->gof);
    for (ga_graph* g : b->slices) ga_graph_destroy(g);
    ga_free(b->alive);
    ga_free(b->tri);
    ga_free(b->rr);
    ga_free(b->pstate);
    ga_free(b->gcnt);
    delete b;
}

extern "C" int ga_batch_info(const ga_batch* b, uint32_t* count, uint64_t* total_n, uint64_t* total_m) {
    if (!b) return GA_ERR_ARG;
    if (count) *count = b->count;
    if (total_n) *total_n = b->total_n;
    if (total_m) *total_m = b->total_m;
    return GA_OK;
}

extern "C" int ga_batch_graph_info(const ga_batch* b, uint32_t i, uint32_t* n, uint32_t* m, uint32_t* max_out_degree) {
    if (!b || i >= b->count) return GA_ERR_ARG;
    if (n) *n = (uint32_t)(b->voff[i + 1] - b->voff[i]);
    if (m) *m = (uint32_t)(b->eoff[i + 1] - b->eoff[i]);
    if (max_out_degree) *max_out_degree = b->gmax[i];
    return GA_OK;
}

extern "C" int ga_batch_download_graph(ga_ctx* ctx, ga_batch* b, uint32_t i, uint32_t* rowptr, uint32_t* col, uint32_t* src) {
    if (!ctx || !b || i >= b->count) return GA_ERR_ARG;
    const uint32_t n = (uint32_t)(b->voff[i + 1] - b->voff[i]), m = (uint32_t)(b->eoff[i + 1] - b->eoff[i]);
    DevBuf r, c, s;
    GA_TRY(ga_malloc(ctx, &r.p, ((size_t)n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &c.p, std::max<size_t>(m, 1) * 4));
    GA_TRY(ga_malloc(ctx, &s.p, std::max<size_t>(m, 1) * 4));
    k_batch_slice<<<grid_for(std::max(n + 1, m), 256, 2048), 256, 0, ctx->stream>>>(b->u->rowptr, b->u->col, b->u->src, (uint32_t)b->voff[i], n,
                                                                                     (uint32_t)b->eoff[i], m, (uint32_t*)r.p, (uint32_t*)c.p,
                                                                                     (uint32_t*)s.p);
    GA_TRY(ga_launch_check(ctx, "k_batch_slice"));
    if (rowptr) GA_HIP(ctx, hipMemcpyAsync(rowptr, r.p, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (col && m) GA_HIP(ctx, hipMemcpyAsync(col, c.p, (size_t)m * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (src && m) GA_HIP(ctx, hipMemcpyAsync(src, s.p, (size_t)m * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

// All graphs counted by one launch set over the union; T has `count` entries. ms = host wall time (launch to result).
extern "C" int ga_batch_count(ga_ctx* ctx, ga_batch* b, uint64_t* T, float* ms) {
    if (!ctx || !b || !T) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    // One copy of the per-graph counters per XCD (workgroups b with the same b mod 8 add to the same copy), so the
    // items of one graph in flight on different XCDs do not meet on one address; folded into copy 0 after the launch.
    if (!b->gcnt) GA_TRY(ga_malloc(ctx, (void**)&b->gcnt, ((size_t)b->count * 8 + 1) * 8));   // kept for later counts
    unsigned long long* gc = b->gcnt;
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double t0 = wall_ms();
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)gc, (size_t)b->count * 8 + 1, 0));
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(ga_tc_launch(ctx, b->u, GA_TC_HASH, gc + (size_t)b->count * 8, nullptr, nullptr, nullptr, nullptr, b->gof, gc, b->count));
    GA_TRY(ga_tc_fold_xcd(ctx, gc, b->count));
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(T, gc, (size_t)b->count * 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ms) *ms = (float)(wall_ms() - t0);
    GA_HIP(ctx, hipEventElapsedTime(&b->count_dev_ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}

namespace {
// Sum over oriented edges of min(d+(u), d+(v)): the probe elements of an exact count (one pass over src / col).
__global__ void k_sum_min_outdeg(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                                 uint64_t m, unsigned long long* __restrict__ out) {
    uint64_t local = 0;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = src[e], v = col[e];
        local += min(rowptr[u + 1] - rowptr[u], rowptr[v + 1] - rowptr[v]);
    }
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(out, (unsigned long long)local);
}
}  // namespace

extern "C" int ga_batch_work(ga_ctx* ctx, ga_batch* b, uint64_t* summin, float* last_count_device_ms) {
    if (!ctx || !b) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (summin) {
        uint64_t* d = ctx->dsmall;
        GA_TRY(ga_fill_u64(ctx, d, 1, 0));
        if (b->u->m)
            k_sum_min_outdeg<<<grid_for(b->u->m, 256, 16384), 256, 0, ctx->stream>>>(b->u->rowptr, b->u->src, b->u->col, b->u->m,
                                                                                    (unsigned long long*)d);
        GA_TRY(ga_launch_check(ctx, "k_sum_min_outdeg"));
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d, 8, hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        *summin = ctx->hstat[0];
    }
    if (last_count_device_ms) *last_count_device_ms = b->count_dev_ms;
    return GA_OK;
}

extern "C" int ga_batch_set_tuning(ga_ctx* ctx, ga_batch* b, const ga_tc_tuning* t) {
    if (!ctx || !b || !t) return GA_ERR_ARG;
    GA_TRY(ga_tc_set_tuning(ctx, b->u, t));
    for (ga_graph* g : b->slices) GA_TRY(ga_tc_set_tuning(ctx, g, t));
    return GA_OK;
}

extern "C" int ga_batch_get_tuning(const ga_batch* b, ga_tc_tuning* out) {
    if (!b || !out) return GA_ERR_ARG;
    return ga_tc_get_tuning(b->u, out);
}

extern "C" int ga_batch_tune(ga_ctx* ctx, ga_batch* b, ga_tc_tuning* best, float* best_ms) {
    if (!ctx || !b) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (!b->gcnt) GA_TRY(ga_malloc(ctx, (void**)&b->gcnt, ((size_t)b->count * 8 + 1) * 8));
    GA_TRY(ga_tc_tune_counts(ctx, b->u, b->gof, b->gcnt, b->count, best, best_ms));   // what ga_batch_count runs
    for (ga_graph* g : b->slices) GA_TRY(ga_tc_set_tuning(ctx, g, &b->u->plan.tune));
    return GA_OK;
}

// Baseline: one count per graph (each graph a separate resident CSR with its own plan).
extern "C" int ga_batch_count_per_graph(ga_ctx* ctx, ga_batch* b, uint64_t* T, float* ms) {
    if (!ctx || !b || !T) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(ensure_slices(ctx, b));
    for (uint32_t i = 0; i < b->count; i++) GA_TRY(ga_tc_ensure_plan(ctx, b->slices[i], true));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double t0 = wall_ms();
    for (uint32_t i = 0; i < b->count; i++) {
        uint64_t t = 0;
        GA_TRY(ga_tc_count(ctx, b->slices[i], GA_TC_HASH, &t, nullptr));
        T[i] = t;
    }
    if (ms) *ms = (float)(wall_ms() - t0);
    return GA_OK;
}

static int peel_error(ga_ctx* ctx, uint32_t g, uint32_t code) {
    switch (code) {
        case BP_ERR_T: return ga_fail(ctx, GA_ERR_STATE, "graph %u: device triangle count in round 0 differs from the given count", g);
        case BP_ERR_ARITH: return ga_fail(ctx, GA_ERR_ARITH, "graph %u: the exact peel arithmetic bounds do not hold", g);
        default: return ga_fail(ctx, GA_ERR_STATE, "graph %u: bad removed count from device", g);
    }
}

// All graphs peeled by shared launches per round. stats: 4 per graph {best_num, best_den, best_round, rounds};
// removed_round: total_n entries (graph i at its vertex offset) or NULL. T_full: per-graph counts or NULL.
extern "C" int ga_batch_peel(ga_ctx* ctx, ga_batch* b, uint64_t num, uint64_t den, const uint64_t* T_full, uint64_t* stats,
                             uint32_t* removed_round, float* ms) {
    if (!ctx || !b || !stats) return GA_ERR_ARG;
    if (den < 1 || den > (1ull << 20) || num > (1ull << 20)) return ga_fail(ctx, GA_ERR_ARITH, "eps = num/den needs 1 <= den <= 2^20 and num <= 2^20");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(ensure_peel_state(ctx, b));
    DevBuf dvoff, dtf, scratch;
    GA_TRY(ga_malloc(ctx, &dvoff.p, ((size_t)b->count + 1) * 8));
    GA_TRY(ga_malloc(ctx, &scratch.p, 8));
    GA_HIP(ctx, hipMemcpyAsync(dvoff.p, b->voff.data(), ((size_t)b->count + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    if (T_full) {
        GA_TRY(ga_malloc(ctx, &dtf.p, (size_t)b->count * 8));
        GA_HIP(ctx, hipMemcpyAsync(dtf.p, T_full, (size_t)b->count * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    const BatchPeelDev& d = b->pd;
    const uint64_t n = b->total_n;
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double t0 = wall_ms();
    GA_TRY(ga_fill_u8(ctx, b->alive, n, 1));
    GA_TRY(ga_fill_u32(ctx, b->rr, n, UINT32_MAX));
    GA_TRY(ga_fill_u32(ctx, d.active, 1, b->count));
    GA_TRY(ga_fill_u32(ctx, d.all_done, 1, 0));
    k_bpeel_init<<<grid_for(b->count, 256), 256, 0, ctx->stream>>>(d, (const unsigned long long*)dvoff.p, b->count);
    GA_TRY(ga_launch_check(ctx, "k_bpeel_init"));
    const unsigned gv = ctx_grid(ctx, grid_for(n, 256, (unsigned)std::max(ctx->cus, 1) * 8u));
    // One round: begin, the participation count (its bucket launches), mark, end. No launch takes a host value that
    // changes between rounds, and a finished batch turns every launch into a no-op through d.all_done, so rounds can be
    // enqueued ahead of the readback, and a block of them can be captured once and replayed.
    auto round = [&]() -> int {
        k_bpeel_begin<<<gv, 256, 0, ctx->stream>>>(d, b->tri, n, b->count);
        GA_TRY(ga_tc_launch(ctx, b->u, GA_TC_HASH, (unsigned long long*)scratch.p, b->alive, b->tri, nullptr, d.all_done, b->gof, d.T));
        k_bpeel_mark<<<gv, 256, 0, ctx->stream>>>(d, b->tri, b->alive, b->rr, b->gof, n, num, den);
        k_bpeel_end<<<grid_for(b->count, 256), 256, 0, ctx->stream>>>(d, (const unsigned long long*)dvoff.p,
                                                                      (const unsigned long long*)dtf.p, b->count, num, den);
        return ga_launch_check(ctx, "batched peel round");
    };
    GraphExec ge;
    if (b->graph_rounds > 0) {
        GA_TRY(ga_tc_ensure_plan(ctx, b->u));   // built (and allocated) before the capture, never inside it
        GA_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        int rc = GA_OK;
        for (int r = 0; r < b->graph_rounds && !rc; r++) rc = round();
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &ge.graph);
        if (rc) return rc;
        if (ce != hipSuccess) return ga_fail(ctx, GA_ERR_HIP, "capture of the peel rounds failed: %s", hipGetErrorString(ce));
        GA_HIP(ctx, hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
    }
    int block = 1;   // rounds (direct) or graph launches (captured) per readback, doubling
    for (uint64_t guard = 0;; guard++) {
        for (int r = 0; r < block; r++) {
            if (ge.exec) GA_HIP(ctx, hipGraphLaunch(ge.exec, ctx->stream));
            else GA_TRY(round());
        }
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d.active, 8, hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (((const uint32_t*)ctx->hstat)[1]) break;
        if (guard > n + 16) return ga_fail(ctx, GA_ERR_STATE, "batched peel did not finish");
        block = std::min(block * 2, ge.exec ? 8 : 64);
    }
    std::vector<unsigned long long> bn(b->count), bd(b->count);
    std::vector<uint32_t> br(b->count), rounds(b->count), err(b->count);
    GA_HIP(ctx, hipMemcpyAsync(bn.data(), d.best_num, (size_t)b->count * 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(bd.data(), d.best_den, (size_t)b->count * 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(br.data(), d.best_round, (size_t)b->count * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(rounds.data(), d.rounds, (size_t)b->count * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(err.data(), d.error, (size_t)b->count * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ms) *ms = (float)(wall_ms() - t0);
    for (uint32_t g = 0; g < b->count; g++) {
        if (err[g]) return peel_error(ctx, g, err[g]);
        stats[4 * g] = bn[g];
        stats[4 * g + 1] = bd[g];
        stats[4 * g + 2] = br[g];
        stats[4 * g + 3] = rounds[g];
    }
    if (removed_round && n) GA_HIP(ctx, hipMemcpy(removed_round, b->rr, n * 4, hipMemcpyDeviceToHost));
    return GA_OK;
}

extern "C" int ga_batch_set_peel_graph(ga_batch* b, int rounds) {
    if (!b || rounds < 0 || rounds > 64) return GA_ERR_ARG;
    b->graph_rounds = rounds;
    return GA_OK;
}

// Baseline: one fused peel (ga_peel_run) per graph.
extern "C" int ga_batch_peel_per_graph(ga_ctx* ctx, ga_batch* b, uint64_t num, uint64_t den, const uint64_t* T_full, uint64_t* stats,
                                       uint32_t* removed_round, float* ms) {
    if (!ctx || !b || !stats) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(ensure_slices(ctx, b));
    std::vector<ga_peel*> peels(b->count, nullptr);
    int rc = GA_OK;
    for (uint32_t i = 0; i < b->count && !rc; i++) rc = ga_peel_create(ctx, b->slices[i], &peels[i]);
    double t = 0;
    for (uint32_t i = 0; i < b->count && !rc; i++) {
        const double t0 = wall_ms();
        rc = ga_peel_run(peels[i], T_full ? T_full[i] : UINT64_MAX, num, den, stats + 4 * i, nullptr);
        t += wall_ms() - t0;
        if (rc) ga_fail(ctx, rc, "graph %u: %s", i, std::string(ctx->err).c_str());
        if (!rc && removed_round) rc = ga_peel_download_removed_round(peels[i], removed_round + b->voff[i]);
    }
    for (ga_peel* p : peels) ga_peel_destroy(p);
    if (ms) *ms = (float)t;
    return rc;
}
