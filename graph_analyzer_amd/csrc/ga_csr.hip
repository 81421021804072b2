// ga_csr.hip - oriented CSR construction on the GPU (graph.h:33-96 and densest.h:162-171 move here).
//
// Raw pairs -> packed 64-bit keys (min << nb | max, self loops mapped to an all-ones sentinel that sorts last)
// -> LSD radix sort over 2*nb bits -> adjacent-unique flags + exclusive scan (dedup) -> undirected degrees (atomics)
// -> orientation key (from << nb | to) by the (degree, id) rule -> second radix sort -> rowptr by histogram + scan,
// col / src unpacked from the sorted keys. Rows come out ascending, so rowptr/col/src are byte-identical to the
// reference's host builder. nb = bits of (n-1), so a key never needs more than 2*nb <= 64 bits.
#include "ga_internal.hpp"

namespace {

inline int bits_for(uint32_t n) {   // bits to represent ids 0..n-1 (>= 1)
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) < (uint64_t)n) b++;
    return b;
}

__global__ void k_max_id(const uint32_t* __restrict__ pairs, uint64_t k, uint32_t* __restrict__ out, uint32_t* __restrict__ any) {
    uint32_t mx = 0, seen = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        if (a != b) { mx = max(mx, max(a, b)); seen = 1; }
    }
    mx = wave_max_u32(mx);
    seen = wave_max_u32(seen);
    if (lane_id() == 0) {
        if (mx) atomicMax(out, mx);
        if (seen) atomicOr(any, 1u);
    }
}

// key = (min << nb) | max; self loops -> sentinel (all ones in 2*nb bits). strict: count loops and ids >= n.
__global__ void k_pack(const uint32_t* __restrict__ pairs, uint64_t k, int nb, uint32_t n, uint64_t* __restrict__ keys,
                       uint32_t* __restrict__ bad) {
    const uint64_t sentinel = (nb == 32) ? ~0ull : ((1ull << (2 * nb)) - 1ull);
    uint32_t nbad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        if (a == b || a >= n || b >= n) {
            keys[i] = sentinel;
            nbad++;
        } else {
            const uint32_t u = min(a, b), v = max(a, b);
            keys[i] = ((uint64_t)u << nb) | v;
        }
    }
    nbad = (uint32_t)wave_sum_u64(nbad);
    if (lane_id() == 0 && nbad) atomicAdd(bad, nbad);
}

// flag[i] = 1 if key[i] is a real pair and differs from key[i-1]; dup counts repeated real pairs.
__global__ void k_unique_flags(const uint64_t* __restrict__ keys, uint64_t k, uint64_t sentinel, uint32_t* __restrict__ flag,
                               uint32_t* __restrict__ dup) {
    uint32_t nd = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = keys[i];
        const bool real = x != sentinel;
        const bool first = i == 0 || keys[i - 1] != x;
        flag[i] = real && first;
        nd += real && !first;
    }
    nd = (uint32_t)wave_sum_u64(nd);
    if (lane_id() == 0 && nd) atomicAdd(dup, nd);
}

// Compacts unique keys and counts undirected degrees. Keys are sorted, so the smaller endpoints of a wave's edges form
// runs: each run adds its length with one atomic (hubs with low ids would otherwise serialise on one address).
__global__ void k_compact_degrees(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flag,
                                  const uint32_t* __restrict__ pos, uint64_t k, int nb, uint64_t* __restrict__ ukeys,
                                  uint32_t* __restrict__ deg) {
    const uint64_t mask = (1ull << nb) - 1ull;
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t base = wave0 * 64; base < k; base += nwaves * 64) {   // wave-uniform loop
        const uint64_t i = base + lane_id();
        const bool f = i < k && flag[i];
        uint32_t u = 0;
        if (f) {
            const uint64_t x = keys[i];
            ukeys[pos[i]] = x;
            u = (uint32_t)(x >> nb);
            atomicAdd(&deg[(uint32_t)(x & mask)], 1u);
        }
        uint64_t todo = __ballot(f);
        while (todo) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)todo) - 1;
            const uint32_t ul = __shfl(u, leader, 64);
            const uint64_t same = __ballot(f && u == ul);
            if (lane_id() == leader) atomicAdd(&deg[ul], (uint32_t)__popcll(same));
            todo &= ~same;
        }
    }
}

// Orientation by (degree, id): from = the endpoint with the smaller (deg, id) (graph.h:43-45).
__global__ void k_orient(uint64_t* __restrict__ keys, uint32_t m, int nb, const uint32_t* __restrict__ deg,
                         uint32_t* __restrict__ outdeg) {
    const uint64_t mask = (1ull << nb) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = keys[i];
        const uint32_t u = (uint32_t)(x >> nb), v = (uint32_t)(x & mask);   // u < v
        const uint32_t du = deg[u], dv = deg[v];
        const bool fwd = du < dv || (du == dv);                               // equal degree: smaller id (u) first
        const uint32_t from = fwd ? u : v, to = fwd ? v : u;
        keys[i] = ((uint64_t)from << nb) | to;
        atomicAdd(&outdeg[from], 1u);
    }
}

__global__ void k_unpack(const uint64_t* __restrict__ keys, uint32_t m, int nb, uint32_t* __restrict__ col,
                         uint32_t* __restrict__ src) {
    const uint64_t mask = (1ull << nb) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = keys[i];
        col[i] = (uint32_t)(x & mask);
        if (src) src[i] = (uint32_t)(x >> nb);
    }
}

// src[e] = row owner, from rowptr (one lane per row, rows are short after orientation).
__global__ void k_fill_src(const uint32_t* __restrict__ rowptr, uint32_t n, uint32_t* __restrict__ src) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
        for (uint32_t e = rowptr[u]; e < rowptr[u + 1]; e++) src[e] = u;
}

__global__ void k_outdeg(const uint32_t* __restrict__ rowptr, uint32_t n, uint32_t* __restrict__ d) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) d[u] = rowptr[u + 1] - rowptr[u];
}

// Undirected keys: both directions of every oriented edge.
__global__ void k_both_directions(const uint32_t* __restrict__ src, const uint32_t* __restrict__ col, uint32_t m, int nb,
                                  uint64_t* __restrict__ keys) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = src[i], v = col[i];
        keys[2 * i] = ((uint64_t)u << nb) | v;
        keys[2 * i + 1] = ((uint64_t)v << nb) | u;
    }
}

__global__ void k_count_rows(const uint32_t* __restrict__ src, uint32_t m, uint32_t* __restrict__ cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) atomicAdd(&cnt[src[i]], 1u);
}

// v9 bucket key: NB-1-floor(log2(min(d+u, d+v)+1)) so ascending key = heaviest bucket first (graph.h:156-171).
__global__ void k_edge_bucket(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                              uint32_t m, uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const uint32_t u = src[e], v = col[e];
        const uint32_t du = rowptr[u + 1] - rowptr[u], dv = rowptr[v + 1] - rowptr[v];
        const uint64_t w = (uint64_t)min(du, dv) + 1ull;
        const uint32_t b = 63u - (uint32_t)__clzll(w);   // floor(log2(w))
        key[e] = 32u - b;
        idx[e] = e;
    }
}

struct DevBuf {   // scoped device allocation
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

int finish_rows(ga_ctx* ctx, ga_graph* g, uint32_t* outdeg) {
    // rowptr = exclusive scan of out-degrees over n+1 entries (outdeg[n] = 0)
    GA_TRY(ga_exclusive_scan_u32(ctx, outdeg, g->rowptr, (size_t)g->n + 1));
    GA_TRY(ga_max_u32(ctx, outdeg, g->n, &g->max_out));
    return GA_OK;
}

}  // namespace

// Device pairs -> oriented CSR (device). n_in < 0: n = max_id + 1 over non-loop pairs. strict: pairs must be unique,
// loop-free and < n (build_graph_from_edges); otherwise loops, duplicates and both orientations are folded.
int ga_build_from_device_pairs(ga_ctx* ctx, const uint32_t* dp, uint64_t k, int64_t n_in, bool strict, ga_graph** out, float* build_ms) {
    *out = nullptr;
    DevBuf keys_a, keys_b, flag, pos, counters;
    GA_TRY(ga_malloc(ctx, &counters.p, 64));
    uint32_t* cnt = (uint32_t*)counters.p;   // [0] max id, [1] any, [2] bad, [3] dup
    GA_TRY(ga_fill_u32(ctx, cnt, 16, 0));
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));

    uint32_t n;
    if (n_in < 0) {
        if (k) {
            k_max_id<<<grid_for(k, 256, 4096), 256, 0, ctx->stream>>>(dp, k, cnt, cnt + 1);
            GA_TRY(ga_launch_check(ctx, "k_max_id"));
        }
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const uint32_t mx = (uint32_t)(ctx->hstat[0] & 0xFFFFFFFFu);
        if (mx == UINT32_MAX) return ga_fail(ctx, GA_ERR_ARG, "vertex id 4294967295 does not fit n = max_id + 1 in uint32");
        n = mx + 1;   // empty list: n = 1 (graph.h:76,94)
    } else {
        n = (uint32_t)n_in;
    }
    const int nb = bits_for(n);
    const uint64_t sentinel = (nb == 32) ? ~0ull : ((1ull << (2 * nb)) - 1ull);

    GA_TRY(ga_malloc(ctx, &keys_a.p, std::max<uint64_t>(k, 1) * 8));
    GA_TRY(ga_malloc(ctx, &keys_b.p, std::max<uint64_t>(k, 1) * 8));
    GA_TRY(ga_malloc(ctx, &flag.p, (std::max<uint64_t>(k, 1) + 1) * 4));
    GA_TRY(ga_malloc(ctx, &pos.p, (std::max<uint64_t>(k, 1) + 1) * 4));
    uint64_t* ka = (uint64_t*)keys_a.p;
    uint64_t* kb = (uint64_t*)keys_b.p;
    uint32_t m = 0;
    ga_graph* g = new ga_graph;
    g->ctx = ctx;
    g->n = n;
    auto fail_g = [&](int rc) { ga_graph_destroy(g); return rc; };
    uint32_t* deg = nullptr;
    if (ga_malloc(ctx, (void**)&g->rowptr, ((size_t)n + 1) * 4)) return fail_g(GA_ERR_OOM);
    DevBuf degb, outdegb;
    if (ga_malloc(ctx, &degb.p, ((size_t)n + 1) * 4) || ga_malloc(ctx, &outdegb.p, ((size_t)n + 1) * 4)) return fail_g(GA_ERR_OOM);
    deg = (uint32_t*)degb.p;
    uint32_t* outdeg = (uint32_t*)outdegb.p;
    int rc;
    if ((rc = ga_fill_u32(ctx, deg, (size_t)n + 1, 0)) || (rc = ga_fill_u32(ctx, outdeg, (size_t)n + 1, 0))) return fail_g(rc);

    if (k) {
        k_pack<<<grid_for(k, 256, 8192), 256, 0, ctx->stream>>>(dp, k, nb, n, ka, cnt + 2);
        if ((rc = ga_launch_check(ctx, "k_pack"))) return fail_g(rc);
        uint64_t* sorted = nullptr;
        if ((rc = ga_sort_u64(ctx, ka, kb, k, 2 * nb, &sorted))) return fail_g(rc);
        uint64_t* other = sorted == ka ? kb : ka;
        uint32_t* fl = (uint32_t*)flag.p;
        uint32_t* ps = (uint32_t*)pos.p;
        k_unique_flags<<<grid_for(k, 256, 8192), 256, 0, ctx->stream>>>(sorted, k, sentinel, fl, cnt + 3);
        if ((rc = ga_launch_check(ctx, "k_unique_flags"))) return fail_g(rc);
        if ((rc = ga_fill_u32(ctx, fl + k, 1, 0))) return fail_g(rc);
        if ((rc = ga_exclusive_scan_u32(ctx, fl, ps, k + 1))) return fail_g(rc);
        if (hipMemcpyAsync(ctx->hstat, ps + k, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(ctx->hstat + 1, cnt + 2, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail_g(ga_fail(ctx, GA_ERR_HIP, "reading the unique-edge count failed"));
        const uint64_t mu = ctx->hstat[0] & 0xFFFFFFFFu;
        const uint32_t bad = (uint32_t)(ctx->hstat[1] & 0xFFFFFFFFu), dup = (uint32_t)(ctx->hstat[1] >> 32);
        if (strict && (bad || dup))
            return fail_g(ga_fail(ctx, GA_ERR_ARG, "build_graph_from_edges needs unique, loop-free pairs with ids < n "
                                  "(%u self loops / out-of-range ids, %u duplicates)", bad, dup));
        if (mu > UINT32_MAX) return fail_g(ga_fail(ctx, GA_ERR_ARG, "more than 2^32-1 unique edges"));
        m = (uint32_t)mu;
        k_compact_degrees<<<grid_for(k, 256, 8192), 256, 0, ctx->stream>>>(sorted, fl, ps, k, nb, other, deg);
        if ((rc = ga_launch_check(ctx, "k_compact_degrees"))) return fail_g(rc);
        if (m) {
            k_orient<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(other, m, nb, deg, outdeg);
            if ((rc = ga_launch_check(ctx, "k_orient"))) return fail_g(rc);
            uint64_t* osorted = nullptr;
            if ((rc = ga_sort_u64(ctx, other, sorted, m, 2 * nb, &osorted))) return fail_g(rc);
            g->m = m;
            if (ga_malloc(ctx, (void**)&g->col, (size_t)m * 4) || ga_malloc(ctx, (void**)&g->src, (size_t)m * 4)) return fail_g(GA_ERR_OOM);
            k_unpack<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(osorted, m, nb, g->col, g->src);
            if ((rc = ga_launch_check(ctx, "k_unpack"))) return fail_g(rc);
        }
    }
    if (!g->col) {
        if (ga_malloc(ctx, (void**)&g->col, 4) || ga_malloc(ctx, (void**)&g->src, 4)) return fail_g(GA_ERR_OOM);
    }
    if ((rc = finish_rows(ctx, g, outdeg))) return fail_g(rc);
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (build_ms) GA_HIP(ctx, hipEventElapsedTime(build_ms, ctx->ev0, ctx->ev1));
    *out = g;
    return GA_OK;
}

// Raw pairs (host) -> oriented CSR (device).
extern "C" int ga_graph_from_pairs(ga_ctx* ctx, const uint32_t* pairs, uint64_t k, int64_t n_in, ga_graph** out, float* build_ms) {
    *out = nullptr;
    if (!ctx) return GA_ERR_ARG;
    if (k > 0 && !pairs) return ga_fail(ctx, GA_ERR_ARG, "pairs is NULL");
    if (n_in > (int64_t)UINT32_MAX) return ga_fail(ctx, GA_ERR_ARG, "n too large");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf dpairs;
    GA_TRY(ga_malloc(ctx, &dpairs.p, std::max<uint64_t>(k, 1) * 8));
    if (k) GA_HIP(ctx, hipMemcpyAsync(dpairs.p, pairs, k * 8, hipMemcpyHostToDevice, ctx->stream));
    return ga_build_from_device_pairs(ctx, (const uint32_t*)dpairs.p, k, n_in, n_in >= 0, out, build_ms);
}

extern "C" int ga_graph_from_csr(ga_ctx* ctx, uint32_t n, uint32_t m, const uint32_t* rowptr, const uint32_t* col, ga_graph** out) {
    *out = nullptr;
    if (!ctx || !rowptr || (m && !col)) return GA_ERR_ARG;
    if (rowptr[0] != 0 || rowptr[n] != m) return ga_fail(ctx, GA_ERR_ARG, "rowptr must start at 0 and end at m");
    for (uint32_t u = 0; u < n; u++)
        if (rowptr[u] > rowptr[u + 1]) return ga_fail(ctx, GA_ERR_ARG, "rowptr decreases at row %u", u);
    for (uint32_t e = 0; e < m; e++)
        if (col[e] >= n) return ga_fail(ctx, GA_ERR_ARG, "col[%u] = %u is not a vertex id < n", e, col[e]);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    ga_graph* g = new ga_graph;
    g->ctx = ctx;
    g->n = n;
    g->m = m;
    auto fail_g = [&](int rc) { ga_graph_destroy(g); return rc; };
    if (ga_malloc(ctx, (void**)&g->rowptr, ((size_t)n + 1) * 4) || ga_malloc(ctx, (void**)&g->col, std::max<size_t>(m, 1) * 4) ||
        ga_malloc(ctx, (void**)&g->src, std::max<size_t>(m, 1) * 4))
        return fail_g(GA_ERR_OOM);
    if (hipMemcpyAsync(g->rowptr, rowptr, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        (m && hipMemcpyAsync(g->col, col, (size_t)m * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        return fail_g(ga_fail(ctx, GA_ERR_HIP, "upload failed"));
    k_fill_src<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, n, g->src);
    int rc;
    if ((rc = ga_launch_check(ctx, "k_fill_src"))) return fail_g(rc);
    DevBuf od;
    if (ga_malloc(ctx, &od.p, ((size_t)n + 1) * 4)) return fail_g(GA_ERR_OOM);
    k_outdeg<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, n, (uint32_t*)od.p);
    if ((rc = ga_launch_check(ctx, "k_outdeg"))) return fail_g(rc);
    if ((rc = ga_max_u32(ctx, (uint32_t*)od.p, n, &g->max_out))) return fail_g(rc);
    *out = g;
    return GA_OK;
}

static void free_plan(TcPlan& p) {
    ga_free(p.items);
    ga_free(p.wcol);
    ga_free(p.wedge);
    ga_free(p.wpl);
    ga_free(p.wh);
    ga_free(p.wu);
    ga_free(p.cstart);
    ga_free(p.frec);
    ga_free(p.perm);
    p = TcPlan();
}

extern "C" void ga_graph_destroy(ga_graph* g) {
    if (!g) return;
    if (g->ctx) {
        (void)hipSetDevice(g->ctx->device);
        (void)hipStreamSynchronize(g->ctx->stream);
    }
    ga_free(g->rowptr);
    ga_free(g->col);
    ga_free(g->src);
    ga_free(g->urow);
    ga_free(g->ucol);
    ga_free(g->crow);
    ga_free(g->ccol);
    ga_free(g->csrc);
    ga_free(g->pst);
    ga_free(g->pcol);
    free_plan(g->plan);
    free_plan(g->cplan);
    delete g;
}

extern "C" int ga_graph_info(const ga_graph* g, uint32_t* n, uint32_t* m, uint32_t* max_out_degree) {
    if (!g) return GA_ERR_ARG;
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (max_out_degree) *max_out_degree = g->max_out;
    return GA_OK;
}

extern "C" int ga_graph_download(ga_ctx* ctx, const ga_graph* g, uint32_t* rowptr, uint32_t* col, uint32_t* src) {
    if (!ctx || !g) return GA_ERR_ARG;
    if (rowptr) GA_HIP(ctx, hipMemcpyAsync(rowptr, g->rowptr, ((size_t)g->n + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (col && g->m) GA_HIP(ctx, hipMemcpyAsync(col, g->col, (size_t)g->m * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (src && g->m) GA_HIP(ctx, hipMemcpyAsync(src, g->src, (size_t)g->m * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

int ga_graph_ensure_undirected(ga_ctx* ctx, ga_graph* g) {
    if (g->urow) return GA_OK;
    const int nb = bits_for(g->n);
    const uint64_t cnt = 2ull * g->m;
    DevBuf ka, kb, rc_cnt;
    GA_TRY(ga_malloc(ctx, &ka.p, std::max<uint64_t>(cnt, 1) * 8));
    GA_TRY(ga_malloc(ctx, &kb.p, std::max<uint64_t>(cnt, 1) * 8));
    GA_TRY(ga_malloc(ctx, &rc_cnt.p, ((size_t)g->n + 1) * 4));
    uint32_t* urow = nullptr;
    uint32_t* ucol = nullptr;
    GA_TRY(ga_malloc(ctx, (void**)&urow, ((size_t)g->n + 1) * 4));
    if (ga_malloc(ctx, (void**)&ucol, std::max<uint64_t>(cnt, 1) * 4)) { ga_free(urow); return GA_ERR_OOM; }
    int rc = GA_OK;
    uint64_t* sorted = (uint64_t*)ka.p;
    if (g->m) {
        k_both_directions<<<grid_for(g->m, 256, 8192), 256, 0, ctx->stream>>>(g->src, g->col, g->m, nb, (uint64_t*)ka.p);
        rc = ga_launch_check(ctx, "k_both_directions");
        if (!rc) rc = ga_sort_u64(ctx, (uint64_t*)ka.p, (uint64_t*)kb.p, cnt, 2 * nb, &sorted);
    }
    uint32_t* c = (uint32_t*)rc_cnt.p;
    if (!rc) rc = ga_fill_u32(ctx, c, (size_t)g->n + 1, 0);
    if (!rc && g->m) {
        k_unpack<<<grid_for(cnt, 256, 8192), 256, 0, ctx->stream>>>(sorted, (uint32_t)cnt, nb, ucol, nullptr);
        rc = ga_launch_check(ctx, "k_unpack");
        if (!rc) {
            k_count_rows<<<grid_for(g->m, 256, 8192), 256, 0, ctx->stream>>>(g->src, g->m, c);
            k_count_rows<<<grid_for(g->m, 256, 8192), 256, 0, ctx->stream>>>(g->col, g->m, c);
            rc = ga_launch_check(ctx, "k_count_rows");
        }
    }
    if (!rc) rc = ga_exclusive_scan_u32(ctx, c, urow, (size_t)g->n + 1);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ga_fail(ctx, GA_ERR_HIP, "undirected build failed");
    if (rc) { ga_free(urow); ga_free(ucol); return rc; }
    g->urow = urow;
    g->ucol = ucol;
    return GA_OK;
}

extern "C" int ga_graph_undirected(ga_ctx* ctx, ga_graph* g, uint32_t* urow, uint32_t* ucol) {
    if (!ctx || !g) return GA_ERR_ARG;
    GA_TRY(ga_graph_ensure_undirected(ctx, g));
    if (urow) GA_HIP(ctx, hipMemcpyAsync(urow, g->urow, ((size_t)g->n + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (ucol && g->m) GA_HIP(ctx, hipMemcpyAsync(ucol, g->ucol, 2ull * g->m * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

int ga_graph_ensure_perm(ga_ctx* ctx, ga_graph* g) {
    if (g->plan.perm || g->m == 0) return GA_OK;
    DevBuf k0, k1, v1;
    uint32_t* v0 = nullptr;
    GA_TRY(ga_malloc(ctx, &k0.p, (size_t)g->m * 4));
    GA_TRY(ga_malloc(ctx, &k1.p, (size_t)g->m * 4));
    GA_TRY(ga_malloc(ctx, &v1.p, (size_t)g->m * 4));
    GA_TRY(ga_malloc(ctx, (void**)&v0, (size_t)g->m * 4));
    k_edge_bucket<<<grid_for(g->m, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, g->src, g->col, g->m, (uint32_t*)k0.p, v0);
    int rc = ga_launch_check(ctx, "k_edge_bucket");
    uint32_t *ko = nullptr, *vo = nullptr;
    if (!rc) rc = ga_sort_pairs_u32(ctx, (uint32_t*)k0.p, (uint32_t*)k1.p, v0, (uint32_t*)v1.p, g->m, 6, &ko, &vo);
    if (!rc && vo != v0) {
        // keep the sorted values in the buffer we own for the graph's lifetime
        rc = hipMemcpyAsync(v0, vo, (size_t)g->m * 4, hipMemcpyDeviceToDevice, ctx->stream) == hipSuccess ? GA_OK : GA_ERR_HIP;
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ga_fail(ctx, GA_ERR_HIP, "edge order failed");
    if (rc) { ga_free(v0); return rc; }
    g->plan.perm = v0;
    return GA_OK;
}

extern "C" int ga_graph_edge_order(ga_ctx* ctx, ga_graph* g, uint32_t* perm) {
    if (!ctx || !g) return GA_ERR_ARG;
    GA_TRY(ga_graph_ensure_perm(ctx, g));
    if (perm && g->m) GA_HIP(ctx, hipMemcpyAsync(perm, g->plan.perm, (size_t)g->m * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

// ---- build_tasks (graph.h:138-150) on the device: every oriented edge's work min(d+u, d+v) cut into chunks ----
// Tasks of edge e: ceil(min(d+u, d+v) / chunk), in edge order, each (e, chunk * i).
__global__ void k_task_counts(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                              uint32_t m, uint32_t chunk, unsigned long long* __restrict__ cnt) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const uint32_t u = src[e], v = col[e];
        const uint32_t s = min(rowptr[u + 1] - rowptr[u], rowptr[v + 1] - rowptr[v]);
        cnt[e] = ((uint64_t)s + chunk - 1) / chunk;
    }
}

__global__ void k_task_fill(const unsigned long long* __restrict__ off, uint32_t m, uint32_t chunk, uint32_t* __restrict__ t_edge,
                            uint32_t* __restrict__ t_start) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const unsigned long long b = off[e], k = off[e + 1] - b;
        for (unsigned long long i = 0; i < k; i++) {
            t_edge[b + i] = e;
            t_start[b + i] = (uint32_t)(i * chunk);
        }
    }
}

extern "C" int ga_graph_tasks(ga_ctx* ctx, ga_graph* g, uint32_t chunk, uint32_t* t_edge, uint32_t* t_start, uint64_t capacity,
                              uint64_t* ntasks) {
    if (!ctx || !g || !ntasks) return ctx ? ga_fail(ctx, GA_ERR_ARG, "ga_graph_tasks: NULL graph or count") : GA_ERR_ARG;
    if (chunk == 0) return ga_fail(ctx, GA_ERR_ARG, "ga_graph_tasks: chunk must be at least 1");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    *ntasks = 0;
    if (g->m == 0) return GA_OK;
    DevBuf cnt, off;
    GA_TRY(ga_malloc(ctx, &cnt.p, ((size_t)g->m + 1) * 8));
    GA_TRY(ga_malloc(ctx, &off.p, ((size_t)g->m + 1) * 8));
    k_task_counts<<<grid_for(g->m, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, g->src, g->col, g->m, chunk, (unsigned long long*)cnt.p);
    GA_TRY(ga_launch_check(ctx, "k_task_counts"));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)cnt.p + g->m, 1, 0));
    GA_TRY(ga_exclusive_scan_u64(ctx, (const uint64_t*)cnt.p, (uint64_t*)off.p, (size_t)g->m + 1));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, (uint64_t*)off.p + g->m, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint64_t nt = ctx->hstat[0];
    *ntasks = nt;
    if (!t_edge && !t_start) return GA_OK;   // size query
    if (!t_edge || !t_start) return ga_fail(ctx, GA_ERR_ARG, "ga_graph_tasks: t_edge and t_start are both needed");
    if (capacity < nt) return ga_fail(ctx, GA_ERR_ARG, "ga_graph_tasks: %llu tasks, capacity %llu", (unsigned long long)nt, (unsigned long long)capacity);
    if (nt == 0) return GA_OK;
    DevBuf te, ts;
    GA_TRY(ga_malloc(ctx, &te.p, nt * 4));
    GA_TRY(ga_malloc(ctx, &ts.p, nt * 4));
    k_task_fill<<<grid_for(g->m, 256, 8192), 256, 0, ctx->stream>>>((const unsigned long long*)off.p, g->m, chunk, (uint32_t*)te.p,
                                                                    (uint32_t*)ts.p);
    GA_TRY(ga_launch_check(ctx, "k_task_fill"));
    GA_HIP(ctx, hipMemcpyAsync(t_edge, te.p, nt * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(t_start, ts.p, nt * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}
