// ga_greedy.hip - Greedy++ for the triangle-density densest subgraph (Flowless Alg 2 with triangles as hyperedges;
// passes = 1 is Charikar's Greedy), reference dsg.h:176-250 (greedy_pp, greedy).
//
// The reference runs it on the graph that prune_to_core leaves (gpu_dsg.cu:344-356): hundreds to a few thousand
// vertices. Each pass deletes the vertices one at a time, so one workgroup of 1,024 lanes runs a whole Greedy++ (all
// passes, one launch) with the per-vertex state in LDS:
//   load[v]  u64  the Greedy++ load l_v (in place: a deleted vertex is not read again in the pass, dsg.h:184-185)
//   cur[v]   u32  tri_H(v), the triangles of the current subgraph H through v (<= C(n-1, 2) < 2^32 at these sizes)
//   alive    bitmap of H
//   nbr      bitmap of the alive neighbours of the vertex being deleted
// One deletion: argmin of (load + cur, id) over H (each lane scans n / 1024 vertices, then a wave and a workgroup
// reduction), the nbr bitmap from v's undirected row, then every triangle {v, a, b} of H once: wave w takes v's
// neighbours a = w, w + 16, ... and its lanes read a's oriented out-row; b is a hit when its nbr bit is set, and
// cur[a], cur[b] drop by one with LDS atomics. The triangles found must equal cur[v] (the reference's self-check).
// The adjacency (rowptr/col and the undirected urow/ucol) stays in HBM and L2.
#include "ga_internal.hpp"

namespace {

constexpr int GR_BLOCK = 1024;
constexpr int GR_WAVES = GR_BLOCK / GA_WAVE;

struct GreedyOut {
    unsigned long long bn, bd;   // best density found: triangles / vertices (0/0 = none)
    uint32_t best_pass;          // pass whose deletion order holds the best subgraph (UINT32_MAX = none)
    uint32_t best_cut;           // the best subgraph = that pass's order[best_cut .. n)
    uint32_t error;              // 1 = triangle bookkeeping inconsistent
    uint32_t pad;
};

// a/ad > b/bd exactly (ad > 0; bd == 0 means density 0), as dens_gt of dsg.h:40-46: 128-bit cross products.
__device__ __forceinline__ bool dens_gt(uint64_t an, uint64_t ad, uint64_t bn, uint64_t bd) {
    if (bd == 0) { bn = 0; bd = 1; }
    const uint64_t lh = __umul64hi(an, bd), ll = an * bd;
    const uint64_t rh = __umul64hi(bn, ad), rl = bn * ad;
    return lh > rh || (lh == rh && ll > rl);
}

__device__ __forceinline__ bool bit(const uint32_t* b, uint32_t x) { return (b[x >> 5] >> (x & 31)) & 1u; }

__global__ __launch_bounds__(GR_BLOCK) void k_greedy_pp(uint32_t n, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ urow, const uint32_t* __restrict__ ucol,
                                                        const uint32_t* __restrict__ tri0, const unsigned long long* __restrict__ dT,
                                                        uint32_t passes,
                                                        uint32_t* __restrict__ orders, unsigned long long* __restrict__ per_pass,
                                                        GreedyOut* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t nw = (n + 31) >> 5;
    unsigned long long* load = (unsigned long long*)smem;
    uint32_t* cur = (uint32_t*)(load + n);
    uint32_t* alive = cur + n;
    uint32_t* nbr = alive + nw;
    __shared__ unsigned long long red_key[GR_WAVES], red_found[GR_WAVES];
    __shared__ uint32_t red_id[GR_WAVES];
    __shared__ uint32_t bad;

    const uint32_t tid = threadIdx.x, lane = tid & (GA_WAVE - 1), wave = tid / GA_WAVE;
    const uint64_t T0 = *dT;
    for (uint32_t i = tid; i < n; i += GR_BLOCK) load[i] = 0;
    if (tid == 0) bad = 0;
    uint64_t bn = 0, bd = 0;
    uint32_t best_pass = UINT32_MAX, best_cut = 0, err = 0;
    for (uint32_t p = 0; p < passes && !err; p++) {
        for (uint32_t i = tid; i < n; i += GR_BLOCK) cur[i] = tri0[i];
        for (uint32_t i = tid; i < nw; i += GR_BLOCK) {
            alive[i] = (i + 1) * 32 <= n ? ~0u : ((1u << (n & 31)) - 1u);
            nbr[i] = 0;
        }
        __syncthreads();
        uint32_t* order = orders + (size_t)p * n;
        uint64_t T = T0;
        bool improved = false;
        uint32_t cut = 0;
        for (uint32_t step = 0; step < n; step++) {
            const uint32_t left = n - step;
            if (dens_gt(T, left, bn, bd)) {   // H is evaluated before every deletion (dsg.h:208)
                bn = T;
                bd = left;
                improved = true;
                cut = step;
            }
            // the alive vertex with the smallest (load + cur, id); each lane's vertices ascend, so a strict < keeps the
            // smaller id on ties
            unsigned long long bk = ~0ull;
            uint32_t bv = UINT32_MAX;
            for (uint32_t x = tid; x < n; x += GR_BLOCK) {
                if (bit(alive, x)) {
                    const unsigned long long k = load[x] + cur[x];
                    if (k < bk) { bk = k; bv = x; }
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned long long ok = __shfl_xor(bk, o, GA_WAVE);
                const uint32_t ov = __shfl_xor(bv, o, GA_WAVE);
                if (ok < bk || (ok == bk && ov < bv)) { bk = ok; bv = ov; }
            }
            if (lane == 0) { red_key[wave] = bk; red_id[wave] = bv; }
            __syncthreads();
            bk = red_key[0];
            bv = red_id[0];
            for (int w = 1; w < GR_WAVES; w++) {
                const unsigned long long ok = red_key[w];
                const uint32_t ov = red_id[w];
                if (ok < bk || (ok == bk && ov < bv)) { bk = ok; bv = ov; }
            }
            const uint32_t v = bv;              // left > 0, so an alive vertex exists and v < n
            const uint32_t cv = cur[v];
            const uint32_t ub = urow[v], ue = urow[v + 1];
            for (uint32_t j = ub + tid; j < ue; j += GR_BLOCK) {
                const uint32_t a = ucol[j];
                if (bit(alive, a)) atomicOr(&nbr[a >> 5], 1u << (a & 31));
            }
            __syncthreads();
            // triangles {v, a, b} of H, each once: from the oriented edge between a and b
            uint64_t found = 0;
            for (uint32_t j = ub + wave; j < ue; j += GR_WAVES) {
                const uint32_t a = ucol[j];
                if (!bit(nbr, a)) continue;
                const uint32_t re = rowptr[a + 1];
                for (uint32_t e = rowptr[a] + lane; e < re; e += GA_WAVE) {
                    const uint32_t b = col[e];
                    if (bit(nbr, b)) {
                        found++;
                        const uint32_t oa = atomicSub(&cur[a], 1u), ob = atomicSub(&cur[b], 1u);
                        if (oa == 0 || ob == 0) bad = 1;
                    }
                }
            }
            found = wave_sum_u64(found);
            if (lane == 0) red_found[wave] = found;
            __syncthreads();
            found = 0;
            for (int w = 0; w < GR_WAVES; w++) found += red_found[w];
            for (uint32_t j = ub + tid; j < ue; j += GR_BLOCK) nbr[ucol[j] >> 5] = 0;
            if (bad || found != cv) err = 1;   // both read after the barrier: the same in every lane
            if (tid == 0) {
                order[step] = v;
                load[v] += cv;                  // l_v of the next pass (dsg.h:234)
                alive[v >> 5] &= ~(1u << (v & 31));
            }
            T -= cv;
            __syncthreads();
            if (err) break;
        }
        if (improved) {
            best_pass = p;
            best_cut = cut;
        }
        if (tid == 0) {
            per_pass[2 * p] = bn;
            per_pass[2 * p + 1] = bd;
        }
    }
    if (tid == 0) {
        out->bn = bn;
        out->bd = bd;
        out->best_pass = best_pass;
        out->best_cut = best_cut;
        out->error = err;
    }
}

// Dense form for graphs whose adjacency bitmap fits LDS (n <= about 1,100 with 160 KB): adj[x] is a row of W = n / 32
// words, padded to an odd stride so that lanes reading consecutive rows hit different banks. Lane t owns the vertices
// t and t + 1024 and keeps their load, tri_H and alive flag in registers. One deletion:
//   argmin over the owned vertices, a wave reduction and a 16-entry reduction through LDS;
//   each owner of an alive neighbour x of v counts c = |N(x) & N(v) & H| with W popcounts and drops tri_H(x) by c:
//   every triangle {v, x, b} of H is counted once at x and once at b, so the counts sum to 2 tri_H(v);
//   the owner of v retires it; lane 0 writes the order and clears v's bit of the shared alive row.
// Two barriers per deletion; nothing but the order leaves the CU.
__global__ __launch_bounds__(GR_BLOCK) void k_greedy_dense(uint32_t n, uint32_t m, uint32_t wpad, const uint32_t* __restrict__ src,
                                                           const uint32_t* __restrict__ col, const uint32_t* __restrict__ tri0,
                                                           const unsigned long long* __restrict__ dT, uint32_t passes,
                                                           uint32_t* __restrict__ orders, unsigned long long* __restrict__ per_pass,
                                                           GreedyOut* __restrict__ out) {
    extern __shared__ __align__(16) uint32_t adj[];
    uint32_t* alive_w = adj + (size_t)n * wpad;
    __shared__ unsigned long long red_key[GR_WAVES];
    __shared__ uint32_t red_id[GR_WAVES], red_found[GR_WAVES];
    __shared__ uint32_t s_cv, s_bad;

    const uint32_t tid = threadIdx.x, lane = tid & (GA_WAVE - 1), wave = tid / GA_WAVE;
    const uint32_t W = (n + 31) >> 5;
    for (uint32_t i = tid; i < n * wpad; i += GR_BLOCK) adj[i] = 0;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (uint32_t e = tid; e < m; e += GR_BLOCK) {
        const uint32_t a = src[e], b = col[e];
        atomicOr(&adj[a * wpad + (b >> 5)], 1u << (b & 31));
        atomicOr(&adj[b * wpad + (a >> 5)], 1u << (a & 31));
    }
    const uint64_t T0 = *dT;
    unsigned long long load[2] = {0, 0};
    uint32_t cur[2] = {0, 0};
    bool alive[2] = {false, false};
    uint64_t bn = 0, bd = 0;
    uint32_t best_pass = UINT32_MAX, best_cut = 0, err = 0;
    for (uint32_t p = 0; p < passes && !err; p++) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const uint32_t x = tid + k * GR_BLOCK;
            alive[k] = x < n;
            cur[k] = x < n ? tri0[x] : 0;
        }
        for (uint32_t i = tid; i < W; i += GR_BLOCK) alive_w[i] = (i + 1) * 32 <= n ? ~0u : ((1u << (n & 31)) - 1u);
        __syncthreads();
        uint32_t* order = orders + (size_t)p * n;
        uint64_t T = T0;
        bool improved = false;
        uint32_t cut = 0;
        for (uint32_t step = 0; step < n; step++) {
            const uint32_t left = n - step;
            if (dens_gt(T, left, bn, bd)) {
                bn = T;
                bd = left;
                improved = true;
                cut = step;
            }
            unsigned long long bk = ~0ull;
            uint32_t bv = UINT32_MAX;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const unsigned long long key = load[k] + cur[k];
                if (alive[k] && key < bk) { bk = key; bv = tid + k * GR_BLOCK; }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned long long ok = __shfl_xor(bk, o, GA_WAVE);
                const uint32_t ov = __shfl_xor(bv, o, GA_WAVE);
                if (ok < bk || (ok == bk && ov < bv)) { bk = ok; bv = ov; }
            }
            if (lane == 0) { red_key[wave] = bk; red_id[wave] = bv; }
            __syncthreads();
            bk = red_key[0];
            bv = red_id[0];
            for (int w = 1; w < GR_WAVES; w++) {
                const unsigned long long ok = red_key[w];
                const uint32_t ov = red_id[w];
                if (ok < bk || (ok == bk && ov < bv)) { bk = ok; bv = ov; }
            }
            const uint32_t v = bv;   // left > 0: some lane owns an alive vertex, so v < n
            const uint32_t* av = adj + v * wpad;
            uint32_t found = 0;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const uint32_t x = tid + k * GR_BLOCK;
                if (x == v) s_cv = cur[k];
                if (alive[k] && ((av[x >> 5] >> (x & 31)) & 1u)) {
                    const uint32_t* ax = adj + x * wpad;
                    uint32_t c = 0;
                    for (uint32_t i = 0; i < W; i++) c += __popc(ax[i] & av[i] & alive_w[i]);
                    if (c > cur[k]) s_bad = 1;
                    cur[k] -= min(c, cur[k]);
                    found += c;
                }
            }
            found = (uint32_t)wave_sum_u64(found);
            if (lane == 0) red_found[wave] = found;
            __syncthreads();
            uint64_t total = 0;
            for (int w = 0; w < GR_WAVES; w++) total += red_found[w];
            const uint32_t cv = s_cv;
            if (s_bad || total != 2ull * cv) err = 1;   // read after the barrier: the same in every lane
#pragma unroll
            for (int k = 0; k < 2; k++)
                if (tid + k * GR_BLOCK == v) {
                    alive[k] = false;
                    load[k] += cv;
                }
            if (tid == 0) {
                order[step] = v;
                alive_w[v >> 5] &= ~(1u << (v & 31));
            }
            T -= cv;
            if (err) break;
        }
        if (improved) {
            best_pass = p;
            best_cut = cut;
        }
        if (tid == 0) {
            per_pass[2 * p] = bn;
            per_pass[2 * p + 1] = bd;
        }
        __syncthreads();   // the alive row is rewritten by the next pass
    }
    if (tid == 0) {
        out->bn = bn;
        out->bd = bd;
        out->best_pass = best_pass;
        out->best_cut = best_cut;
        out->error = err;
    }
}

__global__ void k_greedy_flags(const uint32_t* __restrict__ order, uint32_t n, const GreedyOut* __restrict__ out,
                               uint32_t* __restrict__ flags) {
    const uint32_t cut = out->best_pass == UINT32_MAX ? n : out->best_cut;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (i >= cut) flags[order[i]] = 1;
}

__global__ void k_greedy_set(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, uint32_t n,
                             uint32_t* __restrict__ set) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        if (flags[v]) set[pos[v]] = v;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

size_t greedy_lds(uint32_t n) { return (size_t)n * 12 + (size_t)((n + 31) / 32) * 8; }

uint32_t dense_stride(uint32_t n) { return ((n + 31) / 32) | 1u; }
size_t dense_lds(uint32_t n) { return ((size_t)n + 1) * dense_stride(n) * 4; }

size_t lds_room(const ga_ctx* ctx) { return ctx->lds_max > 1024 ? ctx->lds_max - 1024 : 0; }   // static LDS stays below 1 KB

// The largest n the dense form takes: its bitmap fits LDS and each lane owns at most two vertices.
uint32_t dense_max(const ga_ctx* ctx) {
    uint32_t n = 2 * GR_BLOCK;
    while (n && dense_lds(n) > lds_room(ctx)) n--;
    return n;
}

}  // namespace

extern "C" uint32_t ga_greedy_max_vertices(ga_ctx* ctx) {
    if (!ctx) return 0;
    const size_t room = lds_room(ctx);
    uint32_t n = (uint32_t)(room / 12);
    while (n && greedy_lds(n) > room) n--;
    return n;
}

extern "C" int ga_greedy_pp(ga_ctx* ctx, ga_graph* g, uint32_t passes, uint64_t* tri, uint64_t* verts, uint64_t* per_pass,
                            uint32_t* set, uint32_t* set_size, float* ms) {
    if (!ctx) return GA_ERR_ARG;
    if (!g || g->ctx != ctx) return ga_fail(ctx, GA_ERR_ARG, "graph is NULL or belongs to another context");
    if (passes == 0) return ga_fail(ctx, GA_ERR_ARG, "greedy_pp needs at least one pass");
    const uint32_t n = g->n;
    const uint32_t cap = ga_greedy_max_vertices(ctx);
    if (n > cap)
        return ga_fail(ctx, GA_ERR_ARG, "greedy_pp runs one workgroup per graph with its state in LDS: n = %u exceeds %u vertices "
                       "(run it on the graph prune_to_core leaves)", n, cap);
    if ((uint64_t)passes * n > (1ull << 28))
        return ga_fail(ctx, GA_ERR_ARG, "greedy_pp keeps every pass's deletion order: passes * n = %llu exceeds 2^28",
                       (unsigned long long)passes * n);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(ga_graph_ensure_undirected(ctx, g));
    DevBuf dtri, dorders, dpass, dout, dflags, dpos, dset, dT;
    GA_TRY(ga_malloc(ctx, &dtri.p, (size_t)std::max(n, 1u) * 4));
    GA_TRY(ga_malloc(ctx, &dorders.p, (size_t)std::max<uint64_t>((uint64_t)passes * n, 1) * 4));
    GA_TRY(ga_malloc(ctx, &dpass.p, (size_t)passes * 16));
    GA_TRY(ga_malloc(ctx, &dout.p, sizeof(GreedyOut)));
    GA_TRY(ga_malloc(ctx, &dflags.p, ((size_t)n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &dpos.p, ((size_t)n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &dset.p, (size_t)std::max(n, 1u) * 4));
    GA_TRY(ga_malloc(ctx, &dT.p, 8));
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(ga_fill_u32(ctx, (uint32_t*)dtri.p, std::max(n, 1u), 0));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)dT.p, 1, 0));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)dpass.p, (size_t)passes * 2, 0));
    GA_TRY(ga_fill_u32(ctx, (uint32_t*)dflags.p, (size_t)n + 1, 0));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)dout.p, sizeof(GreedyOut) / 8, 0));
    if (g->m) GA_TRY(ga_tc_launch(ctx, g, GA_TC_HASH, (unsigned long long*)dT.p, nullptr, (uint32_t*)dtri.p, nullptr, nullptr));
    if (n && n <= dense_max(ctx)) {
        k_greedy_dense<<<1, GR_BLOCK, dense_lds(n), ctx->stream>>>(n, g->m, dense_stride(n), g->src, g->col, (const uint32_t*)dtri.p,
                                                                 (const unsigned long long*)dT.p, passes, (uint32_t*)dorders.p,
                                                                 (unsigned long long*)dpass.p, (GreedyOut*)dout.p);
        GA_TRY(ga_launch_check(ctx, "k_greedy_dense"));
    } else if (n) {
        k_greedy_pp<<<1, GR_BLOCK, greedy_lds(n), ctx->stream>>>(n, g->rowptr, g->col, g->urow, g->ucol, (const uint32_t*)dtri.p,
                                                               (const unsigned long long*)dT.p, passes, (uint32_t*)dorders.p, (unsigned long long*)dpass.p,
                                                               (GreedyOut*)dout.p);
        GA_TRY(ga_launch_check(ctx, "k_greedy_pp"));
    }
    GreedyOut h = {};
    GA_HIP(ctx, hipMemcpyAsync(&h, dout.p, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!n) h.best_pass = UINT32_MAX;
    if (h.error) return ga_fail(ctx, GA_ERR_STATE, "greedy_pp: triangle bookkeeping inconsistent");
    uint32_t count = 0;
    if (h.best_pass != UINT32_MAX && h.bd) {
        const uint32_t* order = (const uint32_t*)dorders.p + (size_t)h.best_pass * n;
        k_greedy_flags<<<grid_for(n, 256, 1024), 256, 0, ctx->stream>>>(order, n, (const GreedyOut*)dout.p, (uint32_t*)dflags.p);
        GA_TRY(ga_launch_check(ctx, "k_greedy_flags"));
        GA_TRY(ga_exclusive_scan_u32(ctx, (const uint32_t*)dflags.p, (uint32_t*)dpos.p, (size_t)n + 1));
        k_greedy_set<<<grid_for(n, 256, 1024), 256, 0, ctx->stream>>>((const uint32_t*)dflags.p, (const uint32_t*)dpos.p, n,
                                                                      (uint32_t*)dset.p);
        GA_TRY(ga_launch_check(ctx, "k_greedy_set"));
        GA_HIP(ctx, hipMemcpyAsync(&count, (const uint32_t*)dpos.p + n, 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (count != (h.best_pass != UINT32_MAX && h.bd ? h.bd : 0))
        return ga_fail(ctx, GA_ERR_STATE, "greedy_pp: best set has %u vertices, expected %llu", count, (unsigned long long)h.bd);
    if (set && count) GA_HIP(ctx, hipMemcpy(set, dset.p, (size_t)count * 4, hipMemcpyDeviceToHost));
    if (per_pass) GA_HIP(ctx, hipMemcpy(per_pass, dpass.p, (size_t)passes * 16, hipMemcpyDeviceToHost));
    if (tri) *tri = h.bd ? h.bn : 0;
    if (verts) *verts = h.bd;
    if (set_size) *set_size = count;
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}
