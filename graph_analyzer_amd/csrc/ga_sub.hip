// ga_sub.hip - induced subgraphs and core pruning on the GPU (induced_subgraph, dsg.h:103-114; prune_to_core,
// dsg.h:636-645).
//
// keep flags (u32, n + 1, last 0) -> exclusive scan = new ids, n' = scan[n] -> old_id[new_id[v]] = v -> the oriented
// edges with both ends kept, renamed to new ids and appended with one atomic per wave -> the GPU CSR builder in strict
// mode (build_graph_from_edges semantics: orientation by (degree, id) with the degrees inside the subgraph, rows sorted,
// src[] and max_out_degree). The append order does not matter: the builder sorts. New ids follow the old ids' order, as
// in the reference, because they come from a scan in vertex order.
#include "ga_internal.hpp"

namespace {

struct DevBuf {   // scoped device allocation
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

__global__ void k_keep_from_bytes(const uint8_t* __restrict__ keep, uint32_t n, uint32_t* __restrict__ flag) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += gridDim.x * blockDim.x)
        flag[v] = v < n ? (keep[v] != 0) : 0u;
}

// keep[v] = core[v] >= k (prune_to_core, dsg.h:641); k is 64-bit as in the reference's signature.
__global__ void k_keep_from_core(const uint32_t* __restrict__ core, uint32_t n, uint64_t k, uint32_t* __restrict__ flag) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += gridDim.x * blockDim.x)
        flag[v] = v < n ? ((uint64_t)core[v] >= k) : 0u;
}

__global__ void k_old_ids(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ nid, uint32_t n, uint32_t* __restrict__ old_id) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        if (flag[v]) old_id[nid[v]] = v;
}

// Edges with both ends kept, as (new src, new col) pairs; one atomic per wave for the append position.
__global__ void k_induced_pairs(const uint32_t* __restrict__ src, const uint32_t* __restrict__ col, uint32_t m,
                                const uint32_t* __restrict__ flag, const uint32_t* __restrict__ nid, uint32_t* __restrict__ pairs,
                                unsigned long long* __restrict__ cursor) {
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t base = wave0 * 64; base < m; base += nwaves * 64) {   // wave-uniform loop
        const uint64_t e = base + lane_id();
        uint32_t a = 0, b = 0;
        bool k2 = false;
        if (e < m) {
            a = src[e];
            b = col[e];
            k2 = flag[a] && flag[b];
        }
        const uint64_t mask = __ballot(k2);
        if (!mask) continue;
        unsigned long long at = 0;
        if (lane_id() == 0) at = atomicAdd(cursor, (unsigned long long)__popcll(mask));
        at = __shfl(at, 0, 64);
        if (k2) {
            const uint64_t w = at + mask_rank(mask);
            pairs[2 * w] = nid[a];
            pairs[2 * w + 1] = nid[b];
        }
    }
}

__global__ void k_gather_u32(const uint32_t* __restrict__ in, const uint32_t* __restrict__ idx, uint32_t count, uint32_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) out[i] = in[idx[i]];
}

}  // namespace

// Induced subgraph of g on the vertices whose flag (device u32[n + 1], flag[n] = 0) is set. old_id (host, capacity g->n)
// and core_out (host, capacity g->n: core_dev[old_id[i]]) may be NULL.
static int induced_from_flags(ga_ctx* ctx, const ga_graph* g, uint32_t* flag, const uint32_t* core_dev, ga_graph** out,
                              uint32_t* old_id, uint32_t* core_out) {
    const uint32_t n = g->n, m = g->m;
    DevBuf nidb, oldb, pairsb, curb, coreb;
    GA_TRY(ga_malloc(ctx, &nidb.p, ((size_t)n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &curb.p, 64));
    uint32_t* nid = (uint32_t*)nidb.p;
    unsigned long long* cursor = (unsigned long long*)curb.p;
    GA_TRY(ga_exclusive_scan_u32(ctx, flag, nid, (size_t)n + 1));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)cursor, 1, 0));
    GA_TRY(ga_malloc(ctx, &oldb.p, std::max<size_t>(n, 1) * 4));
    GA_TRY(ga_malloc(ctx, &pairsb.p, std::max<size_t>(m, 1) * 8));
    k_old_ids<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(flag, nid, n, (uint32_t*)oldb.p);
    GA_TRY(ga_launch_check(ctx, "k_old_ids"));
    if (m) {
        k_induced_pairs<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(g->src, g->col, m, flag, nid, (uint32_t*)pairsb.p, cursor);
        GA_TRY(ga_launch_check(ctx, "k_induced_pairs"));
    }
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, nid + n, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat + 1, cursor, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t n2 = (uint32_t)(ctx->hstat[0] & 0xFFFFFFFFu);
    const uint64_t m2 = ctx->hstat[1];
    if (n2 > n || m2 > m) return ga_fail(ctx, GA_ERR_STATE, "induced subgraph sizes out of range (n' %u of %u, m' %llu of %u)", n2, n,
                                         (unsigned long long)m2, m);
    ga_graph* h = nullptr;
    GA_TRY(ga_build_from_device_pairs(ctx, (const uint32_t*)pairsb.p, m2, (int64_t)n2, true, &h, nullptr));
    int rc = GA_OK;
    if (old_id && n2 && hipMemcpyAsync(old_id, oldb.p, (size_t)n2 * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = ga_fail(ctx, GA_ERR_HIP, "old_id download failed");
    if (!rc && core_out && n2) {
        if ((rc = ga_malloc(ctx, &coreb.p, (size_t)n2 * 4)) == GA_OK) {
            k_gather_u32<<<grid_for(n2, 256, 8192), 256, 0, ctx->stream>>>(core_dev, (const uint32_t*)oldb.p, n2, (uint32_t*)coreb.p);
            rc = ga_launch_check(ctx, "k_gather_u32");
            if (!rc && hipMemcpyAsync(core_out, coreb.p, (size_t)n2 * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
                rc = ga_fail(ctx, GA_ERR_HIP, "core_out download failed");
        }
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ga_fail(ctx, GA_ERR_HIP, "induced subgraph download failed");
    if (rc) { ga_graph_destroy(h); return rc; }
    *out = h;
    return GA_OK;
}

// Runs `flags` (a launch that fills the keep flags) and the induced build between two events; *ms (may be NULL) is the
// device time from the flags kernel to the finished CSR, including the two count readbacks the build waits on.
template <class F>
static int induced_timed(ga_ctx* ctx, const ga_graph* g, F&& flags, uint32_t* flag, const uint32_t* core_dev, ga_graph** out,
                         uint32_t* old_id, uint32_t* core_out, float* ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GA_HIP(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return ga_fail(ctx, GA_ERR_HIP, "hipEventCreate failed"); }
    int rc = hipEventRecord(e0, ctx->stream) == hipSuccess ? GA_OK : ga_fail(ctx, GA_ERR_HIP, "event record failed");
    if (!rc) rc = flags();
    if (!rc) rc = induced_from_flags(ctx, g, flag, core_dev, out, old_id, core_out);
    if (!rc && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = ga_fail(ctx, GA_ERR_HIP, "event record failed");
    if (!rc && ms) (void)hipEventElapsedTime(ms, e0, e1);
    if (rc && *out) { ga_graph_destroy(*out); *out = nullptr; }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

static int check_graph_args(ga_ctx* ctx, const ga_graph* g, ga_graph** out) {
    if (!out) return ctx ? ga_fail(ctx, GA_ERR_ARG, "out is NULL") : GA_ERR_ARG;
    *out = nullptr;
    if (!ctx || !g) return GA_ERR_ARG;
    return GA_OK;
}

extern "C" int ga_graph_induced(ga_ctx* ctx, const ga_graph* g, const uint8_t* keep, ga_graph** out, uint32_t* old_id, float* ms) {
    GA_TRY(check_graph_args(ctx, g, out));
    if (g->n && !keep) return ga_fail(ctx, GA_ERR_ARG, "keep is NULL");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf kb, fb;
    GA_TRY(ga_malloc(ctx, &kb.p, std::max<size_t>(g->n, 1)));
    GA_TRY(ga_malloc(ctx, &fb.p, ((size_t)g->n + 1) * 4));
    if (g->n) GA_HIP(ctx, hipMemcpyAsync(kb.p, keep, g->n, hipMemcpyHostToDevice, ctx->stream));
    uint32_t* flag = (uint32_t*)fb.p;
    auto flags = [&]() {
        k_keep_from_bytes<<<grid_for((uint64_t)g->n + 1, 256, 8192), 256, 0, ctx->stream>>>((const uint8_t*)kb.p, g->n, flag);
        return ga_launch_check(ctx, "k_keep_from_bytes");
    };
    return induced_timed(ctx, g, flags, flag, nullptr, out, old_id, nullptr, ms);
}

int ga_prune_core_dev(ga_ctx* ctx, const ga_graph* g, const uint32_t* core_dev, uint64_t k, ga_graph** out, uint32_t* old_id,
                      uint32_t* core_out, float* ms) {
    GA_TRY(check_graph_args(ctx, g, out));
    GA_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf fb;
    GA_TRY(ga_malloc(ctx, &fb.p, ((size_t)g->n + 1) * 4));
    uint32_t* flag = (uint32_t*)fb.p;
    auto flags = [&]() {
        k_keep_from_core<<<grid_for((uint64_t)g->n + 1, 256, 8192), 256, 0, ctx->stream>>>(core_dev, g->n, k, flag);
        return ga_launch_check(ctx, "k_keep_from_core");
    };
    return induced_timed(ctx, g, flags, flag, core_dev, out, old_id, core_out, ms);
}

extern "C" int ga_graph_prune_to_core(ga_ctx* ctx, const ga_graph* g, const uint32_t* core, uint64_t k, ga_graph** out,
                                      uint32_t* old_id, uint32_t* core_out, float* ms) {
    GA_TRY(check_graph_args(ctx, g, out));
    if (g->n && !core) return ga_fail(ctx, GA_ERR_ARG, "core is NULL");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf cb;
    GA_TRY(ga_malloc(ctx, &cb.p, std::max<size_t>(g->n, 1) * 4));
    if (g->n) GA_HIP(ctx, hipMemcpyAsync(cb.p, core, (size_t)g->n * 4, hipMemcpyHostToDevice, ctx->stream));
    return ga_prune_core_dev(ctx, g, (const uint32_t*)cb.p, k, out, old_id, core_out, ms);
}
