// ga_tc_simple.hpp - device code of the count variants that read neither src[] nor a count plan (k_tc_nosrc, the
// role of tc_edge_find_src v7, kernels.cuh:219-248; k_tc_vertex, the role of tc_vertex_per_thread v6,
// kernels.cuh:195-217), shared by the plain count (ga_tc_simple.hip) and the allocation-free count (ga_noalloc.hip).
//
//   k_tc_nosrc<G>   the source u of edge e is recovered from rowptr instead of a stored src[e]. Each wavefront owns a
//                   contiguous range of edges; one binary search over rowptr finds the row of its first edge, and
//                   after that every step of 64 / G consecutive edges takes its rows from a 64-entry window of rowptr
//                   held one entry per lane (a shuffle binary search per edge). G lanes per edge stride the shorter
//                   out-list and binary-search each element in the longer one (the v5 intersection).
//   k_tc_vertex     vertex-centric. Each wavefront takes 64 consecutive vertices; a vertex with at most VTX_LANE_MAX
//                   out-neighbours is merged by its own lane (N+(u) with N+(v) for every out-neighbour v, two
//                   pointers), a vertex with more is deferred and then done by the whole wavefront, one out-edge per
//                   lane, so that no hub serialises on one lane.
//
// Both sum T = sum over oriented edges (u,v) of |N+(u) ∩ N+(v)| (graph.h:110-121) in u64 and end in block_finish:
// one atomic per workgroup into a counter, or (PARTIALS) one partial sum per workgroup into partials[blockIdx.x],
// where the last workgroup to finish (an atomic ticket) adds the partials in workgroup order and stores T into *out.
#pragma once
#include "ga_internal.hpp"

#define VTX_LANE_MAX 16u   // out-degree up to which one lane merges a vertex's edges alone
#define SIMPLE_BLOCK 256u  // threads per workgroup of both kernels

namespace ga_simple {

__device__ __forceinline__ uint32_t contains_sorted_u32(const uint32_t* __restrict__ a, uint32_t lo, uint32_t hi, uint32_t x) {
    const uint32_t p = lower_bound_u32(a, lo, hi, x);
    return (p < hi && a[p] == x) ? 1u : 0u;
}

// Common elements of two sorted lists: the shorter one is walked by `lanes` lanes from `first`, each element
// binary-searched in the longer one.
__device__ __forceinline__ uint32_t strided_common(const uint32_t* __restrict__ col, uint32_t aS, uint32_t aN, uint32_t bS, uint32_t bN,
                                                   uint32_t first, uint32_t lanes) {
    if (aN > bN) {
        const uint32_t t0 = aS, t1 = aN;
        aS = bS;
        aN = bN;
        bS = t0;
        bN = t1;
    }
    uint32_t c = 0;
    for (uint32_t i = first; i < aN; i += lanes) c += contains_sorted_u32(col, bS, bS + bN, col[aS + i]);
    return c;
}

// Two-pointer merge count of two sorted lists.
__device__ __forceinline__ uint32_t merge_common(const uint32_t* __restrict__ col, uint32_t i, uint32_t ie, uint32_t j, uint32_t je) {
    uint32_t c = 0;
    while (i < ie && j < je) {
        const uint32_t a = col[i], b = col[j];
        c += a == b ? 1u : 0u;
        i += a <= b ? 1u : 0u;
        j += b <= a ? 1u : 0u;
    }
    return c;
}

// Sum of `local` over the workgroup; with PARTIALS, the partial goes to partials[blockIdx.x] and the last workgroup
// to finish (counted by *ticket, which that workgroup resets to 0) stores the sum of all partials into *out; without,
// one atomic per workgroup into *total. Every thread of the workgroup must call it.
template <bool PARTIALS>
__device__ __forceinline__ void block_finish(uint64_t local, unsigned long long* total, unsigned long long* partials,
                                             unsigned long long* ticket, uint64_t* out) {
    __shared__ unsigned long long wsum[16];
    __shared__ uint32_t last;
    local = wave_sum_u64(local);
    const uint32_t wv = threadIdx.x / 64, nw = (blockDim.x + 63) / 64;
    if (lane_id() == 0) wsum[wv] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (uint32_t w = 0; w < nw; w++) s += wsum[w];
        if (!PARTIALS) {
            if (s) atomicAdd(total, s);
        } else {
            __hip_atomic_store(&partials[blockIdx.x], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            const unsigned long long t = atomicAdd(ticket, 1ull);
            last = t == gridDim.x - 1 ? 1u : 0u;
        }
    }
    if (!PARTIALS) return;
    __syncthreads();
    if (!last) return;
    __threadfence();
    // the last workgroup: partials in workgroup order, one wavefront at a time per 64 of them
    unsigned long long s = 0;
    if (threadIdx.x < 64)
        for (uint32_t b = threadIdx.x; b < gridDim.x; b += 64) s += __hip_atomic_load(&partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < 64) {
        s = wave_sum_u64(s);
        if (threadIdx.x == 0) {
            *out = s;
            __hip_atomic_store(ticket, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Row of edge e: the last u with rowptr[u] <= e (rows with no edges are skipped because their rowptr equals the
// next one; rowptr[n] = m > e ends every search).
__device__ __forceinline__ uint32_t row_of_edge(const uint32_t* __restrict__ rowptr, uint32_t lo, uint32_t hi, uint32_t e) {
    return upper_bound_u32(rowptr, lo, hi, e) - 1u;   // upper bound over rowptr[lo, hi), lo <= answer
}

template <int G, bool PARTIALS>
__global__ void __launch_bounds__(256) k_tc_nosrc(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                                  uint32_t m, uint32_t per_wave, unsigned long long* __restrict__ total,
                                                  unsigned long long* __restrict__ partials, unsigned long long* ticket,
                                                  uint64_t* __restrict__ out) {
    static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "G must divide 64");
    constexpr uint32_t K = 64 / G;   // edges per step of a wavefront
    const uint32_t lane = lane_id(), sub = lane % G, grp = lane / G;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint64_t c0 = wave * per_wave, c1 = std::min<uint64_t>((uint64_t)m, c0 + per_wave);
    uint64_t local = 0;
    if (c0 < c1) {
        uint32_t u0 = row_of_edge(rowptr, 0, n + 1, (uint32_t)c0);   // same addresses in every lane
        for (uint64_t base = c0; base < c1; base += K) {
            // window: W_l = rowptr[u0 + 1 + l]; rowptr[n] = m bounds it (entries past n read as m)
            const uint32_t wi = u0 + 1u + lane;
            const uint32_t W = wi <= n ? rowptr[wi] : m;
            const uint64_t e64 = base + grp;
            const uint32_t e = (uint32_t)std::min<uint64_t>(e64, (uint64_t)m - 1);
            // r = #{l : W_l <= e} by binary lifting over the lanes' window (W is non-decreasing in l)
            uint32_t r = 0;
#pragma unroll
            for (uint32_t s = 32; s >= 1; s >>= 1) {
                const uint32_t w = (uint32_t)__shfl((int)W, (int)(r + s - 1u), 64);
                r += w <= e ? s : 0u;
            }
            r += (uint32_t)__shfl((int)W, (int)r, 64) <= e ? 1u : 0u;   // r <= 63 here, so lane r exists
            uint32_t u = u0 + r;
            if (r == 64u) u = row_of_edge(rowptr, u0 + 64u, n + 1, e);   // more than 64 rows in this step (empty rows)
            // next step's rows start at or after the last edge's row
            u0 = (uint32_t)__shfl((int)u, 63, 64);
            if (e64 < c1) {
                const uint32_t v = col[e];
                const uint32_t ru = rowptr[u], du = rowptr[u + 1] - ru, rv = rowptr[v], dv = rowptr[v + 1] - rv;
                local += strided_common(col, ru, du, rv, dv, sub, G);
            }
        }
    }
    block_finish<PARTIALS>(local, total, partials, ticket, out);
}

template <bool PARTIALS>
__global__ void __launch_bounds__(256) k_tc_vertex(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                                   unsigned long long* __restrict__ total, unsigned long long* __restrict__ partials,
                                                   unsigned long long* ticket, uint64_t* __restrict__ out) {
    const uint32_t lane = lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    uint64_t local = 0;
    for (uint64_t v0 = wave * 64; v0 < n; v0 += nwaves * 64) {   // wavefront-uniform
        const uint64_t u64 = v0 + lane;
        uint32_t ub = 0, ue = 0;
        if (u64 < n) {
            ub = rowptr[u64];
            ue = rowptr[u64 + 1];
        }
        const bool heavy = ue - ub > VTX_LANE_MAX;
        if (!heavy) {
            for (uint32_t e = ub; e < ue; e++) {   // v6: merge N+(u) with N+(v) for every out-neighbour v
                const uint32_t v = col[e];
                local += merge_common(col, ub, ue, rowptr[v], rowptr[v + 1]);
            }
        }
        uint64_t todo = __ballot(heavy);
        while (todo) {   // long out-lists: one out-edge per lane, shorter list searched in the longer
            const uint32_t l = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
            todo &= todo - 1ull;
            const uint32_t hb = (uint32_t)__shfl((int)ub, (int)l, 64), he = (uint32_t)__shfl((int)ue, (int)l, 64);
            for (uint32_t e = hb + lane; e < he; e += 64) {
                const uint32_t v = col[e];
                const uint32_t rv = rowptr[v];
                local += strided_common(col, hb, he - hb, rv, rowptr[v + 1] - rv, 0, 1);
            }
        }
    }
    block_finish<PARTIALS>(local, total, partials, ticket, out);
}

// Edges per wavefront of k_tc_nosrc for a grid of `blocks` workgroups.
inline uint32_t nosrc_per_wave(uint32_t m, uint32_t blocks) {
    const uint64_t waves = (uint64_t)blocks * (SIMPLE_BLOCK / 64);
    return (uint32_t)std::max<uint64_t>(1, ((uint64_t)m + waves - 1) / waves);
}

// Workgroups for a graph: enough for 4 lanes per edge (NOSRC) or a lane per vertex (VERTEX), at most 8 per CU.
inline uint32_t default_blocks(const ga_ctx* ctx, const ga_graph* g, int variant) {
    const uint64_t work = variant == GA_TC_NOSRC ? (uint64_t)g->m * 4 : (uint64_t)g->n;
    const uint64_t cap = (uint64_t)std::max(ctx->cus, 1) * 8;
    return ctx_grid(ctx, (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cap, (work + SIMPLE_BLOCK - 1) / SIMPLE_BLOCK)));
}

}  // namespace ga_simple
