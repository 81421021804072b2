// ga_tc.hip - triangle counting on the oriented CSR: T = sum over oriented edges (u,v) of |N+(u) ∩ N+(v)|
// (graph.h:15-18, 110-121), with optional per-vertex participation on an alive mask (densest.h:81-101,
// kernels.cuh:372-425) and per-edge support (dsg.h:62-79 indices).
//
// Kernels (all wave64; counts are exact u64, reduced across the wave and added with one atomic per wave):
//   k_tc_merge        one lane per oriented edge, two-pointer merge                       (the v1 role)
//   k_tc_group<G>     G lanes per edge, strided binary search of the shorter list         (the v2/v5 roles)
//   k_tc_sorted       one lane per edge in the device-built heavy-first work order         (the v9 role)
//   k_tc_hash<G,...>  work items grouped by the endpoint h with the LONGER out-list: N+(h) goes into an LDS
//                     open-addressing set, and the shorter lists N+(p) of all of h's work edges are streamed
//                     flat (coalesced) through a lane group of G lanes that probes the set. Every probe is one
//                     element of min(d+(u), d+(v)), so the streamed volume is the algorithmic minimum
//                     sum_e min(d+u, d+v). Work items are bucketed by d+(h); each bucket has its own G and table
//                     size (the tuning table), and hub rows are split into work-bounded chunks.
#include "ga_internal.hpp"

enum { F_ALIVE = 1, F_TRI = 2, F_SUP = 4, F_GRAPH = 8 };

#define GA_EMPTY 0xFFFFFFFFu

// ---------------------------------------------------------------------------------------------------------------
// simple variants
// ---------------------------------------------------------------------------------------------------------------
template <int FLAGS>
__global__ void k_tc_merge(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                           const uint32_t* __restrict__ src, const uint32_t* __restrict__ order, uint32_t m,
                           const uint8_t* __restrict__ alive, uint32_t* __restrict__ tri, uint32_t* __restrict__ sup,
                           unsigned long long* __restrict__ total, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    uint64_t local = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x) {
        const uint32_t e = order ? order[t] : t;
        const uint32_t u = src[e], v = col[e];
        if ((FLAGS & F_ALIVE) && (!alive[u] || !alive[v])) continue;
        uint32_t i = rowptr[u], ie = rowptr[u + 1], j = rowptr[v], je = rowptr[v + 1];
        uint32_t c = 0;
        while (i < ie && j < je) {
            const uint32_t a = col[i], b = col[j];
            if (a < b) { i++; continue; }
            if (a > b) { j++; continue; }
            if (!(FLAGS & F_ALIVE) || alive[a]) {
                c++;
                if (FLAGS & F_TRI) atomicAdd(&tri[a], 1u);
                if (FLAGS & F_SUP) { atomicAdd(&sup[i], 1u); atomicAdd(&sup[j], 1u); }
            }
            i++;
            j++;
        }
        if (c) {
            local += c;
            if (FLAGS & F_TRI) { atomicAdd(&tri[u], c); atomicAdd(&tri[v], c); }
            if (FLAGS & F_SUP) atomicAdd(&sup[e], c);
        }
    }
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(total, (unsigned long long)local);
}

__device__ __forceinline__ uint32_t contains_sorted(const uint32_t* __restrict__ a, uint32_t lo, uint32_t hi, uint32_t x) {
    const uint32_t p = lower_bound_u32(a, lo, hi, x);
    return (p < hi && a[p] == x) ? 1u : 0u;
}

template <int G>
__global__ void k_tc_group(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col, const uint32_t* __restrict__ src,
                           uint32_t m, unsigned long long* __restrict__ total) {
    static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "G must divide 64");
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t ngroups = ((uint64_t)gridDim.x * blockDim.x) / G;
    const uint32_t sub = threadIdx.x % G;
    uint64_t local = 0;
    for (uint64_t e = tid / G; e < m; e += ngroups) {
        const uint32_t u = src[e], v = col[e];
        const uint32_t ru = rowptr[u], du = rowptr[u + 1] - ru, rv = rowptr[v], dv = rowptr[v + 1] - rv;
        const bool us = du <= dv;
        const uint32_t aS = us ? ru : rv, aN = us ? du : dv, bS = us ? rv : ru, bN = us ? dv : du;
        for (uint32_t i = sub; i < aN; i += G) local += contains_sorted(col, bS, bS + bN, col[aS + i]);
    }
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(total, (unsigned long long)local);
}

// ---------------------------------------------------------------------------------------------------------------
// hash variant
// ---------------------------------------------------------------------------------------------------------------
// Items: 3 words (h, begin, end) into wcol/wedge. Each lane group of G lanes owns one item at a time:
//   1. clear a table of ts = next_pow2(d+(h)) << load_shift slots (>= 2x the entries, so every probe chain ends)
//   2. insert N+(h) (LDS compare-and-swap, linear probing; dead vertices are not inserted)
//   3. for each batch of G work edges (h, p): scan the lengths d+(p), then walk the concatenated lists G elements at a
//      time; the owner of flat element f is found by a binary search over the group's inclusive lengths (shuffles),
//      so consecutive lanes read consecutive col[] addresses.
// HASHED=false is the bucket for huge d+(h): no table, each element is binary-searched in N+(h) in global memory.

__device__ __forceinline__ void group_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t hash_slot(uint32_t x, uint32_t tbits) {
    return (x * 2654435769u) >> (32u - tbits);
}

template <int G, bool HASHED, int FLAGS>
__global__ void k_tc_hash(const uint32_t* __restrict__ items, uint32_t nitems, const uint32_t* __restrict__ rowptr,
                          const uint32_t* __restrict__ col, const uint32_t* __restrict__ wcol, const uint32_t* __restrict__ wedge,
                          const uint8_t* __restrict__ alive, uint32_t* __restrict__ tri, uint32_t* __restrict__ sup,
                          unsigned long long* __restrict__ total, uint32_t ts_max, uint32_t load_shift,
                          const uint32_t* __restrict__ done, const uint32_t* __restrict__ gof, unsigned long long* __restrict__ gcnt) {
    static_assert(G >= 4 && G <= 64 && (64 % G) == 0, "G must divide 64");
    extern __shared__ uint32_t lds[];
    if (done && *done) return;
    const uint32_t sub = threadIdx.x % G;
    const uint32_t gib = threadIdx.x / G;
    const uint32_t gpb = blockDim.x / G;
    uint32_t* tab = lds + (size_t)gib * ts_max * ((FLAGS & F_SUP) ? 2u : 1u);
    uint32_t* tix = tab + ts_max;   // position of each key in N+(h) (F_SUP only)
    uint64_t local = 0;
    for (uint32_t it = blockIdx.x * gpb + gib; it < nitems; it += gridDim.x * gpb) {
        const uint32_t h = items[3 * it], b = items[3 * it + 1], e = items[3 * it + 2];
        if ((FLAGS & F_ALIVE) && !alive[h]) continue;
        const uint32_t hs = rowptr[h], dh = rowptr[h + 1] - hs;
        uint32_t tbits = 0, tsz = 1;
        if (HASHED) {
            // ceil(log2(dh)) + shift, at least 2: never more than the bucket's slots (bucket_slots uses the same rule)
            tbits = (dh > 1 ? 32u - (uint32_t)__clz(dh - 1) : 0u) + load_shift;
            if (tbits < 2) tbits = 2;
            tsz = 1u << tbits;
            for (uint32_t s = sub; s < tsz; s += G) tab[s] = GA_EMPTY;
            group_sync();
            for (uint32_t k = sub; k < dh; k += G) {
                const uint32_t x = col[hs + k];
                if ((FLAGS & F_ALIVE) && !alive[x]) continue;
                uint32_t s = hash_slot(x, tbits);
                for (uint32_t probe = 0; probe < tsz; probe++) {
                    const uint32_t old = atomicCAS(&tab[s], GA_EMPTY, x);
                    if (old == GA_EMPTY) {
                        if (FLAGS & F_SUP) tix[s] = k;
                        break;
                    }
                    s = (s + 1) & (tsz - 1);
                }
            }
            group_sync();
        }
        uint32_t hcount = 0;
        for (uint32_t base = b; base < e; base += G) {
            const uint32_t j = base + sub;
            uint32_t p = 0, ps = 0, len = 0;
            if (j < e) {
                p = wcol[j];
                if (!(FLAGS & F_ALIVE) || alive[p]) {
                    ps = rowptr[p];
                    len = rowptr[p + 1] - ps;
                }
            }
            uint32_t incl = len;   // inclusive scan of len over the group
#pragma unroll
            for (uint32_t o = 1; o < G; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o, G);
                if (sub >= o) incl += t;
            }
            const uint32_t tot = __shfl(incl, G - 1, G);
            for (uint32_t f0 = 0; f0 < tot; f0 += G) {
                const uint32_t f = f0 + sub;
                uint32_t lo = 0, hi = G - 1;   // owner = first lane with incl > f
#pragma unroll
                for (uint32_t st = G; st > 1; st >>= 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    const uint32_t v = __shfl(incl, mid, G);
                    if (v <= f) lo = mid + 1;
                    else hi = mid;
                }
                const uint32_t o_ps = __shfl(ps, lo, G);
                const uint32_t o_end = __shfl(incl, lo, G);
                const uint32_t o_len = __shfl(len, lo, G);
                const uint32_t o_p = __shfl(p, lo, G);
                if (f < tot) {
                    const uint32_t addr = o_ps + (f - (o_end - o_len));
                    const uint32_t x = col[addr];
                    bool hit = false;
                    uint32_t hk = 0;
                    if (HASHED) {
                        uint32_t s = hash_slot(x, tbits);
                        for (uint32_t probe = 0; probe < tsz; probe++) {
                            const uint32_t t = tab[s];
                            if (t == x) { hit = true; if (FLAGS & F_SUP) hk = tix[s]; break; }
                            if (t == GA_EMPTY) break;
                            s = (s + 1) & (tsz - 1);
                        }
                    } else {
                        const uint32_t q = lower_bound_u32(col, hs, hs + dh, x);
                        hit = q < hs + dh && col[q] == x && (!(FLAGS & F_ALIVE) || alive[x]);
                        hk = q - hs;
                    }
                    if (hit) {
                        local++;
                        if (FLAGS & (F_TRI | F_GRAPH)) hcount++;
                        if (FLAGS & F_TRI) {
                            atomicAdd(&tri[x], 1u);
                            atomicAdd(&tri[o_p], 1u);
                        }
                        if (FLAGS & F_SUP) {
                            // work entry of this owner lane: its oriented edge is (h,p) or (p,h)
                            const uint32_t oj = base + lo;
                            atomicAdd(&sup[wedge[oj]], 1u);
                            atomicAdd(&sup[hs + hk], 1u);
                            atomicAdd(&sup[addr], 1u);
                        }
                    }
                }
            }
        }
        if (FLAGS & (F_TRI | F_GRAPH)) {   // triangles of this item: h is a corner of each, h's graph owns them
#pragma unroll
            for (uint32_t o = G / 2; o > 0; o >>= 1) hcount += __shfl_xor(hcount, o, G);
            if (sub == 0 && hcount) {
                if (FLAGS & F_TRI) atomicAdd(&tri[h], hcount);
                if (FLAGS & F_GRAPH) atomicAdd(&gcnt[gof[h]], (unsigned long long)hcount);
            }
        }
        group_sync();   // every lane is done with the table before the next item clears it
    }
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(total, (unsigned long long)local);
}

// ---------------------------------------------------------------------------------------------------------------
// count plan: work entries grouped by hashed endpoint, chunked by work, bucketed by d+(h)
// ---------------------------------------------------------------------------------------------------------------
namespace {

const int NBUCKET = 5;                       // 4 hashed buckets + 1 binary-search bucket for huge d+(h)
const uint32_t BUCKET_DEG[4] = {8, 32, 128, 4096};   // upper d+(h) of hashed buckets 0..3 (bucket 3: hash_cap)

__device__ __forceinline__ void hash_side(const uint32_t* __restrict__ rowptr, uint32_t u, uint32_t v, uint32_t& h, uint32_t& p,
                                          uint32_t& dmin) {
    const uint32_t du = rowptr[u + 1] - rowptr[u], dv = rowptr[v + 1] - rowptr[v];
    const bool hu = du >= dv;
    h = hu ? u : v;
    p = hu ? v : u;
    dmin = hu ? dv : du;
}

__global__ void k_plan_count(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                             uint32_t m, uint32_t* __restrict__ hcnt) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        uint32_t h, p, dmin;
        hash_side(rowptr, src[e], col[e], h, p, dmin);
        if (dmin) atomicAdd(&hcnt[h], 1u);
    }
}

__global__ void k_plan_scatter(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                               uint32_t m, const uint32_t* __restrict__ hrow, uint32_t* __restrict__ cursor, uint32_t* __restrict__ wcol,
                               uint32_t* __restrict__ wedge, uint32_t* __restrict__ wh, unsigned long long* __restrict__ wt) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        uint32_t h, p, dmin;
        hash_side(rowptr, src[e], col[e], h, p, dmin);
        if (!dmin) continue;
        const uint32_t pos = hrow[h] + atomicAdd(&cursor[h], 1u);
        wcol[pos] = p;
        wedge[pos] = e;
        wh[pos] = h;
        wt[pos] = dmin;
    }
}

// flag[j] = 1 where a work item starts: at each row start, and where the work prefix crosses a multiple of chunk.
__global__ void k_plan_flags(const uint32_t* __restrict__ hrow, const uint32_t* __restrict__ wh, const unsigned long long* __restrict__ wsum,
                             uint32_t nwork, uint32_t chunk, uint32_t* __restrict__ flag) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x) {
        const uint32_t r0 = hrow[wh[j]];
        uint32_t f = 1;
        if (j != r0) {
            const unsigned long long base = wsum[r0];
            f = (wsum[j] - base) / chunk != (wsum[j - 1] - base) / chunk;
        }
        flag[j] = f;
    }
}

__device__ __forceinline__ uint32_t bucket_of(uint32_t dh, uint32_t cap) {
    if (dh > cap) return 4;
    if (dh <= 8) return 0;
    if (dh <= 32) return 1;
    if (dh <= 128) return 2;
    if (dh <= 4096) return 3;
    return 4;
}

// Item starts -> per-bucket counts.
__global__ void k_plan_item_count(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ wh, const uint32_t* __restrict__ rowptr,
                                  uint32_t nwork, uint32_t cap, uint32_t* __restrict__ bcount) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x) {
        if (!flag[j]) continue;
        const uint32_t h = wh[j];
        atomicAdd(&bcount[bucket_of(rowptr[h + 1] - rowptr[h], cap)], 1u);
    }
}

// Items placed bucket-major: (h, begin, end). end = next item start (rows are contiguous and every row starts an item).
__global__ void k_plan_items(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ fpos, const uint32_t* __restrict__ starts_by_rank,
                             const uint32_t* __restrict__ wh, const uint32_t* __restrict__ rowptr, uint32_t nwork, uint32_t nall,
                             uint32_t cap, const uint32_t* __restrict__ boff, uint32_t* __restrict__ bcur, uint32_t* __restrict__ items) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x) {
        if (!flag[j]) continue;
        const uint32_t r = fpos[j];
        const uint32_t end = r + 1 < nall ? starts_by_rank[r + 1] : nwork;
        const uint32_t h = wh[j];
        const uint32_t bk = bucket_of(rowptr[h + 1] - rowptr[h], cap);
        const uint32_t pos = boff[bk] + atomicAdd(&bcur[bk], 1u);
        items[3 * pos] = h;
        items[3 * pos + 1] = j;
        items[3 * pos + 2] = end;
    }
}

__global__ void k_plan_rank(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ fpos, uint32_t nwork, uint32_t* __restrict__ starts) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x)
        if (flag[j]) starts[fpos[j]] = j;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

ga_tc_tuning default_tuning() {
    ga_tc_tuning t;
    t.lanes[0] = 8;
    t.lanes[1] = 16;
    t.lanes[2] = 32;
    t.lanes[3] = 64;
    t.block = 256;
    t.work_chunk = 4096;
    t.load_shift = 1;
    t.hash_cap = 4096;
    return t;
}

bool tuning_valid(const ga_tc_tuning& t) {
    for (int i = 0; i < 4; i++)
        if (t.lanes[i] != 4 && t.lanes[i] != 8 && t.lanes[i] != 16 && t.lanes[i] != 32 && t.lanes[i] != 64) return false;
    return t.block >= 64 && t.block <= 1024 && t.block % 64 == 0 && t.work_chunk >= 64 && t.load_shift >= 1 && t.load_shift <= 3 &&
           t.hash_cap >= 1 && t.hash_cap <= 4096;
}

}  // namespace

static int build_plan(ga_ctx* ctx, ga_graph* g) {
    TcPlan& P = g->plan;
    if (!P.tune.block) P.tune = default_tuning();
    const uint32_t n = g->n, m = g->m;
    DevBuf hcnt_b, hrow_b, wh_b, wt_b, wsum_b, flag_b, fpos_b, starts_b, bc_b;
    GA_TRY(ga_malloc(ctx, &hcnt_b.p, ((size_t)n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &hrow_b.p, ((size_t)n + 1) * 4));
    uint32_t* hcnt = (uint32_t*)hcnt_b.p;
    uint32_t* hrow = (uint32_t*)hrow_b.p;
    GA_TRY(ga_fill_u32(ctx, hcnt, (size_t)n + 1, 0));
    if (m) {
        k_plan_count<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, g->src, g->col, m, hcnt);
        GA_TRY(ga_launch_check(ctx, "k_plan_count"));
    }
    GA_TRY(ga_exclusive_scan_u32(ctx, hcnt, hrow, (size_t)n + 1));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, hrow + n, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t nwork = (uint32_t)(ctx->hstat[0] & 0xFFFFFFFFu);
    ga_free(P.wcol);
    ga_free(P.wedge);
    ga_free(P.items);
    P.wcol = P.wedge = P.items = nullptr;
    GA_TRY(ga_malloc(ctx, (void**)&P.wcol, std::max<size_t>(nwork, 1) * 4));
    GA_TRY(ga_malloc(ctx, (void**)&P.wedge, std::max<size_t>(nwork, 1) * 4));
    P.nwork = nwork;
    for (int b = 0; b < 4; b++) P.nitems[b] = 0;
    P.nhuge = 0;
    if (nwork == 0) {
        GA_TRY(ga_malloc(ctx, (void**)&P.items, 12));
        P.ready = true;
        return GA_OK;
    }
    GA_TRY(ga_malloc(ctx, &wh_b.p, (size_t)nwork * 4));
    GA_TRY(ga_malloc(ctx, &wt_b.p, ((size_t)nwork + 1) * 8));
    GA_TRY(ga_malloc(ctx, &wsum_b.p, ((size_t)nwork + 1) * 8));
    GA_TRY(ga_malloc(ctx, &flag_b.p, ((size_t)nwork + 1) * 4));
    GA_TRY(ga_malloc(ctx, &fpos_b.p, ((size_t)nwork + 1) * 4));
    GA_TRY(ga_malloc(ctx, &bc_b.p, 64));
    uint32_t* wh = (uint32_t*)wh_b.p;
    unsigned long long* wt = (unsigned long long*)wt_b.p;
    unsigned long long* wsum = (unsigned long long*)wsum_b.p;
    uint32_t* flag = (uint32_t*)flag_b.p;
    uint32_t* fpos = (uint32_t*)fpos_b.p;
    uint32_t* bc = (uint32_t*)bc_b.p;   // [0..5) counts, [8..13) cursors
    GA_TRY(ga_fill_u32(ctx, hcnt, (size_t)n + 1, 0));   // reused as the scatter cursor
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)wt + nwork, 1, 0));
    k_plan_scatter<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, g->src, g->col, m, hrow, hcnt, P.wcol, P.wedge, wh, wt);
    GA_TRY(ga_launch_check(ctx, "k_plan_scatter"));
    GA_TRY(ga_exclusive_scan_u64(ctx, (const uint64_t*)wt, (uint64_t*)wsum, (size_t)nwork + 1));
    k_plan_flags<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(hrow, wh, wsum, nwork, (uint32_t)P.tune.work_chunk, flag);
    GA_TRY(ga_launch_check(ctx, "k_plan_flags"));
    GA_TRY(ga_fill_u32(ctx, flag + nwork, 1, 0));
    GA_TRY(ga_exclusive_scan_u32(ctx, flag, fpos, (size_t)nwork + 1));
    GA_TRY(ga_fill_u32(ctx, bc, 16, 0));
    const uint32_t cap = (uint32_t)P.tune.hash_cap;
    k_plan_item_count<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(flag, wh, g->rowptr, nwork, cap, bc);
    GA_TRY(ga_launch_check(ctx, "k_plan_item_count"));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, bc, 20, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat + 3, fpos + nwork, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t cnt[NBUCKET];
    const uint32_t* hc = (const uint32_t*)ctx->hstat;
    for (int b = 0; b < NBUCKET; b++) cnt[b] = hc[b];
    const uint32_t nall = (uint32_t)(ctx->hstat[3] & 0xFFFFFFFFu);
    uint32_t off[8] = {0};
    for (int b = 0; b < NBUCKET; b++) off[b + 1] = off[b] + cnt[b];
    if (off[NBUCKET] != nall) return ga_fail(ctx, GA_ERR_STATE, "count plan: %u items bucketed, %u expected", off[NBUCKET], nall);
    GA_TRY(ga_malloc(ctx, &starts_b.p, (size_t)nall * 4));
    GA_TRY(ga_malloc(ctx, (void**)&P.items, (size_t)nall * 12));
    k_plan_rank<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(flag, fpos, nwork, (uint32_t*)starts_b.p);
    GA_TRY(ga_launch_check(ctx, "k_plan_rank"));
    GA_HIP(ctx, hipMemcpyAsync(bc, off, 8 * 4, hipMemcpyHostToDevice, ctx->stream));
    k_plan_items<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(flag, fpos, (uint32_t*)starts_b.p, wh, g->rowptr, nwork, nall, cap,
                                                                      bc, bc + 8, P.items);
    GA_TRY(ga_launch_check(ctx, "k_plan_items"));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < 4; b++) P.nitems[b] = cnt[b];
    P.nhuge = cnt[4];
    P.ready = true;
    return GA_OK;
}

int ga_tc_ensure_plan(ga_ctx* ctx, ga_graph* g) {
    if (g->plan.ready) return GA_OK;
    return build_plan(ctx, g);
}

// table slots for the largest d+(h) a bucket may hold (h in bucket b has d+(h) <= bound):
// 2^max(2, ceil(log2(bound)) + shift), the kernel's per-item rule at d+(h) = bound
static uint32_t bucket_slots(uint32_t bound, uint32_t shift) {
    uint32_t lg = 0;
    while ((1u << lg) < bound) lg++;
    uint32_t bits = lg + shift;
    if (bits < 2) bits = 2;
    return 1u << bits;
}

template <int G, bool HASHED, int FLAGS>
static int launch_hash_bucket(ga_ctx* ctx, ga_graph* g, const uint32_t* items, uint32_t nitems, uint32_t bound,
                              unsigned long long* total, const uint8_t* alive, uint32_t* tri, uint32_t* sup, const uint32_t* done,
                              const uint32_t* gof, unsigned long long* gcnt) {
    if (!nitems) return GA_OK;
    const TcPlan& P = g->plan;
    uint32_t block = (uint32_t)P.tune.block;
    const uint32_t ts = HASHED ? bucket_slots(std::min(bound, std::max(g->max_out, 1u)), (uint32_t)P.tune.load_shift) : 0;
    const uint32_t per_group = ts * 4u * ((FLAGS & F_SUP) ? 2u : 1u);
    while (block > 64 && (size_t)(block / G) * per_group > 64 * 1024) block >>= 1;
    const size_t shmem = (size_t)(block / G) * per_group;
    if (shmem > 160 * 1024) return ga_fail(ctx, GA_ERR_ARG, "hash table of %u slots does not fit LDS", ts);
    const uint32_t gpb = block / G;
    const unsigned grid = grid_for((uint64_t)nitems * G, block, (unsigned)std::max(ctx->cus, 1) * 64u);
    (void)gpb;
    k_tc_hash<G, HASHED, FLAGS><<<grid, block, shmem, ctx->stream>>>(items, nitems, g->rowptr, g->col, P.wcol, P.wedge, alive, tri, sup,
                                                                      total, ts, (uint32_t)P.tune.load_shift, done, gof, gcnt);
    return ga_launch_check(ctx, "k_tc_hash");
}

template <int G, int FLAGS>
static int launch_hash_g(ga_ctx* ctx, ga_graph* g, const uint32_t* items, uint32_t nitems, uint32_t bound, unsigned long long* total,
                         const uint8_t* alive, uint32_t* tri, uint32_t* sup, const uint32_t* done, const uint32_t* gof,
                         unsigned long long* gcnt) {
    return launch_hash_bucket<G, true, FLAGS>(ctx, g, items, nitems, bound, total, alive, tri, sup, done, gof, gcnt);
}

template <int FLAGS>
static int launch_hash(ga_ctx* ctx, ga_graph* g, unsigned long long* total, const uint8_t* alive, uint32_t* tri, uint32_t* sup,
                       const uint32_t* done, const uint32_t* gof, unsigned long long* gcnt) {
    const TcPlan& P = g->plan;
    uint32_t off = 0;
    for (int b = 0; b < 4; b++) {
        const uint32_t* items = P.items + 3ull * off;
        const uint32_t ni = P.nitems[b];
        const uint32_t bound = std::min(BUCKET_DEG[b], (uint32_t)P.tune.hash_cap);
        int rc = GA_OK;
        switch (P.tune.lanes[b]) {
            case 4: rc = launch_hash_g<4, FLAGS>(ctx, g, items, ni, bound, total, alive, tri, sup, done, gof, gcnt); break;
            case 8: rc = launch_hash_g<8, FLAGS>(ctx, g, items, ni, bound, total, alive, tri, sup, done, gof, gcnt); break;
            case 16: rc = launch_hash_g<16, FLAGS>(ctx, g, items, ni, bound, total, alive, tri, sup, done, gof, gcnt); break;
            case 32: rc = launch_hash_g<32, FLAGS>(ctx, g, items, ni, bound, total, alive, tri, sup, done, gof, gcnt); break;
            default: rc = launch_hash_g<64, FLAGS>(ctx, g, items, ni, bound, total, alive, tri, sup, done, gof, gcnt); break;
        }
        GA_TRY(rc);
        off += ni;
    }
    return launch_hash_bucket<64, false, FLAGS>(ctx, g, P.items + 3ull * off, P.nhuge, 0, total, alive, tri, sup, done, gof, gcnt);
}

template <int FLAGS>
static int launch_merge(ga_ctx* ctx, ga_graph* g, const uint32_t* order, unsigned long long* total, const uint8_t* alive, uint32_t* tri,
                        uint32_t* sup, const uint32_t* done) {
    if (!g->m) return GA_OK;
    k_tc_merge<FLAGS><<<grid_for(g->m, 256), 256, 0, ctx->stream>>>(g->rowptr, g->col, g->src, order, g->m, alive, tri, sup, total, done);
    return ga_launch_check(ctx, "k_tc_merge");
}

template <int G>
static int launch_group(ga_ctx* ctx, ga_graph* g, unsigned long long* total) {
    if (!g->m) return GA_OK;
    k_tc_group<G><<<grid_for((uint64_t)g->m * G, 256), 256, 0, ctx->stream>>>(g->rowptr, g->col, g->src, g->m, total);
    return ga_launch_check(ctx, "k_tc_group");
}

// Launches one count of `variant` into *d_total (zeroed by the caller). alive/tri/sup as in the flags.
int ga_tc_launch(ga_ctx* ctx, ga_graph* g, int variant, unsigned long long* d_total, const uint8_t* alive, uint32_t* tri, uint32_t* sup,
                 const uint32_t* done, const uint32_t* gof, unsigned long long* gcnt) {
    unsigned long long* total = d_total;
    if (variant == GA_TC_AUTO) variant = GA_TC_HASH;
    const int flags = (alive ? F_ALIVE : 0) | (tri ? F_TRI : 0) | (sup ? F_SUP : 0) | (gcnt ? F_GRAPH : 0);
    if (variant == GA_TC_HASH) {
        GA_TRY(ga_tc_ensure_plan(ctx, g));
        switch (flags) {
            case 0: return launch_hash<0>(ctx, g, total, alive, tri, sup, done, gof, gcnt);
            case F_TRI: return launch_hash<F_TRI>(ctx, g, total, alive, tri, sup, done, gof, gcnt);
            case F_ALIVE | F_TRI: return launch_hash<F_ALIVE | F_TRI>(ctx, g, total, alive, tri, sup, done, gof, gcnt);
            case F_SUP: return launch_hash<F_SUP>(ctx, g, total, alive, tri, sup, done, gof, gcnt);
            case F_GRAPH: return launch_hash<F_GRAPH>(ctx, g, total, alive, tri, sup, done, gof, gcnt);
            case F_ALIVE | F_TRI | F_GRAPH: return launch_hash<F_ALIVE | F_TRI | F_GRAPH>(ctx, g, total, alive, tri, sup, done, gof, gcnt);
            default: return ga_fail(ctx, GA_ERR_ARG, "unsupported flag combination %d for the hash variant", flags);
        }
    }
    if (gcnt) return ga_fail(ctx, GA_ERR_ARG, "per-graph counts need the hash variant");
    if (variant == GA_TC_MERGE || variant == GA_TC_SORTED) {
        const uint32_t* order = nullptr;
        if (variant == GA_TC_SORTED) {
            if (flags) return ga_fail(ctx, GA_ERR_ARG, "the sorted variant only counts");
            GA_TRY(ga_graph_ensure_perm(ctx, g));
            order = g->plan.perm;
        }
        switch (flags) {
            case 0: return launch_merge<0>(ctx, g, order, total, alive, tri, sup, done);
            case F_TRI: return launch_merge<F_TRI>(ctx, g, order, total, alive, tri, sup, done);
            case F_ALIVE | F_TRI: return launch_merge<F_ALIVE | F_TRI>(ctx, g, order, total, alive, tri, sup, done);
            case F_SUP: return launch_merge<F_SUP>(ctx, g, order, total, alive, tri, sup, done);
            default: return ga_fail(ctx, GA_ERR_ARG, "unsupported flag combination %d for the merge variant", flags);
        }
    }
    if (flags) return ga_fail(ctx, GA_ERR_ARG, "variant %d only counts", variant);
    switch (variant) {
        case GA_TC_GROUP4: return launch_group<4>(ctx, g, total);
        case GA_TC_GROUP8: return launch_group<8>(ctx, g, total);
        case GA_TC_GROUP16: return launch_group<16>(ctx, g, total);
        case GA_TC_GROUP64: return launch_group<64>(ctx, g, total);
        default: return ga_fail(ctx, GA_ERR_ARG, "unknown variant %d", variant);
    }
}

extern "C" const char* ga_tc_variant_name(int variant) {
    switch (variant) {
        case GA_TC_AUTO: return "auto";
        case GA_TC_MERGE: return "merge_lane_per_edge";
        case GA_TC_GROUP4: return "group4_bsearch";
        case GA_TC_GROUP8: return "group8_bsearch";
        case GA_TC_GROUP16: return "group16_bsearch";
        case GA_TC_GROUP64: return "group64_bsearch";
        case GA_TC_HASH: return "hash_lds_bucketed";
        case GA_TC_SORTED: return "merge_work_sorted";
        default: return "unknown";
    }
}

extern "C" int ga_tc_prepare(ga_ctx* ctx, ga_graph* g, float* ms) {
    if (!ctx || !g) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    g->plan.ready = false;
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(build_plan(ctx, g));
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}

extern "C" int ga_tc_count(ga_ctx* ctx, ga_graph* g, int variant, uint64_t* T, float* ms) {
    if (!ctx || !g || !T) return GA_ERR_ARG;
    if (variant < 0 || variant >= GA_TC_NVARIANTS) return ga_fail(ctx, GA_ERR_ARG, "unknown variant %d", variant);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (variant == GA_TC_AUTO || variant == GA_TC_HASH) GA_TRY(ga_tc_ensure_plan(ctx, g));
    if (variant == GA_TC_SORTED) GA_TRY(ga_graph_ensure_perm(ctx, g));
    uint64_t* d = ctx->dsmall;
    GA_TRY(ga_fill_u64(ctx, d, 1, 0));
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(ga_tc_launch(ctx, g, variant, (unsigned long long*)d, nullptr, nullptr, nullptr, nullptr));
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *T = ctx->hstat[0];
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}

extern "C" int ga_tc_get_tuning(const ga_graph* g, ga_tc_tuning* out) {
    if (!g || !out) return GA_ERR_ARG;
    *out = g->plan.tune.block ? g->plan.tune : default_tuning();
    return GA_OK;
}

extern "C" int ga_tc_set_tuning(ga_ctx* ctx, ga_graph* g, const ga_tc_tuning* t) {
    if (!ctx || !g || !t) return GA_ERR_ARG;
    if (!tuning_valid(*t)) return ga_fail(ctx, GA_ERR_ARG, "invalid tuning");
    const bool replan = !g->plan.tune.block || g->plan.tune.work_chunk != t->work_chunk || g->plan.tune.hash_cap != t->hash_cap;
    g->plan.tune = *t;
    if (replan) g->plan.ready = false;
    return GA_OK;
}

extern "C" int ga_tc_tune(ga_ctx* ctx, ga_graph* g, ga_tc_tuning* best, float* best_ms) {
    if (!ctx || !g) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    ga_tc_tuning base = g->plan.tune.block ? g->plan.tune : default_tuning();
    GA_TRY(ga_tc_set_tuning(ctx, g, &base));
    uint64_t T0 = 0;
    float ms = 0;
    GA_TRY(ga_tc_count(ctx, g, GA_TC_HASH, &T0, &ms));
    ga_tc_tuning bt = base;
    float bms = 1e30f;
    auto trial = [&](const ga_tc_tuning& t) -> int {
        GA_TRY(ga_tc_set_tuning(ctx, g, &t));
        float best_here = 1e30f;
        for (int r = 0; r < 3; r++) {
            uint64_t T = 0;
            float x = 0;
            GA_TRY(ga_tc_count(ctx, g, GA_TC_HASH, &T, &x));
            if (T != T0) return ga_fail(ctx, GA_ERR_STATE, "tuning trial changed the count (%llu vs %llu)", (unsigned long long)T,
                                        (unsigned long long)T0);
            best_here = std::min(best_here, x);
        }
        if (best_here < bms) { bms = best_here; bt = t; }
        return GA_OK;
    };
    // coordinate search: lanes per bucket, then block, load shift and work chunk
    GA_TRY(trial(base));
    const int lane_opts[5] = {4, 8, 16, 32, 64};
    for (int b = 0; b < 4; b++) {
        for (int li = 0; li < 5; li++) {
            ga_tc_tuning t = bt;
            t.lanes[b] = lane_opts[li];
            if (t.lanes[b] == bt.lanes[b]) continue;
            if ((b == 3 && t.lanes[b] < 32) || (b == 2 && t.lanes[b] < 8)) continue;   // tables would not fit LDS
            GA_TRY(trial(t));
        }
    }
    const int blocks[3] = {128, 256, 512};
    for (int bi = 0; bi < 3; bi++) { ga_tc_tuning t = bt; t.block = blocks[bi]; if (t.block != bt.block) GA_TRY(trial(t)); }
    for (int ls = 1; ls <= 2; ls++) { ga_tc_tuning t = bt; t.load_shift = ls; if (ls != bt.load_shift) GA_TRY(trial(t)); }
    const int chunks[3] = {1024, 4096, 16384};
    for (int ci = 0; ci < 3; ci++) { ga_tc_tuning t = bt; t.work_chunk = chunks[ci]; if (t.work_chunk != bt.work_chunk) GA_TRY(trial(t)); }
    GA_TRY(ga_tc_set_tuning(ctx, g, &bt));
    GA_TRY(ga_tc_ensure_plan(ctx, g));
    if (best) *best = bt;
    if (best_ms) *best_ms = bms;
    return GA_OK;
}

// ---- per-vertex participation and per-edge support ----
extern "C" int ga_tc_vertex(ga_ctx* ctx, ga_graph* g, int variant, const uint8_t* alive, uint32_t* tri, uint64_t* T) {
    if (!ctx || !g || !tri) return GA_ERR_ARG;
    if (variant == GA_TC_AUTO) variant = GA_TC_HASH;
    if (variant != GA_TC_HASH && variant != GA_TC_MERGE) return ga_fail(ctx, GA_ERR_ARG, "variant %d has no per-vertex form", variant);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf a, t, d;
    GA_TRY(ga_malloc(ctx, &t.p, ((size_t)g->n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &d.p, 8));
    if (alive) {
        GA_TRY(ga_malloc(ctx, &a.p, (size_t)g->n + 1));
        GA_HIP(ctx, hipMemcpyAsync(a.p, alive, g->n, hipMemcpyHostToDevice, ctx->stream));
    }
    GA_TRY(ga_fill_u32(ctx, (uint32_t*)t.p, g->n, 0));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)d.p, 1, 0));
    GA_TRY(ga_tc_launch(ctx, g, variant, (unsigned long long*)d.p, (const uint8_t*)a.p, (uint32_t*)t.p, nullptr, nullptr));
    if (g->n) GA_HIP(ctx, hipMemcpyAsync(tri, t.p, (size_t)g->n * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (T) *T = ctx->hstat[0];
    return GA_OK;
}

extern "C" int ga_tc_edge_support(ga_ctx* ctx, ga_graph* g, int variant, uint32_t* sup, uint64_t* T) {
    if (!ctx || !g || !sup) return GA_ERR_ARG;
    if (variant == GA_TC_AUTO) variant = GA_TC_HASH;
    if (variant != GA_TC_HASH && variant != GA_TC_MERGE) return ga_fail(ctx, GA_ERR_ARG, "variant %d has no per-edge form", variant);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf s, d;
    GA_TRY(ga_malloc(ctx, &s.p, ((size_t)g->m + 1) * 4));
    GA_TRY(ga_malloc(ctx, &d.p, 8));
    GA_TRY(ga_fill_u32(ctx, (uint32_t*)s.p, g->m, 0));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)d.p, 1, 0));
    GA_TRY(ga_tc_launch(ctx, g, variant, (unsigned long long*)d.p, nullptr, nullptr, (uint32_t*)s.p, nullptr));
    if (g->m) GA_HIP(ctx, hipMemcpyAsync(sup, s.p, (size_t)g->m * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (T) *T = ctx->hstat[0];
    return GA_OK;
}
