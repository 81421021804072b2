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
// Work entries: every oriented edge (u,v) with min(d+u, d+v) > 0, filed under the endpoint h with the LONGER out-list
// (ties: u), the other endpoint p being the probe side. Entries are sorted by (h, d+(p) descending), so each row of h
// lists its probe lists longest first. Rows are cut into items (h, begin, end) of bounded probe work and bounded
// batch count, and items are bucketed by d+(h) (the LDS table size).
//
// One lane group of G lanes owns an item:
//   1. clear a table of next_pow2(d+(h)) << load_shift slots (at least twice the entries: every probe chain ends)
//   2. insert N+(h) by LDS compare-and-swap with linear probing (dead vertices are not inserted)
//   3. take the item's probe lists G at a time (a batch), compact the non-empty ones in LDS (they stay longest first),
//      and stream them in phases: lists of length in (S/2, S] are read by sub-groups of S lanes (G/S lists at once,
//      one element per lane, consecutive lanes on consecutive addresses); lists longer than G/2 take the whole group.
//      Every element probes the table; a hit is the triangle (h, p, x).
// HASHED = false is the bucket for d+(h) above the LDS cap: the element is binary-searched in N+(h) in HBM instead.

__device__ __forceinline__ void group_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t hash_slot(uint32_t x, uint32_t tbits) {
    return (x * 2654435769u) >> (32u - tbits);
}

// This lane group's bits of a wave ballot, shifted down to bit 0.
template <int G>
__device__ __forceinline__ uint64_t group_bits(uint64_t wave_mask) {
    if (G == 64) return wave_mask;
    const uint32_t goff = lane_id() & ~(uint32_t)(G - 1);
    return (wave_mask >> goff) & ((1ull << G) - 1ull);
}

struct TcArgs {
    const uint32_t* items;
    uint32_t nitems;
    uint32_t ts_max;       // table slots reserved per group
    uint32_t load_shift;
    const uint32_t* rowptr;
    const uint32_t* col;
    const uint32_t* wcol;
    const uint32_t* wedge;
    const uint32_t* wps;     // rowptr[p] of each work entry
    const uint32_t* wlen;    // d+(p) of each work entry
    const uint8_t* alive;
    uint32_t* tri;
    uint32_t* sup;
    unsigned long long* total;
    const uint32_t* done;
    const uint32_t* gof;
    unsigned long long* gcnt;
};

template <int G, bool HASHED, int FLAGS>
__global__ void k_tc_hash(const TcArgs a) {
    static_assert(G >= 4 && G <= 64 && (64 % G) == 0, "G must divide 64");
    extern __shared__ uint32_t lds[];
    if (a.done && *a.done) return;
    const uint32_t* __restrict__ rowptr = a.rowptr;
    const uint32_t* __restrict__ col = a.col;
    const uint8_t* __restrict__ alive = a.alive;
    const uint32_t sub = threadIdx.x % G;
    const uint32_t gib = threadIdx.x / G;
    const uint32_t gpb = blockDim.x / G;
    const uint32_t tab_words = HASHED ? a.ts_max * ((FLAGS & (F_SUP | F_TRI)) ? 2u : 1u) : 0u;
    uint32_t* tab = lds + (size_t)gib * (tab_words + 4u * G);
    uint32_t* tix = tab + a.ts_max;          // F_SUP: position of each key in N+(h); F_TRI: hits per key
    uint32_t* stg = tab + tab_words;         // batch staging: ps[G] | len[G] | p[G] | work index[G]
    uint64_t local = 0;
    for (uint32_t it = blockIdx.x * gpb + gib; it < a.nitems; it += gridDim.x * gpb) {
        const uint32_t h = a.items[3 * it], b = a.items[3 * it + 1], e = a.items[3 * it + 2];
        if ((FLAGS & F_ALIVE) && !alive[h]) continue;
        const uint32_t hs = rowptr[h], dh = rowptr[h + 1] - hs;
        uint32_t tbits = 0, tsz = 1;
        if (HASHED) {
            // ceil(log2(dh)) + shift, at least 2: never more than the bucket's slots (bucket_slots uses the same rule)
            tbits = (dh > 1 ? 32u - (uint32_t)__clz(dh - 1) : 0u) + a.load_shift;
            if (tbits < 2) tbits = 2;
            tsz = 1u << tbits;
            for (uint32_t s = sub; s < tsz; s += G) {
                tab[s] = GA_EMPTY;
                if (FLAGS & F_TRI) tix[s] = 0;
            }
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
        // batch metadata of the first batch; later batches are fetched one batch ahead
        uint32_t n_p = 0, n_ps = 0, n_len = 0;
        {
            const uint32_t j = b + sub;
            if (j < e) {
                n_ps = a.wps[j];
                n_len = a.wlen[j];
                if (FLAGS & (F_ALIVE | F_TRI)) n_p = a.wcol[j];
            }
        }
        for (uint32_t base = b; base < e; base += G) {
            const uint32_t j = base + sub;
            const uint32_t p = n_p, ps = n_ps;
            uint32_t len = n_len;
            if ((FLAGS & F_ALIVE) && len && !alive[p]) len = 0;
            const uint64_t gm = group_bits<G>(__ballot(len > 0));
            const uint32_t cnt = (uint32_t)__popcll(gm);
            if (len > 0) {   // stable compaction keeps the batch longest first
                const uint32_t r = (uint32_t)__popcll(gm & ((1ull << sub) - 1ull));
                stg[r] = ps;
                stg[G + r] = len;
                if (FLAGS & F_TRI) stg[2 * G + r] = p;
                if (FLAGS & F_SUP) stg[3 * G + r] = j;
            }
            {   // prefetch the next batch while this one is streamed
                const uint32_t jn = base + G + sub;
                n_p = n_ps = n_len = 0;
                if (jn < e) {
                    n_ps = a.wps[jn];
                    n_len = a.wlen[jn];
                    if (FLAGS & (F_ALIVE | F_TRI)) n_p = a.wcol[jn];
                }
            }
            group_sync();
            const uint32_t lk = sub < cnt ? stg[G + sub] : 0u;
            uint32_t start = 0;
#pragma unroll
            for (uint32_t S = G; S >= 1; S >>= 1) {
                // entries [start, end) have length in (S/2, S] (for S = G: every length above G/2); each list is read
                // by W = S/U lanes (at least 1), every lane holding up to U loads in flight
                constexpr uint32_t U = 4;
                const uint32_t W = S >= U ? S / U : 1u;
                const uint32_t end = (uint32_t)__popcll(group_bits<G>(__ballot(lk > S / 2)));
                for (uint32_t k0 = start; k0 < end; k0 += G / W) {
                    const uint32_t k = k0 + sub / W, r = sub % W;
                    uint32_t c = 0;
                    if (k < end) {
                        const uint32_t kps = stg[k], klen = stg[G + k];
                        for (uint32_t off = r; off < klen; off += U * W) {
                            uint32_t x[U];
#pragma unroll
                            for (uint32_t u = 0; u < U; u++) x[u] = off + u * W < klen ? col[kps + off + u * W] : 0u;
#pragma unroll
                            for (uint32_t u = 0; u < U; u++) {
                                if (off + u * W >= klen) break;
                                const uint32_t addr = kps + off + u * W;
                                bool hit = false;
                                uint32_t hk = 0;
                                if (HASHED) {
                                    uint32_t s = hash_slot(x[u], tbits);
                                    for (uint32_t probe = 0; probe < tsz; probe++) {
                                        const uint32_t t = tab[s];
                                        if (t == x[u]) {
                                        hit = true;
                                        if (FLAGS & F_SUP) hk = tix[s];
                                        if (FLAGS & F_TRI) atomicAdd(&tix[s], 1u);   // x's hits gather in LDS
                                        break;
                                    }
                                        if (t == GA_EMPTY) break;
                                        s = (s + 1) & (tsz - 1);
                                    }
                                } else {
                                    const uint32_t q = lower_bound_u32(col, hs, hs + dh, x[u]);
                                    hit = q < hs + dh && col[q] == x[u] && (!(FLAGS & F_ALIVE) || alive[x[u]]);
                                    hk = q - hs;
                                }
                                if (hit) {
                                    c++;
                                    if (!HASHED && (FLAGS & F_TRI)) atomicAdd(&a.tri[x[u]], 1u);
                                    if (FLAGS & F_SUP) {
                                        atomicAdd(&a.sup[hs + hk], 1u);
                                        atomicAdd(&a.sup[addr], 1u);
                                    }
                                }
                            }
                        }
                    }
                    local += c;
                    if (FLAGS & (F_TRI | F_GRAPH)) hcount += c;
                    if (FLAGS & (F_TRI | F_SUP)) {   // hits of list k, summed over its W lanes, go to p and edge (h,p)
#pragma unroll
                        for (uint32_t o = W / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, G);
                        if (r == 0 && k < end && c) {
                            if (FLAGS & F_TRI) atomicAdd(&a.tri[stg[2 * G + k]], c);
                            if (FLAGS & F_SUP) atomicAdd(&a.sup[a.wedge[stg[3 * G + k]]], c);
                        }
                    }
                }
                start = end;
                if (S == 1) break;
            }
            group_sync();   // staging is rewritten by the next batch
        }
        if (HASHED && (FLAGS & F_TRI)) {   // one atomic per hit key of N+(h) instead of one per triangle
            group_sync();
            for (uint32_t s = sub; s < tsz; s += G) {
                const uint32_t cx = tix[s];
                if (cx) atomicAdd(&a.tri[tab[s]], cx);
            }
        }
        if (FLAGS & (F_TRI | F_GRAPH)) {   // triangles of this item: h is a corner of each, h's graph owns them
#pragma unroll
            for (uint32_t o = G / 2; o > 0; o >>= 1) hcount += __shfl_xor(hcount, o, G);
            if (sub == 0 && hcount) {
                if (FLAGS & F_TRI) atomicAdd(&a.tri[h], hcount);
                if (FLAGS & F_GRAPH) atomicAdd(&a.gcnt[a.gof[h]], (unsigned long long)hcount);
            }
        }
        group_sync();   // every lane is done with the table before the next item clears it
    }
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(a.total, (unsigned long long)local);
}

// ---------------------------------------------------------------------------------------------------------------
// count plan
// ---------------------------------------------------------------------------------------------------------------
namespace {

const int NHASH = 5;                                        // hashed buckets
const uint32_t BUCKET_DEG[NHASH] = {8, 32, 128, 512, 4096};  // upper d+(h) per hashed bucket (the last one: hash_cap)

__device__ __forceinline__ void hash_side(const uint32_t* __restrict__ rowptr, uint32_t u, uint32_t v, uint32_t& h, uint32_t& p,
                                          uint32_t& dmin) {
    const uint32_t du = rowptr[u + 1] - rowptr[u], dv = rowptr[v + 1] - rowptr[v];
    const bool hu = du >= dv;
    h = hu ? u : v;
    p = hu ? v : u;
    dmin = hu ? dv : du;
}

__device__ __forceinline__ uint32_t bucket_of(uint32_t dh, uint32_t cap) {
    if (dh > cap) return NHASH;
    return dh <= 8 ? 0u : dh <= 32 ? 1u : dh <= 128 ? 2u : dh <= 512 ? 3u : 4u;
}

struct PlanShape {   // entries per item before a row is cut, per bucket (lanes x batches)
    uint32_t pmax[NHASH + 1];
};

// key = (h << 32) | ~d+(p): rows of h, longest probe list first; edges without work get the all-ones key (last).
__global__ void k_plan_keys(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                            uint32_t m, unsigned long long* __restrict__ key, uint32_t* __restrict__ val, uint32_t* __restrict__ nwork) {
    uint32_t cnt = 0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        uint32_t h, p, dmin;
        hash_side(rowptr, src[e], col[e], h, p, dmin);
        key[e] = dmin ? (((unsigned long long)h << 32) | (uint32_t)~dmin) : ~0ull;
        val[e] = e;
        cnt += dmin != 0;
    }
    cnt = (uint32_t)wave_sum_u64(cnt);
    if (lane_id() == 0 && cnt) atomicAdd(nwork, cnt);
}

__global__ void k_plan_fill(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                            const uint32_t* __restrict__ val, uint32_t nwork, uint32_t* __restrict__ wcol, uint32_t* __restrict__ wedge,
                            uint32_t* __restrict__ wps, uint32_t* __restrict__ wlen, uint32_t* __restrict__ wh,
                            unsigned long long* __restrict__ wt) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x) {
        const uint32_t e = val[j];
        uint32_t h, p, dmin;
        hash_side(rowptr, src[e], col[e], h, p, dmin);
        wcol[j] = p;
        wedge[j] = e;
        wps[j] = rowptr[p];
        wlen[j] = dmin;
        wh[j] = h;
        wt[j] = dmin;
    }
}

__global__ void k_plan_rowstart(const uint32_t* __restrict__ wh, uint32_t nwork, uint32_t* __restrict__ hstart) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x)
        if (j == 0 || wh[j - 1] != wh[j]) hstart[wh[j]] = j;
}

// flag[j] = 1 where an item starts: at a row start, where the row's work prefix crosses a multiple of `chunk`, and
// every `pmax` entries of the row (pmax = lanes x batches of h's bucket).
__global__ void k_plan_flags(const uint32_t* __restrict__ hstart, const uint32_t* __restrict__ wh, const unsigned long long* __restrict__ wsum,
                             const uint32_t* __restrict__ rowptr, uint32_t nwork, uint32_t chunk, uint32_t cap, PlanShape shape,
                             uint32_t* __restrict__ flag) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x) {
        const uint32_t h = wh[j];
        const uint32_t r0 = hstart[h];
        uint32_t f = 1;
        if (j != r0) {
            const unsigned long long base = wsum[r0];
            const uint32_t pmax = shape.pmax[bucket_of(rowptr[h + 1] - rowptr[h], cap)];
            f = ((wsum[j] - base) / chunk != (wsum[j - 1] - base) / chunk) || ((j - r0) % pmax == 0);
        }
        flag[j] = f;
    }
}

// Item starts -> per-bucket counts (one atomic per bucket per wave).
__global__ void k_plan_item_count(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ wh, const uint32_t* __restrict__ rowptr,
                                  uint32_t nwork, uint32_t cap, uint32_t* __restrict__ bcount) {
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t base = wave0 * 64; base < nwork; base += nwaves * 64) {
        const uint64_t j = base + lane_id();
        uint32_t bk = 0xFFu;
        if (j < nwork && flag[j]) {
            const uint32_t h = wh[j];
            bk = bucket_of(rowptr[h + 1] - rowptr[h], cap);
        }
#pragma unroll
        for (uint32_t b = 0; b <= NHASH; b++) {
            const uint64_t mk = __ballot(bk == b);
            if (mk && lane_id() == (uint32_t)__ffsll((unsigned long long)mk) - 1) atomicAdd(&bcount[b], (uint32_t)__popcll(mk));
        }
    }
}

// Items placed bucket-major: (h, begin, end). end = next item start (rows are contiguous and every row starts an item).
__global__ void k_plan_items(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ fpos, const uint32_t* __restrict__ starts_by_rank,
                             const uint32_t* __restrict__ wh, const uint32_t* __restrict__ rowptr, uint32_t nwork, uint32_t nall,
                             uint32_t cap, const uint32_t* __restrict__ boff, uint32_t* __restrict__ bcur, uint32_t* __restrict__ items) {
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = ((uint64_t)gridDim.x * blockDim.x) / 64;
    for (uint64_t base = wave0 * 64; base < nwork; base += nwaves * 64) {
        const uint64_t j = base + lane_id();
        uint32_t bk = 0xFFu, h = 0;
        if (j < nwork && flag[j]) {
            h = wh[j];
            bk = bucket_of(rowptr[h + 1] - rowptr[h], cap);
        }
#pragma unroll
        for (uint32_t b = 0; b <= NHASH; b++) {
            const uint64_t mk = __ballot(bk == b);
            if (!mk) continue;
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mk) - 1;
            uint32_t pos0 = 0;
            if (lane_id() == leader) pos0 = atomicAdd(&bcur[b], (uint32_t)__popcll(mk));
            pos0 = __shfl(pos0, leader, 64);
            if (bk == b) {
                const uint32_t r = fpos[j];
                const uint32_t pos = boff[b] + pos0 + mask_rank(mk);
                items[3 * pos] = h;
                items[3 * pos + 1] = (uint32_t)j;
                items[3 * pos + 2] = r + 1 < nall ? starts_by_rank[r + 1] : nwork;
            }
        }
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
    t.lanes[4] = 64;
    t.block = 256;
    t.work_chunk = 8192;
    t.batches = 8;
    t.load_shift = 2;
    t.hash_cap = 4096;
    return t;
}

bool tuning_valid(const ga_tc_tuning& t) {
    for (int i = 0; i < NHASH; i++)
        if (t.lanes[i] != 4 && t.lanes[i] != 8 && t.lanes[i] != 16 && t.lanes[i] != 32 && t.lanes[i] != 64) return false;
    return t.block >= 64 && t.block <= 1024 && t.block % 64 == 0 && t.work_chunk >= 64 && t.batches >= 1 && t.batches <= 1024 &&
           t.load_shift >= 1 && t.load_shift <= 3 && t.hash_cap >= 1 && t.hash_cap <= 4096;
}

PlanShape plan_shape(const ga_tc_tuning& t) {
    PlanShape s;
    for (int b = 0; b < NHASH; b++) s.pmax[b] = (uint32_t)(t.lanes[b] * t.batches);
    s.pmax[NHASH] = (uint32_t)(64 * t.batches);
    return s;
}

}  // namespace

static int build_plan(ga_ctx* ctx, ga_graph* g) {
    TcPlan& P = g->plan;
    if (!P.tune.block) P.tune = default_tuning();
    const uint32_t n = g->n, m = g->m;
    ga_free(P.wcol);
    ga_free(P.wedge);
    ga_free(P.wps);
    ga_free(P.wlen);
    ga_free(P.items);
    P.wcol = P.wedge = P.wps = P.wlen = P.items = nullptr;
    for (int b = 0; b < NHASH; b++) P.nitems[b] = 0;
    P.nhuge = 0;
    P.nwork = 0;
    DevBuf keys, keys2, vals, vals2, hstart_b, wh_b, wt_b, wsum_b, flag_b, fpos_b, starts_b, bc_b;
    GA_TRY(ga_malloc(ctx, &bc_b.p, 128));
    uint32_t* bc = (uint32_t*)bc_b.p;   // [0..6) bucket counts, [8..14) cursors, [16..24) offsets, [31] nwork
    GA_TRY(ga_fill_u32(ctx, bc, 32, 0));
    uint32_t nwork = 0;
    uint32_t* sorted_vals = nullptr;
    if (m) {
        GA_TRY(ga_malloc(ctx, &keys.p, (size_t)m * 8));
        GA_TRY(ga_malloc(ctx, &keys2.p, (size_t)m * 8));
        GA_TRY(ga_malloc(ctx, &vals.p, (size_t)m * 4));
        GA_TRY(ga_malloc(ctx, &vals2.p, (size_t)m * 4));
        k_plan_keys<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, g->src, g->col, m, (unsigned long long*)keys.p,
                                                                     (uint32_t*)vals.p, bc + 31);
        GA_TRY(ga_launch_check(ctx, "k_plan_keys"));
        int nb = 1;
        while (nb < 32 && (1ull << nb) < (uint64_t)n) nb++;
        unsigned long long* ko = nullptr;
        GA_TRY(ga_sort_pairs_u64(ctx, (unsigned long long*)keys.p, (unsigned long long*)keys2.p, (uint32_t*)vals.p, (uint32_t*)vals2.p, m,
                                 32 + nb, &ko, &sorted_vals));
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, bc + 31, 4, hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        nwork = (uint32_t)(ctx->hstat[0] & 0xFFFFFFFFu);
    }
    GA_TRY(ga_malloc(ctx, (void**)&P.wcol, std::max<size_t>(nwork, 1) * 4));
    GA_TRY(ga_malloc(ctx, (void**)&P.wedge, std::max<size_t>(nwork, 1) * 4));
    GA_TRY(ga_malloc(ctx, (void**)&P.wps, std::max<size_t>(nwork, 1) * 4));
    GA_TRY(ga_malloc(ctx, (void**)&P.wlen, std::max<size_t>(nwork, 1) * 4));
    P.nwork = nwork;
    if (nwork == 0) {
        GA_TRY(ga_malloc(ctx, (void**)&P.items, 12));
        P.ready = true;
        return GA_OK;
    }
    GA_TRY(ga_malloc(ctx, &hstart_b.p, ((size_t)n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &wh_b.p, (size_t)nwork * 4));
    GA_TRY(ga_malloc(ctx, &wt_b.p, ((size_t)nwork + 1) * 8));
    GA_TRY(ga_malloc(ctx, &wsum_b.p, ((size_t)nwork + 1) * 8));
    GA_TRY(ga_malloc(ctx, &flag_b.p, ((size_t)nwork + 1) * 4));
    GA_TRY(ga_malloc(ctx, &fpos_b.p, ((size_t)nwork + 1) * 4));
    uint32_t* hstart = (uint32_t*)hstart_b.p;
    uint32_t* wh = (uint32_t*)wh_b.p;
    unsigned long long* wt = (unsigned long long*)wt_b.p;
    unsigned long long* wsum = (unsigned long long*)wsum_b.p;
    uint32_t* flag = (uint32_t*)flag_b.p;
    uint32_t* fpos = (uint32_t*)fpos_b.p;
    k_plan_fill<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, g->src, g->col, sorted_vals, nwork, P.wcol, P.wedge, P.wps,
                                                                     P.wlen, wh, wt);
    GA_TRY(ga_launch_check(ctx, "k_plan_fill"));
    k_plan_rowstart<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(wh, nwork, hstart);
    GA_TRY(ga_launch_check(ctx, "k_plan_rowstart"));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)wt + nwork, 1, 0));
    GA_TRY(ga_exclusive_scan_u64(ctx, (const uint64_t*)wt, (uint64_t*)wsum, (size_t)nwork + 1));
    const uint32_t cap = (uint32_t)P.tune.hash_cap;
    k_plan_flags<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(hstart, wh, wsum, g->rowptr, nwork, (uint32_t)P.tune.work_chunk, cap,
                                                                      plan_shape(P.tune), flag);
    GA_TRY(ga_launch_check(ctx, "k_plan_flags"));
    GA_TRY(ga_fill_u32(ctx, flag + nwork, 1, 0));
    GA_TRY(ga_exclusive_scan_u32(ctx, flag, fpos, (size_t)nwork + 1));
    k_plan_item_count<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(flag, wh, g->rowptr, nwork, cap, bc);
    GA_TRY(ga_launch_check(ctx, "k_plan_item_count"));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, bc, 32, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat + 4, fpos + nwork, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t cnt[NHASH + 1];
    const uint32_t* hc = (const uint32_t*)ctx->hstat;
    for (int b = 0; b <= NHASH; b++) cnt[b] = hc[b];
    const uint32_t nall = (uint32_t)(ctx->hstat[4] & 0xFFFFFFFFu);
    uint32_t off[8] = {0};
    for (int b = 0; b <= NHASH; b++) off[b + 1] = off[b] + cnt[b];
    if (off[NHASH + 1] != nall) return ga_fail(ctx, GA_ERR_STATE, "count plan: %u items bucketed, %u expected", off[NHASH + 1], nall);
    GA_TRY(ga_malloc(ctx, &starts_b.p, (size_t)nall * 4));
    GA_TRY(ga_malloc(ctx, (void**)&P.items, (size_t)nall * 12));
    k_plan_rank<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(flag, fpos, nwork, (uint32_t*)starts_b.p);
    GA_TRY(ga_launch_check(ctx, "k_plan_rank"));
    GA_HIP(ctx, hipMemcpyAsync(bc + 16, off, 8 * 4, hipMemcpyHostToDevice, ctx->stream));
    k_plan_items<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(flag, fpos, (uint32_t*)starts_b.p, wh, g->rowptr, nwork, nall, cap,
                                                                      bc + 16, bc + 8, P.items);
    GA_TRY(ga_launch_check(ctx, "k_plan_items"));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < NHASH; b++) P.nitems[b] = cnt[b];
    P.nhuge = cnt[NHASH];
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
static int launch_hash_bucket(ga_ctx* ctx, ga_graph* g, const uint32_t* items, uint32_t nitems, uint32_t bound, TcArgs a) {
    if (!nitems) return GA_OK;
    const TcPlan& P = g->plan;
    uint32_t block = (uint32_t)P.tune.block;
    const uint32_t ts = HASHED ? bucket_slots(std::min(bound, std::max(g->max_out, 1u)), (uint32_t)P.tune.load_shift) : 0;
    const uint32_t per_group = (ts * ((FLAGS & (F_SUP | F_TRI)) ? 2u : 1u) + 4u * G) * 4u;
    while (block > 64 && (size_t)(block / G) * per_group > 64 * 1024) block >>= 1;
    const size_t shmem = (size_t)(block / G) * per_group;
    if (HASHED && shmem > ctx->lds_max)   // the tables do not fit one workgroup: binary-search this bucket in HBM
        return launch_hash_bucket<64, false, FLAGS>(ctx, g, items, nitems, bound, a);
    const unsigned grid = grid_for((uint64_t)nitems * G, block, (unsigned)std::max(ctx->cus, 1) * 64u);
    a.items = items;
    a.nitems = nitems;
    a.ts_max = ts;
    a.load_shift = (uint32_t)P.tune.load_shift;
    k_tc_hash<G, HASHED, FLAGS><<<grid, block, shmem, ctx->stream>>>(a);
    return ga_launch_check(ctx, "k_tc_hash");
}

template <int FLAGS>
static int launch_hash(ga_ctx* ctx, ga_graph* g, TcArgs a) {
    const TcPlan& P = g->plan;
    uint32_t off = 0;
    for (int b = 0; b < NHASH; b++) {
        const uint32_t* items = P.items + 3ull * off;
        const uint32_t ni = P.nitems[b];
        const uint32_t bound = std::min(BUCKET_DEG[b], (uint32_t)P.tune.hash_cap);
        int rc = GA_OK;
        switch (P.tune.lanes[b]) {
            case 4: rc = launch_hash_bucket<4, true, FLAGS>(ctx, g, items, ni, bound, a); break;
            case 8: rc = launch_hash_bucket<8, true, FLAGS>(ctx, g, items, ni, bound, a); break;
            case 16: rc = launch_hash_bucket<16, true, FLAGS>(ctx, g, items, ni, bound, a); break;
            case 32: rc = launch_hash_bucket<32, true, FLAGS>(ctx, g, items, ni, bound, a); break;
            default: rc = launch_hash_bucket<64, true, FLAGS>(ctx, g, items, ni, bound, a); break;
        }
        GA_TRY(rc);
        off += ni;
    }
    return launch_hash_bucket<64, false, FLAGS>(ctx, g, P.items + 3ull * off, P.nhuge, 0, a);
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
        TcArgs a{};
        a.rowptr = g->rowptr;
        a.col = g->col;
        a.wcol = g->plan.wcol;
        a.wedge = g->plan.wedge;
        a.wps = g->plan.wps;
        a.wlen = g->plan.wlen;
        a.alive = alive;
        a.tri = tri;
        a.sup = sup;
        a.total = total;
        a.done = done;
        a.gof = gof;
        a.gcnt = gcnt;
        switch (flags) {
            case 0: return launch_hash<0>(ctx, g, a);
            case F_TRI: return launch_hash<F_TRI>(ctx, g, a);
            case F_ALIVE | F_TRI: return launch_hash<F_ALIVE | F_TRI>(ctx, g, a);
            case F_SUP: return launch_hash<F_SUP>(ctx, g, a);
            case F_GRAPH: return launch_hash<F_GRAPH>(ctx, g, a);
            case F_ALIVE | F_TRI | F_GRAPH: return launch_hash<F_ALIVE | F_TRI | F_GRAPH>(ctx, g, a);
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

extern "C" int ga_tc_plan_info(ga_ctx* ctx, ga_graph* g, uint32_t* nwork, uint32_t* nitems6) {
    if (!ctx || !g) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(ga_tc_ensure_plan(ctx, g));
    if (nwork) *nwork = g->plan.nwork;
    if (nitems6) {
        for (int b = 0; b < NHASH; b++) nitems6[b] = g->plan.nitems[b];
        nitems6[NHASH] = g->plan.nhuge;
    }
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
    bool replan = !g->plan.tune.block || g->plan.tune.work_chunk != t->work_chunk || g->plan.tune.hash_cap != t->hash_cap ||
                  g->plan.tune.batches != t->batches;
    for (int b = 0; b < NHASH; b++) replan = replan || g->plan.tune.lanes[b] != t->lanes[b];
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
    // coordinate search: lanes per bucket, then block, load shift, batches and work chunk
    GA_TRY(trial(base));
    const int lane_opts[5] = {4, 8, 16, 32, 64};
    for (int b = 0; b < NHASH; b++) {
        if (!g->plan.nitems[b]) continue;
        for (int li = 0; li < 5; li++) {
            ga_tc_tuning t = bt;
            t.lanes[b] = lane_opts[li];
            if (t.lanes[b] == bt.lanes[b]) continue;
            if (b >= 3 && t.lanes[b] < 16) continue;   // large tables per few lanes would not fit LDS
            GA_TRY(trial(t));
        }
    }
    const int blocks[3] = {128, 256, 512};
    for (int bi = 0; bi < 3; bi++) { ga_tc_tuning t = bt; t.block = blocks[bi]; if (t.block != bt.block) GA_TRY(trial(t)); }
    for (int ls = 1; ls <= 3; ls++) { ga_tc_tuning t = bt; t.load_shift = ls; if (ls != bt.load_shift) GA_TRY(trial(t)); }
    const int bat[3] = {2, 8, 32};
    for (int bi = 0; bi < 3; bi++) { ga_tc_tuning t = bt; t.batches = bat[bi]; if (t.batches != bt.batches) GA_TRY(trial(t)); }
    const int chunks[3] = {2048, 8192, 32768};
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
