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

enum { F_ALIVE = 1, F_TRI = 2, F_SUP = 4, F_GRAPH = 8, F_SLOTS = 16 };
#ifndef GA_TC_SU
#define GA_TC_SU 8u   // elements per lane slot of the lane-slot probe
#endif   // F_SLOTS: plain / per-graph counts read in lane slots

#ifndef GA_TC_U8_FROM
#define GA_TC_U8_FROM 32u   // lane-group phase length from which each lane keeps 8 loads in flight (else 4)
#endif

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

// LDS hash tables of the lane-group kernel: 2^bbits buckets of 4 slots (16 bytes). A key lives in the first bucket
// of its probe sequence (its hash bucket, then the next ones cyclically) that had a free slot when it was inserted.
// A lookup reads whole buckets and stops at the first one with a free slot, so one 16-byte read decides almost every
// lookup without a loop (at most a quarter of the slots are used at load_shift 2).
__device__ __forceinline__ uint32_t key_bucket(uint32_t x, uint32_t bbits, uint32_t mul = 2654435769u) {
    return bbits ? (x * mul) >> (32u - bbits) : 0u;
}

// Odd multipliers of the lane-slot tables' rebuilds (below): the first is the one every other table uses.
#define HASH_TRIES 4u
__device__ __forceinline__ uint32_t hash_mul(uint32_t attempt) {
    return attempt == 0 ? 2654435769u : attempt == 1 ? 2246822507u : attempt == 2 ? 3266489909u : 668265263u;
}

// Inserts a key that is not in the table; returns its slot. Ends: a table has at least twice as many slots as keys,
// and a failed compare-and-swap means another lane has just filled that slot.
__device__ __forceinline__ uint32_t bucket_insert(uint32_t* tab, uint32_t bbits, uint32_t x, uint32_t mul = 2654435769u) {
    const uint32_t nbm = (1u << bbits) - 1u;
    uint32_t bk = key_bucket(x, bbits, mul);
    for (;;) {
        const uint4 t = ((const uint4*)tab)[bk];
        const uint32_t j = t.x == GA_EMPTY ? 0u : t.y == GA_EMPTY ? 1u : t.z == GA_EMPTY ? 2u : t.w == GA_EMPTY ? 3u : 4u;
        if (j == 4u) {
            bk = (bk + 1u) & nbm;
            continue;
        }
        if (atomicCAS(&tab[4u * bk + j], GA_EMPTY, x) == GA_EMPTY) return 4u * bk + j;
    }
}

// Position of x in a bucket (keys are unique, so at most one slot matches), 4 when absent: arithmetic, no branch.
__device__ __forceinline__ uint32_t bucket_pos(const uint4 t, uint32_t x) {
    const uint32_t m0 = t.x == x ? 1u : 0u, m1 = t.y == x ? 1u : 0u, m2 = t.z == x ? 1u : 0u, m3 = t.w == x ? 1u : 0u;
    return (m0 | m1 | m2 | m3) ? m1 + 2u * m2 + 3u * m3 : 4u;
}

// Slot of x, or GA_EMPTY when x is not in the table. The loop runs only past a full bucket without x.
__device__ __forceinline__ uint32_t bucket_find(const uint32_t* tab, uint32_t bbits, uint32_t x, uint32_t mul = 2654435769u) {
    const uint32_t nbm = (1u << bbits) - 1u;
    uint32_t bk = key_bucket(x, bbits, mul);
    uint4 t = ((const uint4*)tab)[bk];
    uint32_t j = bucket_pos(t, x);
    bool full = max(max(t.x, t.y), max(t.z, t.w)) != GA_EMPTY;
    if (j == 4u && full) {
        for (uint32_t step = 0; step < nbm && j == 4u && full; step++) {
            bk = (bk + 1u) & nbm;
            t = ((const uint4*)tab)[bk];
            j = bucket_pos(t, x);
            full = max(max(t.x, t.y), max(t.z, t.w)) != GA_EMPTY;
        }
    }
    return j < 4u ? 4u * bk + j : GA_EMPTY;
}

// Membership past a full first bucket (rare): walks the buckets after bk until x or a bucket with a free slot.
__device__ __noinline__ uint32_t bucket_has_after(const uint32_t* tab, uint32_t bbits, uint32_t x, uint32_t bk) {
    const uint32_t nbm = (1u << bbits) - 1u;
    for (uint32_t step = 0; step < nbm; step++) {
        bk = (bk + 1u) & nbm;
        const uint4 t = ((const uint4*)tab)[bk];
        if (t.x == x || t.y == x || t.z == x || t.w == x) return 1u;
        if (t.x == GA_EMPTY || t.y == GA_EMPTY || t.z == GA_EMPTY || t.w == GA_EMPTY) return 0u;
    }
    return 0u;
}

// Participation counts of one workgroup gathered in LDS before they go to tri[] in HBM. Device-scope atomics on one
// address are serialised at the memory side, and a power-law graph sends one per item to each of its hubs (a hub is
// in the out-list of most rows): on R-MAT scale 21 those atomics took three quarters of a participation count. Each
// workgroup keeps TRI_CACHE counters (two probe slots per vertex, claimed by the first vertex to arrive) and flushes
// them once at its end; a vertex finding both slots taken by others adds to tri[] directly.
#define TRI_CACHE 512u

// CACHE false (batch unions: many small graphs, few items per hub) adds straight to tri[]: measured faster there
// (385 vs 425 ms for the 4,096-graph batched peel, with 64 workgroups per CU in both).
template <bool CACHE>
__device__ __forceinline__ void tri_add(uint32_t* __restrict__ ckey, uint32_t* __restrict__ ccnt, uint32_t* __restrict__ tri, uint32_t x,
                                        uint32_t c) {
    if (!CACHE) {
        atomicAdd(&tri[x], c);
        return;
    }
    const uint32_t h0 = (x * 2654435769u) >> 23;   // 9 bits: TRI_CACHE slots
#pragma unroll
    for (uint32_t q = 0; q < 2; q++) {
        const uint32_t sl = (h0 + q) & (TRI_CACHE - 1u);
        uint32_t k = ckey[sl];
        if (k == GA_EMPTY) k = atomicCAS(&ckey[sl], GA_EMPTY, x) == GA_EMPTY ? x : ckey[sl];
        if (k == x) {
            atomicAdd(&ccnt[sl], c);
            return;
        }
    }
    atomicAdd(&tri[x], c);
}

// This lane group's bits of a wave ballot, shifted down to bit 0.
template <int G>
__device__ __forceinline__ uint64_t group_bits(uint64_t wave_mask) {
    if (G == 64) return wave_mask;
    const uint32_t goff = lane_id() & ~(uint32_t)(G - 1);
    return (wave_mask >> goff) & ((1ull << G) - 1ull);
}

// Work item record (16 bytes): {h, first work entry, rowptr[h], entries | d+(h) << ITEM_DSHIFT}. Entries per item are
// at most lanes x batches <= 65536; d+(h) is exact up to 32767 (every hashed bucket; the binary-search bucket reads
// rowptr instead).
#define ITEM_DSHIFT 17u
#define ITEM_NMASK 0x1FFFFu

struct TcArgs {
    const uint32_t* items;
    uint32_t nitems;
    uint32_t ts_max;       // table slots reserved per group
    uint32_t load_shift;
    const uint32_t* rowptr;
    const uint32_t* col;
    const uint32_t* wcol;
    const uint32_t* wedge;
    const uint2* wpl;        // {rowptr[p], d+(p)} of each work entry
    const uint8_t* alive;
    uint32_t* tri;
    uint32_t* sup;
    unsigned long long* total;
    const uint32_t* done;
    const uint32_t* gof;
    unsigned long long* gcnt;
    uint32_t gstride;      // per-graph counts: 0 = gcnt[g]; else each XCD (blockIdx mod 8) adds to its own copy,
                           // gcnt[g + gstride * x] (the copies are folded after the launch)
    uint32_t xcd;          // items in contiguous chunks per workgroup, each XCD (blockIdx mod 8) one eighth of them
};

// Items of this workgroup: with a.xcd, the workgroups the hardware places on XCD x (b mod 8 = x) share the x-th
// contiguous eighth of the bucket's items, grid-strided among them, so each XCD's L2 serves one eighth; else the plain
// grid stride over all items.
struct ItemRange {
    uint64_t first, end, step;
};
// bid / nb: this workgroup and the workgroups of the grid that share the items (the whole grid, or in a launch over
// several graphs the workgroups of this graph).
__device__ __forceinline__ ItemRange item_range(const TcArgs& a, uint32_t step, uint32_t offset, uint32_t bid, uint32_t nb) {
    ItemRange r;
    if (a.xcd && nb >= 8u) {   // (a grid of fewer than 8 workgroups would leave eighths without one)
        // XCD x (workgroups b = x, x + 8, ...) takes the x-th eighth of the items, grid-strided over its workgroups
        const uint32_t x = bid % 8u, j = bid / 8u;
        const uint32_t nx = (nb - x + 7u) / 8u;   // workgroups on XCD x
        const uint64_t lo = (uint64_t)a.nitems * x / 8u, hi = (uint64_t)a.nitems * (x + 1u) / 8u;
        r.first = lo + (uint64_t)j * step + offset;
        r.end = hi;
        r.step = (uint64_t)nx * step;
    } else {
        r.first = (uint64_t)bid * step + offset;
        r.end = a.nitems;
        r.step = (uint64_t)nb * step;
    }
    return r;
}

template <int G, bool HASHED, int FLAGS>
__device__ __forceinline__ void tc_hash_body(const TcArgs& a, uint32_t bid, uint32_t nb) {
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
    constexpr bool TC = (FLAGS & F_TRI) && !(FLAGS & F_GRAPH);   // participation through the counter cache
    // plain and per-graph counts read a batch's lists in lane slots of SU elements (below); participation and supports,
    // which need each hit's slot or list, read them by length class
    constexpr bool SLOTS = HASHED && (FLAGS & F_SLOTS) && !(FLAGS & (F_TRI | F_SUP | F_ALIVE));
    constexpr uint32_t SU = GA_TC_SU;
    uint32_t* ckey = lds + (size_t)gpb * (tab_words + 4u * G);
    uint32_t* ccnt = ckey + TRI_CACHE;
    if (TC) {
        for (uint32_t i = threadIdx.x; i < TRI_CACHE; i += blockDim.x) {
            ckey[i] = GA_EMPTY;
            ccnt[i] = 0;
        }
        __syncthreads();
    }
    uint64_t local = 0;
    const ItemRange ir = item_range(a, gpb, gib, bid, nb);
    // 32-bit item index (one register): the step is taken only while it stays below the end, so it cannot wrap
    const uint32_t it_end = (uint32_t)ir.end, it_step = (uint32_t)ir.step;
    for (uint32_t it = (uint32_t)min(ir.first, ir.end); it < it_end; it = it_end - it > it_step ? it + it_step : it_end) {
        const uint4 im = ((const uint4*)a.items)[it];   // {h, first entry, rowptr[h], entries | d+(h) << 17}
        const uint32_t h = im.x, b = im.y, hs = im.z, e = im.y + (im.w & ITEM_NMASK);
        if ((FLAGS & F_ALIVE) && !alive[h]) continue;
        const uint32_t dh = HASHED ? im.w >> ITEM_DSHIFT : rowptr[h + 1] - rowptr[h];
        // batch metadata of the first batch, issued before the table is built (later batches: one batch ahead)
        uint32_t n_p = 0, n_ps = 0, n_len = 0;
        {
            const uint32_t j = b + sub;
            if (j < e) {
                const uint2 pl = a.wpl[j];
                n_ps = pl.x;
                n_len = pl.y;
                if (FLAGS & (F_ALIVE | F_TRI)) n_p = a.wcol[j];
            }
        }
        uint32_t tbits = 0, tsz = 1, mul = hash_mul(0);
        bool spill = true;   // false: every key of N+(h) is in its home bucket, so a lookup reads that bucket only
        if (HASHED) {
            // ceil(log2(dh)) + shift, at least 2, at most the group's reserved slots: the reservation holds the bucket's
            // largest row at load shift 2 (table at most a quarter full), so the largest rows of a bucket get shift 2
            // and the smaller ones the tuned shift
            tbits = (dh > 1 ? 32u - (uint32_t)__clz(dh - 1) : 0u) + a.load_shift;
            if (tbits < 3) tbits = 3;   // at least two buckets (key_bucket's shift stays below 32; ts_max >= 8 slots)
            const uint32_t tcap = 31u - (uint32_t)__clz(a.ts_max);
            if (tbits > tcap) tbits = tcap;
            // The lane-slot probe reads home buckets only (no walk in its loop, which keeps it at 8 waves per SIMD by
            // registers): a table in which a key left its home bucket is rebuilt with the next multiplier and one more
            // table bit (up to the reservation), at most HASH_TRIES tables in all. Whether a key leaves its home bucket
            // does not depend on the insertion order: it happens exactly when more than four keys share a home bucket.
            for (uint32_t attempt = 0;; attempt++) {
                tsz = 1u << tbits;
                for (uint32_t w = sub; w < tsz / 4u; w += G) {   // 16-byte stores (tab and tix are 16-byte aligned)
                    ((uint4*)tab)[w] = make_uint4(GA_EMPTY, GA_EMPTY, GA_EMPTY, GA_EMPTY);
                    if (FLAGS & F_TRI) ((uint4*)tix)[w] = make_uint4(0, 0, 0, 0);
                }
                group_sync();
                uint32_t moved = 0;   // a key that sits past its home bucket (then lookups must walk on)
                for (uint32_t k = sub; k < dh; k += G) {
                    const uint32_t x = col[hs + k];
                    if ((FLAGS & F_ALIVE) && !alive[x]) continue;
                    const uint32_t s = bucket_insert(tab, tbits - 2u, x, mul);
                    if (FLAGS & F_SUP) tix[s] = k;
                    moved |= (s >> 2) != key_bucket(x, tbits - 2u, mul) ? 1u : 0u;
                }
                spill = group_bits<G>(__ballot(moved != 0)) != 0;
                group_sync();
                if (!SLOTS || !spill || attempt + 1u == HASH_TRIES) break;
                mul = hash_mul(attempt + 1u);
                if (tbits < tcap) tbits++;
            }
        }
        uint32_t hcount = 0;
        if (SLOTS && spill) {
            // every table tried left a key outside its home bucket (rare): one lane per entry, each element looked up
            // with the walk past full buckets
            for (uint32_t j = b + sub; j < e; j += G) {
                const uint2 pl = a.wpl[j];
                const uint32_t ps = pl.x, len = pl.y;
                uint32_t c = 0;
                for (uint32_t k = 0; k < len; k++) c += bucket_find(tab, tbits - 2u, col[ps + k], mul) != GA_EMPTY ? 1u : 0u;
                if (!(FLAGS & F_GRAPH)) local += c;   // per-graph counts: the graphs' counters hold the total
                hcount += c;
            }
        }
        // (an item taken by the walk above skips the batches)
        for (uint32_t base = SLOTS && spill ? e : b; base < e; base += G) {
            const uint32_t j = base + sub;
            const uint32_t p = n_p, ps = n_ps;
            uint32_t len = n_len;
            if ((FLAGS & F_ALIVE) && len && !alive[p]) len = 0;
            if constexpr (SLOTS) {
                // Lane slots: list q of the batch gets w_q = ceil(len_q / SU) lanes, lane r of them reading elements
                // r, r + w_q, ... (consecutive lanes on consecutive addresses), so every list of the batch is in flight
                // in the same round: ceil(sum w_q / G) rounds per batch instead of one per length class.
                const uint32_t w = (len + SU - 1u) / SU;
                uint32_t incl = w;   // inclusive prefix of w over the group
#pragma unroll
                for (uint32_t d = 1; d < G; d <<= 1) {
                    const uint32_t t = __shfl_up(incl, d, G);
                    incl += sub >= d ? t : 0u;
                }
                const uint32_t wt = __shfl(incl, G - 1, G);   // lane slots of the batch
                stg[sub] = ps;
                stg[G + sub] = len;
                stg[2 * G + sub] = incl;
                {   // prefetch the next batch while this one is streamed
                    const uint32_t jn = base + G + sub;
                    n_ps = n_len = 0;
                    if (jn < e) {
                        const uint2 pl = a.wpl[jn];
                        n_ps = pl.x;
                        n_len = pl.y;
                    }
                }
                group_sync();
                for (uint32_t s0 = 0; s0 < wt; s0 += G) {
                    const uint32_t s = s0 + sub;
                    uint32_t c = 0;
                    if (s < wt) {
                        // the list of slot s: the first q with incl_q > s (incl never decreases over the group)
                        uint32_t q = 0;
#pragma unroll
                        for (uint32_t step = G / 2; step >= 1; step >>= 1)
                            if (stg[2 * G + q + step - 1u] <= s) q += step;
                        const uint32_t kps = stg[q], klen = stg[G + q];
                        const uint32_t kw = (klen + SU - 1u) / SU;
                        const uint32_t r = s - (stg[2 * G + q] - kw);   // this lane's rank among the list's kw lanes
                        const uint32_t last = klen - 1u;
                        const uint32_t* __restrict__ lp = col + kps;
                        uint32_t x[SU];
#pragma unroll
                        for (uint32_t u = 0; u < SU; u++) x[u] = lp[min(r + u * kw, last)];   // past the end: masked
                        __builtin_amdgcn_sched_barrier(0);
                        // the home bucket decides membership (no key left its own): one bucket read and four compares
                        // per element, the compares going to wave masks (as in the length-class form below)
                        const uint32_t bb = tbits - 2u;
#pragma unroll
                        for (uint32_t u0 = 0; u0 < SU; u0 += 4) {
                            uint4 tb[4];
#pragma unroll
                            for (uint32_t t = 0; t < 4; t++) tb[t] = ((const uint4*)tab)[(x[u0 + t] * mul) >> (32u - bb)];
#pragma unroll
                            for (uint32_t t = 0; t < 4; t++) {
                                const uint32_t xv = x[u0 + t];
                                const uint64_t hm = (__builtin_amdgcn_ballot_w64(tb[t].x == xv) | __builtin_amdgcn_ballot_w64(tb[t].y == xv) |
                                                     __builtin_amdgcn_ballot_w64(tb[t].z == xv) | __builtin_amdgcn_ballot_w64(tb[t].w == xv)) &
                                                    __builtin_amdgcn_ballot_w64(r + (u0 + t) * kw <= last);
                                c += __builtin_amdgcn_inverse_ballot_w64(hm) ? 1u : 0u;
                            }
                        }
                    }
                    if (!(FLAGS & F_GRAPH)) local += c;   // per-graph counts: the graphs' counters hold the total
                    if (FLAGS & F_GRAPH) hcount += c;
                }
                group_sync();   // staging is rewritten by the next batch
                continue;
            }
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
                    const uint2 pl = a.wpl[jn];
                    n_ps = pl.x;
                    n_len = pl.y;
                    if (FLAGS & (F_ALIVE | F_TRI)) n_p = a.wcol[jn];
                }
            }
            group_sync();
            const uint32_t lk = sub < cnt ? stg[G + sub] : 0u;
            uint32_t start = 0;
#pragma unroll
            for (uint32_t S = G; S >= 1; S >>= 1) {
                // entries [start, end) have length in (S/2, S] (for S = G: every length above G/2); each list is read
                // by W = S/U lanes (at least 1), every lane holding up to U loads in flight: 8 for lists longer than
                // GA_TC_U8_FROM / 2 (more lists in flight at once), 4 for shorter ones, S for lists of 1-3 elements
                constexpr uint32_t UMAX = 8;
                const uint32_t U = S >= GA_TC_U8_FROM ? 8u : S >= 4u ? 4u : S;   // lists of 1-3 elements: S slots
                const uint32_t W = S >= U ? S / U : 1u;
                const uint32_t end = (uint32_t)__popcll(group_bits<G>(__ballot(lk > S / 2)));
                for (uint32_t k0 = start; k0 < end; k0 += G / W) {
                    const uint32_t k = k0 + sub / W, r = sub % W;
                    uint32_t c = 0;
                    if (k < end) {
                        const uint32_t kps = stg[k], klen = stg[G + k];
                        for (uint32_t off = r; off < klen; off += U * W) {
                            // every slot loads (a slot past the list reloads its last element, same line, and is
                            // masked): no branch per load or per probe; all loads issue before the first use
                            uint32_t x[UMAX];
                            const uint32_t last = klen - 1u - off;   // slot u is valid while u * W <= last
                            const uint32_t* __restrict__ lp = col + kps + off;
#pragma unroll
                            for (uint32_t u = 0; u < UMAX; u++)
                                if (u < U) x[u] = lp[min(u * W, last)];
                            __builtin_amdgcn_sched_barrier(0);
                            if (HASHED && !(FLAGS & (F_TRI | F_SUP)) && !spill) {
                                // no key left its home bucket: one bucket read and four compares per element. The
                                // compares go to wave masks (v_cmp into SGPRs, or-ed on the scalar unit) and come back
                                // as one per-lane bit: written as an or of four bools, the compiler packs the bits
                                // into a 4-bit vector with about three times the vector instructions.
                                const uint32_t bb = tbits - 2u;
                                uint4 tb[UMAX];
#pragma unroll
                                for (uint32_t u = 0; u < UMAX; u++)
                                    if (u < U) tb[u] = ((const uint4*)tab)[(x[u] * 2654435769u) >> (32u - bb)];
#pragma unroll
                                for (uint32_t u = 0; u < UMAX; u++)
                                    if (u < U) {
                                        const uint32_t xv = x[u];
                                        const uint64_t hm = (__builtin_amdgcn_ballot_w64(tb[u].x == xv) | __builtin_amdgcn_ballot_w64(tb[u].y == xv) |
                                                             __builtin_amdgcn_ballot_w64(tb[u].z == xv) | __builtin_amdgcn_ballot_w64(tb[u].w == xv)) &
                                                            __builtin_amdgcn_ballot_w64(u * W <= last);
                                        c += __builtin_amdgcn_inverse_ballot_w64(hm) ? 1u : 0u;
                                    }
                                continue;
                            }
                            if (HASHED && !(FLAGS & (F_TRI | F_SUP))) {
                                // plain counts need only membership: every bucket read issued first, then vector
                                // tests (x present: min of the four xors is 0; bucket full: max of its slots is not
                                // GA_EMPTY), and the walk past a full bucket only for the lanes that need it
                                const uint32_t bb = tbits - 2u;
                                uint32_t slow = 0;
#pragma unroll
                                for (uint32_t u0 = 0; u0 < UMAX; u0 += 4) {   // four bucket reads in flight at a time
                                    if (u0 >= U) break;
                                    uint4 tb[4];
#pragma unroll
                                    for (uint32_t q = 0; q < 4; q++)
                                        if (u0 + q < U) tb[q] = ((const uint4*)tab)[key_bucket(x[u0 + q], bb)];
#pragma unroll
                                    for (uint32_t q = 0; q < 4; q++)
                                        if (u0 + q < U) {
                                            const uint32_t u = u0 + q, xv = x[u];
                                            const uint32_t d = min(min(tb[q].x ^ xv, tb[q].y ^ xv), min(tb[q].z ^ xv, tb[q].w ^ xv));
                                            const uint32_t mx = max(max(tb[q].x, tb[q].y), max(tb[q].z, tb[q].w));
                                            const uint32_t v = off + u * W < klen ? 1u : 0u;
                                            c += (d == 0u ? 1u : 0u) & v;
                                            slow |= ((d != 0u && mx != GA_EMPTY) ? v : 0u) << u;
                                        }
                                }
                                if (slow) {
#pragma unroll
                                    for (uint32_t u = 0; u < UMAX; u++)
                                        if (u < U && ((slow >> u) & 1u)) c += bucket_has_after(tab, bb, x[u], key_bucket(x[u], bb));
                                }
                                continue;
                            }
#pragma unroll
                            for (uint32_t u = 0; u < UMAX; u++) {
                                if (u >= U) break;
                                const bool valid = off + u * W < klen;
                                const uint32_t addr = kps + off + u * W;
                                bool hit = false;
                                uint32_t hk = 0;
                                if (HASHED) {
                                    const uint32_t s = bucket_find(tab, tbits - 2u, x[u]);
                                    hit = valid && s != GA_EMPTY;
                                    if (hit && (FLAGS & F_SUP)) hk = tix[s];
                                    if (hit && (FLAGS & F_TRI)) atomicAdd(&tix[s], 1u);   // x's hits gather in LDS
                                } else if (valid) {
                                    const uint32_t q = lower_bound_u32(col, hs, hs + dh, x[u]);
                                    hit = q < hs + dh && col[q] == x[u] && (!(FLAGS & F_ALIVE) || alive[x[u]]);
                                    hk = q - hs;
                                }
                                if (hit) {
                                    c++;
                                    if (!HASHED && (FLAGS & F_TRI)) tri_add<TC>(ckey, ccnt, a.tri, x[u], 1u);
                                    if (FLAGS & F_SUP) {
                                        atomicAdd(&a.sup[hs + hk], 1u);
                                        atomicAdd(&a.sup[addr], 1u);
                                    }
                                }
                            }
                        }
                    }
                    if (!(FLAGS & F_GRAPH)) local += c;   // per-graph counts: the graphs' counters hold the total
                    if (FLAGS & (F_TRI | F_GRAPH)) hcount += c;
                    if (FLAGS & (F_TRI | F_SUP)) {   // hits of list k, summed over its W lanes, go to p and edge (h,p)
#pragma unroll
                        for (uint32_t o = W / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, G);
                        if (r == 0 && k < end && c) {
                            if (FLAGS & F_TRI) tri_add<TC>(ckey, ccnt, a.tri, stg[2 * G + k], c);
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
                if (cx) tri_add<TC>(ckey, ccnt, a.tri, tab[s], cx);
            }
        }
        if (FLAGS & (F_TRI | F_GRAPH)) {   // triangles of this item: h is a corner of each, h's graph owns them
#pragma unroll
            for (uint32_t o = G / 2; o > 0; o >>= 1) hcount += __shfl_xor(hcount, o, G);
            if (sub == 0 && hcount) {
                if (FLAGS & F_TRI) tri_add<TC>(ckey, ccnt, a.tri, h, hcount);
                if (FLAGS & F_GRAPH) atomicAdd(&a.gcnt[a.gof[h] + a.gstride * (blockIdx.x & 7u)], (unsigned long long)hcount);
            }
        }
        group_sync();   // every lane is done with the table before the next item clears it
    }
    if (TC) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < TRI_CACHE; i += blockDim.x)
            if (ccnt[i]) atomicAdd(&a.tri[ckey[i]], ccnt[i]);
    }
    if (FLAGS & F_GRAPH) return;
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(a.total, (unsigned long long)local);
}

// Rows of d+(h) <= 8 (bucket 0), plain and per-graph counts, with lanes[0] = 1: one lane per item. N+(h) sits in 8
// registers, padded with GA_EMPTY (no vertex id equals it), and every probe element is compared with all of them: no
// table, no lane-group synchronisation, 64 items in flight per wavefront. Each entry's metadata is fetched one entry
// ahead. The item loop is wavefront-uniform so that per-graph counts can be summed per wavefront.
template <int FLAGS>
__device__ __forceinline__ void tc_tiny_body(const TcArgs& a, uint32_t bid, uint32_t nb) {
    static_assert(FLAGS == 0 || FLAGS == F_GRAPH, "plain and per-graph counts only");
    if (a.done && *a.done) return;
    const uint32_t* __restrict__ col = a.col;
    const uint32_t lane = lane_id();
    uint64_t local = 0;
    const ItemRange ir = item_range(a, blockDim.x, threadIdx.x & ~63u, bid, nb);
    for (uint64_t base = ir.first; base < ir.end; base += ir.step) {   // wavefront-uniform
        const uint64_t it = base + lane;
        uint32_t c = 0, gr = 0;
        if (it < ir.end) {
            const uint4 im = ((const uint4*)a.items)[it];   // {h, first entry, rowptr[h], entries | d+(h) << 17}
            const uint32_t h = im.x, b = im.y, hs = im.z, e = im.y + (im.w & ITEM_NMASK), dh = im.w >> ITEM_DSHIFT;   // 1..8
            uint2 npl = a.wpl[b];   // every item has an entry; every list has 1..d+(h) elements
            uint32_t hr[8];
#pragma unroll
            for (uint32_t k = 0; k < 8; k++) {
                const uint32_t v = col[hs + min(k, dh - 1u)];
                hr[k] = k < dh ? v : GA_EMPTY;
            }
            if (FLAGS & F_GRAPH) gr = a.gof[h];
            for (uint32_t j = b; j < e; j++) {
                const uint32_t ps = npl.x, len = npl.y;
                if (j + 1u < e) npl = a.wpl[j + 1u];
                uint32_t x[8];
#pragma unroll
                for (uint32_t t = 0; t < 8; t++) x[t] = col[ps + min(t, len - 1u)];
#pragma unroll
                for (uint32_t t = 0; t < 8; t++) {
                    const bool m = (x[t] == hr[0]) | (x[t] == hr[1]) | (x[t] == hr[2]) | (x[t] == hr[3]) | (x[t] == hr[4]) |
                                   (x[t] == hr[5]) | (x[t] == hr[6]) | (x[t] == hr[7]);
                    c += (m && t < len) ? 1u : 0u;
                }
            }
        }
        local += c;
        if (FLAGS & F_GRAPH) {   // one atomic per graph present in the wavefront (items are graph-major: usually one)
            uint64_t todo = __ballot(c != 0);
            while (todo) {
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
                const uint32_t gl = __shfl(gr, (int)leader, 64);
                const bool mine = c != 0 && gr == gl;
                const uint64_t same = __ballot(mine);
                const uint64_t s = wave_sum_u64(mine ? c : 0u);
                if (lane == leader) atomicAdd(&a.gcnt[gl + a.gstride * (blockIdx.x & 7u)], (unsigned long long)s);
                todo &= ~same;
            }
        }
    }
    if (FLAGS & F_GRAPH) return;   // per-graph counts: the graphs' counters hold the total
    local = wave_sum_u64(local);
    if (lane == 0 && local) atomicAdd(a.total, (unsigned long long)local);
}

template <int G, bool HASHED, int FLAGS>
__global__ void k_tc_hash(const TcArgs a) {
    tc_hash_body<G, HASHED, FLAGS>(a, blockIdx.x, gridDim.x);
}

template <int FLAGS>
__global__ void k_tc_tiny(const TcArgs a) {
    tc_tiny_body<FLAGS>(a, blockIdx.x, gridDim.x);
}

// Counts of several resident graphs in one launch per bucket shape: job i holds graph i's arguments for the bucket
// (its items, tables and its own count word) and a range [base, base + nblk) of the launch's workgroups; map[b] is the
// job of workgroup b. Each graph's workgroups run exactly what its own launch would run on a grid of nblk.
struct TcJob {
    TcArgs a;
    uint32_t base, nblk;
};

template <int G, bool HASHED, int FLAGS>
__global__ void k_tc_hash_many(const TcJob* __restrict__ jobs, const uint32_t* __restrict__ map) {
    const TcJob& J = jobs[map[blockIdx.x]];
    tc_hash_body<G, HASHED, FLAGS>(J.a, blockIdx.x - J.base, J.nblk);
}

template <int FLAGS>
__global__ void k_tc_tiny_many(const TcJob* __restrict__ jobs, const uint32_t* __restrict__ map) {
    const TcJob& J = jobs[map[blockIdx.x]];
    tc_tiny_body<FLAGS>(J.a, blockIdx.x - J.base, J.nblk);
}

// ---------------------------------------------------------------------------------------------------------------
// flat kernel for light rows (d+(h) <= FLAT_MAX), plain counts
// ---------------------------------------------------------------------------------------------------------------
// The light work entries (plan order: rows of h, h ascending) are laid out on one line of "units":
//   row of h:   [2^flat_bits(d+(h)) table units]  then for each of its entries j: [1 header unit][wlen[j] element units]
// The line is cut into chunks of UPT x blockDim units; one workgroup takes a chunk, so every workgroup gets the same
// work however the probe lists are distributed, and thread t handles units t, t + blockDim, ... (consecutive threads
// on consecutive units, so element loads of one list are coalesced):
//   1. mark where each overlapping entry starts and prefix-sum the marks in LDS: unit -> entry
//   2. every thread reads its units' entry records (16 bytes, shared by a list's units, so mostly cache hits) and
//      issues all its loads at once: table units load N+(h), element units load their probe element
//   3. table units insert N+(h) into an LDS table that sits exactly on those units (the row whose table began in an
//      earlier chunk is rebuilt in a spare region after the chunk)
//   4. element units probe their row's table.
// Entry record (frec): {first unit, first entry of the row ? rowptr[h] : units back to the row's table, rowptr[p],
// bits | d+(h) << 8 | first << 16}. LDS: chunk + FLAT_TMAX table slots and chunk u16 owner indices.
#define FLAT_MAX 128u
#define FLAT_SHIFT 2u
#define FLAT_TMAX (FLAT_MAX << FLAT_SHIFT)
#define FLAT_UNIT_LIMIT 0xFFF00000ull   // unit offsets are 32-bit

__host__ __device__ __forceinline__ uint32_t flat_bits(uint32_t dh) {
    uint32_t lg = 0;
    while ((1u << lg) < dh) lg++;
    const uint32_t b = lg + FLAT_SHIFT;
    return b < 2u ? 2u : b;
}

struct FlatArgs {
    const uint32_t* rowptr;
    const uint32_t* col;
    const uint32_t* wh;
    const uint4* frec;
    const uint32_t* cstart;
    const uint32_t* chs;     // rowptr[h] of the row holding each chunk's first unit
    uint32_t nch, nlight, utot;
    unsigned long long* total;
    const uint32_t* done;
    const uint32_t* gof;           // GRAPH: graph of each vertex and per-graph counts (a disjoint union of graphs)
    unsigned long long* gcnt;
    uint32_t gstride;              // as TcArgs::gstride
};

static size_t flat_shmem(uint32_t chunk) {
    return (size_t)(chunk + FLAT_TMAX) * 4 + (size_t)chunk * 2 + (size_t)(chunk / 2 + 2) * 16;   // table, owners, records
}

__device__ __forceinline__ void flat_insert(uint32_t* tab, uint32_t tbits, uint32_t x) {
    const uint32_t tsz = 1u << tbits;
    uint32_t s = hash_slot(x, tbits);
    for (uint32_t probe = 0; probe < tsz; probe++) {
        if (atomicCAS(&tab[s], GA_EMPTY, x) == GA_EMPTY) return;
        s = (s + 1) & (tsz - 1);
    }
}

template <int UPT, bool GRAPH>
__global__ void __launch_bounds__(1024) k_tc_flat(const FlatArgs a) {
    static_assert(UPT == 4 || UPT == 8 || UPT == 16, "4, 8 or 16 units per thread");
    constexpr int PF = 2;   // entry records per thread fetched one chunk ahead
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t wtot[16];
    if (a.done && *a.done) return;
    const uint32_t T = blockDim.x, CH = UPT * T, tid = threadIdx.x;
    uint32_t* tab = lds;                                // CH + FLAT_TMAX slots ([CH, CH + FLAT_TMAX) = carried row)
    uint16_t* own = (uint16_t*)(tab + CH + FLAT_TMAX);  // unit -> entry (16-byte aligned: CH + FLAT_TMAX is a multiple of 4)
    uint4* own4 = (uint4*)own;                          // 8 owner slots per uint4
    uint4* tab4 = (uint4*)tab;
    uint4* erec = (uint4*)(own + CH);                   // the chunk's entry records (at most CH / 2 + 2)
    const uint32_t lane = lane_id(), wv = tid / 64;
    uint64_t local = 0;
    // next chunk's entry range, carried-row start and first PF records per thread, loaded a chunk ahead
    uint32_t c = blockIdx.x, nj0 = 0, nne = 0, nhs = 0;
    uint4 pre[PF];
    auto fetch = [&](uint32_t cc) {
        nj0 = a.cstart[cc];
        nne = min(a.cstart[cc + 1] + 1, a.nlight) - nj0;
        nhs = a.chs[cc];
#pragma unroll
        for (int q = 0; q < PF; q++) {
            const uint32_t i = tid + q * T;
            pre[q] = i < nne ? a.frec[nj0 + i] : make_uint4(0, 0, 0, 0);
        }
    };
    if (c < a.nch) fetch(c);
    for (; c < a.nch; c += gridDim.x) {
        const uint32_t u0 = c * CH;
        const uint32_t clen = min(CH, a.utot - u0);
        const uint32_t j0 = nj0, ne = nne, hs0 = nhs;
        for (uint32_t w = tid; w < (CH + FLAT_TMAX) / 4; w += T) tab4[w] = make_uint4(GA_EMPTY, GA_EMPTY, GA_EMPTY, GA_EMPTY);
        for (uint32_t w = tid; w < CH / 8; w += T) own4[w] = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < PF; q++)
            if (tid + q * T < ne) erec[tid + q * T] = pre[q];
        for (uint32_t i = tid + PF * T; i < ne; i += T) erec[i] = a.frec[j0 + i];   // chunks with many short entries
        __syncthreads();
        // the row of entry j0: carried when its table begins before this chunk
        const uint4 r0 = erec[0];
        const bool first0 = (r0.w >> 16) & 1u;
        const uint32_t tu0 = first0 ? r0.x : r0.x - r0.y;
        const bool carried = tu0 < u0;
        for (uint32_t i = tid + 1; i < ne; i += T) {
            const uint32_t U = erec[i].x - u0;   // > 0: entry j0 holds unit u0
            if (U < clen) own[U] = 1;
        }
        __syncthreads();
        {   // inclusive prefix sum of the marks, UPT contiguous units per thread read as 16-byte words
            uint32_t m[UPT];
#pragma unroll
            for (int q = 0; q < UPT / 8 + (UPT == 4 ? 1 : 0); q++) {
                if (UPT == 4) {
                    const uint2 v = ((const uint2*)own)[tid];
                    m[0] = v.x & 0xFFFFu; m[1] = v.x >> 16; m[2] = v.y & 0xFFFFu; m[3] = v.y >> 16;
                } else {
                    const uint4 v = own4[tid * (UPT / 8) + q];
                    m[8 * q + 0] = v.x & 0xFFFFu; m[8 * q + 1] = v.x >> 16; m[8 * q + 2] = v.y & 0xFFFFu; m[8 * q + 3] = v.y >> 16;
                    m[8 * q + 4] = v.z & 0xFFFFu; m[8 * q + 5] = v.z >> 16; m[8 * q + 6] = v.w & 0xFFFFu; m[8 * q + 7] = v.w >> 16;
                }
            }
            uint32_t run = 0;
#pragma unroll
            for (int k = 0; k < UPT; k++) run += m[k];
            uint32_t inc = run;
#pragma unroll
            for (uint32_t o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(inc, o, 64);
                if (lane >= o) inc += t;
            }
            if (lane == 63) wtot[wv] = inc;
            __syncthreads();
            uint32_t acc = inc - run;
            for (uint32_t w = 0; w < wv; w++) acc += wtot[w];
#pragma unroll
            for (int k = 0; k < UPT; k++) { acc += m[k]; m[k] = acc; }
#pragma unroll
            for (int q = 0; q < UPT / 8 + (UPT == 4 ? 1 : 0); q++) {
                if (UPT == 4) {
                    ((uint2*)own)[tid] = make_uint2(m[0] | (m[1] << 16), m[2] | (m[3] << 16));
                } else {
                    own4[tid * (UPT / 8) + q] = make_uint4(m[8 * q] | (m[8 * q + 1] << 16), m[8 * q + 2] | (m[8 * q + 3] << 16),
                                                           m[8 * q + 4] | (m[8 * q + 5] << 16), m[8 * q + 6] | (m[8 * q + 7] << 16));
                }
            }
        }
        __syncthreads();
        // classify this thread's units: owner indices and entry records from LDS, then one element load per unit (col[0]
        // for units that neither insert nor probe), all loads of a phase in flight together, selects instead of branches
        uint32_t x[UPT], meta[UPT];   // meta: bit 31 insert, bit 30 probe | bits << 16 | table base
        {
            uint32_t ix[UPT];
#pragma unroll
            for (int k = 0; k < UPT; k++) {
                const uint32_t u = k * T + tid;
                ix[k] = own[u < clen ? u : 0u];
            }
            uint4 r[UPT];
#pragma unroll
            for (int k = 0; k < UPT; k++) r[k] = erec[ix[k]];
            uint32_t addr[UPT];
#pragma unroll
            for (int k = 0; k < UPT; k++) {
                const uint32_t u = k * T + tid;
                const bool valid = u < clen;
                const uint32_t tb = r[k].w & 31u, dh = (r[k].w >> 8) & 0xFFu;
                const bool first = (r[k].w >> 16) & 1u;
                const uint32_t U = r[k].x - u0;   // unsigned: wraps for the entry that began before the chunk
                const uint32_t q = u - U;
                const uint32_t t = first ? (1u << tb) : 0u;
                const bool ins = valid & (q < t) & (q < dh) & (t <= clen) & (U <= clen - t);   // whole table in the chunk
                const bool prb = valid & (q > t);
                const uint32_t tu = first ? r[k].x : r[k].x - r[k].y;
                // table base: the insert target, the probed row's table (here or the carried region), 0 otherwise
                const uint32_t base = ins ? U : (prb ? (tu >= u0 ? tu - u0 : CH) : 0u);
                addr[k] = ins ? r[k].y + q : (prb ? r[k].z + (q - t - 1) : 0u);
                meta[k] = (ins ? 0x80000000u : 0u) | (prb ? 0x40000000u : 0u) | (tb << 16) | (base & 0xFFFFu);
            }
#pragma unroll
            for (int k = 0; k < UPT; k++) x[k] = a.col[addr[k]];
        }
        if (carried) {
            const uint32_t tb = r0.w & 31u, dh = (r0.w >> 8) & 0xFFu;
            for (uint32_t k = tid; k < dh; k += T) flat_insert(tab + CH, tb, a.col[hs0 + k]);
        }
#pragma unroll
        for (int k = 0; k < UPT; k++)
            if (meta[k] & 0x80000000u) flat_insert(tab + (meta[k] & 0xFFFFu), (meta[k] >> 16) & 31u, x[k]);
        if (c + gridDim.x < a.nch) fetch(c + gridDim.x);   // lands while this chunk probes
        // GRAPH: light entries are in vertex order and each graph holds a contiguous id range, so a chunk lies in one
        // graph except where a graph ends inside it; those chunks credit each hit to its entry's graph
        uint32_t g0 = 0, cc = 0;
        bool split = false;
        if (GRAPH) {
            g0 = a.gof[a.wh[j0]];
            split = a.gof[a.wh[j0 + ne - 1]] != g0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < UPT; k++) {
            const uint32_t tb = (meta[k] >> 16) & 31u, tm = (1u << tb) - 1u;
            const uint32_t* tt = tab + (meta[k] & 0xFFFFu);
            const uint32_t s0 = hash_slot(x[k], tb);
            // two slots unrolled (tables are at most a quarter full), the rare longer chain in a loop
            const uint32_t v0 = (meta[k] & 0x40000000u) ? tt[s0] : GA_EMPTY;
            bool hit = v0 == x[k] && v0 != GA_EMPTY;
            if (v0 != GA_EMPTY && !hit) {
                uint32_t s = (s0 + 1) & tm;
                for (uint32_t probe = 1; probe <= tm; probe++) {
                    const uint32_t v = tt[s];
                    if (v == x[k]) { hit = true; break; }
                    if (v == GA_EMPTY) break;
                    s = (s + 1) & tm;
                }
            }
            local += hit ? 1u : 0u;
            if (GRAPH) {
                if (!split) cc += hit ? 1u : 0u;
                else if (hit) atomicAdd(&a.gcnt[a.gof[a.wh[j0 + own[k * T + tid]]] + a.gstride * (blockIdx.x & 7u)], 1ull);
            }
        }
        if (GRAPH) {   // one atomic per chunk: wave sums in wtot (free since the prefix sum), thread 0 adds them
            cc = (uint32_t)wave_sum_u64(cc);
            if (lane == 0) wtot[wv] = cc;
        }
        __syncthreads();
        if (GRAPH && tid == 0 && !split) {   // wtot is rewritten only after the next chunk's first two barriers
            unsigned long long s = 0;
            for (uint32_t w = 0; w < T / 64; w++) s += wtot[w];
            if (s) atomicAdd(&a.gcnt[g0 + a.gstride * (blockIdx.x & 7u)], s);
        }
    }
    if (GRAPH) return;   // per-graph counts: the graphs' counters hold the total
    local = wave_sum_u64(local);
    if (lane == 0 && local) atomicAdd(a.total, (unsigned long long)local);
}

// Per-graph counts kept per XCD (gcnt[g + n * x], x < 8): gcnt[g] becomes the sum of graph g's eight copies.
__global__ void k_fold_xcd(unsigned long long* __restrict__ gcnt, uint32_t n) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
        unsigned long long s = 0;
#pragma unroll
        for (uint32_t x = 0; x < 8; x++) s += gcnt[g + (size_t)n * x];
        gcnt[g] = s;
    }
}

int ga_tc_fold_xcd(ga_ctx* ctx, unsigned long long* gcnt, uint32_t n) {
    if (!n) return GA_OK;
    k_fold_xcd<<<grid_for(n, 256, 1024), 256, 0, ctx->stream>>>(gcnt, n);
    return ga_launch_check(ctx, "k_fold_xcd");
}

// out += the sum of in[0, n) (out zeroed by the caller).
__global__ void k_sum_u64(const unsigned long long* __restrict__ in, uint32_t n, unsigned long long* __restrict__ out) {
    uint64_t local = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) local += in[i];
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(out, (unsigned long long)local);
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

// key = (light << (32 + nb)) | (h << 32) | ~d+(p): light rows (d+(h) <= FLAT_MAX) first, then rows of h, longest
// probe list first; edges without work get the all-ones key (last). light = 0 for every row when split is false.
__global__ void k_plan_keys(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                            uint32_t m, int nb, bool split, unsigned long long* __restrict__ key, uint32_t* __restrict__ val,
                            uint32_t* __restrict__ nwork) {
    uint32_t cnt = 0, light = 0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        uint32_t h, p, dmin;
        hash_side(rowptr, src[e], col[e], h, p, dmin);
        const bool heavy = split && rowptr[h + 1] - rowptr[h] > FLAT_MAX;
        key[e] = dmin ? (((unsigned long long)heavy << (32 + nb)) | ((unsigned long long)h << 32) | (uint32_t)~dmin) : ~0ull;
        val[e] = e;
        cnt += dmin != 0;
        light += dmin != 0 && !heavy;
    }
    cnt = (uint32_t)wave_sum_u64(cnt);
    light = (uint32_t)wave_sum_u64(light);
    if (lane_id() == 0 && cnt) atomicAdd(nwork, cnt);
    if (lane_id() == 0 && light) atomicAdd(nwork + 1, light);
}

// Flat units of each light entry: its row's table (first entry of the row only), one header unit, its probe list.
__global__ void k_plan_units(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ wh, const uint2* __restrict__ wpl,
                             uint32_t nlight, unsigned long long* __restrict__ wt) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nlight; j += gridDim.x * blockDim.x) {
        const uint32_t h = wh[j];
        const bool first = j == 0 || wh[j - 1] != h;
        wt[j] = (first ? (1ull << flat_bits(rowptr[h + 1] - rowptr[h])) : 0ull) + 1ull + wpl[j].y;
    }
}

// Entry records of the flat kernel (see k_tc_flat).
__global__ void k_plan_frec(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ st, const uint32_t* __restrict__ wh,
                            const uint2* __restrict__ wpl, const unsigned long long* __restrict__ wu, const uint32_t* __restrict__ hstart,
                            uint32_t nlight, uint4* __restrict__ frec) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nlight; j += gridDim.x * blockDim.x) {
        const uint32_t h = wh[j];
        const uint32_t jf = hstart[h];
        const bool first = jf == j;
        const uint32_t dh = rowptr[h + 1] - rowptr[h];
        frec[j] = make_uint4((uint32_t)wu[j], first ? st[h] : (uint32_t)(wu[j] - wu[jf]), wpl[j].x,
                             flat_bits(dh) | (dh << 8) | ((first ? 1u : 0u) << 16));
    }
}

// cstart[c] = the light entry whose units contain unit c * chunk; chs[c] = the row start of that entry's row.
__global__ void k_plan_chunks(const unsigned long long* __restrict__ wu, const uint32_t* __restrict__ wh, const uint32_t* __restrict__ st,
                              uint32_t nlight, uint32_t chunk, uint32_t* __restrict__ cstart, uint32_t* __restrict__ chs) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nlight; j += gridDim.x * blockDim.x) {
        const unsigned long long a = wu[j], b = wu[j + 1];
        for (unsigned long long c = (a + chunk - 1) / chunk; c * chunk < b; c++) {
            cstart[c] = j;
            chs[c] = st[wh[j]];
        }
    }
}

__global__ void k_plan_fill(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                            const uint32_t* __restrict__ val, uint32_t nwork, uint32_t* __restrict__ wcol, uint32_t* __restrict__ wedge,
                            uint2* __restrict__ wpl, uint32_t* __restrict__ wh,
                            unsigned long long* __restrict__ wt) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x) {
        const uint32_t e = val[j];
        uint32_t h, p, dmin;
        hash_side(rowptr, src[e], col[e], h, p, dmin);
        wcol[j] = p;
        wedge[j] = e;
        wpl[j] = make_uint2(rowptr[p], dmin);
        wh[j] = h;
        wt[j] = dmin;
    }
}

// Probe work of each entry, from the kept entries (what k_plan_fill writes into wt).
__global__ void k_plan_work(const uint2* __restrict__ wpl, uint32_t nwork, unsigned long long* __restrict__ wt) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x) wt[j] = wpl[j].y;
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

// Item starts -> per-(bucket, block) counts, bcount[bucket * gridDim.x + block]; block b covers entries
// [b * span, (b + 1) * span) (LDS counters: no global atomics).
__global__ void k_plan_item_count(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ wh, const uint32_t* __restrict__ rowptr,
                                  uint32_t nwork, uint32_t span, uint32_t cap, uint32_t* __restrict__ bcount) {
    __shared__ uint32_t cnt[NHASH + 1];
    if (threadIdx.x <= NHASH) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t j0 = (uint64_t)blockIdx.x * span, j1 = min((uint64_t)nwork, j0 + span);
    for (uint64_t base = j0; base < j1; base += blockDim.x) {   // block-uniform loop
        const uint64_t j = base + threadIdx.x;
        uint32_t bk = 0xFFu;
        if (j < j1 && flag[j]) {
            const uint32_t h = wh[j];
            bk = bucket_of(rowptr[h + 1] - rowptr[h], cap);
        }
#pragma unroll
        for (uint32_t b = 0; b <= NHASH; b++) {
            const uint64_t mk = __ballot(bk == b);
            if (mk && lane_id() == (uint32_t)__ffsll((unsigned long long)mk) - 1) atomicAdd(&cnt[b], (uint32_t)__popcll(mk));
        }
    }
    __syncthreads();
    if (threadIdx.x <= NHASH) bcount[threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

// Items placed bucket-major: (h, begin, end), end = next item start (rows are contiguous and every row starts an
// item). Each block writes its range's items in order from its offset boff[bucket * gridDim.x + block].
__global__ void k_plan_items(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ fpos, const uint32_t* __restrict__ starts_by_rank,
                             const uint32_t* __restrict__ wh, const uint32_t* __restrict__ rowptr, uint32_t nwork, uint32_t nall,
                             uint32_t span, uint32_t cap, const uint32_t* __restrict__ boff, uint32_t* __restrict__ items) {
    __shared__ uint32_t run[NHASH + 1];
    __shared__ uint32_t wc[16][NHASH + 1];
    const uint32_t wv = threadIdx.x / 64, nw = blockDim.x / 64;
    if (threadIdx.x <= NHASH) run[threadIdx.x] = boff[threadIdx.x * gridDim.x + blockIdx.x];
    __syncthreads();
    const uint64_t j0 = (uint64_t)blockIdx.x * span, j1 = min((uint64_t)nwork, j0 + span);
    for (uint64_t base = j0; base < j1; base += blockDim.x) {   // block-uniform loop
        const uint64_t j = base + threadIdx.x;
        uint32_t bk = 0xFFu, h = 0, rank = 0;
        if (j < j1 && flag[j]) {
            h = wh[j];
            bk = bucket_of(rowptr[h + 1] - rowptr[h], cap);
        }
#pragma unroll
        for (uint32_t b = 0; b <= NHASH; b++) {
            const uint64_t mk = __ballot(bk == b);
            if (bk == b) rank = mask_rank(mk);
            if (lane_id() == 0) wc[wv][b] = (uint32_t)__popcll(mk);
        }
        __syncthreads();
        if (bk <= NHASH) {
            uint32_t pos = run[bk] + rank;
            for (uint32_t w = 0; w < wv; w++) pos += wc[w][bk];
            const uint32_t r = fpos[j];
            const uint32_t e = r + 1 < nall ? starts_by_rank[r + 1] : nwork;
            const uint32_t hs = rowptr[h], dh = rowptr[h + 1] - hs;
            ((uint4*)items)[pos] = make_uint4(h, (uint32_t)j, hs, (e - (uint32_t)j) | (min(dh, 32767u) << ITEM_DSHIFT));
        }
        __syncthreads();
        if (threadIdx.x <= NHASH) {
            uint32_t add = 0;
            for (uint32_t w = 0; w < nw; w++) add += wc[w][threadIdx.x];
            run[threadIdx.x] += add;
        }
        __syncthreads();
    }
}

// Shape key of each item: bucket (kept major) | locality group of h (a union's graph id, or h's window of 2^wshift ids
// for the windowed order; none otherwise) |
// entries | longest probe list (the item's first) | d+(h), the last three capped at 10 bits. Consecutive items go to
// the lane groups of one wavefront; equal shapes give those groups loops of equal trip counts (groups of one
// wavefront otherwise wait, masked off, for the longest of their loops). Group-major order keeps the items in flight
// on a few graphs, whose lists then stay in cache.
__global__ void k_plan_item_shape(const uint32_t* __restrict__ items, uint32_t nall, const uint32_t* __restrict__ rowptr,
                                  const uint2* __restrict__ wpl, uint32_t cap, const uint32_t* __restrict__ group, uint32_t gbits,
                                  uint32_t wshift, unsigned long long* __restrict__ key, uint32_t* __restrict__ val) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nall; i += gridDim.x * blockDim.x) {
        const uint4 im = ((const uint4*)items)[i];
        const uint32_t h = im.x, b = im.y, e = im.y + (im.w & ITEM_NMASK);
        const uint32_t dh = rowptr[h + 1] - rowptr[h];
        const unsigned long long bk = bucket_of(dh, cap);
        const unsigned long long gr = group ? group[h] : wshift ? (h >> wshift) : 0u;   // locality group, or h's id window
        key[i] = (bk << (30 + gbits)) | (gr << 30) | ((unsigned long long)min(e - b, 1023u) << 20) |
                 ((unsigned long long)min(wpl[b].y, 1023u) << 10) | min(dh, 1023u);
        val[i] = i;
    }
}

__global__ void k_plan_item_permute(const uint32_t* __restrict__ items, const uint32_t* __restrict__ order, uint32_t nall,
                                    uint32_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nall; i += gridDim.x * blockDim.x) {
        ((uint4*)out)[i] = ((const uint4*)items)[order[i]];
    }
}

__global__ void k_plan_rank(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ fpos, uint32_t nwork, uint32_t* __restrict__ starts) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x)
        if (flag[j]) starts[fpos[j]] = j;
}

// Row starts of the column array the counts read (the line-aligned copy, or back to the plain one): items[].z and
// wpl[].x from st, by the item's h and the entry's probe-side vertex.
__global__ void k_plan_starts(const uint32_t* __restrict__ st, const uint32_t* __restrict__ wcol, uint32_t nwork, uint32_t nall,
                              uint32_t* __restrict__ items, uint2* __restrict__ wpl) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwork; j += gridDim.x * blockDim.x) wpl[j].x = st[wcol[j]];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nall; i += gridDim.x * blockDim.x) items[4 * (size_t)i + 2] = st[items[4 * (size_t)i]];
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

ga_tc_tuning default_tuning() {
    ga_tc_tuning t;
    t.lanes[0] = 1;   // one lane per item for d+(h) <= 8 (4-lane groups in participation and support launches)
    t.lanes[1] = 16;
    t.lanes[2] = 32;
    t.lanes[3] = 64;
    t.lanes[4] = 64;
    t.block = 128;        // block, work chunk and load shift: what the tuner settles on for the 34.7M-edge graph
    t.work_chunk = 2048;
    t.batches = 8;
    t.load_shift = 3;
    t.hash_cap = 4096;
    t.flat_chunk = 0;     // the flat kernel is a tuner option (it wins on some degree mixes, not on all)
    t.flat_block = 512;
    t.item_order = 1;
    t.relabel = 0;
    t.probe = 1;
    t.waves = 0;
    t.streams = 0;
    return t;
}

// The parameters the plan's arrays depend on (the others only change launches).
bool same_plan_params(const ga_tc_tuning& a, const ga_tc_tuning& b) {
    bool same = a.work_chunk == b.work_chunk && a.hash_cap == b.hash_cap && a.batches == b.batches && a.item_order == b.item_order;
    for (int i = 0; i < NHASH; i++) same = same && a.lanes[i] == b.lanes[i];
    return same;
}

bool tuning_valid(const ga_tc_tuning& t) {
    for (int i = 0; i < NHASH; i++)
        if (t.lanes[i] != 4 && t.lanes[i] != 8 && t.lanes[i] != 16 && t.lanes[i] != 32 && t.lanes[i] != 64 && !(i == 0 && t.lanes[i] == 1))
            return false;
    return t.block >= 64 && t.block <= 1024 && t.block % 64 == 0 && t.work_chunk >= 64 && t.batches >= 1 && t.batches <= 1024 &&
           t.load_shift >= 1 && t.load_shift <= 3 && t.hash_cap >= 1 && t.hash_cap <= 4096 && t.flat_block >= 64 &&
           t.flat_block <= 1024 && t.flat_block % 64 == 0 && t.item_order >= 0 && t.item_order <= 5 && t.relabel >= 0 && t.relabel <= 2 && (t.probe == 0 || t.probe == 1) &&
           (t.waves == 0 || (t.waves >= 4 && t.waves <= 8)) && (t.streams == 0 || t.streams == 1) &&
           (t.flat_chunk == 0 || (t.flat_chunk <= 16384 && t.flat_chunk % t.flat_block == 0 &&
                                  (t.flat_chunk / t.flat_block == 4 || t.flat_chunk / t.flat_block == 8 ||
                                   t.flat_chunk / t.flat_block == 16)));
}

PlanShape plan_shape(const ga_tc_tuning& t) {
    PlanShape s;
    for (int b = 0; b < NHASH; b++) s.pmax[b] = (uint32_t)(t.lanes[b] * t.batches);
    s.pmax[NHASH] = (uint32_t)(64 * t.batches);
    return s;
}

}  // namespace

// The CSR and the plan one launch reads: the graph's own (participation, support, alive masks, and counts without
// relabel) or the renumbered copy of the counts. The plan is built from rowptr/col/src; the count kernels read kcol,
// whose rows start at st (the same arrays, or the line-aligned copy of relabel 2).
struct CountView {
    TcPlan* P;
    const uint32_t *rowptr, *col, *src;
    const uint32_t *kcol, *st;
};
static CountView own_view(ga_graph* g) { return CountView{&g->plan, g->rowptr, g->col, g->src, g->col, g->rowptr}; }
static bool relabel_on(const ga_graph* g) { return g->plan.tune.block && g->plan.tune.relabel; }
static CountView count_view(ga_graph* g) {
    if (!relabel_on(g)) return own_view(g);
    const bool pad = g->plan.tune.relabel == 2 && g->pst;
    return CountView{&g->cplan, g->crow, g->ccol, g->csrc, pad ? g->pcol : g->ccol, pad ? g->pst : g->crow};
}

// Flat units and entry records of the light entries (built the first time the flat kernel is switched on).
static int flat_build_units(ga_ctx* ctx, ga_graph* g, const CountView& V) {
    TcPlan& P = *V.P;
    const uint32_t nlight = P.nlight;
    DevBuf hstart_b, wt_b;
    GA_TRY(ga_malloc(ctx, &hstart_b.p, ((size_t)g->n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &wt_b.p, ((size_t)nlight + 1) * 8));
    uint32_t* hstart = (uint32_t*)hstart_b.p;
    unsigned long long* wt = (unsigned long long*)wt_b.p;
    k_plan_rowstart<<<grid_for(nlight, 256, 8192), 256, 0, ctx->stream>>>(P.wh, nlight, hstart);
    GA_TRY(ga_launch_check(ctx, "k_plan_rowstart"));
    k_plan_units<<<grid_for(nlight, 256, 8192), 256, 0, ctx->stream>>>(V.rowptr, P.wh, P.wpl, nlight, wt);
    GA_TRY(ga_launch_check(ctx, "k_plan_units"));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)wt + nlight, 1, 0));
    GA_TRY(ga_malloc(ctx, (void**)&P.wu, ((size_t)nlight + 1) * 8));
    GA_TRY(ga_exclusive_scan_u64(ctx, (const uint64_t*)wt, (uint64_t*)P.wu, (size_t)nlight + 1));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, P.wu + nlight, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    P.utot = ctx->hstat[0];
    if (P.utot >= FLAT_UNIT_LIMIT) return GA_OK;   // too many units for 32-bit offsets: lane groups only
    GA_TRY(ga_malloc(ctx, (void**)&P.frec, (size_t)nlight * 16));
    k_plan_frec<<<grid_for(nlight, 256, 8192), 256, 0, ctx->stream>>>(V.rowptr, V.st, P.wh, P.wpl, P.wu, hstart, nlight, (uint4*)P.frec);
    GA_TRY(ga_launch_check(ctx, "k_plan_frec"));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

// Cuts the light entries' units into chunks of tune.flat_chunk (nothing when the flat kernel is off).
static int flat_rechunk(ga_ctx* ctx, ga_graph* g, const CountView& V) {
    TcPlan& P = *V.P;
    ga_free(P.cstart);
    P.cstart = nullptr;
    P.nch = 0;
    P.chunk_used = (uint32_t)P.tune.flat_chunk;
    if (!P.nlight || !P.tune.flat_chunk) return GA_OK;
    if (!P.wu) GA_TRY(flat_build_units(ctx, g, V));
    if (!P.frec) return GA_OK;
    const uint32_t chunk = (uint32_t)P.tune.flat_chunk;
    const unsigned long long nch = (P.utot + chunk - 1) / chunk;
    GA_TRY(ga_malloc(ctx, (void**)&P.cstart, (nch + 1) * 8));   // cstart (nch + 1) then chs (nch)
    k_plan_chunks<<<grid_for(P.nlight, 256, 8192), 256, 0, ctx->stream>>>(P.wu, P.wh, V.st, P.nlight, chunk, P.cstart,
                                                                          P.cstart + nch + 1);
    GA_TRY(ga_launch_check(ctx, "k_plan_chunks"));
    GA_TRY(ga_fill_u32(ctx, P.cstart + nch, 1, P.nlight));
    P.nch = (uint32_t)nch;
    return GA_OK;
}

static int build_plan(ga_ctx* ctx, ga_graph* g, const CountView& V) {
    TcPlan& P = *V.P;
    if (!P.tune.block) P.tune = default_tuning();
    const uint32_t n = g->n, m = g->m;
    // The sorted work entries depend on the CSR only: a rebuild for other tuning parameters keeps them and redoes the
    // item cutting, bucketing and ordering (on a 2.6B-edge batch union the entry sort is most of a plan build).
    const bool keep = P.entries;
    if (!keep) {
        ga_free(P.wcol);
        ga_free(P.wedge);
        ga_free(P.wpl);
        ga_free(P.wh);
        P.wcol = P.wedge = P.wh = nullptr;
        P.wpl = nullptr;
        P.nwork = P.nlight = 0;
    }
    ga_free(P.items);
    ga_free(P.wu);
    ga_free(P.cstart);
    ga_free(P.frec);
    P.items = P.cstart = P.frec = nullptr;
    P.wu = nullptr;
    for (int b = 0; b < NHASH; b++) P.nitems[b] = 0;
    P.nhuge = 0;
    P.nch = P.chunk_used = 0;
    P.utot = 0;
    DevBuf keys, keys2, vals, vals2, hstart_b, wt_b, wsum_b, flag_b, fpos_b, starts_b, bc_b, bcount_b;
    GA_TRY(ga_malloc(ctx, &bc_b.p, 128));
    uint32_t* bc = (uint32_t*)bc_b.p;   // [30] work entries, [31] light work entries
    GA_TRY(ga_fill_u32(ctx, bc, 32, 0));
    uint32_t nwork = keep ? P.nwork : 0u, nlight = keep ? P.nlight : 0u;
    uint32_t* sorted_vals = nullptr;
    if (m && !keep) {
        GA_TRY(ga_malloc(ctx, &keys.p, (size_t)m * 8));
        GA_TRY(ga_malloc(ctx, &keys2.p, (size_t)m * 8));
        GA_TRY(ga_malloc(ctx, &vals.p, (size_t)m * 4));
        GA_TRY(ga_malloc(ctx, &vals2.p, (size_t)m * 4));
        int nb = 1;
        while (nb < 32 && (1ull << nb) < (uint64_t)n) nb++;
        const bool split = 33 + nb <= 64;   // room for the light/heavy bit above the vertex id
        k_plan_keys<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(V.rowptr, V.src, V.col, m, nb, split,
                                                                     (unsigned long long*)keys.p, (uint32_t*)vals.p, bc + 30);
        GA_TRY(ga_launch_check(ctx, "k_plan_keys"));
        unsigned long long* ko = nullptr;
        GA_TRY(ga_sort_pairs_u64(ctx, (unsigned long long*)keys.p, (unsigned long long*)keys2.p, (uint32_t*)vals.p, (uint32_t*)vals2.p, m,
                                 32 + nb + (split ? 1 : 0), &ko, &sorted_vals));
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, bc + 30, 8, hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        nwork = (uint32_t)(ctx->hstat[0] & 0xFFFFFFFFu);
        nlight = split ? (uint32_t)(ctx->hstat[0] >> 32) : 0u;
    }
    if (!keep) {
        GA_TRY(ga_malloc(ctx, (void**)&P.wcol, std::max<size_t>(nwork, 1) * 4));
        GA_TRY(ga_malloc(ctx, (void**)&P.wedge, std::max<size_t>(nwork, 1) * 4));
        GA_TRY(ga_malloc(ctx, (void**)&P.wpl, std::max<size_t>(nwork, 1) * 8));
        GA_TRY(ga_malloc(ctx, (void**)&P.wh, std::max<size_t>(nwork, 1) * 4));
    }
    P.nwork = nwork;
    P.nlight = nlight;
    if (nwork == 0) {
        GA_TRY(ga_malloc(ctx, (void**)&P.items, 16));
        P.built = P.tune;
        P.ready = true;
        return GA_OK;
    }
    GA_TRY(ga_malloc(ctx, &hstart_b.p, ((size_t)n + 1) * 4));
    GA_TRY(ga_malloc(ctx, &wt_b.p, ((size_t)nwork + 1) * 8));
    GA_TRY(ga_malloc(ctx, &wsum_b.p, ((size_t)nwork + 1) * 8));
    GA_TRY(ga_malloc(ctx, &flag_b.p, ((size_t)nwork + 1) * 4));
    GA_TRY(ga_malloc(ctx, &fpos_b.p, ((size_t)nwork + 1) * 4));
    uint32_t* hstart = (uint32_t*)hstart_b.p;
    uint32_t* wh = P.wh;
    unsigned long long* wt = (unsigned long long*)wt_b.p;
    unsigned long long* wsum = (unsigned long long*)wsum_b.p;
    uint32_t* flag = (uint32_t*)flag_b.p;
    uint32_t* fpos = (uint32_t*)fpos_b.p;
    if (!keep) {
        k_plan_fill<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(V.rowptr, V.src, V.col, sorted_vals, nwork, P.wcol, P.wedge,
                                                                         P.wpl, wh, wt);
        GA_TRY(ga_launch_check(ctx, "k_plan_fill"));
        P.entries = true;
    } else {
        k_plan_work<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(P.wpl, nwork, wt);
        GA_TRY(ga_launch_check(ctx, "k_plan_work"));
    }
    k_plan_rowstart<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(wh, nwork, hstart);
    GA_TRY(ga_launch_check(ctx, "k_plan_rowstart"));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)wt + nwork, 1, 0));
    GA_TRY(ga_exclusive_scan_u64(ctx, (const uint64_t*)wt, (uint64_t*)wsum, (size_t)nwork + 1));
    const uint32_t cap = (uint32_t)P.tune.hash_cap;
    k_plan_flags<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(hstart, wh, wsum, V.rowptr, nwork, (uint32_t)P.tune.work_chunk, cap,
                                                                      plan_shape(P.tune), flag);
    GA_TRY(ga_launch_check(ctx, "k_plan_flags"));
    GA_TRY(ga_fill_u32(ctx, flag + nwork, 1, 0));
    GA_TRY(ga_exclusive_scan_u32(ctx, flag, fpos, (size_t)nwork + 1));
    const uint32_t nblk = (uint32_t)std::min<uint64_t>(2048, std::max<uint64_t>(1, ((uint64_t)nwork + 8191) / 8192));
    const uint32_t span = (uint32_t)((((uint64_t)nwork + nblk - 1) / nblk + 255) / 256 * 256);
    const size_t nbc = (size_t)(NHASH + 1) * nblk;
    GA_TRY(ga_malloc(ctx, &bcount_b.p, (nbc + 1) * 8));
    uint32_t* bcount = (uint32_t*)bcount_b.p;
    uint32_t* boff = bcount + nbc + 1;
    k_plan_item_count<<<nblk, 256, 0, ctx->stream>>>(flag, wh, V.rowptr, nwork, span, cap, bcount);
    GA_TRY(ga_launch_check(ctx, "k_plan_item_count"));
    GA_TRY(ga_fill_u32(ctx, bcount + nbc, 1, 0));
    GA_TRY(ga_exclusive_scan_u32(ctx, bcount, boff, nbc + 1));
    for (int b = 0; b <= NHASH + 1; b++)
        GA_HIP(ctx, hipMemcpyAsync((uint32_t*)ctx->hstat + b, boff + (size_t)b * nblk, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync((uint32_t*)ctx->hstat + 8, fpos + nwork, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t cnt[NHASH + 1];
    const uint32_t* hc = (const uint32_t*)ctx->hstat;
    for (int b = 0; b <= NHASH; b++) cnt[b] = hc[b + 1] - hc[b];
    const uint32_t nall = hc[8];
    if (hc[NHASH + 1] != nall) return ga_fail(ctx, GA_ERR_STATE, "count plan: %u items bucketed, %u expected", hc[NHASH + 1], nall);
    GA_TRY(ga_malloc(ctx, &starts_b.p, (size_t)nall * 4));
    GA_TRY(ga_malloc(ctx, (void**)&P.items, (size_t)nall * 16));
    k_plan_rank<<<grid_for(nwork, 256, 8192), 256, 0, ctx->stream>>>(flag, fpos, nwork, (uint32_t*)starts_b.p);
    GA_TRY(ga_launch_check(ctx, "k_plan_rank"));
    k_plan_items<<<nblk, 256, 0, ctx->stream>>>(flag, fpos, (uint32_t*)starts_b.p, wh, V.rowptr, nwork, nall, span, cap, boff, P.items);
    GA_TRY(ga_launch_check(ctx, "k_plan_items"));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    P.built = P.tune;
    if (P.tune.item_order && nall > 1) {   // shape order within each bucket
        DevBuf k1, k2, v1, v2, it2;
        GA_TRY(ga_malloc(ctx, &k1.p, (size_t)nall * 8));
        GA_TRY(ga_malloc(ctx, &k2.p, (size_t)nall * 8));
        GA_TRY(ga_malloc(ctx, &v1.p, (size_t)nall * 4));
        GA_TRY(ga_malloc(ctx, &v2.p, (size_t)nall * 4));
        GA_TRY(ga_malloc(ctx, &it2.p, (size_t)nall * 16));
        // item_order 2: windows of 2^12 consecutive h ids in id order (locality), shape order inside each window
        const uint32_t wshift = P.tune.item_order >= 4 ? 16u : P.tune.item_order >= 2 ? 12u : 0u;
        const uint32_t* group = wshift ? nullptr : P.group;
        uint32_t gbits = group ? P.group_bits : 0u;
        if (wshift)
            while (gbits < 32 && ((uint64_t)1 << gbits) <= ((uint64_t)n >> wshift)) gbits++;
        k_plan_item_shape<<<grid_for(nall, 256, 8192), 256, 0, ctx->stream>>>(P.items, nall, V.rowptr, P.wpl, cap, group, gbits, wshift,
                                                                              (unsigned long long*)k1.p, (uint32_t*)v1.p);
        GA_TRY(ga_launch_check(ctx, "k_plan_item_shape"));
        unsigned long long* ks = nullptr;
        uint32_t* vs = nullptr;
        GA_TRY(ga_sort_pairs_u64(ctx, (unsigned long long*)k1.p, (unsigned long long*)k2.p, (uint32_t*)v1.p, (uint32_t*)v2.p, nall,
                                 33 + (int)gbits, &ks, &vs));
        k_plan_item_permute<<<grid_for(nall, 256, 8192), 256, 0, ctx->stream>>>(P.items, vs, nall, (uint32_t*)it2.p);
        GA_TRY(ga_launch_check(ctx, "k_plan_item_permute"));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        uint32_t* row_order = P.items;   // freed with it2
        P.items = (uint32_t*)it2.p;
        it2.p = row_order;
    }
    if (V.st != V.rowptr || (P.st_used && P.st_used != V.rowptr)) {   // starts in the line-aligned copy (or back)
        if (nwork || nall) k_plan_starts<<<grid_for(std::max(nwork, nall), 256, 8192), 256, 0, ctx->stream>>>(V.st, P.wcol, nwork, nall, P.items, P.wpl);
        GA_TRY(ga_launch_check(ctx, "k_plan_starts"));
    }
    P.st_used = V.st;
    for (int b = 0; b < NHASH; b++) P.nitems[b] = cnt[b];
    P.nhuge = cnt[NHASH];
    P.ready = true;
    return flat_rechunk(ctx, g, V);
}

static int ensure_relabel(ga_ctx* ctx, ga_graph* g);

int ga_tc_ensure_plan(ga_ctx* ctx, ga_graph* g, bool counts_only) {
    if (!g->plan.tune.block) g->plan.tune = default_tuning();
    if (counts_only && relabel_on(g)) GA_TRY(ensure_relabel(ctx, g));
    const CountView V = counts_only ? count_view(g) : own_view(g);
    TcPlan& P = *V.P;
    if (&P != &g->plan) {   // the renumbered copy's plan follows the graph's tuning and locality groups
        P.tune = g->plan.tune;
        P.group = g->plan.group;
        P.group_bits = g->plan.group_bits;
    }
    if (P.ready && same_plan_params(P.built, P.tune) && P.st_used == V.st)
        return P.chunk_used == (uint32_t)P.tune.flat_chunk ? GA_OK : flat_rechunk(ctx, g, V);
    return build_plan(ctx, g, V);
}

// table slots for the largest d+(h) a bucket may hold (h in bucket b has d+(h) <= bound):
// 2^max(3, ceil(log2(bound)) + shift), the kernel's per-item rule at d+(h) = bound
static uint32_t bucket_slots(uint32_t bound, uint32_t shift) {
    uint32_t lg = 0;
    while ((1u << lg) < bound) lg++;
    uint32_t bits = lg + shift;
    if (bits < 3) bits = 3;   // two buckets at least: the kernel's smallest table
    return 1u << bits;
}

// Launch shape of a lane-group bucket (also used for the shared launches of ga_tc_count_many, so that each graph there
// runs exactly what its own launch runs).
struct HashShape {
    uint32_t block, ts, shift;
    size_t shmem;
    unsigned grid;
    bool fits;   // the tables fit one workgroup's LDS (else the bucket binary-searches N+(h) in HBM)
};

// Dynamic LDS of a plain or per-graph count launch that holds it to at most `waves` waves per SIMD (0: no cap): a CU
// has 4 SIMDs and its LDS is shared by its resident workgroups, so asking each workgroup for more than 1/(k+1) of it
// leaves room for k of them. Fewer items in flight keep more of their probe lists in the L2 (the 4,096-graph batch
// union: 94.5 ms uncapped, 88.9 ms at 6 waves, 88.6 ms at 5, 96.2 ms at 4).
static size_t occupancy_lds(const ga_ctx* ctx, uint32_t waves, uint32_t block, size_t shmem) {
    if (!waves) return shmem;
    const uint32_t wg = std::max(1u, 4u * waves / std::max(1u, block / 64u));   // workgroups per CU at that cap
    const size_t need = ctx->lds_cu / (wg + 1u) + 256u;
    return need > shmem && need <= ctx->lds_max ? need : shmem;
}

static HashShape hash_shape(const ga_ctx* ctx, const ga_graph* g, const TcPlan& P, uint32_t G, bool hashed, int flags, uint32_t nitems,
                            uint32_t bound) {
    HashShape s;
    s.block = (uint32_t)P.tune.block;
    const bool counters = (flags & (F_TRI | F_SUP)) != 0;
    // tables with per-slot counters (participation, supports) take twice the LDS: at most load shift 2 there
    s.shift = counters ? std::min<uint32_t>((uint32_t)P.tune.load_shift, 2u) : (uint32_t)P.tune.load_shift;
    // LDS reserved per lane group: the bucket's largest row at load shift <= 2 (a quarter full at most). Reserving for
    // shift 3 doubled the tables of the 32- and 64-lane buckets and held them to 4 waves per SIMD by LDS.
    s.ts = hashed ? bucket_slots(std::min(bound, std::max(g->max_out, 1u)), std::min<uint32_t>(s.shift, 2u)) : 0;
    const uint32_t per_group = (s.ts * (counters ? 2u : 1u) + 4u * G) * 4u;
    const bool tc = (flags & F_TRI) && !(flags & F_GRAPH);
    const size_t cache_bytes = tc ? 2u * TRI_CACHE * 4u : 0u;   // the workgroup's participation counters
    while (s.block > 64 && (size_t)(s.block / G) * per_group + cache_bytes > 64 * 1024) s.block >>= 1;
    s.shmem = (size_t)(s.block / G) * per_group + cache_bytes;
    s.fits = !hashed || s.shmem <= ctx->lds_max;
    if (s.fits && (flags & ~(F_GRAPH | F_SLOTS)) == 0) s.shmem = occupancy_lds(ctx, (uint32_t)P.tune.waves, s.block, s.shmem);
    // participation: fewer, longer-lived workgroups, so that each counter cache gathers more items before its flush
    // (64 -> 16 per CU: 27.3 -> 23.0 ms for the R-MAT scale-21 peel, 9.5 -> 8.6 ms on the 34.7M-edge graph)
    s.grid = ctx_grid(ctx, grid_for((uint64_t)nitems * G, s.block, (unsigned)std::max(ctx->cus, 1) * (tc ? 16u : 64u)));
    return s;
}

template <int G, bool HASHED, int FLAGS>
static int launch_hash_bucket(ga_ctx* ctx, ga_graph* g, const CountView& V, const uint32_t* items, uint32_t nitems, uint32_t bound,
                              TcArgs a) {
    if (!nitems) return GA_OK;
    const TcPlan& P = *V.P;
    const HashShape sh = hash_shape(ctx, g, P, G, HASHED, FLAGS, nitems, bound);
    if (!sh.fits)   // the tables do not fit one workgroup: binary-search this bucket in HBM
        return launch_hash_bucket<64, false, FLAGS>(ctx, g, V, items, nitems, bound, a);
    a.items = items;
    a.nitems = nitems;
    a.ts_max = sh.ts;
    a.load_shift = sh.shift;
    a.xcd = (P.built.item_order == 3 || P.built.item_order == 5) ? 1u : 0u;
    k_tc_hash<G, HASHED, FLAGS><<<sh.grid, sh.block, sh.shmem, ctx->stream>>>(a);
    return ga_launch_check(ctx, "k_tc_hash");
}

template <int FLAGS>
static int launch_tiny(ga_ctx* ctx, ga_graph* g, const CountView& V, const uint32_t* items, uint32_t nitems, TcArgs a) {
    if (!nitems) return GA_OK;
    const uint32_t block = (uint32_t)V.P->tune.block;
    a.items = items;
    a.nitems = nitems;
    a.xcd = (V.P->built.item_order == 3 || V.P->built.item_order == 5) ? 1u : 0u;
    k_tc_tiny<FLAGS><<<ctx_grid(ctx, grid_for(nitems, block, (unsigned)std::max(ctx->cus, 1) * 64u)), block,
                       occupancy_lds(ctx, (uint32_t)V.P->tune.waves, block, 0), ctx->stream>>>(a);
    return ga_launch_check(ctx, "k_tc_tiny");
}

// Flat kernel on the light rows; their lane-group buckets (d+(h) <= 128: buckets 0-2) are then skipped.
static bool flat_on(const CountView& V) {
    const TcPlan& P = *V.P;
    return P.nch && P.chunk_used && P.tune.hash_cap >= (int)FLAT_MAX;
}

static int launch_flat(ga_ctx* ctx, ga_graph* g, const CountView& V, unsigned long long* total, const uint32_t* done, const uint32_t* gof,
                       unsigned long long* gcnt, uint32_t gstride) {
    const TcPlan& P = *V.P;
    FlatArgs f{};
    f.rowptr = V.rowptr;
    f.col = V.kcol;
    f.wh = P.wh;
    f.frec = (const uint4*)P.frec;
    f.cstart = P.cstart;
    f.chs = P.cstart + P.nch + 1;
    f.nch = P.nch;
    f.nlight = P.nlight;
    f.utot = (uint32_t)P.utot;
    f.total = total;
    f.done = done;
    f.gof = gof;
    f.gcnt = gcnt;
    f.gstride = gstride;
    const size_t shmem = flat_shmem(P.chunk_used);
    if (shmem > ctx->lds_max) return ga_fail(ctx, GA_ERR_ARG, "flat chunk %u needs %zu bytes of LDS", P.chunk_used, shmem);
    const uint32_t block = (uint32_t)P.tune.flat_block;
    const uint32_t grid = ctx_grid(ctx, std::min<uint32_t>(P.nch, (uint32_t)std::max(ctx->cus, 1) * 16u));
    const int upt = (int)(P.chunk_used / block);
    if (upt != 4 && upt != 8 && upt != 16) return ga_fail(ctx, GA_ERR_ARG, "flat chunk / block must be 4, 8 or 16");
    if (gcnt) {
        if (upt == 4) k_tc_flat<4, true><<<grid, block, shmem, ctx->stream>>>(f);
        else if (upt == 8) k_tc_flat<8, true><<<grid, block, shmem, ctx->stream>>>(f);
        else k_tc_flat<16, true><<<grid, block, shmem, ctx->stream>>>(f);
    } else {
        if (upt == 4) k_tc_flat<4, false><<<grid, block, shmem, ctx->stream>>>(f);
        else if (upt == 8) k_tc_flat<8, false><<<grid, block, shmem, ctx->stream>>>(f);
        else k_tc_flat<16, false><<<grid, block, shmem, ctx->stream>>>(f);
    }
    return ga_launch_check(ctx, "k_tc_flat");
}

// A lane-group bucket with its LDS table; plain and per-graph counts read their lists in lane slots when tune.probe.
template <int G, int FLAGS>
static int launch_hashed(ga_ctx* ctx, ga_graph* g, const CountView& V, const uint32_t* items, uint32_t ni, uint32_t bound, TcArgs a) {
    if constexpr (FLAGS == 0 || FLAGS == F_GRAPH)
        if (V.P->tune.probe) return launch_hash_bucket<G, true, FLAGS | F_SLOTS>(ctx, g, V, items, ni, bound, a);
    return launch_hash_bucket<G, true, FLAGS>(ctx, g, V, items, ni, bound, a);
}

// tune.streams: the bucket launches of one count go to four streams (ctx->stream and three side streams that wait for
// it), so that one launch's tail and the one-lane kernel of the shortest rows, which uses no LDS, overlap the others;
// on scope exit ctx->stream is restored and waits for every side stream used (also on an error return).
struct StreamFork {
    ga_ctx* ctx;
    hipStream_t main;
    bool on = false;
    bool used[3] = {false, false, false};
    StreamFork(ga_ctx* c, bool want) : ctx(c), main(c->stream) {
        if (!want) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(main, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;   // captured: one stream
        for (int i = 0; i < 3; i++) {
            if (!c->side[i] && hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking) != hipSuccess) return;
            if (!c->join_ev[i] && hipEventCreateWithFlags(&c->join_ev[i], hipEventDisableTiming) != hipSuccess) return;
        }
        if (!c->fork_ev && hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess) return;
        if (hipEventRecord(c->fork_ev, main) != hipSuccess) return;
        for (int i = 0; i < 3; i++)
            if (hipStreamWaitEvent(c->side[i], c->fork_ev, 0) != hipSuccess) return;
        on = true;
    }
    void to(int i) {   // -1: ctx->stream
        if (!on) return;
        if (i >= 0) used[i] = true;
        ctx->stream = i >= 0 ? ctx->side[i] : main;
    }
    ~StreamFork() {
        if (!on) return;
        ctx->stream = main;
        for (int i = 0; i < 3; i++)
            if (used[i]) {
                (void)hipEventRecord(ctx->join_ev[i], ctx->side[i]);
                (void)hipStreamWaitEvent(main, ctx->join_ev[i], 0);
            }
    }
};

// evs (optional, NHASH + 2 events): evs[b] is recorded before bucket b's launch, evs[NHASH] before the binary-search
// bucket's and evs[NHASH + 1] after it (per-launch timing of ga_tc_bucket_profile, always on one stream).
template <int FLAGS>
static int launch_hash(ga_ctx* ctx, ga_graph* g, const CountView& V, TcArgs a, int first_bucket = 0, hipEvent_t* evs = nullptr) {
    const TcPlan& P = *V.P;
    // bucket 1 (d+(h) <= 32) on ctx->stream, buckets 0 and 2 on side streams 0 and 1, the rest after each other on 2
    StreamFork fork(ctx, (FLAGS == 0 || FLAGS == F_GRAPH) && !evs && P.tune.streams);
    static const int stream_of[NHASH + 1] = {0, -1, 1, 2, 2, 2};
    uint32_t off = 0;
    for (int b = 0; b < NHASH; b++) {
        if (evs) GA_HIP(ctx, hipEventRecord(evs[b], ctx->stream));
        fork.to(stream_of[b]);
        const uint32_t* items = P.items + 4ull * off;
        const uint32_t ni = b < first_bucket ? 0u : P.nitems[b];
        const uint32_t bound = std::min(BUCKET_DEG[b], (uint32_t)P.tune.hash_cap);
        int rc = GA_OK;
        switch (P.tune.lanes[b]) {
            case 1:   // bucket 0 only: one lane per item for plain and per-graph counts, 4-lane groups otherwise
                if constexpr (FLAGS == 0 || FLAGS == F_GRAPH) rc = launch_tiny<FLAGS>(ctx, g, V, items, ni, a);
                else rc = launch_hash_bucket<4, true, FLAGS>(ctx, g, V, items, ni, bound, a);
                break;
            case 4: rc = launch_hashed<4, FLAGS>(ctx, g, V, items, ni, bound, a); break;
            case 8: rc = launch_hashed<8, FLAGS>(ctx, g, V, items, ni, bound, a); break;
            case 16: rc = launch_hashed<16, FLAGS>(ctx, g, V, items, ni, bound, a); break;
            case 32: rc = launch_hashed<32, FLAGS>(ctx, g, V, items, ni, bound, a); break;
            default: rc = launch_hashed<64, FLAGS>(ctx, g, V, items, ni, bound, a); break;
        }
        GA_TRY(rc);
        off += P.nitems[b];
    }
    fork.to(stream_of[NHASH]);
    if (evs) GA_HIP(ctx, hipEventRecord(evs[NHASH], ctx->stream));
    GA_TRY((launch_hash_bucket<64, false, FLAGS>(ctx, g, V, P.items + 4ull * off, P.nhuge, 0, a)));
    fork.to(-1);
    if (evs) GA_HIP(ctx, hipEventRecord(evs[NHASH + 1], ctx->stream));
    return GA_OK;
}

template <int FLAGS>
static int launch_merge(ga_ctx* ctx, ga_graph* g, const uint32_t* order, unsigned long long* total, const uint8_t* alive, uint32_t* tri,
                        uint32_t* sup, const uint32_t* done) {
    if (!g->m) return GA_OK;
    k_tc_merge<FLAGS><<<ctx_grid(ctx, grid_for(g->m, 256)), 256, 0, ctx->stream>>>(g->rowptr, g->col, g->src, order, g->m, alive, tri, sup, total, done);
    return ga_launch_check(ctx, "k_tc_merge");
}

template <int G>
static int launch_group(ga_ctx* ctx, ga_graph* g, unsigned long long* total) {
    if (!g->m) return GA_OK;
    k_tc_group<G><<<ctx_grid(ctx, grid_for((uint64_t)g->m * G, 256)), 256, 0, ctx->stream>>>(g->rowptr, g->col, g->src, g->m, total);
    return ga_launch_check(ctx, "k_tc_group");
}

// ---------------------------------------------------------------------------------------------------------------
// the renumbered copy of plain and per-graph counts (tune.relabel)
// ---------------------------------------------------------------------------------------------------------------
// Vertex v is keyed by lab(v) = the smallest id among v and its neighbours u with sup(u, v) >= 2 (the edge closes two
// or more triangles), and the new ids follow (lab(v), v). Vertices of a dense cluster mostly share a few keys, so their
// out-lists, which the count's items read together, land in nearby lines. Every strong neighbour of v lies in v's own
// graph of a disjoint union and lab(v) <= v, so each graph keeps its id range. The renumbered CSR has the same
// orientation (edge u -> v becomes perm(u) -> perm(v), rows re-sorted), hence the same count and the same work.
namespace {

__global__ void k_relabel_init(uint32_t n, uint32_t* __restrict__ lab) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) lab[v] = v;
}

__global__ void k_relabel_strong(const uint32_t* __restrict__ src, const uint32_t* __restrict__ col, const uint32_t* __restrict__ sup,
                                 uint32_t m, uint32_t* __restrict__ lab) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
        if (sup[e] >= 2) {
            const uint32_t u = src[e], v = col[e];
            atomicMin(&lab[u], v);
            atomicMin(&lab[v], u);
        }
}

__global__ void k_relabel_vkeys(const uint32_t* __restrict__ lab, uint32_t n, unsigned long long* __restrict__ key) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        key[v] = ((unsigned long long)lab[v] << 32) | v;
}

// perm[old] = new, and the new row length of each new id
__global__ void k_relabel_perm(const unsigned long long* __restrict__ sorted, const uint32_t* __restrict__ rowptr, uint32_t n,
                               uint32_t* __restrict__ perm, uint32_t* __restrict__ deg) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)sorted[i];
        perm[v] = i;
        deg[i] = rowptr[v + 1] - rowptr[v];
    }
}

__global__ void k_relabel_ekeys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ col, const uint32_t* __restrict__ perm,
                                uint32_t m, unsigned long long* __restrict__ key) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
        key[e] = ((unsigned long long)perm[src[e]] << 32) | perm[col[e]];
}

__global__ void k_relabel_unpack(const unsigned long long* __restrict__ key, uint32_t m, uint32_t* __restrict__ col,
                                 uint32_t* __restrict__ src) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const unsigned long long k = key[e];
        col[e] = (uint32_t)k;
        src[e] = (uint32_t)(k >> 32);
    }
}

// the line-aligned copy: edge e of row v goes to pst[v] + (e - rowptr[v])
__global__ void k_relabel_pad(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                              const uint32_t* __restrict__ pst, uint32_t m, uint32_t* __restrict__ pcol) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const uint32_t v = src[e];
        pcol[pst[v] + (e - rowptr[v])] = col[e];
    }
}

// Row starts of the line-aligned copy (host, once per graph): rows in id order, a row moved to the next 128-byte line
// (32 ids) when that makes it cross one line fewer, so rows of at most 32 ids cross none. Returns the padded length,
// or 0 when it does not fit in 32 bits.
uint64_t padded_starts(const std::vector<uint32_t>& rowptr, uint32_t n, std::vector<uint32_t>& pst) {
    pst.assign((size_t)n + 1, 0);
    uint64_t off = 0;
    for (uint32_t v = 0; v < n; v++) {
        const uint64_t d = rowptr[v + 1] - rowptr[v];
        const uint64_t r = off & 31u;
        if (d && r && (r + d + 31u) / 32u > (d + 31u) / 32u) off += 32u - r;
        if (off + d > 0xFFFFFFFFull) return 0;
        pst[v] = (uint32_t)off;
        off += d;
    }
    pst[n] = (uint32_t)off;
    return off;
}

}  // namespace

template <int FLAGS>
static int launch_merge(ga_ctx* ctx, ga_graph* g, const uint32_t* order, unsigned long long* total, const uint8_t* alive, uint32_t* tri,
                        uint32_t* sup, const uint32_t* done);

static int ensure_relabel(ga_ctx* ctx, ga_graph* g) {
    if (g->crow) return GA_OK;
    const uint32_t n = g->n, m = g->m;
    uint32_t *crow = nullptr, *ccol = nullptr, *csrc = nullptr;
    auto fail = [&](int rc) {
        ga_free(crow);
        ga_free(ccol);
        ga_free(csrc);
        return rc;
    };
    if (ga_malloc(ctx, (void**)&crow, ((size_t)n + 1) * 4) || ga_malloc(ctx, (void**)&ccol, std::max<size_t>(m, 1) * 4) ||
        ga_malloc(ctx, (void**)&csrc, std::max<size_t>(m, 1) * 4))
        return fail(GA_ERR_OOM);
    int nb = 1;
    while (nb < 32 && (1ull << nb) < (uint64_t)n) nb++;
    int rc = GA_OK;
    {
        DevBuf sup, lab, cnt;
        DevBuf vk0, vk1, perm, deg;
        if ((rc = ga_malloc(ctx, &sup.p, std::max<size_t>(m, 1) * 4)) || (rc = ga_malloc(ctx, &lab.p, ((size_t)n + 1) * 4)) ||
            (rc = ga_malloc(ctx, &cnt.p, 8)) || (rc = ga_malloc(ctx, &vk0.p, (size_t)n * 8)) || (rc = ga_malloc(ctx, &vk1.p, (size_t)n * 8)) ||
            (rc = ga_malloc(ctx, &perm.p, (size_t)n * 4)) || (rc = ga_malloc(ctx, &deg.p, ((size_t)n + 1) * 4)))
            return fail(rc);
        // supports of every oriented edge (plan-free merge count) and the strong-neighbour key of every vertex
        if ((rc = ga_fill_u32(ctx, (uint32_t*)sup.p, m, 0)) || (rc = ga_fill_u64(ctx, (uint64_t*)cnt.p, 1, 0))) return fail(rc);
        if ((rc = launch_merge<F_SUP>(ctx, g, nullptr, (unsigned long long*)cnt.p, nullptr, nullptr, (uint32_t*)sup.p, nullptr))) return fail(rc);
        k_relabel_init<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(n, (uint32_t*)lab.p);
        if ((rc = ga_launch_check(ctx, "k_relabel_init"))) return fail(rc);
        if (m) k_relabel_strong<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(g->src, g->col, (const uint32_t*)sup.p, m, (uint32_t*)lab.p);
        if ((rc = ga_launch_check(ctx, "k_relabel_strong"))) return fail(rc);
        k_relabel_vkeys<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>((const uint32_t*)lab.p, n, (unsigned long long*)vk0.p);
        if ((rc = ga_launch_check(ctx, "k_relabel_vkeys"))) return fail(rc);
        unsigned long long* vs = nullptr;
        uint64_t* vsorted = nullptr;
        if ((rc = ga_sort_u64(ctx, (uint64_t*)vk0.p, (uint64_t*)vk1.p, n, 32 + nb, &vsorted))) return fail(rc);
        vs = (unsigned long long*)vsorted;
        k_relabel_perm<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(vs, g->rowptr, n, (uint32_t*)perm.p, (uint32_t*)deg.p);
        if ((rc = ga_launch_check(ctx, "k_relabel_perm"))) return fail(rc);
        if ((rc = ga_fill_u32(ctx, (uint32_t*)deg.p + n, 1, 0))) return fail(rc);
        if ((rc = ga_exclusive_scan_u32(ctx, (const uint32_t*)deg.p, crow, (size_t)n + 1))) return fail(rc);
        if (m) {
            DevBuf ek0, ek1;
            if ((rc = ga_malloc(ctx, &ek0.p, (size_t)m * 8)) || (rc = ga_malloc(ctx, &ek1.p, (size_t)m * 8))) return fail(rc);
            k_relabel_ekeys<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(g->src, g->col, (const uint32_t*)perm.p, m,
                                                                            (unsigned long long*)ek0.p);
            if ((rc = ga_launch_check(ctx, "k_relabel_ekeys"))) return fail(rc);
            uint64_t* es = nullptr;
            if ((rc = ga_sort_u64(ctx, (uint64_t*)ek0.p, (uint64_t*)ek1.p, m, 32 + nb, &es))) return fail(rc);
            k_relabel_unpack<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>((const unsigned long long*)es, m, ccol, csrc);
            if ((rc = ga_launch_check(ctx, "k_relabel_unpack"))) return fail(rc);
        }
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ga_fail(ctx, GA_ERR_HIP, "relabel failed"));
    }
    g->crow = crow;
    g->ccol = ccol;
    g->csrc = csrc;
    // the line-aligned copy of relabel 2 (left out when it does not fit: relabel 2 then reads the plain copy)
    std::vector<uint32_t> hrow((size_t)n + 1), hpst;
    if (hipMemcpy(hrow.data(), crow, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return ga_fail(ctx, GA_ERR_HIP, "relabel: row offsets");
    const uint64_t plen = padded_starts(hrow, n, hpst);
    uint32_t *pst = nullptr, *pcol = nullptr;
    if (plen && !ga_malloc(ctx, (void**)&pst, ((size_t)n + 1) * 4) && !ga_malloc(ctx, (void**)&pcol, std::max<size_t>(plen, 1) * 4)) {
        int rc2 = hipMemcpy(pst, hpst.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ? ga_fail(ctx, GA_ERR_HIP, "relabel: padded starts")
                                                                                                       : ga_fill_u32(ctx, pcol, plen, GA_EMPTY);
        if (!rc2 && m) {
            k_relabel_pad<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(crow, csrc, ccol, pst, m, pcol);
            rc2 = ga_launch_check(ctx, "k_relabel_pad");
        }
        if (!rc2 && hipStreamSynchronize(ctx->stream) != hipSuccess) rc2 = ga_fail(ctx, GA_ERR_HIP, "relabel: padded copy");
        if (rc2) {
            ga_free(pst);
            ga_free(pcol);
            return rc2;
        }
        g->pst = pst;
        g->pcol = pcol;
    } else {
        ga_free(pst);   // (an allocation failure here leaves only the message; relabel 2 reads the plain copy)
        ga_free(pcol);
    }
    return GA_OK;
}

// Launches one count of `variant` into *d_total (zeroed by the caller). alive/tri/sup as in the flags.
int ga_tc_launch(ga_ctx* ctx, ga_graph* g, int variant, unsigned long long* d_total, const uint8_t* alive, uint32_t* tri, uint32_t* sup,
                 const uint32_t* done, const uint32_t* gof, unsigned long long* gcnt, uint32_t gstride) {
    unsigned long long* total = d_total;
    if (variant == GA_TC_AUTO) variant = GA_TC_HASH;
    const int flags = (alive ? F_ALIVE : 0) | (tri ? F_TRI : 0) | (sup ? F_SUP : 0) | (gcnt ? F_GRAPH : 0);
    if (variant == GA_TC_HASH) {
        const bool counts = flags == 0 || flags == F_GRAPH;   // plain and per-graph counts may run on the renumbered copy
        GA_TRY(ga_tc_ensure_plan(ctx, g, counts));
        const CountView V = counts ? count_view(g) : own_view(g);
        TcArgs a{};
        a.rowptr = V.rowptr;
        a.col = V.kcol;
        a.wcol = V.P->wcol;
        a.wedge = V.P->wedge;
        a.wpl = V.P->wpl;
        a.alive = alive;
        a.tri = tri;
        a.sup = sup;
        a.total = total;
        a.done = done;
        a.gof = gof;
        a.gcnt = gcnt;
        a.gstride = gstride;
        if (counts && flat_on(V)) {
            GA_TRY(launch_flat(ctx, g, V, total, done, gof, gcnt, gstride));
            return flags ? launch_hash<F_GRAPH>(ctx, g, V, a, 3) : launch_hash<0>(ctx, g, V, a, 3);
        }
        switch (flags) {
            case 0: return launch_hash<0>(ctx, g, V, a);
            case F_TRI: return launch_hash<F_TRI>(ctx, g, V, a);
            case F_ALIVE | F_TRI: return launch_hash<F_ALIVE | F_TRI>(ctx, g, V, a);
            case F_SUP: return launch_hash<F_SUP>(ctx, g, V, a);
            case F_GRAPH: return launch_hash<F_GRAPH>(ctx, g, V, a);
            case F_ALIVE | F_TRI | F_GRAPH: return launch_hash<F_ALIVE | F_TRI | F_GRAPH>(ctx, g, V, a);
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
    if (variant == GA_TC_NOSRC || variant == GA_TC_VERTEX) return ga_tc_launch_simple(ctx, g, variant, total);
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
        case GA_TC_NOSRC: return "group4_rowptr_src";
        case GA_TC_VERTEX: return "vertex_merge";
        default: return "unknown";
    }
}

extern "C" int ga_tc_prepare(ga_ctx* ctx, ga_graph* g, float* ms) {
    if (!ctx || !g) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    g->plan.ready = false;
    g->cplan.ready = false;
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(ga_tc_ensure_plan(ctx, g, true));   // the plan the counts use (with the renumbered copy when relabel is set)
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}

extern "C" int ga_tc_count(ga_ctx* ctx, ga_graph* g, int variant, uint64_t* T, float* ms) {
    if (!ctx || !g || !T) return GA_ERR_ARG;
    if (variant < 0 || variant >= GA_TC_NVARIANTS) return ga_fail(ctx, GA_ERR_ARG, "unknown variant %d", variant);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (variant == GA_TC_AUTO || variant == GA_TC_HASH) GA_TRY(ga_tc_ensure_plan(ctx, g, true));
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

// ---- shared launches over several resident graphs (ga_tc_count_many, ga_peel_run_batch) ----
// A launch group: graphs whose bucket runs the same kernel (kind, lanes, flags, block); one launch covers them.
struct ManyGroup {
    int kind;    // 0: one lane per item (k_tc_tiny), 1: lane groups with LDS tables, 2: binary search in HBM
    int G, flags;
    uint32_t block;
    size_t shmem = 0;
    std::vector<TcJob> jobs;
    std::vector<unsigned> grid;   // each job's own grid
    std::vector<uint32_t> map;    // workgroup -> job (index into the whole job table)
    unsigned nwg = 0;
};

// The groups of one set of shared launches; jobs and maps live in the context's grow-only buffer (ctx->dtcm, host copy
// ctx->htcm), so one such set is in use at a time on a context.
struct ga_tc_many {
    std::vector<ManyGroup> groups;
    const TcJob* djobs = nullptr;
    const uint32_t* dmap = nullptr;
};

// Adds graph g's buckets as jobs: plain counts (flags 0; lane slots where the tuning says so, one lane per item for
// bucket 0 when lanes[0] = 1) or participation (F_ALIVE | F_TRI: 4-lane groups for bucket 0, length classes). base
// holds the graph's pointers; the items, tables and grid of each bucket are filled in as its own launch would.
static void many_add(const ga_ctx* ctx, ga_tc_many& M, ga_graph* g, const CountView& V, const TcArgs& base, int flags) {
    const TcPlan& P = *V.P;
    const unsigned cus = (unsigned)std::max(ctx->cus, 1);
    auto group_of = [&](int kind, int G, int fl, uint32_t block) -> ManyGroup& {
        for (auto& gr : M.groups)
            if (gr.kind == kind && gr.G == G && gr.flags == fl && gr.block == block) return gr;
        M.groups.push_back(ManyGroup{kind, G, fl, block});
        return M.groups.back();
    };
    uint32_t off = 0;
    for (int b = 0; b <= NHASH; b++) {   // b = NHASH: the binary-search bucket
        const uint32_t ni = b < NHASH ? P.nitems[b] : P.nhuge;
        TcArgs a = base;
        a.items = P.items + 4ull * off;
        a.nitems = ni;
        a.xcd = (P.built.item_order == 3 || P.built.item_order == 5) ? 1u : 0u;
        off += ni;
        if (!ni) continue;
        int G = b < NHASH ? P.tune.lanes[b] : 64;
        if (G == 1 && flags == 0) {   // bucket 0, one lane per item
            ManyGroup& gr = group_of(0, 1, 0, (uint32_t)P.tune.block);
            gr.jobs.push_back(TcJob{a, 0, 0});
            gr.grid.push_back(ctx_grid(ctx, grid_for(ni, gr.block, cus * 64u)));
            continue;
        }
        if (G == 1) G = 4;   // participation keeps 4-lane groups there
        bool hashed = b < NHASH;
        const uint32_t bound = b < NHASH ? std::min(BUCKET_DEG[b], (uint32_t)P.tune.hash_cap) : 0u;
        int fl = flags | (hashed && flags == 0 && P.tune.probe ? F_SLOTS : 0);
        HashShape sh = hash_shape(ctx, g, P, (uint32_t)G, hashed, fl, ni, bound);
        if (!sh.fits) {   // as the graph's own launch: binary search in HBM
            G = 64;
            hashed = false;
            fl = flags;
            sh = hash_shape(ctx, g, P, 64u, false, fl, ni, bound);
        }
        a.ts_max = sh.ts;
        a.load_shift = sh.shift;
        ManyGroup& gr = group_of(hashed ? 1 : 2, G, fl, sh.block);
        gr.shmem = std::max(gr.shmem, sh.shmem);
        gr.jobs.push_back(TcJob{a, 0, 0});
        gr.grid.push_back(sh.grid);
    }
}

// Workgroup ranges (each job its own grid, scaled down together when a launch would exceed its budget: 64 workgroups
// per CU, 16 for participation) and the upload of jobs and maps into the context's grow-only buffer. The host copy
// stays untouched until the caller's stream sync.
static int many_upload(ga_ctx* ctx, ga_tc_many& M) {
    const unsigned cus = (unsigned)std::max(ctx->cus, 1);
    size_t jobs_total = 0, map_total = 0;
    for (auto& gr : M.groups) {
        uint64_t sum = 0;
        for (unsigned x : gr.grid) sum += x;
        const bool tc = (gr.flags & F_TRI) != 0;
        const uint64_t budget = (tc ? 16ull : 64ull) * cus;
        gr.map.clear();
        for (size_t j = 0; j < gr.jobs.size(); j++) {
            unsigned nb = gr.grid[j];
            if (sum > budget) nb = std::max(1u, (unsigned)((uint64_t)nb * budget / sum));
            gr.jobs[j].base = (uint32_t)gr.map.size();
            gr.jobs[j].nblk = nb;
            gr.map.insert(gr.map.end(), nb, (uint32_t)(jobs_total + j));
        }
        gr.nwg = (unsigned)gr.map.size();
        jobs_total += gr.jobs.size();
        map_total += gr.map.size();
    }
    if (!jobs_total) return GA_OK;
    const size_t jobs_bytes = (jobs_total * sizeof(TcJob) + 255) & ~(size_t)255;
    const size_t need = jobs_bytes + map_total * 4;
    if (need > ctx->dtcm_bytes) {   // grow-only
        ga_free(ctx->dtcm);
        ctx->dtcm = nullptr;
        ctx->dtcm_bytes = 0;
        GA_TRY(ga_malloc(ctx, &ctx->dtcm, need * 2));
        ctx->dtcm_bytes = need * 2;
    }
    std::vector<uint8_t>& host = ctx->htcm;
    host.resize(need);
    size_t jo = 0, mo = jobs_bytes;
    for (const auto& gr : M.groups) {
        memcpy(host.data() + jo, gr.jobs.data(), gr.jobs.size() * sizeof(TcJob));
        jo += gr.jobs.size() * sizeof(TcJob);
        memcpy(host.data() + mo, gr.map.data(), gr.map.size() * 4);
        mo += gr.map.size() * 4;
    }
    GA_HIP(ctx, hipMemcpyAsync(ctx->dtcm, host.data(), need, hipMemcpyHostToDevice, ctx->stream));
    M.djobs = (const TcJob*)ctx->dtcm;
    M.dmap = (const uint32_t*)((const char*)ctx->dtcm + jobs_bytes);
    return GA_OK;
}

template <int FLAGS>
static void launch_many_hash(ga_ctx* ctx, int G, bool hashed, unsigned nwg, uint32_t block, size_t shmem, const TcJob* jobs, const uint32_t* map) {
    if (!hashed) { k_tc_hash_many<64, false, FLAGS & ~F_SLOTS><<<nwg, block, shmem, ctx->stream>>>(jobs, map); return; }
    switch (G) {
        case 4: k_tc_hash_many<4, true, FLAGS><<<nwg, block, shmem, ctx->stream>>>(jobs, map); break;
        case 8: k_tc_hash_many<8, true, FLAGS><<<nwg, block, shmem, ctx->stream>>>(jobs, map); break;
        case 16: k_tc_hash_many<16, true, FLAGS><<<nwg, block, shmem, ctx->stream>>>(jobs, map); break;
        case 32: k_tc_hash_many<32, true, FLAGS><<<nwg, block, shmem, ctx->stream>>>(jobs, map); break;
        default: k_tc_hash_many<64, true, FLAGS><<<nwg, block, shmem, ctx->stream>>>(jobs, map); break;
    }
}

static int many_launch(ga_ctx* ctx, const ga_tc_many& M) {
    const uint32_t* dmap = M.dmap;
    for (const auto& gr : M.groups) {
        const bool hashed = gr.kind == 1;
        if (gr.kind == 0) k_tc_tiny_many<0><<<gr.nwg, gr.block, 0, ctx->stream>>>(M.djobs, dmap);
        else if (gr.flags == F_SLOTS) launch_many_hash<F_SLOTS>(ctx, gr.G, hashed, gr.nwg, gr.block, gr.shmem, M.djobs, dmap);
        else if (gr.flags == (F_ALIVE | F_TRI)) launch_many_hash<F_ALIVE | F_TRI>(ctx, gr.G, hashed, gr.nwg, gr.block, gr.shmem, M.djobs, dmap);
        else launch_many_hash<0>(ctx, gr.G, hashed, gr.nwg, gr.block, gr.shmem, M.djobs, dmap);
        GA_TRY(ga_launch_check(ctx, "k_tc_hash_many"));
        dmap += gr.nwg;
    }
    return GA_OK;
}

// Plain counts of several resident graphs into counts[i] with one launch per bucket shape instead of one launch set per
// graph. Graphs whose tuning runs the flat kernel are counted by their own launches.
static int count_many_shared(ga_ctx* ctx, ga_graph* const* graphs, uint32_t count, unsigned long long* counts) {
    ga_tc_many M;
    std::vector<uint32_t> solo;
    for (uint32_t i = 0; i < count; i++) {
        const CountView V = count_view(graphs[i]);
        if (flat_on(V)) {
            solo.push_back(i);
            continue;
        }
        TcArgs base{};
        base.rowptr = V.rowptr;
        base.col = V.kcol;
        base.wcol = V.P->wcol;
        base.wedge = V.P->wedge;
        base.wpl = V.P->wpl;
        base.total = counts + i;
        many_add(ctx, M, graphs[i], V, base, 0);
    }
    GA_TRY(many_upload(ctx, M));   // ctx->htcm stays untouched until the stream sync that ends ga_tc_count_many
    GA_TRY(many_launch(ctx, M));
    for (uint32_t i : solo) GA_TRY(ga_tc_launch(ctx, graphs[i], GA_TC_HASH, counts + i, nullptr, nullptr, nullptr, nullptr));
    return GA_OK;
}

// Participation counts (alive masks, per-vertex tri[], triangle total, device `done` flag each) of several graphs, for
// ga_peel_run_batch: built and uploaded once, launched every round.
int ga_tc_participation_many(ga_ctx* ctx, ga_graph* const* graphs, const uint8_t* const* alive, uint32_t* const* tri,
                             unsigned long long* const* total, const uint32_t* const* done, uint32_t count, ga_tc_many** out) {
    *out = nullptr;
    ga_tc_many* M = new ga_tc_many;
    for (uint32_t i = 0; i < count; i++) {
        int rc = ga_tc_ensure_plan(ctx, graphs[i]);
        if (rc) { delete M; return rc; }
        const CountView V = own_view(graphs[i]);
        TcArgs base{};
        base.rowptr = V.rowptr;
        base.col = V.kcol;
        base.wcol = V.P->wcol;
        base.wedge = V.P->wedge;
        base.wpl = V.P->wpl;
        base.alive = alive[i];
        base.tri = tri[i];
        base.total = total[i];
        base.done = done[i];
        many_add(ctx, *M, graphs[i], V, base, F_ALIVE | F_TRI);
    }
    int rc = many_upload(ctx, *M);
    if (rc) { delete M; return rc; }
    *out = M;
    return GA_OK;
}

int ga_tc_many_launch(ga_ctx* ctx, const ga_tc_many* M) { return many_launch(ctx, *M); }

void ga_tc_many_free(ga_tc_many* M) { delete M; }

extern "C" int ga_tc_count_many(ga_ctx* ctx, ga_graph* const* graphs, uint32_t count, int variant, uint64_t* T, float* ms) {
    if (!ctx || (count && (!graphs || !T))) return GA_ERR_ARG;
    if (variant < 0 || variant >= GA_TC_NVARIANTS) return ga_fail(ctx, GA_ERR_ARG, "unknown variant %d", variant);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    for (uint32_t i = 0; i < count; i++) {
        if (!graphs[i]) return ga_fail(ctx, GA_ERR_ARG, "graph %u is NULL", i);
        if (variant == GA_TC_AUTO || variant == GA_TC_HASH) GA_TRY(ga_tc_ensure_plan(ctx, graphs[i], true));
        if (variant == GA_TC_SORTED) GA_TRY(ga_graph_ensure_perm(ctx, graphs[i]));
    }
    if (count > ctx->dmany_n) {
        ga_free(ctx->dmany);
        ctx->dmany = nullptr;
        ctx->dmany_n = 0;
        const size_t want = std::max<size_t>(count, 256);
        GA_TRY(ga_malloc(ctx, (void**)&ctx->dmany, want * 8));
        ctx->dmany_n = want;
    }
    if (!count) return GA_OK;
    GA_TRY(ga_fill_u64(ctx, ctx->dmany, count, 0));
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if ((variant == GA_TC_AUTO || variant == GA_TC_HASH) && count > 1) {
        GA_TRY(count_many_shared(ctx, graphs, count, (unsigned long long*)ctx->dmany));
    } else {
        for (uint32_t i = 0; i < count; i++)
            GA_TRY(ga_tc_launch(ctx, graphs[i], variant, (unsigned long long*)ctx->dmany + i, nullptr, nullptr, nullptr, nullptr));
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(T, ctx->dmany, (size_t)count * 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}

extern "C" int ga_tc_plan_info(ga_ctx* ctx, ga_graph* g, uint32_t* nwork, uint32_t* nitems6) {
    if (!ctx || !g) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(ga_tc_ensure_plan(ctx, g, true));
    const TcPlan& P = *count_view(g).P;
    if (nwork) *nwork = P.nwork;
    if (nitems6) {
        for (int b = 0; b < NHASH; b++) nitems6[b] = P.nitems[b];
        nitems6[NHASH] = P.nhuge;
    }
    return GA_OK;
}

struct BucketStarts {   // first item of each bucket of a plan (by value into k_bucket_work)
    uint32_t s[NHASH + 2];
};

namespace {
// Work of each bucket's items: entries and probe elements (each entry's list is min(d+u, d+v) long). bstart[b] is the
// first item of bucket b (NHASH hashed buckets, then the binary-search one; bstart[NHASH + 1] = all items).
__global__ void k_bucket_work(const uint4* __restrict__ items, const uint2* __restrict__ wpl, BucketStarts bs,
                              unsigned long long* __restrict__ out) {
    uint64_t ent[NHASH + 1] = {}, el[NHASH + 1] = {};   // this lane's sums per bucket, reduced per wavefront at the end
    for (uint32_t it = blockIdx.x * blockDim.x + threadIdx.x; it < bs.s[NHASH + 1]; it += gridDim.x * blockDim.x) {
        const uint4 im = items[it];
        const uint32_t ne = im.w & ITEM_NMASK;
        uint64_t s = 0;
        for (uint32_t j = im.y; j < im.y + ne; j++) s += wpl[j].y;
#pragma unroll
        for (uint32_t b = 0; b <= NHASH; b++)
            if (it >= bs.s[b] && it < bs.s[b + 1]) {
                ent[b] += ne;
                el[b] += s;
            }
    }
#pragma unroll
    for (uint32_t b = 0; b <= NHASH; b++) {
        const uint64_t e = wave_sum_u64(ent[b]), l = wave_sum_u64(el[b]);
        if (lane_id() == 0 && (e || l)) {
            atomicAdd(&out[2 * b], (unsigned long long)e);
            atomicAdd(&out[2 * b + 1], (unsigned long long)l);
        }
    }
}
}  // namespace

// Per-launch profile of the plain count as tuned: for each bucket (NHASH hashed buckets, the binary-search bucket, and
// the flat kernel in slot NHASH + 1 when it takes the light rows) its work entries, probe elements and the mean device
// time of its launch over `reps` counts (HIP events between the launches on the library stream). The count itself is
// verified against the first one.
extern "C" int ga_tc_bucket_profile(ga_ctx* ctx, ga_graph* g, int reps, uint64_t* entries7, uint64_t* elems7, float* ms7) {
    if (!ctx || !g || !entries7 || !elems7 || !ms7 || reps < 1) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(ga_tc_ensure_plan(ctx, g, true));
    const CountView V = count_view(g);
    const TcPlan& P = *V.P;
    BucketStarts bs{};
    for (int b = 0; b < NHASH; b++) bs.s[b + 1] = bs.s[b] + P.nitems[b];
    bs.s[NHASH + 1] = bs.s[NHASH] + P.nhuge;
    unsigned long long* d = (unsigned long long*)ctx->dsmall;   // 64 words: [0, 2 * (NHASH + 1)) work, [32] the count
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)d, 2 * (NHASH + 1), 0));
    if (bs.s[NHASH + 1]) {
        k_bucket_work<<<grid_for(bs.s[NHASH + 1], 256, 4096), 256, 0, ctx->stream>>>((const uint4*)P.items, P.wpl, bs, d);
        GA_TRY(ga_launch_check(ctx, "k_bucket_work"));
    }
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d, 2 * (NHASH + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int b = 0; b <= NHASH; b++) {
        entries7[b] = ctx->hstat[2 * b];
        elems7[b] = ctx->hstat[2 * b + 1];
    }
    entries7[NHASH + 1] = elems7[NHASH + 1] = 0;
    hipEvent_t ev[NHASH + 4] = {};
    int rc = GA_OK;
    for (int i = 0; i < NHASH + 4 && rc == GA_OK; i++)
        if (hipEventCreate(&ev[i]) != hipSuccess) rc = ga_fail(ctx, GA_ERR_HIP, "hipEventCreate failed");
    double acc[NHASH + 2] = {};
    uint64_t T0 = 0;
    TcArgs a{};
    a.rowptr = V.rowptr;
    a.col = V.kcol;
    a.wcol = P.wcol;
    a.wedge = P.wedge;
    a.wpl = P.wpl;
    a.total = d + 32;
    const bool flat = flat_on(V);
    for (int r = 0; r <= reps && rc == GA_OK; r++) {   // r = 0: untimed warm-up
        rc = ga_fill_u64(ctx, (uint64_t*)(d + 32), 1, 0);
        if (rc == GA_OK && flat) rc = hipEventRecord(ev[NHASH + 2], ctx->stream) == hipSuccess ? GA_OK : GA_ERR_HIP;
        if (rc == GA_OK && flat) rc = launch_flat(ctx, g, V, d + 32, nullptr, nullptr, nullptr, 0);
        if (rc == GA_OK && flat) rc = hipEventRecord(ev[NHASH + 3], ctx->stream) == hipSuccess ? GA_OK : GA_ERR_HIP;
        if (rc == GA_OK) rc = launch_hash<0>(ctx, g, V, a, flat ? 3 : 0, ev);
        if (rc == GA_OK && hipMemcpyAsync(ctx->hstat, d + 32, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) rc = GA_ERR_HIP;
        if (rc == GA_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = GA_ERR_HIP;
        if (rc != GA_OK) break;
        if (r == 0) T0 = ctx->hstat[0];
        else if (ctx->hstat[0] != T0) rc = ga_fail(ctx, GA_ERR_STATE, "count changed between profiled runs");
        for (int b = 0; r > 0 && rc == GA_OK && b <= NHASH; b++) {
            float x = 0;
            if (hipEventElapsedTime(&x, ev[b], ev[b + 1]) != hipSuccess) rc = GA_ERR_HIP;
            acc[b] += x;
        }
        if (r > 0 && rc == GA_OK && flat) {
            float x = 0;
            if (hipEventElapsedTime(&x, ev[NHASH + 2], ev[NHASH + 3]) != hipSuccess) rc = GA_ERR_HIP;
            acc[NHASH + 1] += x;
        }
    }
    for (int i = 0; i < NHASH + 4; i++)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    GA_TRY(rc);
    for (int b = 0; b < NHASH + 2; b++) ms7[b] = (float)(acc[b] / reps);
    return GA_OK;
}

extern "C" int ga_tc_default_tuning(ga_tc_tuning* out) {
    if (!out) return GA_ERR_ARG;
    *out = default_tuning();
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
    if (t->flat_chunk && flat_shmem((uint32_t)t->flat_chunk) > ctx->lds_max)
        return ga_fail(ctx, GA_ERR_ARG, "flat chunk %d does not fit the workgroup's LDS", t->flat_chunk);
    // each plan's arrays are rebuilt, on its next use, only when a plan parameter differs from the ones they were built
    // with (a tuning that comes back after a trial reuses them)
    g->plan.tune = *t;
    g->cplan.tune = *t;
    return GA_OK;
}

// One timed hash count: plain, or per graph of a disjoint union (gof: graph of each vertex; gcnt: 8 * ngraphs + 1 device
// words, the last one the total) as ga_batch_count runs it.
static int timed_count(ga_ctx* ctx, ga_graph* g, const uint32_t* gof, unsigned long long* gcnt, uint32_t ngraphs, uint64_t* T, float* ms) {
    if (!gcnt) return ga_tc_count(ctx, g, GA_TC_HASH, T, ms);
    GA_TRY(ga_tc_ensure_plan(ctx, g, true));
    // as ga_batch_count: one copy of the counters per XCD (gcnt holds 8 * ngraphs + 1 words), folded after the launch
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)gcnt, (size_t)ngraphs * 8 + 1, 0));
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(ga_tc_launch(ctx, g, GA_TC_HASH, gcnt + (size_t)ngraphs * 8, nullptr, nullptr, nullptr, nullptr, gof, gcnt, ngraphs));
    GA_TRY(ga_tc_fold_xcd(ctx, gcnt, ngraphs));
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    // per-graph counts leave the total to their counters: it is their sum
    k_sum_u64<<<grid_for(ngraphs, 256, 256), 256, 0, ctx->stream>>>(gcnt, ngraphs, gcnt + (size_t)ngraphs * 8);
    GA_TRY(ga_launch_check(ctx, "k_sum_u64"));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, gcnt + (size_t)ngraphs * 8, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *T = ctx->hstat[0];
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}

extern "C" int ga_tc_tune(ga_ctx* ctx, ga_graph* g, ga_tc_tuning* best, float* best_ms) {
    return ga_tc_tune_counts(ctx, g, nullptr, nullptr, 0, best, best_ms);
}

int ga_tc_tune_counts(ga_ctx* ctx, ga_graph* g, const uint32_t* gof, unsigned long long* gcnt, uint32_t ngraphs, ga_tc_tuning* best,
                      float* best_ms) {
    if (!ctx || !g) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    ga_tc_tuning base = g->plan.tune.block ? g->plan.tune : default_tuning();
    GA_TRY(ga_tc_set_tuning(ctx, g, &base));
    uint64_t T0 = 0;
    float ms = 0;
    GA_TRY(timed_count(ctx, g, gof, gcnt, ngraphs, &T0, &ms));
    ga_tc_tuning bt = base;
    float bms = 1e30f;
    // timed counts per trial (the best one is kept): about 15 ms of counting, at least 3 and at most 16, so that a trial
    // on a graph counted in about a millisecond is not decided by one slow launch
    const int reps = std::max(3, std::min(16, (int)std::ceil(15.0f / std::max(ms, 0.01f))));
    auto trial = [&](const ga_tc_tuning& t) -> int {
        GA_TRY(ga_tc_set_tuning(ctx, g, &t));
        float best_here = 1e30f;
        // r = 0 is an untimed warm-up: after a plan rebuild the device has idled on the host's work, and the first
        // counts can run slower than the rest (a trial timed on them alone once kept 8-lane groups for d+(h) <= 128 on
        // the 34.7M-edge graph: 1.58 ms against 1.31 ms)
        for (int r = 0; r <= reps; r++) {
            uint64_t T = 0;
            float x = 0;
            GA_TRY(timed_count(ctx, g, gof, gcnt, ngraphs, &T, &x));
            if (T != T0) return ga_fail(ctx, GA_ERR_STATE, "tuning trial changed the count (%llu vs %llu)", (unsigned long long)T,
                                        (unsigned long long)T0);
            if (r > 0) best_here = std::min(best_here, x);
        }
#ifdef GA_TUNE_TRACE   // build-time switch: print every trial of the coordinate search
        fprintf(stderr, "tune trial lanes %d %d %d %d %d block %d chunk %d batches %d shift %d flat %d/%d order %d relabel %d probe %d waves %d: %.4f ms\n",
                t.lanes[0], t.lanes[1], t.lanes[2], t.lanes[3], t.lanes[4], t.block, t.work_chunk, t.batches, t.load_shift, t.flat_chunk,
                t.flat_block, t.item_order, t.relabel, t.probe, t.waves, best_here);
#endif
        if (best_here < bms) { bms = best_here; bt = t; }
        return GA_OK;
    };
    // search: relabel, item order, block size and occupancy cap jointly, then one at a time the probe form, lanes per
    // bucket, block and cap again, load shift, batches and work chunk.
    // The renumbered copy and its plan need about 60 bytes per edge while they are built (sorts included), and its
    // line-aligned copy about 6 more: it is tried only when that fits in half of the free device memory (a 2.6B-edge
    // batch union does not).
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const bool relabel_fits = g->crow != nullptr || 66ull * g->m + 20ull * g->n < free_b / 2;
    GA_TRY(trial(base));
    // The lane widths that win at the starting block size and load factor need not win at the block size and load
    // factor found after them (from block 256 and load shift 2, 8-lane groups for d+(h) <= 128 led and were kept at block
    // 128 and shift 3: 1.58 ms against 1.31 ms on the 34.7M-edge graph). So when the first pass moves the block size or
    // the load factor, a second pass tries the lane widths, block sizes and load factors again. Every trial that changes
    // the plan rebuilds it (1-3 s on a 2.6B-edge batch union), so the second pass runs only after a first pass of less
    // than 20 s.
    const double t_first = wall_ms();
    for (int pass = 0; pass < 2; pass++) {
        const ga_tc_tuning start = bt;
        // item order, block size and occupancy cap together: they decide which items are in flight at once and so
        // what the L2 holds, and a search one at a time stops short (on the 4,096-graph batch it kept row order, block
        // 512 and no cap at 93 ms; shape order, block 512 and 6 waves take 89 ms)
        const int blocks[3] = {128, 256, 512};
        const int wv[4] = {0, 7, 6, 5};   // waves per SIMD (0: no cap)
        auto order_block_cap = [&](bool relabel_too) -> int {
            for (int rl = 0; rl <= (relabel_too && relabel_fits ? 2 : 0); rl++)
                for (int io = 0; io <= 2; io++)
                    for (int bi = 0; bi < 3; bi++)
                        for (int wi = 0; wi < 4; wi++) {
                            ga_tc_tuning t = bt;
                            t.relabel = relabel_too ? rl : bt.relabel;
                            t.item_order = io;
                            t.block = blocks[bi];
                            t.waves = wv[wi];
                            if (t.relabel == bt.relabel && t.item_order == bt.item_order && t.block == bt.block && t.waves == bt.waves)
                                continue;
                            GA_TRY(trial(t));
                        }
            return GA_OK;
        };
        if (pass == 0) GA_TRY(order_block_cap(true));
        if (pass == 0) {   // how the lane groups read a batch's probe lists: lane slots or length classes
            ga_tc_tuning t = bt;
            t.probe = 1 - bt.probe;
            GA_TRY(trial(t));
        }
        {   // the bucket launches on one stream or overlapped on four
            ga_tc_tuning t = bt;
            t.streams = 1 - bt.streams;
            GA_TRY(trial(t));
        }
        GA_TRY(ga_tc_set_tuning(ctx, g, &bt));
        GA_TRY(ga_tc_ensure_plan(ctx, g, true));
        const CountView V = count_view(g);
        const int lane_opts[6] = {4, 8, 16, 32, 64, 1};   // 1: bucket 0 only (one lane per item)
        for (int b = 0; b < NHASH; b++) {
            if (!V.P->nitems[b]) continue;
            if (b < 3 && flat_on(V)) continue;   // light rows go to the flat kernel
            for (int li = 0; li < (b == 0 ? 6 : 5); li++) {
                ga_tc_tuning t = bt;
                t.lanes[b] = lane_opts[li];
                if (t.lanes[b] == bt.lanes[b]) continue;
                if (b >= 3 && t.lanes[b] < 16) continue;   // large tables per few lanes would not fit LDS
                GA_TRY(trial(t));
            }
        }
        // block size and occupancy cap again under the lane widths found (the cap's steps depend on the block: whole
        // workgroups per CU)
        for (int bi = 0; bi < 3; bi++)
            for (int wi = 0; wi < 4; wi++) {
                ga_tc_tuning t = bt;
                t.block = blocks[bi];
                t.waves = wv[wi];
                if (t.block != bt.block || t.waves != bt.waves) GA_TRY(trial(t));
            }
        for (int ls = 1; ls <= 3; ls++) { ga_tc_tuning t = bt; t.load_shift = ls; if (ls != bt.load_shift) GA_TRY(trial(t)); }
        if (pass == 1) break;
        // batches and work chunk together (both cut long rows into items: on the 4,096-graph batch the best chunk
        // depended on the batches)
        const int bat[3] = {2, 8, 32};
        const int chunks[3] = {2048, 8192, 32768};
        for (int bi = 0; bi < 3; bi++)
            for (int ci = 0; ci < 3; ci++) {
                ga_tc_tuning t = bt;
                t.batches = bat[bi];
                t.work_chunk = chunks[ci];
                if (t.batches != bt.batches || t.work_chunk != bt.work_chunk) GA_TRY(trial(t));
            }
        if (V.P->nlight) {
            // 4 or 8 units per thread (16 keeps too many records and loads in registers)
            const int fsh[6][2] = {{0, 256}, {1024, 256}, {2048, 512}, {2048, 256}, {4096, 512}, {4096, 1024}};
            for (int ci = 0; ci < 6; ci++) {
                ga_tc_tuning t = bt;
                t.flat_chunk = fsh[ci][0];
                t.flat_block = fsh[ci][1];
                if ((t.flat_chunk == bt.flat_chunk && t.flat_block == bt.flat_block) ||
                    (t.flat_chunk && flat_shmem((uint32_t)t.flat_chunk) > ctx->lds_max))
                    continue;
                GA_TRY(trial(t));
            }
        }
        // once more with the batches and work chunk found: shape order gained on the batch union only at the larger
        // work chunks (row order, block 512, no cap led at 2,048 per item)
        if (pass == 0) GA_TRY(order_block_cap(false));
        if ((bt.block == start.block && bt.load_shift == start.load_shift) || wall_ms() - t_first > 20000.0) break;
    }
    GA_TRY(ga_tc_set_tuning(ctx, g, &bt));
    GA_TRY(ga_tc_ensure_plan(ctx, g, true));
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
