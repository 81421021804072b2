// ga_truss.hip - k-truss decomposition on the GPU (edge-support peeling). The reference has no k-truss; it extends the
// per-edge triangle supports (dsg.h:62-94) the way the (k,triangle)-core extends per-vertex participation, and uses the
// same level rule as the core loops (dsg.h:694-751): level = max(level, least support of a present edge); every
// present edge whose support is at most the level leaves in this round with truss = level + 2.
//
// Set-up: per-edge supports from the hash count (F_SUP); an undirected adjacency whose entries carry the oriented edge
// id (keys (u << nb | w) sorted with the edge id as value, rows by counts + scan); for every vertex of degree > 32 an
// open-addressing membership table {neighbour -> position in the adjacency} with load <= 1/2, so that testing a
// neighbour of the longer list costs about one or two loads instead of a binary search.
// The frontier of a round at an unchanged level is exactly the set of edges whose support a decrement of the round
// before took from level + 1 to level (supports fall by one at a time and atomicSub returns the old value, so each
// such edge is appended once, by that decrement). Only when that set is empty does the level rise, and only then is the
// list of present edges scanned: for the least support, then for the edges at it.
// One round, all on the device:
//   k_tr_min      level rise only: least support over the present list (one atomicMin per wave)
//   k_tr_find     level rise only: present edges at that support -> state 2, appended to the frontier
//   k_tr_update   one wave per frontier edge (u, v): walks the shorter of N(u), N(v) (chunks past the first 256
//                 elements go to a queue that k_tr_update_chunks spreads over waves) and binary-searches each w in the
//                 other; for a triangle whose two other edges are present: both present and not in the frontier ->
//                 both lose one; one of them in the frontier -> the other loses one, decremented only by the frontier
//                 edge of smaller id; both in the frontier -> nothing (the triangle is gone with all of them). A
//                 decrement from level + 1 appends the edge to the next round's frontier
//   k_tr_retire   frontier edges -> state 0, truss = level + 2; next frontier -> state 2
//   k_tr_compact  when the present edges are at most half the list: they are copied to the other list
//   k_tr_end      present count, level, round count, buffer switches, sanity, done flag
// Rounds are enqueued in blocks that double up to 64 with one readback per block; every kernel returns at once when
// done is set (and the two scans unless the level rises), so rounds past the end cost only their launches.
#include "ga_internal.hpp"

namespace {

struct DevBuf {   // scoped device allocation
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

inline int bits_for(uint32_t n) {
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) < (uint64_t)n) b++;
    return b;
}

}  // namespace

struct TrussDev {
    unsigned long long present;   // edges still present (state 1 or 2)
    uint32_t level;               // support level of the last round (truss = level + 2)
    uint32_t min_rest;            // least support of the present edges (level-rise rounds)
    uint32_t nf;                  // frontier size of this round
    uint32_t nnext;               // next round's frontier, appended by this round's decrements
    uint32_t rise;                // 1: this round raises the level (the next frontier came out empty)
    uint32_t fcur;                // which frontier buffer holds this round's frontier
    uint32_t rounds;
    uint32_t done;
    uint32_t error;               // 1: empty frontier with edges present, 2: frontier larger than the present set,
                                  // 3: a compacted list that does not hold exactly the present edges
    uint32_t np;                  // length of the present list
    uint32_t ident;               // 1: the present list is the identity (no compaction yet)
    uint32_t cur;                 // which of the two list buffers is current
    uint32_t np_next;             // length of the list being compacted into
    uint32_t compact;             // 1: this round compacts
    unsigned long long nchunk;    // queued chunks of long shorter lists (this round)
};

// support level of this round: the least present support when the level rises, else the last round's
__device__ __forceinline__ uint32_t round_level(const TrussDev* st) { return st->rise ? st->min_rest : st->level; }

struct ga_truss {
    ga_ctx* ctx = nullptr;
    ga_graph* g = nullptr;
    uint32_t* sup = nullptr;      // current support per oriented edge
    uint32_t* truss = nullptr;
    uint32_t* frontier[2] = {nullptr, nullptr};   // this round's frontier and the next one (swapped each round)
    uint32_t* plist[2] = {nullptr, nullptr};   // lists of present edges (compacted when half are gone)
    uint2* chunks = nullptr;                   // queue of (frontier edge, start) chunks of long shorter lists
    uint8_t* state = nullptr;     // 0 removed, 1 present, 2 frontier of this round
    uint32_t* urow = nullptr;     // undirected adjacency with edge ids
    uint32_t* uadj = nullptr;
    uint32_t* ueid = nullptr;
    unsigned long long* toff = nullptr;   // per vertex: offset of its membership table (vertices of degree > 32)
    uint8_t* tbits = nullptr;             // per vertex: log2 of its table size, 0 = no table (binary search)
    uint2* tables = nullptr;              // {neighbour, position in uadj}; key 0xFFFFFFFF = empty
    TrussDev* st = nullptr;
    uint32_t kmax = 0, rounds = 0;
};

namespace {

__global__ void k_tr_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ col, uint32_t m, int nb,
                          unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = src[e], v = col[e];
        keys[2 * e] = ((unsigned long long)u << nb) | v;
        keys[2 * e + 1] = ((unsigned long long)v << nb) | u;
        vals[2 * e] = (uint32_t)e;
        vals[2 * e + 1] = (uint32_t)e;
    }
}

__global__ void k_tr_unpack(const unsigned long long* __restrict__ keys, uint64_t cnt, int nb, uint32_t* __restrict__ adj,
                            uint32_t* __restrict__ rowcnt) {
    const unsigned long long mask = (1ull << nb) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long x = keys[i];
        adj[i] = (uint32_t)(x & mask);
        atomicAdd(&rowcnt[(uint32_t)(x >> nb)], 1u);
    }
}

constexpr uint32_t TR_HASH_MIN = 32;   // a vertex with more neighbours gets a membership table
constexpr uint32_t TR_EMPTY = 0xFFFFFFFFu;   // empty table slot (no vertex id: ids are < n <= 2^32 - 1)

__device__ __forceinline__ uint32_t tr_hash(uint32_t w, uint32_t bits) { return (w * 2654435769u) >> (32u - bits); }

// Table size per vertex: the power of two >= 2 x degree (load <= 1/2) for degrees in (32, 2^30], else none.
__global__ void k_tr_tsize(const uint32_t* __restrict__ urow, uint32_t n, unsigned long long* __restrict__ tsz, uint8_t* __restrict__ tbits) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += gridDim.x * blockDim.x) {
        const uint32_t d = v < n ? urow[v + 1] - urow[v] : 0u;
        uint32_t b = 0;
        if (d > TR_HASH_MIN && d <= (1u << 30)) b = 64u - (uint32_t)__clzll(2ull * d - 1ull);
        tsz[v] = b ? (1ull << b) : 0ull;
        if (v < n) tbits[v] = (uint8_t)b;
    }
}

// Inserts every adjacency entry of a vertex with a table; the sorted keys still carry the row (u << nb | w).
__global__ void k_tr_tinsert(const unsigned long long* __restrict__ keys, uint64_t cnt, int nb, const unsigned long long* __restrict__ toff,
                             const uint8_t* __restrict__ tbits, uint2* __restrict__ tab) {
    const unsigned long long mask = (1ull << nb) - 1ull;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < cnt; p += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long x = keys[p];
        const uint32_t u = (uint32_t)(x >> nb), w = (uint32_t)(x & mask);
        const uint32_t b = tbits[u];
        if (!b) continue;
        const unsigned long long base = toff[u], smask = (1ull << b) - 1ull;
        unsigned long long sl = tr_hash(w, b);
        for (;;) {   // at most half the slots are taken, so an empty one is always reached
            if (atomicCAS(&tab[base + sl].x, TR_EMPTY, w) == TR_EMPTY) {
                tab[base + sl].y = (uint32_t)p;
                break;
            }
            sl = (sl + 1ull) & smask;
        }
    }
}

__global__ void k_tr_init(TrussDev* st, uint8_t* __restrict__ state, uint32_t m) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) state[e] = 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->present = m;
        st->level = 0;
        st->min_rest = 0xFFFFFFFFu;
        st->nf = 0;
        st->nnext = 0;
        st->rise = 1;
        st->fcur = 0;
        st->rounds = 0;
        st->done = m == 0;
        st->error = 0;
        st->np = m;
        st->ident = 1;
        st->cur = 0;
        st->np_next = 0;
        st->compact = 0;
        st->nchunk = 0;
    }
}

struct Bufs {
    uint32_t* p[2];
};

// Level rise: least support over the present edges of the list.
__global__ void k_tr_min(const uint32_t* __restrict__ sup, const uint8_t* __restrict__ state, Bufs L, TrussDev* st) {
    if (st->done || !st->rise) return;
    const uint32_t np = st->np;
    const uint32_t* list = L.p[st->cur];
    const bool ident = st->ident;
    uint32_t mn = 0xFFFFFFFFu;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        const uint32_t e = ident ? i : list[i];
        if (state[e] == 1) mn = min(mn, sup[e]);
    }
    mn = wave_min_u32(mn);
    if (lane_id() == 0 && mn != 0xFFFFFFFFu) atomicMin(&st->min_rest, mn);
}

// Level rise: the present edges at the least support form the frontier.
__global__ void k_tr_find(const uint32_t* __restrict__ sup, uint8_t* __restrict__ state, Bufs L, Bufs F, TrussDev* st) {
    if (st->done || !st->rise) return;
    const uint32_t level = st->min_rest, np = st->np;
    const uint32_t* list = L.p[st->cur];
    uint32_t* frontier = F.p[st->fcur];
    const bool ident = st->ident;
    const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = (gridDim.x * blockDim.x) / 64;
    for (uint32_t base = wave0 * 64; base < np; base += nwaves * 64) {   // wave-uniform loop
        const uint32_t i = base + lane_id();
        bool f = false;
        uint32_t e = 0;
        if (i < np) {
            e = ident ? i : list[i];
            f = state[e] == 1 && sup[e] <= level;
        }
        const uint64_t mask = __ballot(f);
        if (!mask) continue;
        uint32_t at = 0;
        if (lane_id() == 0) at = atomicAdd(&st->nf, (uint32_t)__popcll(mask));
        at = __shfl(at, 0, 64);
        if (f) {
            state[e] = 2;
            frontier[at + mask_rank(mask)] = e;
        }
    }
}

// One support decrement; the one from level + 1 puts the edge in the next frontier.
__device__ __forceinline__ void tr_dec(uint32_t* __restrict__ sup, uint32_t x, uint32_t level, uint32_t* __restrict__ next, TrussDev* st) {
    const uint32_t old = atomicSub(&sup[x], 1u);
    if (old == level + 1u) next[atomicAdd(&st->nnext, 1u)] = x;
}

// Position of w in the longer list [bb, be) of vertex lv, or be: its membership table when it has one, else a binary
// search.
struct TrTables {
    const unsigned long long* toff;
    const uint8_t* tbits;
    const uint2* tab;
};

__device__ __forceinline__ uint32_t tr_find(uint32_t w, uint32_t bb, uint32_t be, uint32_t tb, unsigned long long base,
                                            const uint2* __restrict__ tab, const uint32_t* __restrict__ uadj) {
    if (tb) {
        const unsigned long long smask = (1ull << tb) - 1ull;
        unsigned long long sl = tr_hash(w, tb);
        for (;;) {   // at most half the slots are taken: an empty slot ends every miss
            const uint2 t = tab[base + sl];
            if (t.x == w) return t.y;
            if (t.x == TR_EMPTY) return be;
            sl = (sl + 1ull) & smask;
        }
    }
    const uint32_t q = lower_bound_u32(uadj, bb, be, w);
    return (q < be && uadj[q] == w) ? q : be;
}

// The triangles that frontier edge e closes through the part [p0, p1) of its shorter list [ab, ae), the longer one being
// [bb, be) of vertex lv; one wave, lanes over consecutive elements.
__device__ __forceinline__ void tr_walk(uint32_t e, uint32_t p0, uint32_t p1, uint32_t bb, uint32_t be, uint32_t lv, TrTables T,
                                        const uint32_t* __restrict__ uadj, const uint32_t* __restrict__ ueid,
                                        const uint8_t* __restrict__ state, uint32_t* __restrict__ sup, uint32_t level,
                                        uint32_t* __restrict__ next, TrussDev* st) {
    const uint32_t tb = T.tbits[lv];
    const unsigned long long base = tb ? T.toff[lv] : 0ull;
    for (uint32_t p = p0 + lane_id(); p < p1; p += 64) {
        const uint32_t ea = ueid[p];
        const uint32_t sa = state[ea];
        if (sa == 0 || ea == e) continue;
        const uint32_t w = uadj[p];
        const uint32_t q = tr_find(w, bb, be, tb, base, T.tab, uadj);
        if (q == be) continue;
        const uint32_t eb = ueid[q];
        const uint32_t sb = state[eb];
        if (sb == 0) continue;
        if (sa == 1 && sb == 1) {
            tr_dec(sup, ea, level, next, st);
            tr_dec(sup, eb, level, next, st);
        } else if (sa == 2 && sb == 1) {
            if (e < ea) tr_dec(sup, eb, level, next, st);
        } else if (sa == 1 && sb == 2) {
            if (e < eb) tr_dec(sup, ea, level, next, st);
        }
    }
}

__device__ __forceinline__ void tr_lists(uint32_t e, const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                                         const uint32_t* __restrict__ urow, uint32_t& ab, uint32_t& ae, uint32_t& bb, uint32_t& be,
                                         uint32_t& lv) {
    const uint32_t a = src[e], b = col[e];
    ab = urow[a]; ae = urow[a + 1]; bb = urow[b]; be = urow[b + 1];
    lv = b;
    if (ae - ab > be - bb) {   // stream the shorter list, search the longer
        uint32_t t = ab; ab = bb; bb = t;
        t = ae; ae = be; be = t;
        lv = a;
    }
}


constexpr uint32_t TR_CHUNK = 256;   // shorter-list elements one wave walks per work unit

// One wave per frontier edge; frontier entries are taken wave by wave over a fixed grid. A shorter list longer than
// TR_CHUNK: the wave walks the first chunk and queues the others (edge, start) for k_tr_update_chunks, so the long
// lists of a few frontier edges spread over many waves. A full queue: the wave walks the rest itself.
__global__ void __launch_bounds__(256) k_tr_update(const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                                                   const uint32_t* __restrict__ urow, const uint32_t* __restrict__ uadj,
                                                   const uint32_t* __restrict__ ueid, const uint8_t* __restrict__ state,
                                                   uint32_t* __restrict__ sup, Bufs F, uint2* __restrict__ chunks, uint32_t chunk_cap,
                                                   TrTables T, TrussDev* st) {
    if (st->done) return;
    const uint32_t nf = st->nf, level = round_level(st);
    const uint32_t* frontier = F.p[st->fcur];
    uint32_t* next = F.p[st->fcur ^ 1u];
    const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = (gridDim.x * blockDim.x) / 64;
    for (uint32_t i = wave0; i < nf; i += nwaves) {
        const uint32_t e = frontier[i];
        uint32_t ab, ae, bb, be, lv;
        tr_lists(e, src, col, urow, ab, ae, bb, be, lv);
        uint32_t end = ae;
        if (ae - ab > TR_CHUNK) {
            const uint32_t extra = (ae - ab - 1) / TR_CHUNK;   // chunks after the first
            unsigned long long at = 0;
            if (lane_id() == 0) at = atomicAdd(&st->nchunk, (unsigned long long)extra);
            at = __shfl(at, 0, 64);
            if (at + extra <= chunk_cap) {
                for (uint32_t c = lane_id(); c < extra; c += 64) chunks[at + c] = make_uint2(e, ab + (c + 1) * TR_CHUNK);
                end = ab + TR_CHUNK;
            }
        }
        tr_walk(e, ab, end, bb, be, lv, T, uadj, ueid, state, sup, level, next, st);
    }
}

// One wave per queued chunk of a long shorter list.
__global__ void __launch_bounds__(256) k_tr_update_chunks(const uint32_t* __restrict__ src, const uint32_t* __restrict__ col,
                                                          const uint32_t* __restrict__ urow, const uint32_t* __restrict__ uadj,
                                                          const uint32_t* __restrict__ ueid, const uint8_t* __restrict__ state,
                                                          uint32_t* __restrict__ sup, Bufs F, const uint2* __restrict__ chunks,
                                                          uint32_t chunk_cap, TrTables T, TrussDev* st) {
    if (st->done) return;
    const uint32_t nc = (uint32_t)min(st->nchunk, (unsigned long long)chunk_cap), level = round_level(st);
    uint32_t* next = F.p[st->fcur ^ 1u];
    const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = (gridDim.x * blockDim.x) / 64;
    for (uint32_t i = wave0; i < nc; i += nwaves) {
        const uint2 c = chunks[i];
        uint32_t ab, ae, bb, be, lv;
        tr_lists(c.x, src, col, urow, ab, ae, bb, be, lv);
        tr_walk(c.x, c.y, min(c.y + TR_CHUNK, ae), bb, be, lv, T, uadj, ueid, state, sup, level, next, st);
    }
}

// Retires this round's frontier and marks the next one (after the update, so the owner rule saw this round's states).
__global__ void k_tr_retire(uint8_t* __restrict__ state, uint32_t* __restrict__ truss, Bufs F, TrussDev* st) {
    if (st->done) return;
    const uint32_t nf = st->nf, nn = st->nnext, t = round_level(st) + 2u;
    const uint32_t* frontier = F.p[st->fcur];
    const uint32_t* next = F.p[st->fcur ^ 1u];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) {
        const uint32_t e = frontier[i];
        state[e] = 0;
        truss[e] = t;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) state[next[i]] = 2;
    // compact this round when what stays present is at most half the list
    if (blockIdx.x == 0 && threadIdx.x == 0) st->compact = (st->present - nf) * 2 <= st->np && st->present > nf;
}

// Copies the present edges (state 1 or 2) of the current list to the other one.
__global__ void k_tr_compact(const uint8_t* __restrict__ state, Bufs L, TrussDev* st) {
    if (st->done || !st->compact) return;
    const uint32_t np = st->np;
    const uint32_t* list = L.p[st->cur];
    uint32_t* out = L.p[st->cur ^ 1u];
    const bool ident = st->ident;
    const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) / 64, nwaves = (gridDim.x * blockDim.x) / 64;
    for (uint32_t base = wave0 * 64; base < np; base += nwaves * 64) {   // wave-uniform loop
        const uint32_t i = base + lane_id();
        uint32_t e = 0;
        bool k = false;
        if (i < np) {
            e = ident ? i : list[i];
            k = state[e] != 0;
        }
        const uint64_t mask = __ballot(k);
        if (!mask) continue;
        uint32_t at = 0;
        if (lane_id() == 0) at = atomicAdd(&st->np_next, (uint32_t)__popcll(mask));
        at = __shfl(at, 0, 64);
        if (k) out[at + mask_rank(mask)] = e;
    }
}

__global__ void k_tr_end(TrussDev* st) {
    if (st->done) return;
    const uint32_t nf = st->nf;
    if (nf == 0 || nf > st->present) {
        st->error = nf == 0 ? 1u : 2u;
        st->done = 1;
        return;
    }
    if (st->compact && st->np_next != st->present - nf) {
        st->error = 3;
        st->done = 1;
        return;
    }
    st->level = round_level(st);
    st->present -= nf;
    st->rounds++;
    st->nf = st->nnext;
    st->nnext = 0;
    st->fcur ^= 1u;
    st->rise = st->nf == 0;
    st->min_rest = 0xFFFFFFFFu;
    st->nchunk = 0;
    if (st->compact) {
        st->np = st->np_next;
        st->np_next = 0;
        st->cur ^= 1u;
        st->ident = 0;
        st->compact = 0;
    }
    if (st->present == 0) st->done = 1;
}

}  // namespace

extern "C" int ga_truss_create(ga_ctx* ctx, ga_graph* g, ga_truss** out) {
    if (!out) return ctx ? ga_fail(ctx, GA_ERR_ARG, "out is NULL") : GA_ERR_ARG;
    *out = nullptr;
    if (!ctx || !g) return GA_ERR_ARG;
    if (2ull * g->m >= 0xFFFFFFFFull)
        return ga_fail(ctx, GA_ERR_ARG, "graph too large for the k-truss decomposition (2m must fit 32-bit adjacency positions)");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    ga_truss* t = new ga_truss;
    t->ctx = ctx;
    t->g = g;
    const size_t m = g->m, n = g->n;
    if (ga_malloc(ctx, (void**)&t->sup, (m + 1) * 4) || ga_malloc(ctx, (void**)&t->truss, (m + 1) * 4) ||
        ga_malloc(ctx, (void**)&t->frontier[0], (m + 1) * 4) || ga_malloc(ctx, (void**)&t->frontier[1], (m + 1) * 4) ||
        ga_malloc(ctx, (void**)&t->state, m + 16) || ga_malloc(ctx, (void**)&t->chunks, (m / 4 + 1024) * 8) ||
        ga_malloc(ctx, (void**)&t->plist[0], (m + 1) * 4) || ga_malloc(ctx, (void**)&t->plist[1], (m + 1) * 4) ||
        ga_malloc(ctx, (void**)&t->urow, (n + 1) * 4) || ga_malloc(ctx, (void**)&t->uadj, (2 * m + 1) * 4) ||
        ga_malloc(ctx, (void**)&t->ueid, (2 * m + 1) * 4) || ga_malloc(ctx, (void**)&t->st, sizeof(TrussDev))) {
        ga_truss_destroy(t);
        return GA_ERR_OOM;
    }
    // undirected adjacency with edge ids
    int rc = GA_OK;
    const uint64_t cnt = 2ull * m;
    const int nb = bits_for(g->n);
    DevBuf k0, k1, v1, rc_cnt;
    unsigned long long* ko = nullptr;
    if (!rc && (ga_malloc(ctx, &k0.p, std::max<uint64_t>(cnt, 1) * 8) || ga_malloc(ctx, &k1.p, std::max<uint64_t>(cnt, 1) * 8) ||
                ga_malloc(ctx, &v1.p, std::max<uint64_t>(cnt, 1) * 4) || ga_malloc(ctx, &rc_cnt.p, (n + 1) * 4)))
        rc = GA_ERR_OOM;
    if (!rc) rc = ga_fill_u32(ctx, (uint32_t*)rc_cnt.p, n + 1, 0);
    if (!rc && m) {
        k_tr_keys<<<grid_for(m, 256, 8192), 256, 0, ctx->stream>>>(g->src, g->col, g->m, nb, (unsigned long long*)k0.p, t->ueid);
        rc = ga_launch_check(ctx, "k_tr_keys");
        uint32_t* vo = nullptr;
        if (!rc) rc = ga_sort_pairs_u64(ctx, (unsigned long long*)k0.p, (unsigned long long*)k1.p, t->ueid, (uint32_t*)v1.p, cnt, 2 * nb, &ko, &vo);
        if (!rc && vo != t->ueid && hipMemcpyAsync(t->ueid, vo, cnt * 4, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            rc = ga_fail(ctx, GA_ERR_HIP, "edge id copy failed");
        if (!rc) {
            k_tr_unpack<<<grid_for(cnt, 256, 8192), 256, 0, ctx->stream>>>(ko, cnt, nb, t->uadj, (uint32_t*)rc_cnt.p);
            rc = ga_launch_check(ctx, "k_tr_unpack");
        }
    }
    if (!rc) rc = ga_exclusive_scan_u32(ctx, (const uint32_t*)rc_cnt.p, t->urow, n + 1);
    // membership tables of the vertices with more than TR_HASH_MIN neighbours
    DevBuf tsz;
    if (!rc && (ga_malloc(ctx, &tsz.p, (n + 1) * 8) || ga_malloc(ctx, (void**)&t->toff, (n + 1) * 8) ||
                ga_malloc(ctx, (void**)&t->tbits, n + 16)))
        rc = GA_ERR_OOM;
    if (!rc) {
        k_tr_tsize<<<grid_for(n + 1, 256, 8192), 256, 0, ctx->stream>>>(t->urow, g->n, (unsigned long long*)tsz.p, t->tbits);
        rc = ga_launch_check(ctx, "k_tr_tsize");
    }
    if (!rc) rc = ga_exclusive_scan_u64(ctx, (const uint64_t*)tsz.p, (uint64_t*)t->toff, n + 1);
    if (!rc && hipMemcpyAsync(ctx->hstat, t->toff + n, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = ga_fail(ctx, GA_ERR_HIP, "table size readback failed");
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ga_fail(ctx, GA_ERR_HIP, "truss set-up failed");
    const uint64_t tentries = rc ? 0 : ctx->hstat[0];
    if (!rc && ga_malloc(ctx, (void**)&t->tables, std::max<uint64_t>(tentries, 1) * 8)) rc = GA_ERR_OOM;
    if (!rc) rc = ga_fill_u64(ctx, (uint64_t*)t->tables, std::max<uint64_t>(tentries, 1), ~0ull);
    if (!rc && m && tentries) {
        k_tr_tinsert<<<grid_for(cnt, 256, 8192), 256, 0, ctx->stream>>>(ko, cnt, nb, t->toff, t->tbits, t->tables);
        rc = ga_launch_check(ctx, "k_tr_tinsert");
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ga_fail(ctx, GA_ERR_HIP, "truss set-up failed");
    if (rc) {
        ga_truss_destroy(t);
        return rc;
    }
    *out = t;
    return GA_OK;
}

extern "C" void ga_truss_destroy(ga_truss* t) {
    if (!t) return;
    if (t->ctx) {
        (void)hipSetDevice(t->ctx->device);
        (void)hipStreamSynchronize(t->ctx->stream);
    }
    ga_free(t->sup);
    ga_free(t->truss);
    ga_free(t->chunks);
    ga_free(t->frontier[0]);
    ga_free(t->frontier[1]);
    ga_free(t->plist[0]);
    ga_free(t->plist[1]);
    ga_free(t->state);
    ga_free(t->urow);
    ga_free(t->uadj);
    ga_free(t->ueid);
    ga_free(t->toff);
    ga_free(t->tbits);
    ga_free(t->tables);
    ga_free(t->st);
    delete t;
}

extern "C" int ga_truss_run(ga_truss* t, uint32_t* stats, float* ms) {
    if (!t) return GA_ERR_ARG;
    ga_ctx* ctx = t->ctx;
    ga_graph* g = t->g;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t m = g->m;
    DevBuf d;
    GA_TRY(ga_malloc(ctx, &d.p, 8));
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(ga_fill_u32(ctx, t->sup, m, 0));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)d.p, 1, 0));
    if (m) GA_TRY(ga_tc_launch(ctx, g, GA_TC_HASH, (unsigned long long*)d.p, nullptr, nullptr, t->sup, nullptr));
    k_tr_init<<<grid_for(std::max<uint32_t>(m, 1), 256, 8192), 256, 0, ctx->stream>>>(t->st, t->state, m);
    GA_TRY(ga_launch_check(ctx, "k_tr_init"));
    const unsigned scan_grid = grid_for(std::max<uint32_t>(m, 1), 256, 4u * (unsigned)std::max(ctx->cus, 1));
    const unsigned upd_grid = 4096;   // 16K waves take frontier edges in turn
    const Bufs lists = {{t->plist[0], t->plist[1]}};
    const Bufs fr = {{t->frontier[0], t->frontier[1]}};
    const uint32_t chunk_cap = m / 4 + 1024;
    const TrTables tabs = {t->toff, t->tbits, t->tables};
    int block_rounds = 1;
    for (uint64_t guard = 0;; guard++) {
        for (int r = 0; r < block_rounds; r++) {
            k_tr_min<<<scan_grid, 256, 0, ctx->stream>>>(t->sup, t->state, lists, t->st);
            k_tr_find<<<scan_grid, 256, 0, ctx->stream>>>(t->sup, t->state, lists, fr, t->st);
            k_tr_update<<<upd_grid, 256, 0, ctx->stream>>>(g->src, g->col, t->urow, t->uadj, t->ueid, t->state, t->sup, fr, t->chunks,
                                                           chunk_cap, tabs, t->st);
            k_tr_update_chunks<<<upd_grid, 256, 0, ctx->stream>>>(g->src, g->col, t->urow, t->uadj, t->ueid, t->state, t->sup, fr, t->chunks,
                                                                  chunk_cap, tabs, t->st);
            k_tr_retire<<<1024, 256, 0, ctx->stream>>>(t->state, t->truss, fr, t->st);
            k_tr_compact<<<scan_grid, 256, 0, ctx->stream>>>(t->state, lists, t->st);
            k_tr_end<<<1, 1, 0, ctx->stream>>>(t->st);
        }
        GA_TRY(ga_launch_check(ctx, "truss round"));
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, t->st, sizeof(TrussDev), hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const TrussDev* h = (const TrussDev*)ctx->hstat;
        if (h->error) return ga_fail(ctx, GA_ERR_STATE, "truss round %u: %s", h->rounds,
                                     h->error == 1 ? "empty frontier with edges present"
                                     : h->error == 2 ? "frontier larger than the present set" : "compacted list size differs from the present count");
        if (h->done) break;
        if (guard > (uint64_t)m + 2) return ga_fail(ctx, GA_ERR_STATE, "truss decomposition did not finish");
        block_rounds = std::min(block_rounds * 2, 64);
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    const TrussDev* h = (const TrussDev*)ctx->hstat;
    t->rounds = h->rounds;
    t->kmax = m ? h->level + 2 : 0;
    if (stats) {
        stats[0] = t->rounds;
        stats[1] = t->kmax;
    }
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}

extern "C" int ga_truss_download(ga_truss* t, uint32_t* truss) {
    if (!t || (t->g->m && !truss)) return GA_ERR_ARG;
    ga_ctx* ctx = t->ctx;
    if (t->g->m) GA_HIP(ctx, hipMemcpyAsync(truss, t->truss, (size_t)t->g->m * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}
