// ga_greedy.hip - Greedy++ for the triangle-density densest subgraph (Flowless Alg 2 with triangles as hyperedges;
// passes = 1 is Charikar's Greedy), reference dsg.h:176-250 (greedy_pp, greedy).
//
// The reference runs it on the graph that prune_to_core leaves (gpu_dsg.cu:344-356): hundreds to a few thousand
// vertices. Each pass deletes the vertices one at a time, so one workgroup of 1,024 lanes runs a whole Greedy++ (all
// passes, one launch) with the graph's state on the CU. Two forms:
//   k_greedy_dense  (n <= 1,120 with 160 KB of LDS) the adjacency as a bitmap in LDS, per-vertex state in registers,
//                   one barrier per deletion; also many graphs in one launch, one workgroup each
//   k_greedy_pp     (the CSR form, up to about 13,000 vertices) per-vertex state in LDS, the adjacency in HBM
// Every pass's deletion order is kept; the best set is that pass's order from the cut on.
#include "ga_internal.hpp"

namespace {

constexpr int GR_BLOCK = 1024;
constexpr int GR_WAVES = GR_BLOCK / GA_WAVE;

struct GreedyOut {
    unsigned long long bn, bd;   // best density found: triangles / vertices (0/0 = none)
    uint32_t best_pass;          // pass whose deletion order holds the best subgraph (UINT32_MAX = none)
    uint32_t best_cut;           // the best subgraph = that pass's order[best_cut .. n)
    uint32_t error;              // 1 = triangle bookkeeping inconsistent, 2 = no alive vertex found
    uint32_t pad;
};

// a/ad > b/bd exactly (ad > 0; bd == 0 means density 0), as dens_gt of dsg.h:40-46: 128-bit cross products.
__device__ __forceinline__ bool dens_gt(uint64_t an, uint64_t ad, uint64_t bn, uint64_t bd) {
    if (bd == 0) { bn = 0; bd = 1; }
    const uint64_t lh = __umul64hi(an, bd), ll = an * bd;
    const uint64_t rh = __umul64hi(bn, ad), rl = bn * ad;
    return lh > rh || (lh == rh && ll > rl);
}

// Cross-lane moves inside rows of 16 lanes with DPP (VALU latency, no LDS round trip): quad_perm [1,0,3,2] and
// [2,3,0,1] swap neighbours at distance 1 and 2, row_half_mirror and row_mirror reflect 8- and 16-lane halves. After the
// four steps every lane of a row holds the row's result; readlane combines the four rows.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ unsigned long long dpp64(unsigned long long x) {
    return ((unsigned long long)dpp32<CTRL>((uint32_t)(x >> 32)) << 32) | dpp32<CTRL>((uint32_t)x);
}

__device__ __forceinline__ unsigned long long readlane64(unsigned long long x, int l) {
    return ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), l) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
}

// Minimum over the wavefront, the same in every lane.
__device__ __forceinline__ unsigned long long wave_min_u64_dpp(unsigned long long x) {
    x = min(x, dpp64<0xB1>(x));
    x = min(x, dpp64<0x4E>(x));
    x = min(x, dpp64<0x141>(x));
    x = min(x, dpp64<0x140>(x));
    return min(min(readlane64(x, 0), readlane64(x, 16)), min(readlane64(x, 32), readlane64(x, 48)));
}

// Sum over the wavefront, the same in every lane.
__device__ __forceinline__ uint32_t wave_sum_u32_dpp(uint32_t x) {
    x += dpp32<0xB1>(x);
    x += dpp32<0x4E>(x);
    x += dpp32<0x141>(x);
    x += dpp32<0x140>(x);
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 0) + (uint32_t)__builtin_amdgcn_readlane((int)x, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)x, 32) + (uint32_t)__builtin_amdgcn_readlane((int)x, 48);
}

__device__ __forceinline__ bool bit(const uint32_t* b, uint32_t x) { return (b[x >> 5] >> (x & 31)) & 1u; }

// CSR form. LDS holds load[v] (u64, updated in place: a deleted vertex is not read again in the pass, dsg.h:184-185),
// tri_H[v] (u32, at most C(n-1, 2) < 2^32 at these sizes), the alive bitmap of H and the bitmap of the alive neighbours
// of the vertex being deleted. One deletion: argmin of (load + tri_H, id) over H (each lane scans n / 1024 vertices,
// then a wave and a workgroup reduction), the neighbour bitmap from v's undirected row, then every triangle {v, a, b}
// of H once: wave w takes v's neighbours a = w, w + 16, ... and its lanes read a's oriented out-row; b is a hit when
// its bit is set, and tri_H[a], tri_H[b] drop by one with LDS atomics. The triangles found must equal tri_H[v] (the
// reference's self-check).
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
            if (v >= n) { err = 2; break; }     // uniform: every lane reduced the same values
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

// Dense form for graphs whose adjacency bitmap fits LDS (n <= 1,120 with 160 KB): adj[x] is a row of W = n / 32 words,
// padded to an odd stride so that lanes reading consecutive rows hit different banks. A workgroup of DB lanes runs it;
// lane t owns the vertices t + k * DB (k < OWN) and keeps their load, tri_H and alive flag in registers. 1,024 lanes
// owning two vertices each beat 256 lanes owning five (14.4 vs 24.2 ms for 10 passes on a 729-vertex graph): the
// neighbours' popcount rows are the serial part, and more lanes split them finer. A vertex's key is
// (load + tri_H) << 11 | id, so one 64-bit minimum is the (key, id) order; a key is at most
// passes * C(n-1, 2) + C(n-1, 2) < 2^32 * 2^20, so it fits the 53 bits above the id.
// Before the first pass each owner takes tri(x) from the job's per-vertex count or, for a job without one (many graphs
// in one launch), counts it from the bitmap (popcount of N(x) & N(y) over x's neighbours y, halved); T = the sum / 3.
// One graph is faster with the separate count (0.19 vs 0.5 ms on a 729-vertex graph: the owners of hubs are the
// serial part), many graphs without it (one launch instead of a few count launches per graph).
// One deletion, one barrier:
//   v = the minimum published for this deletion (one LDS word, the same for every lane);
//   each owner of an alive neighbour x of v counts c = |N(x) & N(v)| with W popcounts (rows hold alive neighbours only,
//   so this is the triangles {v, x, b} of H) and drops tri_H(x) by c (every such triangle is counted at x and at b, so
//   the counts sum to 2 tri_H(v)), then clears v from its own row x; the owner of v retires it and publishes tri_H(v);
//   every wave takes the minimum of its lanes' new keys (DPP) and merges it into the next deletion's slot with one LDS
//   atomicMin; wave sums of the counts go into this deletion's counter;
//   after the barrier, the next deletion first checks this one (counts = 2 tri_H(v), no count above tri_H).
// Rows only lose bits inside a pass (v's own row is not read again once v is deleted), so every pass after the first
// rebuilds them from the CSR (m atomics, a few percent of a pass).
// The deletion order stays in registers (lane t keeps the deletions t + k * DB) and is written once per pass, so the
// loop touches no global memory (one global store per deletion measured the same).
// The minimum slots and counters rotate over three and the published tri_H over two, so each is reset one deletion
// before it is written and never written while a slower lane may still read it.

// One graph of a dense launch: workgroup b runs jobs[b].
struct GreedyJob {
    uint32_t n, m, wpad, pad;
    const uint32_t* src;
    const uint32_t* col;
    const uint32_t* tri0;             // per-vertex triangles (the dense form counts them from its bitmap when NULL)
    const unsigned long long* dT;     // T (the CSR form)
    uint32_t* orders;                 // passes * n
    unsigned long long* per_pass;     // 2 * passes
    GreedyOut* out;
};

template <int DB, int OWN>
__global__ __launch_bounds__(DB) void k_greedy_dense(const GreedyJob* __restrict__ jobs, uint32_t passes) {
    const GreedyJob& job = jobs[blockIdx.x];
    const uint32_t n = job.n, m = job.m, wpad = job.wpad;
    const uint32_t* __restrict__ src = job.src;
    const uint32_t* __restrict__ col = job.col;
    uint32_t* __restrict__ orders = job.orders;
    unsigned long long* __restrict__ per_pass = job.per_pass;
    GreedyOut* __restrict__ out = job.out;
    extern __shared__ __align__(16) uint32_t adj[];                  // n rows of wpad words
    __shared__ unsigned long long s_min[3];
    __shared__ uint32_t s_cv[2], s_found[3], s_bad;
    __shared__ unsigned long long s_T0;
    constexpr uint32_t IDB = 11;
    constexpr unsigned long long IDM = (1ull << IDB) - 1;

    const uint32_t tid = threadIdx.x, lane = tid & (GA_WAVE - 1);
    const uint32_t W = (n + 31) >> 5;
    for (uint32_t i = tid; i < n * wpad; i += DB) adj[i] = 0;
    if (tid == 0) {
        s_bad = 0;
        s_T0 = 0;
    }
    __syncthreads();
    for (uint32_t e = tid; e < m; e += DB) {
        const uint32_t a = src[e], b = col[e];
        atomicOr(&adj[a * wpad + (b >> 5)], 1u << (b & 31));
        atomicOr(&adj[b * wpad + (a >> 5)], 1u << (a & 31));
    }
    __syncthreads();
    // tri(x): from the job's per-vertex count when it has one, else from the bitmap (every triangle {x, y, z} is seen
    // twice from x, through y and through z)
    unsigned long long load[OWN];
    uint32_t tri0[OWN], cur[OWN], ord[OWN];
    bool alive[OWN];
    uint32_t tsum = 0;
#pragma unroll
    for (int k = 0; k < OWN; k++) {
        load[k] = 0;
        cur[k] = ord[k] = 0;
        alive[k] = false;
        const uint32_t x = tid + k * DB;
        uint32_t c = 0;
        if (x < n && job.tri0) {
            c = 2u * job.tri0[x];
        } else if (x < n) {
            const uint32_t* ax = adj + x * wpad;
            for (uint32_t i = 0; i < W; i++) {
                for (uint32_t bits = ax[i]; bits; bits &= bits - 1u) {
                    const uint32_t* ay = adj + (32u * i + (uint32_t)__ffs(bits) - 1u) * wpad;
#pragma unroll 8
                    for (uint32_t j = 0; j < W; j++) c += __popc(ax[j] & ay[j]);
                }
            }
        }
        tri0[k] = c / 2u;
        tsum += tri0[k];
    }
    tsum = wave_sum_u32_dpp(tsum);
    if (lane == 0 && tsum) atomicAdd(&s_T0, (unsigned long long)tsum);
    __syncthreads();
    const uint64_t T0 = s_T0 / 3u;   // each triangle counted at its three corners
    uint64_t bn = 0, bd = 0;
    uint32_t best_pass = UINT32_MAX, best_cut = 0, err = 0;
    for (uint32_t p = 0; p < passes && !err; p++) {
        if (p) {   // the rows of the previous pass lost every deleted vertex: rebuild them from the CSR
            for (uint32_t i = tid; i < n * wpad; i += DB) adj[i] = 0;
            __syncthreads();
            for (uint32_t e = tid; e < m; e += DB) {
                const uint32_t a = src[e], b = col[e];
                atomicOr(&adj[a * wpad + (b >> 5)], 1u << (b & 31));
                atomicOr(&adj[b * wpad + (a >> 5)], 1u << (a & 31));
            }
        }
        if (tid < 3) {
            s_found[tid] = 0;
            s_min[tid] = ~0ull;
        }
        __syncthreads();
        unsigned long long mine = ~0ull;
#pragma unroll
        for (int k = 0; k < OWN; k++) {
            const uint32_t x = tid + k * DB;
            alive[k] = x < n;
            cur[k] = tri0[k];
            if (x < n) mine = min(mine, ((load[k] + cur[k]) << IDB) | x);
        }
        mine = wave_min_u64_dpp(mine);
        if (lane == 0) atomicMin(&s_min[0], mine);
        __syncthreads();
        uint32_t* order = orders + (size_t)p * n;
        uint64_t T = T0;
        bool improved = false;
        uint32_t cut = 0;
        uint32_t done = 0;   // deletions of this pass
        for (uint32_t step = 0; step < n; step++) {
            if (step) {   // the previous deletion, read after the barrier: the same in every lane
                const uint32_t cv = s_cv[(step - 1) & 1];
                if (s_bad || s_found[(step - 1) % 3] != 2ull * cv) err = 1;
                T -= cv;
                if (err) break;
            }
            if (tid == 0) {
                s_found[(step + 1) % 3] = 0;
                s_min[(step + 2) % 3] = ~0ull;
            }
            const uint32_t left = n - step;
            if (dens_gt(T, left, bn, bd)) {
                bn = T;
                bd = left;
                improved = true;
                cut = step;
            }
            const uint32_t v = (uint32_t)(s_min[step % 3] & IDM);   // left > 0: an alive key was published, so v < n
            if (v >= n) { err = 2; break; }                           // uniform: every lane read the same slot
            const uint32_t* av = adj + v * wpad;
            uint32_t found = 0;
            mine = ~0ull;
#pragma unroll
            for (int k = 0; k < OWN; k++) {
                const uint32_t x = tid + k * DB;
                if (!alive[k]) continue;
                if (x == v) {
                    s_cv[step & 1] = cur[k];
                    alive[k] = false;
                    load[k] += cur[k];
                    continue;
                }
                if ((av[x >> 5] >> (x & 31)) & 1u) {
                    uint32_t* ax = adj + x * wpad;
                    uint32_t c = 0;
#pragma unroll 8
                    for (uint32_t i = 0; i < W; i++) c += __popc(ax[i] & av[i]);
                    ax[v >> 5] &= ~(1u << (v & 31));   // x's own row drops v; no other lane reads it this step
                    if (c > cur[k]) s_bad = 1;
                    cur[k] -= min(c, cur[k]);
                    found += c;
                }
                mine = min(mine, ((load[k] + cur[k]) << IDB) | x);
            }
            mine = wave_min_u64_dpp(mine);
            found = wave_sum_u32_dpp(found);
            if (lane == 0) {
                if (mine != ~0ull) atomicMin(&s_min[(step + 1) % 3], mine);
                if (found) atomicAdd(&s_found[step % 3], found);
            }
            if (step % DB == tid) {
#pragma unroll
                for (int k = 0; k < OWN; k++)
                    if (step / DB == (uint32_t)k) ord[k] = v;
            }
            done = step + 1;
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < OWN; k++)
            if (tid + k * DB < done) order[tid + k * DB] = ord[k];
        if (!err && n) {
            const uint32_t cv = s_cv[(n - 1) & 1];
            if (s_bad || s_found[(n - 1) % 3] != 2ull * cv) err = 1;
            T -= cv;
        }
        if (improved) {
            best_pass = p;
            best_cut = cut;
        }
        if (tid == 0) {
            per_pass[2 * p] = bn;
            per_pass[2 * p + 1] = bd;
        }
        __syncthreads();   // the rows, the slots and the counters are reset by the next pass
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
        if (i >= cut && order[i] < n) flags[order[i]] = 1;
}

__global__ void k_greedy_set(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, uint32_t n,
                             uint32_t* __restrict__ set) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        if (flags[v]) set[pos[v]] = v;
}

// In-set flags of many graphs: graph y (blockIdx.y + base) marks order[cut .. n) of its best pass at its offset.
__global__ void k_greedy_flags_many(const GreedyJob* __restrict__ jobs, const unsigned long long* __restrict__ voff, uint32_t base,
                                    uint8_t* __restrict__ flags) {
    const uint32_t gi = base + blockIdx.y;
    const GreedyJob& j = jobs[gi];
    const GreedyOut* o = j.out;
    if (o->best_pass == UINT32_MAX || o->bd == 0) return;
    const uint32_t* order = j.orders + (size_t)o->best_pass * j.n;
    uint8_t* f = flags + voff[gi];
    for (uint32_t i = o->best_cut + blockIdx.x * blockDim.x + threadIdx.x; i < j.n; i += gridDim.x * blockDim.x)
        if (order[i] < j.n) f[order[i]] = 1;
}

// Large form, for graphs whose per-vertex state does not fit LDS (the whole graph of gpu_dsg.cu:318-330 before the
// pruning). One workgroup still makes the deletions one at a time, in the exact (load + tri_H, id) order, but the state
// lives in HBM and the minimum comes from a two-level structure instead of a scan of all vertices:
//   bmin[b]   per block of 64 consecutive ids: min of (key << 6 | offset) over its alive vertices (HBM)
//   smin[s]   per super-block of 64 blocks (4,096 ids): min of (key << 12 | offset in the super-block) (LDS)
// A deletion takes the minimum over smin (ties: the lower super-block, so the lower id), enumerates the triangles
// {v, a, b} of H as the CSR form does (alive-neighbour bitmap of v, oriented rows of v's neighbours), and lowers the
// minima of every touched vertex with atomicMin (its key only decreases). Then one wave recomputes v's block and
// super-block. Reads of state that other waves change with atomics go past the L1 cache (agent-scope atomic loads).
// Each pass first deletes, in id order and at once, the vertices whose key is 0 (no triangle and no load): they touch
// no other vertex, and as they leave T / |H| only grows, so of their density checks only the last one can record a new
// best (the check before the next deletion sees a higher density still). Keys must fit 52 bits: T < 2^32 and at most
// 2^19 passes.
constexpr uint32_t GB_BLK = 64, GB_SUPV = 4096;
constexpr unsigned long long GB_NONE = ~0ull;

struct GreedyBig {
    uint32_t n, passes;
    const uint32_t* rowptr;
    const uint32_t* col;
    const uint32_t* urow;
    const uint32_t* ucol;
    const uint32_t* tri0;
    const unsigned long long* dT;
    uint32_t* orders;
    unsigned long long* per_pass;
    GreedyOut* out;
    unsigned long long* load;   // n
    uint32_t* cur;              // n: tri_H
    uint32_t* alive;            // n
    uint32_t* nbr;              // (n + 31) / 32: alive neighbours of the vertex being deleted
    unsigned long long* bmin;   // (n + 63) / 64
};

__device__ __forceinline__ unsigned long long gb_ld64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t gb_ld32(const uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Minimum of block b over its alive vertices other than `skip` (one wave; the same value in every lane).
__device__ __forceinline__ unsigned long long gb_block_min(const GreedyBig& J, uint32_t b, uint32_t skip) {
    const uint32_t l = lane_id(), x = b * GB_BLK + l;
    unsigned long long k = GB_NONE;
    if (x < J.n && x != skip && gb_ld32(J.alive + x)) k = ((gb_ld64(J.load + x) + gb_ld32(J.cur + x)) << 6) | l;
    return wave_min_u64_dpp(k);
}

// Minimum of super-block sb from its blocks' minima; block `fresh_b` reads `fresh` instead of bmin (one wave).
__device__ __forceinline__ unsigned long long gb_super_min(const GreedyBig& J, uint32_t sb, uint32_t fresh_b, unsigned long long fresh) {
    const uint32_t l = lane_id(), b = sb * (GB_SUPV / GB_BLK) + l, nb = (J.n + GB_BLK - 1) / GB_BLK;
    unsigned long long k = GB_NONE;
    if (b < nb) {
        const unsigned long long v = b == fresh_b ? fresh : gb_ld64(J.bmin + b);
        if (v != GB_NONE) k = ((v >> 6) << 12) | (l << 6) | (v & 63u);
    }
    return wave_min_u64_dpp(k);
}

__device__ __forceinline__ void gb_lower(const GreedyBig& J, unsigned long long* smin, uint32_t x, unsigned long long key) {
    atomicMin(J.bmin + x / GB_BLK, (key << 6) | (x % GB_BLK));
    atomicMin(smin + x / GB_SUPV, (key << 12) | (x % GB_SUPV));
}

__global__ __launch_bounds__(GR_BLOCK) void k_greedy_big(const GreedyBig J) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long* smin = (unsigned long long*)smem;
    __shared__ unsigned long long red_key[GR_WAVES], red_found[GR_WAVES];
    __shared__ uint32_t red_s[GR_WAVES], red_cnt[GR_WAVES + 1];
    __shared__ uint32_t bad;
    const uint32_t n = J.n, tid = threadIdx.x, lane = tid & (GA_WAVE - 1), wave = tid / GA_WAVE;
    const uint32_t nb = (n + GB_BLK - 1) / GB_BLK, ns = (n + GB_SUPV - 1) / GB_SUPV;
    const uint64_t T0 = *J.dT;
    for (uint32_t i = tid; i < n; i += GR_BLOCK) J.load[i] = 0;
    for (uint32_t i = tid; i < (n + 31) / 32; i += GR_BLOCK) J.nbr[i] = 0;
    if (tid == 0) bad = 0;
    uint64_t bn = 0, bd = 0;
    uint32_t best_pass = UINT32_MAX, best_cut = 0, err = 0;
    for (uint32_t p = 0; p < J.passes && !err; p++) {
        __threadfence();
        __syncthreads();
        uint32_t* order = J.orders + (size_t)p * n;
        // the vertices of key 0 (no triangle, no load) leave first, in id order; the others start alive
        uint32_t z = 0;
        for (uint32_t base = 0; base < n; base += GR_BLOCK) {
            const uint32_t x = base + tid;
            bool zero = false;
            if (x < n) {
                const uint32_t c = J.tri0[x];
                zero = c == 0 && gb_ld64(J.load + x) == 0;
                J.cur[x] = c;
                J.alive[x] = zero ? 0u : 1u;
            }
            const uint64_t m = __ballot(zero);
            if (lane == 0) red_cnt[wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t before = z;
            for (uint32_t w = 0; w < wave; w++) before += red_cnt[w];
            if (zero) order[before + mask_rank(m)] = x;
            for (uint32_t w = 0; w < GR_WAVES; w++) z += red_cnt[w];
            __syncthreads();
        }
        uint64_t T = T0;
        bool improved = false;
        uint32_t cut = 0;
        if (z && dens_gt(T, n - (z - 1), bn, bd)) {   // the last of the bulk's checks (the only one that can matter)
            bn = T;
            bd = n - (z - 1);
            improved = true;
            cut = z - 1;
        }
        __threadfence();
        __syncthreads();
        for (uint32_t b = wave; b < nb; b += GR_WAVES) {
            const unsigned long long k = gb_block_min(J, b, UINT32_MAX);
            if (lane == 0) __hip_atomic_store(J.bmin + b, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __threadfence();
        __syncthreads();
        for (uint32_t sb = wave; sb < ns; sb += GR_WAVES) {
            const unsigned long long k = gb_super_min(J, sb, UINT32_MAX, 0);
            if (lane == 0) smin[sb] = k;
        }
        __syncthreads();
        for (uint32_t step = z; step < n; step++) {
            const uint32_t left = n - step;
            if (dens_gt(T, left, bn, bd)) {   // H is evaluated before every deletion (dsg.h:208)
                bn = T;
                bd = left;
                improved = true;
                cut = step;
            }
            // the minimum (key, id): key = smin >> 12, then the super-block index; each lane's super-blocks ascend
            unsigned long long bk = GB_NONE;
            uint32_t bs = UINT32_MAX;
            for (uint32_t sb = tid; sb < ns; sb += GR_BLOCK) {
                const unsigned long long k = smin[sb];
                if (k != GB_NONE && (bk == GB_NONE || (k >> 12) < (bk >> 12))) { bk = k; bs = sb; }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned long long ok = __shfl_xor(bk, o, GA_WAVE);
                const uint32_t os = __shfl_xor(bs, o, GA_WAVE);
                if (ok != GB_NONE && (bk == GB_NONE || (ok >> 12) < (bk >> 12) || ((ok >> 12) == (bk >> 12) && os < bs))) { bk = ok; bs = os; }
            }
            if (lane == 0) { red_key[wave] = bk; red_s[wave] = bs; }
            __syncthreads();
            bk = red_key[0];
            bs = red_s[0];
            for (int w = 1; w < GR_WAVES; w++) {
                const unsigned long long ok = red_key[w];
                const uint32_t os = red_s[w];
                if (ok != GB_NONE && (bk == GB_NONE || (ok >> 12) < (bk >> 12) || ((ok >> 12) == (bk >> 12) && os < bs))) { bk = ok; bs = os; }
            }
            if (bk == GB_NONE) { err = 2; break; }   // left > 0, so an alive vertex exists (uniform)
            const uint32_t v = bs * GB_SUPV + (uint32_t)(bk & (GB_SUPV - 1));
            if (v >= n) { err = 2; break; }
            const uint32_t cv = gb_ld32(J.cur + v);
            if (tid == 0) {
                __hip_atomic_store(J.alive + v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                order[step] = v;
            }
            const uint32_t ub = J.urow[v], ue = J.urow[v + 1];
            uint64_t found = 0;
            if (cv) {   // uniform
                for (uint32_t j = ub + tid; j < ue; j += GR_BLOCK) {
                    const uint32_t a = J.ucol[j];
                    if (gb_ld32(J.alive + a)) atomicOr(&J.nbr[a >> 5], 1u << (a & 31));
                }
                __syncthreads();
                // triangles {v, a, b} of H, each once: from the oriented edge between a and b
                for (uint32_t j = ub + wave; j < ue; j += GR_WAVES) {
                    const uint32_t a = J.ucol[j];
                    if (!((gb_ld32(J.nbr + (a >> 5)) >> (a & 31)) & 1u)) continue;
                    const uint32_t re = J.rowptr[a + 1];
                    for (uint32_t e = J.rowptr[a] + lane; e < re; e += GA_WAVE) {
                        const uint32_t b = J.col[e];
                        if ((gb_ld32(J.nbr + (b >> 5)) >> (b & 31)) & 1u) {
                            found++;
                            const uint32_t oa = atomicSub(J.cur + a, 1u), ob = atomicSub(J.cur + b, 1u);
                            if (oa == 0 || ob == 0) {
                                bad = 1;
                            } else {
                                gb_lower(J, smin, a, gb_ld64(J.load + a) + oa - 1u);
                                gb_lower(J, smin, b, gb_ld64(J.load + b) + ob - 1u);
                            }
                        }
                    }
                }
                found = wave_sum_u64(found);
                if (lane == 0) red_found[wave] = found;
                __syncthreads();
                found = 0;
                for (int w = 0; w < GR_WAVES; w++) found += red_found[w];
                for (uint32_t j = ub + tid; j < ue; j += GR_BLOCK) J.nbr[J.ucol[j] >> 5] = 0;   // only v's neighbours' bits were set
                if (bad || found != cv) err = 1;   // read after the barrier: the same in every lane
            }
            if (tid == 0 && cv)   // l_v of the next pass (dsg.h:234); v is not read again in this pass
                __hip_atomic_fetch_add(J.load + v, (unsigned long long)cv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            T -= cv;
            __threadfence();
            __syncthreads();
            if (err) break;
            if (wave == 0) {   // v's block and super-block without v (after every decrement of this deletion)
                const unsigned long long bm = gb_block_min(J, v / GB_BLK, v);
                if (lane == 0) __hip_atomic_store(J.bmin + v / GB_BLK, bm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long sm = gb_super_min(J, v / GB_SUPV, v / GB_BLK, bm);
                if (lane == 0) smin[v / GB_SUPV] = sm;
            }
            __syncthreads();
        }
        if (improved) {
            best_pass = p;
            best_cut = cut;
        }
        if (tid == 0) {
            J.per_pass[2 * p] = bn;
            J.per_pass[2 * p + 1] = bd;
        }
    }
    if (tid == 0) {
        J.out->bn = bn;
        J.out->bd = bd;
        J.out->best_pass = best_pass;
        J.out->best_cut = best_cut;
        J.out->error = err;
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

size_t greedy_lds(uint32_t n) { return (size_t)n * 12 + (size_t)((n + 31) / 32) * 8; }

uint32_t dense_stride(uint32_t n) { return ((n + 31) / 32) | 1u; }
size_t dense_lds(uint32_t n) { return ((size_t)n + 1) * dense_stride(n) * 4; }

size_t lds_room(const ga_ctx* ctx) { return ctx->lds_max > 1024 ? ctx->lds_max - 1024 : 0; }   // static LDS stays below 1 KB

constexpr int DENSE_BLOCK = 1024, DENSE_OWN = 2;

// The largest n the dense form takes: its bitmap fits LDS and each lane owns at most DENSE_OWN vertices.
uint32_t dense_max(const ga_ctx* ctx) {
    uint32_t n = DENSE_BLOCK * DENSE_OWN;
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
    const bool big = n > cap;   // state in HBM (k_greedy_big)
    if (big && ((size_t)(n + GB_SUPV - 1) / GB_SUPV) * 8 > lds_room(ctx))
        return ga_fail(ctx, GA_ERR_ARG, "greedy_pp: n = %u needs more super-block minima than LDS holds", n);
    if (big && passes > (1u << 19)) return ga_fail(ctx, GA_ERR_ARG, "greedy_pp on a large graph: at most 2^19 passes (52-bit keys)");
    if ((uint64_t)passes * n > (1ull << 28))
        return ga_fail(ctx, GA_ERR_ARG, "greedy_pp keeps every pass's deletion order: passes * n = %llu exceeds 2^28",
                       (unsigned long long)passes * n);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    GA_TRY(ga_graph_ensure_undirected(ctx, g));
    DevBuf dtri, dorders, dpass, dout, dflags, dpos, dset, dT, djob;
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
    DevBuf dbig;
    if (big) {
        unsigned long long hT = 0;
        GA_HIP(ctx, hipMemcpyAsync(&hT, dT.p, 8, hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (hT >= (1ull << 32)) return ga_fail(ctx, GA_ERR_ARITH, "greedy_pp on a large graph: T >= 2^32 (u32 tri_H counters)");
        const size_t nw = ((size_t)n + 31) / 32, nb = ((size_t)n + GB_BLK - 1) / GB_BLK;
        GA_TRY(ga_malloc(ctx, &dbig.p, (size_t)n * 16 + nw * 4 + nb * 8 + 64));
        GreedyBig J;
        J.n = n;
        J.passes = passes;
        J.rowptr = g->rowptr;
        J.col = g->col;
        J.urow = g->urow;
        J.ucol = g->ucol;
        J.tri0 = (const uint32_t*)dtri.p;
        J.dT = (const unsigned long long*)dT.p;
        J.orders = (uint32_t*)dorders.p;
        J.per_pass = (unsigned long long*)dpass.p;
        J.out = (GreedyOut*)dout.p;
        J.load = (unsigned long long*)dbig.p;
        J.bmin = J.load + n;
        J.cur = (uint32_t*)(J.bmin + nb);
        J.alive = J.cur + n;
        J.nbr = J.alive + n;
        const size_t ns = ((size_t)n + GB_SUPV - 1) / GB_SUPV;
        k_greedy_big<<<1, GR_BLOCK, ns * 8, ctx->stream>>>(J);
        GA_TRY(ga_launch_check(ctx, "k_greedy_big"));
    } else if (n && n <= dense_max(ctx)) {
        const GreedyJob job = {n, g->m, dense_stride(n), 0, g->src, g->col, (const uint32_t*)dtri.p, (const unsigned long long*)dT.p,
                               (uint32_t*)dorders.p, (unsigned long long*)dpass.p, (GreedyOut*)dout.p};
        GA_TRY(ga_malloc(ctx, &djob.p, sizeof job));
        GA_HIP(ctx, hipMemcpyAsync(djob.p, &job, sizeof job, hipMemcpyHostToDevice, ctx->stream));
        k_greedy_dense<DENSE_BLOCK, DENSE_OWN><<<1, DENSE_BLOCK, dense_lds(n), ctx->stream>>>((const GreedyJob*)djob.p, passes);
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
    if (h.error)
        return ga_fail(ctx, GA_ERR_STATE, "greedy_pp: %s",
                       h.error == 1 ? "triangle bookkeeping inconsistent" : "no alive vertex found for a deletion");
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

// Greedy++ on many graphs: one workgroup per graph in one launch for the graphs the dense form takes, the CSR form one
// graph at a time for the others.
extern "C" int ga_greedy_pp_many(ga_ctx* ctx, ga_graph* const* graphs, uint32_t count, uint32_t passes, uint64_t* stats,
                                 uint8_t* in_set, float* ms) {
    if (!ctx) return GA_ERR_ARG;
    if (count && (!graphs || !stats)) return ga_fail(ctx, GA_ERR_ARG, "graphs and stats are required");
    if (passes == 0) return ga_fail(ctx, GA_ERR_ARG, "greedy_pp needs at least one pass");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t dmax = dense_max(ctx), cap = ga_greedy_max_vertices(ctx);
    std::vector<unsigned long long> voff((size_t)count + 1, 0);
    std::vector<size_t> slab_off((size_t)count + 1, 0);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    for (uint32_t i = 0; i < count; i++) {
        const ga_graph* g = graphs[i];
        if (!g || g->ctx != ctx) return ga_fail(ctx, GA_ERR_ARG, "graph %u is NULL or belongs to another context", i);
        if (g->n > cap) return ga_fail(ctx, GA_ERR_ARG, "graph %u: n = %u exceeds greedy_pp's %u vertices", i, g->n, cap);
        if ((uint64_t)passes * g->n > (1ull << 28))
            return ga_fail(ctx, GA_ERR_ARG, "graph %u: passes * n = %llu exceeds 2^28", i, (unsigned long long)passes * g->n);
        voff[i + 1] = voff[i] + g->n;
        // per graph: tri0 (n), T (u64), orders (passes * n), per_pass (2 * passes u64); the results sit in one array
        slab_off[i + 1] = slab_off[i] + al((size_t)g->n * 4) + al(8) + al((size_t)passes * g->n * 4) + al((size_t)passes * 16);
    }
    if (!count) return GA_OK;
    DevBuf slab, djobs, dvoff, dflags, douts;
    GA_TRY(ga_malloc(ctx, &slab.p, std::max<size_t>(slab_off[count], 256)));
    GA_TRY(ga_malloc(ctx, &douts.p, (size_t)count * sizeof(GreedyOut)));
    GA_TRY(ga_malloc(ctx, &djobs.p, (size_t)count * sizeof(GreedyJob)));
    GA_TRY(ga_malloc(ctx, &dvoff.p, ((size_t)count + 1) * 8));
    if (in_set) GA_TRY(ga_malloc(ctx, &dflags.p, std::max<size_t>(voff[count], 1)));
    std::vector<GreedyJob> jobs(count);
    std::vector<uint32_t> dense_ids, sparse_ids;
    size_t lds = 0;
    for (uint32_t i = 0; i < count; i++) {
        ga_graph* g = graphs[i];
        char* base = (char*)slab.p + slab_off[i];
        GreedyJob& j = jobs[i];
        j.n = g->n;
        j.m = g->m;
        j.wpad = dense_stride(g->n);
        j.pad = 0;
        j.src = g->src;
        j.col = g->col;
        j.tri0 = (const uint32_t*)base;
        base += al((size_t)g->n * 4);
        j.dT = (const unsigned long long*)base;
        base += al(8);
        j.orders = (uint32_t*)base;
        base += al((size_t)passes * g->n * 4);
        j.per_pass = (unsigned long long*)base;
        j.out = (GreedyOut*)douts.p + i;
        if (g->n && g->n <= dmax) {   // counted from its bitmap inside the launch
            j.tri0 = nullptr;
            dense_ids.push_back(i);
            lds = std::max(lds, dense_lds(g->n));
        } else if (g->n) {
            sparse_ids.push_back(i);
            GA_TRY(ga_graph_ensure_undirected(ctx, g));
        }
    }
    std::vector<GreedyJob> djv;
    for (uint32_t i : dense_ids) djv.push_back(jobs[i]);
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(ga_fill_u8(ctx, (uint8_t*)slab.p, slab_off[count], 0));
    GA_TRY(ga_fill_u8(ctx, (uint8_t*)douts.p, (size_t)count * sizeof(GreedyOut), 0));
    GA_HIP(ctx, hipMemcpyAsync(djobs.p, jobs.data(), (size_t)count * sizeof(GreedyJob), hipMemcpyHostToDevice, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(dvoff.p, voff.data(), ((size_t)count + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    for (uint32_t i = 0; i < count; i++)   // per-vertex triangles and T of the graphs the CSR form runs
        if (graphs[i]->m && graphs[i]->n > dmax)
            GA_TRY(ga_tc_launch(ctx, graphs[i], GA_TC_HASH, (unsigned long long*)jobs[i].dT, nullptr, (uint32_t*)jobs[i].tri0, nullptr,
                                nullptr));
    DevBuf ddense;
    if (!djv.empty()) {
        GA_TRY(ga_malloc(ctx, &ddense.p, djv.size() * sizeof(GreedyJob)));
        GA_HIP(ctx, hipMemcpyAsync(ddense.p, djv.data(), djv.size() * sizeof(GreedyJob), hipMemcpyHostToDevice, ctx->stream));
        k_greedy_dense<DENSE_BLOCK, DENSE_OWN><<<(unsigned)djv.size(), DENSE_BLOCK, lds, ctx->stream>>>((const GreedyJob*)ddense.p, passes);
        GA_TRY(ga_launch_check(ctx, "k_greedy_dense (many)"));
    }
    for (uint32_t i : sparse_ids) {
        const ga_graph* g = graphs[i];
        const GreedyJob& j = jobs[i];
        k_greedy_pp<<<1, GR_BLOCK, greedy_lds(g->n), ctx->stream>>>(g->n, g->rowptr, g->col, g->urow, g->ucol, j.tri0, j.dT, passes,
                                                                  j.orders, j.per_pass, j.out);
        GA_TRY(ga_launch_check(ctx, "k_greedy_pp (many)"));
    }
    if (in_set) {
        GA_TRY(ga_fill_u8(ctx, (uint8_t*)dflags.p, std::max<size_t>(voff[count], 1), 0));
        for (uint32_t b0 = 0; b0 < count; b0 += 65535u) {
            const dim3 grid(4, std::min<uint32_t>(65535u, count - b0));
            k_greedy_flags_many<<<grid, 256, 0, ctx->stream>>>((const GreedyJob*)djobs.p, (const unsigned long long*)dvoff.p, b0,
                                                              (uint8_t*)dflags.p);
        }
        GA_TRY(ga_launch_check(ctx, "k_greedy_flags_many"));
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    std::vector<GreedyOut> outs(count);
    GA_HIP(ctx, hipMemcpy(outs.data(), douts.p, (size_t)count * sizeof(GreedyOut), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < count; i++) {
        const GreedyOut& h = outs[i];
        if (h.error)
            return ga_fail(ctx, GA_ERR_STATE, "greedy_pp graph %u: %s", i,
                           h.error == 1 ? "triangle bookkeeping inconsistent" : "no alive vertex found for a deletion");
        stats[2 * i] = h.bd ? h.bn : 0;
        stats[2 * i + 1] = h.bd;
    }
    if (in_set && voff[count]) GA_HIP(ctx, hipMemcpy(in_set, dflags.p, voff[count], hipMemcpyDeviceToHost));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return GA_OK;
}
