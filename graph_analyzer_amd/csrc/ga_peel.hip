// ga_peel.hip - triangle-density batch peeling (densest.h:1-44 algorithm, 265-303 host loop; kernels.cuh:360-459).
//
// Per-round backend calls mirror the reference's backend concept (reset_state / begin_round / run_tc / run_mark /
// read_stats), so the unchanged host loop batch_peel_loop can drive them with one 16-byte readback per round.
// ga_peel_run keeps the whole round state on the device instead (PeelDev): every kernel of a round reads nalive,
// round and eps from it and returns at once when `done` is set, a one-thread epilogue kernel applies the host loop's
// rules (round-0 count check, strict-improvement best, removed-count sanity, nalive update), and the host enqueues
// rounds in growing blocks with one readback per block instead of one per round.
#include "ga_internal.hpp"

struct PeelDev {
    unsigned long long T;          // triangles of the current alive subgraph (written by the count kernel)
    unsigned long long removed;    // vertices removed this round (written by the mark kernel)
    unsigned long long nalive, best_num, best_den, T_full, num, den;
    uint32_t round, best_round, rounds, done, error, pad[3];
};

enum { PEEL_ERR_T = 1, PEEL_ERR_ARITH = 2, PEEL_ERR_REMOVED = 3 };

struct ga_peel {
    ga_ctx* ctx = nullptr;
    ga_graph* g = nullptr;
    uint8_t* alive = nullptr;
    uint32_t* removed_round = nullptr;
    uint32_t* tri = nullptr;
    PeelDev* st = nullptr;
    int variant = GA_TC_HASH;
};

// tri[v] * nalive * den <= 3 * (den + num) * T  (kernels.cuh:427-459, densest.h:151-154): remove v, record the round.
// bid / nb: this workgroup and the workgroups sharing the graph (the grid, or the graph's range of a shared launch)
template <bool FUSED>
__device__ __forceinline__ void mark_peel_body(const uint32_t* __restrict__ tri, uint8_t* __restrict__ alive, uint32_t* __restrict__ removed_round,
                                               uint32_t n, unsigned long long nalive_a, unsigned long long num_a,
                                               unsigned long long den_a, uint32_t round_a, PeelDev* __restrict__ st, uint32_t bid,
                                               uint32_t nb) {
    if (FUSED && st->done) return;
    const unsigned long long nalive = FUSED ? st->nalive : nalive_a;
    const unsigned long long num = FUSED ? st->num : num_a;
    const unsigned long long den = FUSED ? st->den : den_a;
    const uint32_t round = FUSED ? st->round : round_a;
    const unsigned long long rhs = 3ull * (den + num) * st->T;
    const unsigned long long mul = nalive * den;
    uint64_t local = 0;
    // four vertices per lane: one 32-bit load of alive[] and one 128-bit load of tri[] when aligned
    const uint32_t nq = n / 4;
    for (uint32_t q = bid * blockDim.x + threadIdx.x; q < nq; q += nb * blockDim.x) {
        uint32_t a4 = reinterpret_cast<const uint32_t*>(alive)[q];
        if (!a4) continue;
        const uint4 t4 = reinterpret_cast<const uint4*>(tri)[q];
        const uint32_t tv[4] = {t4.x, t4.y, t4.z, t4.w};
        uint32_t na4 = a4;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (((a4 >> (8 * k)) & 0xFF) && (unsigned long long)tv[k] * mul <= rhs) {
                na4 &= ~(0xFFu << (8 * k));
                removed_round[4 * q + k] = round;
                local++;
            }
        }
        if (na4 != a4) reinterpret_cast<uint32_t*>(alive)[q] = na4;
    }
    for (uint32_t v = nq * 4 + bid * blockDim.x + threadIdx.x; v < n; v += nb * blockDim.x) {
        if (alive[v] && (unsigned long long)tri[v] * mul <= rhs) {
            alive[v] = 0;
            removed_round[v] = round;
            local++;
        }
    }
    local = wave_sum_u64(local);
    if (lane_id() == 0 && local) atomicAdd(&st->removed, (unsigned long long)local);
}

template <bool FUSED>
__global__ void k_mark_peel(const uint32_t* __restrict__ tri, uint8_t* __restrict__ alive, uint32_t* __restrict__ removed_round,
                            uint32_t n, unsigned long long nalive_a, unsigned long long num_a, unsigned long long den_a,
                            uint32_t round_a, PeelDev* __restrict__ st) {
    mark_peel_body<FUSED>(tri, alive, removed_round, n, nalive_a, num_a, den_a, round_a, st, blockIdx.x, gridDim.x);
}

// Zeroes tri[] and the round counters (replaces the two cudaMemset calls of gpu_densest.cu:101-104).
template <bool FUSED>
__device__ __forceinline__ void peel_begin_body(uint32_t* __restrict__ tri, uint32_t n, PeelDev* __restrict__ st, uint32_t bid, uint32_t nb) {
    if (FUSED && st->done) return;
    for (uint32_t v = bid * blockDim.x + threadIdx.x; v < n; v += nb * blockDim.x) tri[v] = 0;
    if (bid == 0 && threadIdx.x == 0) {
        st->T = 0;
        st->removed = 0;
    }
}

template <bool FUSED>
__global__ void k_peel_begin(uint32_t* __restrict__ tri, uint32_t n, PeelDev* __restrict__ st) {
    peel_begin_body<FUSED>(tri, n, st, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void peel_init_dev(PeelDev* st, unsigned long long n, unsigned long long T_full, unsigned long long num,
                                              unsigned long long den) {
    st->T = 0;
    st->removed = 0;
    st->nalive = n;
    st->best_num = 0;
    st->best_den = 0;
    st->T_full = T_full;
    st->num = num;
    st->den = den;
    st->round = 0;
    st->best_round = 0;
    st->rounds = 0;
    st->done = n == 0;
    st->error = 0;
}

__global__ void k_peel_init(PeelDev* st, unsigned long long n, unsigned long long T_full, unsigned long long num, unsigned long long den) {
    peel_init_dev(st, n, T_full, num, den);
}

// Device restatement of peel_arith_ok (densest.h:65-72) for the round-0 count.
__device__ bool peel_arith_ok_dev(unsigned long long n, unsigned long long T0, unsigned long long num, unsigned long long den) {
    if (den < 1 || den > (1ull << 20) || num > (1ull << 20)) return false;
    if (T0 >= (1ull << 32)) return false;
    if (n > 0 && T0 > ~0ull / (n * den)) return false;
    if (T0 > ~0ull / (3ull * (den + num))) return false;
    return true;
}

// End of a fused round: the rules of batch_peel_loop (densest.h:288-301).
__device__ void peel_end_dev(PeelDev* st, unsigned long long n) {
    if (st->done) return;
    const unsigned long long T = st->T, removed = st->removed, nalive = st->nalive;
    const uint32_t round = st->round;
    if (round == 0) {
        if (st->T_full != ~0ull && T != st->T_full) { st->error = PEEL_ERR_T; st->done = 1; return; }
        if (!peel_arith_ok_dev(n, T, st->num, st->den)) { st->error = PEEL_ERR_ARITH; st->done = 1; return; }
    }
    if (st->best_den == 0 || T * st->best_den > st->best_num * nalive) {
        st->best_num = T;
        st->best_den = nalive;
        st->best_round = round;
    }
    if (removed == 0 || removed > nalive) { st->error = PEEL_ERR_REMOVED; st->done = 1; return; }
    st->nalive = nalive - removed;
    st->rounds = round + 1;
    st->round = round + 1;
    if (st->nalive == 0) st->done = 1;
}

__global__ void k_peel_end(PeelDev* st, unsigned long long n) { peel_end_dev(st, n); }

// Several graphs' peels in shared launches (ga_peel_run_batch): job i is graph i's state with its ranges of the begin
// (q = 0) and mark (q = 1) launches; every kernel returns at once for a finished state.
struct PeelJob {
    uint32_t* tri;
    uint8_t* alive;
    uint32_t* removed_round;
    PeelDev* st;
    unsigned long long n, T_full, num, den;
    uint32_t base[2], nblk[2];
};

__global__ void k_peel_init_many(const PeelJob* __restrict__ jobs, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        peel_init_dev(jobs[i].st, jobs[i].n, jobs[i].T_full, jobs[i].num, jobs[i].den);
}

__global__ void k_peel_begin_many(const PeelJob* __restrict__ jobs, const uint32_t* __restrict__ map) {
    const PeelJob& J = jobs[map[blockIdx.x]];
    peel_begin_body<true>(J.tri, (uint32_t)J.n, J.st, blockIdx.x - J.base[0], J.nblk[0]);
}

__global__ void k_mark_peel_many(const PeelJob* __restrict__ jobs, const uint32_t* __restrict__ map) {
    const PeelJob& J = jobs[map[blockIdx.x]];
    mark_peel_body<true>(J.tri, J.alive, J.removed_round, (uint32_t)J.n, 0, 0, 0, 0, J.st, blockIdx.x - J.base[1], J.nblk[1]);
}

__global__ void k_peel_end_many(const PeelJob* __restrict__ jobs, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) peel_end_dev(jobs[i].st, jobs[i].n);
}

// Every state's record into one array (one readback per block of rounds).
__global__ void k_peel_gather(const PeelJob* __restrict__ jobs, uint32_t count, PeelDev* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) out[i] = *jobs[i].st;
}

static unsigned vgrid(ga_ctx* ctx, uint32_t n) { return ctx_grid(ctx, grid_for(((uint64_t)n + 3) / 4, 256, (unsigned)std::max(ctx->cus, 1) * 8u)); }

extern "C" int ga_peel_create(ga_ctx* ctx, ga_graph* g, ga_peel** out) {
    *out = nullptr;
    if (!ctx || !g) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    ga_peel* p = new ga_peel;
    p->ctx = ctx;
    p->g = g;
    const size_t n = (size_t)g->n;
    // alive padded to a multiple of 4 bytes and tri to 16 bytes for the vector loads of the mark kernel
    if (ga_malloc(ctx, (void**)&p->alive, n + 16) || ga_malloc(ctx, (void**)&p->removed_round, (n + 1) * 4) ||
        ga_malloc(ctx, (void**)&p->tri, (n + 4) * 4) || ga_malloc(ctx, (void**)&p->st, sizeof(PeelDev))) {
        ga_peel_destroy(p);
        return GA_ERR_OOM;
    }
    int rc = ga_tc_ensure_plan(ctx, g);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ga_fail(ctx, GA_ERR_HIP, "peel state initialisation");
    if (rc) { ga_peel_destroy(p); return rc; }
    *out = p;
    return GA_OK;
}

extern "C" int ga_peel_set_ctx(ga_peel* p, ga_ctx* ctx) {
    if (!p || !ctx || ctx->device != p->ctx->device) return GA_ERR_ARG;
    p->ctx = ctx;   // every call on the state ends with a stream sync, so an idle state has nothing in flight
    return GA_OK;
}

extern "C" void ga_peel_destroy(ga_peel* p) {
    if (!p) return;
    if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
    ga_free(p->alive);
    ga_free(p->removed_round);
    ga_free(p->tri);
    ga_free(p->st);
    delete p;
}

extern "C" int ga_peel_reset_state(ga_peel* p) {
    ga_ctx* ctx = p->ctx;
    GA_TRY(ga_fill_u8(ctx, p->alive, p->g->n, 1));
    GA_TRY(ga_fill_u32(ctx, p->removed_round, p->g->n, UINT32_MAX));
    return GA_OK;
}

extern "C" int ga_peel_begin_round(ga_peel* p) {
    ga_ctx* ctx = p->ctx;
    k_peel_begin<false><<<vgrid(ctx, p->g->n), 256, 0, ctx->stream>>>(p->tri, p->g->n, p->st);
    return ga_launch_check(ctx, "k_peel_begin");
}

extern "C" int ga_peel_run_tc(ga_peel* p) {
    return ga_tc_launch(p->ctx, p->g, p->variant, &p->st->T, p->alive, p->tri, nullptr, nullptr);
}

extern "C" int ga_peel_run_mark(ga_peel* p, uint64_t nalive, uint64_t num, uint64_t den, uint32_t round) {
    ga_ctx* ctx = p->ctx;
    k_mark_peel<false><<<vgrid(ctx, p->g->n), 256, 0, ctx->stream>>>(p->tri, p->alive, p->removed_round, p->g->n, nalive, num, den, round, p->st);
    return ga_launch_check(ctx, "k_mark_peel");
}

extern "C" int ga_peel_read_stats(ga_peel* p, uint64_t* T, uint64_t* removed) {
    ga_ctx* ctx = p->ctx;
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, p->st, 16, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (T) *T = ctx->hstat[0];
    if (removed) *removed = ctx->hstat[1];
    return GA_OK;
}

extern "C" int ga_peel_download_removed_round(ga_peel* p, uint32_t* out) {
    ga_ctx* ctx = p->ctx;
    if (p->g->n) GA_HIP(ctx, hipMemcpyAsync(out, p->removed_round, (size_t)p->g->n * 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

// Host check that mirrors peel_arith_ok (densest.h:65-72).
static const char* peel_arith_why(uint64_t n, uint64_t T0, uint64_t num, uint64_t den) {
    if (den < 1 || den > (1ull << 20) || num > (1ull << 20)) return "eps = num/den needs 1 <= den <= 2^20 and num <= 2^20";
    if (T0 >= (1ull << 32)) return "triangle count >= 2^32: uint32 tri[] counters could overflow";
    if (n > 0 && T0 > UINT64_MAX / (n * den)) return "tri*nalive*den could overflow uint64";
    if (T0 > UINT64_MAX / (3ull * (den + num))) return "3*(den+num)*T could overflow uint64";
    return nullptr;
}

extern "C" int ga_peel_run(ga_peel* p, uint64_t T_full, uint64_t num, uint64_t den, uint64_t* stats, float* ms) {
    ga_ctx* ctx = p->ctx;
    ga_graph* g = p->g;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (T_full != UINT64_MAX) {
        const char* why = peel_arith_why(g->n, T_full, num, den);
        if (why) return ga_fail(ctx, GA_ERR_ARITH, "%s", why);
    } else if (den < 1 || den > (1ull << 20) || num > (1ull << 20)) {
        return ga_fail(ctx, GA_ERR_ARITH, "eps = num/den needs 1 <= den <= 2^20 and num <= 2^20");
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_TRY(ga_peel_reset_state(p));
    k_peel_init<<<1, 1, 0, ctx->stream>>>(p->st, g->n, T_full, num, den);
    GA_TRY(ga_launch_check(ctx, "k_peel_init"));
    const uint32_t* done = &p->st->done;
    int block_rounds = 4;
    for (int guard = 0;; guard++) {
        for (int r = 0; r < block_rounds; r++) {
            k_peel_begin<true><<<vgrid(ctx, g->n), 256, 0, ctx->stream>>>(p->tri, g->n, p->st);
            GA_TRY(ga_launch_check(ctx, "k_peel_begin"));
            GA_TRY(ga_tc_launch(ctx, g, p->variant, &p->st->T, p->alive, p->tri, nullptr, done));
            k_mark_peel<true><<<vgrid(ctx, g->n), 256, 0, ctx->stream>>>(p->tri, p->alive, p->removed_round, g->n, 0, 0, 0, 0, p->st);
            GA_TRY(ga_launch_check(ctx, "k_mark_peel"));
            k_peel_end<<<1, 1, 0, ctx->stream>>>(p->st, g->n);
            GA_TRY(ga_launch_check(ctx, "k_peel_end"));
        }
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, p->st, sizeof(PeelDev), hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const PeelDev* h = (const PeelDev*)ctx->hstat;
        if (h->done) break;
        if (guard > 1000000) return ga_fail(ctx, GA_ERR_STATE, "peel did not finish");
        block_rounds = std::min(block_rounds * 2, 64);
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    const PeelDev* h = (const PeelDev*)ctx->hstat;
    switch (h->error) {
        case 0: break;
        case PEEL_ERR_T: return ga_fail(ctx, GA_ERR_STATE, "device triangle count in round 0 differs from the given count");
        case PEEL_ERR_ARITH: return ga_fail(ctx, GA_ERR_ARITH, "%s", peel_arith_why(g->n, h->T, num, den) ? peel_arith_why(g->n, h->T, num, den) : "arithmetic bound");
        default: return ga_fail(ctx, GA_ERR_STATE, "bad removed count from device (0 or more than alive)");
    }
    if (stats) {
        stats[0] = h->best_num;
        stats[1] = h->best_den;
        stats[2] = h->best_round;
        stats[3] = h->rounds;
    }
    return GA_OK;
}

// Several graphs' fused peels in shared launches: each round is one begin, one participation launch per bucket shape,
// one mark and one end launch for all unfinished states, in blocks of rounds with one readback of every state per
// block. Each state follows exactly the rounds it follows alone.
extern "C" int ga_peel_run_batch(ga_ctx* ctx, ga_peel* const* peels, uint32_t count, const uint64_t* T_full, uint64_t num, uint64_t den,
                                 uint64_t* stats, float* ms) {
    if (!ctx || (count && (!peels || !stats))) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (!count) return GA_OK;
    for (uint32_t i = 0; i < count; i++) {
        ga_peel* p = peels[i];
        if (!p || p->ctx->device != ctx->device) return ga_fail(ctx, GA_ERR_ARG, "peel state %u is NULL or on another device", i);
        if (p->variant != GA_TC_HASH) return ga_fail(ctx, GA_ERR_ARG, "peel state %u: shared launches run the hash variant", i);
        const uint64_t tf = T_full ? T_full[i] : UINT64_MAX;
        const char* why = tf != UINT64_MAX ? peel_arith_why(p->g->n, tf, num, den)
                                           : (den < 1 || den > (1ull << 20) || num > (1ull << 20))
                                                 ? "eps = num/den needs 1 <= den <= 2^20 and num <= 2^20" : nullptr;
        if (why) return ga_fail(ctx, GA_ERR_ARITH, "peel state %u: %s", i, why);
        for (uint32_t j = 0; j < i; j++)
            if (peels[j] == p) return ga_fail(ctx, GA_ERR_ARG, "peel state %u appears twice", i);
    }
    for (uint32_t i = 0; i < count; i++) peels[i]->ctx = ctx;   // idle states: every call on a state ends with a sync
    std::vector<ga_graph*> gs(count);
    std::vector<const uint8_t*> al(count);
    std::vector<uint32_t*> tr(count);
    std::vector<unsigned long long*> tot(count);
    std::vector<const uint32_t*> dn(count);
    for (uint32_t i = 0; i < count; i++) {
        gs[i] = peels[i]->g;
        al[i] = peels[i]->alive;
        tr[i] = peels[i]->tri;
        tot[i] = &peels[i]->st->T;
        dn[i] = &peels[i]->st->done;
    }
    ga_tc_many* M = nullptr;
    GA_TRY(ga_tc_participation_many(ctx, gs.data(), al.data(), tr.data(), tot.data(), dn.data(), count, &M));
    struct Free {
        ga_tc_many* m;
        ~Free() { ga_tc_many_free(m); }
    } free_m{M};
    // begin and mark ranges: each state its own grid, scaled down together above 8 workgroups per CU
    const unsigned cus = (unsigned)std::max(ctx->cus, 1);
    std::vector<PeelJob> jobs(count);
    std::vector<uint32_t> map[2];
    uint64_t sum = 0;
    for (uint32_t i = 0; i < count; i++) sum += vgrid(ctx, peels[i]->g->n);
    for (uint32_t i = 0; i < count; i++) {
        ga_peel* p = peels[i];
        unsigned nb = vgrid(ctx, p->g->n);
        if (sum > 8ull * cus) nb = std::max(1u, (unsigned)((uint64_t)nb * 8ull * cus / sum));
        jobs[i] = PeelJob{p->tri, p->alive, p->removed_round, p->st, p->g->n, T_full ? T_full[i] : UINT64_MAX, num, den, {0, 0}, {0, 0}};
        for (int q = 0; q < 2; q++) {
            jobs[i].base[q] = (uint32_t)map[q].size();
            jobs[i].nblk[q] = nb;
            map[q].insert(map[q].end(), nb, i);
        }
    }
    const size_t jobs_bytes = ((size_t)count * sizeof(PeelJob) + 255) & ~(size_t)255;
    const size_t map_bytes = ((map[0].size() + map[1].size()) * 4 + 255) & ~(size_t)255;
    const size_t need = jobs_bytes + map_bytes + (size_t)count * sizeof(PeelDev);
    if (need > ctx->dcore_bytes) {   // grow-only, shared with the core batches: one call at a time on a context
        ga_free(ctx->dcore);
        ctx->dcore = nullptr;
        ctx->dcore_bytes = 0;
        GA_TRY(ga_malloc(ctx, &ctx->dcore, need * 2));
        ctx->dcore_bytes = need * 2;
    }
    std::vector<uint8_t>& host = ctx->hpeel;   // outlives an early error return while the copy may be in flight
    host.resize(jobs_bytes + map_bytes);
    memcpy(host.data(), jobs.data(), (size_t)count * sizeof(PeelJob));
    memcpy(host.data() + jobs_bytes, map[0].data(), map[0].size() * 4);
    memcpy(host.data() + jobs_bytes + map[0].size() * 4, map[1].data(), map[1].size() * 4);
    char* dbase = (char*)ctx->dcore;
    const PeelJob* djobs = (const PeelJob*)dbase;
    const uint32_t* dmap0 = (const uint32_t*)(dbase + jobs_bytes);
    const uint32_t* dmap1 = dmap0 + map[0].size();
    PeelDev* dout = (PeelDev*)(dbase + jobs_bytes + map_bytes);
    GA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(dbase, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
    for (uint32_t i = 0; i < count; i++) GA_TRY(ga_peel_reset_state(peels[i]));
    const unsigned gj = grid_for(count, 256);
    k_peel_init_many<<<gj, 256, 0, ctx->stream>>>(djobs, count);
    GA_TRY(ga_launch_check(ctx, "k_peel_init_many"));
    std::vector<PeelDev> hst(count);
    int block_rounds = 4;
    for (int guard = 0;; guard++) {
        for (int r = 0; r < block_rounds; r++) {
            k_peel_begin_many<<<(unsigned)map[0].size(), 256, 0, ctx->stream>>>(djobs, dmap0);
            GA_TRY(ga_launch_check(ctx, "k_peel_begin_many"));
            GA_TRY(ga_tc_many_launch(ctx, M));
            k_mark_peel_many<<<(unsigned)map[1].size(), 256, 0, ctx->stream>>>(djobs, dmap1);
            GA_TRY(ga_launch_check(ctx, "k_mark_peel_many"));
            k_peel_end_many<<<gj, 256, 0, ctx->stream>>>(djobs, count);
            GA_TRY(ga_launch_check(ctx, "k_peel_end_many"));
        }
        k_peel_gather<<<gj, 256, 0, ctx->stream>>>(djobs, count, dout);
        GA_TRY(ga_launch_check(ctx, "k_peel_gather"));
        GA_HIP(ctx, hipMemcpyAsync(hst.data(), dout, (size_t)count * sizeof(PeelDev), hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        bool all = true;
        for (uint32_t i = 0; i < count && all; i++) all = hst[i].done != 0;
        if (all) break;
        if (guard > 1000000) return ga_fail(ctx, GA_ERR_STATE, "peel batch did not finish");
        block_rounds = std::min(block_rounds * 2, 64);
    }
    GA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (ms) GA_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    for (uint32_t i = 0; i < count; i++) {
        const PeelDev& h = hst[i];
        switch (h.error) {
            case 0: break;
            case PEEL_ERR_T: return ga_fail(ctx, GA_ERR_STATE, "peel state %u: device triangle count in round 0 differs from the given count", i);
            case PEEL_ERR_ARITH: return ga_fail(ctx, GA_ERR_ARITH, "peel state %u: the exact peel arithmetic bounds do not hold", i);
            default: return ga_fail(ctx, GA_ERR_STATE, "peel state %u: bad removed count from device (0 or more than alive)", i);
        }
        stats[4 * i] = h.best_num;
        stats[4 * i + 1] = h.best_den;
        stats[4 * i + 2] = h.best_round;
        stats[4 * i + 3] = h.rounds;
    }
    return GA_OK;
}
