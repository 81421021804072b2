// densest.h (graph_analyzer_amd) - drop-in for the reference's densest.h API (densest.h:46-78, 254-303):
// PeelResult, density_value, peel_arith_ok, peel_improves, densest_row, build_undirected_csr (on the device) and the
// backend-templated host loop batch_peel_loop, plus GpuBackend (the gpu_densest.cu:70-126 backend) over the C ABI and
// a fused gpu_batch_peel.
// The CPU algorithms of the reference header (cpu_triangle_participation, cpu_batch_peel, cpu_greedy_peel) are the
// CPU baseline and are not part of this library.
#pragma once
#include <string>

#include "graph.h"

struct PeelResult {
    uint64_t best_num = 0;       // triangles of the best subgraph
    uint64_t best_den = 0;       // vertices of the best subgraph (0 = nothing recorded yet)
    uint32_t best_round = 0;     // best subgraph = { v : removed_round[v] >= best_round }
    uint32_t rounds = 0;
    std::vector<uint32_t> removed_round;
    bool ok = true;
    std::string error;
};

inline bool operator==(const PeelResult& a, const PeelResult& b) {
    return a.best_num == b.best_num && a.best_den == b.best_den && a.best_round == b.best_round && a.rounds == b.rounds &&
           a.removed_round == b.removed_round;
}

inline double density_value(uint64_t num, uint64_t den) { return den ? (double)num / (double)den : 0.0; }

// The arithmetic bounds of the exact peel: 1 <= den <= 2^20, num <= 2^20, T0 < 2^32, T0*n*den < 2^64,
// 3(den+num)T0 < 2^64 (densest.h:65-72).
inline bool peel_arith_ok(uint32_t n, uint64_t T0, uint64_t num, uint64_t den, std::string* why = nullptr) {
    const char* msg = nullptr;
    if (den < 1 || den > (1ull << 20) || num > (1ull << 20)) msg = "eps = num/den needs 1 <= den <= 2^20 and num <= 2^20";
    else if (T0 >= (1ull << 32)) msg = "triangle count >= 2^32: uint32 tri[] counters could overflow";
    else if (n > 0 && T0 > UINT64_MAX / ((uint64_t)n * den)) msg = "tri*nalive*den could overflow uint64";
    else if (T0 > UINT64_MAX / (3ull * (den + num))) msg = "3*(den+num)*T could overflow uint64";
    if (msg && why) *why = msg;
    return msg == nullptr;
}

// Strictly better density by cross multiplication; ties keep the earlier, larger set (densest.h:76-78).
inline bool peel_improves(uint64_t T, uint64_t nalive, uint64_t best_num, uint64_t best_den) {
    return best_den == 0 || T * best_den > best_num * nalive;
}

// impl,graph,n,m,best_triangles,best_vertices,density,rounds,time_ms,h2d_ms,total_ms (densest.h:254-263)
inline std::string densest_row(const char* impl, const char* graph, uint32_t n, uint32_t m, uint64_t best_num, uint64_t best_den,
                               uint32_t rounds, double time_ms, double h2d_ms) {
    char buf[1024];
    snprintf(buf, sizeof buf, "%s,%s,%u,%u,%llu,%llu,%.6f,%u,%.3f,%.3f,%.3f", impl, graph, n, m, (unsigned long long)best_num,
             (unsigned long long)best_den, density_value(best_num, best_den), rounds, time_ms, h2d_ms, time_ms + h2d_ms);
    return buf;
}

// Host peeling loop over a backend B (densest.h:265-303 semantics): reset_state, then per round begin_round, run_tc,
// run_mark(nalive, num, den, round), read_stats(T, removed); round-0 T must equal T_full; strict-improvement best;
// removed must be in 1..nalive.
template <class B>
inline PeelResult batch_peel_loop(uint32_t n, uint64_t T_full, uint64_t num, uint64_t den, B& be, bool verbose = false,
                                  FILE* diag = nullptr) {
    PeelResult r;
    if (!peel_arith_ok(n, T_full, num, den, &r.error)) { r.ok = false; return r; }
    be.reset_state();
    uint64_t nalive = n;
    uint32_t round = 0;
    while (nalive > 0) {
        be.begin_round();
        be.run_tc();
        be.run_mark(nalive, num, den, round);
        uint64_t T = 0, removed = 0;
        be.read_stats(T, removed);
        if (round == 0 && T != T_full) { r.ok = false; r.error = "device triangle count in round 0 differs from CPU count"; return r; }
        if (verbose && diag)
            fprintf(diag, "round %4u  alive=%10llu  T=%12llu  removed=%10llu  kernel_ms=%.3f\n", round, (unsigned long long)nalive,
                    (unsigned long long)T, (unsigned long long)removed, be.kernel_ms());
        if (peel_improves(T, nalive, r.best_num, r.best_den)) { r.best_num = T; r.best_den = nalive; r.best_round = round; }
        if (removed == 0 || removed > nalive) { r.ok = false; r.error = "bad removed count from device (0 or more than alive)"; return r; }
        nalive -= removed;
        round++;
        r.rounds = round;
    }
    return r;
}

// The GPU backend of batch_peel_loop (the role of gpu_densest.cu:70-126) over ga_peel_*.
struct GpuBackend {
    ga_ctx* ctx = nullptr;
    ga_peel* p = nullptr;
    uint32_t n = 0;
    bool events = false;   // kept for source compatibility; per-round kernel times are not recorded
    void init(const gaa::DeviceGraph& g) {
        ctx = g.ctx();
        n = g.n();
        gaa::check(ctx, ga_peel_create(ctx, g.handle(), &p), "ga_peel_create");
    }
    ~GpuBackend() { if (p) ga_peel_destroy(p); }
    void reset_state() { gaa::check(ctx, ga_peel_reset_state(p), "ga_peel_reset_state"); }
    void begin_round() { gaa::check(ctx, ga_peel_begin_round(p), "ga_peel_begin_round"); }
    void run_tc() { gaa::check(ctx, ga_peel_run_tc(p), "ga_peel_run_tc"); }
    void run_mark(uint64_t nalive, uint64_t num, uint64_t den, uint32_t round) {
        gaa::check(ctx, ga_peel_run_mark(p, nalive, num, den, round), "ga_peel_run_mark");
    }
    void read_stats(uint64_t& T, uint64_t& removed) { gaa::check(ctx, ga_peel_read_stats(p, &T, &removed), "ga_peel_read_stats"); }
    double kernel_ms() const { return -1.0; }
    void download_removed_round(std::vector<uint32_t>& out) {
        out.assign(n, 0);
        gaa::check(ctx, ga_peel_download_removed_round(p, n ? out.data() : nullptr), "ga_peel_download_removed_round");
    }
};

// Whole peel with the round loop on the device (ga_peel_run); same PeelResult as batch_peel_loop + download.
inline PeelResult gpu_batch_peel(const gaa::DeviceGraph& g, uint64_t T_full, uint64_t num, uint64_t den, float* ms = nullptr) {
    PeelResult r;
    ga_peel* p = nullptr;
    gaa::check(g.ctx(), ga_peel_create(g.ctx(), g.handle(), &p), "ga_peel_create");
    uint64_t st[4] = {0, 0, 0, 0};
    float t = 0;
    const int rc = ga_peel_run(p, T_full, num, den, st, &t);
    if (rc != GA_OK) {
        r.ok = false;
        r.error = ga_ctx_last_error(g.ctx());
    } else {
        r.best_num = st[0];
        r.best_den = st[1];
        r.best_round = (uint32_t)st[2];
        r.rounds = (uint32_t)st[3];
        r.removed_round.assign(g.n(), 0);
        gaa::check(g.ctx(), ga_peel_download_removed_round(p, g.n() ? r.removed_round.data() : nullptr), "download_removed_round");
    }
    ga_peel_destroy(p);
    if (ms) *ms = t;
    return r;
}

// build_undirected_csr (densest.h:162-171): both directions of every edge, rows sorted, built on the device.
inline void build_undirected_csr(const Graph& g, std::vector<uint32_t>& urow, std::vector<uint32_t>& ucol) {
    ga_ctx* ctx = gaa::default_ctx();
    gaa::DeviceGraph d = gaa::DeviceGraph::upload(ctx, g);
    urow.assign((size_t)g.n + 1, 0);
    ucol.assign(2ull * g.m, 0);
    gaa::check(ctx, ga_graph_undirected(ctx, d.handle(), urow.data(), g.m ? ucol.data() : nullptr), "ga_graph_undirected");
}
