// dsg.h (graph_analyzer_amd) - drop-in for the (k,triangle)-core part of the reference's dsg.h API (dsg.h:34-58,
// 647-751): Dens and its exact comparisons, dsg_row, CoreLoopResult, the backend-templated level loops
// core_decomp_loop / core_decomp_loop_inc, the GPU backends of gpu_dsg.cu:88-199 over the C ABI, a fused gpu_core_run,
// core_level_density, induced_subgraph / prune_to_core (dsg.h:103-114, 636-645), Greedy++ / Greedy (dsg.h:189-250) and
// the k-truss decomposition on the device. The max-flow part of the paper pipeline (Exact, CoreExact) is not part of
// this library.
#pragma once
#include "densest.h"

typedef unsigned __int128 u128;

struct Dens {
    uint64_t num = 0, den = 0;   // triangles / vertices; den == 0 = no subgraph (density 0)
};
inline int dens_cmp(Dens a, Dens b) {
    if (a.den == 0) a = Dens{0, 1};
    if (b.den == 0) b = Dens{0, 1};
    const u128 l = (u128)a.num * b.den, r = (u128)b.num * a.den;
    return l < r ? -1 : (l > r ? 1 : 0);
}
inline bool dens_gt(Dens a, Dens b) { return dens_cmp(a, b) > 0; }
inline uint64_t dens_ceil(Dens d) { return d.den == 0 ? 0 : (d.num + d.den - 1) / d.den; }
inline double dens_value(Dens d) { return density_value(d.num, d.den); }

// impl,graph,n,m,triangles,vertices,density,time_ms,h2d_ms,total_ms,rounds,kmax (dsg.h:650-663)
inline std::string dsg_row(const char* impl, const char* graph, uint32_t n, uint32_t m, uint64_t tri, uint64_t verts, double time_ms,
                           double h2d_ms, uint64_t rounds, uint64_t kmax) {
    char buf[1024];
    snprintf(buf, sizeof buf, "%s,%s,%u,%u,%llu,%llu,%.6f,%.3f,%.3f,%.3f,%llu,%llu", impl, graph, n, m, (unsigned long long)tri,
             (unsigned long long)verts, density_value(tri, verts), time_ms, h2d_ms, time_ms + h2d_ms, (unsigned long long)rounds,
             (unsigned long long)kmax);
    return buf;
}

struct CoreLoopResult {
    uint32_t rounds = 0, kmax = 0;
    bool ok = true;
    std::string error;
};

// Recount level loop (dsg.h:694-717 semantics): per round begin_round, run_tc, run_min, run_peel(k),
// read_stats(T, removed, k_new); round-0 T == T_full; removed in 1..nalive and the level never decreases.
template <class B>
inline CoreLoopResult core_decomp_loop(uint32_t n, uint64_t T_full, B& be, bool verbose = false, FILE* diag = nullptr) {
    CoreLoopResult r;
    if (T_full >= (1ull << 32)) { r.ok = false; r.error = "triangle count >= 2^32: uint32 tri[] counters could overflow"; return r; }
    be.reset_state();
    uint64_t nalive = n, k = 0;
    for (uint32_t round = 0; nalive > 0; round++) {
        be.begin_round();
        be.run_tc();
        be.run_min();
        be.run_peel(k);
        uint64_t T = 0, removed = 0, k_new = 0;
        be.read_stats(T, removed, k_new);
        if (round == 0 && T != T_full) { r.ok = false; r.error = "device triangle count in round 0 differs from CPU count"; return r; }
        if (verbose && diag)
            fprintf(diag, "round %6u  alive=%10llu  T=%12llu  level k=%llu  removed=%10llu\n", round, (unsigned long long)nalive,
                    (unsigned long long)T, (unsigned long long)k_new, (unsigned long long)removed);
        if (removed == 0 || removed > nalive || k_new < k) {
            r.ok = false;
            r.error = "bad stats from device (removed 0, removed > alive, or level decreased)";
            return r;
        }
        nalive -= removed;
        k = k_new;
        r.rounds = round + 1;
    }
    r.kmax = (uint32_t)k;
    return r;
}

// Incremental level loop (dsg.h:725-751 semantics): init_tri(T) once, then per round begin_round, run_min,
// run_find(k), run_update, run_retire, read_stats(removed, k_new).
template <class B>
inline CoreLoopResult core_decomp_loop_inc(uint32_t n, uint64_t T_full, B& be, bool verbose = false, FILE* diag = nullptr) {
    CoreLoopResult r;
    if (T_full >= (1ull << 32)) { r.ok = false; r.error = "triangle count >= 2^32: uint32 tri[] counters could overflow"; return r; }
    be.reset_state();
    uint64_t T0 = 0;
    be.init_tri(T0);
    if (T0 != T_full) { r.ok = false; r.error = "device triangle count differs from CPU count"; return r; }
    uint64_t nalive = n, k = 0;
    for (uint32_t round = 0; nalive > 0; round++) {
        be.begin_round();
        be.run_min();
        be.run_find(k);
        be.run_update();
        be.run_retire();
        uint64_t removed = 0, k_new = 0;
        be.read_stats(removed, k_new);
        if (verbose && diag)
            fprintf(diag, "round %6u  alive=%10llu  level k=%llu  removed=%10llu\n", round, (unsigned long long)nalive,
                    (unsigned long long)k_new, (unsigned long long)removed);
        if (removed == 0 || removed > nalive || k_new < k) {
            r.ok = false;
            r.error = "bad stats from device (removed 0, removed > alive, or level decreased)";
            return r;
        }
        nalive -= removed;
        k = k_new;
        r.rounds = round + 1;
    }
    r.kmax = (uint32_t)k;
    return r;
}

// Recount backend (gpu_dsg.cu:88-134) over ga_core_* with incremental = 0.
struct GpuCoreBackend {
    ga_ctx* ctx = nullptr;
    ga_core* c = nullptr;
    uint32_t n = 0;
    void init(const gaa::DeviceGraph& g) {
        ctx = g.ctx();
        n = g.n();
        gaa::check(ctx, ga_core_create(ctx, g.handle(), 0, &c), "ga_core_create");
    }
    ~GpuCoreBackend() { if (c) ga_core_destroy(c); }
    void reset_state() { gaa::check(ctx, ga_core_reset_state(c), "ga_core_reset_state"); }
    void begin_round() { gaa::check(ctx, ga_core_begin_round(c), "ga_core_begin_round"); }
    void run_tc() { gaa::check(ctx, ga_core_run_tc(c), "ga_core_run_tc"); }
    void run_min() { gaa::check(ctx, ga_core_run_min(c), "ga_core_run_min"); }
    void run_peel(uint64_t k) { gaa::check(ctx, ga_core_run_peel(c, k), "ga_core_run_peel"); }
    void read_stats(uint64_t& T, uint64_t& removed, uint64_t& k) { gaa::check(ctx, ga_core_read_stats(c, &T, &removed, &k), "ga_core_read_stats"); }
    void download_core(std::vector<uint32_t>& out) {
        out.assign(n, 0);
        gaa::check(ctx, ga_core_download(c, n ? out.data() : nullptr), "ga_core_download");
    }
};

// Incremental backend (gpu_dsg.cu:138-199) over ga_core_* with incremental = 1.
struct GpuCoreBackendInc {
    ga_ctx* ctx = nullptr;
    ga_core* c = nullptr;
    uint32_t n = 0;
    void init(const gaa::DeviceGraph& g) {
        ctx = g.ctx();
        n = g.n();
        gaa::check(ctx, ga_core_create(ctx, g.handle(), 1, &c), "ga_core_create");
    }
    ~GpuCoreBackendInc() { if (c) ga_core_destroy(c); }
    void reset_state() { gaa::check(ctx, ga_core_reset_state(c), "ga_core_reset_state"); }
    void init_tri(uint64_t& T) { gaa::check(ctx, ga_core_init_tri(c, &T), "ga_core_init_tri"); }
    void begin_round() { gaa::check(ctx, ga_core_begin_round(c), "ga_core_begin_round"); }
    void run_min() { gaa::check(ctx, ga_core_run_min(c), "ga_core_run_min"); }
    void run_find(uint64_t k) { gaa::check(ctx, ga_core_run_find(c, k), "ga_core_run_find"); }
    void run_update() { gaa::check(ctx, ga_core_run_update(c), "ga_core_run_update"); }
    void run_retire() { gaa::check(ctx, ga_core_run_retire(c), "ga_core_run_retire"); }
    void read_stats(uint64_t& removed, uint64_t& k) {
        uint64_t T = 0;
        gaa::check(ctx, ga_core_read_stats(c, &T, &removed, &k), "ga_core_read_stats");
    }
    void download_core(std::vector<uint32_t>& out) {
        out.assign(n, 0);
        gaa::check(ctx, ga_core_download(c, n ? out.data() : nullptr), "ga_core_download");
    }
    // prune_to_core on the core numbers of the last finished loop, without downloading them (ga_core_prune).
    gaa::DeviceGraph prune(uint64_t k, std::vector<uint32_t>& old_id, std::vector<uint32_t>& core_out, float* ms = nullptr) {
        ga_graph* h = nullptr;
        old_id.assign(std::max<uint32_t>(n, 1), 0);
        core_out.assign(std::max<uint32_t>(n, 1), 0);
        float t = 0;
        gaa::check(ctx, ga_core_prune(c, k, &h, old_id.data(), core_out.data(), &t), "ga_core_prune");
        gaa::DeviceGraph d(ctx, h);
        old_id.resize(d.n());
        core_out.resize(d.n());
        if (ms) *ms = t;
        return d;
    }
};

// induced_subgraph (dsg.h:103-114) of a resident graph: the subgraph on {v : keep[v] != 0}, new ids in old-id order,
// oriented by the degrees inside the subgraph; built and kept in HBM (ga_graph_induced).
inline gaa::DeviceGraph induced_subgraph_device(const gaa::DeviceGraph& g, const std::vector<uint8_t>& keep, std::vector<uint32_t>& old_id,
                                                float* ms = nullptr) {
    if (keep.size() != g.n()) gaa::die(g.ctx(), GA_ERR_ARG, "induced_subgraph: keep must have n entries");
    ga_graph* h = nullptr;
    old_id.assign(std::max<uint32_t>(g.n(), 1), 0);
    float t = 0;
    gaa::check(g.ctx(), ga_graph_induced(g.ctx(), g.handle(), g.n() ? keep.data() : nullptr, &h, old_id.data(), &t), "ga_graph_induced");
    gaa::DeviceGraph d(g.ctx(), h);
    old_id.resize(d.n());
    if (ms) *ms = t;
    return d;
}

// prune_to_core (dsg.h:636-645) of a resident graph with host core numbers (ga_graph_prune_to_core).
inline gaa::DeviceGraph prune_to_core_device(const gaa::DeviceGraph& g, const std::vector<uint32_t>& core, uint64_t k,
                                             std::vector<uint32_t>& old_id, std::vector<uint32_t>& core_out, float* ms = nullptr) {
    if (core.size() != g.n()) gaa::die(g.ctx(), GA_ERR_ARG, "prune_to_core: core must have n entries");
    ga_graph* h = nullptr;
    old_id.assign(std::max<uint32_t>(g.n(), 1), 0);
    core_out.assign(std::max<uint32_t>(g.n(), 1), 0);
    float t = 0;
    gaa::check(g.ctx(), ga_graph_prune_to_core(g.ctx(), g.handle(), g.n() ? core.data() : nullptr, k, &h, old_id.data(), core_out.data(), &t),
               "ga_graph_prune_to_core");
    gaa::DeviceGraph d(g.ctx(), h);
    old_id.resize(d.n());
    core_out.resize(d.n());
    if (ms) *ms = t;
    return d;
}

// The reference's host signatures (dsg.h:103, 636): upload, build on the device, download.
inline void induced_subgraph(const Graph& g, const std::vector<uint8_t>& keep, Graph& out, std::vector<uint32_t>& old_id) {
    gaa::DeviceGraph d = gaa::DeviceGraph::upload(gaa::default_ctx(), g);
    induced_subgraph_device(d, keep, old_id).download(out);
}

inline void prune_to_core(const Graph& g, const std::vector<uint32_t>& core, uint64_t k, Graph& out, std::vector<uint32_t>& old_id,
                          std::vector<uint32_t>& core_out) {
    gaa::DeviceGraph d = gaa::DeviceGraph::upload(gaa::default_ctx(), g);
    prune_to_core_device(d, core, k, old_id, core_out).download(out);
}

// Whole incremental decomposition with the level loop on the device (ga_core_run).
inline CoreLoopResult gpu_core_run(const gaa::DeviceGraph& g, uint64_t T_full, std::vector<uint32_t>& core, float* ms = nullptr) {
    CoreLoopResult r;
    ga_core* c = nullptr;
    gaa::check(g.ctx(), ga_core_create(g.ctx(), g.handle(), 1, &c), "ga_core_create");
    uint32_t st[2] = {0, 0};
    float t = 0;
    const int rc = ga_core_run(c, T_full, st, &t);
    if (rc != GA_OK) {
        r.ok = false;
        r.error = ga_ctx_last_error(g.ctx());
    } else {
        r.rounds = st[0];
        r.kmax = st[1];
        core.assign(g.n(), 0);
        gaa::check(g.ctx(), ga_core_download(c, g.n() ? core.data() : nullptr), "ga_core_download");
    }
    ga_core_destroy(c);
    if (ms) *ms = t;
    return r;
}

// k-truss number of every oriented edge (index into col) on the device (ga_truss_*); kmax is 0 for a graph without
// edges. No reference function: it extends the per-edge supports of for_each_triangle (dsg.h:62-94).
inline CoreLoopResult gpu_truss_decomposition(const gaa::DeviceGraph& g, std::vector<uint32_t>& truss, float* ms = nullptr) {
    CoreLoopResult r;
    ga_truss* t = nullptr;
    gaa::check(g.ctx(), ga_truss_create(g.ctx(), g.handle(), &t), "ga_truss_create");
    uint32_t st[2] = {0, 0};
    float tm = 0;
    const int rc = ga_truss_run(t, st, &tm);
    if (rc != GA_OK) {
        r.ok = false;
        r.error = ga_ctx_last_error(g.ctx());
    } else {
        r.rounds = st[0];
        r.kmax = st[1];
        truss.assign(g.m(), 0);
        gaa::check(g.ctx(), ga_truss_download(t, g.m() ? truss.data() : nullptr), "ga_truss_download");
    }
    ga_truss_destroy(t);
    if (ms) *ms = tm;
    return r;
}

// Result record of the paper algorithms (dsg.h:50-60); this library fills it from Greedy++.
struct DsgResult {
    uint64_t tri = 0, verts = 0;          // best subgraph: triangles / vertices (0/0 = graph without triangles)
    std::vector<uint32_t> set;            // its vertices, ascending
    std::vector<Dens> per_pass;           // best density after pass 1, 2, ..., T
    uint32_t flow_calls = 0;              // core_exact / exact_densest: number of max-flow computations
    uint32_t components_searched = 0;     // core_exact: components that entered the binary search
    uint32_t shrinks = 0;                 // core_exact: times the network was rebuilt on a higher core
    bool ok = true;
    std::string error;
    Dens dens() const { return Dens{tri, verts}; }
};

// Greedy++ with `passes` passes on a resident graph (ga_greedy_pp): one workgroup runs every pass, with the per-vertex
// state in LDS for graphs of up to ga_greedy_max_vertices (the pruned graph of gpu_dsg.cu:344-356) and in HBM with
// block minima for larger ones (the whole graph of gpu_dsg.cu:318-330). A failed self-check, or T >= 2^32 on a large
// graph, gives ok = false and the library's message in error.
inline DsgResult greedy_pp_device(const gaa::DeviceGraph& g, uint32_t passes, float* ms = nullptr) {
    DsgResult r;
    std::vector<uint64_t> pp(2 * (size_t)std::max<uint32_t>(passes, 1));
    r.set.resize(std::max<uint32_t>(g.n(), 1));
    uint32_t cnt = 0;
    float t = 0;
    const int rc = ga_greedy_pp(g.ctx(), g.handle(), passes, &r.tri, &r.verts, pp.data(), r.set.data(), &cnt, &t);
    if (rc != GA_OK) {
        r.ok = false;
        r.error = ga_ctx_last_error(g.ctx());
        r.set.clear();
        return r;
    }
    r.set.resize(cnt);
    for (uint32_t i = 0; i < passes; i++) r.per_pass.push_back(Dens{pp[2 * i], pp[2 * i + 1]});
    if (ms) *ms = t;
    return r;
}

// The reference's host signatures (dsg.h:189, 250): upload, run on the device.
inline DsgResult greedy_pp(const Graph& g, uint32_t passes) {
    return greedy_pp_device(gaa::DeviceGraph::upload(gaa::default_ctx(), g), passes);
}
inline DsgResult greedy(const Graph& g) { return greedy_pp(g, 1); }

// Triangles and vertices of {v : core[v] >= k} (dsg.h:169-174), counted on the device.
inline Dens core_level_density(const gaa::DeviceGraph& g, const std::vector<uint32_t>& core, uint32_t k) {
    std::vector<uint8_t> in(g.n());
    uint64_t cnt = 0;
    for (uint32_t v = 0; v < g.n(); v++) { in[v] = core[v] >= k; cnt += in[v]; }
    std::vector<uint32_t> tri(std::max<uint32_t>(g.n(), 1));
    uint64_t T = 0;
    gaa::check(g.ctx(), ga_tc_vertex(g.ctx(), g.handle(), GA_TC_AUTO, g.n() ? in.data() : nullptr, tri.data(), &T), "ga_tc_vertex");
    return Dens{T, cnt};
}

// core_decomposition (dsg.h:116-166): the (k,triangle)-core number of every vertex, computed on the device by the
// incremental level loop (ga_core_run; core numbers are unique, so they equal Alg 3's). Returns false where the device
// path cannot hold the counts in u32 (T >= 2^32; the reference fails only when a core number itself exceeds u32).
inline bool core_decomposition(const Graph& g, std::vector<uint32_t>& core, uint32_t* kmax_out = nullptr) {
    gaa::DeviceGraph d = gaa::DeviceGraph::upload(gaa::default_ctx(), g);
    const CoreLoopResult r = gpu_core_run(d, UINT64_MAX, core);
    if (!r.ok) return false;
    if (kmax_out) *kmax_out = (uint32_t)r.kmax;
    return true;
}

// CoreExact (dsg.h:454-627) on a resident graph: ga_core_exact runs the triangle passes and the induced subgraphs on the
// device and the max-flow on the host. Same arguments, checks and result record as the reference (flow_calls,
// components_searched and shrinks included).
inline DsgResult core_exact_device(const gaa::DeviceGraph& g, const std::vector<uint32_t>& core, bool use_pruning12 = true,
                                   const DsgResult* seed = nullptr) {
    DsgResult r;
    if (core.size() != g.n()) {
        r.ok = false;
        r.error = "core[] has the wrong size";
        return r;
    }
    const bool use_seed = seed && seed->ok && seed->verts > 0;
    if (use_seed && seed->set.size() != seed->verts) {
        r.ok = false;
        r.error = "seed is not a valid subgraph of the graph";
        return r;
    }
    const uint64_t tv[2] = {use_seed ? seed->tri : 0, use_seed ? seed->verts : 0};
    uint64_t out[7] = {};
    r.set.resize(std::max<uint32_t>(g.n(), 1));
    const int rc = ga_core_exact(g.ctx(), g.handle(), core.data(), use_pruning12 ? 1 : 0, use_seed ? tv : nullptr,
                                 use_seed ? seed->set.data() : nullptr, out, r.set.data());
    r.flow_calls = (uint32_t)out[4];
    r.components_searched = (uint32_t)out[5];
    r.shrinks = (uint32_t)out[6];
    if (rc != GA_OK) {
        r.ok = false;
        r.error = ga_ctx_last_error(g.ctx());
        r.set.clear();
        return r;
    }
    r.tri = out[0];
    r.verts = out[1];
    r.set.resize(out[2]);
    return r;
}

// The reference's host signatures (dsg.h:402, 454): upload, run.
inline DsgResult core_exact(const Graph& g, const std::vector<uint32_t>& core, bool use_pruning12 = true,
                            const DsgResult* seed = nullptr) {
    return core_exact_device(gaa::DeviceGraph::upload(gaa::default_ctx(), g), core, use_pruning12, seed);
}

inline DsgResult exact_densest(const Graph& g) {
    const gaa::DeviceGraph d = gaa::DeviceGraph::upload(gaa::default_ctx(), g);
    DsgResult r;
    uint64_t out[7] = {};
    r.set.resize(std::max<uint32_t>(d.n(), 1));
    const int rc = ga_exact_densest(d.ctx(), d.handle(), out, r.set.data());
    r.flow_calls = (uint32_t)out[4];
    if (rc != GA_OK) {
        r.ok = false;
        r.error = ga_ctx_last_error(d.ctx());
        r.set.clear();
        return r;
    }
    r.tri = out[0];
    r.verts = out[1];
    r.set.resize(out[2]);
    return r;
}
