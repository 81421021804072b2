// graph.h (graph_analyzer_amd) - drop-in for the reference's graph.h API (graph.h:19-171), backed by the GPU.
//
// Same names and argument meaning: struct Graph, now_ms, build_graph_from_edges, load_graph, build_tasks,
// build_edge_order.
// Differences a caller sees:
//   * build_graph_from_edges / load_graph do de-duplication, (degree, id) orientation and row sorting on the GPU
//     (ga_graph_from_pairs); load_graph parses the text on the GPU too (ga_graph_load_text). The arrays are
//     byte-identical to the reference's.
//   * the CPU counters cpu_triangles_serial / cpu_triangles_omp are not part of this library (they are the CPU
//     baseline); gpu_triangles() counts on the device.
//   * runtime errors print the HIP message and exit(1), like the reference's CUDA_CHECK (gpu_tc.cu:29-37).
// All calls without an explicit context use one process-wide context on HIP device 0 (gaa::default_ctx()).
#pragma once
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ga_amd.h"

struct Graph {
    uint32_t n = 0;
    uint32_t m = 0;
    std::vector<uint32_t> rowptr;   // n + 1
    std::vector<uint32_t> col;      // m, sorted within each row
    std::vector<uint32_t> src;      // m, row owner of each edge
    uint32_t max_out_degree = 0;
};

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace gaa {

inline void die(ga_ctx* ctx, int rc, const char* what) {
    fprintf(stderr, "graph_analyzer_amd error %d in %s: %s\n", rc, what, ctx ? ga_ctx_last_error(ctx) : "no context");
    exit(1);
}

inline void check(ga_ctx* ctx, int rc, const char* what) {
    if (rc != GA_OK) die(ctx, rc, what);
}

// Process-wide context on device 0 (created on first use, never destroyed).
inline ga_ctx* default_ctx() {
    static ga_ctx* ctx = nullptr;
    if (!ctx) {
        const int rc = ga_ctx_create(0, &ctx);
        if (rc != GA_OK) {
            fprintf(stderr, "graph_analyzer_amd: no usable HIP device (ga_ctx_create returned %d)\n", rc);
            exit(1);
        }
    }
    return ctx;
}

// Oriented CSR resident in HBM. Move-only owner of a ga_graph.
class DeviceGraph {
   public:
    DeviceGraph() = default;
    DeviceGraph(ga_ctx* ctx, ga_graph* h) : ctx_(ctx), h_(h) {
        if (h_) ga_graph_info(h_, &n_, &m_, &max_out_);
    }
    DeviceGraph(const DeviceGraph&) = delete;
    DeviceGraph& operator=(const DeviceGraph&) = delete;
    DeviceGraph(DeviceGraph&& o) noexcept { *this = std::move(o); }
    DeviceGraph& operator=(DeviceGraph&& o) noexcept {
        if (this != &o) {
            reset();
            ctx_ = o.ctx_; h_ = o.h_; n_ = o.n_; m_ = o.m_; max_out_ = o.max_out_; build_ms_ = o.build_ms_;
            o.h_ = nullptr;
        }
        return *this;
    }
    ~DeviceGraph() { reset(); }
    void reset() {
        if (h_) ga_graph_destroy(h_);
        h_ = nullptr;
    }

    // load_graph semantics when n < 0, build_graph_from_edges semantics otherwise (ga_graph_from_pairs).
    static DeviceGraph from_pairs(ga_ctx* ctx, const uint32_t* pairs, uint64_t npairs, int64_t n) {
        ga_graph* h = nullptr;
        float ms = 0;
        check(ctx, ga_graph_from_pairs(ctx, pairs, npairs, n, &h, &ms), "ga_graph_from_pairs");
        DeviceGraph d(ctx, h);
        d.build_ms_ = ms;
        return d;
    }
    // The reference's upload() of rowptr/col (gpu_tc.cu:42-62); src is derived on the device.
    static DeviceGraph upload(ga_ctx* ctx, const Graph& g) {
        ga_graph* h = nullptr;
        check(ctx, ga_graph_from_csr(ctx, g.n, g.m, g.rowptr.data(), g.col.empty() ? nullptr : g.col.data(), &h), "ga_graph_from_csr");
        return DeviceGraph(ctx, h);
    }
    void download(Graph& g) const {
        g.n = n_;
        g.m = m_;
        g.max_out_degree = max_out_;
        g.rowptr.assign((size_t)n_ + 1, 0);
        g.col.assign(m_, 0);
        g.src.assign(m_, 0);
        check(ctx_, ga_graph_download(ctx_, h_, g.rowptr.data(), m_ ? g.col.data() : nullptr, m_ ? g.src.data() : nullptr), "ga_graph_download");
    }

    ga_ctx* ctx() const { return ctx_; }
    ga_graph* handle() const { return h_; }
    uint32_t n() const { return n_; }
    uint32_t m() const { return m_; }
    uint32_t max_out_degree() const { return max_out_; }
    float build_ms() const { return build_ms_; }

   private:
    ga_ctx* ctx_ = nullptr;
    ga_graph* h_ = nullptr;
    uint32_t n_ = 0, m_ = 0, max_out_ = 0;
    float build_ms_ = 0;
};

// T on the device with the chosen variant (ga_tc_count); *ms = device time of the counting launch(es).
inline uint64_t gpu_triangles(const DeviceGraph& g, int variant = GA_TC_AUTO, float* ms = nullptr) {
    uint64_t T = 0;
    float t = 0;
    check(g.ctx(), ga_tc_count(g.ctx(), g.handle(), variant, &T, &t), "ga_tc_count");
    if (ms) *ms = t;
    return T;
}

// Pairs as the flat (a, b, a, b, ...) array the C ABI takes.
inline std::vector<uint32_t> flatten(const std::vector<std::pair<uint32_t, uint32_t>>& edges) {
    std::vector<uint32_t> p(2 * edges.size());
    for (size_t i = 0; i < edges.size(); i++) { p[2 * i] = edges[i].first; p[2 * i + 1] = edges[i].second; }
    return p;
}

// SNAP text parse of the reference's load_graph (graph.h:76-89): skips lines starting '#', '%' or newline, needs two
// base-10 numbers per line, drops self loops, ids cast to uint32.
inline bool parse_edge_file(const char* path, std::vector<uint32_t>& pairs) {
    FILE* f = fopen(path, "r");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); return false; }
    pairs.clear();
    char line[512];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '%' || line[0] == '\n') continue;
        char* e1;
        const unsigned long a = strtoul(line, &e1, 10);
        if (e1 == line) continue;
        char* e2;
        const unsigned long b = strtoul(e1, &e2, 10);
        if (e2 == e1 || a == b) continue;
        pairs.push_back((uint32_t)std::min(a, b));
        pairs.push_back((uint32_t)std::max(a, b));
    }
    fclose(f);
    return true;
}

}  // namespace gaa

// build_graph_from_edges (graph.h:38-67): edges unique and loop-free on vertices 0..n-1.
inline void build_graph_from_edges(uint32_t n, const std::vector<std::pair<uint32_t, uint32_t>>& edges, Graph& g) {
    ga_ctx* ctx = gaa::default_ctx();
    const std::vector<uint32_t> p = gaa::flatten(edges);
    gaa::DeviceGraph d = gaa::DeviceGraph::from_pairs(ctx, p.data(), edges.size(), (int64_t)n);
    d.download(g);
}

// load_graph (graph.h:71-96): text parse, canonicalisation, de-duplication and orientation on the GPU. A file that
// cannot be opened prints "cannot open" and returns false, as in the reference.
inline bool load_graph_device(const char* path, gaa::DeviceGraph& out, ga_ctx* ctx = nullptr) {
    FILE* f = fopen(path, "r");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); return false; }
    fclose(f);
    ctx = ctx ? ctx : gaa::default_ctx();
    ga_graph* h = nullptr;
    float ms = 0;
    gaa::check(ctx, ga_graph_load_text(ctx, path, &h, &ms), "ga_graph_load_text");
    out = gaa::DeviceGraph(ctx, h);
    return true;
}

inline bool load_graph(const char* path, Graph& g) {
    gaa::DeviceGraph d;
    if (!load_graph_device(path, d)) return false;
    d.download(g);
    return true;
}

// The library's binary CSR file (ga_graph_save / ga_graph_load): no text parse and no build on admission; the device
// checks that the arrays are a valid (degree, id)-oriented CSR.
inline void save_graph_binary(const gaa::DeviceGraph& g, const char* path) {
    gaa::check(g.ctx(), ga_graph_save(g.ctx(), g.handle(), path), "ga_graph_save");
}

inline gaa::DeviceGraph load_graph_binary(const char* path, ga_ctx* ctx = nullptr) {
    ctx = ctx ? ctx : gaa::default_ctx();
    ga_graph* h = nullptr;
    gaa::check(ctx, ga_graph_load(ctx, path, &h, nullptr), "ga_graph_load");
    return gaa::DeviceGraph(ctx, h);
}

// build_edge_order (graph.h:156-171), computed on the device.
inline void build_edge_order(const Graph& g, std::vector<uint32_t>& perm) {
    ga_ctx* ctx = gaa::default_ctx();
    gaa::DeviceGraph d = gaa::DeviceGraph::upload(ctx, g);
    perm.assign(g.m, 0);
    gaa::check(ctx, ga_graph_edge_order(ctx, d.handle(), g.m ? perm.data() : nullptr), "ga_graph_edge_order");
}

// build_tasks (graph.h:138-150), computed on the device: every oriented edge's work min(d+u, d+v) in chunks of `chunk`
// elements, task (t_edge[t], t_start[t]) in edge order.
inline void build_tasks(const Graph& g, uint32_t chunk, std::vector<uint32_t>& t_edge, std::vector<uint32_t>& t_start) {
    ga_ctx* ctx = gaa::default_ctx();
    gaa::DeviceGraph d = gaa::DeviceGraph::upload(ctx, g);
    uint64_t nt = 0;
    gaa::check(ctx, ga_graph_tasks(ctx, d.handle(), chunk, nullptr, nullptr, 0, &nt), "ga_graph_tasks");
    t_edge.assign(nt, 0);
    t_start.assign(nt, 0);
    if (nt) gaa::check(ctx, ga_graph_tasks(ctx, d.handle(), chunk, t_edge.data(), t_start.data(), nt, &nt), "ga_graph_tasks");
}
