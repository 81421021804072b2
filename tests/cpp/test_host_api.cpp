// test_host_api.cpp - the drop-in headers' host-signature functions that run on the GPU, called exactly as a reference
// caller calls them and compared with the CPU oracle:
//   build_tasks(g, chunk, t_edge, t_start)     graph.h:138-150
//   build_edge_order(g, perm)                  graph.h:156-171
//   build_undirected_csr(g, urow, ucol)        densest.h:162-171
//   core_decomposition(g, core, &kmax)         dsg.h:116-166
// on the graph of the text file given as the argument (loaded with load_graph, graph.h:71-96) and on seeded random
// graphs with hubs and isolated vertices. Prints ALL OK and exits 0 when every result matches; exits 1 otherwise.
#include <random>

#include "../../graph_analyzer_amd/host/dsg.h"
#include "../../oracle/oracle.hpp"

static int g_fail = 0;
#define CHECK(c, ...) do { if (!(c)) { g_fail++; if (g_fail < 30) { printf("FAIL %s:%d: ", __FILE__, __LINE__); printf(__VA_ARGS__); printf("\n"); } } } while (0)

static orc::Graph to_oracle(const Graph& g) {
    orc::Graph o;
    o.n = g.n;
    o.m = g.m;
    o.rowptr = g.rowptr;
    o.col = g.col;
    o.src = g.src;
    o.max_out_degree = g.max_out_degree;
    return o;
}

static void check_graph(const Graph& g, const char* tag) {
    const orc::Graph o = to_oracle(g);
    for (uint32_t chunk : {1u, 2u, 7u, 128u}) {
        std::vector<uint32_t> te, ts, ote, ots;
        build_tasks(g, chunk, te, ts);
        orc::build_tasks(o, chunk, ote, ots);
        CHECK(te == ote && ts == ots, "%s: build_tasks(chunk %u): %zu vs %zu tasks", tag, chunk, te.size(), ote.size());
    }
    std::vector<uint32_t> perm, operm;
    build_edge_order(g, perm);
    orc::edge_order(o, operm);
    CHECK(perm == operm, "%s: build_edge_order", tag);
    std::vector<uint32_t> ur, uc, our, ouc;
    build_undirected_csr(g, ur, uc);
    orc::undirected_csr(o, our, ouc);
    CHECK(ur == our && uc == ouc, "%s: build_undirected_csr", tag);
    std::vector<uint32_t> core, ocore;
    uint32_t kmax = 0, okmax = 0;
    const bool ok = core_decomposition(g, core, &kmax);
    const bool ook = orc::core_decomposition(o, ocore, &okmax);
    CHECK(ok && ook && core == ocore && kmax == okmax, "%s: core_decomposition (kmax %u vs %u)", tag, kmax, okmax);
}

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s graph.txt\n", argv[0]);
        return 2;
    }
    Graph g;
    if (!load_graph(argv[1], g)) return 1;
    check_graph(g, argv[1]);
    std::mt19937_64 rng(77);
    for (int t = 0; t < 6; t++) {
        const uint32_t n = 50 + (uint32_t)(rng() % 3000);
        std::vector<std::pair<uint32_t, uint32_t>> e;
        std::uniform_real_distribution<double> U(0, 1);
        const double p = 0.002 + 0.03 * U(rng);
        for (uint32_t a = 0; a + 1 < n; a++)
            for (uint32_t b = a + 1; b < n; b++)
                if (U(rng) < p || (a == 3 && b % 3 == 0)) e.emplace_back(a, b);   // vertex 3 is a hub
        Graph r;
        build_graph_from_edges(n + 5, e, r);                                      // 5 isolated vertices at the end
        char tag[64];
        snprintf(tag, sizeof tag, "random %d (n %u, m %u)", t, r.n, r.m);
        check_graph(r, tag);
    }
    Graph empty;
    build_graph_from_edges(4, {}, empty);
    check_graph(empty, "no edges");
    if (g_fail) {
        printf("%d failures\n", g_fail);
        return 1;
    }
    printf("ALL OK\n");
    return 0;
}
