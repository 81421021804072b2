// test_host_api.cpp - the drop-in headers' host-signature functions that run on the GPU, called exactly as a reference
// caller calls them and compared with the CPU oracle:
//   build_tasks(g, chunk, t_edge, t_start)     graph.h:138-150
//   build_edge_order(g, perm)                  graph.h:156-171
//   build_undirected_csr(g, urow, ucol)        densest.h:162-171
//   core_decomposition(g, core, &kmax)         dsg.h:116-166
//   greedy(g), greedy_pp(g, passes)            dsg.h:189-250, on a graph above the one-workgroup (LDS) size as well
// on the graph of the text file given as the argument (loaded with load_graph, graph.h:71-96) and on seeded random
// graphs with hubs and isolated vertices. Prints ALL OK and exits 0 when every result matches; exits 1 otherwise.
#include <random>
#include <set>

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
    {   // Greedy++ through the drop-in signatures: a graph of the pruned size and one of 20,000 vertices (the whole-graph
        // call of gpu_dsg.cu:318-330, above ga_greedy_max_vertices), against the oracle's Greedy++
        for (uint32_t n : {3000u, 20000u}) {
            std::set<std::pair<uint32_t, uint32_t>> es;   // build_graph_from_edges takes unique pairs
            std::uniform_int_distribution<uint32_t> V(0, n - 1);
            for (uint32_t i = 0; i < 5 * n; i++) {
                const uint32_t a = V(rng), b = V(rng);
                if (a != b) es.emplace(std::min(a, b), std::max(a, b));
            }
            for (uint32_t a = 0; a < 25; a++)
                for (uint32_t b = a + 1; b < 25; b++) es.emplace(n - 30 + a, n - 30 + b);   // a planted clique
            const std::vector<std::pair<uint32_t, uint32_t>> e(es.begin(), es.end());
            Graph r;
            build_graph_from_edges(n, e, r);
            const orc::Graph o = to_oracle(r);
            const DsgResult g1 = greedy(r);
            const orc::GreedyResult o1 = orc::greedy_pp(o, 1);
            CHECK(g1.ok && g1.tri == o1.tri && g1.verts == o1.verts && g1.set == o1.set, "greedy(g) on n %u: %s", n, g1.error.c_str());
            const DsgResult g3 = greedy_pp(r, 3);
            const orc::GreedyResult o3 = orc::greedy_pp(o, 3);
            bool pp_ok = g3.ok && g3.per_pass.size() == o3.per_pass.size();
            for (size_t i = 0; pp_ok && i < o3.per_pass.size(); i++)
                pp_ok = g3.per_pass[i].num == o3.per_pass[i].first && g3.per_pass[i].den == o3.per_pass[i].second;
            CHECK(pp_ok && g3.tri == o3.tri && g3.verts == o3.verts && g3.set == o3.set, "greedy_pp(g, 3) on n %u: %s", n, g3.error.c_str());
        }
    }
    if (g_fail) {
        printf("%d failures\n", g_fail);
        return 1;
    }
    printf("ALL OK\n");
    return 0;
}
