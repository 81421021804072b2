// oracle.hpp - CPU ORACLE for the triangle-count / batch-peel / (k,triangle)-core path.
//
// TEST INFRASTRUCTURE ONLY. This header is the checker: only tests/, the smoke check and the CPU-baseline leg of
// bench.py may include, link or call it. The product library (graph_analyzer_amd) never links it and has no CPU
// fallback; a missing HIP library is an error there, never a silent switch to this code.
//
// Everything here is a plain restatement of the reference's own CPU semantics (graph.h, densest.h, dsg.h of
// graph_analyzer), written from the reference text, plus brute-force oracles for tiny graphs. Every function cites the
// reference lines it follows. Namespace `orc` keeps it apart from the product's drop-in names (Graph, load_graph, ...).
//
// Pinning (how each function is checked, see tests/test_oracle.py):
//   loader / builder      structural invariants, closed forms, brute force, and the reference compiled into oracle/_ref
//   triangle counts       closed forms (K_k, K_ab, wheels), brute force (bit-set adjacency), the reference build
//   participation / peel  brute force on random alive masks, u128 threshold oracle, exhaustive densest (n <= 16)
//   core numbers          exhaustive definition (n <= 12), fixpoint definition, closed form K_k, the reference build
//   per-edge support      brute force, sum = 3T, K_k -> k-2 on every edge
//   greedy_pp             the reference build, exhaustive optimum (n <= 16: never above it, never below a third of it
//                         in pass 1), closed forms (K_k, disjoint cliques), the returned set's own triangle count
#pragma once
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <set>
#include <string>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef unsigned __int128 u128;
typedef std::vector<std::pair<uint32_t, uint32_t>> EdgeList;

// Oriented CSR: every undirected edge {a,b} stored once, from the smaller (degree, id) endpoint (graph.h:15-26).
struct Graph {
    uint32_t n = 0, m = 0;
    std::vector<uint32_t> rowptr, col, src;
    uint32_t max_out_degree = 0;
};

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// (degree, id) orientation and per-row sorted CSR from a duplicate-free, loop-free undirected list
// (graph.h:38-67). Degrees are undirected degrees; rows ascending; src[e] = row owner; max_out_degree over rows.
inline void build_graph_from_edges(uint32_t n, const EdgeList& edges, Graph& g) {
    g.n = n;
    g.m = (uint32_t)edges.size();
    std::vector<uint32_t> deg(n, 0);
    for (const auto& e : edges) { deg[e.first]++; deg[e.second]++; }
    auto lower = [&](uint32_t a, uint32_t b) { return deg[a] < deg[b] || (deg[a] == deg[b] && a < b); };
    g.rowptr.assign((size_t)n + 1, 0);
    for (const auto& e : edges) g.rowptr[(lower(e.first, e.second) ? e.first : e.second) + 1]++;
    for (uint32_t i = 0; i < n; i++) g.rowptr[i + 1] += g.rowptr[i];
    g.col.assign(g.m, 0);
    g.src.assign(g.m, 0);
    std::vector<uint32_t> cur(g.rowptr.begin(), g.rowptr.end() - 1);
    for (const auto& e : edges) {
        const bool f = lower(e.first, e.second);
        const uint32_t from = f ? e.first : e.second, to = f ? e.second : e.first;
        g.col[cur[from]++] = to;
    }
    g.max_out_degree = 0;
    for (uint32_t u = 0; u < n; u++) {
        const uint32_t b = g.rowptr[u], en = g.rowptr[u + 1];
        std::sort(g.col.begin() + b, g.col.begin() + en);
        for (uint32_t e = b; e < en; e++) g.src[e] = u;
        g.max_out_degree = std::max(g.max_out_degree, en - b);
    }
}

// Canonicalisation used by the text loader (graph.h:85-92): drop self loops, (min,max) order, sort + unique.
// Returns max_id over the kept pairs (n = max_id + 1, graph.h:88,94; an empty list gives n = 1).
inline uint32_t canonical_pairs(const uint32_t* pairs, uint64_t k, EdgeList& out) {
    out.clear();
    out.reserve(k);
    uint32_t max_id = 0;
    for (uint64_t i = 0; i < k; i++) {
        const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        if (a == b) continue;
        const uint32_t u = std::min(a, b), v = std::max(a, b);
        out.emplace_back(u, v);
        max_id = std::max(max_id, v);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return max_id;
}

// load_graph semantics on an in-memory pair list (graph.h:71-96 minus the text parsing).
inline void graph_from_pairs(const uint32_t* pairs, uint64_t k, Graph& g) {
    EdgeList e;
    const uint32_t max_id = canonical_pairs(pairs, k, e);
    build_graph_from_edges(max_id + 1, e, g);
}

// SNAP text parsing (graph.h:71-89): lines starting '#', '%' or newline are skipped; a line needs two base-10
// numbers (strtoul); ids are cast to uint32. Then graph_from_pairs.
inline bool load_graph(const char* path, Graph& g) {
    FILE* f = fopen(path, "r");
    if (!f) return false;
    std::vector<uint32_t> pairs;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '%' || line[0] == '\n') continue;
        char* e1;
        const unsigned long a = strtoul(line, &e1, 10);
        if (e1 == line) continue;
        char* e2;
        const unsigned long b = strtoul(e1, &e2, 10);
        if (e2 == e1) continue;
        if (a == b) continue;
        pairs.push_back((uint32_t)std::min(a, b));
        pairs.push_back((uint32_t)std::max(a, b));
    }
    fclose(f);
    graph_from_pairs(pairs.data(), pairs.size() / 2, g);
    return true;
}

// |A ∩ B| of two sorted ranges by two-pointer merge (graph.h:99-108).
inline uint64_t merge_count(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb) {
    uint64_t c = 0;
    uint32_t i = 0, j = 0;
    while (i < na && j < nb) {
        if (a[i] == b[j]) { c++; i++; j++; }
        else if (a[i] < b[j]) i++;
        else j++;
    }
    return c;
}

inline uint32_t outdeg(const Graph& g, uint32_t u) { return g.rowptr[u + 1] - g.rowptr[u]; }

// T = sum over oriented edges (u,v) of |N+(u) ∩ N+(v)| (graph.h:110-121).
inline uint64_t triangles_serial(const Graph& g) {
    uint64_t t = 0;
    for (uint32_t u = 0; u < g.n; u++) {
        const uint32_t* nu = g.col.data() + g.rowptr[u];
        const uint32_t du = outdeg(g, u);
        for (uint32_t k = 0; k < du; k++) t += merge_count(nu, du, g.col.data() + g.rowptr[nu[k]], outdeg(g, nu[k]));
    }
    return t;
}

// Same sum, vertex loop split with OpenMP schedule(dynamic,64) (graph.h:123-136).
inline uint64_t triangles_omp(const Graph& g) {
    uint64_t t = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : t)
    for (long long ui = 0; ui < (long long)g.n; ui++) {
        const uint32_t u = (uint32_t)ui;
        const uint32_t* nu = g.col.data() + g.rowptr[u];
        const uint32_t du = outdeg(g, u);
        for (uint32_t k = 0; k < du; k++) t += merge_count(nu, du, g.col.data() + g.rowptr[nu[k]], outdeg(g, nu[k]));
    }
    return t;
}

// Calls f(u, v, w, e_uv, e_uw, e_vw) once per triangle; e_xy indexes g.col (dsg.h:62-79).
template <class F>
inline void for_each_triangle(const Graph& g, F&& f) {
    for (uint32_t e = 0; e < g.m; e++) {
        const uint32_t u = g.src[e], v = g.col[e];
        uint32_t i = g.rowptr[u], ie = g.rowptr[u + 1], j = g.rowptr[v], je = g.rowptr[v + 1];
        while (i < ie && j < je) {
            const uint32_t a = g.col[i], b = g.col[j];
            if (a == b) { f(u, v, a, e, i, j); i++; j++; }
            else if (a < b) i++;
            else j++;
        }
    }
}

// Per-oriented-edge triangle support: sup[e]++ for the three edges of every triangle (dsg.h:62-79 indices; the
// reference has no kernel for it, its natural CPU form is this enumeration).
inline uint64_t edge_support(const Graph& g, std::vector<uint32_t>& sup) {
    sup.assign(g.m, 0);
    uint64_t T = 0;
    for_each_triangle(g, [&](uint32_t, uint32_t, uint32_t, uint32_t a, uint32_t b, uint32_t c) { T++; sup[a]++; sup[b]++; sup[c]++; });
    return T;
}

// Triangles of the alive-induced subgraph and tri[v] = triangles containing v (densest.h:81-101).
inline uint64_t participation(const Graph& g, const std::vector<uint8_t>& alive, std::vector<uint32_t>& tri) {
    tri.assign(g.n, 0);
    uint64_t T = 0;
    for (uint32_t e = 0; e < g.m; e++) {
        const uint32_t u = g.src[e], v = g.col[e];
        if (!alive[u] || !alive[v]) continue;
        uint32_t i = g.rowptr[u], ie = g.rowptr[u + 1], j = g.rowptr[v], je = g.rowptr[v + 1];
        while (i < ie && j < je) {
            const uint32_t a = g.col[i], b = g.col[j];
            if (a < b) { i++; continue; }
            if (a > b) { j++; continue; }
            if (alive[a]) { T++; tri[u]++; tri[v]++; tri[a]++; }
            i++; j++;
        }
    }
    return T;
}

// OpenMP edge-parallel form with atomic increments (densest.h:104-134).
inline uint64_t participation_omp(const Graph& g, const std::vector<uint8_t>& alive, std::vector<uint32_t>& tri) {
    tri.assign(g.n, 0);
    uint64_t T = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : T)
    for (long long ei = 0; ei < (long long)g.m; ei++) {
        const uint32_t e = (uint32_t)ei, u = g.src[e], v = g.col[e];
        if (!alive[u] || !alive[v]) continue;
        uint32_t i = g.rowptr[u], ie = g.rowptr[u + 1], j = g.rowptr[v], je = g.rowptr[v + 1];
        while (i < ie && j < je) {
            const uint32_t a = g.col[i], b = g.col[j];
            if (a < b) { i++; continue; }
            if (a > b) { j++; continue; }
            if (alive[a]) {
                T++;
#pragma omp atomic
                tri[u]++;
#pragma omp atomic
                tri[v]++;
#pragma omp atomic
                tri[a]++;
            }
            i++; j++;
        }
    }
    return T;
}

// ---- batch peeling (densest.h:46-78, 136-160) ----
struct PeelResult {
    uint64_t best_num = 0, best_den = 0;
    uint32_t best_round = 0, rounds = 0;
    std::vector<uint32_t> removed_round;
    bool ok = true;
    std::string error;
};
inline bool same(const PeelResult& a, const PeelResult& b) {   // densest.h:56-59
    return a.best_num == b.best_num && a.best_den == b.best_den && a.best_round == b.best_round && a.rounds == b.rounds &&
           a.removed_round == b.removed_round;
}

// Overflow guards (densest.h:65-72): 1 <= den <= 2^20, num <= 2^20, T0 < 2^32, T0*n*den and 3(den+num)T0 < 2^64.
inline bool peel_arith_ok(uint32_t n, uint64_t T0, uint64_t num, uint64_t den, std::string* why = nullptr) {
    const char* msg = nullptr;
    if (den == 0 || den > (1ull << 20) || num > (1ull << 20)) msg = "eps = num/den needs 1 <= den <= 2^20 and num <= 2^20";
    else if (T0 >= (1ull << 32)) msg = "triangle count >= 2^32: uint32 tri[] counters could overflow";
    else if (n > 0 && T0 > UINT64_MAX / ((uint64_t)n * den)) msg = "tri*nalive*den could overflow uint64";
    else if (T0 > UINT64_MAX / (3ull * (den + num))) msg = "3*(den+num)*T could overflow uint64";
    if (msg && why) *why = msg;
    return msg == nullptr;
}

// Strict improvement by cross multiplication; ties keep the earlier, larger set (densest.h:76-78).
inline bool peel_improves(uint64_t T, uint64_t nalive, uint64_t bn, uint64_t bd) { return bd == 0 || T * bd > bn * nalive; }

// Rounds of: count alive-induced T and tri[]; record density if strictly better; remove every alive v with
// tri[v]*nalive*den <= 3(den+num)T; removed_round[v] = round (densest.h:138-160).
inline PeelResult batch_peel(const Graph& g, uint64_t num, uint64_t den, bool omp = false) {
    PeelResult r;
    r.removed_round.assign(g.n, UINT32_MAX);
    std::vector<uint8_t> alive(g.n, 1);
    std::vector<uint32_t> tri;
    uint64_t nalive = g.n;
    for (uint32_t round = 0; nalive > 0; round++) {
        const uint64_t T = omp ? participation_omp(g, alive, tri) : participation(g, alive, tri);
        if (round == 0 && !peel_arith_ok(g.n, T, num, den, &r.error)) { r.ok = false; return r; }
        if (peel_improves(T, nalive, r.best_num, r.best_den)) { r.best_num = T; r.best_den = nalive; r.best_round = round; }
        const uint64_t rhs = 3ull * (den + num) * T;
        uint64_t removed = 0;
        for (uint32_t v = 0; v < g.n; v++) {
            if (alive[v] && (uint64_t)tri[v] * nalive * den <= rhs) { alive[v] = 0; r.removed_round[v] = round; removed++; }
        }
        if (removed == 0) { r.ok = false; r.error = "no vertex removed in a round"; return r; }
        nalive -= removed;
        r.rounds = round + 1;
    }
    return r;
}

// Both-direction sorted adjacency (densest.h:163-171).
inline void undirected_csr(const Graph& g, std::vector<uint32_t>& urow, std::vector<uint32_t>& ucol) {
    urow.assign((size_t)g.n + 1, 0);
    for (uint32_t e = 0; e < g.m; e++) { urow[g.src[e] + 1]++; urow[g.col[e] + 1]++; }
    for (uint32_t i = 0; i < g.n; i++) urow[i + 1] += urow[i];
    ucol.assign(2ull * g.m, 0);
    std::vector<uint32_t> cur(urow.begin(), urow.end() - 1);
    for (uint32_t e = 0; e < g.m; e++) { ucol[cur[g.src[e]]++] = g.col[e]; ucol[cur[g.col[e]]++] = g.src[e]; }
    for (uint32_t u = 0; u < g.n; u++) std::sort(ucol.begin() + urow[u], ucol.begin() + urow[u + 1]);
}

// Calls f(w) for each common element w > lo of two sorted ranges (densest.h:176-191: gallop when sizes differ by
// more than 16x, else merge; the result is the same set either way).
template <class F>
inline void common_above(const uint32_t* A, uint32_t na, const uint32_t* B, uint32_t nb, uint32_t lo, F&& f) {
    const uint32_t* a = std::upper_bound(A, A + na, lo);
    const uint32_t* ae = A + na;
    const uint32_t* b = std::upper_bound(B, B + nb, lo);
    const uint32_t* be = B + nb;
    const size_t sa = (size_t)(ae - a), sb = (size_t)(be - b);
    if (sa * 16 < sb) {
        for (; a < ae; ++a) { b = std::lower_bound(b, be, *a); if (b == be) break; if (*b == *a) f(*a); }
    } else if (sb * 16 < sa) {
        for (; b < be; ++b) { a = std::lower_bound(a, ae, *b); if (a == ae) break; if (*a == *b) f(*b); }
    } else {
        while (a < ae && b < be) {
            if (*a < *b) ++a;
            else if (*a > *b) ++b;
            else { f(*a); ++a; ++b; }
        }
    }
}

// Triangle degree of every vertex over the whole graph (dsg.h:82-87).
inline uint64_t triangle_degrees(const Graph& g, std::vector<uint64_t>& deg) {
    deg.assign(g.n, 0);
    uint64_t T = 0;
    for_each_triangle(g, [&](uint32_t u, uint32_t v, uint32_t w, uint32_t, uint32_t, uint32_t) { T++; deg[u]++; deg[v]++; deg[w]++; });
    return T;
}

// (k,triangle)-core numbers, DSD Alg 3 with a lazy min-heap on (degree, id) (dsg.h:125-166): pop the minimum,
// core[v] = its degree, and for each triangle {v,a,b} with a,b alive decrement a and b only if their degree exceeds
// core[v]. Returns false if a degree does not fit uint32.
inline bool core_decomposition(const Graph& g, std::vector<uint32_t>& core, uint32_t* kmax_out = nullptr) {
    std::vector<uint32_t> urow, ucol;
    undirected_csr(g, urow, ucol);
    std::vector<uint64_t> deg;
    triangle_degrees(g, deg);
    core.assign(g.n, 0);
    std::vector<uint8_t> alive(g.n, 1), mark(g.n, 0);
    typedef std::pair<uint64_t, uint32_t> Item;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    for (uint32_t v = 0; v < g.n; v++) pq.emplace(deg[v], v);
    uint32_t kmax = 0, done = 0;
    std::vector<uint32_t> touched;
    while (done < g.n) {
        const Item it = pq.top();
        pq.pop();
        const uint32_t v = it.second;
        if (!alive[v] || deg[v] != it.first) continue;
        if (deg[v] > UINT32_MAX) return false;
        core[v] = (uint32_t)deg[v];
        kmax = std::max(kmax, core[v]);
        alive[v] = 0;
        done++;
        touched.clear();
        const uint32_t* nv = ucol.data() + urow[v];
        const uint32_t dv = urow[v + 1] - urow[v];
        for (uint32_t k = 0; k < dv; k++) {
            const uint32_t a = nv[k];
            if (!alive[a]) continue;
            common_above(nv, dv, ucol.data() + urow[a], urow[a + 1] - urow[a], a, [&](uint32_t b) {
                if (!alive[b]) return;
                for (uint32_t x : {a, b}) {
                    if (deg[x] > core[v]) {
                        deg[x]--;
                        if (!mark[x]) { mark[x] = 1; touched.push_back(x); }
                    }
                }
            });
        }
        for (uint32_t x : touched) { mark[x] = 0; pq.emplace(deg[x], x); }
    }
    if (kmax_out) *kmax_out = kmax;
    return true;
}

// Level-synchronous round structure of the GPU core loops (dsg.h:665-751, kernels.cuh:493-517): k = max(k, min alive
// tri); remove every alive v with tri <= k (core = k) from one snapshot; recount. Returns the number of rounds, the
// quantity recount and incremental designs must agree on. Uses incremental CPU updates, which give the same tri[]
// as a full recount (every triangle through a removed vertex is subtracted once).
inline bool core_levels(const Graph& g, std::vector<uint32_t>& core, uint32_t* rounds_out, uint32_t* kmax_out) {
    std::vector<uint32_t> urow, ucol;
    undirected_csr(g, urow, ucol);
    std::vector<uint8_t> state(g.n, 1);        // 1 alive, 2 frontier, 0 removed
    std::vector<uint32_t> tri(g.n, 0);
    participation(g, state, tri);
    core.assign(g.n, UINT32_MAX);
    uint64_t nalive = g.n, k = 0;
    uint32_t rounds = 0;
    std::vector<uint32_t> frontier;
    while (nalive > 0) {
        uint64_t mn = UINT64_MAX;
        for (uint32_t v = 0; v < g.n; v++) if (state[v] == 1) mn = std::min<uint64_t>(mn, tri[v]);
        k = std::max(k, mn);
        frontier.clear();
        for (uint32_t v = 0; v < g.n; v++)
            if (state[v] == 1 && tri[v] <= k) { state[v] = 2; core[v] = (uint32_t)k; frontier.push_back(v); }
        // every triangle with at least one frontier corner is subtracted once from its alive (state 1) corners
        for (uint32_t u : frontier) {
            const uint32_t* nu = ucol.data() + urow[u];
            const uint32_t du = urow[u + 1] - urow[u];
            for (uint32_t p = 0; p < du; p++) {
                const uint32_t v = nu[p];
                if (state[v] == 0) continue;
                common_above(nu, du, ucol.data() + urow[v], urow[v + 1] - urow[v], v, [&](uint32_t w) {
                    if (state[w] == 0) return;
                    const bool owner = !((state[v] == 2 && v < u) || (state[w] == 2 && w < u));
                    if (!owner) return;
                    if (state[v] == 1) tri[v]--;
                    if (state[w] == 1) tri[w]--;
                });
            }
        }
        for (uint32_t u : frontier) state[u] = 0;
        nalive -= frontier.size();
        rounds++;
    }
    if (rounds_out) *rounds_out = rounds;
    if (kmax_out) *kmax_out = (uint32_t)k;
    return true;
}

// Triangles and vertices of {v : core[v] >= k} (dsg.h:169-174).
inline void core_level_density(const Graph& g, const std::vector<uint32_t>& core, uint32_t k, uint64_t& tri, uint64_t& verts) {
    std::vector<uint8_t> in(g.n);
    verts = 0;
    for (uint32_t v = 0; v < g.n; v++) { in[v] = core[v] >= k; verts += in[v]; }
    tri = 0;
    for_each_triangle(g, [&](uint32_t u, uint32_t v, uint32_t w, uint32_t, uint32_t, uint32_t) { tri += in[u] && in[v] && in[w]; });
}

// v9 work order: stable counting sort of edges by floor(log2(min(d+u, d+v) + 1)), heaviest bucket first
// (graph.h:156-171).
inline void edge_order(const Graph& g, std::vector<uint32_t>& perm) {
    const int NB = 33;
    auto bucket = [&](uint32_t e) {
        const uint32_t du = outdeg(g, g.src[e]), dv = outdeg(g, g.col[e]);
        uint64_t w = (uint64_t)std::min(du, dv) + 1;
        int b = 0;
        while (w > 1) { w >>= 1; b++; }
        return NB - 1 - b;
    };
    std::vector<uint32_t> start(NB + 1, 0);
    for (uint32_t e = 0; e < g.m; e++) start[bucket(e) + 1]++;
    for (int b = 0; b < NB; b++) start[b + 1] += start[b];
    perm.assign(g.m, 0);
    for (uint32_t e = 0; e < g.m; e++) perm[start[bucket(e)]++] = e;
}

// v4 work split (graph.h:138-150): edge by edge, chunks of `chunk` elements of the smaller out-list, task (e, start).
inline void build_tasks(const Graph& g, uint32_t chunk, std::vector<uint32_t>& t_edge, std::vector<uint32_t>& t_start) {
    t_edge.clear();
    t_start.clear();
    for (uint32_t e = 0; e < g.m; e++) {
        const uint32_t s = std::min(outdeg(g, g.src[e]), outdeg(g, g.col[e]));
        for (uint32_t off = 0; off < s; off += chunk) {
            t_edge.push_back(e);
            t_start.push_back(off);
        }
    }
}

// Induced subgraph on keep[] with ids compacted in old-id order, rebuilt by build_graph_from_edges (dsg.h:103-114).
inline void induced_subgraph(const Graph& g, const std::vector<uint8_t>& keep, Graph& out, std::vector<uint32_t>& old_id) {
    std::vector<uint32_t> nid(g.n, UINT32_MAX);
    old_id.clear();
    for (uint32_t v = 0; v < g.n; v++) if (keep[v]) { nid[v] = (uint32_t)old_id.size(); old_id.push_back(v); }
    EdgeList e;
    for (uint32_t i = 0; i < g.m; i++)
        if (keep[g.src[i]] && keep[g.col[i]]) e.emplace_back(nid[g.src[i]], nid[g.col[i]]);
    build_graph_from_edges((uint32_t)old_id.size(), e, out);
}

// prune_to_core (dsg.h:636-645): the induced subgraph on {v : core[v] >= k}; core_out[i] = core[old_id[i]].
inline void prune_to_core(const Graph& g, const std::vector<uint32_t>& core, uint64_t k, Graph& out, std::vector<uint32_t>& old_id,
                          std::vector<uint32_t>& core_out) {
    std::vector<uint8_t> keep(g.n);
    for (uint32_t v = 0; v < g.n; v++) keep[v] = (uint64_t)core[v] >= k;
    induced_subgraph(g, keep, out, old_id);
    core_out.resize(old_id.size());
    for (size_t i = 0; i < old_id.size(); i++) core_out[i] = core[old_id[i]];
}

// ---- k-truss decomposition (no reference function: it extends the per-edge supports of dsg.h:62-94) ----
// truss[e] for every oriented edge e (an index into g.col) is the largest k such that e lies in the k-truss, the
// maximal subgraph in which every edge closes at least k - 2 triangles; every edge is in the 2-truss. Computed by the
// min-support-first peeling of Wang & Cheng, "Truss decomposition in massive networks" (PVLDB 5(9), 2012, Alg. 2):
// take an edge of least current support s, truss = max(level, s) + 2 with level raised to s, remove it, and take one
// from the support of the two other edges of every triangle it closes whose edges are all still present.
// Returns kmax (0 for a graph without edges).
inline uint32_t truss_decomposition(const Graph& g, std::vector<uint32_t>& truss) {
    const uint32_t n = g.n, m = g.m;
    std::vector<uint32_t> urow((size_t)n + 1, 0), uadj(2ull * m), ueid(2ull * m);
    for (uint32_t e = 0; e < m; e++) { urow[g.src[e] + 1]++; urow[g.col[e] + 1]++; }
    for (uint32_t i = 0; i < n; i++) urow[i + 1] += urow[i];
    {
        std::vector<std::pair<uint32_t, uint32_t>> tmp(2ull * m);
        std::vector<uint32_t> cur(urow.begin(), urow.end() - 1);
        for (uint32_t e = 0; e < m; e++) {
            tmp[cur[g.src[e]]++] = {g.col[e], e};
            tmp[cur[g.col[e]]++] = {g.src[e], e};
        }
        for (uint32_t u = 0; u < n; u++) std::sort(tmp.begin() + urow[u], tmp.begin() + urow[u + 1]);
        for (size_t i = 0; i < tmp.size(); i++) { uadj[i] = tmp[i].first; ueid[i] = tmp[i].second; }
    }
    std::vector<uint32_t> sup;
    edge_support(g, sup);
    truss.assign(m, 0);
    std::vector<uint8_t> gone(m, 0);
    typedef std::pair<uint32_t, uint32_t> Item;   // (support, edge)
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    for (uint32_t e = 0; e < m; e++) heap.push({sup[e], e});
    uint32_t level = 0, kmax = 0;
    while (!heap.empty()) {
        const Item it = heap.top();
        heap.pop();
        const uint32_t e = it.second;
        if (gone[e] || it.first != sup[e]) continue;   // removed, or a stale entry
        level = std::max(level, sup[e]);
        truss[e] = level + 2;
        kmax = std::max(kmax, truss[e]);
        gone[e] = 1;
        uint32_t a = g.src[e], b = g.col[e];
        if (urow[a + 1] - urow[a] > urow[b + 1] - urow[b]) std::swap(a, b);   // walk the shorter list
        for (uint32_t p = urow[a]; p < urow[a + 1]; p++) {
            const uint32_t w = uadj[p], ea = ueid[p];
            if (gone[ea]) continue;
            const uint32_t* lo = uadj.data() + urow[b];
            const uint32_t* hi = uadj.data() + urow[b + 1];
            const uint32_t* q = std::lower_bound(lo, hi, w);
            if (q == hi || *q != w) continue;
            const uint32_t eb = ueid[q - uadj.data()];
            if (gone[eb]) continue;
            heap.push({--sup[ea], ea});
            heap.push({--sup[eb], eb});
        }
    }
    return kmax;
}

// ---- brute-force oracles (independent of the CSR algorithms above) ----

// Bit-set adjacency: triangle count and per-vertex counts of the alive-induced subgraph, O(n^2 * n/64).
struct Brute {
    uint32_t n, words;
    std::vector<uint64_t> bits;
    Brute(uint32_t n_, const EdgeList& undirected) : n(n_), words((n_ + 63) / 64), bits((size_t)n_ * ((n_ + 63) / 64), 0) {
        for (const auto& e : undirected) { set(e.first, e.second); set(e.second, e.first); }
    }
    void set(uint32_t a, uint32_t b) { bits[(size_t)a * words + b / 64] |= 1ull << (b % 64); }
    bool has(uint32_t a, uint32_t b) const { return bits[(size_t)a * words + b / 64] >> (b % 64) & 1; }
    // triangles {u<v<w} with all corners alive
    uint64_t participation(const std::vector<uint8_t>& alive, std::vector<uint32_t>& tri) const {
        tri.assign(n, 0);
        std::vector<uint64_t> mask(words, 0);
        for (uint32_t v = 0; v < n; v++) if (alive[v]) mask[v / 64] |= 1ull << (v % 64);
        uint64_t T = 0;
        for (uint32_t u = 0; u < n; u++) {
            if (!alive[u]) continue;
            for (uint32_t v = u + 1; v < n; v++) {
                if (!alive[v] || !has(u, v)) continue;
                for (uint32_t w = v + 1; w < n; w++)
                    if ((mask[w / 64] >> (w % 64) & 1) && has(u, w) && has(v, w)) { T++; tri[u]++; tri[v]++; tri[w]++; }
            }
        }
        return T;
    }
    uint64_t triangles() const {
        uint64_t T = 0;
        for (uint32_t u = 0; u < n; u++)
            for (uint32_t v = u + 1; v < n; v++) {
                if (!has(u, v)) continue;
                const uint64_t* a = &bits[(size_t)u * words];
                const uint64_t* b = &bits[(size_t)v * words];
                for (uint32_t w = 0; w < words; w++) {
                    uint64_t x = a[w] & b[w];
                    if (w == (v + 1) / 64) x &= ~0ull << ((v + 1) % 64);   // only w > v
                    else if (w < (v + 1) / 64) x = 0;
                    T += __builtin_popcountll(x);
                }
            }
        return T;
    }
};

// Exhaustive densest subset (n <= 20): max over non-empty S of triangles(S)/|S| as an exact fraction.
inline void brute_densest(uint32_t n, const EdgeList& undirected, uint64_t& bn, uint64_t& bd) {
    Brute b(n, undirected);
    std::vector<uint32_t> tm;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t v = u + 1; v < n; v++)
            if (b.has(u, v))
                for (uint32_t w = v + 1; w < n; w++) if (b.has(u, w) && b.has(v, w)) tm.push_back((1u << u) | (1u << v) | (1u << w));
    bn = 0; bd = 1;
    for (uint32_t S = 1; S < (1u << n); S++) {
        uint64_t t = 0;
        for (uint32_t x : tm) t += (S & x) == x;
        const uint64_t sz = __builtin_popcount(S);
        if (t * bd > bn * sz) { bn = t; bd = sz; }
    }
}

// Core numbers by definition (n <= 16): core[v] = max over subsets S containing v of min_{u in S} tri_S(u).
inline void brute_core(uint32_t n, const EdgeList& undirected, std::vector<uint32_t>& core) {
    Brute b(n, undirected);
    std::vector<uint32_t> tm;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t v = u + 1; v < n; v++)
            if (b.has(u, v))
                for (uint32_t w = v + 1; w < n; w++) if (b.has(u, w) && b.has(v, w)) tm.push_back((1u << u) | (1u << v) | (1u << w));
    core.assign(n, 0);
    std::vector<uint32_t> d(n);
    for (uint32_t S = 1; S < (1u << n); S++) {
        std::fill(d.begin(), d.end(), 0);
        for (uint32_t x : tm)
            if ((S & x) == x)
                for (uint32_t v = 0; v < n; v++) if (x >> v & 1) d[v]++;
        uint32_t mn = UINT32_MAX;
        for (uint32_t v = 0; v < n; v++) if (S >> v & 1) mn = std::min(mn, d[v]);
        for (uint32_t v = 0; v < n; v++) if (S >> v & 1) core[v] = std::max(core[v], mn);
    }
}

// k-truss by its definition: for k = 3, 4, ... keep removing edges that close fewer than k - 2 triangles among the
// edges still kept until none is left to remove; truss of an undirected pair = the largest k it survives (2 at least).
// Supports are recounted from the adjacency bits each pass; for small n.
inline void brute_truss(uint32_t n, const EdgeList& e, std::vector<uint32_t>& truss) {
    truss.assign(e.size(), 2);
    std::vector<uint8_t> in(e.size(), 1);
    for (uint32_t k = 3;; k++) {
        bool changed = true;
        while (changed) {
            changed = false;
            EdgeList kept;
            for (size_t i = 0; i < e.size(); i++) if (in[i]) kept.push_back(e[i]);
            Brute b(n, kept);
            for (size_t i = 0; i < e.size(); i++) {
                if (!in[i]) continue;
                uint32_t s = 0;
                for (uint32_t w = 0; w < n; w++) s += b.has(e[i].first, w) && b.has(e[i].second, w);
                if (s + 2 < k) { in[i] = 0; changed = true; }
            }
        }
        bool any = false;
        for (size_t i = 0; i < e.size(); i++) if (in[i]) { truss[i] = k; any = true; }
        if (!any) return;
    }
}

// Exact density comparison a/ad > b/bd with 0/0 read as 0 (dsg.h:37-46).
inline bool dens_gt(uint64_t an, uint64_t ad, uint64_t bn, uint64_t bd) {
    if (!ad) { an = 0; ad = 1; }
    if (!bd) { bn = 0; bd = 1; }
    return (u128)an * bd > (u128)bn * ad;
}

struct GreedyResult {
    uint64_t tri = 0, verts = 0;                            // best density found (0/0: no triangle)
    std::vector<uint32_t> set;                              // its vertices, ascending
    std::vector<std::pair<uint64_t, uint64_t>> per_pass;    // best (tri, verts) after each pass
    bool ok = true;
};

// Greedy++ with triangles as hyperedges (Flowless Alg 2; dsg.h:176-250 greedy_pp, greedy = one pass). Each pass starts
// from the whole graph and deletes, one at a time, the alive vertex with the smallest (load + tri_H, id), adding tri_H
// to its load; the current subgraph H is compared with the best before every deletion and replaces it only when
// strictly denser. This restatement keeps the alive vertices in an ordered set of (load + tri_H, id), re-keying the
// vertices a deletion touched once it is done (the reference keeps a lazy heap; both take the same vertex), and finds
// the triangles through v as the oriented edges a -> b between two alive neighbours of v.
inline GreedyResult greedy_pp(const Graph& g, uint32_t passes) {
    GreedyResult r;
    std::vector<uint32_t> urow, ucol;
    undirected_csr(g, urow, ucol);
    std::vector<uint64_t> tri0, cur;
    const uint64_t T0 = triangle_degrees(g, tri0);
    std::vector<uint64_t> load(g.n, 0), oldkey(g.n, 0);
    std::vector<uint8_t> alive, nb(g.n, 0), touched(g.n, 0);
    std::vector<uint32_t> order, best_order, tl;
    std::set<std::pair<uint64_t, uint32_t>> keys;   // (load + tri_H, id) of the alive vertices
    size_t best_cut = 0;
    for (uint32_t p = 0; p < passes; p++) {
        cur = tri0;
        alive.assign(g.n, 1);
        order.clear();
        keys.clear();
        for (uint32_t x = 0; x < g.n; x++) keys.emplace(load[x] + cur[x], x);
        uint64_t T = T0;
        bool improved = false;
        size_t cut = 0;
        for (uint32_t left = g.n; left > 0; left--) {
            if (dens_gt(T, left, r.tri, r.verts)) { r.tri = T; r.verts = left; improved = true; cut = order.size(); }
            const uint32_t v = keys.begin()->second;
            keys.erase(keys.begin());
            for (uint32_t j = urow[v]; j < urow[v + 1]; j++) nb[ucol[j]] = alive[ucol[j]];
            uint64_t found = 0;
            auto touch = [&](uint32_t x) {
                if (!touched[x]) { touched[x] = 1; oldkey[x] = load[x] + cur[x]; tl.push_back(x); }
            };
            for (uint32_t j = urow[v]; j < urow[v + 1]; j++) {
                const uint32_t a = ucol[j];
                if (!nb[a]) continue;
                for (uint32_t e = g.rowptr[a]; e < g.rowptr[a + 1]; e++) {
                    const uint32_t b = g.col[e];
                    if (!nb[b]) continue;
                    if (!cur[a] || !cur[b]) { r.ok = false; return r; }
                    touch(a);
                    touch(b);
                    cur[a]--;
                    cur[b]--;
                    found++;
                }
            }
            for (uint32_t j = urow[v]; j < urow[v + 1]; j++) nb[ucol[j]] = 0;
            for (uint32_t x : tl) {
                keys.erase({oldkey[x], x});
                keys.emplace(load[x] + cur[x], x);
                touched[x] = 0;
            }
            tl.clear();
            if (found != cur[v]) { r.ok = false; return r; }
            load[v] += cur[v];
            alive[v] = 0;
            T -= cur[v];
            order.push_back(v);
        }
        if (improved) { best_order = order; best_cut = cut; }
        r.per_pass.push_back({r.tri, r.verts});
    }
    if (r.verts) {
        r.set.assign(best_order.begin() + best_cut, best_order.end());
        std::sort(r.set.begin(), r.set.end());
    }
    return r;
}

// Fixpoint definition: the k-core is what remains after repeatedly deleting vertices with fewer than k triangles
// (emu_dsg.cpp:67-82 describes this check).
inline std::vector<uint32_t> core_by_fixpoint(const Graph& g) {
    std::vector<uint32_t> core(g.n, 0), tri;
    std::vector<uint8_t> alive(g.n, 1);
    for (uint32_t k = 1;; k++) {
        for (;;) {
            participation(g, alive, tri);
            bool changed = false;
            for (uint32_t v = 0; v < g.n; v++) if (alive[v] && tri[v] < k) { alive[v] = 0; changed = true; }
            if (!changed) break;
        }
        bool any = false;
        for (uint32_t v = 0; v < g.n; v++) if (alive[v]) { core[v] = k; any = true; }
        if (!any) break;
    }
    return core;
}

}  // namespace orc
