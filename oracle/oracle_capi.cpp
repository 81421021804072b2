// oracle_capi.cpp - C entry points of the CPU oracle (oracle.hpp) for the Python tests and the bench's CPU-baseline
// leg. TEST INFRASTRUCTURE ONLY: the product library never links this file.
#include "oracle.hpp"

using namespace orc;

extern "C" {

struct orc_graph { Graph g; };

// load_graph semantics (canonicalise, drop loops, dedup, n = max_id + 1) when n < 0; otherwise
// build_graph_from_edges semantics on a caller-guaranteed unique, loop-free list with the given n.
orc_graph* orc_graph_from_pairs(const uint32_t* pairs, uint64_t k, int64_t n) {
    orc_graph* h = new orc_graph;
    if (n < 0) {
        graph_from_pairs(pairs, k, h->g);
    } else {
        EdgeList e(k);
        for (uint64_t i = 0; i < k; i++) e[i] = {pairs[2 * i], pairs[2 * i + 1]};
        build_graph_from_edges((uint32_t)n, e, h->g);
    }
    return h;
}

orc_graph* orc_graph_load_text(const char* path) {
    orc_graph* h = new orc_graph;
    if (!load_graph(path, h->g)) { delete h; return nullptr; }
    return h;
}

// An already oriented CSR (rows sorted); src[] and max_out_degree are derived.
orc_graph* orc_graph_from_csr(uint32_t n, uint32_t m, const uint32_t* rowptr, const uint32_t* col) {
    orc_graph* h = new orc_graph;
    Graph& g = h->g;
    g.n = n; g.m = m;
    g.rowptr.assign(rowptr, rowptr + (size_t)n + 1);
    g.col.assign(col, col + m);
    g.src.assign(m, 0);
    g.max_out_degree = 0;
    for (uint32_t u = 0; u < n; u++) {
        for (uint32_t e = g.rowptr[u]; e < g.rowptr[u + 1]; e++) g.src[e] = u;
        g.max_out_degree = std::max(g.max_out_degree, g.rowptr[u + 1] - g.rowptr[u]);
    }
    return h;
}

void orc_graph_free(orc_graph* h) { delete h; }
uint32_t orc_graph_n(const orc_graph* h) { return h->g.n; }
uint32_t orc_graph_m(const orc_graph* h) { return h->g.m; }
uint32_t orc_graph_max_out_degree(const orc_graph* h) { return h->g.max_out_degree; }

void orc_graph_copy(const orc_graph* h, uint32_t* rowptr, uint32_t* col, uint32_t* src) {
    const Graph& g = h->g;
    if (rowptr) std::copy(g.rowptr.begin(), g.rowptr.end(), rowptr);
    if (col) std::copy(g.col.begin(), g.col.end(), col);
    if (src) std::copy(g.src.begin(), g.src.end(), src);
}

uint64_t orc_triangles_serial(const orc_graph* h) { return triangles_serial(h->g); }

uint64_t orc_triangles_omp(const orc_graph* h, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    return triangles_omp(h->g);
}

int orc_omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

uint64_t orc_participation(const orc_graph* h, const uint8_t* alive, uint32_t* tri, int use_omp) {
    std::vector<uint8_t> a(alive, alive + h->g.n);
    std::vector<uint32_t> t;
    const uint64_t T = use_omp ? participation_omp(h->g, a, t) : participation(h->g, a, t);
    std::copy(t.begin(), t.end(), tri);
    return T;
}

uint64_t orc_edge_support(const orc_graph* h, uint32_t* sup) {
    std::vector<uint32_t> s;
    const uint64_t T = edge_support(h->g, s);
    std::copy(s.begin(), s.end(), sup);
    return T;
}

// Returns 1 on success; stats = {best_num, best_den, best_round, rounds}.
int orc_batch_peel(const orc_graph* h, uint64_t num, uint64_t den, int use_omp, uint64_t* stats, uint32_t* removed_round,
                   char* err, int errlen) {
    PeelResult r = batch_peel(h->g, num, den, use_omp != 0);
    if (!r.ok) {
        if (err && errlen > 0) snprintf(err, errlen, "%s", r.error.c_str());
        return 0;
    }
    stats[0] = r.best_num; stats[1] = r.best_den; stats[2] = r.best_round; stats[3] = r.rounds;
    std::copy(r.removed_round.begin(), r.removed_round.end(), removed_round);
    return 1;
}

int orc_peel_arith_ok(uint32_t n, uint64_t T0, uint64_t num, uint64_t den) { return peel_arith_ok(n, T0, num, den) ? 1 : 0; }

void orc_undirected(const orc_graph* h, uint32_t* urow, uint32_t* ucol) {
    std::vector<uint32_t> r, c;
    undirected_csr(h->g, r, c);
    std::copy(r.begin(), r.end(), urow);
    std::copy(c.begin(), c.end(), ucol);
}

int orc_core_decomposition(const orc_graph* h, uint32_t* core, uint32_t* kmax) {
    std::vector<uint32_t> c;
    if (!core_decomposition(h->g, c, kmax)) return 0;
    std::copy(c.begin(), c.end(), core);
    return 1;
}

int orc_core_levels(const orc_graph* h, uint32_t* core, uint32_t* rounds, uint32_t* kmax) {
    std::vector<uint32_t> c;
    if (!core_levels(h->g, c, rounds, kmax)) return 0;
    std::copy(c.begin(), c.end(), core);
    return 1;
}

void orc_core_fixpoint(const orc_graph* h, uint32_t* core) {
    std::vector<uint32_t> c = core_by_fixpoint(h->g);
    std::copy(c.begin(), c.end(), core);
}

void orc_core_level_density(const orc_graph* h, const uint32_t* core, uint32_t k, uint64_t* tri, uint64_t* verts) {
    std::vector<uint32_t> c(core, core + h->g.n);
    core_level_density(h->g, c, k, *tri, *verts);
}

void orc_edge_order(const orc_graph* h, uint32_t* perm) {
    std::vector<uint32_t> p;
    edge_order(h->g, p);
    std::copy(p.begin(), p.end(), perm);
}

// build_tasks: returns the task count; writes up to `cap` tasks when the buffers are given.
uint64_t orc_build_tasks(const orc_graph* h, uint32_t chunk, uint32_t* t_edge, uint32_t* t_start, uint64_t cap) {
    std::vector<uint32_t> te, ts;
    build_tasks(h->g, chunk, te, ts);
    if (t_edge && t_start && cap >= te.size()) {
        std::copy(te.begin(), te.end(), t_edge);
        std::copy(ts.begin(), ts.end(), t_start);
    }
    return te.size();
}

// Induced subgraph on keep[]; returns the new graph, old ids written to old_id (capacity n).
orc_graph* orc_induced_subgraph(const orc_graph* h, const uint8_t* keep, uint32_t* old_id, uint32_t* n_out) {
    orc_graph* o = new orc_graph;
    std::vector<uint8_t> k(keep, keep + h->g.n);
    std::vector<uint32_t> ids;
    induced_subgraph(h->g, k, o->g, ids);
    std::copy(ids.begin(), ids.end(), old_id);
    *n_out = (uint32_t)ids.size();
    return o;
}

// prune_to_core on caller-given core numbers (n entries); old_id and core_out have capacity n.
orc_graph* orc_prune_to_core(const orc_graph* h, const uint32_t* core, uint64_t k, uint32_t* old_id, uint32_t* core_out, uint32_t* n_out) {
    orc_graph* o = new orc_graph;
    std::vector<uint32_t> c(core, core + h->g.n), ids, co;
    prune_to_core(h->g, c, k, o->g, ids, co);
    std::copy(ids.begin(), ids.end(), old_id);
    std::copy(co.begin(), co.end(), core_out);
    *n_out = (uint32_t)ids.size();
    return o;
}

// k-truss numbers per oriented edge (m entries); returns kmax.
uint32_t orc_truss(const orc_graph* h, uint32_t* truss) {
    std::vector<uint32_t> t;
    const uint32_t kmax = truss_decomposition(h->g, t);
    std::copy(t.begin(), t.end(), truss);
    return kmax;
}

// Greedy++ (passes >= 1): out = {tri, verts, set_size, ok}; set (capacity n), per_pass (2 * passes).
void orc_greedy_pp(const orc_graph* h, uint32_t passes, uint64_t* out, uint32_t* set, uint64_t* per_pass) {
    const GreedyResult r = greedy_pp(h->g, passes);
    out[0] = r.tri;
    out[1] = r.verts;
    out[2] = r.set.size();
    out[3] = r.ok ? 1 : 0;
    std::copy(r.set.begin(), r.set.end(), set);
    for (size_t i = 0; i < r.per_pass.size(); i++) { per_pass[2 * i] = r.per_pass[i].first; per_pass[2 * i + 1] = r.per_pass[i].second; }
}

// k-truss by definition on an undirected unique loop-free pair list (truss per pair).
void orc_brute_truss(uint32_t n, const uint32_t* pairs, uint64_t k, uint32_t* truss) {
    EdgeList e(k);
    for (uint64_t i = 0; i < k; i++) e[i] = {pairs[2 * i], pairs[2 * i + 1]};
    std::vector<uint32_t> t;
    brute_truss(n, e, t);
    std::copy(t.begin(), t.end(), truss);
}

// Brute force on an undirected unique loop-free pair list.
uint64_t orc_brute_participation(uint32_t n, const uint32_t* pairs, uint64_t k, const uint8_t* alive, uint32_t* tri) {
    EdgeList e(k);
    for (uint64_t i = 0; i < k; i++) e[i] = {pairs[2 * i], pairs[2 * i + 1]};
    Brute b(n, e);
    std::vector<uint8_t> a(alive, alive + n);
    std::vector<uint32_t> t;
    const uint64_t T = b.participation(a, t);
    std::copy(t.begin(), t.end(), tri);
    return T;
}

uint64_t orc_brute_triangles(uint32_t n, const uint32_t* pairs, uint64_t k) {
    EdgeList e(k);
    for (uint64_t i = 0; i < k; i++) e[i] = {pairs[2 * i], pairs[2 * i + 1]};
    return Brute(n, e).triangles();
}

void orc_brute_densest(uint32_t n, const uint32_t* pairs, uint64_t k, uint64_t* bn, uint64_t* bd) {
    EdgeList e(k);
    for (uint64_t i = 0; i < k; i++) e[i] = {pairs[2 * i], pairs[2 * i + 1]};
    brute_densest(n, e, *bn, *bd);
}

void orc_brute_core(uint32_t n, const uint32_t* pairs, uint64_t k, uint32_t* core) {
    EdgeList e(k);
    for (uint64_t i = 0; i < k; i++) e[i] = {pairs[2 * i], pairs[2 * i + 1]};
    std::vector<uint32_t> c;
    brute_core(n, e, c);
    std::copy(c.begin(), c.end(), core);
}

}  // extern "C"
