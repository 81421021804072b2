// ref_harness.cpp - C entry points around the reference's OWN CPU code, compiled from the reference headers where they
// lie (graph.h, densest.h, dsg.h under the reference's src/ directory, passed with -I by oracle/Makefile). Used only to
// cross-check the oracle restatement and, optionally, as the CPU baseline (kind "reference"). TEST INFRASTRUCTURE ONLY.
// Nothing of the reference is copied here: this file only calls the reference functions by name.
#include "dsg.h"

extern "C" {

struct ref_graph { Graph g; };

// Reference load_graph on a text file (graph.h:71-96).
ref_graph* ref_load_graph(const char* path) {
    ref_graph* h = new ref_graph;
    if (!load_graph(path, h->g)) { delete h; return nullptr; }
    return h;
}

// Reference build_graph_from_edges on a unique, loop-free pair list (graph.h:38-67).
ref_graph* ref_build_graph(uint32_t n, const uint32_t* pairs, uint64_t k) {
    ref_graph* h = new ref_graph;
    std::vector<std::pair<uint32_t, uint32_t>> e(k);
    for (uint64_t i = 0; i < k; i++) e[i] = {pairs[2 * i], pairs[2 * i + 1]};
    build_graph_from_edges(n, e, h->g);
    return h;
}

void ref_free(ref_graph* h) { delete h; }
uint32_t ref_n(const ref_graph* h) { return h->g.n; }
uint32_t ref_m(const ref_graph* h) { return h->g.m; }
uint32_t ref_max_out_degree(const ref_graph* h) { return h->g.max_out_degree; }

void ref_copy(const ref_graph* h, uint32_t* rowptr, uint32_t* col, uint32_t* src) {
    std::copy(h->g.rowptr.begin(), h->g.rowptr.end(), rowptr);
    std::copy(h->g.col.begin(), h->g.col.end(), col);
    std::copy(h->g.src.begin(), h->g.src.end(), src);
}

uint64_t ref_triangles_serial(const ref_graph* h) { return cpu_triangles_serial(h->g); }

uint64_t ref_triangles_omp(const ref_graph* h, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    return cpu_triangles_omp(h->g);
}

uint64_t ref_participation(const ref_graph* h, const uint8_t* alive, uint32_t* tri) {
    std::vector<uint8_t> a(alive, alive + h->g.n);
    std::vector<uint32_t> t;
    const uint64_t T = cpu_triangle_participation(h->g, a, t);
    std::copy(t.begin(), t.end(), tri);
    return T;
}

uint64_t ref_edge_support(const ref_graph* h, uint32_t* sup) {
    std::fill(sup, sup + h->g.m, 0u);
    uint64_t T = 0;
    for_each_triangle(h->g, [&](uint32_t, uint32_t, uint32_t, uint32_t a, uint32_t b, uint32_t c) { T++; sup[a]++; sup[b]++; sup[c]++; });
    return T;
}

int ref_batch_peel(const ref_graph* h, uint64_t num, uint64_t den, int use_omp, uint64_t* stats, uint32_t* removed_round) {
    PeelResult r = cpu_batch_peel(h->g, num, den, use_omp != 0);
    if (!r.ok) return 0;
    stats[0] = r.best_num; stats[1] = r.best_den; stats[2] = r.best_round; stats[3] = r.rounds;
    std::copy(r.removed_round.begin(), r.removed_round.end(), removed_round);
    return 1;
}

int ref_core_decomposition(const ref_graph* h, uint32_t* core, uint32_t* kmax) {
    std::vector<uint32_t> c;
    if (!core_decomposition(h->g, c, kmax)) return 0;
    std::copy(c.begin(), c.end(), core);
    return 1;
}

void ref_core_level_density(const ref_graph* h, const uint32_t* core, uint32_t k, uint64_t* tri, uint64_t* verts) {
    std::vector<uint32_t> c(core, core + h->g.n);
    const Dens d = core_level_density(h->g, c, k);
    *tri = d.num; *verts = d.den;
}

void ref_edge_order(const ref_graph* h, uint32_t* perm) {
    std::vector<uint32_t> p;
    build_edge_order(h->g, p);
    std::copy(p.begin(), p.end(), perm);
}

uint64_t ref_build_tasks(const ref_graph* h, uint32_t chunk, uint32_t* t_edge, uint32_t* t_start, uint64_t cap) {
    std::vector<uint32_t> te, ts;
    build_tasks(h->g, chunk, te, ts);
    if (t_edge && t_start && cap >= te.size()) {
        std::copy(te.begin(), te.end(), t_edge);
        std::copy(ts.begin(), ts.end(), t_start);
    }
    return te.size();
}

void ref_undirected(const ref_graph* h, uint32_t* urow, uint32_t* ucol) {
    std::vector<uint32_t> r, c;
    build_undirected_csr(h->g, r, c);
    std::copy(r.begin(), r.end(), urow);
    std::copy(c.begin(), c.end(), ucol);
}

// Reference induced_subgraph (dsg.h:103-114); old_id has capacity n.
ref_graph* ref_induced_subgraph(const ref_graph* h, const uint8_t* keep, uint32_t* old_id, uint32_t* n_out) {
    ref_graph* o = new ref_graph;
    std::vector<uint8_t> k(keep, keep + h->g.n);
    std::vector<uint32_t> ids;
    induced_subgraph(h->g, k, o->g, ids);
    std::copy(ids.begin(), ids.end(), old_id);
    *n_out = (uint32_t)ids.size();
    return o;
}

// Reference prune_to_core (dsg.h:636-645); old_id and core_out have capacity n.
ref_graph* ref_prune_to_core(const ref_graph* h, const uint32_t* core, uint64_t k, uint32_t* old_id, uint32_t* core_out, uint32_t* n_out) {
    ref_graph* o = new ref_graph;
    std::vector<uint32_t> c(core, core + h->g.n), ids, co;
    prune_to_core(h->g, c, k, o->g, ids, co);
    std::copy(ids.begin(), ids.end(), old_id);
    std::copy(co.begin(), co.end(), core_out);
    *n_out = (uint32_t)ids.size();
    return o;
}

// Reference greedy_pp (dsg.h:189-248): out = {tri, verts, set_size, ok}; set (capacity n), per_pass (2 * passes).
void ref_greedy_pp(const ref_graph* h, uint32_t passes, uint64_t* out, uint32_t* set, uint64_t* per_pass) {
    const DsgResult r = greedy_pp(h->g, passes);
    out[0] = r.tri;
    out[1] = r.verts;
    out[2] = r.set.size();
    out[3] = r.ok ? 1 : 0;
    std::copy(r.set.begin(), r.set.end(), set);
    for (size_t i = 0; i < r.per_pass.size(); i++) { per_pass[2 * i] = r.per_pass[i].num; per_pass[2 * i + 1] = r.per_pass[i].den; }
}

int ref_peel_arith_ok(uint32_t n, uint64_t T0, uint64_t num, uint64_t den) { return peel_arith_ok(n, T0, num, den) ? 1 : 0; }

// out = {tri, verts, set_size, ok, flow_calls, components_searched, shrinks}; set has room for n ids.
static void put_dsg(const DsgResult& r, uint64_t* out, uint32_t* set) {
    out[0] = r.tri;
    out[1] = r.verts;
    out[2] = r.set.size();
    out[3] = r.ok ? 1 : 0;
    out[4] = r.flow_calls;
    out[5] = r.components_searched;
    out[6] = r.shrinks;
    std::copy(r.set.begin(), r.set.end(), set);
}

// Reference core_exact (dsg.h:454-627) on the given core numbers; seed (tri, verts, its set) or seed_tv == NULL.
void ref_core_exact(const ref_graph* h, const uint32_t* core, int use_pruning12, const uint64_t* seed_tv, const uint32_t* seed_set,
                    uint64_t* out, uint32_t* set) {
    std::vector<uint32_t> c(core, core + h->g.n);
    DsgResult seed;
    if (seed_tv) {
        seed.tri = seed_tv[0];
        seed.verts = seed_tv[1];
        seed.set.assign(seed_set, seed_set + seed_tv[1]);
    }
    put_dsg(core_exact(h->g, c, use_pruning12 != 0, seed_tv ? &seed : nullptr), out, set);
}

// Reference exact_densest (dsg.h:402-428).
void ref_exact_densest(const ref_graph* h, uint64_t* out, uint32_t* set) { put_dsg(exact_densest(h->g), out, set); }

}  // extern "C"
