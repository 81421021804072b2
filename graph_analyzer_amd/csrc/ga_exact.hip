// ga_exact.hip - the exact densest subgraph for the triangle density: CoreExact (DSD paper Alg 4, dsg.h:454-627) and
// Exact (Alg 1, dsg.h:402-428), with the flow network of dsg.h:314-399 and a max-flow that gives the minimal source
// side of a minimum cut.
//
// Where the work runs:
//   - device: every pass over the triangles of a graph (one lane per oriented edge, merge of the two sorted out-lists):
//     the Pruning1 counts per earliest vertex in the (core, id) order, the Pruning2 counts per connected component,
//     the seed check, and the triangle list of each component's subgraph that the network is built from; the induced
//     subgraphs themselves (ga_graph_induced);
//   - host: connected components (one pass over the undirected CSR), the network and its max-flow (Dinic). Max-flow
//     phases are sequential, and the networks left after the pruning are small (three arcs per triangle of a pruned
//     component), so the flows stay on the CPU.
// Results equal the reference's bit for bit: the minimal source side of a minimum cut is the same for every maximum
// flow (the vertices reachable from s in its residual network), so the sets, the binary search and its counters follow.
#include "ga_internal.hpp"

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

namespace {

struct ExBuf {   // scoped device allocation
    void* p = nullptr;
    ~ExBuf() { ga_free(p); }
};

enum { TV_MINPOS = 0, TV_LABEL = 1, TV_LIST = 2 };
constexpr uint32_t NO_LABEL = 0xFFFFFFFFu;

// Row of oriented edge e (the last u with rowptr[u] <= e), for graphs stored without src[].
__device__ __forceinline__ uint32_t row_of_edge(const uint32_t* __restrict__ rowptr, uint32_t n, uint32_t e) {
    uint32_t lo = 0, hi = n;
    while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (rowptr[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Every triangle (u, v, w) of the oriented CSR once, at its edge (u, v): w runs over N+(u) ∩ N+(v) (sorted rows, merged
// with their positions, so the edge ids of (u, w) and (v, w) come for free). One lane per oriented edge.
//   TV_MINPOS: acc[min(key[u], key[v], key[w])] += 1     (key = position in the (core, id) order)
//   TV_LABEL:  acc[key[u]] += 1 when no corner is NO_LABEL (key = component label; the corners of a triangle inside a
//              union of components share one)
//   TV_LIST:   the k-th triangle found writes (u, v, w) and its edge ids (uv, uw, vw) when k < cap; *nlist = the count
template <int MODE>
__global__ void k_tri_visit(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col, const uint32_t* __restrict__ src,
                            uint32_t n, uint32_t m, const uint32_t* __restrict__ key, unsigned long long* __restrict__ acc,
                            uint32_t* __restrict__ tris, uint32_t* __restrict__ tedges, unsigned long long* __restrict__ nlist,
                            unsigned long long cap) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const uint32_t u = src ? src[e] : row_of_edge(rowptr, n, e), v = col[e];
        uint32_t i = rowptr[u], j = rowptr[v];
        const uint32_t ie = rowptr[u + 1], je = rowptr[v + 1];
        while (i < ie && j < je) {
            const uint32_t a = col[i], b = col[j];
            if (a < b) {
                i++;
            } else if (b < a) {
                j++;
            } else {
                if (MODE == TV_MINPOS) {
                    atomicAdd(&acc[min(min(key[u], key[v]), key[a])], 1ull);
                } else if (MODE == TV_LABEL) {
                    const uint32_t l = key[u];
                    if (l != NO_LABEL && key[v] != NO_LABEL && key[a] != NO_LABEL) atomicAdd(&acc[l], 1ull);
                } else {
                    const unsigned long long k = atomicAdd(nlist, 1ull);
                    if (k < cap) {
                        tris[3 * k] = u;
                        tris[3 * k + 1] = v;
                        tris[3 * k + 2] = a;
                        tedges[3 * k] = e;
                        tedges[3 * k + 1] = i;
                        tedges[3 * k + 2] = j;
                    }
                }
                i++;
                j++;
            }
        }
    }
}

template <int MODE>
int tri_visit(ga_ctx* ctx, const ga_graph* g, const uint32_t* key, unsigned long long* acc, uint32_t* tris, uint32_t* tedges,
              unsigned long long* nlist, unsigned long long cap) {
    if (!g->m) return GA_OK;
    k_tri_visit<MODE><<<ctx_grid(ctx, grid_for(g->m, 256, (unsigned)std::max(ctx->cus, 1) * 16u)), 256, 0, ctx->stream>>>(
        g->rowptr, g->col, g->src, g->n, g->m, key, acc, tris, tedges, nlist, cap);
    return ga_launch_check(ctx, "k_tri_visit");
}

// Per-key triangle counts (TV_MINPOS: nacc = n; TV_LABEL: nacc = number of labels) from a host key array.
template <int MODE>
int tri_counts(ga_ctx* ctx, const ga_graph* g, const std::vector<uint32_t>& key, size_t nacc, std::vector<uint64_t>& out) {
    out.assign(nacc, 0);
    if (!g->m || !nacc) return GA_OK;
    ExBuf kb, ab;
    GA_TRY(ga_malloc(ctx, &kb.p, (size_t)g->n * 4));
    GA_TRY(ga_malloc(ctx, &ab.p, nacc * 8));
    GA_HIP(ctx, hipMemcpyAsync(kb.p, key.data(), (size_t)g->n * 4, hipMemcpyHostToDevice, ctx->stream));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)ab.p, nacc, 0));
    GA_TRY(tri_visit<MODE>(ctx, g, (const uint32_t*)kb.p, (unsigned long long*)ab.p, nullptr, nullptr, nullptr, 0));
    GA_HIP(ctx, hipMemcpyAsync(out.data(), ab.p, nacc * 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

// Every triangle of g with its edge ids (two passes: count, then write).
int tri_list(ga_ctx* ctx, const ga_graph* g, std::vector<uint32_t>& tris, std::vector<uint32_t>& tedges) {
    tris.clear();
    tedges.clear();
    if (!g->m) return GA_OK;
    unsigned long long* d = (unsigned long long*)ctx->dsmall;
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)d, 1, 0));
    GA_TRY(tri_visit<TV_LIST>(ctx, g, nullptr, nullptr, nullptr, nullptr, d, 0));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint64_t T = ctx->hstat[0];
    if (!T) return GA_OK;
    ExBuf tb, eb;
    GA_TRY(ga_malloc(ctx, &tb.p, T * 12));
    GA_TRY(ga_malloc(ctx, &eb.p, T * 12));
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)d, 1, 0));
    GA_TRY(tri_visit<TV_LIST>(ctx, g, nullptr, nullptr, (uint32_t*)tb.p, (uint32_t*)eb.p, d, T));
    tris.resize(3 * T);
    tedges.resize(3 * T);
    GA_HIP(ctx, hipMemcpyAsync(tris.data(), tb.p, T * 12, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(tedges.data(), eb.p, T * 12, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->hstat[0] != T) return ga_fail(ctx, GA_ERR_STATE, "triangle list changed between passes (%llu vs %llu)",
                                           (unsigned long long)ctx->hstat[0], (unsigned long long)T);
    return GA_OK;
}

// ---- exact rationals (triangles / vertices; den 0 = no subgraph, density 0), as dsg.h:36-47 ----
typedef unsigned __int128 u128;
struct Frac {
    uint64_t num = 0, den = 0;
};
int frac_cmp(Frac a, Frac b) {
    if (a.den == 0) a = Frac{0, 1};
    if (b.den == 0) b = Frac{0, 1};
    const u128 l = (u128)a.num * b.den, r = (u128)b.num * a.den;
    return l < r ? -1 : (l > r ? 1 : 0);
}
bool frac_gt(Frac a, Frac b) { return frac_cmp(a, b) > 0; }
uint64_t frac_ceil(Frac d) { return d.den == 0 ? 0 : (d.num + d.den - 1) / d.den; }

// ---- max-flow: Dinic on a CSR of arcs; every arc has its reverse ----
struct MaxFlow {
    static constexpr int64_t INF = (int64_t)1 << 60;
    uint32_t nodes = 0;
    std::vector<uint32_t> a_from, a_to;   // arc list while building: arc 2k is forward, 2k + 1 its reverse
    std::vector<int64_t> a_cap;
    std::vector<uint32_t> off, to, rev, pos;   // CSR by tail; pos[arc id] = its CSR slot
    std::vector<int64_t> cap, cap0;
    std::vector<int32_t> level;
    std::vector<uint32_t> cur;

    void reset(uint32_t n) {
        nodes = n;
        a_from.clear();
        a_to.clear();
        a_cap.clear();
    }
    uint32_t add(uint32_t a, uint32_t b, int64_t c) {   // returns the id of the forward arc
        const uint32_t id = (uint32_t)a_from.size();
        a_from.push_back(a);
        a_to.push_back(b);
        a_cap.push_back(c);
        a_from.push_back(b);
        a_to.push_back(a);
        a_cap.push_back(0);
        return id;
    }
    void finish() {   // lay the arcs out by tail
        const size_t na = a_from.size();
        off.assign(nodes + 1, 0);
        for (size_t k = 0; k < na; k++) off[a_from[k] + 1]++;
        for (uint32_t v = 0; v < nodes; v++) off[v + 1] += off[v];
        std::vector<uint32_t> fill(off.begin(), off.end() - 1);
        pos.resize(na);
        to.resize(na);
        cap0.resize(na);
        for (size_t k = 0; k < na; k++) {
            const uint32_t p = fill[a_from[k]]++;
            pos[k] = p;
            to[p] = a_to[k];
            cap0[p] = a_cap[k];
        }
        rev.resize(na);
        for (size_t k = 0; k < na; k++) rev[pos[k]] = pos[k ^ 1];
        a_from.clear();
        a_from.shrink_to_fit();
        a_to.clear();
        a_to.shrink_to_fit();
        a_cap.clear();
        a_cap.shrink_to_fit();
    }
    bool levels(uint32_t s, uint32_t t) {   // BFS over residual arcs; afterwards level[v] >= 0 iff v is reachable from s
        level.assign(nodes, -1);
        std::vector<uint32_t> q;
        q.reserve(nodes);
        level[s] = 0;
        q.push_back(s);
        for (size_t h = 0; h < q.size(); h++) {
            const uint32_t x = q[h];
            for (uint32_t p = off[x]; p < off[x + 1]; p++)
                if (cap[p] > 0 && level[to[p]] < 0) {
                    level[to[p]] = level[x] + 1;
                    q.push_back(to[p]);
                }
        }
        return level[t] >= 0;
    }
    // One blocking flow on the level graph: advance along admissible arcs (current-arc pointers), augment at t and
    // restart from the tail of the first saturated arc, retreat from dead ends (explicit stack: paths can be long).
    int64_t blocking(uint32_t s, uint32_t t) {
        int64_t total = 0;
        cur.assign(off.begin(), off.end() - 1);
        std::vector<uint32_t> path;   // CSR slots of the arcs from s to x
        uint32_t x = s;
        for (;;) {
            if (x == t) {
                int64_t f = INF;
                for (uint32_t p : path) f = std::min(f, cap[p]);
                size_t first = path.size();
                for (size_t k = 0; k < path.size(); k++) {
                    cap[path[k]] -= f;
                    cap[rev[path[k]]] += f;
                    if (cap[path[k]] == 0 && first == path.size()) first = k;
                }
                total += f;
                path.resize(first);
                x = first ? to[path[first - 1]] : s;
                continue;
            }
            uint32_t& c = cur[x];
            while (c < off[x + 1] && !(cap[c] > 0 && level[to[c]] == level[x] + 1)) c++;
            if (c < off[x + 1]) {
                path.push_back(c);
                x = to[c];
                continue;
            }
            if (x == s) break;
            level[x] = -1;   // dead end for the rest of this phase
            const uint32_t p = path.back();
            path.pop_back();
            x = path.empty() ? s : to[path.back()];
            cur[x]++;   // the arc into the dead end
            (void)p;
        }
        return total;
    }
    int64_t run(uint32_t s, uint32_t t) {
        int64_t flow = 0;
        while (levels(s, t)) flow += blocking(s, t);
        return flow;   // level[] now marks the minimal source side of a minimum cut
    }
};

// The flow network of Alg 1 for triangles (dsg.h:314-329): s = 0, t = 1, node 2 + v per vertex of H, one node per edge
// of H that lies in a triangle. s -> v: Q deg(v); v -> t: 3 alpha Q (set per solve); edge node -> its two ends: INF;
// v -> edge node: Q for every triangle made of v and that edge. A cut pays 3 Q per triangle not inside the source side
// and 3 alpha Q per vertex in it, so the source side is non-trivial exactly when some subgraph is denser than alpha.
struct TriFlow {
    MaxFlow F;
    uint32_t n = 0;
    uint64_t Q = 1;
    std::vector<uint32_t> tris;          // 3 per triangle (ids of H)
    std::vector<uint64_t> deg;           // triangles per vertex
    std::vector<uint32_t> sink_slot;     // CSR slot of v -> t, UINT32_MAX for vertices in no triangle

    int build(ga_ctx* ctx, const ga_graph* H, uint64_t Q_, std::string* err) {
        n = H->n;
        Q = Q_;
        std::vector<uint32_t> tedges;
        GA_TRY(tri_list(ctx, H, tris, tedges));
        const uint64_t T = tris.size() / 3;
        deg.assign(n, 0);
        for (uint64_t k = 0; k < 3 * T; k++) deg[tris[k]]++;
        if ((u128)T * 3 * Q >= ((u128)1 << 59)) {
            *err = "flow capacities would exceed 2^59";
            return GA_ERR_ARITH;
        }
        std::vector<uint32_t> enode(H->m, UINT32_MAX);
        uint32_t nodes = 2 + n;
        for (uint64_t k = 0; k < 3 * T; k++)
            if (enode[tedges[k]] == UINT32_MAX) enode[tedges[k]] = nodes++;
        F.reset(nodes);
        std::vector<uint32_t> sink_arc(n, UINT32_MAX);
        for (uint32_t v = 0; v < n; v++) {
            if (!deg[v]) continue;
            F.add(0, 2 + v, (int64_t)(Q * deg[v]));
            sink_arc[v] = F.add(2 + v, 1, 0);
        }
        std::vector<uint8_t> made(H->m, 0);
        for (uint64_t i = 0; i < T; i++) {
            const uint32_t u = tris[3 * i], v = tris[3 * i + 1], w = tris[3 * i + 2];
            // edge (u,v) closes with w, (u,w) with v, (v,w) with u
            const uint32_t ends[3][2] = {{u, v}, {u, w}, {v, w}};
            const uint32_t third[3] = {w, v, u};
            for (int k = 0; k < 3; k++) {
                const uint32_t e = tedges[3 * i + k], pn = enode[e];
                if (!made[e]) {
                    made[e] = 1;
                    F.add(pn, 2 + ends[k][0], MaxFlow::INF);
                    F.add(pn, 2 + ends[k][1], MaxFlow::INF);
                }
                F.add(2 + third[k], pn, (int64_t)Q);
            }
        }
        F.finish();
        sink_slot.assign(n, UINT32_MAX);
        for (uint32_t v = 0; v < n; v++)
            if (sink_arc[v] != UINT32_MAX) sink_slot[v] = F.pos[sink_arc[v]];
        return GA_OK;
    }
    // Vertices of H on the minimal source side for alpha = a / Q (empty: nothing denser than alpha).
    std::vector<uint32_t> solve(int64_t a) {
        F.cap = F.cap0;
        for (uint32_t v = 0; v < n; v++)
            if (sink_slot[v] != UINT32_MAX) F.cap[sink_slot[v]] = 3 * a;
        F.run(0, 1);
        std::vector<uint32_t> S;
        for (uint32_t v = 0; v < n; v++)
            if (deg[v] && F.level[2 + v] >= 0) S.push_back(v);
        return S;
    }
    uint64_t count_in(const std::vector<uint32_t>& S) const {
        std::vector<uint8_t> in(n, 0);
        for (uint32_t v : S) in[v] = 1;
        uint64_t c = 0;
        for (size_t k = 0; k < tris.size(); k += 3) c += in[tris[k]] & in[tris[k + 1]] & in[tris[k + 2]];
        return c;
    }
};

struct Result {
    Frac best;
    std::vector<uint32_t> set;
    uint32_t flow_calls = 0, comps = 0, shrinks = 0;
    bool ok = true;
    std::string err;
};

struct GraphHold {   // scoped device graph
    ga_graph* g = nullptr;
    ~GraphHold() { ga_graph_destroy(g); }
};

int induced(ga_ctx* ctx, const ga_graph* g, const std::vector<uint8_t>& keep, GraphHold& out, std::vector<uint32_t>& old_id) {
    ga_graph_destroy(out.g);
    out.g = nullptr;
    old_id.assign(std::max<uint32_t>(g->n, 1), 0);
    GA_TRY(ga_graph_induced(ctx, g, keep.data(), &out.g, old_id.data(), nullptr));
    old_id.resize(out.g->n);
    return GA_OK;
}

// Connected components of {v : core[v] >= k} over the undirected CSR, labelled in order of their smallest vertex.
uint32_t components(const std::vector<uint32_t>& urow, const std::vector<uint32_t>& ucol, const uint32_t* core, uint32_t n, uint64_t k,
                    std::vector<uint32_t>& label) {
    label.assign(n, NO_LABEL);
    uint32_t nc = 0;
    std::vector<uint32_t> stack;
    for (uint32_t s = 0; s < n; s++) {
        if (core[s] < k || label[s] != NO_LABEL) continue;
        label[s] = nc;
        stack.push_back(s);
        while (!stack.empty()) {
            const uint32_t x = stack.back();
            stack.pop_back();
            for (uint32_t p = urow[x]; p < urow[x + 1]; p++) {
                const uint32_t y = ucol[p];
                if (core[y] >= k && label[y] == NO_LABEL) {
                    label[y] = nc;
                    stack.push_back(y);
                }
            }
        }
        nc++;
    }
    return nc;
}

#define EX_FAIL(msg)      \
    do {                  \
        R.ok = false;     \
        R.err = (msg);    \
        return GA_OK;     \
    } while (0)

// CoreExact (dsg.h:454-627), the steps of Alg 4 with the reference's interpretations: u = kmax reset per component,
// l the exact density of the best subgraph so far, the initial check recorded as a candidate, component cores taken
// from the input core numbers, and the optional seed verified before it becomes the lower bound.
int core_exact_run(ga_ctx* ctx, ga_graph* g, const uint32_t* core, bool pruning12, const Frac* seed_fr, const uint32_t* seed_set,
                   Result& R) {
    const uint32_t n = g->n;
    uint32_t kmax = 0;
    for (uint32_t v = 0; v < n; v++) kmax = std::max(kmax, core[v]);
    if (kmax == 0) return GA_OK;   // no triangle at all
    std::vector<uint32_t> urow(n + 1), ucol(2 * (size_t)g->m);
    GA_TRY(ga_graph_undirected(ctx, g, urow.data(), ucol.data()));
    Frac& best = R.best;
    std::vector<uint32_t>& best_set = R.set;
    std::vector<uint32_t> label;
    uint64_t k2 = 1;
    if (pruning12) {
        // Pruning1: the residual graphs of the peeling are the suffixes of the (core, id) order; each triangle belongs
        // to the suffixes that hold its earliest corner
        std::vector<uint32_t> order(n), pos(n);
        std::iota(order.begin(), order.end(), 0u);
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return core[a] != core[b] ? core[a] < core[b] : a < b; });
        for (uint32_t i = 0; i < n; i++) pos[order[i]] = i;
        std::vector<uint64_t> cnt;
        GA_TRY(tri_counts<TV_MINPOS>(ctx, g, pos, n, cnt));
        uint64_t suffix = 0;
        size_t best_i = n;
        for (uint32_t i = n; i-- > 0;) {
            suffix += cnt[i];
            if (frac_gt(Frac{suffix, n - i}, best)) {
                best = Frac{suffix, n - i};
                best_i = i;
            }
        }
        best_set.assign(order.begin() + best_i, order.end());
        const uint64_t k1 = frac_ceil(best);
        if (k1 > kmax) EX_FAIL("inconsistent core numbers (ceil(rho') > kmax)");
        // Pruning2: the connected components of the k'-core
        const uint32_t nc1 = components(urow, ucol, core, n, k1, label);
        std::vector<uint64_t> ct, cv(nc1, 0);
        for (uint32_t v = 0; v < n; v++)
            if (label[v] != NO_LABEL) cv[label[v]]++;
        GA_TRY(tri_counts<TV_LABEL>(ctx, g, label, nc1, ct));
        for (uint32_t c = 0; c < nc1; c++)
            if (frac_gt(Frac{ct[c], cv[c]}, best)) {
                best = Frac{ct[c], cv[c]};
                best_set.clear();
                for (uint32_t v = 0; v < n; v++)
                    if (label[v] == c) best_set.push_back(v);
            }
        k2 = std::max<uint64_t>(k1, frac_ceil(best));
        if (k2 > kmax) EX_FAIL("inconsistent core numbers (ceil(rho'') > kmax)");
    } else {
        // test hook of the reference: start from one triangle (density 1/3), the first one in the order of the
        // reference's enumeration (smallest u, then v, then w over the oriented CSR)
        std::vector<uint32_t> tris, tedges;
        GA_TRY(tri_list(ctx, g, tris, tedges));
        if (!tris.empty()) {
            size_t f = 0;
            for (size_t k = 3; k < tris.size(); k += 3) {
                // the reference visits u ascending, then v by its position in N+(u), then w by its position: edge ids
                // (uv, then uw) order the triangles the same way
                if (tedges[k] < tedges[f] || (tedges[k] == tedges[f] && tedges[k + 1] < tedges[f + 1])) f = k;
            }
            best = Frac{1, 3};
            best_set = {tris[f], tris[f + 1], tris[f + 2]};
        }
    }
    if (seed_fr && seed_fr->den > 0) {
        // the seed must be a real subgraph with the claimed triangle count before it bounds anything
        std::vector<uint32_t> in(n, NO_LABEL);
        bool okseed = true;
        for (uint64_t i = 0; i < seed_fr->den; i++) {
            const uint32_t v = seed_set[i];
            if (v >= n || in[v] != NO_LABEL) {
                okseed = false;
                break;
            }
            in[v] = 0;
        }
        std::vector<uint64_t> st;
        if (okseed) GA_TRY(tri_counts<TV_LABEL>(ctx, g, in, 1, st));
        if (!okseed || st[0] != seed_fr->num) EX_FAIL("seed is not a valid subgraph of the graph");
        if (frac_gt(*seed_fr, best)) {
            best = *seed_fr;
            best_set.assign(seed_set, seed_set + seed_fr->den);
        }
        k2 = std::max<uint64_t>(k2, frac_ceil(best));
        if (k2 > kmax) EX_FAIL("inconsistent seed (ceil(rho) > kmax)");
    }
    // lines 4-20: one component at a time
    const uint32_t nc = components(urow, ucol, core, n, k2, label);
    for (uint32_t c = 0; c < nc; c++) {
        uint64_t kc = std::max<uint64_t>(k2, frac_ceil(best));   // line 6: C within the ceil(l)-core
        if (kc > kmax) break;
        std::vector<uint8_t> keep(n, 0);
        for (uint32_t v = 0; v < n; v++) keep[v] = label[v] == c && core[v] >= kc;
        GraphHold H;
        std::vector<uint32_t> old_id;
        GA_TRY(induced(ctx, g, keep, H, old_id));
        if (H.g->n < 3) continue;
        const uint64_t Q = 2ull * H.g->n * (H.g->n - 1);   // Pruning3's grid, fixed from C after line 6
        TriFlow net;
        {
            const int rc = net.build(ctx, H.g, Q, &R.err);
            if (rc == GA_ERR_ARITH) {
                R.ok = false;
                return GA_OK;
            }
            GA_TRY(rc);
        }
        if (net.tris.empty()) continue;
        int64_t lo = best.den ? (int64_t)(((u128)best.num * Q + best.den - 1) / best.den) : 0;
        int64_t hi = (int64_t)((u128)kmax * Q);
        if (lo >= hi) continue;
        auto consider = [&](const std::vector<uint32_t>& S, int64_t alpha) -> bool {
            const Frac d{net.count_in(S), S.size()};
            if ((u128)d.num * Q <= (u128)alpha * d.den) {
                R.ok = false;
                R.err = "min cut returned a set that is not denser than alpha";
                return false;
            }
            if (frac_gt(d, best)) {
                best = d;
                best_set.clear();
                for (uint32_t v : S) best_set.push_back(old_id[v]);
            }
            return true;
        };
        std::vector<uint32_t> S = net.solve(lo);   // lines 7-8
        R.flow_calls++;
        if (S.empty()) continue;   // line 9
        if (!consider(S, lo)) return GA_OK;
        R.comps++;
        while (hi - lo >= 2) {   // lines 10-20
            const int64_t mid = lo + (hi - lo) / 2;
            S = net.solve(mid);
            R.flow_calls++;
            if (S.empty()) {
                hi = mid;
                continue;
            }
            lo = mid;
            if (!consider(S, mid)) return GA_OK;
            const uint64_t knew = ((uint64_t)mid + Q - 1) / Q;   // ceil(alpha)
            if (knew > kc) {   // line 17: the network shrinks to the ceil(alpha)-core of C
                kc = knew;
                R.shrinks++;
                if (kc > kmax) break;
                std::vector<uint8_t> m2(H.g->n, 0);
                for (uint32_t v = 0; v < H.g->n; v++) m2[v] = core[old_id[v]] >= kc;
                GraphHold H2;
                std::vector<uint32_t> map2;
                GA_TRY(induced(ctx, H.g, m2, H2, map2));
                if (H2.g->n < 3) EX_FAIL("shrunk network is empty although a denser subgraph exists");
                std::vector<uint32_t> old2(map2.size());
                for (size_t i = 0; i < map2.size(); i++) old2[i] = old_id[map2[i]];
                {
                    const int rc = net.build(ctx, H2.g, Q, &R.err);
                    if (rc == GA_ERR_ARITH) {
                        R.ok = false;
                        return GA_OK;
                    }
                    GA_TRY(rc);
                }
                if (net.tris.empty()) EX_FAIL("shrunk network has no triangle although a denser subgraph exists");
                std::swap(H.g, H2.g);
                old_id.swap(old2);
            }
        }
    }
    std::sort(best_set.begin(), best_set.end());
    return GA_OK;
}

// Exact (dsg.h:402-428): the binary search of Alg 1 on one network of the whole graph.
int exact_run(ga_ctx* ctx, ga_graph* g, Result& R) {
    TriFlow net;
    if (g->n < 2) return GA_OK;
    const uint64_t Q = 2ull * g->n * (g->n - 1);
    {
        const int rc = net.build(ctx, g, Q, &R.err);
        if (rc == GA_ERR_ARITH) {
            R.ok = false;
            return GA_OK;
        }
        GA_TRY(rc);
    }
    if (net.tris.empty()) return GA_OK;
    const uint64_t maxdeg = *std::max_element(net.deg.begin(), net.deg.end());
    int64_t lo = 0, hi = (int64_t)(maxdeg * Q);
    while (hi - lo >= 2) {
        const int64_t mid = lo + (hi - lo) / 2;
        std::vector<uint32_t> S = net.solve(mid);
        R.flow_calls++;
        if (S.empty()) {
            hi = mid;
            continue;
        }
        lo = mid;
        const Frac d{net.count_in(S), S.size()};
        if ((u128)d.num * Q <= (u128)mid * d.den) {
            R.ok = false;
            R.err = "min cut returned a set that is not denser than alpha";
            return GA_OK;
        }
        if (frac_gt(d, R.best)) {
            R.best = d;
            R.set = S;
        }
    }
    return GA_OK;
}

int put_result(ga_ctx* ctx, const Result& R, uint64_t* out7, uint32_t* set) {
    out7[0] = R.ok ? R.best.num : 0;
    out7[1] = R.ok ? R.best.den : 0;
    out7[2] = R.ok ? R.set.size() : 0;
    out7[3] = R.ok ? 1 : 0;
    out7[4] = R.flow_calls;
    out7[5] = R.comps;
    out7[6] = R.shrinks;
    if (!R.ok) return ga_fail(ctx, GA_ERR_STATE, "%s", R.err.c_str());
    if (set) std::copy(R.set.begin(), R.set.end(), set);
    return GA_OK;
}

}  // namespace

extern "C" int ga_core_exact(ga_ctx* ctx, ga_graph* g, const uint32_t* core, int use_pruning12, const uint64_t* seed_tri_verts,
                             const uint32_t* seed_set, uint64_t* out7, uint32_t* set) {
    if (!ctx || !g || !out7 || (g->n && (!core || !set))) return GA_ERR_ARG;
    if (seed_tri_verts && seed_tri_verts[1] && !seed_set) return ga_fail(ctx, GA_ERR_ARG, "seed set is NULL");
    if (seed_tri_verts && seed_tri_verts[1] > g->n) return ga_fail(ctx, GA_ERR_ARG, "seed has more vertices than the graph");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    Result R;
    Frac sf;
    if (seed_tri_verts) sf = Frac{seed_tri_verts[0], seed_tri_verts[1]};
    GA_TRY(core_exact_run(ctx, g, core, use_pruning12 != 0, seed_tri_verts ? &sf : nullptr, seed_set, R));
    return put_result(ctx, R, out7, set);
}

extern "C" int ga_exact_densest(ga_ctx* ctx, ga_graph* g, uint64_t* out7, uint32_t* set) {
    if (!ctx || !g || !out7 || (g->n && !set)) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    Result R;
    GA_TRY(exact_run(ctx, g, R));
    return put_result(ctx, R, out7, set);
}
