// gen.cpp - seeded synthetic edge-list generators (Chung-Lu power law, R-MAT, Erdos-Renyi, planted communities).
//
// Output is a RAW undirected pair list: duplicates, both orientations and self loops may occur. The CSR builder
// (GPU, or the text-loader semantics on the CPU) canonicalises, drops loops and de-duplicates, exactly as the
// reference's load_graph does for a SNAP file (graph.h:71-96). Every generator splits its samples into fixed chunks
// with their own counter-derived RNG stream, so the output is identical for any OpenMP thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const uint64_t CHUNK = 1u << 16;

inline uint64_t mix64(uint64_t x) {   // splitmix64 finaliser
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct Rng {   // splitmix64 stream
    uint64_t s;
    explicit Rng(uint64_t seed) : s(mix64(seed)) {}
    uint64_t next() { s += 0x9E3779B97F4A7C15ull; uint64_t z = s; z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; return z ^ (z >> 31); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint32_t below(uint32_t n) { return (uint32_t)(((unsigned __int128)next() * n) >> 64); }
};

inline uint64_t stream_seed(uint64_t seed, uint64_t tag, uint64_t chunk) { return mix64(seed ^ mix64(tag * 0x100000001B3ull + chunk)); }

// Cumulative weights w_i = (i + 1 + i0)^(-1/(gamma-1)) as 64-bit fixed point, for inverse-CDF sampling.
std::vector<uint64_t> chung_lu_cdf(uint32_t n, double gamma, double i0) {
    std::vector<double> w(n);
    const double ex = -1.0 / (gamma - 1.0);
    double W = 0;
    for (uint32_t i = 0; i < n; i++) { w[i] = std::pow((double)i + 1.0 + i0, ex); W += w[i]; }
    std::vector<uint64_t> cdf(n);
    double acc = 0;
    const double scale = 18446744073709551615.0;
    for (uint32_t i = 0; i < n; i++) {
        acc += w[i];
        const double x = acc / W * scale;
        cdf[i] = x >= scale ? UINT64_MAX : (uint64_t)x;
    }
    cdf[n - 1] = UINT64_MAX;
    return cdf;
}

inline uint32_t sample_cdf(const std::vector<uint64_t>& cdf, uint64_t r) {
    const size_t i = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
    return (uint32_t)std::min(i, cdf.size() - 1);
}

// Walker/Vose alias table for w_i = (i + 1 + i0)^(-1/(gamma-1)): draw i uniformly, keep it when the second draw is
// below prob[i] (64-bit fixed point), else take alias[i]. O(1) per endpoint.
struct Alias {
    std::vector<uint64_t> prob;
    std::vector<uint32_t> alias;
};

Alias alias_table(uint32_t n, double gamma, double i0) {
    std::vector<double> q(n);
    const double ex = -1.0 / (gamma - 1.0);
    double W = 0;
    for (uint32_t i = 0; i < n; i++) { q[i] = std::pow((double)i + 1.0 + i0, ex); W += q[i]; }
    for (uint32_t i = 0; i < n; i++) q[i] = q[i] * n / W;
    Alias a;
    a.prob.assign(n, UINT64_MAX);
    a.alias.resize(n);
    std::vector<uint32_t> small, large;
    for (uint32_t i = 0; i < n; i++) (q[i] < 1.0 ? small : large).push_back(i);
    for (uint32_t i = 0; i < n; i++) a.alias[i] = i;
    while (!small.empty() && !large.empty()) {
        const uint32_t s = small.back(), l = large.back();
        small.pop_back();
        a.prob[s] = (uint64_t)(q[s] * 18446744073709549568.0);
        a.alias[s] = l;
        q[l] = (q[l] + q[s]) - 1.0;
        if (q[l] < 1.0) { large.pop_back(); small.push_back(l); }
    }
    return a;   // leftovers keep probability one
}

inline uint32_t sample_alias(const Alias& a, Rng& r) {
    const uint32_t i = r.below((uint32_t)a.prob.size());
    return r.next() < a.prob[i] ? i : a.alias[i];
}

std::vector<uint32_t> id_permutation(uint32_t n, uint64_t seed) {
    std::vector<uint32_t> p(n);
    for (uint32_t i = 0; i < n; i++) p[i] = i;
    Rng r(stream_seed(seed, 77, 0));
    for (uint32_t i = n; i > 1; i--) std::swap(p[i - 1], p[r.below(i)]);
    return p;
}

}  // namespace

extern "C" {

struct ga_edges {
    uint32_t* pairs;    // 2 * count entries, (a, b) per edge
    uint64_t count;
    uint32_t n;         // id space of the generator (max id + 1 may be smaller)
};

void ga_gen_free(ga_edges* e) {
    if (e && e->pairs) free(e->pairs);
    if (e) { e->pairs = nullptr; e->count = 0; }
}

static int alloc_edges(ga_edges* out, uint64_t count, uint32_t n) {
    out->pairs = (uint32_t*)malloc(std::max<uint64_t>(count, 1) * 2 * sizeof(uint32_t));
    out->count = count;
    out->n = n;
    return out->pairs ? 0 : -1;
}

// Chung-Lu: `samples` independent pairs with both endpoints drawn proportional to w_i; ids shuffled when permute.
int ga_gen_chung_lu(uint32_t n, uint64_t samples, double gamma, double i0, uint64_t seed, int permute, ga_edges* out) {
    if (n < 2 || gamma <= 1.0 || alloc_edges(out, samples, n)) return -1;
    const std::vector<uint64_t> cdf = chung_lu_cdf(n, gamma, i0);
    std::vector<uint32_t> perm;
    if (permute) perm = id_permutation(n, seed);
    const uint64_t nchunks = (samples + CHUNK - 1) / CHUNK;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long c = 0; c < (long long)nchunks; c++) {
        Rng r(stream_seed(seed, 1, (uint64_t)c));
        const uint64_t b = (uint64_t)c * CHUNK, e = std::min(samples, b + CHUNK);
        for (uint64_t i = b; i < e; i++) {
            uint32_t a = sample_cdf(cdf, r.next()), d = sample_cdf(cdf, r.next());
            if (permute) { a = perm[a]; d = perm[d]; }
            out->pairs[2 * i] = a;
            out->pairs[2 * i + 1] = d;
        }
    }
    return 0;
}

// Chung-Lu as above, endpoints drawn with an alias table (a different stream of graphs from ga_gen_chung_lu).
int ga_gen_chung_lu_alias(uint32_t n, uint64_t samples, double gamma, double i0, uint64_t seed, int permute, ga_edges* out) {
    if (n < 2 || gamma <= 1.0 || alloc_edges(out, samples, n)) return -1;
    const Alias at = alias_table(n, gamma, i0);
    std::vector<uint32_t> perm;
    if (permute) perm = id_permutation(n, seed);
    const uint64_t nchunks = (samples + CHUNK - 1) / CHUNK;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long c = 0; c < (long long)nchunks; c++) {
        Rng r(stream_seed(seed, 5, (uint64_t)c));
        const uint64_t b = (uint64_t)c * CHUNK, e = std::min(samples, b + CHUNK);
        for (uint64_t i = b; i < e; i++) {
            uint32_t a = sample_alias(at, r), d = sample_alias(at, r);
            if (permute) { a = perm[a]; d = perm[d]; }
            out->pairs[2 * i] = a;
            out->pairs[2 * i + 1] = d;
        }
    }
    return 0;
}

// Chung-Lu with EXACTLY m unique loop-free undirected edges: draw in generation order, keep the first occurrence of
// each pair until m are kept (serial; meant for graphs of a few hundred thousand edges).
int ga_gen_chung_lu_exact(uint32_t n, uint64_t m, double gamma, double i0, uint64_t seed, ga_edges* out) {
    if (n < 2 || gamma <= 1.0 || m > (uint64_t)n * (n - 1) / 2 || alloc_edges(out, m, n)) return -1;
    const std::vector<uint64_t> cdf = chung_lu_cdf(n, gamma, i0);
    const std::vector<uint32_t> perm = id_permutation(n, seed);
    std::unordered_set<uint64_t> seen;
    seen.reserve(m * 2);
    Rng r(stream_seed(seed, 2, 0));
    uint64_t kept = 0, tries = 0;
    while (kept < m) {
        if (++tries > 1000 * m + 1000000) return -2;
        uint32_t a = perm[sample_cdf(cdf, r.next())], d = perm[sample_cdf(cdf, r.next())];
        if (a == d) continue;
        const uint64_t key = ((uint64_t)std::min(a, d) << 32) | std::max(a, d);
        if (!seen.insert(key).second) continue;
        out->pairs[2 * kept] = a;
        out->pairs[2 * kept + 1] = d;
        kept++;
    }
    return 0;
}

// Expected-degree offset i0 such that the heaviest vertex expects degree dmax in a graph with m edges.
double ga_gen_chung_lu_offset(uint32_t n, uint64_t m, double gamma, double dmax) {
    const double ex = -1.0 / (gamma - 1.0);
    auto top = [&](double i0) {
        double W = 0;
        for (uint32_t i = 0; i < n; i++) W += std::pow(i + 1.0 + i0, ex);
        return 2.0 * (double)m * std::pow(1.0 + i0, ex) / W;
    };
    double lo = 0, hi = (double)n;
    if (top(lo) <= dmax) return 0.0;
    for (int it = 0; it < 60; it++) {
        const double mid = 0.5 * (lo + hi);
        (top(mid) > dmax ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
}

// R-MAT: `samples` pairs on 2^scale ids; each level picks a quadrant with probabilities (a, b, c, 1-a-b-c) from a
// 32-bit uniform, two levels per 64-bit draw. No noise, no id permutation.
int ga_gen_rmat(uint32_t scale, uint64_t samples, double a, double b, double c, uint64_t seed, ga_edges* out) {
    if (scale < 1 || scale > 31 || alloc_edges(out, samples, 1u << scale)) return -1;
    const uint64_t ta = (uint64_t)(a * 4294967296.0), tb = (uint64_t)((a + b) * 4294967296.0), tc = (uint64_t)((a + b + c) * 4294967296.0);
    const uint64_t nchunks = (samples + CHUNK - 1) / CHUNK;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long ch = 0; ch < (long long)nchunks; ch++) {
        Rng r(stream_seed(seed, 3, (uint64_t)ch));
        const uint64_t bgn = (uint64_t)ch * CHUNK, end = std::min(samples, bgn + CHUNK);
        for (uint64_t i = bgn; i < end; i++) {
            uint32_t u = 0, v = 0;
            uint64_t bits = 0;
            for (uint32_t lvl = 0; lvl < scale; lvl++) {
                if ((lvl & 1) == 0) bits = r.next();
                const uint64_t x = (lvl & 1) ? (bits >> 32) : (bits & 0xFFFFFFFFull);
                const uint32_t row = x >= tb, colb = (x >= ta && x < tb) || x >= tc;
                u = (u << 1) | row;
                v = (v << 1) | colb;
            }
            out->pairs[2 * i] = u;
            out->pairs[2 * i + 1] = v;
        }
    }
    return 0;
}

// Erdos-Renyi G(n, M)-style: `samples` uniform pairs (duplicates and loops removed later by the builder).
int ga_gen_er(uint32_t n, uint64_t samples, uint64_t seed, ga_edges* out) {
    if (n < 2 || alloc_edges(out, samples, n)) return -1;
    const uint64_t nchunks = (samples + CHUNK - 1) / CHUNK;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long c = 0; c < (long long)nchunks; c++) {
        Rng r(stream_seed(seed, 4, (uint64_t)c));
        const uint64_t b = (uint64_t)c * CHUNK, e = std::min(samples, b + CHUNK);
        for (uint64_t i = b; i < e; i++) { out->pairs[2 * i] = r.below(n); out->pairs[2 * i + 1] = r.below(n); }
    }
    return 0;
}

// Many graphs into one buffer: graph g gets `samples[g]` pairs generated exactly as ga_gen_er (kind 0),
// ga_gen_chung_lu (kind 1) or ga_gen_chung_lu_alias (kind 2), ids permuted, would generate them alone; offsets[g] = first pair of graph g
// (count + 1 entries). i0[g] < 0 means ga_gen_chung_lu_offset(n, samples, gamma, -i0[g]). Graphs are spread over
// the OpenMP threads (inner loops run serially inside).
int ga_gen_batch(uint32_t count, const uint32_t* n, const uint64_t* samples, const int32_t* kind, const double* gamma,
                 const double* i0, const uint64_t* seed, uint64_t* offsets, ga_edges* out) {
    offsets[0] = 0;
    for (uint32_t g = 0; g < count; g++) offsets[g + 1] = offsets[g] + samples[g];
    if (alloc_edges(out, offsets[count], 0)) return -1;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
    for (long long g = 0; g < (long long)count; g++) {
        ga_edges one{};
        const double off = i0[g] < 0 ? ga_gen_chung_lu_offset(n[g], samples[g], gamma[g], -i0[g]) : i0[g];
        const int rc = kind[g] == 0   ? ga_gen_er(n[g], samples[g], seed[g], &one)
                       : kind[g] == 1 ? ga_gen_chung_lu(n[g], samples[g], gamma[g], off, seed[g], 1, &one)
                                      : ga_gen_chung_lu_alias(n[g], samples[g], gamma[g], off, seed[g], 1, &one);
        if (rc) { bad++; continue; }
        memcpy(out->pairs + 2 * offsets[g], one.pairs, samples[g] * 2 * sizeof(uint32_t));
        ga_gen_free(&one);
    }
    return bad ? -1 : 0;
}

// Planted communities: `count` groups of distinct random ids with log-uniform sizes in [smin, smax]; every pair inside
// a group is an edge with probability p.
int ga_gen_communities(uint32_t n, uint32_t count, uint32_t smin, uint32_t smax, double p, uint64_t seed, ga_edges* out) {
    if (n < smax || smin < 2 || smin > smax) return -1;
    std::vector<std::vector<uint32_t>> edges(count);
    const double lmin = std::log((double)smin), lmax = std::log((double)smax + 1.0);
#pragma omp parallel for schedule(dynamic, 16)
    for (long long ci = 0; ci < (long long)count; ci++) {
        Rng r(stream_seed(seed, 5, (uint64_t)ci));
        uint32_t s = (uint32_t)std::exp(lmin + (lmax - lmin) * r.uniform());
        s = std::min(std::max(s, smin), smax);
        std::vector<uint32_t> mem;
        mem.reserve(s);
        while (mem.size() < s) {
            const uint32_t v = r.below(n);
            if (std::find(mem.begin(), mem.end(), v) == mem.end()) mem.push_back(v);
        }
        auto& out_c = edges[ci];
        for (uint32_t i = 0; i < s; i++)
            for (uint32_t j = i + 1; j < s; j++)
                if (r.uniform() < p) { out_c.push_back(mem[i]); out_c.push_back(mem[j]); }
    }
    uint64_t total = 0;
    for (auto& e : edges) total += e.size() / 2;
    if (alloc_edges(out, total, n)) return -1;
    uint64_t pos = 0;
    for (auto& e : edges) { std::copy(e.begin(), e.end(), out->pairs + pos); pos += e.size(); }
    return 0;
}

// Appends b to a (a grows).
int ga_gen_concat(ga_edges* a, const ga_edges* b) {
    uint32_t* p = (uint32_t*)realloc(a->pairs, std::max<uint64_t>(a->count + b->count, 1) * 2 * sizeof(uint32_t));
    if (!p) return -1;
    memcpy(p + 2 * a->count, b->pairs, b->count * 2 * sizeof(uint32_t));
    a->pairs = p;
    a->count += b->count;
    a->n = std::max(a->n, b->n);
    return 0;
}

// SNAP-style text file, one "a<TAB>b" line per pair.
int ga_gen_write_text(const ga_edges* e, const char* path, const char* comment) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    if (comment) fprintf(f, "# %s\n", comment);
    for (uint64_t i = 0; i < e->count; i++) fprintf(f, "%u\t%u\n", e->pairs[2 * i], e->pairs[2 * i + 1]);
    fclose(f);
    return 0;
}

}  // extern "C"
