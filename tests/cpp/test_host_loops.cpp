// test_host_loops.cpp - CPU-only test of the host loops of the drop-in headers (batch_peel_loop, core_decomp_loop,
// core_decomp_loop_inc) through the backend seam, the way the reference tests them with emulator backends
// (emu_densest.cpp:57-74, emu_dsg.cpp:85-124). The fake backends compute with the CPU oracle and can inject the faults
// the host loops must catch: a wrong round-0 count, zero or too many removals, a decreasing level, and the
// arithmetic guards. Prints ALL OK and exits 0 on success.
#include <random>

#include "../../graph_analyzer_amd/host/dsg.h"
#include "../../oracle/oracle.hpp"

static int g_fail = 0;
#define CHECK(c, ...) do { if (!(c)) { g_fail++; if (g_fail < 30) { printf("FAIL %s:%d: ", __FILE__, __LINE__); printf(__VA_ARGS__); printf("\n"); } } } while (0)

enum Fault { NONE, T_PLUS_ONE, REMOVED_ZERO, REMOVED_TOO_MANY, K_DECREASES };

struct FakePeel {
    const orc::Graph& g;
    Fault fault;
    std::vector<uint8_t> alive;
    std::vector<uint32_t> rr, tri;
    uint64_t T = 0, removed = 0;
    uint32_t rounds_seen = 0;
    FakePeel(const orc::Graph& g_, Fault f) : g(g_), fault(f) {}
    void reset_state() { alive.assign(g.n, 1); rr.assign(g.n, UINT32_MAX); }
    void begin_round() { T = removed = 0; }
    void run_tc() { T = orc::participation(g, alive, tri); }
    void run_mark(uint64_t nalive, uint64_t num, uint64_t den, uint32_t round) {
        const uint64_t rhs = 3ull * (den + num) * T;
        for (uint32_t v = 0; v < g.n; v++)
            if (alive[v] && (uint64_t)tri[v] * nalive * den <= rhs) { alive[v] = 0; rr[v] = round; removed++; }
        if (fault == T_PLUS_ONE && round == 0) T++;
        if (fault == REMOVED_ZERO && round == 1) removed = 0;
        if (fault == REMOVED_TOO_MANY && round == 0) removed = nalive + 1;
        rounds_seen = round + 1;
    }
    void read_stats(uint64_t& t, uint64_t& r) { t = T; r = removed; }
    double kernel_ms() const { return -1; }
};

struct FakeCoreInc {   // incremental backend semantics computed by recounting on the CPU
    const orc::Graph& g;
    Fault fault;
    std::vector<uint8_t> state;
    std::vector<uint32_t> core, tri;
    uint64_t removed = 0, k = 0;
    uint32_t min_tri = UINT32_MAX, round = 0;
    FakeCoreInc(const orc::Graph& g_, Fault f) : g(g_), fault(f) {}
    void reset_state() { state.assign(g.n, 1); core.assign(g.n, UINT32_MAX); round = 0; }
    void init_tri(uint64_t& T) { T = orc::participation(g, state, tri) + (fault == T_PLUS_ONE ? 1 : 0); }
    void begin_round() { removed = 0; k = 0; min_tri = UINT32_MAX; }
    void run_min() { for (uint32_t v = 0; v < g.n; v++) if (state[v] == 1) min_tri = std::min(min_tri, tri[v]); }
    void run_find(uint64_t k_in) {
        k = std::max<uint64_t>(k_in, min_tri);
        for (uint32_t v = 0; v < g.n; v++)
            if (state[v] == 1 && tri[v] <= k) { state[v] = 2; core[v] = (uint32_t)k; removed++; }
    }
    void run_update() {}
    void run_retire() {
        for (uint32_t v = 0; v < g.n; v++) if (state[v] == 2) state[v] = 0;
        std::vector<uint8_t> alive(g.n);
        for (uint32_t v = 0; v < g.n; v++) alive[v] = state[v] == 1;
        orc::participation(g, alive, tri);
    }
    void read_stats(uint64_t& r, uint64_t& kk) {
        r = removed;
        kk = k;
        if (fault == K_DECREASES && round == 2 && k > 0) kk = 0;
        if (fault == REMOVED_ZERO && round == 1) r = 0;
        round++;
    }
};

struct FakeCore {      // recount backend
    FakeCoreInc inner;
    uint64_t T = 0;
    FakeCore(const orc::Graph& g, Fault f) : inner(g, f) {}
    void reset_state() { inner.reset_state(); }
    void begin_round() { inner.begin_round(); }
    void run_tc() {
        std::vector<uint8_t> alive(inner.g.n);
        for (uint32_t v = 0; v < inner.g.n; v++) alive[v] = inner.state[v] == 1;
        T = orc::participation(inner.g, alive, inner.tri) + ((inner.fault == T_PLUS_ONE && inner.round == 0) ? 1 : 0);
    }
    void run_min() { inner.run_min(); }
    void run_peel(uint64_t k) { inner.run_find(k); for (uint32_t v = 0; v < inner.g.n; v++) if (inner.state[v] == 2) inner.state[v] = 0; }
    void read_stats(uint64_t& t, uint64_t& r, uint64_t& k) { t = T; inner.read_stats(r, k); }
};

int main() {
    std::mt19937_64 rng(20261203);
    int graphs = 0;
    for (int it = 0; it < 60; it++) {
        const uint32_t n = 8 + rng() % 60;
        const double p = 0.05 + 0.5 * (double)(rng() % 100) / 100.0;
        std::vector<uint32_t> pairs;
        for (uint32_t a = 0; a < n; a++)
            for (uint32_t b = a + 1; b < n; b++)
                if ((double)(rng() % 10000) / 10000.0 < p) { pairs.push_back(a); pairs.push_back(b); }
        orc::Graph g;
        orc::EdgeList e;
        for (size_t i = 0; i < pairs.size(); i += 2) e.emplace_back(pairs[i], pairs[i + 1]);
        orc::build_graph_from_edges(n, e, g);
        const uint64_t T = orc::triangles_serial(g);
        graphs++;
        const uint64_t eps[][2] = {{1, 10}, {1, 1}, {0, 1}, {3, 7}};
        for (auto& ep : eps) {
            const orc::PeelResult ref = orc::batch_peel(g, ep[0], ep[1]);
            FakePeel fp(g, NONE);
            PeelResult r = batch_peel_loop(g.n, T, ep[0], ep[1], fp);
            r.removed_round = fp.rr;
            CHECK(r.ok && r.best_num == ref.best_num && r.best_den == ref.best_den && r.best_round == ref.best_round &&
                      r.rounds == ref.rounds && r.removed_round == ref.removed_round,
                  "batch_peel_loop != oracle (graph %d eps %llu/%llu)", it, (unsigned long long)ep[0], (unsigned long long)ep[1]);
        }
        {   // faults
            FakePeel f1(g, T_PLUS_ONE);
            PeelResult r1 = batch_peel_loop(g.n, T, 1, 10, f1);
            CHECK(!r1.ok && r1.error.find("round 0") != std::string::npos, "round-0 count fault not caught");
            FakePeel f2(g, REMOVED_TOO_MANY);
            CHECK(!batch_peel_loop(g.n, T, 1, 10, f2).ok, "removed > alive not caught");
            FakePeel f3(g, REMOVED_ZERO);
            PeelResult r3 = batch_peel_loop(g.n, T, 1, 10, f3);
            CHECK(!r3.ok || f3.rounds_seen <= 1, "removed == 0 not caught");
            FakePeel f4(g, NONE);
            CHECK(!batch_peel_loop(g.n, T, 1, 0, f4).ok && !batch_peel_loop(g.n, T, 1, (1ull << 20) + 1, f4).ok &&
                      !batch_peel_loop(g.n, 1ull << 32, 1, 10, f4).ok,
                  "peel_arith_ok guards not applied");
        }
        {   // core loops
            std::vector<uint32_t> core;
            uint32_t kmax = 0, rounds = 0, kmax2 = 0;
            orc::core_decomposition(g, core, &kmax);
            std::vector<uint32_t> core2;
            orc::core_levels(g, core2, &rounds, &kmax2);
            CHECK(core == core2 && kmax == kmax2, "level-synchronous oracle differs from Alg 3 (graph %d)", it);
            FakeCoreInc fi(g, NONE);
            CoreLoopResult ri = core_decomp_loop_inc(g.n, T, fi);
            CHECK(ri.ok && fi.core == core && ri.kmax == kmax && ri.rounds == rounds, "core_decomp_loop_inc != oracle (graph %d)", it);
            FakeCore fr(g, NONE);
            CoreLoopResult rr = core_decomp_loop(g.n, T, fr);
            CHECK(rr.ok && fr.inner.core == core && rr.kmax == kmax && rr.rounds == rounds, "core_decomp_loop != oracle (graph %d)", it);
            FakeCoreInc fk(g, K_DECREASES);
            CoreLoopResult rk = core_decomp_loop_inc(g.n, T, fk);
            CHECK(!rk.ok || rounds <= 2 || kmax == 0, "level decrease not caught");
            FakeCoreInc ft(g, T_PLUS_ONE);
            CHECK(!core_decomp_loop_inc(g.n, T, ft).ok, "incremental T mismatch not caught");
            FakeCore frt(g, T_PLUS_ONE);
            CHECK(!core_decomp_loop(g.n, T, frt).ok, "recount round-0 T mismatch not caught");
            FakeCoreInc fz(g, REMOVED_ZERO);
            CoreLoopResult rz = core_decomp_loop_inc(g.n, T, fz);
            CHECK(!rz.ok || rounds <= 1, "removed == 0 not caught");
            FakeCoreInc fb(g, NONE);
            CHECK(!core_decomp_loop_inc(g.n, 1ull << 32, fb).ok, "T >= 2^32 guard not applied");
        }
    }
    // exact comparisons of Dens
    CHECK(dens_cmp(Dens{1, 3}, Dens{2, 6}) == 0 && dens_gt(Dens{2, 5}, Dens{1, 3}) && dens_cmp(Dens{0, 0}, Dens{0, 7}) == 0, "dens_cmp");
    CHECK(dens_ceil(Dens{7, 2}) == 4 && dens_ceil(Dens{6, 2}) == 3 && dens_ceil(Dens{5, 0}) == 0, "dens_ceil");
    CHECK(peel_improves(3, 2, 0, 0) && !peel_improves(3, 2, 3, 2) && peel_improves(4, 2, 3, 2), "peel_improves");
    printf("%d graphs\n", graphs);
    printf(g_fail ? "FAILED (%d checks)\n" : "ALL OK\n", g_fail);
    return g_fail ? 1 : 0;
}
