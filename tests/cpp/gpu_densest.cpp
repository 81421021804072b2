// gpu_densest.cpp - batch-peeling driver with the reference's contract (gpu_densest.cu:1-28, 128-252):
//   usage: gpu_densest graph.txt [eps_num eps_den] [--no-burnin] [--repeats N] [--verbose]
//   rows impl,graph,n,m,best_triangles,best_vertices,density,rounds,time_ms,h2d_ms,total_ms for cpu_batch_peel,
//   cpu_batch_peel_omp, gpu_batch_peel (host loop batch_peel_loop over GpuBackend, one sync per round) and
//   gpu_batch_peel_fused (ga_peel_run, round loop on the device). Rows are printed only if every GPU run equals the CPU
//   oracle exactly (best_num, best_den, best_round, rounds and the whole removed_round array); exit 1 otherwise.
// Integration test: links the CPU oracle as the checker and as the CPU rows.
#include "../../graph_analyzer_amd/host/densest.h"
#include "../../oracle/oracle.hpp"

static bool same(const PeelResult& a, const orc::PeelResult& b) {
    return a.best_num == b.best_num && a.best_den == b.best_den && a.best_round == b.best_round && a.rounds == b.rounds &&
           a.removed_round == b.removed_round;
}

int main(int argc, char** argv) {
    std::vector<const char*> pos;
    bool burnin = true, verbose = false;
    int repeats = 5;
    for (int i = 1; i < argc; i++) {
        if (strcmp(argv[i], "--no-burnin") == 0) burnin = false;
        else if (strcmp(argv[i], "--verbose") == 0) verbose = true;
        else if (strcmp(argv[i], "--pinned") == 0 || strcmp(argv[i], "--no-greedy") == 0) {}
        else if (strcmp(argv[i], "--repeats") == 0 && i + 1 < argc) repeats = std::max(1, atoi(argv[++i]));
        else pos.push_back(argv[i]);
    }
    if (pos.empty() || pos.size() == 2 || pos.size() > 3) {
        fprintf(stderr, "usage: %s graph.txt [eps_num eps_den] [--no-burnin] [--repeats N] [--verbose]\n", argv[0]);
        return 2;
    }
    const char* path = pos[0];
    const uint64_t num = pos.size() == 3 ? strtoull(pos[1], nullptr, 10) : 1;
    const uint64_t den = pos.size() == 3 ? strtoull(pos[2], nullptr, 10) : 10;

    ga_ctx* ctx = gaa::default_ctx();
    std::vector<uint32_t> pairs;
    if (!gaa::parse_edge_file(path, pairs)) return 1;
    orc::Graph og;
    orc::graph_from_pairs(pairs.data(), pairs.size() / 2, og);
    fprintf(stderr, "loaded n=%u m=%u max_out_deg=%u eps=%llu/%llu\n", og.n, og.m, og.max_out_degree, (unsigned long long)num,
            (unsigned long long)den);
    const uint64_t T_full = orc::triangles_serial(og);

    std::vector<std::string> rows;
    double t0 = now_ms();
    const orc::PeelResult ref = orc::batch_peel(og, num, den, false);
    const double t_ref = now_ms() - t0;
    if (!ref.ok) { fprintf(stderr, "cpu_batch_peel failed: %s\n", ref.error.c_str()); return 1; }
    rows.push_back(densest_row("cpu_batch_peel", path, og.n, og.m, ref.best_num, ref.best_den, ref.rounds, t_ref, 0));
    t0 = now_ms();
    const orc::PeelResult omp = orc::batch_peel(og, num, den, true);
    const double t_omp = now_ms() - t0;
    if (!omp.ok || !orc::same(omp, ref)) { fprintf(stderr, "MISMATCH: OpenMP batch peel differs from serial\n"); return 1; }
    rows.push_back(densest_row("cpu_batch_peel_omp", path, og.n, og.m, omp.best_num, omp.best_den, omp.rounds, t_omp, 0));

    gaa::DeviceGraph dg = gaa::DeviceGraph::from_pairs(ctx, pairs.data(), pairs.size() / 2, -1);
    GpuBackend be;
    be.init(dg);
    if (burnin) {
        be.reset_state();
        const double tb = now_ms();
        do {
            be.begin_round();
            be.run_tc();
            gaa::check(ctx, ga_ctx_sync(ctx), "sync");
        } while (now_ms() - tb < 1500.0);
    }
    bool ok = true;
    auto run_once = [&](bool diag, double* ms_out) {
        gaa::check(ctx, ga_ctx_sync(ctx), "sync");
        const double ts = now_ms();
        PeelResult r = batch_peel_loop(dg.n(), T_full, num, den, be, diag, stderr);
        gaa::check(ctx, ga_ctx_sync(ctx), "sync");
        if (ms_out) *ms_out = now_ms() - ts;
        if (!r.ok) { fprintf(stderr, "GPU pipeline error: %s\n", r.error.c_str()); ok = false; return r; }
        be.download_removed_round(r.removed_round);
        if (!same(r, ref)) { fprintf(stderr, "MISMATCH gpu vs cpu_batch_peel\n"); ok = false; }
        return r;
    };
    run_once(false, nullptr);
    if (verbose) run_once(true, nullptr);
    double best = 1e300, best_fused = 1e300;
    PeelResult last;
    for (int i = 0; i < repeats && ok; i++) {
        double ms = 0;
        last = run_once(false, &ms);
        best = std::min(best, ms);
    }
    for (int i = -1; i < repeats && ok; i++) {
        const double ts = now_ms();
        PeelResult f = gpu_batch_peel(dg, T_full, num, den);
        const double ms = now_ms() - ts;
        if (!f.ok || !same(f, ref)) { fprintf(stderr, "MISMATCH fused gpu peel vs cpu_batch_peel (%s)\n", f.error.c_str()); ok = false; }
        if (i >= 0) best_fused = std::min(best_fused, ms);
    }
    if (!ok) { fprintf(stderr, "VERIFICATION FAILED - no timing rows printed\n"); return 1; }
    rows.push_back(densest_row("gpu_batch_peel", path, og.n, og.m, last.best_num, last.best_den, last.rounds, best, 0));
    rows.push_back(densest_row("gpu_batch_peel_fused", path, og.n, og.m, last.best_num, last.best_den, last.rounds, best_fused, 0));
    for (auto& r : rows) puts(r.c_str());
    fflush(stdout);
    fprintf(stderr, "GPU result verified bit-exact against cpu_batch_peel (best %llu triangles / %llu vertices, %u rounds)\n",
            (unsigned long long)last.best_num, (unsigned long long)last.best_den, last.rounds);
    return 0;
}
