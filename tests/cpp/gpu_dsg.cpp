// gpu_dsg.cpp - (k,triangle)-core driver with the core part of the reference's contract (gpu_dsg.cu:1-46, 201-316):
//   usage: gpu_dsg graph.txt [--T 10] [--repeats N] [--with-recount] [--no-burnin] [--no-greedy] [--no-greedypp] [--verbose]
//   rows impl,graph,n,m,triangles,vertices,density,time_ms,h2d_ms,total_ms,rounds,kmax for cpu_core_decomp (Alg 3),
//   gpu_core_decomp (incremental host loop, one sync per round), gpu_core_decomp_fused (ga_core_run) and, with
//   --with-recount, gpu_core_decomp_recount. Rows = the kmax-core. Then the lower bound (gpu_dsg.cu:318-330): the
//   kmax-core density, raised to Greedy's on the whole graph when that is denser (cpu_greedy, the oracle's one-pass
//   greedy_pp as the reference's CPU row; --no-greedy skips it, as on graphs with huge hubs), and prune_kcore: keep
//   core >= ceil(lb) on the device core numbers (ga_core_prune), row n,m = the pruned graph, rounds = ceil(lb). Printed only if every GPU core array equals CPU Alg 3 element-wise, the round counts of all GPU
//   designs agree, and the pruned graph equals the oracle's prune_to_core with its own core numbers equal to the
//   restricted ones; exit 1 otherwise. Then Greedy++ with T passes on the pruned graph (gpu_dsg.cu:344-356):
//   cpu_greedy_pp_pruned (the oracle) and gpu_greedy_pp_pruned (ga_greedy_pp), rounds = T; the GPU result (density, set
//   and the best after every pass) must equal the oracle's. --no-greedypp skips it, and so does a pruned graph larger
//   than ga_greedy_max_vertices (said on stderr). Then CoreExact on the pruned graph seeded with the Greedy++ result
//   (gpu_dsg.cu:348-361): gpu_core_exact_pruned_seeded (core_exact through ga_core_exact), rounds = flow calls. Checked
//   against the oracle: the set's triangle count, a density no lower than Greedy++'s, and, on pruned graphs of at most
//   2,000 vertices, the same density from Exact (Alg 1) on the whole pruned graph; as gpu_dsg.cu:362-363 the optimum
//   may not lie below the lower bound nor below Greedy++'s. Then the pipeline rows (gpu_dsg.cu:365-369):
//   pipeline_gpu_core (fused GPU core + Greedy + prune + Greedy++ + CoreExact) and pipeline_cpu_core (CPU Alg 3 + the
//   same remainder). --no-exact skips all of that. --verify-full (gpu_dsg.cu:372-385): core_exact on the whole graph
//   with the CPU core numbers, gpu_core_exact_full, whose optimum must equal the pruned pipeline's.
// Integration test: links the CPU oracle as the checker and as the CPU row.
#include "../../graph_analyzer_amd/host/dsg.h"
#include "../../oracle/oracle.hpp"

int main(int argc, char** argv) {
    std::vector<const char*> pos;
    bool recount = false, burnin = true, verbose = false, gpp = true, greedy_on = true, exact_on = true, verify_full = false;
    int repeats = 5;
    uint32_t T_pass = 10;
    for (int i = 1; i < argc; i++) {
        if (strcmp(argv[i], "--with-recount") == 0) recount = true;
        else if (strcmp(argv[i], "--no-burnin") == 0) burnin = false;
        else if (strcmp(argv[i], "--verbose") == 0) verbose = true;
        else if (strcmp(argv[i], "--no-greedypp") == 0) gpp = false;
        else if (strcmp(argv[i], "--no-greedy") == 0) greedy_on = false;
        else if (strcmp(argv[i], "--no-exact") == 0) exact_on = false;
        else if (strcmp(argv[i], "--verify-full") == 0) verify_full = true;
        else if (strcmp(argv[i], "--T") == 0 && i + 1 < argc) T_pass = (uint32_t)std::max(1, atoi(argv[++i]));
        else if (strcmp(argv[i], "--repeats") == 0 && i + 1 < argc) repeats = std::max(1, atoi(argv[++i]));
        else if (argv[i][0] == '-') {}
        else pos.push_back(argv[i]);
    }
    if (pos.size() != 1) {
        fprintf(stderr, "usage: %s graph.txt [--T 10] [--repeats N] [--with-recount] [--no-burnin] [--no-greedy] [--no-greedypp] [--no-exact] [--verify-full] [--verbose]\n",
                argv[0]);
        return 2;
    }
    const char* path = pos[0];
    ga_ctx* ctx = gaa::default_ctx();
    std::vector<uint32_t> pairs;
    if (!gaa::parse_edge_file(path, pairs)) return 1;
    orc::Graph og;
    orc::graph_from_pairs(pairs.data(), pairs.size() / 2, og);
    fprintf(stderr, "loaded n=%u m=%u max_out_deg=%u\n", og.n, og.m, og.max_out_degree);
    const uint64_t T_full = orc::triangles_serial(og);

    std::vector<std::string> rows;
    std::vector<uint32_t> core_cpu;
    uint32_t kmax = 0;
    double t0 = now_ms();
    if (!orc::core_decomposition(og, core_cpu, &kmax)) { fprintf(stderr, "core number does not fit uint32\n"); return 1; }
    const double t_cpu = now_ms() - t0;
    uint64_t ktri = 0, kverts = 0;
    orc::core_level_density(og, core_cpu, kmax, ktri, kverts);
    rows.push_back(dsg_row("cpu_core_decomp", path, og.n, og.m, ktri, kverts, t_cpu, 0, 0, kmax));

    gaa::DeviceGraph dg = gaa::DeviceGraph::from_pairs(ctx, pairs.data(), pairs.size() / 2, -1);
    GpuCoreBackendInc bi;
    bi.init(dg);
    GpuCoreBackend br;
    if (recount) br.init(dg);
    if (burnin) {
        bi.reset_state();
        const double tb = now_ms();
        uint64_t Tb = 0;
        do { bi.init_tri(Tb); } while (now_ms() - tb < 1500.0);
    }
    bool ok = true;
    std::vector<uint32_t> core_gpu;
    auto check = [&](const char* what, const CoreLoopResult& r) {
        if (!r.ok) { fprintf(stderr, "%s error: %s\n", what, r.error.c_str()); ok = false; return; }
        if (core_gpu != core_cpu || r.kmax != kmax) {
            size_t bad = 0;
            while (bad < core_gpu.size() && bad < core_cpu.size() && core_gpu[bad] == core_cpu[bad]) bad++;
            fprintf(stderr, "MISMATCH %s core numbers vs CPU Alg 3: kmax %u vs %u, first differing vertex %zu\n", what, r.kmax, kmax, bad);
            ok = false;
        }
    };
    CoreLoopResult inc;
    double best_inc = 1e300, best_fused = 1e300, best_rc = 1e300;
    for (int i = -1; i < repeats && ok; i++) {
        gaa::check(ctx, ga_ctx_sync(ctx), "sync");
        const double ts = now_ms();
        inc = core_decomp_loop_inc(dg.n(), T_full, bi, verbose && i < 0, stderr);
        const double ms = now_ms() - ts;
        bi.download_core(core_gpu);
        check("gpu incremental", inc);
        if (i >= 0) best_inc = std::min(best_inc, ms);
    }
    uint32_t fused_rounds = 0;
    for (int i = -1; i < repeats && ok; i++) {
        const double ts = now_ms();
        CoreLoopResult f = gpu_core_run(dg, T_full, core_gpu);
        const double ms = now_ms() - ts;
        check("gpu fused", f);
        fused_rounds = f.rounds;
        if (i >= 0) best_fused = std::min(best_fused, ms);
    }
    uint32_t rc_rounds = inc.rounds;
    if (recount) {
        for (int i = -1; i < repeats && ok; i++) {
            const double ts = now_ms();
            CoreLoopResult r = core_decomp_loop(dg.n(), T_full, br);
            const double ms = now_ms() - ts;
            br.download_core(core_gpu);
            check("gpu recount", r);
            rc_rounds = r.rounds;
            if (i >= 0) best_rc = std::min(best_rc, ms);
        }
    }
    if (ok && (fused_rounds != inc.rounds || rc_rounds != inc.rounds)) {
        fprintf(stderr, "MISMATCH round counts: incremental %u, fused %u, recount %u\n", inc.rounds, fused_rounds, rc_rounds);
        ok = false;
    }
    // ---- lower bound and pruning (gpu_dsg.cu:318-330) ----
    uint64_t kprune = 0, T_pruned = 0;
    double t_prune = 1e300, t_greedy = 0;
    Dens lb{ktri, kverts};
    orc::GreedyResult cgr;
    if (ok && greedy_on) {
        const double tg = now_ms();
        cgr = orc::greedy_pp(og, 1);
        t_greedy = now_ms() - tg;
        if (!cgr.ok) { fprintf(stderr, "greedy failed: triangle bookkeeping inconsistent\n"); return 1; }
        if (dens_gt(Dens{cgr.tri, cgr.verts}, lb)) lb = Dens{cgr.tri, cgr.verts};
    }
    Graph gp;
    std::vector<uint32_t> core_pruned;
    if (ok) {
        kprune = dens_ceil(lb);
        std::vector<uint32_t> old_id, core_p;
        for (int i = -1; i < repeats; i++) {
            gaa::check(ctx, ga_ctx_sync(ctx), "sync");
            const double ts = now_ms();
            gaa::DeviceGraph dp = bi.prune(kprune, old_id, core_p);
            const double ms = now_ms() - ts;
            if (i >= 0) t_prune = std::min(t_prune, ms);
            if (i < 0) {
                dp.download(gp);
                T_pruned = gaa::gpu_triangles(dp);
            }
        }
        orc::Graph op;
        std::vector<uint32_t> oold, ocore;
        orc::prune_to_core(og, core_cpu, kprune, op, oold, ocore);
        std::vector<uint32_t> own;
        if (gp.n != op.n || gp.m != op.m || gp.rowptr != op.rowptr || gp.col != op.col || gp.src != op.src || old_id != oold || core_p != ocore) {
            fprintf(stderr, "MISMATCH: GPU prune_to_core differs from the oracle's\n");
            ok = false;
        } else if (!orc::core_decomposition(op, own) || own != core_p || T_pruned != orc::triangles_serial(op)) {
            fprintf(stderr, "MISMATCH: restricted core numbers != core numbers of the pruned graph (or its count differs)\n");
            ok = false;
        } else {
            fprintf(stderr, "lower bound %llu/%llu -> keep core >= %llu: pruned graph n=%u m=%u (of n=%u m=%u)\n", (unsigned long long)lb.num,
                    (unsigned long long)lb.den, (unsigned long long)kprune, gp.n, gp.m, og.n, og.m);
            core_pruned = core_p;
        }
    }
    // ---- Greedy++ on the pruned graph (gpu_dsg.cu:344-356) ----
    DsgResult gpr;
    double t_gpp = 1e300, t_gpp_cpu = 0;
    bool have_gpp = false;
    if (ok && gpp && gp.n > ga_greedy_max_vertices(ctx)) {
        fprintf(stderr, "pruned graph n=%u exceeds greedy_pp's %u vertices: no Greedy++ rows\n", gp.n, ga_greedy_max_vertices(ctx));
    } else if (ok && gpp) {
        gaa::DeviceGraph dgp = gaa::DeviceGraph::upload(ctx, gp);
        for (int i = -1; i < repeats && ok; i++) {
            float ms = 0;
            gpr = greedy_pp_device(dgp, T_pass, &ms);
            if (!gpr.ok) { fprintf(stderr, "greedy_pp failed: %s\n", gpr.error.c_str()); ok = false; }
            if (i >= 0) t_gpp = std::min(t_gpp, (double)ms);
        }
        orc::Graph op;
        std::vector<uint32_t> oold, ocore;
        orc::prune_to_core(og, core_cpu, kprune, op, oold, ocore);
        const double tc = now_ms();
        const orc::GreedyResult cg = orc::greedy_pp(op, T_pass);
        t_gpp_cpu = now_ms() - tc;
        bool same = ok && cg.ok && cg.tri == gpr.tri && cg.verts == gpr.verts && cg.set == gpr.set && cg.per_pass.size() == gpr.per_pass.size();
        for (size_t i = 0; same && i < cg.per_pass.size(); i++)
            same = cg.per_pass[i].first == gpr.per_pass[i].num && cg.per_pass[i].second == gpr.per_pass[i].den;
        if (!same) {
            fprintf(stderr, "MISMATCH: GPU greedy_pp differs from the oracle's\n");
            ok = false;
        } else {
            have_gpp = true;
            fprintf(stderr, "greedy++ (%u passes) on the pruned graph: %llu/%llu\n", T_pass, (unsigned long long)gpr.tri, (unsigned long long)gpr.verts);
        }
    }
    // ---- CoreExact on the pruned graph, seeded with Greedy++ (gpu_dsg.cu:348-361) ----
    DsgResult exr;
    double t_exact = 1e300;
    bool have_exact = false;
    if (ok && exact_on) {
        gaa::DeviceGraph dgp = gaa::DeviceGraph::upload(ctx, gp);
        for (int i = -1; i < repeats && ok; i++) {
            gaa::check(ctx, ga_ctx_sync(ctx), "sync");
            const double te = now_ms();
            exr = core_exact_device(dgp, core_pruned, true, have_gpp ? &gpr : nullptr);
            if (i >= 0) t_exact = std::min(t_exact, now_ms() - te);
            if (!exr.ok) { fprintf(stderr, "core_exact failed: %s\n", exr.error.c_str()); ok = false; }
        }
        if (ok) {
            orc::Graph op;
            std::vector<uint32_t> oold, ocore;
            orc::prune_to_core(og, core_cpu, kprune, op, oold, ocore);
            std::vector<uint8_t> in(op.n, 0);
            for (uint32_t v : exr.set) in[v] = 1;
            std::vector<uint32_t> otri;
            const uint64_t tin = orc::participation(op, in, otri);   // triangles inside the set
            bool good = exr.set.size() == exr.verts && tin == exr.tri && (!have_gpp || !dens_gt(Dens{gpr.tri, gpr.verts}, exr.dens())) &&
                        dens_cmp(exr.dens(), lb) >= 0;   // gpu_dsg.cu:362-363: not below the lower bound, nor below Greedy++
            if (good && gp.n <= 2000) {
                const DsgResult ed = exact_densest(gp);
                good = ed.ok && dens_cmp(ed.dens(), exr.dens()) == 0;
            }
            if (!good) {
                fprintf(stderr, "MISMATCH: core_exact's result fails the oracle's checks\n");
                ok = false;
            } else {
                have_exact = true;
                fprintf(stderr, "core_exact on the pruned graph: %llu/%llu, %u flow calls\n", (unsigned long long)exr.tri,
                        (unsigned long long)exr.verts, exr.flow_calls);
            }
        }
    }
    // ---- --verify-full: CoreExact on the whole graph with the CPU core numbers (gpu_dsg.cu:372-385) ----
    DsgResult exf;
    double t_full = 0;
    if (ok && have_exact && verify_full) {
        const double tf = now_ms();
        exf = core_exact_device(dg, core_cpu);
        t_full = now_ms() - tf;
        if (!exf.ok) { fprintf(stderr, "core_exact (full) failed: %s\n", exf.error.c_str()); ok = false; }
        else if (dens_cmp(exf.dens(), exr.dens()) != 0) {
            fprintf(stderr, "MISMATCH: full-graph CoreExact %llu/%llu vs pruned pipeline %llu/%llu\n", (unsigned long long)exf.tri,
                    (unsigned long long)exf.verts, (unsigned long long)exr.tri, (unsigned long long)exr.verts);
            ok = false;
        } else {
            fprintf(stderr, "verify-full: optimum identical (%llu/%llu) with and without pruning\n", (unsigned long long)exf.tri,
                    (unsigned long long)exf.verts);
        }
    }
    if (!ok) { fprintf(stderr, "VERIFICATION FAILED - no timing rows printed\n"); return 1; }
    rows.push_back(dsg_row("gpu_core_decomp", path, og.n, og.m, ktri, kverts, best_inc, 0, inc.rounds, kmax));
    rows.push_back(dsg_row("gpu_core_decomp_fused", path, og.n, og.m, ktri, kverts, best_fused, 0, fused_rounds, kmax));
    if (recount) rows.push_back(dsg_row("gpu_core_decomp_recount", path, og.n, og.m, ktri, kverts, best_rc, 0, rc_rounds, kmax));
    if (greedy_on) rows.push_back(dsg_row("cpu_greedy", path, og.n, og.m, cgr.tri, cgr.verts, t_greedy, 0, og.n, 0));
    rows.push_back(dsg_row("prune_kcore", path, gp.n, gp.m, T_pruned, gp.n, t_prune, 0, kprune, kmax));
    if (have_gpp) {
        rows.push_back(dsg_row("cpu_greedy_pp_pruned", path, gp.n, gp.m, gpr.tri, gpr.verts, t_gpp_cpu, 0, T_pass, 0));
        rows.push_back(dsg_row("gpu_greedy_pp_pruned", path, gp.n, gp.m, gpr.tri, gpr.verts, t_gpp, 0, T_pass, 0));
    }
    if (have_exact) {
        rows.push_back(dsg_row("gpu_core_exact_pruned_seeded", path, gp.n, gp.m, exr.tri, exr.verts, t_exact, 0, exr.flow_calls, 0));
        const double rest = (greedy_on ? t_greedy : 0.0) + t_prune + (have_gpp ? t_gpp : 0.0) + t_exact;
        rows.push_back(dsg_row("pipeline_gpu_core", path, og.n, og.m, exr.tri, exr.verts, best_fused + rest, 0, fused_rounds, kmax));
        rows.push_back(dsg_row("pipeline_cpu_core", path, og.n, og.m, exr.tri, exr.verts, t_cpu + rest, 0, 0, kmax));
    }
    if (ok && have_exact && verify_full)
        rows.push_back(dsg_row("gpu_core_exact_full", path, og.n, og.m, exf.tri, exf.verts, t_full, 0, exf.flow_calls, 0));
    for (auto& r : rows) puts(r.c_str());
    fflush(stdout);
    fprintf(stderr, "GPU core numbers verified element-wise against CPU Alg 3 (kmax %u, %u rounds)\n", kmax, inc.rounds);
    return 0;
}
