// gpu_tc.cpp - triangle-count driver with the reference's contract (gpu_tc.cu:1-22, 128-164):
//   usage: gpu_tc <edge_list.txt> [repeats] [--no-burnin]
//   CSV rows impl,graph,n,m,triangles,kernel_ms,h2d_ms,gpu_total_ms, one per variant, printed only when every launch
//   of every variant equals the serial CPU count; exit 1 otherwise, exit 2 on usage errors.
// The graph is loaded on the GPU from the text file (load_graph: parse, de-duplication, orientation, sort all on the
// device) and must be byte-identical to the CPU oracle's CSR built from the host parse of the same file.
// kernel_ms = best of `repeats` device-timed launches after one untimed warm-up; h2d_ms = upload of rowptr + col
// (the CSR a caller would hand over); gpu_total_ms = kernel_ms + h2d_ms (+ plan build for the planned variants).
// Integration test: links the CPU oracle as the checker.
#include "../../graph_analyzer_amd/host/graph.h"
#include "../../oracle/oracle.hpp"

int main(int argc, char** argv) {
    std::vector<const char*> pos;
    bool burnin = true;
    for (int i = 1; i < argc; i++) {
        if (strcmp(argv[i], "--no-burnin") == 0) burnin = false;
        else if (strcmp(argv[i], "--pinned") == 0) {}
        else pos.push_back(argv[i]);
    }
    if (pos.empty()) { fprintf(stderr, "usage: %s graph.txt [repeats] [--no-burnin]\n", argv[0]); return 2; }
    const char* path = pos[0];
    const int repeats = std::max(1, pos.size() > 1 ? atoi(pos[1]) : 5);

    ga_ctx* ctx = gaa::default_ctx();
    char name[256];
    int cus = 0, wave = 0;
    uint64_t hbm = 0;
    ga_ctx_device_info(ctx, name, sizeof name, &cus, &wave, &hbm);
    fprintf(stderr, "GPU: %s, CUs=%d, wave=%d\n", name, cus, wave);

    std::vector<uint32_t> pairs;
    if (!gaa::parse_edge_file(path, pairs)) return 1;
    gaa::DeviceGraph dg;
    const double l0 = now_ms();
    if (!load_graph_device(path, dg, ctx)) return 1;
    fprintf(stderr, "GPU load_graph n=%u m=%u max_out_deg=%u in %.3f ms (file read, upload, parse, dedup, orientation, sort)\n", dg.n(),
            dg.m(), dg.max_out_degree(), now_ms() - l0);
    orc::Graph og;
    orc::graph_from_pairs(pairs.data(), pairs.size() / 2, og);
    Graph g;
    dg.download(g);
    if (g.n != og.n || g.m != og.m || g.rowptr != og.rowptr || g.col != og.col || g.src != og.src || g.max_out_degree != og.max_out_degree) {
        fprintf(stderr, "MISMATCH: GPU-built CSR differs from the CPU oracle's\n");
        return 1;
    }
    const uint64_t expected = orc::triangles_serial(og);

    const double h0 = now_ms();
    gaa::DeviceGraph up = gaa::DeviceGraph::upload(ctx, g);
    gaa::check(ctx, ga_ctx_sync(ctx), "sync");
    const double h2d = now_ms() - h0;
    float prep_ms = 0;
    gaa::check(ctx, ga_tc_prepare(ctx, up.handle(), &prep_ms), "ga_tc_prepare");
    fprintf(stderr, "count plan built in %.3f ms\n", prep_ms);

    if (burnin) {
        const double t0 = now_ms();
        do { (void)gaa::gpu_triangles(up, GA_TC_MERGE); } while (now_ms() - t0 < 1500.0);
    }
    bool ok = true;
    const int variants[] = {GA_TC_MERGE, GA_TC_GROUP4, GA_TC_GROUP8, GA_TC_GROUP16, GA_TC_GROUP64, GA_TC_VERTEX, GA_TC_NOSRC, GA_TC_SORTED,
                            GA_TC_HASH};
    for (int v : variants) {
        float best = 1e30f;
        uint64_t T = 0;
        for (int r = -1; r < repeats; r++) {
            float ms = 0;
            T = gaa::gpu_triangles(up, v, &ms);
            if (r >= 0) best = std::min(best, ms);
            if (T != expected) {
                fprintf(stderr, "MISMATCH in %s (launch %d): gpu=%llu cpu=%llu\n", ga_tc_variant_name(v), r, (unsigned long long)T,
                        (unsigned long long)expected);
                ok = false;
            }
        }
        const double extra = (v == GA_TC_HASH) ? prep_ms : 0.0;
        printf("gpu_%s,%s,%u,%u,%llu,%.3f,%.3f,%.3f\n", ga_tc_variant_name(v), path, g.n, g.m, (unsigned long long)T, best, h2d,
               best + h2d + extra);
        fflush(stdout);
    }
    if (!ok) return 1;
    fprintf(stderr, "all kernels verified against CPU count: %llu\n", (unsigned long long)expected);
    return 0;
}
