// ga_tc_simple.hip - GA_TC_NOSRC and GA_TC_VERTEX inside ga_tc_count: the count variants that read neither src[] nor
// a count plan (kernels in ga_tc_simple.hpp), each workgroup adding its sum to the context's counter.
#include "ga_tc_simple.hpp"

using namespace ga_simple;

int ga_tc_launch_simple(ga_ctx* ctx, ga_graph* g, int variant, unsigned long long* total) {
    if (!g->m) return GA_OK;
    const uint32_t blocks = default_blocks(ctx, g, variant);
    if (variant == GA_TC_NOSRC) {
        k_tc_nosrc<4, false><<<blocks, SIMPLE_BLOCK, 0, ctx->stream>>>(g->rowptr, g->col, g->n, g->m, nosrc_per_wave(g->m, blocks), total,
                                                                       nullptr, nullptr, nullptr);
        return ga_launch_check(ctx, "k_tc_nosrc");
    }
    k_tc_vertex<false><<<blocks, SIMPLE_BLOCK, 0, ctx->stream>>>(g->rowptr, g->col, g->n, total, nullptr, nullptr, nullptr);
    return ga_launch_check(ctx, "k_tc_vertex");
}
