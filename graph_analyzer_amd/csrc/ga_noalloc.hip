// ga_noalloc.hip - ga_tc_count_noalloc: the triangle count of GA_TC_NOSRC / GA_TC_VERTEX (kernels in ga_tc_simple.hpp)
// for callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc). The per-workgroup partial sums and the completion ticket live in the file-scope device array
// `scratch` (GA_NOALLOC_MAX_BLOCKS + 1 words per device, held until the process ends). Host code never passes the
// array's name to a kernel: its device address on each device is looked up once, when a context is created on that
// device (ga_noalloc_init), and that address is what the kernels get.
#include <mutex>

#include "ga_tc_simple.hpp"

using namespace ga_simple;

static __device__ __align__(256) unsigned long long scratch[GA_NOALLOC_MAX_BLOCKS + 1];   // partials, then the ticket

namespace {

struct NoallocSlot {
    unsigned long long* ptr = nullptr;   // this device's address of `scratch`
    bool ready = false;
};
constexpr int NOALLOC_DEVICES = 64;
std::mutex noalloc_mu;
NoallocSlot noalloc_slot[NOALLOC_DEVICES];

// Resets the ticket before the count (a fill kernel, so a captured graph holds no memset node).
__global__ void k_noalloc_clear(unsigned long long* __restrict__ ticket) {
    if (threadIdx.x == 0) *ticket = 0ull;
}

}  // namespace

// Called by ga_ctx_create with the context's device current: looks up `scratch` on that device once.
int ga_noalloc_init(ga_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= NOALLOC_DEVICES)
        return ga_fail(ctx, GA_ERR_ARG, "device %d is beyond the %d devices of the static scratch table", ctx->device, NOALLOC_DEVICES);
    std::lock_guard<std::mutex> lk(noalloc_mu);
    NoallocSlot& s = noalloc_slot[ctx->device];
    if (s.ready) return GA_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return ga_fail(ctx, GA_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", ctx->device, hipGetErrorString(e));
    }
    s.ptr = (unsigned long long*)p;
    s.ready = true;
    return GA_OK;
}

extern "C" int ga_tc_count_noalloc(ga_ctx* ctx, const ga_graph* g, int variant, uint32_t blocks, uint64_t* d_T) {
    if (!ctx) return GA_ERR_ARG;
    if (!g || !d_T) return ga_fail(ctx, GA_ERR_ARG, "ga_tc_count_noalloc: NULL graph or output word");
    if (g->ctx && g->ctx->device != ctx->device) return ga_fail(ctx, GA_ERR_ARG, "ga_tc_count_noalloc: the graph lives on another device");
    if (variant != GA_TC_NOSRC && variant != GA_TC_VERTEX)
        return ga_fail(ctx, GA_ERR_ARG, "ga_tc_count_noalloc: variant %d needs src[] or a count plan (GA_TC_NOSRC or GA_TC_VERTEX only)",
                       variant);
    if (blocks > GA_NOALLOC_MAX_BLOCKS)
        return ga_fail(ctx, GA_ERR_ARG, "ga_tc_count_noalloc: %u workgroups asked, the static scratch holds partial sums for at most %u",
                       blocks, (unsigned)GA_NOALLOC_MAX_BLOCKS);
    unsigned long long* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(noalloc_mu);
        if (ctx->device >= 0 && ctx->device < NOALLOC_DEVICES && noalloc_slot[ctx->device].ready) ptr = noalloc_slot[ctx->device].ptr;
    }
    if (!ptr) return ga_fail(ctx, GA_ERR_STATE, "ga_tc_count_noalloc: the static scratch of device %d is not set up", ctx->device);
    GA_HIP(ctx, hipSetDevice(ctx->device));
    if (!blocks) blocks = default_blocks(ctx, g, variant);
    unsigned long long* ticket = ptr + GA_NOALLOC_MAX_BLOCKS;
    k_noalloc_clear<<<1, 64, 0, ctx->stream>>>(ticket);
    GA_TRY(ga_launch_check(ctx, "k_noalloc_clear"));
    if (variant == GA_TC_NOSRC)
        k_tc_nosrc<4, true><<<blocks, SIMPLE_BLOCK, 0, ctx->stream>>>(g->rowptr, g->col, g->n, g->m, nosrc_per_wave(g->m, blocks), nullptr,
                                                                      ptr, ticket, d_T);
    else
        k_tc_vertex<true><<<blocks, SIMPLE_BLOCK, 0, ctx->stream>>>(g->rowptr, g->col, g->n, nullptr, ptr, ticket, d_T);
    return ga_launch_check(ctx, variant == GA_TC_NOSRC ? "k_tc_nosrc" : "k_tc_vertex");
}
