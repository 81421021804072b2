// ga_prims.hip - context, error plumbing, device memory, and the device-wide primitives the CSR builder and the
// count plan use (LSD radix sort and exclusive scans from rocPRIM; fill and max kernels written here).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <chrono>
#include <cstring>

#include "ga_internal.hpp"

int ga_fail(ga_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int ga_malloc(ga_ctx* ctx, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return ga_fail(ctx, GA_ERR_OOM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    return GA_OK;
}

void ga_free(void* p) {
    if (p) (void)hipFree(p);
}

int ga_scratch(ga_ctx* ctx, size_t bytes, void** p) {
    if (bytes > ctx->scratch_bytes) {
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ga_free(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        size_t want = bytes + bytes / 4 + 4096;
        GA_TRY(ga_malloc(ctx, &ctx->scratch, want));
        ctx->scratch_bytes = want;
    }
    *p = ctx->scratch;
    return GA_OK;
}

int ga_launch_check(ga_ctx* ctx, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ga_fail(ctx, GA_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return GA_OK;
}

extern "C" int ga_ctx_create(int device, ga_ctx** out) {
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return GA_ERR_HIP;
    if (device < 0 || device >= count) return GA_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return GA_ERR_HIP;
    ga_ctx* c = new ga_ctx;
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return GA_ERR_HIP; }
    c->cus = prop.multiProcessorCount;
    c->lds_max = prop.sharedMemPerBlock;
    if (prop.maxSharedMemoryPerMultiProcessor) c->lds_cu = prop.maxSharedMemoryPerMultiProcessor;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipHostMalloc((void**)&c->hstat, 512, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&c->dsmall, 512) != hipSuccess) {
        ga_ctx_destroy(c);
        return GA_ERR_HIP;
    }
    if (ga_noalloc_init(c) != GA_OK) {   // this device's address of the allocation-free count's static scratch
        ga_ctx_destroy(c);
        return GA_ERR_HIP;
    }
    *out = c;
    return GA_OK;
}

extern "C" void ga_ctx_destroy(ga_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    ga_free(c->scratch);
    ga_free(c->dmany);
    ga_free(c->dcore);
    ga_free(c->dtcm);
    ga_free(c->dsmall);
    if (c->hstat) (void)hipHostFree(c->hstat);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int i = 0; i < 3; i++) {
        if (c->side[i]) (void)hipStreamDestroy(c->side[i]);
        if (c->join_ev[i]) (void)hipEventDestroy(c->join_ev[i]);
    }
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* ga_ctx_last_error(const ga_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int ga_ctx_mem_info(ga_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!ctx) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    GA_HIP(ctx, hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return GA_OK;
}

double wall_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" int ga_ctx_set_grid_cap(ga_ctx* ctx, uint32_t max_workgroups) {
    if (!ctx) return GA_ERR_ARG;
    ctx->grid_cap = max_workgroups;
    return GA_OK;
}

extern "C" int ga_ctx_sync(ga_ctx* ctx) {
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

extern "C" int ga_ctx_device_info(ga_ctx* ctx, char* name, int len, int* cus, int* wave, uint64_t* hbm_bytes) {
    hipDeviceProp_t prop;
    GA_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
    if (name && len > 0) {
        snprintf(name, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cus) *cus = prop.multiProcessorCount;
    if (wave) *wave = prop.warpSize;
    if (hbm_bytes) *hbm_bytes = (uint64_t)prop.totalGlobalMem;
    return GA_OK;
}

__global__ void k_copy_probe(const uint4* __restrict__ a, uint4* __restrict__ b, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) b[i] = a[i];
}

extern "C" int ga_ctx_copy_probe(ga_ctx* ctx, uint64_t bytes, float* ms) {
    if (!ctx || bytes < 16) return GA_ERR_ARG;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    const uint64_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    GA_TRY(ga_malloc(ctx, &a, n * 16));
    if (ga_malloc(ctx, &b, n * 16)) { ga_free(a); return GA_ERR_OOM; }
    int rc = ga_fill_u32(ctx, (uint32_t*)a, n * 4, 1u);
    if (!rc && hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess) rc = GA_ERR_HIP;
    if (!rc) {
        k_copy_probe<<<grid_for(n, 256, (unsigned)std::max(ctx->cus, 1) * 16u), 256, 0, ctx->stream>>>((const uint4*)a, (uint4*)b, n);
        rc = ga_launch_check(ctx, "k_copy_probe");
    }
    if (!rc && (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess)) rc = GA_ERR_HIP;
    if (!rc && ms) (void)hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    ga_free(a);
    ga_free(b);
    return rc;
}

// Random-segment gather: every group of G lanes reads 8 independent random G-word segments per round (one u32 per
// lane, consecutive lanes on consecutive words), the access pattern of the count's short probe lists.
__device__ __forceinline__ uint32_t probe_mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

template <int G>
__global__ void k_gather_probe(const uint32_t* __restrict__ a, uint32_t nseg, uint32_t rounds, unsigned long long* __restrict__ sink) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x % G, grp = tid / G;
    uint32_t acc = 0;
    for (uint32_t r = 0; r < rounds; r++) {
        uint32_t v[8];
#pragma unroll
        for (uint32_t i = 0; i < 8; i++) {
            const uint32_t seg = probe_mix(grp * 0x9E3779B9u + r * 131u + i * 7919u) % nseg;   // < nseg
            v[i] = a[(size_t)seg * G + lane];
        }
#pragma unroll
        for (uint32_t i = 0; i < 8; i++) acc += v[i];
    }
    if (acc == 0x12345678u) atomicAdd(sink, 1ull);   // keeps the loads live
}

extern "C" int ga_ctx_gather_probe(ga_ctx* ctx, uint64_t bytes, uint32_t seg_bytes, uint64_t* segments, float* ms) {
    if (!ctx || (seg_bytes != 32 && seg_bytes != 64 && seg_bytes != 128 && seg_bytes != 256) || bytes < 4096 ||
        bytes / seg_bytes >= 0xFFFFFFFFull)
        return ga_fail(ctx, GA_ERR_ARG, "gather probe: segments of 32, 64, 128 or 256 bytes in 4 KiB .. 2^32 segments");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t nseg = (uint32_t)(bytes / seg_bytes), G = seg_bytes / 4, rounds = 32, block = 256;
    const uint32_t grid = (uint32_t)std::max(ctx->cus, 1) * 32u;
    void* a = nullptr;
    GA_TRY(ga_malloc(ctx, &a, (size_t)nseg * seg_bytes + 64));
    unsigned long long* sink = (unsigned long long*)((char*)a + (size_t)nseg * seg_bytes);
    int rc = ga_fill_u32(ctx, (uint32_t*)a, (size_t)nseg * G + 16, 1u);
    if (!rc && hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess) rc = GA_ERR_HIP;
    if (!rc) {
        switch (G) {
            case 8: k_gather_probe<8><<<grid, block, 0, ctx->stream>>>((const uint32_t*)a, nseg, rounds, sink); break;
            case 16: k_gather_probe<16><<<grid, block, 0, ctx->stream>>>((const uint32_t*)a, nseg, rounds, sink); break;
            case 32: k_gather_probe<32><<<grid, block, 0, ctx->stream>>>((const uint32_t*)a, nseg, rounds, sink); break;
            default: k_gather_probe<64><<<grid, block, 0, ctx->stream>>>((const uint32_t*)a, nseg, rounds, sink); break;
        }
        rc = ga_launch_check(ctx, "k_gather_probe");
    }
    if (!rc && (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess)) rc = GA_ERR_HIP;
    if (!rc && ms) (void)hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    if (segments) *segments = (uint64_t)grid * block / G * rounds * 8;
    ga_free(a);
    return rc;
}

// ---------------------------------------------------------------------------------------------------------------
// barrier probe: what one device-wide barrier inside a cooperative (all-resident) kernel costs, next to what one
// kernel boundary costs, for the choice between a persistent round loop and one launch per phase
// ---------------------------------------------------------------------------------------------------------------
namespace {
struct ProbeBar {
    unsigned count[8 * 16];   // per-XCD arrival counters, 64 bytes apart
    unsigned top, gen, fail, pad;
};

// Two-level barrier: workgroups arrive on their XCD's counter (blockIdx mod 8), the last of each XCD on the top
// counter, the last of those bumps the generation. Release before arriving, acquire after leaving (agent scope: L2
// write-back and invalidate on this part); the wait polls relaxed and gives up after `limit` polls (fail = 1).
__device__ __forceinline__ bool probe_grid_sync(ProbeBar* b, unsigned& gen, unsigned limit) {
    __shared__ unsigned ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned g = gen, nb = gridDim.x;
        const unsigned x = blockIdx.x % 8u, nx = (nb - x + 7u) / 8u, groups = nb < 8u ? nb : 8u;
        unsigned good = 1;
        bool last = false;
        if (atomicAdd(&b->count[x * 16], 1u) + 1u == nx) {
            __hip_atomic_store(&b->count[x * 16], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = atomicAdd(&b->top, 1u) + 1u == groups;
        }
        if (last) {
            __hip_atomic_store(&b->top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            for (unsigned spins = 0; __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g;) {
                if (++spins > limit) {
                    good = 0;
                    __hip_atomic_store(&b->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __threadfence();
        gen = g + 1;
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

__global__ void k_barrier_probe(ProbeBar* b, unsigned iters, unsigned limit, unsigned* sink) {
    unsigned gen = 0;
    for (unsigned i = 0; i < iters; i++) {
        if (threadIdx.x == 0) sink[blockIdx.x] += 1u;   // one write per workgroup per phase
        if (!probe_grid_sync(b, gen, limit)) return;
    }
}

__global__ void k_launch_probe(unsigned* sink) {
    if (threadIdx.x == 0) sink[blockIdx.x] += 1u;
}
}  // namespace

extern "C" int ga_ctx_barrier_probe(ga_ctx* ctx, uint32_t workgroups, uint32_t iters, float* us_per_barrier, float* us_per_launch) {
    if (!ctx) return GA_ERR_ARG;
    if (workgroups < 1 || workgroups > 65536 || iters < 1 || iters > 100000)
        return ga_fail(ctx, GA_ERR_ARG, "barrier probe: 1..65536 workgroups, 1..100000 barriers");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    int per_cu = 0;
    GA_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_barrier_probe, 256, 0));
    if ((uint64_t)workgroups > (uint64_t)per_cu * (uint64_t)std::max(ctx->cus, 1))
        return ga_fail(ctx, GA_ERR_ARG, "barrier probe: %u workgroups cannot all be resident (%d per CU)", workgroups, per_cu);
    void* mem = nullptr;
    GA_TRY(ga_malloc(ctx, &mem, sizeof(ProbeBar) + (size_t)workgroups * 4));
    ProbeBar* b = (ProbeBar*)mem;
    unsigned* sink = (unsigned*)((char*)mem + sizeof(ProbeBar));
    int rc = ga_fill_u32(ctx, (uint32_t*)mem, sizeof(ProbeBar) / 4 + workgroups, 0);
    const unsigned limit = 1u << 22;   // about 0.1 s of polling before a workgroup gives up
    float ms_b = 0, ms_l = 0;
    if (!rc) {
        void* args[] = {&b, &iters, (void*)&limit, &sink};
        if (hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
            hipLaunchCooperativeKernel((const void*)k_barrier_probe, dim3(workgroups), dim3(256), args, 0, ctx->stream) != hipSuccess ||
            hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
            hipEventElapsedTime(&ms_b, ctx->ev0, ctx->ev1) != hipSuccess)
            rc = ga_fail(ctx, GA_ERR_HIP, "barrier probe: cooperative launch failed");
    }
    if (!rc) {
        ProbeBar h;
        if (hipMemcpy(&h, b, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) rc = GA_ERR_HIP;
        else if (h.fail || h.gen != iters) rc = ga_fail(ctx, GA_ERR_STATE, "barrier probe: a workgroup gave up waiting");
    }
    if (!rc) {
        (void)hipEventRecord(ctx->ev0, ctx->stream);
        for (uint32_t i = 0; i < iters && !rc; i++) {
            k_launch_probe<<<workgroups, 256, 0, ctx->stream>>>(sink);
            if (hipGetLastError() != hipSuccess) rc = GA_ERR_HIP;
        }
        if (!rc && (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
                    hipEventElapsedTime(&ms_l, ctx->ev0, ctx->ev1) != hipSuccess))
            rc = GA_ERR_HIP;
    }
    ga_free(mem);
    if (rc) return rc;
    if (us_per_barrier) *us_per_barrier = ms_b * 1e3f / (float)iters;
    if (us_per_launch) *us_per_launch = ms_l * 1e3f / (float)iters;
    return GA_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// sorts and scans (rocPRIM device primitives, temporary storage from the context scratch)
// ---------------------------------------------------------------------------------------------------------------
int ga_sort_u64(ga_ctx* ctx, uint64_t* keys, uint64_t* tmp, size_t count, int end_bit, uint64_t** result) {
    *result = keys;
    if (count <= 1) return GA_OK;
    rocprim::double_buffer<uint64_t> db(keys, tmp);
    size_t bytes = 0;
    GA_HIP(ctx, rocprim::radix_sort_keys(nullptr, bytes, db, count, 0, end_bit, ctx->stream));
    void* s = nullptr;
    GA_TRY(ga_scratch(ctx, bytes, &s));
    GA_HIP(ctx, rocprim::radix_sort_keys(s, bytes, db, count, 0, end_bit, ctx->stream));
    *result = db.current();
    return GA_OK;
}

int ga_sort_pairs_u32(ga_ctx* ctx, uint32_t* keys, uint32_t* keys_tmp, uint32_t* vals, uint32_t* vals_tmp, size_t count,
                      int end_bit, uint32_t** keys_out, uint32_t** vals_out) {
    *keys_out = keys;
    *vals_out = vals;
    if (count <= 1) return GA_OK;
    rocprim::double_buffer<uint32_t> kb(keys, keys_tmp), vb(vals, vals_tmp);
    size_t bytes = 0;
    GA_HIP(ctx, rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, count, 0, end_bit, ctx->stream));
    void* s = nullptr;
    GA_TRY(ga_scratch(ctx, bytes, &s));
    GA_HIP(ctx, rocprim::radix_sort_pairs(s, bytes, kb, vb, count, 0, end_bit, ctx->stream));
    *keys_out = kb.current();
    *vals_out = vb.current();
    return GA_OK;
}

int ga_sort_pairs_u64(ga_ctx* ctx, unsigned long long* keys, unsigned long long* keys_tmp, uint32_t* vals, uint32_t* vals_tmp,
                      size_t count, int end_bit, unsigned long long** keys_out, uint32_t** vals_out) {
    *keys_out = keys;
    *vals_out = vals;
    if (count <= 1) return GA_OK;
    rocprim::double_buffer<unsigned long long> kb(keys, keys_tmp);
    rocprim::double_buffer<uint32_t> vb(vals, vals_tmp);
    size_t bytes = 0;
    GA_HIP(ctx, rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, count, 0, end_bit, ctx->stream));
    void* s = nullptr;
    GA_TRY(ga_scratch(ctx, bytes, &s));
    GA_HIP(ctx, rocprim::radix_sort_pairs(s, bytes, kb, vb, count, 0, end_bit, ctx->stream));
    *keys_out = kb.current();
    *vals_out = vb.current();
    return GA_OK;
}

int ga_exclusive_scan_u32(ga_ctx* ctx, const uint32_t* in, uint32_t* out, size_t count) {
    if (count == 0) return GA_OK;
    size_t bytes = 0;
    GA_HIP(ctx, rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, count, rocprim::plus<uint32_t>(), ctx->stream));
    void* s = nullptr;
    GA_TRY(ga_scratch(ctx, bytes, &s));
    GA_HIP(ctx, rocprim::exclusive_scan(s, bytes, in, out, 0u, count, rocprim::plus<uint32_t>(), ctx->stream));
    return GA_OK;
}

int ga_exclusive_scan_u64(ga_ctx* ctx, const uint64_t* in, uint64_t* out, size_t count) {
    if (count == 0) return GA_OK;
    size_t bytes = 0;
    GA_HIP(ctx, rocprim::exclusive_scan(nullptr, bytes, in, out, (uint64_t)0, count, rocprim::plus<uint64_t>(), ctx->stream));
    void* s = nullptr;
    GA_TRY(ga_scratch(ctx, bytes, &s));
    GA_HIP(ctx, rocprim::exclusive_scan(s, bytes, in, out, (uint64_t)0, count, rocprim::plus<uint64_t>(), ctx->stream));
    return GA_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// fills (used instead of hipMemsetAsync so every clear is an ordinary kernel node) and a max reduction
// ---------------------------------------------------------------------------------------------------------------
template <class T>
__global__ void k_fill(T* __restrict__ p, size_t count, T v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

template <class T>
static int fill_impl(ga_ctx* ctx, T* p, size_t count, T v) {
    if (count == 0) return GA_OK;
    k_fill<T><<<grid_for(count, 256, 4096), 256, 0, ctx->stream>>>(p, count, v);
    return ga_launch_check(ctx, "k_fill");
}

int ga_fill_u32(ga_ctx* ctx, uint32_t* p, size_t count, uint32_t v) { return fill_impl(ctx, p, count, v); }
int ga_fill_u8(ga_ctx* ctx, uint8_t* p, size_t count, uint8_t v) { return fill_impl(ctx, p, count, v); }
int ga_fill_u64(ga_ctx* ctx, uint64_t* p, size_t count, uint64_t v) { return fill_impl(ctx, p, count, v); }

__global__ void k_max_u32(const uint32_t* __restrict__ p, size_t count, uint32_t* __restrict__ out) {
    uint32_t mx = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        mx = p[i] > mx ? p[i] : mx;
    mx = wave_max_u32(mx);
    if (lane_id() == 0 && mx) atomicMax(out, mx);
}

int ga_max_u32(ga_ctx* ctx, const uint32_t* p, size_t count, uint32_t* out_host) {
    *out_host = 0;
    if (count == 0) return GA_OK;
    uint32_t* d = (uint32_t*)(ctx->dsmall + 63);
    GA_TRY(ga_fill_u32(ctx, d, 1, 0));
    k_max_u32<<<grid_for(count, 256, 1024), 256, 0, ctx->stream>>>(p, count, d);
    GA_TRY(ga_launch_check(ctx, "k_max_u32"));
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, d, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out_host = (uint32_t)(ctx->hstat[0] & 0xFFFFFFFFu);
    return GA_OK;
}
