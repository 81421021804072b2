// ga_io.hip - graph ingest on the GPU: the SNAP text parse of load_graph (graph.h:69-96) and a binary CSR file format.
//
// Text. The host streams the file through two 64 MiB pinned buffers into device memory (reading one chunk while the
// previous one is copied); the parse runs on the device with the exact
// rules of the reference's loop over fgets(line, 512) + strtoul(base 10):
//   * a physical line [s, e) (e just past its '\n', or the end of the file) reaches the loop as fgets chunks of at most
//     511 bytes: [s, s+511), [s+511, s+1022), ..., each handled as a line of its own;
//   * a chunk whose first byte is '#', '%' or '\n' is skipped;
//   * strtoul skips C-locale white space, takes an optional sign and base-10 digits, stops at any other byte (NUL
//     included) or at the chunk's end, saturates to ULONG_MAX on overflow and negates modulo 2^64 after a '-'; a
//     chunk without a first or a second number is skipped;
//   * a == b (as 64-bit values) is skipped; otherwise the pair ((u32)min, (u32)max) is kept.
// The kept pairs go to the GPU builder with load_graph semantics (loops dropped, de-duplication, n = max_id + 1).
// Skipped chunks write the loop (0, 0), which the builder drops, so no compaction pass is needed.
//
// Steps: newline counts per 4 KiB tile -> scan -> newline positions in file order (u64) -> one lane per physical line
// parses its first chunk into pairs[line]; the rare further chunks of lines longer than 511 bytes are appended after
// the per-line slots with one atomic each (the builder sorts, so order does not matter).
//
// Binary CSR. A 64-byte header {magic "GACSR001", n, m, max_out_degree (u64 each), 32 reserved bytes} followed by
// rowptr u32[n+1] and col u32[m]. A load validates on the device that the arrays are exactly what the builder makes
// for their edge set: rowptr starts at 0, never decreases and ends at m; every col < n; rows strictly ascending; every
// edge u -> v oriented by the (degree, id) rule with the undirected degrees of the file's own edges; max_out_degree
// matches. Anything else is rejected with GA_ERR_ARG.
#include <cstring>

#include "ga_internal.hpp"

namespace {

struct DevBuf {   // scoped device allocation
    void* p = nullptr;
    ~DevBuf() { ga_free(p); }
};

constexpr uint32_t FGETS_MAX = 511;  // fgets(line, 512, f) returns at most 511 bytes per call

// Newline bytes in a 32-bit word: x has a zero byte where w had '\n'; ((x & 0x7F..) + 0x7F..) | x has the top bit of a
// byte set exactly when that byte is non-zero (the add cannot carry into the next byte), so the count is exact.
__device__ __forceinline__ uint32_t nl_in_word(uint32_t w) {
    const uint32_t x = w ^ 0x0A0A0A0Au;
    const uint32_t nz = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
    return (uint32_t)__popc(~nz & 0x80808080u);
}

__device__ __forceinline__ uint32_t nl_count16(uint4 q) { return nl_in_word(q.x) + nl_in_word(q.y) + nl_in_word(q.z) + nl_in_word(q.w); }

// Newlines per tile. The buffer is padded with zero bytes to a multiple of 16, so every 16-byte load is in bounds.
__global__ void __launch_bounds__(256) k_nl_count(const uint4* __restrict__ buf, uint64_t nvec, uint32_t* __restrict__ tile_cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t c = i < nvec ? nl_count16(buf[i]) : 0u;
    c = (uint32_t)wave_sum_u64(c);
    __shared__ uint32_t part[4];
    if (lane_id() == 0) part[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Newline positions in file order: tile offset (scan) + the lane's offset inside the tile (workgroup scan).
__global__ void __launch_bounds__(256) k_nl_write(const uint4* __restrict__ buf, uint64_t nvec, const uint32_t* __restrict__ tile_off,
                                                  unsigned long long* __restrict__ nlpos) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint4 q = i < nvec ? buf[i] : make_uint4(0, 0, 0, 0);
    const uint32_t c = nl_count16(q);
    // inclusive scan of c over the wave, then across the four waves
    uint32_t x = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if ((int)lane_id() >= o) x += y;
    }
    __shared__ uint32_t wsum[4];
    if (lane_id() == 63) wsum[threadIdx.x / 64] = x;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < threadIdx.x / 64; w++) before += wsum[w];
    if (!c) return;
    const unsigned long long at = (unsigned long long)tile_off[blockIdx.x] + before + x - c;
    const uint32_t words[4] = {q.x, q.y, q.z, q.w};
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint32_t byte = (words[k >> 2] >> (8 * (k & 3))) & 0xFFu;
        if (byte == 0x0Au && w < c) nlpos[at + w++] = i * 16 + k;   // w < c: never more writes than counted
    }
}

__device__ __forceinline__ bool c_space(uint32_t ch) { return ch == ' ' || (ch >= '\t' && ch <= '\r'); }   // \t \n \v \f \r

// strtoul(base 10) on bytes [p, end), `end` standing for the NUL fgets writes after a chunk. Returns false when no
// digit follows the optional white space and sign (strtoul's "no conversion": endptr = its argument).
__device__ __forceinline__ bool dev_strtoul(const uint8_t* __restrict__ b, uint64_t p, uint64_t end, unsigned long long& val, uint64_t& stop) {
    while (p < end && c_space(b[p])) p++;
    bool neg = false;
    if (p < end && (b[p] == '+' || b[p] == '-')) {
        neg = b[p] == '-';
        p++;
    }
    if (p >= end || (uint32_t)(b[p] - '0') > 9u) return false;
    unsigned long long v = 0;
    bool ovf = false;
    while (p < end) {
        const uint32_t d = (uint32_t)b[p] - '0';
        if (d > 9u) break;
        if (v > (~0ull - d) / 10ull) ovf = true;
        else v = v * 10ull + d;
        p++;
    }
    val = ovf ? ~0ull : (neg ? 0ull - v : v);
    stop = p;
    return true;
}

// One fgets chunk [cs, ce) -> (u, v), or (0, 0) when the loop would skip it.
__device__ __forceinline__ uint2 parse_chunk(const uint8_t* __restrict__ b, uint64_t cs, uint64_t ce) {
    const uint32_t c0 = b[cs];
    if (c0 == '#' || c0 == '%' || c0 == '\n') return make_uint2(0, 0);
    unsigned long long a, bv;
    uint64_t p1, p2;
    if (!dev_strtoul(b, cs, ce, a, p1)) return make_uint2(0, 0);
    if (!dev_strtoul(b, p1, ce, bv, p2)) return make_uint2(0, 0);
    if (a == bv) return make_uint2(0, 0);
    const unsigned long long lo = a < bv ? a : bv, hi = a < bv ? bv : a;
    return make_uint2((uint32_t)lo, (uint32_t)hi);
}

// One lane per physical line: its first chunk into pairs[line]; further chunks appended at pairs[lines + cursor].
__global__ void k_parse_lines(const uint8_t* __restrict__ b, uint64_t len, const unsigned long long* __restrict__ nlpos, uint64_t nnl,
                              uint64_t lines, uint2* __restrict__ pairs, unsigned long long* __restrict__ cursor) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < lines; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = i == 0 ? 0 : nlpos[i - 1] + 1;
        const uint64_t e = i < nnl ? nlpos[i] + 1 : len;
        const uint64_t ce = e - s > FGETS_MAX ? s + FGETS_MAX : e;
        pairs[i] = parse_chunk(b, s, ce);
        for (uint64_t cs = ce; cs < e; cs += FGETS_MAX) {   // lines longer than 511 bytes only
            const uint64_t at = atomicAdd(cursor, 1ull);
            pairs[lines + at] = parse_chunk(b, cs, e - cs > FGETS_MAX ? cs + FGETS_MAX : e);
        }
    }
}

// ---- binary CSR validation ----
__global__ void k_csr_check_rows(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n, uint32_t m,
                                 uint32_t* __restrict__ deg, uint32_t* __restrict__ bad) {
    uint32_t nb = 0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const uint32_t b = rowptr[u], e = rowptr[u + 1];
        if (e < b || e > m) { nb++; continue; }
        atomicAdd(&deg[u], e - b);
        uint32_t prev = 0;
        for (uint32_t p = b; p < e; p++) {
            const uint32_t v = col[p];
            if (v >= n || (p > b && v <= prev)) { nb++; break; }
            prev = v;
            atomicAdd(&deg[v], 1u);
        }
    }
    nb = (uint32_t)wave_sum_u64(nb);
    if (lane_id() == 0 && nb) atomicAdd(bad, nb);
}

// Every edge u -> v must satisfy (deg u, u) < (deg v, v) (graph.h:43-45); rows were checked first.
__global__ void k_csr_check_orient(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                   const uint32_t* __restrict__ deg, uint32_t* __restrict__ bad) {
    uint32_t nb = 0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const uint32_t du = deg[u];
        for (uint32_t p = rowptr[u]; p < rowptr[u + 1]; p++) {
            const uint32_t v = col[p], dv = deg[v];
            nb += !(du < dv || (du == dv && u < v));
        }
    }
    nb = (uint32_t)wave_sum_u64(nb);
    if (lane_id() == 0 && nb) atomicAdd(bad, nb);
}

__global__ void k_rows_src(const uint32_t* __restrict__ rowptr, uint32_t n, uint32_t* __restrict__ src, uint32_t* __restrict__ outdeg) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const uint32_t b = rowptr[u], e = rowptr[u + 1];
        outdeg[u] = e - b;
        for (uint32_t p = b; p < e; p++) src[p] = u;
    }
}

struct FileHeader {
    char magic[8];
    uint64_t n, m, max_out;
    uint8_t reserved[32];
};
static_assert(sizeof(FileHeader) == 64, "header is 64 bytes");
constexpr char MAGIC[8] = {'G', 'A', 'C', 'S', 'R', '0', '0', '1'};

struct HostPinned {   // scoped pinned host allocation
    void* p = nullptr;
    ~HostPinned() {
        if (p) (void)hipHostFree(p);
    }
};

}  // namespace

// Streams len bytes of f (from its current position) to the device through two 64 MiB pinned buffers, reading one
// chunk while the previous one is copied: bytes [0, len1) go to dst1, the rest to dst2. Captures the 4-byte words at
// the file offsets in `at` (multiples of 4, so no word straddles two chunks) and the last byte.
struct Capture {
    uint64_t at;
    uint32_t* out;
};

static int stream_file(ga_ctx* ctx, FILE* f, const char* path, uint64_t len, uint8_t* dst1, uint64_t len1, uint8_t* dst2, Capture* cap,
                       int ncap, char* last) {
    constexpr size_t STAGE = 64u << 20;
    HostPinned hb[2];
    hipEvent_t done[2] = {nullptr, nullptr};
    struct Evs {
        hipEvent_t* e;
        ~Evs() { for (int i = 0; i < 2; i++) if (e[i]) (void)hipEventDestroy(e[i]); }
    } evs{done};
    struct Drain {   // runs before the staging buffers are freed, on every return: no copy out of them is in flight
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{ctx->stream};
    if (!len) return GA_OK;
    for (int i = 0; i < 2; i++) {
        if (hipHostMalloc(&hb[i].p, STAGE) != hipSuccess) return ga_fail(ctx, GA_ERR_OOM, "pinned staging failed");
        GA_HIP(ctx, hipEventCreate(&done[i]));
    }
    uint64_t pos = 0;
    for (int i = 0; pos < len; i ^= 1) {
        GA_HIP(ctx, hipEventSynchronize(done[i]));   // the copy out of this buffer (two chunks back) has finished
        const size_t want = (size_t)std::min<uint64_t>(STAGE, len - pos);
        const uint8_t* hp = (const uint8_t*)hb[i].p;
        if (fread(hb[i].p, 1, want, f) != want) return ga_fail(ctx, GA_ERR_ARG, "short read of %s", path);
        for (int c = 0; c < ncap; c++)
            if (cap[c].at >= pos && cap[c].at + 4 <= pos + want) std::memcpy(cap[c].out, hp + (cap[c].at - pos), 4);
        const uint64_t end = pos + want;
        if (pos < len1) {
            const uint64_t k = std::min<uint64_t>(end, len1) - pos;
            GA_HIP(ctx, hipMemcpyAsync(dst1 + pos, hp, k, hipMemcpyHostToDevice, ctx->stream));
        }
        if (end > len1) {
            const uint64_t from = std::max<uint64_t>(pos, len1);
            GA_HIP(ctx, hipMemcpyAsync(dst2 + (from - len1), hp + (from - pos), end - from, hipMemcpyHostToDevice, ctx->stream));
        }
        GA_HIP(ctx, hipEventRecord(done[i], ctx->stream));
        if (last) *last = (char)hp[want - 1];
        pos = end;
    }
    return GA_OK;
}

// The device text buffer: nvec 16-byte vectors, the last one zeroed first so the padding after the text is zero.
static int text_buffer(ga_ctx* ctx, uint64_t len, DevBuf& db) {
    const uint64_t nvec = (len + 15) / 16;
    GA_TRY(ga_malloc(ctx, &db.p, std::max<uint64_t>(nvec, 1) * 16));
    if (nvec) GA_TRY(ga_fill_u32(ctx, (uint32_t*)((uint8_t*)db.p + (nvec - 1) * 16), 4, 0));
    return GA_OK;
}

// Parse and build from the text in db (len bytes, zero-padded to 16); `tail`: the last byte is not '\n' (a final line
// without newline). Frees the text buffer once the pairs are made. ms: device time from the first parse kernel.
static int text_to_graph(ga_ctx* ctx, DevBuf& db, uint64_t len, bool tail, ga_graph** out, float* ms) {
    const uint64_t nvec = (len + 15) / 16;
    const uint64_t ntiles = std::max<uint64_t>((nvec + 255) / 256, 1);
    if (ntiles > 0xFFFFFFFFull) return ga_fail(ctx, GA_ERR_ARG, "text too large");
    DevBuf tcnt, toff, nlb, pb, curb;
    GA_TRY(ga_malloc(ctx, &tcnt.p, (ntiles + 1) * 4));
    GA_TRY(ga_malloc(ctx, &toff.p, (ntiles + 1) * 4));
    GA_TRY(ga_malloc(ctx, &curb.p, 64));
    unsigned long long* cursor = (unsigned long long*)curb.p;
    GA_TRY(ga_fill_u64(ctx, (uint64_t*)cursor, 1, 0));
    GA_TRY(ga_fill_u32(ctx, (uint32_t*)tcnt.p, ntiles + 1, 0));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GA_HIP(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return ga_fail(ctx, GA_ERR_HIP, "hipEventCreate failed"); }
    struct Ev {
        hipEvent_t a, b;
        ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};
    GA_HIP(ctx, hipEventRecord(e0, ctx->stream));
    uint64_t nnl = 0;
    if (nvec) {
        k_nl_count<<<(unsigned)ntiles, 256, 0, ctx->stream>>>((const uint4*)db.p, nvec, (uint32_t*)tcnt.p);
        GA_TRY(ga_launch_check(ctx, "k_nl_count"));
        GA_TRY(ga_exclusive_scan_u32(ctx, (const uint32_t*)tcnt.p, (uint32_t*)toff.p, ntiles + 1));
        GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, (uint32_t*)toff.p + ntiles, 4, hipMemcpyDeviceToHost, ctx->stream));
        GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        nnl = ctx->hstat[0] & 0xFFFFFFFFu;
    }
    GA_TRY(ga_malloc(ctx, &nlb.p, std::max<uint64_t>(nnl, 1) * 8));
    if (nnl) {
        k_nl_write<<<(unsigned)ntiles, 256, 0, ctx->stream>>>((const uint4*)db.p, nvec, (const uint32_t*)toff.p, (unsigned long long*)nlb.p);
        GA_TRY(ga_launch_check(ctx, "k_nl_write"));
    }
    const uint64_t lines = nnl + (tail ? 1 : 0);
    const uint64_t extra = len / FGETS_MAX + 1;   // bound on the further chunks of long lines
    GA_TRY(ga_malloc(ctx, &pb.p, std::max<uint64_t>(lines + extra, 1) * 8));
    if (lines) {
        k_parse_lines<<<grid_for(lines, 256, 1u << 16), 256, 0, ctx->stream>>>((const uint8_t*)db.p, len, (const unsigned long long*)nlb.p, nnl,
                                                                               lines, (uint2*)pb.p, cursor);
        GA_TRY(ga_launch_check(ctx, "k_parse_lines"));
    }
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, cursor, 8, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint64_t npairs = lines + ctx->hstat[0];
    if (npairs > lines + extra) return ga_fail(ctx, GA_ERR_STATE, "chunk count out of range");
    ga_free(db.p);   // the text is no longer needed
    db.p = nullptr;
    ga_graph* g = nullptr;
    GA_TRY(ga_build_from_device_pairs(ctx, (const uint32_t*)pb.p, npairs, -1, false, &g, nullptr));
    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
        ga_graph_destroy(g);
        return ga_fail(ctx, GA_ERR_HIP, "event record failed");
    }
    if (ms) (void)hipEventElapsedTime(ms, e0, e1);
    *out = g;
    return GA_OK;
}

extern "C" int ga_graph_from_text(ga_ctx* ctx, const char* bytes, uint64_t len, ga_graph** out, float* ms) {
    if (!out) return ctx ? ga_fail(ctx, GA_ERR_ARG, "out is NULL") : GA_ERR_ARG;
    *out = nullptr;
    if (!ctx) return GA_ERR_ARG;
    if (len && !bytes) return ga_fail(ctx, GA_ERR_ARG, "bytes is NULL");
    GA_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf db;
    GA_TRY(text_buffer(ctx, len, db));
    if (len) GA_HIP(ctx, hipMemcpyAsync(db.p, bytes, len, hipMemcpyHostToDevice, ctx->stream));
    return text_to_graph(ctx, db, len, len && bytes[len - 1] != '\n', out, ms);
}

// The file is streamed to the device (stream_file): the host never holds the whole text, and reading overlaps the
// copies.
extern "C" int ga_graph_load_text(ga_ctx* ctx, const char* path, ga_graph** out, float* ms) {
    if (!out) return ctx ? ga_fail(ctx, GA_ERR_ARG, "out is NULL") : GA_ERR_ARG;
    *out = nullptr;
    if (!ctx || !path) return GA_ERR_ARG;
    FILE* f = fopen(path, "rb");
    if (!f) return ga_fail(ctx, GA_ERR_ARG, "cannot open %s", path);
    struct File {
        FILE* f;
        ~File() { fclose(f); }
    } file{f};
    if (fseek(f, 0, SEEK_END) != 0) return ga_fail(ctx, GA_ERR_ARG, "cannot size %s", path);
    const long sz = ftell(f);
    if (sz < 0 || fseek(f, 0, SEEK_SET) != 0) return ga_fail(ctx, GA_ERR_ARG, "cannot size %s", path);
    const uint64_t len = (uint64_t)sz;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf db;
    GA_TRY(text_buffer(ctx, len, db));
    char last = '\n';
    GA_TRY(stream_file(ctx, f, path, len, (uint8_t*)db.p, len, nullptr, nullptr, 0, &last));
    return text_to_graph(ctx, db, len, len && last != '\n', out, ms);
}

extern "C" int ga_graph_save(ga_ctx* ctx, const ga_graph* g, const char* path) {
    if (!ctx || !g || !path) return GA_ERR_ARG;
    FileHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, MAGIC, 8);
    h.n = g->n;
    h.m = g->m;
    h.max_out = g->max_out;
    std::vector<uint32_t> rowptr((size_t)g->n + 1), col(g->m);
    GA_TRY(ga_graph_download(ctx, g, rowptr.data(), g->m ? col.data() : nullptr, nullptr));
    FILE* f = fopen(path, "wb");
    if (!f) return ga_fail(ctx, GA_ERR_ARG, "cannot create %s", path);
    bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(rowptr.data(), 4, rowptr.size(), f) == rowptr.size() &&
              (col.empty() || fwrite(col.data(), 4, col.size(), f) == col.size());
    ok = (fclose(f) == 0) && ok;
    return ok ? GA_OK : ga_fail(ctx, GA_ERR_ARG, "write to %s failed", path);
}

extern "C" int ga_graph_load(ga_ctx* ctx, const char* path, ga_graph** out, float* ms) {
    if (!out) return ctx ? ga_fail(ctx, GA_ERR_ARG, "out is NULL") : GA_ERR_ARG;
    *out = nullptr;
    if (!ctx || !path) return GA_ERR_ARG;
    FILE* f = fopen(path, "rb");
    if (!f) return ga_fail(ctx, GA_ERR_ARG, "cannot open %s", path);
    FileHeader h;
    if (fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, MAGIC, 8) != 0) {
        fclose(f);
        return ga_fail(ctx, GA_ERR_ARG, "%s is not a GACSR001 file", path);
    }
    fseek(f, 0, SEEK_END);
    const long sz = ftell(f);
    if (h.n > UINT32_MAX || h.m > UINT32_MAX || h.max_out > UINT32_MAX || sz < 0 ||
        (uint64_t)sz != sizeof h + 4 * (h.n + 1) + 4 * h.m) {
        fclose(f);
        return ga_fail(ctx, GA_ERR_ARG, "%s: header sizes do not match the file (n %llu, m %llu, %ld bytes)", path,
                       (unsigned long long)h.n, (unsigned long long)h.m, sz);
    }
    const uint32_t n = (uint32_t)h.n, m = (uint32_t)h.m;
    GA_HIP(ctx, hipSetDevice(ctx->device));
    ga_graph* g = new ga_graph;
    g->ctx = ctx;
    g->n = n;
    g->m = m;
    auto fail_g = [&](int rc) { ga_graph_destroy(g); return rc; };
    auto fail_gf = [&](int rc) { fclose(f); return fail_g(rc); };
    DevBuf degb, outb, badb;
    if (ga_malloc(ctx, (void**)&g->rowptr, ((size_t)n + 1) * 4) || ga_malloc(ctx, (void**)&g->col, std::max<size_t>(m, 1) * 4) ||
        ga_malloc(ctx, (void**)&g->src, std::max<size_t>(m, 1) * 4) || ga_malloc(ctx, &degb.p, ((size_t)n + 1) * 4) ||
        ga_malloc(ctx, &outb.p, ((size_t)n + 1) * 4) || ga_malloc(ctx, &badb.p, 64))
        return fail_gf(GA_ERR_OOM);
    if (fseek(f, sizeof h, SEEK_SET) != 0) return fail_gf(ga_fail(ctx, GA_ERR_ARG, "cannot read %s", path));
    uint32_t r0 = 1, rn = 0;
    Capture cap[2] = {{0, &r0}, {4ull * n, &rn}};
    int src = stream_file(ctx, f, path, 4ull * (n + 1) + 4ull * m, (uint8_t*)g->rowptr, 4ull * (n + 1), (uint8_t*)g->col, cap, 2, nullptr);
    fclose(f);
    if (src) return fail_g(src);
    if (r0 != 0 || rn != m) return fail_g(ga_fail(ctx, GA_ERR_ARG, "%s: rowptr must start at 0 and end at m", path));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) return fail_g(ga_fail(ctx, GA_ERR_HIP, "hipEventCreate failed"));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail_g(ga_fail(ctx, GA_ERR_HIP, "hipEventCreate failed")); }
    struct Ev {
        hipEvent_t a, b;
        ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};
    uint32_t* deg = (uint32_t*)degb.p;
    uint32_t* bad = (uint32_t*)badb.p;
    int rc = GA_OK;
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) return fail_g(ga_fail(ctx, GA_ERR_HIP, "event record failed"));
    if ((rc = ga_fill_u32(ctx, deg, (size_t)n + 1, 0)) || (rc = ga_fill_u32(ctx, bad, 2, 0)) || (rc = ga_fill_u32(ctx, (uint32_t*)outb.p, (size_t)n + 1, 0)))
        return fail_g(rc);
    k_csr_check_rows<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, g->col, n, m, deg, bad);
    if ((rc = ga_launch_check(ctx, "k_csr_check_rows"))) return fail_g(rc);
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->hstat[0] & 0xFFFFFFFFu)
        return fail_g(ga_fail(ctx, GA_ERR_ARG, "%s: rowptr decreases, a col is >= n or a row is not strictly ascending", path));
    k_csr_check_orient<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, g->col, n, deg, bad + 1);
    if ((rc = ga_launch_check(ctx, "k_csr_check_orient"))) return fail_g(rc);
    k_rows_src<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(g->rowptr, n, g->src, (uint32_t*)outb.p);
    if ((rc = ga_launch_check(ctx, "k_rows_src"))) return fail_g(rc);
    if ((rc = ga_max_u32(ctx, (const uint32_t*)outb.p, n, &g->max_out))) return fail_g(rc);
    GA_HIP(ctx, hipMemcpyAsync(ctx->hstat, bad + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    GA_HIP(ctx, hipEventRecord(e1, ctx->stream));
    GA_HIP(ctx, hipEventSynchronize(e1));
    if (ctx->hstat[0] & 0xFFFFFFFFu)
        return fail_g(ga_fail(ctx, GA_ERR_ARG, "%s: %u edges are not oriented by the (degree, id) rule", path, (uint32_t)(ctx->hstat[0] & 0xFFFFFFFFu)));
    if (g->max_out != h.max_out)
        return fail_g(ga_fail(ctx, GA_ERR_ARG, "%s: max_out_degree %llu in the header, %u in the rows", path, (unsigned long long)h.max_out, g->max_out));
    if (ms) (void)hipEventElapsedTime(ms, e0, e1);
    *out = g;
    return GA_OK;
}
