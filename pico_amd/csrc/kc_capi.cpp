// kc_capi.cpp -- implementation of the C ABI in include/kcore.h: host graph
// objects, device contexts and graphs, and dispatch to the gfx950 engines.
#include "kc_capi_impl.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace kci {

thread_local std::string g_err;

kc_status fail(kc_status s, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

kc_status hip_fail(hipError_t e, const char *what)
{
    return fail(e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EDEVICE, "%s: %s", what,
                hipGetErrorString(e));
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

kc_status ensure(kc_graph *g, DevBuf &b, size_t bytes, const char *what)
{
    if (b.n >= bytes && b.p)
        return KC_OK;
    if (b.p) {
        hipFree(b.p);
        g->bytes -= b.n;
        b.p = nullptr;
        b.n = 0;
    }
    size_t n = bytes ? bytes : 8;
    hipError_t e = hipMalloc(&b.p, n);
    if (e != hipSuccess)
        return hip_fail(e, what);
    b.n = n;
    g->bytes += n;
    return KC_OK;
}

void ctx_ref(kc_ctx *c)
{
    if (c)
        c->refs.fetch_add(1, std::memory_order_relaxed);
}

void ctx_unref(kc_ctx *c)
{
    if (!c || c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1)
        return;
    hipSetDevice(c->device);
    if (c->ev0)
        hipEventDestroy(c->ev0);
    if (c->ev1)
        hipEventDestroy(c->ev1);
    if (c->stream)
        hipStreamDestroy(c->stream);
    delete c;
}

} // namespace kci

using namespace kci;

extern "C" {

const char *kc_last_error(void) { return g_err.c_str(); }
const char *kc_version(void) { return "kcore-mi355x 0.1 (gfx950)"; }

// ------------------------------------------------------------ host graphs
// Run a host reader: status from its return code, and no C++ exception
// (an edge file whose ids ask for more memory than the host has) leaves
// through the C ABI.
static kc_status run_reader(int reader, const char *path, kc_host_graph &g)
{
    std::string err;
    int rc = 0;
    try {
        rc = reader == KC_READER_COO ? kch::read_coo(path, g, err) : kch::read_dedup(path, g, err);
    } catch (const std::bad_alloc &) {
        g = kc_host_graph{};
        return fail(KC_ENOMEM, "host allocation for the graph of %s failed", path);
    } catch (const std::exception &x) {
        g = kc_host_graph{};
        return fail(KC_EINVAL, "%s: %s", path, x.what());
    }
    if (rc == -1)
        return fail(KC_EIO, "%s", err.c_str());
    if (rc)
        return fail(KC_EINVAL, "%s: %s", path, err.c_str());
    return KC_OK;
}

kc_status kc_hg_read_coo(const char *path, kc_host_graph **out)
{
    if (!path || !out)
        return fail(KC_EINVAL, "null argument");
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    kc_status s = run_reader(KC_READER_COO, path, *g);
    if (s) {
        delete g;
        return s;
    }
    *out = g;
    return KC_OK;
}

kc_status kc_hg_save(const kc_host_graph *g, const char *path)
{
    if (!g || !path)
        return fail(KC_EINVAL, "null argument");
    std::string err;
    if (kch::save_csr(*g, path, kch::KC_NO_READER, 0, 0, err))
        return fail(KC_EIO, "%s", err.c_str());
    return KC_OK;
}

kc_status kc_hg_load(const char *path, kc_host_graph **out)
{
    if (!path || !out)
        return fail(KC_EINVAL, "null argument");
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    std::string err;
    if (kch::load_csr(path, *g, nullptr, nullptr, nullptr, err)) {
        delete g;
        return fail(KC_EIO, "%s", err.c_str());
    }
    *out = g;
    return KC_OK;
}

kc_status kc_hg_read_cached(const char *path, int reader, const char *cache_dir,
                            kc_host_graph **out, int *hit)
{
    if (!path || !out || !cache_dir)
        return fail(KC_EINVAL, "null argument");
    if (reader != KC_READER_COO && reader != KC_READER_DEDUP)
        return fail(KC_EINVAL, "unknown reader %d", reader);
    if (hit)
        *hit = 0;
    std::string err;
    uint64_t hash = 0, size = 0;
    if (kch::file_hash(path, hash, size, err))
        return fail(KC_EIO, "%s", err.c_str());
    char name[64];
    std::snprintf(name, sizeof name, "/kcsr-%016llx-%d.bin", (unsigned long long)hash, reader);
    std::string entry = std::string(cache_dir) + name;
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    uint32_t r = 0;
    uint64_t h2 = 0, s2 = 0;
    if (kch::load_csr(entry.c_str(), *g, &r, &h2, &s2, err) == 0 && r == (uint32_t)reader &&
        h2 == hash && s2 == size) {
        if (hit)
            *hit = 1;
        *out = g;
        return KC_OK;
    }
    // miss (absent, stale or damaged entry): parse, then (re)write the entry
    *g = kc_host_graph{};
    kc_status rs = run_reader(reader, path, *g);
    if (rs) {
        delete g;
        return rs;
    }
    if (kch::save_csr(*g, entry.c_str(), (uint32_t)reader, hash, size, err)) {
        delete g;
        return fail(KC_EIO, "%s", err.c_str());
    }
    *out = g;
    return KC_OK;
}

kc_status kc_hg_read_dedup(const char *path, kc_host_graph **out)
{
    if (!path || !out)
        return fail(KC_EINVAL, "null argument");
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    kc_status s = run_reader(KC_READER_DEDUP, path, *g);
    if (s) {
        delete g;
        return s;
    }
    *out = g;
    return KC_OK;
}

kc_status kc_hg_from_pairs(uint32_t V, uint64_t n, const uint32_t *src, const uint32_t *dst,
                           kc_host_graph **out)
{
    if (!out || (n && (!src || !dst)))
        return fail(KC_EINVAL, "null argument");
    for (uint64_t i = 0; i < n; i++)
        if (src[i] >= V || dst[i] >= V)
            return fail(KC_EINVAL, "pair %llu has an id >= V=%u", (unsigned long long)i, V);
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    kch::csr_from_pairs(V, n, src, dst, *g);
    *out = g;
    return KC_OK;
}

kc_status kc_hg_from_csr(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                         kc_host_graph **out)
{
    if (!out || !rowptr || (rowptr[V] && !col))
        return fail(KC_EINVAL, "null argument");
    if (rowptr[0] != 0)
        return fail(KC_EINVAL, "rowptr[0] != 0");
    for (uint32_t v = 0; v < V; v++)
        if (rowptr[v + 1] < rowptr[v])
            return fail(KC_EINVAL, "rowptr decreases at %u", v);
    for (uint64_t e = 0; e < rowptr[V]; e++)
        if (col[e] >= V)
            return fail(KC_EINVAL, "col[%llu]=%u >= V", (unsigned long long)e, col[e]);
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    g->V = V;
    g->rowptr.assign(rowptr, rowptr + V + 1);
    g->col.assign(col, col + rowptr[V]);
    *out = g;
    return KC_OK;
}

kc_status kc_hg_gen_er(uint32_t n, uint64_t m, uint64_t seed, kc_host_graph **out)
{
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g || !out)
        return fail(KC_EINVAL, "bad argument");
    if (kch::gen_er(n, m, seed, *g)) {
        delete g;
        return fail(KC_EINVAL, "G(n=%u, m=%llu) is not realisable", n, (unsigned long long)m);
    }
    *out = g;
    return KC_OK;
}

kc_status kc_hg_gen_rmat(uint32_t scale, uint64_t m_gen, double a, double b, double c,
                         uint64_t seed, kc_host_graph **out)
{
    if (!out || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(KC_EINVAL, "bad R-MAT parameters");
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    if (kch::gen_rmat(scale, m_gen, a, b, c, seed, *g)) {
        delete g;
        return fail(KC_EINVAL, "R-MAT scale %u out of range", scale);
    }
    *out = g;
    return KC_OK;
}

kc_status kc_hg_gen_chunglu(uint32_t n, double beta, double mean_deg, double cap, uint64_t seed,
                            kc_host_graph **out)
{
    if (!out)
        return fail(KC_EINVAL, "null argument");
    auto *g = new (std::nothrow) kc_host_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    if (kch::gen_chunglu(n, beta, mean_deg, cap, seed, *g)) {
        delete g;
        return fail(KC_EINVAL, "bad Chung-Lu parameters");
    }
    *out = g;
    return KC_OK;
}

kc_status kc_hg_info(const kc_host_graph *g, kc_graph_info *info)
{
    if (!g || !info)
        return fail(KC_EINVAL, "null argument");
    info->V = g->V;
    info->nnz = g->rowptr.empty() ? 0 : g->rowptr[g->V];
    info->edges_read = g->edges_read;
    info->has_zero_id = g->has_zero_id;
    info->multigraph = g->multigraph;
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < g->V; v++)
        dmax = std::max(dmax, (uint32_t)(g->rowptr[v + 1] - g->rowptr[v]));
    info->dmax = dmax;
    info->rowptr = g->rowptr.data();
    info->col = g->col.data();
    return KC_OK;
}

kc_status kc_hg_vgraph(const kc_host_graph *g, uint32_t part_size, uint32_t *node_pointer,
                       uint32_t *edge_list, uint32_t *parts, uint64_t *num_parts)
{
    if (!g || part_size == 0)
        return fail(KC_EINVAL, "bad argument");
    uint64_t P = 0;
    for (uint32_t v = 0; v < g->V; v++) {
        uint64_t d = g->rowptr[v + 1] - g->rowptr[v];
        uint64_t np = (d + part_size - 1) / part_size;
        if (node_pointer)
            node_pointer[v] = (uint32_t)(g->rowptr[v] + v);
        if (edge_list) {
            uint64_t b = g->rowptr[v] + v;
            edge_list[b] = (uint32_t)d;
            if (d)
                std::memcpy(edge_list + b + 1, g->col.data() + g->rowptr[v],
                            d * sizeof(uint32_t));
        }
        if (parts)
            for (uint64_t j = 0; j < np; j++) {
                parts[2 * (P + j)] = v;
                parts[2 * (P + j) + 1] = (uint32_t)j;
            }
        P += np;
    }
    if (num_parts)
        *num_parts = P;
    return KC_OK;
}

kc_status kc_hg_write_edgelist(const kc_host_graph *g, const char *path)
{
    if (!g || !path)
        return fail(KC_EINVAL, "null argument");
    std::string err;
    if (kch::write_edgelist(*g, path, err))
        return fail(KC_EIO, "%s", err.c_str());
    return KC_OK;
}

void kc_hg_free(kc_host_graph *g) { delete g; }

kc_status kc_write_cores(const char *path, const uint32_t *core, uint32_t V)
{
    if (!path || (V && !core))
        return fail(KC_EINVAL, "null argument");
    FILE *f = std::fopen(path, "wb");
    if (!f)
        return fail(KC_EIO, "cannot create %s", path);
    for (uint32_t i = 0; i < V; i++)
        std::fprintf(f, "%u %u\n", i, core[i]);
    std::fclose(f);
    return KC_OK;
}

// ---------------------------------------------------------------- device
kc_status kc_ctx_create(int device, kc_ctx **out)
{
    if (!out)
        return fail(KC_EINVAL, "null argument");
    int n = 0;
    KC_HIP(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n)
        return fail(KC_EINVAL, "device %d not present (%d visible)", device, n);
    KC_HIP(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    KC_HIP(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    auto *c = new (std::nothrow) kc_ctx;
    if (!c)
        return fail(KC_ENOMEM, "host allocation");
    c->device = device;
    c->num_cus = (uint32_t)prop.multiProcessorCount;
    c->hbm = prop.totalGlobalMem;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        c->clock_khz = (uint64_t)khz;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&c->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        if (c->ev0)
            hipEventDestroy(c->ev0);
        if (c->stream)
            hipStreamDestroy(c->stream);
        delete c;
        return hip_fail(e, "context stream / events");
    }
    // every kernel loaded now, not by the first decomposition that needs it
    // (the degree-ordered adjacency's sort alone: ~19 ms once per process)
    if ((e = kc::load_modules(c->stream)) != hipSuccess) {
        hipEventDestroy(c->ev1);
        hipEventDestroy(c->ev0);
        hipStreamDestroy(c->stream);
        delete c;
        return hip_fail(e, "kernel load");
    }
    // the no-allocation entry point's scratch array, located once per device
    // (a failure is reported by that entry point, not here)
    noalloc_setup(device);
    *out = c;
    return KC_OK;
}

void kc_ctx_destroy(kc_ctx *ctx) { kci::ctx_unref(ctx); }

kc_status kc_ctx_info(const kc_ctx *ctx, uint32_t *num_cus, uint64_t *hbm_bytes)
{
    if (!ctx)
        return fail(KC_EINVAL, "null context");
    if (num_cus)
        *num_cus = ctx->num_cus;
    if (hbm_bytes)
        *hbm_bytes = ctx->hbm;
    return KC_OK;
}

kc_status kc_graph_upload(kc_ctx *ctx, uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                          uint32_t rowptr_bytes, kc_graph **out)
{
    if (!ctx || !rowptr || !out)
        return fail(KC_EINVAL, "null argument");
    if (rowptr_bytes != 4 && rowptr_bytes != 8)
        return fail(KC_EINVAL, "rowptr_bytes must be 4 or 8");
    if (V == 0xffffffffu)
        return fail(KC_EINVAL, "V too large");
    uint64_t nnz = rowptr[V];
    if (rowptr[0] != 0)
        return fail(KC_EINVAL, "rowptr[0] != 0");
    if (nnz && !col)
        return fail(KC_EINVAL, "null col");
    for (uint32_t v = 0; v < V; v++)
        if (rowptr[v + 1] < rowptr[v])
            return fail(KC_EINVAL, "rowptr decreases at vertex %u", v);
    for (uint64_t e = 0; e < nnz; e++)
        if (col[e] >= V)
            return fail(KC_EINVAL, "col[%llu]=%u out of range (V=%u)", (unsigned long long)e,
                        col[e], V);
    bool wide = rowptr_bytes == 8 || nnz > 0xffffffffull;
    KC_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto *g = new (std::nothrow) kc_graph;
    if (!g)
        return fail(KC_ENOMEM, "host allocation");
    g->ctx = ctx;
    ctx_ref(ctx);
    g->V = V;
    g->nnz = nnz;
    g->wide = wide;
    size_t rb = (size_t)(V + 1ull) * (wide ? 8 : 4);
    hipError_t e = hipMalloc(&g->rowptr, rb);
    if (e == hipSuccess)
        e = hipMalloc((void **)&g->col, nnz ? nnz * 4 : 4);
    if (e != hipSuccess) {
        kc_graph_free(g);
        return hip_fail(e, "graph allocation");
    }
    g->bytes = rb + (nnz ? nnz * 4 : 4);
    if (wide) {
        e = hipMemcpyAsync(g->rowptr, rowptr, rb, hipMemcpyHostToDevice, ctx->stream);
    } else {
        std::vector<uint32_t> r32(V + 1ull);
        for (uint64_t v = 0; v <= V; v++)
            r32[v] = (uint32_t)rowptr[v];
        e = hipMemcpyAsync(g->rowptr, r32.data(), rb, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
    }
    if (e == hipSuccess && nnz)
        e = hipMemcpyAsync(g->col, col, nnz * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
        e = hipEventCreate(&g->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&g->ev1);
    if (e != hipSuccess) {
        kc_graph_free(g);
        return hip_fail(e, "graph upload");
    }
    *out = g;
    return KC_OK;
}

kc_status kc_graph_upload_hg(kc_ctx *ctx, const kc_host_graph *hg, uint32_t rowptr_bytes,
                             kc_graph **out)
{
    if (!hg)
        return fail(KC_EINVAL, "null graph");
    return kc_graph_upload(ctx, hg->V, hg->rowptr.data(), hg->col.data(), rowptr_bytes, out);
}

void kc_graph_free(kc_graph *g)
{
    if (!g)
        return;
    if (g->ctx)
        hipSetDevice(g->ctx->device);
    if (g->owns_csr) {
        hipFree(g->rowptr);
        hipFree(g->col);
    }
    for (DevBuf *b : {&g->deg, &g->out, &g->done, &g->list0, &g->list1, &g->gs, &g->stats,
                      &g->histo, &g->cnt, &g->flags, &g->hl0, &g->hl1, &g->hl2, &g->kmax,
                      &g->trace, &g->col_ds, &g->nd_ds, &g->hb_ds, &g->hubs, &g->hchg_bm,
                      &g->hchg_rec, &g->hprep, &g->cs_row[0],
                      &g->cs_row[1],
                      &g->cs_col[0], &g->cs_col[1], &g->cs_deg[0], &g->cs_deg[1], &g->cs_done[0],
                      &g->cs_done[1], &g->cs_gmap[0], &g->cs_gmap[1], &g->cs_loc, &g->cs_wpre,
                      &g->cs_tmp, &g->cs_bad, &g->aux})
        if (b->p)
            hipFree(b->p);
    if (g->ev0)
        hipEventDestroy(g->ev0);
    if (g->ev1)
        hipEventDestroy(g->ev1);
    dyn_free(g->dyn);
    kc_ctx *c = g->ctx;
    delete g;
    ctx_unref(c);
}

uint64_t kc_graph_device_bytes(const kc_graph *g) { return g ? g->bytes : 0; }

kc_status kc_graph_histogram(const kc_graph *g, uint32_t *out, uint64_t n)
{
    if (!g || !out)
        return fail(KC_EINVAL, "null argument");
    if (!g->histo.p || n > g->nnz + g->V)
        return fail(KC_EINVAL, "no histogram of %llu words", (unsigned long long)n);
    KC_HIP(hipMemcpy(out, g->histo.p, n * 4, hipMemcpyDeviceToHost), "histogram copy");
    return KC_OK;
}

uint32_t kc_graph_trace(const kc_graph *g, uint64_t *out, uint32_t max_phases,
                        uint64_t *ticks_per_ms)
{
    if (!g || !g->trace.p || !g->trace_phases)
        return 0;
    uint32_t n = std::min(max_phases, g->trace_phases);
    if (ticks_per_ms)
        *ticks_per_ms = g->ctx->clock_khz;
    if (out && n &&
        hipMemcpy(out, g->trace.p, (size_t)n * kc::KC_TRACE_WORDS * 8, hipMemcpyDeviceToHost) !=
            hipSuccess)
        return 0;
    return n;
}

const uint32_t *kc_graph_core_device(const kc_graph *g) { return g ? g->result : nullptr; }

} // extern "C"

// ----------------------------------------------------------- dispatching
// Persistent engines wait at grid barriers for every workgroup of their
// grid; two of them sharing a device from different streams could each hold
// CUs the other's unscheduled workgroups need.  Every persistent launch
// therefore reserves workgroup slots first, in arrival order, from a budget
// of one 1024-thread workgroup per CU: launches of 1024-thread groups reserve
// their grid width (their grids are then all resident together -- a CU holds
// at least one such group), every other launch reserves the whole device.
// Non-persistent kernels are not counted: they finish on their own.
namespace {
struct CoopState {
    std::mutex mu;
    std::condition_variable cv;
    uint64_t next_ticket = 0, serving = 0; // FIFO order of reservations
    uint32_t used = 0;
};

CoopState &coop_state(int device)
{
    static CoopState s[64];
    return s[(unsigned)device & 63];
}
} // namespace

kci::CoopGuard::CoopGuard(int dev, uint32_t need, uint32_t capacity) : device(dev)
{
    if (need == 0 || need > capacity)
        need = capacity;
    CoopState &s = coop_state(dev);
    std::unique_lock<std::mutex> lk(s.mu);
    const uint64_t t = s.next_ticket++;
    s.cv.wait(lk, [&] { return s.serving == t && s.used + need <= capacity; });
    s.serving++;
    s.used += need;
    n = need;
    s.cv.notify_all(); // the next reservation may fit beside this one
}

kci::CoopGuard::~CoopGuard()
{
    CoopState &s = coop_state(device);
    std::lock_guard<std::mutex> lk(s.mu);
    s.used -= n;
    s.cv.notify_all();
}

// Largest value of a[0..n): one atomic per workgroup (256 threads), not per
// wave -- thousands of same-address atomics serialise into tens of us.
__global__ __launch_bounds__(256) void kc_max_kernel(const uint32_t *a, uint32_t n,
                                                     uint32_t *out)
{
    __shared__ uint32_t s_m[256 / 64];
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = a[i] > m ? a[i] : m;
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t y = __shfl_xor(m, o, 64);
        m = y > m ? y : m;
    }
    if ((threadIdx.x & 63) == 0)
        s_m[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / 64; w++)
            m = s_m[w] > m ? s_m[w] : m;
        if (m)
            atomicMax(out, m);
    }
}

// Default grid width for a graph of nnz adjacency entries: a smaller graph's
// phases have less work than a full grid of workgroups can share, and every
// grid barrier and shared counter costs more with more of them.  Measured on
// R-MAT scales 12-20 and the ER configuration, the best width grows about
// with sqrt(nnz): 32 up to ~0.5M entries, 64 at 1.8M, 128 at 7.6M, the full
// grid from ~30M on.
uint32_t kci::size_grid(uint64_t nnz)
{
    double g = 48.0 * std::sqrt((double)nnz / 1e6);
    return g < 32.0 ? 32u : g > 65536.0 ? 65536u : (uint32_t)g;
}

static kc_status run_peel(kc_graph *g, kc_method method, const kc_opts &o, kc_stats &st)
{
    kc_ctx *c = g->ctx;
    uint32_t block = o.block ? o.block : 1024;
    if (block != 256 && block != 512 && block != 1024)
        return fail(KC_EINVAL, "block must be 256, 512 or 1024");
    int occ = kc::peel_occupancy(block, g->wide);
    if (occ <= 0)
        return fail(KC_EDEVICE, "peel kernel cannot be resident with block %u", block);
    // Persistent grids stay well inside the occupancy limit: every workgroup
    // must be resident for the grid barrier, and the occupancy query does not
    // see every per-CU limit (a full 7-per-CU grid of 256-thread groups was
    // observed not to be co-resident).
    uint32_t per_cu = std::min<uint32_t>((uint32_t)occ, block == 1024 ? 1u : 2u);
    uint32_t maxgrid = per_cu * c->num_cus;
    // Default widths measured on R-MAT s22: the level-synchronous GPP and the
    // PeelOne sweeps are latency-bound and run best on half of the CUs; the
    // bucket peel's first stage uses them all.
    uint32_t defgrid = method == KC_GPP || method == KC_PEELONE ? std::max(1u, maxgrid / 2)
                                                                : maxgrid;
    defgrid = std::min(defgrid, size_grid(g->nnz));
    uint32_t grid = o.grid ? std::min(o.grid, maxgrid) : defgrid;
    // hubs' virtual nodes of 4,096 edges: 1,024 / 4,096 / 16,384 ran 58.9 / 58.0 /
    // 58.1 ms on the Chung-Lu 2^24 graph (PeelOne 64.6 / 63.6 / 63.6), no
    // change on R-MAT s22 or for GPP
    uint32_t hub = o.hub_split ? o.hub_split : 4096;
    if (hub < 64 || hub > (1u << 24))
        return fail(KC_EINVAL, "hub_split must be in [64, 2^24]");
    uint64_t cap64 = (uint64_t)g->V + g->nnz / hub + 2ull * grid + 64;
    if (o.queue_cap)
        cap64 = o.queue_cap;
    if (cap64 > 0xffffffffull)
        cap64 = 0xffffffffull;
    uint32_t cap = (uint32_t)cap64;
    kc_status s;
    if ((s = ensure(g, g->deg, (size_t)g->V * 4, "deg")) ||
        (s = ensure(g, g->done, (size_t)((g->V + 63) / 64) * 8, "claim bitmap")) ||
        (s = ensure(g, g->list0, (size_t)cap * 8, "work list")) ||
        (s = ensure(g, g->list1, (size_t)cap * 8, "work list")) ||
        (s = ensure(g, g->gs, kc::grid_sync_bytes(), "grid sync")) ||
        (s = ensure(g, g->stats, kc::KC_ST_COUNT * 8, "stats")) ||
        (s = ensure(g, g->kmax, 8, "kmax")))
        return s;
    uint32_t mode = method == KC_GPP ? 2 : method == KC_PEELONE ? 1 : 0;
    // GPP: final cores (the cascading modes keep theirs in deg)
    if (mode == 2 && (s = ensure(g, g->out, (size_t)(g->V ? g->V : 1) * 4, "vertex list")))
        return s;
    kc::PeelLaunch p{};
    p.rowptr = g->rowptr;
    p.col = g->col;
    p.V = g->V;
    p.hub_split = hub;
    p.local_cap = o.local_cap; // 0: 16 pending edges per thread
    p.deg = (uint32_t *)g->deg.p;
    p.done = (uint64_t *)g->done.p;
    p.list0 = (uint64_t *)g->list0.p;
    p.list1 = (uint64_t *)g->list1.p;
    p.list_cap = cap;
    p.mode = mode;
    p.out = (uint32_t *)g->out.p;
    if (mode == 2) { // GPP keeps its cores in out: its live-vertex list goes elsewhere
        if ((s = ensure(g, g->hl0, (size_t)(g->V ? g->V : 1) * 4, "live list")))
            return s;
        p.live = (uint32_t *)g->hl0.p;
    }
    // switch to the live-vertex list below a quarter of V (GPP: half, as its
    // every level starts with a full scan)
    p.live_thresh = mode == 2 ? g->V / 2 : g->V / 4;
    // the dense core: pre-check + aggregation once at most V/agg_div remain
    uint32_t agg_div = o.agg_div ? o.agg_div : 128;
    p.agg_live = agg_div == 0xffffffffu ? 0xffffffffu : g->V / agg_div;
    p.gs = (kc::GridSync *)g->gs.p;
    p.stats = (uint64_t *)g->stats.p;
    p.grid = grid;
    p.block = block;
    p.wide = g->wide;
    p.gpp_compat = o.gpp_compat != 0;
    p.count_atomics = (o.flags & KC_OPT_COUNT) != 0;
    p.timeout_ticks = c->clock_khz * 1000ull * 20ull;
    g->trace_phases = 0;
    if (o.flags & KC_OPT_TRACE) {
        if ((s = ensure(g, g->trace, (size_t)KC_TRACE_CAP * kc::KC_TRACE_WORDS * 8, "trace")))
            return s;
        KC_HIP(hipMemsetAsync(g->trace.p, 0, g->trace.n, c->stream), "trace clear");
        p.trace = (uint64_t *)g->trace.p;
        p.trace_cap = KC_TRACE_CAP;
    }
    // Bucket peels continue on compacted live subgraphs once few vertices
    // remain (kc_compact.hip); stage 0 is the graph itself.
    const uint32_t cdiv = o.compact_div ? o.compact_div : 8u;
    auto tail_of = [&](uint32_t Vs) -> uint32_t {
        if (mode == 2 || cdiv == 0xffffffffu || cdiv < 2) // not GPP
            return 0u;
        uint32_t t = Vs / cdiv;
        return t >= 4096u ? t : 0u;
    };
    p.tail_live = tail_of(g->V);
    p.start_level = 0;
    // the whole graph's walks sum their decrements in LDS when compacted
    // stages take over the dense levels (kc_peel.hip, walk_items)
    p.agg_walks = mode != 2 && p.tail_live ? 1u : 0u;
    // workgroups of the compacted stages' launches (see below)
    const uint32_t stage_g =
        std::max(1u, std::min(o.stage_grid ? o.stage_grid : maxgrid / 2, maxgrid));
    CoopGuard coop(c->device, block == 1024 ? std::max(grid, p.tail_live ? stage_g : 0u) : 0u,
                   c->num_cus);
    KC_HIP(hipEventRecord(g->ev0, c->stream), "hipEventRecord");
    hipError_t e = kc::launch_peel(p, c->stream);
    if (e != hipSuccess)
        return hip_fail(e, "peel launch");
    uint64_t sv[kc::KC_ST_COUNT];
    uint64_t acc_levels = 0, acc_phases = 0, acc_edges = 0, acc_atomics = 0, acc_iters = 0;
    bool compacted = false;
    uint32_t nstages = 1;
    uint64_t stage_nnz = g->nnz; // bound on the current stage's adjacency
    for (int stage = 0;; stage++) {
        KC_HIP(hipMemcpyAsync(sv, g->stats.p, sizeof sv, hipMemcpyDeviceToHost, c->stream),
               "stats copy");
        KC_HIP(hipStreamSynchronize(c->stream), "peel stage");
        if (sv[kc::KC_ST_ERR] == 1)
            return fail(KC_ETIMEOUT, "device wait timed out");
        if (sv[kc::KC_ST_ERR] == 2)
            return fail(KC_EOVERFLOW, "shared work list overflow (capacity %u items)",
                        p.list_cap);
        if (sv[kc::KC_ST_ERR] == 4) {
            uint32_t bad = 0;
            KC_HIP(hipMemcpy(&bad, g->cs_bad.p, 4, hipMemcpyDeviceToHost), "compaction copy");
            return fail(KC_EDEVICE, "live-subgraph compaction failed its checks (flags %u)", bad);
        }
        if (sv[kc::KC_ST_ERR])
            return fail(KC_EDEVICE, "device consistency check failed (%llu)",
                        (unsigned long long)sv[kc::KC_ST_ERR]);
        acc_levels += sv[kc::KC_ST_LEVELS_RUN];
        acc_phases += sv[kc::KC_ST_PHASES];
        acc_edges += sv[kc::KC_ST_EDGES];
        acc_atomics += sv[kc::KC_ST_ATOMICS];
        acc_iters += sv[kc::KC_ST_ITERS];
        const int cur = stage & 1, nxt = (stage + 1) & 1;
        if (stage > 0)
            KC_HIP(kc::scatter_cores(p.deg, (const uint32_t *)g->cs_gmap[cur].p, p.V,
                                     (uint32_t *)g->deg.p, c->stream),
                   "core scatter");
        if (p.tail_live == 0 || sv[kc::KC_ST_TAIL_LEVEL] == 0xffffffffull)
            break;
        // compact the live subgraph of this stage into the other stage set
        const uint32_t Vn = (uint32_t)sv[kc::KC_ST_LIVE];
        const uint32_t nw = (p.V + 63) / 64;
        const bool wide_n = g->nnz >= (1ull << 32);
        const size_t rb = wide_n ? 8 : 4;
        const size_t tb = kc::compact_tmp_bytes(p.V, Vn);
        if ((s = ensure(g, g->cs_wpre, (size_t)(nw + 1) * 4, "compaction")) ||
            (s = ensure(g, g->cs_loc, (size_t)Vn * 4 + 4, "compaction")) ||
            (s = ensure(g, g->cs_gmap[nxt], (size_t)Vn * 4 + 4, "compaction")) ||
            (s = ensure(g, g->cs_row[nxt], (size_t)(Vn + 1) * rb, "compaction")) ||
            (s = ensure(g, g->cs_tmp, tb, "compaction")) ||
            (s = ensure(g, g->cs_bad, 8, "compaction")))
            return s;
        // flags (kept across stages; checked after each) and the long-row count
        KC_HIP(hipMemsetAsync((char *)g->cs_bad.p + (compacted ? 4 : 0), 0, compacted ? 4 : 8,
                              c->stream),
               "memset");
        compacted = true;
        kc::CompactIn in{p.rowptr, p.wide, p.col, p.V, p.deg, p.done,
                         stage > 0 ? (const uint32_t *)g->cs_gmap[cur].p : nullptr};
        kc::CompactOut co{};
        co.V_in = p.V;
        co.Vn = Vn;
        co.wpre = (uint32_t *)g->cs_wpre.p;
        co.loc = (uint32_t *)g->cs_loc.p;
        co.gmap = (uint32_t *)g->cs_gmap[nxt].p;
        co.rowptr = g->cs_row[nxt].p;
        co.wide = wide_n;
        co.bad = (uint32_t *)g->cs_bad.p;
        co.tmp = g->cs_tmp.p;
        co.tmp_bytes = g->cs_tmp.n;
        // No host round trip: the adjacency gets an upper bound of room, the
        // copy kernels check the renumbering and every row against it, and
        // the next stage leaves at once (error 4) if any check failed.
        const uint64_t nnz_cap = std::min<uint64_t>(stage_nnz, (uint64_t)Vn * Vn);
        if ((s = ensure(g, g->cs_col[nxt], (size_t)nnz_cap * 4 + 4, "compaction")) ||
            (s = ensure(g, g->cs_deg[nxt], (size_t)Vn * 4 + 4, "compaction")) ||
            (s = ensure(g, g->cs_done[nxt], (size_t)((Vn + 63) / 64) * 8 + 8, "compaction")))
            return s;
        co.col = (uint32_t *)g->cs_col[nxt].p;
        co.col_cap = nnz_cap;
        KC_HIP(kc::compact_rows(in, co, c->stream), "compaction rows");
        KC_HIP(kc::compact_cols(in, co, c->stream), "compaction adjacency");
        const uint64_t nnz_n = nnz_cap;
        stage_nnz = nnz_cap;
        p.abort_flag = co.bad;
        // the next stage: same launch shape, the level where this one stopped
        p.rowptr = co.rowptr;
        p.wide = wide_n;
        p.col = co.col;
        p.V = Vn;
        p.deg = (uint32_t *)g->cs_deg[nxt].p;
        p.done = (uint64_t *)g->cs_done[nxt].p;
        uint64_t cap_n = (uint64_t)Vn + nnz_n / hub + 2ull * maxgrid + 64;
        p.list_cap = (uint32_t)std::min<uint64_t>(cap_n, cap);
        p.live_thresh = Vn / 4;
        p.agg_live = agg_div == 0xffffffffu ? 0xffffffffu : Vn / agg_div;
        p.tail_live = tail_of(Vn);
        p.start_level = (uint32_t)sv[kc::KC_ST_TAIL_LEVEL];
        p.agg_walks = 0;
        if (stage == 0) {
            // the compacted stages are latency-bound: fewer workgroups make
            // their barriers and shared-list polling cheaper (R-MAT s22 with
            // 1024-thread groups: 96 / 128 / 160 / 256 of 256 CUs take
            // 11.76 / 11.67 / 11.71 / 12.14 ms)
            p.grid = stage_g;
        }
        if (p.trace) {
            uint64_t used = std::min<uint64_t>(acc_phases, KC_TRACE_CAP);
            p.trace = (uint64_t *)g->trace.p + used * kc::KC_TRACE_WORDS;
            p.trace_cap = (uint32_t)(KC_TRACE_CAP - used);
        }
        if ((e = kc::launch_peel(p, c->stream)) != hipSuccess)
            return hip_fail(e, "peel launch");
        nstages++;
    }
    KC_HIP(hipEventRecord(g->ev1, c->stream), "hipEventRecord");
    KC_HIP(hipMemsetAsync(g->kmax.p, 0, 4, c->stream), "memset");
    g->result = mode == 2 ? (const uint32_t *)g->out.p : (const uint32_t *)g->deg.p;
    if (g->V)
        hipLaunchKernelGGL(kc_max_kernel, dim3(std::min<uint32_t>(512, (g->V + 255) / 256)),
                           dim3(256), 0, c->stream, g->result, g->V, (uint32_t *)g->kmax.p);
    KC_HIP(hipStreamSynchronize(c->stream), "decomposition");
    float ms = 0;
    hipEventElapsedTime(&ms, g->ev0, g->ev1);
    uint32_t km = 0;
    KC_HIP(hipMemcpy(&km, g->kmax.p, 4, hipMemcpyDeviceToHost), "kmax copy");
    st.device_ms = ms;
    st.kmax = km;
    st.levels = g->V ? (uint64_t)km + 1 : 0;
    st.levels_run = acc_levels;
    st.phases = acc_phases;
    st.edges_walked = acc_edges;
    st.atomics = acc_atomics;
    st.stages = nstages;
    st.grid = grid;
    st.block = block;
    if (o.flags & KC_OPT_TRACE)
        g->trace_phases = (uint32_t)std::min<uint64_t>(st.phases, KC_TRACE_CAP);
    if (method == KC_PEEL_DYN || method == KC_PEELONE_DYN)
        st.iterations = st.levels; // peel-dyn.cc:155-173: one trip per level
    else
        // GPP runs in one launch.  PeelOne's stages add up; each launch
        // counts the sweep that ends the host loop, reached only by the last
        st.iterations = acc_iters - (nstages - 1);
    return KC_OK;
}

static kc_status run_hindex(kc_graph *g, kc_method method, const kc_opts &o, kc_stats &st)
{
    kc_ctx *c = g->ctx;
    // HistoCore runs best with 512-thread groups, CntCore / NbrCore with 1,024
    // (R-MAT s22 and the Chung-Lu graph)
    uint32_t block = o.block ? o.block : method == KC_HISTO ? 512u : 1024u;
    if (block != 256 && block != 512 && block != 1024)
        return fail(KC_EINVAL, "block must be 256, 512 or 1024");
    const uint32_t mi = method == KC_HISTO ? 0 : method == KC_CNT ? 1 : 2;
    int occ = kc::hindex_occupancy(block, g->wide, mi);
    if (occ <= 0)
        return fail(KC_EDEVICE, "h-index kernel cannot be resident with block %u", block);
    uint32_t per_cu = std::min<uint32_t>((uint32_t)occ, block == 1024 ? 1u : 2u);
    uint32_t maxgrid = per_cu * c->num_cus;
    // HistoCore runs best with one workgroup per CU, CntCore / NbrCore with
    // as many as fit (R-MAT s22); smaller graphs with fewer (size_grid)
    uint32_t defgrid = method == KC_HISTO ? std::min(maxgrid, c->num_cus) : maxgrid;
    defgrid = std::min(defgrid, size_grid(g->nnz));
    uint32_t grid = o.grid ? std::min(o.grid, maxgrid) : defgrid;
    size_t V4 = (size_t)(g->V ? g->V : 1) * 4;
    kc_status s;
    if ((s = ensure(g, g->deg, V4, "core")) ||
        (s = ensure(g, g->histo, (size_t)(g->nnz + g->V + 1) * 4, "histogram")) ||
        (s = ensure(g, g->hl0, V4, "list")) || (s = ensure(g, g->hl1, V4, "list")) ||
        (s = ensure(g, g->gs, kc::grid_sync_bytes(), "grid sync")) ||
        (s = ensure(g, g->stats, kc::KC_ST_COUNT * 8, "stats")) ||
        (s = ensure(g, g->kmax, 8, "kmax")))
        return s;
    if (method != KC_HISTO &&
        ((s = ensure(g, g->out, V4, "new values")) || (s = ensure(g, g->flags, V4, "marks"))))
        return s;
    // CntCore / NbrCore: a third list; HistoCore: its list of vertices of more
    // than KC_TINY neighbours (the fourth list of the kernel's arguments)
    if ((s = ensure(g, g->hl2, V4, "list")))
        return s;
    if (method == KC_CNT &&
        ((s = ensure(g, g->cnt, V4, "cnt")) || (s = ensure(g, g->list0, V4, "list"))))
        return s;
    // HistoCore queues changed vertices as 64-edge virtual nodes
    uint64_t chunk_cap = (uint64_t)g->V + g->nnz / 64 + 64;
    if (chunk_cap > 0xffffffffull)
        chunk_cap = 0xffffffffull;
    if (method == KC_HISTO && (s = ensure(g, g->list1, (size_t)chunk_cap * 8, "virtual nodes")))
        return s;
    // degree-ordered adjacency, built once: it bounds HistoCore's update walks
    // and CntCore's / NbrCore's marks and recounts, and gives every method its
    // first round's values in row order
    bool use_ds = !(o.flags & KC_OPT_NO_DEGSORT) && g->nnz && g->nnz <= 0x7fffffffull;
    if (use_ds && !g->degsorted) {
        size_t nb = (size_t)g->nnz * 4;
        if ((s = ensure(g, g->col_ds, nb, "ordered adjacency")) ||
            (s = ensure(g, g->nd_ds, nb, "neighbour degrees")) ||
            (g->nnz + g->V < 0xffffffffull &&
             (s = ensure(g, g->hb_ds, nb, "neighbour histogram bases"))))
            return s;
        double tp = now_ms();
        // the histogram buffer (nnz + V + 1 words) is free scratch before the
        // run, and so is the neighbour-base array, filled after the sort
        hipError_t pe = kc::build_degree_sorted(g->rowptr, g->wide, g->col, g->V, g->nnz,
                                                (uint32_t *)g->col_ds.p, (uint32_t *)g->nd_ds.p,
                                                (uint32_t *)g->histo.p,
                                                (uint32_t *)g->hb_ds.p, c->stream);
        if (pe == hipSuccess && g->nnz + g->V < 0xffffffffull)
            pe = kc::build_nbr_hbase(g->rowptr, g->wide, (const uint32_t *)g->col_ds.p, g->nnz,
                                     (uint32_t *)g->hb_ds.p, c->stream);
        if (pe == hipSuccess)
            pe = hipStreamSynchronize(c->stream);
        if (pe != hipSuccess)
            return hip_fail(pe, "degree-ordered adjacency");
        g->degsorted = true;
        st.prep_ms = now_ms() - tp;
    }
    kc::HindexLaunch p{};
    p.rowptr = g->rowptr;
    p.col = g->col;
    p.V = g->V;
    p.method = mi;
    p.core = (uint32_t *)g->deg.p;
    p.histo = (uint32_t *)g->histo.p;
    p.cnt = (uint32_t *)g->cnt.p;
    p.newc = (uint32_t *)g->out.p;
    p.mark = (uint32_t *)g->flags.p;
    p.list0 = (uint32_t *)g->hl0.p;
    p.list1 = (uint32_t *)g->hl1.p;
    p.list2 = (uint32_t *)g->hl2.p;
    p.list3 = method == KC_HISTO ? (uint32_t *)g->hl2.p : (uint32_t *)g->list0.p;
    // every list the method writes must exist before the launch
    if (!p.list0 || !p.list1 || (method != KC_HISTO && !p.list2) ||
        (method != KC_NBR && !p.list3) || (method == KC_CNT && !p.cnt))
        return fail(KC_EDEVICE, "h-index work lists not allocated");
    p.chunks = (uint64_t *)g->list1.p;
    p.chunk_cap = (uint32_t)chunk_cap;
    p.col_ds = use_ds ? (const uint32_t *)g->col_ds.p : nullptr;
    p.nd_ds = use_ds ? (const uint32_t *)g->nd_ds.p : nullptr;
    p.hb_ds = use_ds && g->nnz + g->V < 0xffffffffull ? (const uint32_t *)g->hb_ds.p : nullptr;
    if (method == KC_HISTO) { // the rows above 4,096 neighbours (V entries), then the hubs
        if ((s = ensure(g, g->hubs, 2 * V4, "hub list")) ||
            (s = ensure(g, g->hprep, 64, "initial build")))
            return s;
        p.glist = (uint32_t *)g->hubs.p;
        p.hprep = g->hprep.p;
    }
    if (use_ds && method == KC_HISTO) { // hubs, for the pull rounds
        p.hubs = (uint32_t *)g->hubs.p + (g->V ? g->V : 1);
        p.hub_cap = g->V;
        size_t bmw = ((size_t)g->V + 63) / 64;
        if ((s = ensure(g, g->hchg_bm, 2 * (bmw ? bmw : 1) * 8, "change bitmaps")) ||
            (s = ensure(g, g->hchg_rec, (size_t)(g->V ? g->V : 1) * 8, "change records")))
            return s;
        p.chg_bm = (uint64_t *)g->hchg_bm.p;
        p.chg_rec = (uint64_t *)g->hchg_rec.p;
        p.pull_den = o.pull_div;
    }
    if (method == KC_NBR) { // NbrCore: its list of the largest rows
        if ((s = ensure(g, g->hubs, V4, "giant list")))
            return s;
        p.giants = (uint32_t *)g->hubs.p;
        p.giant_cap = g->V;
    }
    p.gs = (kc::GridSync *)g->gs.p;
    p.stats = (uint64_t *)g->stats.p;
    p.grid = grid;
    p.block = block;
    p.wide = g->wide;
    p.timeout_ticks = c->clock_khz * 1000ull * 20ull;
    g->trace_phases = 0;
    if (o.flags & KC_OPT_TRACE) {
        if ((s = ensure(g, g->trace, (size_t)KC_TRACE_CAP * kc::KC_TRACE_WORDS * 8, "trace")))
            return s;
        KC_HIP(hipMemsetAsync(g->trace.p, 0, g->trace.n, c->stream), "trace clear");
        p.trace = (uint64_t *)g->trace.p;
        p.trace_cap = KC_TRACE_CAP;
    }
    CoopGuard coop(c->device, 0u, c->num_cus); // the whole device
    KC_HIP(hipEventRecord(g->ev0, c->stream), "hipEventRecord");
    hipError_t e = kc::launch_hindex(p, c->stream);
    if (e != hipSuccess)
        return hip_fail(e, "h-index launch");
    KC_HIP(hipEventRecord(g->ev1, c->stream), "hipEventRecord");
    KC_HIP(hipMemsetAsync(g->kmax.p, 0, 4, c->stream), "memset");
    g->result = (const uint32_t *)g->deg.p;
    if (g->V)
        hipLaunchKernelGGL(kc_max_kernel, dim3(std::min<uint32_t>(512, (g->V + 255) / 256)),
                           dim3(256), 0, c->stream, g->result, g->V, (uint32_t *)g->kmax.p);
    KC_HIP(hipStreamSynchronize(c->stream), "decomposition");
    float ms = 0;
    hipEventElapsedTime(&ms, g->ev0, g->ev1);
    uint64_t sv[kc::KC_ST_COUNT];
    uint32_t km = 0;
    KC_HIP(hipMemcpy(sv, g->stats.p, sizeof sv, hipMemcpyDeviceToHost), "stats copy");
    KC_HIP(hipMemcpy(&km, g->kmax.p, 4, hipMemcpyDeviceToHost), "kmax copy");
    if (sv[kc::KC_ST_ERR] == 1)
        return fail(KC_ETIMEOUT, "device wait timed out");
    if (sv[kc::KC_ST_ERR])
        return fail(KC_EOVERFLOW, "h-index list overflow (%llu)",
                    (unsigned long long)sv[kc::KC_ST_ERR]);
    st.device_ms = ms;
    st.kmax = km;
    st.levels = g->V ? (uint64_t)km + 1 : 0;
    st.iterations = sv[kc::KC_ST_ITERS];
    st.phases = sv[kc::KC_ST_PHASES];
    st.edges_walked = sv[kc::KC_ST_EDGES];
    st.changed = sv[kc::KC_ST_CHANGED];
    st.changed_degree = sv[kc::KC_ST_CHG_DEG];
    st.atomics = sv[kc::KC_ST_ATOMICS];
    st.rebuild_rounds = sv[kc::KC_ST_REBUILDS];
    if (p.trace)
        g->trace_phases = (uint32_t)std::min<uint64_t>(st.phases, KC_TRACE_CAP);
    st.grid = grid;
    st.block = block;
    st.stages = 1;
    return KC_OK;
}

kc_status kci::run_method(kc_graph *g, kc_method method, const kc_opts &o, kc_stats &st)
{
    KC_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    if (g->csr_stale) {
        kc_status s = dyn_materialize(g);
        if (s)
            return s;
    }
    switch (method) {
    case KC_GPP:
    case KC_PEELONE:
    case KC_PEEL_DYN:
    case KC_PEELONE_DYN:
        return run_peel(g, method, o, st);
    case KC_HISTO:
    case KC_CNT:
    case KC_NBR:
        return run_hindex(g, method, o, st);
    default:
        return fail(KC_EINVAL, "method %d not available", (int)method);
    }
}

extern "C" kc_status kc_decompose(kc_graph *g, kc_method method, const kc_opts *opts,
                                  uint32_t *core_out, kc_stats *stats)
{
    if (!g)
        return fail(KC_EINVAL, "null graph");
    double t0 = now_ms();
    kc_opts o{};
    if (opts)
        o = *opts;
    o = with_tuned(g, method, o);
    kc_stats st{};
    kc_status s = run_method(g, method, o, st);
    if (s)
        return s;
    if (core_out && g->V)
        KC_HIP(hipMemcpy(core_out, g->result, (size_t)g->V * 4, hipMemcpyDeviceToHost),
               "core copy");
    st.total_ms = now_ms() - t0;
    if (stats)
        *stats = st;
    return KC_OK;
}

// ------------------------------------------------- reference-shaped entry
static std::mutex g_ctx_mu;
static kc_ctx *g_ctx0 = nullptr;

kc_ctx *kci::default_ctx()
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx0 && kc_ctx_create(0, &g_ctx0) != KC_OK)
        g_ctx0 = nullptr;
    return g_ctx0;
}

static int ref_peel(kc_ref_graph *rg, uint32_t *core_out, kc_method m)
{
    if (!rg || !rg->neighbors_offset || (rg->E && !rg->neighbors)) {
        fail(KC_EINVAL, "null graph");
        return -1;
    }
    kc_ctx *c = default_ctx();
    if (!c)
        return -1;
    std::vector<uint64_t> rp(rg->V + 1ull);
    for (uint64_t v = 0; v <= rg->V; v++)
        rp[v] = rg->neighbors_offset[v];
    kc_graph *g = nullptr;
    if (kc_graph_upload(c, rg->V, rp.data(), rg->neighbors, 4, &g) != KC_OK)
        return -1;
    // timed like the reference: after the graph is resident, through the
    // last level (peel-dyn.cc:152-177)
    auto t0 = std::chrono::steady_clock::now();
    kc_stats st{};
    kc_status s = kc_decompose(g, m, nullptr, nullptr, &st);
    auto t1 = std::chrono::steady_clock::now();
    if (s == KC_OK && core_out && rg->V &&
        hipMemcpy(core_out, g->result, (size_t)rg->V * 4, hipMemcpyDeviceToHost) != hipSuccess)
        s = fail(KC_EDEVICE, "core copy");
    kc_graph_free(g);
    if (s)
        return -1;
    rg->kmax = st.kmax;
    return (int)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
}

extern "C" int kc_peel_dyn(kc_ref_graph *g, uint32_t *core_out)
{
    return ref_peel(g, core_out, KC_PEEL_DYN);
}

extern "C" int kc_peelone_dyn(kc_ref_graph *g, uint32_t *core_out)
{
    return ref_peel(g, core_out, KC_PEELONE_DYN);
}

extern "C" kc_status kc_vgraph_decompose(uint32_t num_nodes, uint64_t num_edges2,
                                         const uint32_t *node_pointer, const uint32_t *edge_list,
                                         kc_method method, const kc_opts *opts,
                                         uint32_t *core_out, kc_stats *stats)
{
    if (!node_pointer || !edge_list)
        return fail(KC_EINVAL, "null argument");
    std::vector<uint64_t> rp(num_nodes + 1ull);
    std::vector<uint32_t> col(num_edges2);
    uint64_t pos = 0;
    for (uint32_t v = 0; v < num_nodes; v++) {
        uint64_t np = node_pointer[v];
        if (np != pos + v)
            return fail(KC_EINVAL, "node_pointer[%u] does not follow the degree-prefixed layout", v);
        uint32_t d = edge_list[np];
        if (pos + d > num_edges2)
            return fail(KC_EINVAL, "degree of vertex %u overruns the edge list", v);
        rp[v] = pos;
        if (d)
            std::memcpy(col.data() + pos, edge_list + np + 1, (size_t)d * 4);
        pos += d;
    }
    rp[num_nodes] = pos;
    if (pos != num_edges2)
        return fail(KC_EINVAL, "degrees sum to %llu, expected %llu", (unsigned long long)pos,
                    (unsigned long long)num_edges2);
    kc_ctx *c = default_ctx();
    if (!c)
        return KC_EDEVICE;
    kc_graph *g = nullptr;
    kc_status s = kc_graph_upload(c, num_nodes, rp.data(), col.data(), 4, &g);
    if (s)
        return s;
    s = kc_decompose(g, method, opts, core_out, stats);
    kc_graph_free(g);
    return s;
}
