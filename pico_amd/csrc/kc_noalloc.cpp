// kc_noalloc.cpp -- kc_decompose_noalloc: the bucket peel of kc_decompose on
// a CSR the caller holds in device memory, with every work buffer carved
// from one static device array instead of hipMalloc (no counterpart in the
// reference; for callers that cannot allocate).
//
// The array's name is a host-side stand-in in host code, not a device
// address: it is located once per device with hipGetSymbolAddress when a
// context is created on that device, and only that address is used
// afterwards.  Every call on one GPU shares that GPU's array (its contents
// are whatever the last call left: everything the peel needs zeroed is
// written by its init kernel), so calls that overlap on one GPU race; they
// must be serialized by the caller.
#include "kc_capi_impl.hpp"

#include <algorithm>
#include <chrono>
#include <mutex>

using namespace kci;

namespace {

constexpr size_t SCRATCH_BYTES = KC_NOALLOC_SCRATCH_BYTES;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

std::mutex g_mu;
void *g_ptr[MAX_DEVICES];         // device address of scratch, per device
hipError_t g_err[MAX_DEVICES];    // why it could not be located

inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte layout of the work buffers in the array (offsets 256-aligned).
struct Layout {
    size_t gs, stats, done, list0, list1, total;
};

Layout layout_for(uint32_t V, uint32_t cap)
{
    Layout l{};
    size_t o = 0;
    l.gs = o;
    o += up256(kc::grid_sync_bytes());
    l.stats = o;
    o += up256(kc::KC_ST_COUNT * 8);
    l.done = o;
    o += up256((size_t)((V + 63ull) / 64) * 8);
    l.list0 = o;
    o += up256((size_t)cap * 8);
    l.list1 = o;
    o += up256((size_t)cap * 8);
    l.total = o;
    return l;
}

} // namespace

kc_status kci::noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return fail(KC_EINVAL, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device])
        return KC_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return hip_fail(g_err[device], "scratch array address");
    }
    g_ptr[device] = p;
    return KC_OK;
}

extern "C" kc_status kc_decompose_noalloc(kc_ctx *ctx, uint32_t V, uint64_t nnz,
                                          const uint32_t *rowptr, const uint32_t *col,
                                          kc_method method, const kc_opts *opts, uint32_t *core,
                                          kc_stats *stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !rowptr || (V && !core) || (nnz && !col))
        return fail(KC_EINVAL, "null argument");
    if (method != KC_PEEL_DYN && method != KC_PEELONE_DYN)
        return fail(KC_EINVAL, "kc_decompose_noalloc runs the bucket peels only (method %d)",
                    (int)method);
    if (V == 0xffffffffu || nnz > 0xffffffffull)
        return fail(KC_EINVAL, "graph too large for u32 row offsets (V=%u, nnz=%llu)", V,
                    (unsigned long long)nnz);
    kc_opts o{};
    if (opts)
        o = *opts;
    if (o.flags & KC_OPT_TRACE)
        return fail(KC_EINVAL, "kc_decompose_noalloc keeps no trace");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ctx->device >= 0 && ctx->device < MAX_DEVICES)
            base = g_ptr[ctx->device];
    }
    if (!base)
        return fail(KC_EDEVICE, "the scratch array of device %d was not located", ctx->device);
    const uint32_t block = o.block ? o.block : 1024;
    if (block != 256 && block != 512 && block != 1024)
        return fail(KC_EINVAL, "block must be 256, 512 or 1024");
    const int occ = kc::peel_occupancy(block, false);
    if (occ <= 0)
        return fail(KC_EDEVICE, "peel kernel cannot be resident with block %u", block);
    const uint32_t per_cu = std::min<uint32_t>((uint32_t)occ, block == 1024 ? 1u : 2u);
    const uint32_t maxgrid = per_cu * ctx->num_cus;
    const uint32_t grid =
        o.grid ? std::min(o.grid, maxgrid) : std::min(maxgrid, size_grid(nnz));
    const uint32_t hub = o.hub_split ? o.hub_split : 1024;
    if (hub < 64 || hub > (1u << 24))
        return fail(KC_EINVAL, "hub_split must be in [64, 2^24]");
    // each claimed vertex or hub chunk is queued at most once per phase
    uint64_t cap64 = o.queue_cap ? o.queue_cap : (uint64_t)V + nnz / hub + 2ull * grid + 64;
    if (cap64 > 0xffffffffull)
        cap64 = 0xffffffffull;
    const Layout l = layout_for(V, (uint32_t)cap64);
    if (l.total > SCRATCH_BYTES)
        return fail(KC_ENOMEM,
                    "graph needs %llu bytes of scratch (V=%u, nnz=%llu); the static array "
                    "holds %llu",
                    (unsigned long long)l.total, V, (unsigned long long)nnz,
                    (unsigned long long)SCRATCH_BYTES);
    uint8_t *s8 = (uint8_t *)base;
    kc::PeelLaunch p{};
    p.rowptr = rowptr;
    p.col = col;
    p.V = V;
    p.hub_split = hub;
    p.local_cap = o.local_cap;
    p.deg = core; // the degrees become the core numbers in place
    p.done = (uint64_t *)(s8 + l.done);
    p.list0 = (uint64_t *)(s8 + l.list0);
    p.list1 = (uint64_t *)(s8 + l.list1);
    p.list_cap = (uint32_t)cap64;
    p.mode = 0;
    p.out = nullptr; // the bucket peel keeps its cores in deg and no vertex list
    const uint32_t agg_div = o.agg_div ? o.agg_div : 128;
    p.agg_live = agg_div == 0xffffffffu ? 0xffffffffu : V / agg_div;
    p.tail_live = 0; // one launch: the residual compaction needs allocations
    p.start_level = 0;
    p.gs = (kc::GridSync *)(s8 + l.gs);
    p.stats = (uint64_t *)(s8 + l.stats);
    p.grid = grid;
    p.block = block;
    p.wide = false;
    p.count_atomics = (o.flags & KC_OPT_COUNT) != 0;
    p.timeout_ticks = ctx->clock_khz * 1000ull * 20ull;
    KC_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    uint64_t sv[kc::KC_ST_COUNT];
    float ms = 0;
    {
        CoopGuard coop(ctx->device, block == 1024 ? grid : 0u, ctx->num_cus);
        KC_HIP(hipEventRecord(ctx->ev0, ctx->stream), "hipEventRecord");
        hipError_t e = kc::launch_peel(p, ctx->stream);
        if (e != hipSuccess)
            return hip_fail(e, "peel launch");
        KC_HIP(hipEventRecord(ctx->ev1, ctx->stream), "hipEventRecord");
        KC_HIP(hipMemcpyAsync(sv, p.stats, sizeof sv, hipMemcpyDeviceToHost, ctx->stream),
               "stats copy");
        KC_HIP(hipStreamSynchronize(ctx->stream), "peel");
        hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    }
    if (sv[kc::KC_ST_ERR] == 1)
        return fail(KC_ETIMEOUT, "device wait timed out");
    if (sv[kc::KC_ST_ERR] == 2)
        return fail(KC_EOVERFLOW, "shared work list overflow (capacity %u items)", p.list_cap);
    if (sv[kc::KC_ST_ERR])
        return fail(KC_EDEVICE, "device consistency check failed (%llu)",
                    (unsigned long long)sv[kc::KC_ST_ERR]);
    if (stats) {
        kc_stats st{};
        st.kmax = (uint32_t)sv[kc::KC_ST_KMAX]; // the last level that claimed a vertex
        st.levels = V ? (uint64_t)st.kmax + 1 : 0;
        st.iterations = st.levels; // peel-dyn.cc:155-173: one trip per level
        st.levels_run = sv[kc::KC_ST_LEVELS_RUN];
        st.phases = sv[kc::KC_ST_PHASES];
        st.edges_walked = sv[kc::KC_ST_EDGES];
        st.atomics = sv[kc::KC_ST_ATOMICS];
        st.stages = 1;
        st.grid = grid;
        st.block = block;
        st.device_ms = ms;
        st.total_ms = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0)
                          .count();
        *stats = st;
    }
    return KC_OK;
}
