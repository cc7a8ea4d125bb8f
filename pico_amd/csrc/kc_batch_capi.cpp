// kc_batch_capi.cpp -- C ABI of the batched engine (host batches, device
// batches, the one-launch run and the per-graph plain-launch comparison) and
// of the launch autotuner.
#include "kc_capi_impl.hpp"

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

using namespace kci;

struct kc_batch {
    kc_ctx *ctx = nullptr;
    uint32_t G = 0;
    uint64_t Vtot = 0, nnz = 0;
    std::vector<uint64_t> vtx_off, edge_off; // host copies
    uint64_t *d_voff = nullptr, *d_eoff = nullptr;
    uint32_t *d_rowptr = nullptr, *d_col = nullptr, *d_core = nullptr, *d_queue = nullptr,
             *d_live = nullptr, *d_order = nullptr, *d_counter = nullptr, *d_kmax = nullptr,
             *d_levels = nullptr;
    uint64_t bytes = 0;
    kc_batch_opts tuned{};
    std::vector<uint32_t> scale; // generator draws (device-built batches)
    std::vector<uint64_t> m_gen;
    kc_graph *view = nullptr; // plain launches on one graph's slice
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

// one built shape of the batch kernel (kc_batch.hip)
struct BatchShape {
    uint32_t block, lds_v, bu;
};

template <typename T> hipError_t dalloc(kc_batch *b, T **p, size_t n)
{
    size_t bytes = n ? n * sizeof(T) : sizeof(T);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e == hipSuccess)
        b->bytes += bytes;
    return e;
}

// Validate the concatenated layout; returns the first bad graph or G.
uint32_t first_bad_graph(uint32_t G, const uint64_t *vo, const uint64_t *eo, const uint32_t *rp,
                         const uint32_t *col, std::string &why)
{
    if (vo[0] != 0 || eo[0] != 0) {
        why = "offsets must start at 0";
        return 0;
    }
    for (uint32_t g = 0; g < G; g++) {
        if (vo[g + 1] < vo[g] || eo[g + 1] < eo[g]) {
            why = "offsets decrease";
            return g;
        }
        if (vo[g + 1] - vo[g] >= 0xffffffffull || eo[g + 1] - eo[g] > 0xffffffffull) {
            why = "graph too large for 32-bit local ids";
            return g;
        }
    }
    if (vo[G] + G > (1ull << 40) || eo[G] > (1ull << 40)) {
        why = "batch too large";
        return 0;
    }
    std::atomic<uint32_t> bad{G};
    kch::parallel_for(G, 4, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t g = lo; g < hi; g++) {
            uint32_t V = (uint32_t)(vo[g + 1] - vo[g]);
            const uint32_t *r = rp + vo[g] + g;
            const uint32_t *c = col + eo[g];
            bool ok = r[0] == 0 && r[V] == eo[g + 1] - eo[g];
            for (uint32_t v = 0; ok && v < V; v++)
                ok = r[v + 1] >= r[v];
            for (uint64_t e = 0; ok && e < r[V]; e++)
                ok = c[e] < V;
            if (!ok) {
                uint32_t cur = bad.load();
                while (g < cur && !bad.compare_exchange_weak(cur, (uint32_t)g)) {
                }
            }
        }
    });
    if (bad.load() < G)
        why = "rows not monotone, not ending at the graph's entry count, or a col id >= V_g";
    return bad.load();
}

kc_status batch_grid(const kc_batch *b, const kc_batch_opts &o, BatchShape &sh, uint32_t &grid)
{
    const uint32_t block = o.block ? o.block : (b->tuned.block ? b->tuned.block : 512);
    if (block != 256 && block != 512 && block != 1024)
        return fail(KC_EINVAL, "block must be 256, 512 or 1024");
    // the tuned shape's capacity and split apply to its own block size only
    const bool tb = b->tuned.block == block;
    uint32_t lv = o.lds_vertices ? o.lds_vertices
                                 : (tb && b->tuned.lds_vertices ? b->tuned.lds_vertices
                                                                : kc::batch_lds_vertices(block, true));
    uint32_t bu = o.split ? o.split : (tb && b->tuned.split ? b->tuned.split : 8u);
    if (lv != kc::batch_lds_vertices(block, true) && lv != kc::batch_lds_vertices(block, false))
        return fail(KC_EINVAL, "lds_vertices %u is not built for block %u (%u or %u)", lv, block,
                    kc::batch_lds_vertices(block, false), kc::batch_lds_vertices(block, true));
    if (bu != 4 && bu != 8)
        return fail(KC_EINVAL, "split must be 4 or 8");
    sh = BatchShape{block, lv, bu};
    int occ = kc::batch_occupancy(block, lv, bu);
    if (occ <= 0)
        return fail(KC_EDEVICE, "batch kernel cannot be resident with block %u", block);
    uint32_t per_cu = o.per_cu ? o.per_cu : (b->tuned.per_cu ? b->tuned.per_cu : (uint32_t)occ);
    per_cu = std::min<uint32_t>(per_cu, (uint32_t)occ);
    grid = per_cu * b->ctx->num_cus;
    grid = std::max<uint32_t>(1, std::min<uint32_t>(grid, b->G ? b->G : 1));
    return KC_OK;
}

// one launch over the whole batch; results stay on the device
kc_status batch_launch(kc_batch *b, const BatchShape &sh, uint32_t grid, float &ms)
{
    hipStream_t st = b->ctx->stream;
    kc::BatchArgs a{};
    a.G = b->G;
    a.vtx_off = b->d_voff;
    a.edge_off = b->d_eoff;
    a.rowptr = b->d_rowptr;
    a.col = b->d_col;
    a.core = b->d_core;
    a.queue = b->d_queue;
    a.live = b->d_live;
    a.order = b->d_order;
    a.counter = b->d_counter;
    a.kmax = b->d_kmax;
    a.levels = b->d_levels;
    KC_HIP(hipMemsetAsync(b->d_counter, 0, 4, st), "counter clear");
    KC_HIP(hipEventRecord(b->ev0, st), "hipEventRecord");
    hipError_t e = kc::launch_batch(a, sh.block, sh.lds_v, sh.bu, grid, st);
    if (e != hipSuccess)
        return hip_fail(e, "batch launch");
    KC_HIP(hipEventRecord(b->ev1, st), "hipEventRecord");
    KC_HIP(hipStreamSynchronize(st), "batch decomposition");
    ms = 0;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    return KC_OK;
}

kc_status batch_results(kc_batch *b, uint32_t *core_out, uint32_t *kmax_out, kc_batch_stats &st,
                        bool levels)
{
    std::vector<uint32_t> km(b->G), lv;
    if (b->G)
        KC_HIP(hipMemcpy(km.data(), b->d_kmax, (size_t)b->G * 4, hipMemcpyDeviceToHost),
               "kmax copy");
    if (levels && b->G) {
        lv.resize(b->G);
        KC_HIP(hipMemcpy(lv.data(), b->d_levels, (size_t)b->G * 4, hipMemcpyDeviceToHost),
               "levels copy");
        st.levels_run = std::accumulate(lv.begin(), lv.end(), (uint64_t)0);
    }
    st.kmax = b->G ? *std::max_element(km.begin(), km.end()) : 0;
    if (kmax_out && b->G)
        std::memcpy(kmax_out, km.data(), (size_t)b->G * 4);
    if (core_out && b->Vtot)
        KC_HIP(hipMemcpy(core_out, b->d_core, (size_t)b->Vtot * 4, hipMemcpyDeviceToHost),
               "core copy");
    st.graphs = b->G;
    st.vertices = b->Vtot;
    st.entries = b->nnz;
    return KC_OK;
}

// Allocate and fill everything but the CSR arrays: offsets, the work order
// (largest graphs first, the longest-processing-time order for the device
// work counter), outputs and scratch.
kc_status batch_finish(kc_batch *b)
{
    const uint32_t G = b->G;
    const auto &vo = b->vtx_off, &eo = b->edge_off;
    std::vector<uint32_t> order(G);
    std::iota(order.begin(), order.end(), 0u);
    auto work = [&](uint32_t g) { return (eo[g + 1] - eo[g]) + (vo[g + 1] - vo[g]); };
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return work(x) > work(y); });
    hipStream_t st = b->ctx->stream;
    hipError_t e = dalloc(b, &b->d_voff, G + 1ull);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_eoff, G + 1ull);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_core, b->Vtot);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_queue, b->Vtot);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_live, b->Vtot);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_order, G);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_counter, 1);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_kmax, G);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_levels, G);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_voff, vo.data(), (G + 1ull) * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_eoff, eo.data(), (G + 1ull) * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && G)
        e = hipMemcpyAsync(b->d_order, order.data(), (size_t)G * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e == hipSuccess)
        e = hipEventCreate(&b->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&b->ev1);
    if (e != hipSuccess)
        return hip_fail(e, "batch allocation");
    return KC_OK;
}

// a device graph that takes ownership of device CSR arrays
kc_status graph_adopt(kc_ctx *ctx, uint32_t V, uint64_t nnz, bool wide, void *rowptr,
                      uint32_t *col, kc_graph **out)
{
    auto *g = new (std::nothrow) kc_graph;
    if (!g) {
        hipFree(rowptr);
        hipFree(col);
        return fail(KC_ENOMEM, "host allocation");
    }
    g->ctx = ctx;
    ctx_ref(ctx);
    g->V = V;
    g->nnz = nnz;
    g->wide = wide;
    g->rowptr = rowptr;
    g->col = col;
    g->bytes = (V + 1ull) * (wide ? 8 : 4) + (nnz ? nnz * 4 : 4);
    hipError_t e = hipEventCreate(&g->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&g->ev1);
    if (e != hipSuccess) {
        kc_graph_free(g);
        return hip_fail(e, "hipEventCreate");
    }
    *out = g;
    return KC_OK;
}

kc_status graph_from_build(kc_ctx *ctx, uint32_t V, uint32_t rowptr_bytes, kc::BuildOut &bo,
                           kc_graph **out)
{
    bool wide = rowptr_bytes == 8 || bo.nnz > 0xffffffffull;
    void *rp = wide ? (void *)bo.rowptr64 : (void *)bo.rowptr;
    if (wide)
        hipFree(bo.rowptr);
    else
        hipFree(bo.rowptr64);
    return graph_adopt(ctx, V, bo.nnz, wide, rp, bo.col, out);
}

void free_build(kc::BuildOut &bo)
{
    hipFree(bo.rowptr);
    hipFree(bo.rowptr64);
    hipFree(bo.col);
    bo = kc::BuildOut{};
}

} // namespace

extern "C" {

// ------------------------------------------------------------ host batches
kc_status kc_hb_gen_rmat(uint32_t G, uint32_t scale_min, uint32_t scale_max, double ef_min,
                         double ef_max, double a, double b, double c, uint64_t seed,
                         kc_host_batch **out)
{
    if (!out || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(KC_EINVAL, "bad R-MAT parameters");
    auto *hb = new (std::nothrow) kc_host_batch;
    if (!hb)
        return fail(KC_ENOMEM, "host allocation");
    if (kch::gen_rmat_batch(G, scale_min, scale_max, ef_min, ef_max, a, b, c, seed, *hb)) {
        delete hb;
        return fail(KC_EINVAL, "bad batch parameters (scales %u..%u, edge factor %g..%g)",
                    scale_min, scale_max, ef_min, ef_max);
    }
    *out = hb;
    return KC_OK;
}

kc_status kc_hb_from_graphs(uint32_t G, const kc_host_graph *const *graphs, kc_host_batch **out)
{
    if (!out || (G && !graphs))
        return fail(KC_EINVAL, "null argument");
    std::vector<kc_host_graph> gs(G);
    for (uint32_t g = 0; g < G; g++) {
        if (!graphs[g])
            return fail(KC_EINVAL, "graph %u is null", g);
        gs[g].V = graphs[g]->V;
        gs[g].rowptr = graphs[g]->rowptr;
        gs[g].col = graphs[g]->col;
    }
    auto *hb = new (std::nothrow) kc_host_batch;
    if (!hb)
        return fail(KC_ENOMEM, "host allocation");
    if (kch::batch_from_graphs(gs, *hb, true)) {
        delete hb;
        return fail(KC_EINVAL, "a graph has 2^32 or more adjacency entries");
    }
    *out = hb;
    return KC_OK;
}

kc_status kc_hb_info(const kc_host_batch *b, kc_batch_info *info)
{
    if (!b || !info)
        return fail(KC_EINVAL, "null argument");
    std::memset(info, 0, sizeof *info);
    info->G = b->G;
    info->vertices = b->vtx_off[b->G];
    info->entries = b->edge_off[b->G];
    info->vtx_off = b->vtx_off.data();
    info->edge_off = b->edge_off.data();
    info->rowptr = b->rowptr.data();
    info->col = b->col.data();
    info->scale = b->scale.empty() ? nullptr : b->scale.data();
    info->m_gen = b->m_gen.empty() ? nullptr : b->m_gen.data();
    return KC_OK;
}

kc_status kc_hb_graph(const kc_host_batch *b, uint32_t g, kc_host_graph **out)
{
    if (!b || !out)
        return fail(KC_EINVAL, "null argument");
    if (g >= b->G)
        return fail(KC_EINVAL, "graph %u out of range (G=%u)", g, b->G);
    auto *h = new (std::nothrow) kc_host_graph;
    if (!h)
        return fail(KC_ENOMEM, "host allocation");
    uint32_t V = (uint32_t)(b->vtx_off[g + 1] - b->vtx_off[g]);
    const uint32_t *rp = b->rowptr.data() + b->vtx_off[g] + g;
    h->V = V;
    h->rowptr.assign(rp, rp + V + 1);
    h->col.assign(b->col.begin() + b->edge_off[g], b->col.begin() + b->edge_off[g + 1]);
    *out = h;
    return KC_OK;
}

void kc_hb_free(kc_host_batch *b) { delete b; }

// ----------------------------------------------------------- device batches
kc_status kc_batch_upload(kc_ctx *ctx, uint32_t G, const uint64_t *vtx_off,
                          const uint64_t *edge_off, const uint32_t *rowptr, const uint32_t *col,
                          kc_batch **out)
{
    if (!ctx || !vtx_off || !edge_off || !out || (G && !rowptr))
        return fail(KC_EINVAL, "null argument");
    if (G == 0xffffffffu)
        return fail(KC_EINVAL, "too many graphs");
    std::string why;
    uint32_t bad = first_bad_graph(G, vtx_off, edge_off, rowptr, col, why);
    if (bad < G)
        return fail(KC_EINVAL, "graph %u: %s", bad, why.c_str());
    if (edge_off[G] && !col)
        return fail(KC_EINVAL, "null col");
    KC_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto *b = new (std::nothrow) kc_batch;
    if (!b)
        return fail(KC_ENOMEM, "host allocation");
    b->ctx = ctx;
    ctx_ref(ctx);
    b->G = G;
    b->Vtot = vtx_off[G];
    b->nnz = edge_off[G];
    b->vtx_off.assign(vtx_off, vtx_off + G + 1);
    b->edge_off.assign(edge_off, edge_off + G + 1);
    hipStream_t st = ctx->stream;
    hipError_t e = dalloc(b, &b->d_rowptr, b->Vtot + G);
    if (e == hipSuccess)
        e = dalloc(b, &b->d_col, b->nnz);
    if (e == hipSuccess && b->Vtot + G)
        e = hipMemcpyAsync(b->d_rowptr, rowptr, (b->Vtot + G) * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && b->nnz)
        e = hipMemcpyAsync(b->d_col, col, b->nnz * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        kc_batch_free(b);
        return hip_fail(e, "batch upload");
    }
    kc_status s = batch_finish(b);
    if (s) {
        kc_batch_free(b);
        return s;
    }
    *out = b;
    return KC_OK;
}

// ------------------------------------------------- device-built graphs
kc_status kc_graph_gen_rmat(kc_ctx *ctx, uint32_t scale, uint64_t m_gen, double a, double b,
                            double c, uint64_t seed, uint32_t rowptr_bytes, kc_graph **out)
{
    if (!ctx || !out || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(KC_EINVAL, "bad R-MAT parameters");
    if (scale == 0 || scale > 31)
        return fail(KC_EINVAL, "R-MAT scale %u out of range", scale);
    if (rowptr_bytes != 4 && rowptr_bytes != 8)
        return fail(KC_EINVAL, "rowptr_bytes must be 4 or 8");
    KC_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    uint64_t vo[2] = {0, 1ull << scale};
    kc::GenSpec s;
    s.kind = kc::GEN_RMAT;
    s.G = 1;
    s.scale = &scale;
    s.m_gen = &m_gen;
    s.vtx_off = vo;
    s.seed = seed;
    s.a = a;
    s.ab = a + b;
    s.abc = a + b + c;
    s.keep_rowptr64 = rowptr_bytes == 8 || 2 * m_gen > 0xffffffffull;
    kc::BuildOut bo;
    std::string err;
    hipError_t e = kc::build_csr_device(s, ctx->stream, bo, err);
    if (e != hipSuccess) {
        free_build(bo);
        return fail(e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EINVAL, "R-MAT build: %s",
                    err.c_str());
    }
    return graph_from_build(ctx, (uint32_t)vo[1], rowptr_bytes, bo, out);
}

kc_status kc_graph_read(kc_ctx *ctx, const char *path, int reader, uint32_t rowptr_bytes,
                        kc_graph **out, kc_read_stats *stats)
{
    if (!ctx || !path || !out)
        return fail(KC_EINVAL, "null argument");
    if (reader != KC_READER_COO && reader != KC_READER_DEDUP)
        return fail(KC_EINVAL, "unknown reader %d", reader);
    if (rowptr_bytes != 4 && rowptr_bytes != 8)
        return fail(KC_EINVAL, "rowptr_bytes must be 4 or 8");
    const double t0 = now_ms();
    kch::EdgeFile ef;
    std::string err;
    int rc = 0;
    try {
        rc = reader == KC_READER_COO ? kch::read_coo_pairs(path, ef, err)
                                     : kch::read_dedup_pairs(path, ef, err);
    } catch (const std::bad_alloc &) {
        return fail(KC_ENOMEM, "host allocation for the edges of %s failed", path);
    }
    if (rc == -1)
        return fail(KC_EIO, "%s", err.c_str());
    if (rc)
        return fail(KC_EINVAL, "%s: %s", path, err.c_str());
    const double t1 = now_ms();
    KC_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    const uint64_t m = ef.src.size();
    const bool keep = reader == KC_READER_COO;
    const uint64_t nnz_max = 2 * m;
    const bool wide = rowptr_bytes == 8 || nnz_max > 0xffffffffull;
    kc::PairCsr pc;
    hipError_t e = kc::build_csr_pairs(ef.src.data(), ef.dst.data(), m, ef.V, keep, wide,
                                       ctx->stream, pc, err);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EINVAL, "device CSR build: %s",
                    err.c_str());
    const double t2 = now_ms();
    kc_status s = graph_adopt(ctx, ef.V, pc.nnz, pc.wide, pc.rowptr, pc.col, out);
    if (s)
        return s;
    if (stats) {
        kc_read_stats r{};
        r.V = ef.V;
        r.has_zero_id = ef.has_zero_id;
        r.multigraph = keep;
        r.threads = ef.threads;
        r.nnz = pc.nnz;
        r.edges_read = m;
        r.file_bytes = ef.file_bytes;
        r.parse_ms = t1 - t0;
        r.build_ms = t2 - t1;
        r.total_ms = now_ms() - t0;
        *stats = r;
    }
    return KC_OK;
}

kc_status kc_graph_gen_chunglu(kc_ctx *ctx, uint32_t n, double beta, double mean_deg, double cap,
                               uint64_t seed, uint32_t rowptr_bytes, kc_graph **out)
{
    if (!ctx || !out)
        return fail(KC_EINVAL, "null argument");
    if (rowptr_bytes != 4 && rowptr_bytes != 8)
        return fail(KC_EINVAL, "rowptr_bytes must be 4 or 8");
    std::vector<uint32_t> thr, alias;
    uint64_t m_gen = 0;
    if (kch::chunglu_alias(n, beta, mean_deg, cap, thr, alias, m_gen))
        return fail(KC_EINVAL, "bad Chung-Lu parameters");
    KC_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    uint32_t *d_thr = nullptr, *d_alias = nullptr;
    hipError_t e = hipMalloc((void **)&d_thr, (size_t)n * 4);
    if (e == hipSuccess)
        e = hipMalloc((void **)&d_alias, (size_t)n * 4);
    if (e == hipSuccess)
        e = hipMemcpy(d_thr, thr.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d_alias, alias.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    kc::BuildOut bo;
    std::string err = "alias table upload";
    if (e == hipSuccess) {
        uint64_t vo[2] = {0, n};
        kc::GenSpec s;
        s.kind = kc::GEN_CHUNGLU;
        s.G = 1;
        s.m_gen = &m_gen;
        s.vtx_off = vo;
        s.seed = seed;
        s.thr = d_thr;
        s.alias = d_alias;
        s.keep_rowptr64 = rowptr_bytes == 8 || 2 * m_gen > 0xffffffffull;
        e = kc::build_csr_device(s, ctx->stream, bo, err);
    }
    hipFree(d_thr);
    hipFree(d_alias);
    if (e != hipSuccess) {
        free_build(bo);
        return fail(e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EINVAL, "Chung-Lu build: %s",
                    err.c_str());
    }
    return graph_from_build(ctx, n, rowptr_bytes, bo, out);
}

kc_status kc_graph_download(const kc_graph *g, kc_host_graph **out)
{
    if (!g || !out)
        return fail(KC_EINVAL, "null argument");
    if (g->csr_stale) { // edge updates since the CSR was last laid out
        kc_status s = dyn_materialize(const_cast<kc_graph *>(g));
        if (s)
            return s;
    }
    auto *h = new (std::nothrow) kc_host_graph;
    if (!h)
        return fail(KC_ENOMEM, "host allocation");
    h->V = g->V;
    h->rowptr.resize(g->V + 1ull);
    h->col.resize(g->nnz);
    hipError_t e;
    if (g->wide) {
        e = hipMemcpy(h->rowptr.data(), g->rowptr, (g->V + 1ull) * 8, hipMemcpyDeviceToHost);
    } else {
        std::vector<uint32_t> r32(g->V + 1ull);
        e = hipMemcpy(r32.data(), g->rowptr, (g->V + 1ull) * 4, hipMemcpyDeviceToHost);
        for (size_t v = 0; v < r32.size(); v++)
            h->rowptr[v] = r32[v];
    }
    if (e == hipSuccess && g->nnz)
        e = hipMemcpy(h->col.data(), g->col, g->nnz * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "graph download");
    }
    *out = h;
    return KC_OK;
}

kc_status kc_graph_degree_stats(kc_graph *g, kc_degree_stats *out)
{
    if (!g || !out)
        return fail(KC_EINVAL, "null argument");
    KC_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    if (g->csr_stale) {
        kc_status s = dyn_materialize(g);
        if (s)
            return s;
    }
    hipStream_t st = g->ctx->stream;
    if (kc_status s = ensure(g, g->aux, 64, "degree statistics"))
        return s;
    unsigned long long r[4] = {0, 0, 0, 0};
    KC_HIP(hipMemsetAsync(g->aux.p, 0, 32, st), "memset");
    KC_HIP(kc::degree_stats(g->rowptr, g->wide, g->V, (unsigned long long *)g->aux.p, st),
           "degree statistics");
    KC_HIP(hipMemcpyAsync(r, g->aux.p, 32, hipMemcpyDeviceToHost, st), "copy");
    KC_HIP(hipStreamSynchronize(st), "degree statistics");
    *out = kc_degree_stats{};
    out->V = g->V;
    out->dmax = (uint32_t)r[0];
    out->nnz = r[1];
    if (g->V) {
        // population variance (V * sum d^2 - (sum d)^2) / V^2, exact up to the division
        unsigned __int128 s2 = (unsigned __int128)r[2] + ((unsigned __int128)r[3] << 32);
        unsigned __int128 num = s2 * g->V - (unsigned __int128)r[1] * r[1];
        long double var = (long double)num / ((long double)g->V * (long double)g->V);
        out->mean = (double)((long double)r[1] / g->V);
        out->stddev = (double)std::sqrt(var);
    }
    return KC_OK;
}

kc_status kc_kcore_subgraph(kc_graph *g, uint32_t k, kc_graph **out, uint32_t *ids, uint32_t *n)
{
    if (!g || !out)
        return fail(KC_EINVAL, "null argument");
    if (!g->result)
        return fail(KC_EINVAL, "the graph has no core numbers yet: decompose it first");
    KC_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    if (g->csr_stale) {
        kc_status s = dyn_materialize(g);
        if (s)
            return s;
    }
    hipStream_t st = g->ctx->stream;
    const uint32_t V = g->V, nw = (V + 63) / 64;
    kc_status s;
    if ((s = ensure(g, g->cs_done[0], (size_t)nw * 8 + 8, "subgraph")) ||
        (s = ensure(g, g->cs_deg[0], (size_t)V * 4 + 4, "subgraph")) ||
        (s = ensure(g, g->cs_wpre, (size_t)(nw + 1) * 4, "subgraph")) ||
        (s = ensure(g, g->cs_bad, 8, "subgraph")))
        return s;
    uint64_t *excl = (uint64_t *)g->cs_done[0].p;
    uint32_t *deg = (uint32_t *)g->cs_deg[0].p;
    KC_HIP(kc::kcore_exclusion(g->result, V, k, excl, st), "k-core marks");
    KC_HIP(kc::kcore_row_lengths(g->rowptr, g->wide, g->col, excl, V, deg, st), "k-core rows");
    std::vector<uint64_t> words(nw);
    if (nw)
        KC_HIP(hipMemcpyAsync(words.data(), excl, (size_t)nw * 8, hipMemcpyDeviceToHost, st),
               "copy");
    KC_HIP(hipStreamSynchronize(st), "k-core marks");
    uint32_t Vn = 0;
    for (uint64_t w : words)
        Vn += (uint32_t)__builtin_popcountll(~w);
    const bool wide_n = g->nnz >= (1ull << 32);
    const size_t rb = wide_n ? 8 : 4;
    if ((s = ensure(g, g->cs_loc, (size_t)Vn * 4 + 4, "subgraph")) ||
        (s = ensure(g, g->cs_gmap[0], (size_t)Vn * 4 + 4, "subgraph")) ||
        (s = ensure(g, g->cs_tmp, kc::compact_tmp_bytes(V, Vn), "subgraph")))
        return s;
    void *rp = nullptr;
    uint32_t *col = nullptr;
    KC_HIP(hipMalloc(&rp, (Vn + 1ull) * rb), "hipMalloc subgraph rows");
    KC_HIP(hipMemsetAsync(g->cs_bad.p, 0, 8, st), "memset");
    kc::CompactIn in{g->rowptr, g->wide, g->col, V, deg, excl, nullptr};
    kc::CompactOut co{};
    co.V_in = V;
    co.Vn = Vn;
    co.wpre = (uint32_t *)g->cs_wpre.p;
    co.loc = (uint32_t *)g->cs_loc.p;
    co.gmap = (uint32_t *)g->cs_gmap[0].p;
    co.rowptr = rp;
    co.wide = wide_n;
    co.bad = (uint32_t *)g->cs_bad.p;
    co.tmp = g->cs_tmp.p;
    co.tmp_bytes = g->cs_tmp.n;
    uint64_t nnz_n = 0;
    uint32_t bad[2] = {0, 0};
    hipError_t e = kc::compact_rows(in, co, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&nnz_n, (char *)rp + (size_t)Vn * rb, rb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(bad, co.bad, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e == hipSuccess && bad[0] == 0 && nnz_n <= g->nnz)
        e = hipMalloc(&col, nnz_n ? nnz_n * 4 : 4);
    if (e == hipSuccess && bad[0] == 0 && col) {
        co.col = col;
        co.col_cap = nnz_n;
        e = kc::compact_cols(in, co, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(bad, co.bad, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && ids && Vn)
            e = hipMemcpyAsync(ids, co.gmap, (size_t)Vn * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess || bad[0] || !col) {
        hipFree(rp);
        hipFree(col);
        if (e != hipSuccess)
            return hip_fail(e, "k-core subgraph");
        return fail(KC_EDEVICE, "k-core subgraph: row lengths do not match (flags %u)", bad[0]);
    }
    if (n)
        *n = Vn;
    return graph_adopt(g->ctx, Vn, nnz_n, wide_n, rp, col, out);
}

kc_status kc_batch_gen_rmat_list(kc_ctx *ctx, uint32_t G, const uint32_t *scale,
                                 const uint64_t *m_gen, const uint64_t *seeds, double a, double b,
                                 double c, kc_batch **out)
{
    if (!ctx || !out || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return fail(KC_EINVAL, "bad R-MAT parameters");
    if (G == 0xffffffffu || (G && (!scale || !m_gen || !seeds)))
        return fail(KC_EINVAL, "bad batch parameters");
    for (uint32_t g = 0; g < G; g++)
        if (scale[g] == 0 || scale[g] > 31)
            return fail(KC_EINVAL, "graph %u: R-MAT scale %u out of range", g, scale[g]);
    KC_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    auto *bt = new (std::nothrow) kc_batch;
    if (!bt)
        return fail(KC_ENOMEM, "host allocation");
    bt->ctx = ctx;
    ctx_ref(ctx);
    bt->G = G;
    bt->scale.assign(scale, scale + G);
    bt->m_gen.assign(m_gen, m_gen + G);
    bt->vtx_off.assign(G + 1ull, 0);
    for (uint32_t g = 0; g < G; g++)
        bt->vtx_off[g + 1] = bt->vtx_off[g] + (1ull << bt->scale[g]);
    bt->Vtot = bt->vtx_off[G];
    kc::GenSpec s;
    s.kind = kc::GEN_RMAT;
    s.G = G;
    s.scale = bt->scale.data();
    s.m_gen = bt->m_gen.data();
    s.vtx_off = bt->vtx_off.data();
    s.seeds = seeds;
    s.a = a;
    s.ab = a + b;
    s.abc = a + b + c;
    kc::BuildOut bo;
    std::string err;
    hipError_t e = kc::build_csr_device(s, ctx->stream, bo, err);
    if (e != hipSuccess) {
        free_build(bo);
        delete bt;
        return fail(e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EINVAL, "batch build: %s",
                    err.c_str());
    }
    hipFree(bo.rowptr64);
    bt->d_rowptr = bo.rowptr;
    bt->d_col = bo.col;
    bt->nnz = bo.nnz;
    bt->edge_off = bo.edge_off;
    bt->bytes = (bt->Vtot + G + 1) * 4 + (bo.nnz ? bo.nnz * 4 : 4);
    kc_status st = batch_finish(bt);
    if (st) {
        kc_batch_free(bt);
        return st;
    }
    *out = bt;
    return KC_OK;
}

kc_status kc_batch_gen_rmat(kc_ctx *ctx, uint32_t G, uint32_t scale_min, uint32_t scale_max,
                            double ef_min, double ef_max, double a, double b, double c,
                            uint64_t seed, kc_batch **out)
{
    if (scale_min == 0 || scale_max > 31 || scale_min > scale_max || ef_min < 0 ||
        ef_max < ef_min || G == 0xffffffffu)
        return fail(KC_EINVAL, "bad batch parameters");
    std::vector<uint32_t> sc;
    std::vector<uint64_t> mg, seeds(G);
    kch::rmat_batch_draws(G, scale_min, scale_max, ef_min, ef_max, seed, sc, mg);
    for (uint32_t g = 0; g < G; g++)
        seeds[g] = seed + g;
    return kc_batch_gen_rmat_list(ctx, G, sc.data(), mg.data(), seeds.data(), a, b, c, out);
}

kc_status kc_batch_checksums(kc_batch *b, uint64_t *out)
{
    if (!b || (b->G && !out))
        return fail(KC_EINVAL, "null argument");
    if (!b->G)
        return KC_OK;
    hipStream_t st = b->ctx->stream;
    uint64_t *d = nullptr;
    KC_HIP(hipMallocAsync((void **)&d, (size_t)b->G * 8, st), "checksum buffer");
    hipError_t e = hipSuccess;
    for (uint32_t g = 0; g < b->G && e == hipSuccess; g++)
        e = kc::core_checksum(b->d_core + b->vtx_off[g],
                              (uint32_t)(b->vtx_off[g + 1] - b->vtx_off[g]), d + g, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d, (size_t)b->G * 8, hipMemcpyDeviceToHost, st);
    hipError_t e2 = hipFreeAsync(d, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e == hipSuccess)
        e = e2;
    return e == hipSuccess ? KC_OK : hip_fail(e, "batch checksums");
}

kc_status kc_batch_graph_cores(const kc_batch *b, uint32_t g, uint32_t *out)
{
    if (!b || !out)
        return fail(KC_EINVAL, "null argument");
    if (g >= b->G)
        return fail(KC_EINVAL, "graph %u out of range (G=%u)", g, b->G);
    uint64_t V = b->vtx_off[g + 1] - b->vtx_off[g];
    if (V)
        KC_HIP(hipMemcpy(out, b->d_core + b->vtx_off[g], V * 4, hipMemcpyDeviceToHost),
               "core copy");
    return KC_OK;
}

kc_status kc_batch_download(const kc_batch *b, kc_host_batch **out)
{
    if (!b || !out)
        return fail(KC_EINVAL, "null argument");
    auto *h = new (std::nothrow) kc_host_batch;
    if (!h)
        return fail(KC_ENOMEM, "host allocation");
    h->G = b->G;
    h->vtx_off = b->vtx_off;
    h->edge_off = b->edge_off;
    h->scale = b->scale;
    h->m_gen = b->m_gen;
    h->rowptr.resize(b->Vtot + b->G);
    h->col.resize(b->nnz);
    hipError_t e = hipSuccess;
    if (!h->rowptr.empty())
        e = hipMemcpy(h->rowptr.data(), b->d_rowptr, h->rowptr.size() * 4,
                      hipMemcpyDeviceToHost);
    if (e == hipSuccess && b->nnz)
        e = hipMemcpy(h->col.data(), b->d_col, b->nnz * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "batch download");
    }
    *out = h;
    return KC_OK;
}

kc_status kc_batch_dims(const kc_batch *b, uint32_t *G, uint64_t *vertices, uint64_t *entries)
{
    if (!b)
        return fail(KC_EINVAL, "null batch");
    if (G)
        *G = b->G;
    if (vertices)
        *vertices = b->Vtot;
    if (entries)
        *entries = b->nnz;
    return KC_OK;
}

kc_status kc_batch_upload_hb(kc_ctx *ctx, const kc_host_batch *hb, kc_batch **out)
{
    if (!hb)
        return fail(KC_EINVAL, "null batch");
    return kc_batch_upload(ctx, hb->G, hb->vtx_off.data(), hb->edge_off.data(),
                           hb->rowptr.data(), hb->col.data(), out);
}

kc_status kc_batch_run(kc_batch *b, const kc_batch_opts *opts, uint32_t *core_out,
                       uint32_t *kmax_out, kc_batch_stats *stats)
{
    if (!b)
        return fail(KC_EINVAL, "null batch");
    double t0 = now_ms();
    KC_HIP(hipSetDevice(b->ctx->device), "hipSetDevice");
    kc_batch_opts o{};
    if (opts)
        o = *opts;
    BatchShape sh{};
    uint32_t grid = 0;
    kc_status s = batch_grid(b, o, sh, grid);
    if (s)
        return s;
    kc_batch_stats st{};
    float ms = 0;
    if (b->G && (s = batch_launch(b, sh, grid, ms)))
        return s;
    if ((s = batch_results(b, core_out, kmax_out, st, true)))
        return s;
    st.block = sh.block;
    st.grid = grid;
    st.launches = b->G ? 1 : 0;
    st.device_ms = ms;
    st.total_ms = now_ms() - t0;
    if (stats)
        *stats = st;
    return KC_OK;
}

kc_status kc_batch_run_plain(kc_batch *b, kc_method method, const kc_opts *opts,
                             uint32_t *core_out, uint32_t *kmax_out, kc_batch_stats *stats)
{
    if (!b)
        return fail(KC_EINVAL, "null batch");
    double t0 = now_ms();
    KC_HIP(hipSetDevice(b->ctx->device), "hipSetDevice");
    if (!b->view) {
        b->view = new (std::nothrow) kc_graph;
        if (!b->view)
            return fail(KC_ENOMEM, "host allocation");
        b->view->ctx = b->ctx;
        ctx_ref(b->ctx);
        b->view->owns_csr = false;
        KC_HIP(hipEventCreate(&b->view->ev0), "hipEventCreate");
        KC_HIP(hipEventCreate(&b->view->ev1), "hipEventCreate");
    }
    kc_opts o{};
    if (opts)
        o = *opts;
    kc_graph *g = b->view;
    std::vector<uint32_t> km(b->G, 0);
    kc_batch_stats st{};
    double dev = 0;
    for (uint32_t gi = 0; gi < b->G; gi++) {
        uint32_t V = (uint32_t)(b->vtx_off[gi + 1] - b->vtx_off[gi]);
        if (!V)
            continue;
        g->V = V;
        g->nnz = b->edge_off[gi + 1] - b->edge_off[gi];
        g->wide = false;
        g->rowptr = b->d_rowptr + b->vtx_off[gi] + gi;
        g->col = b->d_col + b->edge_off[gi];
        g->degsorted = false;
        kc_stats gs{};
        kc_status s = run_method(g, method, o, gs);
        if (s)
            return s;
        KC_HIP(hipMemcpyAsync(b->d_core + b->vtx_off[gi], g->result, (size_t)V * 4,
                              hipMemcpyDeviceToDevice, b->ctx->stream),
               "core copy");
        km[gi] = (uint32_t)gs.kmax;
        st.levels_run += gs.levels_run;
        st.launches++;
        dev += gs.device_ms;
        st.block = gs.block;
        st.grid = gs.grid;
    }
    KC_HIP(hipStreamSynchronize(b->ctx->stream), "plain launches");
    double t1 = now_ms();
    if (b->G)
        KC_HIP(hipMemcpy(b->d_kmax, km.data(), (size_t)b->G * 4, hipMemcpyHostToDevice),
               "kmax copy");
    kc_status s = batch_results(b, core_out, kmax_out, st, false);
    if (s)
        return s;
    st.device_ms = dev;
    st.total_ms = t1 - t0;
    if (stats)
        *stats = st;
    return KC_OK;
}

kc_status kc_batch_autotune(kc_batch *b, uint32_t reps, kc_batch_opts *best, double *best_ms)
{
    if (!b)
        return fail(KC_EINVAL, "null batch");
    KC_HIP(hipSetDevice(b->ctx->device), "hipSetDevice");
    if (!reps)
        reps = 1;
    const kc_batch_opts keep = b->tuned;
    b->tuned = kc_batch_opts{}; // candidates are timed with their own fields only
    kc_batch_opts bo{};
    double bms = 1e300;
    kc_status s = KC_OK;
    auto time_one = [&](const kc_batch_opts &o) -> kc_status {
        BatchShape sh{};
        uint32_t grid = 0;
        kc_status r = batch_grid(b, o, sh, grid);
        if (r)
            return r;
        double t = 1e300;
        for (uint32_t k = 0; k < reps && b->G; k++) {
            float ms = 0;
            if ((r = batch_launch(b, sh, grid, ms)))
                return r;
            t = std::min(t, (double)ms);
        }
        if (t < bms) {
            bms = t;
            bo = o;
        }
        return KC_OK;
    };
    // block x workgroups per CU at the default capacity and split ...
    for (uint32_t block : {256u, 512u, 1024u}) {
        int occ = kc::batch_occupancy(block, kc::batch_lds_vertices(block, true), 8);
        for (uint32_t per_cu : {1u, 2u, 4u, 8u}) {
            if (per_cu > (uint32_t)occ)
                continue;
            kc_batch_opts o{};
            o.block = block;
            o.per_cu = per_cu;
            if ((s = time_one(o)))
                break;
        }
        if (s)
            break;
    }
    // ... then the LDS degree capacity and the cascade split at the best of them
    if (!s) {
        const kc_batch_opts base = bo;
        for (bool large : {true, false})
            for (uint32_t split : {8u, 4u}) {
                kc_batch_opts o = base;
                o.lds_vertices = kc::batch_lds_vertices(base.block, large);
                o.split = split;
                if (!s)
                    s = time_one(o);
            }
    }
    if (s) {
        b->tuned = keep;
        return s;
    }
    if (!bo.lds_vertices)
        bo.lds_vertices = kc::batch_lds_vertices(bo.block, true);
    if (!bo.split)
        bo.split = 8;
    b->tuned = bo;
    if (best)
        *best = bo;
    if (best_ms)
        *best_ms = b->G ? bms : 0;
    return KC_OK;
}

kc_status kc_batch_decompose(kc_ctx *ctx, uint32_t G, const uint64_t *vtx_off,
                             const uint64_t *edge_off, const uint32_t *rowptr, const uint32_t *col,
                             const kc_batch_opts *opts, uint32_t *core_out, kc_batch_stats *stats)
{
    kc_batch *b = nullptr;
    kc_status s = kc_batch_upload(ctx, G, vtx_off, edge_off, rowptr, col, &b);
    if (s)
        return s;
    s = kc_batch_run(b, opts, core_out, nullptr, stats);
    kc_batch_free(b);
    return s;
}

uint64_t kc_batch_device_bytes(const kc_batch *b)
{
    return b ? b->bytes + (b->view ? b->view->bytes : 0) : 0;
}

void kc_batch_free(kc_batch *b)
{
    if (!b)
        return;
    if (b->ctx)
        hipSetDevice(b->ctx->device);
    if (b->view)
        kc_graph_free(b->view);
    for (void *p : {(void *)b->d_voff, (void *)b->d_eoff, (void *)b->d_rowptr, (void *)b->d_col,
                    (void *)b->d_core, (void *)b->d_queue, (void *)b->d_live, (void *)b->d_order,
                    (void *)b->d_counter, (void *)b->d_kmax, (void *)b->d_levels})
        if (p)
            hipFree(p);
    if (b->ev0)
        hipEventDestroy(b->ev0);
    if (b->ev1)
        hipEventDestroy(b->ev1);
    kc_ctx *c = b->ctx;
    delete b;
    ctx_unref(c);
}

// ------------------------------------------------------------ calibration
kc_status kc_probe_copy(kc_ctx *ctx, uint64_t bytes, uint32_t reps, double *ms)
{
    if (!ctx || !ms || bytes < 16)
        return fail(KC_EINVAL, "bad argument");
    bytes &= ~15ull;
    KC_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    void *a = nullptr, *b = nullptr;
    KC_HIP(hipMalloc(&a, bytes), "probe buffer");
    hipError_t e = hipMalloc(&b, bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess)
        e = hipMemsetAsync(a, 1, bytes, ctx->stream);
    if (e == hipSuccess)
        e = hipEventCreate(&e0);
    if (e == hipSuccess)
        e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = kc::copy_probe(a, b, bytes, ctx->stream); // warm-up
    if (e == hipSuccess)
        e = hipEventRecord(e0, ctx->stream);
    for (uint32_t r = 0; r < reps && e == hipSuccess; r++)
        e = kc::copy_probe(a, b, bytes, ctx->stream);
    if (e == hipSuccess)
        e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess)
        e = hipEventSynchronize(e1);
    float t = 0;
    if (e == hipSuccess)
        e = hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess)
        return hip_fail(e, "copy probe");
    *ms = reps ? t / reps : 0.0;
    return KC_OK;
}

// ---------------------------------------------------------------- autotuner
kc_status kc_autotune(kc_graph *g, kc_method method, uint32_t reps, kc_opts *best,
                      double *best_ms)
{
    if (!g)
        return fail(KC_EINVAL, "null graph");
    if ((unsigned)method >= 7)
        return fail(KC_EINVAL, "method %d not available", (int)method);
    if (!reps)
        reps = 1;
    bool peel = method <= KC_PEELONE_DYN;
    std::vector<kc_opts> cand;
    for (uint32_t block : {256u, 512u, 1024u}) {
        if (peel) {
            for (uint32_t hub : {256u, 1024u, 4096u})
                for (uint32_t lc : {0u, 4u, 64u}) // pending edges per thread
                    for (uint32_t ad : {32u, 128u}) {
                        kc_opts o{};
                        o.block = block;
                        o.hub_split = hub;
                        o.local_cap = lc ? lc * block : 0;
                        o.agg_div = ad;
                        cand.push_back(o);
                    }
        } else {
            for (uint32_t gdiv : {0u, 2u}) { // all CUs' worth, or half
                kc_opts o{};
                o.block = block;
                o.grid = gdiv ? std::max(1u, g->ctx->num_cus / gdiv) : 0u;
                cand.push_back(o);
            }
        }
    }
    kc_opts bo{};
    double bms = 1e300;
    auto trial = [&](const kc_opts &o) -> kc_status {
        double t = 1e300;
        for (uint32_t r = 0; r < reps; r++) {
            kc_stats st{};
            kc_status s = run_method(g, method, o, st);
            if (s)
                return s;
            t = std::min(t, st.device_ms);
        }
        if (t < bms) {
            bms = t;
            bo = o;
        }
        return KC_OK;
    };
    for (const kc_opts &o : cand)
        if (kc_status s = trial(o))
            return s;
    // the bucket peels' compaction ratio and the compacted stages' grid,
    // with the best of the rest
    if (method == KC_PEEL_DYN || method == KC_PEELONE_DYN) {
        kc_opts base = bo;
        for (uint32_t cd : {4u, 16u, 0xffffffffu}) {
            kc_opts o = base;
            o.compact_div = cd;
            if (kc_status s = trial(o))
                return s;
        }
        base = bo;
        const uint32_t ncu = g->ctx->num_cus;
        for (uint32_t sg : {ncu / 4, ncu / 2, ncu}) {
            kc_opts o = base;
            o.stage_grid = std::max(1u, sg);
            if (kc_status s = trial(o))
                return s;
        }
    }
    g->tuned[method] = bo;
    if (best)
        *best = bo;
    if (best_ms)
        *best_ms = bms;
    return KC_OK;
}

} // extern "C"
