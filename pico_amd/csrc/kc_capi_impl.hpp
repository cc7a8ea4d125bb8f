// kc_capi_impl.hpp -- internals shared by the C-ABI translation units:
// contexts, device graphs with their reusable method scratch, error
// reporting and the per-method runners.
#pragma once
#include "../../include/kcore.h"
#include "../host/kc_host.hpp"
#include "kc_engines.hpp"

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <mutex>

namespace kci {
// record a message for kc_last_error() and return s
kc_status fail(kc_status s, const char *fmt, ...);
kc_status hip_fail(hipError_t e, const char *what);
double now_ms();
} // namespace kci

#define KC_HIP(call, what)                                                                         \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return kci::hip_fail(_e, what);                                                        \
    } while (0)

static constexpr uint32_t KC_TRACE_CAP = 1u << 16; // phases kept by the trace

struct kc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t num_cus = 0;
    uint64_t hbm = 0;
    uint64_t clock_khz = 100000; // wall_clock64 rate
    hipEvent_t ev0 = nullptr, ev1 = nullptr; // timing of calls that own no graph
    // the creator's reference plus one per graph or batch made on it: a
    // context destroyed while they live stays until the last one is freed
    std::atomic<int> refs{1};
};

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

struct DynState; // slack CSR + maintained cores (kc_dyn_capi.cpp)
void dyn_free(DynState *d);

struct kc_graph {
    kc_ctx *ctx = nullptr;
    uint32_t V = 0;
    uint64_t nnz = 0;
    bool wide = false;
    bool owns_csr = true; // false for a view into another object's arrays
    void *rowptr = nullptr;
    uint32_t *col = nullptr;
    uint64_t bytes = 0;
    // method scratch, allocated on first use and kept for reuse
    DevBuf deg, out, done, list0, list1, gs, stats, histo, cnt, flags, hl0, hl1, hl2, kmax, trace;
    DevBuf col_ds, nd_ds, hb_ds; // degree-ordered adjacency (built on first HistoCore run)
    DevBuf hubs;                 // HistoCore's list of high-degree vertices
    DevBuf hchg_bm, hchg_rec;    // HistoCore pull rounds: change bitmaps and records
    DevBuf hprep;                // HistoCore: counters of the initial build
    // bucket peel on compacted live subgraphs: two alternating stage sets
    DevBuf cs_row[2], cs_col[2], cs_deg[2], cs_done[2], cs_gmap[2], cs_loc, cs_wpre, cs_tmp, cs_bad;
    DevBuf aux; // small reductions (degree statistics)
    bool degsorted = false;
    uint32_t trace_phases = 0;
    const uint32_t *result = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    kc_opts tuned[7] = {}; // autotuner's pick per method (block == 0: none)
    DynState *dyn = nullptr; // created by the first kc_update
    bool csr_stale = false;  // rowptr/col predate the last kc_update
};

// fill the zero fields of o from the graph's tuned options for `method`
static inline kc_opts with_tuned(const kc_graph *g, kc_method method, kc_opts o)
{
    if ((unsigned)method >= 7 || !g->tuned[method].block)
        return o;
    const kc_opts &t = g->tuned[method];
    if (!o.block)
        o.block = t.block;
    if (!o.hub_split)
        o.hub_split = t.hub_split;
    if (!o.local_cap)
        o.local_cap = t.local_cap;
    if (!o.agg_div)
        o.agg_div = t.agg_div;
    if (!o.compact_div)
        o.compact_div = t.compact_div;
    if (!o.grid)
        o.grid = t.grid;
    if (!o.stage_grid)
        o.stage_grid = t.stage_grid;
    return o;
}

namespace kci {
// Workgroup slots reserved by a persistent (grid-barrier) launch on a device
// for its lifetime, in arrival order: `need` 1024-thread workgroups out of a
// budget of `capacity` (one per CU); need 0 (or above the budget) reserves
// the whole device.
struct CoopGuard {
    int device;
    uint32_t n = 0;
    CoopGuard(int device, uint32_t need, uint32_t capacity);
    ~CoopGuard();
    CoopGuard(const CoopGuard &) = delete;
    CoopGuard &operator=(const CoopGuard &) = delete;
};
// default persistent grid width for a graph of nnz adjacency entries
uint32_t size_grid(uint64_t nnz);
// locate the static scratch array of kc_decompose_noalloc on the current
// device (kc_noalloc.cpp); called once per context creation
kc_status noalloc_setup(int device);
// take / drop a reference to a context (the last drop destroys it)
void ctx_ref(kc_ctx *c);
void ctx_unref(kc_ctx *c);
// grow b to at least `bytes` (contents are not kept)
kc_status ensure(kc_graph *g, DevBuf &b, size_t bytes, const char *what);
// run one method on the device; the result stays at g->result
kc_status run_method(kc_graph *g, kc_method method, const kc_opts &o, kc_stats &st);
// the process-wide context on device 0 used by the reference-shaped entries
kc_ctx *default_ctx();
// rebuild rowptr/col from the slack CSR after edge updates
kc_status dyn_materialize(kc_graph *g);
} // namespace kci
