// kc_engines.hpp -- host-visible launch descriptors of the gfx950 engines.
// Plain structs of device pointers and sizes; no torch or STL types.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>
#include <stdint.h>

namespace kc {

struct GridSync;

// stats[] slots written by the persistent kernels
enum {
    KC_ST_ITERS = 0,      // iteration / round count (method-specific meaning)
    KC_ST_KMAX = 1,       // max core seen by the kernel
    KC_ST_PHASES = 2,     // grid barriers passed
    KC_ST_LEVELS_RUN = 3, // levels actually visited (after skipping)
    KC_ST_EDGES = 4,      // adjacency entries walked
    KC_ST_CHANGED = 5,    // h-index: total core-value changes
    KC_ST_ERR = 6,
    KC_ST_CHG_DEG = 7,    // h-index: degrees of changed vertices, summed
    KC_ST_ATOMICS = 8,    // atomics issued on graph-sized arrays (degrees, claim
                          // bitmap, histograms, counters, marks)
    KC_ST_TAIL_LEVEL = 9, // peel: level at which the kernel stopped early, or ~0
    KC_ST_LIVE = 10,      // peel: vertices still live when the kernel ended
    KC_ST_REBUILDS = 11,  // HistoCore: rounds that recounted every hub's bins
    KC_ST_COUNT = 12
};

// Optional per-phase trace (profiling aid), KC_TRACE_WORDS u64 per phase:
//  [0] phase start (WG 0)          [1] WG 0 work done
//  [2] WG 0 left the barrier       [3] latest work-done time over all WGs
//  [4] level  [5] claims in phase  [6] items pushed to the shared list
//  [7] phase kind
//  [8] most local drain rounds in one WG   [9] edges walked (all WGs)
//  [10] WG 0 drain rounds                  [11] most edges walked by one WG
//  peel only: [12] latest end of a WG's scan  [13] latest end of a WG's own
//  cascade (before it helps with shared work)  [14] latest WG phase start
//  HistoCore pull rounds: [12] latest end of a WG's hub pulls
// Times are wall_clock64 ticks.
enum { KC_TRACE_WORDS = 16 };

struct PeelLaunch {
    const void *rowptr; // u32[V+1] or u64[V+1] (wide)
    const uint32_t *col;
    uint32_t V;
    uint32_t hub_split;
    uint32_t local_cap; // soft cap of a workgroup's local frontier
    uint32_t *deg;  // in/out
    uint64_t *done; // (V+63)/64 words
    uint64_t *list0, *list1;
    uint32_t list_cap;
    uint32_t mode; // 0 bucket, 1 PeelOne sweeps, 2 GPP
    uint32_t *out; // GPP final core; bucket: the live-vertex list (V entries)
    uint32_t *live; // GPP: its live-vertex list (V entries), or null
    uint32_t live_thresh; // bucket, GPP: scan a live-vertex list once <= this many remain
    uint32_t agg_live;    // pre-check degrees and aggregate hub decrements once <= this many remain
    uint32_t agg_walks;   // bucket, PeelOne: sum every walk's decrements per workgroup in LDS (first stage)
    uint32_t tail_live;   // bucket: stop at a level start once <= this many remain (0: never)
    uint32_t start_level; // bucket: first level to visit
    const uint32_t *abort_flag; // device word; non-zero: the kernel leaves at once (error 4)
    GridSync *gs;
    uint64_t *stats;
    uint32_t grid, block;
    bool wide;
    bool gpp_compat;
    bool count_atomics; // run the variant that counts its atomics (stats)
    uint64_t timeout_ticks;
    uint64_t *trace; // may be null
    uint32_t trace_cap;
};

hipError_t launch_peel(const PeelLaunch &p, hipStream_t st);

// Residual compaction (kc_compact.hip).  The live vertices of a stopped peel
// (claim bit clear) are renumbered in id order and their live-live adjacency
// copied, in row order, into a new CSR.
struct CompactIn {
    const void *rowptr; // the stage graph (u32 or u64 rows)
    bool wide;
    const uint32_t *col;
    uint32_t V;
    const uint32_t *deg;   // residual degrees of the live vertices
    const uint64_t *done;  // claim bitmap
    const uint32_t *gmap;  // stage id -> original id, or null for the original graph
};
struct CompactOut {
    uint32_t V_in;        // vertices of the stage graph
    uint32_t Vn;          // live vertices (known from the peel's stats)
    uint32_t *wpre;       // (V+63)/64 + 1 words: live vertices before each claim word
    uint32_t *loc;        // Vn: stage id of each new vertex
    uint32_t *gmap;       // Vn: original id of each new vertex
    void *rowptr;         // Vn+1, u32 (wide == false) or u64
    bool wide;
    uint32_t *col;        // rowptr[Vn] entries
    uint64_t col_cap;     // entries allocated at col (rows reaching past it are refused)
    uint32_t *bad;        // [0] set non-zero if a row's live count differs from its degree,
                          // [1] long rows listed (zero before compact_rows)
    void *tmp;            // scan scratch
    size_t tmp_bytes;
};
// Phase 1: renumbering and row starts.
hipError_t compact_rows(const CompactIn &in, CompactOut &out, hipStream_t st);
size_t compact_tmp_bytes(uint32_t V, uint32_t Vn);
// Phase 2: the adjacency.  A failed consistency check sets bad[0]; the next
// peel stage takes bad as its abort flag.
hipError_t compact_cols(const CompactIn &in, const CompactOut &out, hipStream_t st);
// k-core subgraph: bit set = vertex excluded (core < k or past V), one bit per
// vertex in (V+63)/64 words; row lengths = live neighbours of live vertices.
hipError_t kcore_exclusion(const uint32_t *core, uint32_t V, uint32_t k, uint64_t *excl,
                           hipStream_t st);
hipError_t kcore_row_lengths(const void *rowptr, bool wide, const uint32_t *col,
                             const uint64_t *excl, uint32_t V, uint32_t *deg, hipStream_t st);
// out[4] (zeroed by the caller): max degree, degree sum, and the sums of the
// low and high 32-bit halves of the squared degrees
hipError_t degree_stats(const void *rowptr, bool wide, uint32_t V, unsigned long long *out,
                        hipStream_t st);
// core[gmap[i]] = sub[i] for i < n
hipError_t scatter_cores(const uint32_t *sub, const uint32_t *gmap, uint32_t n, uint32_t *core,
                         hipStream_t st);
int peel_occupancy(uint32_t block, bool wide);

struct HindexLaunch {
    const void *rowptr;
    const uint32_t *col;
    uint32_t V;
    uint32_t method;  // 0 HistoCore, 1 CntCore, 2 NbrCore
    uint32_t *core;   // out
    uint32_t *histo;  // rowptr[V]+V words (reference layout)
    uint32_t *cnt;    // CntCore counters, V words
    uint32_t *newc;   // deferred new values, V words
    uint32_t *mark;   // epoch stamps, V words
    uint32_t *list0, *list1, *list2, *list3; // frontier / changed lists, V each
    uint64_t *chunks; // HistoCore virtual nodes of changed vertices
    uint32_t chunk_cap;
    const uint32_t *col_ds; // degree-ordered adjacency, or null
    const uint32_t *nd_ds;
    const uint32_t *hb_ds;  // histogram base of each col_ds neighbour, or null
    uint32_t *hubs;         // HistoCore hub lists (2 x hub_cap words), or null
    uint32_t hub_cap;
    uint64_t *chg_bm;       // HistoCore pull rounds: 2 x ceil(V/64) words, or null
    uint64_t *chg_rec;      // and V records (old << 32 | new)
    uint32_t pull_den;      // pull when a round's walk > hub adjacency / pull_den (0: default)
    uint32_t *glist;        // HistoCore: list of the vertices of more than 4,096
                            // neighbours (V words)
    void *hprep;            // HistoCore: the initial build's counters (64 bytes)
    uint32_t *giants;       // NbrCore: list of the vertices of more than 4,096
    uint32_t giant_cap;     // neighbours (giant_cap words), or null
    GridSync *gs;
    uint64_t *stats;
    uint32_t grid, block;
    bool wide;
    uint64_t timeout_ticks;
    uint64_t *trace; // may be null
    uint32_t trace_cap;
};

hipError_t launch_hindex(const HindexLaunch &p, hipStream_t st);
// method: 0 HistoCore, 1 CntCore, 2 NbrCore (one kernel instance each)
int hindex_occupancy(uint32_t block, bool wide, uint32_t method);

size_t grid_sync_bytes();

// Batched decomposition (kc_batch.hip): one workgroup per graph at a time.
struct BatchArgs {
    uint32_t G;
    const uint64_t *vtx_off;  // G+1
    const uint64_t *edge_off; // G+1
    const uint32_t *rowptr;   // per graph V_g+1 local offsets at vtx_off[g] + g
    const uint32_t *col;      // local ids at edge_off[g]
    uint32_t *core;           // V_total
    uint32_t *queue;          // V_total
    uint32_t *live;           // V_total
    const uint32_t *order;    // G, largest first
    uint32_t *counter;        // work counter, zeroed before launch
    uint32_t *kmax;           // G
    uint32_t *levels;         // G: levels visited
};
// Device-built synthetic graphs (kc_gen.hip): the host generators' edges,
// sorted, symmetrized and deduplicated into the batch layout.
enum GenKind { GEN_RMAT = 0, GEN_CHUNGLU = 1 };
struct GenSpec {
    GenKind kind = GEN_RMAT;
    uint32_t G = 0;
    const uint32_t *scale = nullptr;   // host, G (R-MAT)
    const uint64_t *m_gen = nullptr;   // host, G: generated pairs
    const uint64_t *vtx_off = nullptr; // host, G+1
    uint64_t seed = 0;                 // graph g uses seed + g ...
    const uint64_t *seeds = nullptr;   // ... or seeds[g] (host, G) when given
    double a = 0, ab = 0, abc = 0;     // R-MAT thresholds as the host computes them
    const uint32_t *thr = nullptr;     // device, Chung-Lu alias table (G == 1)
    const uint32_t *alias = nullptr;
    bool keep_rowptr64 = false;        // also return the u64 global rowptr
};
struct BuildOut {
    uint32_t *rowptr = nullptr;   // device, local u32 rows, vtx_off[G] + G
    uint64_t *rowptr64 = nullptr; // device, global u64 rows, vtx_off[G] + 1 (optional)
    uint32_t *col = nullptr;      // device, nnz
    uint64_t nnz = 0;
    std::vector<uint64_t> edge_off; // host, G+1
};
hipError_t build_csr_device(const GenSpec &s, hipStream_t st, BuildOut &out, std::string &err);

// CSR of an edge file's kept pairs (host arrays, ids < V) built on the
// device: keep_order (the COO reader) gives both directed entries of every
// pair, duplicates kept, each row in file order; otherwise (the
// deduplicating reader) the symmetric, deduplicated CSR with sorted rows.
// rowptr is u64 when wide, else u32; both outputs are owned by the caller.
struct PairCsr {
    void *rowptr = nullptr;
    uint32_t *col = nullptr;
    uint64_t nnz = 0;
    bool wide = false;
};
hipError_t build_csr_pairs(const uint32_t *h_src, const uint32_t *h_dst, uint64_t m, uint32_t V,
                           bool keep_order, bool wide, hipStream_t st, PairCsr &out,
                           std::string &err);

// Core maintenance under edge updates (kc_dyn.hip) on a slack CSR: vertex v
// owns slots [off[v], off[v+1]) of adj, the first len[v] of them in use.
struct DynView {
    uint32_t V;
    const uint64_t *off; // V+1
    uint32_t *len;       // V
    uint32_t *adj;
    uint32_t *core; // V, maintained core numbers
};
// op 0 insert / 1 delete of a matching; status[i]: 0 skipped, 1 applied, 2 row full
hipError_t dyn_edit(const DynView &d, const uint32_t *eu, const uint32_t *ev, uint32_t n,
                    uint32_t op, uint8_t *status, hipStream_t st);
// every deletion at once: segment j lists (sorted) the neighbours hn[hs[j] ..
// hs[j+1]) to drop from row hv[j] and their edges' indexes he[]; status[e]
// becomes 1 for each edge found (status zeroed by the caller)
hipError_t dyn_delete_rows(const DynView &d, const uint32_t *hv, const uint32_t *hs,
                          uint32_t nseg, const uint32_t *hn, const uint32_t *he, uint8_t *status,
                          hipStream_t st);
hipError_t dyn_roots(const DynView &d, const uint32_t *eu, const uint32_t *ev, uint32_t n,
                     const uint8_t *status, uint32_t *mark, uint32_t stamp, uint32_t *list,
                     uint32_t *n_list, uint32_t cap, hipStream_t st);
hipError_t dyn_bfs(const DynView &d, const uint32_t *in, uint32_t n_in, uint32_t *out,
                   uint32_t *n_out, uint32_t *visit, uint32_t stamp, uint32_t *cand,
                   uint32_t *n_cand, uint32_t cap, hipStream_t st);
hipError_t dyn_raise(uint32_t *core, const uint32_t *cand, uint32_t n, hipStream_t st);
hipError_t dyn_round(const DynView &d, const uint32_t *in, uint32_t n_in, uint32_t *out,
                     uint32_t *n_out, uint32_t *mark, uint32_t stamp, uint32_t *n_drops,
                     uint32_t cap, hipStream_t st);
hipError_t dyn_copy(const uint64_t *off_src, const uint32_t *len, const uint32_t *src,
                    const uint64_t *off_dst, uint32_t *dst, uint32_t V, hipStream_t st);
hipError_t dyn_from_csr(const void *rowptr, bool wide, const uint32_t *col, const uint64_t *off,
                        uint32_t *len, uint32_t *adj, uint32_t V, hipStream_t st);
hipError_t dyn_to_csr(const uint64_t *off, const uint32_t *len, const uint32_t *adj,
                      const void *rowptr, bool wide, uint32_t *col, uint32_t V, hipStream_t st);
hipError_t dyn_changed(const uint32_t *a, const uint32_t *b, uint32_t V, uint32_t *n,
                       hipStream_t st);

// streaming copy of bytes (multiple of 16) for bandwidth calibration
hipError_t copy_probe(const void *src, void *dst, uint64_t bytes, hipStream_t st);
// order-independent checksum of a core vector: sum of mix64(v << 32 | core[v])
hipError_t core_checksum(const uint32_t *core, uint32_t V, uint64_t *out, hipStream_t st);

// The batched engine's launch shapes: block 256 / 512 / 1024 threads, LDS
// degree capacity batch_lds_vertices(block, large) vertices, bu = 4 or 8
// cascade edges per thread per pass.  Occupancy 0: not a built shape.
hipError_t launch_batch(const BatchArgs &a, uint32_t block, uint32_t lds_v, uint32_t bu,
                        uint32_t grid, hipStream_t st);
int batch_occupancy(uint32_t block, uint32_t lds_v, uint32_t bu);
uint32_t batch_lds_vertices(uint32_t block, bool large);

// Rows re-ordered by neighbour degree, largest first (kc_prep.hip); nd_out
// receives the neighbour degrees.
hipError_t build_nbr_hbase(const void *rowptr, bool wide, const uint32_t *adj, uint64_t nnz,
                           uint32_t *hb, hipStream_t st);
// nd_tmp: nnz + 1 words of scratch; col_tmp: nnz words of scratch, or null
// (the sort then allocates its own copy of the keys and values).
hipError_t build_degree_sorted(const void *rowptr, bool wide, const uint32_t *col, uint32_t V,
                               uint64_t nnz, uint32_t *col_out, uint32_t *nd_out,
                               uint32_t *nd_tmp, uint32_t *col_tmp, hipStream_t st);

// An empty launch from each source file's code object (kc_*.hip), and all of
// them in turn: loads every kernel of the library.
hipError_t peel_touch(hipStream_t st);
hipError_t compact_touch(hipStream_t st);
hipError_t hindex_touch(hipStream_t st);
hipError_t prep_touch(hipStream_t st);
hipError_t batch_touch(hipStream_t st);
hipError_t gen_touch(hipStream_t st);
hipError_t dyn_touch(hipStream_t st);
inline hipError_t load_modules(hipStream_t st)
{
    hipError_t (*const fns[])(hipStream_t) = {peel_touch, compact_touch, hindex_touch, prep_touch,
                                              batch_touch, gen_touch, dyn_touch};
    for (auto f : fns) {
        hipError_t e = f(st);
        if (e != hipSuccess)
            return e;
    }
    return hipStreamSynchronize(st);
}

} // namespace kc
