/*
 * kcore.h -- C ABI of libkcore, the MI355X (gfx950) k-core engine.
 *
 * Plain C types only: pointers, sizes, status codes.  No function ever calls
 * exit(); every failure returns a kc_status and leaves a message readable
 * with kc_last_error() (thread-local).
 *
 * What each group replaces in the reference (chriszhao13/pico):
 *  - host graph readers/generators  <- Graph::ReadGraph (peel/shared/graph.cpp:17-99),
 *    VirtualGraph (peel/shared/virtual_graph.cpp:9-183), Graph(string) /
 *    readFile (peel/src/graph.cpp:58-162)
 *  - kc_ctx / kc_graph upload       <- the cudaMalloc/cudaMemcpy blocks of each
 *    main (gppeel.cu:128-176, HistoCore.cu:222-286, ...) and
 *    malloc_graph_gpu_memory / free_graph_gpu_memory
 *    (peel/src/gpu_memory_allocation.cu:4-22)
 *  - kc_decompose                   <- the per-method host loops plus kernels
 *    (gppeel.cu:186-228, peelone.cu:157-186, peel-dyn.cc:133-182,
 *    peelone-dyn.cc:120-168, HistoCore.cu:289-341, CntCore.cu:341-417,
 *    NbrCore.cu:241-297), returning the core vector the reference keeps on
 *    the device
 *  - kc_peel_dyn / kc_peelone_dyn   <- `int peel_dyn(Graph&)` (peel-dyn.cc:133) and
 *    `int peelone_dyn(Graph&)` (peelone-dyn.cc:120), the function pointers
 *    main.cu:28-29 hands to repSimulation
 *  - kc_server_*                     <- no counterpart: the reference has no
 *    request loop
 *  - kc_update                       <- no counterpart: core maintenance under
 *    edge insertions/deletions is new functionality
 *  - kc_batch_*, kc_hb_*, kc_autotune  <- no counterpart: the reference runs
 *    one graph per process with the fixed launch constants of
 *    peel/inc/common.h:4-14 and peel/shared/argument.h:24
 *  - kc_vgraph_decompose            <- the device half of gppeel.cu / peelone.cu /
 *    HistoCore.cu / CntCore.cu / NbrCore.cu main(), taking the
 *    degree-prefixed VirtualGraph layout those mains build
 */
#ifndef KCORE_H
#define KCORE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KC_API __attribute__((visibility("default")))

typedef enum kc_status {
    KC_OK = 0,
    KC_EINVAL = 1,    /* bad argument or malformed graph */
    KC_ENOMEM = 2,    /* host or device allocation failed / over budget */
    KC_EDEVICE = 3,   /* HIP runtime error */
    KC_EOVERFLOW = 4, /* a device work list overflowed its capacity */
    KC_ETIMEOUT = 5,  /* a device-side wait exceeded its time limit */
    KC_EIO = 6        /* file could not be read or written */
} kc_status;

typedef enum kc_method {
    KC_GPP = 0,         /* level-synchronous peel, gppeel.cu */
    KC_PEELONE = 1,     /* PeelOne sweeps, peelone.cu */
    KC_PEEL_DYN = 2,    /* bucket peel "ICDE23_peel", peel-dyn.cc */
    KC_PEELONE_DYN = 3, /* bucket peel "Our_Paper_peel", peelone-dyn.cc */
    KC_HISTO = 4,       /* HistoCore.cu */
    KC_CNT = 5,         /* CntCore.cu */
    KC_NBR = 6          /* NbrCore.cu */
} kc_method;

/* kc_opts.flags bits */
#define KC_OPT_TRACE 1u       /* record a per-phase device trace (kc_graph_trace); the peel
                                 methods record it in the variant that also counts atomics */
#define KC_OPT_NO_DEGSORT 2u  /* h-index methods: walk the adjacency in input order */
#define KC_OPT_COUNT 4u       /* peel methods: run the variant that counts its atomics
                                 (kc_stats.atomics); a little slower, so off by default */

/* Launch options; zero fields select defaults (or the autotuner's pick). */
typedef struct kc_opts {
    uint32_t block;      /* threads per workgroup: 256, 512 or 1024; 0: 512 for HistoCore,
                            1024 for the other methods */
    uint32_t grid;       /* persistent workgroups (<= occupancy * CUs); 0: the method's
                            default, narrowed for graphs of fewer than ~30M adjacency
                            entries to about 48 * sqrt(nnz / 1e6), at least 32 */
    uint32_t hub_split;  /* edges per virtual node of a hub vertex; 0: 4096 (kc_decompose_noalloc:
                            1024) */
    uint32_t queue_cap;  /* shared work-list capacity in items */
    uint32_t gpp_compat; /* KC_GPP: report isolated vertices as 1 (reference quirk) */
    uint32_t part_size;  /* PeelOne reference part width (reporting only) */
    uint32_t local_cap;  /* pending-edge cap of a workgroup's local frontier; the
                            excess is shared with every workgroup */
    uint32_t flags;      /* KC_OPT_* */
    uint32_t agg_div;    /* bucket peels: read each target's degree before decrementing and
                            sum hub decrements in LDS once at most V/agg_div vertices are
                            live (0: 128; 0xffffffff: from the start) */
    uint32_t compact_div; /* bucket peels and PeelOne: once at most V/compact_div vertices are live (and
                             at least 4,096), continue on a compacted copy of the live
                             subgraph, again at each further 1/compact_div; 0: default (8),
                             0xffffffff: never */
    uint32_t stage_grid; /* bucket peels and PeelOne: persistent workgroups for the compacted
                            stages (<= occupancy * CUs); 0: half of the largest resident grid */
    uint32_t pull_div;   /* HistoCore: a round applies the hubs' updates by pulls (each hub
                            scans its row against the round's changed vertices) when the
                            round's walk exceeds 1/pull_div of the hubs' adjacency; 0: 5,
                            0xffffffff: always */
} kc_opts;

typedef struct kc_stats {
    uint64_t iterations; /* the method's "Number of iterations" */
    uint64_t levels;     /* peel levels 0..kmax (= kmax+1) */
    uint64_t levels_run; /* levels actually visited after skipping empty ones */
    uint64_t phases;     /* grid barriers passed */
    uint64_t edges_walked; /* peel: adjacency entries walked; HistoCore: adjacency of
                              changed vertices walked by the update rounds */
    uint32_t kmax;
    uint32_t grid;
    uint32_t block;
    uint32_t stages;  /* peel kernel launches: 1, plus one per compaction of the live
                         subgraph (bucket peels, kc_opts.compact_div) */
    double device_ms; /* HIP-event time of the decomposition on its stream */
    double total_ms;  /* host wall time of the call incl. copies */
    uint64_t changed; /* h-index: core-value changes over all rounds */
    double prep_ms;   /* one-time layout preparation done by this call (HistoCore's
                         degree-ordered adjacency), outside device_ms */
    uint64_t changed_degree; /* HistoCore: degrees of the changed vertices summed over
                                rounds (the adjacency the reference's update walks) */
    uint64_t atomics; /* atomic read-modify-writes the kernel issued on graph-sized
                         arrays: degrees and claim bitmap (peel, with KC_OPT_COUNT;
                         0 without), histograms and marks (h-index, always); queue
                         counters are not included */
    uint64_t rebuild_rounds; /* HistoCore: update rounds whose hubs pulled their changed
                                neighbours' updates instead of receiving patches */
} kc_stats;

KC_API const char *kc_last_error(void);
KC_API const char *kc_version(void);

/* ------------------------------------------------------------ host graphs */
typedef struct kc_host_graph kc_host_graph;

typedef struct kc_graph_info {
    uint32_t V;
    uint32_t dmax;
    uint64_t nnz;         /* adjacency entries (2 x undirected edges) */
    uint64_t edges_read;  /* edge lines kept by a file reader */
    uint32_t has_zero_id; /* COO reader: some id was 0 (no 1-based shift) */
    uint32_t multigraph;  /* COO reader keeps duplicate edges */
    const uint64_t *rowptr; /* owned by the graph */
    const uint32_t *col;
} kc_graph_info;

/* reader of gpp/peelone/HistoCore/CntCore/NbrCore (shared/graph.cpp:17-99) */
KC_API kc_status kc_hg_read_coo(const char *path, kc_host_graph **out);
/* reader of peelone_dyn (src/graph.cpp:58-133): 1-based, deduplicated */
KC_API kc_status kc_hg_read_dedup(const char *path, kc_host_graph **out);
/* symmetrize + drop self-loops + dedup (n directed pairs over V vertices) */
KC_API kc_status kc_hg_from_pairs(uint32_t V, uint64_t n, const uint32_t *src,
                                  const uint32_t *dst, kc_host_graph **out);
/* copy of a symmetric CSR as given (duplicates kept) */
KC_API kc_status kc_hg_from_csr(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                                kc_host_graph **out);
KC_API kc_status kc_hg_gen_er(uint32_t n, uint64_t m, uint64_t seed, kc_host_graph **out);
KC_API kc_status kc_hg_gen_rmat(uint32_t scale, uint64_t m_gen, double a, double b, double c,
                                uint64_t seed, kc_host_graph **out);
KC_API kc_status kc_hg_gen_chunglu(uint32_t n, double beta, double mean_deg, double cap,
                                   uint64_t seed, kc_host_graph **out);
KC_API kc_status kc_hg_info(const kc_host_graph *g, kc_graph_info *info);
/* degree-prefixed layout of virtual_graph.cpp:93-183: node_pointer[V],
 * edge_list[nnz+V], parts[2*num_parts] ({node, part} pairs); any output may
 * be NULL to query *num_parts only */
KC_API kc_status kc_hg_vgraph(const kc_host_graph *g, uint32_t part_size,
                              uint32_t *node_pointer, uint32_t *edge_list, uint32_t *parts,
                              uint64_t *num_parts);
/* example.txt dialect: 2 header lines, 1-based "u v 1" per undirected edge */
KC_API kc_status kc_hg_write_edgelist(const kc_host_graph *g, const char *path);
KC_API void kc_hg_free(kc_host_graph *g);

/* Binary CSR files and a content-keyed reader cache, in place of the text
 * ./output/serialized-<basename> files of Graph::readSerialized /
 * writeSerialized (peel/src/graph.cpp:9-54).  A cache entry records the
 * 64-bit content hash and size of the edge file and the reader that parsed
 * it, and is named after that hash: lookups are by content, not by name.  A
 * damaged, truncated or mismatched entry counts as a miss: the edge file is
 * parsed again and the entry rewritten. */
#define KC_READER_COO 0   /* kc_hg_read_coo */
#define KC_READER_DEDUP 1 /* kc_hg_read_dedup */
KC_API kc_status kc_hg_save(const kc_host_graph *g, const char *path);
KC_API kc_status kc_hg_load(const char *path, kc_host_graph **out);
KC_API kc_status kc_hg_read_cached(const char *path, int reader, const char *cache_dir,
                                   kc_host_graph **out, int *hit);

/* "i core\n" per vertex, the format of utility::SaveResults (util.cpp:70-89) */
KC_API kc_status kc_write_cores(const char *path, const uint32_t *core, uint32_t V);

/* ---------------------------------------------------------- device side */
typedef struct kc_ctx kc_ctx;
typedef struct kc_graph kc_graph;

/* Degree statistics of a resident graph: what peelone_dyn prints before it
 * peels (main.cu:10-63: PrintMax's " Maxdegree = ", STDdegrees' population
 * standard deviation). */
typedef struct kc_degree_stats {
    uint32_t V;
    uint32_t dmax;
    uint64_t nnz;
    double mean;
    double stddev; /* population standard deviation of the degrees */
    uint64_t reserved[2];
} kc_degree_stats;

KC_API kc_status kc_ctx_create(int device, kc_ctx **out);
KC_API void kc_ctx_destroy(kc_ctx *ctx);
KC_API kc_status kc_ctx_info(const kc_ctx *ctx, uint32_t *num_cus, uint64_t *hbm_bytes);

/* Copy a symmetric CSR to the device.  rowptr_bytes is 4 or 8: the width the
 * device keeps (an 8-byte host rowptr may be narrowed when nnz < 2^32).
 * Inputs are caller-owned and not retained.  Every col[] entry is checked
 * against V. */
KC_API kc_status kc_graph_upload(kc_ctx *ctx, uint32_t V, const uint64_t *rowptr,
                                 const uint32_t *col, uint32_t rowptr_bytes, kc_graph **out);
KC_API kc_status kc_graph_upload_hg(kc_ctx *ctx, const kc_host_graph *g,
                                    uint32_t rowptr_bytes, kc_graph **out);
KC_API void kc_graph_free(kc_graph *g);
/* device bytes held by the graph (CSR + method scratch allocated so far) */
KC_API uint64_t kc_graph_device_bytes(const kc_graph *g);

/* Decompose; core_out (host, V entries) may be NULL to leave the result on
 * the device (kc_graph_core_device returns it). */
KC_API kc_status kc_decompose(kc_graph *g, kc_method method, const kc_opts *opts,
                              uint32_t *core_out, kc_stats *stats);
KC_API const uint32_t *kc_graph_core_device(const kc_graph *g);
/* Per-phase trace of the last decomposition run with KC_OPT_TRACE: 8 u64
 * per phase (start, WG0 done, barrier exit, slowest WG done, level, claims,
 * shared items, phase kind), times in device ticks (*ticks_per_ms).
 * Returns the number of phases copied. */
KC_API uint32_t kc_graph_trace(const kc_graph *g, uint64_t *out, uint32_t max_phases,
                               uint64_t *ticks_per_ms);
/* HistoCore's histogram after the last h-index run, in the reference layout
 * (HistoCore.cu:265: nnz+V words parallel to the degree-prefixed edge list);
 * n <= nnz + V words are copied. */
KC_API kc_status kc_graph_histogram(const kc_graph *g, uint32_t *out, uint64_t n);

/* ------------------------------------------------- no device allocation */
/* No counterpart in the reference.  For callers that cannot call hipMalloc
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * device allocation): kc_decompose's bucket peel (KC_PEEL_DYN /
 * KC_PEELONE_DYN, one persistent launch, no residual compaction) on a CSR
 * the caller holds in device memory on ctx's device -- rowptr u32[V+1],
 * col u32[nnz] -- writing the V core numbers to the caller's device buffer
 * core.  Every work buffer (claim bitmap, work lists, barrier and
 * statistics block) lives in one static device array of
 * KC_NOALLOC_SCRATCH_BYTES on each GPU, held until the process ends; the
 * call makes no device allocation.  A graph whose work buffers need more
 * (about 16 bytes per vertex plus 16 per 1,024 adjacency entries: R-MAT
 * scale 22 fits) is refused with KC_ENOMEM.  The array is shared by every
 * call on one GPU: calls that overlap on one GPU race and must be
 * serialized by the caller.  Cores equal kc_decompose's. */
#define KC_NOALLOC_SCRATCH_BYTES (96u << 20)
KC_API kc_status kc_decompose_noalloc(kc_ctx *ctx, uint32_t V, uint64_t nnz,
                                      const uint32_t *rowptr, const uint32_t *col,
                                      kc_method method, const kc_opts *opts, uint32_t *core,
                                      kc_stats *stats);

/* ------------------------------------------------- reference-shaped ABI */
/* Mirrors the data members of the peelone_dyn Graph (peel/inc/graph.h:7-23). */
typedef struct kc_ref_graph {
    uint32_t V;
    uint32_t E; /* adjacency entries, neighbors_offset[V] */
    uint32_t *neighbors;
    uint32_t *neighbors_offset;
    uint32_t *degrees;
    uint32_t kmax;
    uint32_t dmax;
} kc_ref_graph;

/* Drop-in for peel_dyn / peelone_dyn: upload, decompose, free; returns the
 * elapsed milliseconds (integer, as the reference) and sets g->kmax, or -1
 * on error.  core_out may be NULL (the reference never copies cores back). */
KC_API int kc_peel_dyn(kc_ref_graph *g, uint32_t *core_out);
KC_API int kc_peelone_dyn(kc_ref_graph *g, uint32_t *core_out);

/* Device half of the virtual-graph programs on the degree-prefixed layout
 * their mains hold after MakeUGraph(): node_pointer[num_nodes],
 * edge_list[num_edges2 + num_nodes] with edge_list[node_pointer[v]] = deg(v). */
KC_API kc_status kc_vgraph_decompose(uint32_t num_nodes, uint64_t num_edges2,
                                     const uint32_t *node_pointer, const uint32_t *edge_list,
                                     kc_method method, const kc_opts *opts, uint32_t *core_out,
                                     kc_stats *stats);

/* ------------------------------------------------------------- batches */
/* Many independent graphs in one concatenated int32 CSR.  Graph g owns
 * vertices [vtx_off[g], vtx_off[g+1]) and adjacency entries
 * [edge_off[g], edge_off[g+1]); its V_g+1 local row offsets (u32) start at
 * rowptr[vtx_off[g] + g] and its col entries are local ids in [0, V_g).
 * Cores come back concatenated the same way (vtx_off[G] entries). */
typedef struct kc_host_batch kc_host_batch;
typedef struct kc_batch kc_batch;

typedef struct kc_batch_info {
    uint32_t G;
    uint32_t reserved0;
    uint64_t vertices; /* vtx_off[G] */
    uint64_t entries;  /* edge_off[G] */
    const uint64_t *vtx_off;
    const uint64_t *edge_off;
    const uint32_t *rowptr;
    const uint32_t *col;
    const uint32_t *scale; /* generator draws per graph, or NULL */
    const uint64_t *m_gen; /* generated edges per graph, or NULL */
} kc_batch_info;

/* Launch options of the batched engine; zero fields select the autotuner's
 * pick when kc_batch_autotune has run, else defaults. */
typedef struct kc_batch_opts {
    uint32_t block;        /* threads per workgroup: 256, 512 or 1024 */
    uint32_t per_cu;       /* workgroups launched per CU */
    uint32_t lds_vertices; /* graphs of at most this many vertices keep their
                              degrees in LDS: 4096 or 8192 (block 256), 8192 or
                              16384 (512), 16384 or 32768 (1024); default the
                              larger.  Larger graphs use HBM. */
    uint32_t split;        /* cascade edges per thread per pass, the width of
                              the edge-balanced split of a level's frontier: 4
                              or 8 (default) */
    uint32_t reserved[4];
} kc_batch_opts;

typedef struct kc_batch_stats {
    uint64_t graphs;
    uint64_t vertices;
    uint64_t entries;    /* adjacency entries (2x undirected edges) */
    uint64_t levels_run; /* levels visited, summed over graphs */
    uint32_t kmax;       /* largest core number in the batch */
    uint32_t block;
    uint32_t grid;
    uint32_t launches; /* kernel launches issued for the batch */
    double device_ms;  /* GPU time of the launch(es) */
    double total_ms;   /* host wall time of the call */
    uint64_t reserved[4];
} kc_batch_stats;

/* R-MAT batch: graph g uses seed + g, scale uniform over the integers
 * [scale_min, scale_max], edge factor uniform in [ef_min, ef_max]. */
KC_API kc_status kc_hb_gen_rmat(uint32_t G, uint32_t scale_min, uint32_t scale_max,
                                double ef_min, double ef_max, double a, double b, double c,
                                uint64_t seed, kc_host_batch **out);
/* concatenate G host graphs (copied) */
KC_API kc_status kc_hb_from_graphs(uint32_t G, const kc_host_graph *const *graphs,
                                   kc_host_batch **out);
KC_API kc_status kc_hb_info(const kc_host_batch *b, kc_batch_info *info);
/* graph g of a host batch as a stand-alone host graph */
KC_API kc_status kc_hb_graph(const kc_host_batch *b, uint32_t g, kc_host_graph **out);
KC_API void kc_hb_free(kc_host_batch *b);

/* Copy a batch to the device (inputs caller-owned, not retained).  Every
 * graph's rows and col entries are validated. */
KC_API kc_status kc_batch_upload(kc_ctx *ctx, uint32_t G, const uint64_t *vtx_off,
                                 const uint64_t *edge_off, const uint32_t *rowptr,
                                 const uint32_t *col, kc_batch **out);
KC_API kc_status kc_batch_upload_hb(kc_ctx *ctx, const kc_host_batch *hb, kc_batch **out);
/* Bucket-peel core numbers of every graph in ONE persistent launch (one
 * workgroup per graph, largest graphs first from a device work counter).
 * core_out: vtx_off[G] entries, kmax_out: G entries; either may be NULL. */
KC_API kc_status kc_batch_run(kc_batch *b, const kc_batch_opts *opts, uint32_t *core_out,
                              uint32_t *kmax_out, kc_batch_stats *stats);
/* The same batch as per-graph plain launches of kc_decompose's engines on
 * the device-resident slices, one after another (the comparison point). */
KC_API kc_status kc_batch_run_plain(kc_batch *b, kc_method method, const kc_opts *opts,
                                    uint32_t *core_out, uint32_t *kmax_out,
                                    kc_batch_stats *stats);
/* Time every (block, per_cu) candidate on the batch itself, then every
 * (lds_vertices, split) candidate at the fastest of those; keep the fastest
 * for later kc_batch_run calls and return it. */
KC_API kc_status kc_batch_autotune(kc_batch *b, uint32_t reps, kc_batch_opts *best,
                                   double *best_ms);
/* upload + run + free in one call */
KC_API kc_status kc_batch_decompose(kc_ctx *ctx, uint32_t G, const uint64_t *vtx_off,
                                    const uint64_t *edge_off, const uint32_t *rowptr,
                                    const uint32_t *col, const kc_batch_opts *opts,
                                    uint32_t *core_out, kc_batch_stats *stats);
KC_API uint64_t kc_batch_device_bytes(const kc_batch *b);
KC_API void kc_batch_free(kc_batch *b);

/* ------------------------------------------------- device-built graphs */
/* The graphs of kc_hg_gen_rmat / kc_hg_gen_chunglu / kc_hb_gen_rmat (same
 * seeds, same edges), generated, symmetrized, sorted and deduplicated in HBM
 * instead of on the host.  kc_*_download copies one back for inspection. */
KC_API kc_status kc_graph_gen_rmat(kc_ctx *ctx, uint32_t scale, uint64_t m_gen, double a,
                                   double b, double c, uint64_t seed, uint32_t rowptr_bytes,
                                   kc_graph **out);
KC_API kc_status kc_graph_gen_chunglu(kc_ctx *ctx, uint32_t n, double beta, double mean_deg,
                                      double cap, uint64_t seed, uint32_t rowptr_bytes,
                                      kc_graph **out);
KC_API kc_status kc_graph_download(const kc_graph *g, kc_host_graph **out);

/* An edge file straight into a resident graph: the replacement of
 * Graph::ReadGraph + VirtualGraph (peel/shared/graph.cpp:17-99,
 * virtual_graph.cpp:9-157; reader KC_READER_COO) and of Graph(string) /
 * readFile (peel/src/graph.cpp:58-133; KC_READER_DEDUP) followed by the
 * upload.  The file is memory-mapped and parsed by host threads in chunks of
 * whole lines (results as one sequential pass, reader quirks included), and
 * the CSR is built on the device: a stable radix sort of the directed entries
 * for the COO reader (duplicates kept, rows in file order), a sort + unique
 * for the deduplicating reader (rows sorted).  The graph is entry for entry
 * the one kc_hg_read_coo / kc_hg_read_dedup + kc_graph_upload give.
 * stats may be NULL. */
typedef struct kc_read_stats {
    uint32_t V;
    uint32_t has_zero_id; /* COO reader: some id was 0 (no 1-based shift) */
    uint32_t multigraph;  /* COO reader: duplicates kept */
    uint32_t threads;     /* parse threads */
    uint64_t nnz;         /* adjacency entries of the resident graph */
    uint64_t edges_read;  /* edge lines kept (self-loops dropped) */
    uint64_t file_bytes;
    double parse_ms;  /* map + parse on the host */
    double build_ms;  /* upload of the pairs + device CSR build */
    double total_ms;  /* host wall time of the call */
    uint64_t reserved[4];
} kc_read_stats;
KC_API kc_status kc_graph_read(kc_ctx *ctx, const char *path, int reader, uint32_t rowptr_bytes,
                               kc_graph **out, kc_read_stats *stats);
KC_API kc_status kc_batch_gen_rmat(kc_ctx *ctx, uint32_t G, uint32_t scale_min,
                                   uint32_t scale_max, double ef_min, double ef_max, double a,
                                   double b, double c, uint64_t seed, kc_batch **out);
KC_API kc_status kc_batch_download(const kc_batch *b, kc_host_batch **out);
/* graphs with explicit (scale, generated edges, seed) each */
KC_API kc_status kc_batch_gen_rmat_list(kc_ctx *ctx, uint32_t G, const uint32_t *scale,
                                        const uint64_t *m_gen, const uint64_t *seeds, double a,
                                        double b, double c, kc_batch **out);
/* per-graph kc_core_checksum of the last kc_batch_run's cores (G entries) */
KC_API kc_status kc_batch_checksums(kc_batch *b, uint64_t *out);
/* graph g's cores from the last kc_batch_run (V_g entries) */
KC_API kc_status kc_batch_graph_cores(const kc_batch *b, uint32_t g, uint32_t *out);
KC_API kc_status kc_batch_dims(const kc_batch *b, uint32_t *G, uint64_t *vertices,
                               uint64_t *entries);

/* ------------------------------------------------------- edge updates */
/* Core maintenance under batched edge insertions and deletions (no
 * counterpart in the reference, whose graphs are static).  The first call
 * moves the graph into a slack layout on the device and peels it once; each
 * call then applies the batch and leaves the exact core numbers of the
 * updated graph on the device (core_out, V entries, may be NULL).  Pairs are
 * unordered; self-loops, repeated pairs, inserts of present edges and
 * deletes of absent edges are skipped and counted.  kc_decompose and
 * kc_graph_download see the updated graph. */
typedef enum kc_update_op { KC_INSERT = 0, KC_DELETE = 1 } kc_update_op;

typedef struct kc_update_stats {
    uint64_t requested;   /* pairs passed in */
    uint64_t applied;     /* edges inserted / deleted */
    uint64_t skipped;     /* requested - applied */
    uint64_t candidates;  /* vertices raised by one before lowering (inserts) */
    uint64_t evaluations; /* worklist entries evaluated while lowering */
    uint64_t changed;     /* vertices whose core number changed */
    uint32_t matchings;   /* vertex-disjoint sub-batches applied in turn */
    uint32_t bfs_levels;  /* candidate-search levels (inserts) */
    uint32_t rounds;      /* lowering rounds */
    uint32_t grows;       /* row re-layouts for lack of spare slots */
    double device_ms;     /* HIP-event time from first to last device step */
    double total_ms;      /* host wall time of the call */
    uint64_t reserved[4];
} kc_update_stats;

KC_API kc_status kc_update(kc_graph *g, kc_update_op op, const uint32_t *u, const uint32_t *v,
                           uint64_t n, uint32_t *core_out, kc_update_stats *stats);
/* Build the maintenance state (cores by a bucket peel, slack adjacency) now
 * instead of in the first kc_update: a server calls it when a graph becomes
 * resident, so that no request pays for it.  No reference counterpart. */
KC_API kc_status kc_update_prepare(kc_graph *g);
/* current vertex count and adjacency entries (2x undirected edges) */
KC_API kc_status kc_graph_dims(const kc_graph *g, uint32_t *V, uint64_t *nnz);
KC_API kc_status kc_graph_degree_stats(kc_graph *g, kc_degree_stats *out);
/* The k-core of a resident graph as a new resident graph on the same
 * context: the vertices whose core number from the last decomposition of g
 * (kc_decompose, or the cores kc_update maintains) is >= k, renumbered in id
 * order, with the edges among them in their original row order.  k = kmax
 * gives the kmax-core.  ids (room for g's V entries, may be NULL) receives
 * the id in g of each new vertex; *n (may be NULL) the new vertex count.
 * KC_EINVAL if g has not been decomposed. */
KC_API kc_status kc_kcore_subgraph(kc_graph *g, uint32_t k, kc_graph **out, uint32_t *ids,
                                   uint32_t *n);

/* ------------------------------------------------------------- serving */
/* A request loop on one GPU (no counterpart in the reference).  Worker
 * threads, each with its own HIP stream, admit queued requests as slots and
 * HBM budget free up; small fresh-graph decompositions admitted together are
 * fused into one batched launch; requests on a resident graph apply in
 * submission order.  Results are bit-identical to running each request
 * alone. */
typedef struct kc_server kc_server;

typedef struct kc_server_opts {
    uint32_t slots;          /* worker threads / streams (default 4) */
    uint32_t max_batch;      /* most requests fused into one launch (default 256) */
    uint32_t fuse_max_scale; /* fresh R-MAT graphs up to this scale are fused (default 16) */
    uint32_t reserved0;
    uint64_t hbm_budget;     /* bytes of HBM the requests in flight may use (default: half) */
    uint64_t reserved[4];
} kc_server_opts;

typedef enum kc_req_kind {
    KC_REQ_DECOMPOSE_RMAT = 0,  /* build R-MAT(scale, edge_factor, seed) in HBM, decompose */
    KC_REQ_UPDATE = 1,          /* kc_update on a resident graph */
    KC_REQ_DECOMPOSE_GRAPH = 2  /* decompose a resident graph */
} kc_req_kind;

typedef struct kc_request {
    uint32_t kind;  /* kc_req_kind */
    uint32_t graph; /* resident graph id (UPDATE, DECOMPOSE_GRAPH) */
    uint32_t op;    /* kc_update_op (UPDATE) */
    uint32_t scale; /* R-MAT scale (DECOMPOSE_RMAT), a,b,c = .57,.19,.19 */
    double edge_factor;
    uint64_t seed;
    uint64_t n;           /* UPDATE: pairs in u, v (copied at submit) */
    const uint32_t *u, *v;
    uint32_t *core_out;   /* optional host buffer for the V cores; caller keeps it
                             valid until the completion is polled */
    uint64_t tag;         /* caller's, returned in the completion */
    uint64_t reserved[4];
} kc_request;

typedef struct kc_completion {
    uint64_t id;
    uint64_t tag;
    int32_t status; /* kc_status */
    uint32_t kind;
    uint32_t V;
    uint32_t kmax;  /* fresh decompositions */
    uint32_t fused; /* requests in the launch that served this one */
    uint32_t reserved0;
    double submit_ms, start_ms, done_ms; /* server clock */
    uint64_t checksum; /* kc_core_checksum of the resulting core vector */
    uint64_t applied;  /* UPDATE: edges applied */
    uint64_t reserved[4];
} kc_completion;

typedef struct kc_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t jobs;            /* units of work started (a fused launch counts once) */
    uint64_t fused_jobs;      /* jobs serving more than one request */
    uint64_t fused_requests;  /* requests served by those */
    uint64_t admission_waits; /* admission stopped at a request that did not fit */
    uint64_t max_busy;        /* most workers busy at once */
    uint64_t max_inflight_bytes;
    uint64_t hbm_budget;
    uint64_t reserved[4];
} kc_server_stats;

KC_API kc_status kc_server_create(int device, const kc_server_opts *opts, kc_server **out);
/* a resident graph (R-MAT, .57/.19/.19) that UPDATE requests maintain */
KC_API kc_status kc_server_gen_rmat_graph(kc_server *s, uint32_t scale, double edge_factor,
                                          uint64_t seed, uint32_t *graph_id);
/* host copy of a resident graph once its queued requests are done */
KC_API kc_status kc_server_graph_download(kc_server *s, uint32_t graph_id,
                                          kc_host_graph **out);
KC_API kc_status kc_server_submit(kc_server *s, const kc_request *r, uint64_t *id);
/* up to max completions; waits up to wait_ms (-1: until one is available or
 * nothing is outstanding, 0: no wait) */
KC_API kc_status kc_server_poll(kc_server *s, kc_completion *out, uint32_t max, uint32_t *n,
                                int wait_ms);
KC_API kc_status kc_server_drain(kc_server *s);
KC_API double kc_server_now_ms(const kc_server *s);
KC_API kc_status kc_server_get_stats(kc_server *s, kc_server_stats *out);
/* waits for outstanding requests, then frees everything */
KC_API void kc_server_destroy(kc_server *s);

/* order-independent checksum of a core vector, sum over v of
 * splitmix64(v << 32 | core[v]) mod 2^64 (host computation) */
KC_API uint64_t kc_core_checksum(const uint32_t *core, uint32_t V);

/* ---------------------------------------------------------- calibration */
/* Streaming device copy of `bytes` (two buffers), `reps` timed launches;
 * *ms = mean time of one.  Reads and writes `bytes` each: the reference
 * point for achievable HBM bandwidth and for scaling HBM byte counters. */
KC_API kc_status kc_probe_copy(kc_ctx *ctx, uint64_t bytes, uint32_t reps, double *ms);

/* ------------------------------------------------------------ autotuner */
/* Time the launch candidates of `method` on graph g (block size, virtual
 * node split width, local frontier capacity for the peels; block size for
 * the h-index methods), `reps` runs each, and keep the fastest as the
 * graph's default for that method; *best receives it. */
KC_API kc_status kc_autotune(kc_graph *g, kc_method method, uint32_t reps, kc_opts *best,
                             double *best_ms);

#ifdef __cplusplus
}
#endif
#endif
