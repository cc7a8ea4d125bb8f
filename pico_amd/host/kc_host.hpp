// kc_host.hpp -- host-side graph objects shared by the readers, generators,
// CSR builder and the C-ABI.  Plain C++ (no HIP).
#pragma once
#include <stdint.h>

#include <atomic>
#include <string>
#include <thread>
#include <vector>

struct kc_host_graph {
    uint32_t V = 0;
    std::vector<uint64_t> rowptr; // V+1
    std::vector<uint32_t> col;    // rowptr[V]
    // reader provenance (0 when not read from a file)
    uint64_t edges_read = 0; // edge lines kept by the reader (self-loops dropped)
    bool has_zero_id = false;
    bool multigraph = false; // COO reader keeps duplicate edges
};

// Many independent graphs in one concatenated int32 CSR: graph g owns
// vertices [vtx_off[g], vtx_off[g+1]) and adjacency entries
// [edge_off[g], edge_off[g+1]); its V_g+1 local row offsets start at
// rowptr[vtx_off[g] + g] and its col entries are local ids in [0, V_g).
struct kc_host_batch {
    uint32_t G = 0;
    std::vector<uint64_t> vtx_off, edge_off; // G+1 each
    std::vector<uint32_t> rowptr;            // vtx_off[G] + G
    std::vector<uint32_t> col;               // edge_off[G]
    std::vector<uint32_t> scale;             // generator draws (empty when built from arrays)
    std::vector<uint64_t> m_gen;
};

namespace kch {

unsigned worker_threads();
// while set on this thread, parallel_for runs inline (outer loops that are
// already parallel over independent work items)
void set_serial(bool on);
bool serial();

// parallel-for over [0, n) in chunks of `chunk`, on worker_threads() threads
template <typename F> void parallel_for(uint64_t n, uint64_t chunk, F &&f)
{
    unsigned T = serial() ? 1 : worker_threads();
    if (n <= chunk || T == 1) {
        for (uint64_t i = 0; i < n; i += chunk)
            f(i, n < i + chunk ? n : i + chunk);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&] {
            for (;;) {
                uint64_t b = next.fetch_add(chunk);
                if (b >= n)
                    break;
                f(b, n < b + chunk ? n : b + chunk);
            }
        });
    for (auto &x : th)
        x.join();
}


// counter-based RNG (splitmix64 finaliser)
static inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// symmetrize + drop self-loops + dedup + sort each row
int csr_from_pairs(uint32_t V, uint64_t n, const uint32_t *src, const uint32_t *dst,
                   kc_host_graph &g);

int gen_er(uint32_t n, uint64_t m, uint64_t seed, kc_host_graph &g);

// Binary CSR files (kc_cache.cpp).  `reader` records which reader produced
// the graph (KC_NO_READER for none); src_hash / src_size identify the edge
// file it was read from.
constexpr uint32_t KC_NO_READER = 0xffffffffu;
int file_hash(const char *path, uint64_t &hash, uint64_t &size, std::string &err);
int save_csr(const kc_host_graph &g, const char *path, uint32_t reader, uint64_t src_hash,
             uint64_t src_size, std::string &err);
int load_csr(const char *path, kc_host_graph &g, uint32_t *reader, uint64_t *src_hash,
             uint64_t *src_size, std::string &err);
int gen_rmat(uint32_t scale, uint64_t m_gen, double a, double b, double c, uint64_t seed,
             kc_host_graph &g);
int gen_chunglu(uint32_t n, double beta, double mean_deg, double cap, uint64_t seed,
                kc_host_graph &g);
// the Chung-Lu generator's alias table (32-bit thresholds) and pair count
int chunglu_alias(uint32_t n, double beta, double mean_deg, double cap,
                  std::vector<uint32_t> &thr, std::vector<uint32_t> &alias, uint64_t &m_gen);
// per-graph (scale, generated pairs) of an R-MAT batch
void rmat_batch_draws(uint32_t G, uint32_t scale_min, uint32_t scale_max, double ef_min,
                      double ef_max, uint64_t seed, std::vector<uint32_t> &sc,
                      std::vector<uint64_t> &mg);

// 4096-style batch of R-MAT graphs: graph g uses seed + g, scale drawn
// uniformly from the integers [scale_min, scale_max] and edge factor
// uniformly from [ef_min, ef_max] (both from a hash of seed and g).
int gen_rmat_batch(uint32_t G, uint32_t scale_min, uint32_t scale_max, double ef_min,
                   double ef_max, double a, double b, double c, uint64_t seed, kc_host_batch &out);
// concatenate independent host graphs into the batch layout; with release
// each graph's arrays are freed once copied
int batch_from_graphs(std::vector<kc_host_graph> &gs, kc_host_batch &out, bool release);

// the device checksum of a core vector, computed on the host
uint64_t core_checksum(const uint32_t *core, uint32_t V);

// An edge file's kept edges in file order, after the reader's id shift.
struct EdgeFile {
    uint32_t V = 0;
    std::vector<uint32_t> src, dst;
    bool has_zero_id = false;
    uint64_t file_bytes = 0;
    unsigned threads = 1; // parse threads used
};
// The line loops of the two readers, parsed on worker threads (results as
// one sequential pass); 0 ok, -1 file error, -2 malformed (ids out of range).
int read_coo_pairs(const char *path, EdgeFile &ef, std::string &err);
int read_dedup_pairs(const char *path, EdgeFile &ef, std::string &err);

int read_coo(const char *path, kc_host_graph &g, std::string &err);
int read_dedup(const char *path, kc_host_graph &g, std::string &err);
int write_edgelist(const kc_host_graph &g, const char *path, std::string &err);

} // namespace kch
