// kc_host.hpp -- host-side graph objects shared by the readers, generators,
// CSR builder and the C-ABI.  Plain C++ (no HIP).
#pragma once
#include <stdint.h>

#include <string>
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

namespace kch {

unsigned worker_threads();

// parallel-for over [0, n) in chunks
template <typename F> void parallel_for(uint64_t n, uint64_t chunk, F &&f);

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
int gen_rmat(uint32_t scale, uint64_t m_gen, double a, double b, double c, uint64_t seed,
             kc_host_graph &g);
int gen_chunglu(uint32_t n, double beta, double mean_deg, double cap, uint64_t seed,
                kc_host_graph &g);

int read_coo(const char *path, kc_host_graph &g, std::string &err);
int read_dedup(const char *path, kc_host_graph &g, std::string &err);
int write_edgelist(const kc_host_graph &g, const char *path, std::string &err);

} // namespace kch
