// Host-only self test of libkcore's host C++ (readers with their chunked
// parallel parse, CSR builder, generators, binary CSR files, edge-list
// writer), built with AddressSanitizer + UndefinedBehaviorSanitizer by
// tests/test_host_sanitizers.py.  Prints one line per input: V, nnz and a
// hash of the CSR, which the test compares with libkcore's own results.
//
//   host_selftest coo|dedup <edge file>       read and print
//   host_selftest gen <workdir>               generators + save/load + write/read
#include "../../pico_amd/host/kc_host.hpp"

#include <cstdio>
#include <cstring>
#include <string>

static uint64_t csr_hash(const kc_host_graph &g)
{
    uint64_t h = 1469598103934665603ull;
    for (uint64_t x : g.rowptr)
        h = (h ^ x) * 1099511628211ull;
    for (uint32_t x : g.col)
        h = (h ^ x) * 1099511628211ull;
    return h;
}

static int show(const char *what, const kc_host_graph &g)
{
    std::printf("%s V=%u nnz=%llu read=%llu zero=%d hash=%016llx\n", what, g.V,
                (unsigned long long)(g.rowptr.empty() ? 0 : g.rowptr.back()),
                (unsigned long long)g.edges_read, g.has_zero_id ? 1 : 0,
                (unsigned long long)csr_hash(g));
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3)
        return 2;
    std::string err;
    if (!std::strcmp(argv[1], "coo") || !std::strcmp(argv[1], "dedup")) {
        kc_host_graph g;
        int r = !std::strcmp(argv[1], "coo") ? kch::read_coo(argv[2], g, err)
                                             : kch::read_dedup(argv[2], g, err);
        if (r) {
            std::printf("%s error %d\n", argv[1], r);
            return 0;
        }
        return show(argv[1], g);
    }
    if (!std::strcmp(argv[1], "gen")) {
        std::string wd = argv[2];
        kc_host_graph a, b, c, d;
        if (kch::gen_rmat(12, 16 << 12, 0.57, 0.19, 0.19, 7, a) ||
            kch::gen_chunglu(1 << 12, 2.1, 20.0, 2000.0, 3, b) || kch::gen_er(500, 2000, 1, c))
            return 3;
        show("rmat", a);
        show("chunglu", b);
        show("er", c);
        std::string f = wd + "/g.kcsr";
        if (kch::save_csr(a, f.c_str(), kch::KC_NO_READER, 0, 0, err) ||
            kch::load_csr(f.c_str(), d, nullptr, nullptr, nullptr, err))
            return 4;
        show("load", d);
        std::string t = wd + "/g.txt";
        if (kch::write_edgelist(a, t.c_str(), err))
            return 5;
        kc_host_graph e;
        if (kch::read_dedup(t.c_str(), e, err))
            return 6;
        show("reread", e);
        std::vector<uint32_t> s(a.col.size()), q(a.col.size());
        for (uint32_t v = 0; v < a.V; v++)
            for (uint64_t x = a.rowptr[v]; x < a.rowptr[v + 1]; x++) {
                s[x] = v;
                q[x] = a.col[x];
            }
        kc_host_graph p;
        kch::csr_from_pairs(a.V, s.size(), s.data(), q.data(), p);
        show("pairs", p);
        return 0;
    }
    return 2;
}
