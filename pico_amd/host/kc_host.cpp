// kc_host.cpp -- host-side graph construction: synthetic generators for the
// benchmark configurations, the symmetrize/dedup CSR builder, the two edge
// list readers of the reference and the example.txt-dialect writer.
#include "kc_host.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <unordered_set>

namespace kch {

unsigned worker_threads()
{
    const char *e = std::getenv("KC_HOST_THREADS");
    unsigned n = e ? (unsigned)std::atoi(e) : std::thread::hardware_concurrency();
    if (n == 0)
        n = 1;
    return n > 16 ? 16 : n; // the GPU hosts give one GPU a 16-core share
}

static thread_local bool tl_serial = false;
void set_serial(bool on) { tl_serial = on; }
bool serial() { return tl_serial; }

// ---------------------------------------------------------------- CSR -----
// Symmetrize, drop self-loops, dedup and sort every row.  Row order of the
// result does not depend on the thread count (rows are sorted).
int csr_from_pairs(uint32_t V, uint64_t n, const uint32_t *src, const uint32_t *dst,
                   kc_host_graph &g)
{
    std::vector<uint64_t> cnt((size_t)V + 1, 0);
    parallel_for(n, 1 << 20, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t u = src[i], v = dst[i];
            if (u == v)
                continue;
            __atomic_fetch_add(&cnt[u], 1, __ATOMIC_RELAXED);
            __atomic_fetch_add(&cnt[v], 1, __ATOMIC_RELAXED);
        }
    });
    std::vector<uint64_t> off((size_t)V + 1, 0);
    for (uint32_t v = 0; v < V; v++)
        off[v + 1] = off[v] + cnt[v];
    std::vector<uint32_t> tmp(off[V]);
    std::vector<uint64_t> cur(off.begin(), off.end() - 1);
    parallel_for(n, 1 << 20, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t u = src[i], v = dst[i];
            if (u == v)
                continue;
            tmp[__atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED)] = v;
            tmp[__atomic_fetch_add(&cur[v], 1, __ATOMIC_RELAXED)] = u;
        }
    });
    std::vector<uint32_t> ud(V, 0);
    parallel_for(V, 4096, [&](uint64_t b, uint64_t e) {
        for (uint64_t v = b; v < e; v++) {
            uint32_t *r0 = tmp.data() + off[v], *r1 = tmp.data() + off[v + 1];
            std::sort(r0, r1);
            ud[v] = (uint32_t)(std::unique(r0, r1) - r0);
        }
    });
    g.V = V;
    g.rowptr.assign((size_t)V + 1, 0);
    for (uint32_t v = 0; v < V; v++)
        g.rowptr[v + 1] = g.rowptr[v] + ud[v];
    g.col.resize(g.rowptr[V]);
    parallel_for(V, 4096, [&](uint64_t b, uint64_t e) {
        for (uint64_t v = b; v < e; v++)
            std::memcpy(g.col.data() + g.rowptr[v], tmp.data() + off[v], ud[v] * sizeof(uint32_t));
    });
    g.multigraph = false;
    return 0;
}

// ---------------------------------------------------------- generators -----
// Erdos-Renyi G(n, m): exactly m distinct unordered pairs u != v drawn
// uniformly without replacement from one seeded stream.
int gen_er(uint32_t n, uint64_t m, uint64_t seed, kc_host_graph &g)
{
    if (n < 2 || m > (uint64_t)n * (n - 1) / 2)
        return -1;
    std::unordered_set<uint64_t> seen;
    seen.reserve(m * 2);
    std::vector<uint32_t> s, d;
    s.reserve(m);
    d.reserve(m);
    uint64_t key = mix64(seed ^ 0x45522d6772617068ull);
    for (uint64_t i = 0; s.size() < m; i++) {
        uint64_t h = mix64(key ^ (i * 0xD1B54A32D192ED03ull));
        uint32_t u = (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
        uint32_t v = (uint32_t)(((h & 0xffffffffull) * (uint64_t)n) >> 32);
        if (u == v)
            continue;
        if (u > v)
            std::swap(u, v);
        if (seen.insert(((uint64_t)u << 32) | v).second) {
            s.push_back(u);
            d.push_back(v);
        }
    }
    return csr_from_pairs(n, m, s.data(), d.data(), g);
}

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004): per edge and per bit level a
// quadrant draw with probabilities (a, b, c, 1-a-b-c); no noise, no
// relabelling.  Edge i, level l uses hash(seed, i, l), so the edge list does
// not depend on threading.
int gen_rmat(uint32_t scale, uint64_t m_gen, double a, double b, double c, uint64_t seed,
             kc_host_graph &g)
{
    if (scale == 0 || scale > 31)
        return -1;
    uint32_t V = 1u << scale;
    std::vector<uint32_t> s(m_gen), d(m_gen);
    uint64_t key = mix64(seed ^ 0x524d41542d67656eull);
    double ab = a + b, abc = a + b + c;
    parallel_for(m_gen, 1 << 16, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; i++) {
            uint32_t u = 0, v = 0;
            uint64_t base = key ^ (i * 0xD1B54A32D192ED03ull);
            for (uint32_t l = 0; l < scale; l++) {
                double r = u01(mix64(base + l * 0x9E3779B97F4A7C15ull));
                uint32_t bu = r >= ab;                      // c or d
                uint32_t bv = (r >= a && r < ab) || r >= abc; // b or d
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            s[i] = u;
            d[i] = v;
        }
    });
    return csr_from_pairs(V, m_gen, s.data(), d.data(), g);
}

// Chung-Lu with power-law weights w_i ~ (i+1)^(-1/(beta-1)), scaled to the
// requested mean degree and capped; n*mean/2 endpoint pairs drawn
// independently proportional to w (Vose alias table), then dedup/loops.
int chunglu_alias(uint32_t n, double beta, double mean_deg, double cap,
                  std::vector<uint32_t> &thr, std::vector<uint32_t> &alias, uint64_t &m_gen)
{
    if (n < 2 || beta <= 1.0)
        return -1;
    double alpha = 1.0 / (beta - 1.0);
    std::vector<double> w(n);
    for (uint32_t i = 0; i < n; i++)
        w[i] = std::pow((double)i + 1.0, -alpha);
    // scale so that the capped weights have the requested mean
    double target = mean_deg * (double)n;
    double scale = 1.0;
    for (int it = 0; it < 60; it++) {
        double sum = 0;
        for (uint32_t i = 0; i < n; i++)
            sum += std::min(cap, w[i] * scale);
        scale *= target / sum;
    }
    double total = 0;
    for (uint32_t i = 0; i < n; i++) {
        w[i] = std::min(cap, w[i] * scale);
        total += w[i];
    }
    // Vose alias table
    alias.assign(n, 0);
    thr.assign(n, 0);
    std::vector<double> p(n);
    std::vector<uint32_t> small, large;
    for (uint32_t i = 0; i < n; i++) {
        p[i] = w[i] * n / total;
        (p[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t l = small.back(), gi = large.back();
        small.pop_back();
        alias[l] = gi;
        thr[l] = (uint32_t)std::min(4294967295.0, p[l] * 4294967296.0);
        p[gi] = (p[gi] + p[l]) - 1.0;
        if (p[gi] < 1.0) {
            large.pop_back();
            small.push_back(gi);
        }
    }
    for (uint32_t i : large) {
        thr[i] = 0xffffffffu;
        alias[i] = i;
    }
    for (uint32_t i : small) {
        thr[i] = 0xffffffffu;
        alias[i] = i;
    }
    m_gen = (uint64_t)std::llround(mean_deg * (double)n / 2.0);
    return 0;
}

int gen_chunglu(uint32_t n, double beta, double mean_deg, double cap, uint64_t seed,
                kc_host_graph &g)
{
    std::vector<uint32_t> thr, alias;
    uint64_t m_gen = 0;
    if (chunglu_alias(n, beta, mean_deg, cap, thr, alias, m_gen))
        return -1;
    std::vector<uint32_t> s(m_gen), d(m_gen);
    uint64_t key = mix64(seed ^ 0x4368756e674c7521ull);
    auto draw = [&](uint64_t h) -> uint32_t {
        uint32_t i = (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
        return ((uint32_t)h < thr[i]) ? i : alias[i];
    };
    parallel_for(m_gen, 1 << 16, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; i++) {
            uint64_t base = key ^ (i * 0xD1B54A32D192ED03ull);
            s[i] = draw(mix64(base));
            d[i] = draw(mix64(base + 0x9E3779B97F4A7C15ull));
        }
    });
    return csr_from_pairs(n, m_gen, s.data(), d.data(), g);
}

// -------------------------------------------------------------- batches -----
int batch_from_graphs(std::vector<kc_host_graph> &gs, kc_host_batch &out, bool release)
{
    uint32_t G = (uint32_t)gs.size();
    out.G = G;
    out.vtx_off.assign((size_t)G + 1, 0);
    out.edge_off.assign((size_t)G + 1, 0);
    for (uint32_t g = 0; g < G; g++) {
        if (gs[g].rowptr.size() != (size_t)gs[g].V + 1 || gs[g].rowptr[gs[g].V] > 0xffffffffull)
            return -1;
        out.vtx_off[g + 1] = out.vtx_off[g] + gs[g].V;
        out.edge_off[g + 1] = out.edge_off[g] + gs[g].rowptr[gs[g].V];
    }
    out.rowptr.resize(out.vtx_off[G] + G);
    out.col.resize(out.edge_off[G]);
    parallel_for(G, 1, [&](uint64_t b, uint64_t e) {
        for (uint64_t g = b; g < e; g++) {
            kc_host_graph &h = gs[g];
            uint32_t *rp = out.rowptr.data() + out.vtx_off[g] + g;
            for (uint64_t v = 0; v <= h.V; v++)
                rp[v] = (uint32_t)h.rowptr[v];
            if (!h.col.empty())
                std::memcpy(out.col.data() + out.edge_off[g], h.col.data(), h.col.size() * 4);
            if (release) {
                std::vector<uint64_t>().swap(h.rowptr);
                std::vector<uint32_t>().swap(h.col);
            }
        }
    });
    return 0;
}

void rmat_batch_draws(uint32_t G, uint32_t scale_min, uint32_t scale_max, double ef_min,
                      double ef_max, uint64_t seed, std::vector<uint32_t> &sc,
                      std::vector<uint64_t> &mg)
{
    sc.resize(G);
    mg.resize(G);
    uint32_t span = scale_max - scale_min + 1;
    uint64_t key = mix64(seed ^ 0x62617463682d726dull);
    for (uint32_t g = 0; g < G; g++) {
        uint64_t h = mix64(key ^ ((uint64_t)g * 0xD1B54A32D192ED03ull));
        sc[g] = scale_min + (uint32_t)(((h >> 32) * (uint64_t)span) >> 32);
        double ef = ef_min + u01(mix64(h)) * (ef_max - ef_min);
        mg[g] = (uint64_t)std::llround(ef * (double)(1ull << sc[g]));
    }
}

int gen_rmat_batch(uint32_t G, uint32_t scale_min, uint32_t scale_max, double ef_min,
                   double ef_max, double a, double b, double c, uint64_t seed, kc_host_batch &out)
{
    if (scale_min == 0 || scale_max > 31 || scale_min > scale_max || ef_min < 0 ||
        ef_max < ef_min)
        return -1;
    std::vector<uint32_t> sc;
    std::vector<uint64_t> mg;
    rmat_batch_draws(G, scale_min, scale_max, ef_min, ef_max, seed, sc, mg);
    // largest first so the workers finish together
    std::vector<uint32_t> ord(G);
    for (uint32_t g = 0; g < G; g++)
        ord[g] = g;
    std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return mg[x] > mg[y]; });
    std::vector<kc_host_graph> gs(G);
    std::atomic<int> bad{0};
    parallel_for(G, 1, [&](uint64_t lo, uint64_t hi) {
        set_serial(true);
        for (uint64_t i = lo; i < hi; i++) {
            uint32_t g = ord[i];
            if (gen_rmat(sc[g], mg[g], a, b, c, seed + g, gs[g]))
                bad = 1;
        }
        set_serial(false);
    });
    if (bad)
        return -1;
    if (batch_from_graphs(gs, out, true))
        return -1;
    out.scale = sc;
    out.m_gen = mg;
    return 0;
}

uint64_t core_checksum(const uint32_t *core, uint32_t V)
{
    uint64_t h = 0;
    for (uint32_t v = 0; v < V; v++)
        h += mix64(((uint64_t)v << 32) | core[v]);
    return h;
}

// ------------------------------------------------------------- readers -----
static bool slurp(const char *path, std::vector<char> &buf, std::string &err)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path;
        return false;
    }
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize((size_t)(sz > 0 ? sz : 0) + 1);
    size_t rd = sz > 0 ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
    std::fclose(f);
    buf.resize(rd);
    return true;
}

static inline bool is_ws(char c)
{
    return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f' || c == '\n';
}

// Extraction of one unsigned value the way `std::istream >> unsigned` does
// it in the C++11 library: skip whitespace; at end of line the stream fails
// and the target keeps its old value; a field without digits stores 0 and
// fails; '-' negates modulo 2^32; overflow stores UINT_MAX and fails.
// Returns true when the stream is still good afterwards.
static bool extract_u32(const char *&p, const char *end, uint32_t &val)
{
    while (p < end && is_ws(*p))
        p++;
    if (p >= end)
        return false; // sentry fails: value untouched
    bool neg = false;
    if (*p == '-' || *p == '+') {
        neg = *p == '-';
        p++;
    }
    if (p >= end || *p < '0' || *p > '9') {
        val = 0;
        return false;
    }
    uint64_t x = 0;
    bool ovf = false;
    while (p < end && *p >= '0' && *p <= '9') {
        x = x * 10 + (uint64_t)(*p - '0');
        if (x > 0xffffffffull) {
            ovf = true;
            x = 0xffffffffull + 1;
        }
        p++;
    }
    if (ovf) {
        val = 0xffffffffu;
        return false;
    }
    val = neg ? (uint32_t)(0u - (uint32_t)x) : (uint32_t)x;
    return true;
}

// Line splitter shared by both readers: calls fn(line_begin, line_end) for
// every line after the first two (both reference readers discard exactly
// two lines: shared/graph.cpp:47-48, src/graph.cpp:78-79).
template <typename F> static void for_data_lines(const std::vector<char> &buf, F &&fn)
{
    const char *p = buf.data(), *end = p + buf.size();
    int skipped = 0;
    while (p < end) {
        const char *nl = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        const char *le = nl ? nl : end;
        if (skipped < 2)
            skipped++;
        else
            fn(p, le);
        p = nl ? nl + 1 : end;
    }
}

// Reader of gpp / peelone / HistoCore / CntCore / NbrCore
// (shared/graph.cpp:17-99 + virtual_graph.cpp:9-43, 93-157): keeps duplicate
// edges, drops self-loops, num_nodes = max id (+1 when some id is 0), shifts
// to 0-based when no id is 0, neighbours in edge-insertion order.
int read_coo(const char *path, kc_host_graph &g, std::string &err)
{
    std::vector<char> buf;
    if (!slurp(path, buf, err))
        return -1;
    std::vector<uint32_t> es, ed;
    uint32_t s = 0, t = 0, mx = 0;
    bool zero = false;
    for_data_lines(buf, [&](const char *p, const char *le) {
        if (extract_u32(p, le, s))
            extract_u32(p, le, t);
        if (s == t)
            return;
        es.push_back(s);
        ed.push_back(t);
        zero |= (s == 0) | (t == 0);
        mx = std::max(mx, std::max(s, t));
    });
    uint64_t m = es.size();
    uint32_t V = mx + (zero ? 1u : 0u);
    if (!zero)
        for (uint64_t i = 0; i < m; i++) {
            es[i]--;
            ed[i]--;
        }
    g.V = V;
    g.edges_read = m;
    g.has_zero_id = zero;
    g.multigraph = true;
    g.rowptr.assign((size_t)V + 1, 0);
    for (uint64_t i = 0; i < m; i++) {
        g.rowptr[es[i] + 1]++;
        g.rowptr[ed[i] + 1]++;
    }
    for (uint32_t v = 0; v < V; v++)
        g.rowptr[v + 1] += g.rowptr[v];
    g.col.resize(g.rowptr[V]);
    std::vector<uint64_t> cur(g.rowptr.begin(), g.rowptr.end() - 1);
    for (uint64_t i = 0; i < m; i++) {
        g.col[cur[es[i]]++] = ed[i];
        g.col[cur[ed[i]]++] = es[i];
    }
    return 0;
}

// Reader of peelone_dyn (src/graph.cpp:58-133): always subtracts 1 from both
// ids, keeps the shifted values across lines (a line that fails to parse
// re-uses and re-decrements them), drops self-loops, dedups with sorted
// neighbour sets, V = max id + 1.  Inputs with a 0 id wrap to UINT_MAX in the
// reference and index out of bounds; they are rejected here.
int read_dedup(const char *path, kc_host_graph &g, std::string &err)
{
    std::vector<char> buf;
    if (!slurp(path, buf, err))
        return -1;
    std::vector<uint32_t> es, ed;
    uint32_t s = 0, t = 0, val = 0, V = 0;
    uint64_t lines = 0;
    for_data_lines(buf, [&](const char *p, const char *le) {
        if (extract_u32(p, le, s))
            if (extract_u32(p, le, t))
                extract_u32(p, le, val);
        s = s - 1;
        t = t - 1;
        if (s == t)
            return;
        V = std::max(V, std::max(s, t));
        es.push_back(s);
        ed.push_back(t);
        lines++;
    });
    if (V == 0xffffffffu) {
        err = "the deduplicating reader needs 1-based vertex ids (found id 0)";
        return -1;
    }
    uint32_t nv = V + 1u; // V++ of src/graph.cpp:102
    int r = csr_from_pairs(nv, es.size(), es.data(), ed.data(), g);
    g.edges_read = lines;
    return r;
}

int write_edgelist(const kc_host_graph &g, const char *path, std::string &err)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        err = std::string("cannot create ") + path;
        return -1;
    }
    uint64_t m = 0;
    for (uint32_t v = 0; v < g.V; v++)
        for (uint64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; e++)
            m += g.col[e] > v;
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate integer symmetric\n%u %u %llu\n", g.V, g.V,
                 (unsigned long long)m);
    std::vector<char> out;
    out.reserve(1 << 22);
    char tmp[32];
    for (uint32_t v = 0; v < g.V; v++) {
        for (uint64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; e++) {
            uint32_t u = g.col[e];
            if (u <= v)
                continue;
            int k = std::snprintf(tmp, sizeof tmp, "%u %u 1\n", v + 1, u + 1);
            out.insert(out.end(), tmp, tmp + k);
        }
        if (out.size() > (1 << 22) - 64) {
            std::fwrite(out.data(), 1, out.size(), f);
            out.clear();
        }
    }
    std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    return 0;
}

} // namespace kch
