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

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

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
            if (ud[v])
                std::memcpy(g.col.data() + g.rowptr[v], tmp.data() + off[v],
                            ud[v] * sizeof(uint32_t));
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
// The file image: memory-mapped where the file allows it, so that the parse
// threads fault its pages in side by side; read into memory otherwise.
struct FileImage {
    const char *p = nullptr;
    size_t n = 0;
    void *map = nullptr;
    std::vector<char> buf;
    FileImage() = default;
    FileImage(const FileImage &) = delete;
    FileImage &operator=(const FileImage &) = delete;
    ~FileImage()
    {
        if (map)
            munmap(map, n);
    }
};

static bool open_image(const char *path, FileImage &im, std::string &err)
{
    int fd = ::open(path, O_RDONLY);
    if (fd < 0) {
        err = std::string("cannot open ") + path;
        return false;
    }
    struct stat sb;
    if (fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode) && sb.st_size > 0) {
        void *m = mmap(nullptr, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m != MAP_FAILED) {
            ::close(fd);
            im.map = m;
            im.p = (const char *)m;
            im.n = (size_t)sb.st_size;
            return true;
        }
    }
    char tmp[1 << 16];
    for (;;) {
        ssize_t r = ::read(fd, tmp, sizeof tmp);
        if (r < 0) {
            ::close(fd);
            err = std::string("cannot read ") + path;
            return false;
        }
        if (r == 0)
            break;
        im.buf.insert(im.buf.end(), tmp, tmp + r);
    }
    ::close(fd);
    im.p = im.buf.data();
    im.n = im.buf.size();
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
// Returns 0 (target untouched, stream failed), 1 (target written, stream
// failed) or 2 (target written, stream still good).
static int extract_u32(const char *&p, const char *end, uint32_t &val)
{
    while (p < end && is_ws(*p))
        p++;
    if (p >= end)
        return 0; // sentry fails: value untouched
    bool neg = false;
    if (*p == '-' || *p == '+') {
        neg = *p == '-';
        p++;
    }
    if (p >= end || *p < '0' || *p > '9') {
        val = 0;
        return 1;
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
        return 1;
    }
    val = neg ? (uint32_t)(0u - (uint32_t)x) : (uint32_t)x;
    return 2;
}

// What one data line does to the reader's two carried values.  Both
// reference readers extract the source id, and the target id only when that
// succeeded (shared/graph.cpp:50-55, src/graph.cpp:82-88); a value the line
// does not write keeps the previous line's.
struct LineFx {
    uint32_t vs = 0, vt = 0;
    uint8_t ws = 0, wt = 0; // the line writes s / t
};

static inline LineFx line_fx(const char *p, const char *le)
{
    LineFx f;
    int r = extract_u32(p, le, f.vs);
    f.ws = r != 0;
    if (r == 2)
        f.wt = extract_u32(p, le, f.vt) != 0;
    return f;
}

// The line loop of both readers, run on worker threads over chunks of whole
// lines, with results identical to one pass over the file.
//  * COO reader (dedup = false): s, t carry over as written; a line keeps
//    the edge (s, t) unless s == t.
//  * dedup reader (dedup = true): after the extraction both values are
//    decremented on every line, written or not (src/graph.cpp:89-90), so a
//    value carried over is decremented again.
// A chunk's lines until both values have been written by a line of the chunk
// depend on the state the previous chunk leaves; they are kept as line
// effects and replayed, chunk by chunk in file order, once every chunk is
// parsed (usually the chunk's first line writes both and nothing waits).
struct ParsedEdges {
    std::vector<uint32_t> es, ed; // kept edges in file order
    uint32_t mx = 0;              // largest id of a kept edge
    bool zero = false;            // some kept edge has a 0 id
    unsigned threads = 1;
};

namespace {
struct ChunkParse {
    std::vector<LineFx> wait; // lines replayed with the incoming state
    std::vector<uint32_t> es, ed;
    std::vector<uint32_t> pes, ped; // edges of the replayed lines
    bool ks = false, kt = false;
    uint32_t s = 0, t = 0;
    uint32_t mx = 0;
    bool zero = false;
};

struct EdgeSink {
    std::vector<uint32_t> &es, &ed;
    uint32_t &mx;
    bool &zero;
    inline void line(uint32_t s, uint32_t t)
    {
        if (s == t)
            return;
        es.push_back(s);
        ed.push_back(t);
        zero |= (s == 0) | (t == 0);
        uint32_t m = s > t ? s : t;
        mx = m > mx ? m : mx;
    }
};
} // namespace

static inline void apply_fx(const LineFx &f, bool dedup, uint32_t &s, uint32_t &t)
{
    if (f.ws)
        s = f.vs;
    if (f.wt)
        t = f.vt;
    if (dedup) {
        s -= 1;
        t -= 1;
    }
}

static void parse_chunk(const char *p, const char *end, bool dedup, ChunkParse &c)
{
    EdgeSink sink{c.es, c.ed, c.mx, c.zero};
    while (p < end) {
        const char *nl = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        const char *le = nl ? nl : end;
        LineFx f = line_fx(p, le);
        if (c.ks && c.kt) {
            apply_fx(f, dedup, c.s, c.t);
            sink.line(c.s, c.t);
        } else {
            c.wait.push_back(f);
            c.ks |= f.ws != 0;
            c.kt |= f.wt != 0;
            if (c.ks && c.kt) {
                // the values after this line no longer depend on the
                // incoming state: replay the waiting lines from any state
                uint32_t s = 0, t = 0;
                for (const LineFx &w : c.wait)
                    apply_fx(w, dedup, s, t);
                c.s = s;
                c.t = t;
            }
        }
        p = nl ? nl + 1 : end;
    }
}

static bool parse_edges(const char *path, bool dedup, ParsedEdges &out, uint64_t *bytes,
                        std::string &err)
{
    FileImage im;
    if (!open_image(path, im, err))
        return false;
    if (bytes)
        *bytes = im.n;
    // both reference readers discard exactly two lines first
    // (shared/graph.cpp:47-48, src/graph.cpp:78-79)
    const char *p = im.p, *end = im.p + im.n;
    for (int k = 0; k < 2 && p < end; k++) {
        const char *nl = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        p = nl ? nl + 1 : end;
    }
    const uint64_t n = (uint64_t)(end - p);
    unsigned T = serial() ? 1 : worker_threads();
    // chunks of at least 1 MiB, a few per thread for balance; KC_PARSE_CHUNK
    // (bytes) sets the chunk size itself, uncapped (tests cut small files
    // into many chunks with it)
    uint64_t nch;
    const char *ce = std::getenv("KC_PARSE_CHUNK");
    if (ce && std::strtoull(ce, nullptr, 10) > 0)
        nch = std::max<uint64_t>(1, n / std::strtoull(ce, nullptr, 10));
    else
        nch = std::max<uint64_t>(1, std::min<uint64_t>(n >> 20, 4ull * T));
    std::vector<const char *> cut(nch + 1);
    cut[0] = p;
    cut[nch] = end;
    for (uint64_t k = 1; k < nch; k++) {
        const char *q = p + n * k / nch;
        if (q < cut[k - 1])
            q = cut[k - 1];
        const char *nl = (const char *)std::memchr(q, '\n', (size_t)(end - q));
        cut[k] = nl ? nl + 1 : end;
    }
    std::vector<ChunkParse> ch(nch);
    parallel_for(nch, 1, [&](uint64_t b, uint64_t e) {
        for (uint64_t k = b; k < e; k++)
            parse_chunk(cut[k], cut[k + 1], dedup, ch[k]);
    });
    out.threads = nch > 1 ? std::min<unsigned>(T, (unsigned)nch) : 1u;
    // replay the waiting lines in file order; both readers start from s = t = 0
    uint32_t s = 0, t = 0;
    uint64_t total = 0;
    for (auto &c : ch) {
        EdgeSink sink{c.pes, c.ped, c.mx, c.zero};
        for (const LineFx &w : c.wait) {
            apply_fx(w, dedup, s, t);
            sink.line(s, t);
        }
        if (c.ks && c.kt) {
            s = c.s;
            t = c.t;
        }
        total += c.pes.size() + c.es.size();
        out.zero |= c.zero;
        out.mx = std::max(out.mx, c.mx);
    }
    out.es.resize(total);
    out.ed.resize(total);
    std::vector<uint64_t> pos(nch + 1, 0);
    for (uint64_t k = 0; k < nch; k++)
        pos[k + 1] = pos[k] + ch[k].pes.size() + ch[k].es.size();
    parallel_for(nch, 1, [&](uint64_t b, uint64_t e) {
        for (uint64_t k = b; k < e; k++) {
            ChunkParse &c = ch[k];
            uint64_t o = pos[k];
            size_t np = c.pes.size(), nm = c.es.size();
            if (np) {
                std::memcpy(out.es.data() + o, c.pes.data(), np * 4);
                std::memcpy(out.ed.data() + o, c.ped.data(), np * 4);
            }
            if (nm) {
                std::memcpy(out.es.data() + o + np, c.es.data(), nm * 4);
                std::memcpy(out.ed.data() + o + np, c.ed.data(), nm * 4);
            }
            c = ChunkParse{};
        }
    });
    return true;
}

// Vertex count and id shift of the COO reader (shared/graph.cpp:89-92,
// virtual_graph.cpp:13-20): V = max id, plus one when some id is 0; without
// a 0 id every id moves down by one.  Ids that would give 2^32 - 1 or more
// vertices are refused (the reference would index past its arrays).
static bool coo_vertices(const ParsedEdges &pe, uint32_t &V, std::string &err)
{
    if (pe.es.empty()) {
        V = 0;
        return true;
    }
    if (pe.zero ? pe.mx >= 0xfffffffeu : pe.mx == 0xffffffffu) {
        err = "vertex id " + std::to_string(pe.mx) + " is too large (at most 2^32 - 2 vertices)";
        return false;
    }
    V = pe.mx + (pe.zero ? 1u : 0u);
    return true;
}

// Vertex count of the dedup reader: max id + 1 after the shift (src/graph.cpp:102).
static bool dedup_vertices(const ParsedEdges &pe, uint32_t &V, std::string &err)
{
    if (pe.es.empty()) {
        V = 1; // V++ of an empty maximum
        return true;
    }
    if (pe.mx == 0xffffffffu) {
        err = "the deduplicating reader needs 1-based vertex ids (found id 0)";
        return false;
    }
    if (pe.mx >= 0xfffffffeu) {
        err = "vertex id " + std::to_string((uint64_t)pe.mx + 1) +
              " is too large (at most 2^32 - 2 vertices)";
        return false;
    }
    V = pe.mx + 1u;
    return true;
}

// Reader of gpp / peelone / HistoCore / CntCore / NbrCore
// (shared/graph.cpp:17-99 + virtual_graph.cpp:9-43, 93-157): keeps duplicate
// edges, drops self-loops, num_nodes = max id (+1 when some id is 0), shifts
// to 0-based when no id is 0, neighbours in edge-insertion order.
int read_coo_pairs(const char *path, EdgeFile &ef, std::string &err)
{
    ParsedEdges pe;
    if (!parse_edges(path, false, pe, &ef.file_bytes, err))
        return -1;
    uint32_t V = 0;
    if (!coo_vertices(pe, V, err))
        return -2;
    if (!pe.zero) {
        parallel_for(pe.es.size(), 1 << 20, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++) {
                pe.es[i]--;
                pe.ed[i]--;
            }
        });
    }
    ef.V = V;
    ef.has_zero_id = pe.zero;
    ef.threads = pe.threads;
    ef.src.swap(pe.es);
    ef.dst.swap(pe.ed);
    return 0;
}

int read_dedup_pairs(const char *path, EdgeFile &ef, std::string &err)
{
    ParsedEdges pe;
    if (!parse_edges(path, true, pe, &ef.file_bytes, err))
        return -1;
    uint32_t V = 0;
    if (!dedup_vertices(pe, V, err))
        return -2;
    ef.V = V;
    ef.has_zero_id = false;
    ef.threads = pe.threads;
    ef.src.swap(pe.es);
    ef.dst.swap(pe.ed);
    return 0;
}

int read_coo(const char *path, kc_host_graph &g, std::string &err)
{
    EdgeFile ef;
    int r = read_coo_pairs(path, ef, err);
    if (r)
        return r;
    const uint32_t V = ef.V;
    const uint64_t m = ef.src.size();
    const uint32_t *es = ef.src.data(), *ed = ef.dst.data();
    g.V = V;
    g.edges_read = m;
    g.has_zero_id = ef.has_zero_id;
    g.multigraph = true;
    g.rowptr.assign((size_t)V + 1, 0);
    parallel_for(m, 1 << 20, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            __atomic_fetch_add(&g.rowptr[es[i] + 1], 1, __ATOMIC_RELAXED);
            __atomic_fetch_add(&g.rowptr[ed[i] + 1], 1, __ATOMIC_RELAXED);
        }
    });
    for (uint32_t v = 0; v < V; v++)
        g.rowptr[v + 1] += g.rowptr[v];
    g.col.resize(g.rowptr[V]);
    // rows in edge-insertion order: one pass over the edges in file order
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
    EdgeFile ef;
    int r = read_dedup_pairs(path, ef, err);
    if (r)
        return r;
    r = csr_from_pairs(ef.V, ef.src.size(), ef.src.data(), ef.dst.data(), g);
    g.edges_read = ef.src.size();
    return r;
}

static inline char *put_u32(char *o, uint32_t x)
{
    char tmp[10];
    int k = 0;
    do {
        tmp[k++] = (char)('0' + x % 10);
        x /= 10;
    } while (x);
    while (k)
        *o++ = tmp[--k];
    return o;
}

// example.txt dialect: the two header lines, then "u v 1" (1-based) once per
// undirected edge from its smaller endpoint, rows in order.  Rows are
// formatted on worker threads in blocks and written in order.
int write_edgelist(const kc_host_graph &g, const char *path, std::string &err)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        err = std::string("cannot create ") + path;
        return -1;
    }
    const uint32_t V = g.V;
    constexpr uint32_t ROWS = 1u << 16; // rows per formatted block
    const uint32_t nblk = (V + ROWS - 1) / ROWS;
    std::vector<uint64_t> half(nblk, 0);
    parallel_for(nblk, 1, [&](uint64_t b, uint64_t e) {
        for (uint64_t k = b; k < e; k++) {
            uint64_t m = 0;
            uint32_t v1 = (uint32_t)std::min<uint64_t>(V, (k + 1) * ROWS);
            for (uint32_t v = (uint32_t)(k * ROWS); v < v1; v++)
                for (uint64_t x = g.rowptr[v]; x < g.rowptr[v + 1]; x++)
                    m += g.col[x] > v;
            half[k] = m;
        }
    });
    uint64_t m = 0;
    for (uint64_t h : half)
        m += h;
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate integer symmetric\n%u %u %llu\n", V, V,
                 (unsigned long long)m);
    bool ok = true;
    const uint32_t T = std::max(1u, serial() ? 1u : worker_threads());
    std::vector<std::vector<char>> out(T);
    for (uint32_t k0 = 0; k0 < nblk && ok; k0 += T) {
        uint32_t nb = std::min<uint32_t>(T, nblk - k0);
        parallel_for(nb, 1, [&](uint64_t b, uint64_t e) {
            for (uint64_t j = b; j < e; j++) {
                uint64_t k = k0 + j;
                std::vector<char> &o = out[j];
                o.resize(half[k] * 24 + 1);
                char *w = o.data();
                uint32_t v1 = (uint32_t)std::min<uint64_t>(V, (k + 1) * ROWS);
                for (uint32_t v = (uint32_t)(k * ROWS); v < v1; v++)
                    for (uint64_t x = g.rowptr[v]; x < g.rowptr[v + 1]; x++) {
                        uint32_t u = g.col[x];
                        if (u <= v)
                            continue;
                        w = put_u32(w, v + 1);
                        *w++ = ' ';
                        w = put_u32(w, u + 1);
                        *w++ = ' ';
                        *w++ = '1';
                        *w++ = '\n';
                    }
                o.resize((size_t)(w - o.data()));
            }
        });
        for (uint32_t j = 0; j < nb && ok; j++)
            ok = std::fwrite(out[j].data(), 1, out[j].size(), f) == out[j].size();
    }
    if (std::fclose(f) != 0 || !ok) {
        err = std::string("cannot write ") + path;
        return -1;
    }
    return 0;
}

} // namespace kch
