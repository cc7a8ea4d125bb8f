// kc_cache.cpp -- binary CSR files and the content-keyed reader cache.
//
// The reference caches a parsed graph as text under
// ./output/serialized-<basename> (peel/src/graph.cpp:9-54): keyed on the file's
// base name (two different files of one name collide, an edited file is served
// stale) and parsed back number by number.  Here a cache entry is a binary
// CSR whose header records the 64-bit content hash and the size of the edge
// file it came from and which reader parsed it; the entry's name is derived
// from that hash, so a lookup is by content, and a header or payload check
// that fails turns the entry into a miss (it is parsed again and rewritten).
//
// File layout (little endian): 64-byte header
//   magic[8] "KCCSR\0\0\1", u32 version, u32 reader, u64 src_hash, u64 src_size,
//   u32 V, u32 flags (1: has_zero_id, 2: multigraph), u64 nnz, u64 edges_read,
//   u64 payload_hash
// then rowptr u64[V+1] and col u32[nnz].
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <algorithm>
#include <vector>

#include "kc_host.hpp"

namespace kch {

namespace {

const char KC_MAGIC[8] = {'K', 'C', 'C', 'S', 'R', 0, 0, 1};
constexpr uint32_t KC_VERSION = 1;

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t reader;
    uint64_t src_hash;
    uint64_t src_size;
    uint32_t V;
    uint32_t flags;
    uint64_t nnz;
    uint64_t edges_read;
    uint64_t payload_hash;
};
static_assert(sizeof(Header) == 64, "header layout");

// Streaming 64-bit hash: four independent lanes over 32-byte blocks (each
// lane h = mix64(h ^ word)), the tail word by word, then the lanes and the
// length folded together.  Incremental, so files are hashed in chunks.
struct Hasher {
    uint64_t h[4] = {0x243F6A8885A308D3ull, 0x13198A2E03707344ull, 0xA4093822299F31D0ull,
                     0x082EFA98EC4E6C89ull};
    uint64_t len = 0;
    unsigned char pend[32];
    size_t npend = 0;

    void block(const unsigned char *p)
    {
        for (int i = 0; i < 4; i++) {
            uint64_t w;
            std::memcpy(&w, p + 8 * i, 8);
            h[i] = mix64(h[i] ^ w);
        }
    }
    void update(const void *data, size_t n)
    {
        const unsigned char *p = (const unsigned char *)data;
        len += n;
        if (npend) {
            size_t take = std::min(n, 32 - npend);
            std::memcpy(pend + npend, p, take);
            npend += take;
            p += take;
            n -= take;
            if (npend < 32)
                return;
            block(pend);
            npend = 0;
        }
        for (; n >= 32; p += 32, n -= 32)
            block(p);
        std::memcpy(pend, p, n);
        npend = n;
    }
    uint64_t digest() const
    {
        uint64_t a = h[0], b = h[1], c = h[2], d = h[3];
        uint64_t t = mix64(a ^ mix64(b ^ mix64(c ^ mix64(d))));
        for (size_t i = 0; i < npend; i += 8) {
            uint64_t w = 0;
            std::memcpy(&w, pend + i, std::min<size_t>(8, npend - i));
            t = mix64(t ^ w);
        }
        return mix64(t ^ len);
    }
};

uint64_t payload_hash(const kc_host_graph &g)
{
    Hasher hs;
    hs.update(g.rowptr.data(), g.rowptr.size() * 8);
    if (!g.col.empty())
        hs.update(g.col.data(), g.col.size() * 4);
    return hs.digest();
}

bool write_all(FILE *f, const void *p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
bool read_all(FILE *f, void *p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }

} // namespace

int file_hash(const char *path, uint64_t &hash, uint64_t &size, std::string &err)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path + ": " + std::strerror(errno);
        return 1;
    }
    Hasher hs;
    std::vector<unsigned char> buf(1 << 22);
    size_t n;
    while ((n = std::fread(buf.data(), 1, buf.size(), f)) > 0)
        hs.update(buf.data(), n);
    bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) {
        err = std::string("read error on ") + path;
        return 1;
    }
    hash = hs.digest();
    size = hs.len;
    return 0;
}

int save_csr(const kc_host_graph &g, const char *path, uint32_t reader, uint64_t src_hash,
             uint64_t src_size, std::string &err)
{
    if (g.rowptr.size() != (size_t)g.V + 1 || g.col.size() != g.rowptr.back()) {
        err = "inconsistent graph";
        return 1;
    }
    Header h{};
    std::memcpy(h.magic, KC_MAGIC, 8);
    h.version = KC_VERSION;
    h.reader = reader;
    h.src_hash = src_hash;
    h.src_size = src_size;
    h.V = g.V;
    h.flags = (g.has_zero_id ? 1u : 0u) | (g.multigraph ? 2u : 0u);
    h.nnz = g.col.size();
    h.edges_read = g.edges_read;
    h.payload_hash = payload_hash(g);
    // write to a temporary name, then rename: a reader never sees a partial file
    std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        err = "cannot create " + tmp + ": " + std::strerror(errno);
        return 1;
    }
    bool ok = write_all(f, &h, sizeof h) && write_all(f, g.rowptr.data(), g.rowptr.size() * 8) &&
              write_all(f, g.col.data(), g.col.size() * 4);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        err = std::string("cannot write ") + path;
        return 1;
    }
    return 0;
}

// 0: loaded; 1: unusable (err says why)
int load_csr(const char *path, kc_host_graph &g, uint32_t *reader, uint64_t *src_hash,
             uint64_t *src_size, std::string &err)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path + ": " + std::strerror(errno);
        return 1;
    }
    Header h{};
    int rc = 1;
    if (!read_all(f, &h, sizeof h) || std::memcmp(h.magic, KC_MAGIC, 8) != 0 ||
        h.version != KC_VERSION) {
        err = std::string(path) + " is not a CSR file of this version";
    } else {
        struct stat st{};
        uint64_t want = sizeof h + ((uint64_t)h.V + 1) * 8 + h.nnz * 4;
        if (fstat(fileno(f), &st) != 0 || (uint64_t)st.st_size != want) {
            err = std::string(path) + ": size does not match its header";
        } else {
            kc_host_graph t;
            t.V = h.V;
            t.rowptr.resize((size_t)h.V + 1);
            t.col.resize(h.nnz);
            if (!read_all(f, t.rowptr.data(), t.rowptr.size() * 8) ||
                !read_all(f, t.col.data(), t.col.size() * 4)) {
                err = std::string(path) + ": short read";
            } else if (payload_hash(t) != h.payload_hash) {
                err = std::string(path) + ": payload hash mismatch";
            } else if (t.rowptr[0] != 0 || t.rowptr[h.V] != h.nnz) {
                err = std::string(path) + ": row offsets inconsistent";
            } else {
                bool ok = true;
                for (uint32_t v = 0; v < h.V && ok; v++)
                    ok = t.rowptr[v] <= t.rowptr[v + 1];
                for (uint64_t e = 0; e < h.nnz && ok; e++)
                    ok = t.col[e] < h.V;
                if (!ok) {
                    err = std::string(path) + ": rows out of order or ids out of range";
                } else {
                    t.edges_read = h.edges_read;
                    t.has_zero_id = (h.flags & 1u) != 0;
                    t.multigraph = (h.flags & 2u) != 0;
                    g = std::move(t);
                    if (reader)
                        *reader = h.reader;
                    if (src_hash)
                        *src_hash = h.src_hash;
                    if (src_size)
                        *src_size = h.src_size;
                    rc = 0;
                }
            }
        }
    }
    std::fclose(f);
    return rc;
}

} // namespace kch
