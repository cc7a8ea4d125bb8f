// kc_gen.hip -- synthetic graphs built on the device.
//
// The generators draw exactly the edges of the host generators
// (pico_amd/host/kc_host.cpp: same counter-based hashes, same per-level
// quadrant thresholds in double precision, same alias tables), so a graph
// built here equals the host-built one entry for entry.  The CSR builder:
//
//   1. generate: edge i of graph g -> two 64-bit keys (src << 32 | dst) with
//      src numbered from the first vertex of the chunk, self-loops -> ~0
//   2. hipCUB radix sort of the keys (graphs are contiguous in src order, so
//      rows come out grouped per graph and sorted by neighbour)
//   3. hipCUB unique -> symmetric, deduplicated adjacency
//   4. row starts from the sorted keys, col = low halves
//
// Graphs are processed in chunks of consecutive graphs of at most 2^30 keys
// (two 8 GiB key buffers), so a 4,096-graph batch of ~7G adjacency entries
// never needs more than a chunk's sort space beside its output.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "kc_engines.hpp"

namespace kc {

static constexpr uint64_t KEY_NONE = ~0ull;

__device__ __forceinline__ uint64_t d_mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double d_u01(uint64_t h)
{
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// last index j in [0, n) with pre[j] <= x (pre[0] == 0 <= x)
__device__ __forceinline__ uint32_t upper_idx(const uint64_t *pre, uint32_t n, uint64_t x)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

struct RmatKeys {
    uint32_t gc;           // graphs in the chunk
    const uint64_t *mpre;  // gc+1: prefix of generated edges
    const uint32_t *scale; // gc
    const uint64_t *gkey;  // gc: mix64(seed_g ^ tag)
    const uint64_t *vrel;  // gc: first vertex relative to the chunk
    double a, ab, abc;
    uint64_t *keys;        // 2 * mpre[gc]
};

__global__ void rmat_keys_kernel(RmatKeys k)
{
    const uint64_t n = k.mpre[k.gc];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t g = upper_idx(k.mpre, k.gc, i);
        uint64_t e = i - k.mpre[g];
        uint32_t sc = k.scale[g];
        uint64_t base = k.gkey[g] ^ (e * 0xD1B54A32D192ED03ull);
        uint32_t u = 0, v = 0;
        for (uint32_t l = 0; l < sc; l++) {
            double r = d_u01(d_mix64(base + l * 0x9E3779B97F4A7C15ull));
            uint32_t bu = r >= k.ab;
            uint32_t bv = (r >= k.a && r < k.ab) || r >= k.abc;
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        uint64_t k0 = KEY_NONE, k1 = KEY_NONE;
        if (u != v) {
            uint64_t vr = k.vrel[g];
            k0 = ((vr + u) << 32) | v;
            k1 = ((vr + v) << 32) | u;
        }
        k.keys[2 * i] = k0;
        k.keys[2 * i + 1] = k1;
    }
}

struct AliasKeys {
    uint64_t m;            // pairs to draw
    uint32_t n;            // vertices
    uint64_t key;          // mix64(seed ^ tag)
    const uint32_t *thr;   // n: 32-bit acceptance thresholds
    const uint32_t *alias; // n
    uint64_t *keys;        // 2m
};

__device__ __forceinline__ uint32_t alias_draw(const AliasKeys &k, uint64_t h)
{
    uint32_t i = (uint32_t)(((h >> 32) * (uint64_t)k.n) >> 32);
    return ((uint32_t)h < k.thr[i]) ? i : k.alias[i];
}

__global__ void alias_keys_kernel(AliasKeys k)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k.m; i += stride) {
        uint64_t base = k.key ^ (i * 0xD1B54A32D192ED03ull);
        uint32_t s = alias_draw(k, d_mix64(base));
        uint32_t d = alias_draw(k, d_mix64(base + 0x9E3779B97F4A7C15ull));
        uint64_t k0 = KEY_NONE, k1 = KEY_NONE;
        if (s != d) {
            k0 = ((uint64_t)s << 32) | d;
            k1 = ((uint64_t)d << 32) | s;
        }
        k.keys[2 * i] = k0;
        k.keys[2 * i + 1] = k1;
    }
}

// Row starts of the chunk's vertices [0, nv] from n sorted unique keys (the
// last may be KEY_NONE) and the col entries, written at offset `base`.
__global__ void rows_from_keys(const uint64_t *keys, uint64_t n, uint32_t nv, uint64_t base,
                               uint32_t *col, uint64_t *rp)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        uint64_t key = i < n ? keys[i] : KEY_NONE;
        uint64_t s = key == KEY_NONE ? nv : (key >> 32);
        uint64_t p = 0;
        if (i > 0) {
            uint64_t pk = keys[i - 1];
            p = (pk == KEY_NONE ? (uint64_t)nv : (pk >> 32)) + 1;
        }
        for (uint64_t w = p; w <= s && w <= nv; w++)
            rp[w] = base + i;
        if (key != KEY_NONE)
            col[base + i] = (uint32_t)key;
    }
}

// local u32 rows: entry j of graph g (j in [0, V_g]) = rp[vo[g] + j] - rp[vo[g]]
__global__ void local_rows(const uint64_t *rp, const uint64_t *vo, uint32_t G, uint32_t *out)
{
    const uint64_t total = vo[G] + G;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += stride) {
        // graph g owns output slots [vo[g] + g, vo[g+1] + g + 1)
        uint32_t lo = 0, hi = G;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (vo[mid] + mid <= j)
                lo = mid;
            else
                hi = mid;
        }
        uint64_t v = j - vo[lo] - lo;
        out[j] = (uint32_t)(rp[vo[lo] + v] - rp[vo[lo]]);
    }
}

__global__ void gather_edge_off(const uint64_t *rp, const uint64_t *vo, uint32_t G, uint64_t *eo)
{
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x)
        eo[g] = rp[vo[g]];
}

static inline uint64_t h_mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static uint32_t bits_for(uint64_t x)
{
    uint32_t b = 1;
    while (b < 64 && (x >> b))
        b++;
    return b;
}

namespace {
struct Scratch {
    std::vector<void *> ptrs;
    ~Scratch()
    {
        for (void *p : ptrs)
            hipFree(p);
    }
    template <typename T> hipError_t get(T **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, n ? n * sizeof(T) : sizeof(T));
        if (e == hipSuccess)
            ptrs.push_back(*p);
        return e;
    }
};
} // namespace

#define GEN_TRY(x)                                                                                 \
    do {                                                                                           \
        hipError_t _e = (x);                                                                       \
        if (_e != hipSuccess) {                                                                    \
            err = std::string(#x) + ": " + hipGetErrorString(_e);                                  \
            return _e;                                                                             \
        }                                                                                          \
    } while (0)

hipError_t build_csr_device(const GenSpec &s, hipStream_t st, BuildOut &out, std::string &err)
{
    const uint32_t G = s.G;
    const uint64_t *vo = s.vtx_off;
    const uint64_t Vtot = vo[G];
    constexpr uint64_t CHUNK_KEYS = 1ull << 30;
    // chunks of consecutive graphs
    std::vector<uint32_t> cut{0};
    uint64_t acc = 0, total_pairs = 0;
    for (uint32_t g = 0; g < G; g++) {
        uint64_t k = 2 * s.m_gen[g];
        if (k > CHUNK_KEYS) {
            err = "a graph has more than 2^29 generated edges";
            return hipErrorInvalidValue;
        }
        if (vo[g + 1] - vo[g] > 0xffffffffull) {
            err = "a graph has 2^32 or more vertices";
            return hipErrorInvalidValue;
        }
        if (acc + k > CHUNK_KEYS || (cut.size() && g > cut.back() &&
                                     vo[g + 1] - vo[cut.back()] > 0x7fffffffull)) {
            cut.push_back(g);
            acc = 0;
        }
        acc += k;
        total_pairs += s.m_gen[g];
    }
    cut.push_back(G);
    uint64_t max_keys = 0;
    for (size_t c = 0; c + 1 < cut.size(); c++) {
        uint64_t k = 0;
        for (uint32_t g = cut[c]; g < cut[c + 1]; g++)
            k += 2 * s.m_gen[g];
        max_keys = std::max(max_keys, k);
    }
    Scratch sc;
    uint64_t *k0 = nullptr, *k1 = nullptr, *rp = nullptr, *d_vo = nullptr, *d_meta = nullptr;
    uint32_t *col_big = nullptr, *d_scale = nullptr;
    int *d_nsel = nullptr;
    GEN_TRY(sc.get(&k0, max_keys));
    GEN_TRY(sc.get(&k1, max_keys));
    GEN_TRY(sc.get(&rp, Vtot + 1));
    GEN_TRY(sc.get(&d_vo, G + 1ull));
    GEN_TRY(sc.get(&d_meta, 3ull * G + 1));
    GEN_TRY(sc.get(&d_scale, G ? G : 1));
    GEN_TRY(sc.get(&d_nsel, 1));
    GEN_TRY(hipMalloc((void **)&col_big, (2 * total_pairs ? 2 * total_pairs : 1) * 4));
    struct Free {
        uint32_t *&p;
        ~Free()
        {
            if (p)
                hipFree(p);
        }
    } free_big{col_big};
    GEN_TRY(hipMemcpyAsync(d_vo, vo, (G + 1ull) * 8, hipMemcpyHostToDevice, st));
    // sort/unique temp space sized once for the largest chunk
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    if (max_keys) {
        hipcub::DoubleBuffer<uint64_t> db(k0, k1);
        size_t need = 0, need2 = 0;
        GEN_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, need, db, (int)max_keys, 0, 64, st));
        GEN_TRY(hipcub::DeviceSelect::Unique(nullptr, need2, k0, k1, d_nsel, (int)max_keys, st));
        tmp_bytes = std::max(need, need2);
        GEN_TRY(sc.get((uint8_t **)&tmp, tmp_bytes));
    }
    // per-graph metadata for the R-MAT kernel: mpre (chunk-relative, G+1),
    // gkey (G), vrel (G)
    std::vector<uint64_t> meta(3ull * G + 1);
    std::vector<uint32_t> scale_h(G);
    uint64_t base = 0;
    for (size_t c = 0; c + 1 < cut.size(); c++) {
        uint32_t g0 = cut[c], g1 = cut[c + 1], gc = g1 - g0;
        uint64_t v0 = vo[g0];
        uint32_t nv = (uint32_t)(vo[g1] - v0);
        uint64_t npairs = 0;
        uint64_t *mpre = meta.data(), *gkey = meta.data() + gc + 1, *vrel = gkey + gc;
        for (uint32_t j = 0; j < gc; j++) {
            uint32_t g = g0 + j;
            mpre[j] = npairs;
            npairs += s.m_gen[g];
            gkey[j] = h_mix64((s.seeds ? s.seeds[g] : s.seed + g) ^ 0x524d41542d67656eull);
            vrel[j] = vo[g] - v0;
            scale_h[j] = s.kind == GEN_RMAT ? s.scale[g] : 0;
        }
        mpre[gc] = npairs;
        uint64_t nkeys = 2 * npairs;
        if (npairs) {
            if (s.kind == GEN_RMAT) {
                GEN_TRY(hipMemcpyAsync(d_meta, meta.data(), (3ull * gc + 1) * 8,
                                       hipMemcpyHostToDevice, st));
                GEN_TRY(hipMemcpyAsync(d_scale, scale_h.data(), (size_t)gc * 4,
                                       hipMemcpyHostToDevice, st));
                RmatKeys k{};
                k.gc = gc;
                k.mpre = d_meta;
                k.gkey = d_meta + gc + 1;
                k.vrel = d_meta + 2 * gc + 1;
                k.scale = d_scale;
                k.a = s.a;
                k.ab = s.ab;
                k.abc = s.abc;
                k.keys = k0;
                hipLaunchKernelGGL(rmat_keys_kernel, dim3(8192), dim3(256), 0, st, k);
            } else {
                AliasKeys k{};
                k.m = npairs;
                k.n = nv;
                k.key = h_mix64(s.seed ^ 0x4368756e674c7521ull);
                k.thr = s.thr;
                k.alias = s.alias;
                k.keys = k0;
                hipLaunchKernelGGL(alias_keys_kernel, dim3(8192), dim3(256), 0, st, k);
            }
            GEN_TRY(hipGetLastError());
            // sort on the bits in use; KEY_NONE keeps every bit set, so it
            // still sorts last
            int end_bit = (int)std::min<uint32_t>(64, 32 + bits_for(nv));
            hipcub::DoubleBuffer<uint64_t> db(k0, k1);
            GEN_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, db, (int)nkeys, 0, end_bit,
                                                      st));
            uint64_t *sorted = db.Current(), *uniq = db.Alternate();
            GEN_TRY(hipcub::DeviceSelect::Unique(tmp, tmp_bytes, sorted, uniq, d_nsel, (int)nkeys,
                                                 st));
            int nsel = 0;
            GEN_TRY(hipMemcpyAsync(&nsel, d_nsel, 4, hipMemcpyDeviceToHost, st));
            GEN_TRY(hipStreamSynchronize(st));
            hipLaunchKernelGGL(rows_from_keys, dim3(4096), dim3(256), 0, st, uniq, (uint64_t)nsel,
                               nv, base, col_big, rp + v0);
        } else {
            hipLaunchKernelGGL(rows_from_keys, dim3(1), dim3(256), 0, st, k0, (uint64_t)0, nv,
                               base, col_big, rp + v0);
        }
        GEN_TRY(hipGetLastError());
        uint64_t end = 0;
        GEN_TRY(hipMemcpyAsync(&end, rp + v0 + nv, 8, hipMemcpyDeviceToHost, st));
        GEN_TRY(hipStreamSynchronize(st));
        base = end;
    }
    // exact-size outputs
    out.nnz = base;
    GEN_TRY(hipMalloc((void **)&out.col, base ? base * 4 : 4));
    if (base)
        GEN_TRY(hipMemcpyAsync(out.col, col_big, base * 4, hipMemcpyDeviceToDevice, st));
    GEN_TRY(hipMalloc((void **)&out.rowptr, (Vtot + G) * 4 + 4));
    if (Vtot + G)
        hipLaunchKernelGGL(local_rows, dim3(4096), dim3(256), 0, st, rp, d_vo, G, out.rowptr);
    GEN_TRY(hipGetLastError());
    uint64_t *d_eo = d_meta; // >= G+1 words
    hipLaunchKernelGGL(gather_edge_off, dim3((G + 256) / 256), dim3(256), 0, st, rp, d_vo, G,
                       d_eo);
    GEN_TRY(hipGetLastError());
    out.edge_off.resize(G + 1ull);
    GEN_TRY(hipMemcpyAsync(out.edge_off.data(), d_eo, (G + 1ull) * 8, hipMemcpyDeviceToHost, st));
    if (s.keep_rowptr64) {
        GEN_TRY(hipMalloc((void **)&out.rowptr64, (Vtot + 1) * 8));
        GEN_TRY(hipMemcpyAsync(out.rowptr64, rp, (Vtot + 1) * 8, hipMemcpyDeviceToDevice, st));
    }
    GEN_TRY(hipStreamSynchronize(st));
    return hipSuccess;
}

// ------------------------------------------------- edge-file pairs -> CSR
// Both directed entries of every kept pair, keyed by their source row, in
// file order: entry 2i is (src_i -> dst_i), entry 2i+1 (dst_i -> src_i).
__global__ void expand_pairs32(const uint32_t *src, const uint32_t *dst, uint64_t m,
                               uint32_t *key, uint32_t *val)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        uint32_t a = src[i], b = dst[i];
        key[2 * i] = a;
        val[2 * i] = b;
        key[2 * i + 1] = b;
        val[2 * i + 1] = a;
    }
}

__global__ void expand_pairs64(const uint32_t *src, const uint32_t *dst, uint64_t m,
                               uint64_t *keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        uint64_t a = src[i], b = dst[i];
        keys[2 * i] = a == b ? KEY_NONE : (a << 32) | b;
        keys[2 * i + 1] = a == b ? KEY_NONE : (b << 32) | a;
    }
}

// Row starts [0, V] from n row keys sorted ascending (all < V).
template <typename RP>
__global__ void rows_from_sorted32(const uint32_t *keys, uint64_t n, uint32_t V, RP *rp)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        uint64_t k = i < n ? keys[i] : V;
        uint64_t p = i > 0 ? (uint64_t)keys[i - 1] + 1 : 0;
        for (uint64_t w = p; w <= k && w <= V; w++)
            rp[w] = (RP)i;
    }
}

__global__ void narrow_rows(const uint64_t *in, uint64_t n, uint32_t *out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint32_t)in[i];
}

hipError_t build_csr_pairs(const uint32_t *h_src, const uint32_t *h_dst, uint64_t m, uint32_t V,
                           bool keep_order, bool wide, hipStream_t st, PairCsr &out,
                           std::string &err)
{
    out = PairCsr{};
    const uint64_t n2 = 2 * m;
    if (n2 >= (1ull << 31)) {
        err = "more than 2^30 edges: beyond the device reader's sort";
        return hipErrorInvalidValue;
    }
    if (V == 0xffffffffu) {
        err = "too many vertices";
        return hipErrorInvalidValue;
    }
    Scratch sc;
    uint32_t *d_src = nullptr, *d_dst = nullptr;
    GEN_TRY(sc.get(&d_src, m));
    GEN_TRY(sc.get(&d_dst, m));
    if (m) {
        GEN_TRY(hipMemcpyAsync(d_src, h_src, m * 4, hipMemcpyHostToDevice, st));
        GEN_TRY(hipMemcpyAsync(d_dst, h_dst, m * 4, hipMemcpyHostToDevice, st));
    }
    const int nbk = (int)std::min<uint64_t>(8192, (n2 + 255) / 256 + 1);
    void *rp = nullptr;
    uint32_t *col = nullptr;
    uint64_t nnz = 0;
    struct Owned {
        void *&rp;
        uint32_t *&col;
        bool keep = false;
        ~Owned()
        {
            if (!keep) {
                hipFree(rp);
                hipFree(col);
            }
        }
    } own{rp, col};
    if (keep_order) {
        // COO reader: duplicates kept, each row in file order -- a stable
        // radix sort of the directed entries by source row
        uint32_t *k0 = nullptr, *k1 = nullptr, *v1 = nullptr;
        GEN_TRY(sc.get(&k0, n2));
        GEN_TRY(sc.get(&k1, n2));
        GEN_TRY(hipMalloc((void **)&col, n2 ? n2 * 4 : 4));
        GEN_TRY(sc.get(&v1, n2));
        if (m) {
            hipLaunchKernelGGL(expand_pairs32, dim3(nbk), dim3(256), 0, st, d_src, d_dst, m, k0,
                               col);
            GEN_TRY(hipGetLastError());
            hipcub::DoubleBuffer<uint32_t> kb(k0, k1), vb(col, v1);
            size_t need = 0;
            const int end_bit = (int)bits_for(V);
            GEN_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kb, vb, (int)n2, 0, end_bit,
                                                       st));
            void *tmp = nullptr;
            GEN_TRY(sc.get((uint8_t **)&tmp, need));
            GEN_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, need, kb, vb, (int)n2, 0, end_bit,
                                                       st));
            if (vb.Current() != col)
                GEN_TRY(hipMemcpyAsync(col, vb.Current(), n2 * 4, hipMemcpyDeviceToDevice, st));
            k0 = kb.Current();
        }
        nnz = n2;
        GEN_TRY(hipMalloc(&rp, ((uint64_t)V + 1) * (wide ? 8 : 4)));
        if (wide)
            hipLaunchKernelGGL(rows_from_sorted32<uint64_t>, dim3(nbk), dim3(256), 0, st, k0, n2,
                               V, (uint64_t *)rp);
        else
            hipLaunchKernelGGL(rows_from_sorted32<uint32_t>, dim3(nbk), dim3(256), 0, st, k0, n2,
                               V, (uint32_t *)rp);
        GEN_TRY(hipGetLastError());
    } else {
        // dedup reader: symmetric, deduplicated, rows sorted -- sort the
        // (src << 32 | dst) keys of both directions and drop repeats
        uint64_t *k0 = nullptr, *k1 = nullptr, *rp64 = nullptr;
        int *d_nsel = nullptr;
        GEN_TRY(sc.get(&k0, n2));
        GEN_TRY(sc.get(&k1, n2));
        GEN_TRY(sc.get(&d_nsel, 1));
        GEN_TRY(sc.get(&rp64, (uint64_t)V + 1));
        int nsel = 0;
        uint64_t *uniq = k0;
        if (m) {
            hipLaunchKernelGGL(expand_pairs64, dim3(nbk), dim3(256), 0, st, d_src, d_dst, m, k0);
            GEN_TRY(hipGetLastError());
            hipcub::DoubleBuffer<uint64_t> db(k0, k1);
            const int end_bit = (int)std::min<uint32_t>(64, 32 + bits_for(V));
            size_t need = 0, need2 = 0;
            GEN_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, need, db, (int)n2, 0, end_bit, st));
            GEN_TRY(hipcub::DeviceSelect::Unique(nullptr, need2, k0, k1, d_nsel, (int)n2, st));
            void *tmp = nullptr;
            size_t tb = std::max(need, need2);
            GEN_TRY(sc.get((uint8_t **)&tmp, tb));
            GEN_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, need, db, (int)n2, 0, end_bit, st));
            uint64_t *sorted = db.Current();
            uniq = db.Alternate();
            GEN_TRY(hipcub::DeviceSelect::Unique(tmp, tb, sorted, uniq, d_nsel, (int)n2, st));
            GEN_TRY(hipMemcpyAsync(&nsel, d_nsel, 4, hipMemcpyDeviceToHost, st));
            GEN_TRY(hipStreamSynchronize(st));
        }
        // a KEY_NONE (never produced: the reader drops self-loops) would sort last
        nnz = (uint64_t)nsel;
        GEN_TRY(hipMalloc((void **)&col, nnz ? nnz * 4 : 4));
        hipLaunchKernelGGL(rows_from_keys, dim3(nbk), dim3(256), 0, st, uniq, nnz, V,
                           (uint64_t)0, col, rp64);
        GEN_TRY(hipGetLastError());
        GEN_TRY(hipMalloc(&rp, ((uint64_t)V + 1) * (wide ? 8 : 4)));
        if (wide) {
            GEN_TRY(hipMemcpyAsync(rp, rp64, ((uint64_t)V + 1) * 8, hipMemcpyDeviceToDevice, st));
        } else {
            hipLaunchKernelGGL(narrow_rows, dim3(nbk), dim3(256), 0, st, rp64, (uint64_t)V + 1,
                               (uint32_t *)rp);
            GEN_TRY(hipGetLastError());
        }
    }
    GEN_TRY(hipStreamSynchronize(st));
    own.keep = true;
    out.rowptr = rp;
    out.col = col;
    out.nnz = nnz;
    out.wide = wide;
    return hipSuccess;
}

// plain streaming copy (16 B per lane per step): the bandwidth probe the
// HBM counters of the engines are scaled against
__global__ void copy_probe_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                  uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

hipError_t copy_probe(const void *src, void *dst, uint64_t bytes, hipStream_t st)
{
    uint64_t n = bytes / 16;
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(copy_probe_kernel, dim3(8192), dim3(256), 0, st, (const uint4 *)src,
                       (uint4 *)dst, n);
    return hipGetLastError();
}

__global__ void checksum_kernel(const uint32_t *core, uint32_t V, unsigned long long *out)
{
    uint64_t h = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        h += d_mix64(((uint64_t)v << 32) | core[v]);
    for (int o = 32; o >= 1; o >>= 1)
        h += __shfl_xor(h, o, 64);
    if ((threadIdx.x & 63) == 0 && h)
        atomicAdd(out, (unsigned long long)h);
}

hipError_t core_checksum(const uint32_t *core, uint32_t V, uint64_t *out, hipStream_t st)
{
    hipError_t e = hipMemsetAsync(out, 0, 8, st);
    if (e != hipSuccess || !V)
        return e;
    uint32_t blocks = (V + 255) / 256;
    hipLaunchKernelGGL(checksum_kernel, dim3(blocks < 2048 ? blocks : 2048), dim3(256), 0, st,
                       core, V, (unsigned long long *)out);
    return hipGetLastError();
}

// The runtime loads a source file's kernels at the first launch of any of
// them; an empty launch at context creation takes that cost out of the first
// decomposition.
__global__ void gen_touch_kernel() {}

hipError_t gen_touch(hipStream_t st)
{
    hipLaunchKernelGGL(gen_touch_kernel, dim3(1), dim3(64), 0, st);
    return hipGetLastError();
}

} // namespace kc
