// kc_compact.hip -- residual compaction for the bucket peel.
//
// Once few vertices are live, a level of the peel still walks whole rows
// (edges to claimed vertices included) over a degree array spread across the
// full vertex range.  The peel kernel can stop at a level start; the live
// vertices are then renumbered in id order (a prefix over the claim bitmap)
// and their live-live adjacency copied into a compact CSR, on which the peel
// continues from the same level.  The k-cores for k >= that level are
// subgraphs of the live subgraph, so the core numbers it yields are those of
// the whole graph.
//
// Rows are copied in order: short rows by one wave, long rows by a whole
// workgroup with a block-wide prefix per step, so the result is deterministic.
// The live count of each row must equal the vertex's residual degree (every
// claimed neighbour decremented it exactly once); a mismatch is reported, and
// no write ever goes past the row's end.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "kc_device.hpp"
#include "kc_engines.hpp"

namespace kc {

namespace {

constexpr uint32_t CB = 256;          // threads per block in these kernels
constexpr uint32_t LONG_ROW = 1024;   // rows longer than this go to a workgroup

__global__ void live_word_count(const uint64_t *done, uint32_t nw, uint32_t *cnt)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nw)
        cnt[i] = (uint32_t)__popcll(~done[i]);
    else if (i == nw)
        cnt[nw] = 0;
}

template <typename R>
__global__ void renumber(const R *rowptr, const uint64_t *done, const uint32_t *wpre,
                         const uint32_t *deg, const uint32_t *gmap_in, uint32_t V, uint32_t Vn,
                         uint32_t *loc, uint32_t *gmap_out, uint32_t *rowlen, uint32_t *bad,
                         uint32_t *longs)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v == 0)
        rowlen[Vn] = 0;
    if (v >= V)
        return;
    uint64_t freew = ~done[v >> 6];
    uint32_t b = v & 63;
    if (!((freew >> b) & 1ull))
        return;
    uint32_t i = wpre[v >> 6] + (uint32_t)__popcll(freew & ((1ull << b) - 1ull));
    if (i >= Vn) { // live count differs from the peel's
        atomicOr(bad, 1u);
        return;
    }
    loc[i] = v;
    gmap_out[i] = gmap_in ? gmap_in[v] : v;
    uint32_t d = deg[v];
    uint64_t full = (uint64_t)(rowptr[v + 1] - rowptr[v]);
    if (d > full) {
        atomicOr(bad, 2u);
        d = 0;
    }
    rowlen[i] = d;
    if (full > LONG_ROW) // listed for copy_long (bad[1] counts them)
        longs[atomicAdd(&bad[1], 1u)] = i;
}

// Per 64 vertices: the free (live) bits and the live vertices before them,
// side by side so that renumbering a neighbour is one 16-byte gather.
struct alignas(16) LiveWord {
    uint64_t freew;
    uint32_t before;
    uint32_t pad;
};

__global__ void pack_words(const uint64_t *done, const uint32_t *wpre, uint32_t nw, LiveWord *lw)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nw)
        lw[i] = LiveWord{~done[i], wpre[i], 0u};
}

// New id of u if live, else ~0.
__device__ __forceinline__ uint32_t new_id(const LiveWord *lw, uint32_t u)
{
    LiveWord x = lw[u >> 6];
    uint32_t b = u & 63;
    if (!((x.freew >> b) & 1ull))
        return 0xffffffffu;
    return x.before + (uint32_t)__popcll(x.freew & ((1ull << b) - 1ull));
}

// One wave per short row, 4 x 64 entries in flight per step.  A wave takes
// SR consecutive rows at a time and loads their bounds with one access per
// lane, so a row costs only its own adjacency's round trips.
constexpr uint32_t SR = 8;
__device__ __forceinline__ uint64_t shfl64(uint64_t x, uint32_t src)
{
    uint32_t lo = __shfl((uint32_t)x, src, KC_WAVE), hi = __shfl((uint32_t)(x >> 32), src, KC_WAVE);
    return (uint64_t)hi << 32 | lo;
}

template <typename R, typename R2>
__device__ void copy_short(uint32_t bid, uint32_t nblocks, const R *rowptr, const uint32_t *col,
                           const LiveWord *lw, const uint32_t *wpre, const uint32_t *loc,
                           uint32_t Vn, uint32_t nw, const R2 *orow, uint32_t *ocol, uint64_t cap,
                           uint32_t *bad)
{
    constexpr uint32_t U = 4;
    const uint32_t lane = lane_id();
    if (wpre[nw] != Vn) { // the renumbering is not a bijection onto [0, Vn)
        if (threadIdx.x == 0)
            atomicOr(bad, 16u);
        return;
    }
    const uint32_t nwaves = nblocks * (CB / KC_WAVE);
    for (uint32_t i0 = (bid * (CB / KC_WAVE) + threadIdx.x / KC_WAVE) * SR; i0 < Vn;
         i0 += nwaves * SR) {
        uint32_t il = i0 + lane;
        uint64_t ls = 0, le = 0, lob = 0, loe = 0;
        if (lane < SR && il < Vn) {
            uint32_t v = loc[il];
            ls = (uint64_t)rowptr[v];
            le = (uint64_t)rowptr[v + 1];
            lob = (uint64_t)orow[il];
            loe = (uint64_t)orow[il + 1];
        }
        const uint32_t nr = Vn - i0 < SR ? Vn - i0 : SR;
        for (uint32_t r = 0; r < nr; r++) {
            uint64_t s = shfl64(ls, r), e = shfl64(le, r);
            if (e - s > LONG_ROW)
                continue;
            uint64_t pos = shfl64(lob, r), oe = shfl64(loe, r);
            if (oe > cap) {
                if (lane == 0)
                    atomicOr(bad, 32u);
                continue;
            }
            for (uint64_t j0 = s; j0 < e; j0 += (uint64_t)KC_WAVE * U) {
                uint32_t u[U], id[U];
#pragma unroll
                for (uint32_t k = 0; k < U; k++) {
                    uint64_t j = j0 + (uint64_t)k * KC_WAVE + lane;
                    u[k] = j < e ? col[j] : 0xffffffffu;
                }
#pragma unroll
                for (uint32_t k = 0; k < U; k++)
                    id[k] = u[k] != 0xffffffffu ? new_id(lw, u[k]) : 0xffffffffu;
#pragma unroll
                for (uint32_t k = 0; k < U; k++) {
                    uint64_t m = ballot(id[k] != 0xffffffffu);
                    uint64_t p = pos + lane_rank(m);
                    if (id[k] != 0xffffffffu && p < oe)
                        ocol[p] = id[k];
                    pos += (uint32_t)__popcll(m);
                }
            }
            if (lane == 0 && pos != oe)
                atomicOr(bad, 4u);
        }
    }
}

// One workgroup per long row, 16 entries per thread per step.
template <typename R, typename R2>
__device__ void copy_long(uint32_t bid, uint32_t nblocks, const R *rowptr, const uint32_t *col,
                          const LiveWord *lw, const uint32_t *wpre, const uint32_t *loc,
                          const uint32_t *longs, uint32_t Vn, uint32_t nw, const R2 *orow,
                          uint32_t *ocol, uint64_t cap, uint32_t *bad)
{
    if (wpre[nw] != Vn)
        return; // reported by copy_short
    const uint32_t nlong = bad[1];
    constexpr uint32_t U = 16;
    constexpr uint32_t NW = CB / KC_WAVE;
    static_assert(NW * U == KC_WAVE, "one count per lane of a wave");
    __shared__ uint32_t s_w[NW * U];
    __shared__ uint32_t s_pre[NW * U + 1];
    const uint32_t lane = lane_id(), w = threadIdx.x / KC_WAVE;
    for (uint32_t li = bid; li < nlong; li += nblocks) {
        uint32_t i = longs[li];
        uint32_t v = loc[i];
        uint64_t s = (uint64_t)rowptr[v], e = (uint64_t)rowptr[v + 1];
        uint64_t oe = (uint64_t)orow[i + 1];
        uint64_t pos = (uint64_t)orow[i];
        if (oe > cap) {
            if (threadIdx.x == 0)
                atomicOr(bad, 32u);
            continue;
        }
        for (uint64_t j0 = s; j0 < e; j0 += (uint64_t)CB * U) {
            uint32_t id[U];
#pragma unroll
            for (uint32_t k = 0; k < U; k++) {
                uint64_t j = j0 + (uint64_t)k * CB + threadIdx.x;
                id[k] = j < e ? col[j] : 0xffffffffu;
            }
#pragma unroll
            for (uint32_t k = 0; k < U; k++)
                id[k] = id[k] != 0xffffffffu ? new_id(lw, id[k]) : 0xffffffffu;
#pragma unroll
            for (uint32_t k = 0; k < U; k++) {
                uint64_t m = ballot(id[k] != 0xffffffffu);
                if (lane == 0)
                    s_w[k * NW + w] = (uint32_t)__popcll(m);
            }
            __syncthreads();
            // entries are ordered by (k, wave, lane): prefix over s_w in that
            // order, its NW * U = 64 counts one per lane of wave 0
            if (w == 0) {
                const uint32_t x = s_w[lane];
                const uint32_t incl = wave_incl_scan_u32(x);
                s_pre[lane] = incl - x;
                if (lane == KC_WAVE - 1)
                    s_pre[NW * U] = incl;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t k = 0; k < U; k++) {
                uint64_t m = ballot(id[k] != 0xffffffffu);
                uint64_t p = pos + s_pre[k * NW + w] + lane_rank(m);
                if (id[k] != 0xffffffffu && p < oe)
                    ocol[p] = id[k];
            }
            pos += s_pre[NW * U];
            __syncthreads();
        }
        if (threadIdx.x == 0 && pos != oe)
            atomicOr(bad, 8u);
    }
}

// Both row copies in one launch: the first nl workgroups take the long rows
// (dispatched first), the others the short ones, so the two run side by side.
template <typename R, typename R2>
__global__ __launch_bounds__(CB) void copy_rows(uint32_t nl, const R *rowptr, const uint32_t *col,
                                                const LiveWord *lw, const uint32_t *wpre,
                                                const uint32_t *loc, const uint32_t *longs,
                                                uint32_t Vn, uint32_t nw, const R2 *orow,
                                                uint32_t *ocol, uint64_t cap, uint32_t *bad)
{
    if (blockIdx.x < nl)
        copy_long<R, R2>(blockIdx.x, nl, rowptr, col, lw, wpre, loc, longs, Vn, nw, orow, ocol,
                         cap, bad);
    else
        copy_short<R, R2>(blockIdx.x - nl, gridDim.x - nl, rowptr, col, lw, wpre, loc, Vn, nw,
                          orow, ocol, cap, bad);
}

__global__ void scatter_kernel(const uint32_t *sub, const uint32_t *gmap, uint32_t n,
                               uint32_t *core)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        core[gmap[i]] = sub[i];
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// scratch layout: word counts | row lengths | long rows | packed words | scan
inline uint32_t *longs_of(const CompactOut &out)
{
    uint32_t nw = (out.V_in + 63) / 64;
    return (uint32_t *)((char *)out.tmp + align256((size_t)(nw + 1) * 4) +
                        align256((size_t)(out.Vn + 1) * 4));
}
inline LiveWord *words_of(const CompactOut &out)
{
    return (LiveWord *)((char *)longs_of(out) + align256((size_t)(out.Vn + 1) * 4));
}

template <typename R2> size_t scan_bytes(uint32_t n)
{
    size_t b = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const uint32_t *)nullptr, (R2 *)nullptr, n);
    return b;
}

// k-core subgraph: excluded vertices (core < k, and the padding past V) are
// the set bits of a claim-bitmap-shaped word array, so the compaction above
// applies unchanged.
__global__ void kcore_mark(const uint32_t *core, uint32_t V, uint32_t k, uint64_t *excl,
                           uint32_t nw)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; // blockDim is a multiple of 64
    uint64_t m = ballot(v >= V || core[v < V ? v : 0] < k);
    if (lane_id() == 0 && v / 64 < nw)
        excl[v / 64] = m;
}

// Row lengths of the subgraph: live neighbours of each live vertex (0 for
// excluded ones), a wave per row with 4 x 64 entries in flight.
template <typename R>
__global__ __launch_bounds__(CB) void count_live(const R *rowptr, const uint32_t *col,
                                                 const uint64_t *excl, uint32_t V, uint32_t *deg)
{
    constexpr uint32_t U = 4;
    const uint32_t lane = lane_id();
    const uint32_t nwaves = gridDim.x * (CB / KC_WAVE);
    for (uint32_t v = blockIdx.x * (CB / KC_WAVE) + threadIdx.x / KC_WAVE; v < V; v += nwaves) {
        if ((excl[v >> 6] >> (v & 63)) & 1ull) {
            if (lane == 0)
                deg[v] = 0;
            continue;
        }
        uint64_t s = (uint64_t)rowptr[v], e = (uint64_t)rowptr[v + 1];
        uint32_t c = 0;
        for (uint64_t j0 = s; j0 < e; j0 += (uint64_t)KC_WAVE * U) {
            uint32_t u[U];
#pragma unroll
            for (uint32_t k = 0; k < U; k++) {
                uint64_t j = j0 + (uint64_t)k * KC_WAVE + lane;
                u[k] = j < e ? col[j] : 0xffffffffu;
            }
#pragma unroll
            for (uint32_t k = 0; k < U; k++)
                c += u[k] != 0xffffffffu && !((excl[u[k] >> 6] >> (u[k] & 63)) & 1ull);
        }
        c = wave_sum_u32(c);
        if (lane == 0)
            deg[v] = c;
    }
}

// Degree statistics: out[0] max, out[1] sum, out[2] / out[3] the low / high
// 32-bit halves of each squared degree, summed (exact to 2^64 per half).
template <typename R>
__global__ void degree_stats_kernel(const R *rowptr, uint32_t V, unsigned long long *out)
{
    uint32_t mx = 0;
    uint64_t sum = 0, lo = 0, hi = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        uint64_t d = (uint64_t)(rowptr[v + 1] - rowptr[v]);
        uint32_t d32 = d > 0xffffffffull ? 0xffffffffu : (uint32_t)d;
        mx = d32 > mx ? d32 : mx;
        sum += d;
        uint64_t sq = d * d;
        lo += sq & 0xffffffffull;
        hi += sq >> 32;
    }
    mx = wave_max_u32(mx);
    sum = wave_sum_u64(sum);
    lo = wave_sum_u64(lo);
    hi = wave_sum_u64(hi);
    if (lane_id() == 0) {
        atomicMax(&out[0], (unsigned long long)mx);
        atomicAdd(&out[1], (unsigned long long)sum);
        atomicAdd(&out[2], (unsigned long long)lo);
        atomicAdd(&out[3], (unsigned long long)hi);
    }
}

} // namespace

size_t compact_tmp_bytes(uint32_t V, uint32_t Vn)
{
    uint32_t nw = (V + 63) / 64;
    size_t a = scan_bytes<uint32_t>(nw + 1);
    size_t b = scan_bytes<uint64_t>(Vn + 1);
    size_t c = scan_bytes<uint32_t>(Vn + 1);
    return align256((size_t)(nw + 1) * 4) + 2 * align256((size_t)(Vn + 1) * 4) +
           align256((size_t)(nw + 1) * sizeof(LiveWord)) + align256(std::max(a, std::max(b, c)));
}

hipError_t compact_rows(const CompactIn &in, CompactOut &out, hipStream_t st)
{
    uint32_t nw = (in.V + 63) / 64;
    char *t = (char *)out.tmp;
    uint32_t *cnt = (uint32_t *)t;
    uint32_t *rowlen = (uint32_t *)(t + align256((size_t)(nw + 1) * 4));
    uint32_t *longs = longs_of(out);
    void *ctmp = (char *)words_of(out) + align256((size_t)(nw + 1) * sizeof(LiveWord));
    size_t cbytes = out.tmp_bytes - (size_t)((char *)ctmp - t);
    hipLaunchKernelGGL(live_word_count, dim3((nw + 1 + CB - 1) / CB), dim3(CB), 0, st, in.done,
                       nw, cnt);
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(ctmp, cbytes, cnt, out.wpre, nw + 1, st);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(pack_words, dim3((nw + CB - 1) / CB), dim3(CB), 0, st, in.done, out.wpre,
                       nw, words_of(out));
    if (in.V) {
        if (in.wide)
            hipLaunchKernelGGL(renumber<uint64_t>, dim3((in.V + CB - 1) / CB), dim3(CB), 0, st,
                               (const uint64_t *)in.rowptr, in.done, out.wpre, in.deg, in.gmap,
                               in.V, out.Vn, out.loc, out.gmap, rowlen, out.bad, longs);
        else
            hipLaunchKernelGGL(renumber<uint32_t>, dim3((in.V + CB - 1) / CB), dim3(CB), 0, st,
                               (const uint32_t *)in.rowptr, in.done, out.wpre, in.deg, in.gmap,
                               in.V, out.Vn, out.loc, out.gmap, rowlen, out.bad, longs);
    }
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if (out.wide)
        return hipcub::DeviceScan::ExclusiveSum(ctmp, cbytes, rowlen, (uint64_t *)out.rowptr,
                                                out.Vn + 1, st);
    return hipcub::DeviceScan::ExclusiveSum(ctmp, cbytes, rowlen, (uint32_t *)out.rowptr,
                                            out.Vn + 1, st);
}

template <typename R, typename R2>
static void launch_copy(const CompactIn &in, const CompactOut &out, hipStream_t st)
{
    uint32_t gs = std::min<uint32_t>(16384, (out.Vn + SR * (CB / KC_WAVE) - 1) /
                                                (SR * (CB / KC_WAVE)) + 1);
    const uint32_t nw = (in.V + 63) / 64;
    uint32_t gl = std::min<uint32_t>(2048, out.Vn ? out.Vn : 1);
    hipLaunchKernelGGL((copy_rows<R, R2>), dim3(gl + gs), dim3(CB), 0, st, gl,
                       (const R *)in.rowptr, in.col, (const LiveWord *)words_of(out), out.wpre,
                       out.loc, (const uint32_t *)longs_of(out), out.Vn, nw,
                       (const R2 *)out.rowptr, out.col, out.col_cap, out.bad);
}

hipError_t compact_cols(const CompactIn &in, const CompactOut &out, hipStream_t st)
{
    if (out.Vn == 0)
        return hipSuccess;
    if (in.wide) {
        if (out.wide)
            launch_copy<uint64_t, uint64_t>(in, out, st);
        else
            launch_copy<uint64_t, uint32_t>(in, out, st);
    } else {
        if (out.wide)
            launch_copy<uint32_t, uint64_t>(in, out, st);
        else
            launch_copy<uint32_t, uint32_t>(in, out, st);
    }
    return hipGetLastError();
}

hipError_t scatter_cores(const uint32_t *sub, const uint32_t *gmap, uint32_t n, uint32_t *core,
                         hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(scatter_kernel, dim3((n + CB - 1) / CB), dim3(CB), 0, st, sub, gmap, n,
                       core);
    return hipGetLastError();
}

hipError_t kcore_exclusion(const uint32_t *core, uint32_t V, uint32_t k, uint64_t *excl,
                           hipStream_t st)
{
    uint32_t nw = (V + 63) / 64;
    if (nw == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kcore_mark, dim3((nw * 64 + CB - 1) / CB), dim3(CB), 0, st, core, V, k,
                       excl, nw);
    return hipGetLastError();
}

hipError_t kcore_row_lengths(const void *rowptr, bool wide, const uint32_t *col,
                             const uint64_t *excl, uint32_t V, uint32_t *deg, hipStream_t st)
{
    if (V == 0)
        return hipSuccess;
    uint32_t g = std::min<uint32_t>(8192, (V + (CB / KC_WAVE) - 1) / (CB / KC_WAVE));
    if (wide)
        hipLaunchKernelGGL(count_live<uint64_t>, dim3(g), dim3(CB), 0, st,
                           (const uint64_t *)rowptr, col, excl, V, deg);
    else
        hipLaunchKernelGGL(count_live<uint32_t>, dim3(g), dim3(CB), 0, st,
                           (const uint32_t *)rowptr, col, excl, V, deg);
    return hipGetLastError();
}

hipError_t degree_stats(const void *rowptr, bool wide, uint32_t V, unsigned long long *out,
                        hipStream_t st)
{
    if (V == 0)
        return hipSuccess;
    uint32_t g = std::min<uint32_t>(1024, (V + CB - 1) / CB);
    if (wide)
        hipLaunchKernelGGL(degree_stats_kernel<uint64_t>, dim3(g), dim3(CB), 0, st,
                           (const uint64_t *)rowptr, V, out);
    else
        hipLaunchKernelGGL(degree_stats_kernel<uint32_t>, dim3(g), dim3(CB), 0, st,
                           (const uint32_t *)rowptr, V, out);
    return hipGetLastError();
}

// The runtime loads a source file's kernels at the first launch of any of
// them; an empty launch at context creation takes that cost out of the first
// decomposition.
__global__ void compact_touch_kernel() {}

hipError_t compact_touch(hipStream_t st)
{
    hipLaunchKernelGGL(compact_touch_kernel, dim3(1), dim3(64), 0, st);
    return hipGetLastError();
}

} // namespace kc
