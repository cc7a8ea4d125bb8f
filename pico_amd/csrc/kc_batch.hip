// kc_batch.hip -- batched k-core: thousands of independent graphs of uneven
// size decomposed in one launch.
//
// Layout (concatenated int32 CSR with per-graph offsets):
//   vtx_off[G+1], edge_off[G+1]  u64 prefix offsets
//   rowptr   u32, graph g's V_g+1 local offsets start at vtx_off[g] + g
//   col      u32 local neighbour ids, graph g's at edge_off[g]
//   core     u32 per vertex (output), at vtx_off[g]
//   queue    u32 per vertex: the graph's peel order (each vertex enters once)
//   live     u32 per vertex: still-unpeeled vertices, compacted per level
//   order[G] graphs by decreasing size; a device counter hands them out
//
// One workgroup owns one graph at a time and runs the bucket peel of
// peel-dyn.cc:2-131 on it with workgroup barriers only: per level, scan the
// live list for deg == level (the scan and the cascade never overlap, so no
// claim bits are needed), then drain the cascade in edge-balanced rounds with
// clamp-at-level decrements.  Empty levels are skipped via the minimum live
// degree.  Degrees live in LDS for graphs up to LDS_V vertices, otherwise in
// the graph's slice of `core` with workgroup-scope atomics (no other
// workgroup touches it).  No grid barrier, so the launch need not be
// co-resident and graphs of any size mix freely.
#include "kc_device.hpp"
#include "kc_engines.hpp"

namespace kc {


template <int BLOCK, int LDS_V> struct BLds {
    uint32_t deg[LDS_V];
    uint64_t start[BLOCK];
    uint64_t pre[BLOCK + 1];
    uint32_t g, head, tail, nlive, nkeep, minv, levels;
};

__device__ __forceinline__ uint32_t wg_sub(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t wg_add(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int BLOCK>
__device__ __forceinline__ uint32_t block_min_u32(uint32_t x, uint32_t *slot)
{
    x = wave_min_u32(x);
    if (lane_id() == 0)
        atomicMin(slot, x);
    __syncthreads();
    return *slot;
}

// Exclusive prefix of one length per thread into pre[0..m], pre[m] = total.
template <int BLOCK, int LDS_V>
__device__ __forceinline__ void b_stage_prefix(BLds<BLOCK, LDS_V> &L, uint32_t m, uint32_t len)
{
    constexpr int NW = BLOCK / KC_WAVE;
    __shared__ uint64_t s_w[NW];
    uint64_t incl = len;
    uint32_t l = lane_id();
#pragma unroll
    for (int o = 1; o < KC_WAVE; o <<= 1) {
        uint64_t y = __shfl_up(incl, o, KC_WAVE);
        if (l >= (uint32_t)o)
            incl += y;
    }
    uint32_t w = threadIdx.x / KC_WAVE;
    if (l == KC_WAVE - 1)
        s_w[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = 0;
        for (int i = 0; i < NW; i++) {
            uint64_t t = s_w[i];
            s_w[i] = run;
            run += t;
        }
        L.pre[m] = run;
    }
    __syncthreads();
    if (threadIdx.x < m)
        L.pre[threadIdx.x] = incl - len + s_w[w];
    __syncthreads();
}

__device__ __forceinline__ uint32_t b_find(const uint64_t *pre, uint32_t n, uint64_t e)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// wave-aggregated append of `want` lanes' vertices to the graph's queue
template <int BLOCK, int LDS_V>
__device__ __forceinline__ void b_push(BLds<BLOCK, LDS_V> &L, uint32_t *q, bool want, uint32_t v)
{
    uint64_t m = ballot(want);
    if (!m)
        return;
    uint32_t lane = lane_id();
    uint32_t leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(&L.tail, (uint32_t)__popcll(m));
    base = __shfl(base, leader, KC_WAVE);
    if (want)
        q[base + lane_rank(m)] = v;
}

// BU: cascade edges per thread per pass (the edge-balanced split width).
template <int BLOCK, int LDS_V, int BU>
__global__ __launch_bounds__(BLOCK) void batch_kernel(BatchArgs a)
{
    __shared__ BLds<BLOCK, LDS_V> L;
    const uint32_t lane = lane_id();
    for (;;) {
        if (threadIdx.x == 0)
            L.g = atomicAdd(a.counter, 1u);
        __syncthreads();
        uint32_t gi = L.g;
        if (gi >= a.G)
            break;
        const uint32_t g = a.order[gi];
        const uint64_t voff = a.vtx_off[g];
        const uint32_t V = (uint32_t)(a.vtx_off[g + 1] - voff);
        const uint64_t eoff = a.edge_off[g];
        const uint32_t *rp = a.rowptr + voff + g;
        const uint32_t *col = a.col + eoff;
        uint32_t *q = a.queue + voff;
        uint32_t *live = a.live + voff;
        const bool in_lds = V <= (uint32_t)LDS_V;
        uint32_t *deg = in_lds ? L.deg : a.core + voff;
        for (uint32_t v = threadIdx.x; v < V; v += BLOCK) {
            deg[v] = rp[v + 1] - rp[v];
            live[v] = v;
        }
        if (threadIdx.x == 0) {
            L.head = 0;
            L.tail = 0;
            L.nlive = V;
            L.levels = 0;
        }
        __syncthreads();
        uint32_t level = 0, kmax = 0;
        while (L.tail < V) {
            // ---- scan the live list: claim deg == level, compact the rest
            if (threadIdx.x == 0) {
                L.nkeep = 0;
                L.minv = 0xffffffffu;
            }
            __syncthreads();
            const uint32_t nl = L.nlive;
            uint32_t lmin = 0xffffffffu;
            constexpr uint32_t SU = 4; // live entries per thread per step, loads in flight
            for (uint32_t b0 = 0; b0 < nl; b0 += BLOCK * SU) {
                uint32_t v[SU], d[SU];
#pragma unroll
                for (uint32_t k = 0; k < SU; k++) {
                    uint32_t i = b0 + k * BLOCK + threadIdx.x;
                    v[k] = i < nl ? live[i] : 0xffffffffu;
                }
                // exact values: the scan never overlaps the cascade, and HBM
                // degrees are read past the L1 (the cascade's atomics run in L2)
#pragma unroll
                for (uint32_t k = 0; k < SU; k++)
                    d[k] = v[k] != 0xffffffffu ? (in_lds ? deg[v[k]] : ld_relaxed(&deg[v[k]]))
                                               : 0xffffffffu;
                __syncthreads(); // every read of this step done before writes
#pragma unroll
                for (uint32_t k = 0; k < SU; k++) {
                    bool hit = v[k] != 0xffffffffu && d[k] == level;
                    bool keep = v[k] != 0xffffffffu && d[k] > level;
                    if (keep)
                        lmin = d[k] < lmin ? d[k] : lmin;
                    b_push<BLOCK, LDS_V>(L, q, hit, v[k]);
                    uint64_t km = ballot(keep);
                    if (km) {
                        uint32_t leader = __ffsll((unsigned long long)km) - 1;
                        uint32_t base = 0;
                        if (lane == leader)
                            base = atomicAdd(&L.nkeep, (uint32_t)__popcll(km));
                        base = __shfl(base, leader, KC_WAVE);
                        // kept entries <= entries read so far: positions < b0 + BLOCK*SU
                        if (keep)
                            live[base + lane_rank(km)] = v[k];
                    }
                }
            }
            __syncthreads();
            uint32_t mn = block_min_u32<BLOCK>(lmin, &L.minv);
            if (threadIdx.x == 0)
                L.nlive = L.nkeep;
            __syncthreads();
            if (L.tail == L.head) {
                // empty level: jump to the smallest live degree
                level = mn;
                continue;
            }
            // ---- cascade: drain the queue in edge-balanced rounds
            while (L.head < L.tail) {
                uint32_t h = L.head, t = L.tail;
                uint32_t m = t - h < (uint32_t)BLOCK ? t - h : (uint32_t)BLOCK;
                uint32_t len = 0;
                if (threadIdx.x < m) {
                    uint32_t v = q[h + threadIdx.x];
                    uint32_t s = rp[v];
                    len = rp[v + 1] - s;
                    L.start[threadIdx.x] = s;
                }
                b_stage_prefix<BLOCK, LDS_V>(L, m, len);
                if (threadIdx.x == 0)
                    L.head = h + m;
                const uint64_t total = L.pre[m];
                // BU edges per thread per pass, their loads and atomics
                // independent so they are in flight together
                for (uint64_t e0 = 0; e0 < total; e0 += (uint64_t)BLOCK * BU) {
                    uint32_t u[BU], d[BU];
                    bool claim[BU];
#pragma unroll
                    for (int j = 0; j < BU; j++) {
                        uint64_t e = e0 + (uint64_t)j * BLOCK + threadIdx.x;
                        u[j] = 0xffffffffu;
                        if (e < total) {
                            uint32_t i = b_find(L.pre, m, e);
                            u[j] = col[L.start[i] + (e - L.pre[i])];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < BU; j++)
                        d[j] = u[j] != 0xffffffffu ? deg[u[j]] : 0u;
#pragma unroll
                    for (int j = 0; j < BU; j++) {
                        claim[j] = false;
                        if (d[j] > level) {
                            uint32_t old = in_lds ? atomicSub(&deg[u[j]], 1u)
                                                  : wg_sub(&deg[u[j]], 1u);
                            if (old == level + 1)
                                claim[j] = true;
                            else if (old <= level) {
                                if (in_lds)
                                    atomicAdd(&deg[u[j]], 1u);
                                else
                                    wg_add(&deg[u[j]], 1u);
                            }
                        }
                    }
#pragma unroll
                    for (int j = 0; j < BU; j++)
                        b_push<BLOCK, LDS_V>(L, q, claim[j], u[j]);
                }
                __syncthreads();
            }
            kmax = level;
            if (threadIdx.x == 0)
                L.levels++;
            level++;
            __syncthreads();
        }
        if (in_lds)
            for (uint32_t v = threadIdx.x; v < V; v += BLOCK)
                a.core[voff + v] = L.deg[v];
        if (threadIdx.x == 0) {
            a.kmax[g] = V ? kmax : 0;
            a.levels[g] = L.levels;
        }
        __syncthreads();
    }
}

template <int BLOCK, int LDS_V, int BU>
static hipError_t launch_batch_t(const BatchArgs &a, uint32_t grid, hipStream_t st)
{
    hipLaunchKernelGGL((batch_kernel<BLOCK, LDS_V, BU>), dim3(grid), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}

template <int BLOCK, int LDS_V, int BU> static int occupancy_t()
{
    int nb = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<BLOCK, LDS_V, BU>, BLOCK,
                                                        0) == hipSuccess
               ? nb
               : 0;
}

// The built shapes: per block size a small and a large LDS degree capacity
// (the large one is the default), each with 4 or 8 cascade edges per thread
// (8 the default).  Dispatch by the three parameters; 0 when not built.
#define KC_BATCH_SHAPES(X)                                                                        \
    X(256, 4096, 4) X(256, 4096, 8) X(256, 8192, 4) X(256, 8192, 8)                               \
    X(512, 8192, 4) X(512, 8192, 8) X(512, 16384, 4) X(512, 16384, 8)                             \
    X(1024, 16384, 4) X(1024, 16384, 8) X(1024, 32768, 4) X(1024, 32768, 8)

int batch_occupancy(uint32_t block, uint32_t lds_v, uint32_t bu)
{
#define KC_OCC(B, V, U)                                                                           \
    if (block == B && lds_v == V && bu == U)                                                      \
        return occupancy_t<B, V, U>();
    KC_BATCH_SHAPES(KC_OCC)
#undef KC_OCC
    return 0;
}

uint32_t batch_lds_vertices(uint32_t block, bool large)
{
    switch (block) {
    case 256: return large ? 8192u : 4096u;
    case 512: return large ? 16384u : 8192u;
    case 1024: return large ? 32768u : 16384u;
    default: return 0;
    }
}

hipError_t launch_batch(const BatchArgs &a, uint32_t block, uint32_t lds_v, uint32_t bu,
                        uint32_t grid, hipStream_t st)
{
#define KC_LAUNCH(B, V, U)                                                                        \
    if (block == B && lds_v == V && bu == U)                                                      \
        return launch_batch_t<B, V, U>(a, grid, st);
    KC_BATCH_SHAPES(KC_LAUNCH)
#undef KC_LAUNCH
    return hipErrorInvalidValue;
}

// The runtime loads a source file's kernels at the first launch of any of
// them; an empty launch at context creation takes that cost out of the first
// decomposition.
__global__ void batch_touch_kernel() {}

hipError_t batch_touch(hipStream_t st)
{
    hipLaunchKernelGGL(batch_touch_kernel, dim3(1), dim3(64), 0, st);
    return hipGetLastError();
}

} // namespace kc
