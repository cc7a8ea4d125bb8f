// kc_device.hpp -- device-side building blocks shared by the gfx950 k-core
// engines: wave64 helpers, a grid barrier that carries per-phase reductions,
// bounded spin waits, and the edge-balanced batch walker.
//
// Memory-model notes (CDNA3/CDNA4, 8 XCDs with private L2s):
//  * degree / bitmap / histogram updates use HIP's agent-scope relaxed
//    atomics, which are performed past the XCD-private L2s and are coherent
//    across XCDs;
//  * the rule every engine that uses grid_barrier<false> follows: each word
//    that one workgroup writes and another reads -- degrees, claim bits,
//    work-list slots, counters, reduction slots -- is written by an
//    agent-scope atomic or atomic store and read by an atomic or atomic load
//    (ld_relaxed*), never by a plain load or store.  The non-flushing barrier
//    (grid_barrier<false>) only waits for each workgroup's own memory
//    operations to complete (vm_drain) before it arrives; it does not write
//    back or invalidate any L2, so a plain store would stay in its XCD's L2
//    and a plain load could return a line cached before the barrier.  The
//    read-only graph (rowptr, col) is the only data read with plain loads;
//  * two planned exceptions, both plain loads in the peel's edge walk that
//    only decide whether to issue the atomic whose returned value decides
//    everything else: the dense-core pre-check reads a degree (a stale value
//    is never smaller than the true one: degrees only fall), and otherwise
//    the walk reads the target's claim word (a stale word only lacks recent
//    claims: claim bits are only ever set), so neither skips a live target;
//  * grid_barrier<true> (agent-scope release/acquire fences, which write back
//    and invalidate every XCD's L2) is for engines whose cross-workgroup data
//    is written with plain stores.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define KC_WAVE 64

namespace kc {

// ---------------------------------------------------------------- wave64 ---
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of `mask` strictly below this lane
__device__ __forceinline__ uint32_t lane_rank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t y = __shfl_xor(x, o, KC_WAVE);
        x = y < x ? y : x;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t y = __shfl_xor(x, o, KC_WAVE);
        x = y > x ? y : x;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        x += __shfl_xor(x, o, KC_WAVE);
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        x += __shfl_xor(x, o, KC_WAVE);
    return x;
}

// inclusive prefix sum across the wave
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x)
{
    uint32_t l = lane_id();
#pragma unroll
    for (int o = 1; o < KC_WAVE; o <<= 1) {
        uint32_t y = __shfl_up(x, o, KC_WAVE);
        if (l >= (uint32_t)o)
            x += y;
    }
    return x;
}

// -------------------------------------------------------------- atomics ---
__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_relaxed64(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until this thread's outstanding global loads, stores and atomics have
// completed (s_waitcnt vmcnt(0); expcnt and lgkmcnt left alone).  Orders
// device-coherent accesses -- agent-scope atomics and atomic loads/stores --
// against later ones without the L2 write-back and invalidation an
// agent-scope fence costs on gfx950.
__device__ __forceinline__ void vm_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ void st_relaxed64(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t now_ticks() { return wall_clock64(); }

// ----------------------------------------------------------- grid sync ----
// Reduction slots carried by the barrier.  Phase g (the work between barrier
// g-1 and barrier g) accumulates into slot[g % 3], which is read after
// barrier g and during phase g+1; the last WG to arrive at barrier g clears
// slot[(g + 2) % 3] (last read in phase g) after releasing the others, off
// the barrier's critical path, and its own wait at barrier g+1 completes
// those stores before phase g+2 uses the slot.
// Each slot, the arrival counter and the generation word sit on cache lines
// of their own (128 bytes), so that polling one does not contend with
// atomics on another.
// The counters a phase hits continuously (queue word, list tail, list head)
// also get a line each.
struct alignas(128) PhaseSlot {
    uint64_t sum0;  // engine-defined sums
    uint64_t sum1;
    uint64_t sum2;
    uint32_t min0;  // engine-defined minimum (cleared to UINT32_MAX)
    uint32_t flag;  // engine-defined OR
    uint32_t pad0;
    uint32_t pad1;
    alignas(128) uint32_t qtail; // items pushed to the shared list of this phase
    alignas(128) uint32_t qhead; // grab counter for reading that list
    // Work list of the peel's bucket / PeelOne modes: items pushed (high 32
    // bits) and pending work (low 32: workgroups still in their own work,
    // which starts at nblocks, plus pushed items not yet fully processed).
    // One word, so a single atomic load tells whether the phase is over.
    alignas(128) uint64_t qword;
};

struct GridSync {
    alignas(128) uint32_t arrive; // arrivals since the launch: barrier g is complete
                                  // at (g+1) * nblocks
    uint32_t nblocks;
    uint64_t timeout_ticks;
    uint64_t t0;
    alignas(128) uint32_t gen; // barriers completed (polled by the waiting workgroups)
    uint32_t err;              // 0 ok, 1 spin timeout, 2 queue overflow, ...
    alignas(128) uint32_t live_next; // reservations of the peel's live-vertex list
    uint32_t pad;
    PhaseSlot slot[3];
};

__device__ __forceinline__ PhaseSlot *slot_of(GridSync *gs, uint32_t phase)
{
    return &gs->slot[phase % 3];
}

__device__ __forceinline__ void slot_clear(PhaseSlot *s, uint32_t nblocks)
{
    __hip_atomic_store(&s->qword, (uint64_t)nblocks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&s->sum0, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&s->sum1, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&s->sum2, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st_relaxed(&s->min0, 0xffffffffu);
    st_relaxed(&s->qtail, 0u);
    st_relaxed(&s->qhead, 0u);
    st_relaxed(&s->flag, 0u);
}

// Called by one thread of the kernel that precedes a persistent launch.
__device__ __forceinline__ void grid_sync_init(GridSync *gs, uint32_t nblocks, uint64_t timeout)
{
    gs->arrive = 0;
    gs->gen = 0;
    gs->err = 0;
    gs->nblocks = nblocks;
    gs->timeout_ticks = timeout;
    gs->t0 = 0;
    gs->live_next = 0;
    gs->pad = 0;
    for (int p = 0; p < 3; p++) {
        PhaseSlot *s = &gs->slot[p];
        s->sum0 = s->sum1 = s->sum2 = 0;
        s->min0 = 0xffffffffu;
        s->qtail = s->qhead = s->flag = 0;
        s->pad0 = s->pad1 = 0;
        s->qword = nblocks;
    }
}

__device__ __forceinline__ bool sync_failed(GridSync *gs)
{
    return ld_relaxed(&gs->err) != 0;
}

__device__ __forceinline__ void sync_fail(GridSync *gs, uint32_t code)
{
    atomicCAS(&gs->err, 0u, code);
}

// Grid-wide barrier.  `gen` is this WG's count of barriers passed so far
// (identical in every WG).  Returns false if the kernel must abort (timeout
// or error raised by some WG); all WGs then leave through the same path.
// kFlush = true publishes plain (non-atomic) stores across the grid with
// agent-scope fences, which on gfx950 write back and invalidate the L2 of
// every XCD.  An engine whose cross-workgroup data is only ever touched by
// atomics and device-coherent (atomic) loads/stores passes false: the barrier
// then only waits for each workgroup's outstanding memory operations, and the
// read-only graph stays cached in L2 from one phase to the next.
template <bool kFlush = true>
__device__ __forceinline__ bool grid_barrier(GridSync *gs, uint32_t &gen)
{
    __shared__ uint32_t s_ok;
    vm_drain();
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        if (kFlush)
            __threadfence();
        uint32_t nb = gs->nblocks;
        uint32_t arrived = atomicAdd(&gs->arrive, 1u);
        if (arrived == (gen + 1) * nb - 1) {
            // every workgroup's stores completed before its arrival
            atomicAdd(&gs->gen, 1u);
            slot_clear(&gs->slot[(gen + 2) % 3], nb);
        } else {
            uint64_t start = now_ticks();
            while (ld_relaxed(&gs->gen) == gen) {
                if (ld_relaxed(&gs->err)) {
                    ok = 0;
                    break;
                }
                if (now_ticks() - start > gs->timeout_ticks) {
                    sync_fail(gs, 1);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (kFlush)
            __threadfence();
        else
            vm_drain();
        if (ld_relaxed(&gs->err))
            ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    gen++;
    return s_ok != 0;
}

// -------------------------------------------------------- block helpers ---
// Exclusive scan of `n` (<= blockDim) values held one per thread; returns the
// thread's exclusive prefix and writes the total to *total (all threads).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *total)
{
    constexpr int NW = BLOCK / KC_WAVE;
    __shared__ uint32_t s_w[NW];
    __shared__ uint32_t s_tot;
    uint32_t incl = wave_incl_scan_u32(x);
    uint32_t w = threadIdx.x / KC_WAVE;
    if (lane_id() == KC_WAVE - 1)
        s_w[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < NW; i++) {
            uint32_t t = s_w[i];
            s_w[i] = run;
            run += t;
        }
        s_tot = run;
    }
    __syncthreads();
    uint32_t r = incl - x + s_w[w];
    *total = s_tot;
    return r;
}

// Upper-bound search: largest i in [0,n) with pre[i] <= e (pre[0] == 0).
__device__ __forceinline__ uint32_t find_item(const uint32_t *pre, uint32_t n, uint32_t e)
{
    uint32_t lo = 0, hi = n; // invariant: pre[lo] <= e < pre[hi] (pre[n] = total)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Launch a persistent grid-barrier kernel.  The grid is sized by the caller
// to at most the occupancy limit (and at most two workgroups per CU), and
// the barrier is the kernel's own (atomics, not the runtime's cooperative
// groups), so a plain launch suffices; the runtime's cooperative launch is
// not used because it is refused (after a long wait) while another process
// holds the device's cooperative-launch resources.  Should the grid ever not
// be resident at once, the barrier's wall-clock limit reports a time-out
// instead of hanging.  Persistent launches of one process are serialized by
// the caller.
inline hipError_t launch_persistent(const void *fn, dim3 grid, dim3 block, void **args,
                                    hipStream_t st)
{
    return hipLaunchKernel(fn, grid, block, args, 0, st);
}

} // namespace kc
