// kc_prep.hip -- device-side graph preparation.
//
// Degree-ordered adjacency: every row of the CSR re-ordered by neighbour
// degree, largest first, with the neighbour degrees kept beside it.  HistoCore
// uses it to bound the edges a changed vertex must revisit: a neighbour u can
// only be affected when core[u] > the vertex's new value, and core[u] <=
// deg(u), so only the prefix of the row with deg(u) > new value matters.
// Built once per graph with hipCUB's segmented radix sort (rows are the
// segments); results never depend on neighbour order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "kc_engines.hpp"

namespace kc {

template <typename R>
__global__ void gather_nbr_degree(const R *rowptr, const uint32_t *col, uint64_t nnz,
                                  uint32_t *nd)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < nnz; i += stride) {
        uint32_t u = col[i];
        nd[i] = (uint32_t)(rowptr[u + 1] - rowptr[u]);
    }
}

// Largest row length (atomicMax per wave into *out, cleared before).
template <typename R>
__global__ void max_row_len(const R *rowptr, uint32_t V, uint32_t *out)
{
    uint32_t m = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        uint32_t d = (uint32_t)(rowptr[v + 1] - rowptr[v]);
        m = d > m ? d : m;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t y = __shfl_xor(m, o, 64);
        m = y > m ? y : m;
    }
    if ((threadIdx.x & 63) == 0 && m)
        atomicMax(out, m);
}

// The sort keys are degrees: only the bytes up to the largest one's are
// sorted (24 of 32 bits on R-MAT scale 22 and the Chung-Lu 2^24 graph; the
// exact bit count, 18 on R-MAT, ran three times slower).  With a spare
// adjacency-sized buffer (col_tmp) the sort runs on double buffers, whose
// temporary storage is small; without one, hipCUB allocates a copy of the
// keys and values.  R-MAT s22 15.7 -> 7.3 ms, Chung-Lu 55 -> 45 ms.
template <typename R>
static hipError_t degsort_t(const R *rowptr, const uint32_t *col, uint32_t V, uint64_t nnz,
                            uint32_t *col_out, uint32_t *nd_out, uint32_t *nd_tmp,
                            uint32_t *col_tmp, hipStream_t st)
{
    if (nnz == 0)
        return hipSuccess;
    if (nnz > 0x7fffffffull)
        return hipErrorInvalidValue; // hipCUB item counts are int
    // nd_tmp holds at least nnz + 1 words: the last one takes the largest degree
    uint32_t *dmax_d = nd_tmp + nnz;
    hipError_t e = hipMemsetAsync(dmax_d, 0, 4, st);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL((max_row_len<R>), dim3(1024), dim3(256), 0, st, rowptr, V, dmax_d);
    uint32_t dmax = 0;
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(&dmax, dmax_d, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return e;
    const int end_bit = dmax ? ((32 - __builtin_clz(dmax) + 7) / 8) * 8 : 8;
    hipLaunchKernelGGL((gather_nbr_degree<R>), dim3(2048), dim3(256), 0, st, rowptr, col, nnz,
                       nd_tmp);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    size_t tb = 0;
    void *tmp = nullptr;
    if (col_tmp) {
        if ((e = hipMemcpyAsync(col_tmp, col, nnz * 4, hipMemcpyDeviceToDevice, st)) != hipSuccess)
            return e;
        hipcub::DoubleBuffer<uint32_t> dk(nd_tmp, nd_out), dv(col_tmp, col_out);
        e = hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
            nullptr, tb, dk, dv, (int)nnz, (int)V, rowptr, rowptr + 1, 0, end_bit, st);
        if (e != hipSuccess || (e = hipMallocAsync(&tmp, tb ? tb : 8, st)) != hipSuccess)
            return e;
        e = hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
            tmp, tb, dk, dv, (int)nnz, (int)V, rowptr, rowptr + 1, 0, end_bit, st);
        hipError_t e2 = hipFreeAsync(tmp, st);
        if (e != hipSuccess || (e = e2) != hipSuccess)
            return e;
        if (dk.Current() != nd_out &&
            (e = hipMemcpyAsync(nd_out, dk.Current(), nnz * 4, hipMemcpyDeviceToDevice, st)) !=
                hipSuccess)
            return e;
        if (dv.Current() != col_out)
            e = hipMemcpyAsync(col_out, dv.Current(), nnz * 4, hipMemcpyDeviceToDevice, st);
        return e;
    }
    e = hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
        nullptr, tb, nd_tmp, nd_out, col, col_out, (int)nnz, (int)V, rowptr, rowptr + 1, 0,
        end_bit, st);
    if (e != hipSuccess)
        return e;
    e = hipMallocAsync(&tmp, tb ? tb : 8, st);
    if (e != hipSuccess)
        return e;
    e = hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
        tmp, tb, nd_tmp, nd_out, col, col_out, (int)nnz, (int)V, rowptr, rowptr + 1, 0, end_bit,
        st);
    hipError_t e2 = hipFreeAsync(tmp, st);
    return e != hipSuccess ? e : e2;
}

// hb[i] = rowptr[adj[i]] + adj[i]: where neighbour adj[i]'s histogram starts
// (HistoCore's layout), so an update reaches it without a dependent load
template <typename R>
__global__ void gather_nbr_hbase(const R *rowptr, const uint32_t *adj, uint64_t nnz, uint32_t *hb)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < nnz; i += stride) {
        uint32_t u = adj[i];
        hb[i] = (uint32_t)(rowptr[u] + u);
    }
}

hipError_t build_nbr_hbase(const void *rowptr, bool wide, const uint32_t *adj, uint64_t nnz,
                           uint32_t *hb, hipStream_t st)
{
    if (nnz == 0)
        return hipSuccess;
    if (wide)
        hipLaunchKernelGGL(gather_nbr_hbase<uint64_t>, dim3(2048), dim3(256), 0, st,
                           (const uint64_t *)rowptr, adj, nnz, hb);
    else
        hipLaunchKernelGGL(gather_nbr_hbase<uint32_t>, dim3(2048), dim3(256), 0, st,
                           (const uint32_t *)rowptr, adj, nnz, hb);
    return hipGetLastError();
}

hipError_t build_degree_sorted(const void *rowptr, bool wide, const uint32_t *col, uint32_t V,
                               uint64_t nnz, uint32_t *col_out, uint32_t *nd_out,
                               uint32_t *nd_tmp, uint32_t *col_tmp, hipStream_t st)
{
    return wide ? degsort_t((const uint64_t *)rowptr, col, V, nnz, col_out, nd_out, nd_tmp,
                            col_tmp, st)
                : degsort_t((const uint32_t *)rowptr, col, V, nnz, col_out, nd_out, nd_tmp,
                            col_tmp, st);
}

// The runtime loads a source file's kernels at the first launch of any of
// them; an empty launch at context creation takes that cost out of the first
// decomposition.
__global__ void prep_touch_kernel() {}

hipError_t prep_touch(hipStream_t st)
{
    hipLaunchKernelGGL(prep_touch_kernel, dim3(1), dim3(64), 0, st);
    return hipGetLastError();
}

} // namespace kc
