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

template <typename R>
static hipError_t degsort_t(const R *rowptr, const uint32_t *col, uint32_t V, uint64_t nnz,
                            uint32_t *col_out, uint32_t *nd_out, uint32_t *nd_tmp,
                            hipStream_t st)
{
    if (nnz == 0)
        return hipSuccess;
    if (nnz > 0x7fffffffull)
        return hipErrorInvalidValue; // hipCUB item counts are int
    hipLaunchKernelGGL((gather_nbr_degree<R>), dim3(2048), dim3(256), 0, st, rowptr, col, nnz,
                       nd_tmp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    size_t tb = 0;
    e = hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
        nullptr, tb, nd_tmp, nd_out, col, col_out, (int)nnz, (int)V, rowptr, rowptr + 1, 0,
        32, st);
    if (e != hipSuccess)
        return e;
    void *tmp = nullptr;
    e = hipMallocAsync(&tmp, tb ? tb : 8, st);
    if (e != hipSuccess)
        return e;
    e = hipcub::DeviceSegmentedRadixSort::SortPairsDescending(
        tmp, tb, nd_tmp, nd_out, col, col_out, (int)nnz, (int)V, rowptr, rowptr + 1, 0, 32, st);
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
                               uint32_t *nd_tmp, hipStream_t st)
{
    return wide ? degsort_t((const uint64_t *)rowptr, col, V, nnz, col_out, nd_out, nd_tmp, st)
                : degsort_t((const uint32_t *)rowptr, col, V, nnz, col_out, nd_out, nd_tmp, st);
}

} // namespace kc
