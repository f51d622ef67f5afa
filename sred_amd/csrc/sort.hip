// sort.hip -- order of the valid blocks of a global problem.
//
// The global kernels weight every pair by exp2(-(t_i - t_j)^2) with
// t = sigma * tscale (Pearson, SSIM and the weight sums).  Sorting the valid
// blocks by sigma makes every panel, sweep and column tile of the contraction
// cover a narrow, ascending range of t, so whole panel pairs and tiles whose
// weights are all below 2^-40 can be recognised from their first and last
// entries and skipped.  Sums per reference pixel do not depend on the order
// beyond rounding.  A stable LSD radix sort on the float bits (sigma >= 0, so
// the unsigned order is the float order) keeps ties in pixel order, so the
// result is deterministic.
#include <hipcub/hipcub.hpp>

#include "sred_internal.h"

namespace sred {

__global__ __launch_bounds__(256) void k_sigma_keys(const float *__restrict__ stds, const int *__restrict__ idx, int Nv,
                                                    unsigned int *__restrict__ keys)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v < Nv) keys[v] = __float_as_uint(stds[idx[v]]);
}

int sort_valid_by_sigma(sred_ctx *ctx, DevBuf &buf, const float *d_stds, const int *d_idx, int Nv, int **d_sorted)
{
    size_t tmp = 0;
    SRED_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const unsigned int *)nullptr,
                                                      (unsigned int *)nullptr, (const int *)nullptr, (int *)nullptr,
                                                      Nv, 0, 32, ctx->stream));
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bk = al(sizeof(unsigned int) * (size_t)Nv), bv = al(sizeof(int) * (size_t)Nv + 256);
    SRED_TRY(buf.ensure(2 * bk + bv + al(tmp)));
    char *base = buf.as<char>();
    unsigned int *kin = (unsigned int *)base, *kout = (unsigned int *)(base + bk);
    int *vout = (int *)(base + 2 * bk);
    void *d_tmp = base + 2 * bk + bv;
    hipLaunchKernelGGL(k_sigma_keys, dim3((Nv + 255) / 256), dim3(256), 0, ctx->stream, d_stds, d_idx, Nv, kin);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp, kin, kout, d_idx, vout, Nv, 0, 32, ctx->stream));
    *d_sorted = vout;
    return SRED_OK;
}

}  // namespace sred
