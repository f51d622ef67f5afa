// stack.hip -- min-max normalisation of stacks, whole or frame by frame.
//
// Utils.normalizeArray (Utils.java:1642-1660) and the normalisation of
// Utils.getInputImage2D(float[], ...) (Utils.java:760-775) both map every
// value to (v - min) / (max - min + EPSILON) in float, min and max over the
// array (java.lang.Math.min / max).  The time-lapse loop of
// BlockRepetition2D_ (BlockRepetition2D_.java:95-137) applies the first to
// the whole stack, the second to each frame, and the first again to the
// stacked maps.  seg_normalize_dev does nseg such normalisations of
// consecutive segments of seg elements in two launches: k_seg_minmax
// (grid y = segment; workgroup reductions of order-preserving integer keys,
// then one atomic min/max per workgroup) and k_seg_remap.  Min and max are
// exact, so the result equals the sequential one bit for bit (finite inputs).
// stack_frames_normalize_dev does the stack normalisation followed by the
// per-frame one in one min/max pass and one remap pass: the map
// v -> fl(fl(v - min) / den) is monotone, so a frame's min and max after the
// first normalisation are the images of its raw min and max.
#include <cfloat>

#include "sred_internal.h"

namespace sred {

namespace {

__global__ void k_seg_mm_init(unsigned int *mm, int nseg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nseg) {
        mm[2 * i] = 0xffffffffu;  // min key
        mm[2 * i + 1] = 0u;       // max key
    }
}

__global__ __launch_bounds__(256) void k_seg_minmax(const float *__restrict__ x, int64_t seg,
                                                    unsigned int *__restrict__ mm)
{
    __shared__ unsigned int smin[4], smax[4];
    const float *p = x + (int64_t)blockIdx.y * seg;
    unsigned int lo = 0xffffffffu, hi = 0u;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < seg; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned int k = f2ord(p[i]);
        lo = min(lo, k);
        hi = max(hi, k);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned int)__shfl_xor((int)lo, off));
        hi = max(hi, (unsigned int)__shfl_xor((int)hi, off));
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smin[wv] = lo;
        smax[wv] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
            lo = min(lo, smin[k]);
            hi = max(hi, smax[k]);
        }
        atomicMin(mm + 2 * blockIdx.y, lo);
        atomicMax(mm + 2 * blockIdx.y + 1, hi);
    }
}

__global__ __launch_bounds__(256) void k_seg_remap(const float *__restrict__ x, float *__restrict__ y, int64_t seg,
                                                   const unsigned int *__restrict__ mm)
{
#pragma clang fp contract(off)
    const float mn = ord2f(mm[2 * blockIdx.y]), mx = ord2f(mm[2 * blockIdx.y + 1]);
    const float den = mx - mn + kEps;
    const int64_t base = (int64_t)blockIdx.y * seg;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < seg; i += (int64_t)gridDim.x * blockDim.x)
        y[base + i] = (x[base + i] - mn) / den;
}

// mm holds the frames' raw (min, max) keys; fm receives, per frame, the
// stack's (min, den) and the frame's normalised (min, den) after it.
__global__ __launch_bounds__(256) void k_stack_frames_mm(const unsigned int *__restrict__ mm, int nseg,
                                                         float4 *__restrict__ fm)
{
#pragma clang fp contract(off)
    __shared__ unsigned int slo[256], shi[256];
    unsigned int lo = 0xffffffffu, hi = 0u;
    for (int i = threadIdx.x; i < nseg; i += 256) {
        lo = min(lo, mm[2 * i]);
        hi = max(hi, mm[2 * i + 1]);
    }
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + s]);
            shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    const float MN = ord2f(slo[0]), MX = ord2f(shi[0]);
    const float DEN = MX - MN + kEps;
    for (int i = threadIdx.x; i < nseg; i += 256) {
        const float mn = (ord2f(mm[2 * i]) - MN) / DEN, mx = (ord2f(mm[2 * i + 1]) - MN) / DEN;
        fm[i] = make_float4(MN, DEN, mn, mx - mn + kEps);
    }
}

__global__ __launch_bounds__(256) void k_stack_frames_remap(const float *__restrict__ x, float *__restrict__ y,
                                                            int64_t seg, const float4 *__restrict__ fm)
{
#pragma clang fp contract(off)
    const float4 f = fm[blockIdx.y];
    const int64_t base = (int64_t)blockIdx.y * seg;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < seg; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = (x[base + i] - f.x) / f.y;
        y[base + i] = (v - f.z) / f.w;
    }
}

}  // namespace

int seg_minmax_dev(sred_ctx *ctx, const float *d_in, int64_t seg, int nseg, unsigned int *mm)
{
    hipLaunchKernelGGL(k_seg_mm_init, dim3((nseg + 255) / 256), dim3(256), 0, ctx->stream, mm, nseg);
    const int per = (int)std::max<int64_t>(1, std::min<int64_t>((seg + 1023) / 1024, std::max(1, 2048 / nseg)));
    hipLaunchKernelGGL(k_seg_minmax, dim3(per, nseg), dim3(256), 0, ctx->stream, d_in, seg, mm);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

int stack_frames_normalize_dev(sred_ctx *ctx, const float *d_in, float *d_out, int64_t seg, int nseg)
{
    if (seg <= 0 || nseg <= 0) return SRED_OK;
    SRED_TRY(ctx->seg_mm.ensure(sizeof(unsigned int) * 2 * (size_t)nseg + sizeof(float4) * (size_t)nseg + 512));
    unsigned int *mm = ctx->seg_mm.as<unsigned int>();
    float4 *fm = reinterpret_cast<float4 *>(ctx->seg_mm.as<char>() + ((sizeof(unsigned int) * 2 * nseg + 255) & ~255));
    hipLaunchKernelGGL(k_seg_mm_init, dim3((nseg + 255) / 256), dim3(256), 0, ctx->stream, mm, nseg);
    const int per = (int)std::max<int64_t>(1, std::min<int64_t>((seg + 1023) / 1024, std::max(1, 2048 / nseg)));
    hipLaunchKernelGGL(k_seg_minmax, dim3(per, nseg), dim3(256), 0, ctx->stream, d_in, seg, mm);
    hipLaunchKernelGGL(k_stack_frames_mm, dim3(1), dim3(256), 0, ctx->stream, mm, nseg, fm);
    hipLaunchKernelGGL(k_stack_frames_remap, dim3(per, nseg), dim3(256), 0, ctx->stream, d_in, d_out, seg, fm);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

int seg_normalize_dev(sred_ctx *ctx, const float *d_in, float *d_out, int64_t seg, int nseg)
{
    if (seg <= 0 || nseg <= 0) return SRED_OK;
    SRED_TRY(ctx->seg_mm.ensure(sizeof(unsigned int) * 2 * (size_t)nseg + 256));
    unsigned int *mm = ctx->seg_mm.as<unsigned int>();
    hipLaunchKernelGGL(k_seg_mm_init, dim3((nseg + 255) / 256), dim3(256), 0, ctx->stream, mm, nseg);
    // about 2048 workgroups in all, at least one per segment
    const int per = (int)std::max<int64_t>(1, std::min<int64_t>((seg + 1023) / 1024, std::max(1, 2048 / nseg)));
    hipLaunchKernelGGL(k_seg_minmax, dim3(per, nseg), dim3(256), 0, ctx->stream, d_in, seg, mm);
    hipLaunchKernelGGL(k_seg_remap, dim3(per, nseg), dim3(256), 0, ctx->stream, d_in, d_out, seg, mm);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

}  // namespace sred
