// chain.h -- in-order float chains on one wave, shared by the quadtree and
// octree GAT estimates (device code; each including file gets its own copy).
#pragma once

#include "sred_internal.h"

namespace sred {
namespace {

// java.lang.Math.max(float, float): NaN wins, -0 < +0.
__device__ __host__ __forceinline__ float java_maxf(float a, float b)
{
    if (a != a) return a;
    if (a == 0.0f && b == 0.0f && __builtin_signbitf(a)) return b;
    return a >= b ? a : b;
}
__host__ __forceinline__ float java_minf(float a, float b)
{
    if (a != a) return a;
    if (a == 0.0f && b == 0.0f && __builtin_signbitf(b)) return b;
    return a <= b ? a : b;
}

constexpr int kSeg = 512;          // elements per pipeline stage of a node's stream (8 per lane)
constexpr int kStride = kSeg + 4;  // LDS floats per term stream (streams start 4 banks apart)
constexpr int kPad = 32;           // LDS floats after the last stream (read-ahead of a full stage)

// Streams the n elements of a node through a one-wave workgroup: fetch(e)
// loads element e's raw values (e < 0: none), term(raw, t) turns them into NT
// terms (all lanes, 64 elements at a time, parked in LDS), and lane c < NT
// then runs op(c, x) over stream c's terms in element order -- the NT
// in-order chains advance together, one per lane, in the same instructions.
// The next stage's loads are in flight while the current stage's chains run.
template <int NT, class Fetch, class Term, class Op>
__device__ __forceinline__ void stream_chain(int n, Fetch fetch, Term term, Op &op, float *sb)
{
    using Raw = decltype(fetch(0));
    constexpr int Q = kSeg / 64;
    const int lane = threadIdx.x;
    Raw raw[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int e = lane + 64 * q;
        raw[q] = fetch(e < n ? e : -1);
    }
    for (int base = 0; base < n; base += kSeg) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            float t[NT];
            term(raw[q], t);
#pragma unroll
            for (int c = 0; c < NT; c++) sb[c * kStride + lane + 64 * q] = t[c];
        }
        const int nb = base + kSeg;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int e = nb + lane + 64 * q;
            raw[q] = fetch(e < n ? e : -1);
        }
        __syncthreads();
        if (lane < NT) {
            const float *s = sb + lane * kStride;
            const float4 *s4 = reinterpret_cast<const float4 *>(s);
            const int cnt = min(kSeg, n - base);
            auto eat = [&](const float4 &v) {
                op(lane, v.x);
                op(lane, v.y);
                op(lane, v.z);
                op(lane, v.w);
            };
            if (cnt == kSeg) {
                // full stage: the next 32 elements' reads are issued before the
                // current 32 are consumed (the last prefetch reads the padding)
                float4 A[8], B[8];
#pragma unroll
                for (int u = 0; u < 8; u++) A[u] = s4[u];
#pragma unroll 1
                for (int g = 0; g < kSeg / 4; g += 16) {
#pragma unroll
                    for (int u = 0; u < 8; u++) B[u] = s4[g + 8 + u];
#pragma unroll
                    for (int u = 0; u < 8; u++) eat(A[u]);
#pragma unroll
                    for (int u = 0; u < 8; u++) A[u] = s4[g + 16 + u];
#pragma unroll
                    for (int u = 0; u < 8; u++) eat(B[u]);
                }
            } else {
                const int ng = cnt >> 2;
#pragma unroll 8
                for (int g = 0; g < ng; g++) eat(s4[g]);
                for (int t = ng << 2; t < cnt; t++) op(lane, s[t]);
            }
        }
        __syncthreads();
    }
}

// QuadTree_.calculateLTSVariance (QuadTree_.java:326-356), OcTree_.calculateLTSVariance
// (OcTree_.java:382-420) on the sorted squared residuals of each leaf segment.
__global__ __launch_bounds__(64) void k_qt_lts_sum(const float *__restrict__ sorted, const int *__restrict__ off,
                                                   float frac, float *__restrict__ vars)
{
#pragma clang fp contract(off)
    __shared__ __align__(16) float sb[kStride + kPad];
    const int o = off[blockIdx.x], n = off[blockIdx.x + 1] - o;
    const int hh = (int)floorf(frac * (float)n);
    float acc = 0.0f;
    auto fetch = [&](int e) { return e >= 0 && e < hh ? sorted[o + e] : 0.0f; };
    auto term = [](float v, float (&t)[1]) { t[0] = v; };
    auto add = [&](int, float x) { acc += x; };
    stream_chain<1>(hh, fetch, term, add, sb);
    if (threadIdx.x == 0) vars[blockIdx.x] = acc / ((float)hh + kEps);
}

}  // namespace
}  // namespace sred
