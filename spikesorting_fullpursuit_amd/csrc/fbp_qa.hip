// Template QA: the shift search of find_overlap_templates (SF/c_cython/sort_cython.pyx:333-462,
// SF = spikesorting_fullpursuit/), which asks whether a unit's template is the sum of two
// other units' templates at some pair of shifts.
//
// For a (test, n1, n2) triple and shifts s1, s2 in [-n_pre, n_post]:
//   R(s1, s2) = sum_e (test_e - (a_e + b_e))^2,   a = t_n1 shifted by s1, b = t_n2 by s2
// where a template shifted by s holds t[c][j - s] at sample j of channel c when 0 <= j - s < L,
// else 0.  The reference evaluates R in FP64 with numpy: the sum a + b, the difference, the
// square, then np.sum, which is numpy's pairwise summation (blocks of at most 128 elements
// summed by 8 interleaved accumulators, halves split at multiples of 8).  The kernel repeats
// every one of those roundings in the same order, so R is bit-identical and the minimum (the
// first one in s1-major, s2-minor order) is the reference's.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/fbp.h"

namespace fbp {
namespace qa {

constexpr int kMaxDepth = 16;

struct Elem {
    const float* t;   // test template [C*L]
    const float* a;   // template n1 [C*L]
    const float* b;   // template n2 [C*L]
    int L, s1, s2;
    __device__ __forceinline__ double operator()(int e) const
    {
        const int c = e / L, j = e - c * L;
        const int ja = j - s1, jb = j - s2;
        const double va = (ja >= 0 && ja < L) ? (double)a[c * L + ja] : 0.0;
        const double vb = (jb >= 0 && jb < L) ? (double)b[c * L + jb] : 0.0;
        const double d = (double)t[e] - (va + vb);
        return d * d;
    }
};

// numpy's pairwise sum of f(e0 .. e0+n-1).
template <int D>
__device__ double pairwise(const Elem& f, int e0, int n)
{
    if (n < 8) {
        double r = -0.0;
        for (int i = 0; i < n; i++) r += f(e0 + i);
        return r;
    }
    if (n <= 128) {
        double r[8];
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = f(e0 + j);
        int i = 8;
        for (; i < n - (n % 8); i += 8)
#pragma unroll
            for (int j = 0; j < 8; j++) r[j] += f(e0 + i + j);
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; i++) res += f(e0 + i);
        return res;
    }
    int n2 = n / 2;
    n2 -= n2 % 8;
    if constexpr (D + 1 < kMaxDepth) {
        return pairwise<D + 1>(f, e0, n2) + pairwise<D + 1>(f, e0 + n2, n - n2);
    } else {
        return NAN;  // unreachable: the host limits C * L
    }
}

// One block per triple; thread q takes shift pairs q, q + blockDim, ... (q = s1i * S + s2i) and
// keeps the first minimum; the block then takes the lexicographic minimum (value, q).
__global__ __launch_bounds__(256) void k_overlap_residuals(const float* __restrict__ T, int C, int L, int n_pre, int S,
                                                           const int32_t* __restrict__ triples, int64_t n_triples,
                                                           double* __restrict__ best_v, int32_t* __restrict__ best_q)
{
    extern __shared__ float tl[];  // [3][C*L]: test, n1, n2
    __shared__ double rv[256];
    __shared__ int rq[256];
    const int CL = C * L;
    for (int64_t k = blockIdx.x; k < n_triples; k += gridDim.x) {
        __syncthreads();
        for (int r = 0; r < 3; r++) {
            const int u = triples[k * 3 + r];
            for (int e = threadIdx.x; e < CL; e += blockDim.x) tl[r * CL + e] = T[(int64_t)u * CL + e];
        }
        __syncthreads();
        double bv = INFINITY;
        int bq = 0x7fffffff;
        for (int q = threadIdx.x; q < S * S; q += blockDim.x) {
            const Elem f{tl, tl + CL, tl + 2 * CL, L, q / S - n_pre, q % S - n_pre};
            const double v = pairwise<0>(f, 0, CL);
            if (v < bv) { bv = v; bq = q; }
        }
        rv[threadIdx.x] = bv;
        rq[threadIdx.x] = bq;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                const double ov = rv[threadIdx.x + w];
                const int oq = rq[threadIdx.x + w];
                if (ov < rv[threadIdx.x] || (ov == rv[threadIdx.x] && oq < rq[threadIdx.x])) {
                    rv[threadIdx.x] = ov;
                    rq[threadIdx.x] = oq;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            best_v[k] = rv[0];
            best_q[k] = rq[0];
        }
    }
}

thread_local std::string t_err;

}  // namespace qa
}  // namespace fbp

using namespace fbp::qa;

extern "C" const char* fbp_qa_last_error(void) { return t_err.c_str(); }

extern "C" int fbp_overlap_residuals(const float* T, int32_t U, int32_t C, int32_t L, int32_t n_pre, int32_t n_post,
                                     const int32_t* triples, int64_t n_triples, double* best_value,
                                     int32_t* best_shift_pair, void* stream)
{
    if (U <= 0 || C <= 0 || L <= 0 || n_pre < 0 || n_post < 0 || n_triples < 0) {
        t_err = "U, C, L must be positive and n_pre, n_post, n_triples >= 0";
        return FBP_E_INVALID;
    }
    if ((int64_t)C * L > 16384) {
        t_err = "C * L above 16384 (LDS holds three templates)";
        return FBP_E_UNSUPPORTED;
    }
    if (n_triples == 0) return FBP_OK;
    if (!T || !triples || !best_value || !best_shift_pair) {
        t_err = "null device pointer";
        return FBP_E_INVALID;
    }
    const int S = n_pre + n_post + 1;
    const size_t lds = (size_t)3 * C * L * 4;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_overlap_residuals),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const unsigned g = (unsigned)std::min<int64_t>(n_triples, 1 << 20);
    hipLaunchKernelGGL(k_overlap_residuals, dim3(g), dim3(256), lds, static_cast<hipStream_t>(stream), T, (int)C,
                       (int)L, (int)n_pre, S, triples, n_triples, best_value, best_shift_pair);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        t_err = std::string("fbp_overlap_residuals: ") + hipGetErrorString(e);
        return FBP_E_HIP;
    }
    return FBP_OK;
}
