// Voltage preprocessing before the pursuit (SF = spikesorting_fullpursuit/):
//
//   thresholds (SF/threshold/threshold.py:6-70): per channel nanmedian(|V_c|) (numpy float32
//     median: the mean of the two middle values for an even count, rounded to float32) and the
//     count of samples with |V_c| above a threshold;
//   ZCA whitening (SF/processing/zca.py:5-31 and the caller's `zca_matrix @ voltage`,
//     SF/parallel/spikesorting_parallel.py:295): the channel covariance np.cov(V) (FP64,
//     ddof = 1) and the product Z @ V in FP64 rounded to float32.
//
// Medians are exact order statistics by a grid-wide radix select: 4 passes of 8 bits over
// order-preserving keys, each a histogram kernel (integer atomics: exact counts in any order)
// and a one-block pick kernel.  Sums have fixed orders.  Device pointers, caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/fbp.h"

namespace fbp {
namespace preproc {

constexpr int kHistChunk = 16384;  // samples per block of the histogram kernel
constexpr int kCovSub = 64;        // samples per LDS stage of the covariance kernel (64 x 65 doubles)
constexpr int kCovChunk = 16384;   // samples per partial sum of the covariance

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

__device__ __forceinline__ uint32_t order_key(float v)
{
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Per-row select state (in the workspace): [C] prefix, rank, count, [C][256] histogram.
struct SelState {
    uint32_t* prefix;
    uint32_t* rank;
    uint32_t* count;
    uint32_t* hist;
    float* lo;   // lower middle of each row
    float* out;  // medians
};

__global__ __launch_bounds__(256) void k_count_finite(const float* __restrict__ V, int64_t ld, int64_t N,
                                                      uint32_t* __restrict__ count)
{
    const int c = blockIdx.y;
    const int64_t x0 = (int64_t)blockIdx.x * kHistChunk;
    const int64_t x1 = min(N, x0 + kHistChunk);
    uint32_t k = 0;
    for (int64_t x = x0 + threadIdx.x; x < x1; x += blockDim.x) k += isnan(V[(int64_t)c * ld + x]) ? 0u : 1u;
    for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o, 64);
    if ((threadIdx.x & 63) == 0 && k) atomicAdd(&count[c], k);
}

__global__ __launch_bounds__(256) void k_sel_reset(SelState s, int C)
{
    for (int c = threadIdx.x; c < C; c += blockDim.x) s.count[c] = 0u;
    for (int e = threadIdx.x; e < C * 256; e += blockDim.x) s.hist[e] = 0u;
}

__global__ __launch_bounds__(256) void k_sel_init(SelState s, int C, int upper)
{
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        const uint32_t n = s.count[c];
        s.prefix[c] = 0u;
        s.rank[c] = n == 0u ? 0u : (upper ? n / 2u : (n - 1u) / 2u);
    }
    for (int e = threadIdx.x; e < C * 256; e += blockDim.x) s.hist[e] = 0u;
}

// Digit histogram of |V| over the keys that match the prefix so far.
__global__ __launch_bounds__(256) void k_sel_hist(const float* __restrict__ V, int64_t ld, int64_t N, SelState s,
                                                  int pass)
{
    __shared__ uint32_t h[256];
    const int c = blockIdx.y;
    const int shift = 24 - 8 * pass;
    const uint32_t hmask = pass == 0 ? 0u : (0xffffffffu << (shift + 8));
    const uint32_t pre = s.prefix[c];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t x0 = (int64_t)blockIdx.x * kHistChunk;
    const int64_t x1 = min(N, x0 + kHistChunk);
    for (int64_t x = x0 + threadIdx.x; x < x1; x += blockDim.x) {
        const float v = V[(int64_t)c * ld + x];
        if (isnan(v)) continue;
        const uint32_t k = order_key(fabsf(v));
        if ((k & hmask) == pre) atomicAdd(&h[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&s.hist[c * 256 + threadIdx.x], h[threadIdx.x]);
}

// Bucket of each row's rank; clears the histogram for the next pass.
__global__ __launch_bounds__(64) void k_sel_pick(SelState s, int C, int pass, int last, int upper)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int shift = 24 - 8 * pass;
    uint32_t k = s.rank[c], cum = 0;
    int b = 0;
    for (; b < 255; b++) {
        const uint32_t h = s.hist[c * 256 + b];
        if (cum + h > k) break;
        cum += h;
    }
    s.rank[c] = k - cum;
    s.prefix[c] |= (uint32_t)b << shift;
    for (int e = 0; e < 256; e++) s.hist[c * 256 + e] = 0u;
    if (last) {
        const float v = key_value(s.prefix[c]);
        const uint32_t n = s.count[c];
        if (!upper) {
            s.lo[c] = v;
        } else {
            // numpy float32 median: the middle value, or fl32(fl32(lo + hi) / 2)
            s.out[c] = n == 0u ? NAN : ((n & 1u) ? v : (s.lo[c] + v) * 0.5f);
        }
    }
}

__global__ __launch_bounds__(256) void k_count_above(const float* __restrict__ V, int64_t ld, int64_t N,
                                                     const double* __restrict__ thr, unsigned long long* __restrict__ cnt)
{
    const int c = blockIdx.y;
    const double t = thr[c];
    const int64_t x0 = (int64_t)blockIdx.x * kHistChunk;
    const int64_t x1 = min(N, x0 + kHistChunk);
    unsigned long long k = 0;
    for (int64_t x = x0 + threadIdx.x; x < x1; x += blockDim.x) k += ((double)fabsf(V[(int64_t)c * ld + x]) > t) ? 1ull : 0ull;
    for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o, 64);
    if ((threadIdx.x & 63) == 0 && k) atomicAdd(&cnt[c], k);
}

// Row sums of one sample chunk (FP64, samples ascending).
__global__ __launch_bounds__(256) void k_row_part(const float* __restrict__ V, int64_t ld, int64_t N, int C,
                                                  double* __restrict__ part)
{
    __shared__ double red[256];
    const int c = blockIdx.y;
    const int64_t x0 = (int64_t)blockIdx.x * kCovChunk, x1 = min(N, x0 + kCovChunk);
    double s = 0.0;
    for (int64_t x = x0 + threadIdx.x; x < x1; x += blockDim.x) s += (double)V[(int64_t)c * ld + x];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * C + c] = red[0];
}

__global__ __launch_bounds__(256) void k_row_mean(const double* __restrict__ part, int64_t nchunk, int C, int64_t N,
                                                  double* __restrict__ mean)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    double s = 0.0;
    for (int64_t k = 0; k < nchunk; k++) s += part[k * C + c];
    mean[c] = s / (double)N;
}

// Centred cross products of channel pairs over one sample chunk: a 16 x 16 thread grid owns
// 4 x 4 register blocks of the C x C matrix (C <= 64); samples are staged in LDS, centred.
__global__ __launch_bounds__(256) void k_rowcov_part(const float* __restrict__ V, int64_t ld, int64_t N, int C,
                                                     const double* __restrict__ mean, double* __restrict__ part)
{
    __shared__ double xs[kCovSub][65];
    const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
    const int64_t x0 = (int64_t)blockIdx.x * kCovChunk, x1 = min(N, x0 + kCovChunk);
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
    for (int64_t xb = x0; xb < x1; xb += kCovSub) {
        const int nb = (int)min((int64_t)kCovSub, x1 - xb);
        __syncthreads();
        for (int e = threadIdx.x; e < kCovSub * 64; e += blockDim.x) {
            const int c = e / kCovSub, j = e - c * kCovSub;
            xs[j][c] = (c < C && j < nb) ? (double)V[(int64_t)c * ld + xb + j] - mean[c] : 0.0;
        }
        __syncthreads();
        for (int j = 0; j < nb; j++) {
            double u[4], w[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                u[a] = xs[j][ti * 4 + a];
                w[a] = xs[j][tj * 4 + a];
            }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) acc[a][b] = fma(u[a], w[b], acc[a][b]);
        }
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int i = ti * 4 + a, j = tj * 4 + b;
            if (i < C && j < C) part[((int64_t)blockIdx.x * C + i) * C + j] = acc[a][b];
        }
}

__global__ __launch_bounds__(256) void k_rowcov_final(const double* __restrict__ part, int64_t nchunk, int C,
                                                      double denom, double* __restrict__ cov)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= C * C) return;
    double s = 0.0;
    for (int64_t k = 0; k < nchunk; k++) s += part[k * C * C + e];
    cov[e] = s / denom;
}

// out[c][x] = fl32(sum_k Z[c][k] V[k][x]), FP64 with k ascending; Z staged in LDS.
__global__ __launch_bounds__(256) void k_zca_apply(const double* __restrict__ Z, const float* __restrict__ V,
                                                   int64_t ldv, int64_t N, int C, float* __restrict__ out,
                                                   int64_t ldo)
{
    __shared__ double zs[64 * 64];
    for (int e = threadIdx.x; e < C * C; e += blockDim.x) zs[e] = Z[e];
    __syncthreads();
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    double v[64];
#pragma unroll
    for (int k = 0; k < 64; k++) v[k] = k < C ? (double)V[(int64_t)k * ldv + x] : 0.0;
    for (int c = 0; c < C; c++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 64; k++)
            if (k < C) s = fma(zs[c * C + k], v[k], s);
        out[(int64_t)c * ldo + x] = (float)s;
    }
}

thread_local std::string t_err;

int fail(int code, const char* msg)
{
    t_err = msg;
    return code;
}

int status(const char* where)
{
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return FBP_OK;
    t_err = std::string(where) + ": " + hipGetErrorString(e);
    return FBP_E_HIP;
}

int64_t sel_bytes(int64_t C) { return align256(C * 4) * 3 + align256(C * 256 * 4) + align256(C * 4) * 2; }

}  // namespace preproc
}  // namespace fbp

using namespace fbp::preproc;

extern "C" const char* fbp_preproc_last_error(void) { return t_err.c_str(); }

extern "C" int fbp_preproc_workspace_bytes(int64_t N, int32_t C, int64_t* bytes)
{
    if (N <= 0 || C <= 0 || !bytes) return fail(FBP_E_INVALID, "N and C must be positive");
    const int64_t nchunk = ceil_div(N, kCovChunk);
    const int64_t cov = align256((int64_t)C * 8) + align256(nchunk * (int64_t)C * C * 8);
    *bytes = std::max(sel_bytes(C), cov);
    return FBP_OK;
}

extern "C" int fbp_abs_medians(const float* V, int64_t ld, int64_t N, int32_t C, float* medians, void* work,
                               int64_t work_bytes, void* stream)
{
    if (N <= 0 || C <= 0 || ld < N) return fail(FBP_E_INVALID, "N and C must be positive with ld >= N");
    if (N >= (int64_t)0xffffffffLL) return fail(FBP_E_UNSUPPORTED, "rows of 2^32-1 samples or more");
    if (!V || !medians || !work) return fail(FBP_E_INVALID, "null device pointer");
    int64_t need = 0;
    fbp_preproc_workspace_bytes(N, C, &need);
    if (work_bytes < need) return fail(FBP_E_INVALID, "workspace smaller than fbp_preproc_workspace_bytes()");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* w = static_cast<char*>(work);
    SelState st;
    st.prefix = reinterpret_cast<uint32_t*>(w);
    st.rank = reinterpret_cast<uint32_t*>(w + align256(C * 4));
    st.count = reinterpret_cast<uint32_t*>(w + 2 * align256(C * 4));
    st.hist = reinterpret_cast<uint32_t*>(w + 3 * align256(C * 4));
    st.lo = reinterpret_cast<float*>(w + 3 * align256(C * 4) + align256((int64_t)C * 256 * 4));
    st.out = medians;
    const dim3 gh((unsigned)ceil_div(N, kHistChunk), (unsigned)C);
    // non-NaN samples per row, then the lower and the upper middle order statistic
    hipLaunchKernelGGL(k_sel_reset, dim3(1), dim3(256), 0, s, st, (int)C);
    hipLaunchKernelGGL(k_count_finite, gh, dim3(256), 0, s, V, ld, N, st.count);
    for (int upper = 0; upper < 2; upper++) {
        hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, s, st, (int)C, upper);
        for (int pass = 0; pass < 4; pass++) {
            hipLaunchKernelGGL(k_sel_hist, gh, dim3(256), 0, s, V, ld, N, st, pass);
            hipLaunchKernelGGL(k_sel_pick, dim3((unsigned)ceil_div(C, 64)), dim3(64), 0, s, st, (int)C, pass,
                               pass == 3 ? 1 : 0, upper);
        }
    }
    return status("fbp_abs_medians");
}

extern "C" int fbp_count_above(const float* V, int64_t ld, int64_t N, int32_t C, const double* thresholds,
                               uint64_t* counts, void* stream)
{
    if (N <= 0 || C <= 0 || ld < N) return fail(FBP_E_INVALID, "N and C must be positive with ld >= N");
    if (!V || !thresholds || !counts) return fail(FBP_E_INVALID, "null device pointer");
    hipLaunchKernelGGL(k_count_above, dim3((unsigned)ceil_div(N, kHistChunk), (unsigned)C), dim3(256), 0,
                       static_cast<hipStream_t>(stream), V, ld, N, thresholds,
                       reinterpret_cast<unsigned long long*>(counts));
    return status("fbp_count_above");
}

extern "C" int fbp_channel_cov_rows(const float* V, int64_t ld, int64_t N, int32_t C, int32_t ddof, double* cov,
                                    void* work, int64_t work_bytes, void* stream)
{
    if (N <= 0 || C <= 0 || ld < N) return fail(FBP_E_INVALID, "N and C must be positive with ld >= N");
    if (C > 64) return fail(FBP_E_UNSUPPORTED, "more than 64 channels");
    if (N - ddof <= 0) return fail(FBP_E_INVALID, "N - ddof must be positive");
    if (!V || !cov || !work) return fail(FBP_E_INVALID, "null device pointer");
    int64_t need = 0;
    fbp_preproc_workspace_bytes(N, C, &need);
    if (work_bytes < need) return fail(FBP_E_INVALID, "workspace smaller than fbp_preproc_workspace_bytes()");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* w = static_cast<char*>(work);
    double* mean = reinterpret_cast<double*>(w);
    double* part = reinterpret_cast<double*>(w + align256((int64_t)C * 8));
    const int64_t nchunk = ceil_div(N, kCovChunk);
    hipLaunchKernelGGL(k_row_part, dim3((unsigned)nchunk, (unsigned)C), dim3(256), 0, s, V, ld, N, (int)C, part);
    hipLaunchKernelGGL(k_row_mean, dim3(1), dim3(256), 0, s, part, nchunk, (int)C, N, mean);
    hipLaunchKernelGGL(k_rowcov_part, dim3((unsigned)nchunk), dim3(256), 0, s, V, ld, N, (int)C, mean, part);
    hipLaunchKernelGGL(k_rowcov_final, dim3((unsigned)ceil_div((int64_t)C * C, 256)), dim3(256), 0, s, part, nchunk,
                       (int)C, (double)(N - ddof), cov);
    return status("fbp_channel_cov_rows");
}

extern "C" int fbp_zca_apply(const double* Z, const float* V, int64_t ldv, int64_t N, int32_t C, float* out,
                             int64_t ldo, void* stream)
{
    if (N <= 0 || C <= 0 || ldv < N || ldo < N) return fail(FBP_E_INVALID, "N and C must be positive, strides >= N");
    if (C > 64) return fail(FBP_E_UNSUPPORTED, "more than 64 channels");
    if (!Z || !V || !out) return fail(FBP_E_INVALID, "null device pointer");
    hipLaunchKernelGGL(k_zca_apply, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), Z,
                       V, ldv, N, (int)C, out, ldo);
    return status("fbp_zca_apply");
}
