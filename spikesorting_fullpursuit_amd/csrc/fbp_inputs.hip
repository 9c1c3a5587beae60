// Inputs of the pursuit computed on the GPU: the per-channel noise covariance of clip
// residuals and the templates' quadratic forms (SF = spikesorting_fullpursuit/).
//
//   get_channel_covariance_matrix (SF/overlap/full_binary_pursuit.py:863-923):
//     cov[c] = np.cov(X_c, rowvar=False, ddof=0),  X_c = R[rows][:, c*L : (c+1)*L]
//            = (X_c - mean)^T (X_c - mean) / n           (FP64, as numpy promotes float32 clips)
//   compute_separability_metrics (SF/overlap/neuron_separability.py:158-183):
//     var[u] = sum_c t_uc^T cov[c] t_uc                    (FP64)
//
// Every sum has a fixed order: row chunks of kRowChunk rows are summed in ascending row order
// by one thread each, the chunk partials in ascending chunk order.  Results are identical run
// to run and agree with numpy to FP64 rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/fbp.h"

namespace fbp {
namespace inputs {

constexpr int kRowChunk = 1024;  // rows per partial sum
constexpr int kColTile = 256;    // columns per block of the mean kernel

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

struct Layout {
    int64_t nchunk, off_mean, off_part, bytes;
};

inline Layout layout(int64_t n_rows, int64_t C, int64_t L)
{
    Layout o;
    o.nchunk = std::max<int64_t>(1, ceil_div(n_rows, kRowChunk));
    o.off_mean = 0;                                                  // f64 [C*L] column means
    o.off_part = align256(C * L * 8);                                // f64 [nchunk][C][L][L] (means: [nchunk][C*L])
    o.bytes = o.off_part + align256(o.nchunk * C * L * L * 8);
    return o;
}

__device__ inline int64_t row_of(const int64_t* __restrict__ rows, int64_t i) { return rows ? rows[i] : i; }

// Column sums of one row chunk: thread = column, rows ascending.
__global__ __launch_bounds__(kColTile) void k_col_part(const float* __restrict__ R, int64_t ld, int64_t n_rows,
                                                       const int64_t* __restrict__ rows, int64_t CL,
                                                       double* __restrict__ part)
{
    const int64_t col = (int64_t)blockIdx.x * kColTile + threadIdx.x;
    const int64_t ch = blockIdx.y;
    if (col >= CL) return;
    const int64_t r0 = ch * kRowChunk, r1 = std::min<int64_t>(n_rows, r0 + kRowChunk);
    double s = 0.0;
    for (int64_t i = r0; i < r1; i++) s += (double)R[row_of(rows, i) * ld + col];
    part[ch * CL + col] = s;
}

__global__ __launch_bounds__(256) void k_col_mean(const double* __restrict__ part, int64_t nchunk, int64_t CL,
                                                  int64_t n_rows, double* __restrict__ mean)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= CL) return;
    double s = 0.0;
    for (int64_t k = 0; k < nchunk; k++) s += part[k * CL + col];
    mean[col] = s / (double)n_rows;
}

// Centred cross products of one (channel, row chunk): thread owns entries e = i*L + j of the
// L x L matrix (e = tid, tid + blockDim, ...); the chunk's rows are staged in LDS, centred.
template <int kPer>
__global__ __launch_bounds__(256) void k_cov_part(const float* __restrict__ R, int64_t ld, int64_t n_rows,
                                                  const int64_t* __restrict__ rows, int C, int L,
                                                  const double* __restrict__ mean, double* __restrict__ part)
{
    extern __shared__ double xs[];  // [kStage][L]
    constexpr int kStage = 16;
    const int c = blockIdx.x;
    const int64_t ch = blockIdx.y;
    const int64_t r0 = ch * kRowChunk, r1 = std::min<int64_t>(n_rows, r0 + kRowChunk);
    const int LL = L * L;
    // entries e = tid + 256 q, q < kPer (L * L <= 256 kPer); constant indices keep the
    // accumulators in registers
    double acc[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++) acc[q] = 0.0;
    for (int64_t b = r0; b < r1; b += kStage) {
        const int nb = (int)std::min<int64_t>(kStage, r1 - b);
        __syncthreads();
        for (int e = threadIdx.x; e < nb * L; e += blockDim.x) {
            const int r = e / L, j = e - r * L;
            xs[e] = (double)R[row_of(rows, b + r) * ld + (int64_t)c * L + j] - mean[(int64_t)c * L + j];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const int e = threadIdx.x + q * 256;
            if (e < LL) {
                const int i = e / L, j = e - i * L;
                double a = acc[q];
                for (int r = 0; r < nb; r++) a = fma(xs[r * L + i], xs[r * L + j], a);
                acc[q] = a;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kPer; q++) {
        const int e = threadIdx.x + q * 256;
        if (e < LL) part[(ch * C + c) * (int64_t)LL + e] = acc[q];
    }
}

__global__ __launch_bounds__(256) void k_cov_final(const double* __restrict__ part, int64_t nchunk, int64_t CLL,
                                                   int64_t n_rows, double* __restrict__ cov)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= CLL) return;
    double s = 0.0;
    for (int64_t k = 0; k < nchunk; k++) s += part[k * CLL + e];
    cov[e] = s / (double)n_rows;
}

// var[u] = sum_c t_uc^T cov[c] t_uc: one block per template; thread owns (c, i) pairs and
// forms t_uc[i] * (cov[c] t_uc)[i] with j ascending; the block sums them in a fixed tree.
__global__ __launch_bounds__(256) void k_quad_forms(const double* __restrict__ T, const double* __restrict__ cov,
                                                    int C, int L, double* __restrict__ var)
{
    __shared__ double red[256];
    const int u = blockIdx.x;
    const double* t = T + (int64_t)u * C * L;
    double s = 0.0;
    for (int e = threadIdx.x; e < C * L; e += blockDim.x) {
        const int c = e / L, i = e - c * L;
        const double* row = cov + ((int64_t)c * L + i) * L;
        const double* tc = t + (int64_t)c * L;
        double m = 0.0;
        for (int j = 0; j < L; j++) m = fma(row[j], tc[j], m);
        s = fma(t[e], m, s);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) var[u] = red[0];
}

// ---------------------------------------------------------------------------
// Robust template (calculate_robust_template, SF/processing/clip_utils.py:202-247), per sample
// column x of a unit's clips, with numpy's float32 semantics:
//   med  = median(x)                        (even count: fl32(fl32(a + b) / 2) of the middle two)
//   mad  = fl32(median(|x - med|) / fl32(0.6745))
//   keep = fl32(med - mad) < x < fl32(med + mad)
//   out  = median(x[keep]), or med when nothing is kept
// Exact order statistics by radix select on order-preserving 32-bit keys (4 passes of 8 bits).
// A block owns kRtCols consecutive columns; its threads read whole row segments (coalesced)
// and count key digits per column in LDS.
// ---------------------------------------------------------------------------
constexpr int kRtCols = 32;
constexpr int kRtGroups = 8;  // row groups: blockDim = kRtCols * kRtGroups

__device__ __forceinline__ uint32_t order_key(float v)
{
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

struct RtShared {
    uint32_t hist[kRtCols][256];
    uint32_t prefix[kRtCols], rank[kRtCols], cnt[kRtCols];
    float med[kRtCols], lo[kRtCols], hi[kRtCols], res[kRtCols];
};

// MODE 0: values x; 1: |x - med|; 2: x restricted to lo < x < hi.  Leaves in sh.res[c] the
// k-th smallest value (k = sh.rank[c] on entry, 0-based) of column c.
template <int MODE>
__device__ void rt_select(const float* __restrict__ X, int64_t n, int64_t ld, int64_t col0, int ncol, RtShared& sh)
{
    const int c = threadIdx.x % kRtCols, g = threadIdx.x / kRtCols;
    if (g == 0) sh.prefix[c] = 0u;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        const uint32_t hmask = pass == 0 ? 0u : (0xffffffffu << (shift + 8));
        for (int e = threadIdx.x; e < kRtCols * 256; e += blockDim.x) (&sh.hist[0][0])[e] = 0u;
        __syncthreads();
        if (c < ncol) {
            const uint32_t pre = sh.prefix[c];
            const float m = sh.med[c], lo = sh.lo[c], hi = sh.hi[c];
            for (int64_t r = g; r < n; r += kRtGroups) {
                const float x = X[r * ld + col0 + c];
                float v = x;
                if (MODE == 1) v = fabsf(x - m);
                if (MODE == 2 && !(x > lo && x < hi)) continue;
                const uint32_t k = order_key(v);
                if ((k & hmask) == pre) atomicAdd(&sh.hist[c][(k >> shift) & 255u], 1u);
            }
        }
        __syncthreads();
        if (g == 0 && c < ncol) {
            uint32_t k = sh.rank[c], cum = 0;
            int b = 0;
            for (; b < 255; b++) {
                const uint32_t h = sh.hist[c][b];
                if (cum + h > k) break;
                cum += h;
            }
            sh.rank[c] = k - cum;
            sh.prefix[c] |= (uint32_t)b << shift;
        }
        __syncthreads();
    }
    if (g == 0 && c < ncol) sh.res[c] = key_value(sh.prefix[c]);
    __syncthreads();
}

// numpy median of MODE's values of each column, count sh.cnt[c] (> 0): result in out[c].
template <int MODE>
__device__ void rt_median(const float* __restrict__ X, int64_t n, int64_t ld, int64_t col0, int ncol, RtShared& sh,
                          float* out)
{
    const int c = threadIdx.x % kRtCols, g = threadIdx.x / kRtCols;
    if (g == 0 && c < ncol) sh.rank[c] = (sh.cnt[c] - 1u) / 2u;  // the lower middle
    __syncthreads();
    rt_select<MODE>(X, n, ld, col0, ncol, sh);
    float a = 0.0f;
    if (g == 0 && c < ncol) {
        a = sh.res[c];
        sh.rank[c] = sh.cnt[c] / 2u;  // the upper middle (the same element for an odd count)
    }
    __syncthreads();
    rt_select<MODE>(X, n, ld, col0, ncol, sh);
    if (g == 0 && c < ncol) {
        const float b = sh.res[c];
        out[c] = (sh.cnt[c] & 1u) ? b : (a + b) * 0.5f;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kRtCols* kRtGroups) void k_robust_template(const float* __restrict__ X, int64_t n,
                                                                        int64_t ld, int64_t W,
                                                                        float* __restrict__ out)
{
    __shared__ RtShared sh;
    __shared__ float tmp[kRtCols];
    const int64_t col0 = (int64_t)blockIdx.x * kRtCols;
    const int ncol = (int)min((int64_t)kRtCols, W - col0);
    const int c = threadIdx.x % kRtCols, g = threadIdx.x / kRtCols;
    if (g == 0) {
        sh.cnt[c] = (uint32_t)n;
        sh.med[c] = 0.0f;
        sh.lo[c] = 0.0f;
        sh.hi[c] = 0.0f;
    }
    __syncthreads();
    rt_median<0>(X, n, ld, col0, ncol, sh, tmp);  // sample medians
    if (g == 0 && c < ncol) sh.med[c] = tmp[c];
    __syncthreads();
    rt_median<1>(X, n, ld, col0, ncol, sh, tmp);  // median absolute deviation
    if (g == 0 && c < ncol) {
        const float mad = tmp[c] / 0.6745f;
        sh.lo[c] = sh.med[c] - mad;
        sh.hi[c] = sh.med[c] + mad;
        sh.cnt[c] = 0u;
    }
    __syncthreads();
    // values within one MAD
    if (c < ncol) {
        uint32_t k = 0;
        const float lo = sh.lo[c], hi = sh.hi[c];
        for (int64_t r = g; r < n; r += kRtGroups) {
            const float x = X[r * ld + col0 + c];
            k += (x > lo && x < hi) ? 1u : 0u;
        }
        atomicAdd(&sh.cnt[c], k);
    }
    __syncthreads();
    bool any = (c < ncol) && sh.cnt[c] > 0u;
    if (g == 0 && c < ncol && !any) sh.cnt[c] = 1u;  // placeholder count; the median is used instead
    __syncthreads();
    rt_median<2>(X, n, ld, col0, ncol, sh, tmp);
    if (g == 0 && c < ncol) out[col0 + c] = any ? tmp[c] : sh.med[c];
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

}  // namespace inputs
}  // namespace fbp

using namespace fbp::inputs;

extern "C" const char* fbp_inputs_last_error(void) { return t_err.c_str(); }

extern "C" int fbp_covariance_workspace_bytes(int64_t n_rows, int32_t C, int32_t L, int64_t* bytes)
{
    if (n_rows <= 0 || C <= 0 || L <= 0 || !bytes) return fail(FBP_E_INVALID, "n_rows, C and L must be positive");
    if (L > 128) return fail(FBP_E_UNSUPPORTED, "L above 128 samples");
    *bytes = layout(n_rows, C, L).bytes;
    return FBP_OK;
}

extern "C" int fbp_channel_covariance(const float* R, int64_t ld, int64_t n_total, int32_t C, int32_t L,
                                      const int64_t* rows, int64_t n_rows, double* cov, void* work,
                                      int64_t work_bytes, void* stream)
{
    if (n_rows <= 0 || C <= 0 || L <= 0 || n_total <= 0) return fail(FBP_E_INVALID, "sizes must be positive");
    if (L > 128) return fail(FBP_E_UNSUPPORTED, "L above 128 samples");
    if (ld < (int64_t)C * L) return fail(FBP_E_INVALID, "row stride below C * L");
    if (!rows && n_rows > n_total) return fail(FBP_E_INVALID, "n_rows above n_total without a row list");
    if (!R || !cov || !work) return fail(FBP_E_INVALID, "null device pointer");
    const Layout lay = layout(n_rows, C, L);
    if (work_bytes < lay.bytes) return fail(FBP_E_INVALID, "workspace smaller than fbp_covariance_workspace_bytes()");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* w = static_cast<char*>(work);
    double* mean = reinterpret_cast<double*>(w + lay.off_mean);
    double* part = reinterpret_cast<double*>(w + lay.off_part);
    const int64_t CL = (int64_t)C * L;
    hipLaunchKernelGGL(k_col_part, dim3((unsigned)ceil_div(CL, kColTile), (unsigned)lay.nchunk), dim3(kColTile), 0, s,
                       R, ld, n_rows, rows, CL, part);
    hipLaunchKernelGGL(k_col_mean, dim3((unsigned)ceil_div(CL, 256)), dim3(256), 0, s, part, lay.nchunk, CL, n_rows,
                       mean);
    const dim3 gc((unsigned)C, (unsigned)lay.nchunk);
    if (L <= 64)
        hipLaunchKernelGGL(k_cov_part<16>, gc, dim3(256), (size_t)16 * L * 8, s, R, ld, n_rows, rows, (int)C, (int)L,
                           mean, part);
    else
        hipLaunchKernelGGL(k_cov_part<64>, gc, dim3(256), (size_t)16 * L * 8, s, R, ld, n_rows, rows, (int)C, (int)L,
                           mean, part);
    const int64_t CLL = CL * L;
    hipLaunchKernelGGL(k_cov_final, dim3((unsigned)ceil_div(CLL, 256)), dim3(256), 0, s, part, lay.nchunk, CLL, n_rows,
                       cov);
    return status("fbp_channel_covariance");
}

extern "C" int fbp_robust_template(const float* clips, int64_t n, int64_t ld, int64_t W, float* out, void* stream)
{
    if (n <= 0 || W <= 0 || ld < W) return fail(FBP_E_INVALID, "n and W must be positive with ld >= W");
    if (n >= (int64_t)0xffffffffLL) return fail(FBP_E_UNSUPPORTED, "more than 2^32-2 clips");
    if (!clips || !out) return fail(FBP_E_INVALID, "null device pointer");
    hipLaunchKernelGGL(k_robust_template, dim3((unsigned)ceil_div(W, kRtCols)), dim3(kRtCols * kRtGroups), 0,
                       static_cast<hipStream_t>(stream), clips, n, ld, W, out);
    return status("fbp_robust_template");
}

extern "C" int fbp_template_variances(const double* T, int32_t U, int32_t C, int32_t L, const double* cov,
                                      double* var, void* stream)
{
    if (U <= 0 || C <= 0 || L <= 0) return fail(FBP_E_INVALID, "U, C and L must be positive");
    if (!T || !cov || !var) return fail(FBP_E_INVALID, "null device pointer");
    hipLaunchKernelGGL(k_quad_forms, dim3((unsigned)U), dim3(256), 0, static_cast<hipStream_t>(stream), T, cov, (int)C,
                       (int)L, var);
    return status("fbp_template_variances");
}
