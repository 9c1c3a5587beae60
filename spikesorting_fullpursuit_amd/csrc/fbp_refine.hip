// Residual energy and its gradients at fixed spike times (template refinement).
//
// The reference has no counterpart: its residual pass (SF/kernels/binary_pursuit.cl:57-102,
// SF = spikesorting_fullpursuit/) subtracts each found template once and never differentiates.
// Here every spike k carries an amplitude a_k:
//
//   R[c][x]        = V[c][x] - sum_k a_k T[lab_k][c][x - p_k]     (0 <= x - p_k < L, x < N)
//   E              = 0.5 * sum R^2                                (FP64 partial sums)
//   dE/dT[u][c][i] = - sum_{k in unit u, p_k + i < N} a_k R[c][p_k + i]
//   dE/da_k        = - sum_{c, i : p_k + i < N} T[lab_k][c][i] R[c][p_k + i]
//
// Every sum has a fixed order (no floating-point atomics), so results are run-to-run
// identical.  All pointers are device pointers; work is ordered on the caller's stream and
// nothing here allocates, synchronises or clears memory with hipMemsetAsync, so a caller may
// capture the calls in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/fbp.h"

namespace fbp {
namespace refine {

constexpr int kTile = 256;   // samples per block of the residual kernel
constexpr int kSpan = 512;   // spikes staged in LDS per round of the residual kernel
constexpr int kChunk = 256;  // spikes per partial sum of the template gradient
constexpr int kRed = 1024;   // threads of the final reduction
constexpr int kCG = 16;      // channels held in registers per thread by the residual kernel

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

struct Layout {
    int64_t nblk;      // residual blocks
    int64_t nchunk;    // upper bound of template-gradient chunks: ceil(n / kChunk) + U
    int64_t off_flag;  // u32 [nblk] per-block input check
    int64_t off_cs;    // u32 [U + 1] first chunk of each unit
    int64_t off_part;  // f32 [nchunk][C * L] partial sums
    int64_t bytes;
};

inline Layout layout(int64_t N, int64_t n, int64_t U, int64_t C, int64_t L)
{
    Layout o;
    o.nblk = ceil_div(N, kTile);
    o.nchunk = ceil_div(n, kChunk) + U;
    const int64_t fwd = align256(o.nblk * 8) + align256(o.nblk * 4);
    o.off_flag = align256(o.nblk * 8);
    o.off_cs = 0;
    o.off_part = align256((U + 1) * 4);
    const int64_t bwd = o.off_part + align256(o.nchunk * C * L * 4);
    o.bytes = fwd > bwd ? fwd : bwd;
    return o;
}

// First k in [0, n) with a[k] >= key (n when none); a is ascending.
__device__ inline int64_t lower_bound_u32(const uint32_t* __restrict__ a, int64_t n, int64_t key)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ inline double wave_sum_f64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline float wave_sum_f32(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One block per 256 consecutive samples, all channels.  Each (c, x) is owned by one thread,
// which subtracts the covering spikes in list order (fmaf(-a, T, r)).  The spikes that can
// cover the tile, [first p >= x0 - L + 1, first p >= x0 + 256), are staged in LDS in rounds of
// kSpan; between rounds the partial residual lives in R.  Each block also checks a slice of
// the spike list (ascending onsets, labels below U) and writes the outcome to flag[b].
__global__ __launch_bounds__(kTile) void k_residual_energy(const float* __restrict__ V, int64_t ldv, int64_t N, int C,
                                                           const float* __restrict__ T, int U, int L,
                                                           const uint32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ lab,
                                                           const float* __restrict__ amp, int64_t nsp,
                                                           float* __restrict__ R, int64_t ldr,
                                                           double* __restrict__ epart, uint32_t* __restrict__ flag)
{
    __shared__ int64_t s_range[2];
    __shared__ uint32_t s_p[kSpan];
    __shared__ uint32_t s_l[kSpan];
    __shared__ float s_a[kSpan];
    __shared__ double s_red[kTile / 64];

    const int tid = threadIdx.x;
    const int64_t x0 = (int64_t)blockIdx.x * kTile;
    const int64_t x = x0 + tid;

    // input check of this block's slice of the spike list
    bool bad = false;
    {
        const int64_t per = ceil_div(nsp, (int64_t)gridDim.x);
        const int64_t k0 = (int64_t)blockIdx.x * per;
        const int64_t k1 = min(nsp, k0 + per);
        for (int64_t k = k0 + tid; k < k1; k += kTile) {
            if (lab[k] >= (uint32_t)U) bad = true;
            if (k + 1 < nsp && idx[k] > idx[k + 1]) bad = true;
        }
    }
    const int any_bad = __syncthreads_or(bad ? 1 : 0);

    if (tid == 0) {
        s_range[0] = lower_bound_u32(idx, nsp, x0 - L + 1);
        s_range[1] = lower_bound_u32(idx, nsp, x0 + kTile);
    }
    __syncthreads();
    const int64_t klo = s_range[0];
    const int64_t khi = max(klo, s_range[1]);

    double e = 0.0;
    const int64_t nround = max((int64_t)1, ceil_div(khi - klo, kSpan));
    for (int64_t r = 0; r < nround; r++) {
        const int64_t kb = klo + r * kSpan;
        const int cnt = (int)min((int64_t)kSpan, khi - kb);
        __syncthreads();
        for (int j = tid; j < cnt; j += kTile) {
            s_p[j] = idx[kb + j];
            s_l[j] = lab[kb + j];
            s_a[j] = amp[kb + j];
        }
        __syncthreads();
        if (x < N) {
            const bool last = (r == nround - 1);
            // kCG channels at a time in registers: their loads are issued together and each
            // covering spike is tested once per group
            for (int c0 = 0; c0 < C; c0 += kCG) {
                float v[kCG];
#pragma unroll
                for (int cc = 0; cc < kCG; cc++)
                    if (c0 + cc < C) v[cc] = (r == 0) ? V[(int64_t)(c0 + cc) * ldv + x] : R[(int64_t)(c0 + cc) * ldr + x];
                for (int j = 0; j < cnt; j++) {
                    const int64_t d = x - (int64_t)s_p[j];
                    const uint32_t u = s_l[j];
                    if (d < 0 || d >= L || u >= (uint32_t)U) continue;
                    const float na = -s_a[j];
                    const float* t = T + ((int64_t)u * C + c0) * L + d;
#pragma unroll
                    for (int cc = 0; cc < kCG; cc++)
                        if (c0 + cc < C) v[cc] = fmaf(na, t[(int64_t)cc * L], v[cc]);
                }
#pragma unroll
                for (int cc = 0; cc < kCG; cc++) {
                    if (c0 + cc < C) {
                        R[(int64_t)(c0 + cc) * ldr + x] = v[cc];
                        if (last) e += (double)v[cc] * (double)v[cc];
                    }
                }
            }
        }
    }

    e = wave_sum_f64(e);
    if ((tid & 63) == 0) s_red[tid >> 6] = e;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < kTile / 64; w++) s += s_red[w];
        epart[blockIdx.x] = s;
        flag[blockIdx.x] = any_bad ? 1u : 0u;
    }
}

// E = 0.5 * sum of the block partials in a fixed order; NaN when any block found the spike
// list unsorted or a label out of range.
__global__ __launch_bounds__(kRed) void k_energy_final(const double* __restrict__ epart,
                                                       const uint32_t* __restrict__ flag, int64_t nblk,
                                                       double* __restrict__ energy)
{
    __shared__ double s_sum[kRed / 64];
    const int tid = threadIdx.x;
    double s = 0.0;
    bool bad = false;
    for (int64_t b = tid; b < nblk; b += kRed) {
        s += epart[b];
        bad = bad || flag[b] != 0u;
    }
    const int any_bad = __syncthreads_or(bad ? 1 : 0);
    s = wave_sum_f64(s);
    if ((tid & 63) == 0) s_sum[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < kRed / 64; w++) t += s_sum[w];
        energy[0] = any_bad ? (double)NAN : 0.5 * t;
    }
}

// dE/da_k: one wave per spike, lanes over the template's samples, channel by channel.
__global__ __launch_bounds__(256) void k_grad_amp(const float* __restrict__ R, int64_t ldr, int64_t N, int C,
                                                  const float* __restrict__ T, int U, int L,
                                                  const uint32_t* __restrict__ idx, const uint32_t* __restrict__ lab,
                                                  int64_t nsp, float* __restrict__ ga)
{
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= nsp) return;
    const uint32_t u = lab[k];
    const int64_t p = idx[k];
    float acc = 0.0f;
    if (u < (uint32_t)U) {
        const int n = (int)min((int64_t)L, N - p);
        for (int c = 0; c < C; c++) {
            const float* t = T + ((int64_t)u * C + c) * L;
            const float* r = R + (int64_t)c * ldr + p;
            for (int i = lane; i < n; i += 64) acc = fmaf(t[i], r[i], acc);
        }
    }
    acc = wave_sum_f32(acc);
    if (lane == 0) ga[k] = (u < (uint32_t)U) ? -acc : (float)NAN;
}

__device__ inline uint32_t unit_chunks(const uint32_t* __restrict__ off, int u)
{
    return off[u + 1] > off[u] ? (off[u + 1] - off[u] + kChunk - 1) / kChunk : 0u;
}

// First chunk of each unit: cs[u] = sum_{u' < u} ceil(n_u' / kChunk), cs[U] = total.
__global__ __launch_bounds__(kRed) void k_chunk_table(const uint32_t* __restrict__ off, int U, uint32_t* __restrict__ cs)
{
    __shared__ uint32_t s_tot[kRed];
    const int tid = threadIdx.x;
    const int per = (U + kRed - 1) / kRed;
    const int u0 = min(U, tid * per), u1 = min(U, u0 + per);
    uint32_t loc = 0;
    for (int u = u0; u < u1; u++) loc += unit_chunks(off, u);
    s_tot[tid] = loc;
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int t = 0; t < kRed; t++) {
            const uint32_t v = s_tot[t];
            s_tot[t] = run;
            run += v;
        }
        cs[U] = run;
    }
    __syncthreads();
    uint32_t run = s_tot[tid];
    for (int u = u0; u < u1; u++) {
        cs[u] = run;
        run += unit_chunks(off, u);
    }
}

// Partial template gradients: block (chunk g, element tile).  Chunk g is the j-th group of
// kChunk spikes of unit u (in the caller's by-unit order); thread t of the tile owns element
// e = (c, i) of the flattened [C][L] template and sums a_k R[c][p_k + i] over the chunk.
__global__ __launch_bounds__(256) void k_grad_tmpl_part(const float* __restrict__ R, int64_t ldr, int64_t N, int C,
                                                        int U, int L, const uint32_t* __restrict__ idx,
                                                        const float* __restrict__ amp, int64_t nsp,
                                                        const uint32_t* __restrict__ by_unit,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ cs, float* __restrict__ part)
{
    __shared__ uint32_t s_p[kChunk];
    __shared__ float s_a[kChunk];
    __shared__ int s_u;
    const uint32_t g = blockIdx.x;
    const int tid = threadIdx.x;
    if (g >= cs[U]) return;
    if (tid == 0) {
        // last unit u with cs[u] <= g (units without spikes have cs[u] == cs[u + 1])
        int lo = 0, hi = U;  // invariant: cs[lo] <= g < cs[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cs[mid] <= g) lo = mid;
            else hi = mid;
        }
        s_u = lo;
    }
    __syncthreads();
    const int u = s_u;
    const int64_t k0 = (int64_t)off[u] + (int64_t)(g - cs[u]) * kChunk;
    const int cnt = (int)max((int64_t)0, min((int64_t)kChunk, min((int64_t)off[u + 1], nsp) - k0));
    for (int j = tid; j < cnt; j += blockDim.x) {
        const uint32_t k = by_unit[k0 + j];
        const bool ok = (int64_t)k < nsp;
        s_p[j] = ok ? idx[k] : 0xffffffffu;
        s_a[j] = ok ? amp[k] : 0.0f;
    }
    __syncthreads();
    const int64_t CL = (int64_t)C * L;
    const int64_t e = (int64_t)blockIdx.y * blockDim.x + tid;
    if (e >= CL) return;
    const int c = (int)(e / L), i = (int)(e - (int64_t)c * L);
    const float* r = R + (int64_t)c * ldr + i;
    float acc = 0.0f;
    for (int j = 0; j < cnt; j++) {
        const int64_t x = (int64_t)s_p[j] + i;
        if (x < N) acc = fmaf(s_a[j], r[s_p[j]], acc);
    }
    part[(int64_t)g * CL + e] = acc;
}

// dE/dT[u][e] = -(sum of unit u's chunk partials in chunk order).
__global__ __launch_bounds__(256) void k_grad_tmpl_final(const float* __restrict__ part,
                                                         const uint32_t* __restrict__ cs, int U, int64_t CL,
                                                         uint32_t nchunk, float* __restrict__ gT)
{
    const int64_t tot = (int64_t)U * CL;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int u = (int)(q / CL);
        const int64_t e = q - (int64_t)u * CL;
        float s = 0.0f;
        const uint32_t g1 = min(cs[u + 1], nchunk);
        for (uint32_t g = cs[u]; g < g1; g++) s += part[(int64_t)g * CL + e];
        gT[q] = 0.0f - s;
    }
}

thread_local std::string t_err;

// Scratch of fbp_refine_forward_noalloc: the residual kernel's per-block FP64 energy
// partials (kNoAllocMaxBlocks doubles) followed by its per-block input-check flags
// (kNoAllocMaxBlocks u32), i.e. recordings of up to 2^26 samples.  Every block writes both
// entries before k_energy_final reads them, so nothing needs clearing between calls.  Each
// device holds its own copy until the process ends; host code uses the per-device address
// from hipGetSymbolAddress, never the array's name.
constexpr int64_t kNoAllocMaxBlocks = (int64_t)1 << 18;
constexpr int64_t kNoAllocFlagOff = kNoAllocMaxBlocks * 8;
constexpr int64_t kNoAllocBytes = kNoAllocFlagOff + kNoAllocMaxBlocks * 4;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int fail(int code, const char* msg)
{
    t_err = msg;
    return code;
}

int check_shape(int64_t N, int32_t C, int32_t U, int32_t L, int64_t nsp)
{
    if (N <= 0 || C <= 0 || U <= 0 || L <= 0 || nsp < 0) return fail(FBP_E_INVALID, "N, C, U, L must be positive and n_spikes >= 0");
    if (N >= (int64_t)0xffffffffLL) return fail(FBP_E_UNSUPPORTED, "recording longer than 2^32-1 samples");
    if (nsp >= (int64_t)0xffffffffLL) return fail(FBP_E_UNSUPPORTED, "more than 2^32-2 spikes");
    // the template-gradient grid has ceil(C * L / 256) blocks in y (at most 65535)
    if ((int64_t)C * L > ((int64_t)1 << 24)) return fail(FBP_E_UNSUPPORTED, "C * L above 2^24");
    return FBP_OK;
}

int launch_status(const char* where)
{
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return FBP_OK;
    t_err = std::string(where) + ": " + hipGetErrorString(e);
    return FBP_E_HIP;
}

}  // namespace refine
}  // namespace fbp

using namespace fbp::refine;

extern "C" const char* fbp_refine_last_error(void) { return t_err.c_str(); }

extern "C" int fbp_refine_workspace_bytes(int64_t N, int64_t n_spikes, int32_t U, int32_t C, int32_t L, int64_t* bytes)
{
    const int rc = check_shape(N, C, U, L, n_spikes);
    if (rc != FBP_OK) return rc;
    if (!bytes) return fail(FBP_E_INVALID, "null output");
    *bytes = layout(N, n_spikes, U, C, L).bytes;
    return FBP_OK;
}

extern "C" int fbp_refine_forward(const float* V, int64_t ldv, int64_t N, int32_t C, const float* T, int32_t U,
                                  int32_t L, const uint32_t* idx, const uint32_t* lab, const float* amp,
                                  int64_t n_spikes, float* R, int64_t ldr, double* energy, void* work,
                                  int64_t work_bytes, void* stream)
{
    int rc = check_shape(N, C, U, L, n_spikes);
    if (rc != FBP_OK) return rc;
    if (!V || !T || !R || !energy || !work) return fail(FBP_E_INVALID, "null device pointer");
    if (n_spikes > 0 && (!idx || !lab || !amp)) return fail(FBP_E_INVALID, "null spike arrays");
    if (ldv < N || ldr < N) return fail(FBP_E_INVALID, "row strides must be >= N");
    const Layout lay = layout(N, n_spikes, U, C, L);
    if (work_bytes < lay.bytes) return fail(FBP_E_INVALID, "workspace smaller than fbp_refine_workspace_bytes()");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* w = static_cast<char*>(work);
    double* epart = reinterpret_cast<double*>(w);
    uint32_t* flag = reinterpret_cast<uint32_t*>(w + lay.off_flag);
    hipLaunchKernelGGL(k_residual_energy, dim3((unsigned)lay.nblk), dim3(kTile), 0, s, V, ldv, N, (int)C, T, (int)U,
                       (int)L, idx, lab, amp, n_spikes, R, ldr, epart, flag);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kRed), 0, s, epart, flag, lay.nblk, energy);
    return launch_status("fbp_refine_forward");
}

// Device address of `scratch` on the current device, looked up once per device.
static int scratch_address(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        t_err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return FBP_E_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) return fail(FBP_E_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            t_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return FBP_E_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FBP_OK;
}

extern "C" int fbp_refine_forward_noalloc(const float* V, int64_t ldv, int64_t N, int32_t C, const float* T,
                                          int32_t U, int32_t L, const uint32_t* idx, const uint32_t* lab,
                                          const float* amp, int64_t n_spikes, float* R, int64_t ldr, double* energy,
                                          void* stream)
{
    int rc = check_shape(N, C, U, L, n_spikes);
    if (rc != FBP_OK) return rc;
    if (ceil_div(N, kTile) > kNoAllocMaxBlocks)
        return fail(FBP_E_UNSUPPORTED, "fbp_refine_forward_noalloc: N above 67108864 samples (its static scratch "
                                       "holds 262144 blocks of 256 samples); use fbp_refine_forward");
    if (!V || !T || !R || !energy) return fail(FBP_E_INVALID, "null device pointer");
    if (n_spikes > 0 && (!idx || !lab || !amp)) return fail(FBP_E_INVALID, "null spike arrays");
    if (ldv < N || ldr < N) return fail(FBP_E_INVALID, "row strides must be >= N");
    void* base = nullptr;
    rc = scratch_address(&base);
    if (rc != FBP_OK) return rc;
    const int64_t nblk = ceil_div(N, kTile);
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* w = static_cast<char*>(base);
    double* epart = reinterpret_cast<double*>(w);
    uint32_t* flag = reinterpret_cast<uint32_t*>(w + kNoAllocFlagOff);
    hipLaunchKernelGGL(k_residual_energy, dim3((unsigned)nblk), dim3(kTile), 0, s, V, ldv, N, (int)C, T, (int)U,
                       (int)L, idx, lab, amp, n_spikes, R, ldr, epart, flag);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kRed), 0, s, epart, flag, nblk, energy);
    return launch_status("fbp_refine_forward_noalloc");
}

extern "C" int fbp_refine_backward(const float* R, int64_t ldr, int64_t N, int32_t C, const float* T, int32_t U,
                                   int32_t L, const uint32_t* idx, const uint32_t* lab, const float* amp,
                                   int64_t n_spikes, const uint32_t* by_unit, const uint32_t* unit_offsets,
                                   float* grad_T, float* grad_amp, void* work, int64_t work_bytes, void* stream)
{
    int rc = check_shape(N, C, U, L, n_spikes);
    if (rc != FBP_OK) return rc;
    if (!R || !T || !unit_offsets || !work) return fail(FBP_E_INVALID, "null device pointer");
    if (!grad_T && !grad_amp) return fail(FBP_E_INVALID, "no gradient requested");
    if (n_spikes > 0 && (!idx || !lab || !amp || (grad_T && !by_unit))) return fail(FBP_E_INVALID, "null spike arrays");
    if (ldr < N) return fail(FBP_E_INVALID, "row stride must be >= N");
    const Layout lay = layout(N, n_spikes, U, C, L);
    if (work_bytes < lay.bytes) return fail(FBP_E_INVALID, "workspace smaller than fbp_refine_workspace_bytes()");
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* w = static_cast<char*>(work);
    const int64_t CL = (int64_t)C * L;
    if (grad_amp && n_spikes > 0)
        hipLaunchKernelGGL(k_grad_amp, dim3((unsigned)ceil_div(n_spikes, 4)), dim3(256), 0, s, R, ldr, N, (int)C, T,
                           (int)U, (int)L, idx, lab, n_spikes, grad_amp);
    if (grad_T) {
        uint32_t* cs = reinterpret_cast<uint32_t*>(w + lay.off_cs);
        float* part = reinterpret_cast<float*>(w + lay.off_part);
        hipLaunchKernelGGL(k_chunk_table, dim3(1), dim3(kRed), 0, s, unit_offsets, (int)U, cs);
        hipLaunchKernelGGL(k_grad_tmpl_part, dim3((unsigned)lay.nchunk, (unsigned)ceil_div(CL, 256)), dim3(256), 0, s,
                           R, ldr, N, (int)C, (int)U, (int)L, idx, amp, n_spikes, by_unit, unit_offsets, cs, part);
        const int64_t tot = (int64_t)U * CL;
        const int64_t nb = std::min<int64_t>(ceil_div(tot, 256), 1 << 16);
        hipLaunchKernelGGL(k_grad_tmpl_final, dim3((unsigned)nb), dim3(256), 0, s, part, cs, (int)U, CL,
                           (uint32_t)lay.nchunk, grad_T);
    }
    return launch_status("fbp_refine_backward");
}
