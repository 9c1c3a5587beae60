// Device helpers shared by the pass-loop kernels and the overlap screen.
#pragma once
#include "fbp_internal.h"

namespace fbp {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// ---------------------------------------------------------------------------
// Block helpers (256 threads = 4 wave64)
// ---------------------------------------------------------------------------

// Exclusive prefix of `flag` over the block in thread order; `total` receives
// the block count.  lds must hold 4 uint32.
__device__ __forceinline__ uint32_t block_excl_scan_flag(bool flag, uint32_t* lds, uint32_t& total)
{
    const uint64_t m = __ballot(flag);
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) lds[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; i++) {
        const uint32_t c = lds[i];
        woff += (i < wv) ? c : 0u;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return woff + pre;
}

// (value, key) arg-max with ties to the smaller key.  key == kNoItem means "none".
__device__ __forceinline__ void argmax_merge(float& v, uint32_t& k, float v2, uint32_t k2)
{
    if (k2 == kNoItem) return;
    if (k == kNoItem || v2 > v || (v2 == v && k2 < k)) { v = v2; k = k2; }
}

__device__ __forceinline__ void wave_argmax(float& v, uint32_t& k)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float v2 = __shfl_xor(v, off, 64);
        const uint32_t k2 = __shfl_xor(k, off, 64);
        argmax_merge(v, k, v2, k2);
    }
}

__device__ __forceinline__ bool recheck_item_skipped(const float* __restrict__ LIK, int64_t N, int L, uint32_t lab,
                                                     uint32_t u, int64_t b, int64_t f, int64_t s, int64_t n)
{
    if (f + b < n || f + b >= N - L + n) return true;
    if (s + b < n || s + b >= N - L + n) return true;
    if (f > s && f - s > L) return true;
    if (s > f && s - f > L) return true;
    const int64_t af = b + f - n, as = b + s - n;
    return LIK[(size_t)lab * N + af] <= 0.0f && LIK[(size_t)u * N + as] <= 0.0f;
}

// Exact FP32 value of item (u, f, s), binary_pursuit.cl:578-632.  vw points at
// the window's voltage: vw[c * cstride + j * jstride] = V[c, b + j].
__device__ __forceinline__ float recheck_item_value(const TemplateSet& ts, const float* vw, int cstride, int jstride,
                                                    uint32_t lab, uint32_t u, int64_t f, int64_t s, int64_t n)
{
    const int C = ts.C, L = ts.L;
    const int fa = (int)(n - f), sa = (int)(n - s);
    const int ffj = fa < 0 ? -fa : 0, sfj = sa < 0 ? -sa : 0;
    float lik = 0.0f;
    for (int c = 0; c < C; c++) {
        const bool nzf = ts.nz[(size_t)lab * C + c] != 0;
        const bool nzs = ts.nz[(size_t)u * C + c] != 0;
        if (!nzf && !nzs) continue;
        const float* tf = ts.T + ((size_t)lab * C + c) * L;
        const float* tsr = ts.T + ((size_t)u * C + c) * L;
        const float* vr = vw + (size_t)c * cstride;
        float ss = 0.0f;
        for (int j = 0; j < L; j++) {
            const float fv = (j >= ffj && fa + j < L) ? tf[fa + j] : 0.0f;
            const float sv = (j >= sfj && sa + j < L) ? tsr[sa + j] : 0.0f;
            const float sum = fv + sv;
            ss = fmaf(sum, sum, ss);
            lik = fmaf(sum, vr[(size_t)j * jstride], lik);
        }
        if (ss > 0.0f) lik = fmaf(-0.5f, ss, lik);
    }
    return lik;
}


// parse_overlap_recheck_indices accept cascade, binary_pursuit.cl:693-814, for
// window `id` whose best overlap item is k with value v (kNoItem: none).
__device__ inline void apply_parse(const TemplateSet& ts, const PassState& st, int64_t n_max, uint32_t id, float v,
                                   uint32_t k)
{
    const int64_t N = st.N;
    const int64_t U = ts.U, S = 2 * n_max + 1;
    const float bl = st.best_lik[id];
    if (k == kNoItem || !(v > bl) || v <= 0.0f) { st.orc[id] = 0; return; }
    const uint32_t lab = st.best_lab[id];
    const int64_t b = st.best_idx[id];
    const uint32_t u = (uint32_t)(k % U);
    const int64_t f = (int64_t)((k / U) % S), s = (int64_t)(k / (S * U));
    const int64_t af = b + f - n_max, as = b + s - n_max;
    const float A = st.LIK[(size_t)lab * N + af];
    const float B = st.LIK[(size_t)u * N + as];
    if (v < A || v < B || A <= 0.0f || B <= 0.0f) { st.orc[id] = 0; return; }
    if (A >= B && A > ts.thr[lab]) {
        st.best_lik[id] = A; st.obest_lab[id] = lab; st.obest_idx[id] = (uint32_t)af;
    } else if (B > A && B > ts.thr[u]) {
        st.best_lik[id] = B; st.obest_lab[id] = u; st.obest_idx[id] = (uint32_t)as;
    } else if (A > ts.thr[lab]) {
        st.best_lik[id] = A; st.obest_lab[id] = lab; st.obest_idx[id] = (uint32_t)af;
    } else {
        st.obest_idx[id] = st.best_idx[id];
        st.obest_lab[id] = st.best_lab[id];
        st.orc[id] = 0;
    }
}

// Sliding-register correlation of one template row against a voltage row:
// acc[q] = fma chain over taps i of trow[i] * vrow[i + q], taps ascending.
constexpr int kPos = 8;  // positions per work item

// Taps ascending over registers r[i + q] (r holds LT + kPos - 1 samples).
template <int LT>
__device__ __forceinline__ void lik_row_core(const float* __restrict__ trow, const float* r, float* acc)
{
#pragma unroll
    for (int i = 0; i < LT; i += 4) {
        const float4 tv = *reinterpret_cast<const float4*>(trow + i);
#pragma unroll
        for (int q = 0; q < kPos; q++) acc[q] = fmaf(tv.x, r[i + q], acc[q]);
#pragma unroll
        for (int q = 0; q < kPos; q++) acc[q] = fmaf(tv.y, r[i + 1 + q], acc[q]);
#pragma unroll
        for (int q = 0; q < kPos; q++) acc[q] = fmaf(tv.z, r[i + 2 + q], acc[q]);
#pragma unroll
        for (int q = 0; q < kPos; q++) acc[q] = fmaf(tv.w, r[i + 3 + q], acc[q]);
    }
}

// A 16-byte aligned voltage row with at least (LT + kPos + 2) & ~3 readable
// floats, read four floats at a time into registers, then the chain above.
template <int LT>
__device__ __forceinline__ void lik_row_fixed_a4(const float* __restrict__ trow, const float* vrow, float* acc)
{
    constexpr int NR = (LT + kPos - 1 + 3) & ~3;
    float r[NR];
#pragma unroll
    for (int j = 0; j < NR / 4; j++) {
        const float4 q = reinterpret_cast<const float4*>(vrow)[j];
        r[4 * j] = q.x;
        r[4 * j + 1] = q.y;
        r[4 * j + 2] = q.z;
        r[4 * j + 3] = q.w;
    }
    lik_row_core<LT>(trow, r, acc);
}

// Runtime L: the same chain with a sliding window of kPos registers.
__device__ __forceinline__ void lik_row_generic(const float* __restrict__ trow, const float* vrow, int L, float* acc)
{
    float r[kPos];
#pragma unroll
    for (int q = 0; q < kPos; q++) r[q] = vrow[q];
    for (int i = 0; i < L; i++) {
        const float tv = trow[i];
#pragma unroll
        for (int q = 0; q < kPos; q++) acc[q] = fmaf(tv, r[q], acc[q]);
#pragma unroll
        for (int q = 0; q < kPos - 1; q++) r[q] = r[q + 1];
        r[kPos - 1] = vrow[i + kPos];
    }
}

// A flagged window whose best spike fell below its threshold or too near the
// chunk end takes no overlap recheck.
__device__ __forceinline__ bool window_skipped(const TemplateSet& ts, const PassState& st, uint32_t id)
{
    const uint32_t lab = st.best_lab[id];
    const int64_t b = st.best_idx[id];
    const float lb = st.LIK[(size_t)lab * st.N + b];
    return (lb < ts.thr[lab]) || (b >= st.N - ts.L);  // binary_pursuit.cl:518-527
}

}  // namespace fbp
