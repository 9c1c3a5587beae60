// Screened overlap recheck: the same decisions as overlap_recheck_indices +
// parse_overlap_recheck_indices (SF/kernels/binary_pursuit.cl:444-815, SF =
// spikesorting_fullpursuit/) at a fraction of the work.
//
// In exact arithmetic the value of item (u, f, s) of a window whose best spike
// is (l, b) splits into
//     E(f, u, s) = A(f) + B(u, s) - G(f, u, s)
// with B(u, s) = <b_s, v> - |b_s|^2 / 2 for the window-truncated template b_s
// of unit u shifted by s - n (A(f) = B(l, f)) and G the window-truncated
// cross product of the two templates, which is voltage independent and zero
// when the units share no channel.  The screen evaluates E-hat (the FP32
// decomposed value) with a per-window bound delta >= |E-hat - E_fp32|,
// keeps every item whose E-hat could still be the maximum, and rescores those
// candidates with the exact FP32 operation order of the literal kernel.  The
// winner (value, then lowest item id) is therefore the literal kernel's.
#include <math.h>

#include <algorithm>

#include "fbp_device.h"

namespace fbp {

// Template value of unit u on channel c at window sample j when the template
// onset sits `sh` samples after the window start (zero outside the template).
__device__ __forceinline__ float shifted(const TemplateSet& ts, int u, int c, int j, int sh)
{
    const int i = j - sh;
    return (i >= 0 && i < ts.L) ? ts.T[((size_t)u * ts.C + c) * ts.L + i] : 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_tab_gram(TemplateSet ts, ScreenTables tb, int64_t npairs)
{
    const int S = tb.S, SG = tb.sg, n = tb.n, L = ts.L, C = ts.C;
    const int64_t total = npairs * S * SG;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = e / ((int64_t)S * SG);
        const int rem = (int)(e - p * S * SG);
        const int s = rem / SG, f = rem - s * SG;
        if (f >= S) {  // row padding
            tb.G[e] = 0.0f;
            continue;
        }
        const int l = tb.pair_l[p], u = tb.pair_u[p];
        const int fsh = f - n, ssh = s - n;
        const int jlo = max(0, max(fsh, ssh)), jhi = min(L, min(L + fsh, L + ssh));
        float g = 0.0f;
        for (int c = 0; c < C; c++) {
            if (!ts.nz[(size_t)l * C + c] || !ts.nz[(size_t)u * C + c]) continue;
            for (int j = jlo; j < jhi; j++) g = fmaf(shifted(ts, l, c, j, fsh), shifted(ts, u, c, j, ssh), g);
        }
        tb.G[e] = g;
    }
}

__global__ __launch_bounds__(kBlock) void k_tab_min_ntr(TemplateSet ts, ScreenTables tb, int64_t npairs)
{
    const int S = tb.S, n = tb.n, L = ts.L, C = ts.C;
    const int64_t total_min = npairs * S;
    const int64_t total = total_min + (int64_t)ts.U * S;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        if (e < total_min) {
            const float* row = tb.G + e * tb.sg;
            float m = INFINITY;
            for (int g = 0; g < tb.nbg; g++) {
                float mb = INFINITY;
                for (int f = g * kGB; f < min(S, (g + 1) * kGB); f++) mb = fminf(mb, row[f]);
                tb.minGb[e * tb.nbg + g] = mb;
                m = fminf(m, mb);
            }
            tb.minG[e] = m;
        } else {
            const int64_t r = e - total_min;
            const int u = (int)(r / S), s = (int)(r - (int64_t)u * S);
            const int ssh = s - n;
            const int jlo = max(0, ssh), jhi = min(L, L + ssh);
            float a = 0.0f;
            for (int c = 0; c < C; c++) {
                if (!ts.nz[(size_t)u * C + c]) continue;
                for (int j = jlo; j < jhi; j++) {
                    const float t = shifted(ts, u, c, j, ssh);
                    a = fmaf(t, t, a);
                }
            }
            tb.Ntr[r] = a;
        }
    }
}

// The unit-interleaved copy of the template rows (ScreenTables::Til).
__global__ __launch_bounds__(kBlock) void k_tab_interleave(TemplateSet ts, ScreenTables tb)
{
    const int Lp = ts.Lp, up = tb.upad;
    const int64_t total = (int64_t)tb.cmax * Lp * up;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e & 3);
        const int64_t r = e >> 2;  // (k * Lp / 4 + i / 4) * up + u
        const int u = (int)(r % up);
        const int64_t kq = r / up;
        const int k = (int)(kq / (Lp / 4)), iq = (int)(kq - (int64_t)k * (Lp / 4));
        float v = 0.0f;
        if (u < ts.U && k < ts.nch[u]) v = ts.rows[(size_t)(ts.off[u] + k) * Lp + 4 * iq + i];
        tb.Til[e] = v;
    }
}

// Taps [ilo, ihi) of a sliding correlation (4-tap chunks outside are skipped;
// they would only add products with the zero padding).  t4 steps over the
// unit-interleaved rows (ustride float4 per 4 taps); vrow is 16-byte aligned
// with LT + kPos readable samples.
template <int LT>
__device__ __forceinline__ void corr_row_range(const float4* __restrict__ t4, int ustride, const float* vrow, float* acc,
                                               int ilo, int ihi)
{
    static_assert(LT % 4 == 0, "whole 4-tap chunks");
    float r[LT + kPos];
#pragma unroll
    for (int j = 0; j < LT + kPos; j += 4) {
        const float4 v = *reinterpret_cast<const float4*>(vrow + j);
        r[j] = v.x;
        r[j + 1] = v.y;
        r[j + 2] = v.z;
        r[j + 3] = v.w;
    }
#pragma unroll
    for (int i = 0; i < LT; i += 4) {
        if (i + 3 < ilo || i >= ihi) continue;
        const float4 tv = t4[(size_t)(i >> 2) * ustride];
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

// Runtime L: the same taps with a sliding window of kPos registers.
__device__ __forceinline__ void corr_row_generic(const float* __restrict__ tf, int ustride, const float* vrow, int L,
                                                 float* acc)
{
    float r[kPos];
#pragma unroll
    for (int q = 0; q < kPos; q++) r[q] = vrow[q];
    for (int i = 0; i < L; i++) {
        const float tv = tf[(size_t)(i >> 2) * ustride * 4 + (i & 3)];
#pragma unroll
        for (int q = 0; q < kPos; q++) acc[q] = fmaf(tv, r[q], acc[q]);
#pragma unroll
        for (int q = 0; q < kPos - 1; q++) r[q] = r[q + 1];
        r[kPos - 1] = vrow[i + kPos];
    }
}

// B(u, s) for every unit and shift of the flagged windows F[f0 .. f0+nbatch).
// One block per window; the window's voltage (zero outside [b, b+L)) is staged
// in LDS as rows of rp samples, sample p <-> v[p - n].
template <int LT>
__device__ __forceinline__ void screen_corr_body(TemplateSet ts, PassState st, ScreenTables tb, uint32_t f0,
                                                        uint32_t nbatch, int rp, float* __restrict__ Bbuf, size_t bstride, uint32_t bx, uint32_t gx)
{
    extern __shared__ __align__(16) float sP[];
    const int C = ts.C, U = ts.U;
    const int L = (LT > 0) ? LT : ts.L;
    const int S = tb.S, n = tb.n;
    const int ngr = (S + kPos - 1) / kPos;
    const int64_t N = st.N;
    const uint32_t nF = st.counters[CNT_F];
    // every window of the block fills the same samples [n, n + L) of each row:
    // the zero padding around them is written once
    if (bx < nbatch)
        for (int e = threadIdx.x; e < C * rp; e += blockDim.x) sP[e] = 0.0f;
    for (uint32_t wi = bx; wi < nbatch; wi += gx) {
        const uint32_t w = f0 + wi;
        if (w >= nF) break;
        const uint32_t id = st.F[w];
        const bool skip = window_skipped(ts, st, id);
        const int64_t b = st.best_idx[id];
        __syncthreads();
        // the window's C x L samples are one contiguous run of the sample-major
        // voltage (b + L <= N for a window that is not skipped): four float4
        // loads in flight per thread, then the transposed LDS stores
        if (!skip) {
            if (C % 4 == 0) {
                const float4* src = reinterpret_cast<const float4*>(st.V + (size_t)b * C);
                const int n4 = C * L / 4;
                for (int e0 = threadIdx.x; e0 < n4; e0 += 4 * blockDim.x) {
                    float4 v[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int e = e0 + q * (int)blockDim.x;
                        v[q] = src[e < n4 ? e : 0];
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int e = e0 + q * (int)blockDim.x;
                        if (e >= n4) continue;
                        const int j = (4 * e) / C, c = 4 * e - j * C;
                        float* d = sP + c * rp + n + j;
                        d[0] = v[q].x;
                        d[rp] = v[q].y;
                        d[2 * rp] = v[q].z;
                        d[3 * rp] = v[q].w;
                    }
                }
            } else {
                for (int e = threadIdx.x; e < C * L; e += blockDim.x) {
                    const int j = e / C, c = e - j * C;
                    sP[c * rp + n + j] = st.V[(size_t)b * C + e];
                }
            }
        }
        __syncthreads();
        if (skip) continue;
        float* Bw = Bbuf + (size_t)wi * bstride;
        // item = (shift group g, unit u) with units padded to a multiple of 64,
        // so every wave works on one shift group and skips the same taps
        const int Up = (U + 63) & ~63;
        for (int it = threadIdx.x; it < Up * ngr; it += blockDim.x) {
            const int g = it / Up, u = it - g * Up;
            if (u >= U) continue;
            const int s0 = g * kPos;
            float acc[kPos];
#pragma unroll
            for (int q = 0; q < kPos; q++) acc[q] = 0.0f;
            const int ilo = n - s0 - (kPos - 1), ihi = n + L - s0;
            const int r0 = ts.off[u], nr = ts.nch[u];
            const int lq = ts.Lp >> 2;
            for (int k = 0; k < nr; k++) {
                const int c = ts.chan[r0 + k];
                const float4* t4 = reinterpret_cast<const float4*>(tb.Til) + (size_t)k * lq * tb.upad + u;
                const float* vrow = sP + c * rp + s0;
                if (LT > 0) corr_row_range<(LT > 0 ? LT : 4)>(t4, tb.upad, vrow, acc, ilo, ihi);
                else corr_row_generic(reinterpret_cast<const float*>(t4), tb.upad, vrow, L, acc);
            }
#pragma unroll
            for (int q = 0; q < kPos; q++) {
                const int s = s0 + q;
                if (s < S) Bw[(size_t)u * S + s] = fmaf(-0.5f, tb.Ntr[(size_t)u * S + s], acc[q]);
            }
        }
    }
}

template <int LT>
__global__ __launch_bounds__(kBlock) void k_screen_corr(TemplateSet ts, PassState st, ScreenTables tb, uint32_t f0,
                                                        uint32_t nbatch, int rp, float* __restrict__ Bbuf, size_t bstride)
{
    screen_corr_body<LT>(ts, st, tb, f0, nbatch, rp, Bbuf, bstride, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ float block_max(float v, float* red)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float m = red[0];
#pragma unroll
    for (int i = 1; i < kBlock / 64; i++) m = fmaxf(m, red[i]);
    return m;
}

// Exact FP32 value of item (u, f, s): rows of the window's label l in LDS
// (stride L), rows of unit u in the compact template set (stride Lp).  Same
// operation order as recheck_item_value: channels of either template
// ascending, then samples.
template <int LT>
__device__ __forceinline__ float exact_item_rows(const float* vW, int Lrt, const float* Tl, const int* cl, int nl,
                                        const float* Tu, int ldu, const int* cu, int nu, int f, int s, int n)
{
    const int L = LT > 0 ? LT : Lrt;
    const int fa = n - f, sa = n - s;
    const int ffj = fa < 0 ? -fa : 0, sfj = sa < 0 ? -sa : 0;
    float lik = 0.0f;
    int il = 0, iu = 0;
    while (il < nl || iu < nu) {
        const int c1 = il < nl ? cl[il] : 0x7fffffff;
        const int c2 = iu < nu ? cu[iu] : 0x7fffffff;
        const int c = min(c1, c2);
        const bool hf = c1 == c, hs = c2 == c;
        const float* tf = Tl + (hf ? il : 0) * L;
        const float* tsr = Tu + (size_t)(hs ? iu : 0) * ldu;
        il += hf;
        iu += hs;
        const float* vr = vW + c * L;
        float ss = 0.0f;
        // 8 samples at a time: all loads (from clamped, valid addresses) are
        // issued before the two FMA chains consume them
        for (int j0 = 0; j0 < L; j0 += 8) {
            float a[8], c2[8], vv[8];
            bool okf[8], oks[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int j = j0 + q, jf = fa + j, js = sa + j;
                okf[q] = hf && j >= ffj && jf < L;
                oks[q] = hs && j >= sfj && js < L;
                a[q] = tf[okf[q] ? jf : 0];
                c2[q] = tsr[oks[q] ? js : 0];
                vv[q] = vr[j < L ? j : 0];
            }
            __builtin_amdgcn_sched_barrier(0);  // keep the 24 loads ahead of the chains
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (LT == 0 && j0 + q >= L) break;
                const float sum = (okf[q] ? a[q] : 0.0f) + (oks[q] ? c2[q] : 0.0f);
                ss = fmaf(sum, sum, ss);
                lik = fmaf(sum, vv[q], lik);
            }
        }
        if (ss > 0.0f) lik = fmaf(-0.5f, ss, lik);
    }
    return lik;
}

// exact_item_rows over zero-padded rows (row length LP = L + 2n, P[i] = T[i - n]):
// the label's term at sample j is TlP[row][2n - f + j], the unit's
// Tu[row][2n - s + j], an absent channel reads the zero row; window channel c
// starts at vW[c * vs].  Same values and operation order as exact_item_rows,
// without per-sample range tests.
template <int LT>
__device__ __forceinline__ float exact_item_padded(const float* vW, int Lrt, int vs, const float* TlP, const int* cl, int nl,
                                                   const float* Tu, const int* cu, int nu, const float* zrow, int LP,
                                                   int f, int s, int n)
{
    const int L = LT > 0 ? LT : Lrt;
    const int of = 2 * n - f, os = 2 * n - s;
    float lik = 0.0f;
    int il = 0, iu = 0;
    while (il < nl || iu < nu) {
        const int c1 = il < nl ? cl[il] : 0x7fffffff;
        const int c2 = iu < nu ? cu[iu] : 0x7fffffff;
        const int c = min(c1, c2);
        const bool hf = c1 == c, hs = c2 == c;
        const float* tf = (hf ? TlP + il * LP : zrow) + of;
        const float* tsr = (hs ? Tu + iu * LP : zrow) + os;
        il += hf;
        iu += hs;
        const float* vr = vW + c * vs;
        float ss = 0.0f;
        // eight samples at a time, the next eight read while these are consumed
        float a[8], b[8], vv[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int j = (LT > 0 || q < L) ? q : 0;
            a[q] = tf[j];
            b[q] = tsr[j];
            vv[q] = vr[j];
        }
#pragma unroll
        for (int j0 = 0; j0 < L; j0 += 8) {
            float na[8], nb[8], nv[8];
            if (j0 + 8 < L) {
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int j = (LT > 0 || j0 + 8 + q < L) ? j0 + 8 + q : 0;
                    na[q] = tf[j];
                    nb[q] = tsr[j];
                    nv[q] = vr[j];
                }
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (LT == 0 && j0 + q >= L) break;
                const float sum = a[q] + b[q];
                ss = fmaf(sum, sum, ss);
                lik = fmaf(sum, vv[q], lik);
            }
            if (j0 + 8 < L) {
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    a[q] = na[q];
                    b[q] = nb[q];
                    vv[q] = nv[q];
                }
            }
        }
        if (ss > 0.0f) lik = fmaf(-0.5f, ss, lik);
    }
    return lik;
}

constexpr int kCandCap = 1024;

// e = q * S + r for 0 <= e < 2^24 with a float reciprocal and one correction.
__device__ __forceinline__ void divmod_small(int e, int S, float invS, int& q, int& r)
{
    q = (int)((float)e * invS);
    r = e - q * S;
    if (r < 0) { q--; r += S; }
    else if (r >= S) { q++; r -= S; }
}
constexpr int kListChunk = 2048;  // (unit, shift) pairs examined per work-list round

// B(u, s) of the current window: element e = u*S + s.
struct BView {
    const float* Bg;
    __device__ __forceinline__ float operator()(int e, int, int) const { return Bg[e]; }
};

// Selection, exact rescoring and the parse cascade for one window per block.
template <int LT>
__global__ __launch_bounds__(kBlock) void k_screen_select(TemplateSet ts, PassState st, ScreenTables tb, uint32_t f0,
                                                          uint32_t nbatch, const float* __restrict__ Bbuf)
{
    extern __shared__ float sm[];
    __shared__ float red[kBlock / 64];
    __shared__ uint32_t redk[kBlock / 64];
    __shared__ float cand_e[kCandCap];    // E-hat of each candidate
    __shared__ uint32_t cand_k[kCandCap]; // item id
    __shared__ uint32_t wl[kListChunk];
    __shared__ uint32_t nwl;
    __shared__ uint32_t ncand_s;
    const int C = ts.C, U = ts.U;
    const int L = LT > 0 ? LT : ts.L;
    const int S = tb.S, n = tb.n, CM = tb.cmax;
    const int US = U * S;
    const int64_t N = st.N;
    const int nwords = (US + 63) / 64 * 2;
    float* vW = sm;                       // [C][L]
    float* vmax = vW + C * L;             // [C]
    float* A = vmax + C;                  // [S]
    float* RAall = A + S;                 // [S]
    float* RApos = RAall + S;             // [S]
    float* PMa = RApos + S;               // [S] prefix / suffix maxima of A
    float* PMp = PMa + S;
    float* SMa = PMp + S;
    float* SMp = SMa + S;
    float* Q = SMp + S;                   // [U]
    float* D = Q + U;                     // [U] per-unit bound delta
    float* Tl = D + U;                    // [CM][L] rows of the window's label
    int* cl = reinterpret_cast<int*>(Tl + CM * L);  // [CM]
    uint32_t* posu = reinterpret_cast<uint32_t*>(cl + CM);  // [nwords] bit (u*S+s): LIK[u][b+s-n] > 0
    uint8_t* fin = reinterpret_cast<uint8_t*>(posu + nwords);  // [S] b+f-n in [0, N-L)
    uint8_t* posl = fin + S;              // [S] fin && LIK[l][b+f-n] > 0
    const int wv = threadIdx.x >> 6, lane = lane_id();
    const float invS = 1.0f / (float)S;
    const uint32_t nF = st.counters[CNT_F];
    for (uint32_t wi = blockIdx.x; wi < nbatch; wi += gridDim.x) {
        const uint32_t w = f0 + wi;
        if (w >= nF) break;
        const uint32_t id = st.F[w];
        const uint32_t lab = st.best_lab[id];
        const int64_t b = st.best_idx[id];
        const float bl = st.best_lik[id];
        __syncthreads();
        if (window_skipped(ts, st, id)) {
            if (threadIdx.x == 0) apply_parse(ts, st, n, id, 0.0f, kNoItem);
            continue;
        }
        BView Bv{Bbuf + (size_t)wi * US};
        for (int e = threadIdx.x; e < C * L; e += blockDim.x) {
            const int j = e / C, c = e - j * C;
            vW[c * L + j] = st.V[(size_t)(b + j) * C + c];
        }
        const int nl = ts.nch[lab];
        for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
            const int k = e / L, j = e - k * L;
            Tl[e] = ts.rows[(size_t)(ts.off[lab] + k) * ts.Lp + j];
        }
        for (int k = threadIdx.x; k < nl; k += blockDim.x) cl[k] = ts.chan[ts.off[lab] + k];
        for (int f = threadIdx.x; f < S; f += blockDim.x) {
            const int64_t af = b + f - n;
            const bool in = af >= 0 && af < N - L;
            fin[f] = in;
            posl[f] = in && st.LIK[(size_t)lab * N + af] > 0.0f;
            A[f] = Bv((int)lab * S + f, (int)lab, f);
        }
        // sign bitmap of LIK[u][b+s-n], one ballot per 64 consecutive pairs;
        // each wave keeps 4 loads in flight
        for (int base0 = wv * 64; base0 < US; base0 += 4 * kBlock) {
            float lv[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int e = base0 + r * kBlock + lane;
                lv[r] = 0.0f;
                if (e < US) {
                    int u, s;
                divmod_small(e, S, invS, u, s);
                    const int64_t as = b + s - n;
                    if (as >= 0 && as < N - L) lv[r] = st.LIK[(size_t)u * N + as];
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int base = base0 + r * kBlock;
                const uint64_t m = __ballot(lv[r] > 0.0f);
                if (lane == 0 && base < US) {
                    posu[base >> 5] = (uint32_t)m;
                    posu[(base >> 5) + 1] = (uint32_t)(m >> 32);
                }
            }
        }
        if (threadIdx.x == 0) ncand_s = 0;
        __syncthreads();
        for (int c = threadIdx.x; c < C; c += blockDim.x) {
            float m = 0.0f;
            for (int j = 0; j < L; j++) m = fmaxf(m, fabsf(vW[c * L + j]));
            vmax[c] = m;
        }
        // range maxima of A over f in [s-L, s+L]: all in-bound f / those with LIK > 0.
        // With n <= L every range starts at 0 or ends at S-1, so prefix maxima
        // (wave 0) and suffix maxima (wave 1), each a wave-wide scan, answer them.
        if (n <= L) {
            if (wv < 2) {
                const bool suffix = wv == 1;
                float* Xa = suffix ? SMa : PMa;
                float* Xp = suffix ? SMp : PMp;
                float ca = -INFINITY, cp = -INFINITY;
                for (int c0 = 0; c0 < S; c0 += 64) {
                    const int f = suffix ? S - 1 - (c0 + lane) : c0 + lane;
                    const bool ok = c0 + lane < S;
                    float va = -INFINITY, vp = -INFINITY;
                    if (ok) {
                        if (fin[f]) va = A[f];
                        if (posl[f]) vp = A[f];
                    }
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) {
                        const float ta = __shfl_up(va, off, 64), tq = __shfl_up(vp, off, 64);
                        if (lane >= off) { va = fmaxf(va, ta); vp = fmaxf(vp, tq); }
                    }
                    va = fmaxf(va, ca);
                    vp = fmaxf(vp, cp);
                    if (ok) { Xa[f] = va; Xp[f] = vp; }
                    ca = __shfl(va, 63, 64);
                    cp = __shfl(vp, 63, 64);
                }
            }
            __syncthreads();
            for (int s = threadIdx.x; s < S; s += blockDim.x) {
                const int fl = max(0, s - L), fh = min(S - 1, s + L);
                RAall[s] = (fl == 0) ? PMa[fh] : SMa[fl];
                RApos[s] = (fl == 0) ? PMp[fh] : SMp[fl];
            }
        } else {
            for (int s = threadIdx.x; s < S; s += blockDim.x) {
                float ma = -INFINITY, mp = -INFINITY;
                const int fl = max(0, s - L), fh = min(S - 1, s + L);
                for (int f = fl; f <= fh; f++) {
                    if (fin[f]) ma = fmaxf(ma, A[f]);
                    if (posl[f]) mp = fmaxf(mp, A[f]);
                }
                RAall[s] = ma;
                RApos[s] = mp;
            }
        }
        __syncthreads();
        for (int u = threadIdx.x; u < U; u += blockDim.x) {
            float q = 0.0f, t1 = 0.0f, vsum = 0.0f;
            const int r0 = ts.off[u], nr = ts.nch[u];
            for (int k = 0; k < nr; k++) {
                const float vm = vmax[ts.chan[r0 + k]];
                q += tb.rowL1[r0 + k] * vm;
                t1 += tb.rowL1[r0 + k];
                vsum += vm;
            }
            float z = 0.0f;
            if (tb.kappa3 != 0.0f) {  // fp16 operands: absolute residue of the scaled split
                float mw = 0.0f;
                for (int c = 0; c < C; c++) mw = fmaxf(mw, vmax[c]);
                z = tb.kappa3 * mw * t1 + tb.kappa4 * vsum;
            }
            Q[u] = (tb.kappa + tb.kappa2) * q + 2.0f * tb.kappa * tb.SS[u] + z;
        }
        __syncthreads();
        // delta(l, u) = P_l + P_u: the bound is a sum of one term per template
        for (int u = threadIdx.x; u < U; u += blockDim.x) D[u] = Q[lab] + Q[u];
        float mA = -INFINITY;
        for (int f = threadIdx.x; f < S; f += blockDim.x)
            if (fin[f]) mA = fmaxf(mA, A[f]);
        const float maxA = block_max(mA, red);
        // phase 1: pairs sharing no channel, E-hat = A(f) + B(u, s) separable.
        // Tracks M' = max(E-hat - delta_u), a lower bound of the exact maximum.
        const int* pidl = tb.pid + (size_t)lab * U;
        float mns = -INFINITY;
        for (int e0 = threadIdx.x; e0 < US; e0 += 4 * kBlock) {
            float bv4[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int e = e0 + r * kBlock;
                bv4[r] = 0.0f;
                if (e < US) {
                    int u, sx;
                    divmod_small(e, S, invS, u, sx);
                    bv4[r] = Bv(e, u, sx);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int e = e0 + r * kBlock;
                if (e >= US) continue;
                int u, s;
                divmod_small(e, S, invS, u, s);
                if (pidl[u] >= 0) continue;
                const int64_t as = b + s - n;
                if (as < 0 || as >= N - L) continue;
                const bool pu = (posu[e >> 5] >> (e & 31)) & 1u;
                const float ra = pu ? RAall[s] : RApos[s];
                if (ra == -INFINITY) continue;
                mns = fmaxf(mns, (ra + bv4[r]) - D[u]);
            }
        }
        const float Mns = block_max(mns, red);
        // phase 2: pairs sharing a channel whose bound max_f A + B - min_f G
        // reaches tp_u = max(M' - delta_u, bl - delta_u) go to a work list;
        // one wave per pair enumerates f, tracks M' and keeps items with
        // E-hat >= tp_u as candidates.
        float msh = -INFINITY;
        bool overflow = false;
        for (int e0 = 0; e0 < US; e0 += kListChunk) {
            if (threadIdx.x == 0) nwl = 0;
            __syncthreads();
            for (int e = e0 + threadIdx.x; e < min(US, e0 + kListChunk); e += blockDim.x) {
                int u, s;
                divmod_small(e, S, invS, u, s);
                const int p = pidl[u];
                if (p < 0) continue;
                const int64_t as = b + s - n;
                if (as < 0 || as >= N - L) continue;
                const float tp = fmaxf(Mns - D[u], bl - D[u]);
                if ((maxA + Bv(e, u, s)) - tb.minG[(size_t)p * S + s] < tp) continue;
                wl[atomicAdd(&nwl, 1u)] = (uint32_t)e;
            }
            __syncthreads();
            if (threadIdx.x == 0 && nwl) atomicAdd(st.counters + CNT_SURV, nwl);
            for (uint32_t i = wv; i < nwl; i += kBlock / 64) {
                const int e = (int)wl[i];
                int u, s;
                divmod_small(e, S, invS, u, s);
                const float Bu = Bv(e, u, s);
                const float Du = D[u];
                const float tp = fmaxf(Mns - Du, bl - Du);
                const bool pu = (posu[e >> 5] >> (e & 31)) & 1u;
                const float* Gr = tb.G + ((size_t)pidl[u] * S + s) * tb.sg;
                for (int f = lane; f < S; f += 64) {
                    if (!fin[f] || abs(f - s) > L || !(posl[f] || pu)) continue;
                    const float eh = (A[f] + Bu) - Gr[f];
                    msh = fmaxf(msh, eh - Du);
                    if (eh < tp) continue;
                    const uint32_t slot = atomicAdd(&ncand_s, 1u);
                    if (slot < kCandCap) {
                        cand_e[slot] = eh;
                        cand_k[slot] = (uint32_t)u + (uint32_t)U * ((uint32_t)f + (uint32_t)S * (uint32_t)s);
                    } else {
                        overflow = true;
                    }
                }
            }
            __syncthreads();
        }
        const float Mp = fmaxf(Mns, block_max(msh, red));
        // phase 3: pairs sharing no channel whose exact best over f reaches
        // theta_u = max(M' - delta_u, bl - delta_u) (few) add their items
        for (int e0 = 0; e0 < US; e0 += kListChunk) {
            if (threadIdx.x == 0) nwl = 0;
            __syncthreads();
            for (int e = e0 + threadIdx.x; e < min(US, e0 + kListChunk); e += blockDim.x) {
                int u, s;
                divmod_small(e, S, invS, u, s);
                if (pidl[u] >= 0) continue;
                const int64_t as = b + s - n;
                if (as < 0 || as >= N - L) continue;
                const bool pu = (posu[e >> 5] >> (e & 31)) & 1u;
                const float ra = pu ? RAall[s] : RApos[s];
                if (ra == -INFINITY || ra + Bv(e, u, s) < fmaxf(Mp - D[u], bl - D[u])) continue;
                wl[atomicAdd(&nwl, 1u)] = (uint32_t)e;
            }
            __syncthreads();
            if (threadIdx.x == 0 && nwl) atomicAdd(st.counters + CNT_NSP, nwl);
            for (uint32_t i = wv; i < nwl; i += kBlock / 64) {
                const int e = (int)wl[i];
                int u, s;
                divmod_small(e, S, invS, u, s);
                const float Bu = Bv(e, u, s);
                const float theta = fmaxf(Mp - D[u], bl - D[u]);
                const bool pu = (posu[e >> 5] >> (e & 31)) & 1u;
                for (int f = lane; f < S; f += 64) {
                    if (!fin[f] || abs(f - s) > L || !(posl[f] || pu)) continue;
                    const float eh = A[f] + Bu;
                    if (eh < theta) continue;
                    const uint32_t slot = atomicAdd(&ncand_s, 1u);
                    if (slot < kCandCap) {
                        cand_e[slot] = eh;
                        cand_k[slot] = (uint32_t)u + (uint32_t)U * ((uint32_t)f + (uint32_t)S * (uint32_t)s);
                    } else {
                        overflow = true;
                    }
                }
            }
            __syncthreads();
        }
        const uint32_t nc = min(ncand_s, (uint32_t)kCandCap);
        float bv = 0.0f;
        uint32_t bk = kNoItem;
        if (__syncthreads_or(overflow)) {
            // more candidates than the list holds: enumerate and rescore in place
            for (int e = threadIdx.x; e < US; e += blockDim.x) {
                int u, s;
                divmod_small(e, S, invS, u, s);
                const int64_t as = b + s - n;
                if (as < 0 || as >= N - L) continue;
                const float Bu = Bv(e, u, s);
                const int p = pidl[u];
                const bool pu = (posu[e >> 5] >> (e & 31)) & 1u;
                const float theta = fmaxf(Mp - D[u], bl - D[u]);
                const float* Gr = p >= 0 ? tb.G + ((size_t)p * S + s) * tb.sg : nullptr;
                for (int f = 0; f < S; f++) {
                    if (!fin[f] || abs(f - s) > L || !(posl[f] || pu)) continue;
                    const float eh = p >= 0 ? (A[f] + Bu) - Gr[f] : A[f] + Bu;
                    if (eh < theta) continue;
                    const float v = recheck_item_value(ts, vW, L, 1, lab, (uint32_t)u, f, s, n);
                    argmax_merge(bv, bk, v, (uint32_t)u + (uint32_t)U * ((uint32_t)f + (uint32_t)S * (uint32_t)s));
                }
            }
        } else {
            // candidates reaching theta_u, each rescored exactly by one thread
            for (uint32_t i = threadIdx.x; i < nc; i += blockDim.x) {
                const uint32_t k = cand_k[i];
                const int u = (int)(k % (uint32_t)U);
                const int f = (int)((k / (uint32_t)U) % (uint32_t)S), s = (int)(k / ((uint32_t)U * S));
                if (cand_e[i] < fmaxf(Mp - D[u], bl - D[u])) continue;
                const int r0 = ts.off[u];
                const float v = exact_item_rows<LT>(vW, L, Tl, cl, nl, ts.rows + (size_t)r0 * ts.Lp, ts.Lp,
                                                    ts.chan + r0, ts.nch[u], f, s, n);
                argmax_merge(bv, bk, v, k);
            }
        }
        wave_argmax(bv, bk);
        __syncthreads();
        if (lane == 0) { red[wv] = bv; redk[wv] = bk; }
        __syncthreads();
        if (threadIdx.x == 0) {
            float v = red[0];
            uint32_t k = redk[0];
            for (int i = 1; i < kBlock / 64; i++) argmax_merge(v, k, red[i], redk[i]);
            apply_parse(ts, st, n, id, v, k);
            atomicAdd(st.counters + CNT_EVAL, ncand_s);
            if (ncand_s > (uint32_t)kCandCap) atomicAdd(st.counters + CNT_CAND, 1u);
        }
    }
}

// ---------------------------------------------------------------------------
// Register-resident selection: the same bounds, candidates and rescoring as
// k_screen_select, organised so that every (unit, shift) pair of the window is
// read once.  The block's first KU * S threads own the pairs (KU = ceil(U /
// kSelR) unit slices): thread t keeps the shift s = t mod S and the units
// u = t / S + KU * i, i < kSelR, so its shift-indexed bounds are two registers
// and its pairs' unit-indexed values lie at a fixed LDS stride.  It keeps
// B(u, s) (-inf when b+s-n is out of bounds or u >= U) and two bit masks
// (LIK[u][b+s-n] > 0; u shares a channel with the label) in registers through
// the three phases.  Each wave enumerates the shifts f of its own surviving
// pairs one pair at a time, so no phase needs a block-wide list.
// ---------------------------------------------------------------------------
// Deferred exact rescoring (k_screen_rescore): a window whose candidates fit
// one wave (at most 64, of at most kRU units, kCV channels with the label's)
// leaves them in a per-window record [count, keys...]; every other window is
// finished by the selection kernel itself and marks its record kXDone.
constexpr uint32_t kXDone = 0xffffffffu;
constexpr int kRU = 3;
constexpr int kCV = 32;
constexpr int kSelR = 24;           // pairs per thread (bits of a 32-bit mask)
constexpr int kQR = 2;              // (pair, f-quad) items per thread per round of G loads
constexpr int kSelLR = 12;          // pairs per round of the per-pair loads (divides kSelR)
static_assert(kSelR % kSelLR == 0, "whole rounds");
static_assert(kSelR <= 32, "pair masks are 32-bit");
constexpr int kSelMaxBlock = 1024;  // threads per block at most (KU * S <= kSelMaxBlock)
constexpr size_t kSelMaxLds = 96 * 1024;  // dynamic LDS of the register-resident selection at most
constexpr int kStageRegion = 8192;  // floats of the selection's phase-dependent LDS region at least

// Unit slices of the register-resident selection's pair layout.
__host__ __device__ inline int sel_slices(int U) { return (U + kSelR - 1) / kSelR; }

__device__ __forceinline__ float block_max_n(float v, float* red, int nw)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float m = red[0];
    for (int i = 1; i < nw; i++) m = fmaxf(m, red[i]);
    return m;
}

__device__ __forceinline__ int wave_sum_u32(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Advance (q, r) of e = q * D + r by bd = dq * D + dr (dr < D).
__device__ __forceinline__ void qr_step(int& q, int& r, int dq, int dr, int D)
{
    r += dr;
    q += dq;
    if (r >= D) { r -= D; q++; }
}

// apply_parse for a window whose best spike (lab, b) and likelihood bl are
// already at hand: the same decisions (parse_overlap_recheck_indices,
// binary_pursuit.cl:657-815), with the item decoded in 32 bits and the four
// values it needs read together.
__device__ __forceinline__ void apply_parse_at(const TemplateSet& ts, const PassState& st, int n, uint32_t id, float v,
                                               uint32_t k, float bl, uint32_t lab, int64_t b, int S)
{
    if (k == kNoItem || !(v > bl) || v <= 0.0f) { st.orc[id] = 0; return; }
    const uint32_t U = (uint32_t)ts.U;
    const uint32_t u = k % U, r = k / U;
    const int64_t f = (int64_t)(r % (uint32_t)S), s = (int64_t)(r / (uint32_t)S);
    const int64_t af = b + f - n, as = b + s - n;
    const float A = st.LIK[(size_t)lab * st.N + af];
    const float B = st.LIK[(size_t)u * st.N + as];
    const float tl = ts.thr[lab], tu = ts.thr[u];
    if (v < A || v < B || A <= 0.0f || B <= 0.0f) { st.orc[id] = 0; return; }
    if (A >= B && A > tl) {
        st.best_lik[id] = A; st.obest_lab[id] = lab; st.obest_idx[id] = (uint32_t)af;
    } else if (B > A && B > tu) {
        st.best_lik[id] = B; st.obest_lab[id] = u; st.obest_idx[id] = (uint32_t)as;
    } else if (A > tl) {
        st.best_lik[id] = A; st.obest_lab[id] = lab; st.obest_idx[id] = (uint32_t)af;
    } else {
        st.obest_idx[id] = (uint32_t)b;
        st.obest_lab[id] = lab;
        st.orc[id] = 0;
    }
}

template <int LT>
__device__ __forceinline__ void screen_select_reg_body(TemplateSet ts, PassState st, ScreenTables tb,
                                                                    uint32_t f0, uint32_t nbatch,
                                                                    const float* __restrict__ Bbuf, size_t bstride, uint32_t bx, uint32_t gx,
                                                                    uint32_t* __restrict__ xcand)
{
    extern __shared__ float sm[];
    __shared__ float red[kSelMaxBlock / 64];
    __shared__ uint32_t redk[kSelMaxBlock / 64];
    __shared__ float cand_e[kCandCap];
    __shared__ uint32_t cand_k[kCandCap];
    __shared__ uint32_t ncand_s;
    __shared__ int plo_sh, phi_sh;  // first / last f with LIK[l][b+f-n] > 0 (S, -1: none)
    __shared__ int pfirst_sh, npl_sh;  // pair ids of the label: pfirst .. pfirst + npl - 1
    __shared__ uint32_t nsurv_s, nsurvp_s, nrow_s, nresc_s;
    __shared__ unsigned long long cun_s;  // channels of the candidates' units (deferred rescoring)
    __shared__ uint32_t nstat_s[2];  // the window's shared-channel / separable survivors (statistics)
    const int C = ts.C, U = ts.U;
    const int L = LT > 0 ? LT : ts.L;
    const int S = tb.S, n = tb.n, CM = tb.cmax;
    const int US = U * S;
    const int64_t N = st.N;
    const int bd = (int)blockDim.x, nw = bd >> 6;
    const int KU = sel_slices(U);
    const int UP = KU * kSelR;            // unit slots of the pair layout (>= U)
    // window rows at an odd stride: lanes rescoring different units read
    // different channels at one sample without sharing a bank
    const int VS = L | 1;
    const int LP = L + 2 * n;             // zero-padded row: P[i] = T[i - n], i - n outside [0, L) -> 0
    const int RG = tb.sel_region;         // floats of the phase-dependent region
    float* vW = sm;                       // [C][VS]
    float* vmax = vW + C * VS;            // [C]
    float* A = vmax + C;                  // [S]
    float* RAall = A + S;                 // [S]
    float* RApos = RAall + S;             // [S]
    float* PMa = RApos + S;               // [S] prefix / suffix maxima of A
    float* PMp = PMa + S;
    float* SMa = PMp + S;
    float* SMp = SMa + S;
    float* Q = SMp + S;                   // [U]
    const int dpo = (C * VS + C + 7 * S + U + 1) & ~1;
    // [UP] (delta(l, u), pair id of (l, u) as bits, -1: no shared channel); slots u >= U hold (0, -1)
    float2* Dp = reinterpret_cast<float2*>(sm + dpo);
    float* TlP = sm + dpo + 2 * UP;       // [CM][LP] padded rows of the window's label
    float* zrow = TlP + CM * LP;          // [LP] zeros (a template channel that is absent)
    int* cl = reinterpret_cast<int*>(zrow + LP);  // [CM]
    float* reg = reinterpret_cast<float*>(cl + CM);  // [RG], in turn:
    float* gmA = reg + RG / 2;            //   [npl][S] least G over every f
    uint32_t* surv_e = reinterpret_cast<uint32_t*>(reg);  //   survivor lists: pair ...
    float* surv_b = reg + RG / 2;                          //   ... and its B
    const uint32_t survcap = (uint32_t)(RG / 2);
    float* stgP = reg;                    //   [srows][LP] padded rows of the candidates' units
    const int srows = RG / LP;
    int* stg_ch = reinterpret_cast<int*>(reg + RG);  // [RG / LP] channel of each staged row
    int* slot_of = stg_ch + RG / LP;      // [U] order of unit u among the staged units (-1: none)
    int* slot_nr = slot_of + U;           // [U] non-zero channels of the unit in each staging slot
    int* slot_off = slot_nr + U;          // [U] its first compact row
    uint8_t* fin = reinterpret_cast<uint8_t*>(slot_off + U);  // [S] b+f-n in [0, N-L)
    uint8_t* posl = fin + S;              // [S] fin && LIK[l][b+f-n] > 0
    const int wv = threadIdx.x >> 6, lane = lane_id();
    const float invS = 1.0f / (float)S;
    const uint32_t nF = st.counters[CNT_F];
    // one window per block (the grid covers the batch; blocks beyond it return)
    {
        const uint32_t wi = bx;
        const uint32_t w = f0 + wi;
        if (wi >= nbatch || w >= nF) return;
        const int tid = (int)threadIdx.x;
        const bool owner = tid < KU * S;
        int u0, s_t;  // the thread's first unit and its shift
        divmod_small(tid, S, invS, u0, s_t);
        if (!owner) u0 = 0;
        const uint32_t id = st.F[w];
        const uint32_t lab = st.best_lab[id];
        const int64_t b = st.best_idx[id];
        const float bl = st.best_lik[id];
        // the skip test's values, the label's row count and first row, and the
        // loads that need only the window's label and onset go out together; a
        // window whose best spike fell below its threshold or too near the
        // chunk end takes no overlap recheck (binary_pursuit.cl:518-527)
        const float lik_b = st.LIK[(size_t)lab * N + b];
        const float thr_l = ts.thr[lab];
        const int nl = ts.nch[lab];
        const int offl = ts.off[lab];
        // first row and row count of unit u = tid (the bound terms below)
        const int q_r0 = tid < U ? ts.off[tid] : 0;
        const int q_nr = tid < U ? ts.nch[tid] : 0;
        const float* Bw = Bbuf + (size_t)wi * bstride;
        // Staging: every global load below is independent of the others, and
        // each load's first round is issued before any value is used, so they
        // are all in flight together.
        // per-pair state: B, sign of LIK[u][b+s-n]; loads from clamped (always
        // valid) addresses, B straight into the pair's register
        const int64_t as_t = b + s_t - n;
        const bool sin = owner && as_t >= 0 && as_t < N - L;
        float bv[kSelR];
        float lv[kSelR];
        uint32_t pum = 0;
        const int us = u0;
#pragma unroll
        for (int i = 0; i < kSelR; i++) {
            const int u = us + KU * i;
            const bool ok = sin && u < U;
            const int uc = ok ? u : 0;
            lv[i] = st.LIK[(size_t)uc * N + (ok ? as_t : 0)];
            bv[i] = Bw[ok ? uc * S + s_t : 0];
        }
        // first round of the window's samples (eight per thread), its pair ids
        // and its A row with the sign of LIK[l][b+f-n] (b + L <= N - 1 for a
        // window that is not skipped; clamped otherwise)
        const int64_t bs = b < N - L ? b : 0;
        int wj = tid / C, wc = tid - wj * C;
        const int wdj = bd / C, wdc = bd - wdj * C;
        float wvv[8];
        int wdst[8], wcq[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const bool ok = tid + q * bd < C * L;
            wdst[q] = ok ? wc * VS + wj : -1;
            wcq[q] = wc;
            wvv[q] = st.V[ok && bs + wj < N ? (size_t)(bs + wj) * C + wc : 0];
            qr_step(wj, wc, wdj, wdc, C);
        }
        const int pid0 = tid < U ? tb.pid[(size_t)lab * U + tid] : -1;
        const int64_t af0 = b + tid - n;
        const bool fin0 = tid < S && af0 >= 0 && af0 < N - L;
        const float la0 = st.LIK[(size_t)lab * N + (fin0 ? af0 : b)];
        const float A0 = Bw[(int)lab * S + (tid < S ? tid : 0)];
        if (tid == 0) {
            ncand_s = 0; plo_sh = S; phi_sh = -1; pfirst_sh = 0x7fffffff; npl_sh = 0;
            nstat_s[0] = 0u; nstat_s[1] = 0u;
            nrow_s = 0; nresc_s = 0; cun_s = 0ull;  // the rescoring's staging slots (its own LDS)
        }
        for (int c = tid; c < C; c += bd) vmax[c] = 0.0f;
        for (int u = tid; u < U; u += bd) slot_of[u] = -1;
        __syncthreads();
        if (lik_b < thr_l || b >= N - L) {
            if (tid == 0) {
                st.orc[id] = 0;
                if (xcand != nullptr) xcand[(size_t)wi * kXStride] = kXDone;
            }
            return;
        }
        // the label's zero-padded rows (eight per thread) and channels
        int rk = tid / LP, ri = tid - rk * LP;
        const int rdk = bd / LP, rdi = bd - rdk * LP;
        float rvv[8];
        int rdst[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int e = tid + q * bd;
            const bool in = e < nl * LP && ri >= n && ri < n + L;
            rdst[q] = e < nl * LP ? e : -1;
            rvv[q] = ts.rows[in ? (size_t)(offl + rk) * ts.Lp + (ri - n) : (size_t)offl * ts.Lp];
            rvv[q] = in ? rvv[q] : 0.0f;
            qr_step(rk, ri, rdk, rdi, LP);
        }
        const int cl0 = tid < nl ? ts.chan[offl + tid] : 0;
        __builtin_amdgcn_sched_barrier(0);  // every first-round load goes out before any is used
#pragma unroll
        for (int i = 0; i < kSelR; i++) {
            const int u = us + KU * i;
            const bool ok = sin && u < U;
            if (!ok) bv[i] = -INFINITY;
            pum |= (ok && lv[i] > 0.0f) ? 1u << i : 0u;
        }
        // vmax[c] = max_j |V[b+j][c]| (non-negative floats order as their bit
        // patterns; a NaN is skipped as fmaxf skips it)
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (wdst[q] < 0) continue;
            vW[wdst[q]] = wvv[q];
            const float a = fabsf(wvv[q]);
            if (a == a) atomicMax(reinterpret_cast<unsigned int*>(vmax) + wcq[q], __float_as_uint(a));
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (rdst[q] >= 0) TlP[rdst[q]] = rvv[q];
        if (tid < UP) {
            Dp[tid] = make_float2(0.0f, __int_as_float(pid0));
            if (pid0 >= 0) { atomicMin(&pfirst_sh, pid0); atomicAdd(&npl_sh, 1); }
        }
        if (tid < S) {
            const bool pl = fin0 && la0 > 0.0f;
            fin[tid] = fin0;
            posl[tid] = pl;
            if (pl) { atomicMin(&plo_sh, tid); atomicMax(&phi_sh, tid); }
            A[tid] = A0;
        }
        if (tid < nl) cl[tid] = cl0;
        // later rounds (large windows, label rows, unit counts or shift ranges only)
        for (int e0 = tid + 8 * bd; e0 < C * L; e0 += 8 * bd) {
            float v[8];
            int dst[8], cq[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const bool ok = e0 + q * bd < C * L;
                dst[q] = ok ? wc * VS + wj : -1;
                cq[q] = wc;
                v[q] = st.V[ok ? (size_t)(b + wj) * C + wc : (size_t)b * C];
                qr_step(wj, wc, wdj, wdc, C);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (dst[q] < 0) continue;
                vW[dst[q]] = v[q];
                const float a = fabsf(v[q]);
                if (a == a) atomicMax(reinterpret_cast<unsigned int*>(vmax) + cq[q], __float_as_uint(a));
            }
        }
        for (int e0 = tid + 8 * bd; e0 < nl * LP; e0 += 8 * bd) {
            float v[8];
            int dst[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int e = e0 + q * bd;
                const bool in = e < nl * LP && ri >= n && ri < n + L;
                dst[q] = e < nl * LP ? e : -1;
                v[q] = ts.rows[in ? (size_t)(offl + rk) * ts.Lp + (ri - n) : (size_t)offl * ts.Lp];
                v[q] = in ? v[q] : 0.0f;
                qr_step(rk, ri, rdk, rdi, LP);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (dst[q] >= 0) TlP[dst[q]] = v[q];
        }
        for (int i = tid; i < LP; i += bd) zrow[i] = 0.0f;
        for (int k = tid + bd; k < nl; k += bd) cl[k] = ts.chan[offl + k];
        for (int u = tid + bd; u < UP; u += bd) {
            const int p = u < U ? tb.pid[(size_t)lab * U + u] : -1;
            Dp[u] = make_float2(0.0f, __int_as_float(p));
            if (p >= 0) { atomicMin(&pfirst_sh, p); atomicAdd(&npl_sh, 1); }
        }
        for (int f = tid + bd; f < S; f += bd) {
            const int64_t af = b + f - n;
            const bool in = af >= 0 && af < N - L;
            const bool pl = in && st.LIK[(size_t)lab * N + af] > 0.0f;
            fin[f] = in;
            posl[f] = pl;
            if (pl) { atomicMin(&plo_sh, f); atomicMax(&phi_sh, f); }
            A[f] = Bw[(int)lab * S + f];
        }
        // channels and L1 norms of the first eight rows of unit u = tid (the
        // bound terms below): in flight across the barrier
        int qch[8];
        float ql1[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int r = q_r0 + min(j, max(q_nr, 1) - 1);
            const bool ok = tid < U && q_nr > 0;
            qch[j] = ok ? ts.chan[r] : 0;
            ql1[j] = ok ? tb.rowL1[r] : 0.0f;
        }
        __syncthreads();  // staging, pair ids, plo / phi, pfirst / npl are complete
        // Gram minima, the range maxima of A and the per-unit bound terms are
        // independent of each other: one phase
        // least G of every pair of the label over every f (gmA)
        {
            const int pf = pfirst_sh, np = npl_sh;
#pragma unroll 4
            for (int k = tid; k < np * S; k += bd) gmA[k] = tb.minG[(size_t)pf * S + k];
        }
        // range maxima of A over f in [s-L, s+L]: all in-bound f / those with LIK > 0
        if (n <= L) {
            if (wv < 2) {
                const bool suffix = wv == 1;
                float* Xa = suffix ? SMa : PMa;
                float* Xp = suffix ? SMp : PMp;
                float ca = -INFINITY, cp = -INFINITY;
                for (int c0 = 0; c0 < S; c0 += 64) {
                    const int f = suffix ? S - 1 - (c0 + lane) : c0 + lane;
                    const bool ok = c0 + lane < S;
                    float va = -INFINITY, vp = -INFINITY;
                    if (ok) {
                        if (fin[f]) va = A[f];
                        if (posl[f]) vp = A[f];
                    }
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) {
                        const float ta = __shfl_up(va, off, 64), tq = __shfl_up(vp, off, 64);
                        if (lane >= off) { va = fmaxf(va, ta); vp = fmaxf(vp, tq); }
                    }
                    va = fmaxf(va, ca);
                    vp = fmaxf(vp, cp);
                    if (ok) { Xa[f] = va; Xp[f] = vp; }
                    ca = __shfl(va, 63, 64);
                    cp = __shfl(vp, 63, 64);
                }
            }
        } else {
            for (int s = tid; s < S; s += bd) {
                float ma = -INFINITY, mp = -INFINITY;
                const int fl = max(0, s - L), fh = min(S - 1, s + L);
                for (int f = fl; f <= fh; f++) {
                    if (fin[f]) ma = fmaxf(ma, A[f]);
                    if (posl[f]) mp = fmaxf(mp, A[f]);
                }
                RAall[s] = ma;
                RApos[s] = mp;
            }
        }
        // the window maximum over every channel, in every wave
        {
            float mw = 0.0f;
            for (int c = lane; c < C; c += 64) mw = fmaxf(mw, vmax[c]);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) mw = fmaxf(mw, __shfl_xor(mw, off, 64));
            for (int u = tid; u < U; u += bd) {
                float q = 0.0f, t1 = 0.0f, vsum = 0.0f;
                const int r0 = u == tid ? q_r0 : ts.off[u], nr = u == tid ? q_nr : ts.nch[u];
                // eight rows at a time: their channel and L1 loads go out together
                for (int k0 = 0; k0 < nr; k0 += 8) {
                    int ch[8];
                    float l1[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        if (k0 == 0 && u == tid) {
                            ch[j] = qch[j];
                            l1[j] = ql1[j];
                        } else {
                            const int r = r0 + min(k0 + j, nr - 1);
                            ch[j] = ts.chan[r];
                            l1[j] = tb.rowL1[r];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        if (k0 + j < nr) {
                            const float vm = vmax[ch[j]];
                            q += l1[j] * vm;
                            t1 += l1[j];
                            vsum += vm;
                        }
                    }
                }
                float z = 0.0f;
                if (tb.kappa3 != 0.0f)  // fp16 operands: absolute residue of the scaled split
                    z = tb.kappa3 * mw * t1 + tb.kappa4 * vsum;
                Q[u] = (tb.kappa + tb.kappa2) * q + 2.0f * tb.kappa * tb.SS[u] + z;
            }
        }
        __syncthreads();
        if (n <= L) {
            for (int s = tid; s < S; s += bd) {
                const int fl = max(0, s - L), fh = min(S - 1, s + L);
                RAall[s] = (fl == 0) ? PMa[fh] : SMa[fl];
                RApos[s] = (fl == 0) ? PMp[fh] : SMp[fl];
            }
        }
        // delta(l, u) = P_l + P_u: the bound is a sum of one term per template
        {
            const float ql = Q[lab];
            for (int u = tid; u < U; u += bd) Dp[u].x = ql + Q[u];
        }
        __syncthreads();
        const float rall = RAall[s_t], rpos = RApos[s_t];
        // pair i of the thread reads dpu[KU * i]; every phase recomputes these
        // addresses from an opaque copy of u0 rather than keeping 24 of them live
        auto pair_base = [&]() {
            int ub = u0;
            asm volatile("" : "+v"(ub));
            return Dp + ub;
        };
        // phase 1: pairs sharing no channel, E-hat = A(f) + B(u, s) separable;
        // M' = max(E-hat - delta_u) is a lower bound of the exact maximum
        float mns = -INFINITY;
        uint32_t shm = 0;
        const float2* dpu = pair_base();
#pragma unroll
        for (int i = 0; i < kSelR; i++) {
            const float2 dp = dpu[KU * i];
            const bool sh = __float_as_int(dp.y) >= 0;
            shm |= sh ? 1u << i : 0u;
            const float ra = ((pum >> i) & 1u) ? rall : rpos;
            const bool ok = bv[i] != -INFINITY && !sh && ra != -INFINITY;
            mns = fmaxf(mns, ok ? (ra + bv[i]) - dp.x : -INFINITY);
        }
        const float Mns = block_max_n(mns, red, nw);
        // phase 2: pairs sharing a channel whose bound reaches
        // tp_u = max(M' - delta_u, bl - delta_u); the wave enumerates f of
        // each, tracks M' and keeps items with E-hat >= tp_u.  Bound: the
        // largest A over the allowed f within L of s (any in-bound f when
        // LIK[u][b+s-n] > 0, else those with LIK[l][b+f-n] > 0) plus B minus
        // the least G over the blocks of f that can hold an allowed f.  That
        // least G is the least over every f (gmA) when every f is allowed or
        // the f with LIK[l][b+f-n] > 0 span more than 3 blocks, else the least
        // over their (at most 3) blocks; gmA <= it, so the bound with gmA is
        // tested first and the block minima are read only for the pairs that
        // pass it (the same survivors as the block bound alone).
        uint32_t sv = 0;
        {
            const int pf = pfirst_sh, plo = plo_sh, phi = phi_sh;
            const int j1 = max(plo, 0) / kGB, j3 = max(phi, 0) / kGB, j2 = min(j1 + 1, j3);
            const bool narrow = phi >= 0 && j3 - j1 <= 2;
            const float2* dpu = pair_base();
#pragma unroll
            for (int i = 0; i < kSelR; i++) {
                const float2 dp = dpu[KU * i];
                const int pw = __float_as_int(dp.y);
                const bool pu = (pum >> i) & 1u;
                const float ra = pu ? rall : rpos;
                const int kk = pw >= 0 ? (pw - pf) * S + s_t : 0;  // row of (label, u) in the tables
                const float tp = fmaxf(Mns - dp.x, bl - dp.x);
                bool hit = pw >= 0 && ra != -INFINITY && !((ra + bv[i]) - gmA[kk] < tp);
                if (hit && !pu && narrow) {
                    const float* g = tb.minGb + ((size_t)pf * S + kk) * tb.nbg;
                    hit = !((ra + bv[i]) - fminf(fminf(g[j1], g[j2]), g[j3]) < tp);
                }
                sv |= hit ? 1u << i : 0u;
            }
        }
        float msh = -INFINITY;
        bool overflow = false;
        auto push = [&](float eh, int u, int f, int s) {
            const uint32_t slot = atomicAdd(&ncand_s, 1u);
            if (slot < kCandCap) {
                cand_e[slot] = eh;
                cand_k[slot] = (uint32_t)u + (uint32_t)U * ((uint32_t)f + (uint32_t)S * (uint32_t)s);
            } else {
                overflow = true;
            }
        };
        const int plo = plo_sh, phi = phi_sh;
        {
            const int c = wave_sum_u32(__popc(sv));
            if (lane == 0 && c) atomicAdd(&nstat_s[0], (uint32_t)c);
        }
        // Surviving pairs go to two lists with their B: pairs with
        // LIK[u][b+s-n] <= 0 allow only the f with LIK[l][b+f-n] > 0, all in
        // [plo, phi] (one thread per pair); pairs with LIK[u][b+s-n] > 0 allow
        // every in-bound f within L of s (one thread per (pair, f)).
        // Phase 2 (shared channel, G term, threshold tp_u, tracks M') and
        // phase 3 (no shared channel, threshold theta_u) use the same code.
        const uint32_t capN = survcap * 3 / 4, capP = survcap - capN;
        float Msh = -INFINITY;  // max over the enumerated shared-channel items of E-hat - delta_u (every thread)
#pragma unroll 1
        for (int ph = 2; ph <= 3; ph++) {
            const bool shared = ph == 2;
            const float Mref = shared ? Mns : fmaxf(Mns, Msh);
            auto thr_u = [&](int u) { return fmaxf(Mref - Dp[u].x, bl - Dp[u].x); };
            auto gterm = [&](int u, int s, int f) -> float {
                return shared ? tb.G[((size_t)__float_as_int(Dp[u].y) * S + s) * tb.sg + f] : 0.0f;
            };
            auto eval_pl = [&](int e, float Bu) {
                int u, s;
                divmod_small(e, S, invS, u, s);
                const float Du = Dp[u].x;
                const float tp = thr_u(u);
                const int fa = max(plo, s - L), fb = min(phi, s + L);
                for (int f0 = fa; f0 <= fb; f0 += 4) {
                    float g[4], a[4];
                    bool pl[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int f = min(f0 + q, fb);
                        g[q] = gterm(u, s, f);
                        a[q] = A[f];
                        pl[q] = posl[f] != 0;
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        if (f0 + q > fb || !pl[q]) continue;
                        const float eh = shared ? (a[q] + Bu) - g[q] : a[q] + Bu;
                        if (shared) msh = fmaxf(msh, eh - Du);
                        if (!(eh < tp)) push(eh, u, f0 + q, s);
                    }
                }
            };
            auto eval_all = [&](int e, float Bu, int f) {
                int u, s;
                divmod_small(e, S, invS, u, s);
                const float g = gterm(u, s, f);
                if (!fin[f] || abs(f - s) > L) return;
                const float eh = shared ? (A[f] + Bu) - g : A[f] + Bu;
                if (shared) msh = fmaxf(msh, eh - Dp[u].x);
                if (!(eh < thr_u(u))) push(eh, u, f, s);
            };
            if (!shared) {
                // pairs sharing no channel whose exact best over f reaches theta_u
                uint32_t pm = pum, sm2 = shm;
                asm volatile("" : "+v"(pm), "+v"(sm2));  // do not keep per-pair selections live across the phases
                const float2* dpu = pair_base();
#pragma unroll
                for (int i = 0; i < kSelR; i++) {
                    const float2 dp = dpu[KU * i];
                    const float ra = ((pm >> i) & 1u) ? rall : rpos;
                    const float th = fmaxf(Mref - dp.x, bl - dp.x);
                    const bool hit = bv[i] != -INFINITY && !((sm2 >> i) & 1u) &&
                                     !(ra == -INFINITY || ra + bv[i] < th);
                    sv |= hit ? 1u << i : 0u;
                }
                const int c = wave_sum_u32(__popc(sv));
                if (lane == 0 && c) atomicAdd(&nstat_s[1], (uint32_t)c);
            }
            if (tid == 0) { nsurv_s = 0; nsurvp_s = 0; }
            __syncthreads();
            uint32_t spill = 0;
            uint32_t pm = pum;
            int ub = u0;
            asm volatile("" : "+v"(pm), "+v"(ub));  // do not keep per-pair selections live across the phases
#pragma unroll
            for (int i = 0; i < kSelR; i++) {
                if ((sv >> i) & 1u) {
                    const bool pu = (pm >> i) & 1u;
                    const uint32_t slot = atomicAdd(pu ? &nsurvp_s : &nsurv_s, 1u);
                    if (slot < (pu ? capP : capN)) {
                        const uint32_t o = pu ? capN + slot : slot;
                        surv_e[o] = (uint32_t)((ub + KU * i) * S + s_t);
                        surv_b[o] = bv[i];
                    } else {
                        spill |= 1u << i;
                    }
                }
            }
            sv = 0;
            __syncthreads();
            const uint32_t nN = min(nsurv_s, capN), nP = min(nsurvp_s, capP);
            // (pair, f-quad) items of a survivor list: the f in [flo, fhi] taken four at a
            // time from fo = flo rounded down to a multiple of 4, so that a quad's G terms
            // are one aligned float4 of the pair's G row; spread over the block kQR
            // quads per thread per round, with their loads issued together (the order of
            // candidates and of the running maximum does not matter).  An f counts when
            // it lies in [flo, fhi], is allowed (posl for the first list, fin for the
            // second) and lies within L of s.
            auto run_items = [&](uint32_t base, uint32_t np, int flo, int fhi, bool plm) {
                const int fo = flo & ~3;
                const int nq = (fhi - fo) / 4 + 1;
                const uint32_t nI = np * (uint32_t)nq;
                const float invq = 1.0f / (float)nq;
                for (uint32_t j0 = tid; j0 < nI; j0 += (uint32_t)kQR * (uint32_t)bd) {
                    int uq[kQR], sq[kQR], fq[kQR];
                    float bq[kQR];
                    float4 gq[kQR];
                    bool okq[kQR];
#pragma unroll
                    for (int q = 0; q < kQR; q++) {
                        const uint32_t j = j0 + (uint32_t)q * (uint32_t)bd;
                        okq[q] = j < nI;
                        int pi, t;
                        divmod_small(okq[q] ? (int)j : 0, nq, invq, pi, t);
                        const int e = (int)surv_e[base + pi];
                        bq[q] = surv_b[base + pi];
                        divmod_small(e, S, invS, uq[q], sq[q]);
                        fq[q] = fo + 4 * t;
                        gq[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if (okq[q] && shared)
                            gq[q] = *reinterpret_cast<const float4*>(
                                tb.G + ((size_t)__float_as_int(Dp[uq[q]].y) * S + sq[q]) * tb.sg + fq[q]);
                    }
                    __builtin_amdgcn_sched_barrier(0);  // the quads' G loads are in flight together
#pragma unroll
                    for (int q = 0; q < kQR; q++) {
                        if (!okq[q]) continue;
                        const int u = uq[q], s = sq[q];
                        const float gv[4] = {gq[q].x, gq[q].y, gq[q].z, gq[q].w};
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int f = fq[q] + r;
                            if (f < flo || f > fhi || !(plm ? posl[f] : fin[f]) || abs(f - s) > L) continue;
                            const float eh = shared ? (A[f] + bq[q]) - gv[r] : A[f] + bq[q];
                            if (shared) msh = fmaxf(msh, eh - Dp[u].x);
                            if (!(eh < thr_u(u))) push(eh, u, f, s);
                        }
                    }
                }
            };
            // pairs with LIK[u][b+s-n] <= 0: f in [plo, phi] (none when phi < plo)
            if (phi >= plo) run_items(0, nN, plo, phi, true);
            // pairs with LIK[u][b+s-n] > 0: every f
            run_items(capN, nP, 0, S - 1, false);
            for (uint32_t m = spill; m; m &= m - 1) {
                const int i = __builtin_ctz(m);
                const int e = (u0 + KU * i) * S + s_t;
                const float Bu = Bw[e];
                if ((pum >> i) & 1u) {
                    for (int f = 0; f < S; f++) eval_all(e, Bu, f);
                } else {
                    eval_pl(e, Bu);
                }
            }
            if (shared) Msh = block_max_n(msh, red, nw);  // every thread gets the block's maximum
        }
        const float Mp = fmaxf(Mns, Msh);
        // the barrier of the overflow vote also completes the candidate list and the statistics
        const bool ovf = __syncthreads_or(overflow);
        // one global add per window and statistic (not one per wave: every block adds to the same words)
        if (tid == 0) {
            if (nstat_s[0]) atomicAdd(st.counters + CNT_SURV, nstat_s[0]);
            if (nstat_s[1]) atomicAdd(st.counters + CNT_NSP, nstat_s[1]);
        }
        const uint32_t nc = min(ncand_s, (uint32_t)kCandCap);
        float bv_best = 0.0f;
        uint32_t bk = kNoItem;
        if (ovf) {
            // more candidates than the list holds: enumerate and rescore in place
            if (xcand != nullptr && tid == 0) xcand[(size_t)wi * kXStride] = kXDone;
            for (int e = tid; e < US; e += bd) {
                int u, s;
                divmod_small(e, S, invS, u, s);
                const int64_t as = b + s - n;
                if (as < 0 || as >= N - L) continue;
                const float Bu = Bw[e];
                const float2 dp = Dp[u];
                const int p = __float_as_int(dp.y);
                const bool pu = st.LIK[(size_t)u * N + as] > 0.0f;
                const float theta = fmaxf(Mp - dp.x, bl - dp.x);
                const float* Gr = p >= 0 ? tb.G + ((size_t)p * S + s) * tb.sg : nullptr;
                for (int f = 0; f < S; f++) {
                    if (!fin[f] || abs(f - s) > L || !(posl[f] || pu)) continue;
                    const float eh = p >= 0 ? (A[f] + Bu) - Gr[f] : A[f] + Bu;
                    if (eh < theta) continue;
                    const float v = recheck_item_value(ts, vW, VS, 1, lab, (uint32_t)u, f, s, n);
                    argmax_merge(bv_best, bk, v, (uint32_t)u + (uint32_t)U * ((uint32_t)f + (uint32_t)S * (uint32_t)s));
                }
            }
        } else {
            // candidates reaching theta_u, each rescored exactly by one thread;
            // the zero-padded rows of their units are staged in LDS, upr units
            // per round
            // (slot_of, nrow_s, nresc_s and cun_s were reset with the window's other state)
            uint32_t* xw = xcand != nullptr ? xcand + (size_t)wi * kXStride : nullptr;
            for (uint32_t i = tid; i < nc; i += bd) {
                const int u = (int)(cand_k[i] % (uint32_t)U);
                const float du = Dp[u].x;
                if (cand_e[i] < fmaxf(Mp - du, bl - du)) continue;
                if (xw != nullptr) {
                    const uint32_t q = atomicAdd(&nresc_s, 1u);
                    if (q < 64u) xw[1 + q] = cand_k[i];
                }
                if (atomicCAS(&slot_of[u], -1, -2) == -1) {
                    const int o = (int)atomicAdd(&nrow_s, 1u);
                    slot_nr[o] = ts.nch[u];
                    slot_off[o] = ts.off[u];
                    slot_of[u] = o;
                    if (xw != nullptr && C <= 64) atomicOr(&cun_s, (unsigned long long)ts.cmask[u]);
                }
            }
            __syncthreads();
            const int nun = (int)nrow_s;
            if (xw != nullptr) {
                // the candidates fit one wave of the rescoring kernel: it finishes the window
                const uint32_t nr = nresc_s;
                const bool defer = nr >= 1u && nr <= 64u && nun <= kRU && C <= 64 &&
                                   __popcll(cun_s | ts.cmask[lab]) <= kCV;
                if (tid == 0) xw[0] = defer ? nr : kXDone;
                if (defer) return;  // uniform over the block
            }
            const int upr = max(1, srows / CM);
            const float invCM = 1.0f / (float)CM;
            for (int r0u = 0; r0u < nun; r0u += upr) {
                if (r0u > 0) __syncthreads();  // the previous round's chains are done
                // the round's units staged by the whole block, eight loads in flight per
                // thread: element e = rr * LP + i of row rr = o * CM + k (unit o, row k)
                const int tot = min(upr, nun - r0u) * CM * LP;
                int rr = tid / LP, ii = tid - rr * LP;
                const int drr = bd / LP, dii = bd - drr * LP;
                for (int e0 = tid; e0 < tot; e0 += 8 * bd) {
                    float v[8];
                    int dst[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int e = e0 + q * bd;
                        int o, k;
                        divmod_small(rr, CM, invCM, o, k);
                        const bool row = e < tot && k < slot_nr[r0u + min(o, upr - 1)];
                        const bool in = row && ii >= n && ii < n + L;
                        dst[q] = row ? e : -1;
                        v[q] = ts.rows[in ? (size_t)(slot_off[r0u + o] + k) * ts.Lp + (ii - n) : 0];
                        v[q] = in ? v[q] : 0.0f;
                        qr_step(rr, ii, drr, dii, LP);
                    }
                    __builtin_amdgcn_sched_barrier(0);  // the eight loads are in flight together
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        if (dst[q] >= 0) stgP[dst[q]] = v[q];
                }
                for (int e = tid; e < min(upr, nun - r0u) * CM; e += bd) {
                    int o, k;
                    divmod_small(e, CM, invCM, o, k);
                    if (k < slot_nr[r0u + o]) stg_ch[e] = ts.chan[slot_off[r0u + o] + k];
                }
                __syncthreads();
                // candidates packed into the first waves: a wave runs the longest chain of its lanes
                for (uint32_t i = tid; i < nc; i += bd) {
                    const uint32_t k = cand_k[i];
                    const int u = (int)(k % (uint32_t)U);
                    const int f = (int)((k / (uint32_t)U) % (uint32_t)S), s = (int)(k / ((uint32_t)U * S));
                    const float du = Dp[u].x;
                    if (cand_e[i] < fmaxf(Mp - du, bl - du)) continue;
                    const int o = slot_of[u];
                    if (o < r0u || o >= r0u + upr) continue;
                    const int rb = (o - r0u) * CM;
                    const float v = exact_item_padded<LT>(vW, L, VS, TlP, cl, nl, stgP + rb * LP, stg_ch + rb, ts.nch[u],
                                                          zrow, LP, f, s, n);
                    argmax_merge(bv_best, bk, v, k);
                }
            }
        }
        wave_argmax(bv_best, bk);
        // (red was last read in the phase-2 maximum, several barriers ago)
        if (lane == 0) { red[wv] = bv_best; redk[wv] = bk; }
        __syncthreads();
        if (tid == 0) {
            float v = red[0];
            uint32_t k = redk[0];
            for (int i = 1; i < nw; i++) argmax_merge(v, k, red[i], redk[i]);
            apply_parse_at(ts, st, n, id, v, k, bl, lab, b, S);
            atomicAdd(st.counters + CNT_EVAL, ncand_s);
            if (ncand_s > (uint32_t)kCandCap) atomicAdd(st.counters + CNT_CAND, 1u);
        }
    }
}

template <int LT>
__global__ __launch_bounds__(kSelMaxBlock) void k_screen_select_reg(TemplateSet ts, PassState st, ScreenTables tb,
                                                                    uint32_t f0, uint32_t nbatch,
                                                                    const float* __restrict__ Bbuf, size_t bstride,
                                                                    uint32_t* __restrict__ xcand)
{
    screen_select_reg_body<LT>(ts, st, tb, f0, nbatch, Bbuf, bstride, blockIdx.x, gridDim.x, xcand);
}

// ---------------------------------------------------------------------------
// Deferred exact rescoring: one wave per window whose selection left its
// candidates (at most 64, of at most kRU units) in its record.  The wave stages
// the label's and the units' zero-padded rows and the window's samples of
// their channels (compact slots, ascending channel order), runs one exact FP32
// chain per candidate (exact_item_padded, the order of the literal kernel),
// takes the arg-max (value, then lowest item id) and applies the parse cascade:
// the same result as the selection kernel's own rescoring.
// ---------------------------------------------------------------------------
template <int LT>
__global__ __launch_bounds__(64) void k_screen_rescore(TemplateSet ts, PassState st, ScreenTables tb, uint32_t f0,
                                                       uint32_t nbatch, const uint32_t* __restrict__ xcand)
{
    extern __shared__ __align__(16) float sx[];
    const uint32_t wi = blockIdx.x, w = f0 + wi;
    if (wi >= nbatch || w >= st.counters[CNT_F]) return;
    const uint32_t* xw = xcand + (size_t)wi * kXStride;
    const uint32_t cnt = xw[0];
    if (cnt == kXDone || cnt == 0u || cnt > 64u) return;  // uniform over the block
    const int lane = (int)threadIdx.x;
    const int U = ts.U, C = ts.C;
    const int L = LT > 0 ? LT : ts.L;
    const int S = tb.S, n = tb.n, CM = tb.cmax;
    const int LP = L + 2 * n, VS = L | 1;
    float* TlP = sx;                      // [CM][LP] the label's zero-padded rows
    float* zrow = TlP + CM * LP;          // [LP]
    float* Tu = zrow + LP;                // [kRU][CM][LP] the units' zero-padded rows
    float* vW = Tu + kRU * CM * LP;       // [kCV][VS] window samples of the slotted channels
    int* clS = reinterpret_cast<int*>(vW + kCV * VS);  // [CM] slots of the label's channels
    int* cuS = clS + CM;                  // [kRU][CM] slots of each unit's channels
    int* uid = cuS + kRU * CM;            // [kRU] unit of each staging slot
    const uint32_t id = st.F[w];
    const uint32_t lab = st.best_lab[id];
    const int64_t b = st.best_idx[id];
    const float bl = st.best_lik[id];
    const int nl = ts.nch[lab], offl = ts.off[lab];
    const uint32_t key = lane < (int)cnt ? xw[1 + lane] : kNoItem;
    const int u = key != kNoItem ? (int)(key % (uint32_t)U) : -1;
    const uint32_t r = key != kNoItem ? key / (uint32_t)U : 0u;
    const int f = (int)(r % (uint32_t)S), s = (int)(r / (uint32_t)S);
    // distinct units in lane order: the first lane of each takes a staging slot
    bool first = u >= 0;
    for (int j = 0; j < 64; j++) {
        const int uj = __shfl(u, j, 64);
        if (j < lane && uj == u) first = false;
    }
    const uint64_t fm = __ballot(first);
    const int nun = min(__popcll(fm), kRU);
    if (first) {
        const int o = __popcll(fm & ((1ull << lane) - 1ull));
        if (o < kRU) uid[o] = u;
    }
    // channel slots: the label's and the units' channels, ascending
    unsigned long long m = u >= 0 ? (unsigned long long)ts.cmask[u] : 0ull;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off, 64);
    m |= (unsigned long long)ts.cmask[lab];
    __syncthreads();
    // the zero padding of every staged row first (LDS stores only), then the
    // rows' L samples and the window's samples, eight loads in flight per lane
    const int nrow_u = nun;  // units staged
    for (int e = lane; e < (1 + kRU) * CM * LP + LP; e += 64) TlP[e] = 0.0f;  // TlP, zrow, Tu are contiguous
    if (lane < nl) clS[lane] = __popcll(m & ((1ull << ts.chan[offl + lane]) - 1ull));
    for (int o = 0; o < nrow_u; o++) {
        const int uo = uid[o], ro = ts.off[uo];
        if (lane < ts.nch[uo]) cuS[o * CM + lane] = __popcll(m & ((1ull << ts.chan[ro + lane]) - 1ull));
    }
    __syncthreads();
    {
        // template samples: rows of the label, then of each staged unit; element e = row * L + i
        int rows_tot = nl, row_base[kRU + 1];
        row_base[0] = nl;
        for (int o = 0; o < nrow_u; o++) {
            rows_tot += ts.nch[uid[o]];
            row_base[o + 1] = rows_tot;
        }
        const int tot = rows_tot * L;
        for (int e0 = lane; e0 < tot; e0 += 8 * 64) {
            float v[8];
            int dst[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int e = e0 + q * 64;
                dst[q] = -1;
                v[q] = 0.0f;
                if (e < tot) {
                    const int row = e / L, i = e - row * L;
                    int src_row, d0;
                    if (row < nl) {
                        src_row = offl + row;
                        d0 = row * LP;
                    } else {
                        int o = 0;
                        while (o < nrow_u - 1 && row >= row_base[o + 1]) o++;
                        const int k = row - row_base[o];
                        src_row = ts.off[uid[o]] + k;
                        d0 = (int)(Tu - TlP) + o * CM * LP + k * LP;
                    }
                    dst[q] = d0 + n + i;
                    v[q] = ts.rows[(size_t)src_row * ts.Lp + i];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (dst[q] >= 0) TlP[dst[q]] = v[q];
        }
    }
    {
        // the window's C x L samples are one contiguous run of the sample-major voltage
        const float* src = st.V + (size_t)b * C;
        const int tot = C * L;
        for (int e0 = lane; e0 < tot; e0 += 8 * 64) {
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int e = e0 + q * 64;
                v[q] = src[e < tot ? e : 0];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int e = e0 + q * 64;
                if (e >= tot) continue;
                const int j = e / C, c = e - j * C;
                if ((m >> c) & 1ull) vW[__popcll(m & ((1ull << c) - 1ull)) * VS + j] = v[q];
            }
        }
    }
    __syncthreads();
    float v = 0.0f;
    uint32_t k = kNoItem;
    if (u >= 0) {
        int o = 0;
        while (o < kRU - 1 && uid[o] != u) o++;
        v = exact_item_padded<LT>(vW, L, VS, TlP, clS, nl, Tu + (size_t)o * CM * LP, cuS + o * CM, ts.nch[u], zrow, LP,
                                  f, s, n);
        k = key;
    }
    wave_argmax(v, k);
    if (lane == 0) apply_parse_at(ts, st, n, id, v, k, bl, lab, b, S);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned sgrid(int64_t n, int64_t per, unsigned cap)
{
    int64_t g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

void launch_screen_tables(const TemplateSet& ts, const ScreenTables& tb, int64_t npairs, hipStream_t s)
{
    if (npairs > 0)
        hipLaunchKernelGGL(k_tab_gram, dim3(sgrid(npairs * tb.S * tb.S, kBlock, 1u << 20)), dim3(kBlock), 0, s, ts, tb,
                           npairs);
    hipLaunchKernelGGL(k_tab_min_ntr, dim3(sgrid(npairs * tb.S + (int64_t)ts.U * tb.S, kBlock, 1u << 20)),
                       dim3(kBlock), 0, s, ts, tb, npairs);
    hipLaunchKernelGGL(k_tab_interleave, dim3(sgrid((int64_t)tb.cmax * ts.Lp * tb.upad, kBlock, 1u << 20)), dim3(kBlock),
                       0, s, ts, tb);
}

size_t screen_corr_lds(const TemplateSet& ts, int S, int* rp)
{
    const int ngr = (S + kPos - 1) / kPos;
    // rows of 4k floats with k odd: float4 reads, and lanes on different
    // channels start in different banks; (ngr - 1) kPos + L + kPos samples read
    int q = (ngr * kPos + ts.L + 3) / 4;
    if (q % 2 == 0) q++;
    *rp = 4 * q;
    return (size_t)ts.C * (*rp) * sizeof(float);
}

size_t screen_select_lds(const TemplateSet& ts, int S, int cmax)
{
    const int nwords = (ts.U * S + 63) / 64 * 2;
    return (size_t)(ts.C * ts.L + ts.C + 7 * S + 2 * ts.U + cmax * ts.L + cmax) * 4 + (size_t)nwords * 4 +
           2 * (size_t)S + 16;
}

// Floats of the selection's phase-dependent region: the two tables of Gram
// minima of one label (one in each half), the survivor lists, or the staged
// padded rows.
int screen_select_region(const TemplateSet& ts, int S, int cmax, int npl_max)
{
    const int LP = ts.L + S - 1;
    int rg = std::max(2 * npl_max * S, kStageRegion);  // both tables of Gram minima
    rg = std::max(rg, cmax * LP);  // one unit's rows at least
    return (rg + 3) & ~3;
}

size_t screen_select_reg_lds(const TemplateSet& ts, int S, int cmax, int npl_max)
{
    const int LP = ts.L + S - 1;
    const int rg = screen_select_region(ts, S, cmax, npl_max);
    const size_t dpo = ((size_t)ts.C * (ts.L | 1) + ts.C + 7 * (size_t)S + ts.U + 1) & ~(size_t)1;
    const size_t up = (size_t)sel_slices(ts.U) * kSelR;
    return (dpo + 2 * up + (size_t)cmax * LP + LP + cmax + (size_t)rg + rg / LP + 3 * (size_t)ts.U) * 4 +
           2 * (size_t)S + 16;
}

// Dynamic LDS above 64 KiB needs the per-kernel attribute (gfx950 allows 160 KiB).
template <class K>
static void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

// The register-resident selection: one block per flagged window of the batch.
template <class K>
static void launch_select_reg(K kernel, int bd, size_t lds, uint32_t nbatch, hipStream_t s, const TemplateSet& ts,
                              const PassState& st, const ScreenTables& tb, uint32_t f0, const float* Bbuf,
                              uint32_t* xcand)
{
    allow_lds(kernel, lds);
    hipLaunchKernelGGL(kernel, dim3(nbatch), dim3(bd), lds, s, ts, st, tb, f0, nbatch, Bbuf, (size_t)ts.U * tb.S,
                       xcand);
}

// Dynamic LDS of the deferred rescoring; 0 when it cannot take this template
// set (more than 64 channels, or rows that do not fit 64 KiB).
size_t screen_rescore_lds(const TemplateSet& ts, const ScreenTables& tb)
{
    if (ts.C > 64) return 0;
    const size_t LP = (size_t)ts.L + 2 * (size_t)tb.n, VS = (size_t)(ts.L | 1), CM = (size_t)tb.cmax;
    const size_t bytes = (CM * LP * (1 + kRU) + LP + (size_t)kCV * VS + CM * (1 + kRU) + kRU) * 4;
    return bytes <= 64 * 1024 ? bytes : 0;
}

void launch_screen_rescore(const TemplateSet& ts, const PassState& st, const ScreenTables& tb, uint32_t f0,
                           uint32_t nbatch, const uint32_t* xcand, hipStream_t s)
{
    const size_t lds = screen_rescore_lds(ts, tb);
    if (nbatch == 0 || lds == 0) return;
    const dim3 g(nbatch), b(64);
    if (ts.L == 48) hipLaunchKernelGGL(k_screen_rescore<48>, g, b, lds, s, ts, st, tb, f0, nbatch, xcand);
    else if (ts.L == 64) hipLaunchKernelGGL(k_screen_rescore<64>, g, b, lds, s, ts, st, tb, f0, nbatch, xcand);
    else if (ts.L == 32) hipLaunchKernelGGL(k_screen_rescore<32>, g, b, lds, s, ts, st, tb, f0, nbatch, xcand);
    else hipLaunchKernelGGL(k_screen_rescore<0>, g, b, lds, s, ts, st, tb, f0, nbatch, xcand);
}

// Returns whether the selection left windows to the deferred rescoring (xcand
// given, the register-resident selection taken and the rescoring's LDS fits).
bool launch_screen_select(const TemplateSet& ts, const PassState& st, const ScreenTables& tb, uint32_t f0,
                          uint32_t nbatch, const float* Bbuf, bool streaming, hipStream_t s, uint32_t* xcand)
{
    if (nbatch == 0) return false;
    const int64_t owners = (int64_t)sel_slices(ts.U) * tb.S;  // threads owning pairs
    if (!streaming && owners <= kSelMaxBlock &&
        screen_select_reg_lds(ts, tb.S, tb.cmax, tb.npl_max) <= kSelMaxLds) {
        const int bd = (int)std::max<int64_t>(kBlock, (owners + 63) / 64 * 64);
        const size_t lds = screen_select_reg_lds(ts, tb.S, tb.cmax, tb.npl_max);
        uint32_t* xc = screen_rescore_lds(ts, tb) > 0 ? xcand : nullptr;
        if (ts.L == 48) launch_select_reg(k_screen_select_reg<48>, bd, lds, nbatch, s, ts, st, tb, f0, Bbuf, xc);
        else if (ts.L == 64) launch_select_reg(k_screen_select_reg<64>, bd, lds, nbatch, s, ts, st, tb, f0, Bbuf, xc);
        else if (ts.L == 32) launch_select_reg(k_screen_select_reg<32>, bd, lds, nbatch, s, ts, st, tb, f0, Bbuf, xc);
        else launch_select_reg(k_screen_select_reg<0>, bd, lds, nbatch, s, ts, st, tb, f0, Bbuf, xc);
        return xc != nullptr;
    }
    const dim3 g(sgrid(nbatch, 1, 1u << 20)), b(kBlock);
    const size_t lds2 = screen_select_lds(ts, tb.S, tb.cmax);
    if (ts.L == 48) {
        allow_lds(k_screen_select<48>, lds2);
        hipLaunchKernelGGL(k_screen_select<48>, g, b, lds2, s, ts, st, tb, f0, nbatch, Bbuf);
    } else if (ts.L == 64) {
        allow_lds(k_screen_select<64>, lds2);
        hipLaunchKernelGGL(k_screen_select<64>, g, b, lds2, s, ts, st, tb, f0, nbatch, Bbuf);
    } else if (ts.L == 32) {
        allow_lds(k_screen_select<32>, lds2);
        hipLaunchKernelGGL(k_screen_select<32>, g, b, lds2, s, ts, st, tb, f0, nbatch, Bbuf);
    } else {
        allow_lds(k_screen_select<0>, lds2);
        hipLaunchKernelGGL(k_screen_select<0>, g, b, lds2, s, ts, st, tb, f0, nbatch, Bbuf);
    }
    return false;
}

void launch_screen_corr(const TemplateSet& ts, const PassState& st, const ScreenTables& tb, uint32_t f0, uint32_t nbatch,
                        float* Bbuf, hipStream_t s)
{
    if (nbatch == 0) return;
    int rp;
    const size_t lds1 = screen_corr_lds(ts, tb.S, &rp);
    const dim3 g(sgrid(nbatch, 1, 1u << 20)), b(kBlock);
    if (ts.L == 48) hipLaunchKernelGGL(k_screen_corr<48>, g, b, lds1, s, ts, st, tb, f0, nbatch, rp, Bbuf, (size_t)ts.U * tb.S);
    else if (ts.L == 64) hipLaunchKernelGGL(k_screen_corr<64>, g, b, lds1, s, ts, st, tb, f0, nbatch, rp, Bbuf, (size_t)ts.U * tb.S);
    else if (ts.L == 32) hipLaunchKernelGGL(k_screen_corr<32>, g, b, lds1, s, ts, st, tb, f0, nbatch, rp, Bbuf, (size_t)ts.U * tb.S);
    else hipLaunchKernelGGL(k_screen_corr<0>, g, b, lds1, s, ts, st, tb, f0, nbatch, rp, Bbuf, (size_t)ts.U * tb.S);
}


// ---------------------------------------------------------------------------
// Segmented launches of the screened recheck: the flagged windows of every
// active segment in one grid.  Entry e covers the segment's flagged windows
// [a0, a0 + n), whose B rows are rows [g0, g0 + n) of Bbuf (bstride floats each).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_fill_eob_scr(const SegLaunch* __restrict__ tab, uint32_t ntab, uint16_t* eob)
{
    for (uint32_t e = blockIdx.x; e < ntab; e += gridDim.x) {
        const uint32_t b0 = tab[e].blk0, nb = tab[e].nblk;
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) eob[b0 + i] = (uint16_t)e;
    }
}

template <int LT>
__global__ __launch_bounds__(kBlock) void k_screen_corr_seg(const SegDesc* __restrict__ sd, const SegLaunch* __restrict__ tab,
                                                            const uint16_t* __restrict__ eob, float* __restrict__ Bbuf,
                                                            size_t bstride)
{
    const SegLaunch e = tab[eob[blockIdx.x]];
    const SegDesc& d = sd[e.seg];
    screen_corr_body<LT>(d.ts, d.st, d.tb, e.a0, e.n, d.rp, Bbuf + (size_t)e.g0 * bstride, bstride, blockIdx.x - e.blk0,
                         e.nblk);
}

template <int LT>
__global__ __launch_bounds__(kSelMaxBlock) void k_screen_select_reg_seg(const SegDesc* __restrict__ sd,
                                                                        const SegLaunch* __restrict__ tab,
                                                                        const uint16_t* __restrict__ eob,
                                                                        const float* __restrict__ Bbuf, size_t bstride)
{
    const SegLaunch e = tab[eob[blockIdx.x]];
    const SegDesc& d = sd[e.seg];
    screen_select_reg_body<LT>(d.ts, d.st, d.tb, e.a0, e.n, Bbuf + (size_t)e.g0 * bstride, bstride,
                               blockIdx.x - e.blk0, e.nblk, nullptr);
}

// Whether the register-resident selection takes this segment, and its block
// size and dynamic LDS.
bool screen_select_reg_plan(const TemplateSet& ts, const ScreenTables& tb, int* bd, size_t* lds)
{
    const int64_t owners = (int64_t)sel_slices(ts.U) * tb.S;
    *lds = screen_select_reg_lds(ts, tb.S, tb.cmax, tb.npl_max);
    *bd = (int)std::max<int64_t>(kBlock, (owners + 63) / 64 * 64);
    return owners <= kSelMaxBlock && *lds <= kSelMaxLds;
}

// corr: entries' nblk blocks each (grid-stride over their windows); select: one
// block per window (nblk >= n; the blocks beyond n return at once), bd threads.
void launch_seg_screen_corr(int L, const SegDesc* sd, const SegLaunch* tab, uint32_t ntab, uint32_t nblocks,
                            uint16_t* eob, size_t lds, float* Bbuf, size_t bstride, hipStream_t s, GraphRec* rec)
{
    if (ntab == 0) return;
    const unsigned ge = std::max<uint32_t>(1, std::min<uint32_t>(ntab, 1024));
    if (rec == nullptr) hipLaunchKernelGGL(k_fill_eob_scr, dim3(ge), dim3(kBlock), 0, s, tab, ntab, eob);
    const uint16_t* m = eob;
    const dim3 gc(nblocks), bc(kBlock);
    if (L == 48) launch_or_record(rec, k_screen_corr_seg<48>, gc, bc, lds, s, sd, tab, m, Bbuf, bstride);
    else if (L == 64) launch_or_record(rec, k_screen_corr_seg<64>, gc, bc, lds, s, sd, tab, m, Bbuf, bstride);
    else if (L == 32) launch_or_record(rec, k_screen_corr_seg<32>, gc, bc, lds, s, sd, tab, m, Bbuf, bstride);
    else launch_or_record(rec, k_screen_corr_seg<0>, gc, bc, lds, s, sd, tab, m, Bbuf, bstride);
}

// The selection of one class of segments: bd threads and lds bytes of dynamic
// LDS (the class's largest) per block.
void launch_seg_select(int L, const SegDesc* sd, const SegLaunch* tab, uint32_t ntab, uint32_t nblocks, uint16_t* eob,
                       int bd, size_t lds, const float* Bbuf, size_t bstride, hipStream_t s, GraphRec* rec)
{
    if (ntab == 0 || nblocks == 0) return;
    const unsigned ge = std::max<uint32_t>(1, std::min<uint32_t>(ntab, 1024));
    if (rec == nullptr) hipLaunchKernelGGL(k_fill_eob_scr, dim3(ge), dim3(kBlock), 0, s, tab, ntab, eob);
    const uint16_t* m = eob;
    const dim3 gs(nblocks), bs(bd);
    if (L == 48) {
        allow_lds(k_screen_select_reg_seg<48>, lds);
        launch_or_record(rec, k_screen_select_reg_seg<48>, gs, bs, lds, s, sd, tab, m, Bbuf, bstride);
    } else if (L == 64) {
        allow_lds(k_screen_select_reg_seg<64>, lds);
        launch_or_record(rec, k_screen_select_reg_seg<64>, gs, bs, lds, s, sd, tab, m, Bbuf, bstride);
    } else if (L == 32) {
        allow_lds(k_screen_select_reg_seg<32>, lds);
        launch_or_record(rec, k_screen_select_reg_seg<32>, gs, bs, lds, s, sd, tab, m, Bbuf, bstride);
    } else {
        allow_lds(k_screen_select_reg_seg<0>, lds);
        launch_or_record(rec, k_screen_select_reg_seg<0>, gs, bs, lds, s, sd, tab, m, Bbuf, bstride);
    }
}

}  // namespace fbp
