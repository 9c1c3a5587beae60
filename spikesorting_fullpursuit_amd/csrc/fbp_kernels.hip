// HIP kernels of the full-binary-pursuit pass loop for gfx950 (CDNA4).
//
// Reference behaviour: SF/kernels/binary_pursuit.cl (SF = spikesorting_fullpursuit/),
// restated with a phase-separated meaning where the reference races.  Every
// floating-point reduction that feeds a decision uses the same operation order
// as the CPU restatement (explicit fmaf, ascending channel then sample), so
// decisions are bit-identical to it.  Built with -ffp-contract=off.
#include <algorithm>

#include "fbp_device.h"

namespace fbp {

// ---------------------------------------------------------------------------
// Likelihood: compute_full_likelihood + compute_maximum_likelihood,
// binary_pursuit.cl:216-300.  One block handles `wb` listed windows; a work
// item is (window, template, group of 8 consecutive positions) and slides a
// register window over the voltage row staged in LDS.
// ---------------------------------------------------------------------------

constexpr int kLikMaxWb = 16;  // windows per likelihood block at most (pair entries keep the window in 4 bits)
static_assert(kLikMaxWb <= 16, "pair entries t * 16 + b");

template <int LT>
__device__ __forceinline__ void likelihood_body(TemplateSet ts, const float* __restrict__ V, int64_t N,
                                                       const uint32_t* __restrict__ W, uint32_t nW, int wb, int sv,
                                                       float* __restrict__ LIK, float* __restrict__ wmax_v,
                                                       uint16_t* __restrict__ wmax_p,
                                                       const uint64_t* __restrict__ dirty,
                                                       uint32_t* __restrict__ cnt_pairs,
                                                       const uint8_t* __restrict__ gmask, uint32_t bx, uint32_t gx)
{
    // [wb][C][sv] voltage rows, then per-item maxima [wb][U][ngr] (value, position),
    // then the block's (window b, template t) pairs to compute [wb][U] (entries
    // t * 16 + b), then the channels each window stages [wb][C]
    extern __shared__ __align__(16) float sV[];
    __shared__ uint32_t npair_s;
    __shared__ unsigned long long cun_s[kLikMaxWb];
    __shared__ int ncl_s[kLikMaxWb];
    const int C = ts.C, U = ts.U;
    const int L = (LT > 0) ? LT : ts.L;
    const int ngr = (L + kPos - 1) / kPos;
    const int per_win = U * ngr;
    const uint32_t nbatch = (nW + wb - 1) / wb;
    float* pv = sV + wb * C * sv;
    uint32_t* pp = reinterpret_cast<uint32_t*>(pv + wb * per_win);
    uint32_t* plist = pp + wb * per_win;
    int* clist = reinterpret_cast<int*>(plist + wb * U);
    for (uint32_t bt = bx; bt < nbatch; bt += gx) {
        const uint32_t w0 = bt * wb;
        const int nb = (int)min((uint32_t)wb, nW - w0);
        __syncthreads();
        if (threadIdx.x == 0) npair_s = 0;
        if (threadIdx.x < kLikMaxWb) cun_s[threadIdx.x] = 0ull;
        __syncthreads();
        // pairs whose channels changed since the window was last computed
        // (every pair without a dirty map); the others keep LIK and maxima
        // gmask (exact pass after a matrix-core approximation): only the pairs
        // with 8-position groups to rescore, and only those groups
        // (b, t) of pr = b * U + t advanced by the block size: one division per thread
        const int bd = blockDim.x, db = bd / U, dt = bd - db * U;
        int rb = threadIdx.x / U, rt = threadIdx.x - (int)(threadIdx.x / U) * U;
        for (int pr = threadIdx.x; pr < nb * U; pr += bd) {
            const int bb = rb, tt = rt;
            rt += dt;
            rb += db;
            if (rt >= U) {
                rt -= U;
                ++rb;
            }
            if (gmask != nullptr && gmask[(size_t)W[w0 + bb] * U + tt] == 0) continue;
            const uint32_t ent = (uint32_t)tt * 16u + (uint32_t)bb;
            if (dirty == nullptr) {
                plist[atomicAdd(&npair_s, 1u)] = ent;
                if (gmask != nullptr && C <= 64) atomicOr(&cun_s[bb], (unsigned long long)ts.cmask[tt]);
            } else {
                const uint64_t m = ts.cmask[tt];
                if ((m & dirty[W[w0 + bb]]) != 0ull) {
                    plist[atomicAdd(&npair_s, 1u)] = ent;
                    atomicOr(&cun_s[bb], (unsigned long long)m);
                }
            }
        }
        __syncthreads();
        const int npair = (int)npair_s;
        if (npair == 0) continue;  // uniform over the block
        if (threadIdx.x == 0 && cnt_pairs != nullptr) atomicAdd(cnt_pairs, (uint32_t)npair);
        // channels to stage per window: those of its computed templates; with C a
        // multiple of 4, the channel quads that hold any of them (clist then lists quads)
        const bool quads = C % 4 == 0;
        if ((int)threadIdx.x < nb) {
            const int b = threadIdx.x;
            int n = 0;
            // channels of the block's computed templates (every channel without a mask)
            const bool all = dirty == nullptr && (gmask == nullptr || C > 64);
            const uint64_t m = all ? ~0ull : cun_s[b];
            if (quads) {
                for (int q = 0; q < C / 4; q++)
                    if (all || ((m >> (4 * q)) & 0xFull) != 0ull) clist[b * C + n++] = q;
            } else {
                for (int c = 0; c < C; c++)
                    if (all || ((m >> c) & 1ull) != 0ull) clist[b * C + n++] = c;
            }
            ncl_s[b] = n;
        }
        __syncthreads();
        // V is sample-major ([N][C]): read along the staged channels (float4 per quad),
        // store channel rows; several loads in flight per thread
        for (int b = 0; b < nb; b++) {
            const int ncl = ncl_s[b];
            if (ncl == 0) continue;  // uniform over the block
            const int tot = ncl * sv;
            const int64_t xb = (int64_t)W[w0 + b] * L;
            // element e = j * ncl + qi (sample j, listed quad or channel qi), advanced by the
            // block size one step at a time: two divisions per window, none per element
            const int dj = bd / ncl, dq = bd - dj * ncl;
            int j = threadIdx.x / ncl, qi = threadIdx.x - (int)(threadIdx.x / ncl) * ncl;
            auto step = [&]() {
                qi += dq;
                j += dj;
                if (qi >= ncl) {
                    qi -= ncl;
                    ++j;
                }
            };
            if (quads) {
                for (int e0 = threadIdx.x; e0 < tot; e0 += 4 * bd) {
                    float4 v[4];
                    int c0[4], jj[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int e = e0 + q * bd;
                        v[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        c0[q] = -1;
                        jj[q] = 0;
                        if (e < tot) {
                            const int64_t x = xb + j;
                            c0[q] = 4 * clist[b * C + qi];
                            jj[q] = j;
                            if (x < N) v[q] = *reinterpret_cast<const float4*>(V + (size_t)x * C + c0[q]);
                        }
                        step();
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        if (c0[q] < 0) continue;
                        float* dst = sV + (size_t)(b * C + c0[q]) * sv + jj[q];
                        dst[0] = v[q].x;
                        dst[sv] = v[q].y;
                        dst[2 * sv] = v[q].z;
                        dst[3 * sv] = v[q].w;
                    }
                }
                continue;
            }
            for (int e0 = threadIdx.x; e0 < tot; e0 += 8 * bd) {
                float v[8];
                int dst[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int e = e0 + q * bd;
                    v[q] = 0.0f;
                    dst[q] = -1;
                    if (e < tot) {
                        const int c = clist[b * C + qi];
                        const int64_t x = xb + j;
                        dst[q] = (b * C + c) * sv + j;
                        if (x < N) v[q] = V[(size_t)x * C + c];
                    }
                    step();
                }
#pragma unroll
                for (int q = 0; q < 8; q++)
                    if (dst[q] >= 0) sV[dst[q]] = v[q];
            }
        }
        __syncthreads();
        for (int it = threadIdx.x; it < npair * ngr; it += blockDim.x) {
            const int j = it / ngr, g = it - j * ngr;
            const uint32_t pr = plist[j];
            const int b = (int)(pr & 15u), t = (int)(pr >> 4);
            if (gmask != nullptr && ((gmask[(size_t)W[w0 + b] * U + t] >> g) & 1) == 0) {
                pv[it] = -INFINITY;  // its approximations are negative: no window maximum there
                pp[it] = 0xffffu;
                continue;
            }
            const int p0 = g * kPos;
            const int64_t x0 = (int64_t)W[w0 + b] * L + p0;
            float acc[kPos];
            const float bt0 = ts.bias[t];
#pragma unroll
            for (int q = 0; q < kPos; q++) acc[q] = bt0;
            const int r0 = ts.off[t], nr = ts.nch[t];
            for (int k = 0; k < nr; k++) {
                const int c = ts.chan[r0 + k];
                const float* trow = ts.rows + (size_t)(r0 + k) * ts.Lp;
                const float* vrow = sV + (b * C + c) * sv + p0;  // 16-byte aligned: sv, p0 multiples of 4
                if (LT > 0) lik_row_fixed_a4<(LT > 0 ? LT : 4)>(trow, vrow, acc);
                else lik_row_generic(trow, vrow, L, acc);
            }
            float* out = LIK + (size_t)t * N;
            float mv = -INFINITY;
            uint32_t mp = 0xffffu;
#pragma unroll
            for (int q = 0; q < kPos; q++) {
                const int64_t x = x0 + q;
                if (p0 + q < L && x < N) {
                    const float v = (x + L >= N) ? 0.0f : acc[q];
                    out[x] = v;
                    if (v > mv) { mv = v; mp = (uint32_t)(p0 + q); }
                }
            }
            pv[it] = mv;
            pp[it] = mp;
        }
        __syncthreads();
        // per computed (window, template): first position of the window maximum
        for (int j = threadIdx.x; j < npair; j += blockDim.x) {
            const uint32_t pr = plist[j];
            const int b = (int)(pr & 15u), t = (int)(pr >> 4);
            float mv = -INFINITY;
            uint32_t mp = 0xffffu;
            for (int g = 0; g < ngr; g++) {
                const int it = j * ngr + g;
                if (pv[it] > mv) { mv = pv[it]; mp = pp[it]; }
            }
            const size_t o = (size_t)W[w0 + b] * U + t;
            wmax_v[o] = mv;
            wmax_p[o] = (uint16_t)mp;
        }
    }
}

template <int LT>
__global__ __launch_bounds__(kBlock) void k_likelihood(TemplateSet ts, const float* __restrict__ V, int64_t N,
                                                       const uint32_t* __restrict__ W, uint32_t nW, int wb, int sv,
                                                       float* __restrict__ LIK, float* __restrict__ wmax_v,
                                                       uint16_t* __restrict__ wmax_p,
                                                       const uint64_t* __restrict__ dirty,
                                                       uint32_t* __restrict__ cnt_pairs,
                                                       const uint8_t* __restrict__ gmask)
{
    likelihood_body<LT>(ts, V, N, W, nW, wb, sv, LIK, wmax_v, wmax_p, dirty, cnt_pairs, gmask, blockIdx.x, gridDim.x);
}

// Marks, for every spike subtracted this pass, the windows whose likelihood
// positions read any of its samples with the spike's channels: window w
// computes positions [wL, wL + L), which read samples [wL, wL + 2L - 1), and
// that range meets the spike's [p, p + L) exactly for
// w in [ceil((p - 2L + 2) / L), floor((p + L - 1) / L)].  Integer OR is order
// independent.
__device__ __forceinline__ void mark_dirty_body(TemplateSet ts, PassState st, const uint32_t* __restrict__ sidx,
                                                       const uint32_t* __restrict__ slab, uint32_t nsp, uint32_t bx, uint32_t gx)
{
    const int64_t L = ts.L;
    for (uint32_t k = bx * blockDim.x + threadIdx.x; k < nsp; k += gx * blockDim.x) {
        const uint32_t u = slab[k];
        if (u >= (uint32_t)ts.U) continue;
        const int64_t p = sidx[k];
        const uint64_t m = ts.cmask[u];
        const int64_t lo = p + 1 >= L ? (p + 1 - L) / L : 0, hi = min(st.nwin - 1, (p + L - 1) / L);
        for (int64_t v = lo; v <= hi; v++) atomicOr((unsigned long long*)&st.dirty[v], (unsigned long long)m);
    }
}

__global__ __launch_bounds__(kBlock) void k_mark_dirty(TemplateSet ts, PassState st, const uint32_t* __restrict__ sidx,
                                                       const uint32_t* __restrict__ slab, uint32_t nsp)
{
    mark_dirty_body(ts, st, sidx, slab, nsp, blockIdx.x, gridDim.x);
}

// The listed windows' likelihoods are current: clear their dirty channels.
__device__ __forceinline__ void clear_dirty_body(PassState st, uint32_t nW, uint32_t bx, uint32_t gx)
{
    for (uint32_t i = bx * blockDim.x + threadIdx.x; i < nW; i += gx * blockDim.x)
        st.dirty[st.W[i]] = 0ull;
}

__global__ __launch_bounds__(kBlock) void k_clear_dirty(PassState st, uint32_t nW)
{
    clear_dirty_body(st, nW, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Window maximum: compute_template_maximum_likelihood launched for every
// template in order, binary_pursuit.cl:350-441.  One wave per listed window;
// lexicographic arg-max (value, then template, then position).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void window_max_body(int U, int L, int64_t N, const float* __restrict__ wmax_v,
                                                       const uint16_t* __restrict__ wmax_p,
                                                       const uint32_t* __restrict__ W, uint32_t nW, PassState st, uint32_t bx, uint32_t gx)
{
    const uint32_t wpb = kBlock / 64;
    for (uint32_t w = bx * wpb + (threadIdx.x >> 6); w < nW; w += gx * wpb) {
        const int64_t id = W[w];
        const int64_t sw = id * L, ew = (id + 1) * L;
        if (ew >= N - L) continue;  // binary_pursuit.cl:380-383
        // the scanned range [max(0, sw-L), min(N, ew+L)) is windows id-1 .. id+1
        const int64_t wlo = id > 0 ? id - 1 : 0;
        const int64_t whi = min(st.nwin - 1, id + 1);
        const int64_t start = wlo * L;
        const uint32_t span = (uint32_t)((whi + 1) * L - start);
        float best = 0.0f;
        uint32_t key = kNoItem;  // t * span + (x - start)
        bool cw = false;
        for (int t = lane_id(); t < U; t += 64) {
            for (int64_t wd = wlo; wd <= whi; wd++) {
                const float v = wmax_v[(size_t)wd * U + t];
                if (v > best) {
                    best = v;
                    key = (uint32_t)t * span + (uint32_t)(wd * L - start) + wmax_p[(size_t)wd * U + t];
                }
                cw |= (wd == id) && (v > 0.0f);
            }
        }
        wave_argmax(best, key);
        const bool any_cw = __ballot(cw) != 0ull;
        if (lane_id() == 0) {
            uint32_t lab = st.best_lab[id], idx = st.best_idx[id];
            float bl = 0.0f;
            if (key != kNoItem) {
                bl = best;
                lab = key / span;
                idx = (uint32_t)(start + (key - lab * span));
            }
            if (any_cw) st.check[id] = 1;
            st.orc[id] = (bl > 0.0f && (int64_t)idx >= sw && (int64_t)idx < ew) ? 1 : 0;
            st.best_lik[id] = bl;
            st.best_lab[id] = lab;
            st.best_idx[id] = idx;
            st.obest_idx[id] = idx;
            st.obest_lab[id] = lab;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_window_max(int U, int L, int64_t N, const float* __restrict__ wmax_v,
                                                       const uint16_t* __restrict__ wmax_p,
                                                       const uint32_t* __restrict__ W, uint32_t nW, PassState st)
{
    window_max_body(U, L, N, wmax_v, wmax_p, W, nW, st, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Ordered stream compaction (three kernels: count, scan, scatter).  Output
// order equals input order, so lists stay sorted by window.
// ---------------------------------------------------------------------------
struct SelFlagged {  // F = [W[i] : orc[W[i]]]
    const uint32_t* W; const uint8_t* orc; uint32_t* F;
    __device__ bool pred(uint32_t e) const { return orc[W[e]] != 0; }
    __device__ void emit(uint32_t e, uint32_t pos) const { F[pos] = W[e]; }
    __device__ void always(uint32_t) const {}
};
struct SelSpikes {   // spikes of this pass, appended at `base`
    const uint8_t* has; const uint32_t* ci; const uint32_t* cl; uint32_t* oi; uint32_t* ol; uint64_t base;
    __device__ bool pred(uint32_t e) const { return has[e] != 0; }
    __device__ void emit(uint32_t e, uint32_t pos) const { oi[base + pos] = ci[e]; ol[base + pos] = cl[e]; }
    __device__ void always(uint32_t) const {}
};
struct SelNextWindows {  // W' = nonzero(check); resets check and overlap flags
    uint8_t* check; uint8_t* orc; uint32_t* W;
    __device__ bool pred(uint32_t e) const { return check[e] != 0; }
    __device__ void emit(uint32_t e, uint32_t pos) const { W[pos] = e; }
    __device__ void always(uint32_t e) const { check[e] = 0; orc[e] = 0; }
};

template <class Sel>
__device__ __forceinline__ void compact_count_body(Sel sel, uint32_t n, uint32_t* blockcnt, uint32_t bx, uint32_t gx)
{
    __shared__ uint32_t lds[kBlock / 64];
    const uint32_t base = bx * kCompactTile;
    uint32_t tot_all = 0;
    for (int j = 0; j < kCompactTile / kBlock; j++) {
        const uint32_t e = base + j * kBlock + threadIdx.x;
        const bool f = e < n && sel.pred(e);
        uint32_t tot;
        block_excl_scan_flag(f, lds, tot);
        tot_all += tot;
    }
    if (threadIdx.x == 0) blockcnt[bx] = tot_all;
}

template <class Sel>
__global__ __launch_bounds__(kBlock) void k_compact_count(Sel sel, uint32_t n, uint32_t* blockcnt)
{
    compact_count_body<Sel>(sel, n, blockcnt, blockIdx.x, gridDim.x);
}

// Exclusive scan of nb block counts in place; total into *out_total.
__device__ __forceinline__ void compact_scan_body(uint32_t* blockcnt, uint32_t nb, uint32_t* out_total)
{
    __shared__ uint32_t part[1024];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = (i < nb) ? blockcnt[i] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t a = (threadIdx.x >= (uint32_t)off) ? part[threadIdx.x - off] : 0u;
            __syncthreads();
            part[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nb) blockcnt[i] = carry + part[threadIdx.x] - v;
        const uint32_t chunk_total = part[1023];
        __syncthreads();
        carry += chunk_total;
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(1024) void k_compact_scan(uint32_t* blockcnt, uint32_t nb, uint32_t* out_total)
{
    compact_scan_body(blockcnt, nb, out_total);
}

template <class Sel>
__device__ __forceinline__ void compact_scatter_body(Sel sel, uint32_t n, const uint32_t* blockoff, uint32_t bx, uint32_t gx)
{
    __shared__ uint32_t lds[kBlock / 64];
    const uint32_t base = bx * kCompactTile;
    uint32_t run = blockoff[bx];
    for (int j = 0; j < kCompactTile / kBlock; j++) {
        const uint32_t e = base + j * kBlock + threadIdx.x;
        const bool inb = e < n;
        const bool f = inb && sel.pred(e);
        uint32_t tot;
        const uint32_t pos = block_excl_scan_flag(f, lds, tot);
        if (f) sel.emit(e, run + pos);
        if (inb) sel.always(e);
        run += tot;
    }
}

template <class Sel>
__global__ __launch_bounds__(kBlock) void k_compact_scatter(Sel sel, uint32_t n, const uint32_t* blockoff)
{
    compact_scatter_body<Sel>(sel, n, blockoff, blockIdx.x, gridDim.x);
}

template <class Sel>
static void compact(Sel sel, uint32_t n, uint32_t* blockbuf, uint32_t* out_total, hipStream_t s)
{
    const uint32_t nb = (n + kCompactTile - 1) / kCompactTile;
    if (nb == 0) {
        hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(1024), 0, s, blockbuf, 0u, out_total);
        return;
    }
    hipLaunchKernelGGL(k_compact_count<Sel>, dim3(nb), dim3(kBlock), 0, s, sel, n, blockbuf);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(1024), 0, s, blockbuf, nb, out_total);
    hipLaunchKernelGGL(k_compact_scatter<Sel>, dim3(nb), dim3(kBlock), 0, s, sel, n, blockbuf);
}

// ---------------------------------------------------------------------------
// Overlap recheck, literal form: overlap_recheck_indices,
// binary_pursuit.cl:444-654.  A block evaluates 256*kItemsPerThread items of
// one flagged window and writes its arg-max (lowest item id on ties).
// ---------------------------------------------------------------------------
constexpr int kItemsPerThread = 4;

__global__ __launch_bounds__(kBlock) void k_recheck_literal(TemplateSet ts, PassState st, int64_t n_max,
                                                            uint32_t nchunk, int use_lds,
                                                            float* __restrict__ part_v, uint32_t* __restrict__ part_k)
{
    extern __shared__ float sW[];  // [C][L] voltage of the window
    __shared__ float red_v[kBlock / 64];
    __shared__ uint32_t red_k[kBlock / 64];
    const uint32_t nF = st.counters[CNT_F];
    const int64_t N = st.N;
    const int L = ts.L, C = ts.C;
    const int64_t U = ts.U, S = 2 * n_max + 1;
    const uint64_t items = (uint64_t)S * S * U;
    const uint64_t total_blocks = (uint64_t)nF * nchunk;
    for (uint64_t blk = blockIdx.x; blk < total_blocks; blk += gridDim.x) {
        const uint32_t w = (uint32_t)(blk / nchunk);
        const uint32_t ch = (uint32_t)(blk - (uint64_t)w * nchunk);
        const uint32_t id = st.F[w];
        const uint32_t lab = st.best_lab[id];
        const int64_t b = st.best_idx[id];
        const float lb = st.LIK[(size_t)lab * N + b];
        const bool wskip = (lb < ts.thr[lab]) || (b >= N - L);
        const float* vw;
        int cstride, jstride;
        __syncthreads();
        if (use_lds) {
            if (!wskip)
                for (int e = threadIdx.x; e < C * L; e += blockDim.x) {
                    const int j = e / C, c = e - j * C;
                    sW[c * L + j] = st.V[(size_t)(b + j) * C + c];
                }
            __syncthreads();
            vw = sW; cstride = L; jstride = 1;
        } else {
            vw = st.V + (size_t)b * C; cstride = 1; jstride = C;
        }
        float bv = 0.0f;
        uint32_t bk = kNoItem;
        if (!wskip) {
            const uint64_t k0 = (uint64_t)ch * kBlock * kItemsPerThread;
            for (int r = 0; r < kItemsPerThread; r++) {
                const uint64_t k = k0 + (uint64_t)r * kBlock + threadIdx.x;
                if (k >= items) break;
                const uint32_t u = (uint32_t)(k % U);
                const int64_t f = (int64_t)((k / U) % S), s = (int64_t)(k / (S * U));
                if (recheck_item_skipped(st.LIK, N, L, lab, u, b, f, s, n_max)) continue;
                const float v = recheck_item_value(ts, vw, cstride, jstride, lab, u, f, s, n_max);
                argmax_merge(bv, bk, v, (uint32_t)k);
            }
        }
        wave_argmax(bv, bk);
        if (lane_id() == 0) { red_v[threadIdx.x >> 6] = bv; red_k[threadIdx.x >> 6] = bk; }
        __syncthreads();
        if (threadIdx.x == 0) {
            float v = red_v[0];
            uint32_t k = red_k[0];
            for (int i = 1; i < kBlock / 64; i++) argmax_merge(v, k, red_v[i], red_k[i]);
            part_v[blk] = v;
            part_k[blk] = k;
        }
    }
}

// parse_overlap_recheck_indices, binary_pursuit.cl:657-815: one wave per
// flagged window reduces its partial maxima and applies the accept cascade.
__global__ __launch_bounds__(kBlock) void k_recheck_parse(TemplateSet ts, PassState st, int64_t n_max, uint32_t nchunk,
                                                          const float* __restrict__ part_v,
                                                          const uint32_t* __restrict__ part_k)
{
    const uint32_t nF = st.counters[CNT_F];
    const uint32_t wpb = kBlock / 64;
    const int64_t N = st.N;
    const int64_t U = ts.U, S = 2 * n_max + 1;
    for (uint32_t w = blockIdx.x * wpb + (threadIdx.x >> 6); w < nF; w += gridDim.x * wpb) {
        float v = 0.0f;
        uint32_t k = kNoItem;
        for (uint32_t c = lane_id(); c < nchunk; c += 64)
            argmax_merge(v, k, part_v[(size_t)w * nchunk + c], part_k[(size_t)w * nchunk + c]);
        wave_argmax(v, k);
        if (lane_id() != 0) continue;
        apply_parse(ts, st, n_max, st.F[w], v, k);
    }
}

// check_overlap_reassignments, binary_pursuit.cl:822-875.  orc is only read
// here (the post-parse snapshot); giving up is recorded in `giveup`.
__device__ __forceinline__ void reassign_body(int L, PassState st, uint32_t bx, uint32_t gx)
{
    const uint32_t nF = st.counters[CNT_F];
    const int64_t nwf = st.N / L;
    for (uint32_t w = bx * blockDim.x + threadIdx.x; w < nF; w += gx * blockDim.x) {
        const int64_t id = st.F[w];
        uint8_t gu = 0;
        if (st.orc[id] != 0) {
            const int64_t sw = id * L, ew = (id + 1) * L;
            const uint32_t oi = st.obest_idx[id];
            st.best_idx[id] = oi;
            st.best_lab[id] = st.obest_lab[id];
            if ((int64_t)oi < sw && id > 1) {
                if (st.orc[id - 2] == 1) {
                    gu = 1;
                    st.best_idx[id] = 0;
                    st.check[id - 1] = 2;
                    st.check[id] = 2;
                    st.check[id + 1] = 2;
                } else {
                    st.check[id - 2] = 2;
                }
            }
            if ((int64_t)oi >= ew && id + 2 < nwf) st.check[id + 2] = 2;
        }
        st.giveup[id] = gu;
    }
}

__global__ __launch_bounds__(kBlock) void k_reassign(int L, PassState st)
{
    reassign_body(L, st, blockIdx.x, gridDim.x);
}

// binary_pursuit kernel phase A (binary_pursuit.cl:960-1004): add decisions
// and the "=3" flags.  `use_giveup` is 0 when no recheck ran this pass.
__device__ __forceinline__ void pursuit_decide_body(int L, const float* __restrict__ thr, PassState st,
                                                           uint32_t nW, int use_giveup, uint32_t bx, uint32_t gx)
{
    const int64_t N = st.N, nwf = N / L;
    for (uint32_t w = bx * blockDim.x + threadIdx.x; w < nW; w += gx * blockDim.x) {
        const int64_t id = st.W[w];
        const int64_t sw = id * L, ew = (id + 1) * L;
        float ml = 0.0f;
        uint32_t lab = 0, idx = 0;
        if (ew < N - L) { ml = st.best_lik[id]; lab = st.best_lab[id]; idx = st.best_idx[id]; }
        const bool orc = st.orc[id] == 1 && !(use_giveup && st.giveup[id]);
        uint8_t h = 0;
        if ((ml > thr[lab]) || orc) {
            if (((int64_t)idx >= sw && (int64_t)idx < ew) || orc) {
                h = 1;
                st.check[id] = 3;
                if (id > 0) st.check[id - 1] = 3;
                if (id + 1 < nwf) st.check[id + 1] = 3;
            }
        }
        st.has[w] = h;
        st.cand_idx[w] = idx;
        st.cand_lab[w] = lab;
    }
}

__global__ __launch_bounds__(kBlock) void k_pursuit_decide(int L, const float* __restrict__ thr, PassState st,
                                                           uint32_t nW, int use_giveup)
{
    pursuit_decide_body(L, thr, st, nW, use_giveup, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void copy_bytes_body(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n, uint32_t bx, uint32_t gx)
{
    for (int64_t i = bx * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gx * blockDim.x)
        dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_copy_bytes(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n)
{
    copy_bytes_body(src, dst, n, blockIdx.x, gridDim.x);
}

// binary_pursuit kernel phase B (binary_pursuit.cl:1009-1045): pruning of
// windows that added nothing, reading neighbour flags from the snapshot.
__device__ __forceinline__ void pursuit_prune_body(int L, const float* __restrict__ thr, PassState st, uint32_t nW, uint32_t bx, uint32_t gx)
{
    const int64_t N = st.N, nwf = N / L;
    for (uint32_t w = bx * blockDim.x + threadIdx.x; w < nW; w += gx * blockDim.x) {
        const int64_t id = st.W[w];
        const int64_t sw = id * L, ew = (id + 1) * L;
        if (!(ew < N - L) || st.has[w]) continue;
        const float ml = st.best_lik[id];
        const uint32_t lab = st.best_lab[id];
        const int64_t idx = st.best_idx[id];
        uint8_t own = st.check_snap[id];
        if (idx < sw && own == 1 && id > 0 && st.check_snap[id - 1] == 0) own = 0;
        if (idx >= ew && own == 1 && id + 1 < nwf && st.check_snap[id + 1] == 0) own = 0;
        if (ml <= thr[lab] && own < 3) own = 0;
        st.check[id] = own;
    }
}

__global__ __launch_bounds__(kBlock) void k_pursuit_prune(int L, const float* __restrict__ thr, PassState st, uint32_t nW)
{
    pursuit_prune_body(L, thr, st, nW, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Residual: compute_residual, binary_pursuit.cl:57-102, without atomics.
// A spike changes only the channels its template holds (a channel that is
// zero throughout subtracts zeros).  Every (sample, channel) covered by
// several spikes of the pass whose templates hold that channel is written
// once, by the first of them in list order, which subtracts all covering
// templates in list order.  Covering spikes lie within `reach` list entries.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void residual_body(TemplateSet ts, float* __restrict__ V, int64_t N,
                                                     const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ slab,
                                                     uint32_t nsp, int reach, uint32_t bx, uint32_t gx)
{
    const int C = ts.C, L = ts.L;
    for (uint32_t k = bx; k < nsp; k += gx) {
        const int64_t pk = sidx[k];
        const uint32_t lk = slab[k];
        if (lk >= (uint32_t)ts.U) continue;
        const uint32_t jlo = (k > (uint32_t)reach) ? k - reach : 0u;
        const uint32_t jhi = min(nsp - 1, k + (uint32_t)reach);
        const int nr = ts.nch[lk], r0 = ts.off[lk];
        // one thread per (sample, channel of the template), channels fastest:
        // V is sample-major
        for (int e = threadIdx.x; e < L * nr; e += blockDim.x) {
            const int i = e / nr, r = e - i * nr;
            const int c = ts.chan[r0 + r];
            const int64_t x = pk + i;
            if (x >= N) continue;
            bool owner = true;
            for (uint32_t j = jlo; j < k; j++) {
                const int64_t pj = sidx[j];
                const uint32_t lj = slab[j];
                if (lj < (uint32_t)ts.U && x >= pj && x < pj + L && ts.nz[(size_t)lj * C + c]) {
                    owner = false;
                    break;
                }
            }
            if (!owner) continue;
            float v = V[(size_t)x * C + c];
            for (uint32_t j = k; j <= jhi; j++) {
                const int64_t pj = sidx[j];
                const uint32_t lj = slab[j];
                if (lj >= (uint32_t)ts.U || x < pj || x >= pj + L) continue;
                v = v - ts.T[((size_t)lj * C + c) * L + (x - pj)];
            }
            V[(size_t)x * C + c] = v;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_residual(TemplateSet ts, float* __restrict__ V, int64_t N,
                                                     const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ slab,
                                                     uint32_t nsp, int reach)
{
    residual_body(ts, V, N, sidx, slab, nsp, reach, blockIdx.x, gridDim.x);
}

// get_adjusted_clips, binary_pursuit.cl:1088-1130.
__global__ __launch_bounds__(kBlock) void k_adjusted_clips(TemplateSet ts, const float* __restrict__ V, int64_t N,
                                                           const uint32_t* __restrict__ sidx,
                                                           const uint32_t* __restrict__ slab, uint64_t nsp,
                                                           float* __restrict__ clips)
{
    const int C = ts.C, L = ts.L;
    const uint64_t per = (uint64_t)C * L;
    const uint64_t total = nsp * per;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = e / per;
        const int rem = (int)(e - k * per);
        const int c = rem / L, i = rem - c * L;
        const int64_t x = (int64_t)sidx[k] + i;
        float out = 0.0f;
        if (slab[k] < (uint32_t)ts.U && x < N)
            out = V[(size_t)x * C + c] + ts.T[((size_t)slab[k] * C + c) * L + i];
        clips[e] = out;
    }
}

// Zero fill (used instead of hipMemsetAsync so the sequence stays capturable).
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, uint32_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ __launch_bounds__(kBlock) void k_fill_u8(uint8_t* p, uint8_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (uint32_t)i;
}
// dst[x][c] = src[c][x] (row stride ld), 64 x 64 tiles through LDS.
__global__ __launch_bounds__(kBlock) void k_to_sample_major(const float* __restrict__ src, int64_t ld,
                                                            float* __restrict__ dst, int64_t C, int64_t N)
{
    __shared__ float tile[64][65];
    const int64_t tx = (N + 63) / 64, tc = (C + 63) / 64;
    for (int64_t tI = blockIdx.x; tI < tx * tc; tI += gridDim.x) {
        const int64_t x0 = (tI % tx) * 64, c0 = (tI / tx) * 64;
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
            const int r = e / 64, q = e - r * 64;  // r: channel, q: sample
            const int64_t c = c0 + r, x = x0 + q;
            tile[r][q] = (c < C && x < N) ? src[c * ld + x] : 0.0f;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
            const int q = e / 64, r = e - q * 64;
            const int64_t c = c0 + r, x = x0 + q;
            if (c < C && x < N) dst[x * C + c] = tile[r][q];
        }
    }
}

// dst[c][x] = src[x][c].
__global__ __launch_bounds__(kBlock) void k_to_channel_major(const float* __restrict__ src, float* __restrict__ dst,
                                                             int64_t C, int64_t N)
{
    __shared__ float tile[64][65];
    const int64_t tx = (N + 63) / 64, tc = (C + 63) / 64;
    for (int64_t tI = blockIdx.x; tI < tx * tc; tI += gridDim.x) {
        const int64_t x0 = (tI % tx) * 64, c0 = (tI / tx) * 64;
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
            const int q = e / 64, r = e - q * 64;
            const int64_t c = c0 + r, x = x0 + q;
            tile[r][q] = (c < C && x < N) ? src[x * C + c] : 0.0f;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
            const int r = e / 64, q = e - r * 64;
            const int64_t c = c0 + r, x = x0 + q;
            if (c < C && x < N) dst[c * N + x] = tile[r][q];
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int64_t per_block, unsigned cap = 65535u * 16u)
{
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, p, v, n);
}
void launch_fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, p, v, n);
}
void launch_iota(uint32_t* p, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_iota, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, p, n);
}
void launch_to_sample_major(const float* src, int64_t ld, float* dst, int64_t C, int64_t N, hipStream_t s)
{
    const int64_t tiles = ((N + 63) / 64) * ((C + 63) / 64);
    hipLaunchKernelGGL(k_to_sample_major, dim3((unsigned)std::min<int64_t>(tiles, 1 << 20)), dim3(kBlock), 0, s, src,
                       ld, dst, C, N);
}
void launch_to_channel_major(const float* src, float* dst, int64_t C, int64_t N, hipStream_t s)
{
    const int64_t tiles = ((N + 63) / 64) * ((C + 63) / 64);
    hipLaunchKernelGGL(k_to_channel_major, dim3((unsigned)std::min<int64_t>(tiles, 1 << 20)), dim3(kBlock), 0, s, src,
                       dst, C, N);
}

// Windows per likelihood block and LDS row length.
static int64_t likelihood_lds_per_window(const TemplateSet& ts, int sv)
{
    const int ngr = (ts.L + kPos - 1) / kPos;
    return (int64_t)ts.C * sv * 4 + (int64_t)ts.U * ngr * 8 + (int64_t)ts.U * 4 + (int64_t)ts.C * 4;
}

void likelihood_plan(const TemplateSet& ts, int* wb, int* sv)
{
    const int L = ts.L;
    *sv = (2 * L + kPos + 1 + 3) & ~3;  // a multiple of 4: rows are read as float4
    const int ngr = (L + kPos - 1) / kPos;
    const int per_win = ts.U * ngr;
    int w = (2 * kBlock + per_win - 1) / per_win;
    if (w < 1) w = 1;
    if (w > kLikMaxWb) w = kLikMaxWb;
    const int64_t lds_per = likelihood_lds_per_window(ts, *sv);
    while (w > 1 && w * lds_per > 48 * 1024) w--;
    *wb = w;
}

bool likelihood_fits(const TemplateSet& ts)
{
    int wb, sv;
    likelihood_plan(ts, &wb, &sv);
    return wb * likelihood_lds_per_window(ts, sv) <= 64 * 1024;
}

void launch_likelihood(const TemplateSet& ts, const PassState& st, uint32_t nW, const uint64_t* dirty, hipStream_t s,
                       const uint8_t* gmask)
{
    if (nW == 0) return;
    int wb, sv;
    likelihood_plan(ts, &wb, &sv);
    const size_t lds = (size_t)wb * likelihood_lds_per_window(ts, sv);
    const uint32_t nbatch = (nW + wb - 1) / wb;
    const dim3 g(grid_for(nbatch, 1, 1u << 20)), b(kBlock);
    const float* V = st.V;
    const int64_t N = st.N;
    const uint32_t* W = st.W;
    if (ts.L == 48)
        hipLaunchKernelGGL(k_likelihood<48>, g, b, lds, s, ts, V, N, W, nW, wb, sv, st.LIK, st.wmax_v, st.wmax_p,
                           dirty, gmask ? nullptr : st.counters + CNT_PAIRS, gmask);
    else if (ts.L == 64)
        hipLaunchKernelGGL(k_likelihood<64>, g, b, lds, s, ts, V, N, W, nW, wb, sv, st.LIK, st.wmax_v, st.wmax_p,
                           dirty, gmask ? nullptr : st.counters + CNT_PAIRS, gmask);
    else if (ts.L == 32)
        hipLaunchKernelGGL(k_likelihood<32>, g, b, lds, s, ts, V, N, W, nW, wb, sv, st.LIK, st.wmax_v, st.wmax_p,
                           dirty, gmask ? nullptr : st.counters + CNT_PAIRS, gmask);
    else
        hipLaunchKernelGGL(k_likelihood<0>, g, b, lds, s, ts, V, N, W, nW, wb, sv, st.LIK, st.wmax_v, st.wmax_p,
                           dirty, gmask ? nullptr : st.counters + CNT_PAIRS, gmask);
}

void launch_mark_dirty(const TemplateSet& ts, const PassState& st, const uint32_t* sidx, const uint32_t* slab,
                       uint32_t nsp, hipStream_t s)
{
    if (nsp == 0 || st.dirty == nullptr) return;
    hipLaunchKernelGGL(k_mark_dirty, dim3(grid_for(nsp, kBlock, 1u << 16)), dim3(kBlock), 0, s, ts, st, sidx, slab, nsp);
}

void launch_clear_dirty(const PassState& st, uint32_t nW, hipStream_t s)
{
    if (nW == 0 || st.dirty == nullptr) return;
    hipLaunchKernelGGL(k_clear_dirty, dim3(grid_for(nW, kBlock, 1u << 16)), dim3(kBlock), 0, s, st, nW);
}

void launch_window_max(const TemplateSet& ts, const PassState& st, uint32_t nW, hipStream_t s)
{
    if (nW == 0) return;
    hipLaunchKernelGGL(k_window_max, dim3(grid_for(nW, kBlock / 64, 1u << 20)), dim3(kBlock), 0, s, ts.U, ts.L, st.N,
                       st.wmax_v, st.wmax_p, st.W, nW, st);
}

// ---------------------------------------------------------------------------
// Flagged windows grouped by the label of their best spike (counting sort,
// order within a label unspecified): the overlap screen of consecutive blocks
// then reads the same label's Gram rows and minima, which stay in the L2.  The
// screen, selection and reassignment give every window the same result in any
// order (each writes only its own window's state).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_label_hist(const uint32_t* __restrict__ F, const uint32_t* __restrict__ cnt_f,
                                                       const uint32_t* __restrict__ lab, uint32_t* __restrict__ hist)
{
    const uint32_t nF = *cnt_f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nF; i += gridDim.x * blockDim.x)
        atomicAdd(hist + lab[F[i]], 1u);
}

// The same with a block-local histogram in LDS (U words): one global atomic
// per (block, label) instead of one per window.
__global__ __launch_bounds__(kBlock) void k_label_hist_lds(const uint32_t* __restrict__ F,
                                                           const uint32_t* __restrict__ cnt_f,
                                                           const uint32_t* __restrict__ lab, uint32_t* __restrict__ hist,
                                                           int U)
{
    extern __shared__ uint32_t lh[];
    const uint32_t nF = *cnt_f;
    for (int u = threadIdx.x; u < U; u += blockDim.x) lh[u] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nF; i += gridDim.x * blockDim.x)
        atomicAdd(lh + lab[F[i]], 1u);
    __syncthreads();
    for (int u = threadIdx.x; u < U; u += blockDim.x)
        if (lh[u] != 0u) atomicAdd(hist + u, lh[u]);
}

// hist[U] -> exclusive offsets in place (one block), the labels laid out in the
// order order[0 .. U) (label ids ascending without it).
__global__ __launch_bounds__(1024) void k_label_offsets(uint32_t* __restrict__ hist, int U,
                                                        const uint32_t* __restrict__ order)
{
    __shared__ uint32_t part[1024];
    const int t = (int)threadIdx.x;
    const int per = (U + 1023) / 1024;
    uint32_t sum = 0;
    for (int k = 0; k < per; k++) {
        const int r = t * per + k;
        if (r < U) sum += hist[order ? order[r] : (uint32_t)r];
    }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (int k = 0; k < per; k++) {
        const int r = t * per + k;
        if (r < U) {
            const uint32_t u = order ? order[r] : (uint32_t)r;
            const uint32_t h = hist[u];
            hist[u] = run;
            run += h;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_label_scatter(const uint32_t* __restrict__ F, const uint32_t* __restrict__ cnt_f,
                                                          const uint32_t* __restrict__ lab, uint32_t* __restrict__ off,
                                                          uint32_t* __restrict__ out)
{
    const uint32_t nF = *cnt_f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nF; i += gridDim.x * blockDim.x) {
        const uint32_t id = F[i];
        out[atomicAdd(off + lab[id], 1u)] = id;
    }
}

// The same with block-local counts in LDS: each block counts its windows per
// label, reserves one range per label with a single global atomic, and places
// its windows within the ranges by LDS atomics.  The block takes the same
// windows in both of its passes (grid-stride, fixed grid).
__global__ __launch_bounds__(kBlock) void k_label_scatter_lds(const uint32_t* __restrict__ F,
                                                              const uint32_t* __restrict__ cnt_f,
                                                              const uint32_t* __restrict__ lab,
                                                              uint32_t* __restrict__ off, uint32_t* __restrict__ out,
                                                              int U)
{
    extern __shared__ uint32_t lh[];  // [U] counts, then each label's next slot
    const uint32_t nF = *cnt_f;
    for (int u = threadIdx.x; u < U; u += blockDim.x) lh[u] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nF; i += gridDim.x * blockDim.x)
        atomicAdd(lh + lab[F[i]], 1u);
    __syncthreads();
    for (int u = threadIdx.x; u < U; u += blockDim.x)
        if (lh[u] != 0u) lh[u] = atomicAdd(off + u, lh[u]);
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nF; i += gridDim.x * blockDim.x) {
        const uint32_t id = F[i];
        out[atomicAdd(lh + lab[id], 1u)] = id;
    }
}

void launch_sort_by_label(const PassState& st, uint32_t nF, int U, uint32_t* out, uint32_t* hist, hipStream_t s,
                          const uint32_t* order)
{
    if (nF == 0) return;
    launch_fill_u32(hist, 0u, U, s);
    if (U <= 16384) {
        // about 16 windows per thread: few blocks, so few global atomics per label
        const unsigned g = (unsigned)std::min<uint32_t>((nF + 16 * kBlock - 1) / (16 * kBlock), 1024u);
        const size_t lds = (size_t)U * 4;
        hipLaunchKernelGGL(k_label_hist_lds, dim3(g), dim3(kBlock), lds, s, st.F, st.counters + CNT_F, st.best_lab,
                           hist, U);
        hipLaunchKernelGGL(k_label_offsets, dim3(1), dim3(1024), 0, s, hist, U, order);
        hipLaunchKernelGGL(k_label_scatter_lds, dim3(g), dim3(kBlock), lds, s, st.F, st.counters + CNT_F, st.best_lab,
                           hist, out, U);
        return;
    }
    const unsigned g = (unsigned)std::min<uint32_t>((nF + kBlock - 1) / kBlock, 4096u);
    hipLaunchKernelGGL(k_label_hist, dim3(g), dim3(kBlock), 0, s, st.F, st.counters + CNT_F, st.best_lab, hist);
    hipLaunchKernelGGL(k_label_offsets, dim3(1), dim3(1024), 0, s, hist, U, order);
    hipLaunchKernelGGL(k_label_scatter, dim3(g), dim3(kBlock), 0, s, st.F, st.counters + CNT_F, st.best_lab, hist, out);
}

// The first pass's flagged windows, rechecked region by region while the
// voltage is still being copied (region r: windows [r * rwin, (r + 1) * rwin),
// its flagged windows grouped by label at Fs + r * rwin, its counters at
// rcnt + r * CNT_N): one list F in region order, the pass's flagged count, and
// the regions' recheck statistics added to the run's counters.  Block r copies
// region r.
__global__ __launch_bounds__(kBlock) void k_merge_regions(uint32_t* __restrict__ F, const uint32_t* __restrict__ Fs,
                                                          const uint32_t* __restrict__ rcnt, int nreg, uint32_t rwin,
                                                          uint32_t* __restrict__ cnt)
{
    const int r = blockIdx.x;
    uint32_t o = 0;
    for (int q = 0; q < r; q++) o += rcnt[q * CNT_N + CNT_F];
    const uint32_t c = rcnt[r * CNT_N + CNT_F];
    for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) F[o + i] = Fs[(size_t)r * rwin + i];
    if (r == nreg - 1 && threadIdx.x == 0) {
        cnt[CNT_F] = o + c;
        const int add[] = {CNT_CAND, CNT_EVAL, CNT_SURV, CNT_NSP};
        for (int k : add) {
            uint32_t t = 0;
            for (int q = 0; q < nreg; q++) t += rcnt[q * CNT_N + k];
            cnt[k] += t;
        }
    }
}

void launch_merge_regions(uint32_t* F, const uint32_t* Fs, const uint32_t* rcnt, int nreg, uint32_t rwin, uint32_t* cnt,
                          hipStream_t s)
{
    if (nreg > 0) hipLaunchKernelGGL(k_merge_regions, dim3(nreg), dim3(kBlock), 0, s, F, Fs, rcnt, nreg, rwin, cnt);
}

void launch_flag_compact(const PassState& st, uint32_t nW, uint32_t* blockbuf, hipStream_t s)
{
    SelFlagged sel{st.W, st.orc, st.F};
    compact(sel, nW, blockbuf, st.counters + CNT_F, s);
}

void launch_spike_compact(const PassState& st, uint32_t nW, uint32_t* out_idx, uint32_t* out_lab, uint64_t base,
                          uint32_t* blockbuf, hipStream_t s)
{
    SelSpikes sel{st.has, st.cand_idx, st.cand_lab, out_idx, out_lab, base};
    compact(sel, nW, blockbuf, st.counters + CNT_SPIKES, s);
}

void launch_next_windows(const PassState& st, uint32_t* blockbuf, hipStream_t s)
{
    SelNextWindows sel{st.check, st.orc, st.W};
    compact(sel, (uint32_t)st.nwin, blockbuf, st.counters + CNT_W, s);
}

uint32_t recheck_literal_chunks(const TemplateSet& ts, int64_t n_max)
{
    const uint64_t S = 2 * (uint64_t)n_max + 1;
    const uint64_t items = S * S * (uint64_t)ts.U;
    return (uint32_t)((items + kBlock * kItemsPerThread - 1) / (kBlock * kItemsPerThread));
}

void launch_recheck_literal(const TemplateSet& ts, const PassState& st, int64_t n_max, uint32_t nF, float* part_v,
                            uint32_t* part_k, hipStream_t s)
{
    if (nF == 0) return;
    const uint32_t nchunk = recheck_literal_chunks(ts, n_max);
    const size_t lds_need = (size_t)ts.C * ts.L * sizeof(float);
    const int use_lds = lds_need <= 48 * 1024;
    const uint64_t blocks = (uint64_t)nF * nchunk;
    hipLaunchKernelGGL(k_recheck_literal, dim3(grid_for((int64_t)blocks, 1, 1u << 21)), dim3(kBlock),
                       use_lds ? lds_need : 0, s, ts, st, n_max, nchunk, use_lds, part_v, part_k);
    hipLaunchKernelGGL(k_recheck_parse, dim3(grid_for(nF, kBlock / 64, 1u << 20)), dim3(kBlock), 0, s, ts, st, n_max,
                       nchunk, part_v, part_k);
}

void launch_reassign(const TemplateSet& ts, const PassState& st, uint32_t nF, hipStream_t s)
{
    if (nF == 0) return;
    hipLaunchKernelGGL(k_reassign, dim3(grid_for(nF, kBlock, 1u << 20)), dim3(kBlock), 0, s, ts.L, st);
}

void launch_pursuit(const TemplateSet& ts, const PassState& st, uint32_t nW, int use_giveup, hipStream_t s)
{
    if (nW == 0) return;
    const unsigned g = grid_for(nW, kBlock, 1u << 20);
    hipLaunchKernelGGL(k_pursuit_decide, dim3(g), dim3(kBlock), 0, s, ts.L, ts.thr, st, nW, use_giveup);
    hipLaunchKernelGGL(k_copy_bytes, dim3(grid_for(st.nwin, kBlock * 4)), dim3(kBlock), 0, s, st.check, st.check_snap,
                       st.nwin);
    hipLaunchKernelGGL(k_pursuit_prune, dim3(g), dim3(kBlock), 0, s, ts.L, ts.thr, st, nW);
}

void launch_residual(const TemplateSet& ts, float* V, int64_t N, const uint32_t* sidx, const uint32_t* slab, uint32_t nsp,
                     int reach, hipStream_t s)
{
    if (nsp == 0) return;
    hipLaunchKernelGGL(k_residual, dim3(grid_for(nsp, 1, 1u << 20)), dim3(kBlock), 0, s, ts, V, N,
                       sidx, slab, nsp, reach);
}

void launch_adjusted_clips(const TemplateSet& ts, const float* V, int64_t N, const uint32_t* sidx, const uint32_t* slab,
                           uint64_t nsp, float* clips, hipStream_t s)
{
    if (nsp == 0) return;
    hipLaunchKernelGGL(k_adjusted_clips, dim3(grid_for((int64_t)(nsp * ts.C * ts.L), kBlock * 4)), dim3(kBlock), 0, s,
                       ts, V, N, sidx, slab, nsp, clips);
}


// ---------------------------------------------------------------------------
// Segmented launches of the pass kernels: one grid over every active segment
// of a batch.  Block b works on entry eob[b] of the launch table with the
// entry's blocks as its grid; each body is the single-segment kernel's.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_fill_eob(const SegLaunch* __restrict__ tab, uint32_t ntab, uint16_t* eob)
{
    for (uint32_t e = blockIdx.x; e < ntab; e += gridDim.x) {
        const uint32_t b0 = tab[e].blk0, nb = tab[e].nblk;
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) eob[b0 + i] = (uint16_t)e;
    }
}

#define SEG_ENTRY()                                      \
    const SegLaunch e = tab[eob[blockIdx.x]];            \
    const SegDesc& d = sd[e.seg];                        \
    const uint32_t bx = blockIdx.x - e.blk0, gx = e.nblk

template <int LT>
__global__ __launch_bounds__(kBlock) void k_likelihood_seg(const SegDesc* __restrict__ sd, const SegLaunch* __restrict__ tab,
                                                           const uint16_t* __restrict__ eob, int use_dirty)
{
    SEG_ENTRY();
    likelihood_body<LT>(d.ts, d.st.V, d.st.N, d.st.W, e.n, d.wb, d.sv, d.st.LIK, d.st.wmax_v, d.st.wmax_p,
                        use_dirty ? d.st.dirty : nullptr, d.st.counters + CNT_PAIRS, nullptr, bx, gx);
}

__global__ __launch_bounds__(kBlock) void k_clear_dirty_seg(const SegDesc* __restrict__ sd,
                                                            const SegLaunch* __restrict__ tab,
                                                            const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    clear_dirty_body(d.st, e.n, bx, gx);
}

__global__ __launch_bounds__(kBlock) void k_window_max_seg(const SegDesc* __restrict__ sd,
                                                           const SegLaunch* __restrict__ tab,
                                                           const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    window_max_body(d.ts.U, d.ts.L, d.st.N, d.st.wmax_v, d.st.wmax_p, d.st.W, e.n, d.st, bx, gx);
}

// The three ordered compactions of a pass, by kind: 0 flagged windows (into F,
// count CNT_F), 1 this pass's spikes (appended to the segment's list, count
// CNT_SPIKES), 2 the next window list (into W, count CNT_W).
template <int K>
struct SegSel;
template <>
struct SegSel<0> {
    static __device__ SelFlagged make(const SegDesc& d, const SegDyn&) { return SelFlagged{d.st.W, d.st.orc, d.st.F}; }
    static constexpr int cnt = CNT_F;
};
template <>
struct SegSel<1> {
    static __device__ SelSpikes make(const SegDesc& d, const SegDyn& y)
    {
        return SelSpikes{d.st.has, d.st.cand_idx, d.st.cand_lab, y.sp_idx, y.sp_lab, y.total};
    }
    static constexpr int cnt = CNT_SPIKES;
};
template <>
struct SegSel<2> {
    static __device__ SelNextWindows make(const SegDesc& d, const SegDyn&) { return SelNextWindows{d.st.check, d.st.orc, d.st.W}; }
    static constexpr int cnt = CNT_W;
};

template <int K>
__global__ __launch_bounds__(kBlock) void k_compact_count_seg(const SegDesc* __restrict__ sd, const SegDyn* __restrict__ dy,
                                                              const SegLaunch* __restrict__ tab,
                                                              const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    (void)gx;
    compact_count_body(SegSel<K>::make(d, dy[e.seg]), e.n, d.blockbuf, bx, gx);
}

template <int K>
__global__ __launch_bounds__(1024) void k_compact_scan_seg(const SegDesc* __restrict__ sd, const SegLaunch* __restrict__ tab)
{
    const SegLaunch e = tab[blockIdx.x];
    const SegDesc& d = sd[e.seg];
    compact_scan_body(d.blockbuf, (e.n + kCompactTile - 1) / kCompactTile, d.st.counters + SegSel<K>::cnt);
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_compact_scatter_seg(const SegDesc* __restrict__ sd,
                                                                const SegDyn* __restrict__ dy,
                                                                const SegLaunch* __restrict__ tab,
                                                                const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    compact_scatter_body(SegSel<K>::make(d, dy[e.seg]), e.n, d.blockbuf, bx, gx);
}

__global__ __launch_bounds__(kBlock) void k_reassign_seg(const SegDesc* __restrict__ sd, const SegLaunch* __restrict__ tab,
                                                         const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    if (e.n == 0) return;  // a segment of a graph pass with no flagged windows
    reassign_body(d.ts.L, d.st, bx, gx);
}

// a0: 1 when the segment's recheck ran this pass (its give-up flags are current)
__global__ __launch_bounds__(kBlock) void k_pursuit_decide_seg(const SegDesc* __restrict__ sd,
                                                               const SegLaunch* __restrict__ tab,
                                                               const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    pursuit_decide_body(d.ts.L, d.ts.thr, d.st, e.n, (int)e.a0, bx, gx);
}

__global__ __launch_bounds__(kBlock) void k_check_snapshot_seg(const SegDesc* __restrict__ sd,
                                                               const SegLaunch* __restrict__ tab,
                                                               const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    if (e.n == 0) return;  // a segment of a graph pass that has converged
    copy_bytes_body(d.st.check, d.st.check_snap, d.st.nwin, bx, gx);
}

__global__ __launch_bounds__(kBlock) void k_pursuit_prune_seg(const SegDesc* __restrict__ sd,
                                                              const SegLaunch* __restrict__ tab,
                                                              const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    pursuit_prune_body(d.ts.L, d.ts.thr, d.st, e.n, bx, gx);
}

__global__ __launch_bounds__(kBlock) void k_residual_seg(const SegDesc* __restrict__ sd, const SegDyn* __restrict__ dy,
                                                         const SegLaunch* __restrict__ tab,
                                                         const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    const SegDyn y = dy[e.seg];
    residual_body(d.ts, d.st.V, d.st.N, y.sp_idx + y.total, y.sp_lab + y.total, e.n, d.reach, bx, gx);
}

__global__ __launch_bounds__(kBlock) void k_mark_dirty_seg(const SegDesc* __restrict__ sd, const SegDyn* __restrict__ dy,
                                                           const SegLaunch* __restrict__ tab,
                                                           const uint16_t* __restrict__ eob)
{
    SEG_ENTRY();
    const SegDyn y = dy[e.seg];
    mark_dirty_body(d.ts, d.st, y.sp_idx + y.total, y.sp_lab + y.total, e.n, bx, gx);
}

// Per segment: zero the pass's computed-pair count, or copy the counters out.
__global__ __launch_bounds__(kBlock) void k_seg_counters(const SegDesc* __restrict__ sd, uint32_t nseg,
                                                         uint32_t* __restrict__ out)
{
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < nseg; g += gridDim.x * blockDim.x) {
        uint32_t* c = sd[g].st.counters;
        if (out == nullptr) {
            c[CNT_PAIRS] = 0u;
        } else {
            for (int k = 0; k < CNT_N; k++) out[(size_t)g * CNT_N + k] = c[k];
        }
    }
}

// Item counts of a graph pass's entries (phase 0 at the start of the pass,
// 1 after the flagged windows are counted, 2 after this pass's spikes are):
// segments that have converged get none.
__global__ __launch_bounds__(kBlock) void k_seg_plan(const SegDesc* __restrict__ sd, const SegDyn* __restrict__ dy,
                                                     SegLaunch* __restrict__ tab, SegPlan p, int phase)
{
    for (int r = 0; r < p.nr; r++) {
        if (p.phase[r] != phase) continue;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.cnt[r]; i += gridDim.x * blockDim.x) {
            SegLaunch& e = tab[p.off[r] + i];
            const SegDyn y = dy[e.seg];
            const uint32_t* c = sd[e.seg].st.counters;
            uint32_t n = 0, a0 = e.a0;
            if (y.active) {
                switch (p.src[r]) {
                case SRC_NW: n = y.nW; break;
                case SRC_NWIN: n = (uint32_t)sd[e.seg].st.nwin; break;
                case SRC_NF: n = c[CNT_F]; break;
                case SRC_DECIDE:
                    n = y.nW;
                    a0 = (p.recheck && c[CNT_F] > 0) ? 1u : 0u;
                    break;
                default: n = c[CNT_SPIKES]; break;
                }
            }
            e.n = n;
            e.a0 = a0;
        }
    }
}
#undef SEG_ENTRY

// Launchers of the segmented kernels.  `tab` (device) holds ntab entries whose
// blocks tile [0, nblocks); `eob` has room for nblocks entries.
static void fill_eob(const SegLaunch* tab, uint32_t ntab, uint16_t* eob, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_eob, dim3(std::max<uint32_t>(1, std::min<uint32_t>(ntab, 1024))), dim3(kBlock), 0, s, tab,
                       ntab, eob);
}

void seg_likelihood_lds(const TemplateSet& ts, int* wb, int* sv, size_t* lds)
{
    likelihood_plan(ts, wb, sv);
    *lds = (size_t)(*wb) * likelihood_lds_per_window(ts, *sv);
}

void launch_seg_fill_eob(const SegLaunch* tab, uint32_t ntab, uint16_t* eob, hipStream_t s)
{
    if (ntab > 0) fill_eob(tab, ntab, eob, s);
}

void launch_seg_plan(const SegDesc* sd, const SegDyn* dy, SegLaunch* tab, const SegPlan& p, int phase, hipStream_t s,
                     GraphRec* rec)
{
    launch_or_record(rec, k_seg_plan, dim3(4), dim3(kBlock), 0, s, sd, dy, tab, p, phase);
}

// rec: record the launch in a graph being built (eob then already holds this
// table's block-to-entry map); otherwise fill eob and launch on s.
void launch_seg_pass_kernel(int kind, int L, const SegDesc* sd, const SegDyn* dy, const SegLaunch* tab, uint32_t ntab,
                            uint32_t nblocks, uint16_t* eob, size_t lds, int flag, hipStream_t s, GraphRec* rec)
{
    if (ntab == 0 || nblocks == 0) return;
    if (rec == nullptr) fill_eob(tab, ntab, eob, s);
    const uint16_t* m = eob;
    const dim3 g(nblocks), b(kBlock);
    switch (kind) {
    case SEG_LIKELIHOOD:
        if (L == 48) launch_or_record(rec, k_likelihood_seg<48>, g, b, lds, s, sd, tab, m, flag);
        else if (L == 64) launch_or_record(rec, k_likelihood_seg<64>, g, b, lds, s, sd, tab, m, flag);
        else if (L == 32) launch_or_record(rec, k_likelihood_seg<32>, g, b, lds, s, sd, tab, m, flag);
        else launch_or_record(rec, k_likelihood_seg<0>, g, b, lds, s, sd, tab, m, flag);
        break;
    case SEG_CLEAR_DIRTY: launch_or_record(rec, k_clear_dirty_seg, g, b, 0, s, sd, tab, m); break;
    case SEG_WINDOW_MAX: launch_or_record(rec, k_window_max_seg, g, b, 0, s, sd, tab, m); break;
    case SEG_REASSIGN: launch_or_record(rec, k_reassign_seg, g, b, 0, s, sd, tab, m); break;
    case SEG_DECIDE: launch_or_record(rec, k_pursuit_decide_seg, g, b, 0, s, sd, tab, m); break;
    case SEG_SNAPSHOT: launch_or_record(rec, k_check_snapshot_seg, g, b, 0, s, sd, tab, m); break;
    case SEG_PRUNE: launch_or_record(rec, k_pursuit_prune_seg, g, b, 0, s, sd, tab, m); break;
    case SEG_RESIDUAL: launch_or_record(rec, k_residual_seg, g, b, 0, s, sd, dy, tab, m); break;
    case SEG_MARK_DIRTY: launch_or_record(rec, k_mark_dirty_seg, g, b, 0, s, sd, dy, tab, m); break;
    default: break;
    }
}

// Ordered compaction of kind K over every entry: one block per tile of
// kCompactTile items (the entries' nblk), one scan block per entry.
template <int K>
static void seg_compact(const SegDesc* sd, const SegDyn* dy, const SegLaunch* tab, uint32_t ntab, uint32_t nblocks,
                        const uint16_t* eob, hipStream_t s, GraphRec* rec)
{
    const dim3 g(nblocks), b(kBlock);
    if (nblocks > 0) launch_or_record(rec, k_compact_count_seg<K>, g, b, 0, s, sd, dy, tab, eob);
    launch_or_record(rec, k_compact_scan_seg<K>, dim3(ntab), dim3(1024), 0, s, sd, tab);
    if (nblocks > 0) launch_or_record(rec, k_compact_scatter_seg<K>, g, b, 0, s, sd, dy, tab, eob);
}

void launch_seg_compact(int kind, const SegDesc* sd, const SegDyn* dy, const SegLaunch* tab, uint32_t ntab,
                        uint32_t nblocks, uint16_t* eob, hipStream_t s, GraphRec* rec)
{
    if (ntab == 0) return;
    if (rec == nullptr && nblocks > 0) fill_eob(tab, ntab, eob, s);
    if (kind == 0) seg_compact<0>(sd, dy, tab, ntab, nblocks, eob, s, rec);
    else if (kind == 1) seg_compact<1>(sd, dy, tab, ntab, nblocks, eob, s, rec);
    else seg_compact<2>(sd, dy, tab, ntab, nblocks, eob, s, rec);
}

void launch_seg_counters(const SegDesc* sd, uint32_t nseg, uint32_t* out, hipStream_t s, GraphRec* rec)
{
    if (nseg == 0) return;
    launch_or_record(rec, k_seg_counters, dim3(grid_for(nseg, kBlock, 1024)), dim3(kBlock), 0, s, sd, nseg, out);
}

}  // namespace fbp
