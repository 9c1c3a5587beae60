// Likelihood on the gfx950 matrix cores in BF16 or FP8 (e4m3), one template per tile.
//
// compute_full_likelihood (SF/kernels/binary_pursuit.cl:216-300, SF =
// spikesorting_fullpursuit/) gives, for every position x of a listed window and
// every template t,
//     LIK(t, x) = bias_t + sum_{c, j} T[t][c][j] * V[c][x + j].
// Per template, with rows m = 16-position groups (x_m a multiple of 16) and
// columns n = 0..15,
//     D[m][n] = sum_c sum_k Vq[c][x_m + k] * Bt[c][k][n],   Bt[c][k][n] = T[t][c][k - n],
// is LIK(t, x_m + n) - bias_t: the B operand is the template shifted by n samples
// (constant, built once), the A operand is 8 consecutive samples of one channel of
// the channel-major quantized voltage (one aligned 16- or 8-byte load), and the
// reduction runs over the template's own non-zero channels, k < 32 * KB with
// KB = ceil((L + 15) / 32).  So the matrix cores do nz(t) * 32 KB multiply-adds
// per position (64 for nz(t) * 48 useful at L = 48), whatever the probe, and the
// same kernel takes the dense first pass and the changed (window, template)
// pairs of later passes.
//
// Decisions stay those of FP32: only likelihoods above 0, or their sign, decide
// anything.  Per 16-position row, delta bounds |approximation - exact FP32
// chain|; an 8-position group whose every approximation is <= -delta keeps the
// approximations (negative, as are the exact values), every other group is
// marked in gmask and recomputed by k_likelihood with the exact FP32 chain.  A
// NaN approximation (an FP8 operand out of range) marks its group too.
#include <math.h>

#include <algorithm>

#include "fbp_device.h"

namespace fbp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMmaMaxFrag = 16;  // B fragments held in registers (channels x k-blocks)
constexpr int kMmaWb = 64;       // listed windows per block

// Voltage scale: the power of two that maps the largest |V| into [2^e, 2^(e+1)).
__global__ void k_vscale_to(const uint32_t* __restrict__ mbits, float* __restrict__ vs, int e_target)
{
    const float m = __uint_as_float(*mbits);
    int e = 0;
    const bool ok = m > 0.0f && isfinite(m);
    if (ok) frexpf(m, &e);
    *vs = ok ? ldexpf(1.0f, e_target + 1 - e) : 1.0f;
}

template <int FMT>
__device__ __forceinline__ void quant_block(const float* __restrict__ V, int64_t N, int C, int c, int64_t b,
                                            const MmaVoltage& mv, float vs)
{
    float v[16];
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int64_t x = b * 16 + i;
        v[i] = x < N ? V[(size_t)x * C + c] : 0.0f;
        m = fmaxf(m, fabsf(v[i]));
    }
    mv.bmax[(size_t)c * mv.nbq + b] = m;
    if (FMT == 0) {
        bf16x8 h0, h1;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            h0[i] = (__bf16)(v[i] * vs);
            h1[i] = (__bf16)(v[8 + i] * vs);
        }
        bf16x8* dst = reinterpret_cast<bf16x8*>(static_cast<__bf16*>(mv.q) + (size_t)c * mv.ldq + b * 16);
        dst[0] = h0;
        dst[1] = h1;
    } else {
        // round to nearest even; out of range (|x| >= 464) gives NaN, which lists its groups for the exact pass
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int r = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * q] * vs, v[4 * q + 1] * vs, 0, false);
            r = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * q + 2] * vs, v[4 * q + 3] * vs, r, true);
            w[q] = (uint32_t)r;
        }
        uint4* dst = reinterpret_cast<uint4*>(static_cast<uint8_t*>(mv.q) + (size_t)c * mv.ldq + b * 16);
        *dst = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Every 16-sample block of every channel.
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_quantize(const float* __restrict__ V, int64_t N, int C, MmaVoltage mv)
{
    const float vs = *mv.vs;
    const int64_t tot = mv.nbq * C;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = e / C;
        const int c = (int)(e - b * C);
        quant_block<FMT>(V, N, C, c, b, mv, vs);
    }
}

// The blocks a pass's spikes changed: [p, p + L) for each onset p, every channel
// (blocks hit by several spikes are written with the same values).
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_requantize(const float* __restrict__ V, int64_t N, int C, int L,
                                                       MmaVoltage mv, const uint32_t* __restrict__ sidx, uint32_t nsp)
{
    const float vs = *mv.vs;
    const int nbs = (L + 15) / 16 + 1;
    const int64_t tot = (int64_t)nsp * nbs * C;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const int64_t r = e / C;
        const int j = (int)(r % nbs);
        const int64_t k = r / nbs;
        const int64_t p = sidx[k];
        const int64_t b = p / 16 + j;
        if (b * 16 > p + L - 1 || b >= mv.nbq) continue;
        quant_block<FMT>(V, N, C, c, b, mv, vs);
    }
}

template <int FMT>
struct FragT;
template <>
struct FragT<0> {
    typedef bf16x8 type;
};
template <>
struct FragT<1> {
    typedef long type;
};

template <int FMT>
__device__ __forceinline__ floatx4 mma(const typename FragT<FMT>::type& a, const typename FragT<FMT>::type& b,
                                       floatx4 c)
{
    if constexpr (FMT == 0) return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
}

// One block (4 waves) per chunk of kMmaWb listed windows; wave w takes the
// templates t = w, w + 4, ...  For template t the wave compacts the windows whose
// pair must be computed, stages t's B fragments in its own LDS slots (each lane
// reads back only what it wrote), and runs tiles of P windows (16 rows of 16
// positions), issuing the next tile's loads before the current tile's products.
// The reduction index k of lane
// group q = lane / 16 covers the 8 KB consecutive samples from 8 KB q (k-block
// kb: the kb-th 8 of them), so each lane reads one contiguous run of a channel.
// gmask[window][template] receives the pair's 8-position groups to rescore.
// Templates with more channels than one register group (G) are rescored whole.
template <int FMT, int KB>
__global__ __launch_bounds__(256) void k_likmma(TemplateSet ts, PassState st, MmaTables mt, MmaVoltage mv, uint32_t nW,
                                                uint32_t chunk0, uint8_t* __restrict__ gmask,
                                                const uint64_t* __restrict__ dirty)
{
    typedef typename FragT<FMT>::type frag_t;
    constexpr int G = (kMmaMaxFrag / KB < 8) ? kMmaMaxFrag / KB : 8;  // channels held on the matrix-core path
    constexpr int NF = G * KB;
    __shared__ uint32_t wsh[kMmaWb];
    __shared__ unsigned long long dsh[kMmaWb];
    __shared__ uint32_t wact[4][kMmaWb];
    __shared__ frag_t Bs[4][NF * 64];
    const int U = ts.U, L = ts.L;
    const int64_t N = st.N;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t a0 = ((uint64_t)chunk0 + blockIdx.x) * kMmaWb;
    if (a0 >= nW) return;  // uniform over the block
    const int nb = (int)min((uint64_t)kMmaWb, (uint64_t)nW - a0);
    if (tid < kMmaWb) {
        const uint32_t w = tid < nb ? st.W[a0 + tid] : 0u;
        wsh[tid] = w;
        dsh[tid] = tid < nb ? (dirty != nullptr ? dirty[w] : ~0ull) : 0ull;
    }
    __syncthreads();  // the only block-wide barrier: the waves are independent from here
    const int R = mt.R, P = mt.P;
    const int ngr = (L + kPos - 1) / kPos;
    const float vs = *mv.vs;
    const int q4 = lane >> 4, col = lane & 15;
    const int p_row = col / R, rr_row = col - p_row * R;  // this lane's A row: window and row within it
    const int m_b = lane >> 2, j_b = lane & 3;             // bound loads: row m_b, its block j_b
    const int p_b = m_b / R, rr_b = m_b - p_b * R;
    const unsigned long long my_d = dsh[lane];
    const uint32_t my_w = wsh[lane];
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t* wa = wact[wv];
    frag_t* Bw = Bs[wv];
    uint32_t rescored = 0, pairs = 0;
    for (int t = wv; t < U; t += 4) {
        // pairs to compute: every listed window without a dirty map, else those where
        // a channel of t changed since the window was last computed
        const bool act = (ts.cmask[t] & my_d) != 0ull;
        const uint64_t am = __ballot(act);
        if (am == 0ull) continue;  // uniform over the wave
        const int nact = __popcll(am);
        pairs += (uint32_t)nact;
        const int r0 = ts.off[t], nr = ts.nch[t];
        if (nr > G) {
            // the exact pass computes every group of these pairs
            if (act) gmask[(size_t)my_w * U + t] = (uint8_t)((1u << ngr) - 1u);
            rescored += (lane == 0) ? (uint32_t)nact * (uint32_t)(ngr * kPos) : 0u;
            continue;
        }
        if (act) wa[__popcll(am & below)] = my_w;
        {
            const frag_t* F = static_cast<const frag_t*>(mt.frag) + (size_t)r0 * KB * 64 + lane;
#pragma unroll
            for (int i = 0; i < NF; i++)
                if (i < nr * KB) Bw[i * 64 + lane] = F[(size_t)i * 64];  // this lane's own slots
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int chs[G];
        float rl1[G];
#pragma unroll
        for (int i = 0; i < G; i++) {
            chs[i] = i < nr ? ts.chan[r0 + i] : 0;
            rl1[i] = i < nr ? mt.rowL1[r0 + i] : 0.0f;
        }
        const float ts_t = mt.tsc[t];
        const float inv = 1.0f / (ts_t * vs);
        const float bt = ts.bias[t];
        const float dconst = mt.kb * fabsf(bt) + mt.ka * mt.T1[t] / vs + mt.kd * inv;
        const float ka_s = mt.ka * (float)(32 * KB) / ts_t;  // products per channel and position
        const int nbr = (L + 14) / 16 + 1;                   // 16-sample blocks a row's 16 positions read
        const int ntile = (nact + P - 1) / P;
        auto row_x = [&](int tile, int prow, int rrow) -> int64_t {
            const bool ok = prow < P && tile * P + prow < nact;
            return ok ? (int64_t)wa[tile * P + prow] * L + 16 * rrow : 0;
        };
        auto load_a = [&](int64_t xrow, frag_t* a) {
#pragma unroll
            for (int i = 0; i < G; i++) {
                if (i >= nr) break;
                const int64_t base = (int64_t)chs[i] * mv.ldq + xrow + 8 * KB * q4;
#pragma unroll
                for (int kb = 0; kb < KB; kb++) {
                    if constexpr (FMT == 0)
                        a[i * KB + kb] =
                            *reinterpret_cast<const bf16x8*>(static_cast<const __bf16*>(mv.q) + base + 8 * kb);
                    else
                        a[i * KB + kb] = *reinterpret_cast<const long*>(static_cast<const uint8_t*>(mv.q) + base + 8 * kb);
                }
            }
        };
        frag_t acur[NF], anxt[NF];
        load_a(row_x(0, p_row, rr_row), acur);
        for (int tile = 0; tile < ntile; tile++) {
            if (tile + 1 < ntile) load_a(row_x(tile + 1, p_row, rr_row), anxt);
            // bound inputs: lane (row m_b, block j_b) loads blocks j_b, j_b + 4, ... of
            // the nbr 16-sample blocks row m_b reads; the row maximum is a 4-lane reduction
            const int64_t xb = row_x(tile, p_b, rr_b);
            float vb[G];
#pragma unroll
            for (int i = 0; i < G; i++) {
                vb[i] = 0.0f;
                if (i < nr) {
                    const float* bm = mv.bmax + (size_t)chs[i] * mv.nbq + (xb >> 4);
                    vb[i] = bm[j_b];
                    for (int jj = j_b + 4; jj < nbr; jj += 4) vb[i] = fmaxf(vb[i], bm[jj]);
                }
            }
            floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int i = 0; i < G; i++) {
                if (i >= nr) break;
#pragma unroll
                for (int kb = 0; kb < KB; kb++) acc = mma<FMT>(acur[i * KB + kb], Bw[(i * KB + kb) * 64 + lane], acc);
            }
            float Q = 0.0f, S = 0.0f;  // of row m_b (every lane of the 4-lane group)
#pragma unroll
            for (int i = 0; i < G; i++) {
                float v = vb[i];
                v = fmaxf(v, __shfl_xor(v, 1, 64));
                v = fmaxf(v, __shfl_xor(v, 2, 64));
                if (i < nr) {
                    Q = fmaf(rl1[i], v, Q);
                    S += v;
                }
            }
            uint64_t bal[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                // this lane holds D[mo][col] with mo = 4 * q4 + r
                const int mo = 4 * q4 + r;
                const int po = mo / R, ro = mo - po * R;
                const float Qm = __shfl(Q, 4 * mo, 64), Sm = __shfl(S, 4 * mo, 64);
                const bool ok = po < P && tile * P + po < nact;
                const int pos = 16 * ro + col;
                const int64_t x = (ok ? (int64_t)wa[tile * P + po] * L : 0) + pos;
                const bool live = ok && pos < L && x < N;
                bool need = false;
                float v = 0.0f;
                if (live && x + L < N) {
                    v = bt + acc[r] * inv;
                    const float d = mt.kl * Qm + dconst + ka_s * Sm;
                    need = !(v <= -d);
                }
                if (live) st.LIK[(size_t)t * N + x] = v;
                bal[r] = __ballot(need);
            }
            // 8-position groups to rescore, per window of the tile
            if (lane < P && tile * P + lane < nact) {
                uint32_t gm = 0;
                for (int ro = 0; ro < R; ro++) {
                    const int mo = lane * R + ro;
                    const uint32_t bits = (uint32_t)((bal[mo & 3] >> (16 * (mo >> 2))) & 0xffffull);
                    if (bits & 0xffu) gm |= 1u << (2 * ro);
                    if (bits >> 8) gm |= 1u << (2 * ro + 1);
                }
                const size_t o = (size_t)wa[tile * P + lane] * U + t;
                gmask[o] = (uint8_t)gm;
                if (gm == 0u) {
                    st.wmax_v[o] = -INFINITY;  // every value of the pair is negative
                    st.wmax_p[o] = 0xffffu;
                }
                rescored += (uint32_t)__popc(gm) * kPos;
            }
#pragma unroll
            for (int i = 0; i < NF; i++) acur[i] = anxt[i];
        }
        // wa and Bw are rewritten for the wave's next template
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    for (int off = 32; off > 0; off >>= 1) rescored += __shfl_xor(rescored, off, 64);
    if (lane == 0 && rescored) atomicAdd(st.counters + CNT_LIKX, rescored);
    if (lane == 0 && pairs) atomicAdd(st.counters + CNT_PAIRS, pairs);  // the same count in every lane
}

// ---------------------------------------------------------------------------
// Dense first pass (every window, every template; L a multiple of 16).
// A block takes a quad of templates (one per wave; the quads group templates
// by first channel, so a quad's channels are few) and a run of spans of
// kDenseW consecutive windows.  Per span the block stages the quantized
// samples [x0, x0 + kDenseW L + 64) of the quad's channels and their 16-sample
// block maxima in LDS; each wave holds its template's B fragments in registers
// for the whole run and takes the span's R row tiles (16 rows of 16 positions,
// rows = consecutive 16-position groups), A fragments read from LDS.  Bound,
// LIK values and the 8-position groups to rescore are those of k_likmma.
// ---------------------------------------------------------------------------
constexpr int kDenseW = 16;  // windows per span

// Samples staged per channel: rows reach 16 (kDenseW L / 16 - 1) + 32 KB.
__host__ __device__ inline int dense_span_samples(int L, int KB) { return kDenseW * L + 32 * KB; }

size_t likmma_dense_lds(int L, int KB, int qc, int fmt)
{
    const size_t sps = (size_t)dense_span_samples(L, KB);
    const size_t rowb = ((sps * (fmt == 0 ? 2 : 1)) + 15) & ~(size_t)15;
    return (size_t)qc * rowb + (size_t)qc * (sps / 16) * 4;
}

template <int FMT, int KB>
__global__ __launch_bounds__(256) void k_likmma_dense(TemplateSet ts, PassState st, MmaTables mt, MmaVoltage mv,
                                                      int64_t nspan, uint8_t* __restrict__ gmask)
{
    typedef typename FragT<FMT>::type frag_t;
    constexpr int G = (kMmaMaxFrag / KB < 8) ? kMmaMaxFrag / KB : 8;  // channels on the matrix-core path
    constexpr int NF = G * KB;
    constexpr int ES = FMT == 0 ? 2 : 1;  // bytes per quantized sample
    extern __shared__ __align__(16) unsigned char lsm[];
    const int U = ts.U, L = ts.L, R = L / 16;
    const int64_t N = st.N, nwin = st.nwin;
    const int sps = dense_span_samples(L, KB);  // samples per staged channel row
    const int rowb = (sps * ES + 15) & ~15;
    const int nbk = sps / 16;             // block maxima per staged row
    const int qc = mt.qc;
    float* lbm = reinterpret_cast<float*>(lsm + (size_t)qc * rowb);  // [qc][nbk]
    const int q = (int)(blockIdx.x % (unsigned)mt.nq);
    const int64_t sb = blockIdx.x / (unsigned)mt.nq, nsb = gridDim.x / (unsigned)mt.nq;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int t = mt.qtmp[q * 4 + wv];
    const int* qch = mt.qch + (size_t)q * kQuadCh;
    int nqc = 0;
    for (int i = 0; i < kQuadCh; i++) nqc += qch[i] >= 0 ? 1 : 0;
    const int nr = t >= 0 ? ts.nch[t] : 0, r0 = t >= 0 ? ts.off[t] : 0;
    const bool mm = t >= 0 && nr <= G;    // on the matrix cores (else the exact pass takes every group)
    const int ngr = L / kPos;
    const float vs = *mv.vs;
    // this wave's template: B fragments, channel slots, bound terms
    frag_t Bf[NF];
    int slot[G];
    float rl1[G];
#pragma unroll
    for (int i = 0; i < G; i++) {
        slot[i] = (mm && i < nr) ? mt.rslot[r0 + i] : 0;
        rl1[i] = (mm && i < nr) ? mt.rowL1[r0 + i] : 0.0f;
    }
    {
        const frag_t* F = static_cast<const frag_t*>(mt.frag) + (size_t)r0 * KB * 64 + lane;
#pragma unroll
        for (int i = 0; i < NF; i++) Bf[i] = (mm && i < nr * KB) ? F[(size_t)i * 64] : frag_t{};
    }
    const float ts_t = t >= 0 ? mt.tsc[t] : 1.0f;
    const float inv = 1.0f / (ts_t * vs);
    const float bt = t >= 0 ? ts.bias[t] : 0.0f;
    const float dconst = t >= 0 ? mt.kb * fabsf(bt) + mt.ka * mt.T1[t] / vs + mt.kd * inv : 0.0f;
    const float ka_s = mt.ka * (float)(32 * KB) / ts_t;  // products per channel and position
    const int q4 = lane >> 4, col = lane & 15;
    const int m_b = lane >> 2, j_b = lane & 3;
    const int nbr = (L + 14) / 16 + 1;  // 16-sample blocks a row's 16 positions read
    uint32_t rescored = 0, pairs = 0;
    for (int64_t sp = sb; sp < nspan; sp += nsb) {
        const int64_t w0 = sp * kDenseW;
        const int64_t x0 = w0 * L;
        __syncthreads();  // the previous span's rows are consumed
        // stage: the quad's channels, 16-byte pieces of each row, and the block maxima
        {
            // (the last span's rows run past the quantized rows' end: zeros there)
            // (four 16-byte pieces and four block maxima in flight per thread)
            const int pieces = rowb / 16;
            const int64_t rowend = mv.ldq * ES;  // bytes of a quantized channel row
            const int totp = nqc * pieces, totb = nqc * nbk;
            for (int e0 = tid; e0 < max(totp, totb); e0 += 4 * 256) {
                uint4 v[4];
                float bv[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int e = e0 + j * 256;
                    v[j] = make_uint4(0u, 0u, 0u, 0u);
                    bv[j] = 0.0f;
                    if (e < totp) {
                        const int ci = e / pieces, pc = e - ci * pieces;
                        const int64_t boff = x0 * ES + (int64_t)pc * 16;
                        if (boff + 16 <= rowend)
                            v[j] = *reinterpret_cast<const uint4*>(static_cast<const unsigned char*>(mv.q) +
                                                                   (size_t)qch[ci] * rowend + boff);
                    }
                    if (e < totb) {
                        const int ci = e / nbk, bk = e - ci * nbk;
                        const int64_t gb = (x0 >> 4) + bk;
                        if (gb < mv.nbq) bv[j] = mv.bmax[(size_t)qch[ci] * mv.nbq + gb];
                    }
                }
                __builtin_amdgcn_sched_barrier(0);  // the loads go out together
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int e = e0 + j * 256;
                    if (e < totp) {
                        const int ci = e / pieces, pc = e - ci * pieces;
                        *reinterpret_cast<uint4*>(lsm + (size_t)ci * rowb + (size_t)pc * 16) = v[j];
                    }
                    if (e < totb) lbm[e] = bv[j];
                }
            }
        }
        __syncthreads();
        if (t < 0) continue;
        const int nvw = (int)min((int64_t)kDenseW, nwin - w0);  // windows of this span
        if (lane == 0) pairs += (uint32_t)nvw;
        if (!mm) {
            if (lane < nvw) {
                gmask[(size_t)(w0 + lane) * U + t] = (uint8_t)((1u << ngr) - 1u);
                rescored += (uint32_t)(ngr * kPos);
            }
            continue;
        }
        uint32_t gm = 0;  // lanes < kDenseW: the 8-position groups of window w0 + lane to rescore
        for (int k = 0; k < R; k++) {
            // A: row m = col of this tile, group row gr = 16 k + m, samples from 16 gr
            floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            const int arow = 16 * (16 * k + col) + 8 * KB * q4;
#pragma unroll
            for (int i = 0; i < G; i++) {
                if (i >= nr) break;
                const unsigned char* rb = lsm + (size_t)slot[i] * rowb;
#pragma unroll
                for (int kb = 0; kb < KB; kb++) {
                    const frag_t a = *reinterpret_cast<const frag_t*>(rb + (size_t)(arow + 8 * kb) * ES);
                    acc = mma<FMT>(a, Bf[i * KB + kb], acc);
                }
            }
            // bound inputs of row m_b: its nbr 16-sample blocks, a 4-lane maximum
            float Q = 0.0f, S = 0.0f;
#pragma unroll
            for (int i = 0; i < G; i++) {
                if (i >= nr) break;
                const float* bm = lbm + slot[i] * nbk + 16 * k + m_b;
                float v = bm[j_b];
                for (int jj = j_b + 4; jj < nbr; jj += 4) v = fmaxf(v, bm[jj]);
                v = fmaxf(v, __shfl_xor(v, 1, 64));
                v = fmaxf(v, __shfl_xor(v, 2, 64));
                Q = fmaf(rl1[i], v, Q);
                S += v;
            }
            uint64_t bal[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int mo = 4 * q4 + r;
                const float Qm = __shfl(Q, 4 * mo, 64), Sm = __shfl(S, 4 * mo, 64);
                const int64_t x = x0 + 16 * (16 * k + mo) + col;
                bool need = false;
                float v = 0.0f;
                if (x + L < N) {
                    v = bt + acc[r] * inv;
                    const float d = mt.kl * Qm + dconst + ka_s * Sm;
                    need = !(v <= -d);
                }
                if (x < N) st.LIK[(size_t)t * N + x] = v;
                bal[r] = __ballot(need);
            }
            if (lane < nvw) {
                // rows gr of window `lane` in this tile: [lane R, lane R + R) within [16 k, 16 k + 16)
                const int g0 = max(lane * R, 16 * k), g1 = min(lane * R + R, 16 * k + 16);
                for (int gr = g0; gr < g1; gr++) {
                    const int mo = gr - 16 * k;
                    const uint32_t bits = (uint32_t)((bal[mo & 3] >> (16 * (mo >> 2))) & 0xffffull);
                    const int ro = gr - lane * R;
                    if (bits & 0xffu) gm |= 1u << (2 * ro);
                    if (bits >> 8) gm |= 1u << (2 * ro + 1);
                }
            }
        }
        if (lane < nvw) {
            const size_t o = (size_t)(w0 + lane) * U + t;
            gmask[o] = (uint8_t)gm;
            if (gm == 0u) {
                st.wmax_v[o] = -INFINITY;  // every value of the pair is negative
                st.wmax_p[o] = 0xffffu;
            }
            rescored += (uint32_t)__popc(gm) * kPos;
        }
    }
    for (int off = 32; off > 0; off >>= 1) rescored += __shfl_xor(rescored, off, 64);
    if (lane == 0 && rescored) atomicAdd(st.counters + CNT_LIKX, rescored);
    if (lane == 0 && pairs) atomicAdd(st.counters + CNT_PAIRS, pairs);
}

// Returns false when the dense first pass cannot take this template set.
bool launch_likmma_dense(const TemplateSet& ts, const PassState& st, const MmaTables& mt, const MmaVoltage& mv, int fmt,
                         uint8_t* gmask, hipStream_t s)
{
    if (mt.nq <= 0 || ts.L % 16 != 0 || ts.L > 64) return false;
    const size_t lds = likmma_dense_lds(ts.L, mt.KB, mt.qc, fmt);
    if (lds > 64 * 1024) return false;
    const int64_t nspan = (st.nwin + kDenseW - 1) / kDenseW;
    // a block takes a run of spans; quads vary fastest so the blocks of one span run together
    // about eight blocks per CU in all: each block keeps its template's B fragments for many spans
    const int64_t runs = std::max<int64_t>(1, std::min<int64_t>(nspan, (int64_t)2048 / mt.nq + 1));
    const dim3 g((unsigned)(runs * mt.nq)), b(256);
#define FBP_MMAD(F, K) hipLaunchKernelGGL((k_likmma_dense<F, K>), g, b, lds, s, ts, st, mt, mv, nspan, gmask)
    if (fmt == 0) {
        if (mt.KB == 1) FBP_MMAD(0, 1);
        else if (mt.KB == 2) FBP_MMAD(0, 2);
        else FBP_MMAD(0, 3);
    } else {
        if (mt.KB == 1) FBP_MMAD(1, 1);
        else if (mt.KB == 2) FBP_MMAD(1, 2);
        else FBP_MMAD(1, 3);
    }
#undef FBP_MMAD
    return true;
}

// Fragment values (already scaled, 8 per lane) to the operand format, with the
// conversions the voltage uses.
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_frag_convert(const float* __restrict__ src, void* __restrict__ dst, int64_t n8)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n8; e += (int64_t)gridDim.x * blockDim.x) {
        const float* f = src + e * 8;
        if (FMT == 0) {
            bf16x8 h;
#pragma unroll
            for (int i = 0; i < 8; i++) h[i] = (__bf16)f[i];
            static_cast<bf16x8*>(dst)[e] = h;
        } else {
            int lo = __builtin_amdgcn_cvt_pk_fp8_f32(f[0], f[1], 0, false);
            lo = __builtin_amdgcn_cvt_pk_fp8_f32(f[2], f[3], lo, true);
            int hi = __builtin_amdgcn_cvt_pk_fp8_f32(f[4], f[5], 0, false);
            hi = __builtin_amdgcn_cvt_pk_fp8_f32(f[6], f[7], hi, true);
            static_cast<long*>(dst)[e] = (long)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
        }
    }
}

void launch_frag_convert(const float* src, void* dst, int64_t n8, int fmt, hipStream_t s)
{
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n8 + kBlock - 1) / kBlock, 1 << 16));
    if (fmt == 0) hipLaunchKernelGGL(k_frag_convert<0>, dim3(g), dim3(kBlock), 0, s, src, dst, n8);
    else hipLaunchKernelGGL(k_frag_convert<1>, dim3(g), dim3(kBlock), 0, s, src, dst, n8);
}

void launch_quantize(const float* V, int64_t N, int C, const MmaVoltage& mv, int fmt, hipStream_t s)
{
    const int64_t tot = mv.nbq * C;
    const unsigned g = (unsigned)std::min<int64_t>((tot + kBlock - 1) / kBlock, 1 << 20);
    if (fmt == 0) hipLaunchKernelGGL(k_quantize<0>, dim3(g), dim3(kBlock), 0, s, V, N, C, mv);
    else hipLaunchKernelGGL(k_quantize<1>, dim3(g), dim3(kBlock), 0, s, V, N, C, mv);
}

void launch_requantize(const float* V, int64_t N, int C, int L, const MmaVoltage& mv, int fmt, const uint32_t* sidx,
                       uint32_t nsp, hipStream_t s)
{
    if (nsp == 0) return;
    const int64_t tot = (int64_t)nsp * ((L + 15) / 16 + 1) * C;
    const unsigned g = (unsigned)std::min<int64_t>((tot + kBlock - 1) / kBlock, 1 << 20);
    if (fmt == 0) hipLaunchKernelGGL(k_requantize<0>, dim3(g), dim3(kBlock), 0, s, V, N, C, L, mv, sidx, nsp);
    else hipLaunchKernelGGL(k_requantize<1>, dim3(g), dim3(kBlock), 0, s, V, N, C, L, mv, sidx, nsp);
}

void launch_vscale_to(const uint32_t* mbits, float* vs, int e_target, hipStream_t s)
{
    hipLaunchKernelGGL(k_vscale_to, dim3(1), dim3(1), 0, s, mbits, vs, e_target);
}

void launch_likmma(const TemplateSet& ts, const PassState& st, const MmaTables& mt, const MmaVoltage& mv, int fmt,
                   uint32_t nW, uint8_t* gmask, const uint64_t* dirty, hipStream_t s)
{
    if (nW == 0) return;
    const uint64_t nchunk = (nW + kMmaWb - 1) / kMmaWb;
    const uint64_t per = (uint64_t)1 << 30;  // chunks per launch
    for (uint64_t c0 = 0; c0 < nchunk; c0 += per) {
        const unsigned nb = (unsigned)std::min(per, nchunk - c0);
        const uint32_t cc = (uint32_t)c0;
#define FBP_MMA(F, K) \
    hipLaunchKernelGGL((k_likmma<F, K>), dim3(nb), dim3(256), 0, s, ts, st, mt, mv, nW, cc, gmask, dirty)
        if (fmt == 0) {
            if (mt.KB == 1) FBP_MMA(0, 1);
            else if (mt.KB == 2) FBP_MMA(0, 2);
            else FBP_MMA(0, 3);
        } else {
            if (mt.KB == 1) FBP_MMA(1, 1);
            else if (mt.KB == 2) FBP_MMA(1, 2);
            else FBP_MMA(1, 3);
        }
#undef FBP_MMA
    }
}

}  // namespace fbp
