// Mixed-precision likelihood on the gfx950 matrix cores.
//
// compute_full_likelihood (SF/kernels/binary_pursuit.cl:216-300, SF =
// spikesorting_fullpursuit/) gives, for every position x of a listed window
// and every template t,
//     LIK(t, x) = bias_t + sum_{c, j} T[t][c][j] * V[c][x + j].
// Here the sum runs as an fp16 MFMA product (v_mfma_f32_16x16x32_f16, FP32
// accumulation): rows of the A operand are 16 templates (a tile), columns of
// the B operand are 16 positions, and the reduction index k = (j, channel)
// covers the 8-channel blocks that hold the tile's non-zero channels.  A is
// constant and stored in fragment order; B is read straight from the window's
// voltage staged in LDS as [sample][channel] fp16, where the 8 channels of a
// block at one sample are 16 contiguous, 16-byte aligned bytes.
//
// |approximation - exact FP32 chain| <= delta(t, window), a bound built from
// the template L1 norms and the window's channel maxima.  Only likelihoods
// above 0, or their sign, ever decide anything (window maxima start at 0,
// flags and skips test > 0 or <= 0, the parse cascade rejects values <= 0), so
// a 16-position tile of a (window, template) pair whose every approximation is
// <= -delta keeps the approximations (all negative, as are the exact values),
// and every other tile is recomputed by the same block with the exact FP32
// chain of k_likelihood.
#include <math.h>

#include <algorithm>

#include "fbp_device.h"

namespace fbp {

void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMfWin = 4;      // windows per block iteration
constexpr int kMfMaxBlk = 32;  // 8-channel blocks per tile at most (C <= 256)
constexpr int kMfThreads = 256;

struct MfTables {
    const half8* frag;  // A fragments: tile, k-block, lane
    const int* units;   // [ntiles * 16] unit of each tile row (-1: padding)
    const int* nkb;     // [ntiles] k-blocks of 32
    const int* nblk;    // [ntiles] 8-channel blocks
    const int* aoff;    // [ntiles] first k-block of the tile in frag
    const int* chb;     // [ntiles] first 8-channel block (the tile's blocks are contiguous)
    const int* nunit;   // [ntiles] templates of the tile (its first rows)
    const float* T1;    // [U] sum of |T| over the template (rounded up)
    const int* tileof;  // [U] tile holding each template
    int ntiles;
    float inv_tscale;   // 1 / template scale (a power of two)
    float kl, kb, ka;   // delta = kl * Q + kb * |bias| + ka * (T1 / vs + L * nch * max|V| / ts)
};

// Largest |V| of the chunk as float bits (non-negative floats order as integers).
__global__ __launch_bounds__(kBlock) void k_absmax(const float* __restrict__ V, int64_t n, uint32_t* __restrict__ out)
{
    float m = 0.0f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(V[i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (lane_id() == 0 && m > 0.0f) atomicMax(out, __float_as_uint(m));
}

// Voltage scale: the power of two that maps the largest |V| into [2^13, 2^14).
__global__ void k_vscale(const uint32_t* __restrict__ mbits, float* __restrict__ vs)
{
    const float m = __uint_as_float(*mbits);
    int e = 0;
    if (m > 0.0f && isfinite(m)) frexpf(m, &e);
    *vs = (m > 0.0f && isfinite(m)) ? ldexpf(1.0f, 14 - e) : 1.0f;
}

// (value, position) -> 64-bit key whose maximum is the largest value, then the
// smallest position.
__device__ __forceinline__ unsigned long long max_key(float v, uint32_t p)
{
    uint32_t b = __float_as_uint(v);
    b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ((unsigned long long)b << 32) | (0xffffffffu - p);
}
__device__ __forceinline__ void key_max(unsigned long long key, float& v, uint32_t& p)
{
    uint32_t b = (uint32_t)(key >> 32);
    b = (b & 0x80000000u) ? (b & 0x7fffffffu) : ~b;
    v = __uint_as_float(b);
    p = 0xffffffffu - (uint32_t)(key & 0xffffffffu);
}

template <int NP, int LT>
__global__ __launch_bounds__(kMfThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_likmfma(TemplateSet ts, PassState st, MfTables mt,
                                                        const float* __restrict__ rowL1, const float* __restrict__ vsp,
                                                        const uint32_t* __restrict__ mbits,
                                                        const uint64_t* __restrict__ dirty, uint32_t w0, uint32_t nb)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ uint32_t nit_s[kMfWin];
    __shared__ int ntact_s;
    __shared__ int64_t wid_s[kMfWin];
    constexpr int KW = kMfWin * NP;  // accumulator tiles per wave
    const int C = ts.C, U = ts.U, L = ts.L, R = ts.R;
    const int Cp = (C + 7) & ~7, RS = Cp + 8;  // halfs per staged sample (8 spare: bank spread)
    const int NS = NP * 16 + L - 1;            // samples read per window
    const int NSP = (NS + 15) & ~15;           // staged samples per window (16-sample blocks)
    const int NB = NSP / 16;
    const int64_t N = st.N;
    constexpr int NGX = (NP * 16 + kPos - 1) / kPos;
    _Float16* Vh = reinterpret_cast<_Float16*>(smem);                        // [kMfWin][NSP][RS]
    const size_t region = max((size_t)kMfWin * NSP * RS * 2,
                              (size_t)kMfWin * U * (NGX * 4 + 8) + (size_t)C * ((2 * L + kPos + 4) & ~3) * 4);
    float* bmx = reinterpret_cast<float*>(smem + region);                   // [kMfWin][NB][Cp] block max |v|
    float* Qs = bmx + kMfWin * NB * Cp;                                      // [kMfWin][U][NP] sum rowL1 * vmax
    uint32_t* flag = reinterpret_cast<uint32_t*>(Qs + kMfWin * U * NP);     // [kMfWin][U] bit p: tile p rescored
    int* rch = reinterpret_cast<int*>(flag + kMfWin * U);                   // [R] channel of each compact row
    float* rl1 = reinterpret_cast<float*>(rch + R);                          // [R] its L1 norm
    int* tlist = reinterpret_cast<int*>(rl1 + R);                            // [ntiles] tiles computed this group
    int* tact = tlist + mt.ntiles;                                           // [ntiles] 1: computed this group
    const int lane = lane_id(), wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int q4 = lane >> 4, col = lane & 15;
    const float vs = *vsp;
    const float inv = mt.inv_tscale / vs;
    const float ivs = 1.0f / vs;
    const float gmax = __uint_as_float(*mbits) * (1.0f + 0x1p-20f);  // largest |V| of the chunk
    for (int e = threadIdx.x; e < R; e += blockDim.x) {
        rch[e] = ts.chan[e];
        rl1[e] = rowL1[e];
    }
    const uint32_t ngroups = (nb + kMfWin - 1) / kMfWin;
    for (uint32_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        __syncthreads();
        if (threadIdx.x < kMfWin) {
            const uint32_t w = gi * kMfWin + threadIdx.x;
            wid_s[threadIdx.x] = w < nb ? (int64_t)st.W[w0 + w] : -1;
        }
        for (int e = threadIdx.x; e < kMfWin * U; e += blockDim.x) flag[e] = 0u;
        __syncthreads();
        // tiles to compute: those whose 8-channel blocks meet a channel that
        // changed in one of the group's windows since it was last computed (all
        // tiles without a dirty map).  The other tiles' templates would come out
        // identical: their LIK, flags and window maxima are left as they are.
        if (threadIdx.x == 0) {
            uint64_t gm = 0ull;
            for (int k = 0; k < kMfWin; k++)
                if (wid_s[k] >= 0) gm |= dirty ? dirty[wid_s[k]] : ~0ull;
            int na = 0;
            for (int tl = 0; tl < mt.ntiles; tl++) {
                int on = 1;
                if (dirty != nullptr) {  // C <= 64 here
                    const int c0 = mt.chb[tl] * 8, c1 = min(C, c0 + mt.nblk[tl] * 8);
                    const uint64_t tm = (c1 - c0 >= 64) ? ~0ull : (((1ull << (c1 - c0)) - 1ull) << c0);
                    on = (tm & gm) != 0ull ? 1 : 0;
                }
                tact[tl] = on;
                if (on) tlist[na++] = tl;
            }
            ntact_s = na;
            // (window, template) pairs computed, for the statistics
            int nwv_ok = 0, units_on = 0;
            for (int k = 0; k < kMfWin; k++) nwv_ok += wid_s[k] >= 0 ? 1 : 0;
            for (int i = 0; i < na; i++) units_on += mt.nunit[tlist[i]];
            if (na) atomicAdd(st.counters + CNT_PAIRS, (uint32_t)(units_on * nwv_ok));
        }
        __syncthreads();
        const int ntact = ntact_s;
        if (ntact == 0) continue;  // uniform over the block
        // stage: one thread per (window, sample): the sample's channels are
        // contiguous in V ([N][C]), 8 channels per 16-byte LDS store; 16
        // consecutive lanes hold one 16-sample block, so the per-channel block
        // maxima are 4 shuffle steps
        for (int col2 = threadIdx.x; col2 < kMfWin * NSP; col2 += blockDim.x) {
            const int k = col2 / NSP, j = col2 - k * NSP;
            const int64_t wd = wid_s[k];
            const int64_t x = wd * L + j;
            const bool in = wd >= 0 && j < NS && x < N;
            _Float16* dst = Vh + (size_t)col2 * RS;
            const float* src = st.V + (size_t)(in ? x : 0) * C;
            for (int c0 = 0; c0 < Cp; c0 += 32) {
                float v[32];
                if (C % 4 == 0) {
#pragma unroll
                    for (int i = 0; i < 32; i += 4) {
                        float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if (in && c0 + i < C) q = *reinterpret_cast<const float4*>(src + c0 + i);
                        v[i] = q.x; v[i + 1] = q.y; v[i + 2] = q.z; v[i + 3] = q.w;
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < 32; i++) v[i] = (in && c0 + i < C) ? src[c0 + i] : 0.0f;
                }
                __builtin_amdgcn_sched_barrier(0);  // the loads go out together
#pragma unroll
                for (int h8 = 0; h8 < 4; h8++) {
                    if (c0 + h8 * 8 >= Cp) break;
                    half8 h;
#pragma unroll
                    for (int i = 0; i < 8; i++) h[i] = (_Float16)(v[h8 * 8 + i] * vs);
                    *reinterpret_cast<half8*>(dst + c0 + h8 * 8) = h;
                }
#pragma unroll
                for (int i = 0; i < 32; i++) {
                    float m = fabsf(v[i]);
                    m = fmaxf(m, __shfl_xor(m, 1, 64));
                    m = fmaxf(m, __shfl_xor(m, 2, 64));
                    m = fmaxf(m, __shfl_xor(m, 4, 64));
                    m = fmaxf(m, __shfl_xor(m, 8, 64));
                    if (col == 0 && c0 + i < Cp) bmx[((size_t)k * NB + j / 16) * Cp + c0 + i] = m;
                }
            }
        }
        __syncthreads();
        // per (window, template): Q[p] = sum rowL1 * max|v| over the samples
        // that 16-position tile p reads (blocks p .. p + (15 + L - 1) / 16)
        for (int e = threadIdx.x; e < kMfWin * U; e += blockDim.x) {
            const int k = e / U, t = e - k * U;
            if (!tact[mt.tileof[t]]) continue;
            const int r0 = ts.off[t], nr = ts.nch[t];
            float q[NP];
#pragma unroll
            for (int p = 0; p < NP; p++) q[p] = 0.0f;
            const float* bk = bmx + (size_t)k * NB * Cp;
            for (int i = 0; i < nr; i++) {
                const int c = rch[r0 + i];
                const float w1 = rl1[r0 + i];
                float bm[NP + 4];
#pragma unroll
                for (int bq = 0; bq < NP + 4; bq++) bm[bq] = bq < NB ? bk[bq * Cp + c] : 0.0f;
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const int bhi = min(NB - 1, (p * 16 + 15 + L - 1) / 16);
                    float vm = 0.0f;
#pragma unroll
                    for (int bq = 0; bq < NP + 4; bq++)
                        if (bq >= p && bq <= bhi) vm = fmaxf(vm, bm[bq]);
                    q[p] += w1 * vm;
                }
            }
#pragma unroll
            for (int p = 0; p < NP; p++) Qs[e * NP + p] = q[p];
        }
        __syncthreads();
        for (int ti = wv; ti < ntact; ti += nwv) {
            const int tile = tlist[ti];
            const int nkb = mt.nkb[tile], nbt = mt.nblk[tile];
            const half8* Af = mt.frag + (size_t)mt.aoff[tile] * 64 + lane;
            const _Float16* vbase = Vh + col * RS + mt.chb[tile] * 8;
            floatx4 acc[KW];
#pragma unroll
            for (int i = 0; i < KW; i++) acc[i] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            // run of this lane's 8 k-values in k-block kb: r = 4 kb + q4 = (j, b)
            // with r = j * nbt + b; blocks are contiguous from chb
            int rj = q4 / nbt, rb = q4 - rj * nbt;
            auto boff = [&](int j, int b) -> int { return j < L ? j * RS + b * 8 : 0; };
            auto step = [&]() {
                rb += 4;
                while (rb >= nbt) { rb -= nbt; rj++; }
            };
            // B fragments of the current k-block loaded, the next in flight;
            // A fragments four k-blocks ahead
            half8 bc[KW], bn[KW];
            {
                const int o = boff(rj, rb);
#pragma unroll
                for (int k = 0; k < kMfWin; k++)
#pragma unroll
                    for (int p = 0; p < NP; p++)
                        bc[k * NP + p] = *reinterpret_cast<const half8*>(vbase + o + (k * NSP + p * 16) * RS);
            }
            half8 acur[4], anxt[4];
#pragma unroll
            for (int q = 0; q < 4; q++) acur[q] = Af[(size_t)min(q, nkb - 1) * 64];
            for (int kb0 = 0; kb0 < nkb; kb0 += 4) {
#pragma unroll
                for (int q = 0; q < 4; q++) anxt[q] = Af[(size_t)min(kb0 + 4 + q, nkb - 1) * 64];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (kb0 + q >= nkb) break;
                    step();
                    const int o = boff(rj, rb);
#pragma unroll
                    for (int k = 0; k < kMfWin; k++)
#pragma unroll
                        for (int p = 0; p < NP; p++)
                            bn[k * NP + p] = *reinterpret_cast<const half8*>(vbase + o + (k * NSP + p * 16) * RS);
                    __builtin_amdgcn_sched_barrier(0);  // next fragments in flight during these MFMAs
#pragma unroll
                    for (int i = 0; i < KW; i++)
                        acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(acur[q], bc[i], acc[i], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < KW; i++) bc[i] = bn[i];
                }
#pragma unroll
                for (int q = 0; q < 4; q++) acur[q] = anxt[q];
            }
        // lane holds rows 4*q4 + i (templates), column col (position)
            int unit[4];
            float bt[4], ka_t[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                unit[i] = mt.units[tile * 16 + q4 * 4 + i];
                bt[i] = unit[i] >= 0 ? ts.bias[unit[i]] : 0.0f;
                // sub-normal fp16 residue: sum |T| / vs + L * (channels * max|V|) / ts
                ka_t[i] = unit[i] >= 0 ? mt.ka * (mt.T1[unit[i]] * ivs +
                                                  (float)L * (float)ts.nch[unit[i]] * gmax * mt.inv_tscale)
                                       : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < kMfWin; k++) {
                const int64_t wd = wid_s[k];
                if (wd < 0) continue;  // uniform over the block
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const int pos = p * 16 + col;
                    const int64_t x = wd * L + pos;
                    const bool inpos = pos < L && x < N;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int t = unit[i];
                        const bool live = inpos && t >= 0;
                        float v = 0.0f;
                        bool need = false;
                        if (live && x + L < N) {
                            v = bt[i] + acc[k * NP + p][i] * inv;
                            const int qi = (k * U + t) * NP + p;
                            const float d = mt.kl * Qs[qi] + mt.kb * fabsf(bt[i]) + ka_t[i];
                            need = v > -d;
                        }
                        if (live) st.LIK[(size_t)t * N + x] = v;
                        // the 16 lanes of a row group share (template, tile): one LDS OR per group
                        const uint64_t bm = __ballot(need);
                        if (col == 0 && ((bm >> (16 * q4)) & 0xffffull) != 0ull)
                            atomicOr(&flag[k * U + t], 1u << p);
                    }
                }
            }
        }
        __syncthreads();
        // exact FP32 pass over the flagged tiles of this group's windows, with
        // the operation order of k_likelihood (bias, the template's rows in
        // channel order, taps ascending, 8 positions per item), one window at a
        // time over its FP32 voltage rows staged in LDS as k_likelihood stages
        // them (the fp16 windows are dead by now).
        // Window maxima of flagged pairs cover their flagged groups; the other
        // pairs' maxima are negative, which the window maximum treats as none.
        constexpr int NG = (NP * 16 + kPos - 1) / kPos;  // 8-position groups per window (upper bound)
        const int ngr = (L + kPos - 1) / kPos;
        const int sv = (2 * L + kPos + 4) & ~3;  // a multiple of 4: rows are read as float4
        uint32_t* ilist = reinterpret_cast<uint32_t*>(Vh);  // [kMfWin][U * NG] items (template, group) per window
        unsigned long long* keys = reinterpret_cast<unsigned long long*>(ilist + kMfWin * U * NG);  // [kMfWin * U]
        float* sVf = reinterpret_cast<float*>(keys + kMfWin * U);  // [C][sv] one window's rows, 16-byte aligned
        if (threadIdx.x < kMfWin) nit_s[threadIdx.x] = 0u;
        for (int e = threadIdx.x; e < kMfWin * U; e += blockDim.x) keys[e] = 0ull;
        __syncthreads();
        for (int e = threadIdx.x; e < kMfWin * U * ngr; e += blockDim.x) {
            const int k = e / (U * ngr);
            const int rem = e - k * U * ngr;
            const int t = rem / ngr, g = rem - t * ngr;
            if (wid_s[k] < 0) continue;
            if ((flag[k * U + t] >> ((g * kPos) / 16)) & 1u)
                ilist[k * U * NG + atomicAdd(&nit_s[k], 1u)] = (uint32_t)rem;
        }
        __syncthreads();
        uint32_t nall = 0;
#pragma unroll 1
        for (int k = 0; k < kMfWin; k++) {
            const int nit = (int)nit_s[k];
            const int64_t wd = wid_s[k];
            nall += (uint32_t)nit;
            if (wd < 0 || nit == 0) continue;  // uniform over the block
            for (int e = threadIdx.x; e < C * sv; e += blockDim.x) {
                const int j = e / C, c = e - j * C;
                const int64_t x = wd * L + j;
                sVf[c * sv + j] = x < N ? st.V[(size_t)x * C + c] : 0.0f;
            }
            __syncthreads();
            for (int ii = threadIdx.x; ii < nit; ii += blockDim.x) {
                const int rem = (int)ilist[k * U * NG + ii];
                const int t = rem / ngr, g = rem - t * ngr;
                const int p0 = g * kPos;
                const int64_t x0 = wd * L + p0;
                float acc[kPos];
                const float b0 = ts.bias[t];
#pragma unroll
                for (int q = 0; q < kPos; q++) acc[q] = b0;
                const int r0 = ts.off[t], nr = ts.nch[t];
                for (int r = 0; r < nr; r++) {
                    const float* trow = ts.rows + (size_t)(r0 + r) * ts.Lp;
                    const float* vrow = sVf + rch[r0 + r] * sv + p0;
                    if (LT > 0) lik_row_fixed_a4<(LT > 0 ? LT : 4)>(trow, vrow, acc);
                    else lik_row_generic(trow, vrow, L, acc);
                }
                float* out = st.LIK + (size_t)t * N;
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
                if (mv != -INFINITY) atomicMax(&keys[k * U + t], max_key(mv, mp));
            }
            __syncthreads();  // the rows are staged again for the next window
        }
        if (threadIdx.x == 0 && nall) atomicAdd(st.counters + CNT_LIKX, nall * kPos);
        __syncthreads();
        for (int e = threadIdx.x; e < kMfWin * U; e += blockDim.x) {
            const int k = e / U, t = e - k * U;
            const int64_t wd = wid_s[k];
            if (wd < 0 || !tact[mt.tileof[t]]) continue;
            float mv = -INFINITY;
            uint32_t mp = 0xffffu;
            if (flag[e] != 0u && keys[e] != 0ull) key_max(keys[e], mv, mp);
            st.wmax_v[(size_t)wd * U + t] = mv;
            st.wmax_p[(size_t)wd * U + t] = (uint16_t)mp;
        }
    }
}

// ---------------------------------------------------------------------------
// Overlap-screen correlations on the matrix cores:
//     B(u, s) = sum_{c, i} T[u][c][i] * w_c[s + i - n] - Ntr(u, s) / 2
// for every unit u and shift s < S of the flagged windows F[f0 .. f0+nbatch),
// where w is the window's voltage (samples b .. b+L-1, zero elsewhere) and Ntr
// the energy of the window-truncated template (k_screen_corr computes the same
// values on the vector units).  Both operands are split into fp16 hi + lo parts
// (templates scaled by ts, each window by a power of two vs_w with
// max |v| vs_w in [2^13, 2^14)); the products hi*hi and hi*lo + lo*hi per
// k-block accumulate in FP32 in two chains, which leaves about 2^-22 relative
// of the operands' rounding plus the accumulation error (the screen bound's
// kappa2) and an absolute residue of the sub-normal parts (kappa3, kappa4).
// Rows of a window absent from it read a zero row of LDS; a k-block whose
// 16-shift tile reads only absent rows is skipped.
// ---------------------------------------------------------------------------
constexpr int kSwWin = 2;  // flagged windows per block iteration

template <int NPS>
__global__ __launch_bounds__(kMfThreads) __attribute__((amdgpu_waves_per_eu(NPS <= 6 ? 3 : 2, NPS <= 6 ? 3 : 2))) void k_screen_mfma(
    TemplateSet ts, PassState st, MfTables mt, const half8* __restrict__ fragl, const float* __restrict__ Ntr, int S,
    int n, uint32_t f0, uint32_t nbatch, float* __restrict__ Bbuf)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ int64_t b_s[kSwWin];
    __shared__ float vs_s[kSwWin];
    __shared__ float red[kSwWin * (kMfThreads / 64)];
    const int C = ts.C, U = ts.U, L = ts.L;
    const int Cp = (C + 7) & ~7, RS = Cp + 8;
    const int LR = L + 1;  // window rows + one zero row
    const int US = U * S;
    _Float16* Wh = reinterpret_cast<_Float16*>(smem);  // [kSwWin][LR][RS]
    _Float16* Wl = Wh + kSwWin * LR * RS;              // [kSwWin][LR][RS]
    // the wave index is uniform: tile, nkb, nbt and the k-block skips stay scalar
    const int lane = lane_id(), wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwv = blockDim.x >> 6;
    const int q4 = lane >> 4, col = lane & 15;
    const uint32_t nF = st.counters[CNT_F];
    const uint32_t ngroups = (nbatch + kSwWin - 1) / kSwWin;
    for (uint32_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        __syncthreads();
        if (threadIdx.x < kSwWin) {
            const uint32_t wi = gi * kSwWin + threadIdx.x, w = f0 + wi;
            int64_t b = -1;
            if (wi < nbatch && w < nF) {
                const uint32_t id = st.F[w];
                if (!window_skipped(ts, st, id)) b = st.best_idx[id];
            }
            b_s[threadIdx.x] = b;
        }
        __syncthreads();
        // window maxima -> per-window scale, both windows in one pass (float4 loads when
        // the channels come in fours)
        {
            float m[kSwWin];
#pragma unroll
            for (int k = 0; k < kSwWin; k++) {
                const int64_t b = b_s[k];
                m[k] = 0.0f;
                if (b < 0) continue;
                if (C % 4 == 0) {
                    const float4* src = reinterpret_cast<const float4*>(st.V + (size_t)b * C);
                    for (int e = threadIdx.x; e < L * C / 4; e += blockDim.x) {
                        const float4 q = src[e];
                        m[k] = fmaxf(fmaxf(m[k], fmaxf(fabsf(q.x), fabsf(q.y))), fmaxf(fabsf(q.z), fabsf(q.w)));
                    }
                } else {
                    for (int e = threadIdx.x; e < L * C; e += blockDim.x)
                        m[k] = fmaxf(m[k], fabsf(st.V[(size_t)b * C + e]));
                }
            }
#pragma unroll
            for (int k = 0; k < kSwWin; k++) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) m[k] = fmaxf(m[k], __shfl_xor(m[k], off, 64));
                if (lane == 0) red[k * nwv + wv] = m[k];
            }
            __syncthreads();
            if (threadIdx.x < kSwWin) {
                const int k = threadIdx.x;
                float mm = red[k * nwv];
                for (int i = 1; i < nwv; i++) mm = fmaxf(mm, red[k * nwv + i]);
                int ex = 0;
                if (mm > 0.0f) frexpf(mm, &ex);
                vs_s[k] = mm > 0.0f ? ldexpf(1.0f, 14 - ex) : 1.0f;
            }
            __syncthreads();
        }
        // stage the windows as split fp16 rows [sample][channel]; row L is zero
        if (C % 8 == 0) {
            // eight channels per item: two float4 loads, two 16-byte LDS stores (Cp == C)
            const int C8 = C / 8;
            for (int e = threadIdx.x; e < kSwWin * LR * C8; e += blockDim.x) {
                const int k = e / (LR * C8);
                const int rem = e - k * LR * C8;
                const int j = rem / C8, g = rem - j * C8;
                const int64_t b = b_s[k];
                float v[8];
#pragma unroll
                for (int i = 0; i < 8; i++) v[i] = 0.0f;
                if (b >= 0 && j < L) {
                    const float4* src = reinterpret_cast<const float4*>(st.V + (size_t)(b + j) * C + 8 * g);
                    const float4 q0 = src[0], q1 = src[1];
                    const float sc = vs_s[k];
                    v[0] = q0.x * sc; v[1] = q0.y * sc; v[2] = q0.z * sc; v[3] = q0.w * sc;
                    v[4] = q1.x * sc; v[5] = q1.y * sc; v[6] = q1.z * sc; v[7] = q1.w * sc;
                }
                half8 hh, hl;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const _Float16 hi = (_Float16)v[i];
                    hh[i] = hi;
                    hl[i] = (_Float16)(v[i] - (float)hi);
                }
                *reinterpret_cast<half8*>(Wh + (k * LR + j) * RS + 8 * g) = hh;
                *reinterpret_cast<half8*>(Wl + (k * LR + j) * RS + 8 * g) = hl;
            }
        } else {
            for (int e = threadIdx.x; e < kSwWin * LR * Cp; e += blockDim.x) {
                const int k = e / (LR * Cp);
                const int rem = e - k * LR * Cp;
                const int j = rem / Cp, c = rem - j * Cp;
                const int64_t b = b_s[k];
                float v = 0.0f;
                if (b >= 0 && j < L && c < C) v = st.V[(size_t)(b + j) * C + c] * vs_s[k];
                const _Float16 hi = (_Float16)v;
                Wh[(k * LR + j) * RS + c] = hi;
                Wl[(k * LR + j) * RS + c] = (_Float16)(v - (float)hi);
            }
        }
        __syncthreads();
        for (int tile = wv; tile < mt.ntiles; tile += nwv) {
            const int nkb = mt.nkb[tile], nbt = mt.nblk[tile];
            const half8* Ah = mt.frag + (size_t)mt.aoff[tile] * 64 + lane;
            const half8* Al = fragl + (size_t)mt.aoff[tile] * 64 + lane;
            const int cb = mt.chb[tile] * 8;
            // hi*hi and the two cross products accumulate apart: the long chain
            // then holds one product per template value, not three
            floatx4 acc[kSwWin][NPS], acs[kSwWin][NPS];
#pragma unroll
            for (int k = 0; k < kSwWin; k++)
#pragma unroll
                for (int p = 0; p < NPS; p++) {
                    acc[k][p] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                    acs[k][p] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                }
            // (i, bk) = (template sample, channel block) of the lane's k-group r = 4 kb + q4,
            // advanced by 4 groups per k-block (nbt is uniform: one division per tile)
            int i = q4 / nbt, bk = q4 - (q4 / nbt) * nbt;
            auto kblock = [&](const half8 ah, const half8 al) __attribute__((always_inline)) {
                const bool pad = i >= L;  // padding run: its A values are zero
                // template samples of the k-block (uniform): lanes 0 (q4 = 0) and 48 (q4 = 3)
                const int ilo = __builtin_amdgcn_readlane(i, 0);
                const int ihi = min(L - 1, __builtin_amdgcn_readlane(i, 48));
                // row offset of window sample j = s + i - n at the tile's first shift s = col;
                // a padding run reads the zero row at every shift (offset far below zero)
                const int jo = pad ? INT_MIN / 2 : (col + i - n) * RS;
                const int ob = cb + (pad ? 0 : bk) * 8;
#pragma unroll
                for (int p = 0; p < NPS; p++) {
                    // window samples j = s + i - n over the tile's shifts s
                    const int jlo = 16 * p + ilo - n, jhi = 16 * p + 15 + ihi - n;
                    if (jhi < 0 || jlo >= L) continue;  // every row absent: all-zero B
                    // 0 <= j < L  <=>  0 <= j * RS < L * RS: one unsigned compare
                    const int jr = jo + 16 * p * RS;
                    const int ro = (unsigned)jr < (unsigned)(L * RS) ? jr : L * RS;
                    // both windows' fragments in flight before the first products
                    half8 bh[kSwWin], bl[kSwWin];
#pragma unroll
                    for (int k = 0; k < kSwWin; k++) {
                        const int o = ro + k * LR * RS + ob;
                        bh[k] = *reinterpret_cast<const half8*>(Wh + o);
                        bl[k] = *reinterpret_cast<const half8*>(Wl + o);
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k = 0; k < kSwWin; k++) {
                        acc[k][p] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[k], acc[k][p], 0, 0, 0);
                        acs[k][p] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[k], acs[k][p], 0, 0, 0);
                        acs[k][p] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[k], acs[k][p], 0, 0, 0);
                    }
                }
                bk += 4;
                while (bk >= nbt) {
                    bk -= nbt;
                    ++i;
                }
            };
            // A fragments double-buffered two k-blocks ahead; loads past the end repeat the
            // last k-block (in bounds, unused)
            const int last = max(nkb - 1, 0);
            half8 a0h = Ah[0], a0l = Al[0];
            half8 a1h = Ah[(size_t)min(1, last) * 64], a1l = Al[(size_t)min(1, last) * 64];
            for (int kb = 0; kb < nkb; kb += 2) {
                kblock(a0h, a0l);
                a0h = Ah[(size_t)min(kb + 2, last) * 64];
                a0l = Al[(size_t)min(kb + 2, last) * 64];
                if (kb + 1 < nkb) kblock(a1h, a1l);
                a1h = Ah[(size_t)min(kb + 3, last) * 64];
                a1l = Al[(size_t)min(kb + 3, last) * 64];
            }
            // lane holds rows 4*q4 + i (units), column col (shift within the tile)
#pragma unroll
            for (int k = 0; k < kSwWin; k++) {
                if (b_s[k] < 0) continue;  // uniform
                const float sc = mt.inv_tscale / vs_s[k];
                float* Bw = Bbuf + (size_t)(gi * kSwWin + k) * US;
#pragma unroll
                for (int p = 0; p < NPS; p++) {
                    const int sft = 16 * p + col;
#pragma unroll
                    for (int ii = 0; ii < 4; ii++) {
                        const int u = mt.units[tile * 16 + q4 * 4 + ii];
                        if (u < 0 || sft >= S) continue;
                        const int e = u * S + sft;
                        Bw[e] = fmaf(-0.5f, Ntr[e], (acc[k][p][ii] + acs[k][p][ii]) * sc);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
size_t likmfma_lds(const TemplateSet& ts, int ntiles)
{
    const int Cp = (ts.C + 7) & ~7, RS = Cp + 8;
    const int NP = (ts.L + 15) / 16;
    const int NS = NP * 16 + ts.L - 1;
    const int NSP = (NS + 15) & ~15, NB = NSP / 16;
    // the staged windows' region later holds the exact pass's item lists, keys
    // and one window's FP32 voltage rows
    const size_t NG = (NP * 16 + kPos - 1) / kPos;
    const size_t region = std::max((size_t)kMfWin * NSP * RS * 2,
                                   (size_t)kMfWin * ts.U * (NG * 4 + 8) + (size_t)ts.C * ((2 * ts.L + kPos + 4) & ~3) * 4);
    return region + (size_t)kMfWin * Cp * NB * 4 + (size_t)kMfWin * ts.U * NP * 4 +
           (size_t)kMfWin * ts.U * 4 + (size_t)ts.R * 8 + (size_t)ntiles * 8;
}

void launch_absmax(const float* V, int64_t n, uint32_t* mbits, float* vs, hipStream_t s)
{
    launch_fill_u32(mbits, 0u, 1, s);
    int64_t g = (n + kBlock * 8 - 1) / (kBlock * 8);
    g = std::max<int64_t>(1, std::min<int64_t>(g, 8192));
    hipLaunchKernelGGL(k_absmax, dim3((unsigned)g), dim3(kBlock), 0, s, V, n, mbits);
    hipLaunchKernelGGL(k_vscale, dim3(1), dim3(1), 0, s, mbits, vs);
}

template <class K>
static void allow_lds_mf(K kernel, size_t bytes)
{
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

template <int NP, int LT>
static void launch_mf(const TemplateSet& ts, const PassState& st, const MfTables& mt, const float* rowL1,
                      const float* vs, const uint32_t* mbits, const uint64_t* dirty, uint32_t w0, uint32_t nb,
                      size_t lds, dim3 g, hipStream_t s)
{
    allow_lds_mf(k_likmfma<NP, LT>, lds);
    hipLaunchKernelGGL((k_likmfma<NP, LT>), g, dim3(kMfThreads), lds, s, ts, st, mt, rowL1, vs, mbits, dirty, w0, nb);
}

// MFMA likelihood of the listed windows W[w0 .. w0+nb), with the exact FP32
// chain for the 16-position tiles the bound flags.
void launch_likmfma(const TemplateSet& ts, const PassState& st, const MfTables& mt, const float* rowL1,
                    const float* vs, const uint32_t* mbits, const uint64_t* dirty, uint32_t w0, uint32_t nb,
                    hipStream_t s)
{
    if (nb == 0) return;
    const size_t lds = likmfma_lds(ts, mt.ntiles);
    const uint32_t ngroups = (nb + kMfWin - 1) / kMfWin;
    const dim3 g(std::min<uint32_t>(ngroups, 4096u));
    const int NP = (ts.L + 15) / 16;
    if (ts.L == 48) launch_mf<3, 48>(ts, st, mt, rowL1, vs, mbits, dirty, w0, nb, lds, g, s);
    else if (ts.L == 64) launch_mf<4, 64>(ts, st, mt, rowL1, vs, mbits, dirty, w0, nb, lds, g, s);
    else if (ts.L == 32) launch_mf<2, 32>(ts, st, mt, rowL1, vs, mbits, dirty, w0, nb, lds, g, s);
    else if (NP == 4) launch_mf<4, 0>(ts, st, mt, rowL1, vs, mbits, dirty, w0, nb, lds, g, s);
    else if (NP == 3) launch_mf<3, 0>(ts, st, mt, rowL1, vs, mbits, dirty, w0, nb, lds, g, s);
    else if (NP == 2) launch_mf<2, 0>(ts, st, mt, rowL1, vs, mbits, dirty, w0, nb, lds, g, s);
    else launch_mf<1, 0>(ts, st, mt, rowL1, vs, mbits, dirty, w0, nb, lds, g, s);
}

size_t screen_mfma_lds(const TemplateSet& ts)
{
    const int Cp = (ts.C + 7) & ~7, RS = Cp + 8;
    return (size_t)2 * kSwWin * (ts.L + 1) * RS * 2;
}

// B(u, s) of the flagged windows F[f0 .. f0+nbatch) into Bbuf[w][u * S + s].
void launch_screen_mfma(const TemplateSet& ts, const PassState& st, const MfTables& mt, const void* fragl,
                        const float* Ntr, int S, int n, uint32_t f0, uint32_t nbatch, float* Bbuf, hipStream_t s)
{
    if (nbatch == 0) return;
    const size_t lds = screen_mfma_lds(ts);
    const uint32_t ngroups = (nbatch + kSwWin - 1) / kSwWin;
    const dim3 g(std::min<uint32_t>(ngroups, 8192u)), b(kMfThreads);
    const half8* fl = reinterpret_cast<const half8*>(fragl);
    const int NPS = (S + 15) / 16;
#define FBP_SCREEN_MFMA(NPV)                                                                                      \
    do {                                                                                                        \
        allow_lds_mf(k_screen_mfma<NPV>, lds);                                                                  \
        hipLaunchKernelGGL(k_screen_mfma<NPV>, g, b, lds, s, ts, st, mt, fl, Ntr, S, n, f0, nbatch, Bbuf);      \
    } while (0)
    switch (NPS) {
    case 1: FBP_SCREEN_MFMA(1); break;
    case 2: FBP_SCREEN_MFMA(2); break;
    case 3: FBP_SCREEN_MFMA(3); break;
    case 4: FBP_SCREEN_MFMA(4); break;
    case 5: FBP_SCREEN_MFMA(5); break;
    case 6: FBP_SCREEN_MFMA(6); break;
    case 7: FBP_SCREEN_MFMA(7); break;
    default: FBP_SCREEN_MFMA(8); break;
    }
#undef FBP_SCREEN_MFMA
}

}  // namespace fbp
