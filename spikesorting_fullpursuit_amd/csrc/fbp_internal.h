// Internal declarations shared by the HIP kernels and the C-ABI host loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <tuple>
#include <utility>
#include <vector>

namespace fbp {

constexpr int kBlock = 256;          // threads per block for the streaming kernels
constexpr int kCompactTile = 1024;   // elements per block in ordered compaction
constexpr uint32_t kNoItem = 0xffffffffu;
constexpr int kGB = 8;                 // shifts f per block of the blocked Gram minima
constexpr int kXStride = 65;           // u32 per flagged window of the deferred-rescoring record

// Device view of the template set.  Rows of the compact layout are the
// non-zero channels of each template, ascending, zero padded to Lp floats.
struct TemplateSet {
    const float* T;        // [U][C][L] dense
    const float* rows;     // [R][Lp] non-zero channel rows
    const int* nch;        // [U] non-zero channels per template
    const int* off;        // [U] first row of template u
    const int* chan;       // [R] channel of row r
    const uint8_t* nz;     // [U][C] 1 where the template row is not all zero
    const float* bias;     // [U]
    const float* thr;      // [U]
    const uint64_t* cmask; // [U] bit c: template row c is not all zero (C <= 64)
    int U, C, L, Lp;
    int R;                 // rows of the compact layout
};

// Per-chunk pursuit state (all device pointers).
struct PassState {
    float* V;              // [N][C] residual, sample-major (V[x * C + c])
    float* LIK;            // [U][N]
    float* wmax_v;         // [nwin][U] maximum of LIK over each window
    uint16_t* wmax_p;      // [nwin][U] first position of that maximum in the window
    uint32_t* best_idx;    // [nwin]
    uint32_t* best_lab;
    float* best_lik;
    uint32_t* obest_idx;
    uint32_t* obest_lab;
    uint8_t* check;        // [nwin] check_window_on_next_pass
    uint8_t* check_snap;   // [nwin]
    uint8_t* orc;          // [nwin] overlap_recheck
    uint8_t* giveup;       // [nwin] set by the reassignment step
    uint8_t* has;          // [nwin] per listed window: spike added this pass
    uint32_t* W;           // window list
    uint32_t* F;           // flagged (overlap) list
    uint32_t* cand_idx;    // [nwin] per listed window: candidate spike
    uint32_t* cand_lab;
    uint32_t* counters;    // [8] device counters (see Counter)
    // [nwin] channels (bit c) whose residual changed, since the window's
    // likelihood was last computed, somewhere in the samples its positions
    // read ([wL, wL + 2L - 1)).  A template whose channels miss that mask would
    // be recomputed from identical inputs, i.e. to identical values, so the
    // likelihood kernels leave it as it is.  nullptr: compute every template.
    uint64_t* dirty;
    int64_t N;
    int64_t nwin;
};

// Voltage-independent tables of the screened overlap recheck.
struct ScreenTables {
    const int* pid;       // [U*U] pair index of (l, u) sharing a channel, or -1
    const int* pair_l;    // [npairs]
    const int* pair_u;    // [npairs]
    float* G;             // [npairs][S][sg] G[p][s][f]: window-truncated cross product (f >= S: 0)
    float* minG;          // [npairs][S]     min over f
    float* minGb;         // [npairs][S][nbg] min over each block of kGB consecutive f
    int nbg;              // blocks of f: ceil(S / kGB)
    int npl_max;          // most units sharing a channel with one unit
    int sel_region;       // floats of the register-resident selection's phase-dependent LDS region
    float* Ntr;           // [U][S]          energy of the window-truncated template
    // Template rows interleaved by unit for the FP32 screen correlation:
    // Til[((k * Lp / 4 + i / 4) * upad + u) * 4 + i % 4] = row k of unit u, tap i
    // (zero for k >= nch[u] or u >= U), so that lanes on consecutive units read
    // consecutive 16-byte words.
    float* Til;           // [cmax][Lp / 4][upad][4]
    int upad;             // U rounded up to a multiple of 64
    const float* rowL1;   // [R] L1 norm of each compact template row (rounded up)
    const float* SS;      // [U] sum of squares of each template (rounded up)
    int S;
    int sg;               // floats per row of G: S rounded up to a multiple of 4 (rows read as float4)
    int n;
    int cmax;             // most non-zero channels of any template
    // delta(l, u) = P_l + P_u bounds |E-hat - E_fp32| of a pair, with
    // P_x = (kappa + kappa2) Q_x + 2 kappa SS_x + kappa3 maxw T1_x + kappa4 VS_x,
    // Q_x = sum over x's channels of rowL1 * window max, T1_x = sum of rowL1,
    // VS_x = sum of the window maxima over x's channels, maxw = window maximum
    float kappa;          // FP32 chains of B, A and G
    float kappa2;         // 0 when B is the FP32 VALU correlation; > 0 for split-fp16 matrix-core B
    float kappa3, kappa4; // absolute residue of the fp16 split (0 for the FP32 B)
};

// Per-template matrix-core likelihood in BF16 or FP8 (fbp_likmma.hip).  delta =
// kl Q + kb |bias| + ka (T1 / vs + 64 S / ts) + kd / (ts vs) bounds |approximation -
// exact FP32 chain| on a 16-position row (Q = sum_c rowL1 * max|v|, S = sum_c max|v|).
struct MmaTables {
    const void* frag;   // [sum_t nch_t * KB][64 lanes]: 16 bytes (bf16 x 8) or 8 bytes (fp8 x 8) per lane
    const float* tsc;   // [U] template scale (a power of two)
    const float* T1;    // [U] sum |T| over the template (rounded up)
    const float* rowL1; // [R] L1 norm of each compact template row (rounded up)
    int KB;             // k-blocks of 32 per channel
    int R;              // rows per window: ceil(L / 16)
    int P;              // windows per tile: 16 / R
    float kl, kb, ka, kd;
    // dense first pass (k_likmma_dense): templates in quads ordered by first
    // channel; a quad's channels (union of its templates', ascending) are staged
    const int* qtmp;    // [nq * 4] template of each quad slot (-1: none)
    const int* qch;     // [nq][kQuadCh] channels of each quad (-1: none)
    const int* rslot;   // [R rows] slot of each compact template row in its quad's channel list
    int nq;             // quads (0: the dense first pass is not available)
    int qc;             // most channels of any quad
};
constexpr int kQuadCh = 16;  // channels of a template quad of the dense first pass at most

struct MmaVoltage {
    void* q;            // [C][ldq] quantized voltage, channel-major (zero beyond N)
    float* bmax;        // [C][nbq] max |v| over each 16-sample block
    const float* vs;    // voltage scale (device scalar, a power of two)
    int64_t ldq, nbq;
};

// One segment of a batched pursuit, as the segmented kernels read it.
struct SegDesc {
    TemplateSet ts;
    PassState st;
    ScreenTables tb;
    uint32_t* blockbuf;  // per-tile counts of the ordered compactions
    int wb, sv;          // likelihood: windows per block, staged voltage row length
    int rp;              // screen correlation: staged window row length
    int reach;           // residual: list entries whose samples a spike can share
};

// Per-pass values of a segment: its spike list (it grows) and the spikes
// listed before this pass (where this pass's spikes start).
struct SegDyn {
    uint32_t* sp_idx;
    uint32_t* sp_lab;
    uint64_t total;
    uint32_t nW;      // windows listed for this pass (read by k_seg_plan)
    uint32_t active;  // 1 while the segment converges (read by k_seg_plan)
};

// One entry of a segmented launch: blocks [blk0, blk0 + nblk) of the grid work
// on segment `seg`, on its n items (windows, flagged windows or spikes) from
// a0; g0 is the entry's first row in a buffer shared by the launch's entries.
struct SegLaunch {
    uint32_t seg, blk0, nblk, n, a0, g0;
};

// A pass replayed as a hipGraph: launch tables whose blocks are fixed when the
// graph is built and whose item counts a plan kernel writes on the device
// before each launch reads them.  Range r is cnt[r] entries from off[r]; src[r] says where their n
// comes from; phase[r] which plan launch of the pass writes it.
enum SegSrc { SRC_NW = 0, SRC_NWIN, SRC_NF, SRC_DECIDE, SRC_NADD };
constexpr int kPlanRanges = 16;
struct SegPlan {
    uint32_t off[kPlanRanges], cnt[kPlanRanges];
    uint8_t src[kPlanRanges], phase[kPlanRanges];
    int nr;
    int recheck;
};

// Kernel launches recorded as the nodes of a hipGraph being built, instead of
// launched on a stream: each node depends on the nodes of `tail` (the last
// launch, or every launch of a fork once it is joined).  Between fork() and
// join() the nodes all depend on the fork's base and run side by side.
struct GraphRec {
    hipGraph_t graph = nullptr;
    std::vector<hipGraphNode_t> tail, base, branch;
    bool forking = false;
    hipError_t err = hipSuccess;
    void fork()
    {
        base = tail;
        branch.clear();
        forking = true;
    }
    void join()
    {
        if (!branch.empty()) tail = branch;
        forking = false;
    }
    void add(const hipKernelNodeParams& p)
    {
        if (err != hipSuccess) return;
        const std::vector<hipGraphNode_t>& deps = forking ? base : tail;
        hipGraphNode_t n = nullptr;
        err = hipGraphAddKernelNode(&n, graph, deps.empty() ? nullptr : deps.data(), deps.size(), &p);
        if (err != hipSuccess) return;
        if (forking) {
            branch.push_back(n);
        } else {
            tail.assign(1, n);
        }
    }
};

template <class Tup, size_t... I>
inline void arg_addrs(Tup& t, void** a, std::index_sequence<I...>)
{
    ((a[I] = (void*)&std::get<I>(t)), ...);
}

// Launch kernel k on stream s, or record it in r (arguments converted to the
// kernel's parameter types; the graph copies their values).
template <class... P, class... A>
inline void launch_or_record(GraphRec* r, void (*k)(P...), dim3 g, dim3 b, size_t lds, hipStream_t s, A&&... a)
{
    if (r == nullptr) {
        hipLaunchKernelGGL(k, g, b, lds, s, std::forward<A>(a)...);
        return;
    }
    std::tuple<P...> vals(std::forward<A>(a)...);
    void* args[sizeof...(P)];
    arg_addrs(vals, args, std::index_sequence_for<P...>{});
    hipKernelNodeParams p{};
    p.func = reinterpret_cast<void*>(k);
    p.gridDim = g;
    p.blockDim = b;
    p.sharedMemBytes = (unsigned)lds;
    p.kernelParams = args;
    p.extra = nullptr;
    r->add(p);
}

// Kernels of a segmented pass launch (launch_seg_pass_kernel).
enum SegKind { SEG_LIKELIHOOD = 0, SEG_CLEAR_DIRTY, SEG_WINDOW_MAX, SEG_REASSIGN, SEG_DECIDE, SEG_SNAPSHOT, SEG_PRUNE,
               SEG_RESIDUAL, SEG_MARK_DIRTY };

enum Counter { CNT_F = 0, CNT_SPIKES = 1, CNT_W = 2, CNT_CAND = 3, CNT_EVAL = 4, CNT_LIKX = 5, CNT_SURV = 6, CNT_NSP = 7,
               CNT_PAIRS = 8, CNT_N = 9 };

}  // namespace fbp
