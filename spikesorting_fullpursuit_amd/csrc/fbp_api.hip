// C-ABI and host pass loop of the MI355X binary pursuit.
//
// Replaces the pyopencl driver of SF/overlap/binary_pursuit_parallel.py:296-1500
// (SF = spikesorting_fullpursuit/) for one chunk.  The pass loop keeps the
// window lists on the device; the host reads two small counters per pass
// (overlap windows, then spikes added and next list length).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <condition_variable>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fbp.h"
#include "fbp_internal.h"

namespace fbp {
void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
void launch_fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t s);
void launch_iota(uint32_t* p, int64_t n, hipStream_t s);
void launch_to_sample_major(const float* src, int64_t ld, float* dst, int64_t C, int64_t N, hipStream_t s);
void launch_to_channel_major(const float* src, float* dst, int64_t C, int64_t N, hipStream_t s);
bool likelihood_fits(const TemplateSet& ts);
void launch_likelihood(const TemplateSet& ts, const PassState& st, uint32_t nW, const uint64_t* dirty, hipStream_t s,
                       const uint8_t* gmask = nullptr);
void launch_mark_dirty(const TemplateSet& ts, const PassState& st, const uint32_t* sidx, const uint32_t* slab,
                       uint32_t nsp, hipStream_t s);
void launch_clear_dirty(const PassState& st, uint32_t nW, hipStream_t s);
void launch_window_max(const TemplateSet& ts, const PassState& st, uint32_t nW, hipStream_t s);
void launch_flag_compact(const PassState& st, uint32_t nW, uint32_t* blockbuf, hipStream_t s);
void launch_merge_regions(uint32_t* F, const uint32_t* Fs, const uint32_t* rcnt, int nreg, uint32_t rwin, uint32_t* cnt,
                          hipStream_t s);
void launch_sort_by_label(const PassState& st, uint32_t nF, int U, uint32_t* out, uint32_t* hist, hipStream_t s,
                          const uint32_t* order = nullptr);
void launch_spike_compact(const PassState& st, uint32_t nW, uint32_t* out_idx, uint32_t* out_lab, uint64_t base,
                          uint32_t* blockbuf, hipStream_t s);
void launch_next_windows(const PassState& st, uint32_t* blockbuf, hipStream_t s);
uint32_t recheck_literal_chunks(const TemplateSet& ts, int64_t n_max);
void launch_recheck_literal(const TemplateSet& ts, const PassState& st, int64_t n_max, uint32_t nF, float* part_v,
                            uint32_t* part_k, hipStream_t s);
void launch_reassign(const TemplateSet& ts, const PassState& st, uint32_t nF, hipStream_t s);
void launch_pursuit(const TemplateSet& ts, const PassState& st, uint32_t nW, int use_giveup, hipStream_t s);
void launch_residual(const TemplateSet& ts, float* V, int64_t N, const uint32_t* sidx, const uint32_t* slab, uint32_t nsp,
                     int reach, hipStream_t s);
void launch_adjusted_clips(const TemplateSet& ts, const float* V, int64_t N, const uint32_t* sidx, const uint32_t* slab,
                           uint64_t nsp, float* clips, hipStream_t s);
void launch_screen_tables(const TemplateSet& ts, const ScreenTables& tb, int64_t npairs, hipStream_t s);
size_t screen_corr_lds(const TemplateSet& ts, int S, int* rp);
void seg_likelihood_lds(const TemplateSet& ts, int* wb, int* sv, size_t* lds);
void launch_seg_pass_kernel(int kind, int L, const SegDesc* sd, const SegDyn* dy, const SegLaunch* tab, uint32_t ntab,
                            uint32_t nblocks, uint16_t* eob, size_t lds, int flag, hipStream_t s,
                            GraphRec* rec = nullptr);
void launch_seg_compact(int kind, const SegDesc* sd, const SegDyn* dy, const SegLaunch* tab, uint32_t ntab,
                        uint32_t nblocks, uint16_t* eob, hipStream_t s, GraphRec* rec = nullptr);
void launch_seg_fill_eob(const SegLaunch* tab, uint32_t ntab, uint16_t* eob, hipStream_t s);
void launch_seg_plan(const SegDesc* sd, const SegDyn* dy, SegLaunch* tab, const SegPlan& p, int phase, hipStream_t s,
                     GraphRec* rec = nullptr);
void launch_seg_counters(const SegDesc* sd, uint32_t nseg, uint32_t* out, hipStream_t s, GraphRec* rec = nullptr);
bool screen_select_reg_plan(const TemplateSet& ts, const ScreenTables& tb, int* bd, size_t* lds);
void launch_seg_screen_corr(int L, const SegDesc* sd, const SegLaunch* tab, uint32_t ntab, uint32_t nblocks,
                            uint16_t* eob, size_t lds, float* Bbuf, size_t bstride, hipStream_t s,
                            GraphRec* rec = nullptr);
void launch_seg_select(int L, const SegDesc* sd, const SegLaunch* tab, uint32_t ntab, uint32_t nblocks, uint16_t* eob,
                       int bd, size_t lds, const float* Bbuf, size_t bstride, hipStream_t s, GraphRec* rec = nullptr);
size_t screen_select_lds(const TemplateSet& ts, int S, int cmax);
int screen_select_region(const TemplateSet& ts, int S, int cmax, int npl_max);
void launch_screen_corr(const TemplateSet& ts, const PassState& st, const ScreenTables& tb, uint32_t f0, uint32_t nbatch,
                        float* Bbuf, hipStream_t s);
bool launch_screen_select(const TemplateSet& ts, const PassState& st, const ScreenTables& tb, uint32_t f0,
                          uint32_t nbatch, const float* Bbuf, bool streaming, hipStream_t s, uint32_t* xcand);
void launch_screen_rescore(const TemplateSet& ts, const PassState& st, const ScreenTables& tb, uint32_t f0,
                           uint32_t nbatch, const uint32_t* xcand, hipStream_t s);
// MFMA likelihood tables (fbp_likmfma.hip).
constexpr int kMfMaxBlk = 32;
constexpr double kMfPadMax = 2.0;  // tile channel rows per template channel row up to which the matrix cores pay
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
struct MfTables {
    const half8* frag;
    const int* units;
    const int* nkb;
    const int* nblk;
    const int* aoff;
    const int* chb;
    const int* nunit;
    const float* T1;
    const int* tileof;
    int ntiles;
    float inv_tscale;
    float kl, kb, ka;
};
size_t likmfma_lds(const TemplateSet& ts, int ntiles);
void launch_absmax(const float* V, int64_t n, uint32_t* mbits, float* vs, hipStream_t s);
void launch_likmfma(const TemplateSet& ts, const PassState& st, const MfTables& mt, const float* rowL1,
                    const float* vs, const uint32_t* mbits, const uint64_t* dirty, uint32_t w0, uint32_t nb,
                    hipStream_t s);
constexpr int kScreenMfmaMaxS = 128;  // shifts of the matrix-core screen (8 tiles of 16)
// Per-template BF16 / FP8 matrix-core likelihood (fbp_likmma.hip).
void launch_frag_convert(const float* src, void* dst, int64_t n8, int fmt, hipStream_t s);
void launch_quantize(const float* V, int64_t N, int C, const MmaVoltage& mv, int fmt, hipStream_t s);
void launch_requantize(const float* V, int64_t N, int C, int L, const MmaVoltage& mv, int fmt, const uint32_t* sidx,
                       uint32_t nsp, hipStream_t s);
void launch_vscale_to(const uint32_t* mbits, float* vs, int e_target, hipStream_t s);
bool launch_likmma_dense(const TemplateSet& ts, const PassState& st, const MmaTables& mt, const MmaVoltage& mv, int fmt,
                         uint8_t* gmask, hipStream_t s);
void launch_likmma(const TemplateSet& ts, const PassState& st, const MmaTables& mt, const MmaVoltage& mv, int fmt,
                   uint32_t nW, uint8_t* gmask, const uint64_t* dirty, hipStream_t s);
void launch_screen_mfma(const TemplateSet& ts, const PassState& st, const MfTables& mt, const void* fragl,
                        const float* Ntr, int S, int n, uint32_t f0, uint32_t nbatch, float* Bbuf, hipStream_t s);
}  // namespace fbp

// Limits of the screened recheck (beyond them the literal recheck runs).
static const size_t kMaxGramBytes = (size_t)8 << 30;
static const size_t kMaxBbufBytes = (size_t)2 << 30;
// One recording's screen batch: flagged windows whose B rows a batch holds at most (each
// batch ends its three kernels on a partial round of blocks, and the label-grouped order
// puts a label's costly windows together, so fewer, larger batches pay)
static const size_t kScreenBatchBytes = (size_t)8 << 30;
static const size_t kScreenBatchWindows = 262144;
static const size_t kMaxLds = 64 * 1024;
static const size_t kMaxLdsSelect = 160 * 1024;

using namespace fbp;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Growable buffer that keeps its contents (stream-ordered copy).
hipError_t grow_keep(DevBuf& b, size_t bytes, size_t used, hipStream_t s)
{
    if (bytes <= b.cap) return hipSuccess;
    size_t want = std::max(bytes, b.cap * 2);
    void* np = nullptr;
    hipError_t e = hipMalloc(&np, want);
    if (e != hipSuccess) return e;
    if (used && b.p) {
        e = hipMemcpyAsync(np, b.p, used, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
    }
    if (b.p) (void)hipFree(b.p);
    b.p = np;
    b.cap = want;
    return hipSuccess;
}

// Host threads that copy the rows of one voltage piece between the caller's
// (pageable) array and a pinned staging buffer.  run() hands out jobs 0..n-1,
// the calling thread takes part, and it returns when every job is done.
class CopyPool {
public:
    explicit CopyPool(int n)
    {
        for (int i = 0; i < n; i++) th_.emplace_back([this] { loop(); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void run(int njobs, const std::function<void(int)>& f)
    {
        std::unique_lock<std::mutex> lk(mu_);
        job_ = &f;
        njobs_ = njobs;
        next_ = 0;
        done_ = 0;
        gen_++;
        cv_.notify_all();
        while (next_ < njobs_) {
            const int j = next_++;
            lk.unlock();
            f(j);
            lk.lock();
            ++done_;
        }
        cv_done_.wait(lk, [&] { return done_ == njobs_; });
        job_ = nullptr;
        njobs_ = 0;
    }

private:
    void loop()
    {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || (gen_ != seen && next_ < njobs_); });
            if (stop_) return;
            seen = gen_;
            while (next_ < njobs_) {
                const int j = next_++;
                const std::function<void(int)>* f = job_;
                lk.unlock();
                (*f)(j);
                lk.lock();
                if (++done_ == njobs_) cv_done_.notify_all();
            }
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
    const std::function<void(int)>* job_ = nullptr;
    int njobs_ = 0, next_ = 0, done_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Host voltage moves through kStageBufs pieces of at most kStageBytes: the
// rows of piece k are copied into pinned buffer k % kStageBufs by the pool
// while the DMA of piece k-1 runs; each piece is transposed into the
// sample-major working copy on the device (no full-size staging copy).
constexpr int kStageBufs = 3;
constexpr size_t kStageBytes = (size_t)32 << 20;
constexpr int kCopyThreads = 8;
constexpr size_t kCopySegBytes = (size_t)1 << 20;  // rows split into jobs of at most 1 MiB

}  // namespace

struct fbp_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // templates
    int U = 0, C = 0, L = 0, Lp = 0;
    bool have_templates = false;
    DevBuf T, rows, nch, off, chan, nz, bias, thr, cmask;
    std::vector<int> h_nch;
    std::vector<uint8_t> h_nz;
    std::vector<float> h_rowL1, h_SS;
    int Cmax = 0;
    // screened-recheck tables (valid for tab_n)
    int64_t tab_n = -1;
    int npl_max = 0;  // most units sharing a channel with one unit
    int64_t npairs = 0;
    bool screen_ok = false;
    DevBuf pid, pair_l, pair_u, G, minG, minGb, Ntr, Til, rowL1, SS, Bbuf, xcand, Fs, lhist, rcnt, lorder;
    // matrix-core tables: fp16 fragments of 16-template tiles (hi parts for the
    // likelihood; hi and lo parts for the overlap screen's B)
    bool mf_built = false;
    DevBuf mf_frag, mf_fragl, mf_units, mf_nkb, mf_nblk, mf_aoff, mf_chb, mf_nunit, mf_T1, mf_mbits, mf_vs, mf_tileof;
    int mf_ntiles = 0;
    float mf_tscale = 1.0f, mf_kl = 0.0f, mf_kb = 0.0f, mf_ka = 0.0f;
    float scr_k2 = 0.0f, scr_k3 = 0.0f, scr_k4 = 0.0f;  // screen bound terms of the split-fp16 B
    bool mf_ok = false;
    bool mf_pays = false;  // the matrix-core first pass is expected to beat the FP32 kernel
    // per-template BF16 / FP8 matrix-core likelihood: tables for mma_fmt (-1: none built)
    int mma_fmt = -1;
    int mma_KB = 0;
    float mma_kl = 0.0f, mma_kb = 0.0f, mma_ka = 0.0f, mma_kd = 0.0f;
    DevBuf mma_frag, mma_tsc, mma_T1, mma_scal, mma_q, mma_bmax, mma_gmask;
    DevBuf mma_qtmp, mma_qch, mma_rslot;  // template quads of the dense first pass
    int mma_nq = 0, mma_qc = 0;
    std::vector<float> h_rows;
    std::vector<int> h_off;
    // chunk state
    int64_t N = 0, nwin = 0;
    DevBuf V, LIK, wmax_v, wmax_p, best_idx, best_lab, best_lik, obest_idx, obest_lab;
    DevBuf check, check_snap, orc, giveup, has, W, F, cand_idx, cand_lab, counters, blockbuf, dirty;
    DevBuf part_v, part_k;
    DevBuf sp_idx, sp_lab, clips;
    uint32_t* h_counters = nullptr;  // pinned
    int64_t n_spikes = 0;
    bool have_result = false, have_clips = false;
    fbp_stats stats{};
    std::vector<int64_t> pass_log;  // per pass: spikes added, windows listed, windows flagged
    // host <-> device staging of the voltage (created on first use)
    float* pinned[kStageBufs] = {};
    DevBuf dstage[kStageBufs];
    hipEvent_t stage_ev[kStageBufs] = {};  // the copy that read pinned[b] is done
    hipEvent_t stage_tr[kStageBufs] = {};  // the transpose that read dstage[b] is done
    hipEvent_t stage_mark = nullptr;
    bool stage_busy[kStageBufs] = {};
    bool stage_tr_busy[kStageBufs] = {};
    hipStream_t cstream = nullptr;         // host-to-device copies of the voltage (default priority)
    std::unique_ptr<CopyPool> pool;
};

// Pursuit calls in flight in the process.  A batch with FBP_FLAG_BATCH_GRAPH
// (it builds, launches and destroys hipGraphs) runs only while no other
// fbp_run / fbp_run_batch is in flight, and none starts until it returns:
// graph calls are kept on one host thread at a time.
std::mutex g_run_mu;
int g_runs = 0;
bool g_graph_run = false;

struct RunGuard {
    bool ok = false, graph = false;
    explicit RunGuard(bool graph_run) : graph(graph_run)
    {
        std::lock_guard<std::mutex> lk(g_run_mu);
        if (g_graph_run || (graph && g_runs > 0)) return;
        g_runs++;
        if (graph) g_graph_run = true;
        ok = true;
    }
    ~RunGuard()
    {
        if (!ok) return;
        std::lock_guard<std::mutex> lk(g_run_mu);
        g_runs--;
        if (graph) g_graph_run = false;
    }
};

static int set_err(fbp_handle* h, int code, const std::string& msg)
{
    if (h) h->err = msg;
    return code;
}

static int hip_err(fbp_handle* h, hipError_t e, const char* where)
{
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    return set_err(h, e == hipErrorOutOfMemory ? FBP_E_NOMEM : FBP_E_HIP, m);
}

#define HIP_TRY(expr, where)                               \
    do {                                                   \
        hipError_t _e = (expr);                            \
        if (_e != hipSuccess) return hip_err(h, _e, where); \
    } while (0)

static TemplateSet template_view(const fbp_handle* h)
{
    TemplateSet ts;
    ts.T = h->T.as<float>();
    ts.rows = h->rows.as<float>();
    ts.nch = h->nch.as<int>();
    ts.off = h->off.as<int>();
    ts.chan = h->chan.as<int>();
    ts.nz = h->nz.as<uint8_t>();
    ts.bias = h->bias.as<float>();
    ts.thr = h->thr.as<float>();
    ts.cmask = h->cmask.as<uint64_t>();
    ts.U = h->U;
    ts.C = h->C;
    ts.L = h->L;
    ts.Lp = h->Lp;
    ts.R = (int)h->h_rowL1.size();
    return ts;
}

static PassState pass_view(const fbp_handle* h)
{
    PassState st;
    st.V = h->V.as<float>();
    st.LIK = h->LIK.as<float>();
    st.wmax_v = h->wmax_v.as<float>();
    st.wmax_p = h->wmax_p.as<uint16_t>();
    st.best_idx = h->best_idx.as<uint32_t>();
    st.best_lab = h->best_lab.as<uint32_t>();
    st.best_lik = h->best_lik.as<float>();
    st.obest_idx = h->obest_idx.as<uint32_t>();
    st.obest_lab = h->obest_lab.as<uint32_t>();
    st.check = h->check.as<uint8_t>();
    st.check_snap = h->check_snap.as<uint8_t>();
    st.orc = h->orc.as<uint8_t>();
    st.giveup = h->giveup.as<uint8_t>();
    st.has = h->has.as<uint8_t>();
    st.W = h->W.as<uint32_t>();
    st.F = h->F.as<uint32_t>();
    st.cand_idx = h->cand_idx.as<uint32_t>();
    st.cand_lab = h->cand_lab.as<uint32_t>();
    st.counters = h->counters.as<uint32_t>();
    // channel masks are 64-bit: with more channels every template is recomputed
    st.dirty = h->C <= 64 ? h->dirty.as<uint64_t>() : nullptr;
    st.N = h->N;
    st.nwin = h->nwin;
    return st;
}

// Event-pair timer for one kernel class.
struct Timer {
    hipStream_t s;
    bool on;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[8];
    Timer(hipStream_t st, bool enable) : s(st), on(enable) {}
    void begin(int cls)
    {
        if (!on) return;
        hipEvent_t a, b;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, s);
        ev[cls].push_back({a, b});
    }
    void end(int cls)
    {
        if (!on) return;
        hipEventRecord(ev[cls].back().second, s);
    }
    double total(int cls)
    {
        double t = 0;
        for (auto& p : ev[cls]) {
            float ms = 0;
            hipEventElapsedTime(&ms, p.first, p.second);
            t += ms;
        }
        return t;
    }
    ~Timer()
    {
        for (auto& v : ev)
            for (auto& p : v) {
                hipEventDestroy(p.first);
                hipEventDestroy(p.second);
            }
    }
};
// Two events owned for the length of a scope.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
enum { TM_LIK = 0, TM_MAX, TM_RECHECK, TM_PURSUIT, TM_RESID, TM_OTHER, TM_SELECT, TM_RESCORE };

static ScreenTables tables_view(const fbp_handle* h)
{
    ScreenTables tb;
    tb.pid = h->pid.as<int>();
    tb.pair_l = h->pair_l.as<int>();
    tb.pair_u = h->pair_u.as<int>();
    tb.G = h->G.as<float>();
    tb.minG = h->minG.as<float>();
    tb.minGb = h->minGb.as<float>();
    tb.nbg = (int)((2 * h->tab_n + 1 + kGB - 1) / kGB);
    tb.npl_max = h->npl_max;
    tb.Ntr = h->Ntr.as<float>();
    tb.Til = h->Til.as<float>();
    tb.upad = (h->U + 63) & ~63;
    tb.rowL1 = h->rowL1.as<float>();
    tb.SS = h->SS.as<float>();
    tb.n = (int)h->tab_n;
    tb.S = (int)(2 * h->tab_n + 1);
    tb.sg = (tb.S + 3) & ~3;
    tb.cmax = std::max(h->Cmax, 1);
    // delta = kappa * (Q_l + Q_u + 2 (SS_l + SS_u)) bounds |E-hat - E_fp32|:
    // both are sums of at most K rounded terms (gamma_K each), doubled for safety.
    const double K = 2.0 * h->Cmax * (h->L + 1) + 2.0 * h->L + 16.0;
    const double u = std::ldexp(1.0, -24);
    tb.kappa = (float)(4.0 * K * u / (1.0 - K * u));
    // FP32 B (the vector-unit correlation); the matrix-core B sets these
    tb.kappa2 = 0.0f;
    tb.kappa3 = 0.0f;
    tb.kappa4 = 0.0f;
    tb.sel_region = screen_select_region(template_view(h), tb.S, tb.cmax, tb.npl_max);
    return tb;
}

// Matrix-core tables (likelihood and overlap screen): templates ordered by
// their first non-zero channel, cut into tiles of at most 16 where the k-blocks
// summed over the tiles are fewest (dynamic programming over the cut points);
// each tile covers the contiguous 8-channel blocks from its first to its last
// non-zero channel; the
// A operand in fragment order (tile, k-block of 32, lane, 8 values), k =
// (sample j, block b, channel in block) with run 4 * kb + lane / 16 =
// j * blocks + b, as fp16 hi parts and the lo parts of the remainder.  mf_ok
// stays false (FP32 kernels) when L > 64 or the probe has more than 256
// channels.
static int build_mf_tables(fbp_handle* h)
{
    h->mf_built = true;
    h->mf_ok = false;
    const int U = h->U, C = h->C, L = h->L;
    if (L > 64 || (C + 7) / 8 > kMfMaxBlk) return FBP_OK;
    std::vector<float> hT((size_t)U * C * L);
    HIP_TRY(hipMemcpy(hT.data(), h->T.p, hT.size() * 4, hipMemcpyDeviceToHost), "copy templates");
    std::vector<int> first(U, C), lastc(U, -1), order(U);
    for (int u = 0; u < U; u++) {
        for (int c = 0; c < C; c++)
            if (h->h_nz[(size_t)u * C + c]) {
                if (first[u] == C) first[u] = c;
                lastc[u] = c;
            }
        order[u] = u;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return first[a] < first[b]; });
    // cost[j]: fewest k-blocks of the first j sorted templates; cut[j]: start of the last tile
    // (ties go to the larger tile)
    std::vector<int> cost(U + 1, INT_MAX), cut(U + 1, 0);
    cost[0] = 0;
    for (int j = 1; j <= U; j++) {
        int blo = kMfMaxBlk, bhi = -1;
        for (int i = j - 1; i >= std::max(0, j - 16); i--) {
            const int u = order[i];
            if (lastc[u] >= 0) {
                blo = std::min(blo, first[u] / 8);
                bhi = std::max(bhi, lastc[u] / 8);
            }
            const int nb = bhi < 0 ? 1 : bhi - blo + 1;
            const int c = cost[i] + (L * nb + 3) / 4;
            if (c <= cost[j]) {
                cost[j] = c;
                cut[j] = i;
            }
        }
    }
    std::vector<std::pair<int, int>> runs;  // [begin, end) of each tile in the sorted order
    for (int j = U; j > 0; j = cut[j]) runs.emplace_back(cut[j], j);
    std::reverse(runs.begin(), runs.end());
    const int ntiles = (int)runs.size();
    if (likmfma_lds(template_view(h), ntiles) > kMaxLdsSelect) return FBP_OK;  // beyond 160 KiB: FP32 kernels
    std::vector<int> units((size_t)ntiles * 16, -1), nkb(ntiles), nblk(ntiles), aoff(ntiles), chb(ntiles, 0),
        nunit(ntiles);
    int total_kb = 0, maxnb = 1;
    for (int t = 0; t < ntiles; t++) {
        int blo = kMfMaxBlk, bhi = -1;
        nunit[t] = runs[t].second - runs[t].first;
        for (int r = 0; r < nunit[t]; r++) {
            const int u = order[runs[t].first + r];
            units[(size_t)t * 16 + r] = u;
            for (int c = 0; c < C; c++)
                if (h->h_nz[(size_t)u * C + c]) {
                    blo = std::min(blo, c / 8);
                    bhi = std::max(bhi, c / 8);
                }
        }
        if (bhi < 0) blo = bhi = 0;
        chb[t] = blo;  // the tile covers the contiguous 8-channel blocks blo .. bhi
        const int nb = bhi - blo + 1;
        nblk[t] = nb;
        nkb[t] = (L * nb + 3) / 4;
        aoff[t] = total_kb;
        total_kb += nkb[t];
        maxnb = std::max(maxnb, nb);
    }
    float tmax = 0.0f;
    for (float v : hT) tmax = std::max(tmax, std::fabs(v));
    int ex = 0;
    std::frexp(tmax > 0.0f ? tmax : 1.0f, &ex);
    h->mf_tscale = std::ldexp(1.0f, 14 - ex);
    std::vector<_Float16> frag((size_t)total_kb * 64 * 8), fragl(frag.size());
    for (int t = 0; t < ntiles; t++)
        for (int kb = 0; kb < nkb[t]; kb++)
            for (int lane = 0; lane < 64; lane++)
                for (int i = 0; i < 8; i++) {
                    const int u = units[(size_t)t * 16 + (lane & 15)];
                    const int r = 4 * kb + (lane >> 4);
                    float v = 0.0f;
                    if (u >= 0 && r < L * nblk[t]) {
                        const int j = r / nblk[t], b = r - j * nblk[t];
                        const int c = (chb[t] + b) * 8 + i;
                        if (c < C) v = hT[((size_t)u * C + c) * L + j] * h->mf_tscale;
                    }
                    const _Float16 hi = (_Float16)v;
                    frag[(((size_t)aoff[t] + kb) * 64 + lane) * 8 + i] = hi;
                    fragl[(((size_t)aoff[t] + kb) * 64 + lane) * 8 + i] = (_Float16)(v - (float)hi);
                }
    std::vector<int> tileof(U, 0);
    for (int t = 0; t < ntiles; t++)
        for (int r = 0; r < 16; r++)
            if (units[(size_t)t * 16 + r] >= 0) tileof[units[(size_t)t * 16 + r]] = t;
    std::vector<float> T1(U);
    for (int u = 0; u < U; u++) {
        double a = 0.0;
        for (size_t k = 0; k < (size_t)C * L; k++) a += std::fabs((double)hT[(size_t)u * C * L + k]);
        T1[u] = (float)(a * (1.0 + 1e-6));
    }
    HIP_TRY(h->mf_frag.ensure(frag.size() * sizeof(_Float16)), "alloc likelihood tables");
    HIP_TRY(h->mf_fragl.ensure(fragl.size() * sizeof(_Float16)), "alloc likelihood tables");
    HIP_TRY(h->mf_units.ensure(units.size() * 4), "alloc likelihood tables");
    HIP_TRY(h->mf_nkb.ensure((size_t)ntiles * 4), "alloc likelihood tables");
    HIP_TRY(h->mf_nblk.ensure((size_t)ntiles * 4), "alloc likelihood tables");
    HIP_TRY(h->mf_aoff.ensure((size_t)ntiles * 4), "alloc likelihood tables");
    HIP_TRY(h->mf_chb.ensure(chb.size() * 4), "alloc likelihood tables");
    HIP_TRY(h->mf_nunit.ensure(nunit.size() * 4), "alloc likelihood tables");
    HIP_TRY(h->mf_T1.ensure((size_t)U * 4), "alloc likelihood tables");
    HIP_TRY(h->mf_tileof.ensure((size_t)U * 4), "alloc likelihood tables");
    HIP_TRY(h->mf_mbits.ensure(4), "alloc likelihood tables");
    HIP_TRY(h->mf_vs.ensure(4), "alloc likelihood tables");
    HIP_TRY(hipMemcpy(h->mf_frag.p, frag.data(), frag.size() * sizeof(_Float16), hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_fragl.p, fragl.data(), fragl.size() * sizeof(_Float16), hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_units.p, units.data(), units.size() * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_nkb.p, nkb.data(), (size_t)ntiles * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_nblk.p, nblk.data(), (size_t)ntiles * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_aoff.p, aoff.data(), (size_t)ntiles * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_chb.p, chb.data(), chb.size() * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_nunit.p, nunit.data(), nunit.size() * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_T1.p, T1.data(), (size_t)U * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mf_tileof.p, tileof.data(), (size_t)U * 4, hipMemcpyHostToDevice), "copy");
    h->mf_ntiles = ntiles;
    // Likelihood: delta = kl * Q + kb * |bias| + ka * (T1 / vs + L * nch * max|V| / ts).
    // An operand x (scaled) enters as fp16 x' with |x - x'| <= 2^-11 |x| + 2^-14
    // (2^-14 covers a sub-normal part, kept or flushed by the matrix core), so a
    // product errs by at most (1 + 2^-10) (2^-10 |t v| + 2^-14 (|t| + |v|)); add
    // FP32 accumulation of K products (gamma_K, doubled for an unknown order),
    // the exact chain's own error (gamma_K2 on the products and the bias) and
    // the final scaling and bias add; all doubled.
    const double u = std::ldexp(1.0, -24);
    const double K = (double)L * 8.0 * maxnb, K2 = (double)h->Cmax * L + 1.0;
    const double gK = K * u / (1.0 - K * u), gK2 = K2 * u / (1.0 - K2 * u);
    h->mf_kl = (float)(2.0 * (std::ldexp(1.0, -10) * (1.0 + std::ldexp(1.0, -10)) + 2.0 * gK + gK2 +
                              std::ldexp(1.0, -23)));
    h->mf_kb = (float)(2.0 * (gK2 + std::ldexp(1.0, -24)));
    h->mf_ka = (float)(2.0 * std::ldexp(1.0, -14) * (1.0 + std::ldexp(1.0, -10)));
    // Overlap screen, B from split operands x = hi + lo + e with
    // |e| <= 2^-22 |x| + 2^-14 and |e| <= |x| (scaled): three products
    // hi*hi + hi*lo + lo*hi err by at most 4 * 2^-22 |t v| + 2^-14 (2 |t| + |v|)
    // each, i.e. 4 * 2^-22 |t v| + 2^-27 (2 maxw |t| + Tmax |v|) unscaled (both
    // scales put the operand maximum in [2^13, 2^14)).  FP32 accumulation in
    // two chains (unit roundoff doubled for an unknown order inside the matrix
    // core): hi*hi of at most Cmax * L non-zero products, the cross products
    // (at most 2^-10 of |t v| together) of twice as many, and their sum.  The
    // relative part rides on Q (kappa2), the absolute part on sum |T| and the
    // window maxima (kappa3, kappa4); all doubled.
    {
        const double K3 = (double)h->Cmax * L + 8.0;
        const double u2 = 2.0 * u;
        const double g1 = K3 * u2 / (1.0 - K3 * u2), g2 = 2.0 * K3 * u2 / (1.0 - 2.0 * K3 * u2);
        h->scr_k2 = (float)(2.0 * ((1.0 + u) * (g1 + std::ldexp(1.0, -10) * g2) + u + 4.0 * std::ldexp(1.0, -22)));
        h->scr_k3 = (float)(2.0 * 2.0 * std::ldexp(1.0, -27) * (1.0 + std::ldexp(1.0, -10)));
        h->scr_k4 = (float)(2.0 * std::ldexp(1.0, -27) * (1.0 + std::ldexp(1.0, -10)) * (double)tmax * (double)L);
    }
    // First-pass choice: the matrix-core kernel multiplies every tile over the union of
    // its templates' 8-channel blocks, the FP32 kernel only over each template's own
    // channels.  Measured per row, the matrix cores win while the tiles carry at most
    // about 2.5 times the templates' own channel rows: dense templates (ratio 1.0) take
    // 0.4x the FP32 time, 7-of-64-channel templates (3.5) 1.15x, 5 units on 4 channels
    // (6.4) 3.4x.
    {
        double rows_mf = 0.0, rows_fp = 0.0;
        for (int t = 0; t < ntiles; t++) rows_mf += 16.0 * 8.0 * nblk[t];
        for (size_t i = 0; i < h->h_nz.size(); i++) rows_fp += h->h_nz[i] ? 1.0 : 0.0;
        h->mf_pays = rows_mf <= kMfPadMax * std::max(rows_fp, 1.0);
    }
    h->mf_ok = true;
    return FBP_OK;
}

static MfTables mf_view(const fbp_handle* h)
{
    MfTables mt;
    mt.frag = h->mf_frag.as<half8>();
    mt.units = h->mf_units.as<int>();
    mt.nkb = h->mf_nkb.as<int>();
    mt.nblk = h->mf_nblk.as<int>();
    mt.aoff = h->mf_aoff.as<int>();
    mt.chb = h->mf_chb.as<int>();
    mt.nunit = h->mf_nunit.as<int>();
    mt.T1 = h->mf_T1.as<float>();
    mt.tileof = h->mf_tileof.as<int>();
    mt.ntiles = h->mf_ntiles;
    mt.inv_tscale = 1.0f / h->mf_tscale;
    mt.kl = h->mf_kl;
    mt.kb = h->mf_kb;
    mt.ka = h->mf_ka;
    return mt;
}

// Tables of the per-template BF16 (fmt 0) / FP8 e4m3 (fmt 1) likelihood: B
// fragments of each template row (the row shifted by n = lane % 16 samples,
// sample j = 8 KB (lane / 16) + 8 kb + e of k-block kb), each template scaled by a power of two
// (FP8: its largest |value| into [128, 256); BF16: 1), and the bound constants.
// Operand x (scaled) enters as x' with |x - x'| <= eps |x| + a (BF16: eps =
// 2^-8, a = 2^-134, half its smallest sub-normal; FP8 e4m3: eps = 2^-4, a =
// 2^-10; neither format's sub-normals are flushed by the matrix cores), so a
// product errs by (2 eps + eps^2) |t v| + a (1 + eps) (|t| + |v|) + a^2; the
// products are exact in FP32 and summed in FP32 (K = Cmax * 32 KB terms, gamma_K
// doubled for an unknown order); the exact chain errs by gamma_K2 (|bias| + sum
// |t v|), K2 = Cmax * L + 1; the bias add rounds once.  All doubled.
static int build_mma_tables(fbp_handle* h, int fmt)
{
    const int U = h->U, L = h->L, Lp = h->Lp;
    const int KB = (L + 15 + 31) / 32;
    const int Rr = (int)h->h_rowL1.size();
    std::vector<float> tsc(U, 1.0f), T1(U, 0.0f);
    for (int u = 0; u < U; u++) {
        double a = 0.0;
        float mx = 0.0f;
        for (int i = 0; i < h->h_nch[u]; i++)
            for (int j = 0; j < L; j++) {
                const float v = h->h_rows[(size_t)(h->h_off[u] + i) * Lp + j];
                a += std::fabs((double)v);
                mx = std::max(mx, std::fabs(v));
            }
        T1[u] = (float)(a * (1.0 + 1e-6));
        if (fmt == 1 && mx > 0.0f) {
            int e = 0;
            std::frexp(mx, &e);
            tsc[u] = std::ldexp(1.0f, 8 - e);
        }
    }
    const size_t nfr = (size_t)std::max(Rr, 1) * KB;
    std::vector<float> vals(nfr * 64 * 8, 0.0f);
    for (int u = 0; u < U; u++)
        for (int i = 0; i < h->h_nch[u]; i++) {
            const int r = h->h_off[u] + i;
            for (int kb = 0; kb < KB; kb++)
                for (int lane = 0; lane < 64; lane++)
                    for (int e = 0; e < 8; e++) {
                        const int j = 8 * KB * (lane >> 4) + 8 * kb + e - (lane & 15);
                        const float v = (j >= 0 && j < L) ? h->h_rows[(size_t)r * Lp + j] * tsc[u] : 0.0f;
                        vals[(((size_t)r * KB + kb) * 64 + lane) * 8 + e] = v;
                    }
        }
    DevBuf tmp;
    HIP_TRY(tmp.ensure(vals.size() * 4), "alloc fragment staging");
    HIP_TRY(h->mma_frag.ensure(nfr * 64 * (fmt == 0 ? 16 : 8)), "alloc likelihood fragments");
    HIP_TRY(h->mma_tsc.ensure((size_t)U * 4), "alloc likelihood tables");
    HIP_TRY(h->mma_T1.ensure((size_t)U * 4), "alloc likelihood tables");
    HIP_TRY(h->mma_scal.ensure(16), "alloc likelihood tables");
    HIP_TRY(hipMemcpy(tmp.p, vals.data(), vals.size() * 4, hipMemcpyHostToDevice), "copy fragments");
    HIP_TRY(hipMemcpy(h->mma_tsc.p, tsc.data(), (size_t)U * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->mma_T1.p, T1.data(), (size_t)U * 4, hipMemcpyHostToDevice), "copy");
    launch_frag_convert(tmp.as<float>(), h->mma_frag.p, (int64_t)(nfr * 64), fmt, h->stream);
    HIP_TRY(hipGetLastError(), "fragment conversion");
    HIP_TRY(hipStreamSynchronize(h->stream), "fragment conversion");
    tmp.release();
    const double u = std::ldexp(1.0, -24);
    const double eps = fmt == 0 ? std::ldexp(1.0, -8) : std::ldexp(1.0, -4);
    const double a = fmt == 0 ? std::ldexp(1.0, -134) : std::ldexp(1.0, -10);
    const double K = (double)std::max(h->Cmax, 1) * 32.0 * KB, K2 = (double)std::max(h->Cmax, 1) * L + 1.0;
    const double gK = K * u / (1.0 - K * u), gK2 = K2 * u / (1.0 - K2 * u);
    h->mma_kl = (float)(2.0 * ((2.0 * eps + eps * eps) + 2.0 * gK * (1.0 + eps) * (1.0 + eps) + gK2 +
                               std::ldexp(1.0, -23)));
    h->mma_kb = (float)(2.0 * (gK2 + std::ldexp(1.0, -24)));
    h->mma_ka = (float)(2.0 * a * (1.0 + eps) * (1.0 + 2.0 * gK));
    h->mma_kd = (float)(2.0 * K * a * a * (1.0 + 2.0 * gK));
    h->mma_KB = KB;
    h->mma_fmt = fmt;
    // quads of templates for the dense first pass: templates ordered by first
    // channel, four at a time, each quad's channels (union, ascending) staged
    // together; none when L is not a multiple of 16 or a quad spans more than
    // kQuadCh channels
    h->mma_nq = 0;
    h->mma_qc = 0;
    if (L % 16 == 0 && L <= 64) {
        const int C = h->C;
        std::vector<int> first(U, C), ord(U);
        for (int u = 0; u < U; u++)
            for (int c = 0; c < C; c++)
                if (h->h_nz[(size_t)u * C + c]) { first[u] = c; break; }
        for (int u = 0; u < U; u++) ord[u] = u;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return first[a] < first[b]; });
        const int nq = (U + 3) / 4;
        std::vector<int> qtmp((size_t)nq * 4, -1), qch((size_t)nq * kQuadCh, -1), rslot(std::max(Rr, 1), 0);
        bool ok = true;
        int qc = 0;
        for (int q = 0; q < nq && ok; q++) {
            std::vector<int> chs;
            for (int j = 0; j < 4 && 4 * q + j < U; j++) {
                const int u = ord[4 * q + j];
                qtmp[(size_t)q * 4 + j] = u;
                for (int c = 0; c < C; c++)
                    if (h->h_nz[(size_t)u * C + c]) chs.push_back(c);
            }
            std::sort(chs.begin(), chs.end());
            chs.erase(std::unique(chs.begin(), chs.end()), chs.end());
            if ((int)chs.size() > kQuadCh) { ok = false; break; }
            qc = std::max(qc, (int)chs.size());
            for (size_t i = 0; i < chs.size(); i++) qch[(size_t)q * kQuadCh + i] = chs[i];
            for (int j = 0; j < 4 && 4 * q + j < U; j++) {
                const int u = ord[4 * q + j];
                int k = 0;
                for (int c = 0; c < C; c++)
                    if (h->h_nz[(size_t)u * C + c])
                        rslot[h->h_off[u] + k++] = (int)(std::lower_bound(chs.begin(), chs.end(), c) - chs.begin());
            }
        }
        if (ok && qc > 0) {
            HIP_TRY(h->mma_qtmp.ensure(qtmp.size() * 4), "alloc quads");
            HIP_TRY(h->mma_qch.ensure(qch.size() * 4), "alloc quads");
            HIP_TRY(h->mma_rslot.ensure(rslot.size() * 4), "alloc quads");
            HIP_TRY(hipMemcpy(h->mma_qtmp.p, qtmp.data(), qtmp.size() * 4, hipMemcpyHostToDevice), "copy");
            HIP_TRY(hipMemcpy(h->mma_qch.p, qch.data(), qch.size() * 4, hipMemcpyHostToDevice), "copy");
            HIP_TRY(hipMemcpy(h->mma_rslot.p, rslot.data(), rslot.size() * 4, hipMemcpyHostToDevice), "copy");
            h->mma_nq = nq;
            h->mma_qc = qc;
        }
    }
    return FBP_OK;
}

static MmaTables mma_view(const fbp_handle* h)
{
    MmaTables mt;
    mt.frag = h->mma_frag.p;
    mt.tsc = h->mma_tsc.as<float>();
    mt.T1 = h->mma_T1.as<float>();
    mt.rowL1 = h->rowL1.as<float>();
    mt.KB = h->mma_KB;
    mt.R = (h->L + 15) / 16;
    mt.P = 16 / mt.R;
    mt.kl = h->mma_kl;
    mt.kb = h->mma_kb;
    mt.ka = h->mma_ka;
    mt.kd = h->mma_kd;
    mt.qtmp = h->mma_qtmp.as<int>();
    mt.qch = h->mma_qch.as<int>();
    mt.rslot = h->mma_rslot.as<int>();
    mt.nq = h->mma_nq;
    mt.qc = h->mma_qc;
    return mt;
}

// Builds the voltage-independent tables of the screened recheck for shift
// range n.  Returns false (literal recheck) when they do not fit the limits.
static int build_screen_tables(fbp_handle* h, int64_t n, bool* ok)
{
    *ok = false;
    h->tab_n = n;
    h->screen_ok = false;
    const int U = h->U, C = h->C, L = h->L;
    const int64_t S = 2 * n + 1;
    TemplateSet ts = template_view(h);
    int rp;
    if (screen_corr_lds(ts, (int)S, &rp) > kMaxLds || screen_select_lds(ts, (int)S, std::max(h->Cmax, 1)) > kMaxLdsSelect - 16384)
        return FBP_OK;
    if ((size_t)U * S * 4 > kMaxBbufBytes) return FBP_OK;
    std::vector<int> pid((size_t)U * U, -1), pl, pu;
    for (int l = 0; l < U; l++)
        for (int u = 0; u < U; u++) {
            bool sh = false;
            for (int c = 0; c < C && !sh; c++) sh = h->h_nz[(size_t)l * C + c] && h->h_nz[(size_t)u * C + c];
            if (sh) {
                pid[(size_t)l * U + u] = (int)pl.size();
                pl.push_back(l);
                pu.push_back(u);
            }
        }
    const int64_t np = (int64_t)pl.size();
    if ((size_t)np * S * S * 4 > kMaxGramBytes) return FBP_OK;
    h->npl_max = 0;
    std::vector<int> partners(U);
    for (int l = 0; l < U; l++) {
        int c = 0;
        for (int u = 0; u < U; u++) c += pid[(size_t)l * U + u] >= 0;
        partners[l] = c;
        h->npl_max = std::max(h->npl_max, c);
    }
    // Label order of the grouped flagged windows: most channel-sharing partners first, so
    // that a screen launch ends on the labels whose windows enumerate the fewest items
    std::vector<uint32_t> lorder(U);
    for (int l = 0; l < U; l++) lorder[l] = (uint32_t)l;
    std::stable_sort(lorder.begin(), lorder.end(), [&](uint32_t a, uint32_t b) { return partners[a] > partners[b]; });
    if (pl.empty()) { pl.push_back(0); pu.push_back(0); }
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    HIP_TRY(h->pid.ensure(pid.size() * 4), "alloc pid");
    HIP_TRY(h->lorder.ensure((size_t)U * 4), "alloc label order");
    HIP_TRY(hipMemcpy(h->lorder.p, lorder.data(), (size_t)U * 4, hipMemcpyHostToDevice), "copy label order");
    HIP_TRY(h->pair_l.ensure(pl.size() * 4), "alloc pairs");
    HIP_TRY(h->pair_u.ensure(pu.size() * 4), "alloc pairs");
    HIP_TRY(h->G.ensure((size_t)std::max<int64_t>(np, 1) * S * ((S + 3) & ~3) * 4), "alloc gram table");
    HIP_TRY(h->minG.ensure((size_t)std::max<int64_t>(np, 1) * S * 4), "alloc gram minima");
    HIP_TRY(h->minGb.ensure((size_t)std::max<int64_t>(np, 1) * S * ((S + kGB - 1) / kGB) * 4), "alloc gram minima");
    HIP_TRY(h->Ntr.ensure((size_t)U * S * 4), "alloc energies");
    HIP_TRY(h->Til.ensure((size_t)std::max(h->Cmax, 1) * h->Lp * ((U + 63) & ~63) * 4), "alloc interleaved rows");

    HIP_TRY(hipMemcpy(h->pid.p, pid.data(), pid.size() * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->pair_l.p, pl.data(), pl.size() * 4, hipMemcpyHostToDevice), "copy");
    HIP_TRY(hipMemcpy(h->pair_u.p, pu.data(), pu.size() * 4, hipMemcpyHostToDevice), "copy");

    h->npairs = np;
    launch_screen_tables(ts, tables_view(h), np, h->stream);
    HIP_TRY(hipGetLastError(), "screen tables");
    HIP_TRY(hipStreamSynchronize(h->stream), "screen tables");
    h->screen_ok = true;
    *ok = true;
    return FBP_OK;
}

static int ensure_staging(fbp_handle* h)
{
    if (!h->pool) h->pool.reset(new CopyPool(kCopyThreads));
    for (int b = 0; b < kStageBufs; b++) {
        if (!h->pinned[b])
            HIP_TRY(hipHostMalloc((void**)&h->pinned[b], kStageBytes, hipHostMallocDefault), "pinned staging");
        HIP_TRY(h->dstage[b].ensure(kStageBytes), "alloc device staging");
        if (!h->stage_ev[b]) HIP_TRY(hipEventCreateWithFlags(&h->stage_ev[b], hipEventDisableTiming), "event");
        if (!h->stage_tr[b]) HIP_TRY(hipEventCreateWithFlags(&h->stage_tr[b], hipEventDisableTiming), "event");
    }
    if (!h->stage_mark) HIP_TRY(hipEventCreateWithFlags(&h->stage_mark, hipEventDisableTiming), "event");
    if (!h->cstream) HIP_TRY(hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking), "copy stream");
    return FBP_OK;
}

// Samples per staged piece for C channels.
static int64_t stage_piece(int64_t C) { return std::max<int64_t>(64, (int64_t)(kStageBytes / 4) / C); }

// Row copies of one piece, split into jobs of at most kCopySegBytes:
// dst[c * dld + j] = src[c * sld + j] for j < n.
static void pool_copy_rows(CopyPool& pool, float* dst, int64_t dld, const float* src, int64_t sld, int64_t C,
                           int64_t n)
{
    const int64_t seg = std::max<int64_t>(1, (int64_t)(kCopySegBytes / 4));
    const int64_t nseg = (n + seg - 1) / seg;
    pool.run((int)(C * nseg), [&](int j) {
        const int64_t c = j / nseg, g = j - c * nseg;
        const int64_t a = g * seg, len = std::min(seg, n - a);
        std::memcpy(dst + c * dld + a, src + c * sld + a, (size_t)len * 4);
    });
}

// Host voltage [C][N] (row stride ld) into the sample-major device copy dst [N][C],
// through h's staging buffers on h's stream (C: h's channel count unless given).
// Host voltage V [C][N] (row stride ld) into the sample-major device copy dst,
// piece by piece: the copies run on the copy stream, each piece's transpose on
// the handle's stream after its copy; landed(x) (when given) is called after
// the transpose of every piece with the samples [0, x) in place, to queue work
// on them on the handle's stream while the later pieces are still copied.
static int stage_in(fbp_handle* h, const float* V, int64_t ld, int64_t N, float* dst, int64_t Cg = -1,
                    const std::function<int(int64_t)>& landed = nullptr, bool marked = false)
{
    int rc = ensure_staging(h);
    if (rc != FBP_OK) return rc;
    const int64_t C = Cg > 0 ? Cg : h->C, P = stage_piece(C);
    hipStream_t s = h->stream, cs = h->cstream;
    // the copies come after everything queued on the handle's stream (up to
    // the caller's own record of stage_mark when `marked`)
    if (!marked) HIP_TRY(hipEventRecord(h->stage_mark, s), "event");
    HIP_TRY(hipStreamWaitEvent(cs, h->stage_mark, 0), "stream wait");
    int64_t k = 0;
    for (int64_t x0 = 0; x0 < N; x0 += P, k++) {
        const int64_t np = std::min(P, N - x0);
        const int b = (int)(k % kStageBufs);
        if (h->stage_busy[b]) HIP_TRY(hipEventSynchronize(h->stage_ev[b]), "staging wait");
        pool_copy_rows(*h->pool, h->pinned[b], np, V + x0, ld, C, np);
        if (h->stage_tr_busy[b]) HIP_TRY(hipStreamWaitEvent(cs, h->stage_tr[b], 0), "stream wait");
        HIP_TRY(hipMemcpyAsync(h->dstage[b].p, h->pinned[b], (size_t)C * np * 4, hipMemcpyHostToDevice, cs),
                "copy voltage");
        HIP_TRY(hipEventRecord(h->stage_ev[b], cs), "event");
        h->stage_busy[b] = true;
        HIP_TRY(hipStreamWaitEvent(s, h->stage_ev[b], 0), "stream wait");
        launch_to_sample_major(h->dstage[b].as<float>(), np, dst + x0 * C, C, np, s);
        HIP_TRY(hipEventRecord(h->stage_tr[b], s), "event");
        h->stage_tr_busy[b] = true;
        if (landed) {
            rc = landed(x0 + np);
            if (rc != FBP_OK) return rc;
        }
    }
    return FBP_OK;
}

// Sample-major device voltage src [N][C] into the host array out [C][N] (row stride ldo).
static int stage_out(fbp_handle* h, const float* src, int64_t N, float* out, int64_t ldo)
{
    int rc = ensure_staging(h);
    if (rc != FBP_OK) return rc;
    const int64_t C = h->C, P = stage_piece(C);
    const int64_t npieces = (N + P - 1) / P;
    hipStream_t s = h->stream;
    auto issue = [&](int64_t k) -> int {
        const int64_t x0 = k * P, np = std::min(P, N - x0);
        const int b = (int)(k % kStageBufs);
        launch_to_channel_major(src + x0 * C, h->dstage[b].as<float>(), C, np, s);
        HIP_TRY(hipMemcpyAsync(h->pinned[b], h->dstage[b].p, (size_t)C * np * 4, hipMemcpyDeviceToHost, s),
                "copy residual");
        HIP_TRY(hipEventRecord(h->stage_ev[b], s), "event");
        h->stage_busy[b] = true;
        return FBP_OK;
    };
    for (int64_t k = 0; k < std::min<int64_t>(npieces, kStageBufs - 1); k++) {
        rc = issue(k);
        if (rc != FBP_OK) return rc;
    }
    for (int64_t k = 0; k < npieces; k++) {
        if (k + kStageBufs - 1 < npieces) {
            rc = issue(k + kStageBufs - 1);
            if (rc != FBP_OK) return rc;
        }
        const int b = (int)(k % kStageBufs);
        HIP_TRY(hipEventSynchronize(h->stage_ev[b]), "staging wait");
        const int64_t x0 = k * P, np = std::min(P, N - x0);
        pool_copy_rows(*h->pool, out + x0, ldo, h->pinned[b], np, C, np);
    }
    return FBP_OK;
}


// ---------------------------------------------------------------------------
// Batched pursuit over many segments (fbp_run_batch).
// ---------------------------------------------------------------------------
constexpr int kLikMaxWbHost = 16;         // windows per likelihood block at most (the kernel's limit)
constexpr uint32_t kSegBlockCap = 1024;  // default blocks per segment and launch at most (bodies grid-stride)

// Allocations, sample-major voltage and initial state of one segment, as the
// FP32 path of fbp_run sets them up (binary_pursuit_parallel.py:528-655), on
// the segment's own stream (synchronized before return).
static int batch_prepare(fbp_handle* h, fbp_handle* stager, const float* Vsrc, int64_t ld, int64_t N,
                         int32_t v_on_device, uint32_t n_max_shift, bool recheck)
{
    if (!h->have_templates) return set_err(h, FBP_E_STATE, "fbp_set_templates must be called first");
    if (!Vsrc || N <= 0 || ld < N) return set_err(h, FBP_E_INVALID, "voltage must be non-empty with ld >= N");
    const int64_t L = h->L, C = h->C, U = h->U;
    if (N >= (int64_t)0xffffffffLL) return set_err(h, FBP_E_UNSUPPORTED, "segment longer than 2^32-1 samples");
    if (C > 64) return set_err(h, FBP_E_UNSUPPORTED, "batched segments hold at most 64 channels");
    const uint64_t S = 2ull * n_max_shift + 1ull;
    if (recheck && S * S * (uint64_t)U >= 0xffffffffull)
        return set_err(h, FBP_E_UNSUPPORTED, "(2*n_max_shift_inds+1)^2 * n_templates must be < 2^32-1");
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = h->stream;
    TemplateSet ts = template_view(h);
    if (!likelihood_fits(ts))
        return set_err(h, FBP_E_UNSUPPORTED, "channels x template length exceed the likelihood tile (64 KiB LDS)");
    const int64_t nwin = (N + L - 1) / L;
    h->N = N;
    h->nwin = nwin;
    h->have_result = false;
    h->have_clips = false;
    HIP_TRY(h->V.ensure((size_t)C * N * 4), "alloc voltage");
    HIP_TRY(h->LIK.ensure((size_t)U * N * 4), "alloc likelihood");
    HIP_TRY(h->wmax_v.ensure((size_t)U * nwin * 4), "alloc window maxima");
    HIP_TRY(h->wmax_p.ensure((size_t)U * nwin * 2), "alloc window maxima");
    HIP_TRY(h->best_idx.ensure(nwin * 4), "alloc");
    HIP_TRY(h->best_lab.ensure(nwin * 4), "alloc");
    HIP_TRY(h->best_lik.ensure(nwin * 4), "alloc");
    HIP_TRY(h->obest_idx.ensure(nwin * 4), "alloc");
    HIP_TRY(h->obest_lab.ensure(nwin * 4), "alloc");
    HIP_TRY(h->check.ensure(nwin), "alloc");
    HIP_TRY(h->check_snap.ensure(nwin), "alloc");
    HIP_TRY(h->orc.ensure(nwin), "alloc");
    HIP_TRY(h->giveup.ensure(nwin), "alloc");
    HIP_TRY(h->has.ensure(nwin), "alloc");
    HIP_TRY(h->W.ensure(nwin * 4), "alloc");
    HIP_TRY(h->F.ensure(nwin * 4), "alloc");
    HIP_TRY(h->cand_idx.ensure(nwin * 4), "alloc");
    HIP_TRY(h->cand_lab.ensure(nwin * 4), "alloc");
    HIP_TRY(h->counters.ensure(CNT_N * 4), "alloc");
    HIP_TRY(h->blockbuf.ensure(((nwin + kCompactTile - 1) / kCompactTile + 1) * 4), "alloc");
    HIP_TRY(h->dirty.ensure(nwin * 8), "alloc");
    HIP_TRY(grow_keep(h->sp_idx, (size_t)std::max<int64_t>(nwin, 1024) * 4, 0, s), "alloc spikes");
    HIP_TRY(grow_keep(h->sp_lab, (size_t)std::max<int64_t>(nwin, 1024) * 4, 0, s), "alloc spikes");
    if (recheck && h->tab_n != (int64_t)n_max_shift) {
        bool ok = false;
        int rc = build_screen_tables(h, n_max_shift, &ok);
        if (rc != FBP_OK) return rc;
    }
    if (recheck && !h->screen_ok)
        return set_err(h, FBP_E_UNSUPPORTED, "the screened recheck's tables do not fit this segment");
    PassState st = pass_view(h);
    if (v_on_device) {
        launch_to_sample_major(Vsrc, ld, st.V, C, N, s);
    } else {
        // one set of pinned staging buffers (the stager's) serves every segment
        int rc = stage_in(stager, Vsrc, ld, N, st.V, C);
        if (rc != FBP_OK) return rc;
        if (stager != h) HIP_TRY(hipStreamSynchronize(stager->stream), "sync");
    }
    launch_fill_u32(reinterpret_cast<uint32_t*>(st.LIK), 0u, U * N, s);
    launch_fill_u32(st.best_idx, 0u, nwin, s);
    launch_fill_u32(st.best_lab, 0u, nwin, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(st.best_lik), 0u, nwin, s);
    launch_fill_u32(st.obest_idx, 0u, nwin, s);
    launch_fill_u32(st.obest_lab, 0u, nwin, s);
    launch_fill_u8(st.check, 0, nwin, s);
    launch_fill_u8(st.orc, 0, nwin, s);
    launch_fill_u8(st.giveup, 0, nwin, s);
    launch_fill_u32(st.counters, 0u, CNT_N, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(st.dirty), 0u, 2 * nwin, s);
    launch_iota(st.W, nwin, s);
    HIP_TRY(hipStreamSynchronize(s), "sync");
    return FBP_OK;
}

// Side streams (default priority) and their events, owned for the length of a
// batch: the selection classes of one screen round run on them side by side.
struct SideStreams {
    hipStream_t st[4] = {};
    hipEvent_t ev[4] = {};
    hipEvent_t fork = nullptr;
    ~SideStreams()
    {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& x : st)
            if (x) (void)hipStreamDestroy(x);
        if (fork) (void)hipEventDestroy(fork);
    }
};

// Launch tables of one upload, built on the host in pinned memory and copied
// to the device in one transfer.
struct SegTables {
    SegLaunch* host = nullptr;  // pinned
    SegLaunch* dev = nullptr;
    size_t cap = 0, used = 0;
    // entries of one launch start at the returned offset
    size_t begin() const { return used; }
};

struct BatchBufs {
    std::vector<SegDesc> desc;
    SegDesc* d_desc = nullptr;
    SegDyn* h_dyn = nullptr;  // pinned
    SegDyn* d_dyn = nullptr;
    SegTables tabs[3];
    uint16_t* eob[2] = {};
    uint32_t* d_cnt = nullptr;
    uint32_t* h_cnt = nullptr;  // pinned
    float* Bbuf = nullptr;
    ~BatchBufs()
    {
        if (d_desc) (void)hipFree(d_desc);
        if (h_dyn) (void)hipHostFree(h_dyn);
        if (d_dyn) (void)hipFree(d_dyn);
        for (auto& t : tabs) {
            if (t.host) (void)hipHostFree(t.host);
            if (t.dev) (void)hipFree(t.dev);
        }
        for (auto& e : eob)
            if (e) (void)hipFree(e);
        if (d_cnt) (void)hipFree(d_cnt);
        if (h_cnt) (void)hipHostFree(h_cnt);
        if (Bbuf) (void)hipFree(Bbuf);
    }
};

// One pass of a batch as a hipGraph (FBP_FLAG_BATCH_GRAPH): launch tables
// whose blocks are fixed when the graph is built, for the segments converging then
// (at most R[g] listed windows each), their block-to-entry maps, and the
// executable graph.  k_seg_plan writes the entries' item counts on the device
// as the pass goes, so a replay runs the whole pass without a host read.
struct PassGraph {
    hipGraphExec_t exec = nullptr;
    SegLaunch* host = nullptr;  // pinned
    SegLaunch* dev = nullptr;
    size_t cap = 0;
    uint16_t* eob = nullptr;
    size_t eob_cap = 0;
    std::vector<uint32_t> R;  // listed windows a segment's launches hold (0: not in the graph)
    uint64_t sumR = 0;
    SegPlan plan{};
    int64_t builds = 0, replays = 0;
    void drop()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    ~PassGraph()
    {
        drop();
        if (host) (void)hipHostFree(host);
        if (dev) (void)hipFree(dev);
        if (eob) (void)hipFree(eob);
    }
};

// Append one launch's entries: segment g gets nblk(g) blocks for n(g) items
// (segments with no items are left out).  Returns the entry count.
template <class F>
static uint32_t seg_table(SegTables& t, const std::vector<uint32_t>& segs, F&& f, uint32_t* nblocks)
{
    uint32_t b = 0, k = 0;
    for (uint32_t g : segs) {
        SegLaunch e{};
        if (!f(g, e) || e.n == 0 || e.nblk == 0) continue;
        e.seg = g;
        e.blk0 = b;
        b += e.nblk;
        t.host[t.used + k++] = e;
    }
    *nblocks = b;
    return k;
}

extern "C" {

int fbp_create(int device, fbp_handle** out)
{
    if (!out) return FBP_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FBP_E_NODEVICE;
    if (device < 0 || device >= ndev) return FBP_E_INVALID;
    fbp_handle* h = new fbp_handle();
    h->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_counters, CNT_N * sizeof(uint32_t), hipHostMallocDefault);
    if (e != hipSuccess) {
        delete h;
        return FBP_E_HIP;
    }
    *out = h;
    return FBP_OK;
}

int fbp_destroy(fbp_handle* h)
{
    if (!h) return FBP_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    DevBuf* bufs[] = {&h->T, &h->rows, &h->nch, &h->off, &h->chan, &h->nz, &h->bias, &h->thr, &h->V, &h->LIK, &h->wmax_v, &h->wmax_p,
                      &h->best_idx, &h->best_lab, &h->best_lik, &h->obest_idx, &h->obest_lab, &h->check,
                      &h->check_snap, &h->orc, &h->giveup, &h->has, &h->W, &h->F, &h->cand_idx, &h->cand_lab,
                      &h->counters, &h->blockbuf, &h->part_v, &h->part_k, &h->sp_idx, &h->sp_lab, &h->clips,
                      &h->pid, &h->pair_l, &h->pair_u, &h->G, &h->minG, &h->minGb, &h->Ntr, &h->rowL1, &h->SS, &h->Bbuf,
                      &h->mf_frag, &h->mf_fragl, &h->mf_units, &h->mf_nkb, &h->mf_nblk, &h->mf_aoff, &h->mf_chb, &h->mf_T1,
                      &h->mf_mbits, &h->mf_vs, &h->mf_tileof, &h->mf_nunit, &h->cmask, &h->dirty, &h->mma_frag, &h->mma_tsc,
                      &h->mma_T1, &h->mma_scal, &h->mma_q, &h->mma_bmax, &h->mma_gmask, &h->Til, &h->xcand, &h->Fs, &h->lhist, &h->rcnt, &h->lorder,
                      &h->mma_qtmp, &h->mma_qch, &h->mma_rslot};
    for (DevBuf* b : bufs) b->release();
    for (int b = 0; b < kStageBufs; b++) {
        h->dstage[b].release();
        if (h->pinned[b]) (void)hipHostFree(h->pinned[b]);
        if (h->stage_ev[b]) (void)hipEventDestroy(h->stage_ev[b]);
        if (h->stage_tr[b]) (void)hipEventDestroy(h->stage_tr[b]);
    }
    if (h->stage_mark) (void)hipEventDestroy(h->stage_mark);
    if (h->cstream) (void)hipStreamDestroy(h->cstream);
    h->pool.reset();
    if (h->h_counters) (void)hipHostFree(h->h_counters);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return FBP_OK;
}

const char* fbp_last_error(const fbp_handle* h) { return h ? h->err.c_str() : "null handle"; }

int fbp_device_memory(fbp_handle* h, int64_t* total_bytes)
{
    if (!h || !total_bytes) return FBP_E_INVALID;
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    *total_bytes = (int64_t)tot;
    return FBP_OK;
}

int fbp_set_templates(fbp_handle* h, const float* templates, int32_t U, int32_t C, int32_t L, const float* bias,
                      const float* thresholds)
{
    if (!h) return FBP_E_INVALID;
    if (!templates || !bias || !thresholds || U <= 0 || C <= 0 || L <= 0)
        return set_err(h, FBP_E_INVALID, "templates, bias and thresholds must be non-empty");
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    const int Lp = (L + 3) & ~3;
    std::vector<uint8_t> nz((size_t)U * C, 0);
    std::vector<int> nch(U, 0), off(U, 0), chan;
    std::vector<float> rows;
    for (int u = 0; u < U; u++) {
        off[u] = (int)chan.size();
        for (int c = 0; c < C; c++) {
            const float* r = templates + ((size_t)u * C + c) * L;
            bool any = false;
            for (int i = 0; i < L; i++)
                if (r[i] != 0.0f) { any = true; break; }
            if (!any) continue;
            nz[(size_t)u * C + c] = 1;
            nch[u]++;
            chan.push_back(c);
            size_t base = rows.size();
            rows.resize(base + Lp, 0.0f);
            std::memcpy(rows.data() + base, r, L * sizeof(float));
        }
    }
    if (chan.empty()) {  // keep buffers non-empty
        chan.push_back(0);
        rows.resize(Lp, 0.0f);
    }
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    HIP_TRY(h->T.ensure((size_t)U * C * L * 4), "alloc templates");
    HIP_TRY(h->rows.ensure(rows.size() * 4), "alloc rows");
    HIP_TRY(h->nch.ensure(U * 4), "alloc nch");
    HIP_TRY(h->off.ensure(U * 4), "alloc off");
    HIP_TRY(h->chan.ensure(chan.size() * 4), "alloc chan");
    HIP_TRY(h->nz.ensure(nz.size()), "alloc nz");
    HIP_TRY(h->bias.ensure(U * 4), "alloc bias");
    HIP_TRY(h->thr.ensure(U * 4), "alloc thr");
    HIP_TRY(hipMemcpy(h->T.p, templates, (size_t)U * C * L * 4, hipMemcpyHostToDevice), "copy templates");
    HIP_TRY(hipMemcpy(h->rows.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice), "copy rows");
    HIP_TRY(hipMemcpy(h->nch.p, nch.data(), U * 4, hipMemcpyHostToDevice), "copy nch");
    HIP_TRY(hipMemcpy(h->off.p, off.data(), U * 4, hipMemcpyHostToDevice), "copy off");
    HIP_TRY(hipMemcpy(h->chan.p, chan.data(), chan.size() * 4, hipMemcpyHostToDevice), "copy chan");
    HIP_TRY(hipMemcpy(h->nz.p, nz.data(), nz.size(), hipMemcpyHostToDevice), "copy nz");
    HIP_TRY(hipMemcpy(h->bias.p, bias, U * 4, hipMemcpyHostToDevice), "copy bias");
    HIP_TRY(hipMemcpy(h->thr.p, thresholds, U * 4, hipMemcpyHostToDevice), "copy thr");
    {
        // channel bit masks of the templates (used when C <= 64)
        std::vector<uint64_t> cm(U, 0ull);
        for (int u = 0; u < U; u++)
            for (int c = 0; c < C && c < 64; c++)
                if (nz[(size_t)u * C + c]) cm[u] |= 1ull << c;
        HIP_TRY(h->cmask.ensure((size_t)U * 8), "alloc channel masks");
        HIP_TRY(hipMemcpy(h->cmask.p, cm.data(), (size_t)U * 8, hipMemcpyHostToDevice), "copy channel masks");
    }
    h->U = U;
    h->C = C;
    h->L = L;
    h->Lp = Lp;
    h->h_nch = nch;
    h->h_nz = nz;
    h->h_rows = rows;
    h->h_off = off;
    h->mma_fmt = -1;
    h->Cmax = 0;
    for (int u = 0; u < U; u++) h->Cmax = std::max(h->Cmax, nch[u]);
    h->h_rowL1.assign(chan.size(), 0.0f);
    h->h_SS.assign(U, 0.0f);
    for (int u = 0; u < U; u++) {
        double ss = 0.0;
        for (int k = 0; k < nch[u]; k++) {
            const int r = off[u] + k;
            double l1 = 0.0;
            for (int i = 0; i < L; i++) {
                const double t = rows[(size_t)r * Lp + i];
                l1 += std::fabs(t);
                ss += t * t;
            }
            h->h_rowL1[r] = (float)(l1 * (1.0 + 1e-6));
        }
        h->h_SS[u] = (float)(ss * (1.0 + 1e-6));
    }
    HIP_TRY(h->rowL1.ensure(std::max<size_t>(h->h_rowL1.size(), 1) * 4), "alloc norms");
    HIP_TRY(h->SS.ensure((size_t)U * 4), "alloc norms");
    HIP_TRY(hipMemcpy(h->rowL1.p, h->h_rowL1.data(), h->h_rowL1.size() * 4, hipMemcpyHostToDevice), "copy norms");
    HIP_TRY(hipMemcpy(h->SS.p, h->h_SS.data(), (size_t)U * 4, hipMemcpyHostToDevice), "copy norms");
    h->tab_n = -1;
    h->mf_built = false;
    h->have_templates = true;
    h->have_result = false;
    return FBP_OK;
}

int fbp_run(fbp_handle* h, const float* Vsrc, int64_t ld, int64_t N, int32_t v_on_device, uint32_t n_max_shift,
            uint32_t flags)
{
    if (!h) return FBP_E_INVALID;
    if (!h->have_templates) return set_err(h, FBP_E_STATE, "fbp_set_templates must be called first");
    if (!Vsrc || N <= 0 || ld < N) return set_err(h, FBP_E_INVALID, "voltage must be non-empty with ld >= N");
    RunGuard guard(false);
    if (!guard.ok) return set_err(h, FBP_E_STATE, "a batch with FBP_FLAG_BATCH_GRAPH is running in this process");
    const int64_t L = h->L, C = h->C, U = h->U;
    if (N >= (int64_t)0xffffffffLL) return set_err(h, FBP_E_UNSUPPORTED, "chunk longer than 2^32-1 samples");
    if (L > 65535) return set_err(h, FBP_E_UNSUPPORTED, "template length above 65535 samples");
    const uint64_t S = 2ull * n_max_shift + 1ull;
    if ((flags & FBP_FLAG_OVERLAP_RECHECK) && S * S * (uint64_t)U >= 0xffffffffull)
        return set_err(h, FBP_E_UNSUPPORTED, "(2*n_max_shift_inds+1)^2 * n_templates must be < 2^32-1");
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = h->stream;
    TemplateSet ts = template_view(h);
    if (!likelihood_fits(ts))
        return set_err(h, FBP_E_UNSUPPORTED, "channels x template length exceed the likelihood tile (64 KiB LDS)");

    const int64_t nwin = (N + L - 1) / L;
    h->N = N;
    h->nwin = nwin;
    h->have_result = false;
    h->have_clips = false;
    HIP_TRY(h->V.ensure((size_t)C * N * 4), "alloc voltage");
    HIP_TRY(h->LIK.ensure((size_t)U * N * 4), "alloc likelihood");
    HIP_TRY(h->wmax_v.ensure((size_t)U * nwin * 4), "alloc window maxima");
    HIP_TRY(h->wmax_p.ensure((size_t)U * nwin * 2), "alloc window maxima");
    HIP_TRY(h->best_idx.ensure(nwin * 4), "alloc");
    HIP_TRY(h->best_lab.ensure(nwin * 4), "alloc");
    HIP_TRY(h->best_lik.ensure(nwin * 4), "alloc");
    HIP_TRY(h->obest_idx.ensure(nwin * 4), "alloc");
    HIP_TRY(h->obest_lab.ensure(nwin * 4), "alloc");
    HIP_TRY(h->check.ensure(nwin), "alloc");
    HIP_TRY(h->check_snap.ensure(nwin), "alloc");
    HIP_TRY(h->orc.ensure(nwin), "alloc");
    HIP_TRY(h->giveup.ensure(nwin), "alloc");
    HIP_TRY(h->has.ensure(nwin), "alloc");
    HIP_TRY(h->W.ensure(nwin * 4), "alloc");
    HIP_TRY(h->F.ensure(nwin * 4), "alloc");
    HIP_TRY(h->cand_idx.ensure(nwin * 4), "alloc");
    HIP_TRY(h->cand_lab.ensure(nwin * 4), "alloc");
    HIP_TRY(h->counters.ensure(CNT_N * 4), "alloc");
    HIP_TRY(h->blockbuf.ensure(((nwin + kCompactTile - 1) / kCompactTile + 1) * 4), "alloc");
    HIP_TRY(h->dirty.ensure(nwin * 8), "alloc");
    HIP_TRY(grow_keep(h->sp_idx, (size_t)std::max<int64_t>(nwin, 1024) * 4, 0, s), "alloc spikes");
    HIP_TRY(grow_keep(h->sp_lab, (size_t)std::max<int64_t>(nwin, 1024) * 4, 0, s), "alloc spikes");

    bool mixed = (flags & FBP_FLAG_PRECISION_MIXED) != 0;
    const uint32_t lik_batch = 65536;
    const bool recheck = (flags & FBP_FLAG_OVERLAP_RECHECK) != 0;
    // split-fp16 overlap screen on the matrix cores unless FP32 B is asked for
    bool screen_mf = recheck && !(flags & (FBP_FLAG_RECHECK_LITERAL | FBP_FLAG_SCREEN_FP32)) &&
                     S <= (uint64_t)kScreenMfmaMaxS;
    if (mixed || screen_mf) {
        if (!h->mf_built) {
            int rc = build_mf_tables(h);
            if (rc != FBP_OK) return rc;
        }
        // outside the matrix-core tables' limits the FP32 kernels run (same results),
        // and the first pass takes the FP32 kernel where the tiles' padding makes it faster
        mixed = mixed && h->mf_ok && (h->mf_pays || (flags & FBP_FLAG_LIKELIHOOD_MFMA) != 0);
        screen_mf = screen_mf && h->mf_ok;
    }
    // per-template BF16 / FP8 matrix-core likelihood on every pass; a window's
    // 8-position groups to rescore fit one byte (L <= 64), else the FP32 kernel runs
    const int mma_req = (flags & FBP_FLAG_PRECISION_FP8) ? 1 : ((flags & FBP_FLAG_PRECISION_BF16) ? 0 : -1);
    const bool mma = mma_req >= 0 && L <= 64;
    const bool mma_every = mma && (flags & FBP_FLAG_MMA_EVERY_PASS) != 0;
    if (mma) {
        mixed = false;
        if (h->mma_fmt != mma_req) {
            int rc = build_mma_tables(h, mma_req);
            if (rc != FBP_OK) return rc;
        }
    }
    Timer tm(s, (flags & FBP_FLAG_TIMING) != 0);
    EventPair evp;  // destroyed on every return path
    HIP_TRY(hipEventCreate(&evp.a), "event");
    HIP_TRY(hipEventCreate(&evp.b), "event");
    hipEvent_t ev0 = evp.a, ev1 = evp.b;
    HIP_TRY(hipEventRecord(ev0, s), "event");

    fbp_stats stats{};
    int64_t nz_sum = 0;
    for (int u = 0; u < h->U; u++) nz_sum += h->h_nch[u];
    const int reach = (int)(3 + (2 * (int64_t)n_max_shift + L) / L);
    const uint32_t nchunk = recheck ? recheck_literal_chunks(ts, n_max_shift) : 0;
    bool use_screen = false;
    if (recheck && !(flags & FBP_FLAG_RECHECK_LITERAL)) {
        if (h->tab_n != (int64_t)n_max_shift) {
            int rc = build_screen_tables(h, n_max_shift, &use_screen);
            if (rc != FBP_OK) return rc;
        }
        use_screen = h->screen_ok;
    }
    // (no more rows than the chunk has windows)
    const uint32_t screen_batch =
        use_screen ? (uint32_t)std::max<size_t>(1, std::min<size_t>({kScreenBatchWindows, (size_t)nwin,
                                                                     kScreenBatchBytes / ((size_t)U * S * 4)}))
                   : 0;
    screen_mf = screen_mf && use_screen;
    if (use_screen) HIP_TRY(h->Bbuf.ensure((size_t)screen_batch * U * S * 4), "alloc screen buffer");
    if (use_screen) HIP_TRY(h->xcand.ensure((size_t)screen_batch * kXStride * 4), "alloc rescoring records");
    if (use_screen) HIP_TRY(h->Fs.ensure((size_t)nwin * 4), "alloc flagged windows by label");
    if (use_screen) HIP_TRY(h->lhist.ensure((size_t)U * 4), "alloc label counts");
    PassState st = pass_view(h);
    if (!v_on_device) {
        // the voltage copies need not wait for the state fills below
        int rc = ensure_staging(h);
        if (rc != FBP_OK) return rc;
        HIP_TRY(hipEventRecord(h->stage_mark, s), "event");
    }
    // Initial state: binary_pursuit_parallel.py:528-655 (likelihood zeros,
    // best_* zeros, flags zeros, window list 0..nwin-1).
    launch_fill_u32(reinterpret_cast<uint32_t*>(st.LIK), 0u, U * N, s);
    launch_fill_u32(st.best_idx, 0u, nwin, s);
    launch_fill_u32(st.best_lab, 0u, nwin, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(st.best_lik), 0u, nwin, s);
    launch_fill_u32(st.obest_idx, 0u, nwin, s);
    launch_fill_u32(st.obest_lab, 0u, nwin, s);
    launch_fill_u8(st.check, 0, nwin, s);
    launch_fill_u8(st.orc, 0, nwin, s);
    launch_fill_u8(st.giveup, 0, nwin, s);
    launch_fill_u32(st.counters, 0u, CNT_N, s);
    if (st.dirty) launch_fill_u32(reinterpret_cast<uint32_t*>(st.dirty), 0u, 2 * nwin, s);
    launch_iota(st.W, nwin, s);
    // With host voltage and the vector-unit first pass, the first pass's
    // likelihood runs on each staged piece's complete windows while the later
    // pieces are still copied: window w reads the samples [wL, wL + 2L - 1).
    const bool early_lik = !v_on_device && !mma && !mixed && !(flags & FBP_FLAG_FIRST_PASS_ONLY);
    int64_t early_w = 0;  // windows whose first-pass likelihood is queued
    // The screened recheck of the flagged windows sp.F[0 .. n) (n read on the
    // device from sp.counters[CNT_F], at most nF_hi): grouped by label into Fs,
    // then correlations, selection and deferred rescoring in batches.
    auto screened = [&](const PassState& sp, uint32_t nF_hi, uint32_t* Fs, hipStream_t q) {
        launch_sort_by_label(sp, nF_hi, U, Fs, h->lhist.as<uint32_t>(), q, h->lorder.as<uint32_t>());
        PassState sf = sp;
        sf.F = Fs;
        ScreenTables tb = tables_view(h);
        if (screen_mf) {
            tb.kappa2 = h->scr_k2;
            tb.kappa3 = h->scr_k3;
            tb.kappa4 = h->scr_k4;
        }
        for (uint32_t f0 = 0; f0 < nF_hi; f0 += screen_batch) {
            const uint32_t nb = std::min(screen_batch, nF_hi - f0);
            if (screen_mf) {
                launch_screen_mfma(ts, sf, mf_view(h), h->mf_fragl.p, tb.Ntr, tb.S, tb.n, f0, nb, h->Bbuf.as<float>(), q);
            } else {
                launch_screen_corr(ts, sf, tb, f0, nb, h->Bbuf.as<float>(), q);
            }
            tm.begin(TM_SELECT);
            const bool deferred = launch_screen_select(ts, sf, tb, f0, nb, h->Bbuf.as<float>(),
                                                       (flags & FBP_FLAG_SELECT_STREAMING) != 0, q,
                                                       h->xcand.as<uint32_t>());
            tm.end(TM_SELECT);
            stats.n_select_launches++;
            if (deferred) {
                tm.begin(TM_RESCORE);
                launch_screen_rescore(ts, sf, tb, f0, nb, h->xcand.as<uint32_t>(), q);
                tm.end(TM_RESCORE);
                stats.n_rescore_launches++;
            }
        }
    };
    // With host voltage, the first pass's window maxima and screened recheck
    // also run region by region (kRegionWindows windows each) as soon as a
    // region's windows and `reach` windows past it have their likelihood; the
    // regions' flagged lists are merged after the last one.  (On a stream of
    // their own beside the likelihood and the transposes they made the step
    // slower: 571 -> 596 ms at 64 channels, 600 s.)
    constexpr int64_t kRegionWindows = 16384;
    const bool early_rc = early_lik && use_screen;
    int64_t rdone = 0;  // windows whose region is queued
    int nreg = 0;
    if (early_rc) {
        HIP_TRY(h->rcnt.ensure((size_t)((nwin + kRegionWindows - 1) / kRegionWindows) * CNT_N * 4), "alloc region counters");
        launch_fill_u32(h->rcnt.as<uint32_t>(), 0u, ((nwin + kRegionWindows - 1) / kRegionWindows) * CNT_N, s);
    }
    auto region = [&](int64_t w1) -> int {
        const uint32_t n = (uint32_t)(w1 - rdone);
        PassState sp = st;
        sp.W = st.W + rdone;
        sp.F = st.F + rdone;
        sp.counters = h->rcnt.as<uint32_t>() + (size_t)nreg * CNT_N;
        const hipStream_t q = s;
        tm.begin(TM_MAX);
        launch_window_max(ts, sp, n, q);
        tm.end(TM_MAX);
        tm.begin(TM_OTHER);
        launch_flag_compact(sp, n, h->blockbuf.as<uint32_t>(), q);
        tm.end(TM_OTHER);
        tm.begin(TM_RECHECK);
        screened(sp, n, h->Fs.as<uint32_t>() + rdone, q);
        tm.end(TM_RECHECK);
        nreg++;
        rdone = w1;
        return FBP_OK;
    };
    auto early = [&](int64_t x) -> int {
        const int64_t whi = x >= N ? nwin : (x >= 2 * L - 1 ? std::min<int64_t>(nwin, (x - 2 * L + 1) / L + 1) : 0);
        if (whi <= early_w) return FBP_OK;
        PassState sp = st;
        sp.W = st.W + early_w;
        tm.begin(TM_LIK);
        launch_likelihood(ts, sp, (uint32_t)(whi - early_w), nullptr, s);
        tm.end(TM_LIK);
        early_w = whi;
        // regions whose windows, and `reach` windows past them, have their likelihood
        while (early_rc && rdone < nwin) {
            const int64_t w1 = std::min<int64_t>(nwin, rdone + kRegionWindows);
            if (early_w < std::min<int64_t>(nwin, w1 + reach)) break;
            const int rc = region(w1);
            if (rc != FBP_OK) return rc;
        }
        return FBP_OK;
    };
    // the residual is kept sample-major ([N][C]): a window is one contiguous block
    if (v_on_device) {
        launch_to_sample_major(Vsrc, ld, st.V, C, N, s);
    } else {
        int rc = stage_in(h, Vsrc, ld, N, st.V, -1, early_lik ? early : std::function<int(int64_t)>(), true);
        if (rc != FBP_OK) return rc;
        if (early_rc) {
            if (rdone != nwin) return set_err(h, FBP_E_HIP, "first-pass regions incomplete");
            tm.begin(TM_OTHER);
            launch_merge_regions(st.F, h->Fs.as<uint32_t>(), h->rcnt.as<uint32_t>(), nreg, (uint32_t)kRegionWindows,
                                 st.counters, s);
            tm.end(TM_OTHER);
        }
    }
    if (mixed) launch_absmax(st.V, (int64_t)C * N, h->mf_mbits.as<uint32_t>(), h->mf_vs.as<float>(), s);
    MmaVoltage mv{};
    if (mma) {
        // channel-major quantized copy of the residual (rows read up to x + 63 + 32 past a window start)
        const int64_t ldq = ((N + L + 96 + 63) / 64) * 64;
        HIP_TRY(h->mma_q.ensure((size_t)C * ldq * (mma_req == 0 ? 2 : 1)), "alloc quantized voltage");
        HIP_TRY(h->mma_bmax.ensure((size_t)C * (ldq / 16) * 4), "alloc block maxima");
        HIP_TRY(h->mma_gmask.ensure((size_t)nwin * U), "alloc rescore groups");
        // [0] max |V| bits, [1] unused scale, [2] voltage scale
        float* scal = h->mma_scal.as<float>();
        if (mma_req == 1) {
            // FP8: the largest |V| into [64, 128), room below 448 for what the residual adds
            launch_absmax(st.V, (int64_t)C * N, reinterpret_cast<uint32_t*>(scal), scal + 1, s);
            launch_vscale_to(reinterpret_cast<uint32_t*>(scal), scal + 2, 6, s);
        } else {
            launch_fill_u32(reinterpret_cast<uint32_t*>(scal + 2), 0x3f800000u, 1, s);  // BF16: 1.0
        }
        mv.q = h->mma_q.p;
        mv.bmax = h->mma_bmax.as<float>();
        mv.vs = scal + 2;
        mv.ldq = ldq;
        mv.nbq = ldq / 16;
        launch_quantize(st.V, N, (int)C, mv, mma_req, s);
    }

    uint32_t nW = (uint32_t)nwin;
    int64_t total = 0;
    int passes = 0;
    int64_t pairs_total = 0;
    h->pass_log.clear();
    for (;;) {
        // (the staged first pass has counted its pairs since the counters were cleared)
        const bool staged_pass = passes == 0 && early_lik;
        if (!staged_pass) launch_fill_u32(st.counters + CNT_PAIRS, 0u, 1, s);
        stats.windows_listed += nW;
        stats.likelihood_positions += (int64_t)nW * L * U;
        stats.likelihood_flops += 2.0 * (double)nW * (double)L * (double)nz_sum * (double)L;
        tm.begin(TM_LIK);
        // the first pass computes every window; later passes only the templates
        // whose channels changed in a window's samples since it was computed
        const uint64_t* dirty = passes > 0 ? st.dirty : nullptr;
        // The matrix cores take the dense first pass.  Later passes recompute a
        // few templates per window, which the vector-unit kernel skips one by one
        // (the matrix-core kernel only by 16-template tiles over 4 windows).
        // Both kernels leave exact values wherever a decision can depend on them.
        if (mma && (dirty == nullptr || mma_every)) {
            // matrix-core approximation of every computed pair (the dense kernel on the
            // first pass where it applies), then the exact FP32 chain on the 8-position
            // groups it could not settle
            const MmaTables mt = mma_view(h);
            if (dirty != nullptr || mma_every || !launch_likmma_dense(ts, st, mt, mv, mma_req, h->mma_gmask.as<uint8_t>(), s))
                launch_likmma(ts, st, mt, mv, mma_req, nW, h->mma_gmask.as<uint8_t>(), dirty, s);
            stats.n_likelihood_mfma_launches++;
            if (!(flags & FBP_FLAG_NO_RESCORE)) launch_likelihood(ts, st, nW, dirty, s, h->mma_gmask.as<uint8_t>());
        } else if (mma) {
            // later passes: the few changed pairs per window on the vector units
            launch_likelihood(ts, st, nW, dirty, s);
        } else if (mixed && dirty == nullptr) {
            const MfTables mt = mf_view(h);
            for (uint32_t w0 = 0; w0 < nW; w0 += lik_batch) {
                launch_likmfma(ts, st, mt, h->rowL1.as<float>(), h->mf_vs.as<float>(), h->mf_mbits.as<uint32_t>(),
                               dirty, w0, std::min(lik_batch, nW - w0), s);
                stats.n_likelihood_mfma_launches++;
            }
        } else if (!staged_pass) {
            launch_likelihood(ts, st, nW, dirty, s);
        }
        launch_clear_dirty(st, nW, s);
        tm.end(TM_LIK);
        stats.n_likelihood_launches++;
        if (flags & FBP_FLAG_FIRST_PASS_ONLY) {
            passes = 1;
            break;
        }
        // (the staged first pass has run its window maxima and screened recheck by regions)
        const bool region_pass = staged_pass && early_rc;
        if (!region_pass) {
            tm.begin(TM_MAX);
            launch_window_max(ts, st, nW, s);
            tm.end(TM_MAX);
        }
        uint32_t nF = 0;
        if (recheck) {
            if (!region_pass) {
                tm.begin(TM_OTHER);
                launch_flag_compact(st, nW, h->blockbuf.as<uint32_t>(), s);
                tm.end(TM_OTHER);
            }
            HIP_TRY(hipMemcpyAsync(h->h_counters, st.counters, CNT_N * 4, hipMemcpyDeviceToHost, s), "read counters");
            HIP_TRY(hipStreamSynchronize(s), "sync");
            nF = h->h_counters[CNT_F];
            stats.windows_flagged += nF;
            if (nF > 0) {
                tm.begin(TM_RECHECK);
                // the screened recheck and the reassignment take the flagged windows grouped by label
                // (by region, then label, after the region pass)
                PassState sf = st;
                if (use_screen) {
                    if (!region_pass) {
                        sf.F = h->Fs.as<uint32_t>();
                        screened(st, nF, sf.F, s);
                    }
                } else {
                    HIP_TRY(h->part_v.ensure((size_t)nF * nchunk * 4), "alloc recheck partials");
                    HIP_TRY(h->part_k.ensure((size_t)nF * nchunk * 4), "alloc recheck partials");
                    launch_recheck_literal(ts, st, n_max_shift, nF, h->part_v.as<float>(), h->part_k.as<uint32_t>(), s);
                }
                tm.end(TM_RECHECK);
                stats.n_recheck_launches++;
                tm.begin(TM_OTHER);
                launch_reassign(ts, sf, nF, s);
                tm.end(TM_OTHER);
            }
        }
        tm.begin(TM_PURSUIT);
        launch_pursuit(ts, st, nW, recheck && nF > 0 ? 1 : 0, s);
        tm.end(TM_PURSUIT);
        const size_t need = (size_t)(total + nW) * 4;
        HIP_TRY(grow_keep(h->sp_idx, need, (size_t)total * 4, s), "grow spikes");
        HIP_TRY(grow_keep(h->sp_lab, need, (size_t)total * 4, s), "grow spikes");
        tm.begin(TM_OTHER);
        launch_spike_compact(st, nW, h->sp_idx.as<uint32_t>(), h->sp_lab.as<uint32_t>(), (uint64_t)total,
                             h->blockbuf.as<uint32_t>(), s);
        launch_next_windows(st, h->blockbuf.as<uint32_t>(), s);
        tm.end(TM_OTHER);
        HIP_TRY(hipMemcpyAsync(h->h_counters, st.counters, CNT_N * 4, hipMemcpyDeviceToHost, s), "read counters");
        HIP_TRY(hipStreamSynchronize(s), "sync");
        passes++;
        pairs_total += h->h_counters[CNT_PAIRS];
        const uint32_t nadd = h->h_counters[CNT_SPIKES];
        h->pass_log.push_back(nadd);
        h->pass_log.push_back(nW);
        h->pass_log.push_back(nF);
        if (nadd == 0) break;
        tm.begin(TM_RESID);
        launch_residual(ts, st.V, N, h->sp_idx.as<uint32_t>() + total, h->sp_lab.as<uint32_t>() + total, nadd, reach, s);
        launch_mark_dirty(ts, st, h->sp_idx.as<uint32_t>() + total, h->sp_lab.as<uint32_t>() + total, nadd, s);
        if (mma && mma_every)
            launch_requantize(st.V, N, (int)C, (int)L, mv, mma_req, h->sp_idx.as<uint32_t>() + total, nadd, s);
        tm.end(TM_RESID);
        total += nadd;
        nW = h->h_counters[CNT_W];
    }
    if (flags & FBP_FLAG_ADJUSTED_CLIPS) {
        HIP_TRY(h->clips.ensure((size_t)std::max<int64_t>(total, 1) * C * L * 4), "alloc clips");
        launch_adjusted_clips(ts, st.V, N, h->sp_idx.as<uint32_t>(), h->sp_lab.as<uint32_t>(), (uint64_t)total,
                              h->clips.as<float>(), s);
        h->have_clips = true;
    }
    HIP_TRY(hipEventRecord(ev1, s), "event");
    HIP_TRY(hipGetLastError(), "kernel launch");
    HIP_TRY(hipEventSynchronize(ev1), "sync");
    float ms = 0;
    hipEventElapsedTime(&ms, ev0, ev1);
    stats.ms_total = ms;
    stats.ms_likelihood = tm.total(TM_LIK);
    stats.ms_window_max = tm.total(TM_MAX);
    stats.ms_recheck = tm.total(TM_RECHECK);
    stats.ms_pursuit = tm.total(TM_PURSUIT);
    stats.ms_residual = tm.total(TM_RESID);
    stats.ms_other = tm.total(TM_OTHER);
    stats.ms_select = tm.total(TM_SELECT);
    stats.ms_rescore = tm.total(TM_RESCORE);
    HIP_TRY(hipMemcpy(h->h_counters, st.counters, CNT_N * 4, hipMemcpyDeviceToHost), "read counters");
    stats.recheck_items = h->h_counters[CNT_EVAL];
    stats.recheck_overflow = h->h_counters[CNT_CAND];
    stats.likelihood_rescored = h->h_counters[CNT_LIKX];
    stats.screen_shared_pairs = h->h_counters[CNT_SURV];
    stats.screen_separable_pairs = h->h_counters[CNT_NSP];
    stats.passes = passes;
    stats.spikes = total;
    stats.likelihood_pairs_computed = pairs_total;
    h->stats = stats;
    h->n_spikes = total;
    h->have_result = true;
    return FBP_OK;
}

int fbp_run_batch(fbp_handle* const* hs, int32_t nseg, const float* const* V, const int64_t* ld, const int64_t* N,
                  int32_t v_on_device, const uint32_t* n_max_shift, uint32_t flags, int32_t block_cap,
                  int32_t lik_windows)
{
    if (!hs || nseg <= 0 || !V || !ld || !N || !n_max_shift) return FBP_E_INVALID;
    fbp_handle* h = hs[0];
    if (!h) return FBP_E_INVALID;
    if (nseg > 65535) return set_err(h, FBP_E_UNSUPPORTED, "at most 65535 segments per batch");
    for (int32_t g = 0; g < nseg; g++) {
        if (!hs[g]) return set_err(h, FBP_E_INVALID, "null segment handle");
        if (hs[g]->device != h->device) return set_err(h, FBP_E_INVALID, "segment handles on different devices");
        if (!hs[g]->have_templates) return set_err(hs[g], FBP_E_STATE, "fbp_set_templates must be called first");
        if (hs[g]->L != h->L) return set_err(h, FBP_E_UNSUPPORTED, "segments of one batch share the template length");
        for (int32_t k = 0; k < g; k++)
            if (hs[k] == hs[g]) return set_err(h, FBP_E_INVALID, "a handle appears twice in the batch");
    }
    // Graph replays are switched off: a one-lane batch with them ended in a host
    // segmentation fault whose cause is not known (the code below stays for when it is).
    if (flags & FBP_FLAG_BATCH_GRAPH)
        return set_err(h, FBP_E_UNSUPPORTED, "FBP_FLAG_BATCH_GRAPH is disabled in this build");
    RunGuard guard((flags & FBP_FLAG_BATCH_GRAPH) != 0);
    if (!guard.ok)
        return set_err(h, FBP_E_STATE, (flags & FBP_FLAG_BATCH_GRAPH)
                                           ? "FBP_FLAG_BATCH_GRAPH: another pursuit call is running in this process"
                                           : "a batch with FBP_FLAG_BATCH_GRAPH is running in this process");
    const bool recheck = (flags & FBP_FLAG_OVERLAP_RECHECK) != 0;
    const int L = h->L;
    for (int32_t g = 0; g < nseg; g++) {
        int rc = batch_prepare(hs[g], h, V[g], ld[g], N[g], v_on_device, n_max_shift[g], recheck);
        if (rc != FBP_OK) {
            if (hs[g] != h) h->err = hs[g]->err;
            return rc;
        }
    }
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = h->stream;
    EventPair evp;
    HIP_TRY(hipEventCreate(&evp.a), "event");
    HIP_TRY(hipEventCreate(&evp.b), "event");
    HIP_TRY(hipEventRecord(evp.a, s), "event");

    // per-segment descriptors and launch plans
    BatchBufs bb;
    bb.desc.resize(nseg);
    size_t lds_lik = 0, lds_corr = 0, bstride = 1;
    // selection classes by dynamic LDS (at most 4, 3, 2 and 1 blocks per CU of
    // 160 KiB): each class launches with its own largest block and LDS, so the
    // small segments are not held to the largest one's occupancy
    constexpr int kSelClasses = 4;
    const size_t cls_cap[kSelClasses] = {40 * 1024, 53 * 1024, 80 * 1024, SIZE_MAX};
    int cls_bd[kSelClasses] = {};
    size_t cls_lds[kSelClasses] = {};
    std::vector<int> scls(nseg, 0);
    std::vector<int64_t> nwin(nseg);
    for (int32_t g = 0; g < nseg; g++) {
        fbp_handle* hg = hs[g];
        SegDesc& d = bb.desc[g];
        d.ts = template_view(hg);
        d.st = pass_view(hg);
        d.tb = recheck ? tables_view(hg) : ScreenTables{};
        d.blockbuf = hg->blockbuf.as<uint32_t>();
        size_t ll = 0;
        seg_likelihood_lds(d.ts, &d.wb, &d.sv, &ll);
        if (lik_windows > 0 && lik_windows != d.wb) {
            // windows per likelihood block as asked, within the block's limit and 64 KiB of LDS
            const size_t per = ll / (size_t)d.wb;
            d.wb = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)lik_windows, (int64_t)kLikMaxWbHost,
                                                                 (int64_t)(kMaxLds / per)}));
            ll = per * (size_t)d.wb;
        }
        lds_lik = std::max(lds_lik, ll);
        d.reach = (int)(3 + (2 * (int64_t)n_max_shift[g] + L) / L);
        d.rp = 0;
        if (recheck) {
            lds_corr = std::max(lds_corr, screen_corr_lds(d.ts, d.tb.S, &d.rp));
            int bd = 0;
            size_t ls = 0;
            if (!screen_select_reg_plan(d.ts, d.tb, &bd, &ls))
                return set_err(h, FBP_E_UNSUPPORTED, "a segment is too large for the register-resident selection");
            int k = 0;
            while (ls > cls_cap[k]) k++;
            scls[g] = k;
            cls_bd[k] = std::max(cls_bd[k], bd);
            cls_lds[k] = std::max(cls_lds[k], ls);
            bstride = std::max(bstride, (size_t)d.ts.U * d.tb.S);
        }
        nwin[g] = hg->nwin;
    }
    if (lds_corr > kMaxLds) return set_err(h, FBP_E_UNSUPPORTED, "a segment's window exceeds the screen tile");
    HIP_TRY(hipMalloc((void**)&bb.d_desc, sizeof(SegDesc) * nseg), "alloc batch");
    HIP_TRY(hipMemcpy(bb.d_desc, bb.desc.data(), sizeof(SegDesc) * nseg, hipMemcpyHostToDevice), "copy batch");
    HIP_TRY(hipHostMalloc((void**)&bb.h_dyn, sizeof(SegDyn) * nseg, hipHostMallocDefault), "alloc batch");
    HIP_TRY(hipMalloc((void**)&bb.d_dyn, sizeof(SegDyn) * nseg), "alloc batch");
    for (auto& t : bb.tabs) {
        t.cap = (size_t)nseg * 16 + 64;
        HIP_TRY(hipHostMalloc((void**)&t.host, sizeof(SegLaunch) * t.cap, hipHostMallocDefault), "alloc batch");
        HIP_TRY(hipMalloc((void**)&t.dev, sizeof(SegLaunch) * t.cap), "alloc batch");
    }
    // screened recheck rows: flagged windows per screen round
    int64_t sum_nwin = 0;
    for (int64_t w : nwin) sum_nwin += w;
    const uint32_t rows_cap =
        recheck ? (uint32_t)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(sum_nwin, 1 << 20),
                                                                  (int64_t)(kScreenBatchBytes / (bstride * 4))))
                : 0;
    int64_t max_tiles = 0;
    for (int64_t w : nwin) max_tiles = std::max(max_tiles, (w + kCompactTile - 1) / kCompactTile);
    const uint32_t seg_cap = block_cap > 0 ? (uint32_t)std::min<int32_t>(block_cap, 1 << 16) : kSegBlockCap;
    const size_t eob_n = std::max<size_t>({(size_t)nseg * seg_cap, (size_t)rows_cap, (size_t)nseg * max_tiles, 1});
    HIP_TRY(hipMalloc((void**)&bb.eob[0], eob_n * 2), "alloc batch");
    HIP_TRY(hipMalloc((void**)&bb.eob[1], eob_n * 2), "alloc batch");
    HIP_TRY(hipMalloc((void**)&bb.d_cnt, (size_t)nseg * CNT_N * 4), "alloc batch");
    HIP_TRY(hipHostMalloc((void**)&bb.h_cnt, (size_t)nseg * CNT_N * 4, hipHostMallocDefault), "alloc batch");
    if (recheck) HIP_TRY(hipMalloc((void**)&bb.Bbuf, (size_t)rows_cap * bstride * 4), "alloc screen buffer");
    SideStreams side;
    if (recheck) {
        HIP_TRY(hipEventCreateWithFlags(&side.fork, hipEventDisableTiming), "event");
        for (int c = 1; c < kSelClasses; c++) {
            HIP_TRY(hipStreamCreateWithFlags(&side.st[c], hipStreamNonBlocking), "stream");
            HIP_TRY(hipEventCreateWithFlags(&side.ev[c], hipEventDisableTiming), "event");
        }
    }

    std::vector<uint32_t> nW(nseg), nF(nseg, 0), nadd(nseg, 0), active;
    std::vector<uint64_t> total(nseg, 0);
    std::vector<int> passes(nseg, 0);
    std::vector<fbp_stats> stats(nseg);
    for (int32_t g = 0; g < nseg; g++) {
        nW[g] = (uint32_t)nwin[g];
        active.push_back((uint32_t)g);
        hs[g]->pass_log.clear();
    }
    auto cdiv = [](uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); };
    auto capb = [&](uint64_t blocks) {
        return std::max<uint32_t>(1, (uint32_t)std::min<uint64_t>(blocks, seg_cap));
    };
    auto upload = [&](SegTables& t) -> hipError_t {
        if (t.used == 0) return hipSuccess;
        return hipMemcpyAsync(t.dev, t.host, sizeof(SegLaunch) * t.used, hipMemcpyHostToDevice, s);
    };
    auto read_counters = [&]() -> hipError_t {
        launch_seg_counters(bb.d_desc, (uint32_t)nseg, bb.d_cnt, s);
        hipError_t e = hipMemcpyAsync(bb.h_cnt, bb.d_cnt, (size_t)nseg * CNT_N * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e;
    };
    // ---- graph passes: every pass after the first replays one hipGraph ----
    const bool use_graph = (flags & FBP_FLAG_BATCH_GRAPH) != 0;
    PassGraph pg;
    pg.R.assign(nseg, 0);
    // Build the graph of one pass for the segments in `active`, segment g's
    // launches sized for R[g] listed windows (nodes added one by one: no stream
    // capture, so other host threads' calls cannot interleave with it).
    auto graph_build = [&](const std::vector<uint32_t>& R, uint64_t sumR) -> int {
        pg.drop();
        pg.R = R;
        pg.sumR = sumR;
        if (pg.host == nullptr) {
            pg.cap = (size_t)nseg * 16 + 64;
            HIP_TRY(hipHostMalloc((void**)&pg.host, sizeof(SegLaunch) * pg.cap, hipHostMallocDefault), "alloc graph");
            HIP_TRY(hipMalloc((void**)&pg.dev, sizeof(SegLaunch) * pg.cap), "alloc graph");
        }
        pg.plan = SegPlan{};
        pg.plan.recheck = recheck ? 1 : 0;
        // rows of each segment's flagged windows in Bbuf
        std::vector<uint32_t> row0(nseg, 0);
        uint32_t rows = 0;
        for (uint32_t g : active) {
            row0[g] = rows;
            rows += R[g];
        }
        struct GL {
            size_t o;
            uint32_t n, nb;
            size_t eo;
        };
        size_t used = 0, eused = 0;
        bool over = false;
        auto add = [&](int src, int phase, const std::vector<uint32_t>& segs, auto&& nblk_of) -> GL {
            GL l{used, 0, 0, eused};
            if (used + segs.size() > pg.cap) {
                over = true;
                return l;
            }
            for (uint32_t g : segs) {
                SegLaunch e{};
                e.seg = g;
                e.blk0 = l.nb;
                e.nblk = std::max<uint32_t>(1, nblk_of(g));
                e.g0 = row0[g];
                pg.host[used + l.n++] = e;
                l.nb += e.nblk;
            }
            if (l.n == 0) return l;
            SegPlan& p = pg.plan;
            if (p.nr > 0 && p.src[p.nr - 1] == src && p.phase[p.nr - 1] == phase &&
                p.off[p.nr - 1] + p.cnt[p.nr - 1] == (uint32_t)used) {
                p.cnt[p.nr - 1] += l.n;
            } else if (p.nr < kPlanRanges) {
                p.off[p.nr] = (uint32_t)used;
                p.cnt[p.nr] = l.n;
                p.src[p.nr] = (uint8_t)src;
                p.phase[p.nr] = (uint8_t)phase;
                p.nr++;
            } else {
                over = true;
            }
            used += l.n;
            eused += l.nb;
            return l;
        };
        // the screen's blocks per segment: at most R[g], about 8192 in all
        auto scr_blocks = [&](uint32_t g) {
            return std::min<uint32_t>(R[g], std::max<uint32_t>(64, cdiv((uint64_t)R[g] * 8192, std::max<uint64_t>(sumR, 1))));
        };
        const GL lik = add(SRC_NW, 0, active, [&](uint32_t g) { return capb(cdiv(R[g], (uint32_t)bb.desc[g].wb)); });
        const GL clr = add(SRC_NW, 0, active, [&](uint32_t g) { return capb(cdiv(R[g], kBlock)); });
        const GL mx = add(SRC_NW, 0, active, [&](uint32_t g) { return capb(cdiv(R[g], kBlock / 64)); });
        GL flg{}, corr{}, sel[kSelClasses]{}, re{};
        if (recheck) flg = add(SRC_NW, 0, active, [&](uint32_t g) { return cdiv(R[g], kCompactTile); });
        const GL pr = add(SRC_NW, 0, active, [&](uint32_t g) { return capb(cdiv(R[g], kBlock)); });
        const GL sp = add(SRC_NW, 0, active, [&](uint32_t g) { return cdiv(R[g], kCompactTile); });
        const GL snap = add(SRC_NWIN, 0, active, [&](uint32_t g) { return capb(cdiv(nwin[g], 4 * kBlock)); });
        const GL nx = add(SRC_NWIN, 0, active, [&](uint32_t g) { return cdiv(nwin[g], kCompactTile); });
        if (recheck) {
            corr = add(SRC_NF, 1, active, scr_blocks);
            for (int c = 0; c < kSelClasses; c++) {
                std::vector<uint32_t> segs;
                for (uint32_t g : active)
                    if (scls[g] == c) segs.push_back(g);
                sel[c] = add(SRC_NF, 1, segs, [&](uint32_t g) { return R[g]; });  // one block per window
            }
            re = add(SRC_NF, 1, active, [&](uint32_t g) { return capb(cdiv(R[g], kBlock)); });
        }
        const GL dec = add(SRC_DECIDE, 1, active, [&](uint32_t g) { return capb(cdiv(R[g], kBlock)); });
        const GL res = add(SRC_NADD, 2, active, [&](uint32_t g) { return capb(R[g]); });
        const GL mk = add(SRC_NADD, 2, active, [&](uint32_t g) { return capb(cdiv(R[g], kBlock)); });
        if (over) return set_err(h, FBP_E_UNSUPPORTED, "graph pass table overflow");
        if (eused > pg.eob_cap) {
            if (pg.eob) HIP_TRY(hipFree(pg.eob), "free graph");
            pg.eob = nullptr;
            pg.eob_cap = eused + eused / 4;
            HIP_TRY(hipMalloc((void**)&pg.eob, pg.eob_cap * 2), "alloc graph");
        }
        HIP_TRY(hipMemcpyAsync(pg.dev, pg.host, sizeof(SegLaunch) * used, hipMemcpyHostToDevice, s), "copy graph tables");
        const GL* all[] = {&lik, &clr, &mx, &flg, &pr, &sp, &snap, &nx, &corr, &sel[0], &sel[1], &sel[2], &sel[3],
                           &re, &dec, &res, &mk};
        for (const GL* l : all) launch_seg_fill_eob(pg.dev + l->o, l->n, pg.eob + l->eo, s);
        HIP_TRY(hipGetLastError(), "kernel launch");
        HIP_TRY(hipStreamSynchronize(s), "sync");
        auto T = [&](const GL& l) { return pg.dev + l.o; };
        auto E = [&](const GL& l) { return pg.eob + l.eo; };
        const SegDesc* sd = bb.d_desc;
        SegDyn* dy = bb.d_dyn;
        // the pass's launches as graph nodes, in order; the selection classes
        // side by side between the correlations and the reassignment
        GraphRec rec;
        HIP_TRY(hipGraphCreate(&rec.graph, 0), "create pass graph");
        GraphRec* r = &rec;
        launch_seg_plan(sd, dy, pg.dev, pg.plan, 0, s, r);
        launch_seg_counters(sd, (uint32_t)nseg, nullptr, s, r);  // CNT_PAIRS = 0
        launch_seg_pass_kernel(SEG_LIKELIHOOD, L, sd, dy, T(lik), lik.n, lik.nb, E(lik), lds_lik, 1, s, r);
        launch_seg_pass_kernel(SEG_CLEAR_DIRTY, L, sd, dy, T(clr), clr.n, clr.nb, E(clr), 0, 0, s, r);
        launch_seg_pass_kernel(SEG_WINDOW_MAX, L, sd, dy, T(mx), mx.n, mx.nb, E(mx), 0, 0, s, r);
        if (recheck) launch_seg_compact(0, sd, dy, T(flg), flg.n, flg.nb, E(flg), s, r);
        launch_seg_plan(sd, dy, pg.dev, pg.plan, 1, s, r);
        if (recheck) {
            launch_seg_screen_corr(L, sd, T(corr), corr.n, corr.nb, E(corr), lds_corr, bb.Bbuf, bstride, s, r);
            rec.fork();
            for (int c = 0; c < kSelClasses; c++)
                launch_seg_select(L, sd, T(sel[c]), sel[c].n, sel[c].nb, E(sel[c]), cls_bd[c], cls_lds[c], bb.Bbuf,
                                  bstride, s, r);
            rec.join();
            launch_seg_pass_kernel(SEG_REASSIGN, L, sd, dy, T(re), re.n, re.nb, E(re), 0, 0, s, r);
        }
        launch_seg_pass_kernel(SEG_DECIDE, L, sd, dy, T(dec), dec.n, dec.nb, E(dec), 0, 0, s, r);
        launch_seg_pass_kernel(SEG_SNAPSHOT, L, sd, dy, T(snap), snap.n, snap.nb, E(snap), 0, 0, s, r);
        launch_seg_pass_kernel(SEG_PRUNE, L, sd, dy, T(pr), pr.n, pr.nb, E(pr), 0, 0, s, r);
        launch_seg_compact(1, sd, dy, T(sp), sp.n, sp.nb, E(sp), s, r);
        launch_seg_compact(2, sd, dy, T(nx), nx.n, nx.nb, E(nx), s, r);
        launch_seg_plan(sd, dy, pg.dev, pg.plan, 2, s, r);
        launch_seg_pass_kernel(SEG_RESIDUAL, L, sd, dy, T(res), res.n, res.nb, E(res), 0, 0, s, r);
        launch_seg_pass_kernel(SEG_MARK_DIRTY, L, sd, dy, T(mk), mk.n, mk.nb, E(mk), 0, 0, s, r);
        launch_seg_counters(sd, (uint32_t)nseg, bb.d_cnt, s, r);
        hipError_t ge = rec.err;
        if (ge == hipSuccess) ge = hipGraphInstantiate(&pg.exec, rec.graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(rec.graph);
        if (ge != hipSuccess) {
            pg.exec = nullptr;
            return hip_err(h, ge, "build pass graph");
        }
        pg.builds++;
        return FBP_OK;
    };
    // One pass by a replay (building the graph first when a segment lists more windows
    // than the graph holds, or the graph holds over 3x the listed windows);
    // *ran = false when the flagged windows might not fit one screen round (the
    // pass is then planned on the host).
    auto graph_pass = [&](bool* ran) -> int {
        *ran = false;
        uint64_t want = 0;
        bool need = pg.exec == nullptr;
        for (uint32_t g : active) {
            want += nW[g];
            if (nW[g] > pg.R[g]) need = true;
        }
        if (!need && pg.sumR > 3 * want + 2048) need = true;
        if (need) {
            std::vector<uint32_t> R(nseg, 0);
            uint64_t sumR = 0;
            for (uint32_t g : active) {
                R[g] = (uint32_t)std::min<uint64_t>((uint64_t)nwin[g], (uint64_t)nW[g] + nW[g] / 8 + 16);
                sumR += R[g];
            }
            if (recheck && sumR > rows_cap) return FBP_OK;
            int rc = graph_build(R, sumR);
            if (rc != FBP_OK) return rc;
        }
        // spike lists: room for one more spike per listed window
        for (uint32_t g : active) {
            fbp_handle* hg = hs[g];
            const size_t need_b = (size_t)(total[g] + nW[g]) * 4;
            HIP_TRY(grow_keep(hg->sp_idx, need_b, (size_t)total[g] * 4, s), "grow spikes");
            HIP_TRY(grow_keep(hg->sp_lab, need_b, (size_t)total[g] * 4, s), "grow spikes");
        }
        for (int32_t g = 0; g < nseg; g++)
            bb.h_dyn[g] = SegDyn{hs[g]->sp_idx.as<uint32_t>(), hs[g]->sp_lab.as<uint32_t>(), total[g], nW[g], 0u};
        for (uint32_t g : active) bb.h_dyn[g].active = 1u;
        HIP_TRY(hipMemcpyAsync(bb.d_dyn, bb.h_dyn, sizeof(SegDyn) * nseg, hipMemcpyHostToDevice, s), "copy batch");
        HIP_TRY(hipGraphLaunch(pg.exec, s), "graph launch");
        HIP_TRY(hipMemcpyAsync(bb.h_cnt, bb.d_cnt, (size_t)nseg * CNT_N * 4, hipMemcpyDeviceToHost, s), "copy counters");
        HIP_TRY(hipStreamSynchronize(s), "sync");
        pg.replays++;
        std::vector<uint32_t> next;
        for (uint32_t g : active) {
            const uint32_t* c = bb.h_cnt + (size_t)g * CNT_N;
            nF[g] = recheck ? c[CNT_F] : 0u;
            nadd[g] = c[CNT_SPIKES];
            if (nF[g] > nW[g] || nadd[g] > nW[g])
                return set_err(h, FBP_E_HIP, "graph pass: counts beyond the listed windows");
            fbp_stats& sg = stats[g];
            sg.windows_listed += nW[g];
            sg.windows_flagged += nF[g];
            sg.likelihood_positions += (int64_t)nW[g] * L * hs[g]->U;
            sg.likelihood_pairs_computed += c[CNT_PAIRS];
            passes[g]++;
            hs[g]->pass_log.push_back(nadd[g]);
            hs[g]->pass_log.push_back(nW[g]);
            hs[g]->pass_log.push_back(nF[g]);
            if (nadd[g] > 0) next.push_back(g);
        }
        for (uint32_t g : next) {
            total[g] += nadd[g];
            nW[g] = bb.h_cnt[(size_t)g * CNT_N + CNT_W];
        }
        active.swap(next);
        *ran = true;
        return FBP_OK;
    };
    int pass = 0;
    while (!active.empty()) {
        if (use_graph && pass > 0) {
            bool ran = false;
            int rc = graph_pass(&ran);
            if (rc != FBP_OK) return rc;
            if (ran) {
                pass++;
                continue;
            }
        }
        // ---- likelihood, window maxima, flagged windows ----
        SegTables& A = bb.tabs[0];
        A.used = 0;
        uint32_t nb_lik, nb_clr, nb_max, nb_flag = 0;
        const size_t o_lik = A.begin();
        const uint32_t n_lik = seg_table(A, active, [&](uint32_t g, SegLaunch& e) {
            e.n = nW[g];
            e.nblk = capb(cdiv(nW[g], (uint32_t)bb.desc[g].wb));
            return true;
        }, &nb_lik);
        A.used += n_lik;
        const size_t o_clr = A.begin();
        const uint32_t n_clr = seg_table(A, active, [&](uint32_t g, SegLaunch& e) {
            e.n = nW[g];
            e.nblk = capb(cdiv(nW[g], kBlock));
            return true;
        }, &nb_clr);
        A.used += n_clr;
        const size_t o_max = A.begin();
        const uint32_t n_max = seg_table(A, active, [&](uint32_t g, SegLaunch& e) {
            e.n = nW[g];
            e.nblk = capb(cdiv(nW[g], kBlock / 64));
            return true;
        }, &nb_max);
        A.used += n_max;
        const size_t o_flag = A.begin();
        uint32_t n_flag = 0;
        if (recheck) {
            n_flag = seg_table(A, active, [&](uint32_t g, SegLaunch& e) {
                e.n = nW[g];
                e.nblk = cdiv(nW[g], kCompactTile);
                return true;
            }, &nb_flag);
            A.used += n_flag;
        }
        HIP_TRY(upload(A), "copy batch tables");
        launch_seg_counters(bb.d_desc, (uint32_t)nseg, nullptr, s);  // CNT_PAIRS = 0
        launch_seg_pass_kernel(SEG_LIKELIHOOD, L, bb.d_desc, bb.d_dyn, A.dev + o_lik, n_lik, nb_lik, bb.eob[0], lds_lik,
                               pass > 0 ? 1 : 0, s);
        launch_seg_pass_kernel(SEG_CLEAR_DIRTY, L, bb.d_desc, bb.d_dyn, A.dev + o_clr, n_clr, nb_clr, bb.eob[0], 0, 0, s);
        launch_seg_pass_kernel(SEG_WINDOW_MAX, L, bb.d_desc, bb.d_dyn, A.dev + o_max, n_max, nb_max, bb.eob[0], 0, 0, s);
        if (recheck) {
            launch_seg_compact(0, bb.d_desc, bb.d_dyn, A.dev + o_flag, n_flag, nb_flag, bb.eob[0], s);
            HIP_TRY(read_counters(), "read counters");
            for (uint32_t g : active) nF[g] = bb.h_cnt[(size_t)g * CNT_N + CNT_F];
        }
        // ---- screened recheck, reassignment, pursuit, compactions ----
        SegTables& B = bb.tabs[1];
        B.used = 0;
        // spike lists: room for one more spike per listed window
        for (uint32_t g : active) {
            fbp_handle* hg = hs[g];
            const size_t need = (size_t)(total[g] + nW[g]) * 4;
            HIP_TRY(grow_keep(hg->sp_idx, need, (size_t)total[g] * 4, s), "grow spikes");
            HIP_TRY(grow_keep(hg->sp_lab, need, (size_t)total[g] * 4, s), "grow spikes");
        }
        for (int32_t g = 0; g < nseg; g++)
            bb.h_dyn[g] = SegDyn{hs[g]->sp_idx.as<uint32_t>(), hs[g]->sp_lab.as<uint32_t>(), total[g]};
        HIP_TRY(hipMemcpyAsync(bb.d_dyn, bb.h_dyn, sizeof(SegDyn) * nseg, hipMemcpyHostToDevice, s), "copy batch");
        // screen rounds: the flagged windows of the active segments, rows_cap at a time
        struct Round {
            size_t o;
            uint32_t n, nb_corr;
            size_t so[kSelClasses];  // select entries of each class: table offset, count, blocks
            uint32_t sn[kSelClasses], snb[kSelClasses];
        };
        std::vector<Round> rounds;
        if (recheck) {
            std::vector<uint32_t> done(nseg, 0);
            for (;;) {
                uint32_t rows = 0;
                Round r{};
                r.o = B.begin();
                // corr entries, then select entries (same items, one block per window)
                for (uint32_t g : active) {
                    if (done[g] >= nF[g] || rows >= rows_cap) continue;
                    const uint32_t take = std::min(nF[g] - done[g], rows_cap - rows);
                    SegLaunch e{};
                    e.seg = g;
                    e.n = take;
                    e.a0 = done[g];
                    e.g0 = rows;
                    e.blk0 = r.nb_corr;
                    e.nblk = capb(take);
                    r.nb_corr += e.nblk;
                    B.host[B.used + r.n++] = e;
                    done[g] += take;
                    rows += take;
                }
                if (r.n == 0) break;
                if (B.used + 2 * (size_t)r.n > B.cap) return set_err(h, FBP_E_UNSUPPORTED, "batch table overflow");
                // the same items, one block per window, grouped by selection class
                size_t o = B.used + r.n;
                for (int c = 0; c < kSelClasses; c++) {
                    r.so[c] = o;
                    for (uint32_t k = 0; k < r.n; k++) {
                        SegLaunch e = B.host[B.used + k];
                        if (scls[e.seg] != c) continue;
                        e.blk0 = r.snb[c];
                        e.nblk = e.n;
                        r.snb[c] += e.n;
                        r.sn[c]++;
                        B.host[o++] = e;
                    }
                }
                B.used += 2 * (size_t)r.n;
                rounds.push_back(r);
            }
        }
        uint32_t nb_re = 0, nb_dec, nb_snap, nb_pr, nb_sp, nb_nx;
        const size_t o_re = B.begin();
        uint32_t n_re = 0;
        if (recheck) {
            n_re = seg_table(B, active, [&](uint32_t g, SegLaunch& e) {
                e.n = nF[g];
                e.nblk = capb(cdiv(nF[g], kBlock));
                return true;
            }, &nb_re);
            B.used += n_re;
        }
        const size_t o_dec = B.begin();
        const uint32_t n_dec = seg_table(B, active, [&](uint32_t g, SegLaunch& e) {
            e.n = nW[g];
            e.a0 = (recheck && nF[g] > 0) ? 1u : 0u;
            e.nblk = capb(cdiv(nW[g], kBlock));
            return true;
        }, &nb_dec);
        B.used += n_dec;
        const size_t o_snap = B.begin();
        const uint32_t n_snap = seg_table(B, active, [&](uint32_t g, SegLaunch& e) {
            e.n = (uint32_t)nwin[g];
            e.nblk = capb(cdiv(nwin[g], 4 * kBlock));
            return true;
        }, &nb_snap);
        B.used += n_snap;
        const size_t o_pr = B.begin();
        const uint32_t n_pr = seg_table(B, active, [&](uint32_t g, SegLaunch& e) {
            e.n = nW[g];
            e.nblk = capb(cdiv(nW[g], kBlock));
            return true;
        }, &nb_pr);
        B.used += n_pr;
        const size_t o_sp = B.begin();
        const uint32_t n_sp = seg_table(B, active, [&](uint32_t g, SegLaunch& e) {
            e.n = nW[g];
            e.nblk = cdiv(nW[g], kCompactTile);
            return true;
        }, &nb_sp);
        B.used += n_sp;
        const size_t o_nx = B.begin();
        const uint32_t n_nx = seg_table(B, active, [&](uint32_t g, SegLaunch& e) {
            e.n = (uint32_t)nwin[g];
            e.nblk = cdiv(nwin[g], kCompactTile);
            return true;
        }, &nb_nx);
        B.used += n_nx;
        if (B.used > B.cap) return set_err(h, FBP_E_UNSUPPORTED, "batch table overflow");
        HIP_TRY(upload(B), "copy batch tables");
        for (const Round& r : rounds) {
            launch_seg_screen_corr(L, bb.d_desc, B.dev + r.o, r.n, r.nb_corr, bb.eob[0], lds_corr, bb.Bbuf, bstride, s);
            // the classes side by side (class 0 on the batch's stream, the others on
            // side streams), joined before the next round's correlations
            HIP_TRY(hipEventRecord(side.fork, s), "event");
            size_t eo = 0;  // each class's block-to-entry map in its own part of eob[1]
            for (int c = 0; c < kSelClasses; c++) {
                if (r.sn[c] == 0) continue;
                hipStream_t sc = s;
                if (c > 0) {
                    sc = side.st[c];
                    HIP_TRY(hipStreamWaitEvent(sc, side.fork, 0), "stream wait");
                }
                launch_seg_select(L, bb.d_desc, B.dev + r.so[c], r.sn[c], r.snb[c], bb.eob[1] + eo, cls_bd[c], cls_lds[c],
                                  bb.Bbuf, bstride, sc);
                eo += r.snb[c];
                if (c > 0) {
                    HIP_TRY(hipEventRecord(side.ev[c], sc), "event");
                    HIP_TRY(hipStreamWaitEvent(s, side.ev[c], 0), "stream wait");
                }
            }
        }
        if (recheck)
            launch_seg_pass_kernel(SEG_REASSIGN, L, bb.d_desc, bb.d_dyn, B.dev + o_re, n_re, nb_re, bb.eob[0], 0, 0, s);
        launch_seg_pass_kernel(SEG_DECIDE, L, bb.d_desc, bb.d_dyn, B.dev + o_dec, n_dec, nb_dec, bb.eob[0], 0, 0, s);
        launch_seg_pass_kernel(SEG_SNAPSHOT, L, bb.d_desc, bb.d_dyn, B.dev + o_snap, n_snap, nb_snap, bb.eob[0], 0, 0, s);
        launch_seg_pass_kernel(SEG_PRUNE, L, bb.d_desc, bb.d_dyn, B.dev + o_pr, n_pr, nb_pr, bb.eob[0], 0, 0, s);
        launch_seg_compact(1, bb.d_desc, bb.d_dyn, B.dev + o_sp, n_sp, nb_sp, bb.eob[0], s);
        launch_seg_compact(2, bb.d_desc, bb.d_dyn, B.dev + o_nx, n_nx, nb_nx, bb.eob[0], s);
        HIP_TRY(read_counters(), "read counters");
        // ---- residual of this pass's spikes; converged segments leave ----
        std::vector<uint32_t> next;
        for (uint32_t g : active) {
            const uint32_t* c = bb.h_cnt + (size_t)g * CNT_N;
            nadd[g] = c[CNT_SPIKES];
            fbp_stats& sg = stats[g];
            sg.windows_listed += nW[g];
            sg.windows_flagged += nF[g];
            sg.likelihood_positions += (int64_t)nW[g] * L * hs[g]->U;
            sg.likelihood_pairs_computed += c[CNT_PAIRS];
            passes[g]++;
            hs[g]->pass_log.push_back(nadd[g]);
            hs[g]->pass_log.push_back(nW[g]);
            hs[g]->pass_log.push_back(nF[g]);
            if (nadd[g] > 0) next.push_back(g);
        }
        SegTables& Ct = bb.tabs[2];
        Ct.used = 0;
        uint32_t nb_res, nb_mk;
        const size_t o_res = Ct.begin();
        const uint32_t n_res = seg_table(Ct, next, [&](uint32_t g, SegLaunch& e) {
            e.n = nadd[g];
            e.nblk = capb(nadd[g]);
            return true;
        }, &nb_res);
        Ct.used += n_res;
        const size_t o_mk = Ct.begin();
        const uint32_t n_mk = seg_table(Ct, next, [&](uint32_t g, SegLaunch& e) {
            e.n = nadd[g];
            e.nblk = capb(cdiv(nadd[g], kBlock));
            return true;
        }, &nb_mk);
        Ct.used += n_mk;
        HIP_TRY(upload(Ct), "copy batch tables");
        launch_seg_pass_kernel(SEG_RESIDUAL, L, bb.d_desc, bb.d_dyn, Ct.dev + o_res, n_res, nb_res, bb.eob[0], 0, 0, s);
        launch_seg_pass_kernel(SEG_MARK_DIRTY, L, bb.d_desc, bb.d_dyn, Ct.dev + o_mk, n_mk, nb_mk, bb.eob[0], 0, 0, s);
        for (uint32_t g : next) {
            total[g] += nadd[g];
            nW[g] = bb.h_cnt[(size_t)g * CNT_N + CNT_W];
        }
        active.swap(next);
        pass++;
    }
    HIP_TRY(hipEventRecord(evp.b, s), "event");
    HIP_TRY(hipGetLastError(), "kernel launch");
    HIP_TRY(hipEventSynchronize(evp.b), "sync");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, evp.a, evp.b);
    if (flags & FBP_FLAG_ADJUSTED_CLIPS) {
        for (int32_t g = 0; g < nseg; g++) {
            fbp_handle* hg = hs[g];
            HIP_TRY(hg->clips.ensure((size_t)std::max<uint64_t>(total[g], 1) * hg->C * L * 4), "alloc clips");
            launch_adjusted_clips(template_view(hg), hg->V.as<float>(), hg->N, hg->sp_idx.as<uint32_t>(),
                                  hg->sp_lab.as<uint32_t>(), total[g], hg->clips.as<float>(), s);
            hg->have_clips = true;
        }
    }
    HIP_TRY(read_counters(), "read counters");
    for (int32_t g = 0; g < nseg; g++) {
        fbp_handle* hg = hs[g];
        const uint32_t* c = bb.h_cnt + (size_t)g * CNT_N;
        fbp_stats& sg = stats[g];
        sg.ms_total = ms;
        sg.passes = passes[g];
        sg.spikes = (int64_t)total[g];
        sg.recheck_items = c[CNT_EVAL];
        sg.recheck_overflow = c[CNT_CAND];
        sg.screen_shared_pairs = c[CNT_SURV];
        sg.screen_separable_pairs = c[CNT_NSP];
        sg.n_graph_builds = pg.builds;
        sg.n_graph_replays = pg.replays;
        hg->stats = sg;
        hg->n_spikes = (int64_t)total[g];
        hg->have_result = true;
    }
    return FBP_OK;
}

int fbp_result_count(fbp_handle* h, int64_t* n)
{
    if (!h || !n) return FBP_E_INVALID;
    if (!h->have_result) return set_err(h, FBP_E_STATE, "no result: call fbp_run first");
    *n = h->n_spikes;
    return FBP_OK;
}

int fbp_copy_result(fbp_handle* h, uint32_t* indices, uint32_t* labels)
{
    if (!h) return FBP_E_INVALID;
    if (!h->have_result) return set_err(h, FBP_E_STATE, "no result: call fbp_run first");
    if (h->n_spikes == 0) return FBP_OK;
    if (!indices || !labels) return set_err(h, FBP_E_INVALID, "null output");
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    HIP_TRY(hipMemcpyAsync(indices, h->sp_idx.p, h->n_spikes * 4, hipMemcpyDeviceToHost, h->stream), "copy");
    HIP_TRY(hipMemcpyAsync(labels, h->sp_lab.p, h->n_spikes * 4, hipMemcpyDeviceToHost, h->stream), "copy");
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    return FBP_OK;
}

int fbp_copy_clips(fbp_handle* h, float* clips)
{
    if (!h) return FBP_E_INVALID;
    if (!h->have_result || !h->have_clips)
        return set_err(h, FBP_E_STATE, "no clips: run with FBP_FLAG_ADJUSTED_CLIPS first");
    if (h->n_spikes == 0) return FBP_OK;
    if (!clips) return set_err(h, FBP_E_INVALID, "null output");
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    HIP_TRY(hipMemcpyAsync(clips, h->clips.p, (size_t)h->n_spikes * h->C * h->L * 4, hipMemcpyDeviceToHost, h->stream),
            "copy");
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    return FBP_OK;
}

int fbp_copy_residual(fbp_handle* h, float* out)
{
    if (!h) return FBP_E_INVALID;
    if (!h->have_result) return set_err(h, FBP_E_STATE, "no result: call fbp_run first");
    if (!out) return set_err(h, FBP_E_INVALID, "null output");
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    int rc = stage_out(h, h->V.as<float>(), h->N, out, h->N);
    if (rc != FBP_OK) return rc;
    HIP_TRY(hipGetLastError(), "residual layout");
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    return FBP_OK;
}

int fbp_copy_likelihood(fbp_handle* h, float* out)
{
    if (!h) return FBP_E_INVALID;
    if (!h->have_result) return set_err(h, FBP_E_STATE, "no result: call fbp_run first");
    if (!out) return set_err(h, FBP_E_INVALID, "null output");
    HIP_TRY(hipSetDevice(h->device), "hipSetDevice");
    HIP_TRY(hipMemcpyAsync(out, h->LIK.p, (size_t)h->U * h->N * 4, hipMemcpyDeviceToHost, h->stream), "copy");
    HIP_TRY(hipStreamSynchronize(h->stream), "sync");
    return FBP_OK;
}

int fbp_get_stats(fbp_handle* h, fbp_stats* out)
{
    if (!h || !out) return FBP_E_INVALID;
    *out = h->stats;
    return FBP_OK;
}

int fbp_get_pass_log(fbp_handle* h, int64_t* rows, int32_t max_rows, int32_t* n_rows)
{
    if (!h || !n_rows || max_rows < 0 || (max_rows > 0 && !rows)) return FBP_E_INVALID;
    if (!h->have_result) return set_err(h, FBP_E_STATE, "no result: call fbp_run first");
    const int32_t n = (int32_t)(h->pass_log.size() / 3);
    *n_rows = n;
    const size_t k = (size_t)std::min(n, max_rows) * 3;
    if (k) std::memcpy(rows, h->pass_log.data(), k * sizeof(int64_t));
    return FBP_OK;
}

}  // extern "C"
