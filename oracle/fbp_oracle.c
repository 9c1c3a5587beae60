/*
 * fbp_oracle.c -- CPU restatement of the full-binary-pursuit pass loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * implementation in spikesorting_fullpursuit_amd/csrc.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never calls it and has no CPU fallback.
 *
 * What it restates (reference paths relative to the reference repository):
 *   SF = spikesorting_fullpursuit/
 *   - SF/kernels/binary_pursuit.cl       all eight kernels and helpers
 *   - SF/overlap/binary_pursuit_parallel.py:907-1355   the pass loop
 * for ONE chunk of voltage.  Chunk planning and output assembly live in the
 * Python drop-in (spikesorting_fullpursuit_amd/overlap/binary_pursuit_parallel.py).
 *
 * Floating-point semantics (the contract the GPU path reproduces bit for bit):
 *   * likelihood  lik = bias[t]; for each channel c (ascending) whose template
 *     row is not identically zero: for i = 0..L-1: lik = fmaf(T[t,c,i], V[c,x+i], lik).
 *     Skipping an all-zero channel changes nothing for finite voltage
 *     (fmaf(0, v, lik) == lik up to the sign of a zero, which no comparison sees).
 *     lik = 0 when x + L >= N   (binary_pursuit.cl:231-234).
 *   * overlap combination (binary_pursuit.cl:585-631): per channel c of the
 *     union of both templates' non-zero channels, ascending:
 *       ss = 0; for j: sum = a_j + b_j; ss = fmaf(sum, sum, ss); lik = fmaf(sum, V[c,b+j], lik);
 *       if (ss > 0) lik = fmaf(-0.5f, ss, lik);
 *   * residual  V[c,p+i] = V[c,p+i] - T[u,c,i] in spike-list order (binary_pursuit.cl:93-97).
 *
 * Phase-separated meaning adopted where the reference races or is
 * scheduling dependent:
 *   * overlap arg-max ties go to the lowest linear item id k (the reference's
 *     work-group tree reduction, binary_pursuit.cl:167-186, picks a
 *     work-group-size dependent winner among exact ties);
 *   * check_overlap_reassignments reads overlap_recheck[id-2] as it stood
 *     after the parse kernel (binary_pursuit.cl:850-864);
 *   * binary_pursuit: every "=3" flag write happens before any pruning, and
 *     pruning reads neighbour flags as they stood after the "=3" writes
 *     (binary_pursuit.cl:982-1045);
 *   * padding work-items are no-ops (binary_pursuit.cl:939-944 reads past the
 *     end of overlap_recheck);
 *   * spikes of a pass are listed in ascending window order (the reference
 *     appends per work group through an atomic, binary_pursuit.cl:1047-1061);
 *   * the residual subtracts every spike, also where supports overlap
 *     (binary_pursuit.cl:35-43 documents a race there).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FBPO_OK 0
#define FBPO_EINVAL -1
#define FBPO_ENOMEM -2
#define FBPO_ECAP -3

typedef struct {
    const float *T;   /* [U][C][L] */
    int64_t C, N;
    int32_t U, L;
    uint8_t *chan_nz; /* [U][C] 1 if template row is not identically zero */
} tset;

static void tset_init(tset *ts, const float *T, int64_t C, int64_t N, int32_t U, int32_t L)
{
    ts->T = T; ts->C = C; ts->N = N; ts->U = U; ts->L = L;
    ts->chan_nz = (uint8_t *)calloc((size_t)U * (size_t)C, 1);
    for (int32_t t = 0; t < U; t++)
        for (int64_t c = 0; c < C; c++) {
            const float *row = T + ((size_t)t * C + c) * L;
            uint8_t nz = 0;
            for (int32_t i = 0; i < L; i++) if (row[i] != 0.0f) { nz = 1; break; }
            ts->chan_nz[(size_t)t * C + c] = nz;
        }
}

static void tset_free(tset *ts) { free(ts->chan_nz); ts->chan_nz = NULL; }

/* compute_maximum_likelihood, binary_pursuit.cl:216-252 */
static float lik_at(const tset *ts, const float *V, const float *bias, int32_t t, int64_t x)
{
    const int64_t N = ts->N, C = ts->C;
    const int32_t L = ts->L;
    if (x + L >= N) return 0.0f;
    float lik = bias[t];
    for (int64_t c = 0; c < C; c++) {
        if (!ts->chan_nz[(size_t)t * C + c]) continue;
        const float *tr = ts->T + ((size_t)t * C + c) * L;
        const float *vr = V + (size_t)c * N + x;
        for (int32_t i = 0; i < L; i++) lik = fmaf(tr[i], vr[i], lik);
    }
    return lik;
}

/* Public single-position likelihood (for closed-form tests). */
float fbpo_likelihood_at(const float *V, int64_t C, int64_t N, const float *T, int32_t U,
                         int32_t L, const float *bias, int32_t t, int64_t x)
{
    tset ts; tset_init(&ts, T, C, N, U, L);
    float r = lik_at(&ts, V, bias, t, x);
    tset_free(&ts);
    return r;
}

/* compute_full_likelihood, binary_pursuit.cl:255-300, for a list of windows. */
static void full_likelihood(const tset *ts, const float *V, const float *bias,
                            const uint32_t *win, int64_t nwin_list, float *LIK)
{
    const int64_t N = ts->N;
    const int32_t L = ts->L, U = ts->U;
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < nwin_list; w++) {
        int64_t id = win[w];
        int64_t x0 = id * L, x1 = (id + 1) * L;
        if (x1 > N) x1 = N;
        for (int32_t t = 0; t < U; t++)
            for (int64_t x = x0; x < x1; x++)
                LIK[(size_t)t * N + x] = lik_at(ts, V, bias, t, x);
    }
}

/* Overlap combination value for item (u, f, s) of a flagged window whose best
 * spike is (lab, b).  binary_pursuit.cl:578-632. */
static float combo_value(const tset *ts, const float *V, uint32_t lab, uint32_t u, int64_t b,
                         int64_t f, int64_t s, int64_t n)
{
    const int64_t N = ts->N, C = ts->C;
    const int32_t L = ts->L;
    const int64_t fa = n - f;  /* curr_fixed_align = best - absolute_fixed */
    const int64_t sa = n - s;  /* curr_shift_align */
    const int64_t ffj = fa < 0 ? -fa : 0;
    const int64_t sfj = sa < 0 ? -sa : 0;
    float lik = 0.0f;
    for (int64_t c = 0; c < C; c++) {
        const int nzf = ts->chan_nz[(size_t)lab * C + c];
        const int nzs = ts->chan_nz[(size_t)u * C + c];
        if (!nzf && !nzs) continue;
        const float *tf = ts->T + ((size_t)lab * C + c) * L;
        const float *tsr = ts->T + ((size_t)u * C + c) * L;
        const float *vr = V + (size_t)c * N + b;
        float ss = 0.0f;
        for (int64_t j = 0; j < L; j++) {
            float fv = 0.0f, sv = 0.0f;
            if (j >= ffj && fa + j < L) fv = tf[fa + j];
            if (j >= sfj && sa + j < L) sv = tsr[sa + j];
            float sum = fv + sv;
            ss = fmaf(sum, sum, ss);
            lik = fmaf(sum, vr[j], lik);
        }
        if (ss > 0.0f) lik = fmaf(-0.5f, ss, lik);
    }
    return lik;
}

float fbpo_combo_value(const float *V, int64_t C, int64_t N, const float *T, int32_t U, int32_t L,
                       uint32_t lab, uint32_t u, int64_t b, int64_t f, int64_t s, int64_t n)
{
    tset ts; tset_init(&ts, T, C, N, U, L);
    float r = combo_value(&ts, V, lab, u, b, f, s, n);
    tset_free(&ts);
    return r;
}

/* Skip test of overlap_recheck_indices for one item, binary_pursuit.cl:518-576.
 * Returns 1 when the item is skipped (contributes 0 to the reduction). */
static int combo_skipped(const tset *ts, const float *LIK, uint32_t lab, uint32_t u, int64_t b,
                         int64_t f, int64_t s, int64_t n)
{
    const int64_t N = ts->N;
    const int32_t L = ts->L;
    if (f + b < n || f + b >= N - L + n) return 1;
    if (s + b < n || s + b >= N - L + n) return 1;
    if (f > s && f - s > L) return 1;
    if (s > f && s - f > L) return 1;
    const int64_t af = b + f - n, as = b + s - n;
    if (LIK[(size_t)lab * N + af] <= 0.0f && LIK[(size_t)u * N + as] <= 0.0f) return 1;
    return 0;
}

typedef struct {
    int64_t nwin;       /* ceil(N / L) */
    uint32_t *best_idx, *best_lab, *obest_idx, *obest_lab;
    float *best_lik;
    uint8_t *check, *orc;
} pstate;

/* compute_template_maximum_likelihood launched for t = 0..U-1 in order,
 * binary_pursuit.cl:350-441 and binary_pursuit_parallel.py:943-964. */
static void window_max(const tset *ts, const float *LIK, const uint32_t *win, int64_t nw, pstate *st)
{
    const int64_t N = ts->N;
    const int64_t L = ts->L;
    const int32_t U = ts->U;
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < nw; w++) {
        const int64_t id = win[w];
        const int64_t sw = id * L, ew = (id + 1) * L;
        if (ew >= N - L) continue;
        const int64_t start = (L > sw) ? 0 : sw - L;
        const int64_t stop = (ew + L > N) ? N : ew + L;
        float best = 0.0f;
        uint32_t lab = st->best_lab[id], idx = st->best_idx[id];
        int cw = 0;
        for (int32_t t = 0; t < U; t++) {
            const float *lr = LIK + (size_t)t * N;
            for (int64_t x = start; x < stop; x++) {
                float v = lr[x];
                if (v > best) { best = v; lab = (uint32_t)t; idx = (uint32_t)x; }
                if (v > 0.0f && x >= sw && x < ew) cw = 1;
            }
        }
        if (cw) st->check[id] = 1;
        st->orc[id] = (best > 0.0f && (int64_t)idx >= sw && (int64_t)idx < ew) ? 1 : 0;
        st->best_lik[id] = best;
        st->best_lab[id] = lab;
        st->best_idx[id] = idx;
        st->obest_idx[id] = idx;
        st->obest_lab[id] = lab;
    }
}

/* overlap_recheck_indices + parse_overlap_recheck_indices for one flagged
 * window, binary_pursuit.cl:444-815. */
static void recheck_window(const tset *ts, const float *V, const float *LIK, const float *thr,
                           int64_t n, int64_t id, pstate *st, int64_t *n_items_eval)
{
    const int64_t N = ts->N;
    const int32_t L = ts->L;
    const int64_t U = ts->U;
    const int64_t S = 2 * n + 1;
    const uint32_t lab = st->best_lab[id];
    const int64_t b = st->best_idx[id];
    const float lb = LIK[(size_t)lab * N + b];
    int64_t evald = 0;
    float gbest = 0.0f;
    int64_t gk = -1;
    if (!(lb < thr[lab]) && !(b >= N - L)) {
        const int64_t items = S * S * U;
        for (int64_t k = 0; k < items; k++) {
            const int64_t u = k % U, f = (k / U) % S, s = k / (S * U);
            if (combo_skipped(ts, LIK, lab, (uint32_t)u, b, f, s, n)) continue;
            float v = combo_value(ts, V, lab, (uint32_t)u, b, f, s, n);
            evald++;
            if (gk < 0 || v > gbest) { gbest = v; gk = k; }
        }
    }
    *n_items_eval += evald;
    /* parse_overlap_recheck_indices, binary_pursuit.cl:693-814 */
    const float bl = st->best_lik[id];
    if (gk < 0 || !(gbest > bl) || gbest <= 0.0f) { st->orc[id] = 0; return; }
    const uint32_t u = (uint32_t)(gk % U);
    const int64_t f = (gk / U) % S, s = gk / (S * U);
    const int64_t af = b + f - n, as = b + s - n;
    const float A = LIK[(size_t)lab * N + af];
    const float B = LIK[(size_t)u * N + as];
    if (gbest < A || gbest < B) { st->orc[id] = 0; return; }
    if (A <= 0.0f || B <= 0.0f) { st->orc[id] = 0; return; }
    if (A >= B && A > thr[lab]) {
        st->best_lik[id] = A; st->obest_lab[id] = lab; st->obest_idx[id] = (uint32_t)af;
    } else if (B > A && B > thr[u]) {
        st->best_lik[id] = B; st->obest_lab[id] = u; st->obest_idx[id] = (uint32_t)as;
    } else if (A > thr[lab]) {
        st->best_lik[id] = A; st->obest_lab[id] = lab; st->obest_idx[id] = (uint32_t)af;
    } else {
        st->obest_idx[id] = st->best_idx[id];
        st->obest_lab[id] = st->best_lab[id];
        st->orc[id] = 0;
    }
}

/* check_overlap_reassignments, binary_pursuit.cl:822-875, reading
 * overlap_recheck as it stood before the kernel. */
static void reassign(const tset *ts, const uint32_t *flist, int64_t nf, pstate *st, uint8_t *orc_snap)
{
    const int64_t N = ts->N;
    const int64_t L = ts->L;
    const int64_t nwf = N / L; /* voltage_length / template_length */
    memcpy(orc_snap, st->orc, (size_t)st->nwin);
    for (int64_t w = 0; w < nf; w++) {
        const int64_t id = flist[w];
        if (orc_snap[id] == 0) continue;
        const int64_t sw = id * L, ew = (id + 1) * L;
        st->best_idx[id] = st->obest_idx[id];
        st->best_lab[id] = st->obest_lab[id];
        if ((int64_t)st->obest_idx[id] < sw && id > 1) {
            if (orc_snap[id - 2] == 1) {
                st->orc[id] = 0;
                st->best_idx[id] = 0;
                st->check[id - 1] = 2;
                st->check[id] = 2;
                st->check[id + 1] = 2;
            } else {
                st->check[id - 2] = 2;
            }
        }
        if ((int64_t)st->obest_idx[id] >= ew && id + 2 < nwf) st->check[id + 2] = 2;
    }
}

/* binary_pursuit kernel, binary_pursuit.cl:918-1063, phase separated.
 * Appends the pass's spikes (ascending window order) to out. */
static int64_t pursuit(const tset *ts, const float *thr, const uint32_t *win, int64_t nw, pstate *st,
                       uint8_t *has, uint8_t *check_snap, uint32_t *sp_idx, uint32_t *sp_lab)
{
    const int64_t N = ts->N;
    const int64_t L = ts->L;
    const int64_t nwf = N / L;
    int64_t nadd = 0;
    /* phase A: decisions and "=3" flags */
    for (int64_t w = 0; w < nw; w++) {
        const int64_t id = win[w];
        const int64_t sw = id * L, ew = (id + 1) * L;
        float ml = 0.0f; uint32_t lab = 0, idx = 0;
        if (ew < N - L) { ml = st->best_lik[id]; lab = st->best_lab[id]; idx = st->best_idx[id]; }
        has[w] = 0;
        if ((ml > thr[lab]) || st->orc[id] == 1) {
            if (((int64_t)idx >= sw && (int64_t)idx < ew) || st->orc[id] == 1) {
                has[w] = 1;
                st->check[id] = 3;
                if (id > 0) st->check[id - 1] = 3;
                if (id + 1 < nwf) st->check[id + 1] = 3;
                sp_idx[nadd] = idx; sp_lab[nadd] = lab; nadd++;
            }
        }
    }
    /* phase B: pruning against the post-phase-A flags */
    memcpy(check_snap, st->check, (size_t)st->nwin);
    for (int64_t w = 0; w < nw; w++) {
        const int64_t id = win[w];
        const int64_t sw = id * L, ew = (id + 1) * L;
        if (!(ew < N - L) || has[w]) continue;
        const float ml = st->best_lik[id];
        const uint32_t lab = st->best_lab[id];
        const int64_t idx = st->best_idx[id];
        uint8_t own = check_snap[id];
        if (idx < sw && own == 1 && id > 0 && check_snap[id - 1] == 0) own = 0;
        if (idx >= ew && own == 1 && id + 1 < nwf && check_snap[id + 1] == 0) own = 0;
        if (ml <= thr[lab] && own < 3) own = 0;
        st->check[id] = own;
    }
    return nadd;
}

/* compute_residual, binary_pursuit.cl:57-102, every spike in list order. */
static void residual(const tset *ts, float *V, const uint32_t *idx, const uint32_t *lab, int64_t n)
{
    const int64_t N = ts->N, C = ts->C;
    const int32_t L = ts->L, U = ts->U;
    for (int64_t k = 0; k < n; k++) {
        if (lab[k] >= (uint32_t)U) continue;
        for (int64_t c = 0; c < C; c++) {
            const float *tr = ts->T + ((size_t)lab[k] * C + c) * L;
            float *vr = V + (size_t)c * N;
            for (int64_t i = 0; i < L; i++) {
                if (idx[k] + i >= N) break;
                vr[idx[k] + i] = vr[idx[k] + i] - tr[i];
            }
        }
    }
}

/*
 * Run the pass loop of binary_pursuit_parallel.py:907-1355 on one chunk.
 *   V        [C][N] float32, copied (the caller's array is not modified)
 *   T        [U][C][L] float32
 *   bias     [U]  float32(-0.5 * template_SS)
 *   thr      [U]  float32 neuron_lower_thresholds
 * Outputs: spikes (window-start indices, labels) in pass order; optionally the
 * final residual [C][N] and per-pass counters (spikes added, windows listed,
 * windows flagged for overlap recheck, items evaluated).
 */
int fbpo_pursuit(const float *V_in, int64_t C, int64_t N, const float *T, int32_t U, int32_t L,
                 const float *bias, const float *thr, int64_t n_max, int32_t do_recheck,
                 int32_t nthreads, int64_t cap, uint32_t *out_idx, uint32_t *out_lab, int64_t *n_out,
                 float *resid_out, int32_t max_pass_log, int64_t *pass_log, int32_t *n_passes)
{
    if (C <= 0 || N <= 0 || U <= 0 || L <= 0 || n_max < 0 || !V_in || !T || !bias || !thr)
        return FBPO_EINVAL;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    tset ts; tset_init(&ts, T, C, N, U, L);
    const int64_t nwin = (N + L - 1) / L;
    float *V = (float *)malloc((size_t)C * N * sizeof(float));
    float *LIK = (float *)calloc((size_t)U * N, sizeof(float));
    pstate st;
    st.nwin = nwin;
    st.best_idx = (uint32_t *)calloc(nwin, 4); st.best_lab = (uint32_t *)calloc(nwin, 4);
    st.obest_idx = (uint32_t *)calloc(nwin, 4); st.obest_lab = (uint32_t *)calloc(nwin, 4);
    st.best_lik = (float *)calloc(nwin, 4);
    st.check = (uint8_t *)calloc(nwin, 1); st.orc = (uint8_t *)calloc(nwin, 1);
    uint32_t *win = (uint32_t *)malloc(nwin * 4), *flist = (uint32_t *)malloc(nwin * 4);
    uint8_t *has = (uint8_t *)malloc(nwin), *snap = (uint8_t *)malloc(nwin);
    uint32_t *sp_idx = (uint32_t *)malloc(nwin * 4), *sp_lab = (uint32_t *)malloc(nwin * 4);
    int rc = FBPO_OK;
    if (!V || !LIK || !st.best_idx || !st.best_lab || !st.obest_idx || !st.obest_lab || !st.best_lik ||
        !st.check || !st.orc || !win || !flist || !has || !snap || !sp_idx || !sp_lab) {
        rc = FBPO_ENOMEM; goto done;
    }
    memcpy(V, V_in, (size_t)C * N * sizeof(float));
    for (int64_t i = 0; i < nwin; i++) win[i] = (uint32_t)i;
    int64_t nw = nwin, total = 0;
    int32_t pass = 0;
    for (;;) {
        full_likelihood(&ts, V, bias, win, nw, LIK);
        window_max(&ts, LIK, win, nw, &st);
        int64_t nf = 0;
        for (int64_t i = 0; i < nwin; i++) if (st.orc[i]) flist[nf++] = (uint32_t)i;
        int64_t items_eval = 0;
        if (nf > 0 && do_recheck) {
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : items_eval)
            for (int64_t w = 0; w < nf; w++)
                recheck_window(&ts, V, LIK, thr, n_max, flist[w], &st, &items_eval);
            reassign(&ts, flist, nf, &st, snap);
        }
        int64_t nadd = pursuit(&ts, thr, win, nw, &st, has, snap, sp_idx, sp_lab);
        if (pass < max_pass_log && pass_log) {
            pass_log[4 * pass + 0] = nadd;
            pass_log[4 * pass + 1] = nw;
            pass_log[4 * pass + 2] = nf;
            pass_log[4 * pass + 3] = items_eval;
        }
        pass++;
        if (nadd == 0) break;
        if (total + nadd > cap) { rc = FBPO_ECAP; goto done; }
        memcpy(out_idx + total, sp_idx, (size_t)nadd * 4);
        memcpy(out_lab + total, sp_lab, (size_t)nadd * 4);
        total += nadd;
        memset(st.orc, 0, (size_t)nwin);
        nw = 0;
        for (int64_t i = 0; i < nwin; i++) if (st.check[i]) win[nw++] = (uint32_t)i;
        memset(st.check, 0, (size_t)nwin);
        residual(&ts, V, sp_idx, sp_lab, nadd);
    }
    *n_out = total;
    if (n_passes) *n_passes = pass;
    if (resid_out) memcpy(resid_out, V, (size_t)C * N * sizeof(float));
done:
    free(V); free(LIK); free(st.best_idx); free(st.best_lab); free(st.obest_idx); free(st.obest_lab);
    free(st.best_lik); free(st.check); free(st.orc); free(win); free(flist); free(has); free(snap);
    free(sp_idx); free(sp_lab);
    tset_free(&ts);
    return rc;
}

/*
 * Single steps of the pass loop on state the caller holds (arrays of nwin =
 * ceil(N / L) windows), so that known-answer tests can pin the window
 * bookkeeping rule by rule:
 *   fbpo_window_max_step  compute_template_maximum_likelihood for t = 0..U-1
 *                         over the listed windows (binary_pursuit.cl:350-441)
 *   fbpo_reassign_step    check_overlap_reassignments over the flagged
 *                         windows (binary_pursuit.cl:822-875)
 *   fbpo_pursuit_step     binary_pursuit over the listed windows
 *                         (binary_pursuit.cl:918-1063); returns the number of
 *                         spikes written to out_idx / out_lab
 */
static tset step_tset(int64_t N, int32_t U, int32_t L)
{
    tset ts;
    ts.T = NULL; ts.C = 0; ts.N = N; ts.U = U; ts.L = L; ts.chan_nz = NULL;
    return ts;
}

int fbpo_window_max_step(const float *LIK, int64_t N, int32_t U, int32_t L, const uint32_t *win, int64_t nw,
                         uint32_t *best_idx, uint32_t *best_lab, float *best_lik, uint32_t *obest_idx,
                         uint32_t *obest_lab, uint8_t *check, uint8_t *orc)
{
    if (N <= 0 || U <= 0 || L <= 0 || nw < 0) return FBPO_EINVAL;
    tset ts = step_tset(N, U, L);
    pstate st;
    st.nwin = (N + L - 1) / L;
    st.best_idx = best_idx; st.best_lab = best_lab; st.best_lik = best_lik;
    st.obest_idx = obest_idx; st.obest_lab = obest_lab; st.check = check; st.orc = orc;
    window_max(&ts, LIK, win, nw, &st);
    return FBPO_OK;
}

int fbpo_reassign_step(int64_t N, int32_t L, const uint32_t *flist, int64_t nf, uint32_t *best_idx,
                       uint32_t *best_lab, uint32_t *obest_idx, uint32_t *obest_lab, uint8_t *check, uint8_t *orc)
{
    if (N <= 0 || L <= 0 || nf < 0) return FBPO_EINVAL;
    tset ts = step_tset(N, 1, L);
    pstate st;
    st.nwin = (N + L - 1) / L;
    st.best_idx = best_idx; st.best_lab = best_lab; st.best_lik = NULL;
    st.obest_idx = obest_idx; st.obest_lab = obest_lab; st.check = check; st.orc = orc;
    uint8_t *snap = (uint8_t *)malloc((size_t)st.nwin);
    if (!snap) return FBPO_ENOMEM;
    reassign(&ts, flist, nf, &st, snap);
    free(snap);
    return FBPO_OK;
}

int64_t fbpo_pursuit_step(int64_t N, int32_t L, const float *thr, const uint32_t *win, int64_t nw,
                          uint32_t *best_idx, uint32_t *best_lab, float *best_lik, uint8_t *check, uint8_t *orc,
                          uint32_t *out_idx, uint32_t *out_lab)
{
    if (N <= 0 || L <= 0 || nw < 0) return FBPO_EINVAL;
    tset ts = step_tset(N, 1, L);
    pstate st;
    st.nwin = (N + L - 1) / L;
    st.best_idx = best_idx; st.best_lab = best_lab; st.best_lik = best_lik;
    st.obest_idx = NULL; st.obest_lab = NULL; st.check = check; st.orc = orc;
    uint8_t *has = (uint8_t *)malloc((size_t)(nw > 0 ? nw : 1));
    uint8_t *snap = (uint8_t *)malloc((size_t)st.nwin);
    if (!has || !snap) { free(has); free(snap); return FBPO_ENOMEM; }
    int64_t n = pursuit(&ts, thr, win, nw, &st, has, snap, out_idx, out_lab);
    free(has); free(snap);
    return n;
}

/* get_adjusted_clips, binary_pursuit.cl:1088-1130: clip = residual + template. */
int fbpo_adjusted_clips(const float *R, int64_t C, int64_t N, const float *T, int32_t U, int32_t L,
                        const uint32_t *idx, const uint32_t *lab, int64_t n, float *clips)
{
    for (int64_t k = 0; k < n; k++) {
        float *ck = clips + (size_t)k * C * L;
        memset(ck, 0, (size_t)C * L * sizeof(float));
        if (lab[k] >= (uint32_t)U) continue;
        for (int64_t c = 0; c < C; c++) {
            const float *tr = T + ((size_t)lab[k] * C + c) * L;
            const float *vr = R + (size_t)c * N;
            for (int64_t i = 0; i < L; i++) {
                if (idx[k] + i >= N) break;
                ck[c * L + i] = vr[idx[k] + i] + tr[i];
            }
        }
    }
    return FBPO_OK;
}

/* Full likelihood of every position for every template (diagnostics/tests). */
int fbpo_full_likelihood(const float *V, int64_t C, int64_t N, const float *T, int32_t U, int32_t L,
                         const float *bias, float *LIK)
{
    tset ts; tset_init(&ts, T, C, N, U, L);
    const int64_t nwin = (N + L - 1) / L;
    uint32_t *win = (uint32_t *)malloc(nwin * 4);
    if (!win) { tset_free(&ts); return FBPO_ENOMEM; }
    for (int64_t i = 0; i < nwin; i++) win[i] = (uint32_t)i;
    full_likelihood(&ts, V, bias, win, nwin, LIK);
    free(win);
    tset_free(&ts);
    return FBPO_OK;
}
