/*
 * fbp.h -- C ABI of the MI355X full-binary-pursuit library (libfbp_hip.so).
 *
 * This boundary replaces the device half of the reference's
 *   binary_pursuit(voltage, v_dtype, sort_info, separability_metrics,
 *                  n_max_shift_inds=None, kernels_path=None, max_gpu_memory=None)
 * (spikesorting_fullpursuit/overlap/binary_pursuit_parallel.py:175-183), i.e. the
 * pyopencl program build, buffers and the pass loop of :296-1500 together with
 * the eight OpenCL kernels of spikesorting_fullpursuit/kernels/binary_pursuit.cl.
 * Chunk planning and output assembly stay in the Python drop-in.
 *
 * Conventions: plain pointers and sizes; every function returns an int status
 * (FBP_OK or a negative FBP_E_*); fbp_last_error() gives the message.  One
 * handle owns one HIP stream and its device buffers; a handle must not be used
 * from two threads at once.  Voltage is [C][N] float32 with row stride `ld`
 * elements; templates are [U][C][L] float32 (the reference's flattened
 * templates_vector, binary_pursuit_parallel.py:239-246).
 */
#ifndef FBP_H
#define FBP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FBP_OK 0
#define FBP_E_INVALID -1     /* bad argument           -> ValueError   */
#define FBP_E_NODEVICE -2    /* no usable GPU          -> RuntimeError("No GPU available"), binary_pursuit_parallel.py:292-293 */
#define FBP_E_HIP -3         /* HIP runtime failure    -> RuntimeError */
#define FBP_E_NOMEM -4       /* device allocation      -> MemoryError  */
#define FBP_E_UNSUPPORTED -5 /* shape outside limits   -> ValueError   */
#define FBP_E_STATE -6       /* call order             -> RuntimeError */

/* fbp_run flags */
#define FBP_FLAG_OVERLAP_RECHECK 1u  /* sort_info["do_overlap_recheck"], binary_pursuit_parallel.py:980-983 */
#define FBP_FLAG_ADJUSTED_CLIPS 2u   /* sort_info["get_adjusted_clips"], binary_pursuit_parallel.py:1404-1486 */
#define FBP_FLAG_RECHECK_LITERAL 4u  /* evaluate every overlap item literally (no screening) */
#define FBP_FLAG_TIMING 8u           /* record per-kernel device time with HIP events */
#define FBP_FLAG_SCREEN_FP32 16u     /* overlap screen correlations on FP32 vector units instead of split-fp16 GEMMs */
#define FBP_FLAG_PRECISION_MIXED 32u /* likelihood by fp16 GEMM + exact FP32 rescoring near and above zero (same spikes) */
#define FBP_FLAG_SELECT_STREAMING 64u /* overlap screen selection without register-resident pair state (taken anyway when U*(2n+1) > 24576) */
#define FBP_FLAG_LIKELIHOOD_MFMA 128u /* with FBP_FLAG_PRECISION_MIXED: the matrix-core first pass even where the tiles' channel padding predicts the FP32 kernel faster (same spikes either way) */
#define FBP_FLAG_PRECISION_BF16 256u /* likelihood of every pass by BF16 matrix-core products (one template per tile, its own channels) + exact FP32 rescoring near and above zero (same spikes) */
#define FBP_FLAG_PRECISION_FP8 512u  /* the same with FP8 (e4m3) operands */
/* Diagnostics (not for sorting): stop after the first likelihood pass (no spikes; the likelihood
 * of every position stays for fbp_copy_likelihood), and keep every BF16 / FP8 approximation
 * (no exact rescoring, so decisions are no longer guaranteed to be those of FP32). */
#define FBP_FLAG_FIRST_PASS_ONLY 1024u
#define FBP_FLAG_NO_RESCORE 2048u
/* with FBP_FLAG_PRECISION_BF16 / _FP8: the matrix cores take every pass (one template per tile over the listed
 * windows), not only the dense first pass (later passes then run the FP32 kernel on the changed pairs) */
#define FBP_FLAG_MMA_EVERY_PASS 4096u
/* fbp_run_batch only: every pass after the first is one replay of a hipGraph of the pass (its launch
 * sizes fixed when the graph is built, its item counts written on the device), with one host read per
 * replay; the graph is rebuilt when a segment lists more windows than it holds or far fewer in all.
 * Such a call runs only while no other fbp_run / fbp_run_batch is in flight in the process, and other
 * calls fail with FBP_E_STATE until it returns (graph calls stay on one host thread at a time).
 * Switched off in this build: fbp_run_batch returns FBP_E_UNSUPPORTED for it (a one-lane graph batch ended
 * in a host segmentation fault whose cause is not known). */
#define FBP_FLAG_BATCH_GRAPH 8192u

typedef struct fbp_handle fbp_handle;

typedef struct fbp_stats {
    int32_t passes;               /* iterations of the pass loop (binary_pursuit_parallel.py:907) */
    int32_t reserved0;
    int64_t spikes;               /* spikes found in the last run */
    int64_t windows_listed;       /* sum over passes of listed windows */
    int64_t windows_flagged;      /* sum over passes of overlap-recheck windows */
    int64_t likelihood_positions; /* positions x templates recomputed */
    double likelihood_flops;      /* 2 * sum over recomputed (x, t) of nz_channels(t) * L */
    double ms_total;              /* device time of the run (events around the whole run) */
    double ms_likelihood;         /* per kernel class, only with FBP_FLAG_TIMING */
    double ms_window_max;
    double ms_recheck;
    double ms_pursuit;
    double ms_residual;
    double ms_other;
    int64_t n_likelihood_launches;
    int64_t n_recheck_launches;
    int64_t recheck_items;        /* overlap items kept by the screen as candidates */
    int64_t recheck_overflow;     /* windows whose candidates exceeded the on-chip list (slow path) */
    int64_t likelihood_rescored;  /* mixed precision: (position, template) pairs rescored in exact FP32 */
    int64_t screen_shared_pairs;  /* (unit, shift) pairs sharing a channel enumerated over f by the screen */
    int64_t screen_separable_pairs; /* (unit, shift) pairs sharing no channel enumerated over f */
    double ms_select;             /* overlap screen selection kernel (part of ms_recheck), with FBP_FLAG_TIMING */
    int64_t n_select_launches;
    int64_t likelihood_pairs_computed; /* (window, template) pairs whose likelihood was computed; the others
                                          had no channel change in the window's samples since their last
                                          computation and kept their (identical) values */
    int64_t n_likelihood_mfma_launches; /* launches of the matrix-core likelihood pass (mixed precision) */
    double ms_rescore;            /* deferred exact rescoring of the overlap candidates (part of ms_recheck),
                                     with FBP_FLAG_TIMING */
    int64_t n_rescore_launches;
    int64_t n_graph_builds;       /* fbp_run_batch with FBP_FLAG_BATCH_GRAPH: pass graphs built ... */
    int64_t n_graph_replays;      /* ... and replayed (counts of the whole batch) */
} fbp_stats;

/* Create a handle on HIP device `device` (replaces the device search and
 * program build of binary_pursuit_parallel.py:282-300). */
int fbp_create(int device, fbp_handle **out);
int fbp_destroy(fbp_handle *h);
const char *fbp_last_error(const fbp_handle *h);

/* Total device memory in bytes (GLOBAL_MEM_SIZE for the chunk plan,
 * binary_pursuit_parallel.py:387-409). */
int fbp_device_memory(fbp_handle *h, int64_t *total_bytes);

/* Upload templates [U][C][L], bias[U] = float32(-0.5*template_SS) and
 * thresholds[U] = float32(neuron_lower_thresholds); host pointers
 * (binary_pursuit_parallel.py:239-264, :480-496). */
int fbp_set_templates(fbp_handle *h, const float *templates, int32_t U, int32_t C, int32_t L,
                      const float *bias, const float *thresholds);

/* Run the pass loop on one chunk V[c][x], c < C, x < N, row stride ld.
 * v_on_device != 0: V is a device pointer (it is copied, never modified).
 * Replaces one iteration of the chunk loop, binary_pursuit_parallel.py:510-1500. */
int fbp_run(fbp_handle *h, const float *V, int64_t ld, int64_t N, int32_t v_on_device,
            uint32_t n_max_shift_inds, uint32_t flags);

/* Batched pursuit: the pass loops of nseg segments run together, each kernel
 * launched once per pass over every segment still converging, with two host
 * reads of all segments' counters per pass (one per pass, by hipGraph replays,
 * with FBP_FLAG_BATCH_GRAPH).  Segment g is the chunk V[g]
 * (row stride ld[g], N[g] samples; device pointers when v_on_device != 0)
 * pursued with the templates set on hs[g] and n_max_shift[g]; its spikes,
 * residual, statistics and pass log are read back through hs[g] exactly as
 * after fbp_run, and equal fbp_run's on the same segment (the likelihood and
 * the screened overlap recheck run in FP32; precision flags are ignored;
 * FBP_FLAG_ADJUSTED_CLIPS gives every segment its clips).
 * The handles are distinct, on one device, with one template length and at
 * most 64 channels each; launches go on hs[0]'s stream.  Tile shape (results
 * do not depend on it; 0 = default): block_cap = blocks per segment and
 * launch at most (1024), lik_windows = windows per likelihood block (the
 * single-segment plan's).
 * Replaces, for many independent segments, the per-chunk loop of
 * binary_pursuit_parallel.py:508-1500 called once per segment
 * (spikesorting_parallel.py:2161-2184). */
int fbp_run_batch(fbp_handle *const *hs, int32_t nseg, const float *const *V, const int64_t *ld, const int64_t *N,
                  int32_t v_on_device, const uint32_t *n_max_shift, uint32_t flags, int32_t block_cap,
                  int32_t lik_windows);

/* Results of the last fbp_run: spike onsets within the chunk and labels in
 * pass order (ascending window within a pass). */
int fbp_result_count(fbp_handle *h, int64_t *n);
int fbp_copy_result(fbp_handle *h, uint32_t *indices, uint32_t *labels);
/* Adjusted clips [n][C*L] of the last run (requires FBP_FLAG_ADJUSTED_CLIPS). */
int fbp_copy_clips(fbp_handle *h, float *clips);
/* Residual voltage [C][N] of the last run into a host buffer. */
int fbp_copy_residual(fbp_handle *h, float *out);
int fbp_get_stats(fbp_handle *h, fbp_stats *out);
/* Likelihood [U][N] of the last fbp_run into a host buffer (the values the last pass left:
 * with FBP_FLAG_FIRST_PASS_ONLY, the first pass over every position). */
int fbp_copy_likelihood(fbp_handle *h, float *out);
/* Per-pass trace of the last fbp_run, one row of three int64 per pass:
 * spikes added (num_additional_spikes, binary_pursuit_parallel.py:1210-1217),
 * windows listed (new_window_indices, :1242) and windows flagged for the
 * overlap recheck (overlap_window_indices, :977-979; 0 when the recheck is
 * off).  Copies min(n_rows, max_rows) rows; *n_rows = number of passes. */
int fbp_get_pass_log(fbp_handle *h, int64_t *rows, int32_t max_rows, int32_t *n_rows);

/*
 * Template refinement at fixed spike times.  No counterpart in the reference:
 * it subtracts each found template once (binary_pursuit.cl:57-102) and never
 * differentiates.  These calls serve a refinement loop that moves templates
 * and per-spike amplitudes a_k downhill on the residual energy
 *
 *   R[c][x] = V[c][x] - sum_k a_k T[lab_k][c][x - p_k]   (0 <= x - p_k < L, x < N)
 *   E       = 0.5 * sum_{c,x} R[c][x]^2                  (FP64 sums)
 *
 * with p_k = idx[k].  Every pointer is a DEVICE pointer; the work is queued on
 * `stream` (a hipStream_t, NULL = the default stream) and the calls neither
 * allocate nor synchronise.  All sums have a fixed order (no floating-point
 * atomics): results are identical run to run.  Handle-free: on failure the
 * message is fbp_refine_last_error() of the calling thread.
 */

/* Size in bytes of the `work` buffer both calls below need. */
int fbp_refine_workspace_bytes(int64_t N, int64_t n_spikes, int32_t U, int32_t C, int32_t L, int64_t *bytes);

/* R (row stride ldr) and energy[0] (double).  Spikes must be sorted by onset
 * (idx ascending) with labels below U; otherwise energy[0] is NaN.  Spikes
 * covering a sample are subtracted in list order, r = fmaf(-a_k, T, r). */
int fbp_refine_forward(const float *V, int64_t ldv, int64_t N, int32_t C, const float *T, int32_t U, int32_t L,
                       const uint32_t *idx, const uint32_t *lab, const float *amp, int64_t n_spikes,
                       float *R, int64_t ldr, double *energy, void *work, int64_t work_bytes, void *stream);

/* fbp_refine_forward without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * per-block partials live in a static device array of 3 MiB per GPU (held
 * until the process ends), so N is limited to 67108864 samples (larger N:
 * FBP_E_UNSUPPORTED).  Allocates nothing; the stream must belong to the
 * current device.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other threads) race.  Results equal
 * fbp_refine_forward's bit for bit. */
int fbp_refine_forward_noalloc(const float *V, int64_t ldv, int64_t N, int32_t C, const float *T, int32_t U,
                               int32_t L, const uint32_t *idx, const uint32_t *lab, const float *amp,
                               int64_t n_spikes, float *R, int64_t ldr, double *energy, void *stream);

/* Gradients of E from the residual R of fbp_refine_forward:
 *   grad_T[u][c][i] = -sum_{k of unit u, p_k+i < N} a_k R[c][p_k+i]     ([U][C][L], may be NULL)
 *   grad_amp[k]     = -sum_{c, i: p_k+i < N} T[lab_k][c][i] R[c][p_k+i] ([n_spikes], may be NULL)
 * by_unit lists the spike ids grouped by label and unit_offsets[U+1] is its
 * CSR index (unit u owns by_unit[unit_offsets[u] .. unit_offsets[u+1]));
 * grad_T sums each unit's spikes in chunks of 256 in by_unit order. */
int fbp_refine_backward(const float *R, int64_t ldr, int64_t N, int32_t C, const float *T, int32_t U, int32_t L,
                        const uint32_t *idx, const uint32_t *lab, const float *amp, int64_t n_spikes,
                        const uint32_t *by_unit, const uint32_t *unit_offsets, float *grad_T, float *grad_amp,
                        void *work, int64_t work_bytes, void *stream);

const char *fbp_refine_last_error(void);

/* Multi-channel clips of get_clips (SF/processing/clip_utils.py:501-632, run by
 * full binary pursuit over every found spike and channel,
 * SF/overlap/full_binary_pursuit.py:495-502):
 *   clips[k][q*(w1-w0) + j] = V[chans[q]][event_indices[k] + w0 + j]
 * with [w0, w1) the sample window of time_window_to_samples (w0 <= 0).  V and
 * clips hold elements of elem_bytes (2, 4 or 8) bytes, moved bit for bit (any
 * 16-, 32- or 64-bit sample type).  Device pointers, queued on `stream`, no
 * allocation.  The caller passes the events get_clips keeps (clips inside
 * [0, N)); samples outside read as 0.  Handle-free: the message of a failure
 * is fbp_clips_last_error(). */
int fbp_get_clips(const void *V, int32_t elem_bytes, int64_t ldv, int64_t N, int32_t C, const int32_t *chans,
                  int32_t nchan, const int64_t *event_indices, int64_t n_events, int32_t w0, int32_t w1, void *clips,
                  void *stream);
const char *fbp_clips_last_error(void);

/*
 * Inputs of the pursuit (device pointers, queued on `stream`, no allocation,
 * fixed summation orders).  Handle-free: fbp_inputs_last_error() on failure.
 */

/* Noise covariance of clip residuals per channel, as
 * get_channel_covariance_matrix (SF/overlap/full_binary_pursuit.py:863-923)
 * computes it with np.cov(rowvar=False, ddof=0):
 *   X_c = R[rows[i]][c*L .. c*L+L)   (i < n_rows; rows NULL: rows 0 .. n_rows-1)
 *   cov[c][i][j] = sum (X_c - mean)_i (X_c - mean)_j / n_rows        (FP64, [C][L][L])
 * R is [n_total][>= C*L] float32 with row stride ld; L <= 128. */
int fbp_covariance_workspace_bytes(int64_t n_rows, int32_t C, int32_t L, int64_t *bytes);
int fbp_channel_covariance(const float *R, int64_t ld, int64_t n_total, int32_t C, int32_t L, const int64_t *rows,
                           int64_t n_rows, double *cov, void *work, int64_t work_bytes, void *stream);

/* Robust template of one unit's clips [n][W] (row stride ld), as
 * calculate_robust_template (SF/processing/clip_utils.py:202-247) with numpy's
 * float32 semantics, per column: the median of the values within one MAD of
 * the median (the median when none is).  Exact order statistics. */
int fbp_robust_template(const float *clips, int64_t n, int64_t ld, int64_t W, float *out, void *stream);

/* var[u] = sum_c t_uc^T cov[c] t_uc for templates T [U][C][L] and cov
 * [C][L][L], FP64 (compute_separability_metrics,
 * SF/overlap/neuron_separability.py:158-183). */
int fbp_template_variances(const double *T, int32_t U, int32_t C, int32_t L, const double *cov, double *var,
                           void *stream);
const char *fbp_inputs_last_error(void);

/*
 * Voltage preprocessing (device pointers, queued on `stream`, no allocation).
 * Handle-free: fbp_preproc_last_error() on failure.  V is [C][N] float32 with
 * row stride ld.
 */
int fbp_preproc_workspace_bytes(int64_t N, int32_t C, int64_t *bytes);
/* medians[c] = numpy float32 nanmedian(|V[c]|) (SF/threshold/threshold.py:6-70):
 * the middle value, or fl32(fl32(lo + hi) / 2) for an even count; NaN when a
 * row is all NaN.  Exact order statistics. */
int fbp_abs_medians(const float *V, int64_t ld, int64_t N, int32_t C, float *medians, void *work, int64_t work_bytes,
                    void *stream);
/* counts[c] = number of x with |V[c][x]| > thresholds[c] (compared in FP64;
 * threshold.py:51-70). */
int fbp_count_above(const float *V, int64_t ld, int64_t N, int32_t C, const double *thresholds, uint64_t *counts,
                    void *stream);
/* cov = np.cov(V, ddof=ddof) (rows are channels), FP64 [C][C], C <= 64
 * (get_full_zca_matrix, SF/processing/zca.py:5-31). */
int fbp_channel_cov_rows(const float *V, int64_t ld, int64_t N, int32_t C, int32_t ddof, double *cov, void *work,
                         int64_t work_bytes, void *stream);
/* out[c][x] = float32(sum_k Z[c][k] V[k][x]) with FP64 products and sums, k
 * ascending: the caller's (zca_matrix @ voltage).astype(float32)
 * (SF/parallel/spikesorting_parallel.py:295), C <= 64. */
int fbp_zca_apply(const double *Z, const float *V, int64_t ldv, int64_t N, int32_t C, float *out, int64_t ldo,
                  void *stream);
const char *fbp_preproc_last_error(void);

/* Shift search of find_overlap_templates (SF/c_cython/sort_cython.pyx:333-462):
 * for each (test, n1, n2) of triples [n][3] and every shift pair (s1, s2) in
 * [-n_pre, n_post]^2, R = sum_e (test_e - (a_e + b_e))^2 with a, b the shifted
 * templates of n1, n2 (T [U][C*L] float32), in numpy's FP64 operation order
 * (pairwise np.sum), bit-identical.  best_value[k] is the first minimum in
 * s1-major order and best_shift_pair[k] = (s1 + n_pre) * S + (s2 + n_pre),
 * S = n_pre + n_post + 1.  C * L <= 16384.  Device pointers, `stream`;
 * handle-free: fbp_qa_last_error() on failure. */
int fbp_overlap_residuals(const float *T, int32_t U, int32_t C, int32_t L, int32_t n_pre, int32_t n_post,
                          const int32_t *triples, int64_t n_triples, double *best_value, int32_t *best_shift_pair,
                          void *stream);
const char *fbp_qa_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FBP_H */
