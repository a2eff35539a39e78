/*
 * C ABI of the MI355X (gfx950) BLS engine.
 *
 * Plain pointers and sizes only. Every array argument that names a device buffer is a
 * HIP device pointer on the context's device; `stream` is a hipStream_t (NULL = the
 * null stream). Calls enqueue work on `stream` and return without waiting, unless
 * stated otherwise. A context is not thread-safe: use one per thread or process
 * (as cuPeriod's engines, src/cuperiod/methods/base.py:388-396). Distinct contexts may
 * be used concurrently.
 *
 * Status codes: CPB_OK, or one of the CPB_E_* values; cpb_last_error() returns the
 * message of the last failure on a context.
 */
#ifndef CUPERIOD_BLS_H
#define CUPERIOD_BLS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPB_OK 0
#define CPB_E_INVALID 1     /* bad argument (maps to ValueError) */
#define CPB_E_DEVICE 2      /* HIP runtime failure (maps to DeviceError) */
#define CPB_E_RESOURCE 3    /* a resource limit, e.g. too many durations (BackendUnavailableError) */
#define CPB_E_NODEVICE 4    /* no usable GPU (BackendUnavailableError) */

#define CPB_N_OUTPUTS 7     /* power, depth, depth_err, depth_snr, duration, transit_time, log_likelihood */

typedef struct cpb_context cpb_context;

/* One duration-homogeneous period block of a light curve: the `_SearchSpec` of the
 * reference (src/cuperiod/methods/_bls_core.py:687-706). Periods are explicit
 * (mode 0: `periods[period_offset + k]`) or the reciprocals of a parametric frequency
 * grid f_0..f_{num-1}, taken in ascending-period order, period k = 1 / f_{num-1-(k0+k)}:
 *   mode 1, np.linspace(start, stop, num): f_i = i*step + start, f_{num-1} = stop,
 *           step = (stop - start)/(num - 1) (see cpb_linspace_grid);
 *   mode 2, np.arange(start, ., step'):   f_i = start + i*step', with step' =
 *           (start + step) - start as numpy forms it (the BLS default-grid segments,
 *           src/cuperiod/methods/bls.py:96-104). */
typedef struct cpb_segment {
    double bin_duration;    /* min(durations) / oversample */
    int64_t period_offset;  /* mode 0: index of the first period in `periods` */
    int64_t n_periods;      /* periods in this segment */
    int64_t out_offset;     /* column of the first period in the outputs */
    int32_t width;          /* ceil(max period / bin_duration) + oversample + 1 */
    int32_t dur_offset;     /* first box width of this segment in `dur_bins` */
    int32_t n_dur;          /* box widths (bins) of this segment */
    int32_t curve;          /* light-curve index (0 for single-curve calls) */
    int32_t mode;           /* 0: explicit; 1: linspace; 2: arange (frequencies) */
    int32_t group;          /* the reference segment this block belongs to: blocks of one
                               curve with equal `group` share the band of
                               cpb_set_peak_rounding */
    double lin_start;       /* modes 1-2: first frequency */
    double lin_step;        /* modes 1-2: frequency step as numpy applies it */
    double lin_stop;        /* mode 1: last frequency (np.linspace sets it exactly) */
    int64_t lin_num;        /* modes 1-2: frequencies in the grid */
    int64_t lin_k0;         /* modes 1-2: ascending-period index of the first period */
} cpb_segment;

/* Per light curve inputs of a batched search. */
typedef struct cpb_curve {
    int64_t point_offset;   /* first point in the concatenated tau/yw/wv arrays */
    int64_t n_points;
    double t_min;           /* time origin subtracted from tau (added to transit_time) */
    double ivar_floor;      /* `_ivar_floor(N, sum(ivar))` (src/cuperiod/methods/_bls_core.py:75-94) */
    double sum_y;           /* sum(y*ivar) and sum(ivar) as the reference's host forms them */
    double sum_ivar;        /*   (numpy sums, src/cuperiod/methods/_bls_core.py:483-484) */
} cpb_curve;

/* Library version (major*10000 + minor*100 + patch). */
int cpb_version(void);

/* Number of HIP devices visible to the process (0 when none). */
int cpb_device_count(void);

/* Create a context bound to HIP device `device`. */
int cpb_create(int device, cpb_context** out);

/* Release a context and its device buffers. NULL is accepted. */
void cpb_destroy(cpb_context* ctx);

/* Message of the last failure on `ctx` ("" when none). */
const char* cpb_last_error(const cpb_context* ctx);

/* Static description of a status code. */
const char* cpb_status_string(int status);

/*
 * BLS box search of one light curve over several period segments, with one set of
 * device inputs and one (7, n_out) float64 output.
 *
 * Replaces `_bls_search_cuda_specs` (src/cuperiod/methods/_bls_core.py:741-840) and the
 * `bls_block` kernel it launches per segment (:505-656): the same inputs (tau = t - t_min,
 * yw = y*ivar, wv = ivar; the segment binning; oversample; objective; ivar floor) and
 * the same seven outputs per period. Results are exact-arithmetic box statistics
 * (64-bit fixed-point phase bins), rounded once to float64, except at the periods whose
 * power lies within 2^-30 of the curve's largest (see cpb_set_peak_rounding).
 *
 * tau, yw, wv   device float64[n_points]
 * sum_y, sum_ivar  the reference's host sums of yw and wv (`sum_y`, `sum_ivar` of the
 *               kernel's arguments, :514)
 * periods       device float64[...], explicit periods of mode-0 segments (may be NULL)
 * dur_bins      device int32[...], box widths of every segment, concatenated
 * segments      HOST array of n_segments descriptors (curve must be 0)
 * obj_flag      0 = depth SNR, 1 = log-likelihood
 * out           device float64[CPB_N_OUTPUTS * n_out], row k at out + k*n_out
 */
int cpb_bls_search(cpb_context* ctx, void* stream, const double* tau, const double* yw,
                   const double* wv, int64_t n_points, double t_min, double ivar_floor,
                   double sum_y, double sum_ivar, const double* periods, const int32_t* dur_bins,
                   const cpb_segment* segments, int32_t n_segments, int32_t oversample,
                   int32_t obj_flag, double* out, int64_t n_out);

/*
 * Batched search: many light curves, each with its own segments, in one launch
 * sequence (the reference loops over curves, src/cuperiod/batch/runner.py:214-250).
 * Points of all curves are concatenated; `curves` and `segments` are HOST arrays;
 * segment.curve indexes `curves`; outputs of all segments share one (7, n_out) array.
 */
int cpb_bls_search_batch(cpb_context* ctx, void* stream, const double* tau,
                         const double* yw, const double* wv, int64_t total_points,
                         const cpb_curve* curves, int32_t n_curves,
                         const double* periods, const int32_t* dur_bins,
                         const cpb_segment* segments, int32_t n_segments,
                         int32_t oversample, int32_t obj_flag, double* out,
                         int64_t n_out);

/*
 * cpb_bls_search_batch without allocating: no hipMalloc, hipHostMalloc or hipFree (nor
 * any other allocating HIP routine) in the call. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation). Same arguments, kernels and
 * outputs. The search's scratch (fixed-point copies of the points, 16 bytes per point;
 * curve parameters; descriptor tables; the peak rounding's candidate lists) is carved
 * from a static device array of 64 MiB per GPU, held until the process ends; a search
 * that needs more (about 4 million points, or phase spans kept in global memory)
 * returns CPB_E_RESOURCE and should use cpb_bls_search_batch. Descriptors are copied
 * from pageable host memory. Every call on one GPU shares the array: calls that overlap
 * on one GPU (other streams, other contexts, other threads) race.
 */
int cpb_bls_search_batch_static(cpb_context* ctx, void* stream, const double* tau,
                                const double* yw, const double* wv, int64_t total_points,
                                const cpb_curve* curves, int32_t n_curves,
                                const double* periods, const int32_t* dur_bins,
                                const cpb_segment* segments, int32_t n_segments,
                                int32_t oversample, int32_t obj_flag, double* out,
                                int64_t n_out);

/*
 * Per-curve summary of power rows left on the device by a batched search (the SDE
 * statistics and best period of `_assemble` / `Periodogram.best_period`,
 * src/cuperiod/methods/bls.py:274-301, src/cuperiod/core/result.py:285-438):
 * for curve c, the powers power[curve_base[c] .. + curve_count[c]) in ascending-period
 * order, with NaN/+-inf read as 0, give
 *   summary[4c+0] = mean, summary[4c+1] = std (ddof 0),
 *   summary[4c+2] = index (within the curve) of the best period, summary[4c+3] = its power.
 * curve_base, curve_count: DEVICE int64[n_curves]; summary: DEVICE float64[4*n_curves].
 */
int cpb_bls_summarize(cpb_context* ctx, void* stream, const double* power,
                      const int64_t* curve_base, const int64_t* curve_count,
                      int32_t n_curves, double* summary);

/* Peak candidates of each curve's power row of `out` (columns curve_base[c] ..
 * curve_base[c] + curve_count[c], ascending period), as `Periodogram._peak_candidates`
 * defines them on the sanitised power in ascending frequency
 * (src/cuperiod/core/result.py:369-406): position x = curve_count[c]-1-k for column k.
 * Every candidate at or above a per-curve threshold that keeps at least `want` of them
 * (all when there are fewer) is written, at most `cap` per curve, in no particular order:
 * cand[c*cap + i] = x and values[(c*cap + i)*5 + 0..4] = (sanitised power, depth,
 * depth_snr, duration, transit_time). counts[2c] = all candidates of the curve,
 * counts[2c+1] = candidates at or above the threshold (may exceed `cap`). Replaces the
 * (7, F) download per curve of cuPeriod's batch rows (src/cuperiod/batch/io.py:507-571)
 * with `cap` values per curve. */
int cpb_peak_candidates(cpb_context* ctx, void* stream, const double* out, int64_t n_out,
                        const int64_t* curve_base, const int64_t* curve_count,
                        int32_t n_curves, int32_t want, int32_t cap, int32_t* cand,
                        int64_t* counts, double* values);

/* One band of a multi-band star (its points in the concatenated tau/yw/wv arrays). */
typedef struct cpb_band {
    int64_t point_offset;   /* first point of the band */
    int64_t n_points;
    double ivar_floor;      /* `_ivar_floor(n_points, sum(ivar))` of the band alone */
} cpb_band;

/* One star: bands[band_offset .. band_offset + n_bands), their times measured from the
 * star's common origin t_min (`joint_bands`, src/cuperiod/multiband/_bls_joint.py:94-128). */
typedef struct cpb_star {
    int32_t band_offset;
    int32_t n_bands;        /* 1 .. 64 */
    double t_min;
} cpb_star;

#define CPB_N_JOINT_OUTPUTS 6  /* power, depth, depth_snr, duration, transit_time, band_index */

/*
 * Joint-ephemeris multi-band box search of one or more stars: per trial period, one box
 * (one duration, one start on the phase grid every band of the star shares), each band
 * fitting its own depth inside it; the box maximises the sum over the bands that dip in
 * it of depth^2/(1/W_in + 1/W_out) (obj_flag 0) or W_in*depth^2/2 (obj_flag 1). Outputs
 * per period: the quadrature power sum_k max(SNR_k, 0)^2 at that box, the depth of the
 * band with the largest SNR, sqrt(power), the box duration and mid-transit time, and
 * that band's index within the star (all 0 where no box is admissible).
 *
 * Replaces `_joint_cuda` (src/cuperiod/multiband/_bls_joint.py:703-796), its
 * `bls_joint_block` kernel (:531-685) and the host `_finish` (:359-406).
 *
 * tau, yw, wv  device float64[total_points]: times minus the star's t_min, median-centred
 *              y*ivar of each band, ivar
 * bands, stars, segments  HOST arrays; segment.curve indexes `stars`
 * periods, dur_bins       as for cpb_bls_search_batch
 * out          device float64[CPB_N_JOINT_OUTPUTS * n_out], row k at out + k*n_out
 */
int cpb_bls_joint_search(cpb_context* ctx, void* stream, const double* tau, const double* yw,
                         const double* wv, int64_t total_points, const cpb_band* bands,
                         int32_t n_bands, const cpb_star* stars, int32_t n_stars,
                         const double* periods, const int32_t* dur_bins,
                         const cpb_segment* segments, int32_t n_segments, int32_t oversample,
                         int32_t obj_flag, double* out, int64_t n_out);

/* Kernel of later joint searches on `ctx`: 0 (default) the band-group kernel whenever it
 * covers the segments' phase spans (1,024 threads; bands pass through LDS a group at a
 * time, box totals stay in registers), else every band of a period at once; 1 always
 * every band at once (bins in LDS, or in global scratch when they do not fit); 2 always
 * band groups (CPB_E_RESOURCE when they cannot cover the spans). Outputs agree to
 * round-off across modes: mode 1 forms box sums from exact integers, the band-group
 * kernel from FP64 prefix sums as the reference does. */
int cpb_set_joint_kernel(cpb_context* ctx, int mode);

/* One light curve of a TLS search. */
typedef struct cpb_tls_curve {
    int64_t point_offset;   /* first point in the concatenated tau/yw/wv arrays */
    int64_t n_points;
    int64_t period_offset;  /* first trial period of the curve in `periods` */
    int64_t n_periods;
    int64_t out_offset;     /* output column of its first period */
    double w_total;         /* sum of the curve's weights (the W of the level fit) */
} cpb_tls_curve;

#define CPB_N_TLS_OUTPUTS 4  /* sr, depth, duration, t0 (phase time from tau = 0) */

/*
 * TLS matched-filter search of one or more light curves: per trial period the points
 * are folded into `n_bins` phase bins (bin = int(frac(tau/P)*n_bins)) and every
 * limb-darkened template (widths[i] bins, samples templates[sum(widths[:i]) ..]) is slid
 * over every start phase (circularly), fitting level and depth jointly; outputs per
 * period: the best signal residue SR = num^2/D with D = den - cwg^2/W, its depth -num/D,
 * duration width/n_bins*P and the phase time of its centre, frac((start +
 * width/2)/n_bins)*P (0 and 0 where no dip qualifies).
 *
 * Replaces the `tls_block` kernel and `_tls_cuda` (src/cuperiod/methods/tls.py:270-479).
 *
 * tau, yw, wv  device float64[total_points]: times minus the curve's first time,
 *              w*(y - weighted mean), weights w (any common unit)
 * curves       HOST array; periods: device float64[...]
 * widths, templates  HOST arrays (1..32 widths, each in [1, n_bins])
 * n_bins       2 .. 4096; den_rel: the relative floor on D (1e-9 in FP64)
 * out          device float64[CPB_N_TLS_OUTPUTS * n_out], row k at out + k*n_out
 */
int cpb_tls_search(cpb_context* ctx, void* stream, const double* tau, const double* yw,
                   const double* wv, int64_t total_points, const cpb_tls_curve* curves,
                   int32_t n_curves, const double* periods, int32_t n_bins,
                   const int32_t* widths, int32_t n_widths, const double* templates,
                   double den_rel, double* out, int64_t n_out);

/* Frequencies of np.linspace(start, stop, num) by numpy's recipe (value_i = i*step +
 * start, last = stop), computed on the HOST: the values mode-1 segments use. */
int cpb_linspace_grid(double start, double stop, int64_t num, double* out);

/* Peak candidates of one spectrum on the HOST (replaces the candidate rule of
 * src/cuperiod/core/result.py:369-406, Periodogram._peak_candidates, which best_periods,
 * :285, applies to a BLS spectrum whose non-finite powers _assemble has set to 0,
 * src/cuperiod/methods/bls.py:280). power[0..n) is one output row of the engine, in ascending
 * period, i.e. descending frequency; the rule runs on the ascending-frequency spectrum
 * s[i] = power[n-1-i] (0 where not finite): interior local maxima (s[i] > s[i-1] and
 * s[i] >= s[i+1]) in ascending i, then i = 0 if s[0] >= s[1], then i = n-1 if
 * s[n-1] > s[n-2] (n == 1: i = 0); with none, the first largest sample.
 * index[0..*count) receives the ascending-frequency indices i in that order and
 * score[0..*count) their values s[i]; both hold up to n/2 + 2 entries (no two interior
 * maxima are adjacent, and an end that qualifies excludes its neighbour). */
int cpb_spectrum_candidates(const double* power, int64_t n, int64_t* index, double* score,
                            int64_t* count);

/* Greedy alias-diverse peak choice on the HOST (replaces the loop of
 * src/cuperiod/core/peaks.py:169-231, select_diverse_peaks, with its alias test
 * is_alias_related, :121-166): candidates are visited in the given order (most
 * significant first); one is chosen unless it is related, either way round, to a peak
 * already chosen -- |a - b| < separation, |a - m*b| < m*tol or |a - b/m| < tol for each
 * harmonic order m in orders[0..n_orders), |a - |b - o|| < tol or |a - (b + o)| < tol
 * for each alias offset o in offsets[0..n_offsets) -- until n_peaks are chosen.
 * freq[i] is candidate i's frequency; chosen[0..*n_chosen) receives their indices
 * into freq, *examined the number of candidates visited. Every comparison is the
 * same IEEE double expression the Python rule evaluates. */
int cpb_select_diverse_peaks(const double* freq, const int64_t* order, int64_t n,
                             int32_t n_peaks, double tol, double separation,
                             const double* orders, int32_t n_orders, const double* offsets,
                             int32_t n_offsets, int64_t* chosen, int32_t* n_chosen,
                             int64_t* examined);

/* Enable (1) or disable (0) HIP-event timing of the main search kernel of later
 * searches on this context (the events are recorded on the search's stream). */
int cpb_set_timing(cpb_context* ctx, int enable);

/* Milliseconds of the main search kernel of the last timed search (waits for it). */
int cpb_main_kernel_ms(cpb_context* ctx, double* ms);

/* Precision of later searches on `ctx`: 0 (default) FP64 phase fold and exact 64-bit
 * fixed-point bins; 1 FP64 phase fold, FP32 bins with compensated (TwoSum) atomic adds
 * and double-float prefix sums; 2 all FP32: phase fold, bins and prefix sums (the
 * reference's float32 CUDA kernel, src/cuperiod/methods/_bls_core.py:505-656 with
 * REAL=float). Modes 1 and 2 evaluate every box in FP32, serve the precision study, and
 * need the phase bins in LDS (CPB_E_RESOURCE otherwise). */
int cpb_set_accumulation(cpb_context* ctx, int mode);

/* Reference rounding of the near-maximum periods of later searches on `ctx` (1, the
 * default) or none (0). The engine's box sums are exact, so boxes holding the same points
 * tie exactly and the first one wins; the reference's FP64 kernel orders such ties (the
 * samples of a flat-topped peak, boxes shifted across empty bins) by the rounding of its
 * sequential bin and prefix sums (src/cuperiod/methods/_bls_core.py:369-442). With 1,
 * every period whose power lies within 2^-30 (relative) of its segment's largest power is
 * searched again with that FP64 sequence -- bins summed in point order, the wrap copy,
 * sequential prefix sums, every box's objective, the first strict maximum -- and its
 * seven outputs are the reference's. The band is taken per reference segment (the
 * segments of a curve with equal `group`), so a segment's outputs do not depend on the
 * other segments searched with it, and it holds every period within 2^-30 of the curve's
 * largest power. Applies to the exact (accumulation 0) searches; at most 64 periods per
 * reference segment are searched again. */
int cpb_set_peak_rounding(cpb_context* ctx, int enable);

/* Launch shape of later searches on `ctx` (chosen by the autotuner,
 * cuperiod_amd/batch/autotune.py): `threads` per workgroup (128, 256 or 512), `tile`
 * periods per work item (1..32; 0 = chosen from the batch size), and `grid_per_cu`
 * workgroups per compute unit of a persistent grid that loops over the work items
 * (0 = one workgroup per item). Results do not depend on the shape. */
int cpb_set_launch(cpb_context* ctx, int threads, int tile, int grid_per_cu);

/* A captured batched search: the arguments of cpb_bls_search_batch plus, when `summary`
 * is not NULL, those of cpb_bls_summarize, recorded once into a HIP graph with buffers
 * of its own (the context's accumulation and launch shape at capture time apply).
 * Every cpb_graph_launch replays the whole sequence on `stream`; inputs and outputs are
 * read and written at the captured addresses, which must stay valid. With timing on
 * (cpb_set_timing), cpb_main_kernel_ms afterwards reports the replay's duration.
 * Replaces the per-curve launch loop of cuPeriod's batch driver
 * (src/cuperiod/batch/runner.py:214-250) with one launch per chunk. */
typedef struct cpb_graph cpb_graph;
int cpb_graph_create(cpb_context* ctx, const double* tau, const double* yw, const double* wv,
                     int64_t total_points, const cpb_curve* curves, int32_t n_curves,
                     const double* periods, const int32_t* dur_bins,
                     const cpb_segment* segments, int32_t n_segments, int32_t oversample,
                     int32_t obj_flag, double* out, int64_t n_out, const int64_t* curve_base,
                     const int64_t* curve_count, double* summary, cpb_graph** graph);
int cpb_graph_launch(cpb_context* ctx, cpb_graph* graph, void* stream);
void cpb_graph_destroy(cpb_graph* graph);

/* Block until all work queued on `stream` has finished. */
int cpb_synchronize(cpb_context* ctx, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CUPERIOD_BLS_H */
