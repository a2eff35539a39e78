// FP64 CPU restatement of cuPeriod's numba BLS kernel -- TEST INFRASTRUCTURE ONLY.
//
// This file is the correctness oracle and the timed CPU baseline. It is never part of
// the product path: only tests, the smoke check and the benchmark's CPU-baseline leg
// load it, as the checker.
//
// It restates `_numba_bls_kernel._kernel` (src/cuperiod/methods/_bls_core.py:341-462)
// operation for operation, so that its results are the numba backend's bits:
//   * per period: n_bins = ceil(P/bd) + oversample, capped at width-1 (:358-360);
//   * fold w = x - P*floor(x/P), ind = int(w/bd) + 1 clipped to [0, width-1] (:369-378);
//   * wrap COPY of `oversample` bins, bin[n_bins-os+k] = bin[1+k] (:379-383);
//   * inclusive prefix sums over [0, n_bins] (:384-390);
//   * the division-free screen and the exact second pass, first-wins on
//     (duration index, start) (:398-442);
//   * the FP64 finalize of the seven outputs (:443-461).
// Built with -ffp-contract=off: numba compiles with fastmath=False, so LLVM forms no
// fused multiply-adds, and neither may this file.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Python's float `%` (numba follows CPython's float_rem): the result takes the sign
// of the divisor. Used for the transit time, `(ln*bd + 0.5*dur) % period` (:460).
inline double py_mod(double a, double b) {
    double r = std::fmod(a, b);
    if (r != 0.0) {
        if ((r < 0.0) != (b < 0.0)) r += b;
    } else {
        r = std::copysign(0.0, b);
    }
    return r;
}

}  // namespace

extern "C" {

// Output layout: out[k * n_periods + p] for k in
// (power, depth, depth_err, depth_snr, duration, transit_time, log_likelihood),
// the order of `_OUTPUT_NAMES` (src/cuperiod/methods/_bls_core.py:676-684).
int oracle_bls_search(const double* tau, const double* yw, const double* wv,
                      int64_t n_points, const double* periods, int64_t n_periods,
                      const int64_t* dur_bins, int64_t n_dur, double bin_duration,
                      int64_t oversample, int64_t width, double sum_y,
                      double sum_ivar, double t_min, int obj_flag, double ivar_floor,
                      double inv_a, double inv_s, double slack, int n_threads,
                      double* out) {
    if (n_periods < 0 || n_points < 0 || width < 2) return 1;
    const double neg = -std::numeric_limits<double>::infinity();
    const double ybar = sum_ivar > 0.0 ? sum_y / sum_ivar : 0.0;
    double* o_power = out;
    double* o_depth = out + n_periods;
    double* o_depth_err = out + 2 * n_periods;
    double* o_depth_snr = out + 3 * n_periods;
    double* o_duration = out + 4 * n_periods;
    double* o_tt = out + 5 * n_periods;
    double* o_ll = out + 6 * n_periods;
    for (int64_t p = 0; p < n_periods; ++p) {
        o_power[p] = neg;
        o_depth[p] = 0.0;
        o_depth_err[p] = 0.0;
        o_depth_snr[p] = 0.0;
        o_duration[p] = 0.0;
        o_tt[p] = 0.0;
        o_ll[p] = 0.0;
    }
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif

#pragma omp parallel
    {
        std::vector<double> my(width), mi(width), num(width), den(width);
#pragma omp for schedule(dynamic, 16)
        for (int64_t pidx = 0; pidx < n_periods; ++pidx) {
            const double period = periods[pidx];
            int64_t n_bins = (int64_t)std::ceil(period / bin_duration) + oversample;
            if (n_bins > width - 1) n_bins = width - 1;
            for (int64_t i = 0; i < n_bins + 1; ++i) {
                my[i] = 0.0;
                mi[i] = 0.0;
            }
            for (int64_t j = 0; j < n_points; ++j) {
                const double x = tau[j];
                const double w = x - period * std::floor(x / period);
                int64_t ind = (int64_t)(w / bin_duration) + 1;
                if (ind < 0)
                    ind = 0;
                else if (ind > width - 1)
                    ind = width - 1;
                my[ind] += yw[j];
                mi[ind] += wv[j];
            }
            for (int64_t k = 0; k < oversample; ++k) {
                const int64_t dst = n_bins - oversample + k;
                if (0 <= dst && dst < width) {
                    my[dst] = my[1 + k];
                    mi[dst] = mi[1 + k];
                }
            }
            double ay = 0.0, ai = 0.0;
            for (int64_t i = 0; i < n_bins + 1; ++i) {
                ay += my[i];
                my[i] = ay;
                ai += mi[i];
                mi[i] = ai;
            }
            double loc = neg;
            int64_t ln = 0, ld = 0;
            double best_t2 = 0.0, best_den = 1.0;
            for (int64_t di = 0; di < n_dur; ++di) {
                const int64_t d = dur_bins[di];
                if (d < 1 || d >= n_bins) continue;
                const int64_t m = n_bins - d + 1;
                // Pass 1: division-free optimistic bound on every box (-1 = never).
                for (int64_t n = 0; n < m; ++n) {
                    const double y_in = my[n + d] - my[n];
                    const double iv_in = mi[n + d] - mi[n];
                    const double iv_out = sum_ivar - iv_in;
                    const double t = (y_in - ybar * iv_in) * inv_a;
                    const double t_hi = std::fabs(t) + slack;
                    double dn = (iv_in * inv_s) * (iv_out * inv_s);
                    if (obj_flag != 0) dn = dn * (iv_out * inv_s);
                    const bool ok =
                        (iv_in >= ivar_floor) & (iv_out >= ivar_floor) & (t <= slack);
                    num[n] = ok ? t_hi * t_hi : -1.0;
                    den[n] = ok ? dn : 1.0;
                }
                // Pass 2: the original box evaluation, where the bound can win.
                for (int64_t n = 0; n < m; ++n) {
                    if (num[n] * best_den < best_t2 * den[n]) continue;
                    const double y_in = my[n + d] - my[n];
                    const double iv_in = mi[n + d] - mi[n];
                    const double iv_out = sum_ivar - iv_in;
                    if (iv_in < ivar_floor || iv_out < ivar_floor) continue;
                    const double yin = y_in / iv_in;
                    const double yout = (sum_y - y_in) / iv_out;
                    if (yout < yin) continue;
                    const double depth = yout - yin;
                    double obj;
                    if (obj_flag == 0)
                        obj = depth / std::sqrt(1.0 / iv_in + 1.0 / iv_out);
                    else
                        obj = 0.5 * iv_in * depth * depth;
                    if (obj > loc) {
                        loc = obj;
                        ln = n;
                        ld = d;
                        double t_lo = std::fabs((y_in - ybar * iv_in) * inv_a) - slack;
                        if (!(t_lo > 0.0)) t_lo = 0.0;
                        best_t2 = t_lo * t_lo;
                        best_den = den[n];
                    }
                }
            }
            if (loc <= neg) continue;
            const double y_in = my[ln + ld] - my[ln];
            const double iv_in = mi[ln + ld] - mi[ln];
            const double iv_out = sum_ivar - iv_in;
            const double yin = y_in / iv_in;
            const double yout = (sum_y - y_in) / iv_out;
            const double depth = yout - yin;
            const double derr = std::sqrt(1.0 / iv_in + 1.0 / iv_out);
            const double dsnr = depth / derr;
            const double ll = 0.5 * iv_in * depth * depth;
            const double dur = (double)ld * bin_duration;
            o_power[pidx] = obj_flag == 0 ? dsnr : ll;
            o_depth[pidx] = depth;
            o_depth_err[pidx] = derr;
            o_depth_snr[pidx] = dsnr;
            o_duration[pidx] = dur;
            o_tt[pidx] = py_mod((double)ln * bin_duration + 0.5 * dur, period) + t_min;
            o_ll[pidx] = ll;
        }
    }
    return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
