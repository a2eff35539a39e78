// FP64 CPU restatement of cuPeriod's numba TLS matched filter -- TEST INFRASTRUCTURE
// ONLY (the correctness oracle of the engine's TLS search; never part of the product).
//
// It restates `_numba_tls_kernel._kernel` (src/cuperiod/methods/tls.py:513-595)
// operation for operation:
//   * per period: q = tau/P, ph = q - floor(q), bin = int(ph*n_bins) clipped to
//     [0, n_bins-1]; a[bin] += y'w, b[bin] += w, in point order;
//   * one circular pad of max_width-1 bins;
//   * per template width (duration index) and start s: num = sum_k g_k a[s+k],
//     cwg = sum_k g_k b[s+k], den = sum_k g_k^2 b[s+k], accumulated k = 0, 1, ...;
//     a dip (num < 0, den > w_eps) whose level-corrected energy
//     D = den - cwg^2/W exceeds den_rel*den scores SR = num^2/D; first strict maximum
//     in (duration index, start) order;
//   * outputs: SR, depth -num/D, duration width/n_bins*P, t0 = frac((s + width/2)/n_bins)*P.
// Built with -ffp-contract=off (numba compiles with fastmath=False).

#include <cmath>
#include <cstdint>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" int oracle_tls(const double* tau, const double* yw, const double* wv,
                          int64_t n_points, const double* periods, int64_t n_periods,
                          const int64_t* dur_bins, const int64_t* tmpl_off, const double* tmpl,
                          int32_t n_dur, int32_t n_bins, double w_eps, double w_total,
                          double den_rel, int32_t n_threads, double* o_sr, double* o_depth,
                          double* o_duration, double* o_t0) {
    if (n_bins < 1 || n_dur < 0) return 1;
    int64_t max_width = 1;
    for (int di = 0; di < n_dur; ++di)
        if (dur_bins[di] > max_width) max_width = dur_bins[di];
    const int64_t n_ext = n_bins + max_width - 1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
    {
        std::vector<double> a(n_ext), b(n_ext), num(n_bins), cwg(n_bins), den(n_bins);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t p = 0; p < n_periods; ++p) {
            const double period = periods[p];
            for (int64_t i = 0; i < n_ext; ++i) {
                a[i] = 0.0;
                b[i] = 0.0;
            }
            for (int64_t j = 0; j < n_points; ++j) {
                const double q = tau[j] / period;
                const double ph = q - std::floor(q);
                int64_t bb = (int64_t)(ph * n_bins);
                if (bb > n_bins - 1)
                    bb = n_bins - 1;
                else if (bb < 0)
                    bb = 0;
                a[bb] += yw[j];
                b[bb] += wv[j];
            }
            for (int64_t j = 0; j < max_width - 1; ++j) {
                a[n_bins + j] = a[j];
                b[n_bins + j] = b[j];
            }
            double loc_sr = 0.0, loc_depth = 0.0;
            int64_t loc_start = 0, loc_width = 0;
            for (int di = 0; di < n_dur; ++di) {
                const int64_t width = dur_bins[di];
                const int64_t off = tmpl_off[di];
                const double g0 = tmpl[off];
                const double g00 = g0 * g0;
                for (int s = 0; s < n_bins; ++s) {
                    num[s] = g0 * a[s];
                    cwg[s] = g0 * b[s];
                    den[s] = g00 * b[s];
                }
                for (int64_t k = 1; k < width; ++k) {
                    const double gk = tmpl[off + k];
                    const double gkk = gk * gk;
                    for (int s = 0; s < n_bins; ++s) {
                        num[s] += gk * a[k + s];
                        cwg[s] += gk * b[k + s];
                        den[s] += gkk * b[k + s];
                    }
                }
                for (int s = 0; s < n_bins; ++s) {
                    const double nm = num[s], dn = den[s];
                    if (nm < 0.0 && dn > w_eps) {
                        const double de = dn - cwg[s] * cwg[s] / w_total;
                        if (de > den_rel * dn) {
                            const double sr = nm * nm / de;
                            if (sr > loc_sr) {
                                loc_sr = sr;
                                loc_start = s;
                                loc_width = width;
                                loc_depth = -nm / de;
                            }
                        }
                    }
                }
            }
            o_sr[p] = loc_sr;
            o_depth[p] = loc_depth;
            o_duration[p] = (double)loc_width / (double)n_bins * period;
            const double centre = ((double)loc_start + (double)loc_width / 2.0) / (double)n_bins;
            o_t0[p] = (centre - std::floor(centre)) * period;
        }
    }
    return 0;
}
