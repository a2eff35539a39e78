// FP64 CPU restatement of cuPeriod's numba joint-ephemeris multi-band box search --
// TEST INFRASTRUCTURE ONLY (the correctness oracle of the engine's joint search; never
// part of the product path).
//
// It restates `_numba_joint_kernel._kernel` (src/cuperiod/multiband/_bls_joint.py:
// 419-503) operation for operation:
//   * per period: n_bins = ceil(P/bd) + oversample, capped at width-1;
//   * per band k: fold w = x - P*floor(x/P), ind = int(w/bd) + 1 clipped to
//     [0, width-1]; wrap COPY of `oversample` bins (dst clipped to [0, width-1]);
//     inclusive prefix sums over every one of the `width` bins;
//   * per box (duration index, start n in [0, n_bins-d]): J = sum over the bands whose
//     box is admissible (iv_in, iv_out >= floor_k) and dips (yout >= yin) of
//     depth^2/(1/iv_in + 1/iv_out) (snr) or 0.5*iv_in*depth^2 (likelihood); a box counts
//     when at least one band dips; first strict maximum in (duration index, start) order;
//   * outputs per period: found, start, duration bins, and every band's in-box sums.
// Built with -ffp-contract=off (numba compiles with fastmath=False).

#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" int oracle_bls_joint(const double* tau, const double* yw, const double* wv,
                                const int64_t* starts, int32_t n_bands, const double* sum_y,
                                const double* sum_ivar, const double* floors,
                                const double* periods, int64_t n_periods,
                                const int64_t* dur_bins, int32_t n_dur, double bin_duration,
                                int32_t oversample, int32_t width, int32_t obj_flag,
                                int32_t n_threads, int64_t* o_n, int64_t* o_d,
                                uint8_t* o_found, double* o_yin, double* o_ivin) {
    if (n_bands < 1 || width < 2 || oversample < 1) return 1;
    const double neg = -std::numeric_limits<double>::infinity();
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
    {
        std::vector<double> my((size_t)n_bands * width), mi((size_t)n_bands * width);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t p = 0; p < n_periods; ++p) {
            const double period = periods[p];
            int64_t n_bins = (int64_t)std::ceil(period / bin_duration) + oversample;
            if (n_bins > width - 1) n_bins = width - 1;
            for (int k = 0; k < n_bands; ++k) {
                double* y = &my[(size_t)k * width];
                double* w = &mi[(size_t)k * width];
                for (int i = 0; i < width; ++i) {
                    y[i] = 0.0;
                    w[i] = 0.0;
                }
                for (int64_t j = starts[k]; j < starts[k + 1]; ++j) {
                    const double x = tau[j];
                    const double ph = x - period * std::floor(x / period);
                    int64_t ind = (int64_t)(ph / bin_duration) + 1;
                    if (ind < 0)
                        ind = 0;
                    else if (ind > width - 1)
                        ind = width - 1;
                    y[ind] += yw[j];
                    w[ind] += wv[j];
                }
                for (int o = 0; o < oversample; ++o) {
                    int64_t dst = n_bins - oversample + o;
                    if (dst < 0)
                        dst = 0;
                    else if (dst > width - 1)
                        dst = width - 1;
                    y[dst] = y[1 + o];
                    w[dst] = w[1 + o];
                }
                double ay = 0.0, ai = 0.0;
                for (int i = 0; i < width; ++i) {
                    ay += y[i];
                    y[i] = ay;
                    ai += w[i];
                    w[i] = ai;
                }
            }
            double best = neg;
            int64_t bn = 0, bd = 0;
            for (int di = 0; di < n_dur; ++di) {
                const int64_t d = dur_bins[di];
                if (d >= width || d >= n_bins) continue;
                for (int64_t n = 0; n < n_bins - d + 1; ++n) {
                    double total = 0.0;
                    bool hit = false;
                    for (int k = 0; k < n_bands; ++k) {
                        const double* y = &my[(size_t)k * width];
                        const double* w = &mi[(size_t)k * width];
                        const double y_in = y[n + d] - y[n];
                        const double iv_in = w[n + d] - w[n];
                        const double iv_out = sum_ivar[k] - iv_in;
                        if (iv_in < floors[k] || iv_out < floors[k]) continue;
                        const double yin = y_in / iv_in;
                        const double yout = (sum_y[k] - y_in) / iv_out;
                        if (!(yout >= yin)) continue;
                        const double depth = yout - yin;
                        hit = true;
                        if (obj_flag == 0)
                            total += depth * depth / (1.0 / iv_in + 1.0 / iv_out);
                        else
                            total += 0.5 * iv_in * depth * depth;
                    }
                    if (hit && total > best) {
                        best = total;
                        bn = n;
                        bd = d;
                    }
                }
            }
            o_found[p] = best > neg ? 1 : 0;
            o_n[p] = best > neg ? bn : 0;
            o_d[p] = best > neg ? bd : 0;
            for (int k = 0; k < n_bands; ++k) {
                const double* y = &my[(size_t)k * width];
                const double* w = &mi[(size_t)k * width];
                o_yin[(size_t)k * n_periods + p] = best > neg ? y[bn + bd] - y[bn] : 0.0;
                o_ivin[(size_t)k * n_periods + p] = best > neg ? w[bn + bd] - w[bn] : 0.0;
            }
        }
    }
    return 0;
}
