// MI355X (gfx950) joint-ephemeris multi-band box search: one box shared by every band
// of a star, each band fitting its own depth inside it (cuPeriod
// src/cuperiod/multiband/_bls_joint.py:1-26, kernels :419-503 and :531-685).
//
// One workgroup takes one (star, period segment, tile of periods) work item and, within
// it, one trial period after another:
//
//   (a) fold    every point of every band into that band's phase bins, against the
//               star's common time origin, with the reference's FP64 bin expression
//               (fold_bin: bit for bit the FP64 division's index) and 64-bit
//               fixed-point sums per band (ds_add_u64: exact, order-free);
//   (a') wrap   each band's copy of `oversample` bins (:450-457), in order;
//   (b) scan    per band, inclusive prefix sums over [0, n_bins] in rounds of 64 bins
//               (one wave per band, DPP scans carried across rounds);
//   (c) sweep   every box (duration index, start n) of the period: each band's in-box
//               sums are exact integer differences; a band adds
//               max(depth,0)^2/(1/W_in + 1/W_out) (snr) or W_in*depth^2/2 (likelihood)
//               when both of its sides clear its floor; the box counts when a band
//               dips; bands are summed in band order; the first strict maximum in
//               (duration index, start) order wins (:469-494);
//   (d) finish  each band's depth and SNR at the winning box in FP64, the quadrature
//               power, the strongest band and the box geometry (:359-406).
//
// The band prefix arrays of a period stay in LDS (K bands x width x 16 B) when they fit,
// else in a per-workgroup slice of global scratch (an L2-resident tile). The sweep reads
// 16-byte (y, w) pairs at consecutive starts across the lanes: conflict-free LDS reads.
// In the sweep the box value is formed from the integer sums as
// t^2 / (W_in W_out W_tot) * 2^(2 e_y - e_w) with t = Y_tot W_in - Y_in W_tot, which is
// the reference's depth^2/(1/W_in + 1/W_out) without its divisions and cancellations.

#include <climits>
#include <type_traits>

#include "engine_internal.h"

using namespace cpb_engine;

namespace {

constexpr int kJointMaxBands = 64;

// Per band of the star being searched, in LDS: the quantities the sweep reads for every
// box (fixed-point units of the band).
struct BandSweep {
    double y_tot;   // sum of fixed-point y*ivar
    double w_tot;   // sum of fixed-point ivar
    double floor_q; // ivar floor in fixed-point units
    int64_t wq_tot; // exact w_tot
    int32_t scale;  // 2*y_exp - w_exp
    int32_t pad;
};

__device__ __forceinline__ bool first_wins(double a, uint32_t ka, double b, uint32_t kb) {
    return a > b || (a == b && ka < kb);
}

// One DPP step of the wave's first-wins maximum: each lane compares its (value, key)
// with the pair CTRL brings in (rows outside ROW_MASK keep their own pair).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void dpp_first_wins(double& v, uint32_t& k) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
    const uint32_t olo = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
    const uint32_t ohi = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
    const uint32_t ok = (uint32_t)__builtin_amdgcn_update_dpp((int)k, (int)k, CTRL, ROW_MASK, 0xF, false);
    const double ov = __longlong_as_double((long long)(((uint64_t)ohi << 32) | olo));
    if (first_wins(ov, ok, v, k)) {
        v = ov;
        k = ok;
    }
}

// First-wins maximum (largest value, smallest key among equals) over the wave, returned
// uniformly: pairs within quads, half rows and rows by DPP permutations, then across
// rows by row broadcasts; lane 63 ends with the whole wave's pair. The rule is
// commutative and associative, so the order of the combinations does not matter.
__device__ __forceinline__ void wave_first_wins(double& v, uint32_t& k) {
    dpp_first_wins<0xB1, 0xF>(v, k);   // quad_perm [1,0,3,2]
    dpp_first_wins<0x4E, 0xF>(v, k);   // quad_perm [2,3,0,1]
    dpp_first_wins<0x141, 0xF>(v, k);  // row_half_mirror
    dpp_first_wins<0x140, 0xF>(v, k);  // row_mirror
    dpp_first_wins<0x142, 0xA>(v, k);  // row_bcast:15
    dpp_first_wins<0x143, 0xC>(v, k);  // row_bcast:31
    v = readlane_f64(v, kWave - 1);
    k = (uint32_t)__builtin_amdgcn_readlane((int)k, kWave - 1);
}

template <int T, int DMAX, int OBJ, bool GLOBAL>
__global__ __launch_bounds__(T) void bls_joint_kernel(
    const double* __restrict__ tau, const int64_t* __restrict__ yq,
    const int64_t* __restrict__ wq, const CurveParams* __restrict__ bp,
    const cpb_star* __restrict__ stars, const cpb_segment* __restrict__ segs,
    const WorkItem* __restrict__ items, int64_t n_items, const double* __restrict__ periods,
    const int32_t* __restrict__ dur_bins, int32_t oversample, int32_t wl, int64_t slice,
    int64_t* __restrict__ scratch, double* __restrict__ out, int64_t n_out) {
    extern __shared__ int64_t dyn_lds[];
    constexpr int NW = T / kWave;
    __shared__ BandSweep s_band[kJointMaxBands];
    __shared__ int64_t s_cum[kJointMaxBands + 1];
    __shared__ int32_t s_dur[kMaxDurations];
    __shared__ double s_red_v[NW];
    __shared__ uint32_t s_red_k[NW];
    __shared__ double s_snr[kJointMaxBands];
    __shared__ double s_depth[kJointMaxBands];

    const int tid = threadIdx.x;
    const int lane = tid % kWave;
    const int wave = tid / kWave;

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const WorkItem it = items[item];
        const cpb_segment sg = segs[it.seg];
        const cpb_star star = stars[sg.curve];
        const int K = star.n_bands;
        const int n_dur = sg.n_dur;
        const double bd = sg.bin_duration;
        const double inv_bd = 1.0 / bd;
        const int width = sg.width;
        int64_t* S;  // bins of band k: S[2*(k*wl + i)] = y, S[2*(k*wl + i) + 1] = w
        if constexpr (GLOBAL)
            S = scratch + (int64_t)blockIdx.x * slice;
        else
            S = dyn_lds;

        __syncthreads();
        if (tid < n_dur) s_dur[tid] = dur_bins[sg.dur_offset + tid];
        if (tid < K) {
            const CurveParams c = bp[star.band_offset + tid];
            BandSweep b;
            b.y_tot = (double)c.y_tot;
            b.w_tot = (double)c.w_tot;
            b.floor_q = ldexp(c.ivar_floor, -c.w_exp);
            b.wq_tot = c.w_tot;
            b.scale = 2 * c.y_exp - c.w_exp;
            b.pad = 0;
            s_band[tid] = b;
        }
        if (tid == 0) {
            int64_t acc = 0;
            for (int k = 0; k < K; ++k) {
                s_cum[k] = acc;
                acc += bp[star.band_offset + k].n_points;
            }
            s_cum[K] = acc;
        }
        __syncthreads();
        const int64_t n_total = s_cum[K];

        for (int q = 0; q < it.count; ++q) {
            const int64_t col = sg.out_offset + it.begin + q;
            const double P = segment_period(sg, periods, it.begin + q);
            const double invP = 1.0 / P;
            int64_t nb64 = (int64_t)ceil(P / bd) + oversample;
            if (nb64 > width - 1) nb64 = width - 1;
            const int nb = (int)nb64;

            // (a) zero bins [0, nb] of every band, then fold every point.
            {
                longlong2* S2 = reinterpret_cast<longlong2*>(S);
                const int span = nb + 1;
                for (int i = tid; i < K * span; i += T) {
                    const int k = i / span;
                    S2[k * wl + (i - k * span)] = make_longlong2(0, 0);
                }
            }
            if constexpr (GLOBAL) __threadfence_block();
            __syncthreads();
            {
                int k = 0;
                for (int64_t j = tid; j < n_total; j += T) {
                    while (j >= s_cum[k + 1]) ++k;
                    const int64_t p = bp[star.band_offset + k].point_offset + (j - s_cum[k]);
                    const int b = fold_bin(tau[p], P, invP, bd, inv_bd, width);
                    int64_t* slot = S + 2 * ((int64_t)k * wl + b);
                    atomicAdd(reinterpret_cast<unsigned long long*>(slot),
                              (unsigned long long)yq[p]);
                    atomicAdd(reinterpret_cast<unsigned long long*>(slot + 1),
                              (unsigned long long)wq[p]);
                }
            }
            if constexpr (GLOBAL) __threadfence_block();
            __syncthreads();
            // (a') wrap copy per band, in order (a copy may read a bin an earlier one wrote).
            for (int k = tid; k < K; k += T) {
                longlong2* B = reinterpret_cast<longlong2*>(S) + (int64_t)k * wl;
                for (int o = 0; o < oversample; ++o) {
                    int dst = nb - oversample + o;
                    dst = dst < 0 ? 0 : (dst > width - 1 ? width - 1 : dst);
                    B[dst] = B[1 + o];
                }
            }
            if constexpr (GLOBAL) __threadfence_block();
            __syncthreads();
            // (b) inclusive prefix over [0, nb], one wave per band, rounds of 64 bins.
            for (int k = wave; k < K; k += NW) {
                longlong2* B = reinterpret_cast<longlong2*>(S) + (int64_t)k * wl;
                int64_t cy = 0, cw = 0;
                for (int base = 0; base <= nb; base += kWave) {
                    const int i = base + lane;
                    longlong2 v = make_longlong2(0, 0);
                    if (i <= nb) v = B[i];
                    int64_t y = wave_incl_scan_dpp(v.x) + cy;
                    int64_t w = wave_incl_scan_dpp(v.y) + cw;
                    if (i <= nb) B[i] = make_longlong2(y, w);
                    cy = readlane_i64(y, kWave - 1);
                    cw = readlane_i64(w, kWave - 1);
                }
            }
            if constexpr (GLOBAL) __threadfence_block();
            __syncthreads();

            // (c) sweep: thread-strided starts, every duration per start.
            double best = -INFINITY;
            uint32_t bkey = 0xFFFFFFFFu;
            for (int n = tid; n < nb; n += T) {
                double tot[DMAX];
#pragma unroll
                for (int di = 0; di < DMAX; ++di) tot[di] = 0.0;
                uint32_t hit = 0;
                for (int k = 0; k < K; ++k) {
                    const BandSweep b = s_band[k];
                    const longlong2* B = reinterpret_cast<const longlong2*>(S) + (int64_t)k * wl;
                    const longlong2 p0 = B[n];
#pragma unroll
                    for (int di = 0; di < DMAX; ++di) {
                        if (di < n_dur) {
                            const int d = s_dur[di];
                            if (d < nb && n + d <= nb) {
                                const longlong2 p1 = B[n + d];
                                const int64_t yi = p1.x - p0.x;
                                const int64_t wi = p1.y - p0.y;
                                const double Wi = (double)wi;
                                const double Wo = (double)(b.wq_tot - wi);
                                if (Wi >= b.floor_q && Wo >= b.floor_q) {
                                    const double t = fma(b.y_tot, Wi, -((double)yi * b.w_tot));
                                    if (t >= 0.0) {
                                        hit |= 1u << di;
                                        const double den =
                                            OBJ == 0 ? Wi * Wo * b.w_tot : 2.0 * Wi * Wo * Wo;
                                        tot[di] += ldexp(t * t / den, b.scale);
                                    }
                                }
                            }
                        }
                    }
                }
#pragma unroll
                for (int di = 0; di < DMAX; ++di) {
                    if ((hit >> di) & 1u) {
                        const uint32_t key = ((uint32_t)di << 24) | (uint32_t)n;
                        if (first_wins(tot[di], key, best, bkey)) {
                            best = tot[di];
                            bkey = key;
                        }
                    }
                }
            }
            // first-wins maximum over the workgroup
            wave_first_wins(best, bkey);
            if (lane == 0) {
                s_red_v[wave] = best;
                s_red_k[wave] = bkey;
            }
            __syncthreads();
            best = s_red_v[0];
            bkey = s_red_k[0];
            for (int w = 1; w < NW; ++w)
                if (first_wins(s_red_v[w], s_red_k[w], best, bkey)) {
                    best = s_red_v[w];
                    bkey = s_red_k[w];
                }
            const bool found = best > -INFINITY;
            const int bn = found ? (int)(bkey & 0xFFFFFFu) : 0;
            const int bdur = found ? s_dur[bkey >> 24] : 0;

            // (d) each band's depth and SNR at the shared box, then the fields.
            if (found) {
                for (int k = tid; k < K; k += T) {
                    const CurveParams c = bp[star.band_offset + k];
                    const longlong2* B = reinterpret_cast<const longlong2*>(S) + (int64_t)k * wl;
                    const longlong2 p0 = B[bn], p1 = B[bn + bdur];
                    const int64_t yi = p1.x - p0.x, wi = p1.y - p0.y;
                    const double Y = ldexp((double)yi, c.y_exp);
                    const double W = ldexp((double)wi, c.w_exp);
                    const double Yo = ldexp((double)(c.y_tot - yi), c.y_exp);
                    const double Wo = ldexp((double)(c.w_tot - wi), c.w_exp);
                    double snr = 0.0, depth = 0.0;
                    if (W >= c.ivar_floor && Wo >= c.ivar_floor) {
                        const double yin = Y / W;
                        const double yout = Yo / Wo;
                        if (yout >= yin) {
                            depth = yout - yin;
                            snr = depth / sqrt(1.0 / W + 1.0 / Wo);
                        }
                    }
                    s_snr[k] = snr;
                    s_depth[k] = depth;
                }
            }
            __syncthreads();
            if (tid == 0) {
                double power = 0.0, depth = 0.0, dsnr = 0.0, dur = 0.0, t0 = 0.0, band = 0.0;
                if (found) {
                    int arg = 0;
                    for (int k = 0; k < K; ++k) {
                        power = power + s_snr[k] * s_snr[k];
                        if (s_snr[k] > s_snr[arg]) arg = k;
                    }
                    depth = s_depth[arg];
                    dsnr = sqrt(power);
                    dur = (double)bdur * bd;
                    t0 = fmod((double)bn * bd + 0.5 * dur, P) + star.t_min;
                    band = (double)arg;
                }
                out[0 * n_out + col] = power;
                out[1 * n_out + col] = depth;
                out[2 * n_out + col] = dsnr;
                out[3 * n_out + col] = dur;
                out[4 * n_out + col] = t0;
                out[5 * n_out + col] = band;
            }
            __syncthreads();
        }
    }
}

template <int T, int DMAX, int OBJ, bool GLOBAL>
hipError_t launch_joint(dim3 grid, size_t lds, hipStream_t st, const double* tau,
                        const int64_t* yq, const int64_t* wq, const CurveParams* bp,
                        const cpb_star* stars, const cpb_segment* segs, const WorkItem* items,
                        int64_t n_items, const double* periods, const int32_t* dur,
                        int32_t os, int32_t wl, int64_t slice, int64_t* scratch, double* out,
                        int64_t n_out) {
    auto fn = bls_joint_kernel<T, DMAX, OBJ, GLOBAL>;
    if (!GLOBAL && lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(fn, grid, dim3(T), GLOBAL ? 0 : lds, st, tau, yq, wq, bp, stars, segs,
                       items, n_items, periods, dur, os, wl, slice, scratch, out, n_out);
    return hipGetLastError();
}

template <int OBJ, bool GLOBAL>
hipError_t dispatch_joint(int dmax, dim3 grid, size_t lds, hipStream_t st, const double* tau,
                          const int64_t* yq, const int64_t* wq, const CurveParams* bp,
                          const cpb_star* stars, const cpb_segment* segs,
                          const WorkItem* items, int64_t n_items, const double* periods,
                          const int32_t* dur, int32_t os, int32_t wl, int64_t slice,
                          int64_t* scratch, double* out, int64_t n_out) {
    if (dmax <= 8)
        return launch_joint<256, 8, OBJ, GLOBAL>(grid, lds, st, tau, yq, wq, bp, stars, segs,
                                                 items, n_items, periods, dur, os, wl, slice, scratch,
                                                 out, n_out);
    if (dmax <= 16)
        return launch_joint<256, 16, OBJ, GLOBAL>(grid, lds, st, tau, yq, wq, bp, stars, segs,
                                                  items, n_items, periods, dur, os, wl,
                                                  slice, scratch, out, n_out);
    return launch_joint<256, 32, OBJ, GLOBAL>(grid, lds, st, tau, yq, wq, bp, stars, segs,
                                              items, n_items, periods, dur, os, wl, slice, scratch,
                                              out, n_out);
}

// ---------------------------------------------------------------------------------
// Band-group kernel (1,024 threads): the bands of a star pass through LDS a group at a
// time (as many as fit), while every thread keeps the running box totals of its starts
// (tid + r*1024, r < RMAX) and durations in registers; bands are still added in band
// order. After each group's exact int64 prefix, the sums are rounded once to FP64 in
// place, so a band's box sums are FP64 differences of prefix sums -- the reference's own
// arithmetic (`_bls_joint.py:474-476`) -- and a term costs one 16-byte LDS read, a few
// FP64 operations and a Newton-refined reciprocal. The winning box's per-band sums are
// read from the resident group and from re-folds of the others.

// Workgroup sizes of the band-group kernel. Up to 8 durations: 512 threads, for spans up
// to 1,024 bins with two starts per thread and band groups of at most a third of the LDS
// (three workgroups per CU), for spans up to 2,048 bins with four starts per thread and
// half of the LDS (two per CU). More durations: 256 threads over spans up to 512 bins
// (a third of the LDS), else 1,024 (one workgroup per CU, the whole LDS).
constexpr int kGroupThreads = 1024;
template <int V>
using IntC = std::integral_constant<int, V>;
constexpr int kGroupThreadsMid = 512;
constexpr int kGroupThreadsNarrow = 256;

struct BandTerm {
    double y_tot;    // sum of fixed-point y*ivar
    double w_tot;    // sum of fixed-point ivar
    double floor_q;  // ivar floor in fixed-point units
    double c;        // 2^(2 e_y - e_w) / w_tot (snr) or 2^(2 e_y - e_w) / 2 (likelihood)
};

// n / d for finite d > 0: hardware reciprocal, two Newton steps and a residual
// correction (within an ulp or two of the correctly rounded quotient).
__device__ __forceinline__ double div_newton(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    const double q = n * r;
    return fma(fma(-d, q, n), r, q);
}

// Zero, fold, wrap-copy and prefix the bins of bands g0 .. g0+gk-1 into LDS slots
// 0 .. gk-1 (FP64 prefix sums of the exact integer bins). Ends with a barrier.
template <int T>
__device__ __forceinline__ void fold_group(int64_t* S, int wl, int g0, int gk, int nb,
                                           int width, int oversample, double P, double invP,
                                           double bd, double inv_bd,
                                           const int64_t* __restrict__ s_cum,
                                           const int64_t* __restrict__ s_off,
                                           const double* __restrict__ tau,
                                           const int64_t* __restrict__ yq,
                                           const int64_t* __restrict__ wq) {
    constexpr int NW = T / kWave;
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    longlong2* S2 = reinterpret_cast<longlong2*>(S);
    const int span = nb + 1;
    for (int i = tid; i < gk * span; i += T) {
        const int kk = i / span;
        S2[kk * wl + (i - kk * span)] = make_longlong2(0, 0);
    }
    __syncthreads();
    {
        int k = g0;
        for (int64_t j = s_cum[g0] + tid; j < s_cum[g0 + gk]; j += T) {
            while (j >= s_cum[k + 1]) ++k;
            const int64_t p = s_off[k] + (j - s_cum[k]);
            const int b = fold_bin(tau[p], P, invP, bd, inv_bd, width);
            int64_t* slot = S + 2 * ((int64_t)(k - g0) * wl + b);
            atomicAdd(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)yq[p]);
            atomicAdd(reinterpret_cast<unsigned long long*>(slot + 1), (unsigned long long)wq[p]);
        }
    }
    __syncthreads();
    // Wrap copy (bins nb-os+o take bin 1+o): when the copied and the overwritten bins are
    // disjoint the prefix reads the source bin in place of each overwritten one;
    // otherwise the copies run in order first (a copy may read a bin an earlier one wrote).
    const int wrap_lo = nb - oversample;
    const bool fused_wrap = wrap_lo > oversample && oversample <= kWave;
    if (!fused_wrap) {
        for (int kk = tid; kk < gk; kk += T) {
            longlong2* B = S2 + (int64_t)kk * wl;
            for (int o = 0; o < oversample; ++o) {
                int dst = nb - oversample + o;
                dst = dst < 0 ? 0 : (dst > width - 1 ? width - 1 : dst);
                B[dst] = B[1 + o];
            }
        }
        __syncthreads();
    }
    // Chunked prefix: `wpb` waves per band, each lane sums a contiguous chunk of c bins,
    // one DPP scan per wave and one exchange between the band's waves give every chunk its
    // offset, and a second pass over the chunk writes the FP64 prefixes in place. The
    // fused wrap copy is made by the first pass: the lanes owning an overwritten bin read
    // its source (bins 1..os, which nobody writes in that pass) and store it in place, so
    // the second pass reads every bin as copied.
    // (Used by workgroups of 512 or more threads, which have waves to spare; with 256
    // threads one wave per band scanning rounds of 64 bins, and no exchange, measured
    // faster.)
    constexpr int CMAX = 16;
    const int wpb = gk <= NW ? NW / gk : 1;
    const int c = (span + wpb * kWave - 1) / (wpb * kWave);
    if (T >= 512 && gk <= NW && c <= CMAX) {
        __shared__ int64_t s_part[2 * NW];
        const int w_u = __builtin_amdgcn_readfirstlane(wave);
        const int kk = w_u / wpb, part = w_u - kk * wpb;
        const bool active = kk < gk;
        const int base = (part * kWave + lane) * c;
        longlong2* B = S2 + (int64_t)kk * wl;
        int64_t sy = 0, sw = 0;
        if (active) {
            for (int k = 0; k < c; ++k) {
                const int i = base + k;
                if (i <= nb) {
                    longlong2 v;
                    if (fused_wrap && i >= wrap_lo && i < nb) {
                        v = B[1 + (i - wrap_lo)];
                        B[i] = v;
                    } else {
                        v = B[i];
                    }
                    sy += v.x;
                    sw += v.y;
                }
            }
        }
        const int64_t iy = wave_incl_scan_dpp(sy);
        const int64_t iw = wave_incl_scan_dpp(sw);
        if (lane == kWave - 1) {
            s_part[2 * w_u] = iy;
            s_part[2 * w_u + 1] = iw;
        }
        __syncthreads();
        if (active) {
            int64_t oy = iy - sy, ow = iw - sw;
            for (int p = 0; p < part; ++p) {
                oy += s_part[2 * (kk * wpb + p)];
                ow += s_part[2 * (kk * wpb + p) + 1];
            }
            double2* D = reinterpret_cast<double2*>(B);
            for (int k = 0; k < c; ++k) {
                const int i = base + k;
                if (i <= nb) {
                    const longlong2 v = B[i];
                    oy += v.x;
                    ow += v.y;
                    D[i] = make_double2((double)oy, (double)ow);
                }
            }
        }
        __syncthreads();
        return;
    }
    for (int kk = wave; kk < gk; kk += NW) {
        longlong2* B = S2 + (int64_t)kk * wl;
        double2* D = reinterpret_cast<double2*>(B);
        int64_t cy = 0, cw = 0;
        // the copy's sources (bins 1..os), held in lanes 0..os-1 before the scan
        // overwrites them with prefix sums
        int64_t sy = 0, sw = 0;
        if (fused_wrap && lane < oversample) {
            const longlong2 v = B[1 + lane];
            sy = v.x;
            sw = v.y;
        }
        for (int base = 0; base <= nb; base += kWave) {
            const int i = base + lane;
            longlong2 v = make_longlong2(0, 0);
            if (i <= nb) v = B[i];
            if (fused_wrap && base + kWave > wrap_lo) {
                const int src = (i - wrap_lo) & (kWave - 1);
                const int64_t wy = shfl_i64(sy, src), ww = shfl_i64(sw, src);
                if (i >= wrap_lo && i < nb) v = make_longlong2(wy, ww);
            }
            const int64_t y = wave_incl_scan_dpp(v.x) + cy;
            const int64_t w = wave_incl_scan_dpp(v.y) + cw;
            if (i <= nb) D[i] = make_double2((double)y, (double)w);
            cy = readlane_i64(y, kWave - 1);
            cw = readlane_i64(w, kWave - 1);
        }
    }
    __syncthreads();
}

template <int T, int DMAX, int RMAX, int OBJ>
__global__ __launch_bounds__(T, T == 256 ? 3 : (T == 512 ? (RMAX == 4 ? 4 : 6) : 1)) void
bls_joint_group_kernel(
    const double* __restrict__ tau, const int64_t* __restrict__ yq,
    const int64_t* __restrict__ wq, const CurveParams* __restrict__ bp,
    const cpb_star* __restrict__ stars, const cpb_segment* __restrict__ segs,
    const WorkItem* __restrict__ items, int64_t n_items, const double* __restrict__ periods,
    const int32_t* __restrict__ dur_bins, int32_t oversample, int32_t wl, int32_t group,
    double* __restrict__ out, int64_t n_out) {
    extern __shared__ int64_t dyn_lds[];
    constexpr int NW = T / kWave;
    __shared__ BandTerm s_term[kJointMaxBands];
    __shared__ int64_t s_cum[kJointMaxBands + 1];
    __shared__ int64_t s_off[kJointMaxBands];
    __shared__ int32_t s_dur[kMaxDurations];
    __shared__ double s_red_v[NW];
    __shared__ uint32_t s_red_k[NW];
    __shared__ double s_snr[kJointMaxBands];
    __shared__ double s_depth[kJointMaxBands];

    const int tid = threadIdx.x;
    const int lane = tid % kWave;
    const int wave = tid / kWave;
    int64_t* S = dyn_lds;
    const double2* D2 = reinterpret_cast<const double2*>(dyn_lds);

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const WorkItem it = items[item];
        const cpb_segment sg = segs[it.seg];
        const cpb_star star = stars[sg.curve];
        const int K = star.n_bands;
        const int n_dur = sg.n_dur;
        const double bd = sg.bin_duration;
        const double inv_bd = 1.0 / bd;
        const int width = sg.width;

        __syncthreads();
        if (tid < n_dur) s_dur[tid] = dur_bins[sg.dur_offset + tid];
        if (tid < K) {
            const CurveParams c = bp[star.band_offset + tid];
            BandTerm b;
            b.y_tot = (double)c.y_tot;
            b.w_tot = (double)c.w_tot;
            b.floor_q = ldexp(c.ivar_floor, -c.w_exp);
            const int sc = 2 * c.y_exp - c.w_exp;
            b.c = OBJ == 0 ? ldexp(1.0 / b.w_tot, sc) : ldexp(0.5, sc);
            s_term[tid] = b;
            s_off[tid] = c.point_offset;
        }
        if (tid == 0) {
            int64_t acc = 0;
            for (int k = 0; k < K; ++k) {
                s_cum[k] = acc;
                acc += bp[star.band_offset + k].n_points;
            }
            s_cum[K] = acc;
        }
        __syncthreads();

        for (int q = 0; q < it.count; ++q) {
            const int64_t col = sg.out_offset + it.begin + q;
            const double P = segment_period(sg, periods, it.begin + q);
            const double invP = 1.0 / P;
            int64_t nb64 = (int64_t)ceil(P / bd) + oversample;
            if (nb64 > width - 1) nb64 = width - 1;
            const int nb = (int)nb64;

            double tot[RMAX][DMAX];
            uint32_t hit[RMAX];
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                hit[r] = 0;
#pragma unroll
                for (int di = 0; di < DMAX; ++di) tot[r][di] = 0.0;
            }
            int last_g0 = 0;
            for (int g0 = 0; g0 < K; g0 += group) {
                const int gk = min(group, K - g0);
                fold_group<T>(S, wl, g0, gk, nb, width, oversample, P, invP, bd, inv_bd,
                              s_cum, s_off, tau, yq, wq);
                for (int kk = 0; kk < gk; ++kk) {
                    const BandTerm bt = s_term[g0 + kk];
                    const double2* B = D2 + (int64_t)kk * wl;
#pragma unroll
                    for (int r = 0; r < RMAX; ++r) {
                        if (r * T >= nb) break;  // uniform: no start of this round is in use
                        const int n = tid + r * T;
                        if (n < nb) {
                            const double2 p0 = B[n];
#pragma unroll
                            for (int di = 0; di < DMAX; ++di) {
                                if (di < n_dur) {
                                    const int d = s_dur[di];
                                    if (d < nb && n + d <= nb) {
                                        const double2 p1 = B[n + d];
                                        const double wi = p1.y - p0.y;
                                        const double wo = bt.w_tot - wi;
                                        if (wi >= bt.floor_q && wo >= bt.floor_q) {
                                            const double yi = p1.x - p0.x;
                                            const double t = fma(bt.y_tot, wi, -(yi * bt.w_tot));
                                            if (t >= 0.0) {
                                                hit[r] |= 1u << di;
                                                const double den = OBJ == 0 ? wi * wo : wi * wo * wo;
                                                tot[r][di] += div_newton(t * t, den) * bt.c;
                                            }
                                        }
                                    }
                                }
                            }
                        }
                    }
                }
                last_g0 = g0;
                __syncthreads();
            }

            // first-wins maximum in (duration index, start) order
            double best = -INFINITY;
            uint32_t bkey = 0xFFFFFFFFu;
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                const int n = tid + r * T;
#pragma unroll
                for (int di = 0; di < DMAX; ++di) {
                    if ((hit[r] >> di) & 1u) {
                        const uint32_t key = ((uint32_t)di << 24) | (uint32_t)n;
                        if (first_wins(tot[r][di], key, best, bkey)) {
                            best = tot[r][di];
                            bkey = key;
                        }
                    }
                }
            }
            wave_first_wins(best, bkey);
            if (lane == 0) {
                s_red_v[wave] = best;
                s_red_k[wave] = bkey;
            }
            __syncthreads();
            best = s_red_v[0];
            bkey = s_red_k[0];
            for (int w = 1; w < NW; ++w)
                if (first_wins(s_red_v[w], s_red_k[w], best, bkey)) {
                    best = s_red_v[w];
                    bkey = s_red_k[w];
                }
            const bool found = best > -INFINITY;
            const int bn = found ? (int)(bkey & 0xFFFFFFu) : 0;
            const int bdur = found ? s_dur[bkey >> 24] : 0;

            // In workgroups that hold the whole LDS or half of it (1,024 threads; 512 with
            // four starts per thread), with more than one band group and the wrap copy's
            // sources apart from the bins it overwrites, the winner's prefix sums come from
            // one pass over the star's points instead of re-folding every other group (in
            // the third-of-LDS ones the LDS atomics of that pass, all lanes of a wave on
            // one band's four sums, cost more than the re-fold): with W = [nb - os, nb) the
            // overwritten bins, a point in bin b adds to the prefix at X once for
            // b <= X, b outside W, and once more for 1 <= b <= os with its copy's bin
            // nb - os + b - 1 <= X (exact int64 sums, rounded to FP64 as the folded
            // prefixes are, so the outputs are the same).
            constexpr bool kWide = T >= 1024 || (T == 512 && RMAX == 4);
            const bool direct = kWide && found && K > group && nb - oversample > oversample;
            if (direct) {
                __shared__ unsigned long long s_acc[4 * kJointMaxBands];
                for (int i = tid; i < 4 * K; i += T) s_acc[i] = 0ull;
                __syncthreads();
                const int wlo = nb - oversample, x0 = bn, x1 = bn + bdur;
                int k = 0;
                for (int64_t j = s_cum[0] + tid; j < s_cum[K]; j += T) {
                    while (j >= s_cum[k + 1]) ++k;
                    const int64_t p = s_off[k] + (j - s_cum[k]);
                    const int b = fold_bin(tau[p], P, invP, bd, inv_bd, width);
                    const bool kept = b < wlo || b >= nb;
                    const int mirror = (b >= 1 && b <= oversample) ? wlo + b - 1 : INT_MAX;
                    const int m0 = (kept && b <= x0) + (mirror <= x0);
                    const int m1 = (kept && b <= x1) + (mirror <= x1);
                    if (m1 != 0) {
                        const uint64_t y = (uint64_t)yq[p], w = (uint64_t)wq[p];
                        atomicAdd(&s_acc[4 * k + 2], (unsigned long long)(y * (uint64_t)m1));
                        atomicAdd(&s_acc[4 * k + 3], (unsigned long long)(w * (uint64_t)m1));
                        if (m0 != 0) {
                            atomicAdd(&s_acc[4 * k], (unsigned long long)(y * (uint64_t)m0));
                            atomicAdd(&s_acc[4 * k + 1], (unsigned long long)(w * (uint64_t)m0));
                        }
                    }
                }
                __syncthreads();
                for (int kk = tid; kk < K; kk += T) {
                    const CurveParams c = bp[star.band_offset + kk];
                    const double yi = (double)(int64_t)s_acc[4 * kk + 2] - (double)(int64_t)s_acc[4 * kk];
                    const double wi = (double)(int64_t)s_acc[4 * kk + 3] - (double)(int64_t)s_acc[4 * kk + 1];
                    const double Y = ldexp(yi, c.y_exp);
                    const double W = ldexp(wi, c.w_exp);
                    const double Yo = ldexp((double)c.y_tot - yi, c.y_exp);
                    const double Wo = ldexp((double)c.w_tot - wi, c.w_exp);
                    double snr = 0.0, depth = 0.0;
                    if (W >= c.ivar_floor && Wo >= c.ivar_floor) {
                        const double yin = Y / W;
                        const double yout = Yo / Wo;
                        if (yout >= yin) {
                            depth = yout - yin;
                            snr = depth / sqrt(1.0 / W + 1.0 / Wo);
                        }
                    }
                    s_snr[kk] = snr;
                    s_depth[kk] = depth;
                }
            } else if (found) {
                // the resident group first, then re-folds of the others
                for (int pass = 0; pass * group < K; ++pass) {
                    const int g0 = pass == 0 ? last_g0 : (pass - 1) * group;
                    if (pass > 0 && g0 == last_g0) continue;
                    const int gk = min(group, K - g0);
                    if (pass > 0)
                        fold_group<T>(S, wl, g0, gk, nb, width, oversample, P, invP, bd,
                                      inv_bd, s_cum, s_off, tau, yq, wq);
                    for (int kk = tid; kk < gk; kk += T) {
                        const CurveParams c = bp[star.band_offset + g0 + kk];
                        const double2* B = D2 + (int64_t)kk * wl;
                        const double2 p0 = B[bn], p1 = B[bn + bdur];
                        const double yi = p1.x - p0.x, wi = p1.y - p0.y;
                        const double Y = ldexp(yi, c.y_exp);
                        const double W = ldexp(wi, c.w_exp);
                        const double Yo = ldexp((double)c.y_tot - yi, c.y_exp);
                        const double Wo = ldexp((double)c.w_tot - wi, c.w_exp);
                        double snr = 0.0, depth = 0.0;
                        if (W >= c.ivar_floor && Wo >= c.ivar_floor) {
                            const double yin = Y / W;
                            const double yout = Yo / Wo;
                            if (yout >= yin) {
                                depth = yout - yin;
                                snr = depth / sqrt(1.0 / W + 1.0 / Wo);
                            }
                        }
                        s_snr[g0 + kk] = snr;
                        s_depth[g0 + kk] = depth;
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
            if (tid == 0) {
                double power = 0.0, depth = 0.0, dsnr = 0.0, dur = 0.0, t0 = 0.0, band = 0.0;
                if (found) {
                    int arg = 0;
                    for (int k = 0; k < K; ++k) {
                        power = power + s_snr[k] * s_snr[k];
                        if (s_snr[k] > s_snr[arg]) arg = k;
                    }
                    depth = s_depth[arg];
                    dsnr = sqrt(power);
                    dur = (double)bdur * bd;
                    t0 = fmod((double)bn * bd + 0.5 * dur, P) + star.t_min;
                    band = (double)arg;
                }
                out[0 * n_out + col] = power;
                out[1 * n_out + col] = depth;
                out[2 * n_out + col] = dsnr;
                out[3 * n_out + col] = dur;
                out[4 * n_out + col] = t0;
                out[5 * n_out + col] = band;
            }
            __syncthreads();
        }
    }
}

template <int T, int DMAX, int RMAX, int OBJ>
hipError_t launch_group(dim3 grid, size_t lds, hipStream_t st, const double* tau,
                        const int64_t* yq, const int64_t* wq, const CurveParams* bp,
                        const cpb_star* stars, const cpb_segment* segs, const WorkItem* items,
                        int64_t n_items, const double* periods, const int32_t* dur, int32_t os,
                        int32_t wl, int32_t group, double* out, int64_t n_out) {
    auto fn = bls_joint_group_kernel<T, DMAX, RMAX, OBJ>;
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(fn, grid, dim3(T), lds, st, tau, yq, wq, bp, stars, segs,
                       items, n_items, periods, dur, os, wl, group, out, n_out);
    return hipGetLastError();
}

// Starts per thread of a group-kernel variant (RMAX * DMAX <= 32): it covers spans of up to
// RMAX * threads + 1 bins.
__host__ __forceinline__ int group_rmax(int dmax) { return dmax <= 8 ? 4 : (dmax <= 16 ? 2 : 1); }

template <int T, int OBJ>
hipError_t dispatch_group(int dmax, int span, dim3 grid, size_t lds, hipStream_t st,
                          const double* tau, const int64_t* yq, const int64_t* wq,
                          const CurveParams* bp, const cpb_star* stars, const cpb_segment* segs,
                          const WorkItem* items, int64_t n_items, const double* periods,
                          const int32_t* dur, int32_t os, int32_t wl, int32_t group,
                          double* out, int64_t n_out) {
    auto go = [&](auto d, auto r) {
        return launch_group<T, decltype(d)::value, decltype(r)::value, OBJ>(
            grid, lds, st, tau, yq, wq, bp, stars, segs, items, n_items, periods, dur, os, wl,
            group, out, n_out);
    };
    // 512-thread workgroups (launched for at most 8 durations): two starts per thread over
    // spans of up to 2T bins, four up to 4T. 1,024-thread ones over spans of up to 2T with
    // at most 8 durations also take two: the smaller tile of running totals leaves
    // registers for the sweep's chains.
    if constexpr (T == kGroupThreadsMid) {
        return span <= 2 * T ? go(IntC<8>{}, IntC<2>{}) : go(IntC<8>{}, IntC<4>{});
    } else {
        if constexpr (T >= 1024)
            if (span <= 2 * T && dmax <= 8) return go(IntC<8>{}, IntC<2>{});
        if (dmax <= 8) return go(IntC<8>{}, IntC<4>{});
        if (dmax <= 16) return go(IntC<16>{}, IntC<2>{});
        return go(IntC<32>{}, IntC<1>{});
    }
}

}  // namespace

extern "C" int cpb_set_joint_kernel(cpb_context* ctx, int mode) {
    if (!ctx) return CPB_E_INVALID;
    if (mode < 0 || mode > 2) return fail(ctx, CPB_E_INVALID, "joint kernel mode must be 0, 1 or 2");
    ctx->joint_kernel = mode;
    return CPB_OK;
}

extern "C" int cpb_bls_joint_search(cpb_context* ctx, void* stream, const double* tau,
                                    const double* yw, const double* wv, int64_t total_points,
                                    const cpb_band* bands, int32_t n_bands,
                                    const cpb_star* stars, int32_t n_stars,
                                    const double* periods, const int32_t* dur_bins,
                                    const cpb_segment* segments, int32_t n_segments,
                                    int32_t oversample, int32_t obj_flag, double* out,
                                    int64_t n_out) {
    if (!ctx) return CPB_E_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (obj_flag != 0 && obj_flag != 1) return fail(ctx, CPB_E_INVALID, "obj_flag must be 0 or 1");
    if (oversample < 1) return fail(ctx, CPB_E_INVALID, "oversample must be >= 1");
    if (n_bands < 1 || n_stars < 1 || n_segments < 0 || total_points < 0 || n_out < 0)
        return fail(ctx, CPB_E_INVALID, "negative or empty size");
    if (!bands || !stars || (n_segments > 0 && !segments))
        return fail(ctx, CPB_E_INVALID, "null descriptor array");
    for (int32_t b = 0; b < n_bands; ++b) {
        const cpb_band& bn = bands[b];
        if (bn.point_offset < 0 || bn.n_points < 0 || bn.point_offset + bn.n_points > total_points)
            return fail(ctx, CPB_E_INVALID, "band points out of range");
        if (bn.n_points > (int64_t(1) << 30))
            return fail(ctx, CPB_E_RESOURCE, "more than 2^30 points in one band");
    }
    int kmax = 1;
    for (int32_t s = 0; s < n_stars; ++s) {
        const cpb_star& sr = stars[s];
        if (sr.n_bands < 1 || sr.band_offset < 0 || sr.band_offset + sr.n_bands > n_bands)
            return fail(ctx, CPB_E_INVALID, "star bands out of range");
        if (sr.n_bands > kJointMaxBands)
            return fail(ctx, CPB_E_RESOURCE, "more than 64 bands in one star");
        kmax = std::max(kmax, (int)sr.n_bands);
    }
    int dmax = 1, wmax = 2;
    for (int32_t s = 0; s < n_segments; ++s) {
        const cpb_segment& sg = segments[s];
        if (sg.curve < 0 || sg.curve >= n_stars)
            return fail(ctx, CPB_E_INVALID, "segment star index out of range");
        if (sg.n_periods < 0 || sg.out_offset < 0 || sg.out_offset + sg.n_periods > n_out)
            return fail(ctx, CPB_E_INVALID, "segment output range out of bounds");
        if (sg.n_dur < 0 || sg.n_dur > kMaxDurations)
            return fail(ctx, CPB_E_RESOURCE, "more than 32 distinct box widths in one segment");
        if (sg.width < 2 || sg.width < oversample + 2)
            return fail(ctx, CPB_E_INVALID, "segment width must be at least oversample + 2");
        if (!(sg.bin_duration > 0.0))
            return fail(ctx, CPB_E_INVALID, "bin_duration must be positive");
        if (sg.mode == 0 && sg.n_periods > 0 && !periods)
            return fail(ctx, CPB_E_INVALID, "explicit periods needed");
        if (sg.mode < 0 || sg.mode > 2) return fail(ctx, CPB_E_INVALID, "unknown segment mode");
        if (sg.mode != 0 && (sg.lin_num < 1 || sg.lin_k0 < 0 || sg.lin_k0 + sg.n_periods > sg.lin_num))
            return fail(ctx, CPB_E_INVALID, "bad parametric segment");
        if (sg.n_dur > 0 && sg.n_periods > 0) {
            dmax = std::max(dmax, (int)sg.n_dur);
            wmax = std::max(wmax, (int)sg.width);
        }
    }
    if (wmax > (1 << 24)) return fail(ctx, CPB_E_RESOURCE, "more than 2^24 phase bins per period");

    CPB_HIP(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure(ctx, ctx->curves, sizeof(CurveParams) * n_bands))) return rc;
    if ((rc = ensure(ctx, ctx->yq, sizeof(int64_t) * std::max<int64_t>(1, total_points)))) return rc;
    if ((rc = ensure(ctx, ctx->wq, sizeof(int64_t) * std::max<int64_t>(1, total_points)))) return rc;

    // Work items: tiles of consecutive periods of one segment (a segment without box
    // widths still gets items: its periods find no box and report zeros).
    int64_t live = 0;
    for (int32_t s = 0; s < n_segments; ++s) live += segments[s].n_periods;
    const int64_t target_items = (int64_t)ctx->n_cu * 64;
    int64_t tile = live / std::max<int64_t>(1, target_items);
    tile = std::max<int64_t>(1, std::min<int64_t>(tile, 32));
    // Span classes of the band-group kernel: 0 up to 1,024 bins, 1 wider.
    struct SpanClass {
        int64_t begin = 0, count = 0;
        int wmax = 2, dmax = 1;
    } cls[2];
    std::vector<WorkItem> by_class[2];
    for (int32_t s = 0; s < n_segments; ++s) {
        const cpb_segment& sg = segments[s];
        const int c = (sg.width + 63) / 64 * 64 <= 1024 ? 0 : 1;
        if (sg.n_periods > 0) {
            cls[c].wmax = std::max(cls[c].wmax, (int)sg.width);
            cls[c].dmax = std::max(cls[c].dmax, (int)std::max(sg.n_dur, 1));
        }
        for (int64_t b = 0; b < sg.n_periods; b += tile) {
            WorkItem w;
            w.seg = s;
            w.begin = b;
            w.count = (int32_t)std::min<int64_t>(tile, sg.n_periods - b);
            by_class[c].push_back(w);
        }
    }
    ctx->h_items.clear();
    for (int c = 0; c < 2; ++c) {
        cls[c].begin = (int64_t)ctx->h_items.size();
        cls[c].count = (int64_t)by_class[c].size();
        ctx->h_items.insert(ctx->h_items.end(), by_class[c].begin(), by_class[c].end());
    }
    const int64_t n_items = (int64_t)ctx->h_items.size();

    // Descriptor tables: [band curves | stars | segments | work items], 256-byte parts.
    std::vector<cpb_curve> band_curves(n_bands);
    for (int32_t b = 0; b < n_bands; ++b)
        band_curves[b] = {bands[b].point_offset, bands[b].n_points, 0.0, bands[b].ivar_floor};
    auto align = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t off_star = align(sizeof(cpb_curve) * n_bands);
    const size_t off_seg = off_star + align(sizeof(cpb_star) * n_stars);
    const size_t off_item = off_seg + align(sizeof(cpb_segment) * std::max(1, n_segments));
    const size_t tbytes = off_item + align(sizeof(WorkItem) * std::max<int64_t>(1, n_items));
    if ((rc = ensure(ctx, ctx->tables, tbytes))) return rc;
    if (ctx->copy_pending) {
        CPB_HIP(ctx, hipEventSynchronize(ctx->copied));
        ctx->copy_pending = false;
    }
    if (ctx->h_tables_bytes < tbytes) {
        if (ctx->h_tables) CPB_HIP(ctx, hipHostFree(ctx->h_tables));
        ctx->h_tables = nullptr;
        ctx->h_tables_bytes = 0;
        CPB_HIP(ctx, hipHostMalloc(&ctx->h_tables, tbytes + tbytes / 4, hipHostMallocDefault));
        ctx->h_tables_bytes = tbytes + tbytes / 4;
    }
    char* hb = static_cast<char*>(ctx->h_tables);
    std::memcpy(hb, band_curves.data(), sizeof(cpb_curve) * n_bands);
    std::memcpy(hb + off_star, stars, sizeof(cpb_star) * n_stars);
    if (n_segments > 0) std::memcpy(hb + off_seg, segments, sizeof(cpb_segment) * n_segments);
    if (n_items > 0) std::memcpy(hb + off_item, ctx->h_items.data(), sizeof(WorkItem) * n_items);
    CPB_HIP(ctx, hipMemcpyAsync(ctx->tables.ptr, hb, tbytes, hipMemcpyHostToDevice, st));
    if (!ctx->capturing) {
        CPB_HIP(ctx, hipEventRecord(ctx->copied, st));
        ctx->copy_pending = true;
    }
    char* db = static_cast<char*>(ctx->tables.ptr);
    const auto* curve_in = reinterpret_cast<const cpb_curve*>(db);
    const auto* sp = reinterpret_cast<const cpb_star*>(db + off_star);
    const auto* gp = reinterpret_cast<const cpb_segment*>(db + off_seg);
    const auto* ip = reinterpret_cast<const WorkItem*>(db + off_item);
    auto* params = static_cast<CurveParams*>(ctx->curves.ptr);
    auto* yq = static_cast<int64_t*>(ctx->yq.ptr);
    auto* wq = static_cast<int64_t*>(ctx->wq.ptr);

    hipError_t e = launch_prep_curves(st, nullptr, yw, wv, curve_in, n_bands, obj_flag, params, yq, wq);
    if (e != hipSuccess) return hip_fail(ctx, e, "prep_curves_kernel launch");
    if (n_items == 0) return CPB_OK;

    const int wl = (wmax + 63) / 64 * 64;
    const size_t static_lds = 8192;
    const bool timed = ctx->timing && !ctx->capturing;
    // Band-group kernel when, in every span class, its register tiles cover every start
    // and one band fits in LDS; otherwise every band of a period at once, in LDS or
    // global scratch.
    int threads_of[2] = {0, 0};
    bool group_fits = true;
    for (int c = 0; c < 2; ++c) {
        if (cls[c].count == 0) continue;
        const size_t bb = (size_t)16 * ((cls[c].wmax + 63) / 64 * 64);
        const int span = cls[c].wmax - 1;
        const int rmax = group_rmax(cls[c].dmax);
        if (cls[c].dmax <= 8 && span <= (c == 0 ? 2 : 4) * kGroupThreadsMid)
            threads_of[c] = kGroupThreadsMid;
        else if (c == 0 && span <= rmax * kGroupThreadsNarrow)
            threads_of[c] = kGroupThreadsNarrow;
        else if (span <= rmax * kGroupThreads)
            threads_of[c] = kGroupThreads;
        if (threads_of[c] == 0 || bb + static_lds > (size_t)ctx->lds_limit) group_fits = false;
    }
    const int mode = ctx->joint_kernel;
    if (mode == 2 && !group_fits)
        return fail(ctx, CPB_E_RESOURCE, "the band-group joint kernel does not cover this span");
    if (mode != 1 && group_fits) {
        if (timed) CPB_HIP(ctx, hipEventRecord(ctx->t0, st));
        for (int c = 0; c < 2; ++c) {
            if (cls[c].count == 0) continue;
            const int T = threads_of[c];
            const int wlc = (cls[c].wmax + 63) / 64 * 64;
            const size_t bb = (size_t)16 * wlc;
            // Band groups of a third of the LDS in the narrow class (three workgroups per
            // CU), of half of it for 512-thread workgroups in the wide class (two per CU),
            // of all of it for 1,024-thread ones.
            const int share = T == kGroupThreads ? 1 : (c == 0 ? 3 : 2);
            const size_t budget = (size_t)ctx->lds_limit / share - static_lds;
            const int gmax = std::max<int>(1, (int)(budget / bb));
            const int n_groups = (kmax + gmax - 1) / gmax;
            const int group = (kmax + n_groups - 1) / n_groups;
            const size_t lds = bb * group;
            const dim3 grid((unsigned)std::min<int64_t>(cls[c].count, ((int64_t)1 << 31) / T));
            const WorkItem* items = ip + cls[c].begin;
            auto run = [&](auto threads, auto obj) {
                return dispatch_group<decltype(threads)::value, decltype(obj)::value>(
                    cls[c].dmax, cls[c].wmax - 1, grid, lds, st, tau, yq, wq, params, sp, gp,
                    items, cls[c].count, periods, dur_bins, oversample, wlc, group, out, n_out);
            };
            auto run_obj = [&](auto threads) {
                return obj_flag == 0 ? run(threads, IntC<0>{})
                                     : run(threads, IntC<1>{});
            };
            if (T == kGroupThreadsNarrow)
                e = run_obj(IntC<kGroupThreadsNarrow>{});
            else if (T == kGroupThreadsMid)
                e = run_obj(IntC<kGroupThreadsMid>{});
            else
                e = run_obj(IntC<kGroupThreads>{});
            if (e != hipSuccess) return hip_fail(ctx, e, "bls_joint_group_kernel launch");
        }
        if (timed) {
            CPB_HIP(ctx, hipEventRecord(ctx->t1, st));
            ctx->timed = true;
        }
        return CPB_OK;
    }
    const size_t lds_need = (size_t)16 * kmax * wl;
    const bool global = lds_need + static_lds > (size_t)ctx->lds_limit;
    const int64_t slice = 2 * (int64_t)kmax * wl;  // int64 words of one workgroup's bins
    int64_t* scratch = nullptr;
    dim3 grid;
    if (global) {
        const int64_t g = std::min<int64_t>(n_items, (int64_t)ctx->n_cu * 4);
        if ((rc = ensure(ctx, ctx->wide, (size_t)g * lds_need))) return rc;
        scratch = static_cast<int64_t*>(ctx->wide.ptr);
        grid = dim3((unsigned)g);
    } else {
        grid = dim3((unsigned)std::min<int64_t>(n_items, ((int64_t)1 << 31) / 256));
    }
    if (timed) CPB_HIP(ctx, hipEventRecord(ctx->t0, st));
    if (obj_flag == 0)
        e = global ? dispatch_joint<0, true>(dmax, grid, lds_need, st, tau, yq, wq, params, sp, gp,
                                             ip, n_items, periods, dur_bins, oversample, wl, slice,
                                             scratch, out, n_out)
                   : dispatch_joint<0, false>(dmax, grid, lds_need, st, tau, yq, wq, params, sp,
                                              gp, ip, n_items, periods, dur_bins, oversample, wl, slice,
                                              scratch, out, n_out);
    else
        e = global ? dispatch_joint<1, true>(dmax, grid, lds_need, st, tau, yq, wq, params, sp, gp,
                                             ip, n_items, periods, dur_bins, oversample, wl, slice,
                                             scratch, out, n_out)
                   : dispatch_joint<1, false>(dmax, grid, lds_need, st, tau, yq, wq, params, sp,
                                              gp, ip, n_items, periods, dur_bins, oversample, wl, slice,
                                              scratch, out, n_out);
    if (e != hipSuccess) return hip_fail(ctx, e, "bls_joint_kernel launch");
    if (timed) {
        CPB_HIP(ctx, hipEventRecord(ctx->t1, st));
        ctx->timed = true;
    }
    return CPB_OK;
}
