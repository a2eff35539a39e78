// MI355X (gfx950, CDNA4) BLS engine: HIP kernels and the C ABI of include/cuperiod_bls.h.
//
// One workgroup processes one (light curve, period segment, frequency tile) work item
// at a time and, within it, one trial period after another:
//
//   (a) fold    every point's phase bin from the reference's FP64 expression
//               w = x - P*floor(x/P), ind = int(w/bd) + 1 (cuPeriod
//               src/cuperiod/methods/_bls_core.py:369-378), evaluated through
//               reciprocal multiplies with an exact-division fallback whenever the
//               quotient lies within a few ulps of an integer, so every bin index is
//               the one the FP64 division gives; the point's weighted flux and weight
//               are added to LDS bins as 64-bit fixed-point integers (ds_add_u64):
//               exact, so the bins do not depend on the atomics' order;
//   (a') wrap   the reference's wrap COPY of `oversample` bins (:379-383);
//   (b) scan    wave-level inclusive prefix sums (cross-lane shuffles, one barrier
//               between waves), stored split into 32-bit high and low words;
//               then the box sweep over (duration, start): an FP32 screen on the high
//               words finds the best box value per (thread, duration), and only boxes
//               the screen cannot rule out are re-evaluated exactly in FP64 from the
//               full 64-bit sums, keeping the first strict maximum in (duration index,
//               start) order (:398-442, :603-625);
//   (c) final   the seven outputs in FP64 from the winning box (:443-461, :627-654),
//               written to HBM.
//
// Compiled with -ffp-contract=off: the fold must round every product and difference
// exactly as the FP64 reference does.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "cuperiod_bls.h"
#include "engine_internal.h"

using namespace cpb_engine;

namespace {
// ---------------------------------------------------------------------------------
// Per-curve preparation: fixed-point conversion and the screen's bounds.

constexpr int kPrepThreads = 256;

__device__ __forceinline__ double block_sum_f64(double v, double* red) {
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    v = wave_sum_f64(v);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < kPrepThreads / kWave; ++w) s += red[w];
    return s;
}

__device__ __forceinline__ double block_max_f64(double v, double* red) {
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    v = wave_max_f64(v);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < kPrepThreads / kWave; ++w) s = fmax(s, red[w]);
    return s;
}

__device__ __forceinline__ int64_t block_sum_i64(int64_t v, int64_t* red) {
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    v = wave_sum_i64(v);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    int64_t s = 0;
    for (int w = 0; w < kPrepThreads / kWave; ++w) s += red[w];
    return s;
}

__global__ __launch_bounds__(kPrepThreads) void prep_curves_kernel(
    const double* __restrict__ tau, const double* __restrict__ yw,
    const double* __restrict__ wv,
    const cpb_curve* __restrict__ curves, int32_t obj_flag,
    CurveParams* __restrict__ params, int64_t* __restrict__ yq,
    int64_t* __restrict__ wq) {
    __shared__ double red_d[kPrepThreads / kWave];
    __shared__ int64_t red_i[kPrepThreads / kWave];
    const cpb_curve c = curves[blockIdx.x];
    const int64_t off = c.point_offset;
    const int64_t n = c.n_points;
    const int tid = threadIdx.x;

    double sa = 0.0, sw = 0.0;
    for (int64_t i = tid; i < n; i += kPrepThreads) {
        sa += fabs(yw[off + i]);
        sw += wv[off + i];
    }
    sa = block_sum_f64(sa, red_d);
    sw = block_sum_f64(sw, red_d);
    const bool finite = isfinite(sa) && isfinite(sw);
    int ey = 0, ew = 0;
    if (finite && sa > 0.0) (void)frexp(sa, &ey);
    if (finite && sw > 0.0) (void)frexp(sw, &ew);

    int64_t ty = 0, tw = 0;
    for (int64_t i = tid; i < n; i += kPrepThreads) {
        int64_t a = 0, b = 0;
        if (finite) {
            // ldexp scales exactly, whatever the data's exponent range.
            a = llrint(ldexp(yw[off + i], kFixBits - ey));
            b = llrint(ldexp(wv[off + i], kFixBits - ew));
        }
        yq[off + i] = a;
        wq[off + i] = b;
        ty += a;
        tw += b;
    }
    ty = block_sum_i64(ty, red_i);
    tw = block_sum_i64(tw, red_i);
    const double ybar = tw > 0 ? (double)ty / (double)tw : 0.0;

    double cmax = 0.0;
    for (int64_t i = tid; i < n; i += kPrepThreads) {
        const int64_t b = wq[off + i];
        if (b > 0) cmax = fmax(cmax, fabs((double)yq[off + i] / (double)b - ybar));
    }
    cmax = block_max_f64(cmax, red_d);
    double xmax = 0.0;
    if (tau != nullptr)
        for (int64_t i = tid; i < n; i += kPrepThreads) xmax = fmax(xmax, tau[off + i]);
    xmax = tau != nullptr ? block_max_f64(xmax, red_d) : INFINITY;

    if (tid == 0) {
        CurveParams p;
        p.point_offset = off;
        p.y_tot = ty;
        p.w_tot = tw;
        p.y_exp = ey - kFixBits;
        p.w_exp = ew - kFixBits;
        p.t_min = c.t_min;
        p.ivar_floor = c.ivar_floor;
        p.ybar = (float)ybar;
        p.n_points = (int32_t)n;
        const int64_t s_hi = tw >> 32;
        p.s_hi = (int32_t)s_hi;
        // s_hi >= 2^26 - 1 (the weights map to [2^58, 2^59) units), so e_unc >= 511.
        const int64_t e = s_hi >> 17 > 1 ? s_hi >> 17 : 1;
        p.e_unc = (int32_t)e;
        // Bound on |t_screen - t| in high-word units, t = Y - ybar*W. The screen reads
        // t as the difference of two encoded prefixes U = Yh - mean_units(Wh) (Yh, Wh
        // the high words): the truncated high words (1 unit of Y, 1 of W times |ybar|)
        // and the two roundings of mean_units (1/2 + 2^-23 each) make 2 + |ybar| +
        // 2^-22; the conversion of the integer difference to float (|t| 2^-24) is left
        // to the margin.
        const double s = (double)s_hi, ed = (double)e;
        const double terr = 2.0 + fabs(ybar) + 0x1p-22;
        // A box the snr screen leaves uncertified (W*Wout <= dmin) has a side of at most
        // e + 1 high-word units, so at most e + 3 units of weight: the bounds below
        // cover e + 3.
        double vdeg, vbnd;
        if (obj_flag == 0) {
            vdeg = 65536.0 * terr * terr / (ed * (s - ed));
            vbnd = 2.0 * cmax * cmax * (ed + 3.0) / (s - ed - 3.0);
        } else {
            vdeg = 65536.0 * terr * terr / (ed * ed * (s - ed));
            vbnd = 2.0 * cmax * cmax * (ed + 3.0) / ((s - ed - 3.0) * (s - ed - 3.0));
        }
        p.vdeg = (float)fmin(vdeg, 3.0e38);
        p.vbnd = (float)fmin(vbnd, 3.0e38);
        // W*Wout > e*s*(1 + 2^-20) (in float, each factor within 2^-24 of its integer)
        // leaves no side at or below e.
        p.dmin = (float)(ed * s * (1.0 + 0x1p-20));
        p.ybar_d = ybar;
        p.x_max = xmax;
        int flags = 0;
        if (!finite) flags |= 2;
        if (e < 511) flags |= 1;  // too little weight for the screen's bounds: exact sweep
        // |ybar| * 2^29 < 1/2: mean_units is 0 for every high word
        if (fabs(ybar) < 0x1p-31) flags |= 4;
        p.flags = flags;
        params[blockIdx.x] = p;
    }
}

// ---------------------------------------------------------------------------------
// Main kernel.
//
// LDS (or, in MODE 2, global scratch) of one workgroup. During the fold, the
// fixed-point sums (y, w) of phase bin i:
//   MODE 0 (span <= T*kChunk, LDS, 16*wl bytes): y in Y[i] from byte 0, w in W[i] from
//          byte 8*wl. After the prefix, H[i] = (U, cw_hi) over Y and L[i] = (cy_lo,
//          cw_lo) over W;
//   MODE 1 (LDS) and MODE 2 (global scratch): 16-byte slots, slot_of(i) = i + i/8 holds
//          (y, w) (the pad slot after every 8 bins puts the 8 bins a thread scans on
//          distinct banks), and after the prefix (U, cw_hi, cy_lo, cw_lo) in place.
// cy/cw are the inclusive prefix sums, hi/lo their 32-bit words, and U = cy_hi -
// mean_units(cw_hi, ybar) the encoded y high word.

constexpr int kChunk = 8;
// Periods per work item (the item's period table lives in LDS).
constexpr int kMaxTile = 32;

__device__ __forceinline__ int slot_of(int i) { return i + (i >> 3); }

template <int MODE>
struct BinView {
    // MODE 0: dense H/L; MODE 1/2: slots.
    const char* base;
    int wl;
    __device__ __forceinline__ int2 hi(int n) const {
        if constexpr (MODE == 0)
            return reinterpret_cast<const int2*>(base)[n];
        else
            return *reinterpret_cast<const int2*>(base + 16 * slot_of(n));
    }
    __device__ __forceinline__ uint2 lo(int n) const {
        if constexpr (MODE == 0)
            return reinterpret_cast<const uint2*>(base + 8 * (size_t)wl)[n];
        else
            return *reinterpret_cast<const uint2*>(base + 16 * slot_of(n) + 8);
    }
};

// Exact ranking key of box (n, d) from the 64-bit prefix sums, in the curve's fixed-point
// units: with Yi, Wi (Yo, Wo) the in-box (out-of-box) sums and num = Yo*Wi - Yi*Wo =
// depth*Wi*Wo, the objective is a fixed positive multiple of sqrt(num^2/den) with
// den = Wi*Wo*(Wi+Wo) (SNR) or Wi*Wo^2 (log-likelihood), so boxes compare by
// num_a^2*den_b vs num_b^2*den_a without a division. Boxes with equal sums get equal
// keys, so exact ties stay ties for the first-wins rule. Returns false for an
// inadmissible box: a side below the floor, or not a dip (num < 0, i.e. yout < yin).
struct BoxKey {
    double num2;
    double den;
    int64_t yin;
    int64_t win;
};

template <int OBJ, int MODE>
__device__ __forceinline__ bool exact_key(const BinView<MODE>& bv, int n, int d,
                                          const CurveParams& cv, BoxKey& k) {
    const int2 h0 = bv.hi(n), h1 = bv.hi(n + d);
    const uint2 l0 = bv.lo(n), l1 = bv.lo(n + d);
    // the y high words from their encoded form (see the prefix stores)
    int32_t yh0 = h0.x, yh1 = h1.x;
    if (!(cv.flags & 4)) {
        yh0 += mean_units(h0.y, cv.ybar_d);
        yh1 += mean_units(h1.y, cv.ybar_d);
    }
    const int64_t yin_q = join64(yh1, l1.x) - join64(yh0, l0.x);
    const int64_t win_q = join64(h1.y, l1.y) - join64(h0.y, l0.y);
    const int64_t wout_q = cv.w_tot - win_q;
    const double Wi = (double)win_q, Wo = (double)wout_q;
    if (ldexp(Wi, cv.w_exp) < cv.ivar_floor || ldexp(Wo, cv.w_exp) < cv.ivar_floor)
        return false;
    const double Yi = (double)yin_q, Yo = (double)(cv.y_tot - yin_q);
    const double num = fma(Yo, Wi, -(Yi * Wo));
    if (num < 0.0) return false;
    k.num2 = num * num;
    // converted at the use (an empty asm hides the value from loop-invariant code
    // motion), not held in registers across the whole work item
    int64_t wt = cv.w_tot;
    asm volatile("" : "+s"(wt));
    k.den = OBJ == 0 ? Wi * Wo * (double)wt : Wi * Wo * Wo;
    k.yin = yin_q;
    k.win = win_q;
    return true;
}

__device__ __forceinline__ bool key_better(double an, double ad, uint32_t ak, double bn,
                                           double bd, uint32_t bk) {
    const double l = an * bd, r = bn * ad;
    return l > r || (l == r && ak < bk);
}

// First-wins best box of a wave, uniform: the lanes holding a candidate are few, so
// they are visited one by one.
__device__ __forceinline__ void wave_best_ballot(BoxKey& k, uint32_t& key) {
    unsigned long long m = __ballot(k.num2 >= 0.0);
    BoxKey w{-1.0, 1.0, 0, 0};
    uint32_t wk = 0xFFFFFFFFu;
    while (m) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const double n2 = readlane_f64(k.num2, l);
        const double dn = readlane_f64(k.den, l);
        const uint32_t kk = (uint32_t)__builtin_amdgcn_readlane((int)key, l);
        if (key_better(n2, dn, kk, w.num2, w.den, wk)) {
            w.num2 = n2;
            w.den = dn;
            w.yin = readlane_i64(k.yin, l);
            w.win = readlane_i64(k.win, l);
            wk = kk;
        }
    }
    k = w;
    key = wk;
}

// FP32 screen value of box (n, d): -t|t| / den with t = Y - ybar*W (a dip has t <= 0),
// den = W*Wout (snr) or W*Wout^2 (likelihood), from the prefixes' high words: their first
// word is encoded as U = Yh - mean_units(Wh), so the difference of two is t within
// 2 + |ybar| units, and the second is Wh. W is within 1 unit and Wout within 2, so den
// is within a few parts in e_unc of its value when both sides hold more than `e_unc`
// units; for snr W*Wout > dmin says so, for the likelihood the box must hold more than
// e_unc and the rest more than 4*e_unc (Wout enters squared). `cert` reports it. A
// likelihood box with little weight outside can still score high: `unc` marks those
// (the rest at most 4*e_unc units) for exact evaluation.
template <int OBJ>
__device__ __forceinline__ float screen_value(int2 h0, int2 h1, int s_hi, int e_unc,
                                              float dmin, bool& cert, bool& unc) {
    const int dw = h1.y - h0.y;
    const int dwo = s_hi - dw;
    const float t = (float)(h1.x - h0.x);
    const float W = (float)dw;
    const float Wo = (float)dwo;
    float den;
    if constexpr (OBJ == 0) {
        den = W * Wo;
        cert = den > dmin;
        unc = false;
    } else {
        den = W * Wo * Wo;
        cert = (dw > e_unc) & (dwo > 4 * e_unc);
        unc = (dw > e_unc) & (dwo <= 4 * e_unc) & (dwo > -2);
    }
    return (-t) * fabsf(t) * __builtin_amdgcn_rcpf(den);
}

// KMODE 3 is MODE 0 for spans of up to 1,152 bins: its 19 KiB of LDS admit 8 waves per
// SIMD, and the register budget is capped for them; KMODE 4 is MODE 0 capped at 7 waves
// per SIMD (spans of up to 1,344 bins: 22 KiB, seven workgroups per CU); KMODE 5 is MODE 0
// for spans of up to 640 bins in 128-thread workgroups (11 KiB: fourteen of them, two
// waves each, per CU, the 7 waves per SIMD its register cap allows; the per-wave work of
// a period, its scans, maxima and winner, is then shared by two waves instead of four);
// MODE 0 is capped at 6 waves per SIMD.
template <int T, int DMAX, int OBJ, int KMODE>
__global__ __launch_bounds__(T, KMODE == 3 ? 8
                                : KMODE == 4 || KMODE == 5 ? 7
                                : KMODE == 0             ? 6
                                                         : 1) void
bls_main_kernel(
    const double* __restrict__ tau, const int64_t* __restrict__ yq,
    const int64_t* __restrict__ wq, const CurveParams* __restrict__ curves,
    const cpb_segment* __restrict__ segs, const WorkItem* __restrict__ items,
    int64_t n_items, const double* __restrict__ periods,
    const int32_t* __restrict__ dur_bins, int32_t oversample, int32_t wl,
    int64_t* __restrict__ wide_scratch, double* __restrict__ out, int64_t n_out) {
    extern __shared__ int64_t dyn_lds[];
    constexpr int MODE = (KMODE == 3 || KMODE == 4 || KMODE == 5) ? 0 : KMODE;
    // bins per thread of the MODE 0 prefix: spans of up to 640 bins (KMODE 5) need at
    // most the odd ceil(640/T), spans of up to 1,152 bins (KMODE 3) at most 1,024/T + 1
    // (CHUNK + 1 registers each), the wider ones kChunk
    constexpr int CHUNK = KMODE == 5   ? (((640 + T - 1) / T) | 1) - 1
                          : KMODE == 3 ? (1024 / T < kChunk ? 1024 / T : kChunk)
                                       : kChunk;
    constexpr int NW = T / kWave;
    constexpr bool GLOBAL = MODE == 2;
    __shared__ int32_t s_dur[DMAX];
    // Per period of the work item: P, 1/P and the bin count, read at the start of the
    // period; once it is done, the same slots hold its winner (in-box sums, key) until
    // the item's outputs are written. No LDS is added for the winners, so each launch
    // class keeps its workgroups per CU.
    __shared__ double s_per[kMaxTile];
    __shared__ double s_invp[kMaxTile];
    __shared__ int32_t s_nb[kMaxTile];
    __shared__ int64_t s_red_i[2 * NW];
    __shared__ int s_red_v[NW];
    __shared__ BoxKey s_red_b[NW];
    __shared__ uint32_t s_red_k[NW];

    const int tid = threadIdx.x;
    const int lane = tid % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);

    const int n_slots = slot_of(wl - 1) + 1;
    int64_t* S;  // slots: S[2*s] = y, S[2*s + 1] = w
    if constexpr (GLOBAL)
        S = wide_scratch + (int64_t)blockIdx.x * 2 * (int64_t)n_slots;
    else
        S = dyn_lds;
    const BinView<MODE> bv{reinterpret_cast<const char*>(S), wl};

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const WorkItem it = items[item];
        const cpb_segment sg = segs[it.seg];
        const CurveParams cv = curves[sg.curve];
        const int n_dur = sg.n_dur;
        const double bd = sg.bin_duration;
        const double inv_bd = 1.0 / bd;
        const int width = sg.width;
        const int64_t poff = cv.point_offset;
        const int npts = cv.n_points;
        __syncthreads();
        if (tid < n_dur) s_dur[tid] = dur_bins[sg.dur_offset + tid];
        // The item's periods, their reciprocals and bin counts, one lane per period.
        // lane-dependent values are formed where they are used (the empty asm keeps them
        // from being hoisted and held, or spilled, across the work item)
        int ltid = tid;
        asm volatile("" : "+v"(ltid));
        if (ltid >= kWave && ltid < kWave + it.count) {
            const int q = ltid - kWave;
            const double P = segment_period(sg, periods, it.begin + q);
            int64_t nb = (int64_t)ceil(P / bd) + oversample;
            if (nb > width - 1) nb = width - 1;
            s_per[q] = P;
            s_invp[q] = 1.0 / P;
            s_nb[q] = (int32_t)nb;
        }
        __syncthreads();

        for (int q = 0; q < it.count; ++q) {
            const int64_t k = it.begin + q;
            const double P = s_per[q];
            const double invP = s_invp[q];
            const int n_bins = s_nb[q];
            const int64_t col = sg.out_offset + k;

            if (cv.flags & 2) {
                if (tid == 0) {
                    out[col] = -INFINITY;
                    for (int o = 1; o < CPB_N_OUTPUTS; ++o) out[o * n_out + col] = 0.0;
                }
                continue;
            }

            // (a) zero the bins [0, n_bins], then fold every point.
            const int span = n_bins + 1;
            if constexpr (MODE == 0) {
                // pairs of bins; 2*ceil(span/2) <= wl, as wl is a multiple of 64
                longlong2* Y2 = reinterpret_cast<longlong2*>(S);
                longlong2* W2 = reinterpret_cast<longlong2*>(S + wl);
                for (int i = tid; i < (span + 1) / 2; i += T) {
                    Y2[i] = make_longlong2(0, 0);
                    W2[i] = make_longlong2(0, 0);
                }
            } else {
                longlong2* S2 = reinterpret_cast<longlong2*>(S);
                const int last = slot_of(n_bins);
                for (int i = tid; i <= last; i += T) S2[i] = make_longlong2(0, 0);
            }
            if constexpr (GLOBAL) __threadfence();
            __syncthreads();
            {
                // this period's near-integer tolerances (fold_bin_tol): x/P <= x_max/P
                // and w/bd <= width, up to roundings well inside the factor 2 of slack
                const double tol_q = cv.x_max * invP * 0x1p-48;
                const double tol_r = (double)width * 0x1p-48;
                // point loads at 32-bit byte offsets from the curve's (uniform) bases
                const char* __restrict__ tc = reinterpret_cast<const char*>(tau + poff);
                const char* __restrict__ yc = reinterpret_cast<const char*>(yq + poff);
                const char* __restrict__ wc = reinterpret_cast<const char*>(wq + poff);
                auto point = [&](int j, double& x, int64_t& a, int64_t& b) {
                    const uint32_t off = (uint32_t)j * 8u;
                    x = *reinterpret_cast<const double*>(tc + off);
                    a = *reinterpret_cast<const int64_t*>(yc + off);
                    b = *reinterpret_cast<const int64_t*>(wc + off);
                };
                // bin i's y and w accumulators
                unsigned long long* const acc = reinterpret_cast<unsigned long long*>(S);
                auto add_point = [&](int bin, int64_t a, int64_t b) {
                    if constexpr (MODE == 0) {
                        atomicAdd(acc + bin, (unsigned long long)a);
                        atomicAdd(acc + wl + bin, (unsigned long long)b);
                    } else {
                        const int sl = slot_of(bin);
                        atomicAdd(acc + 2 * sl, (unsigned long long)a);
                        atomicAdd(acc + 2 * sl + 1, (unsigned long long)b);
                    }
                };
                // two points per thread per pass, both loaded before either is folded
                auto fold_range = [&](int j0, int step, int end) {
                    for (; j0 + step < end; j0 += 2 * step) {
                        double x0, x1;
                        int64_t a0, b0, a1, b1;
                        point(j0, x0, a0, b0);
                        point(j0 + step, x1, a1, b1);
                        const int i0 = fold_bin_tol(x0, P, invP, bd, inv_bd, width, tol_q, tol_r);
                        const int i1 = fold_bin_tol(x1, P, invP, bd, inv_bd, width, tol_q, tol_r);
                        add_point(i0, a0, b0);
                        add_point(i1, a1, b1);
                    }
                    if (j0 < end) {
                        double x0;
                        int64_t a0, b0;
                        point(j0, x0, a0, b0);
                        add_point(fold_bin_tol(x0, P, invP, bd, inv_bd, width, tol_q, tol_r), a0, b0);
                    }
                };
                // the lane's first point is formed per period (the empty asm keeps it
                // from being hoisted and held in a register across the work item)
                int j0 = tid;
                asm volatile("" : "+v"(j0));
                // Consecutive points are consecutive lanes. When they advance the phase by
                // less than a bin on average (periods long against the point spacing, as
                // in the chunked coarse search of the learned proposer), a wave's lanes
                // fall into the same few bins and their LDS atomics serialise; then pairs
                // of lanes walk T/2 contiguous blocks of the curve instead (each pair's
                // loads are still 16 contiguous bytes; groups of 8 and 4 lanes were 20%
                // and 4% slower on the coarse search, single lanes 62%). (Only in the
                // 128-thread class, where the coarse search runs: in the wider classes the
                // test and the second loop cost more than they saved.)
                bool spread = false;
                if constexpr (KMODE == 5)
                    spread = (double)cv.x_max * (double)n_bins < (double)npts * P;
                if (spread) {
                    constexpr int GL = 2, NG = T / GL;
                    const int blk = ((npts + NG - 1) / NG + GL - 1) / GL * GL;
                    const int g0 = (j0 / GL) * blk;
                    fold_range(g0 + j0 % GL, GL, min(g0 + blk, npts));
                } else {
                    fold_range(j0, T, npts);
                }
            }
            if constexpr (GLOBAL) __threadfence();
            __syncthreads();
            // (a') wrap copy of `oversample` bins, bin[n_bins-os+k] = bin[1+k]. When the
            // copied and the overwritten bins are disjoint, MODE 0's prefix reads the
            // source bin in place of each overwritten one; otherwise the copies run in
            // order here (a copy may read a bin an earlier one wrote).
            const int wrap_lo = n_bins - oversample;
            const bool fused_wrap = MODE == 0 && wrap_lo > oversample;
            if (!fused_wrap) {
                if (tid == 0) {
                    longlong2* S2 = reinterpret_cast<longlong2*>(S);
                    for (int kk = 0; kk < oversample; ++kk) {
                        const int dst = n_bins - oversample + kk;
                        if (dst >= 0 && dst < width) {
                            if constexpr (MODE == 0) {
                                S[dst] = S[1 + kk];
                                S[wl + dst] = S[wl + 1 + kk];
                            } else {
                                S2[slot_of(dst)] = S2[slot_of(1 + kk)];
                            }
                        }
                    }
                }
                if constexpr (GLOBAL) __threadfence();
                __syncthreads();
            }

            // (b1) inclusive prefix over [0, n_bins]: every thread sums a contiguous
            // chunk of c bins, one DPP scan per wave and one exchange between waves give
            // each chunk its offset. MODE 0 makes c odd: the lanes' 8-byte stores, c
            // entries apart, then fall on distinct LDS banks. The y high word is stored
            // as U = Yh - mean_units(Wh, ybar): the difference of two such words is the
            // box's Y - ybar*W to within 2 + |ybar| units.
            {
                const longlong2* S2 = reinterpret_cast<const longlong2*>(S);
                const bool plain = (cv.flags & 4) != 0;
                int64_t sy = 0, sw = 0;
                int64_t ry[CHUNK + 1], rw[CHUNK + 1];
                const int c = MODE == 0 ? ((span + T - 1) / T) | 1 : (span + T - 1) / T;
                const int base = tid * c;
                // bins [wave_lo, wave_hi) are this wave's (uniform)
                const int wave_lo = __builtin_amdgcn_readfirstlane(wave) * kWave * c;
                const int wave_hi = wave_lo + kWave * c;
                if constexpr (MODE == 0) {
                    if (wave_hi <= (fused_wrap ? wrap_lo : span)) {
                        // every bin of the wave exists and is read in place: loads at
                        // fixed offsets from the lane's first bin, no per-bin tests
                        const int64_t* Sy = S + base;
                        const int64_t* Sw = S + wl + base;
#pragma unroll
                        for (int kk = 0; kk <= CHUNK; ++kk) {
                            if (kk < c) {
                                sy += Sy[kk];
                                sw += Sw[kk];
                            }
                            ry[kk] = sy;
                            rw[kk] = sw;
                        }
                    } else {
                        // the overwritten bins of a fused wrap copy read their source
                        // bins (only the one or two lanes whose chunk holds them)
                        const bool wrap_here = fused_wrap && base + c > wrap_lo && base < n_bins;
#pragma unroll
                        for (int kk = 0; kk <= CHUNK; ++kk) {
                            const int i = base + kk;
                            if (kk < c && i < span) {
                                const int src = (wrap_here && i >= wrap_lo && i < n_bins)
                                                    ? 1 + (i - wrap_lo)
                                                    : i;
                                sy += S[src];
                                sw += S[wl + src];
                            }
                            ry[kk] = sy;
                            rw[kk] = sw;
                        }
                    }
                } else {
                    for (int kk = 0; kk < c; ++kk) {
                        const int i = base + kk;
                        if (i < span) {
                            const longlong2 v = S2[slot_of(i)];
                            sy += v.x;
                            sw += v.y;
                        }
                    }
                }
                const int64_t iy = wave_incl_scan_dpp(sy);
                const int64_t iw = wave_incl_scan_dpp(sw);
                if (lane == kWave - 1) {
                    s_red_i[2 * wave] = iy;
                    s_red_i[2 * wave + 1] = iw;
                }
                if constexpr (GLOBAL) __threadfence();
                __syncthreads();
                int64_t oy = iy - sy, ow = iw - sw;
                for (int w = 0; w < wave; ++w) {
                    oy += s_red_i[2 * w];
                    ow += s_red_i[2 * w + 1];
                }
                if constexpr (MODE == 0) {
                    int2* Hw = reinterpret_cast<int2*>(S);
                    uint2* Lw = reinterpret_cast<uint2*>(reinterpret_cast<char*>(S) + 8 * (size_t)wl);
                    // mean_units is 0 for every high word of a `plain` curve (|ybar| <
                    // 2^-31, high words < 2^27), so it needs no test here
                    const double ybar = cv.ybar_d;
                    // the lane's chunk as 32-bit words: word pairs at fixed offsets
                    int32_t* Hb = reinterpret_cast<int32_t*>(Hw + base);
                    uint32_t* Lb = reinterpret_cast<uint32_t*>(Lw + base);
                    auto store = [&](int kk) {
                        const uint64_t vy = (uint64_t)(ry[kk] + oy);
                        const uint64_t vw = (uint64_t)(rw[kk] + ow);
                        const int32_t wh = (int32_t)(vw >> 32);
                        Hb[2 * kk] = (int32_t)(vy >> 32) - mean_units(wh, ybar);
                        Hb[2 * kk + 1] = wh;
                        Lb[2 * kk] = (uint32_t)vy;
                        Lb[2 * kk + 1] = (uint32_t)vw;
                    };
                    if (wave_hi <= span) {
#pragma unroll
                        for (int kk = 0; kk <= CHUNK; ++kk)
                            if (kk < c) store(kk);
                    } else {
#pragma unroll
                        for (int kk = 0; kk <= CHUNK; ++kk)
                            if (kk < c && base + kk < span) store(kk);
                    }
                } else {
                    int4* S4 = reinterpret_cast<int4*>(S);
                    for (int kk = 0; kk < c; ++kk) {
                        const int i = base + kk;
                        if (i < span) {
                            const int s = slot_of(i);
                            const longlong2 v = reinterpret_cast<const longlong2*>(S4)[s];
                            oy += v.x;
                            ow += v.y;
                            const int32_t wh = (int32_t)((uint64_t)ow >> 32);
                            int32_t u = (int32_t)((uint64_t)oy >> 32);
                            if (!plain) u -= mean_units(wh, cv.ybar_d);
                            S4[s] = make_int4(u, wh, (int32_t)(uint32_t)oy, (int32_t)(uint32_t)ow);
                        }
                    }
                }
            }
            if constexpr (GLOBAL) __threadfence();
            __syncthreads();

            // (b2) FP32 screen: best value (as non-negative float bits) per (thread,
            // duration); boxes a screen value cannot rank are flagged for (b3).
            const int s_hi = cv.s_hi;
            const int e_unc = cv.e_unc;
            const float dmin = cv.dmin;
            int cm[DMAX];
            int dd[DMAX];
            uint32_t unc_bits = 0;
            uint32_t dok = 0;  // durations with 1 <= d < n_bins (uniform)
            int dmax_ok = 0;   // the longest duration in use (uniform)
#pragma unroll
            for (int di = 0; di < DMAX; ++di) {
                cm[di] = 0;
                dd[di] = di < n_dur ? s_dur[di] : 0;
                if (di < n_dur && dd[di] >= 1 && dd[di] < n_bins) {
                    dok |= 1u << di;
                    dmax_ok = max(dmax_ok, dd[di]);
                }
            }
            auto screen_box = [&](int2 h0, int2 h1, int di) {
                bool cert, unc;
                const float v = screen_value<OBJ>(h0, h1, s_hi, e_unc, dmin, cert, unc);
                cm[di] = max(cm[di], cert ? __float_as_int(v) : 0);
                if (OBJ == 1 && unc) unc_bits |= 1u << di;
            };
            // Start-outer, duration-inner: one read of the prefix at n per start, and the
            // durations' boxes are independent work for the scheduler. Rounds of starts
            // whose every box ends within the bins skip the per-box bound check.
            int n_first = 0;  // first start of the checked rounds
            if (MODE == 0 && n_dur == DMAX && dok == (DMAX >= 32 ? 0xFFFFFFFFu : (1u << (DMAX & 31)) - 1u)) {
                // every duration of the template in use: the full rounds are unrolled
                // (span <= T*CHUNK), reading all durations' prefixes before the arithmetic
#pragma unroll
                for (int r = 0; r < CHUNK; ++r) {
                    if (r * T + T - 1 + dmax_ok > n_bins) break;
                    const int n = r * T + tid;
                    const int2 h0 = bv.hi(n);
                    int2 h1[DMAX];
#pragma unroll
                    for (int di = 0; di < DMAX; ++di) h1[di] = bv.hi(n + dd[di]);
#pragma unroll
                    for (int di = 0; di < DMAX; ++di) screen_box(h0, h1[di], di);
                    n_first = r * T + T;
                }
            }
            for (int n0 = n_first; n0 < n_bins; n0 += T) {
                const int n = n0 + tid;
                if (n0 + T - 1 + dmax_ok <= n_bins) {
                    const int2 h0 = bv.hi(n);
#pragma unroll
                    for (int di = 0; di < DMAX; ++di)
                        if ((dok >> di) & 1u) screen_box(h0, bv.hi(n + dd[di]), di);
                } else if (n < n_bins) {
                    const int2 h0 = bv.hi(n);
#pragma unroll
                    for (int di = 0; di < DMAX; ++di)
                        if (((dok >> di) & 1u) && n + dd[di] <= n_bins)
                            screen_box(h0, bv.hi(n + dd[di]), di);
                }
            }
            int lane_max = 0;
#pragma unroll
            for (int di = 0; di < DMAX; ++di) lane_max = max(lane_max, cm[di]);
            const int vmax = wave_max_nonneg_dpp(lane_max);
            if (lane == 0) s_red_v[wave] = vmax;
            __syncthreads();
            int Vb = s_red_v[0];
#pragma unroll
            for (int w = 1; w < NW; ++w) Vb = max(Vb, s_red_v[w]);
            const float V = __int_as_float(Vb);
            const bool force = (cv.flags & 1) || !(V >= cv.vdeg) || !(V >= cv.vbnd);
            const float thr = V * kKeep;
            const int thr_bits = __float_as_int(thr);

            // (b3) exact re-evaluation of what the screen kept, first wins: the wanted
            // durations of this thread, in ascending order. Waves none of whose lanes
            // reached the threshold (or holds an uncertain box) have nothing to do.
            BoxKey best{-1.0, 1.0, 0, 0};
            uint32_t bkey = 0xFFFFFFFFu;
            if (__ballot(force || lane_max >= thr_bits || unc_bits != 0) != 0) {
                uint32_t want_bits = 0;
#pragma unroll
                for (int di = 0; di < DMAX; ++di)
                    if (force || (cm[di] >= thr_bits) || ((unc_bits >> di) & 1u))
                        want_bits |= 1u << di;
                want_bits &= dok;
                while (want_bits) {
                    const int di = __builtin_ctz(want_bits);
                    want_bits &= want_bits - 1;
                    const int d = s_dur[di];
                    const int m = n_bins - d + 1;
                    for (int n = tid; n < m; n += T) {
                        bool cert, unc;
                        const float v = screen_value<OBJ>(bv.hi(n), bv.hi(n + d), s_hi, e_unc,
                                                          dmin, cert, unc);
                        if (force || (cert && v >= thr) || unc) {
                            BoxKey kb;
                            if (exact_key<OBJ, MODE>(bv, n, d, cv, kb) &&
                                kb.num2 * best.den > best.num2 * kb.den) {
                                best = kb;
                                bkey = ((uint32_t)di << 24) | (uint32_t)n;
                            }
                        }
                    }
                }
                wave_best_ballot(best, bkey);
            }
            if (lane == 0) {
                s_red_b[wave] = best;
                s_red_k[wave] = bkey;
            }
            __syncthreads();

            // the period's winner (first wins across the waves), kept for (c). Nothing
            // below reads the bins, so the next period's fold can start while thread 0
            // finishes.
            if (tid == 0) {
                BoxKey b = s_red_b[0];
                uint32_t bk = s_red_k[0];
                for (int w = 1; w < NW; ++w) {
                    const BoxKey o = s_red_b[w];
                    if (key_better(o.num2, o.den, s_red_k[w], b.num2, b.den, bk)) {
                        b = o;
                        bk = s_red_k[w];
                    }
                }
                // every thread read slot q at the start of this period
                s_per[q] = __longlong_as_double((long long)b.yin);
                s_invp[q] = __longlong_as_double((long long)b.win);
                s_nb[q] = (int32_t)(b.num2 >= 0.0 ? bk : 0xFFFFFFFFu);
            }
            if constexpr (GLOBAL) {
                __threadfence();
                __syncthreads();
            }
        }

        // (c) finalize in FP64 from each winner's exact sums and write the seven outputs:
        // one lane per period of the item (coalesced rows).
        __syncthreads();
        int ftid = tid;
        asm volatile("" : "+v"(ftid));
        if (!(cv.flags & 2) && ftid < it.count) {
            const int q = ftid;
            const double P = segment_period(sg, periods, it.begin + q);
            const int64_t col = sg.out_offset + it.begin + q;
            const uint32_t bk = (uint32_t)s_nb[q];
            if (bk == 0xFFFFFFFFu) {
                out[col] = -INFINITY;
                for (int o = 1; o < CPB_N_OUTPUTS; ++o) out[o * n_out + col] = 0.0;
            } else {
                const int64_t byin = (int64_t)__double_as_longlong(s_per[q]);
                const int64_t bwin = (int64_t)__double_as_longlong(s_invp[q]);
                const int n = (int)(bk & 0xFFFFFFu);
                const int d = s_dur[bk >> 24];
                const double iv_in = ldexp((double)bwin, cv.w_exp);
                const double iv_out = ldexp((double)(cv.w_tot - bwin), cv.w_exp);
                const double y_in = ldexp((double)byin, cv.y_exp);
                const double y_rest = ldexp((double)(cv.y_tot - byin), cv.y_exp);
                const double yin = y_in / iv_in;
                const double yout = y_rest / iv_out;
                const double depth = yout - yin;
                const double derr = sqrt(1.0 / iv_in + 1.0 / iv_out);
                const double dsnr = depth / derr;
                const double ll = 0.5 * iv_in * depth * depth;
                const double dur = (double)d * bd;
                // (n*bd + dur/2) lies in [0, 2P): fmod by one exact subtraction.
                const double a = (double)n * bd + 0.5 * dur;
                const double ph = a < P ? a : (a < 2.0 * P ? a - P : fmod(a, P));
                out[col] = OBJ == 0 ? dsnr : ll;
                out[1 * n_out + col] = depth;
                out[2 * n_out + col] = derr;
                out[3 * n_out + col] = dsnr;
                out[4 * n_out + col] = dur;
                out[5 * n_out + col] = ph + cv.t_min;
                out[6 * n_out + col] = ll;
            }
        }
    }
}

// Reduced precision, for the precision study (cpb_set_accumulation): phase bins in FP32,
// filled by LDS float atomics.
//   ACC 1, FP64 phase + compensated FP32: the FP64 fold of the main kernel; each atomic
//          add returns the bin's previous value, from which TwoSum recovers the add's
//          rounding error exactly; the error (and the FP32 residual of the input value)
//          is added to a second FP32 bin, and prefix sums and box differences are carried
//          in double-float (hi, lo) pairs;
//   ACC 2, all FP32: the fold x - P*floor(x/P) in FP32 from FP32 tau and period (as the
//          reference's float32 kernel does), plain FP32 bins, prefix sums and box
//          differences.
// The objective of every box is evaluated in FP32; the winning box is finalised in FP64
// from its FP32 (collapsed double-float) sums. One workgroup per work item, one period
// after another; LDS holds 2 (ACC 2) or 4 (ACC 1) FP32 arrays of the phase span.

struct DF {
    float hi, lo;
};

__device__ __forceinline__ DF two_sum(float a, float b) {
    const float s = a + b;
    const float bp = s - a;
    return {s, (a - (s - bp)) + (b - bp)};
}

__device__ __forceinline__ DF df_add(DF a, DF b) {
    const DF s = two_sum(a.hi, b.hi);
    const float lo = s.lo + (a.lo + b.lo);
    const float hi = s.hi + lo;
    return {hi, lo - (hi - s.hi)};
}

__device__ __forceinline__ DF df_neg(DF a) { return {-a.hi, -a.lo}; }

template <int ACC>
struct PrecBins {
    float* yh;
    float* wh;
    float* yl;  // ACC 1 only
    float* wl;
    __device__ __forceinline__ DF y(int i) const { return {yh[i], ACC == 1 ? yl[i] : 0.0f}; }
    __device__ __forceinline__ DF w(int i) const { return {wh[i], ACC == 1 ? wl[i] : 0.0f}; }
    __device__ __forceinline__ void set(int i, DF yv, DF wv) {
        yh[i] = yv.hi;
        wh[i] = wv.hi;
        if (ACC == 1) {
            yl[i] = yv.lo;
            wl[i] = wv.lo;
        }
    }
};

template <int ACC>
__device__ __forceinline__ DF acc_add(DF a, DF b) {
    if constexpr (ACC == 1) return df_add(a, b);
    return {a.hi + b.hi, 0.0f};
}

template <int ACC>
__device__ __forceinline__ float acc_diff(DF a, DF b) {
    if constexpr (ACC == 1) {
        const DF d = df_add(a, df_neg(b));
        return d.hi + d.lo;
    }
    return a.hi - b.hi;
}

template <int ACC, int OBJ>
__global__ __launch_bounds__(256) void bls_precision_kernel(
    const double* __restrict__ tau, const double* __restrict__ yw64,
    const double* __restrict__ wv64, const CurveParams* __restrict__ curves,
    const cpb_segment* __restrict__ segs, const WorkItem* __restrict__ items,
    int64_t n_items, const double* __restrict__ periods,
    const int32_t* __restrict__ dur_bins, int32_t oversample, int32_t wl,
    double* __restrict__ out, int64_t n_out) {
    constexpr int T = 256;
    constexpr int NW = T / kWave;
    extern __shared__ float prec_lds[];
    __shared__ int32_t s_dur[kMaxDurations];
    __shared__ DF s_tot[2 * T];
    __shared__ float s_bv[NW];
    __shared__ uint32_t s_bk[NW];
    __shared__ float s_by[NW], s_bw[NW];
    PrecBins<ACC> bins{prec_lds, prec_lds + wl, prec_lds + 2 * wl, prec_lds + 3 * wl};
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const WorkItem it = items[item];
        const cpb_segment sg = segs[it.seg];
        const CurveParams cv = curves[sg.curve];
        const int n_dur = sg.n_dur;
        const double bd = sg.bin_duration;
        const double inv_bd = 1.0 / bd;
        const int width = sg.width;
        const int64_t poff = cv.point_offset;
        const int npts = cv.n_points;
        const double ytot64 = ldexp((double)cv.y_tot, cv.y_exp);
        const double wtot64 = ldexp((double)cv.w_tot, cv.w_exp);
        const DF ytot{(float)ytot64, (float)(ytot64 - (double)(float)ytot64)};
        const DF wtot{(float)wtot64, (float)(wtot64 - (double)(float)wtot64)};
        const float floor32 = (float)cv.ivar_floor;
        __syncthreads();
        if (tid < n_dur) s_dur[tid] = dur_bins[sg.dur_offset + tid];
        __syncthreads();
        for (int q = 0; q < it.count; ++q) {
            const int64_t k = it.begin + q;
            const double P = segment_period(sg, periods, k);
            const double invP = 1.0 / P;
            const float Pf = (float)P, bdf = (float)bd;
            int64_t nb64 = (int64_t)ceil(P / bd) + oversample;
            if (nb64 > width - 1) nb64 = width - 1;
            const int n_bins = (int)nb64;
            const int64_t col = sg.out_offset + k;
            if (cv.flags & 2) {
                if (tid == 0) {
                    out[col] = -INFINITY;
                    for (int o = 1; o < CPB_N_OUTPUTS; ++o) out[o * n_out + col] = 0.0;
                }
                continue;
            }
            for (int i = tid; i <= n_bins; i += T) bins.set(i, DF{0.0f, 0.0f}, DF{0.0f, 0.0f});
            __syncthreads();
            for (int j = tid; j < npts; j += T) {
                int b;
                if constexpr (ACC == 2) {
                    // all-FP32: the phase too, from FP32 tau and period
                    const float xf = (float)tau[poff + j];
                    const float kf = floorf(xf / Pf);
                    const float wf = xf - Pf * kf;
                    b = (int)(wf / bdf) + 1;
                    b = b < 0 ? 0 : (b > width - 1 ? width - 1 : b);
                } else {
                    b = fold_bin(tau[poff + j], P, invP, bd, inv_bd, width);
                }
                const double a64 = yw64[poff + j], w64 = wv64[poff + j];
                const float a = (float)a64, w = (float)w64;
                const float ao = atomicAdd(&bins.yh[b], a);
                const float wo = atomicAdd(&bins.wh[b], w);
                if (ACC == 1) {
                    atomicAdd(&bins.yl[b], two_sum(ao, a).lo + (float)(a64 - (double)a));
                    atomicAdd(&bins.wl[b], two_sum(wo, w).lo + (float)(w64 - (double)w));
                }
            }
            __syncthreads();
            if (tid == 0) {
                for (int kk = 0; kk < oversample; ++kk) {
                    const int dst = n_bins - oversample + kk;
                    if (dst >= 0 && dst < width) bins.set(dst, bins.y(1 + kk), bins.w(1 + kk));
                }
            }
            __syncthreads();
            // inclusive prefix over [0, n_bins]: per-thread chunks, then chunk offsets
            const int span = n_bins + 1;
            const int c = (span + T - 1) / T;
            const int base = tid * c;
            DF sy{0.0f, 0.0f}, sw{0.0f, 0.0f};
            for (int i = base; i < base + c && i < span; ++i) {
                sy = acc_add<ACC>(sy, bins.y(i));
                sw = acc_add<ACC>(sw, bins.w(i));
            }
            s_tot[2 * tid] = sy;
            s_tot[2 * tid + 1] = sw;
            __syncthreads();
            if (tid == 0) {
                DF ry{0.0f, 0.0f}, rw{0.0f, 0.0f};
                for (int t = 0; t < T; ++t) {
                    const DF ty = s_tot[2 * t], tw = s_tot[2 * t + 1];
                    s_tot[2 * t] = ry;
                    s_tot[2 * t + 1] = rw;
                    ry = acc_add<ACC>(ry, ty);
                    rw = acc_add<ACC>(rw, tw);
                }
            }
            __syncthreads();
            sy = s_tot[2 * tid];
            sw = s_tot[2 * tid + 1];
            for (int i = base; i < base + c && i < span; ++i) {
                sy = acc_add<ACC>(sy, bins.y(i));
                sw = acc_add<ACC>(sw, bins.w(i));
                bins.set(i, sy, sw);
            }
            __syncthreads();
            // box sweep in FP32, first strict maximum in (duration index, start) order
            float best = -INFINITY, by = 0.0f, bw = 0.0f;
            uint32_t bk = 0xFFFFFFFFu;
            for (int di = 0; di < n_dur; ++di) {
                const int d = s_dur[di];
                if (d < 1 || d >= n_bins) continue;
                for (int n = tid; n <= n_bins - d; n += T) {
                    const float win = acc_diff<ACC>(bins.w(n + d), bins.w(n));
                    const float wout = acc_diff<ACC>(wtot, DF{win, 0.0f});
                    if (!(win >= floor32) || !(wout >= floor32)) continue;
                    const float yin = acc_diff<ACC>(bins.y(n + d), bins.y(n));
                    const float yrest = acc_diff<ACC>(ytot, DF{yin, 0.0f});
                    const float m_in = yin / win, m_out = yrest / wout;
                    if (m_out < m_in) continue;
                    const float depth = m_out - m_in;
                    const float obj = OBJ == 0 ? depth / sqrtf(1.0f / win + 1.0f / wout)
                                               : 0.5f * win * depth * depth;
                    if (obj > best) {
                        best = obj;
                        bk = ((uint32_t)di << 24) | (uint32_t)n;
                        by = yin;
                        bw = win;
                    }
                }
            }
            for (int off = kWave / 2; off > 0; off >>= 1) {
                const float ov = __shfl_xor(best, off, kWave);
                const uint32_t ok = __shfl_xor(bk, off, kWave);
                const float oy = __shfl_xor(by, off, kWave);
                const float ow = __shfl_xor(bw, off, kWave);
                if (ov > best || (ov == best && ok < bk)) {
                    best = ov;
                    bk = ok;
                    by = oy;
                    bw = ow;
                }
            }
            if (lane == 0) {
                s_bv[wave] = best;
                s_bk[wave] = bk;
                s_by[wave] = by;
                s_bw[wave] = bw;
            }
            __syncthreads();
            if (tid == 0) {
                best = s_bv[0];
                bk = s_bk[0];
                by = s_by[0];
                bw = s_bw[0];
                for (int w = 1; w < NW; ++w) {
                    if (s_bv[w] > best || (s_bv[w] == best && s_bk[w] < bk)) {
                        best = s_bv[w];
                        bk = s_bk[w];
                        by = s_by[w];
                        bw = s_bw[w];
                    }
                }
                if (!(best > -INFINITY)) {
                    out[col] = -INFINITY;
                    for (int o = 1; o < CPB_N_OUTPUTS; ++o) out[o * n_out + col] = 0.0;
                } else {
                    const int n = (int)(bk & 0xFFFFFFu);
                    const int d = s_dur[bk >> 24];
                    const double iv_in = (double)bw;
                    const double iv_out = wtot64 - iv_in;
                    const double yin = (double)by / iv_in;
                    const double yout = (ytot64 - (double)by) / iv_out;
                    const double depth = yout - yin;
                    const double derr = sqrt(1.0 / iv_in + 1.0 / iv_out);
                    const double dsnr = depth / derr;
                    const double ll = 0.5 * iv_in * depth * depth;
                    const double dur = (double)d * bd;
                    const double a = (double)n * bd + 0.5 * dur;
                    const double ph = a < P ? a : (a < 2.0 * P ? a - P : fmod(a, P));
                    out[col] = OBJ == 0 ? dsnr : ll;
                    out[1 * n_out + col] = depth;
                    out[2 * n_out + col] = derr;
                    out[3 * n_out + col] = dsnr;
                    out[4 * n_out + col] = dur;
                    out[5 * n_out + col] = ph + cv.t_min;
                    out[6 * n_out + col] = ll;
                }
            }
            __syncthreads();
        }
    }
}

// Per-curve summary of a search's power row, for batched runs that keep spectra on
// the device: with p = power with NaN and +-inf set to 0 (`_assemble`,
// src/cuperiod/methods/bls.py:280-282), summary = (mean(p), std(p) (two passes, ddof 0),
// period-order index of the best period, its p). The best period is the first maximum
// in ascending frequency, i.e. the last maximum in the ascending-period order the
// engine writes; the first maximum of a spectrum is always a peak candidate of
// `Periodogram.best_periods` (src/cuperiod/core/result.py:369-406).
__device__ __forceinline__ double sanitize(double v) { return isfinite(v) ? v : 0.0; }

__global__ __launch_bounds__(256) void curve_summary_kernel(
    const double* __restrict__ power, const int64_t* __restrict__ base,
    const int64_t* __restrict__ count, double* __restrict__ summary) {
    __shared__ double red_d[4];
    __shared__ int64_t red_i[4];
    const int64_t b = base[blockIdx.x];
    const int64_t n = count[blockIdx.x];
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    double s = 0.0;
    for (int64_t i = tid; i < n; i += 256) s += sanitize(power[b + i]);
    s = wave_sum_f64(s);
    if (lane == 0) red_d[wave] = s;
    __syncthreads();
    const double mean = n > 0 ? (red_d[0] + red_d[1] + red_d[2] + red_d[3]) / (double)n : 0.0;
    __syncthreads();
    double q = 0.0, bv = -INFINITY;
    int64_t bk = -1;
    for (int64_t i = tid; i < n; i += 256) {
        const double v = sanitize(power[b + i]);
        q += (v - mean) * (v - mean);
        if (v > bv || (v == bv && i > bk)) {
            bv = v;
            bk = i;
        }
    }
    q = wave_sum_f64(q);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off, kWave);
        const int64_t ok = shfl_i64(bk, lane ^ off);
        if (ov > bv || (ov == bv && ok > bk)) {
            bv = ov;
            bk = ok;
        }
    }
    __shared__ double red_v[4];
    if (lane == 0) {
        red_d[wave] = q;
        red_v[wave] = bv;
        red_i[wave] = bk;
    }
    __syncthreads();
    if (tid == 0) {
        const double var = n > 0 ? (red_d[0] + red_d[1] + red_d[2] + red_d[3]) / (double)n : 0.0;
        double v = red_v[0];
        int64_t k = red_i[0];
        for (int w = 1; w < 4; ++w)
            if (red_v[w] > v || (red_v[w] == v && red_i[w] > k)) {
                v = red_v[w];
                k = red_i[w];
            }
        summary[4 * blockIdx.x + 0] = mean;
        summary[4 * blockIdx.x + 1] = sqrt(var);
        summary[4 * blockIdx.x + 2] = (double)k;
        summary[4 * blockIdx.x + 3] = n > 0 ? v : 0.0;
    }
}

// Columns of a segment with no box width: power -inf, every other output 0.
// Peak candidates of each curve's sanitised power row, for batch rows that do not need
// the whole spectrum. Position x counts in ascending frequency (x = count-1-k for the
// engine's ascending-period column k). A candidate is an interior maximum
// (s[x] > s[x-1] and s[x] >= s[x+1]), x = 0 when s[0] >= s[1], x = n-1 when
// s[n-1] > s[n-2] (x = 0 alone when n = 1): the rule of `Periodogram._peak_candidates`
// (src/cuperiod/core/result.py:369-406). Sanitised power is never negative, so its bit
// pattern orders like its value; a two-pass radix select (11 bits each) finds a threshold
// holding at least `want` candidates, and every candidate at or above it is written (at
// most `cap`, in no particular order) with its power, depth, depth_snr, duration and
// transit_time. counts[c] = (all candidates, candidates at or above the threshold).
constexpr int kPeakBins = 2048;

__device__ __forceinline__ uint64_t peak_bits(const double* __restrict__ power, int64_t b,
                                              int64_t n, int64_t x) {
    const double v = sanitize(power[b + (n - 1 - x)]);
    return v > 0.0 ? (uint64_t)__double_as_longlong(v) : 0ull;
}

__device__ __forceinline__ bool peak_candidate(const double* __restrict__ power, int64_t b,
                                               int64_t n, int64_t x, uint64_t& bits) {
    bits = peak_bits(power, b, n, x);
    if (n == 1) return true;
    const double s = sanitize(power[b + (n - 1 - x)]);
    if (x == 0) return s >= sanitize(power[b + (n - 2)]);
    const double left = sanitize(power[b + (n - x)]);
    if (x == n - 1) return s > left;
    const double right = sanitize(power[b + (n - 2 - x)]);
    return s > left && s >= right;
}

__global__ __launch_bounds__(256) void peak_candidates_kernel(
    const double* __restrict__ out, int64_t n_out, const int64_t* __restrict__ base,
    const int64_t* __restrict__ count, int32_t want, int32_t cap,
    int32_t* __restrict__ cand, int64_t* __restrict__ counts, double* __restrict__ values) {
    __shared__ uint32_t hist[kPeakBins];
    __shared__ uint64_t s_prefix;
    __shared__ int s_level;
    __shared__ uint32_t s_total, s_written;
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t b = base[c];
    const int64_t n = count[c];
    const double* power = out;
    // level 0: no threshold yet; 1: top 11 bits fixed; 2: top 22 bits fixed
    if (tid == 0) {
        s_prefix = 0;
        s_level = 0;
        s_total = 0;
        s_written = 0;
    }
    __syncthreads();
    uint32_t above = 0;  // candidates strictly above the current prefix bucket
    for (int level = 0; level < 2; ++level) {
        for (int i = tid; i < kPeakBins; i += 256) hist[i] = 0;
        __syncthreads();
        const int shift = 53 - 11 * level;
        const uint64_t pre = s_prefix;
        uint32_t local = 0;
        for (int64_t x = tid; x < n; x += 256) {
            uint64_t bits;
            if (!peak_candidate(power, b, n, x, bits)) continue;
            if (level == 0) ++local;
            if (level == 0 || (bits >> (shift + 11)) == pre) atomicAdd(&hist[(bits >> shift) & (kPeakBins - 1)], 1u);
        }
        if (level == 0 && local) atomicAdd(&s_total, local);
        __syncthreads();
        const uint32_t total = s_total;
        if (total <= (uint32_t)want) break;  // every candidate is kept
        if (tid == 0) {
            uint32_t cum = above;
            int bucket = kPeakBins - 1;
            for (; bucket > 0; --bucket) {
                if (cum + hist[bucket] >= (uint32_t)want) break;
                cum += hist[bucket];
            }
            s_prefix = (pre << 11) | (uint64_t)bucket;
            s_level = level + 1;
            hist[0] = cum;  // candidates above the chosen bucket
        }
        __syncthreads();
        above = hist[0];
        __syncthreads();
    }
    const int lvl = s_level;
    const uint64_t thr = lvl == 0 ? 0ull : (s_prefix << (64 - 11 * lvl));
    for (int64_t x = tid; x < n; x += 256) {
        uint64_t bits;
        if (!peak_candidate(power, b, n, x, bits) || bits < thr) continue;
        const uint32_t i = atomicAdd(&s_written, 1u);
        if (i < (uint32_t)cap) {
            const int64_t col = b + (n - 1 - x);
            cand[(int64_t)c * cap + i] = (int32_t)x;
            double* v = values + ((int64_t)c * cap + i) * 5;
            v[0] = sanitize(out[col]);
            v[1] = out[1 * n_out + col];
            v[2] = out[3 * n_out + col];
            v[3] = out[4 * n_out + col];
            v[4] = out[5 * n_out + col];
        }
    }
    __syncthreads();
    if (tid == 0) {
        counts[2 * c] = s_total;
        counts[2 * c + 1] = s_written;
    }
}

__global__ void fill_empty_kernel(double* __restrict__ out, int64_t n_out, int64_t off,
                                  int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    out[off + i] = -INFINITY;
    for (int o = 1; o < CPB_N_OUTPUTS; ++o) out[o * n_out + off + i] = 0.0;
}

// ---------------------------------------------------------------------------------
// Reference rounding of the near-maximum periods (cpb_set_peak_rounding).
//
// The main kernel's box sums are exact: boxes that hold the same points tie exactly and
// the first one in (duration, start) order wins, and the adjacent periods of a
// flat-topped peak, which fold the best box's points identically, have equal powers.
// The reference's FP64 kernel orders the same ties by the rounding of its sequential
// bin sums and prefix sums (src/cuperiod/methods/_bls_core.py:369-442). The periods that
// can hold a curve's best power are those within kPeakBand of its largest power (the
// reference's rounding moves a power by about 1e-13 relative); they are searched again
// with the reference's FP64 sequence, so the best period and its box are the ones it
// reports. The band is taken per reference segment (a group of the engine's segments),
// so a segment's outputs do not depend on the other segments searched with it; every
// period within kPeakBand of the curve's largest power is within it of its own
// segment's largest.

constexpr double kPeakBand = 0x1p-30;
constexpr int kPeakCap = 64;       // re-searched periods per reference segment, at most
constexpr int kPeakThreads = 256;

// One workgroup per reference segment g (engine segments seg_idx[seg_start[g] ..
// seg_start[g+1]), all of one curve): the largest sanitised power, then the periods
// within kPeakBand of it, as (engine segment, index in segment) pairs. When more than
// kPeakCap qualify, the choice is the same on every run: first the periods at the
// maximum itself, then the others, each in (segment, index) order; slots are handed out
// by a block-wide count of each 256-period stretch, not by atomics.
__global__ __launch_bounds__(kPeakThreads) void peak_band_kernel(
    const double* __restrict__ out, const cpb_segment* __restrict__ segs,
    const int32_t* __restrict__ seg_start, const int32_t* __restrict__ seg_idx,
    const cpb_curve* __restrict__ curves, int2* __restrict__ cand,
    int32_t* __restrict__ count) {
    constexpr int NW = kPeakThreads / kWave;
    __shared__ double red[NW];
    __shared__ int32_t s_cnt[NW];
    __shared__ int32_t s_n;
    const int c = blockIdx.x;  // the group
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    const int s0 = seg_start[c], s1 = seg_start[c + 1];
    double m = 0.0;
    for (int s = s0; s < s1; ++s) {
        const cpb_segment& sg = segs[seg_idx[s]];
        for (int64_t k = tid; k < sg.n_periods; k += kPeakThreads)
            m = fmax(m, sanitize(out[sg.out_offset + k]));
    }
    m = wave_max_f64(m);
    if (lane == 0) red[wave] = m;
    if (tid == 0) s_n = 0;
    __syncthreads();
    m = red[0];
    for (int w = 1; w < NW; ++w) m = fmax(m, red[w]);
    // appends the periods with lo <= power <= hi in order until kPeakCap are held
    auto take = [&](double lo, double hi) {
        for (int s = s0; s < s1; ++s) {
            const int si = seg_idx[s];
            const cpb_segment& sg = segs[si];
            for (int64_t k0 = 0; k0 < sg.n_periods; k0 += kPeakThreads) {
                const int taken = s_n;  // uniform: written before the last barrier
                if (taken >= kPeakCap) return;
                const int64_t k = k0 + tid;
                bool hit = false;
                if (k < sg.n_periods) {
                    const double p = sanitize(out[sg.out_offset + k]);
                    hit = p >= lo && p <= hi;
                }
                const unsigned long long bal = __ballot(hit);
                if (lane == 0) s_cnt[wave] = __popcll(bal);
                __syncthreads();
                int slot = taken + __popcll(bal & ((1ull << lane) - 1ull));
                int total = 0;
                for (int w = 0; w < NW; ++w) {
                    if (w < wave) slot += s_cnt[w];
                    total += s_cnt[w];
                }
                if (hit && slot < kPeakCap)
                    cand[(int64_t)c * kPeakCap + slot] = make_int2(si, (int)k);
                __syncthreads();  // s_cnt and s_n are read
                if (tid == 0) s_n = taken + total;
                __syncthreads();
            }
        }
    };
    const cpb_curve cv = curves[segs[seg_idx[s0]].curve];
    if (m > 0.0 && cv.sum_ivar > 0.0 && isfinite(cv.sum_y) && isfinite(cv.sum_ivar)) {
        take(m, m);
        take(m - m * kPeakBand, nextafter(m, 0.0));
    }
    __syncthreads();
    if (tid == 0) count[c] = min(s_n, kPeakCap);
}

// One wave per (reference segment, candidate): the reference's FP64 search of that
// period.
//   bins:   my[ind] += yw[j], mi[ind] += wv[j] over the points in order (:369-378), with
//           ind = int((x - P*floor(x/P))/bd) + 1 clipped, by IEEE divisions; lanes compute
//           the bins of 64 points at once, lane 0 adds them in point order, keeping a run
//           of points in one bin in registers (the same additions in the same order);
//   wrap:   bin[n_bins-os+k] = bin[1+k] in order (:379-383);
//   prefix: sequential inclusive sums over [0, n_bins] (:384-390);
//   boxes:  every (duration, start) box's objective with the reference's formulas and
//           admissibility tests (:398-442; its screen skips only boxes that cannot win),
//           lanes taking strided starts; the first strict maximum in (duration index,
//           start) order wins;
//   finish: the seven outputs (:443-461).
// GLOBAL: the bins live in a per-workgroup slice of global memory instead of LDS.
template <bool GLOBAL>
__global__ __launch_bounds__(kWave) void peak_rounding_kernel(
    const double* __restrict__ tau, const double* __restrict__ yw,
    const double* __restrict__ wv, const cpb_curve* __restrict__ curves,
    const cpb_segment* __restrict__ segs, const double* __restrict__ periods,
    const int32_t* __restrict__ dur_bins, int32_t oversample, int32_t obj_flag,
    int32_t n_groups, const int2* __restrict__ cand, const int32_t* __restrict__ count,
    int32_t wl, double* __restrict__ wide, double* __restrict__ out, int64_t n_out) {
    extern __shared__ double peak_lds[];
    const int lane = threadIdx.x;
    double* my = GLOBAL ? wide + (int64_t)blockIdx.x * 2 * wl : peak_lds;
    double* mi = my + wl;
    const int64_t total = (int64_t)n_groups * kPeakCap;
    for (int64_t i = blockIdx.x; i < total; i += gridDim.x) {
        // slot-major order: every group's first candidate, then every second one, ...
        const int slot = (int)(i / n_groups);
        const int g = (int)(i % n_groups);
        if (slot >= min(count[g], kPeakCap)) continue;
        const int2 ck = cand[(int64_t)g * kPeakCap + slot];
        const cpb_segment sg = segs[ck.x];
        const cpb_curve cv = curves[sg.curve];
        const double P = segment_period(sg, periods, ck.y);
        const double bd = sg.bin_duration;
        const int width = sg.width;
        // a candidate has an admissible box, so its period is finite and positive
        if (!(P > 0.0 && P / bd < (double)width)) continue;
        int64_t nb64 = (int64_t)ceil(P / bd) + oversample;
        if (nb64 > width - 1) nb64 = width - 1;
        const int nb = (int)nb64;
        __syncthreads();  // the previous candidate's bins are no longer read
        for (int b = lane; b <= nb; b += kWave) {
            my[b] = 0.0;
            mi[b] = 0.0;
        }
        if constexpr (GLOBAL) __threadfence();
        __syncthreads();
        const int64_t off = cv.point_offset;
        const int64_t n = cv.n_points;
        int cur = -1;
        double acc_y = 0.0, acc_w = 0.0;
        for (int64_t j0 = 0; j0 < n; j0 += kWave) {
            const int64_t j = j0 + lane;
            int ind = 0;
            double a = 0.0, b = 0.0;
            if (j < n) {
                const double x = tau[off + j];
                const double w = x - P * floor(x / P);
                int64_t q = (int64_t)(w / bd) + 1;
                q = q < 0 ? 0 : (q > width - 1 ? width - 1 : q);
                ind = (int)q;
                a = yw[off + j];
                b = wv[off + j];
            }
            const int m = (int)min<int64_t>(kWave, n - j0);
            for (int q = 0; q < m; ++q) {
                const int iq = __builtin_amdgcn_readlane(ind, q);
                const double aq = readlane_f64(a, q);
                const double bq = readlane_f64(b, q);
                if (lane == 0) {
                    if (iq != cur) {
                        if (cur >= 0) {
                            my[cur] = acc_y;
                            mi[cur] = acc_w;
                        }
                        cur = iq;
                        acc_y = my[cur];
                        acc_w = mi[cur];
                    }
                    acc_y += aq;
                    acc_w += bq;
                }
            }
        }
        if (lane == 0) {
            if (cur >= 0) {
                my[cur] = acc_y;
                mi[cur] = acc_w;
            }
            for (int k = 0; k < oversample; ++k) {
                const int dst = nb - oversample + k;
                if (dst >= 0 && dst < width) {
                    my[dst] = my[1 + k];
                    mi[dst] = mi[1 + k];
                }
            }
            double ay = 0.0, ai = 0.0;
            for (int b = 0; b <= nb; ++b) {
                ay += my[b];
                my[b] = ay;
                ai += mi[b];
                mi[b] = ai;
            }
        }
        if constexpr (GLOBAL) __threadfence();
        __syncthreads();
        const double sum_y = cv.sum_y, sum_ivar = cv.sum_ivar, floor_iv = cv.ivar_floor;
        double loc = -INFINITY;
        uint32_t key = 0xFFFFFFFFu;
        for (int di = 0; di < sg.n_dur; ++di) {
            const int d = dur_bins[sg.dur_offset + di];
            if (d < 1 || d >= nb) continue;
            const int mbox = nb - d + 1;
            for (int s = lane; s < mbox; s += kWave) {
                const double y_in = my[s + d] - my[s];
                const double iv_in = mi[s + d] - mi[s];
                const double iv_out = sum_ivar - iv_in;
                if (iv_in < floor_iv || iv_out < floor_iv) continue;
                const double yin = y_in / iv_in;
                const double yout = (sum_y - y_in) / iv_out;
                if (yout < yin) continue;
                const double depth = yout - yin;
                const double obj = obj_flag == 0 ? depth / sqrt(1.0 / iv_in + 1.0 / iv_out)
                                                 : 0.5 * iv_in * depth * depth;
                if (obj > loc) {
                    loc = obj;
                    key = ((uint32_t)di << 24) | (uint32_t)s;
                }
            }
        }
        // first strict maximum: the largest objective, the smallest key among equals
#pragma unroll
        for (int offx = kWave / 2; offx > 0; offx >>= 1) {
            const double ov = __shfl_xor(loc, offx, kWave);
            const uint32_t ok = (uint32_t)__shfl_xor((int)key, offx, kWave);
            if (ov > loc || (ov == loc && ok < key)) {
                loc = ov;
                key = ok;
            }
        }
        if (lane == 0) {
            const int64_t col = sg.out_offset + ck.y;
            if (!(loc > -INFINITY)) {
                out[col] = -INFINITY;
                for (int o = 1; o < CPB_N_OUTPUTS; ++o) out[o * n_out + col] = 0.0;
            } else {
                const int ln = (int)(key & 0xFFFFFFu);
                const int ld = dur_bins[sg.dur_offset + (int)(key >> 24)];
                const double y_in = my[ln + ld] - my[ln];
                const double iv_in = mi[ln + ld] - mi[ln];
                const double iv_out = sum_ivar - iv_in;
                const double yin = y_in / iv_in;
                const double yout = (sum_y - y_in) / iv_out;
                const double depth = yout - yin;
                const double derr = sqrt(1.0 / iv_in + 1.0 / iv_out);
                const double dsnr = depth / derr;
                const double ll = 0.5 * iv_in * depth * depth;
                const double dur = (double)ld * bd;
                // (ln*bd + dur/2) >= 0 and P > 0: Python's float % is fmod here
                const double ph = fmod((double)ln * bd + 0.5 * dur, P);
                out[col] = obj_flag == 0 ? dsnr : ll;
                out[1 * n_out + col] = depth;
                out[2 * n_out + col] = derr;
                out[3 * n_out + col] = dsnr;
                out[4 * n_out + col] = dur;
                out[5 * n_out + col] = ph + cv.t_min;
                out[6 * n_out + col] = ll;
            }
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------------
// Host side.

namespace cpb_engine {

hipError_t launch_prep_curves(hipStream_t st, const double* tau, const double* yw,
                              const double* wv, const cpb_curve* curves, int32_t n_curves,
                              int32_t obj_flag, CurveParams* params, int64_t* yq,
                              int64_t* wq) {
    hipLaunchKernelGGL(prep_curves_kernel, dim3(n_curves), dim3(kPrepThreads), 0, st, tau, yw,
                       wv, curves, obj_flag, params, yq, wq);
    return hipGetLastError();
}

}  // namespace cpb_engine

namespace {

struct LaunchArgs {
    dim3 grid;
    size_t lds;
    hipStream_t st;
    const double* tau;
    const int64_t* yq;
    const int64_t* wq;
    const CurveParams* curves;
    const cpb_segment* segs;
    const WorkItem* items;
    int64_t n_items;
    const double* periods;
    const int32_t* dur;
    int32_t os;
    int32_t wl;
    int64_t* wide;
    double* out;
    int64_t n_out;
};

template <int T, int DMAX, int OBJ, int MODE>
hipError_t launch_main(const LaunchArgs& a) {
    auto fn = bls_main_kernel<T, DMAX, OBJ, MODE>;
    if (MODE != 2 && a.lds > 65536) {
        // Opt in to more than 64 KiB of dynamic LDS where the runtime asks for it; a
        // refusal surfaces at launch.
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds);
    }
    hipLaunchKernelGGL(fn, a.grid, dim3(T), a.lds, a.st, a.tau, a.yq, a.wq, a.curves,
                       a.segs, a.items, a.n_items, a.periods, a.dur, a.os, a.wl, a.wide,
                       a.out, a.n_out);
    return hipGetLastError();
}

template <int T, int OBJ, int MODE>
hipError_t dispatch_d(int dmax, const LaunchArgs& a) {
    if (dmax <= 6) return launch_main<T, 6, OBJ, MODE>(a);
    if (dmax <= 8) return launch_main<T, 8, OBJ, MODE>(a);
    if (dmax <= 16) return launch_main<T, 16, OBJ, MODE>(a);
    return launch_main<T, 32, OBJ, MODE>(a);
}

template <int T, int OBJ>
hipError_t dispatch_mode(int mode, int dmax, const LaunchArgs& a) {
    if (mode == 5) return dispatch_d<128, OBJ, 5>(dmax, a);
    if (mode == 3) return dispatch_d<T, OBJ, 3>(dmax, a);
    if (mode == 4) {
        if constexpr (T == 256) return dispatch_d<T, OBJ, 4>(dmax, a);
        else return dispatch_d<T, OBJ, 0>(dmax, a);
    }
    if (mode == 0) return dispatch_d<T, OBJ, 0>(dmax, a);
    if (mode == 1) return dispatch_d<T, OBJ, 1>(dmax, a);
    return dispatch_d<T, OBJ, 2>(dmax, a);
}

template <int OBJ>
hipError_t dispatch_threads(int threads, int mode, int dmax, const LaunchArgs& a) {
    if (threads == 128) return dispatch_mode<128, OBJ>(mode, dmax, a);
    if (threads == 512) return dispatch_mode<512, OBJ>(mode, dmax, a);
    return dispatch_mode<256, OBJ>(mode, dmax, a);
}

constexpr int kMaxDur = kMaxDurations;

// The reference rounding of every reference segment's near-maximum periods, after the
// main kernel: one workgroup per reference segment selects them, one wave per period
// searches it again.
int launch_peak_rounding(cpb_context* ctx, hipStream_t st, const double* tau, const double* yw,
                         const double* wv, const cpb_curve* curves, int32_t n_groups,
                         const cpb_segment* segs, const int32_t* g_start,
                         const int32_t* g_idx, const double* periods,
                         const int32_t* dur_bins, int32_t oversample, int32_t obj_flag,
                         int wmax, double* out, int64_t n_out) {
    if (n_groups == 0) return CPB_OK;
    const size_t cand_bytes = sizeof(int2) * (size_t)n_groups * kPeakCap;
    int rc;
    void* peak_buf = nullptr;
    if ((rc = device_buffer(ctx, ctx->peak, cand_bytes + sizeof(int32_t) * n_groups, &peak_buf)))
        return rc;
    auto* cand = static_cast<int2*>(peak_buf);
    auto* count = reinterpret_cast<int32_t*>(static_cast<char*>(peak_buf) + cand_bytes);
    hipLaunchKernelGGL(peak_band_kernel, dim3(n_groups), dim3(kPeakThreads), 0, st, out, segs,
                       g_start, g_idx, curves, cand, count);
    CPB_HIP(ctx, hipGetLastError());
    const int wl = wmax;  // n_bins <= width - 1 < wmax
    const size_t bins_bytes = (size_t)16 * wl;
    const int64_t slots = (int64_t)n_groups * kPeakCap;
    if (bins_bytes <= 65536) {
        const int64_t grid = std::min<int64_t>(slots, (int64_t)ctx->n_cu * 8);
        hipLaunchKernelGGL(peak_rounding_kernel<false>, dim3((unsigned)grid), dim3(kWave),
                           bins_bytes, st, tau, yw, wv, curves, segs, periods, dur_bins,
                           oversample, obj_flag, n_groups, cand, count, wl, nullptr, out,
                           n_out);
    } else {
        // bins in global memory: at most 1 GiB of them
        const int64_t fit = std::max<int64_t>(1, ((int64_t)1 << 30) / (int64_t)bins_bytes);
        const int64_t grid = std::min<int64_t>({slots, (int64_t)ctx->n_cu * 2, fit});
        void* wide_bins = nullptr;
        if ((rc = device_buffer(ctx, ctx->peak_wide, (size_t)grid * bins_bytes, &wide_bins)))
            return rc;
        hipLaunchKernelGGL(peak_rounding_kernel<true>, dim3((unsigned)grid), dim3(kWave), 0, st,
                           tau, yw, wv, curves, segs, periods, dur_bins, oversample, obj_flag,
                           n_groups, cand, count, wl,
                           static_cast<double*>(wide_bins), out, n_out);
    }
    CPB_HIP(ctx, hipGetLastError());
    return CPB_OK;
}

int run_search(cpb_context* ctx, hipStream_t st, const double* tau, const double* yw,
               const double* wv, int64_t total_points, const cpb_curve* curves,
               int32_t n_curves, const double* periods, const int32_t* dur_bins,
               const cpb_segment* segments, int32_t n_segments, int32_t oversample,
               int32_t obj_flag, double* out, int64_t n_out) {
    if (!ctx) return CPB_E_INVALID;
    if (obj_flag != 0 && obj_flag != 1) return fail(ctx, CPB_E_INVALID, "obj_flag must be 0 or 1");
    if (oversample < 1) return fail(ctx, CPB_E_INVALID, "oversample must be >= 1");
    if (n_curves < 1 || n_segments < 0 || total_points < 0 || n_out < 0)
        return fail(ctx, CPB_E_INVALID, "negative size");
    if (!curves || (n_segments > 0 && !segments))
        return fail(ctx, CPB_E_INVALID, "null descriptor array");
    for (int32_t c = 0; c < n_curves; ++c) {
        const cpb_curve& cv = curves[c];
        if (cv.point_offset < 0 || cv.n_points < 0 ||
            cv.point_offset + cv.n_points > total_points)
            return fail(ctx, CPB_E_INVALID, "curve points out of range");
        if (cv.n_points > (int64_t(1) << 30))
            return fail(ctx, CPB_E_RESOURCE, "more than 2^30 points in one light curve");
    }
    int dmax = 1, wmax = 2;
    int64_t total_periods = 0;
    for (int32_t s = 0; s < n_segments; ++s) {
        const cpb_segment& sg = segments[s];
        if (sg.curve < 0 || sg.curve >= n_curves)
            return fail(ctx, CPB_E_INVALID, "segment curve index out of range");
        if (sg.n_periods < 0 || sg.out_offset < 0 || sg.out_offset + sg.n_periods > n_out)
            return fail(ctx, CPB_E_INVALID, "segment output range out of bounds");
        if (sg.n_dur < 0 || sg.n_dur > kMaxDur)
            return fail(ctx, CPB_E_RESOURCE,
                        "more than 32 distinct box widths in one segment");
        if (sg.width < 2 || sg.width < oversample + 2)
            return fail(ctx, CPB_E_INVALID, "segment width must be at least oversample + 2");
        if (!(sg.bin_duration > 0.0))
            return fail(ctx, CPB_E_INVALID, "bin_duration must be positive");
        if (sg.mode == 0 && sg.n_periods > 0 && !periods)
            return fail(ctx, CPB_E_INVALID, "explicit periods needed");
        if (sg.mode < 0 || sg.mode > 2) return fail(ctx, CPB_E_INVALID, "unknown segment mode");
        if (sg.mode != 0 && (sg.lin_num < 1 || sg.lin_k0 < 0 ||
                             sg.lin_k0 + sg.n_periods > sg.lin_num))
            return fail(ctx, CPB_E_INVALID, "bad parametric segment");
        if (sg.n_dur > 0 && sg.n_periods > 0) {
            dmax = std::max(dmax, (int)sg.n_dur);
            wmax = std::max(wmax, (int)sg.width);
        }
        total_periods += sg.n_periods;
    }
    (void)total_periods;

    if (wmax > (1 << 24))
        return fail(ctx, CPB_E_RESOURCE, "more than 2^24 phase bins per period");

    CPB_HIP(ctx, hipSetDevice(ctx->device));
    int rc;
    void *curves_buf = nullptr, *yq_buf = nullptr, *wq_buf = nullptr;
    const size_t point_bytes = sizeof(int64_t) * std::max<int64_t>(1, total_points);
    if ((rc = device_buffer(ctx, ctx->curves, sizeof(CurveParams) * n_curves, &curves_buf)))
        return rc;
    if ((rc = device_buffer(ctx, ctx->yq, point_bytes, &yq_buf))) return rc;
    if ((rc = device_buffer(ctx, ctx->wq, point_bytes, &wq_buf))) return rc;
    auto* cparams = static_cast<CurveParams*>(curves_buf);
    auto* yq_d = static_cast<int64_t*>(yq_buf);
    auto* wq_d = static_cast<int64_t*>(wq_buf);

    // Segments with no periods or no box widths: power -inf, the rest 0 (the
    // reference's `_empty_output`, src/cuperiod/methods/_bls_core.py:843-847).
    // Work items: tiles of consecutive periods of one segment.
    int64_t live = 0;
    for (int32_t s = 0; s < n_segments; ++s)
        if (segments[s].n_dur > 0) live += segments[s].n_periods;
    const int64_t target_items = (int64_t)ctx->n_cu * 64;
    int64_t tile = live / std::max<int64_t>(1, target_items);
    tile = std::max<int64_t>(1, std::min<int64_t>(tile, kMaxTile));
    if (ctx->tile > 0) tile = ctx->tile;
    // Work items grouped by the phase-bin layout their segment's width needs (MODE 0:
    // register-chunk prefix in LDS; 1: in-place slots in LDS; 2: global scratch), one
    // launch per layout, so a wide segment does not push narrow ones out of LDS. (The
    // reduced-precision kernel takes every item in one launch.)
    const int threads = ctx->threads;
    // Classes: 0 MODE 0 up to 640 bins in 128-thread workgroups whatever the launch
    // shape (10 KiB of LDS, 8 waves per SIMD), 1 up to 1,152 bins (18 KiB; its own
    // instantiation for 8 waves per SIMD: 1,152 is the widest span whose odd prefix chunk
    // fits its registers at 128, 256 and 512 threads), 2 up to 1,344 bins (21 KiB, 7
    // waves per SIMD with 256 threads), 3-5 MODE 0 up to 1,600 / 1,920 / threads*8 bins
    // (16 bytes a bin: 25, 30, 32 KiB, the most bins that leave room for 6, 5, 4
    // workgroups in a CU's 160 KiB with up to 1.25 KiB of static LDS each), 6 MODE 1,
    // 7 MODE 2.
    constexpr int kClasses = 8;
    auto layout_of = [&](int width) {
        if (ctx->accumulation != 0) return 0;
        const int w = (width + 63) / 64 * 64;
        const size_t need = (size_t)16 * (w + (w - 1) / 8 + 1);
        if (need + 2048 > (size_t)ctx->lds_limit) return 7;
        if (w > threads * kChunk) return 6;
        if (w <= 640) return 0;
        // class 1's prefix holds an odd chunk of at most 1,024/threads + 1 bins a thread
        if (w <= 1152 && (((w + threads - 1) / threads) | 1) <= 1024 / threads + 1) return 1;
        return w <= 1344 ? 2 : (w <= 1600 ? 3 : (w <= 1920 ? 4 : 5));
    };
    constexpr int kModeOf[kClasses] = {5, 3, 4, 0, 0, 0, 1, 2};
    struct Layout {
        int64_t begin = 0, count = 0;
        int wmax = 2, dmax = 1;
    } layouts[kClasses];
    std::vector<WorkItem> by_layout[kClasses];
    std::vector<cpb_segment> empties;
    for (int32_t s = 0; s < n_segments; ++s) {
        const cpb_segment& sg = segments[s];
        if (sg.n_periods == 0) continue;
        if (sg.n_dur == 0) {
            empties.push_back(sg);
            continue;
        }
        const int L = layout_of(sg.width);
        layouts[L].wmax = std::max(layouts[L].wmax, (int)sg.width);
        layouts[L].dmax = std::max(layouts[L].dmax, (int)sg.n_dur);
        for (int64_t b = 0; b < sg.n_periods; b += tile) {
            WorkItem w;
            w.seg = s;
            w.begin = b;
            w.count = (int32_t)std::min<int64_t>(tile, sg.n_periods - b);
            by_layout[L].push_back(w);
        }
    }
    ctx->h_items.clear();
    for (int L = 0; L < kClasses; ++L) {
        layouts[L].begin = (int64_t)ctx->h_items.size();
        layouts[L].count = (int64_t)by_layout[L].size();
        ctx->h_items.insert(ctx->h_items.end(), by_layout[L].begin(), by_layout[L].end());
    }
    const int64_t n_items = (int64_t)ctx->h_items.size();

    // The searched segments of each reference segment (equal curve and group), in
    // order of first appearance, for the reference rounding of its peak (CSR).
    std::vector<int32_t> cs_start(1, 0), cs_idx;
    {
        std::vector<std::pair<int64_t, int32_t>> keyed;
        keyed.reserve(n_segments);
        for (int32_t s = 0; s < n_segments; ++s)
            if (segments[s].n_periods > 0 && segments[s].n_dur > 0)
                keyed.emplace_back(((int64_t)segments[s].curve << 32) |
                                       (int64_t)(uint32_t)segments[s].group, s);
        std::stable_sort(keyed.begin(), keyed.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t i = 0; i < keyed.size(); ++i) {
            if (i > 0 && keyed[i].first != keyed[i - 1].first)
                cs_start.push_back((int32_t)i);
            cs_idx.push_back(keyed[i].second);
        }
        cs_start.push_back((int32_t)keyed.size());
        if (keyed.empty()) cs_start.assign(1, 0);
    }
    const int32_t n_groups = (int32_t)cs_start.size() - 1;
    if (cs_idx.empty()) cs_idx.push_back(0);

    // Descriptor tables: [curves | segments | work items | curve segment lists],
    // 256-byte aligned parts.
    auto align = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t off_seg = align(sizeof(cpb_curve) * n_curves);
    const size_t off_item = off_seg + align(sizeof(cpb_segment) * std::max(1, n_segments));
    const size_t off_cs = off_item + align(sizeof(WorkItem) * std::max<int64_t>(1, n_items));
    const size_t off_ci = off_cs + align(sizeof(int32_t) * cs_start.size());
    const size_t tbytes = off_ci + align(sizeof(int32_t) * cs_idx.size());
    void* tables_buf = nullptr;
    if ((rc = device_buffer(ctx, ctx->tables, tbytes, &tables_buf))) return rc;
    if (ctx->copy_pending) {
        CPB_HIP(ctx, hipEventSynchronize(ctx->copied));
        ctx->copy_pending = false;
    }
    char* hb;
    if (ctx->carve_base) {
        // pageable staging: the copy below has read it when it returns
        if (ctx->h_stage.size() < tbytes) ctx->h_stage.resize(tbytes);
        hb = ctx->h_stage.data();
    } else {
        if (ctx->h_tables_bytes < tbytes) {
            if (ctx->h_tables) CPB_HIP(ctx, hipHostFree(ctx->h_tables));
            ctx->h_tables = nullptr;
            ctx->h_tables_bytes = 0;
            CPB_HIP(ctx, hipHostMalloc(&ctx->h_tables, tbytes + tbytes / 4, hipHostMallocDefault));
            ctx->h_tables_bytes = tbytes + tbytes / 4;
        }
        hb = static_cast<char*>(ctx->h_tables);
    }
    std::memcpy(hb, curves, sizeof(cpb_curve) * n_curves);
    if (n_segments > 0) std::memcpy(hb + off_seg, segments, sizeof(cpb_segment) * n_segments);
    if (n_items > 0)
        std::memcpy(hb + off_item, ctx->h_items.data(), sizeof(WorkItem) * n_items);
    std::memcpy(hb + off_cs, cs_start.data(), sizeof(int32_t) * cs_start.size());
    std::memcpy(hb + off_ci, cs_idx.data(), sizeof(int32_t) * cs_idx.size());
    CPB_HIP(ctx, hipMemcpyAsync(tables_buf, hb, tbytes, hipMemcpyHostToDevice, st));
    if (!ctx->capturing && !ctx->carve_base) {
        CPB_HIP(ctx, hipEventRecord(ctx->copied, st));
        ctx->copy_pending = true;
    }
    char* db = static_cast<char*>(tables_buf);
    const auto* curve_in = reinterpret_cast<const cpb_curve*>(db);
    const auto* sp = reinterpret_cast<const cpb_segment*>(db + off_seg);
    const auto* ip = reinterpret_cast<const WorkItem*>(db + off_item);
    const auto* csp = reinterpret_cast<const int32_t*>(db + off_cs);
    const auto* cip = reinterpret_cast<const int32_t*>(db + off_ci);

    hipLaunchKernelGGL(prep_curves_kernel, dim3(n_curves), dim3(kPrepThreads), 0, st, tau,
                       yw, wv, curve_in, obj_flag,
                       cparams, yq_d, wq_d);
    CPB_HIP(ctx, hipGetLastError());

    for (const cpb_segment& sg : empties) {
        const unsigned blocks = (unsigned)((sg.n_periods + 255) / 256);
        hipLaunchKernelGGL(fill_empty_kernel, dim3(blocks), dim3(256), 0, st, out, n_out,
                           sg.out_offset, sg.n_periods);
        CPB_HIP(ctx, hipGetLastError());
    }
    if (n_items == 0) return CPB_OK;

    const int wl = (wmax + 63) / 64 * 64;
    if (ctx->accumulation != 0) {
        const size_t plds = (size_t)(ctx->accumulation == 1 ? 4 : 2) * wl * sizeof(float);
        if (plds + 8192 > (size_t)ctx->lds_limit)
            return fail(ctx, CPB_E_RESOURCE,
                        "reduced-precision accumulation keeps the phase bins in LDS; the "
                        "phase span is too wide");
        // a launch (and a graph kernel node) holds fewer than 2^32 work-items; the
        // kernel loops over the items its grid does not cover
        const dim3 grid((unsigned)std::min<int64_t>(n_items, ((int64_t)1 << 31) / 256));
        const auto* cp = cparams;
        const bool timed = ctx->timing && !ctx->capturing;
        if (timed) CPB_HIP(ctx, hipEventRecord(ctx->t0, st));
        auto launch = [&](auto fn) {
            if (plds > 65536)
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds);
            hipLaunchKernelGGL(fn, grid, dim3(256), plds, st, tau, yw, wv, cp, sp, ip, n_items,
                               periods, dur_bins, oversample, wl, out, n_out);
        };
        if (ctx->accumulation == 1)
            obj_flag == 0 ? launch(bls_precision_kernel<1, 0>) : launch(bls_precision_kernel<1, 1>);
        else
            obj_flag == 0 ? launch(bls_precision_kernel<2, 0>) : launch(bls_precision_kernel<2, 1>);
        CPB_HIP(ctx, hipGetLastError());
        if (timed) {
            CPB_HIP(ctx, hipEventRecord(ctx->t1, st));
            ctx->timed = true;
        }
        return CPB_OK;
    }
    const size_t static_lds = 2048;
    const int64_t persistent = (int64_t)ctx->n_cu * ctx->grid_per_cu;
    const bool timed = ctx->timing && !ctx->capturing;
    if (timed) CPB_HIP(ctx, hipEventRecord(ctx->t0, st));
    for (int L = 0; L < kClasses; ++L) {
        const Layout& lay = layouts[L];
        if (lay.count == 0) continue;
        const int wlL = (lay.wmax + 63) / 64 * 64;
        const int mode = kModeOf[L];
        const int n_slots = wlL + (wlL - 1) / 8 + 1;
        const size_t lds_need = (size_t)16 * (mode == 0 || mode >= 3 ? wlL : n_slots);
        (void)static_lds;
        int64_t* wide_ptr = nullptr;
        LaunchArgs la;
        la.lds = 0;
        if (mode == 2) {
            const int64_t g = std::min<int64_t>(
                lay.count, persistent > 0 ? persistent : (int64_t)ctx->n_cu * 4);
            void* wide_buf = nullptr;
            if ((rc = device_buffer(ctx, ctx->wide, (size_t)g * lds_need, &wide_buf))) return rc;
            wide_ptr = static_cast<int64_t*>(wide_buf);
            la.grid = dim3((unsigned)g);
        } else {
            // fewer than 2^32 work-items per launch (a graph kernel node refuses more);
            // the kernel loops over the items its grid does not cover
            const int tl = mode == 5 ? 128 : threads;
            const int64_t cap = ((int64_t)1 << 31) / tl;
            // a persistent grid holds no more workgroups per CU than can be resident at
            // once (the class's LDS, with up to 1.25 KiB of static LDS each, and the
            // waves its register cap admits); more would run after the others in a tail
            int64_t g = cap;
            if (persistent > 0) {
                const int waves_cap = mode == 3 ? 8 : (mode == 4 || mode == 5 ? 7 : 6);
                const int by_waves = std::max(1, 4 * waves_cap / (tl / kWave));
                const int by_lds = std::max<int>(1, ctx->lds_per_cu / (int)(lds_need + 1280));
                g = (int64_t)ctx->n_cu * std::min(ctx->grid_per_cu, std::min(by_waves, by_lds));
            }
            la.grid = dim3((unsigned)std::min<int64_t>(lay.count, g));
            la.lds = lds_need;
        }
        la.st = st;
        la.tau = tau;
        la.yq = yq_d;
        la.wq = wq_d;
        la.curves = cparams;
        la.segs = sp;
        la.items = ip + lay.begin;
        la.n_items = lay.count;
        la.periods = periods;
        la.dur = dur_bins;
        la.os = oversample;
        la.wl = wlL;
        la.wide = wide_ptr;
        la.out = out;
        la.n_out = n_out;
        const hipError_t e = obj_flag == 0 ? dispatch_threads<0>(threads, mode, lay.dmax, la)
                                           : dispatch_threads<1>(threads, mode, lay.dmax, la);
        if (e != hipSuccess) return hip_fail(ctx, e, "bls_main_kernel launch");
    }
    if (timed) {
        CPB_HIP(ctx, hipEventRecord(ctx->t1, st));
        ctx->timed = true;
    }
    if (ctx->peak_rounding)
        return launch_peak_rounding(ctx, st, tau, yw, wv, curve_in, n_groups, sp, csp, cip,
                                    periods, dur_bins, oversample, obj_flag, wmax, out, n_out);
    return CPB_OK;
}

// Scratch of cpb_bls_search_batch_static, for callers that cannot allocate device
// memory: the fixed-point points, curve parameters, descriptor tables and the peak
// rounding's candidate lists of one search are carved from it. Every GPU the library
// runs on holds its own copy until the process ends. Host code never passes the array's
// name as an address: each device's copy is found once with hipGetSymbolAddress (with
// that device current) and that pointer is what kernels and copies receive. Calls that
// overlap on one GPU share the array and race. Nothing in it needs to start at zero:
// every part is written by the call before it is read.
constexpr size_t kStaticScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kStaticScratchBytes];

constexpr int kMaxStaticDevices = 64;
std::mutex g_static_mu;
void* g_static_ptr[kMaxStaticDevices] = {};
hipError_t g_static_err[kMaxStaticDevices] = {};
bool g_static_done[kMaxStaticDevices] = {};

// This device's copy of `scratch` (the device must be current); resolved once per device.
void* static_scratch_address(int device, std::string& error) {
    if (device < 0 || device >= kMaxStaticDevices) {
        error = "device index beyond the static scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (!g_static_done[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_static_err[device] = e;
        g_static_ptr[device] = e == hipSuccess ? p : nullptr;
        g_static_done[device] = true;
    }
    if (!g_static_ptr[device])
        error = std::string("hipGetSymbolAddress(scratch): ") +
                hipGetErrorString(g_static_err[device]);
    return g_static_ptr[device];
}

}  // namespace

extern "C" {

int cpb_version(void) { return 100; }

int cpb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cpb_create(int device, cpb_context** out) {
    if (!out) return CPB_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CPB_E_NODEVICE;
    if (device < 0 || device >= n) return CPB_E_INVALID;
    auto* ctx = new cpb_context();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return CPB_E_DEVICE;
    }
    if (hipEventCreateWithFlags(&ctx->copied, hipEventDisableTiming) != hipSuccess) {
        delete ctx;
        return CPB_E_DEVICE;
    }
    int lds = 0, cus = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
            hipSuccess && lds > 0)
        ctx->lds_limit = lds;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) ==
            hipSuccess && cus > 0)
        ctx->n_cu = cus;
    int lds_cu = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                              device) == hipSuccess && lds_cu > 0)
        ctx->lds_per_cu = lds_cu;
    ctx->static_scratch = static_scratch_address(device, ctx->static_error);
    *out = ctx;
    return CPB_OK;
}

void cpb_destroy(cpb_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->copy_pending) (void)hipEventSynchronize(ctx->copied);
    for (DevBuf* b : {&ctx->tables, &ctx->curves, &ctx->yq, &ctx->wq, &ctx->wide, &ctx->peak,
                      &ctx->peak_wide})
        if (b->ptr) (void)hipFree(b->ptr);
    if (ctx->h_tables) (void)hipHostFree(ctx->h_tables);
    if (ctx->copied) (void)hipEventDestroy(ctx->copied);
    if (ctx->t0) (void)hipEventDestroy(ctx->t0);
    if (ctx->t1) (void)hipEventDestroy(ctx->t1);
    delete ctx;
}

const char* cpb_last_error(const cpb_context* ctx) {
    return ctx ? ctx->error.c_str() : "null context";
}

const char* cpb_status_string(int status) {
    switch (status) {
        case CPB_OK: return "ok";
        case CPB_E_INVALID: return "invalid argument";
        case CPB_E_DEVICE: return "HIP runtime error";
        case CPB_E_RESOURCE: return "resource limit exceeded";
        case CPB_E_NODEVICE: return "no HIP device";
        default: return "unknown status";
    }
}

int cpb_bls_search(cpb_context* ctx, void* stream, const double* tau, const double* yw,
                   const double* wv, int64_t n_points, double t_min, double ivar_floor,
                   double sum_y, double sum_ivar, const double* periods, const int32_t* dur_bins,
                   const cpb_segment* segments, int32_t n_segments, int32_t oversample,
                   int32_t obj_flag, double* out, int64_t n_out) {
    if (!ctx) return CPB_E_INVALID;
    for (int32_t s = 0; s < n_segments; ++s)
        if (segments[s].curve != 0)
            return fail(ctx, CPB_E_INVALID, "single-curve search: segment.curve must be 0");
    cpb_curve c;
    c.point_offset = 0;
    c.n_points = n_points;
    c.t_min = t_min;
    c.ivar_floor = ivar_floor;
    c.sum_y = sum_y;
    c.sum_ivar = sum_ivar;
    return run_search(ctx, static_cast<hipStream_t>(stream), tau, yw, wv, n_points, &c, 1,
                      periods, dur_bins, segments, n_segments, oversample, obj_flag, out,
                      n_out);
}

int cpb_bls_search_batch(cpb_context* ctx, void* stream, const double* tau,
                         const double* yw, const double* wv, int64_t total_points,
                         const cpb_curve* curves, int32_t n_curves,
                         const double* periods, const int32_t* dur_bins,
                         const cpb_segment* segments, int32_t n_segments,
                         int32_t oversample, int32_t obj_flag, double* out,
                         int64_t n_out) {
    return run_search(ctx, static_cast<hipStream_t>(stream), tau, yw, wv, total_points,
                      curves, n_curves, periods, dur_bins, segments, n_segments,
                      oversample, obj_flag, out, n_out);
}

int cpb_bls_search_batch_static(cpb_context* ctx, void* stream, const double* tau,
                                const double* yw, const double* wv, int64_t total_points,
                                const cpb_curve* curves, int32_t n_curves,
                                const double* periods, const int32_t* dur_bins,
                                const cpb_segment* segments, int32_t n_segments,
                                int32_t oversample, int32_t obj_flag, double* out,
                                int64_t n_out) {
    if (!ctx) return CPB_E_INVALID;
    if (!ctx->static_scratch)
        return fail(ctx, CPB_E_DEVICE, "static scratch unavailable: " + ctx->static_error);
    ctx->carve_base = static_cast<char*>(ctx->static_scratch);
    ctx->carve_cap = kStaticScratchBytes;
    ctx->carve_used = 0;
    const int rc = run_search(ctx, static_cast<hipStream_t>(stream), tau, yw, wv,
                              total_points, curves, n_curves, periods, dur_bins, segments,
                              n_segments, oversample, obj_flag, out, n_out);
    ctx->carve_base = nullptr;
    ctx->carve_cap = ctx->carve_used = 0;
    return rc;
}

int cpb_bls_summarize(cpb_context* ctx, void* stream, const double* power,
                      const int64_t* curve_base, const int64_t* curve_count,
                      int32_t n_curves, double* summary) {
    if (!ctx) return CPB_E_INVALID;
    if (n_curves < 0 || (n_curves > 0 && (!power || !curve_base || !curve_count || !summary)))
        return fail(ctx, CPB_E_INVALID, "summarize: bad arguments");
    if (n_curves == 0) return CPB_OK;
    CPB_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(curve_summary_kernel, dim3(n_curves), dim3(256), 0,
                       static_cast<hipStream_t>(stream), power, curve_base, curve_count,
                       summary);
    CPB_HIP(ctx, hipGetLastError());
    return CPB_OK;
}

int cpb_peak_candidates(cpb_context* ctx, void* stream, const double* out, int64_t n_out,
                        const int64_t* curve_base, const int64_t* curve_count,
                        int32_t n_curves, int32_t want, int32_t cap, int32_t* cand,
                        int64_t* counts, double* values) {
    if (!ctx) return CPB_E_INVALID;
    if (n_curves < 0 || want < 1 || cap < want ||
        (n_curves > 0 && (!out || !curve_base || !curve_count || !cand || !counts || !values)))
        return fail(ctx, CPB_E_INVALID, "peak candidates: bad arguments");
    if (n_curves == 0) return CPB_OK;
    CPB_HIP(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(peak_candidates_kernel, dim3(n_curves), dim3(256), 0,
                       static_cast<hipStream_t>(stream), out, n_out, curve_base, curve_count,
                       want, cap, cand, counts, values);
    CPB_HIP(ctx, hipGetLastError());
    return CPB_OK;
}

int cpb_linspace_grid(double start, double stop, int64_t num, double* out) {
    if (num < 0 || (num > 0 && !out)) return CPB_E_INVALID;
    if (num == 0) return CPB_OK;
    if (num == 1) {
        out[0] = start;
        return CPB_OK;
    }
    const double step = (stop - start) / (double)(num - 1);
    for (int64_t i = 0; i < num - 1; ++i) out[i] = (double)i * step + start;
    out[num - 1] = stop;
    return CPB_OK;
}

namespace {

// Does a duplicate, harmonise or alias b (one direction of the alias rule).
bool alias_related(double a, double b, double tol, double separation, const double* orders,
                   int32_t n_orders, const double* offsets, int32_t n_offsets) {
    if (std::fabs(a - b) < separation) return true;
    for (int32_t k = 0; k < n_orders; ++k) {
        const double m = orders[k];
        if (std::fabs(a - m * b) < m * tol || std::fabs(a - b / m) < tol) return true;
    }
    for (int32_t k = 0; k < n_offsets; ++k) {
        const double o = offsets[k];
        if (std::fabs(a - std::fabs(b - o)) < tol || std::fabs(a - (b + o)) < tol) return true;
    }
    return false;
}

}  // namespace

int cpb_select_diverse_peaks(const double* freq, const int64_t* order, int64_t n,
                             int32_t n_peaks, double tol, double separation,
                             const double* orders, int32_t n_orders, const double* offsets,
                             int32_t n_offsets, int64_t* chosen, int32_t* n_chosen,
                             int64_t* examined) {
    if (n < 0 || n_peaks < 0 || n_orders < 0 || n_offsets < 0 || !n_chosen || !examined ||
        (n > 0 && (!freq || !order)) || (n_peaks > 0 && !chosen) ||
        (n_orders > 0 && !orders) || (n_offsets > 0 && !offsets))
        return CPB_E_INVALID;
    int32_t count = 0;
    int64_t visited = 0;
    for (int64_t i = 0; i < n && count < n_peaks; ++i) {
        ++visited;
        const double f = freq[order[i]];
        bool free = true;
        for (int32_t c = 0; c < count && free; ++c) {
            const double g = freq[chosen[c]];
            free = !alias_related(f, g, tol, separation, orders, n_orders, offsets, n_offsets) &&
                   !alias_related(g, f, tol, separation, orders, n_orders, offsets, n_offsets);
        }
        if (free) chosen[count++] = order[i];
    }
    *n_chosen = count;
    *examined = visited;
    return CPB_OK;
}

int cpb_spectrum_candidates(const double* power, int64_t n, int64_t* index, double* score,
                            int64_t* count) {
    if (n < 0 || !count || (n > 0 && (!power || !index || !score))) return CPB_E_INVALID;
    // sample i of the ascending-frequency spectrum, non-finite values read as 0
    auto s = [&](int64_t i) {
        const double v = power[n - 1 - i];
        return std::isfinite(v) ? v : 0.0;
    };
    int64_t m = 0;
    auto keep = [&](int64_t i, double v) {
        index[m] = i;
        score[m] = v;
        ++m;
    };
    if (n >= 3) {
        // interior local maxima, ascending: above the left neighbour, at least the right
        double left = s(0), mid = s(1);
        for (int64_t i = 1; i + 1 < n; ++i) {
            const double right = s(i + 1);
            if (mid > left && mid >= right) keep(i, mid);
            left = mid;
            mid = right;
        }
    }
    if (n == 1) {
        keep(0, s(0));
    } else if (n >= 2) {
        if (s(0) >= s(1)) keep(0, s(0));
        if (s(n - 1) > s(n - 2)) keep(n - 1, s(n - 1));
    }
    if (m == 0 && n > 0) {
        // no candidate: the first largest sample
        int64_t best = 0;
        for (int64_t i = 1; i < n; ++i)
            if (s(i) > s(best)) best = i;
        keep(best, s(best));
    }
    *count = m;
    return CPB_OK;
}

int cpb_set_timing(cpb_context* ctx, int enable) {
    if (!ctx) return CPB_E_INVALID;
    CPB_HIP(ctx, hipSetDevice(ctx->device));
    if (enable && !ctx->t0) {
        CPB_HIP(ctx, hipEventCreate(&ctx->t0));
        CPB_HIP(ctx, hipEventCreate(&ctx->t1));
    }
    ctx->timing = enable != 0;
    ctx->timed = false;
    return CPB_OK;
}

int cpb_main_kernel_ms(cpb_context* ctx, double* ms) {
    if (!ctx || !ms) return CPB_E_INVALID;
    if (!ctx->timed) return fail(ctx, CPB_E_INVALID, "no timed launch");
    CPB_HIP(ctx, hipSetDevice(ctx->device));
    CPB_HIP(ctx, hipEventSynchronize(ctx->t1));
    float f = 0.0f;
    CPB_HIP(ctx, hipEventElapsedTime(&f, ctx->t0, ctx->t1));
    *ms = (double)f;
    return CPB_OK;
}

int cpb_set_launch(cpb_context* ctx, int threads, int tile, int grid_per_cu) {
    if (!ctx) return CPB_E_INVALID;
    if (threads != 128 && threads != 256 && threads != 512)
        return fail(ctx, CPB_E_INVALID, "threads must be 128, 256 or 512");
    if (tile < 0 || tile > kMaxTile) return fail(ctx, CPB_E_INVALID, "tile must be in [0, 32]");
    if (grid_per_cu < 0 || grid_per_cu > 64)
        return fail(ctx, CPB_E_INVALID, "grid_per_cu must be in [0, 64]");
    ctx->threads = threads;
    ctx->tile = tile;
    ctx->grid_per_cu = grid_per_cu;
    return CPB_OK;
}

struct cpb_graph {
    cpb_context* priv = nullptr;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

void cpb_graph_destroy(cpb_graph* g) {
    if (!g) return;
    if (g->priv) (void)hipSetDevice(g->priv->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->cap) (void)hipStreamDestroy(g->cap);
    if (g->priv) cpb_destroy(g->priv);
    delete g;
}

int cpb_graph_create(cpb_context* ctx, const double* tau, const double* yw, const double* wv,
                     int64_t total_points, const cpb_curve* curves, int32_t n_curves,
                     const double* periods, const int32_t* dur_bins,
                     const cpb_segment* segments, int32_t n_segments, int32_t oversample,
                     int32_t obj_flag, double* out, int64_t n_out, const int64_t* curve_base,
                     const int64_t* curve_count, double* summary, cpb_graph** graph) {
    if (!ctx || !graph) return CPB_E_INVALID;
    *graph = nullptr;
    auto* g = new cpb_graph();
    int rc = cpb_create(ctx->device, &g->priv);
    if (rc != CPB_OK) {
        delete g;
        return fail(ctx, rc, "graph: cannot create its private context");
    }
    g->priv->accumulation = ctx->accumulation;
    g->priv->threads = ctx->threads;
    g->priv->tile = ctx->tile;
    g->priv->grid_per_cu = ctx->grid_per_cu;
    g->priv->peak_rounding = ctx->peak_rounding;
    g->priv->capturing = true;
    auto bail = [&](int status, const std::string& msg) {
        std::string m = msg;
        if (g->priv && !g->priv->error.empty()) m += ": " + g->priv->error;
        cpb_graph_destroy(g);
        return fail(ctx, status, m);
    };
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(CPB_E_DEVICE, std::string("graph stream: ") + hipGetErrorString(e));
    e = hipStreamBeginCapture(g->cap, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return bail(CPB_E_DEVICE, std::string("begin capture: ") + hipGetErrorString(e));
    rc = run_search(g->priv, g->cap, tau, yw, wv, total_points, curves, n_curves, periods,
                    dur_bins, segments, n_segments, oversample, obj_flag, out, n_out);
    if (rc == CPB_OK && summary) {
        if (!curve_base || !curve_count) {
            rc = CPB_E_INVALID;
        } else {
            hipLaunchKernelGGL(curve_summary_kernel, dim3(n_curves), dim3(256), 0, g->cap, out,
                               curve_base, curve_count, summary);
            if (hipGetLastError() != hipSuccess) rc = CPB_E_DEVICE;
        }
    }
    hipGraph_t captured = nullptr;
    e = hipStreamEndCapture(g->cap, &captured);
    g->graph = captured;
    if (rc != CPB_OK) return bail(rc, "graph: capture of the search failed");
    if (e != hipSuccess || !captured)
        return bail(CPB_E_DEVICE, std::string("end capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return bail(CPB_E_DEVICE, std::string("instantiate: ") + hipGetErrorString(e));
    g->priv->capturing = false;
    *graph = g;
    return CPB_OK;
}

int cpb_graph_launch(cpb_context* ctx, cpb_graph* g, void* stream) {
    if (!ctx || !g || !g->exec) return CPB_E_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    CPB_HIP(ctx, hipSetDevice(ctx->device));
    if (ctx->timing) CPB_HIP(ctx, hipEventRecord(ctx->t0, st));
    CPB_HIP(ctx, hipGraphLaunch(g->exec, st));
    if (ctx->timing) {
        CPB_HIP(ctx, hipEventRecord(ctx->t1, st));
        ctx->timed = true;
    }
    return CPB_OK;
}

int cpb_set_peak_rounding(cpb_context* ctx, int enable) {
    if (!ctx) return CPB_E_INVALID;
    if (enable != 0 && enable != 1) return fail(ctx, CPB_E_INVALID, "peak rounding must be 0 or 1");
    ctx->peak_rounding = enable != 0;
    return CPB_OK;
}

int cpb_set_accumulation(cpb_context* ctx, int mode) {
    if (!ctx) return CPB_E_INVALID;
    if (mode < 0 || mode > 2) return fail(ctx, CPB_E_INVALID, "accumulation mode must be 0, 1 or 2");
    ctx->accumulation = mode;
    return CPB_OK;
}

int cpb_synchronize(cpb_context* ctx, void* stream) {
    if (!ctx) return CPB_E_INVALID;
    CPB_HIP(ctx, hipSetDevice(ctx->device));
    CPB_HIP(ctx, hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return CPB_OK;
}

}  // extern "C"
