// MI355X (gfx950) Transit Least Squares matched filter (cuPeriod
// src/cuperiod/methods/tls.py:1-26; kernels :270-355 and :491-595).
//
// One workgroup takes one (light curve, tile of trial periods) work item and, within it,
// G trial periods at a time (G = 8 while their bins fit a quarter of the LDS):
//
//   (a) fold    every point's phase bin from the reference's FP64 expressions
//               q = tau/P, ph = q - floor(q), bin = int(ph*n_bins) (clipped), bit for
//               bit: q is formed through the reciprocal of P and, where it or ph*n_bins
//               lies too close to an integer for that to decide the bin, by the
//               correctly rounded division. Each point is loaded once for the G
//               periods. Its weighted residual and weight enter LDS bins as 64-bit
//               fixed-point integers: the points of a wave are 64 consecutive times, so
//               runs of equal bins are common (short periods aside); the wave's
//               inclusive prefix sums (the same for all G periods) give each run's sum as
//               a difference, added by the run's last lane, less the prefix before it,
//               subtracted by its first lane (ds_add_u64; exact and independent of the
//               atomics' order, as integer sums modulo 2^64);
//   (b) bins    the integer bins rounded once to FP64, with one circular pad of
//               max_width-1 bins;
//   (c) sweep   every (template width, start) window: num = sum g_k a, cwg = sum g_k b,
//               den = sum g_k^2 b accumulated k = 0, 1, ... in FP64 as the reference
//               does; a dip (num < 0, den > eps) whose level-corrected energy
//               D = den - cwg^2/W exceeds den_rel*den scores SR = num^2/D; the first
//               strict maximum in (width index, start) order wins;
//   (d) final   SR, depth -num/D, duration width/n_bins*P and the mid-transit phase time
//               frac((start + width/2)/n_bins)*P, written to HBM.
//
// Compiled with -ffp-contract=off: the fold and the correlation round every product
// and sum as the FP64 reference does.

#include "engine_internal.h"

using namespace cpb_engine;

namespace {

constexpr double kWEps = 1e-300;
constexpr int kTlsThreads = 256;
constexpr int kTlsMaxBins = 4096;
constexpr int kTlsMaxWidths = 32;

struct TlsItem {
    int32_t curve;
    int32_t count;
    int64_t begin;
};

__device__ __forceinline__ bool tls_better(double a, uint32_t ka, double b, uint32_t kb) {
    return a > b || (a == b && ka < kb);
}

// Adds the wave's 64 consecutive points to the bins of one period: `bb` the point's bin
// (-1 for a lane past the curve's end), (a, b) its values and (pa, pb) the wave's
// inclusive prefix sums of them.
__device__ __forceinline__ void fold_runs(int64_t* bins, int bb, int64_t a, int64_t b,
                                          int64_t pa, int64_t pb) {
    // the neighbours' bins by wavefront shifts (DPP wave_shr:1 / wave_shl:1); lane 0 and
    // lane 63 get INT_MIN, which no bin equals, so they open and close a run
    const int prev = __builtin_amdgcn_update_dpp(INT_MIN, bb, 0x138, 0xF, 0xF, false);
    const int next = __builtin_amdgcn_update_dpp(INT_MIN, bb, 0x130, 0xF, 0xF, false);
    const bool head = bb != prev;
    const bool tail = bb != next;
    if (bb >= 0 && (head || tail)) {
        using u64 = unsigned long long;
        const u64 da = (tail ? (u64)pa : 0ull) - (head ? (u64)pa - (u64)a : 0ull);
        const u64 db = (tail ? (u64)pb : 0ull) - (head ? (u64)pb - (u64)b : 0ull);
        atomicAdd(reinterpret_cast<u64*>(bins + 2 * bb), da);
        atomicAdd(reinterpret_cast<u64*>(bins + 2 * bb + 1), db);
    }
}

template <int G>
__global__ __launch_bounds__(kTlsThreads) void tls_kernel(
    const double* __restrict__ tau, const int64_t* __restrict__ yq,
    const int64_t* __restrict__ wq, const CurveParams* __restrict__ cp,
    const cpb_tls_curve* __restrict__ curves, const TlsItem* __restrict__ items,
    int64_t n_items, const double* __restrict__ periods, int32_t n_bins,
    const int32_t* __restrict__ widths, const int32_t* __restrict__ tmpl_off,
    const double* __restrict__ tmpl, int32_t n_widths, int32_t max_width, int32_t tmpl_len,
    double den_rel, double* __restrict__ out, int64_t n_out) {
    extern __shared__ int64_t dyn_lds[];
    constexpr int T = kTlsThreads;
    constexpr int NW = T / kWave;
    __shared__ int32_t s_w[kTlsMaxWidths];
    __shared__ int32_t s_off[kTlsMaxWidths];
    __shared__ double s_red_v[G][NW];
    __shared__ double s_red_d[G][NW];
    __shared__ uint32_t s_red_k[G][NW];

    const int tid = threadIdx.x;
    const int lane = tid % kWave;
    const int wave = tid / kWave;
    const int n_ext = n_bins + max_width - 1;
    const double dn = (double)n_bins;
    // LDS: G arrays of n_ext (a, b) pairs, int64 during the fold and FP64 in place after
    // it, then the templates.
    int64_t* P2 = dyn_lds;
    double* s_tmpl = reinterpret_cast<double*>(dyn_lds + 2 * (int64_t)G * n_ext);

    if (tid < n_widths) {
        s_w[tid] = widths[tid];
        s_off[tid] = tmpl_off[tid];
    }
    for (int i = tid; i < tmpl_len; i += T) s_tmpl[i] = tmpl[i];
    __syncthreads();

    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const TlsItem it = items[item];
        const cpb_tls_curve cv = curves[it.curve];
        const CurveParams pr = cp[it.curve];
        const int64_t poff = cv.point_offset;
        const int64_t npts = cv.n_points;
        const double W = cv.w_total;

        for (int q0 = 0; q0 < it.count; q0 += G) {
            const int ng = min(G, it.count - q0);
            double Pg[G], invPg[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                Pg[g] = g < ng ? periods[cv.period_offset + it.begin + q0 + g] : 1.0;
                invPg[g] = 1.0 / Pg[g];
            }

            // (a) zero the bins, fold every point into each period's bins.
            longlong2* B2 = reinterpret_cast<longlong2*>(P2);
            __syncthreads();
            for (int i = tid; i < ng * n_ext; i += T) B2[i] = make_longlong2(0, 0);
            __syncthreads();
            for (int64_t jb = 0; jb < npts; jb += T) {
                const int64_t j = jb + tid;
                const bool valid = j < npts;
                const double x = valid ? tau[poff + j] : 0.0;
                const int64_t a = valid ? yq[poff + j] : 0;
                const int64_t b = valid ? wq[poff + j] : 0;
                const int64_t pa = wave_incl_scan_dpp(a);
                const int64_t pb = wave_incl_scan_dpp(b);
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    if (g < ng) {
                        // q within 2^-51.4 |q| of tau/P through the reciprocal; where
                        // ph*n_bins lies within (2^-49 |q| + 2^-50) n_bins of an integer
                        // (q near an integer puts it near 0 or n_bins), the division
                        // decides instead
                        double qq = x * invPg[g];
                        double ph = qq - floor(qq);
                        double r = ph * dn;
                        const double tq = fabs(qq) * 0x1p-49;
                        if (fabs(r - rint(r)) <= (tq + 0x1p-50) * dn) {
                            qq = x / Pg[g];
                            ph = qq - floor(qq);
                            r = ph * dn;
                        }
                        int bb = (int)r;
                        bb = bb > n_bins - 1 ? n_bins - 1 : (bb < 0 ? 0 : bb);
                        fold_runs(P2 + 2 * (int64_t)g * n_ext, valid ? bb : -1, a, b, pa, pb);
                    }
                }
            }
            __syncthreads();
            // (b) FP64 bins in place, then the circular pads.
            double2* D2 = reinterpret_cast<double2*>(P2);
            for (int i = tid; i < ng * n_bins; i += T) {
                const int g = i / n_bins;
                const int k = g * n_ext + (i - g * n_bins);
                const longlong2 v = B2[k];
                D2[k] = make_double2(ldexp((double)v.x, pr.y_exp), ldexp((double)v.y, pr.w_exp));
            }
            __syncthreads();
            const int pad = max_width - 1;
            for (int i = tid; i < ng * pad; i += T) {
                const int g = i / pad;
                const int k = i - g * pad;
                D2[g * n_ext + n_bins + k] = D2[g * n_ext + k];
            }
            __syncthreads();

            // (c) sweep every (width, start) window of every period.
            const int n_win = n_widths * n_bins;
            double best[G], best_depth[G];
            uint32_t bkey[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                best[g] = 0.0;
                best_depth[g] = 0.0;
                bkey[g] = 0xFFFFFFFFu;
                if (g < ng) {
                    const double2* Dg = D2 + (int64_t)g * n_ext;
                    for (int idx = tid; idx < n_win; idx += T) {
                        const int di = idx / n_bins;
                        const int s = idx - di * n_bins;
                        const int width = s_w[di];
                        const double* tg = s_tmpl + s_off[di];
                        const double g0 = tg[0];
                        const double2 v0 = Dg[s];
                        double num = g0 * v0.x;
                        double cwg = g0 * v0.y;
                        double den = (g0 * g0) * v0.y;
                        for (int k = 1; k < width; ++k) {
                            const double gk = tg[k];
                            const double gkk = gk * gk;
                            const double2 v = Dg[s + k];
                            num += gk * v.x;
                            cwg += gk * v.y;
                            den += gkk * v.y;
                        }
                        if (num < 0.0 && den > kWEps) {
                            const double de = den - cwg * cwg / W;
                            if (de > den_rel * den) {
                                const double sr = num * num / de;
                                const uint32_t key = (uint32_t)idx;
                                if (sr > 0.0 && tls_better(sr, key, best[g], bkey[g])) {
                                    best[g] = sr;
                                    bkey[g] = key;
                                    best_depth[g] = -num / de;
                                }
                            }
                        }
                    }
                }
            }
            // first-wins maximum over the workgroup, per period
#pragma unroll
            for (int g = 0; g < G; ++g) {
#pragma unroll
                for (int off = kWave / 2; off > 0; off >>= 1) {
                    const double ov = __shfl_xor(best[g], off, kWave);
                    const double od = __shfl_xor(best_depth[g], off, kWave);
                    const uint32_t ok = (uint32_t)__shfl_xor((int)bkey[g], off, kWave);
                    if (tls_better(ov, ok, best[g], bkey[g])) {
                        best[g] = ov;
                        bkey[g] = ok;
                        best_depth[g] = od;
                    }
                }
                if (lane == 0) {
                    s_red_v[g][wave] = best[g];
                    s_red_d[g][wave] = best_depth[g];
                    s_red_k[g][wave] = bkey[g];
                }
            }
            __syncthreads();
            if (tid < ng) {
                const int g = tid;
                double bv = s_red_v[g][0], d = s_red_d[g][0];
                uint32_t k = s_red_k[g][0];
                for (int w = 1; w < NW; ++w)
                    if (tls_better(s_red_v[g][w], s_red_k[g][w], bv, k)) {
                        bv = s_red_v[g][w];
                        d = s_red_d[g][w];
                        k = s_red_k[g][w];
                    }
                int start = 0, width = 0;
                if (bv > 0.0) {
                    const int di = (int)(k / (uint32_t)n_bins);
                    start = (int)(k - (uint32_t)di * (uint32_t)n_bins);
                    width = s_w[di];
                } else {
                    d = 0.0;
                }
                const double P = periods[cv.period_offset + it.begin + q0 + g];
                const int64_t col = cv.out_offset + it.begin + q0 + g;
                const double dur = (double)width / (double)n_bins * P;
                const double centre = ((double)start + (double)width / 2.0) / (double)n_bins;
                out[0 * n_out + col] = bv;
                out[1 * n_out + col] = d;
                out[2 * n_out + col] = dur;
                out[3 * n_out + col] = (centre - floor(centre)) * P;
            }
        }
    }
}

}  // namespace

extern "C" int cpb_tls_search(cpb_context* ctx, void* stream, const double* tau,
                              const double* yw, const double* wv, int64_t total_points,
                              const cpb_tls_curve* curves, int32_t n_curves,
                              const double* periods, int32_t n_bins, const int32_t* widths,
                              int32_t n_widths, const double* templates, double den_rel,
                              double* out, int64_t n_out) {
    if (!ctx) return CPB_E_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n_curves < 1 || total_points < 0 || n_out < 0)
        return fail(ctx, CPB_E_INVALID, "negative or empty size");
    if (!curves || !widths || !templates) return fail(ctx, CPB_E_INVALID, "null argument");
    if (n_bins < 2 || n_bins > kTlsMaxBins)
        return fail(ctx, CPB_E_RESOURCE, "n_bins must be in [2, 4096]");
    if (n_widths < 1 || n_widths > kTlsMaxWidths)
        return fail(ctx, CPB_E_RESOURCE, "between 1 and 32 template widths");
    int max_width = 1, tmpl_len = 0;
    std::vector<int32_t> offs(n_widths);
    for (int i = 0; i < n_widths; ++i) {
        if (widths[i] < 1 || widths[i] > n_bins)
            return fail(ctx, CPB_E_INVALID, "template width outside [1, n_bins]");
        offs[i] = tmpl_len;
        tmpl_len += widths[i];
        max_width = std::max(max_width, (int)widths[i]);
    }
    int64_t live = 0;
    for (int32_t c = 0; c < n_curves; ++c) {
        const cpb_tls_curve& cv = curves[c];
        if (cv.point_offset < 0 || cv.n_points < 0 || cv.point_offset + cv.n_points > total_points)
            return fail(ctx, CPB_E_INVALID, "curve points out of range");
        if (cv.n_periods < 0 || cv.period_offset < 0 || cv.out_offset < 0 ||
            cv.out_offset + cv.n_periods > n_out)
            return fail(ctx, CPB_E_INVALID, "curve output range out of bounds");
        if (cv.n_periods > 0 && !periods) return fail(ctx, CPB_E_INVALID, "periods needed");
        if (!(cv.w_total > 0.0)) return fail(ctx, CPB_E_INVALID, "w_total must be positive");
        live += cv.n_periods;
    }
    const size_t per_period = (size_t)16 * (n_bins + max_width - 1);
    const size_t tmpl_bytes = (size_t)8 * tmpl_len;
    if (per_period + tmpl_bytes + 1024 > (size_t)ctx->lds_limit)
        return fail(ctx, CPB_E_RESOURCE, "phase bins and templates exceed the LDS");
    // periods per pass: 8 or 4 (2) while their bins and the templates stay within a
    // quarter (a half) of the LDS, so that several workgroups share a CU
    const int n_pass = per_period * 8 + tmpl_bytes <= (size_t)ctx->lds_limit / 4   ? 8
                       : per_period * 4 + tmpl_bytes <= (size_t)ctx->lds_limit / 4   ? 4
                       : per_period * 2 + tmpl_bytes <= (size_t)ctx->lds_limit / 2 ? 2
                                                                                   : 1;
    const size_t lds = per_period * n_pass + tmpl_bytes;

    CPB_HIP(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure(ctx, ctx->curves, sizeof(CurveParams) * n_curves))) return rc;
    if ((rc = ensure(ctx, ctx->yq, sizeof(int64_t) * std::max<int64_t>(1, total_points)))) return rc;
    if ((rc = ensure(ctx, ctx->wq, sizeof(int64_t) * std::max<int64_t>(1, total_points)))) return rc;

    const int64_t target = (int64_t)ctx->n_cu * 64;
    int64_t tile = std::max<int64_t>(1, std::min<int64_t>(live / std::max<int64_t>(1, target), 32));
    if (tile > n_pass) tile -= tile % n_pass;
    std::vector<TlsItem> items;
    for (int32_t c = 0; c < n_curves; ++c)
        for (int64_t b = 0; b < curves[c].n_periods; b += tile)
            items.push_back({c, (int32_t)std::min<int64_t>(tile, curves[c].n_periods - b), b});
    const int64_t n_items = (int64_t)items.size();

    // Tables: [prep curves | tls curves | items | widths | offsets | templates].
    std::vector<cpb_curve> prep(n_curves);
    for (int32_t c = 0; c < n_curves; ++c)
        prep[c] = {curves[c].point_offset, curves[c].n_points, 0.0, 0.0};
    auto align = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t o_tc = align(sizeof(cpb_curve) * n_curves);
    const size_t o_it = o_tc + align(sizeof(cpb_tls_curve) * n_curves);
    const size_t o_w = o_it + align(sizeof(TlsItem) * std::max<int64_t>(1, n_items));
    const size_t o_off = o_w + align(sizeof(int32_t) * n_widths);
    const size_t o_tm = o_off + align(sizeof(int32_t) * n_widths);
    const size_t tbytes = o_tm + align(sizeof(double) * tmpl_len);
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
    std::memcpy(hb, prep.data(), sizeof(cpb_curve) * n_curves);
    std::memcpy(hb + o_tc, curves, sizeof(cpb_tls_curve) * n_curves);
    if (n_items) std::memcpy(hb + o_it, items.data(), sizeof(TlsItem) * n_items);
    std::memcpy(hb + o_w, widths, sizeof(int32_t) * n_widths);
    std::memcpy(hb + o_off, offs.data(), sizeof(int32_t) * n_widths);
    std::memcpy(hb + o_tm, templates, sizeof(double) * tmpl_len);
    CPB_HIP(ctx, hipMemcpyAsync(ctx->tables.ptr, hb, tbytes, hipMemcpyHostToDevice, st));
    if (!ctx->capturing) {
        CPB_HIP(ctx, hipEventRecord(ctx->copied, st));
        ctx->copy_pending = true;
    }
    char* db = static_cast<char*>(ctx->tables.ptr);
    auto* params = static_cast<CurveParams*>(ctx->curves.ptr);
    auto* yq = static_cast<int64_t*>(ctx->yq.ptr);
    auto* wq = static_cast<int64_t*>(ctx->wq.ptr);
    hipError_t e = launch_prep_curves(st, nullptr, yw, wv, reinterpret_cast<const cpb_curve*>(db),
                                      n_curves, 0, params, yq, wq);
    if (e != hipSuccess) return hip_fail(ctx, e, "prep_curves_kernel launch");
    if (n_items == 0) return CPB_OK;

    const dim3 grid((unsigned)std::min<int64_t>(n_items, ((int64_t)1 << 31) / kTlsThreads));
    const bool timed = ctx->timing && !ctx->capturing;
    if (timed) CPB_HIP(ctx, hipEventRecord(ctx->t0, st));
    auto launch = [&](auto kernel) {
        if (lds > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(kernel, grid, dim3(kTlsThreads), lds, st, tau, yq, wq, params,
                           reinterpret_cast<const cpb_tls_curve*>(db + o_tc),
                           reinterpret_cast<const TlsItem*>(db + o_it), n_items, periods,
                           n_bins, reinterpret_cast<const int32_t*>(db + o_w),
                           reinterpret_cast<const int32_t*>(db + o_off),
                           reinterpret_cast<const double*>(db + o_tm), n_widths, max_width,
                           tmpl_len, den_rel, out, n_out);
    };
    if (n_pass == 8)
        launch(tls_kernel<8>);
    else if (n_pass == 4)
        launch(tls_kernel<4>);
    else if (n_pass == 2)
        launch(tls_kernel<2>);
    else
        launch(tls_kernel<1>);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "tls_kernel launch");
    if (timed) {
        CPB_HIP(ctx, hipEventRecord(ctx->t1, st));
        ctx->timed = true;
    }
    return CPB_OK;
}
