// Private declarations shared by the engine's HIP sources (the public C ABI is
// include/cuperiod_bls.h): fixed-point constants, per-curve parameters, work items,
// device helpers (cross-lane moves, the exact phase fold, parametric periods), the
// context and its device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "cuperiod_bls.h"

namespace cpb_engine {


constexpr int kWave = 64;
// Fixed point: a curve's sum of |y*ivar| (and of ivar) maps to at most 2^59 units, so
// prefix sums that also count the `oversample` wrapped bins twice stay below 2^61 and
// their high words (value >> 32) below 2^29 in magnitude.
constexpr int kFixBits = 59;
// Screen margin: boxes whose FP32 value reaches 0.97 of the best FP32 value are
// re-evaluated exactly. When the best screen value is above the curve's `vdeg` bound,
// the boxes that matter have |t| at least 256 times its error bound and a denominator
// within 2/e_unc <= 2/511 of exact, so a screen value lies within [0.9883, 1.0118] of
// its box's exact objective and the best box reaches at least 0.9767 of the best screen
// value; below `vdeg` every box is evaluated exactly.
constexpr float kKeep = 0.97f;
// Distinct box widths (duration bins) per segment.
constexpr int kMaxDurations = 32;

struct CurveParams {
    int64_t point_offset;
    int64_t y_tot;    // sum of fixed-point y*ivar
    int64_t w_tot;    // sum of fixed-point ivar
    int32_t y_exp;    // one fixed-point unit of y*ivar is 2^y_exp (ey - 59)
    int32_t w_exp;    // same for ivar
    double t_min;
    double ivar_floor;
    float ybar;       // y_tot / w_tot
    float vdeg;       // below: the screen's error may exceed the margin -> exact sweep
    float vbnd;       // bound on near-empty / near-full boxes the screen leaves out
    int32_t n_points;
    int32_t s_hi;     // w_tot >> 32
    int32_t e_unc;    // high-word weight below which a box side is "uncertain"
    int32_t flags;    // bit 0: every box exactly; bit 1: non-finite input (no box);
                      // bit 2: mean_units(wh, ybar_d) is 0 for every high word
    double ybar_d;    // y_tot / w_tot in FP64 (the screen's prefix encoding)
    double x_max;     // largest time offset of the curve (infinity when not known)
    float dmin;       // snr screen: W*Wout above this means both sides hold > e_unc
};

struct WorkItem {
    int32_t seg;
    int32_t count;
    int64_t begin;
};

// ---------------------------------------------------------------------------------
// Device helpers

__device__ __forceinline__ int64_t shfl_i64(int64_t v, int src) {
    const uint32_t lo = __shfl((uint32_t)(uint64_t)v, src, kWave);
    const uint32_t hi = __shfl((uint32_t)((uint64_t)v >> 32), src, kWave);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)(uint64_t)v, off, kWave);
        const uint32_t hi = __shfl_xor((uint32_t)((uint64_t)v >> 32), off, kWave);
        v += (int64_t)(((uint64_t)hi << 32) | lo);
    }
    return v;
}

// Phase bin of one point: `int(w/bd) + 1` with `w = x - P*floor(x/P)`, clipped to
// [0, width-1], bit for bit as the FP64 divisions give it. The quotients are formed
// with reciprocals (|error| <= 2^-52 |q|); when a quotient lies within 2^-49 |q| of
// the integer that decides floor/trunc, the exact division is done instead.
__device__ __forceinline__ int fold_bin(double x, double P, double invP, double bd,
                                        double inv_bd, int width) {
    const double q = x * invP;
    double k = floor(q);
    const double fr = q - k;
    const double tol = q * 0x1p-49;
    if ((fr < tol && k >= 1.0) || (1.0 - fr) <= tol) k = floor(x / P);
    const double w = x - P * k;
    const double r = w * inv_bd;
    double j = trunc(r);
    const double rn = rint(r);
    if (rn != 0.0 && fabs(r - rn) <= fabs(r) * 0x1p-49) j = trunc(w / bd);
    int ind = (int)j + 1;
    ind = ind < 0 ? 0 : ind;
    ind = ind > width - 1 ? width - 1 : ind;
    return ind;
}

// fold_bin with the two near-integer tests against fixed tolerances of one period:
// `tol_q` >= 2^-49 max(x/P) and `tol_r` >= 2^-49 max(w/bd). A reciprocal quotient is
// within 2^-51 |q| of the FP64 division's, so floor (trunc) of the two can differ only
// when an integer lies within that distance; the test sends those points, and a few
// more, to the exact division. The result is fold_bin's.
__device__ __forceinline__ int fold_bin_tol(double x, double P, double invP, double bd,
                                            double inv_bd, int width, double tol_q,
                                            double tol_r) {
    const double q = x * invP;
    double k = floor(q);
    if (fabs(q - rint(q)) <= tol_q) k = floor(x / P);
    const double w = x - P * k;
    const double r = w * inv_bd;
    int j = (int)r;
    if (fabs(r - rint(r)) <= tol_r) j = (int)trunc(w / bd);
    return min(max(j + 1, 0), width - 1);
}

// The part of a prefix's high word that the curve's mean accounts for: rint(ybar * wh).
// |ybar| < 2 and |wh| < 2^29, so the product is within 2^-23 of exact and the result
// fits an int32; the same expression encodes and decodes a prefix, so both agree.
__device__ __forceinline__ int32_t mean_units(int32_t wh, double ybar) {
    return (int32_t)rint(ybar * (double)wh);
}

__device__ __forceinline__ double segment_period(const cpb_segment& sg,
                                                 const double* __restrict__ periods,
                                                 int64_t k) {
    if (sg.mode == 0) return periods[sg.period_offset + k];
    const int64_t i = sg.lin_num - 1 - (sg.lin_k0 + k);
    double f;
    if (sg.mode == 1 && i == sg.lin_num - 1)
        f = sg.lin_stop;
    else if (sg.mode == 2 && i == 0)
        f = sg.lin_start;
    else
        f = (double)i * sg.lin_step + sg.lin_start;
    return 1.0 / f;
}

__device__ __forceinline__ int64_t join64(int32_t hi, uint32_t lo) {
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)lo);
}

// Cross-lane moves within a wave (GFX9 DPP): row_shr:n shifts within 16-lane rows,
// row_bcast:15/31 carry a row's last lane into the next row(s); lanes without a
// source read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, true);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int64_t dpp_i64(int64_t v) {
    const uint32_t lo = dpp_u32<CTRL, ROW_MASK>((uint32_t)(uint64_t)v);
    const uint32_t hi = dpp_u32<CTRL, ROW_MASK>((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// One step of a 64-bit DPP scan as 32-bit halves: the low words added with a carry out,
// the high words with the carry in (an add and an add-with-carry, each taking its
// shifted operand straight from the DPP move).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void dpp_add_i64(uint32_t& lo, uint32_t& hi) {
    const uint32_t dlo = dpp_u32<CTRL, ROW_MASK>(lo);
    const uint32_t dhi = dpp_u32<CTRL, ROW_MASK>(hi);
    const uint32_t s = lo + dlo;
    hi = hi + dhi + (uint32_t)(s < dlo);
    lo = s;
}

// Inclusive prefix sum over the 64 lanes of a wave (all lanes active).
__device__ __forceinline__ int64_t wave_incl_scan_dpp(int64_t v) {
    uint32_t lo = (uint32_t)(uint64_t)v, hi = (uint32_t)((uint64_t)v >> 32);
    dpp_add_i64<0x111, 0xF>(lo, hi);
    dpp_add_i64<0x112, 0xF>(lo, hi);
    dpp_add_i64<0x114, 0xF>(lo, hi);
    dpp_add_i64<0x118, 0xF>(lo, hi);
    dpp_add_i64<0x142, 0xA>(lo, hi);
    dpp_add_i64<0x143, 0xC>(lo, hi);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Maximum over the wave of non-negative int32 values, returned uniformly.
__device__ __forceinline__ int wave_max_nonneg_dpp(int v) {
    v = max(v, (int)dpp_u32<0x111, 0xF>((uint32_t)v));
    v = max(v, (int)dpp_u32<0x112, 0xF>((uint32_t)v));
    v = max(v, (int)dpp_u32<0x114, 0xF>((uint32_t)v));
    v = max(v, (int)dpp_u32<0x118, 0xF>((uint32_t)v));
    v = max(v, (int)dpp_u32<0x142, 0xA>((uint32_t)v));
    v = max(v, (int)dpp_u32<0x143, 0xC>((uint32_t)v));
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace cpb_engine

struct cpb_context {
    int device = 0;
    std::string error;
    int lds_limit = 65536;
    int lds_per_cu = 163840;
    int n_cu = 1;
    // `tables` holds the curve, segment and work-item descriptors of one call; they
    // are staged in pinned host memory (`h_tables`) and sent in one copy. The event
    // marks that copy: the staging buffer is not rewritten before it has completed.
    cpb_engine::DevBuf tables, curves, yq, wq, wide;
    void* h_tables = nullptr;
    size_t h_tables_bytes = 0;
    hipEvent_t copied = nullptr;
    bool copy_pending = false;
    // Optional timing of the main kernel (cpb_set_timing / cpb_main_kernel_ms).
    bool timing = false;
    bool timed = false;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::vector<cpb_engine::WorkItem> h_items;
    // Bin accumulation: 0 exact 64-bit fixed point, 1 compensated FP32, 2 plain FP32.
    int accumulation = 0;
    // Launch shape (cpb_set_launch): workgroup size, periods per work item (0: chosen
    // from the batch), workgroups per CU of a persistent grid (0: one per work item).
    int threads = 256;
    int tile = 0;
    int grid_per_cu = 0;
    // Set while a graph captures a search: no event is recorded on the stream.
    bool capturing = false;
    // Joint multi-band search (cpb_set_joint_kernel): 0 automatic, 1 every band of a
    // period at once, 2 band groups with register-resident box totals.
    int joint_kernel = 0;
    // Reference rounding of each curve's near-maximum periods (cpb_set_peak_rounding):
    // `peak` holds the per-curve candidate lists and counts, `peak_wide` the bins of
    // periods too wide for LDS.
    bool peak_rounding = true;
    cpb_engine::DevBuf peak, peak_wide;
    // While cpb_bls_search_batch_static runs: device buffers are carved from the
    // library's static scratch array at `carve_base` (carve_cap bytes) instead of being
    // allocated, and descriptors are staged in pageable host memory (`h_stage`).
    char* carve_base = nullptr;
    size_t carve_cap = 0, carve_used = 0;
    std::vector<char> h_stage;
    // This device's copy of the static scratch array (hipGetSymbolAddress at creation;
    // NULL, with the reason in `static_error`, when it could not be resolved).
    void* static_scratch = nullptr;
    std::string static_error;
};

namespace cpb_engine {

inline int fail(cpb_context* ctx, int status, const std::string& msg) {
    if (ctx) ctx->error = msg;
    return status;
}

inline int hip_fail(cpb_context* ctx, hipError_t e, const char* what) {
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, CPB_E_DEVICE, msg);
}

#define CPB_HIP(ctx, expr)                                          \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return hip_fail((ctx), _e, #expr);    \
    } while (0)

inline int ensure(cpb_context* ctx, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes && b.ptr) return CPB_OK;
    if (b.ptr) {
        hipError_t e = hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
    }
    size_t want = std::max<size_t>(bytes, 256);
    want = want + want / 4;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    b.ptr = p;
    b.bytes = want;
    return CPB_OK;
}

// A device buffer of at least `bytes` for this call: carved from the static scratch
// (256-byte aligned) while a static search runs, else `b`, grown with hipMalloc.
inline int device_buffer(cpb_context* ctx, DevBuf& b, size_t bytes, void** out) {
    if (ctx->carve_base) {
        const size_t off = (ctx->carve_used + 255) / 256 * 256;
        if (off + bytes > ctx->carve_cap)
            return fail(ctx, CPB_E_RESOURCE,
                        "the search needs more scratch than the static scratch array holds "
                        "(" + std::to_string(ctx->carve_cap >> 20) + " MiB); use the "
                        "allocating entry point");
        *out = ctx->carve_base + off;
        ctx->carve_used = off + bytes;
        return CPB_OK;
    }
    const int rc = ensure(ctx, b, bytes);
    if (rc == CPB_OK) *out = b.ptr;
    return rc;
}

// Fixed-point conversion of every curve's points and its parameters (one workgroup per
// curve); defined with the kernel in bls_engine.hip.
// `tau` (the time offsets) may be NULL: x_max is then infinity.
hipError_t launch_prep_curves(hipStream_t st, const double* tau, const double* yw,
                              const double* wv, const cpb_curve* curves, int32_t n_curves,
                              int32_t obj_flag, CurveParams* params, int64_t* yq,
                              int64_t* wq);

}  // namespace cpb_engine
