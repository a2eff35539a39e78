// Learned candidate proposer for gfx950: features, the 1-D convolutional network's
// forward and backward passes, and its Adam step (C ABI of include/cuperiod_proposer.h).
//
// Activations are channels-last BF16 [curve][position][64]. The 64 -> 64 convolutions
// are implicit GEMMs on v_mfma_f32_16x16x32_bf16: output channels are the M dimension,
// positions the N dimension and (tap, input channel) the K dimension (K = 7*64 = 448).
// For an MFMA 16x16x32 the lane holding (row r, k-group g) is lane = 16*g + r: A supplies
// A[r][8g .. 8g+7], B supplies B[8g .. 8g+7][r] (r as the column), and the accumulator
// holds D[4g + j][r] for j = 0..3.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "cuperiod_bls.h"
#include "cuperiod_proposer.h"

namespace {

constexpr int kC = CPP_CHANNELS;
constexpr int kTaps = CPP_TAPS;
constexpr int kL = CPP_POSITIONS;
constexpr int kKdim = kC * kTaps;  // 448
constexpr int kRow = 72;           // LDS row of 64 (+8 pad) BF16: 144 B, so 8 lanes
                                   // reading 16 B from consecutive rows hit distinct banks
constexpr int kTN = 256;           // positions per forward workgroup
constexpr int kWX = 64;            // positions per staged weight-gradient sub-chunk
constexpr int kWSub = 8;           // sub-chunks per weight-gradient workgroup

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even (NaN stays NaN), as torch converts float32 to bfloat16.
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) {
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

__device__ __forceinline__ float block_max(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    float r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmaxf(r, red[w]);
    return r;
}

__device__ __forceinline__ float block_sum(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    float r = 0.0f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += red[w];
    return r;
}

// ---------------------------------------------------------------------------------
// Features: standardised coarse power and the cadence histogram, one workgroup a curve.

__global__ __launch_bounds__(256) void features_kernel(
    const double* __restrict__ power, const int64_t* __restrict__ out_base,
    const double* __restrict__ summary, const double* __restrict__ tau,
    const int64_t* __restrict__ point_base, const int32_t* __restrict__ n_points,
    const double* __restrict__ baseline, float clip, float* __restrict__ x) {
    __shared__ int hist[CPP_HIST_BINS];
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    for (int i = tid; i < CPP_HIST_BINS; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int n = n_points[c];
    const double base = baseline[c];
    const double scale = base > 0.0 ? CPP_HIST_BINS / base : 0.0;
    const double* t = tau + point_base[c];
    for (int j = tid; j < n; j += blockDim.x) {
        int b = (int)(t[j] * scale);
        b = b < 0 ? 0 : (b >= CPP_HIST_BINS ? CPP_HIST_BINS - 1 : b);
        atomicAdd(&hist[b], 1);
    }
    __syncthreads();
    const double mean = summary[4 * c], sd = summary[4 * c + 1];
    const float hs = n > 0 ? (float)CPP_HIST_BINS / (float)n : 0.0f;
    const double* p = power + out_base[c];
    float* xc = x + (size_t)c * kL * 2;
    for (int i = tid; i < kL; i += blockDim.x) {
        double v = p[kL - 1 - i];
        v = isfinite(v) ? v : 0.0;
        float s = sd > 0.0 ? (float)((v - mean) / sd) : 0.0f;
        s = fminf(fmaxf(s, -clip), clip);
        xc[2 * i] = s;
        xc[2 * i + 1] = (float)hist[i >> 3] * hs;
    }
}

// Stacked features: the coarse power of each curve summed over its time chunks, one
// workgroup a curve. Chunk k of curve c has its coarse outputs at out_base[k] (ascending
// period), so cell x reads power[out_base[k] + L-1-x]; S[x] = sum_k snr_k(x)^2 over the
// chunks (a power that is not finite and positive counts 0). Channel 0 is S standardised
// over the curve's cells and clipped; channel 1 the cadence histogram, as features_kernel.
__global__ __launch_bounds__(256) void stack_features_kernel(
    const double* __restrict__ power, const int32_t* __restrict__ chunk_start,
    const int64_t* __restrict__ out_base, const double* __restrict__ tau,
    const int64_t* __restrict__ point_base, const int32_t* __restrict__ n_points,
    const double* __restrict__ baseline, float clip, float* __restrict__ x) {
    __shared__ int hist[CPP_HIST_BINS];
    __shared__ float sv[kL];
    __shared__ double red[4];
    const int c = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < CPP_HIST_BINS; i += blockDim.x) hist[i] = 0;
    const int k0 = chunk_start[c], k1 = chunk_start[c + 1];
    double part = 0.0;
    for (int i = tid; i < kL; i += blockDim.x) {
        double acc = 0.0;
        for (int k = k0; k < k1; ++k) {
            const double v = power[out_base[k] + (kL - 1 - i)];
            if (isfinite(v) && v > 0.0) acc += v * v;
        }
        sv[i] = (float)acc;
        part += acc;
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if (lane == 0) red[wave] = part;
    __syncthreads();
    const double mean = (red[0] + red[1] + red[2] + red[3]) / kL;
    __syncthreads();
    double q = 0.0;
    for (int i = tid; i < kL; i += blockDim.x) q += ((double)sv[i] - mean) * ((double)sv[i] - mean);
    for (int o = 32; o > 0; o >>= 1) q += __shfl_xor(q, o);
    if (lane == 0) red[wave] = q;
    const int n = n_points[c];
    const double base = baseline[c];
    const double scale = base > 0.0 ? CPP_HIST_BINS / base : 0.0;
    const double* t = tau + point_base[c];
    for (int j = tid; j < n; j += blockDim.x) {
        int b = (int)(t[j] * scale);
        b = b < 0 ? 0 : (b >= CPP_HIST_BINS ? CPP_HIST_BINS - 1 : b);
        atomicAdd(&hist[b], 1);
    }
    __syncthreads();
    const double sd = sqrt((red[0] + red[1] + red[2] + red[3]) / kL);
    const float hs = n > 0 ? (float)CPP_HIST_BINS / (float)n : 0.0f;
    float* xc = x + (size_t)c * kL * 2;
    for (int i = tid; i < kL; i += blockDim.x) {
        float z = sd > 0.0 ? (float)(((double)sv[i] - mean) / sd) : 0.0f;
        z = fminf(fmaxf(z, -clip), clip);
        xc[2 * i] = z;
        xc[2 * i + 1] = (float)hist[i >> 3] * hs;
    }
}

// ---------------------------------------------------------------------------------
// Layer 1 (2 -> 64 channels) on the vector ALUs: one thread per position.

__global__ __launch_bounds__(256) void layer1_forward_kernel(
    const float* __restrict__ x, const float* __restrict__ w1, const float* __restrict__ b1,
    int64_t n_pos, uint16_t* __restrict__ z1, uint16_t* __restrict__ h1) {
    __shared__ float sw[kC * kTaps * 2];
    __shared__ float sb[kC];
    for (int i = threadIdx.x; i < kC * kTaps * 2; i += blockDim.x) sw[i] = w1[i];
    if (threadIdx.x < kC) sb[threadIdx.x] = b1[threadIdx.x];
    __syncthreads();
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pos) return;
    const int64_t c = p / kL;
    const int xx = (int)(p % kL);
    float in[kTaps * 2];
#pragma unroll
    for (int k = 0; k < kTaps; ++k) {
        const int xs = xx + k - 3;
        const bool ok = xs >= 0 && xs < kL;
        in[2 * k] = ok ? x[(c * kL + xs) * 2] : 0.0f;
        in[2 * k + 1] = ok ? x[(c * kL + xs) * 2 + 1] : 0.0f;
    }
    uint16_t* zr = z1 ? z1 + p * kC : nullptr;
    uint16_t* hr = h1 + p * kC;
    for (int o0 = 0; o0 < kC; o0 += 8) {
        uint16_t zb[8], hb[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int o = o0 + e;
            float z = sb[o];
#pragma unroll
            for (int j = 0; j < kTaps * 2; ++j) z = fmaf(sw[o * kTaps * 2 + j], in[j], z);
            zb[e] = f2bf(z);
            hb[e] = f2bf(gelu(z));
        }
        if (zr) *reinterpret_cast<uint4*>(zr + o0) = *reinterpret_cast<const uint4*>(zb);
        *reinterpret_cast<uint4*>(hr + o0) = *reinterpret_cast<const uint4*>(hb);
    }
}

// dw1[o][k][c] += sum_x dz1[x][o] * x[x + k - 3][c]; db1[o] += sum_x dz1[x][o].
// One workgroup per 512 positions of a curve; thread = (channel o, position group g).
__global__ __launch_bounds__(256) void layer1_wgrad_kernel(
    const uint16_t* __restrict__ dz1, const float* __restrict__ x, float* __restrict__ dw1,
    float* __restrict__ db1) {
    constexpr int kSpan = 512;
    __shared__ float red[4][kC][kTaps * 2 + 1];
    const int blocks = kL / kSpan;
    const int64_t c = blockIdx.x / blocks;
    const int xb = (blockIdx.x % blocks) * kSpan;
    const int o = threadIdx.x & 63, g = threadIdx.x >> 6;
    float acc[kTaps * 2 + 1];
#pragma unroll
    for (int j = 0; j <= kTaps * 2; ++j) acc[j] = 0.0f;
    for (int xx = xb + g; xx < xb + kSpan; xx += 4) {
        const float d = bf2f(dz1[(c * kL + xx) * kC + o]);
#pragma unroll
        for (int k = 0; k < kTaps; ++k) {
            const int xs = xx + k - 3;
            if (xs >= 0 && xs < kL) {
                acc[2 * k] = fmaf(d, x[(c * kL + xs) * 2], acc[2 * k]);
                acc[2 * k + 1] = fmaf(d, x[(c * kL + xs) * 2 + 1], acc[2 * k + 1]);
            }
        }
        acc[kTaps * 2] += d;
    }
#pragma unroll
    for (int j = 0; j <= kTaps * 2; ++j) red[g][o][j] = acc[j];
    __syncthreads();
    for (int i = threadIdx.x; i < kC * (kTaps * 2 + 1); i += blockDim.x) {
        const int oo = i / (kTaps * 2 + 1), j = i % (kTaps * 2 + 1);
        const float s = red[0][oo][j] + red[1][oo][j] + red[2][oo][j] + red[3][oo][j];
        if (j < kTaps * 2)
            atomicAdd(&dw1[oo * kTaps * 2 + j], s);
        else
            atomicAdd(&db1[oo], s);
    }
}

// ---------------------------------------------------------------------------------
// 64 -> 64 convolution on MFMA. A workgroup computes all 64 output channels of kTN
// positions of one curve; wave w owns output channels 16w .. 16w+15 and keeps its
// 16 x 448 weight slice in registers, the input rows x0-3 .. x0+kTN+2 sit in LDS.

template <int EPI>
__global__ __launch_bounds__(256) void conv_mfma_kernel(
    const uint16_t* __restrict__ in, const uint16_t* __restrict__ w,
    const float* __restrict__ bias, const uint16_t* __restrict__ zsrc,
    uint16_t* __restrict__ zsave, uint16_t* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t tile[(kTN + 6) * kRow];
    constexpr int tiles = kL / kTN;
    const int64_t c = blockIdx.x / tiles;
    const int x0 = (blockIdx.x % tiles) * kTN;
    const int tid = threadIdx.x;
    const uint16_t* inb = in + c * kL * kC;
    for (int q = tid; q < (kTN + 6) * 8; q += blockDim.x) {
        const int r = q >> 3, part = q & 7;
        const int xs = x0 - 3 + r;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (xs >= 0 && xs < kL) v = *reinterpret_cast<const uint4*>(inb + (int64_t)xs * kC + part * 8);
        *reinterpret_cast<uint4*>(tile + r * kRow + part * 8) = v;
    }
    const int lane = tid & 63, wave = tid >> 6;
    const int kg = lane >> 4, l16 = lane & 15;
    bf16x8 a[kKdim / 32];
    const uint16_t* wrow = w + (int64_t)(16 * wave + l16) * kKdim;
#pragma unroll
    for (int cc = 0; cc < kKdim / 32; ++cc)
        a[cc] = *reinterpret_cast<const bf16x8*>(wrow + 32 * cc + 8 * kg);
    __syncthreads();

    f32x4 acc[kTN / 16];
#pragma unroll
    for (int t = 0; t < kTN / 16; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int cc = 0; cc < kKdim / 32; ++cc) {
        const int kk0 = 32 * cc + 8 * kg;
        const int k = kk0 >> 6, ci = kk0 & 63;
#pragma unroll
        for (int t = 0; t < kTN / 16; ++t) {
            const int r = t * 16 + l16 + k;  // input row of position x0 + 16t + l16, tap k
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(tile + r * kRow + ci);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[cc], b, acc[t], 0, 0, 0);
        }
    }

    const int o0 = 16 * wave + 4 * kg;
#pragma unroll
    for (int t = 0; t < kTN / 16; ++t) {
        const int64_t base = (c * kL + x0 + 16 * t + l16) * kC + o0;
        uint16_t ob[4], zb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float v = acc[t][j];
            if constexpr (EPI == 0) {
                const float z = v + bias[o0 + j];
                zb[j] = f2bf(z);
                ob[j] = f2bf(gelu(z));
            } else if constexpr (EPI == 1) {
                ob[j] = f2bf(v * gelu_grad(bf2f(zsrc[base + j])));
            } else {
                ob[j] = f2bf(v + bias[o0 + j]);
            }
        }
        *reinterpret_cast<uint2*>(out + base) = *reinterpret_cast<const uint2*>(ob);
        if constexpr (EPI == 0) {
            if (zsave) *reinterpret_cast<uint2*>(zsave + base) = *reinterpret_cast<const uint2*>(zb);
        }
    }
}

// Elements K .. K+7 of the 16-element concatenation (lo, hi).
template <int K>
__device__ __forceinline__ bf16x8 take8(bf16x8 lo, bf16x8 hi) {
    return __builtin_shufflevector(lo, hi, K, K + 1, K + 2, K + 3, K + 4, K + 5, K + 6, K + 7);
}

template <int NT>
__device__ __forceinline__ void wgrad_tiles(const uint16_t* inT_kc, const bf16x8& a, int l16,
                                            f32x4* acc) {
    if constexpr (NT < 4 * kTaps) {
        constexpr int k = NT >> 2;
        const int ci = (NT & 3) * 16 + l16;
        const uint16_t* row = inT_kc + ci * kRow;
        const bf16x8 lo = *reinterpret_cast<const bf16x8*>(row);
        const bf16x8 hi = *reinterpret_cast<const bf16x8*>(row + 8);
        acc[NT] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, take8<k>(lo, hi), acc[NT], 0, 0, 0);
        wgrad_tiles<NT + 1>(inT_kc, a, l16, acc);
    }
}

// Weight gradient: M = output channel, N = (tap, input channel), K = position. Tiles of
// kWX positions are staged transposed in LDS (channel-major) so both operands read 8
// consecutive positions; a workgroup covers kWX*kWSub positions of one curve and adds its
// partial sums to dw/db with atomics.
__global__ __launch_bounds__(256) void conv_wgrad_kernel(
    const uint16_t* __restrict__ dz, const uint16_t* __restrict__ in, float* __restrict__ dw,
    float* __restrict__ db) {
    __shared__ __attribute__((aligned(16))) uint16_t dzT[kC * kRow];  // [o][xx], xx < 64
    __shared__ __attribute__((aligned(16))) uint16_t inT[kC * kRow];  // [c][rr], x = xs-3+rr
    constexpr int blocks = kL / (kWX * kWSub);
    const int64_t c = blockIdx.x / blocks;
    const int xb = (blockIdx.x % blocks) * (kWX * kWSub);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int kg = lane >> 4, l16 = lane & 15;
    f32x4 acc[4 * kTaps];
#pragma unroll
    for (int i = 0; i < 4 * kTaps; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    for (int sc = 0; sc < kWSub; ++sc) {
        const int xs = xb + sc * kWX;
        __syncthreads();
        for (int q = tid; q < kWX * 8; q += blockDim.x) {
            const int xx = q >> 3, part = q & 7;
            const uint4 v = *reinterpret_cast<const uint4*>(dz + (c * kL + xs + xx) * kC + part * 8);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int i = 0; i < 8; ++i) dzT[(part * 8 + i) * kRow + xx] = e[i];
        }
        for (int q = tid; q < kRow * 8; q += blockDim.x) {
            const int rr = q >> 3, part = q & 7;
            const int x = xs - 3 + rr;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (x >= 0 && x < kL && rr < kWX + 6)
                v = *reinterpret_cast<const uint4*>(in + (c * kL + x) * kC + part * 8);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int i = 0; i < 8; ++i) inT[(part * 8 + i) * kRow + rr] = e[i];
        }
        __syncthreads();
        if (tid < kC) {
            for (int xx = 0; xx < kWX; ++xx) bsum += bf2f(dzT[tid * kRow + xx]);
        }
#pragma unroll
        for (int kc = 0; kc < kWX / 32; ++kc) {
            const bf16x8 a =
                *reinterpret_cast<const bf16x8*>(dzT + (16 * wave + l16) * kRow + 32 * kc + 8 * kg);
            wgrad_tiles<0>(inT + 32 * kc + 8 * kg, a, l16, acc);
        }
    }
#pragma unroll
    for (int nt = 0; nt < 4 * kTaps; ++nt) {
        const int k = nt >> 2, ci = (nt & 3) * 16 + l16;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = 16 * wave + 4 * kg + j;
            atomicAdd(&dw[(o * kTaps + k) * kC + ci], acc[nt][j]);
        }
    }
    if (tid < kC) atomicAdd(&db[tid], bsum);
}

__global__ void flip_weights_kernel(const uint16_t* __restrict__ w, uint16_t* __restrict__ wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // index into wt[c][k][o]
    if (i >= kC * kTaps * kC) return;
    const int c = i / (kTaps * kC), k = (i / kC) % kTaps, o = i % kC;
    wt[i] = w[(o * kTaps + (kTaps - 1 - k)) * kC + c];
}

// ---------------------------------------------------------------------------------
// Head.

__device__ __forceinline__ void head_logits(const uint16_t* __restrict__ h4c,
                                            const float* swh, float bh, float* logit) {
    for (int x = threadIdx.x; x < kL; x += blockDim.x) {
        const uint4* r = reinterpret_cast<const uint4*>(h4c + (int64_t)x * kC);
        float s = bh;
#pragma unroll
        for (int q = 0; q < kC / 8; ++q) {
            const uint4 v = r[q];
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int i = 0; i < 8; ++i) s = fmaf(swh[8 * q + i], bf2f(e[i]), s);
        }
        logit[x] = s;
    }
}

// Training head: softmax over the positions, cross-entropy against a uniform target on
// the label range [lo, hi] (the cells whose frequency lies within a refinement window of
// the full-grid best period).
__global__ __launch_bounds__(256) void head_train_kernel(
    const uint16_t* __restrict__ h4, const uint16_t* __restrict__ z4,
    const float* __restrict__ wh, const float* __restrict__ bh,
    const int32_t* __restrict__ label, float scale, uint16_t* __restrict__ dz4,
    float* __restrict__ dwh, float* __restrict__ dbh, float* __restrict__ loss) {
    __shared__ float swh[kC];
    __shared__ float logit[kL];
    __shared__ float red[16];
    __shared__ float wred[4][kC];
    const int64_t c = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < kC) swh[tid] = wh[tid];
    __syncthreads();
    const uint16_t* h4c = h4 + c * kL * kC;
    head_logits(h4c, swh, bh[0], logit);
    __syncthreads();
    int lo = label[2 * c], hi = label[2 * c + 1];
    lo = lo < 0 ? 0 : (lo >= kL ? kL - 1 : lo);
    hi = hi < lo ? lo : (hi >= kL ? kL - 1 : hi);
    const float q = 1.0f / (float)(hi - lo + 1);
    float m = -INFINITY, tl = 0.0f;
    for (int x = tid; x < kL; x += blockDim.x) {
        m = fmaxf(m, logit[x]);
        if (x >= lo && x <= hi) tl += logit[x];
    }
    m = block_max(m, red);
    tl = block_sum(tl, red);
    float s = 0.0f;
    for (int x = tid; x < kL; x += blockDim.x) s += expf(logit[x] - m);
    s = block_sum(s, red);
    __syncthreads();
    float dsum = 0.0f;
    for (int x = tid; x < kL; x += blockDim.x) {
        const float p = expf(logit[x] - m) / s;
        const float d = scale * (p - (x >= lo && x <= hi ? q : 0.0f));
        logit[x] = d;
        dsum += d;
    }
    dsum = block_sum(dsum, red);
    if (tid == 0) {
        atomicAdd(loss, scale * (logf(s) + m - tl * q));
        atomicAdd(dbh, dsum);
    }
    __syncthreads();
    const int o = tid & 63, g = tid >> 6;
    const float who = swh[o];
    float aw = 0.0f;
    for (int x = g; x < kL; x += 4) {
        const int64_t i = (c * kL + x) * kC + o;
        const float d = logit[x];
        aw = fmaf(d, bf2f(h4[i]), aw);
        dz4[i] = f2bf(d * who * gelu_grad(bf2f(z4[i])));
    }
    wred[g][o] = aw;
    __syncthreads();
    if (tid < kC) atomicAdd(&dwh[tid], wred[0][tid] + wred[1][tid] + wred[2][tid] + wred[3][tid]);
}

// Inference head: softmax probabilities, then greedy windows. A window centred on cell x
// covers the cells whose frequency f_lo * exp(x' * dlog) lies within half_width of x's
// (at least `radius` cells each side); it claims the probability of the cells it covers
// that no earlier window claimed.
__global__ __launch_bounds__(256) void head_windows_kernel(
    const uint16_t* __restrict__ h4, const float* __restrict__ wh, const float* __restrict__ bh,
    const double* __restrict__ f_lo, const double* __restrict__ dlog,
    const double* __restrict__ half_width, int n_windows, int radius,
    int32_t* __restrict__ centre, float* __restrict__ mass, int32_t* __restrict__ span) {
    __shared__ float swh[kC];
    __shared__ float p[kL];
    __shared__ unsigned char avail[kL];
    __shared__ float red[16];
    __shared__ int redi[4];
    const int64_t c = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < kC) swh[tid] = wh[tid];
    __syncthreads();
    head_logits(h4 + c * kL * kC, swh, bh[0], p);
    __syncthreads();
    float m = -INFINITY;
    for (int x = tid; x < kL; x += blockDim.x) m = fmaxf(m, p[x]);
    m = block_max(m, red);
    float s = 0.0f;
    for (int x = tid; x < kL; x += blockDim.x) s += expf(p[x] - m);
    s = block_sum(s, red);
    __syncthreads();
    for (int x = tid; x < kL; x += blockDim.x) {
        p[x] = expf(p[x] - m) / s;
        avail[x] = 1;
    }
    __syncthreads();
    const double fl = f_lo[c], dl = dlog[c], hw = half_width[c];
    const int lane = tid & 63, wave = tid >> 6;
    for (int w = 0; w < n_windows; ++w) {
        // most probable available position, lowest index on ties
        float bv = -1.0f;
        int bi = kL;
        for (int x = tid; x < kL; x += blockDim.x) {
            if (avail[x] && p[x] > bv) {
                bv = p[x];
                bi = x;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(bv, o);
            const int oi = __shfl_xor(bi, o);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            red[wave] = bv;
            redi[wave] = bi;
        }
        __syncthreads();
        float gv = red[0];
        int gi = redi[0];
        for (int q = 1; q < 4; ++q) {
            if (red[q] > gv || (red[q] == gv && redi[q] < gi)) {
                gv = red[q];
                gi = redi[q];
            }
        }
        __syncthreads();
        if (gi >= kL) {
            if (tid == 0) {
                centre[c * n_windows + w] = -1;
                mass[c * n_windows + w] = 0.0f;
                if (span) {
                    span[2 * (c * n_windows + w)] = -1;
                    span[2 * (c * n_windows + w) + 1] = -1;
                }
            }
            continue;
        }
        const double fx = fl * exp((double)gi * dl);
        int lo = gi - radius, hi = gi + radius;
        if (dl > 0.0) {
            const int a = fx - hw > fl ? (int)ceil(log((fx - hw) / fl) / dl) : 0;
            const int b = (int)floor(log((fx + hw) / fl) / dl);
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
        }
        lo = lo < 0 ? 0 : lo;
        hi = hi >= kL ? kL - 1 : hi;
        float part = 0.0f;
        for (int x = lo + tid; x <= hi; x += blockDim.x) part += p[x];
        part = block_sum(part, red);
        __syncthreads();
        for (int x = lo + tid; x <= hi; x += blockDim.x) {
            p[x] = 0.0f;
            avail[x] = 0;
        }
        if (tid == 0) {
            centre[c * n_windows + w] = gi;
            mass[c * n_windows + w] = part;
            if (span) {
                span[2 * (c * n_windows + w)] = lo;
                span[2 * (c * n_windows + w) + 1] = hi;
            }
        }
        __syncthreads();
    }
}

__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ pb, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    g[i] = 0.0f;
    if (pb) pb[i] = f2bf(pi);
}

int launch_status() { return hipGetLastError() == hipSuccess ? CPB_OK : CPB_E_DEVICE; }

}  // namespace

extern "C" {

int cpp_features(void* stream, const double* power, const int64_t* out_base,
                 const double* summary, const double* tau, const int64_t* point_base,
                 const int32_t* n_points, const double* baseline, int32_t n_curves,
                 float clip, float* x) {
    if (n_curves <= 0 || !power || !out_base || !summary || !tau || !point_base || !n_points ||
        !baseline || !x || !(clip > 0.0f))
        return CPB_E_INVALID;
    hipLaunchKernelGGL(features_kernel, dim3(n_curves), dim3(256), 0, (hipStream_t)stream,
                       power, out_base, summary, tau, point_base, n_points, baseline, clip, x);
    return launch_status();
}

int cpp_stack_features(void* stream, const double* power, const int32_t* chunk_start,
                       const int64_t* out_base, const double* tau, const int64_t* point_base,
                       const int32_t* n_points, const double* baseline, int32_t n_curves,
                       float clip, float* x) {
    if (n_curves <= 0 || !power || !chunk_start || !out_base || !tau || !point_base ||
        !n_points || !baseline || !x || !(clip > 0.0f))
        return CPB_E_INVALID;
    hipLaunchKernelGGL(stack_features_kernel, dim3(n_curves), dim3(256), 0, (hipStream_t)stream,
                       power, chunk_start, out_base, tau, point_base, n_points, baseline, clip,
                       x);
    return launch_status();
}

int cpp_layer1_forward(void* stream, const float* x, const float* w1, const float* b1,
                       int32_t n_curves, uint16_t* z1, uint16_t* h1) {
    if (n_curves <= 0 || !x || !w1 || !b1 || !h1) return CPB_E_INVALID;
    const int64_t n_pos = (int64_t)n_curves * kL;
    hipLaunchKernelGGL(layer1_forward_kernel, dim3((unsigned)((n_pos + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, x, w1, b1, n_pos, z1, h1);
    return launch_status();
}

int cpp_conv_forward(void* stream, const uint16_t* in, const uint16_t* w, const float* bias,
                     const uint16_t* zsrc, uint16_t* zsave, uint16_t* out, int32_t n_curves,
                     int32_t epilogue) {
    if (n_curves <= 0 || !in || !w || !out) return CPB_E_INVALID;
    const dim3 grid((unsigned)((int64_t)n_curves * (kL / kTN)));
    hipStream_t s = (hipStream_t)stream;
    switch (epilogue) {
    case 0:
        if (!bias) return CPB_E_INVALID;
        hipLaunchKernelGGL(conv_mfma_kernel<0>, grid, dim3(256), 0, s, in, w, bias, zsrc, zsave, out);
        break;
    case 1:
        if (!zsrc) return CPB_E_INVALID;
        hipLaunchKernelGGL(conv_mfma_kernel<1>, grid, dim3(256), 0, s, in, w, bias, zsrc, zsave, out);
        break;
    case 2:
        if (!bias) return CPB_E_INVALID;
        hipLaunchKernelGGL(conv_mfma_kernel<2>, grid, dim3(256), 0, s, in, w, bias, zsrc, zsave, out);
        break;
    default:
        return CPB_E_INVALID;
    }
    return launch_status();
}

int cpp_conv_wgrad(void* stream, const uint16_t* dz, const uint16_t* in, int32_t n_curves,
                   float* dw, float* db) {
    if (n_curves <= 0 || !dz || !in || !dw || !db) return CPB_E_INVALID;
    const dim3 grid((unsigned)((int64_t)n_curves * (kL / (kWX * kWSub))));
    hipLaunchKernelGGL(conv_wgrad_kernel, grid, dim3(256), 0, (hipStream_t)stream, dz, in, dw, db);
    return launch_status();
}

int cpp_layer1_wgrad(void* stream, const uint16_t* dz1, const float* x, int32_t n_curves,
                     float* dw1, float* db1) {
    if (n_curves <= 0 || !dz1 || !x || !dw1 || !db1) return CPB_E_INVALID;
    const dim3 grid((unsigned)((int64_t)n_curves * (kL / 512)));
    hipLaunchKernelGGL(layer1_wgrad_kernel, grid, dim3(256), 0, (hipStream_t)stream, dz1, x, dw1, db1);
    return launch_status();
}

int cpp_flip_weights(void* stream, const uint16_t* w, uint16_t* wt) {
    if (!w || !wt || w == wt) return CPB_E_INVALID;
    hipLaunchKernelGGL(flip_weights_kernel, dim3((kC * kTaps * kC + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, w, wt);
    return launch_status();
}

int cpp_head_train(void* stream, const uint16_t* h4, const uint16_t* z4, const float* wh,
                   const float* bh, const int32_t* label, int32_t n_curves, float scale,
                   uint16_t* dz4, float* dwh, float* dbh, float* loss) {
    if (n_curves <= 0 || !h4 || !z4 || !wh || !bh || !label || !dz4 || !dwh || !dbh || !loss)
        return CPB_E_INVALID;
    hipLaunchKernelGGL(head_train_kernel, dim3(n_curves), dim3(256), 0, (hipStream_t)stream, h4, z4,
                       wh, bh, label, scale, dz4, dwh, dbh, loss);
    return launch_status();
}

int cpp_head_windows(void* stream, const uint16_t* h4, const float* wh, const float* bh,
                     const double* f_lo, const double* dlog, const double* half_width,
                     int32_t n_curves, int32_t n_windows, int32_t radius, int32_t* centre,
                     float* mass, int32_t* span) {
    if (n_curves <= 0 || n_windows <= 0 || n_windows > kL || radius < 0 || radius >= kL || !h4 ||
        !wh || !bh || !f_lo || !dlog || !half_width || !centre || !mass)
        return CPB_E_INVALID;
    hipLaunchKernelGGL(head_windows_kernel, dim3(n_curves), dim3(256), 0, (hipStream_t)stream, h4,
                       wh, bh, f_lo, dlog, half_width, n_windows, radius, centre, mass, span);
    return launch_status();
}

int cpp_adam(void* stream, float* p, float* g, float* m, float* v, uint16_t* pb, int64_t n,
             float lr, float beta1, float beta2, float eps, int32_t t) {
    if (n <= 0 || !p || !g || !m || !v || t < 1) return CPB_E_INVALID;
    const float c1 = 1.0f - powf(beta1, (float)t);
    const float c2 = 1.0f - powf(beta2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, p, g, m, v, pb, n, lr, beta1, beta2, eps, c1, c2);
    return launch_status();
}

}  // extern "C"
