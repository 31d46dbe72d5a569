// neural.hip -- learned boundary SDF: a Fourier-feature MLP trained and
// evaluated on the GPU with BF16 MFMA (v_mfma_f32_16x16x32_bf16) and FP32
// accumulation.
//
// No reference counterpart: the reference samples its boundary from three
// dense 3D textures (particle_simulation.cu:48-53, 2959-3011) built by
// SphInComplexShapes (SphInComplexShapes.cpp:203-313).  The network is a
// compact stand-in for the distance texture, trained on (point, signed
// distance) samples of the same boundary construction (scene.cpp).
//
// Network: u = 2(x - lo)/(hi - lo) - 1; features g = [u, sin(pi f_k u_a),
// cos(pi f_k u_a)] (k < 16, f_k = 2^(7k/15); 99 used, zero-padded to 128);
// 128 -> 64 -> 64 -> 64 (SiLU) -> 1; d = out_scale * y.
//
// MFMA operand layout (16x16x32 bf16): lane l holds A[l%16][8(l/16)+0..7],
// B[8(l/16)+0..7][l%16] and produces D[4(l/16)+j][l%16], j < 4.  A tile of 64
// samples is worked by four wavefronts, wave w owning sample rows 16w..16w+15
// of every row-parallel product.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.h"
#include "sph.h"

namespace sph {
namespace nn {

constexpr int kFreqs = 16;
constexpr int kFeat = 3 + 6 * kFreqs;   // 99
constexpr int kK1 = 128;
constexpr int kH = 64;
constexpr int kTile = 64;
constexpr int kThreads = 256;
constexpr int kLds = 72;                // bf16 row stride of LDS tiles (16-byte aligned, padded)

// flat FP32 parameter layout
constexpr int oW1 = 0;
constexpr int ob1 = oW1 + kK1 * kH;
constexpr int oW2 = ob1 + kH;
constexpr int ob2 = oW2 + kH * kH;
constexpr int oW3 = ob2 + kH;
constexpr int ob3 = oW3 + kH * kH;
constexpr int ow4 = ob3 + kH;
constexpr int ob4 = ow4 + kH;
constexpr int kParams = ob4 + 1;
static_assert(kParams == SPH_MLP_NUM_PARAMS, "parameter count");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// BF16 copies of the weights in the two layouts the products read.
struct Bf16Weights {
    uint16_t* W1T;   // [kH][kK1]  (n-major: forward operand)
    uint16_t* W1;    // [kK1][kH]  (k-major: input-gradient operand)
    uint16_t* W2T;   // [kH][kH]
    uint16_t* W2;    // [kH][kH]
    uint16_t* W3T;
    uint16_t* W3;
};

struct NetView {
    const float* theta;      // FP32 master parameters
    Bf16Weights w;
    float lo[3], inv_ext[3];
    float out_scale;         // d = out_scale * y (band == 0)
    float band;              // d = band * atanh(y) (band > 0)
    float freq[kFreqs];      // f_k = 2^(7k/15) rounded to FP32
};

// Training target of a signed distance and the decoding of an output.
__device__ __forceinline__ float target_of(const NetView& net, float d) {
    return net.band > 0.0f ? tanhf(d / net.band) : d / net.out_scale;
}

// Distance and d(distance)/dy of an output y.
__device__ __forceinline__ float decode(const NetView& net, float y, float& dd_dy) {
    if (net.band > 0.0f) {
        float yc = fminf(fmaxf(y, -0.999999f), 0.999999f);
        dd_dy = net.band / (1.0f - yc * yc);
        return net.band * atanhf(yc);
    }
    dd_dy = net.out_scale;
    return net.out_scale * y;
}

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 frag_ld(const uint16_t* p) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

// sin(pi x) and cos(pi x): exact reduction to r = x - 2 rint(x/2) in
// [-1, 1] (|x| < 2^22 here), then the hardware sine/cosine of r/2
// revolutions (v_sin_f32 / v_cos_f32 take revolutions).
__device__ __forceinline__ void sincospi_hw(float x, float& s, float& c) {
    const float t = 0.5f * (x - 2.0f * rintf(0.5f * x));
    s = __builtin_amdgcn_sinf(t);
    c = __builtin_amdgcn_cosf(t);
}

// Feature f of normalized coordinates u.
__device__ __forceinline__ float feature(const NetView& net, int f, const float u[3]) {
    if (f < 3) return u[f];
    if (f >= kFeat) return 0.0f;
    int i = f - 3, k = i / 6, a = (i % 6) >> 1;
    float s, c;
    sincospi_hw(net.freq[k] * u[a], s, c);
    return (i & 1) ? c : s;
}

// d feature f / d u_a(f) (the axis it depends on is returned in axis).
__device__ __forceinline__ float feature_du(const NetView& net, int f, const float u[3], int& axis) {
    if (f < 3) {
        axis = f;
        return 1.0f;
    }
    if (f >= kFeat) {
        axis = 0;
        return 0.0f;
    }
    int i = f - 3, k = i / 6, a = (i % 6) >> 1;
    axis = a;
    float w = net.freq[k], s, c;
    sincospi_hw(w * u[a], s, c);
    float pw = 3.14159265358979f * w;
    return (i & 1) ? -pw * s : pw * c;
}

__device__ __forceinline__ void normalized(const NetView& net, float x, float y, float z, float u[3]) {
    const float p[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = 2.0f * ((p[a] - net.lo[a]) * net.inv_ext[a]) - 1.0f;
}

// Row-parallel product for the 64-sample tile: acc[c] (c < NC col tiles of
// 16) += A[rows of wave w][K] . B[K][16c..16c+15], A from LDS (row-major,
// stride kLds), B^T from global (n-major, stride K).
template <int K, int NC>
__device__ __forceinline__ void rows_x_wt(const uint16_t* sA, const uint16_t* __restrict__ BT, int lane, int w,
                                          f32x4 acc[NC]) {
#pragma unroll
    for (int ks = 0; ks < K / 32; ++ks) {
        bf16x8 a = frag_ld(sA + (16 * w + (lane & 15)) * kLds + 32 * ks + 8 * (lane >> 4));
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            bf16x8 b = frag_ld(BT + (size_t)(16 * c + (lane & 15)) * K + 32 * ks + 8 * (lane >> 4));
            acc[c] = mfma(a, b, acc[c]);
        }
    }
}

// Input layer of a tile: features of sample row 16w + lane%16 computed in
// registers in A-fragment form (k = 32ks + 8(lane/16) + 0..7).  Optionally
// also stored transposed ([k][s]) for the weight-gradient product.
__device__ __forceinline__ void input_layer(const NetView& net, const float u[3], int lane, int w, f32x4 acc[4],
                                            uint16_t* sA0T) {
#pragma unroll
    for (int ks = 0; ks < kK1 / 32; ++ks) {
        // features f0..f0+7 (f0 = 8q) use the (sin, cos) pairs
        // m = 4q - 2 .. 4q + 2 (pair m = 3k + a: features 3 + 2m, 4 + 2m)
        const int q = 4 * ks + (lane >> 4), f0 = 8 * q;
        float ps[5], pc[5];
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            const int m = 4 * q - 2 + t;
            ps[t] = pc[t] = 0.0f;
            if (m >= 0 && m < 3 * kFreqs) sincospi_hw(net.freq[m / 3] * u[m % 3], ps[t], pc[t]);
        }
        uint16_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int f = f0 + j, t = (j + 1) >> 1;
            const float g = f < 3 ? u[f < 3 ? f : 0] : (f >= kFeat ? 0.0f : ((j & 1) ? ps[t] : pc[t]));
            v[j] = f2bf(g);
            if (sA0T) sA0T[f * kLds + 16 * w + (lane & 15)] = v[j];
        }
        uint4 pk;
        pk.x = v[0] | ((uint32_t)v[1] << 16);
        pk.y = v[2] | ((uint32_t)v[3] << 16);
        pk.z = v[4] | ((uint32_t)v[5] << 16);
        pk.w = v[6] | ((uint32_t)v[7] << 16);
        bf16x8 a = __builtin_bit_cast(bf16x8, pk);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            bf16x8 b = frag_ld(net.w.W1T + (size_t)(16 * c + (lane & 15)) * kK1 + 32 * ks + 8 * (lane >> 4));
            acc[c] = mfma(a, b, acc[c]);
        }
    }
}

// Bias + SiLU of a 64-wide layer output held in D layout; writes the BF16
// activation (row-major and, optionally, transposed) and keeps silu'(z).
__device__ __forceinline__ void activate(const float* __restrict__ bias, f32x4 acc[4], f32x4 dact[4], int lane,
                                         int w, uint16_t* sA, uint16_t* sAT, float* a_out = nullptr) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int n = 16 * c + (lane & 15);
        float bn = bias[n];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int s = 16 * w + 4 * (lane >> 4) + j;
            float z = acc[c][j] + bn;
            float sg = sigmoid(z);
            float a = z * sg;
            dact[c][j] = sg * (1.0f + z * (1.0f - sg));
            if (a_out) a_out[4 * c + j] = a;
            uint16_t ab = f2bf(a);
            if (sA) sA[s * kLds + n] = ab;
            if (sAT) sAT[n * kLds + s] = ab;
        }
    }
}

// Writes a back-propagated dz (D layout, FP32) as BF16, row-major and
// transposed.
__device__ __forceinline__ void store_dz(const f32x4 dz[4], int lane, int w, uint16_t* sdZ, uint16_t* sdZT) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int n = 16 * c + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int s = 16 * w + 4 * (lane >> 4) + j;
            uint16_t v = f2bf(dz[c][j]);
            sdZ[s * kLds + n] = v;
            if (sdZT) sdZT[n * kLds + s] = v;
        }
    }
}

// dW^T[n_out rows of wave w][k_in tiles] += dZ^T . A over the 64 samples.
template <int NC>
__device__ __forceinline__ void weight_grad(const uint16_t* sdZT, const uint16_t* sAT, int lane, int w,
                                            f32x4 acc[NC]) {
#pragma unroll
    for (int ks = 0; ks < kTile / 32; ++ks) {
        bf16x8 a = frag_ld(sdZT + (16 * w + (lane & 15)) * kLds + 32 * ks + 8 * (lane >> 4));
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            bf16x8 b = frag_ld(sAT + (16 * c + (lane & 15)) * kLds + 32 * ks + 8 * (lane >> 4));
            acc[c] = mfma(a, b, acc[c]);
        }
    }
}

// Sum over the 16 lanes that share lane/16 (one sample row group).
__device__ __forceinline__ float row_sum16(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

// Sum over the four lane groups (lane/16) with the same lane%16.
__device__ __forceinline__ float col_sum4(float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// One training pass over `batch` samples starting at `first` (wrapping at
// n): forward, loss, backward; each block accumulates the gradient of its
// tiles in registers and writes one partial gradient vector.
__global__ __launch_bounds__(kThreads) void k_mlp_train(const float4* __restrict__ samples, uint32_t n,
                                                        uint32_t first, uint32_t batch, NetView net,
                                                        float* __restrict__ partial,
                                                        float* __restrict__ loss_partial) {
    __shared__ __align__(16) uint16_t sA0T[kK1 * kLds];
    __shared__ __align__(16) uint16_t sA1[kTile * kLds], sA1T[kH * kLds];
    __shared__ __align__(16) uint16_t sA2[kTile * kLds], sA2T[kH * kLds];
    __shared__ __align__(16) uint16_t sdZ[kTile * kLds], sdZT[kH * kLds];
    __shared__ float sred[4][4 * kH + 4];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const float* th = net.theta;
    const float gscale = 2.0f / (float)batch;

    f32x4 gW1[8], gW2[4], gW3[4];
#pragma unroll
    for (int c = 0; c < 8; ++c) gW1[c] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; ++c) gW2[c] = gW3[c] = f32x4{0, 0, 0, 0};
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3[4] = {0, 0, 0, 0}, gw4[4] = {0, 0, 0, 0};
    float gb4 = 0.0f, lsum = 0.0f;
    float w4n[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) w4n[c] = th[ow4 + 16 * c + (lane & 15)];
    const float b4 = th[ob4];

    const uint32_t ntiles = batch / kTile;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // ---- forward ----
        const uint32_t sidx = (first + t * kTile + 16 * w + (lane & 15)) % n;
        const float4 smp = samples[sidx];
        float u[3];
        normalized(net, smp.x, smp.y, smp.z, u);
        f32x4 acc[4], D1[4], D2[4], D3[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
        input_layer(net, u, lane, w, acc, sA0T);
        activate(th + ob1, acc, D1, lane, w, sA1, sA1T);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
        rows_x_wt<kH, 4>(sA1, net.w.W2T, lane, w, acc);
        activate(th + ob2, acc, D2, lane, w, sA2, sA2T);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
        rows_x_wt<kH, 4>(sA2, net.w.W3T, lane, w, acc);
        float a3[16];
        activate(th + ob3, acc, D3, lane, w, nullptr, nullptr, a3);
        // ---- output layer and loss (rows s_j = 16w + 4(lane/16) + j) ----
        float g[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float part = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) part += a3[4 * c + j] * w4n[c];
            float y = row_sum16(part) + b4;
            // target of row s_j: held by lane 16*(s_j%16 / 16)... fetch it
            // from the lane that loaded that sample (lane%16 == row in tile)
            int row = 4 * (lane >> 4) + j;   // row within the wave's 16
            float d_true = __shfl(smp.w, row, 64);
            float e = y - target_of(net, d_true);
            g[j] = gscale * e;
            if ((lane & 15) == 0) lsum += e * e;
        }
        f32x4 dz[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                gw4[c] += g[j] * a3[4 * c + j];
                float v = g[j] * w4n[c] * D3[c][j];
                dz[c][j] = v;
                gb3[c] += v;
            }
        }
        if ((lane & 15) == 0) gb4 += g[0] + g[1] + g[2] + g[3];
        store_dz(dz, lane, w, sdZ, sdZT);
        __syncthreads();
        // ---- layer 3 gradient, back-propagate to layer 2 ----
        weight_grad<4>(sdZT, sA2T, lane, w, gW3);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
        rows_x_wt<kH, 4>(sdZ, net.w.W3, lane, w, acc);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dz[c][j] = acc[c][j] * D2[c][j];
                gb2[c] += dz[c][j];
            }
        store_dz(dz, lane, w, sdZ, sdZT);
        __syncthreads();
        weight_grad<4>(sdZT, sA1T, lane, w, gW2);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
        rows_x_wt<kH, 4>(sdZ, net.w.W2, lane, w, acc);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dz[c][j] = acc[c][j] * D1[c][j];
                gb1[c] += dz[c][j];
            }
        store_dz(dz, lane, w, sdZ, sdZT);
        __syncthreads();
        weight_grad<8>(sdZT, sA0T, lane, w, gW1);
        __syncthreads();   // LDS tiles are rewritten by the next tile
    }

    // ---- write this block's partial gradient ----
    float* out = partial + (size_t)blockIdx.x * kParams;
    // dW^T accumulators: row n_out = 16w + 4(lane/16) + j, col k_in = 16c + lane%16
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int no = 16 * w + 4 * (lane >> 4) + j, ki = 16 * c + (lane & 15);
            out[oW1 + ki * kH + no] = gW1[c][j];
        }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int no = 16 * w + 4 * (lane >> 4) + j, ki = 16 * c + (lane & 15);
            out[oW2 + ki * kH + no] = gW2[c][j];
            out[oW3 + ki * kH + no] = gW3[c][j];
        }
    // column sums: reduce over lane groups, then over the four waves
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float v1 = col_sum4(gb1[c]), v2 = col_sum4(gb2[c]), v3 = col_sum4(gb3[c]), v4 = col_sum4(gw4[c]);
        if (lane < 16) {
            int n = 16 * c + lane;
            sred[w][n] = v1;
            sred[w][kH + n] = v2;
            sred[w][2 * kH + n] = v3;
            sred[w][3 * kH + n] = v4;
        }
    }
    float l4 = col_sum4(gb4), ls = col_sum4(lsum);
    if (lane == 0) {
        sred[w][4 * kH] = l4;
        sred[w][4 * kH + 1] = ls;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4 * kH + 2; i += kThreads) {
        float v = sred[0][i] + sred[1][i] + sred[2][i] + sred[3][i];
        if (i < kH) out[ob1 + i] = v;
        else if (i < 2 * kH) out[ob2 + i - kH] = v;
        else if (i < 3 * kH) out[ob3 + i - 2 * kH] = v;
        else if (i < 4 * kH) out[ow4 + i - 3 * kH] = v;
        else if (i == 4 * kH) out[ob4] = v;
        else loss_partial[blockIdx.x] = v;
    }
}

// Sums the partial gradients; with adam != 0 applies one Adam step
// (Kingma & Ba 2015, Algorithm 1) to the FP32 master parameters and refreshes
// the BF16 copies.  grad_out (optional) receives the summed gradient.  A
// block owns 64 parameters; its four wavefronts sum interleaved quarters of
// the partials (four independent accumulators each) and meet in LDS.
constexpr int kAdamParams = 64;
__global__ __launch_bounds__(256) void k_mlp_adam(const float* __restrict__ partial, int nparts,
                                                  const float* __restrict__ loss_partial, float* theta, float* m,
                                                  float* v, Bf16Weights wb, float lr, float b1, float b2, float eps,
                                                  float c1, float c2, int adam, float* grad_out,
                                                  float* loss_out, float inv_batch) {
    __shared__ float sg[4][kAdamParams];
    const int pl = threadIdx.x & (kAdamParams - 1), grp = threadIdx.x >> 6;
    const int p = blockIdx.x * kAdamParams + pl;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) {
        float l = 0.0f;
        for (int b = 0; b < nparts; ++b) l += loss_partial[b];
        *loss_out = l * inv_batch;
    }
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (p < kParams) {
        int b = grp;
        for (; b + 12 < nparts; b += 16) {
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] += partial[(size_t)(b + 4 * u) * kParams + p];
        }
        for (; b < nparts; b += 4) acc[0] += partial[(size_t)b * kParams + p];
    }
    sg[grp][pl] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (grp != 0 || p >= kParams) return;
    const float g = (sg[0][pl] + sg[1][pl]) + (sg[2][pl] + sg[3][pl]);
    if (grad_out) grad_out[p] = g;
    if (!adam) return;
    float mm = b1 * m[p] + (1.0f - b1) * g;
    float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    float mh = mm / c1, vh = vv / c2;
    float th = theta[p] - lr * mh / (sqrtf(vh) + eps);
    theta[p] = th;
    uint16_t hb = f2bf(th);
    if (p < ob1) {
        int k = (p - oW1) / kH, nn = (p - oW1) % kH;
        wb.W1T[nn * kK1 + k] = hb;
        wb.W1[k * kH + nn] = hb;
    } else if (p >= oW2 && p < ob2) {
        int k = (p - oW2) / kH, nn = (p - oW2) % kH;
        wb.W2T[nn * kH + k] = hb;
        wb.W2[k * kH + nn] = hb;
    } else if (p >= oW3 && p < ob3) {
        int k = (p - oW3) / kH, nn = (p - oW3) % kH;
        wb.W3T[nn * kH + k] = hb;
        wb.W3[k * kH + nn] = hb;
    }
}

// BF16 copies of all weights from the FP32 master (after set_params).
__global__ void k_mlp_copies(const float* theta, Bf16Weights wb) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    uint16_t hb = f2bf(theta[p]);
    if (p < ob1) {
        int k = (p - oW1) / kH, nn = (p - oW1) % kH;
        wb.W1T[nn * kK1 + k] = hb;
        wb.W1[k * kH + nn] = hb;
    } else if (p >= oW2 && p < ob2) {
        int k = (p - oW2) / kH, nn = (p - oW2) % kH;
        wb.W2T[nn * kH + k] = hb;
        wb.W2[k * kH + nn] = hb;
    } else if (p >= oW3 && p < ob3) {
        int k = (p - oW3) / kH, nn = (p - oW3) % kH;
        wb.W3T[nn * kH + k] = hb;
        wb.W3[k * kH + nn] = hb;
    }
}

// Network output y and dy/du of the 64-point tile whose row 16w + lane%16
// has normalized coordinates u (forward, then back-propagation to the
// input through the three hidden layers and the feature map).  Rows
// s_j = 16w + 4(lane/16) + j: y[j], gu[j][0..2] (every lane of a row group
// holds them).  Uses three 64 x kLds BF16 LDS tiles.
__device__ __forceinline__ void mlp_value_grad(const NetView& net, const float u[3], int lane, int w, uint16_t* sA1,
                                               uint16_t* sA2, uint16_t* sdZ, float y[4], float gu[4][3]) {
    const float* th = net.theta;
    f32x4 acc[4], D1[4], D2[4], D3[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
    input_layer(net, u, lane, w, acc, nullptr);
    activate(th + ob1, acc, D1, lane, w, sA1, nullptr);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
    rows_x_wt<kH, 4>(sA1, net.w.W2T, lane, w, acc);
    activate(th + ob2, acc, D2, lane, w, sA2, nullptr);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
    rows_x_wt<kH, 4>(sA2, net.w.W3T, lane, w, acc);
    float a3[16];
    activate(th + ob3, acc, D3, lane, w, nullptr, nullptr, a3);
    f32x4 dz[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float part = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) part += a3[4 * c + j] * th[ow4 + 16 * c + (lane & 15)];
        y[j] = row_sum16(part) + th[ob4];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) dz[c][j] = th[ow4 + 16 * c + (lane & 15)] * D3[c][j];
    store_dz(dz, lane, w, sdZ, nullptr);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
    rows_x_wt<kH, 4>(sdZ, net.w.W3, lane, w, acc);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) dz[c][j] = acc[c][j] * D2[c][j];
    store_dz(dz, lane, w, sdZ, nullptr);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
    rows_x_wt<kH, 4>(sdZ, net.w.W2, lane, w, acc);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) dz[c][j] = acc[c][j] * D1[c][j];
    store_dz(dz, lane, w, sdZ, nullptr);
    __syncthreads();
    // dy/dg = dz1 . W1^T (64 x 128), then dy/du through the feature map
    f32x4 dg[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) dg[c] = f32x4{0, 0, 0, 0};
    rows_x_wt<kH, 8>(sdZ, net.w.W1, lane, w, dg);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int row = 4 * (lane >> 4) + j;
        float uj[3];
        uj[0] = __shfl(u[0], row, 64);
        uj[1] = __shfl(u[1], row, 64);
        uj[2] = __shfl(u[2], row, 64);
        gu[j][0] = gu[j][1] = gu[j][2] = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            int axis;
            float jac = feature_du(net, 16 * c + (lane & 15), uj, axis);
            float v = dg[c][j] * jac;
            gu[j][0] += axis == 0 ? v : 0.0f;
            gu[j][1] += axis == 1 ? v : 0.0f;
            gu[j][2] += axis == 2 ? v : 0.0f;
        }
        gu[j][0] = row_sum16(gu[j][0]);
        gu[j][1] = row_sum16(gu[j][1]);
        gu[j][2] = row_sum16(gu[j][2]);
    }
}

// Network output y of the 64-point tile (forward only).  Rows
// s_j = 16w + 4(lane/16) + j: y[j] (every lane of a row group holds them).
// Uses two 64 x kLds BF16 LDS tiles; starts with a barrier so that a tile
// can follow another in the same block.
__device__ __forceinline__ void mlp_value(const NetView& net, const float u[3], int lane, int w, uint16_t* sA1,
                                          uint16_t* sA2, float y[4]) {
    const float* th = net.theta;
    f32x4 acc[4], dact[4];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
    input_layer(net, u, lane, w, acc, nullptr);
    activate(th + ob1, acc, dact, lane, w, sA1, nullptr);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
    rows_x_wt<kH, 4>(sA1, net.w.W2T, lane, w, acc);
    activate(th + ob2, acc, dact, lane, w, sA2, nullptr);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0, 0, 0, 0};
    rows_x_wt<kH, 4>(sA2, net.w.W3T, lane, w, acc);
    float a3[16];
    activate(th + ob3, acc, dact, lane, w, nullptr, nullptr, a3);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float part = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) part += a3[4 * c + j] * th[ow4 + 16 * c + (lane & 15)];
        y[j] = row_sum16(part) + th[ob4];
    }
}

// The wall normal the step takes from the network: the central difference
// of the network's distance at +-spacing along each axis, in world units --
// the stencil of the grid's normal (particle_simulation.cu:429-439: lookups
// +-spacing/L apart in normalized coordinates) applied to the network's
// field, so that the network stands in for exactly the quantity the grid
// gives.  Six forward passes of the tile (point x of row 16w + lane%16);
// g[j] for rows s_j as in mlp_value.
__device__ __forceinline__ void net_cd_normal(const NetView& net, float4 x, float spacing, int lane, int w,
                                              uint16_t* sA1, uint16_t* sA2, float g[4][3]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) g[j][0] = g[j][1] = g[j][2] = 0.0f;
    const float inv2s = 0.5f / spacing;
    for (int a = 0; a < 3; ++a)
        for (int sgn = 1; sgn >= -1; sgn -= 2) {
            const float o = (float)sgn * spacing;
            float u[3], y[4];
            normalized(net, x.x + (a == 0 ? o : 0.0f), x.y + (a == 1 ? o : 0.0f), x.z + (a == 2 ? o : 0.0f), u);
            mlp_value(net, u, lane, w, sA1, sA2, y);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float sc;
                g[j][a] += (float)sgn * decode(net, y[j], sc) * inv2s;
            }
        }
}

// Distance and its gradient at n points (float4 xyz_): one tile of 64
// points per block.
__global__ __launch_bounds__(kThreads) void k_mlp_eval(const float4* __restrict__ pts, uint32_t n, NetView net,
                                                       float4* __restrict__ out) {
    __shared__ __align__(16) uint16_t sA1[kTile * kLds], sA2[kTile * kLds], sdZ[kTile * kLds];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t base = blockIdx.x * kTile;
    const uint32_t my = base + 16 * w + (lane & 15);
    const float4 x = pts[my < n ? my : n - 1];
    float u[3], y[4], gu[4][3];
    normalized(net, x.x, x.y, x.z, u);
    mlp_value_grad(net, u, lane, w, sA1, sA2, sdZ, y, gu);
    if ((lane & 15) < 4) {
        const int j = lane & 15;
        const uint32_t s = base + 16 * w + 4 * (lane >> 4) + j;
        if (s < n) {
            float sc;
            float d = decode(net, y[j], sc);
            out[s] = make_float4(d, sc * gu[j][0] * 2.0f * net.inv_ext[0], sc * gu[j][1] * 2.0f * net.inv_ext[1],
                                 sc * gu[j][2] * 2.0f * net.inv_ext[2]);
        }
    }
}

// ---- learned boundary in the step ------------------------------------------

// Per-step distance records (gradient, value) of the particles in
// learned-source cells that the density pass found in the wall-penalty
// zone, one tile of 64 sorted particles per block; tiles without such a
// particle exit at once.
// Particles that take the network's normal this step (the density pass
// flagged them in the penalty zone, their cell's source is the network),
// packed per scene: scene s's entries at list[p_begin_s ..] with count
// cnt[s] (zeroed by the density pass).  A wavefront whose slots lie in one
// scene appends with one
// atomic; the list order does not matter (each particle's row of the
// network is computed independently of the other rows of its tile).
__global__ __launch_bounds__(256) void k_wall_select(uint32_t n, const uint32_t* __restrict__ skey,
                                                     const uint8_t* __restrict__ mode,
                                                     const uint8_t* __restrict__ need,
                                                     const DevScene* __restrict__ scenes,
                                                     const uint32_t* __restrict__ begins, int nscenes,
                                                     const char* __restrict__ enabled, uint32_t* __restrict__ cnt,
                                                     uint32_t* __restrict__ list) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if ((s & ~63u) >= n) return;   // the whole wavefront lies past the end
    const int lane = threadIdx.x & 63;
    const bool valid = s < n;
    const int sc = valid ? scene_of(begins, nscenes, s) : -1;
    const bool want = valid && enabled[sc] && need[s] && mode[skey[s]] == kWallNeural;
    const uint64_t wants = __ballot(want);
    if (!wants) return;
    const int sc0 = __shfl(sc, 0, 64);
    const uint64_t same = __ballot(!valid || sc == sc0);
    if (same == ~0ull) {
        // one scene: the first wanting lane reserves the run
        const int leader = __ffsll((unsigned long long)wants) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cnt[sc0], (uint32_t)__popcll(wants));
        base = __shfl(base, leader, 64);
        if (want) {
            const uint64_t below = lane == 0 ? 0ull : (wants & (~0ull >> (64 - lane)));
            list[scenes[sc0].p_begin + base + (uint32_t)__popcll(below)] = s;
        }
    } else if (want) {
        list[scenes[sc].p_begin + atomicAdd(&cnt[sc], 1u)] = s;
    }
}

// Network value and gradient for the listed particles: block b takes tile
// (scene, p0) of the scene's slot range, i.e. list entries
// [p0 - p_begin, p0 - p_begin + 64) of that scene; past the count it exits.
__global__ __launch_bounds__(kThreads) void k_wall_mlp(const uint2* __restrict__ tiles,
                                                       const float4* __restrict__ pos,
                                                       const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ cnt,
                                                       const DevScene* __restrict__ scenes,
                                                       const NetView* __restrict__ nets,
                                                       float4* __restrict__ rec) {
    __shared__ __align__(16) uint16_t sA1[kTile * kLds], sA2[kTile * kLds];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint2 t = tiles[blockIdx.x];
    const DevScene& S = scenes[t.x];
    const uint32_t k0 = t.y - S.p_begin, count = cnt[t.x];
    if (k0 >= count) return;   // uniform over the block
    const uint32_t* L = list + S.p_begin;
    const uint32_t k = k0 + 16 * w + (lane & 15);
    const float4 x = pos[L[k < count ? k : k0]];
    const NetView net = nets[t.x];
    float g[4][3];
    net_cd_normal(net, x, S.bspacing, lane, w, sA1, sA2, g);
    if ((lane & 15) < 4) {
        const int j = lane & 15;
        const uint32_t kj = k0 + 16 * w + 4 * (lane >> 4) + j;
        if (kj < count) rec[L[kj]] = make_float4(g[j][0], g[j][1], g[j][2], 0.0f);
    }
}

// Validation: per brick of cells, the sum of squared errors of the network
// (and their count) over the given (x, y, z, distance texture) samples
// where the label exceeds -band (the field is the constant -r0 farther out
// on the fluid side); and, where the label exceeds -r0 (the wall-penalty
// zone, where the step takes the network's normal), the squared difference
// between the network's normal and the one it replaces (both central
// differences at +-spacing, world units), weighted by the penetration.
__global__ __launch_bounds__(kThreads) void k_validate(const float4* __restrict__ smp, uint32_t n, DevScene S,
                                                       NetView net, float band, int brick, int nbx, int nby,
                                                       int nbz, const float* __restrict__ gridd,
                                                       float* __restrict__ sq_tight,
                                                       uint32_t* __restrict__ count, float* __restrict__ sq_grad,
                                                       uint32_t* __restrict__ count_grad, float* __restrict__ sums) {
    __shared__ __align__(16) uint16_t sA1[kTile * kLds], sA2[kTile * kLds], sdZ[kTile * kLds];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t base = blockIdx.x * kTile;
    const uint32_t my = base + 16 * w + (lane & 15);
    const float4 x = smp[my < n ? my : n - 1];
    float u[3], y[4], gu[4][3], gn[4][3];
    normalized(net, x.x, x.y, x.z, u);
    mlp_value_grad(net, u, lane, w, sA1, sA2, sdZ, y, gu);
    net_cd_normal(net, x, S.bspacing, lane, w, sA1, sA2, gn);
    if ((lane & 15) < 4) {
        const int j = lane & 15;
        const int row = 4 * (lane >> 4) + j;
        const uint32_t sj = base + 16 * w + row;
        const float4 xs = smp[sj < n ? sj : n - 1];
        if (sj < n) {
            float sc;
            const float d = decode(net, y[j], sc);
            const float e = fabsf(d - xs.w);
            int cx = (int)floorf((xs.x - S.ox) / S.cell), cy = (int)floorf((xs.y - S.oy) / S.cell),
                cz = (int)floorf((xs.z - S.oz) / S.cell);
            cx = min(max(cx, 0), S.nx - 1);
            cy = min(max(cy, 0), S.ny - 1);
            cz = min(max(cz, 0), S.nz - 1);
            const uint32_t b = (uint32_t)((cz / brick) * nby + cy / brick) * nbx + cx / brick;
            if (xs.w > -band) {
                atomicAdd(&sq_tight[b], e * e);
                atomicAdd(&count[b], 1u);
                atomicMax(reinterpret_cast<uint32_t*>(&sums[2]), __float_as_uint(e));
                atomicAdd(&sums[0], e);
                atomicAdd(&sums[1], 1.0f);
            }
            if (xs.w > -S.r0) {
                // the step's two normals: the central difference of the
                // network's distance (as k_wall_mlp records it) and the
                // grid's normalized-coordinate central difference divided by
                // L (as the collision uses it), both in world units
                float tx, ty, tz;
                texel_coords(S, xs.x, xs.y, xs.z, tx, ty, tz);
                float3 gg = dist_grad_normalized(gridd, S, tx, ty, tz);
                gg.x *= S.bscale_x * (1.0f / (float)S.bnx);
                gg.y *= S.bscale_y * (1.0f / (float)S.bny);
                gg.z *= S.bscale_z * (1.0f / (float)S.bnz);
                const float nx = gn[j][0], ny = gn[j][1], nz = gn[j][2];
                // the normal scales the penalty rho/dt^2 (r0 - distWall) G
                // (cu:422-445): weighted by the penetration (r0 + D)/r0 the
                // error is that of the penalty force relative to its value
                // at a penetration of r0
                const float wgt = fminf(fmaxf((xs.w + S.r0) / S.r0, 0.0f), 1.0f);
                const float e2 = wgt * wgt *
                                 ((nx - gg.x) * (nx - gg.x) + (ny - gg.y) * (ny - gg.y) + (nz - gg.z) * (nz - gg.z));
                atomicAdd(&sq_grad[b], e2);
                atomicAdd(&count_grad[b], 1u);
                atomicAdd(&sums[3], sqrtf(e2));
                atomicAdd(&sums[4], 1.0f);
            }
        }
    }
}

// Cells of bricks whose RMS distance error exceeds tol, or whose RMS normal
// error exceeds tol_n (tol_n > 0), take the grid.
__device__ __forceinline__ bool brick_fails(uint32_t b, const float* sq_tight, const uint32_t* count,
                                            const float* sq_grad, const uint32_t* count_grad, float tol,
                                            float tol_n) {
    const uint32_t k = count[b], kg = count_grad[b];
    return (k > 0 && sq_tight[b] > tol * tol * (float)k) ||
           (tol_n > 0.0f && kg > 0 && sq_grad[b] > tol_n * tol_n * (float)kg);
}

__global__ void k_apply_bricks(DevScene S, const float* __restrict__ sq_tight, const uint32_t* __restrict__ count,
                               const float* __restrict__ sq_grad, const uint32_t* __restrict__ count_grad, float tol,
                               float tol_n, int brick, int nbx, int nby, uint8_t* __restrict__ mode) {
    const uint32_t ncell = (uint32_t)S.nx * S.ny * S.nz;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int cx = (int)(c % S.nx), cy = (int)((c / S.nx) % S.ny), cz = (int)(c / ((uint32_t)S.nx * S.ny));
    const uint32_t b = (uint32_t)((cz / brick) * nby + cy / brick) * nbx + cx / brick;
    if (mode[c] == kWallNeural && brick_fails(b, sq_tight, count, sq_grad, count_grad, tol, tol_n))
        mode[c] = kWallGrid;
}


}  // namespace nn
}  // namespace sph

struct sph_mlp {
    sph_mlp_config cfg{};
    float* theta = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* wbuf = nullptr;
    sph::nn::Bf16Weights wb{};
    float4* samples = nullptr;
    uint32_t nsamples = 0;
    float* partial = nullptr;
    float* loss_partial = nullptr;
    float* grad = nullptr;
    float* loss = nullptr;       // [loss_cap] per-step losses
    int loss_cap = 0;
    int nblocks = 0;
    uint32_t cursor = 0;         // next sample of the training stream
    uint32_t step = 0;           // Adam steps taken
    hipStream_t stream = nullptr;
    std::string err;

    sph_status fail(sph_status s, const std::string& m) {
        err = m;
        return s;
    }
    sph::nn::NetView view() const {
        sph::nn::NetView nv;
        nv.theta = theta;
        nv.w = wb;
        for (int a = 0; a < 3; ++a) {
            nv.lo[a] = cfg.box_lo[a];
            nv.inv_ext[a] = 1.0f / (cfg.box_hi[a] - cfg.box_lo[a]);
        }
        nv.out_scale = cfg.out_scale;
        nv.band = cfg.band;
        for (int k = 0; k < sph::nn::kFreqs; ++k) nv.freq[k] = (float)std::pow(2.0, 7.0 * k / 15.0);
        return nv;
    }
};

namespace sph {

// Engine-side state of the learned boundary.
struct NeuralBoundary {
    float4* rec = nullptr;        // [N]
    uint8_t* need = nullptr;      // [N]
    uint32_t* list = nullptr;     // [N] particles for the network, packed per scene
    uint32_t* cnt = nullptr;      // [nscenes] list lengths
    char* d_enabled = nullptr;    // [nscenes] scenes with a network
    nn::NetView* nets = nullptr;  // [nscenes]
    uint2* tiles = nullptr;
    uint32_t ntiles = 0;
    std::vector<nn::NetView> hnets;
    std::vector<char> enabled;
    size_t bytes = 0;
};

sph_status neural_destroy(Engine& e, bool restore) {
    NeuralBoundary* nb = e.nb;
    if (!nb) return SPH_OK;
    (void)hipStreamSynchronize(e.stream);
    void* ptrs[] = {nb->rec, nb->need, nb->list, nb->cnt, nb->d_enabled, nb->nets, nb->tiles};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    // the cells of the scenes that had a network go back to the grid
    for (int s = 0; s < e.nscenes; ++s)
        if (nb->enabled[s]) launch_wall_modes(e.stream, e.hscenes[s], e.grid4, kWallGrid, e.wall_mode);
    (void)hipStreamSynchronize(e.stream);
    delete nb;
    e.nb = nullptr;
    return restore ? e.own_distance() : SPH_OK;
}

size_t neural_bytes(const Engine& e) { return e.nb ? e.nb->bytes : 0; }

const float4* neural_records(const Engine& e) { return e.nb && e.nb->ntiles ? e.nb->rec : nullptr; }

uint8_t* neural_need(const Engine& e) { return e.nb && e.nb->ntiles ? e.nb->need : nullptr; }
uint32_t* neural_list_count(const Engine& e) { return e.nb && e.nb->ntiles ? e.nb->cnt : nullptr; }

void neural_enqueue(Engine& e) {
    NeuralBoundary* nb = e.nb;
    if (!nb || !nb->ntiles) return;
    hipLaunchKernelGGL(nn::k_wall_select, dim3((e.N + 255) / 256), dim3(256), 0, e.stream, e.N, e.skey, e.wall_mode,
                       nb->need, e.d_scenes, e.d_begins, e.nscenes, nb->d_enabled, nb->cnt, nb->list);
    hipLaunchKernelGGL(nn::k_wall_mlp, dim3(nb->ntiles), dim3(nn::kThreads), 0, e.stream, nb->tiles, e.pos_b,
                       nb->list, nb->cnt, e.d_scenes, nb->nets, nb->rec);
}

static sph_status rebuild_tiles(Engine& e) {
    NeuralBoundary* nb = e.nb;
    std::vector<uint2> t;
    for (int s = 0; s < e.nscenes; ++s) {
        if (!nb->enabled[s]) continue;
        for (uint32_t p = e.hbegins[s]; p < e.hbegins[s + 1]; p += nn::kTile) t.push_back(make_uint2((unsigned)s, p));
    }
    if (nb->tiles) (void)hipFree(nb->tiles);
    nb->tiles = nullptr;
    nb->ntiles = 0;
    if (t.empty()) return SPH_OK;
    if (hipMalloc(&nb->tiles, sizeof(uint2) * t.size()) != hipSuccess ||
        hipMemcpyAsync(nb->tiles, t.data(), sizeof(uint2) * t.size(), hipMemcpyHostToDevice, e.stream) != hipSuccess)
        return e.fail(SPH_OUT_OF_MEMORY, "tile list");
    nb->ntiles = (uint32_t)t.size();
    return SPH_OK;
}

sph_status neural_attach(Engine& e, int scene, sph_mlp* m, const float* val, size_t nval,
                         const sph_neural_config& cfg, sph_neural_report* rep) {
    if (scene < 0 || scene >= e.nscenes) return e.fail(SPH_INVALID_ARG, "scene index out of range");
    if (!m || (nval && !val)) return e.fail(SPH_INVALID_ARG, "network required");
    if (!(cfg.tolerance > 0.0f) || !(cfg.band > 0.0f) || cfg.brick < 1)
        return e.fail(SPH_INVALID_ARG, "tolerance and band must be positive and brick >= 1");
    if (!e.nb) {
        NeuralBoundary* nb = new (std::nothrow) NeuralBoundary();
        if (!nb) return SPH_OUT_OF_MEMORY;
        e.nb = nb;
        nb->hnets.resize(e.nscenes);
        nb->enabled.assign(e.nscenes, 0);
        bool ok = hipMalloc(&nb->rec, sizeof(float4) * (size_t)e.N) == hipSuccess &&
                  hipMalloc(&nb->need, (size_t)e.N) == hipSuccess &&
                  hipMalloc(&nb->list, sizeof(uint32_t) * (size_t)e.N) == hipSuccess &&
                  hipMalloc(&nb->cnt, sizeof(uint32_t) * e.nscenes) == hipSuccess &&
                  hipMalloc(&nb->d_enabled, e.nscenes) == hipSuccess &&
                  hipMemset(nb->d_enabled, 0, e.nscenes) == hipSuccess &&
                  hipMalloc(&nb->nets, sizeof(nn::NetView) * e.nscenes) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            (void)neural_destroy(e);
            return e.fail(SPH_OUT_OF_MEMORY, "learned boundary buffers");
        }
        nb->bytes = (sizeof(float4) + 1 + sizeof(uint32_t)) * (size_t)e.N;
    }
    NeuralBoundary* nb = e.nb;
    const DevScene& S = e.hscenes[scene];
    nn::NetView nv = m->view();
    nb->hnets[scene] = nv;
    if (hipMemcpyAsync(nb->nets + scene, &nv, sizeof(nv), hipMemcpyHostToDevice, e.stream) != hipSuccess)
        return e.fail(SPH_DEVICE_ERROR, "network upload");
    // cell modes: far-field cells from the grid, the rest take the network
    const uint32_t ncell = (uint32_t)S.nx * S.ny * S.nz;
    uint8_t* mode = e.wall_mode + S.cell_offset;
    launch_wall_modes(e.stream, S, e.grid4, kWallNeural, e.wall_mode);
    // validation per brick
    const int br = cfg.brick;
    const int nbx = (S.nx + br - 1) / br, nby = (S.ny + br - 1) / br, nbz = (S.nz + br - 1) / br;
    const size_t nbrick = (size_t)nbx * nby * nbz;
    float* et = nullptr;
    uint32_t* cnt = nullptr;
    float* eg = nullptr;
    uint32_t* cntg = nullptr;
    float* sums = nullptr;
    float4* dval = nullptr;
    sph_status st = SPH_OK;
    if (hipMalloc(&et, sizeof(float) * nbrick) != hipSuccess ||
        hipMalloc(&cnt, sizeof(uint32_t) * nbrick) != hipSuccess ||
        hipMalloc(&eg, sizeof(float) * nbrick) != hipSuccess ||
        hipMalloc(&cntg, sizeof(uint32_t) * nbrick) != hipSuccess || hipMalloc(&sums, sizeof(float) * 5) != hipSuccess ||
        (nval && hipMalloc(&dval, sizeof(float4) * nval) != hipSuccess)) {
        (void)hipGetLastError();
        st = e.fail(SPH_OUT_OF_MEMORY, "validation buffers");
    }
    if (st == SPH_OK) {
        (void)hipMemsetAsync(et, 0, sizeof(float) * nbrick, e.stream);
        (void)hipMemsetAsync(cnt, 0, sizeof(uint32_t) * nbrick, e.stream);
        (void)hipMemsetAsync(eg, 0, sizeof(float) * nbrick, e.stream);
        (void)hipMemsetAsync(cntg, 0, sizeof(uint32_t) * nbrick, e.stream);
        (void)hipMemsetAsync(sums, 0, sizeof(float) * 5, e.stream);
        if (nval) {
            (void)hipMemcpyAsync(dval, val, sizeof(float4) * nval, hipMemcpyHostToDevice, e.stream);
            hipLaunchKernelGGL(nn::k_validate, dim3((unsigned)((nval + nn::kTile - 1) / nn::kTile)), dim3(nn::kThreads),
                               0, e.stream, dval, (uint32_t)nval, S, nv, cfg.band, br, nbx, nby, nbz, e.gridd, et, cnt,
                               eg, cntg, sums);
        }
        hipLaunchKernelGGL(nn::k_apply_bricks, dim3((ncell + 255) / 256), dim3(256), 0, e.stream, S, et, cnt, eg, cntg,
                           cfg.tolerance, cfg.normal_tolerance, br, nbx, nby, mode);
        if (hipStreamSynchronize(e.stream) != hipSuccess) st = e.fail(SPH_DEVICE_ERROR, "attach kernels failed");
    }
    if (st == SPH_OK && rep) {
        std::vector<uint8_t> hm(ncell);
        std::vector<float> het(nbrick), heg(nbrick);
        std::vector<uint32_t> hc(nbrick), hcg(nbrick);
        float hs[5] = {0, 0, 0, 0, 0};
        if (hipMemcpy(hm.data(), mode, ncell, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(het.data(), et, sizeof(float) * nbrick, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hc.data(), cnt, sizeof(uint32_t) * nbrick, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(heg.data(), eg, sizeof(float) * nbrick, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hcg.data(), cntg, sizeof(uint32_t) * nbrick, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hs, sums, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess) {
            st = e.fail(SPH_DEVICE_ERROR, "report read");
        } else {
            std::memset(rep, 0, sizeof(*rep));
            for (uint8_t v : hm) {
                if (v == kWallNone) ++rep->cells_none;
                else if (v == kWallNeural) ++rep->cells_neural;
                else ++rep->cells_grid;
            }
            const float tn = cfg.normal_tolerance;
            for (size_t b = 0; b < nbrick; ++b) {
                if (hc[b] || hcg[b]) ++rep->bricks_checked;
                const bool bad_d = hc[b] && het[b] > cfg.tolerance * cfg.tolerance * (float)hc[b];
                const bool bad_n = tn > 0.0f && hcg[b] && heg[b] > tn * tn * (float)hcg[b];
                if (bad_d || bad_n) ++rep->bricks_fallback;
                if (bad_n && !bad_d) ++rep->bricks_fallback_normal;
            }
            std::memcpy(&rep->max_error, &hs[2], 4);
            rep->mean_error = hs[1] > 0.0f ? hs[0] / hs[1] : 0.0f;
            rep->samples_in_band = (uint64_t)hs[1];
            rep->mean_normal_error = hs[4] > 0.0f ? hs[3] / hs[4] : 0.0f;
        }
    }
    void* tmp[] = {et, cnt, eg, cntg, sums, dval};
    for (void* p : tmp)
        if (p) (void)hipFree(p);
    if (st != SPH_OK) return st;
    nb->enabled[scene] = 1;
    const char on = 1;
    if (hipMemcpy(nb->d_enabled + scene, &on, 1, hipMemcpyHostToDevice) != hipSuccess)
        return e.fail(SPH_DEVICE_ERROR, "network flag upload");
    st = rebuild_tiles(e);
    if (st != SPH_OK) return st;
    // every scene's normals come from a network: the float distance copy goes
    bool all = true;
    for (int s = 0; s < e.nscenes; ++s) all = all && nb->enabled[s];
    if (all && (st = e.alias_distance()) != SPH_OK) return st;
    e.drop_graph();
    return hipStreamSynchronize(e.stream) == hipSuccess ? SPH_OK : e.fail(SPH_DEVICE_ERROR, "attach sync");
}

}  // namespace sph

// ---------------------------------------------------------------------------
// Host side

using namespace sph::nn;

static std::string g_mlp_error;

static uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static float uniform(uint64_t& s) { return (float)((splitmix(s) >> 40) * (1.0 / 16777216.0)); }

extern "C" {

sph_status sph_mlp_default_config(const float box_lo[3], const float box_hi[3], sph_mlp_config* out) {
    if (!box_lo || !box_hi || !out) return SPH_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    float ext = 0.0f;
    for (int a = 0; a < 3; ++a) {
        out->box_lo[a] = box_lo[a];
        out->box_hi[a] = box_hi[a];
        if (!(box_hi[a] > box_lo[a])) return SPH_INVALID_ARG;
        ext = std::max(ext, box_hi[a] - box_lo[a]);
    }
    out->out_scale = 0.5f * ext;
    out->seed = 1;
    out->lr = 1e-3f;
    out->beta1 = 0.9f;
    out->beta2 = 0.999f;
    out->eps = 1e-8f;
    return SPH_OK;
}

sph_status sph_mlp_create(const sph_mlp_config* cfg, sph_mlp** out) {
    if (!cfg || !out) return SPH_INVALID_ARG;
    *out = nullptr;
    for (int a = 0; a < 3; ++a)
        if (!(cfg->box_hi[a] > cfg->box_lo[a])) {
            g_mlp_error = "empty box";
            return SPH_INVALID_ARG;
        }
    if (!(cfg->out_scale > 0.0f) || !(cfg->band >= 0.0f) || cfg->decay_steps < 0) {
        g_mlp_error = "out_scale must be positive, band and decay_steps non-negative";
        return SPH_INVALID_ARG;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        g_mlp_error = "no HIP device";
        return SPH_NO_DEVICE;
    }
    sph_mlp* h = new (std::nothrow) sph_mlp();
    if (!h) return SPH_OUT_OF_MEMORY;
    h->cfg = *cfg;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    h->nblocks = 2 * cus;
    const size_t nb = (size_t)2 * kK1 * kH + (size_t)4 * kH * kH;
    bool ok = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&h->theta, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&h->m, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&h->v, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&h->wbuf, sizeof(uint16_t) * nb) == hipSuccess &&
              hipMalloc(&h->partial, sizeof(float) * kParams * (size_t)h->nblocks) == hipSuccess &&
              hipMalloc(&h->loss_partial, sizeof(float) * (size_t)h->nblocks) == hipSuccess &&
              hipMalloc(&h->grad, sizeof(float) * kParams) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        g_mlp_error = "device allocation failed";
        sph_mlp_destroy(h);
        return SPH_OUT_OF_MEMORY;
    }
    h->wb.W1T = h->wbuf;
    h->wb.W1 = h->wb.W1T + kK1 * kH;
    h->wb.W2T = h->wb.W1 + kK1 * kH;
    h->wb.W2 = h->wb.W2T + kH * kH;
    h->wb.W3T = h->wb.W2 + kH * kH;
    h->wb.W3 = h->wb.W3T + kH * kH;
    // Xavier-uniform weights from the seed, zero biases, zero padding rows
    std::vector<float> th(kParams, 0.0f);
    uint64_t s = cfg->seed;
    auto fill = [&](int off, int rows, int fan_in, int fan_out) {
        float lim = std::sqrt(6.0f / (float)(fan_in + fan_out));
        for (int k = 0; k < rows; ++k)
            for (int n = 0; n < fan_out; ++n) th[off + k * fan_out + n] = (2.0f * uniform(s) - 1.0f) * lim;
    };
    fill(oW1, kFeat, kFeat, kH);
    fill(oW2, kH, kH, kH);
    fill(oW3, kH, kH, kH);
    fill(ow4, 1, kH, 1);   // w4 as a 1 x 64 row with fan (64, 1)
    sph_status st = sph_mlp_set_params(h, th.data(), kParams);
    if (st != SPH_OK) {
        g_mlp_error = h->err;
        sph_mlp_destroy(h);
        return st;
    }
    *out = h;
    return SPH_OK;
}

sph_status sph_mlp_destroy(sph_mlp* h) {
    if (!h) return SPH_INVALID_ARG;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* ptrs[] = {h->theta, h->m, h->v, h->wbuf, h->samples, h->partial, h->loss_partial, h->grad, h->loss};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SPH_OK;
}

sph_status sph_mlp_set_params(sph_mlp* h, const float* params, size_t n) {
    if (!h || !params || n != (size_t)kParams) return SPH_INVALID_ARG;
    std::vector<float> zero(kParams, 0.0f);
    if (hipMemcpyAsync(h->theta, params, sizeof(float) * kParams, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(h->m, zero.data(), sizeof(float) * kParams, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(h->v, zero.data(), sizeof(float) * kParams, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        return h->fail(SPH_DEVICE_ERROR, "parameter upload failed");
    hipLaunchKernelGGL(k_mlp_copies, dim3((kParams + 255) / 256), dim3(256), 0, h->stream, h->theta, h->wb);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return h->fail(SPH_DEVICE_ERROR, "parameter upload sync");
    h->step = 0;
    return SPH_OK;
}

sph_status sph_mlp_get_params(sph_mlp* h, float* params, size_t n) {
    if (!h || !params || n != (size_t)kParams) return SPH_INVALID_ARG;
    if (hipMemcpyAsync(params, h->theta, sizeof(float) * kParams, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return h->fail(SPH_DEVICE_ERROR, "parameter read failed");
    return SPH_OK;
}

sph_status sph_mlp_set_samples(sph_mlp* h, const float* xyzd, size_t n) {
    if (!h || !xyzd || n < (size_t)kTile || n > 0xFFFFFFFFull) return h ? h->fail(SPH_INVALID_ARG, "need >= 64 samples") : SPH_INVALID_ARG;
    if (h->samples) (void)hipFree(h->samples);
    h->samples = nullptr;
    if (hipMalloc(&h->samples, sizeof(float4) * n) != hipSuccess) {
        (void)hipGetLastError();
        return h->fail(SPH_OUT_OF_MEMORY, "sample allocation failed");
    }
    if (hipMemcpyAsync(h->samples, xyzd, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return h->fail(SPH_DEVICE_ERROR, "sample upload failed");
    h->nsamples = (uint32_t)n;
    h->cursor = 0;
    return SPH_OK;
}

static sph_status enqueue_pass(sph_mlp* h, uint32_t first, int32_t batch, bool adam, float* grad, float* loss_dev) {
    NetView nv = h->view();
    const uint32_t tiles = (uint32_t)batch / kTile;
    const int blocks = (int)std::min<uint32_t>(tiles, (uint32_t)h->nblocks);
    hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(kThreads), 0, h->stream, h->samples, h->nsamples, first,
                       (uint32_t)batch, nv, h->partial, h->loss_partial);
    float c1 = 1.0f, c2 = 1.0f, lr = h->cfg.lr;
    if (adam) {
        const uint32_t t = h->step + 1;
        c1 = 1.0f - (float)std::pow((double)h->cfg.beta1, (double)t);
        c2 = 1.0f - (float)std::pow((double)h->cfg.beta2, (double)t);
        if (h->cfg.decay_steps > 0) {
            double f = std::min(1.0, (double)h->step / (double)h->cfg.decay_steps);
            lr = (float)(h->cfg.lr_min + 0.5 * (h->cfg.lr - h->cfg.lr_min) * (1.0 + std::cos(3.141592653589793 * f)));
        }
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + kAdamParams - 1) / kAdamParams), dim3(256), 0, h->stream,
                       h->partial, blocks,
                       h->loss_partial, h->theta, h->m, h->v, h->wb, lr, h->cfg.beta1, h->cfg.beta2,
                       h->cfg.eps, c1, c2, adam ? 1 : 0, grad, loss_dev, 1.0f / (float)batch);
    if (hipGetLastError() != hipSuccess) return h->fail(SPH_DEVICE_ERROR, "training launch failed");
    if (adam) ++h->step;
    return SPH_OK;
}

sph_status sph_mlp_gradient(sph_mlp* h, uint32_t first, int32_t batch, float* grad, float* loss) {
    if (!h || !grad || !loss) return SPH_INVALID_ARG;
    if (!h->samples) return h->fail(SPH_INVALID_ARG, "no samples");
    if (batch < kTile || batch % kTile) return h->fail(SPH_INVALID_ARG, "batch must be a positive multiple of 64");
    float* dloss = nullptr;
    if (hipMalloc(&dloss, sizeof(float)) != hipSuccess) return h->fail(SPH_OUT_OF_MEMORY, "alloc");
    sph_status st = enqueue_pass(h, first % h->nsamples, batch, false, h->grad, dloss);
    if (st == SPH_OK) {
        if (hipMemcpyAsync(grad, h->grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipMemcpyAsync(loss, dloss, sizeof(float), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            st = h->fail(SPH_DEVICE_ERROR, "gradient read failed");
    }
    (void)hipFree(dloss);
    return st;
}

sph_status sph_mlp_train(sph_mlp* h, int32_t steps, int32_t batch, float* losses) {
    if (!h || steps < 0) return SPH_INVALID_ARG;
    if (!h->samples) return h->fail(SPH_INVALID_ARG, "no samples");
    if (batch < kTile || batch % kTile) return h->fail(SPH_INVALID_ARG, "batch must be a positive multiple of 64");
    if (steps > h->loss_cap) {
        if (h->loss) (void)hipFree(h->loss);
        h->loss = nullptr;
        h->loss_cap = 0;
        if (hipMalloc(&h->loss, sizeof(float) * (size_t)std::max(steps, 1)) != hipSuccess)
            return h->fail(SPH_OUT_OF_MEMORY, "loss buffer");
        h->loss_cap = std::max(steps, 1);
    }
    for (int32_t i = 0; i < steps; ++i) {
        sph_status st = enqueue_pass(h, h->cursor, batch, true, nullptr, h->loss + i);
        if (st != SPH_OK) return st;
        h->cursor = (uint32_t)(((uint64_t)h->cursor + (uint64_t)batch) % h->nsamples);
    }
    if (losses && steps > 0 &&
        hipMemcpyAsync(losses, h->loss, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
        return h->fail(SPH_DEVICE_ERROR, "loss read failed");
    if (hipStreamSynchronize(h->stream) != hipSuccess) return h->fail(SPH_DEVICE_ERROR, "training failed");
    return SPH_OK;
}

sph_status sph_mlp_eval(sph_mlp* h, const float* xyz, size_t n, float* dgrad) {
    if (!h || (n && (!xyz || !dgrad))) return SPH_INVALID_ARG;
    if (n == 0) return SPH_OK;
    if (n > 0xFFFFFFFFull) return h->fail(SPH_INVALID_ARG, "too many points");
    std::vector<float> p4(4 * n);
    for (size_t i = 0; i < n; ++i) {
        p4[4 * i] = xyz[3 * i];
        p4[4 * i + 1] = xyz[3 * i + 1];
        p4[4 * i + 2] = xyz[3 * i + 2];
        p4[4 * i + 3] = 0.0f;
    }
    float4 *dp = nullptr, *dout = nullptr;
    if (hipMalloc(&dp, sizeof(float4) * n) != hipSuccess || hipMalloc(&dout, sizeof(float4) * n) != hipSuccess) {
        (void)hipGetLastError();
        if (dp) (void)hipFree(dp);
        return h->fail(SPH_OUT_OF_MEMORY, "eval buffers");
    }
    sph_status st = SPH_OK;
    if (hipMemcpyAsync(dp, p4.data(), sizeof(float4) * n, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        st = h->fail(SPH_DEVICE_ERROR, "eval upload");
    if (st == SPH_OK) {
        hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)((n + kTile - 1) / kTile)), dim3(kThreads), 0, h->stream, dp,
                           (uint32_t)n, h->view(), dout);
        if (hipMemcpyAsync(dgrad, dout, sizeof(float4) * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            st = h->fail(SPH_DEVICE_ERROR, "eval failed");
    }
    (void)hipFree(dp);
    (void)hipFree(dout);
    return st;
}

sph_status sph_mlp_get_config(const sph_mlp* h, sph_mlp_config* out) {
    if (!h || !out) return SPH_INVALID_ARG;
    *out = h->cfg;
    return SPH_OK;
}

sph_status sph_mlp_time_train(sph_mlp* h, int32_t steps, int32_t batch, float* ms) {
    if (!h || !ms || steps < 1) return SPH_INVALID_ARG;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return SPH_DEVICE_ERROR;
    (void)hipEventRecord(a, h->stream);
    sph_status st = SPH_OK;
    for (int32_t i = 0; i < steps && st == SPH_OK; ++i) {
        st = enqueue_pass(h, h->cursor, batch, true, nullptr, nullptr);
        h->cursor = (uint32_t)(((uint64_t)h->cursor + (uint64_t)batch) % h->nsamples);
    }
    (void)hipEventRecord(b, h->stream);
    if (hipEventSynchronize(b) != hipSuccess) st = SPH_DEVICE_ERROR;
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t;
    return st;
}

const char* sph_mlp_last_error_message(const sph_mlp* h) { return h ? h->err.c_str() : g_mlp_error.c_str(); }

}  // extern "C"
