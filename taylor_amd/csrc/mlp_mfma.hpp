// mlp_mfma.hpp -- the 13-64-64-64-7 SiLU network of the J2 return-map
// surrogate on gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: bf16 operands,
// fp32 accumulation).  Shared by the training kernels and by the node kernel
// of the explicit step.
//
// Points are the MFMA columns: a layer computes Z^T (out x points) = W (out x
// in) . A^T (in x points), 16 points per column group.  Operand layouts of
// the instruction (lane l, g = l / 16):
//   A: row l % 16, k = 8g .. 8g+7          (8 bf16 per lane)
//   B: column l % 16, k = 8g .. 8g+7       (8 bf16 per lane)
//   C: column l % 16, rows 4g .. 4g+3      (4 fp32 per lane)
// A result tile therefore leaves lane l holding features 16 rt + 4g + t of
// point l % 16.  Rather than transposing through LDS, the next layer takes
// the eight values of row tiles 2kc and 2kc+1 as its B operand for k chunk
// kc and the weights are packed with the same permutation of k (cfeat), so
// activations stay in registers from layer to layer.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mlp {

constexpr int kIn = 13, kHid = 64, kOut = 7;

// FP32 parameter blob: per layer W[out][in] row-major, then b[out]
constexpr int kW1 = 0, kB1 = kW1 + kHid * kIn;
constexpr int kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid;
constexpr int kW3 = kB2 + kHid, kB3 = kW3 + kHid * kHid;
constexpr int kW4 = kB3 + kHid, kB4 = kW4 + kOut * kHid;
constexpr int kParams = kB4 + kOut;   // 9671

// normalisation block (FP32): xhat = (x - in_mean) * in_scale, network output
// y maps to the plastic correction t = y * out_scale; envelope [env_lo, env_hi]
constexpr int kInMean = 0, kInScale = 13, kOutScale = 26, kEnvLo = 33, kEnvHi = 46, kNorm = 59;

// packed bf16 operand tiles, 64 lanes x 8 values each (1 KB)
constexpr int kT1 = 0, kT2 = 4, kT3 = 12, kT4 = 20, kTilesFwd = 22;   // W tiles (forward)
constexpr int kU4 = 0, kU3 = 4, kU2 = 12, kTilesBwd = 20;            // W^T tiles (backward data)
constexpr int kBiasPack = 3 * kHid + 16;                              // b1 | b2 | b3 | b4 (padded)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// feature carried by k slot j of lane group g in k chunk kc when the operand is
// assembled from two result tiles (slots 0-3 from tile 2kc, 4-7 from 2kc+1)
__host__ __device__ inline int cfeat(int kc, int g, int j)
{
    return 32 * kc + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

// float -> bf16 bits, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t bf16_bits(float f)
{
    union { float f; uint32_t u; } v;
    v.f = f;
    const uint32_t r = 0x7fffu + ((v.u >> 16) & 1u);
    return (uint16_t)((v.u + r) >> 16);
}

// value of forward tile `tile`, lane, slot j from the FP32 parameter blob
__host__ __device__ inline float fwd_tile_value(const float *P, int tile, int lane, int j)
{
    const int g = lane >> 4, r16 = lane & 15;
    int W, in, out, rt, kc;
    bool direct = false;
    if (tile < kT2) { W = kW1; in = kIn; out = kHid; rt = tile - kT1; kc = 0; direct = true; }
    else if (tile < kT3) { W = kW2; in = kHid; out = kHid; rt = (tile - kT2) >> 1; kc = (tile - kT2) & 1; }
    else if (tile < kT4) { W = kW3; in = kHid; out = kHid; rt = (tile - kT3) >> 1; kc = (tile - kT3) & 1; }
    else { W = kW4; in = kHid; out = kOut; rt = 0; kc = tile - kT4; }
    const int row = 16 * rt + r16;
    const int f = direct ? 8 * g + j : cfeat(kc, g, j);
    return (row < out && f < in) ? P[W + row * in + f] : 0.0f;
}

// value of backward tile (W^T: rows = input features, k = output features)
__host__ __device__ inline float bwd_tile_value(const float *P, int tile, int lane, int j)
{
    const int g = lane >> 4, r16 = lane & 15;
    int W, in, out, rt, kc;
    if (tile < kU3) { W = kW4; in = kHid; out = kOut; rt = tile - kU4; kc = 0; }
    else if (tile < kU2) { W = kW3; in = kHid; out = kHid; rt = (tile - kU3) >> 1; kc = (tile - kU3) & 1; }
    else { W = kW2; in = kHid; out = kHid; rt = (tile - kU2) >> 1; kc = (tile - kU2) & 1; }
    const int row = 16 * rt + r16;
    const int f = cfeat(kc, g, j);
    return (row < in && f < out) ? P[W + f * in + row] : 0.0f;
}

__host__ __device__ inline float bias_pack_value(const float *P, int k)
{
    if (k < kHid) return P[kB1 + k];
    if (k < 2 * kHid) return P[kB2 + k - kHid];
    if (k < 3 * kHid) return P[kB3 + k - 2 * kHid];
    const int o = k - 3 * kHid;
    return o < kOut ? P[kB4 + o] : 0.0f;
}

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// logistic sigma(z) = 1 / (1 + e^-z) with the hardware reciprocal (1 ulp);
// e^-z = inf for very negative z gives sigma = 0
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ float silu(float z) { return z * sigm(z); }

__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ bf16x8 pack8(f32x4 lo, f32x4 hi)
{
    bf16x8 b;
    b[0] = (__bf16)lo[0]; b[1] = (__bf16)lo[1]; b[2] = (__bf16)lo[2]; b[3] = (__bf16)lo[3];
    b[4] = (__bf16)hi[0]; b[5] = (__bf16)hi[1]; b[6] = (__bf16)hi[2]; b[7] = (__bf16)hi[3];
    return b;
}

__device__ __forceinline__ f32x4 silu4(f32x4 z)
{
    f32x4 a;
    a[0] = silu(z[0]); a[1] = silu(z[1]); a[2] = silu(z[2]); a[3] = silu(z[3]);
    return a;
}

__device__ __forceinline__ f32x4 bias4(const float *b, int row0)
{
    f32x4 c;
    c[0] = b[row0]; c[1] = b[row0 + 1]; c[2] = b[row0 + 2]; c[3] = b[row0 + 3];
    return c;
}

// Forward pass of one column group of 16 points (whole wave).  x0 is the B
// operand of layer 1 (features 8g..8g+7 of point lane % 16; zero for g >= 2).
// T: forward tiles [kTilesFwd][64] (LDS or global), bias: packed biases.
// Pre-activations z1..z3 are returned for the backward pass; y holds outputs
// 4g..4g+3 of point lane % 16 (valid for g = 0, 1; rows >= 7 are zero).
template <class TP, class BP>
__device__ __forceinline__ void forward16(const TP *T, const BP *bias, bf16x8 x0, int lane,
                                          f32x4 *z1, f32x4 *z2, f32x4 *z3, f32x4 &y)
{
    const int g4 = 4 * (lane >> 4);
    bf16x8 b[2];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) z1[rt] = mfma(T[(kT1 + rt) * 64 + lane], x0, bias4(bias, 16 * rt + g4));
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) b[kc] = pack8(silu4(z1[2 * kc]), silu4(z1[2 * kc + 1]));
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        f32x4 c = bias4(bias, kHid + 16 * rt + g4);
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) c = mfma(T[(kT2 + 2 * rt + kc) * 64 + lane], b[kc], c);
        z2[rt] = c;
    }
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) b[kc] = pack8(silu4(z2[2 * kc]), silu4(z2[2 * kc + 1]));
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        f32x4 c = bias4(bias, 2 * kHid + 16 * rt + g4);
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) c = mfma(T[(kT3 + 2 * rt + kc) * 64 + lane], b[kc], c);
        z3[rt] = c;
    }
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) b[kc] = pack8(silu4(z3[2 * kc]), silu4(z3[2 * kc + 1]));
    f32x4 c = bias4(bias, 3 * kHid + g4);   // b4 padded with zeros to 16 rows
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) c = mfma(T[(kT4 + kc) * 64 + lane], b[kc], c);
    y = c;
}

#endif

}  // namespace mlp
