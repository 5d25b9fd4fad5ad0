// mega_mlp.hip -- training and batch inference of the learned J2 return map
// (13-64-64-64-7 SiLU network, bf16 operands on MFMA, fp32 accumulation and
// fp32 master weights), behind the mega_mlp_* functions of include/mega.h.
//
// The network has no counterpart in the reference: it stands in for
// VON_MISES (CONSTITUTE_VONMISES.F90:21-158) at the seam FINT.F90:800, maps
// (strain increment[6], rotated stress[6], eqps[1]) to the plastic correction
// (sigma_trial - sigma[6], eqps increment[1]) in normalised units, and is
// trained on tuples harvested from the exact method's own runs.
//
// Kernels:
//   k_pack     FP32 master weights -> bf16 operand tiles (forward W and
//              backward W^T, k permuted as mlp::cfeat) + packed biases
//   k_forward  batch inference, one wave per 16 points
//   k_train    forward + backward of 32 points per wave; writes the wave's
//              gradient (all 9671 parameters) and loss to its own slot
//   k_adam     fixed-order reduction of the per-wave gradients (deterministic)
//              and the Adam update
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mega.h"
#include "mlp_mfma.hpp"

using namespace mlp;

namespace {

#define MLPCHK(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            std::fprintf(stderr, "mega_mlp: %s failed: %s (%s:%d)\n", #expr, \
                         hipGetErrorString(e_), __FILE__, __LINE__);          \
            return MEGA_ERR_HIP;                                              \
        }                                                                     \
    } while (0)

constexpr int kXs = 16;   // floats per normalised input row (13 + 3 zero)
constexpr int kTs = 8;    // floats per normalised target row (7 + 1 zero)
constexpr int kRow = 40;  // LDS row stride (bf16) of the [feature][32 points] arrays

__global__ void k_pack(const float *P, uint16_t *fwd, uint16_t *bwd, float *bias)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = kTilesFwd * 512, nb = kTilesBwd * 512;
    if (t < nf) {
        const int tile = t / 512, r = t % 512, lane = r / 8, j = r % 8;
        fwd[t] = bf16_bits(fwd_tile_value(P, tile, lane, j));
    } else if (t < nf + nb) {
        const int u = t - nf, tile = u / 512, r = u % 512, lane = r / 8, j = r % 8;
        bwd[u] = bf16_bits(bwd_tile_value(P, tile, lane, j));
    } else if (t < nf + nb + kBiasPack) {
        const int k = t - nf - nb;
        bias[k] = bias_pack_value(P, k);
    }
}

__device__ __forceinline__ bf16x8 load_x8(const float *row)
{
    const float4 a = *reinterpret_cast<const float4 *>(row);
    const float4 b = *reinterpret_cast<const float4 *>(row + 4);
    return pack8(f32x4{a.x, a.y, a.z, a.w}, f32x4{b.x, b.y, b.z, b.w});
}

// y[n][8] = network(x[n][16]); blocks of 256 threads, a wave per 16 points
__global__ __launch_bounds__(256) void k_forward(const bf16x8 *T, const float *bias, const float *X, float *Y,
                                                 int64_t n)
{
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t base = wave * 16; base < n; base += nw * 16) {
        const int64_t pt = base + (lane & 15);
        bf16x8 x0 = {};
        if (g < 2 && pt < n) x0 = load_x8(X + pt * kXs + 8 * g);
        f32x4 z1[4], z2[4], z3[4], y;
        forward16(T, bias, x0, lane, z1, z2, z3, y);
        if (g < 2 && pt < n) *reinterpret_cast<float4 *>(Y + pt * kTs + 4 * g) = float4{y[0], y[1], y[2], y[3]};
    }
}

// one wave (block of 64) per 32 batch positions: positions pos = 32 b + q map
// to rows (off + pos) % n; positions >= batch contribute nothing.
// part[b][kParams] receives the wave's gradient of
//   loss = (1/batch) sum_pos sum_o (y_o - t_o)^2,
// lpart[b] its share of the loss.
__global__ __launch_bounds__(64) void k_train(const bf16x8 *T, const bf16x8 *U, const float *bias, const float *X,
                                              const float *Tg, int64_t n, int64_t off, int batch, float *part,
                                              float *lpart)
{
    // activations a_l and back-propagated dz_l as bf16 [feature][32 points]
    __shared__ __attribute__((aligned(16))) __bf16 sA[208 * kRow];   // a0 16 | a1 64 | a2 64 | a3 64
    __shared__ __attribute__((aligned(16))) __bf16 sD[208 * kRow];   // dz1 64 | dz2 64 | dz3 64 | dy 16
    constexpr int A0 = 0, A1 = 16, A2 = 80, A3 = 144;
    constexpr int D1 = 0, D2 = 64, D3 = 128, DY = 192;
    const int lane = threadIdx.x, g = lane >> 4, c16 = lane & 15, g4 = 4 * g;
    const float inv_b = 1.0f / (float)batch;
    float loss = 0.0f;
    f32x4 db1[4], db2[4], db3[4], db4;
#pragma unroll
    for (int r = 0; r < 4; ++r) db1[r] = db2[r] = db3[r] = f32x4{0.f, 0.f, 0.f, 0.f};
    db4 = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int cg = 0; cg < 2; ++cg) {
        const int q = 16 * cg + c16;                      // column within the 32
        const int64_t pos = 32 * (int64_t)blockIdx.x + q;
        const bool live = pos < batch;
        const int64_t row = (off + (live ? pos : 0)) % n;
        bf16x8 x0 = {};
        if (g < 2) x0 = load_x8(X + row * kXs + 8 * g);
        if (g < 2)
#pragma unroll
            for (int j = 0; j < 8; ++j) sA[(A0 + 8 * g + j) * kRow + q] = x0[j];
        f32x4 z1[4], z2[4], z3[4], y;
        forward16(T, bias, x0, lane, z1, z2, z3, y);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int f = 16 * rt + g4 + t;
                sA[(A1 + f) * kRow + q] = (__bf16)silu(z1[rt][t]);
                sA[(A2 + f) * kRow + q] = (__bf16)silu(z2[rt][t]);
                sA[(A3 + f) * kRow + q] = (__bf16)silu(z3[rt][t]);
            }
        // output error: rows o = 4g + t, o < 7
        f32x4 dy = {0.f, 0.f, 0.f, 0.f};
        if (g < 2 && live) {
            const float4 tv = *reinterpret_cast<const float4 *>(Tg + row * kTs + g4);
            const float tt[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (g4 + t < kOut) {
                    const float e = y[t] - tt[t];
                    loss += e * e;
                    dy[t] = 2.0f * inv_b * e;
                }
        }
        db4 += dy;
#pragma unroll
        for (int t = 0; t < 4; ++t) sD[(DY + g4 + t) * kRow + q] = (__bf16)dy[t];
        // backward data: da3 = W4^T dy, dz3 = da3 * silu'(z3), ...
        const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
        const bf16x8 dyb = pack8(dy, zero);
        f32x4 dz[4];
        bf16x8 b[2];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const f32x4 da = mfma(U[(kU4 + rt) * 64 + lane], dyb, zero);
#pragma unroll
            for (int t = 0; t < 4; ++t) dz[rt][t] = da[t] * dsilu(z3[rt][t]);
            db3[rt] += dz[rt];
#pragma unroll
            for (int t = 0; t < 4; ++t) sD[(D3 + 16 * rt + g4 + t) * kRow + q] = (__bf16)dz[rt][t];
        }
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) b[kc] = pack8(dz[2 * kc], dz[2 * kc + 1]);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            f32x4 da = zero;
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) da = mfma(U[(kU3 + 2 * rt + kc) * 64 + lane], b[kc], da);
#pragma unroll
            for (int t = 0; t < 4; ++t) dz[rt][t] = da[t] * dsilu(z2[rt][t]);
            db2[rt] += dz[rt];
#pragma unroll
            for (int t = 0; t < 4; ++t) sD[(D2 + 16 * rt + g4 + t) * kRow + q] = (__bf16)dz[rt][t];
        }
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) b[kc] = pack8(dz[2 * kc], dz[2 * kc + 1]);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            f32x4 da = zero;
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) da = mfma(U[(kU2 + 2 * rt + kc) * 64 + lane], b[kc], da);
#pragma unroll
            for (int t = 0; t < 4; ++t) dz[rt][t] = da[t] * dsilu(z1[rt][t]);
            db1[rt] += dz[rt];
#pragma unroll
            for (int t = 0; t < 4; ++t) sD[(D1 + 16 * rt + g4 + t) * kRow + q] = (__bf16)dz[rt][t];
        }
    }
    __syncthreads();

    // weight gradients dW[o][i] = sum over the 32 points of dz[o] a[i]:
    // A operand row o = 16 rt + c16 (points 8g..8g+7), B operand column
    // i = 16 ct + c16 (points 8g..8g+7); result lane: column i, rows 16 rt + 4g + t
    float *out = part + (int64_t)blockIdx.x * kParams;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    auto op = [&](const __bf16 *base, int r) {
        return *reinterpret_cast<const bf16x8 *>(base + (r)*kRow + 8 * g);
    };
    auto put = [&](int W, int in, int nout, int rt, int ct, f32x4 c) {
        const int i = 16 * ct + c16;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int o = 16 * rt + g4 + t;
            if (o < nout && i < in) out[W + o * in + i] = c[t];
        }
    };
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) put(kW1, kIn, kHid, rt, 0, mfma(op(sD, D1 + 16 * rt + c16), op(sA, A0 + c16), zero));
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            put(kW2, kHid, kHid, rt, ct, mfma(op(sD, D2 + 16 * rt + c16), op(sA, A1 + 16 * ct + c16), zero));
            put(kW3, kHid, kHid, rt, ct, mfma(op(sD, D3 + 16 * rt + c16), op(sA, A2 + 16 * ct + c16), zero));
        }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) put(kW4, kHid, kOut, 0, ct, mfma(op(sD, DY + c16), op(sA, A3 + 16 * ct + c16), zero));

    // bias gradients: sum the 16 lanes of each lane group
    auto red16 = [&](f32x4 v) {
#pragma unroll
        for (int m = 1; m < 16; m <<= 1)
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t] += __shfl_xor(v[t], m, 64);
        return v;
    };
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        const f32x4 a = red16(db1[rt]), b = red16(db2[rt]), c = red16(db3[rt]);
        if (c16 == 0)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int o = 16 * rt + g4 + t;
                out[kB1 + o] = a[t];
                out[kB2 + o] = b[t];
                out[kB3 + o] = c[t];
            }
    }
    const f32x4 d4 = red16(db4);
    if (c16 == 0)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (g4 + t < kOut) out[kB4 + g4 + t] = d4[t];
    for (int m = 32; m >= 1; m >>= 1) loss += __shfl_xor(loss, m, 64);
    if (lane == 0) lpart[blockIdx.x] = loss * inv_b;
}

// gradient = fixed-order sum of the per-wave parts; optional Adam update
__global__ void k_adam(const float *part, int nparts, const float *lpart, float *P, float *m, float *v, float *grad,
                       float *loss_out, float lr, float b1, float b2, float eps, float bc1, float bc2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kParams) {
        float gsum = 0.0f;
        for (int b = 0; b < nparts; ++b) gsum += part[(int64_t)b * kParams + k];
        if (grad) grad[k] = gsum;
        if (lr > 0.0f) {
            const float mk = b1 * m[k] + (1.0f - b1) * gsum;
            const float vk = b2 * v[k] + (1.0f - b2) * gsum * gsum;
            m[k] = mk;
            v[k] = vk;
            P[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
        }
    }
    if (blockIdx.x == 0 && loss_out) {
        __shared__ float s[256];
        float a = 0.0f;
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) a += lpart[b];
        s[threadIdx.x] = a;
        __syncthreads();
        if (threadIdx.x == 0) {
            float t = 0.0f;
            for (int j = 0; j < (int)blockDim.x; ++j) t += s[j];
            *loss_out = t;
        }
    }
}

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double unit(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct mega_mlp {
    hipStream_t stream = nullptr;
    uint64_t rng = 0;
    float *P = nullptr, *m = nullptr, *v = nullptr;   // FP32 master weights, Adam moments
    uint16_t *fwd = nullptr, *bwd = nullptr;         // packed bf16 tiles
    float *bias = nullptr;
    float *part = nullptr, *lpart = nullptr, *gbuf = nullptr, *lbuf = nullptr;
    int part_cap = 0;
    float *X = nullptr, *Tg = nullptr;               // normalised training set
    int64_t count = 0;
    int64_t cursor = 0;
    int64_t adam_t = 0;
    float norm[kNorm] = {};
    bool has_norm = false;
};

namespace {

int pack(mega_mlp *m)
{
    const int total = (kTilesFwd + kTilesBwd) * 512 + kBiasPack;
    k_pack<<<(total + 255) / 256, 256, 0, m->stream>>>(m->P, m->fwd, m->bwd, m->bias);
    MLPCHK(hipGetLastError());
    return MEGA_OK;
}

int ensure_parts(mega_mlp *m, int nparts)
{
    if (nparts <= m->part_cap) return MEGA_OK;
    if (m->part) (void)hipFree(m->part);
    if (m->lpart) (void)hipFree(m->lpart);
    m->part = m->lpart = nullptr;
    m->part_cap = 0;
    if (hipMalloc(&m->part, sizeof(float) * (size_t)kParams * nparts) != hipSuccess) return MEGA_ERR_NOMEM;
    if (hipMalloc(&m->lpart, sizeof(float) * nparts) != hipSuccess) return MEGA_ERR_NOMEM;
    m->part_cap = nparts;
    return MEGA_OK;
}

// one gradient evaluation (and Adam step when lr > 0) on rows
// [off, off + batch) mod count of X / Tg
int grad_step(mega_mlp *m, const float *X, const float *Tg, int64_t count, int64_t off, int batch, float lr,
              float *grad_dev, float *loss_dev)
{
    const int nparts = (batch + 31) / 32;
    int st = ensure_parts(m, nparts);
    if (st) return st;
    k_train<<<nparts, 64, 0, m->stream>>>(reinterpret_cast<const bf16x8 *>(m->fwd),
                                          reinterpret_cast<const bf16x8 *>(m->bwd), m->bias, X, Tg, count, off,
                                          batch, m->part, m->lpart);
    MLPCHK(hipGetLastError());
    float bc1 = 1.0f, bc2 = 1.0f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (lr > 0.0f) {
        m->adam_t += 1;
        bc1 = (float)(1.0 - std::pow((double)b1, (double)m->adam_t));
        bc2 = (float)(1.0 - std::pow((double)b2, (double)m->adam_t));
    }
    k_adam<<<(kParams + 255) / 256, 256, 0, m->stream>>>(m->part, nparts, m->lpart, m->P, m->m, m->v, grad_dev,
                                                        loss_dev, lr, b1, b2, 1e-8f, bc1, bc2);
    MLPCHK(hipGetLastError());
    if (lr > 0.0f) return pack(m);
    return MEGA_OK;
}

// upload host rows (n x w_in, row-major) into a padded device matrix (n x w_dev)
int upload_rows(hipStream_t s, const float *host, int64_t n, int w_in, int w_dev, float **dev)
{
    std::vector<float> pad((size_t)n * w_dev, 0.0f);
    for (int64_t r = 0; r < n; ++r)
        for (int k = 0; k < w_in; ++k) pad[(size_t)r * w_dev + k] = host[(size_t)r * w_in + k];
    if (hipMalloc(dev, sizeof(float) * pad.size() + 64) != hipSuccess) return MEGA_ERR_NOMEM;
    MLPCHK(hipMemcpyAsync(*dev, pad.data(), sizeof(float) * pad.size(), hipMemcpyHostToDevice, s));
    MLPCHK(hipStreamSynchronize(s));
    return MEGA_OK;
}

}  // namespace

extern "C" {

int mega_mlp_create(uint64_t seed, mega_mlp **out)
{
    if (!out) return MEGA_ERR_ARG;
    *out = nullptr;
    mega_mlp *m = new mega_mlp();
    m->rng = seed;
    if (hipStreamCreate(&m->stream) != hipSuccess) { delete m; return MEGA_ERR_HIP; }
    bool ok = hipMalloc(&m->P, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&m->m, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&m->v, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&m->fwd, sizeof(uint16_t) * kTilesFwd * 512) == hipSuccess &&
              hipMalloc(&m->bwd, sizeof(uint16_t) * kTilesBwd * 512) == hipSuccess &&
              hipMalloc(&m->bias, sizeof(float) * kBiasPack) == hipSuccess &&
              hipMalloc(&m->gbuf, sizeof(float) * kParams) == hipSuccess &&
              hipMalloc(&m->lbuf, sizeof(float) * 1) == hipSuccess;
    if (!ok) { mega_mlp_destroy(m); return MEGA_ERR_NOMEM; }
    // Glorot-uniform weights, zero biases
    std::vector<float> P(kParams, 0.0f);
    const int W[4] = {kW1, kW2, kW3, kW4}, IN[4] = {kIn, kHid, kHid, kHid}, OUT[4] = {kHid, kHid, kHid, kOut};
    for (int l = 0; l < 4; ++l) {
        const double a = std::sqrt(6.0 / (IN[l] + OUT[l]));
        for (int k = 0; k < IN[l] * OUT[l]; ++k) P[W[l] + k] = (float)((2.0 * unit(m->rng) - 1.0) * a);
    }
    int st = mega_mlp_set_params(m, P.data());
    if (st) { mega_mlp_destroy(m); return st; }
    *out = m;
    return MEGA_OK;
}

int mega_mlp_destroy(mega_mlp *m)
{
    if (!m) return MEGA_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *p : {(void *)m->P, (void *)m->m, (void *)m->v, (void *)m->fwd, (void *)m->bwd, (void *)m->bias,
                    (void *)m->part, (void *)m->lpart, (void *)m->gbuf, (void *)m->lbuf, (void *)m->X,
                    (void *)m->Tg})
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return MEGA_OK;
}

int mega_mlp_set_params(mega_mlp *m, const float *params)
{
    if (!m || !params) return MEGA_ERR_ARG;
    MLPCHK(hipMemcpyAsync(m->P, params, sizeof(float) * kParams, hipMemcpyHostToDevice, m->stream));
    MLPCHK(hipMemsetAsync(m->m, 0, sizeof(float) * kParams, m->stream));
    MLPCHK(hipMemsetAsync(m->v, 0, sizeof(float) * kParams, m->stream));
    m->adam_t = 0;
    int st = pack(m);
    if (st) return st;
    MLPCHK(hipStreamSynchronize(m->stream));
    return MEGA_OK;
}

int mega_mlp_get_params(mega_mlp *m, float *params)
{
    if (!m || !params) return MEGA_ERR_ARG;
    MLPCHK(hipMemcpyAsync(params, m->P, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
    MLPCHK(hipStreamSynchronize(m->stream));
    return MEGA_OK;
}

int mega_mlp_get_norm(mega_mlp *m, float *norm)
{
    if (!m || !norm) return MEGA_ERR_ARG;
    if (!m->has_norm) return MEGA_ERR_ARG;
    std::memcpy(norm, m->norm, sizeof(m->norm));
    return MEGA_OK;
}

int mega_mlp_set_data(mega_mlp *m, const float *tuples, int64_t count, const double *prop)
{
    if (!m || !tuples || count <= 0 || !prop) return MEGA_ERR_ARG;
    const double lam = prop[27], mu = prop[28], l2m = prop[29];
    // inputs x = (de[6], sigma_rot[6], eqps_n); targets t = (trial - sigma[6], eqps - eqps_n)
    std::vector<float> x((size_t)count * kIn), t((size_t)count * kOut);
    for (int64_t r = 0; r < count; ++r) {
        const float *u = tuples + (size_t)r * 20;
        float *xr = &x[(size_t)r * kIn], *tr = &t[(size_t)r * kOut];
        for (int k = 0; k < 13; ++k) xr[k] = u[k];
        const double de0 = u[0], de1 = u[1], de2 = u[2];
        double trial[6];
        trial[0] = u[6] + (l2m * de0 + lam * de1 + lam * de2);
        trial[1] = u[7] + (lam * de0 + l2m * de1 + lam * de2);
        trial[2] = u[8] + (lam * de0 + lam * de1 + l2m * de2);
        trial[3] = u[9] + mu * u[3];
        trial[4] = u[10] + mu * u[4];
        trial[5] = u[11] + mu * u[5];
        for (int k = 0; k < 6; ++k) tr[k] = (float)(trial[k] - (double)u[13 + k]);
        tr[6] = (float)((double)u[19] - (double)u[12]);
    }
    // normalisation and envelope
    float *nm = m->norm;
    for (int k = 0; k < kIn; ++k) {
        double s = 0, s2 = 0, lo = 1e300, hi = -1e300;
        for (int64_t r = 0; r < count; ++r) {
            const double a = x[(size_t)r * kIn + k];
            s += a; s2 += a * a; lo = std::min(lo, a); hi = std::max(hi, a);
        }
        const double mean = s / count, var = std::max(s2 / count - mean * mean, 0.0);
        nm[kInMean + k] = (float)mean;
        nm[kInScale + k] = (float)(var > 0 ? 1.0 / std::sqrt(var) : 1.0);
        nm[kEnvLo + k] = (float)lo;
        nm[kEnvHi + k] = (float)hi;
    }
    for (int k = 0; k < kOut; ++k) {
        double s2 = 0;
        for (int64_t r = 0; r < count; ++r) { const double a = t[(size_t)r * kOut + k]; s2 += a * a; }
        const double rms = std::sqrt(s2 / count);
        nm[kOutScale + k] = (float)(rms > 0 ? rms : 1.0);
    }
    m->has_norm = true;
    // normalise, shuffle (Fisher-Yates, splitmix64), upload
    std::vector<int64_t> perm((size_t)count);
    for (int64_t r = 0; r < count; ++r) perm[(size_t)r] = r;
    for (int64_t r = count - 1; r > 0; --r) {
        const int64_t j = (int64_t)(splitmix64(m->rng) % (uint64_t)(r + 1));
        std::swap(perm[(size_t)r], perm[(size_t)j]);
    }
    std::vector<float> xh((size_t)count * kIn), th((size_t)count * kOut);
    for (int64_t r = 0; r < count; ++r) {
        const int64_t s = perm[(size_t)r];
        for (int k = 0; k < kIn; ++k)
            xh[(size_t)r * kIn + k] = (x[(size_t)s * kIn + k] - nm[kInMean + k]) * nm[kInScale + k];
        for (int k = 0; k < kOut; ++k) th[(size_t)r * kOut + k] = t[(size_t)s * kOut + k] / nm[kOutScale + k];
    }
    if (m->X) (void)hipFree(m->X);
    if (m->Tg) (void)hipFree(m->Tg);
    m->X = m->Tg = nullptr;
    int st = upload_rows(m->stream, xh.data(), count, kIn, kXs, &m->X);
    if (st) return st;
    st = upload_rows(m->stream, th.data(), count, kOut, kTs, &m->Tg);
    if (st) return st;
    m->count = count;
    m->cursor = 0;
    return MEGA_OK;
}

int mega_mlp_train(mega_mlp *m, int32_t iters, int32_t batch, double lr0, double lr1, float *loss_history)
{
    if (!m || iters < 0 || batch <= 0 || !m->X) return MEGA_ERR_ARG;
    std::vector<float> hist(loss_history ? (size_t)iters : 0);
    float *ldev = nullptr;
    if (loss_history && iters > 0 && hipMalloc(&ldev, sizeof(float) * iters) != hipSuccess) return MEGA_ERR_NOMEM;
    int st = MEGA_OK;
    for (int it = 0; it < iters && st == MEGA_OK; ++it) {
        // cosine decay from lr0 to lr1
        const double c = 0.5 * (1.0 + std::cos(M_PI * (double)it / std::max(1, iters)));
        const float lr = (float)(lr1 + (lr0 - lr1) * c);
        st = grad_step(m, m->X, m->Tg, m->count, m->cursor, batch, lr, nullptr, ldev ? ldev + it : nullptr);
        m->cursor = (m->cursor + batch) % m->count;
    }
    if (st == MEGA_OK && ldev) {
        if (hipMemcpyAsync(loss_history, ldev, sizeof(float) * iters, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
            st = MEGA_ERR_HIP;
    }
    if (hipStreamSynchronize(m->stream) != hipSuccess) st = MEGA_ERR_HIP;
    if (ldev) (void)hipFree(ldev);
    return st;
}

int mega_mlp_forward(mega_mlp *m, const float *xhat, float *y, int64_t n)
{
    if (!m || !xhat || !y || n <= 0) return MEGA_ERR_ARG;
    float *X = nullptr, *Y = nullptr;
    int st = upload_rows(m->stream, xhat, n, kIn, kXs, &X);
    if (st) return st;
    if (hipMalloc(&Y, sizeof(float) * n * kTs) != hipSuccess) { (void)hipFree(X); return MEGA_ERR_NOMEM; }
    const int64_t waves = (n + 15) / 16;
    const int blocks = (int)std::min<int64_t>((waves + 3) / 4, 65536);
    k_forward<<<blocks, 256, 0, m->stream>>>(reinterpret_cast<const bf16x8 *>(m->fwd), m->bias, X, Y, n);
    st = hipGetLastError() == hipSuccess ? MEGA_OK : MEGA_ERR_HIP;
    std::vector<float> yh((size_t)n * kTs);
    if (st == MEGA_OK &&
        hipMemcpyAsync(yh.data(), Y, sizeof(float) * n * kTs, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        st = MEGA_ERR_HIP;
    if (hipStreamSynchronize(m->stream) != hipSuccess) st = MEGA_ERR_HIP;
    if (st == MEGA_OK)
        for (int64_t r = 0; r < n; ++r)
            for (int k = 0; k < kOut; ++k) y[(size_t)r * kOut + k] = yh[(size_t)r * kTs + k];
    (void)hipFree(X);
    (void)hipFree(Y);
    return st;
}

int mega_mlp_gradient(mega_mlp *m, const float *xhat, const float *that, int64_t n, float *grad, float *loss)
{
    if (!m || !xhat || !that || !grad || n <= 0 || n > (1 << 30)) return MEGA_ERR_ARG;
    float *X = nullptr, *Tg = nullptr;
    int st = upload_rows(m->stream, xhat, n, kIn, kXs, &X);
    if (st == MEGA_OK) st = upload_rows(m->stream, that, n, kOut, kTs, &Tg);
    if (st == MEGA_OK) st = grad_step(m, X, Tg, n, 0, (int)n, 0.0f, m->gbuf, m->lbuf);
    if (st == MEGA_OK &&
        (hipMemcpyAsync(grad, m->gbuf, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
         (loss && hipMemcpyAsync(loss, m->lbuf, sizeof(float), hipMemcpyDeviceToHost, m->stream) != hipSuccess)))
        st = MEGA_ERR_HIP;
    if (hipStreamSynchronize(m->stream) != hipSuccess) st = MEGA_ERR_HIP;
    if (X) (void)hipFree(X);
    if (Tg) (void)hipFree(Tg);
    return st;
}

}  // extern "C"
