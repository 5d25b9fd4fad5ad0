// mega_hip.hip -- MI355X (gfx950) explicit meshfree step behind the C-ABI in
// include/mega.h.
//
// Kernels (each cites the reference stage it replaces):
//   k_predict      predictor                          MAIN.F90:348-356
//   k_interp       DO_INTERP + coordinate update      HANDELER.F90:114-133, MAIN.F90:377-386
//   k_bin_*        GET_BINS as a counting sort        BIN_SEARCH.F90:2-64
//   k_search       SOFT_SEARCH (+BIN_INODE)           BIN_SEARCH.F90:69-133,153-328
//   k_rk           RK1 shape functions/gradients      RK.f90:16-453, KERNEL.F90:3-140
//   k_fint         CONSTRUCT_FINT + CONSTITUTION      FINT.F90:16-1247, CONSTITUTION.F90:18-89
//   k_correct      EBC, damping, corrector, energies  MAIN.F90:432-488,524-553
//
// Layout in HBM: node vectors interleaved (3*i+k), stress/strain Voigt 6 per
// node, neighbor lists as 0-based CSR (row_ptr int32[n+1], col int32[nnz]),
// shape values phi[nnz] and gradients dphi[3*nnz] in neighbor-list order.
// One explicit step reads each node array once and streams the pair arrays
// once for the gather phase and once (from cache) for the force scatter.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "mega.h"
#include "mega_math.hpp"

using mega::Material;

namespace {

constexpr int kBlock = 256;
constexpr int kGroup = 16;           // lanes cooperating on one node's neighbor list
constexpr int kNodesPerBlock = 128;  // nodes per k_fint block (LDS-staged phases)

struct Dev {
    int n = 0;
    int64_t nnz = 0, nnz_cap = 0;
    // node arrays (FP64)
    double *X0 = nullptr, *x = nullptr, *v = nullptr, *a = nullptr, *d = nullptr, *dtot = nullptr;
    double *uphy = nullptr, *fint = nullptr, *fext = nullptr, *fext_nmo = nullptr;
    double *stress = nullptr, *strain = nullptr, *eps = nullptr;
    double *win = nullptr, *vol = nullptr, *mass = nullptr;
    double *strain_eq = nullptr, *wave_vel = nullptr, *vphy = nullptr, *scratch3 = nullptr;
    int32_t *ebc = nullptr, *ebc_node = nullptr, *mat = nullptr;
    Material *mats = nullptr;
    // CSR + shapes
    int32_t *row_ptr = nullptr, *col = nullptr;
    double *shp = nullptr, *dshp = nullptr;
    // bins
    int32_t *ijk = nullptr, *cell_start = nullptr, *cell_nodes = nullptr, *cell_cursor = nullptr;
    int32_t *counts = nullptr;
    // accumulators: [0] work, [1] ke, [2] ie (small strain), [3] ee
    double *acc = nullptr;
    int32_t *flags = nullptr;   // [0] return-map failures, [1] unsupported material
    void *cub_tmp = nullptr;
    size_t cub_tmp_bytes = 0;
};

struct Params {   // kernel arguments
    int n;
    const int32_t *row_ptr, *col;
    double *shp, *dshp;
    const double *X0, *win, *vol, *mass;
    double *x, *v, *a, *d, *dtot, *uphy, *fint, *fext, *fext_nmo;
    double *stress, *strain, *eps, *strain_eq, *wave_vel, *vphy;
    const int32_t *ebc, *ebc_node, *mat;
    const Material *mats;
    double *acc;
    int32_t *flags;
    int ql;
    double ql_coef, ql_len;
    int shsup, lagrangian, finite_strain, has_gravity;
    double grav[3];
    double dt;
};

// ---------------------------------------------------------------------------
// group reductions over kGroup lanes (xor butterfly: every lane gets the sum)

template <int W>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int m = W / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, W);
    return v;
}

__device__ __forceinline__ double block_sum(double v, double *red)
{
    // one value per thread -> thread 0 gets the block sum
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// predictor (MAIN.F90:353-355): d = dt v + dt^2/2 a; dtot += d; v += dt/2 a.
// Also clears the force accumulators for this step's scatter.

__global__ void k_predict(Params p)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) { p.acc[1] = 0.0; p.acc[2] = 0.0; }
    if (q >= 3 * (int64_t)p.n) return;
    const double dt = p.dt;
    const double v = p.v[q], a = p.a[q];
    const double d = dt * v + dt * dt * 0.5 * a;
    p.d[q] = d;
    p.dtot[q] = p.dtot[q] + d;
    p.v[q] = v + dt * 0.5 * a;
    p.fint[q] = 0.0;
    if (p.has_gravity) p.fext[q] = 0.0;
}

// ---------------------------------------------------------------------------
// DO_INTERP (HANDELER.F90:114-133): out_i = sum_j phi_ij coef_j, kGroup lanes
// per node.  UPDATE: coef is the increment d and the result updates the
// physical total displacement and the coordinates (MAIN.F90:377-386).

template <bool UPDATE>
__global__ __launch_bounds__(kBlock) void k_interp(Params p, const double *coef, double *out)
{
    const int node = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup;
    const int lane = threadIdx.x & (kGroup - 1);
    if (node >= p.n) return;
    const int p0 = p.row_ptr[node], p1 = p.row_ptr[node + 1];
    const double *src = UPDATE ? p.d : coef;   // the update always interpolates this step's increment
    double u0 = 0, u1 = 0, u2 = 0;
    for (int q = p0 + lane; q < p1; q += kGroup) {
        const int j = p.col[q];
        const double f = p.shp[q];
        u0 += f * src[3 * j]; u1 += f * src[3 * j + 1]; u2 += f * src[3 * j + 2];
    }
    u0 = group_sum<kGroup>(u0); u1 = group_sum<kGroup>(u1); u2 = group_sum<kGroup>(u2);
    if (lane < 3) {
        const double u = lane == 0 ? u0 : (lane == 1 ? u1 : u2);
        const int q = 3 * node + lane;
        if (UPDATE) {
            const double t = p.uphy[q] + u;
            p.uphy[q] = t;
            p.x[q] = p.X0[q] + t;
        } else {
            out[q] = u;
        }
    }
}

// ---------------------------------------------------------------------------
// RK1 at point x_i with nodal positions xs (RK.f90:16-453).  kGroup lanes per
// node accumulate the symmetric moment matrix M and its derivatives (FP64),
// then every lane forms M^-1 and evaluates phi, grad phi for its neighbors.

struct KernelEval {
    double phi, px, py, pz;
};

__device__ __forceinline__ KernelEval rk_window(const Params &p, const double *xi, const double *xs, int j)
{
    const double dx = xi[0] - xs[3 * j], dy = xi[1] - xs[3 * j + 1], dz = xi[2] - xs[3 * j + 2];
    const double w1 = p.win[3 * j], w2 = p.win[3 * j + 1], w3 = p.win[3 * j + 2];
    const double zx = dx / w1, zy = dy / w2, zz = dz / w3;
    KernelEval e;
    if (p.shsup) {
        const double dia = sqrt(zx * zx + zy * zy + zz * zz);
        double k0, k1;
        mega::cubic_kernel(dia, k0, k1);
        double rx = 0, ry = 0, rz = 0;
        if (!(dia <= 1.0e-13)) { rx = dx / (w1 * w1) / dia; ry = dy / (w2 * w2) / dia; rz = dz / (w3 * w3) / dia; }
        e.phi = k0; e.px = k1 * rx; e.py = k1 * ry; e.pz = k1 * rz;
    } else {
        double fx, fxx, fy, fyy, fz, fzz;
        mega::cubic_kernel(fabs(zx), fx, fxx);
        mega::cubic_kernel(fabs(zy), fy, fyy);
        mega::cubic_kernel(fabs(zz), fz, fzz);
        const double iw = 1.0 / w1;   // RK.f90:183-205 divides all three by the x window
        const double rx = zx == 0.0 ? 0.0 : (zx >= 0.0 ? iw : -iw);
        const double ry = zy == 0.0 ? 0.0 : (zy >= 0.0 ? iw : -iw);
        const double rz = zz == 0.0 ? 0.0 : (zz >= 0.0 ? iw : -iw);
        double ph = fx * fy * fz;
        double px = fxx * fy * fz * rx, py = fx * fyy * fz * ry, pz = fx * fy * fzz * rz;
        if (p.ebc_node[j]) {   // singular kernel at essential-BC nodes (RK.f90:217-228)
            const double wm = 1.0 / (sqrt(zx * zx + zy * zy + zz * zz) + mega::kSingularEps);
            const double pw3 = ph * wm * wm * wm;
            px = px * wm - pw3 * zx / w1;
            py = py * wm - pw3 * zy / w2;
            pz = pz * wm - pw3 * zz / w3;
            ph = ph * wm;
        }
        e.phi = ph; e.px = px; e.py = py; e.pz = pz;
    }
    return e;
}

__global__ __launch_bounds__(kBlock) void k_rk(Params p, const double *xs)
{
    const int node = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup;
    const int lane = threadIdx.x & (kGroup - 1);
    if (node >= p.n) return;
    const int p0 = p.row_ptr[node], p1 = p.row_ptr[node + 1];
    const double xi[3] = {xs[3 * node], xs[3 * node + 1], xs[3 * node + 2]};
    // 10 unique entries each of M, Mx, My, Mz (symmetric) + sum of phi
    double m[41];
#pragma unroll
    for (int k = 0; k < 41; ++k) m[k] = 0.0;
    for (int q = p0 + lane; q < p1; q += kGroup) {
        const int j = p.col[q];
        const KernelEval e = rk_window(p, xi, xs, j);
        const double H[4] = {1.0, xi[0] - xs[3 * j], xi[1] - xs[3 * j + 1], xi[2] - xs[3 * j + 2]};
        int t = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = r; c < 4; ++c, ++t) {
                const double hh = H[r] * H[c];
                m[t] += hh * e.phi;
                // M_alpha(r,c) = (dH_r H_c + H_r dH_c) phi + H_r H_c phi_alpha, dH = unit vector
                m[10 + t] += (double)((r == 1) * 1) * H[c] * e.phi + (double)(c == 1) * H[r] * e.phi + hh * e.px;
                m[20 + t] += (double)(r == 2) * H[c] * e.phi + (double)(c == 2) * H[r] * e.phi + hh * e.py;
                m[30 + t] += (double)(r == 3) * H[c] * e.phi + (double)(c == 3) * H[r] * e.phi + hh * e.pz;
            }
        m[40] += e.phi;
    }
#pragma unroll
    for (int k = 0; k < 41; ++k) m[k] = group_sum<kGroup>(m[k]);
    // quasi-linear augmentation of M (RK.f90:251-303)
    double hs[4] = {0, 0, 0, 0};
    if (p.ql) {
        const double L = p.ql_len;
        double ms[10] = {0};
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            double q0 = xi[0], q1 = xi[1], q2 = xi[2];
            if (s == 0) q0 = xi[0] + L;
            if (s == 1) q0 = xi[0] - L;
            if (s == 2) q1 = xi[1] + L;
            if (s == 3) q1 = xi[1] - L;
            if (s == 4) q2 = xi[2] + L;
            if (s == 5) q2 = xi[2] - L;
            const double H[4] = {1.0, xi[0] - q0, xi[1] - q1, xi[2] - q2};
            int t = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = r; c < 4; ++c, ++t) ms[t] += H[r] * H[c];
#pragma unroll
            for (int k = 0; k < 4; ++k) hs[k] += H[k];
        }
#pragma unroll
        for (int t = 0; t < 10; ++t) m[t] = m[t] + ms[t] * m[40] * p.ql_coef;
    }
    double Mi[16];
    mega::m44inv_sym(m, Mi);
    // B = e1^T M^-1; B_alpha = -(B M_alpha) M^-1
    double B[4] = {Mi[0], Mi[4], Mi[8], Mi[12]};
    double Bd[3][4];
#pragma unroll
    for (int al = 0; al < 3; ++al) {
        const double *s = m + 10 * (al + 1);
        const double Ma[16] = {s[0], s[1], s[2], s[3], s[1], s[4], s[5], s[6],
                               s[2], s[5], s[7], s[8], s[3], s[6], s[8], s[9]};
        double u[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) u[c] = B[0] * Ma[4 * c] + B[1] * Ma[1 + 4 * c] + B[2] * Ma[2 + 4 * c] + B[3] * Ma[3 + 4 * c];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            Bd[al][c] = -(u[0] * Mi[4 * c] + u[1] * Mi[1 + 4 * c] + u[2] * Mi[2 + 4 * c] + u[3] * Mi[3 + 4 * c]);
    }
    for (int q = p0 + lane; q < p1; q += kGroup) {
        const int j = p.col[q];
        const KernelEval e = rk_window(p, xi, xs, j);
        double H[4] = {1.0, xi[0] - xs[3 * j], xi[1] - xs[3 * j + 1], xi[2] - xs[3 * j + 2]};
        if (p.ql)
#pragma unroll
            for (int k = 0; k < 4; ++k) H[k] = H[k] + hs[k] * p.ql_coef;
        const double c0 = B[0] * H[0] + B[1] * H[1] + B[2] * H[2] + B[3] * H[3];
        const double cx = Bd[0][0] * H[0] + Bd[0][1] * H[1] + Bd[0][2] * H[2] + Bd[0][3] * H[3] + B[1];
        const double cy = Bd[1][0] * H[0] + Bd[1][1] * H[1] + Bd[1][2] * H[2] + Bd[1][3] * H[3] + B[2];
        const double cz = Bd[2][0] * H[0] + Bd[2][1] * H[1] + Bd[2][2] * H[2] + Bd[2][3] * H[3] + B[3];
        p.shp[q] = c0 * e.phi;
        p.dshp[3 * q] = cx * e.phi + c0 * e.px;
        p.dshp[3 * q + 1] = cy * e.phi + c0 * e.py;
        p.dshp[3 * q + 2] = cz * e.phi + c0 * e.pz;
    }
}

// ---------------------------------------------------------------------------
// CONSTRUCT_FINT for one block of kNodesPerBlock nodes, in three phases:
//  A  kGroup lanes per node stream the node's neighbor list once (coalesced
//     over the CSR) and gather d, dtot of the neighbors:
//       F = I + sum grad phi0 (x) dtot   (FINT.F90:363-376, Lagrangian)
//       L0 = sum grad phi0 (x) d         (FINT.F90:673-679)
//       u  = sum phi d                   (DO_INTERP, fused when Lagrangian)
//     the 21 sums are staged in LDS;
//  B  one thread per node: J = |det F|, F^-1 = INV3(F) (bug-compatible),
//     L = L0 F^-1, Hughes-Winget, rotation, elastic trial, radial return,
//     state write-back, work, and P = V J sigma F^-T staged in LDS;
//  C  kGroup lanes per node re-read grad phi0 and scatter f_j += P grad phi0_ij
//     (= B^T sigma V J of FINT.F90:1007-1112) with FP64 atomics.

template <bool PUSH, bool FUSE_INTERP>
__global__ __launch_bounds__(kBlock) void k_fint(Params p)
{
    __shared__ double s_sum[kNodesPerBlock][22];
    __shared__ double s_red[kBlock / 64];
    const int nb0 = blockIdx.x * kNodesPerBlock;
    const int grp = threadIdx.x / kGroup, lane = threadIdx.x & (kGroup - 1);
    constexpr int kGroups = kBlock / kGroup;

    // ---- phase A
    for (int ln = grp; ln < kNodesPerBlock; ln += kGroups) {
        const int i = nb0 + ln;
        double F[9], L[9], u[3];
#pragma unroll
        for (int k = 0; k < 9; ++k) { F[k] = 0.0; L[k] = 0.0; }
        u[0] = u[1] = u[2] = 0.0;
        if (i < p.n) {
            const int p0 = p.row_ptr[i], p1 = p.row_ptr[i + 1];
            for (int q = p0 + lane; q < p1; q += kGroup) {
                const int j = p.col[q];
                const double g0 = p.dshp[3 * q], g1 = p.dshp[3 * q + 1], g2 = p.dshp[3 * q + 2];
                const double d0 = p.d[3 * j], d1 = p.d[3 * j + 1], d2 = p.d[3 * j + 2];
                L[0] += g0 * d0; L[1] += g0 * d1; L[2] += g0 * d2;
                L[3] += g1 * d0; L[4] += g1 * d1; L[5] += g1 * d2;
                L[6] += g2 * d0; L[7] += g2 * d1; L[8] += g2 * d2;
                if (PUSH) {
                    const double t0 = p.dtot[3 * j], t1 = p.dtot[3 * j + 1], t2 = p.dtot[3 * j + 2];
                    F[0] += g0 * t0; F[1] += g0 * t1; F[2] += g0 * t2;
                    F[3] += g1 * t0; F[4] += g1 * t1; F[5] += g1 * t2;
                    F[6] += g2 * t0; F[7] += g2 * t1; F[8] += g2 * t2;
                }
                if (FUSE_INTERP) {
                    const double f = p.shp[q];
                    u[0] += f * d0; u[1] += f * d1; u[2] += f * d2;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) L[k] = group_sum<kGroup>(L[k]);
        if (PUSH) {
#pragma unroll
            for (int k = 0; k < 9; ++k) F[k] = group_sum<kGroup>(F[k]);
        }
        if (FUSE_INTERP) {
#pragma unroll
            for (int k = 0; k < 3; ++k) u[k] = group_sum<kGroup>(u[k]);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 9; ++k) { s_sum[ln][k] = L[k]; s_sum[ln][9 + k] = F[k]; }
            s_sum[ln][18] = u[0]; s_sum[ln][19] = u[1]; s_sum[ln][20] = u[2];
        }
    }
    __syncthreads();

    // ---- phase B
    double work = 0.0;
    if (threadIdx.x < kNodesPerBlock) {
        const int ln = threadIdx.x, i = nb0 + ln;
        if (i < p.n) {
            double L0[9], Fm[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) { L0[k] = s_sum[ln][k]; Fm[k] = s_sum[ln][9 + k]; }
            const Material &mt = p.mats[p.mat[i]];
            const double vol = p.vol[i];
            double det = 1.0, IF[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            if (PUSH) {
                // sums are stored as F_{k,l} = sum g_l t_k at index 3*l + k (column-major)
                Fm[0] += 1.0; Fm[4] += 1.0; Fm[8] += 1.0;
                det = mega::det3_abs(Fm);
                mega::inv3_ref(Fm, IF);
            }
            // L(k,l) = sum_m L0(k,m) IF(m,l)
            double Lm[9];
            if (PUSH) mega::mm3(L0, IF, Lm);
            else
#pragma unroll
                for (int k = 0; k < 9; ++k) Lm[k] = L0[k];
            double sig[6], eps6[6], de[6], dd[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) { sig[k] = p.stress[6 * i + k]; eps6[k] = p.strain[6 * i + k]; }
            double old_s[6], old_e[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) { old_s[k] = sig[k]; old_e[k] = eps6[k]; }
            if (p.finite_strain) {
                double R[9];
                mega::hughes_winget(Lm, R, de, dd);
                mega::rotate_voigt(R, sig);
                mega::rotate_voigt(R, eps6);
            } else {
                de[0] = Lm[0]; de[1] = Lm[4]; de[2] = Lm[8];
                de[3] = Lm[5] + Lm[7]; de[4] = Lm[2] + Lm[6]; de[5] = Lm[1] + Lm[3];
#pragma unroll
                for (int k = 0; k < 6; ++k) dd[k] = 0.0;
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) eps6[k] = eps6[k] + de[k];
            double pred[6];
            mega::elastic_trial(sig, de, mt, pred);
            double ep = p.eps[i];
            if (mt.type == 2) {
                if (mega::von_mises(sig, pred, ep, mt) < 0) atomicAdd(&p.flags[0], 1);
            } else if (mt.type == 1) {
#pragma unroll
                for (int k = 0; k < 6; ++k) sig[k] = pred[k];
            } else {
                atomicAdd(&p.flags[1], 1);
            }
            p.eps[i] = ep;
            double si[6], ei[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                si[k] = sig[k] - old_s[k];
                ei[k] = eps6[k] - old_e[k];
                p.stress[6 * i + k] = sig[k];
                p.strain[6 * i + k] = eps6[k];
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                if (p.finite_strain) work += 0.5 * dd[k] * (2.0 * sig[k] - si[k]) * vol * det;
                else work += 0.5 * ei[k] * si[k] * vol * det;
            }
            if (p.wave_vel) p.wave_vel[i] = mega::wave_speed(si, ei, mt);
            if (p.strain_eq) {
                const double qv = eps6[0] * eps6[0] + eps6[1] * eps6[1] + eps6[2] * eps6[2] +
                                  2.0 * (eps6[3] * eps6[3] + eps6[4] * eps6[4] + eps6[5] * eps6[5]);
                p.strain_eq[i] = sqrt(qv * 2.0 / 3.0);
            }
            if (FUSE_INTERP) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double t = p.uphy[3 * i + k] + s_sum[ln][18 + k];
                    p.uphy[3 * i + k] = t;
                    p.x[3 * i + k] = p.X0[3 * i + k] + t;
                }
            }
            // P = V J sigma IF^T :  f_j = P grad phi0_j
            const double S[9] = {sig[0], sig[5], sig[4], sig[5], sig[1], sig[3], sig[4], sig[3], sig[2]};
            const double w = vol * det;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    s_sum[ln][r + 3 * c] =
                        w * (S[r] * IF[c] + S[r + 3] * IF[c + 3] + S[r + 6] * IF[c + 6]);
            if (p.has_gravity) s_sum[ln][9] = mt.rho * w;
        }
    }
    const double wsum = block_sum(work, s_red);
    if (threadIdx.x == 0 && wsum != 0.0) atomicAdd(&p.acc[0], wsum);
    // block_sum synchronised the block: phase B results are visible

    // ---- phase C
    for (int ln = grp; ln < kNodesPerBlock; ln += kGroups) {
        const int i = nb0 + ln;
        if (i >= p.n) break;
        double P[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) P[k] = s_sum[ln][k];
        const double rw = p.has_gravity ? s_sum[ln][9] : 0.0;
        const int p0 = p.row_ptr[i], p1 = p.row_ptr[i + 1];
        for (int q = p0 + lane; q < p1; q += kGroup) {
            const int j = p.col[q];
            const double g0 = p.dshp[3 * q], g1 = p.dshp[3 * q + 1], g2 = p.dshp[3 * q + 2];
            atomicAdd(&p.fint[3 * j], P[0] * g0 + P[3] * g1 + P[6] * g2);
            atomicAdd(&p.fint[3 * j + 1], P[1] * g0 + P[4] * g1 + P[7] * g2);
            atomicAdd(&p.fint[3 * j + 2], P[2] * g0 + P[5] * g1 + P[8] * g2);
            if (p.has_gravity) {
                const double f = p.shp[q] * rw;
                atomicAdd(&p.fext[3 * j], f * p.grav[0]);
                atomicAdd(&p.fext[3 * j + 1], f * p.grav[1]);
                atomicAdd(&p.fext[3 * j + 2], f * p.grav[2]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// EBC, mass-proportional damping and corrector (MAIN.F90:432-488), energies
// (MAIN.F90:524-553).  One thread per degree of freedom.

__global__ __launch_bounds__(kBlock) void k_correct(Params p)
{
    __shared__ double red[kBlock / 64];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0, ie = 0.0, ee = 0.0;
    if (q < 3 * (int64_t)p.n) {
        const int i = (int)(q / 3);
        const double f = p.fint[q];
        const double fx_old = p.has_gravity ? p.fext_nmo[q] : 0.0;
        double fx = p.has_gravity ? p.fext[q] : 0.0;
        double v = p.v[q];
        if (p.ebc[q] == 1) {
            v = 0.0;
            p.a[q] = 0.0;
            fx = -f;
        } else {
            const double mass = p.mass[q];
            const double c = p.mats[p.mat[i]].damping * mass;
            const double acc = (fx - f - c * v) / (mass + 0.5 * p.dt * c);
            p.a[q] = acc;
            v = v + 0.5 * p.dt * acc;
        }
        p.v[q] = v;
        p.fext[q] = fx;
        if (p.has_gravity) p.fext_nmo[q] = fx;
        ke = 0.5 * p.mass[q] * v * v;
        if (!p.finite_strain) ie = 0.5 * f * p.dtot[q];
        ee = 0.5 * (fx + fx_old) * p.d[q];
    }
    const double sk = block_sum(ke, red);
    const double si = block_sum(ie, red);
    const double se = block_sum(ee, red);
    if (threadIdx.x == 0) {
        if (sk != 0.0) atomicAdd(&p.acc[1], sk);
        if (si != 0.0) atomicAdd(&p.acc[2], si);
        if (se != 0.0) atomicAdd(&p.acc[3], se);
    }
}

// ---------------------------------------------------------------------------
// bins: cell of each node (GET_BINS floor/clamp, 1-based), counting sort of
// nodes into cells, node order restored inside each cell.

struct BinGeom {
    double mn[3], bin[3];
    int nb[3];
};

__global__ void k_bin_assign(int n, const double *x, BinGeom g, int32_t *ijk, int32_t *cell_count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int s = (int)floor((x[3 * i + k] - g.mn[k]) / g.bin[k]) + 1;
        s = s < 1 ? 1 : (s > g.nb[k] ? g.nb[k] : s);
        c[k] = s;
        ijk[3 * i + k] = s;
    }
    const int64_t m = (int64_t)(c[2] - 1) * g.nb[0] * g.nb[1] + (int64_t)(c[1] - 1) * g.nb[0] + (c[0] - 1);
    atomicAdd(&cell_count[m], 1);
}

__global__ void k_bin_fill(int n, const int32_t *ijk, BinGeom g, int32_t *cursor, int32_t *cell_nodes)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t m = (int64_t)(ijk[3 * i + 2] - 1) * g.nb[0] * g.nb[1] + (int64_t)(ijk[3 * i + 1] - 1) * g.nb[0] +
                      (ijk[3 * i] - 1);
    const int slot = atomicAdd(&cursor[m], 1);
    cell_nodes[slot] = i;
}

__global__ void k_bin_sort(int64_t ncell, const int32_t *cell_start, int32_t *cell_nodes)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ncell) return;
    const int s = cell_start[m], e = cell_start[m + 1];
    for (int a = s + 1; a < e; ++a) {   // cells hold a few tens of nodes
        const int v = cell_nodes[a];
        int b = a - 1;
        while (b >= s && cell_nodes[b] > v) { cell_nodes[b + 1] = cell_nodes[b]; --b; }
        cell_nodes[b + 1] = v;
    }
}

// SOFT_SEARCH: one wavefront per node.  The 27 surrounding cells are visited in
// the reference's order (ix slowest, then iy, iz; BIN_SEARCH.F90:105-127), each
// cell's nodes in node order; 64 candidates are tested at a time and the kept
// ones are compacted with a ballot, so the list comes out in the reference's
// order.  Criterion: x_i inside node j's support, |x_j - x_i|_k / a_jk < 1.
template <bool FILL>
__global__ __launch_bounds__(kBlock) void k_search(int n, const double *x, const double *win, const int32_t *ijk,
                                                   BinGeom g, const int32_t *cell_start, const int32_t *cell_nodes,
                                                   int shsup, int32_t *counts, const int32_t *row_ptr, int32_t *col)
{
    __shared__ int s_beg[kBlock / 64][28], s_off[kBlock / 64][28];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= n) return;
    const int ci = ijk[3 * i], cj = ijk[3 * i + 1], ck = ijk[3 * i + 2];
    int cnt = 0;
    if (lane < 27) {
        const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
        const int ix = ci + dx, iy = cj + dy, iz = ck + dz;
        int b = 0;
        if (ix >= 1 && ix <= g.nb[0] && iy >= 1 && iy <= g.nb[1] && iz >= 1 && iz <= g.nb[2]) {
            const int64_t m = (int64_t)(iz - 1) * g.nb[0] * g.nb[1] + (int64_t)(iy - 1) * g.nb[0] + (ix - 1);
            b = cell_start[m];
            cnt = cell_start[m + 1] - b;
        }
        s_beg[wv][lane] = b;
    }
    // exclusive prefix of the 27 counts across the wave
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane < 27) s_off[wv][lane] = incl - cnt;
    const int total = __shfl(incl, 26, 64);
    if (lane == 0) s_off[wv][27] = total;
    __builtin_amdgcn_wave_barrier();
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    int out = FILL ? row_ptr[i] : 0;
    int kept = 0;
    for (int base = 0; base < total; base += 64) {
        const int c = base + lane;
        bool add = false;
        int j = 0;
        if (c < total) {
            int cell = 0;
            while (cell < 26 && s_off[wv][cell + 1] <= c) ++cell;
            j = cell_nodes[s_beg[wv][cell] + (c - s_off[wv][cell])];
            const double ax = fabs(x[3 * j] - xi) / win[3 * j];
            const double ay = fabs(x[3 * j + 1] - yi) / win[3 * j + 1];
            const double az = fabs(x[3 * j + 2] - zi) / win[3 * j + 2];
            if (shsup) add = sqrt(ax * ax + ay * ay + az * az) < 1.0;
            else add = (ax < 1.0) && (ay < 1.0) && (az < 1.0);
        }
        const unsigned long long mask = __ballot(add);
        if (FILL && add) {
            const int pos = __popcll(mask & ((1ull << lane) - 1ull));
            col[out + pos] = j;
        }
        const int nadd = __popcll(mask);
        out += nadd;
        kept += nadd;
    }
    if (!FILL && lane == 0) counts[i] = kept;
}

// ---------------------------------------------------------------------------
// per-point constitution (context-free entry point)

__global__ void k_constitute(int n, const double *lmat, double *stress, double *strain, double *state,
                             Material mt, int finite, int32_t *fail)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double L[9], sig[6], e6[6], de[6], dd[6];
    for (int k = 0; k < 9; ++k) L[k] = lmat[9 * i + k];
    for (int k = 0; k < 6; ++k) { sig[k] = stress[6 * i + k]; e6[k] = strain[6 * i + k]; }
    if (finite) {
        double R[9];
        mega::hughes_winget(L, R, de, dd);
        mega::rotate_voigt(R, sig);
        mega::rotate_voigt(R, e6);
    } else {
        de[0] = L[0]; de[1] = L[4]; de[2] = L[8];
        de[3] = L[5] + L[7]; de[4] = L[2] + L[6]; de[5] = L[1] + L[3];
    }
    for (int k = 0; k < 6; ++k) e6[k] = e6[k] + de[k];
    double pred[6];
    mega::elastic_trial(sig, de, mt, pred);
    double ep = state[20 * i + 2];
    if (mt.type == 2) {
        if (mega::von_mises(sig, pred, ep, mt) < 0) atomicAdd(fail, 1);
    } else {
        for (int k = 0; k < 6; ++k) sig[k] = pred[k];
    }
    state[20 * i + 2] = ep;
    for (int k = 0; k < 6; ++k) { stress[6 * i + k] = sig[k]; strain[6 * i + k] = e6[k]; }
}

// ---------------------------------------------------------------------------
// host side

#define HIPCHK(expr)                                                      \
    do {                                                                  \
        hipError_t e_ = (expr);                                           \
        if (e_ != hipSuccess) {                                           \
            std::fprintf(stderr, "mega: %s failed: %s (%s:%d)\n", #expr, \
                         hipGetErrorString(e_), __FILE__, __LINE__);      \
            return MEGA_ERR_HIP;                                          \
        }                                                                 \
    } while (0)

Material make_material(const double *prop, int type)
{
    Material m{};
    m.nu = prop[0]; m.E = prop[1]; m.rho = prop[2]; m.sig0 = prop[3]; m.B = prop[4]; m.CE = prop[5];
    m.damping = prop[20];
    m.lam = prop[27]; m.mu = prop[28]; m.l2m = prop[29];
    m.bulk = m.E / (3.0 * (1.0 - 2.0 * m.nu));
    m.shear = m.E / (2.0 * (1.0 + m.nu));
    m.type = type;
    return m;
}

}  // namespace

struct mega_ctx {
    mega_control c{};
    int n = 0;
    Dev dv;
    BinGeom geom{};
    int64_t ncell = 0;
    hipStream_t stream = nullptr;
    int64_t step = 0;
    int search_count = 0;
    int has_shapes = 0;
    double dt = 0.0;
    int has_gravity = 0;
    std::vector<Material> mats;
    // profiling
    int profile = 0;
    struct Timer { const char *name; hipEvent_t a, b; };
    std::vector<Timer> timers;
    size_t timer_used = 0;
};

namespace {

template <class T>
int dalloc(T **p, size_t count)
{
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, sizeof(T) * count) != hipSuccess) return MEGA_ERR_NOMEM;
    return MEGA_OK;
}

Params params(mega_ctx *c)
{
    Dev &d = c->dv;
    Params p{};
    p.n = c->n;
    p.row_ptr = d.row_ptr; p.col = d.col; p.shp = d.shp; p.dshp = d.dshp;
    p.X0 = d.X0; p.win = d.win; p.vol = d.vol; p.mass = d.mass;
    p.x = d.x; p.v = d.v; p.a = d.a; p.d = d.d; p.dtot = d.dtot; p.uphy = d.uphy;
    p.fint = d.fint; p.fext = d.fext; p.fext_nmo = d.fext_nmo;
    p.stress = d.stress; p.strain = d.strain; p.eps = d.eps;
    p.strain_eq = d.strain_eq; p.wave_vel = d.wave_vel; p.vphy = d.vphy;
    p.ebc = d.ebc; p.ebc_node = d.ebc_node; p.mat = d.mat; p.mats = d.mats;
    p.acc = d.acc; p.flags = d.flags;
    p.ql = c->c.ql; p.ql_coef = c->c.ql_coef; p.ql_len = c->c.ql_len;
    p.shsup = c->c.shsup; p.lagrangian = c->c.lagrangian; p.finite_strain = c->c.finite_strain;
    p.has_gravity = c->has_gravity;
    for (int k = 0; k < 3; ++k) p.grav[k] = c->c.gravity[k];
    p.dt = c->dt;
    return p;
}

// event-bracketed launch for the per-kernel timing of mega_step
struct Scope {
    mega_ctx *c;
    mega_ctx::Timer *t = nullptr;
    Scope(mega_ctx *ctx, const char *name) : c(ctx)
    {
        if (!c->profile) return;
        if (c->timer_used == c->timers.size()) {
            mega_ctx::Timer nt{name, nullptr, nullptr};
            hipEventCreate(&nt.a);
            hipEventCreate(&nt.b);
            c->timers.push_back(nt);
        }
        t = &c->timers[c->timer_used++];
        t->name = name;
        hipEventRecord(t->a, c->stream);
    }
    ~Scope()
    {
        if (t) hipEventRecord(t->b, c->stream);
    }
};

unsigned grid_for(int64_t threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }

int ensure_pairs(mega_ctx *c, int64_t nnz)
{
    Dev &d = c->dv;
    if (nnz > INT32_MAX) return MEGA_ERR_ARG;
    if (nnz <= d.nnz_cap) return MEGA_OK;
    hipFree(d.col); hipFree(d.shp); hipFree(d.dshp);
    d.col = nullptr; d.shp = nullptr; d.dshp = nullptr;
    const int64_t cap = nnz + nnz / 8 + 1024;
    if (dalloc(&d.col, cap) || dalloc(&d.shp, cap) || dalloc(&d.dshp, 3 * cap)) return MEGA_ERR_NOMEM;
    d.nnz_cap = cap;
    c->has_shapes = 0;
    return MEGA_OK;
}

int cub_scan(mega_ctx *c, const int32_t *in, int32_t *out, int64_t count)
{
    size_t bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)count, c->stream);
    if (bytes > c->dv.cub_tmp_bytes) {
        hipFree(c->dv.cub_tmp);
        if (hipMalloc(&c->dv.cub_tmp, bytes) != hipSuccess) return MEGA_ERR_NOMEM;
        c->dv.cub_tmp_bytes = bytes;
    }
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->dv.cub_tmp, bytes, in, out, (int)count, c->stream));
    return MEGA_OK;
}

int do_bins(mega_ctx *c, const double *x)
{
    Dev &d = c->dv;
    const int n = c->n;
    Scope s(c, "k_bin");
    HIPCHK(hipMemsetAsync(d.cell_cursor, 0, sizeof(int32_t) * (c->ncell + 1), c->stream));
    k_bin_assign<<<grid_for(n), kBlock, 0, c->stream>>>(n, x, c->geom, d.ijk, d.cell_cursor);
    int r = cub_scan(c, d.cell_cursor, d.cell_start, c->ncell + 1);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(d.cell_cursor, d.cell_start, sizeof(int32_t) * c->ncell, hipMemcpyDeviceToDevice, c->stream));
    k_bin_fill<<<grid_for(n), kBlock, 0, c->stream>>>(n, d.ijk, c->geom, d.cell_cursor, d.cell_nodes);
    k_bin_sort<<<grid_for(c->ncell), kBlock, 0, c->stream>>>(c->ncell, d.cell_start, d.cell_nodes);
    HIPCHK(hipGetLastError());
    return MEGA_OK;
}

int do_search(mega_ctx *c, const double *x)
{
    Dev &d = c->dv;
    const int n = c->n;
    {
        Scope s(c, "k_search_count");
        k_search<false><<<grid_for((int64_t)n * 64), kBlock, 0, c->stream>>>(
            n, x, d.win, d.ijk, c->geom, d.cell_start, d.cell_nodes, c->c.shsup, d.counts, nullptr, nullptr);
    }
    HIPCHK(hipMemsetAsync(d.counts + n, 0, sizeof(int32_t), c->stream));
    int r = cub_scan(c, d.counts, d.row_ptr, n + 1);
    if (r) return r;
    int32_t nnz32 = 0;
    HIPCHK(hipMemcpyAsync(&nnz32, d.row_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (nnz32 < 0) return MEGA_ERR_ARG;
    r = ensure_pairs(c, nnz32);
    if (r) return r;
    d.nnz = nnz32;
    {
        Scope s(c, "k_search_fill");
        k_search<true><<<grid_for((int64_t)n * 64), kBlock, 0, c->stream>>>(
            n, x, d.win, d.ijk, c->geom, d.cell_start, d.cell_nodes, c->c.shsup, nullptr, d.row_ptr, d.col);
    }
    HIPCHK(hipGetLastError());
    return MEGA_OK;
}

// the HANDELER search-count rule (HANDELER.F90:256-268): rebin when the count
// has reached the period; during the first step the period is not yet set
int search_step(mega_ctx *c, const double *x, int first)
{
    const int period = first ? 0 : c->c.pdsearch;
    if (c->search_count < period) c->search_count++;
    else {
        int r = do_bins(c, x);
        if (r) return r;
        c->search_count = 1;
    }
    return do_search(c, x);
}

int do_shapes(mega_ctx *c, const double *xs)
{
    Params p = params(c);
    Scope s(c, "k_rk");
    k_rk<<<grid_for((int64_t)c->n * kGroup), kBlock, 0, c->stream>>>(p, xs);
    HIPCHK(hipGetLastError());
    c->has_shapes = 1;
    return MEGA_OK;
}

int do_fint(mega_ctx *c, int push, int fuse_interp)
{
    Params p = params(c);
    const unsigned grid = (unsigned)((c->n + kNodesPerBlock - 1) / kNodesPerBlock);
    Scope s(c, "k_fint");
    if (push && fuse_interp) k_fint<true, true><<<grid, kBlock, 0, c->stream>>>(p);
    else if (push) k_fint<true, false><<<grid, kBlock, 0, c->stream>>>(p);
    else if (fuse_interp) k_fint<false, true><<<grid, kBlock, 0, c->stream>>>(p);
    else k_fint<false, false><<<grid, kBlock, 0, c->stream>>>(p);
    HIPCHK(hipGetLastError());
    return MEGA_OK;
}

int check_flags(mega_ctx *c)
{
    int32_t f[2];
    HIPCHK(hipMemcpyAsync(f, c->dv.flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (f[1]) return MEGA_ERR_MATERIAL;
    if (f[0]) return MEGA_ERR_RETURN_MAP;
    return MEGA_OK;
}

double *field_ptr(mega_ctx *c, int f, int *width)
{
    Dev &d = c->dv;
    *width = 3;
    switch (f) {
    case MEGA_F_X: return d.x;
    case MEGA_F_V: return d.v;
    case MEGA_F_A: return d.a;
    case MEGA_F_D: return d.d;
    case MEGA_F_DTOT: return d.dtot;
    case MEGA_F_UPHY: return d.uphy;
    case MEGA_F_FINT: return d.fint;
    case MEGA_F_FEXT: return d.fext;
    case MEGA_F_STRESS: *width = 6; return d.stress;
    case MEGA_F_STRAIN: *width = 6; return d.strain;
    case MEGA_F_EPS: *width = 1; return d.eps;
    case MEGA_F_STRAIN_EQ: *width = 1; return d.strain_eq;
    case MEGA_F_WAVE_VEL: *width = 1; return d.wave_vel;
    case MEGA_F_VPHY: return d.vphy;
    default: return nullptr;
    }
}

}  // namespace

// ===========================================================================
// C-ABI

extern "C" {

const char *mega_status_string(int s)
{
    switch (s) {
    case MEGA_OK: return "ok";
    case MEGA_ERR_RETURN_MAP: return "plasticity material model did not converge";
    case MEGA_ERR_SINGULAR: return "problem inverting matrix";
    case MEGA_ERR_BINS: return "bin table inconsistent";
    case MEGA_ERR_NOMEM: return "device allocation failed";
    case MEGA_ERR_HIP: return "HIP runtime error";
    case MEGA_ERR_ARG: return "invalid argument or unsupported option";
    case MEGA_ERR_MATERIAL: return "invalid material type on this path";
    default: return "unknown status";
    }
}

int mega_create(const mega_control *ctrl, int32_t n, const double *X0, const double *win, const double *vol,
                const double *mass, const int32_t *ebc, const int32_t *ebc_node, const int32_t *mat_type,
                const double *prop, const double *vinit, mega_ctx **out)
{
    if (!ctrl || !out || n <= 0 || !X0 || !win || !vol || !mass || !ebc || !ebc_node || !mat_type || !prop)
        return MEGA_ERR_ARG;
    if (ctrl->rk_degree != 1 || ctrl->rk_cont != 3 || ctrl->itype_int != 0) return MEGA_ERR_ARG;
    if (ctrl->precision != MEGA_FP64) return MEGA_ERR_ARG;
    *out = nullptr;
    mega_ctx *c = new mega_ctx();
    c->c = *ctrl;
    if (c->c.pdsearch < 1) c->c.pdsearch = 1;
    c->n = n;
    c->has_gravity = (ctrl->gravity[0] != 0.0 || ctrl->gravity[1] != 0.0 || ctrl->gravity[2] != 0.0);
    // materials: nodes with identical property rows share a slot
    std::vector<int32_t> mid(n);
    std::map<std::vector<double>, int> seen;
    for (int i = 0; i < n; ++i) {
        std::vector<double> row(prop + 30 * (int64_t)i, prop + 30 * (int64_t)i + 30);
        row.push_back((double)mat_type[i]);
        auto it = seen.find(row);
        if (it == seen.end()) {
            const int k = (int)c->mats.size();
            seen.emplace(row, k);
            c->mats.push_back(make_material(prop + 30 * (int64_t)i, mat_type[i]));
            mid[i] = k;
        } else {
            mid[i] = it->second;
        }
    }
    // bin geometry from initial positions and windows (HANDELER.F90:187-252)
    double bin[3] = {0, 0, 0}, mx[3], mn[3];
    for (int k = 0; k < 3; ++k) mx[k] = mn[k] = X0[k];
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const double w = win[3 * (int64_t)i + k], xv = X0[3 * (int64_t)i + k];
            if (w > bin[k]) bin[k] = w;
            if (w + ctrl->dist_max[k] > bin[k]) bin[k] = w + ctrl->dist_max[k];
            if (xv > mx[k]) mx[k] = xv;
            if (xv < mn[k]) mn[k] = xv;
        }
    for (int k = 0; k < 3; ++k) {
        mn[k] = mn[k] - bin[k] * 10.0;
        mx[k] = mx[k] + bin[k] * 10.0;
        c->geom.nb[k] = (int)std::ceil((mx[k] - mn[k]) / bin[k]);
        c->geom.mn[k] = mn[k];
        c->geom.bin[k] = bin[k];
    }
    c->ncell = (int64_t)c->geom.nb[0] * c->geom.nb[1] * c->geom.nb[2];
    if (c->ncell <= 0 || c->ncell > INT32_MAX) { delete c; return MEGA_ERR_ARG; }

    Dev &d = c->dv;
    d.n = n;
    const size_t n3 = 3 * (size_t)n, n6 = 6 * (size_t)n;
    int r = 0;
    r |= dalloc(&d.X0, n3); r |= dalloc(&d.x, n3); r |= dalloc(&d.v, n3); r |= dalloc(&d.a, n3);
    r |= dalloc(&d.d, n3); r |= dalloc(&d.dtot, n3); r |= dalloc(&d.uphy, n3); r |= dalloc(&d.fint, n3);
    r |= dalloc(&d.fext, n3); r |= dalloc(&d.vphy, n3); r |= dalloc(&d.scratch3, n3);
    if (c->has_gravity) r |= dalloc(&d.fext_nmo, n3);
    r |= dalloc(&d.stress, n6); r |= dalloc(&d.strain, n6); r |= dalloc(&d.eps, n);
    r |= dalloc(&d.win, n3); r |= dalloc(&d.vol, n); r |= dalloc(&d.mass, n3);
    r |= dalloc(&d.strain_eq, n); r |= dalloc(&d.wave_vel, n);
    r |= dalloc(&d.ebc, n3); r |= dalloc(&d.ebc_node, n); r |= dalloc(&d.mat, n);
    r |= dalloc(&d.mats, c->mats.size());
    r |= dalloc(&d.row_ptr, n + 1); r |= dalloc(&d.counts, n + 1);
    r |= dalloc(&d.ijk, n3); r |= dalloc(&d.cell_start, c->ncell + 1); r |= dalloc(&d.cell_cursor, c->ncell + 1);
    r |= dalloc(&d.cell_nodes, n);
    r |= dalloc(&d.acc, 8); r |= dalloc(&d.flags, 4);
    if (r) { mega_destroy(c); return MEGA_ERR_NOMEM; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { mega_destroy(c); return MEGA_ERR_HIP; }
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = up(d.X0, X0, n3 * 8) && up(d.x, X0, n3 * 8) && up(d.win, win, n3 * 8) && up(d.vol, vol, (size_t)n * 8) &&
              up(d.mass, mass, n3 * 8) && up(d.ebc, ebc, n3 * 4) && up(d.ebc_node, ebc_node, (size_t)n * 4) &&
              up(d.mat, mid.data(), (size_t)n * 4) && up(d.mats, c->mats.data(), c->mats.size() * sizeof(Material));
    if (vinit) ok = ok && up(d.v, vinit, n3 * 8);
    else ok = ok && hipMemset(d.v, 0, n3 * 8) == hipSuccess;
    ok = ok && hipMemset(d.a, 0, n3 * 8) == hipSuccess && hipMemset(d.d, 0, n3 * 8) == hipSuccess &&
         hipMemset(d.dtot, 0, n3 * 8) == hipSuccess && hipMemset(d.uphy, 0, n3 * 8) == hipSuccess &&
         hipMemset(d.fint, 0, n3 * 8) == hipSuccess && hipMemset(d.fext, 0, n3 * 8) == hipSuccess &&
         hipMemset(d.vphy, 0, n3 * 8) == hipSuccess && hipMemset(d.stress, 0, n6 * 8) == hipSuccess &&
         hipMemset(d.strain, 0, n6 * 8) == hipSuccess && hipMemset(d.eps, 0, (size_t)n * 8) == hipSuccess &&
         hipMemset(d.strain_eq, 0, (size_t)n * 8) == hipSuccess && hipMemset(d.wave_vel, 0, (size_t)n * 8) == hipSuccess &&
         hipMemset(d.acc, 0, 64) == hipSuccess && hipMemset(d.flags, 0, 16) == hipSuccess;
    if (ok && d.fext_nmo) ok = hipMemset(d.fext_nmo, 0, n3 * 8) == hipSuccess;
    if (!ok) { mega_destroy(c); return MEGA_ERR_HIP; }
    *out = c;
    return MEGA_OK;
}

int mega_destroy(mega_ctx *c)
{
    if (!c) return MEGA_OK;
    Dev &d = c->dv;
    void *ptrs[] = {d.X0, d.x, d.v, d.a, d.d, d.dtot, d.uphy, d.fint, d.fext, d.fext_nmo, d.stress, d.strain,
                    d.eps, d.win, d.vol, d.mass, d.strain_eq, d.wave_vel, d.vphy, d.scratch3, d.ebc, d.ebc_node,
                    d.mat, d.mats, d.row_ptr, d.col, d.shp, d.dshp, d.ijk, d.cell_start, d.cell_nodes,
                    d.cell_cursor, d.counts, d.acc, d.flags, d.cub_tmp};
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto &t : c->timers) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return MEGA_OK;
}

int mega_search(mega_ctx *c, const double *x_cur, int32_t rebuild_bins)
{
    if (!c) return MEGA_ERR_ARG;
    const double *xs = c->dv.x;
    if (x_cur) {
        HIPCHK(hipMemcpyAsync(c->dv.scratch3, x_cur, sizeof(double) * 3 * c->n, hipMemcpyHostToDevice, c->stream));
        xs = c->dv.scratch3;
    }
    if (rebuild_bins || c->search_count == 0) {
        int r = do_bins(c, xs);
        if (r) return r;
        c->search_count = 1;
    }
    int r = do_search(c, xs);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    return MEGA_OK;
}

int mega_get_neighbors(mega_ctx *c, int32_t *gn, int32_t *gstart, int32_t *gstack, int64_t cap, int64_t *nnz)
{
    if (!c) return MEGA_ERR_ARG;
    const int n = c->n;
    std::vector<int32_t> rp(n + 1);
    HIPCHK(hipMemcpyAsync(rp.data(), c->dv.row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (nnz) *nnz = rp[n];
    if (gn)
        for (int i = 0; i < n; ++i) gn[i] = rp[i + 1] - rp[i];
    if (gstart)
        for (int i = 0; i < n; ++i) gstart[i] = rp[i] + 1;
    if (gstack) {
        if (cap < rp[n]) return MEGA_ERR_ARG;
        HIPCHK(hipMemcpyAsync(gstack, c->dv.col, sizeof(int32_t) * rp[n], hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int64_t q = 0; q < rp[n]; ++q) gstack[q] += 1;
    }
    return MEGA_OK;
}

int mega_shapes(mega_ctx *c, const double *x_eval)
{
    if (!c || c->dv.nnz_cap == 0) return MEGA_ERR_ARG;
    const double *xs = c->dv.X0;
    if (x_eval) {
        HIPCHK(hipMemcpyAsync(c->dv.scratch3, x_eval, sizeof(double) * 3 * c->n, hipMemcpyHostToDevice, c->stream));
        xs = c->dv.scratch3;
    }
    int r = do_shapes(c, xs);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    return MEGA_OK;
}

int mega_get_shapes(mega_ctx *c, double *shp, double *dshp, int64_t cap)
{
    if (!c || cap < c->dv.nnz) return MEGA_ERR_ARG;
    if (shp) HIPCHK(hipMemcpyAsync(shp, c->dv.shp, sizeof(double) * c->dv.nnz, hipMemcpyDeviceToHost, c->stream));
    if (dshp) HIPCHK(hipMemcpyAsync(dshp, c->dv.dshp, sizeof(double) * 3 * c->dv.nnz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MEGA_OK;
}

int mega_interp(mega_ctx *c, const double *coef, double *phys)
{
    if (!c || !coef || !phys || !c->has_shapes) return MEGA_ERR_ARG;
    Dev &d = c->dv;
    const size_t b = sizeof(double) * 3 * c->n;
    double *out = nullptr;
    if (dalloc(&out, 3 * (size_t)c->n)) return MEGA_ERR_NOMEM;
    int st = MEGA_OK;
    if (hipMemcpyAsync(d.scratch3, coef, b, hipMemcpyHostToDevice, c->stream) != hipSuccess) st = MEGA_ERR_HIP;
    if (!st) {
        Params p = params(c);
        k_interp<false><<<grid_for((int64_t)c->n * kGroup), kBlock, 0, c->stream>>>(p, d.scratch3, out);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(phys, out, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            st = MEGA_ERR_HIP;
    }
    hipFree(out);
    return st;
}

int mega_fint(mega_ctx *c, int32_t linit, const double *dinc, const double *dinc_tot, const double *x_cur,
              double *stress, double *strain, double *state, double *fint, double *fext, double *int_work)
{
    if (!c || !dinc || !dinc_tot || !stress || !strain || !state || !fint || !fext || c->dv.nnz_cap == 0)
        return MEGA_ERR_ARG;
    Dev &d = c->dv;
    const int n = c->n;
    const size_t b3 = sizeof(double) * 3 * n, b6 = sizeof(double) * 6 * n;
    std::vector<double> eps(n);
    for (int i = 0; i < n; ++i) eps[i] = state[20 * (int64_t)i + 2];
    HIPCHK(hipMemcpyAsync(d.d, dinc, b3, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d.dtot, dinc_tot, b3, hipMemcpyHostToDevice, c->stream));
    if (x_cur) HIPCHK(hipMemcpyAsync(d.x, x_cur, b3, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d.stress, stress, b6, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d.strain, strain, b6, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d.eps, eps.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(d.fint, 0, b3, c->stream));
    HIPCHK(hipMemsetAsync(d.fext, 0, b3, c->stream));
    HIPCHK(hipMemsetAsync(d.acc, 0, sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(d.flags, 0, 16, c->stream));
    int r = 0;
    const bool lag = c->c.lagrangian;
    if (linit || !lag) {
        r = do_shapes(c, lag ? d.X0 : d.x);
        if (r) return r;
    } else if (!c->has_shapes) {
        return MEGA_ERR_ARG;
    }
    r = do_fint(c, lag && c->c.finite_strain, 0);
    if (r) return r;
    const int st = check_flags(c);
    double w = 0.0;
    HIPCHK(hipMemcpyAsync(&w, d.acc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(stress, d.stress, b6, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(strain, d.strain, b6, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(eps.data(), d.eps, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fint, d.fint, b3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fext, d.fext, b3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) state[20 * (int64_t)i + 2] = eps[i];
    if (int_work) *int_work += w;
    return st;
}

int mega_constitute(int32_t n, const double *lmat, double *stress, double *strain, double *state, const double *prop,
                    int32_t mat_type, int32_t finite_strain, int32_t precision)
{
    if (n <= 0 || !lmat || !stress || !strain || !state || !prop) return MEGA_ERR_ARG;
    if (mat_type != 1 && mat_type != 2) return MEGA_ERR_MATERIAL;
    if (precision != MEGA_FP64) return MEGA_ERR_ARG;
    const Material mt = make_material(prop, mat_type);
    double *dl = nullptr, *ds = nullptr, *de = nullptr, *dst = nullptr;
    int32_t *fail = nullptr;
    const size_t nn = (size_t)n;
    if (dalloc(&dl, 9 * nn) || dalloc(&ds, 6 * nn) || dalloc(&de, 6 * nn) || dalloc(&dst, 20 * nn) || dalloc(&fail, 1)) {
        hipFree(dl); hipFree(ds); hipFree(de); hipFree(dst); hipFree(fail);
        return MEGA_ERR_NOMEM;
    }
    int st = MEGA_OK;
    int32_t f = 0;
    if (hipMemcpy(dl, lmat, 72 * nn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ds, stress, 48 * nn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(de, strain, 48 * nn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dst, state, 160 * nn, hipMemcpyHostToDevice) != hipSuccess || hipMemset(fail, 0, 4) != hipSuccess) {
        st = MEGA_ERR_HIP;
    } else {
        k_constitute<<<grid_for(n), kBlock>>>(n, dl, ds, de, dst, mt, finite_strain, fail);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(stress, ds, 48 * nn, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(strain, de, 48 * nn, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(state, dst, 160 * nn, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&f, fail, 4, hipMemcpyDeviceToHost) != hipSuccess)
            st = MEGA_ERR_HIP;
        else if (f)
            st = MEGA_ERR_RETURN_MAP;
    }
    hipFree(dl); hipFree(ds); hipFree(de); hipFree(dst); hipFree(fail);
    return st;
}

int mega_step(mega_ctx *c, double dt, int32_t nsteps)
{
    if (!c || nsteps < 0 || !(dt > 0.0)) return MEGA_ERR_ARG;
    c->dt = dt;
    if (c->profile) c->timer_used = 0;
    Dev &d = c->dv;
    const int n = c->n;
    const bool lag = c->c.lagrangian;
    const bool push = lag && c->c.finite_strain;
    HIPCHK(hipMemsetAsync(d.flags, 0, 16, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        const bool first = (c->step == 0);
        int r = 0;
        if (first) {
            // LINIT call of HANDELER with zero increments (MAIN.F90:320-334)
            HIPCHK(hipMemsetAsync(d.d, 0, sizeof(double) * 3 * n, c->stream));
            r = search_step(c, d.x, 1);
            if (!r) r = do_shapes(c, lag ? d.X0 : d.x);
            if (!r) r = do_fint(c, push, 0);
            if (r) return r;
        }
        Params p = params(c);
        {
            Scope sc(c, "k_predict");
            k_predict<<<grid_for(3 * (int64_t)n), kBlock, 0, c->stream>>>(p);
        }
        if (lag) {
            r = do_fint(c, push, 1);   // DO_INTERP fused into the gather phase
        } else {
            {
                Scope sc(c, "k_interp");
                k_interp<true><<<grid_for((int64_t)n * kGroup), kBlock, 0, c->stream>>>(p, d.d, nullptr);
            }
            r = search_step(c, d.x, first);
            if (!r) r = do_shapes(c, d.x);
            if (!r) r = do_fint(c, false, 0);
        }
        if (r) return r;
        {
            Scope sc(c, "k_correct");
            k_correct<<<grid_for(3 * (int64_t)n), kBlock, 0, c->stream>>>(p);
        }
        HIPCHK(hipGetLastError());
        c->step++;
    }
    return check_flags(c);
}

int mega_download(mega_ctx *c, int32_t field, double *host)
{
    if (!c || !host) return MEGA_ERR_ARG;
    int w = 0;
    double *src = field_ptr(c, field, &w);
    if (!src) return MEGA_ERR_ARG;
    if (field == MEGA_F_VPHY && c->has_shapes) {   // sum_j phi_ij v_j of the current velocity
        Params p = params(c);
        k_interp<false><<<grid_for((int64_t)c->n * kGroup), kBlock, 0, c->stream>>>(p, c->dv.v, c->dv.vphy);
    }
    HIPCHK(hipMemcpyAsync(host, src, sizeof(double) * w * c->n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MEGA_OK;
}

int mega_upload(mega_ctx *c, int32_t field, const double *host)
{
    if (!c || !host) return MEGA_ERR_ARG;
    int w = 0;
    double *dst = field_ptr(c, field, &w);
    if (!dst) return MEGA_ERR_ARG;
    HIPCHK(hipMemcpyAsync(dst, host, sizeof(double) * w * c->n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MEGA_OK;
}

int mega_energies(mega_ctx *c, double *e_int, double *e_kin, double *e_ext)
{
    if (!c) return MEGA_ERR_ARG;
    double acc[4];
    HIPCHK(hipMemcpyAsync(acc, c->dv.acc, sizeof(acc), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (e_int) *e_int = c->c.finite_strain ? acc[0] : acc[2];
    if (e_kin) *e_kin = acc[1];
    if (e_ext) *e_ext = acc[3];
    return MEGA_OK;
}

int mega_stats(mega_ctx *c, int64_t *nnz, int32_t *max_neighbors, int64_t *fallbacks)
{
    if (!c) return MEGA_ERR_ARG;
    const int n = c->n;
    std::vector<int32_t> rp(n + 1);
    HIPCHK(hipMemcpyAsync(rp.data(), c->dv.row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int32_t mx = 0;
    for (int i = 0; i < n; ++i) mx = std::max(mx, rp[i + 1] - rp[i]);
    if (nnz) *nnz = rp[n];
    if (max_neighbors) *max_neighbors = mx;
    if (fallbacks) *fallbacks = 0;
    return MEGA_OK;
}

int mega_profile(mega_ctx *c, int32_t enable)
{
    if (!c) return MEGA_ERR_ARG;
    c->profile = enable ? 1 : 0;
    c->timer_used = 0;
    return MEGA_OK;
}

int mega_kernel_times(mega_ctx *c, const char **names, double *ms_avg, int32_t *launches, int32_t cap)
{
    if (!c) return MEGA_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int>> agg;
    std::map<std::string, const char *> cname;
    for (size_t k = 0; k < c->timer_used; ++k) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->timers[k].a, c->timers[k].b) != hipSuccess) continue;
        std::string nm = c->timers[k].name;
        if (!agg.count(nm)) { order.push_back(nm); cname[nm] = c->timers[k].name; }
        agg[nm].first += ms;
        agg[nm].second += 1;
    }
    int k = 0;
    for (auto &nm : order) {
        if (k >= cap) break;
        if (names) names[k] = cname[nm];
        if (ms_avg) ms_avg[k] = agg[nm].first / agg[nm].second;
        if (launches) launches[k] = agg[nm].second;
        ++k;
    }
    return k;
}

}  // extern "C"
