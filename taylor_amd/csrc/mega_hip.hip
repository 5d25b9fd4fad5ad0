// mega_hip.hip -- MI355X (gfx950) explicit meshfree step behind the C-ABI in
// include/mega.h.
//
// One explicit step (MAIN.F90:304-488) runs as three kernels over HBM-resident
// arrays, plus the search/shape kernels when the formulation needs them:
//
//   k_gather   16 lanes per node (8 over block source sets) stream the node's
//              neighbor list (CSR, coalesced) and gather the neighbors' increments:
//              F = I + sum grad phi (x) dtot, L0 = sum grad phi (x) d,
//              u = sum phi d (DO_INTERP, fused when Lagrangian)
//                                            FINT.F90:310-391,669-679; HANDELER.F90:114-133
//   k_node     one lane per node: |det F|, INV3 push-forward, Hughes-Winget,
//              rotation, elastic trial, J2 radial return (k_node_surr: the
//              learned return map first), state write-back, work, and
//              P = V J sigma F^-T
//                                            FINT.F90:758-870; CONSTITUTION.F90:18-89
//   k_force    kFGroup lanes per node over the transposed neighbor list:
//              f_j = sum_i P_i grad phi_ij (= B^T sigma V J scattered by the
//              reference, FINT.F90:997-1116, gathered here without atomics,
//              deterministic), then EBC + corrector (MAIN.F90:432-488), energies
//              (MAIN.F90:524-553) and the next step's predictor (MAIN.F90:353-355)
//   k_predict  predictor when it is not fused into the previous k_force
//   k_bin_*    GET_BINS as a counting sort     BIN_SEARCH.F90:2-64
//   k_search   SOFT_SEARCH (+BIN_INODE)        BIN_SEARCH.F90:69-133,153-328
//              and, in the same two passes, the transposed neighbor lists
//   k_rk       RK1 shape functions/gradients   RK.f90:16-453, KERNEL.F90:3-140
//              (lists rebuilt every step: with the gather sums of k_gather,
//              which then does not run)
//
// Precision: R = double (MEGA_FP64) or float (MEGA_MIXED) for shape data,
// stress/strain/plastic strain, the return map and the force; positions,
// velocities, displacements, the moment matrix and its inverse, F, F^-1 and
// the Hughes-Winget rotation are FP64 in both modes.
//
// Layout in HBM: node vectors interleaved (3*i+k); stress/strain 6 per node
// (Voigt XX,YY,ZZ,YZ,XZ,XY); neighbor lists as 0-based CSR (row_ptr int32[n+1],
// col int32[nnz]) with phi[nnz], grad phi[3 nnz] in list order; the transposed
// lists (trow[n+1], tsrc[nnz]; sorted by source when the lists stay for the
// run, else grouped by the source's bin) carry {grad phi, source} records
// (TPair) in transposed order.
// Pair kernels map blocks to XCDs in contiguous node ranges so that each
// XCD's L2 holds the neighbor data its blocks gather.  For lists that stay
// for the run (Lagrangian), each block of consecutive (Morton-ordered) nodes
// has its set of distinct sources (k_blk_set, about a fifth of its list
// entries): k_force stages their P rows and k_gather their increments in
// LDS once, and the list entries name sources by set position (the gather
// reads {grad phi, phi} of an entry as one record, 16 B mixed, 32 B FP64).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "mega.h"
#include "mega_math.hpp"
#include "mlp_mfma.hpp"

using mega::Material;

namespace {

#ifndef MEGA_KGROUP
#define MEGA_KGROUP 16
#endif
#ifndef MEGA_RK_GROUP
#define MEGA_RK_GROUP 16
#endif
#ifndef MEGA_RK_CACHE
#define MEGA_RK_CACHE 4
#endif
static_assert(MEGA_RK_CACHE >= 1, "the shape-function kernel keeps at least one neighbor per lane");
#ifndef MEGA_SET_CAP
#define MEGA_SET_CAP 512
#endif
#ifndef MEGA_GGROUP_SETS
#define MEGA_GGROUP_SETS 8
#endif
#ifndef MEGA_UNROLL_SETS
#define MEGA_UNROLL_SETS 6
#endif
#ifndef MEGA_UNROLL_SETS_F64
#define MEGA_UNROLL_SETS_F64 4
#endif
#ifndef MEGA_UNROLL_MIX
#define MEGA_UNROLL_MIX 4
#endif
#ifndef MEGA_UNROLL_F64
#define MEGA_UNROLL_F64 4
#endif
#ifndef MEGA_FGROUP
#define MEGA_FGROUP 8
#endif
#ifndef MEGA_FUNROLL_MIX
#define MEGA_FUNROLL_MIX 5
#endif
#ifndef MEGA_FUNROLL_F64
#define MEGA_FUNROLL_F64 5
#endif
constexpr int kBlock = 256;
constexpr int kGroup = MEGA_KGROUP;   // lanes cooperating on one node's neighbor list
constexpr int kFGroup = MEGA_FGROUP;  // the same for the force gather over the transposed list
constexpr int kGGroupSets = MEGA_GGROUP_SETS;   // the mixed RK gather over block source sets
constexpr int kRkGroup = MEGA_RK_GROUP;   // the same for the shape-function kernel
constexpr int kRkCache = MEGA_RK_CACHE;   // kernel evaluations per lane kept between its two passes
constexpr int kXcd = 8;               // XCDs on MI355X
// neighbor-list entries each lane keeps in flight in the gathers
constexpr int kUnrollMix = MEGA_UNROLL_MIX, kUnrollF64 = MEGA_UNROLL_F64, kUnrollSets = MEGA_UNROLL_SETS;
constexpr int kUnrollSetsF64 = MEGA_UNROLL_SETS_F64;
constexpr int kFUnrollMix = MEGA_FUNROLL_MIX, kFUnrollF64 = MEGA_FUNROLL_F64;
constexpr int kNodesPerBlock = kBlock / kGroup;    // nodes of one block in the gathers
// lanes per node and nodes per block of the RK gather (k_gather)
// (the packed records and LDS-staged increments need fewer lanes per node)
constexpr int gather_group(bool sets) { return sets ? kGGroupSets : kGroup; }
constexpr int gather_npb(bool sets) { return kBlock / gather_group(sets); }
constexpr int kNodesPerFBlock = kBlock / kFGroup;  // nodes of one block in the force gather
static_assert(kFGroup <= 16 && kGroup <= 16 && kGGroupSets <= 16, "groups inside one DPP row");
// Source sets: a block stages the values its rows gather (P rows, increments)
// once in LDS when it has at most kSetCap distinct sources.  512 P rows of
// 40 B (mixed) plus the force kernel's own 1 KB let 7 of its blocks (its
// register limit) share a CU's 160 KB of LDS; larger sets read global memory.
constexpr int kSetCap = MEGA_SET_CAP;
// peridynamics (mixed): 32 particles per block, 8 lanes each, and sets of up
// to kPdSetCap sources (a 32-particle brick names ~600-900 within a horizon of
// three spacings); the force stages 40 B T rows, the gather u and d (32 B)
#ifndef MEGA_PD_SET_CAP
#define MEGA_PD_SET_CAP 1024
#endif
#ifndef MEGA_PD_GROUP
#define MEGA_PD_GROUP 8
#endif
#ifndef MEGA_PD_UNROLL
#define MEGA_PD_UNROLL 4
#endif
constexpr int kPdSetCap = MEGA_PD_SET_CAP;
// peridynamics in FP64: the kernels' blocks of kNodesPerBlock particles (16 lanes
// each) and source sets: up to kPdSetCap64 80 B rows (X0, x, d) for the gather,
// up to kPdFSetCap64 96 B rows (T, X0) for the force (four force blocks' sets
// fill a CU's 160 KB; a larger set reads global memory)
#ifndef MEGA_PD_SET_CAP64
#define MEGA_PD_SET_CAP64 640
#endif
#ifndef MEGA_PD_FSET_CAP64
#define MEGA_PD_FSET_CAP64 416
#endif
constexpr int kPdSetCap64 = MEGA_PD_SET_CAP64, kPdFSetCap64 = MEGA_PD_FSET_CAP64;
constexpr int kPdG = MEGA_PD_GROUP;       // lanes per particle
constexpr int kPdNpb = kBlock / kPdG;     // particles per block (and per source set)
constexpr int kPdUnroll = MEGA_PD_UNROLL; // bonds per lane in flight
constexpr int kSetSort = 4096;   // list entries of one block the set builder sorts in LDS
static_assert(kSetCap <= 65535, "set positions are 16-bit");

// ---------------------------------------------------------------------------
// kernel arguments

// One transposed-list entry as the force gather reads it: grad phi_ij of the
// pair (i -> j) and its source i, in one aligned record (16 B mixed, 32 B FP64).
template <class R> struct TPair;
template <> struct alignas(16) TPair<float> {
    float gx, gy, gz;
    int32_t src;
};
template <> struct alignas(32) TPair<double> {
    double gx, gy, gz;
    int32_t src, pad;
};

struct Params {
    int n;
    int64_t cap;                                     // component stride of the pair arrays
    // lists
    const int32_t *row_ptr, *col, *trow, *tsrc, *tpos;
    // per-block source sets of lists that stay for the run (null: the gathers
    // read global memory): set of block b at [ptr[b], ptr[b+1]) of set
    const int32_t *gset_ptr, *gset, *fset_ptr, *fset;
    const uint16_t *gloc;                            // forward slot -> position in its block's set
    const float4 *fpair;                             // mixed, with gloc: {grad phi, phi} per forward slot
    const float2 *pd8;                               // peridynamics, mixed: 8 B bond records, gather
                                                     // [0, cap) and force [cap, 2 cap)
    const double *pd_qb;                             // per block of kPdNpb particles: the fixed-point
                                                     // quantum of its bond vectors
    const double2 *pd16;                             // peridynamics, FP64: 16 B bond records {w V, set
                                                     // position}, gather [0, cap) and force [cap, 2 cap)
    const double2 *fpair2;                           // FP64, with gloc: {gx, gy}, {gz, phi} per forward slot
    void *shp, *dshp, *tpair, *tshp;                 // R; grad phi as x, y, z blocks at +0, +cap, +2 cap;
                                                     // tpair: TPair<R> per transposed slot
    // node arrays, FP64 masters
    const double *X0, *win, *winv, *vol, *mass;   // winv: 1 / win
    double *x, *v, *a, *d, *dtot, *uphy, *fint, *fext, *fext_nmo, *wave_vel;
    double *Fst;                                     // deformation gradient per node (9, column-major, + 1 pad)
    float4 *d4, *t4;                                 // FP32 gather copies of d and dtot (mixed)
    float4 *u4;                                      // peridynamics, mixed: FP32 x - X0 for the gather
    // state and scratch in R
    void *stress, *strain, *eps, *red, *P;
    void *st4;   // R[4n]: the damage models' EPS_SHEAR, EPS_VOL, D, DAM_SV (STATE(1),(2),(4),(5))
    const int32_t *ebc, *ebc_node, *mat;
    const Material *mats;
    double *wpart;                                   // per-block cumulative work (k_node)
    double *epart;                                   // per-block [ke, ie, ee cumulative] (k_force)
    int32_t *flags;                                  // [0] return-map failures [1] bad material
    // learned return map (k_node, mixed precision)
    const uint16_t *surr_tiles;                      // packed bf16 forward tiles
    const float *surr_bias, *surr_norm;
    int surr_mat;                                    // material slot it was trained for
    float surr_tol;                                  // yield residual / sigma_y accepted
    float surr_env[2 * mlp::kIn];                    // training envelope: per-input min, then max
                                                     // (kernel arguments: scalar loads)
    unsigned long long *sstat;                       // per k_node block: [plastic, outside envelope,
                                                     // evaluated, rejected], summed on the host
    // tuple harvest (k_node): 20 floats per plastic point
    float *hv;
    unsigned long long *hv_count;
    int64_t hv_cap;
    int ql;
    double ql_coef, ql_len;
    int shsup, finite_strain, has_gravity;
    double grav[3];
    double dt;
    const double *dt_dev;                            // automatic time step: this step's DLT (device), else dt
    double *char_dist;                               // automatic time step: the PD gather's shortest bond
                                                     // per particle (GCHAR_DIST), else null
    // stores of values that no kernel of the step reads: out_dof, fint / a / fext
    // in the corrector; out_x, x = X0 + u in the Lagrangian gather.  Within
    // mega_step they are made on the steps whose values are read (a by the
    // last step's DO_INTERP, the rest by the caller after the call)
    int out_dof, out_x;
    int mass_node;                                   // mass[n]: one value per node (all three equal)
    // correspondence peridynamics
    int pd, pd_linit;
    double *Kinv;                                    // 9 per node, inverse shape tensor
    double *gvec;                                    // 3 per node, sum_j w V_j (X_i - X_j)
};

template <class R> __device__ __forceinline__ R *as(void *p) { return static_cast<R *>(p); }
template <class R> constexpr int pstride() { return std::is_same<R, float>::value ? 12 : 10; }

// Per-node sums of the RK gather for the node update, in whole 16 B pieces:
// mixed [L0 (9) | S (3)], FP64 [L0 (9) | 0]; with F_DIRECT followed by
// [F (9) | 0 (3)] (mixed) or [F (9) | 0] (FP64).
template <class R, bool DIR> constexpr int red_stride()
{
    return std::is_same<R, float>::value ? (DIR ? 24 : 12) : (DIR ? 20 : 10);
}

// N values (a multiple of 16 B) moved as 16 B vector loads and stores
template <class R, int N>
__device__ __forceinline__ void store_vec(R *dst, const R *v)
{
    static_assert(N * sizeof(R) % 16 == 0, "whole 16 B pieces");
    if constexpr (std::is_same<R, float>::value) {
#pragma unroll
        for (int k = 0; k < N / 4; ++k)
            reinterpret_cast<float4 *>(dst)[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < N / 2; ++k) reinterpret_cast<double2 *>(dst)[k] = make_double2(v[2 * k], v[2 * k + 1]);
    }
}
// six values (Voigt stress or strain) in three pieces of two
template <class R>
__device__ __forceinline__ void load6(const R *src, R *v)
{
    using V2 = typename std::conditional<std::is_same<R, float>::value, float2, double2>::type;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const V2 a = reinterpret_cast<const V2 *>(src)[k];
        v[2 * k] = a.x; v[2 * k + 1] = a.y;
    }
}
template <class R>
__device__ __forceinline__ void store6(R *dst, const R *v)
{
    using V2 = typename std::conditional<std::is_same<R, float>::value, float2, double2>::type;
#pragma unroll
    for (int k = 0; k < 3; ++k) reinterpret_cast<V2 *>(dst)[k] = V2{v[2 * k], v[2 * k + 1]};
}

template <class R, int N>
__device__ __forceinline__ void load_vec(const R *src, R *v)
{
    static_assert(N * sizeof(R) % 16 == 0, "whole 16 B pieces");
    if constexpr (std::is_same<R, float>::value) {
#pragma unroll
        for (int k = 0; k < N / 4; ++k) {
            const float4 a = reinterpret_cast<const float4 *>(src)[k];
            v[4 * k] = a.x; v[4 * k + 1] = a.y; v[4 * k + 2] = a.z; v[4 * k + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < N / 2; ++k) {
            const double2 a = reinterpret_cast<const double2 *>(src)[k];
            v[2 * k] = a.x; v[2 * k + 1] = a.y;
        }
    }
}

// deformation-gradient handling of the gather/node pair of kernels
enum FMode { F_NONE = 0, F_DIRECT = 1, F_INCR = 2 };

// ---------------------------------------------------------------------------
// helpers

// Sum over aligned groups of W lanes (W <= 16: groups inside one DPP row).
// Partners per step: lane ^ 1, lane ^ 2 (quad_perm), then the mirrored lane in
// the half row and in the row.  After the first two steps all lanes of a quad
// hold the same value, after the third all lanes of a half row, so every step
// adds the same two values in every lane and all lanes end with the same bits.
namespace dpp {
constexpr int kXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kHalfMirror = 0x141; // row_half_mirror
constexpr int kMirror = 0x140;     // row_mirror
template <int CTRL>
__device__ __forceinline__ float mov(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double mov(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
}  // namespace dpp

template <int W, class T>
__device__ __forceinline__ T group_sum(T v)
{
    static_assert(W == 1 || W == 2 || W == 4 || W == 8 || W == 16, "group of at most one DPP row");
    if (W >= 2) v += dpp::mov<dpp::kXor1>(v);
    if (W >= 4) v += dpp::mov<dpp::kXor2>(v);
    if (W >= 8) v += dpp::mov<dpp::kHalfMirror>(v);
    if (W >= 16) v += dpp::mov<dpp::kMirror>(v);
    return v;
}

template <int W>
__device__ __forceinline__ double group_min(double v)
{
    static_assert(W == 1 || W == 2 || W == 4 || W == 8 || W == 16, "group of at most one DPP row");
    if (W >= 2) v = fmin(v, dpp::mov<dpp::kXor1>(v));
    if (W >= 4) v = fmin(v, dpp::mov<dpp::kXor2>(v));
    if (W >= 8) v = fmin(v, dpp::mov<dpp::kHalfMirror>(v));
    if (W >= 16) v = fmin(v, dpp::mov<dpp::kMirror>(v));
    return v;
}

// (XJ-XI)**2 + (YJ-YI)**2 + (ZJ-ZI)**2 as FINT_PD.F90:724 evaluates it: unfused, left to right
__device__ __forceinline__ double sq_len_unfused(double a, double b, double c)
{
#pragma clang fp contract(off)
    return a * a + b * b + c * c;
}

// block of kBlock threads -> thread 0 gets the sum
__device__ __forceinline__ double block_sum(double v, double *red)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlock / 64; ++k) s += red[k];
    __syncthreads();
    return s;
}

// Peridynamic bond vectors in mixed precision: reference positions relative
// to a block's first particle X_b in fixed point, integer units of the
// block's quantum q_b (pd_qb, chosen so that every offset within the block's
// source sets stays below 2^29 units); the difference of two is exact, and
// the bond vector X_j - X_i = q_b (n_j - n_i) is rounded once, converting it
// to FP32.  Each bond record belongs to one block (that of its row's
// particle), and q_b is in its weight.
__device__ __forceinline__ int pd_fix(double x, double xb, double iq) { return (int)rint((x - xb) * iq); }

// Blocks are handed to the XCDs round-robin; this maps them so that XCD x
// processes the contiguous range [x*per, (x+1)*per) of logical blocks.  The
// grid is padded to a multiple of kXcd; logical blocks beyond the real count
// exit at once.
__device__ __forceinline__ int xcd_block()
{
    const int per = gridDim.x / kXcd;
    return (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
}

// ---------------------------------------------------------------------------
// predictor (MAIN.F90:353-355): d = dt v + dt^2/2 a; dtot += d; v += dt/2 a

__device__ __forceinline__ double step_dt(const Params &p) { return p.dt_dev ? *p.dt_dev : p.dt; }

// the predictor's increment DLT*V + DLT**2*0.5*A and half-step velocity
// V + DLT*0.5*A, unfused as the reference writes them: the standalone
// predictor (the first step of a call) and the one fused into the force
// kernel then give the same bits
__device__ __forceinline__ double predicted_d(double dt, double v, double a)
{
#pragma clang fp contract(off)
    return dt * v + dt * dt * 0.5 * a;
}
__device__ __forceinline__ double predicted_v(double dt, double v, double a)
{
#pragma clang fp contract(off)
    return v + dt * 0.5 * a;
}

__device__ __forceinline__ void predict_dof(const Params &p, int64_t q)
{
    const double dt = step_dt(p);
    const double v = p.v[q], a = p.a[q];
    const double d = predicted_d(dt, v, a);
    const double t = p.dtot[q] + d;
    p.d[q] = d;
    p.dtot[q] = t;
    p.v[q] = predicted_v(dt, v, a);
    if (p.pd) {   // peridynamics: physical = generalized, so x moves with d
        const double u = p.uphy[q] + d;
        p.uphy[q] = u;
        p.x[q] = p.X0[q] + u;
    }
    if (p.d4) {   // t4 is only read by F_DIRECT launches, which sync_d4 precedes
        const int64_t i = q / 3, k = q - 3 * i;
        reinterpret_cast<float *>(p.d4)[4 * i + k] = (float)d;
    }
}

__global__ void k_predict(Params p)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * (int64_t)p.n) return;
    predict_dof(p, q);
}

__global__ void k_cast4(int n, const double *a, float4 *b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = make_float4((float)a[3 * i], (float)a[3 * i + 1], (float)a[3 * i + 2], 0.f);
}

template <class R>
__global__ void k_widen(int64_t m, const R *a, double *b)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m) b[q] = (double)a[q];
}

template <class R>
__global__ void k_narrow(int64_t m, const double *a, R *b)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m) b[q] = (R)a[q];
}

// GSTRAIN_EQ from the stored strain (FINT.F90:831-834)
template <class R>
__global__ void k_strain_eq(int n, const R *strain, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double e[6];
    for (int k = 0; k < 6; ++k) e[k] = (double)strain[6 * i + k];
    const double q = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + 2.0 * (e[3] * e[3] + e[4] * e[4] + e[5] * e[5]);
    out[i] = sqrt(q * 2.0 / 3.0);
}

// ---------------------------------------------------------------------------
// DO_INTERP (HANDELER.F90:114-133): out_i = sum_j phi_ij coef_j.  UPDATE: the
// increment d is interpolated into the physical total displacement and the
// coordinates (MAIN.F90:377-386).

template <class R, bool UPDATE>
__global__ __launch_bounds__(kBlock) void k_interp(Params p, const double *coef, double *out)
{
    const int node = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup;
    const int lane = threadIdx.x & (kGroup - 1);
    if (node >= p.n) return;
    const int p0 = p.row_ptr[node], p1 = p.row_ptr[node + 1];
    const double *src = UPDATE ? p.d : coef;
    const R *shp = as<R>(p.shp);
    double u0 = 0, u1 = 0, u2 = 0;
    // each lane issues the loads of up to U entries before using any of them;
    // the sums still run over its entries in list order
    constexpr int U = kUnrollF64;
    for (int base = p0 + lane; base < p1; base += kGroup * U) {
        int j[U];
        double f[U], s[U][3];
#pragma unroll
        for (int e = 0; e < U; ++e) {
            const int q = base + kGroup * e;
            j[e] = q < p1 ? p.col[q] : node;
            f[e] = q < p1 ? (double)shp[q] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < U; ++e) {
            s[e][0] = src[3 * j[e]]; s[e][1] = src[3 * j[e] + 1]; s[e][2] = src[3 * j[e] + 2];
        }
#pragma unroll
        for (int e = 0; e < U; ++e)
            if (base + kGroup * e < p1) { u0 += f[e] * s[e][0]; u1 += f[e] * s[e][1]; u2 += f[e] * s[e][2]; }
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
// node accumulate the symmetric moment matrix M and its derivatives in FP64,
// then every lane forms M^-1 and evaluates phi, grad phi for its neighbors,
// written in list order and, through tpos, in transposed order.

struct KernelEval {
    double phi, px, py, pz;
};

// window of neighbor j (position xj, essential-BC flag ebc) at x_i; w holds
// its windows a_j, or with RCP and the box support their reciprocals 1/a_j
// (rk_windows<RCP>), so that the box kernel divides nowhere.  RCP is taken in
// mixed precision only: FP64 keeps the reference's divisions, operation for
// operation (products with 1/a differ in the last bit).
template <bool RCP>
__device__ __forceinline__ const double *rk_windows(const Params &p) { return (RCP && !p.shsup) ? p.winv : p.win; }

template <bool RCP>
__device__ __forceinline__ KernelEval rk_window_at(const Params &p, const double *xi, const double *xj,
                                                   const double *w, int ebc)
{
    const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
    KernelEval e;
    if (p.shsup) {
        const double w1 = w[0], w2 = w[1], w3 = w[2];
        const double zx = dx / w1, zy = dy / w2, zz = dz / w3;
        const double dia = sqrt(zx * zx + zy * zy + zz * zz);
        double k0, k1;
        mega::cubic_kernel(dia, k0, k1);
        double rx = 0, ry = 0, rz = 0;
        if (!(dia <= 1.0e-13)) { rx = dx / (w1 * w1) / dia; ry = dy / (w2 * w2) / dia; rz = dz / (w3 * w3) / dia; }
        e.phi = k0; e.px = k1 * rx; e.py = k1 * ry; e.pz = k1 * rz;
    } else {
        const double w1 = w[0], w2 = w[1], w3 = w[2];
        const double zx = RCP ? dx * w1 : dx / w1, zy = RCP ? dy * w2 : dy / w2, zz = RCP ? dz * w3 : dz / w3;
        double fx, fxx, fy, fyy, fz, fzz;
        mega::cubic_kernel(fabs(zx), fx, fxx);
        mega::cubic_kernel(fabs(zy), fy, fyy);
        mega::cubic_kernel(fabs(zz), fz, fzz);
        const double iw = RCP ? w1 : 1.0 / w1;   // RK.f90:183-205 divides all three by the x window
        const double rx = zx == 0.0 ? 0.0 : (zx >= 0.0 ? iw : -iw);
        const double ry = zy == 0.0 ? 0.0 : (zy >= 0.0 ? iw : -iw);
        const double rz = zz == 0.0 ? 0.0 : (zz >= 0.0 ? iw : -iw);
        double ph = fx * fy * fz;
        double px = fxx * fy * fz * rx, py = fx * fyy * fz * ry, pz = fx * fy * fzz * rz;
        if (ebc) {   // singular kernel at essential-BC nodes (RK.f90:217-228)
            const double wm = 1.0 / (sqrt(zx * zx + zy * zy + zz * zz) + mega::kSingularEps);
            const double pw3 = ph * wm * wm * wm;
            px = px * wm - (RCP ? pw3 * zx * w1 : pw3 * zx / w1);
            py = py * wm - (RCP ? pw3 * zy * w2 : pw3 * zy / w2);
            pz = pz * wm - (RCP ? pw3 * zz * w3 : pw3 * zz / w3);
            ph = ph * wm;
        }
        e.phi = ph; e.px = px; e.py = py; e.pz = pz;
    }
    return e;
}

template <bool RCP>
__device__ __forceinline__ KernelEval rk_window(const Params &p, const double *xi, const double *xs, int j)
{
    const double xj[3] = {xs[3 * j], xs[3 * j + 1], xs[3 * j + 2]};
    const double *wp = rk_windows<RCP>(p);
    const double w[3] = {wp[3 * j], wp[3 * j + 1], wp[3 * j + 2]};
    return rk_window_at<RCP>(p, xi, xj, w, p.ebc_node[j]);
}

// one neighbor's terms of the moment matrix M (m[0..9]) and its derivatives
// M_a (m[10a..10a+9]); H = (1, x_i - x_j), so m[0] is also sum phi.
//   M_a(r,c) += (dH_r H_c + H_r dH_c) phi + H_r H_c phi_a, dH the unit vector e_a;
// the product terms of dH vanish except on row or column a and are left out
// there (their zero addends do not change the sum).
__device__ __forceinline__ void rk_moments(double *m, const double *H, const KernelEval &e)
{
    int t = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = r; c < 4; ++c, ++t) {
            const double hh = H[r] * H[c];
            m[t] += hh * e.phi;
            const double pa[3] = {e.px, e.py, e.pz};
#pragma unroll
            for (int a = 1; a <= 3; ++a) {
                double d;
                if (r == a && c == a) d = (H[c] * e.phi + H[r] * e.phi) + hh * pa[a - 1];
                else if (r == a) d = H[c] * e.phi + hh * pa[a - 1];
                else if (c == a) d = H[r] * e.phi + hh * pa[a - 1];
                else d = hh * pa[a - 1];
                m[10 * a + t] += d;
            }
        }
}

// Stabilized non-conforming nodal integration (ITYPE_INT = 1, Lagrangian): the
// gradient at node L is replaced by its divergence over L's smoothing box,
//   grad~ phi_J(x_L) = sum_k e_k (phi_J(x_L + l_k e_k) - phi_J(x_L - l_k e_k)) A_k / V,
// with RK1 evaluated at the six face centres over L's own neighbor list, as
// the reference's smoothing block does (FINT.F90:441-496, commented out
// there; SM_LEN, SM_AREA, SM_VOL of PRE_MODEL_SMOOTHING_LENGTHS.F90:49-90).
// The values keep linear fields exact (each face value reproduces x), so the
// smoothed gradient does too.  Run once, after k_rk has written the node's own
// phi and the lists' transposed slots: it overwrites the list-order and
// transposed gradients.  FP64 throughout, with the reference's divisions.
template <class R>
__global__ __launch_bounds__(kBlock) void k_snni(Params p, const double *xs, const double *sml, const double *aov,
                                                 int nblocks)
{
    constexpr int G = kRkGroup;
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int node = blk * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x & (G - 1);
    if (node >= p.n) return;   // whole groups leave together
    const int p0 = p.row_ptr[node], p1 = p.row_ptr[node + 1];
    const double xi[3] = {xs[3 * (int64_t)node], xs[3 * (int64_t)node + 1], xs[3 * (int64_t)node + 2]};
    R *dshp = as<R>(p.dshp);
    TPair<R> *tpair = as<TPair<R>>(p.tpair);
    // quasi-linear augmentation: its sums over the six points x +- L e_k do not depend on x
    double ms[10] = {0}, hs[4] = {0, 0, 0, 0};
    if (p.ql) {
        const double L = p.ql_len;
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            double d[3] = {0.0, 0.0, 0.0};
            d[s / 2] = (s & 1) ? L : -L;   // x - q_s
            const double H[4] = {1.0, d[0], d[1], d[2]};
            int t = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = r; c < 4; ++c, ++t) ms[t] += H[r] * H[c];
#pragma unroll
            for (int k = 0; k < 4; ++k) hs[k] += H[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        const double l = sml[3 * (int64_t)node + k], a = aov[3 * (int64_t)node + k];
        double B[2][4];   // e1^T M^-1 at x_L + l e_k and x_L - l e_k
        for (int sg = 0; sg < 2; ++sg) {
            double pt[3] = {xi[0], xi[1], xi[2]};
            pt[k] = sg ? xi[k] - l : xi[k] + l;
            double m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int q = p0 + lane; q < p1; q += G) {
                const int j = p.col[q];
                const KernelEval e = rk_window<false>(p, pt, xs, j);
                const double H[4] = {1.0, pt[0] - xs[3 * (int64_t)j], pt[1] - xs[3 * (int64_t)j + 1],
                                     pt[2] - xs[3 * (int64_t)j + 2]};
                int t = 0;
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = r; c < 4; ++c, ++t) m[t] += H[r] * H[c] * e.phi;
            }
#pragma unroll
            for (int t = 0; t < 10; ++t) m[t] = group_sum<G>(m[t]);
            if (p.ql) {
                const double sphi = m[0];
#pragma unroll
                for (int t = 0; t < 10; ++t) m[t] = m[t] + ms[t] * sphi * p.ql_coef;
            }
            double Mi[16];
            mega::m44inv_sym(m, Mi);
            B[sg][0] = Mi[0]; B[sg][1] = Mi[4]; B[sg][2] = Mi[8]; B[sg][3] = Mi[12];
        }
        for (int q = p0 + lane; q < p1; q += G) {
            const int j = p.col[q];
            double f[2];
#pragma unroll
            for (int sg = 0; sg < 2; ++sg) {
                double pt[3] = {xi[0], xi[1], xi[2]};
                pt[k] = sg ? xi[k] - l : xi[k] + l;
                const KernelEval e = rk_window<false>(p, pt, xs, j);
                double H[4] = {1.0, pt[0] - xs[3 * (int64_t)j], pt[1] - xs[3 * (int64_t)j + 1],
                               pt[2] - xs[3 * (int64_t)j + 2]};
                if (p.ql)
#pragma unroll
                    for (int c = 0; c < 4; ++c) H[c] = H[c] + hs[c] * p.ql_coef;
                f[sg] = (B[sg][0] * H[0] + B[sg][1] * H[1] + B[sg][2] * H[2] + B[sg][3] * H[3]) * e.phi;
            }
            const R gk = (R)((f[0] - f[1]) * a);
            dshp[(int64_t)k * p.cap + q] = gk;
            if (p.tpos) {
                TPair<R> *t = tpair + p.tpos[q];
                if (k == 0) t->gx = gk;
                else if (k == 1) t->gy = gk;
                else t->gz = gk;
            }
        }
    }
}

// Nodes are taken in bin order (order = the bins' node lists), blocks mapped to
// XCDs in contiguous ranges, kRkGroup lanes per node.  A lane's first kRkCache
// neighbors are loaded together, and their kernel values and H are kept in
// registers for the second pass; longer rows continue in a plain loop.
// GATHER (lists rebuilt every step): the gather sums of CONSTRUCT_FINT without
// the deformation gradient -- L0 = sum grad phi (x) d, mixed over d_j - d_i
// with S = sum grad phi -- are taken here from the gradients as they are
// formed, and the list-order gradients are not stored.  Each lane visits its
// entries in the order k_gather's lanes do over the same list (kGroup ==
// kRkGroup lanes per node, entry q on lane q mod 16), and the group sums are
// the same: the record is k_gather's, in mixed precision bit for bit (in FP64
// the gradients' own products may be contracted differently, a last-bit
// difference).  Two waves per SIMD are kept by bound: the fused sums' few
// extra registers would otherwise move the kernel to one (4-8 VGPRs spill).
template <class R, bool GATHER>
#ifndef MEGA_RK_WAVES
#define MEGA_RK_WAVES 2
#endif
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MEGA_RK_WAVES, MEGA_RK_WAVES))) void k_rk(Params p, const double *xs, const int32_t *order, int nblocks)
{
    static_assert(!GATHER || kRkGroup == kGroup, "the fused gather visits entries as k_gather does");
    constexpr bool MIX = std::is_same<R, float>::value;
    constexpr int G = kRkGroup, U = kRkCache;
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int g = blk * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x & (G - 1);
    if (g >= p.n) return;
    const int node = order ? order[g] : g;
    const int p0 = p.row_ptr[node], p1 = p.row_ptr[node + 1];
    const double xi[3] = {xs[3 * node], xs[3 * node + 1], xs[3 * node + 2]};
    double m[40];   // 10 unique entries each of M, Mx, My, Mz
#pragma unroll
    for (int k = 0; k < 40; ++k) m[k] = 0.0;
    int jv[U], tp[U];
    double c_ph[U], c_px[U], c_py[U], c_pz[U], c_h[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int q = p0 + lane + u * G;
        jv[u] = q < p1 ? p.col[q] : -1;
        tp[u] = (q < p1 && p.tpos) ? p.tpos[q] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int jj = jv[u] < 0 ? node : jv[u];
        const double xj[3] = {xs[3 * jj], xs[3 * jj + 1], xs[3 * jj + 2]};
        const double *wp = rk_windows<MIX>(p);
        const double w[3] = {wp[3 * jj], wp[3 * jj + 1], wp[3 * jj + 2]};
        const KernelEval e = rk_window_at<MIX>(p, xi, xj, w, p.ebc_node[jj]);
        c_ph[u] = e.phi; c_px[u] = e.px; c_py[u] = e.py; c_pz[u] = e.pz;
        c_h[u][0] = xi[0] - xj[0]; c_h[u][1] = xi[1] - xj[1]; c_h[u][2] = xi[2] - xj[2];
        if (jv[u] >= 0) {
            const double H[4] = {1.0, c_h[u][0], c_h[u][1], c_h[u][2]};
            rk_moments(m, H, e);
        }
    }
    for (int q = p0 + lane + U * G; q < p1; q += G) {
        const int j = p.col[q];
        const KernelEval e = rk_window<MIX>(p, xi, xs, j);
        const double H[4] = {1.0, xi[0] - xs[3 * j], xi[1] - xs[3 * j + 1], xi[2] - xs[3 * j + 2]};
        rk_moments(m, H, e);
    }
#pragma unroll
    for (int k = 0; k < 40; ++k) m[k] = group_sum<kRkGroup>(m[k]);
    double hs[4] = {0, 0, 0, 0};
    if (p.ql) {   // quasi-linear augmentation (RK.f90:251-303, 357-377)
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
        const double sphi = m[0];   // sum of phi: H_0 = 1
#pragma unroll
        for (int t = 0; t < 10; ++t) m[t] = m[t] + ms[t] * sphi * p.ql_coef;
    }
    double Mi[16];
    mega::m44inv_sym(m, Mi);
    const double B[4] = {Mi[0], Mi[4], Mi[8], Mi[12]};   // e1^T M^-1
    double Bd[3][4];                                     // -(B M_a) M^-1
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
    R *shp = as<R>(p.shp), *dshp = as<R>(p.dshp), *tshp = as<R>(p.tshp);
    TPair<R> *tpair = as<TPair<R>>(p.tpair);
    R L[9], S[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) L[k] = R(0);
    S[0] = S[1] = S[2] = R(0);
    float4 di = make_float4(0.f, 0.f, 0.f, 0.f);
    if (GATHER && MIX) di = p.d4[node];
    // phi_j = B.H phi_w, grad phi_j = (Bd_a.H + B_a) phi_w + (B.H) grad phi_w
    auto emit = [&](int q, int t, int j, const double *h, const KernelEval &e) {
        double H[4] = {1.0, h[0], h[1], h[2]};
        if (p.ql)
#pragma unroll
            for (int k = 0; k < 4; ++k) H[k] = H[k] + hs[k] * p.ql_coef;
        const double c0 = B[0] * H[0] + B[1] * H[1] + B[2] * H[2] + B[3] * H[3];
        const double cx = Bd[0][0] * H[0] + Bd[0][1] * H[1] + Bd[0][2] * H[2] + Bd[0][3] * H[3] + B[1];
        const double cy = Bd[1][0] * H[0] + Bd[1][1] * H[1] + Bd[1][2] * H[2] + Bd[1][3] * H[3] + B[2];
        const double cz = Bd[2][0] * H[0] + Bd[2][1] * H[1] + Bd[2][2] * H[2] + Bd[2][3] * H[3] + B[3];
        const R f = (R)(c0 * e.phi);
        const R gx = (R)(cx * e.phi + c0 * e.px), gy = (R)(cy * e.phi + c0 * e.py), gz = (R)(cz * e.phi + c0 * e.pz);
        shp[q] = f;
        if constexpr (GATHER) {   // k_gather's sums, same operands and order
            R dj[3];
            if constexpr (MIX) {
                const float4 dv = p.d4[j];
                dj[0] = dv.x - di.x; dj[1] = dv.y - di.y; dj[2] = dv.z - di.z;
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) dj[k] = p.d[3 * j + k];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) { L[k] += gx * dj[k]; L[3 + k] += gy * dj[k]; L[6 + k] += gz * dj[k]; }
            if (MIX) { S[0] += gx; S[1] += gy; S[2] += gz; }
        } else {
            dshp[q] = gx; dshp[p.cap + q] = gy; dshp[2 * p.cap + q] = gz;
        }
        if (p.tpos) {
            TPair<R> v;
            v.gx = gx; v.gy = gy; v.gz = gz; v.src = node;
            if (sizeof(TPair<R>) == 32) reinterpret_cast<int32_t *>(&v)[7] = 0;
            tpair[t] = v;
            if (tshp) tshp[t] = f;
        }
    };
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (jv[u] >= 0) {
            KernelEval e;
            e.phi = c_ph[u]; e.px = c_px[u]; e.py = c_py[u]; e.pz = c_pz[u];
            emit(p0 + lane + u * G, tp[u], jv[u], c_h[u], e);
        }
    for (int q = p0 + lane + U * G; q < p1; q += G) {
        const int j = p.col[q];
        const KernelEval e = rk_window<MIX>(p, xi, xs, j);
        const double h[3] = {xi[0] - xs[3 * j], xi[1] - xs[3 * j + 1], xi[2] - xs[3 * j + 2]};
        emit(q, p.tpos ? p.tpos[q] : 0, j, h, e);
    }
    if constexpr (GATHER) {
#pragma unroll
        for (int k = 0; k < 9; ++k) L[k] = group_sum<G>(L[k]);
        if (MIX) {
#pragma unroll
            for (int k = 0; k < 3; ++k) S[k] = group_sum<G>(S[k]);
        }
        constexpr int W = red_stride<R, false>();
        if (lane == 0) {
            R rec[W];
#pragma unroll
            for (int k = 0; k < W; ++k) rec[k] = R(0);
#pragma unroll
            for (int k = 0; k < 9; ++k) rec[k] = L[k];
            if constexpr (MIX) { rec[9] = S[0]; rec[10] = S[1]; rec[11] = S[2]; }
            store_vec<R, W>(as<R>(p.red) + (int64_t)W * node, rec);
        }
    }
}

// ---------------------------------------------------------------------------
// gather phase of CONSTRUCT_FINT, kGroup lanes per node.  Sums per node:
//   L0(k,l) = sum_j g_l d_jk                        (FINT.F90:673-679)
//   F(k,l)  = sum_j g_l dtot_jk  (F_DIRECT only)    (FINT.F90:363-376)
// With F_INCR the deformation gradient is carried per node instead,
// F_n = F_{n-1} + L0_n (dtot_n = dtot_{n-1} + d_n with the same fixed
// gradients), so dtot is not gathered.  Mixed: the FP32 sums are taken over
// differences to the node's own value plus S = sum g, and k_node adds
// S (x) own value back; this keeps the rigid translation out of the FP32 sums.
// red per node: FP64 [L0 | F], mixed [L0 | S | F].  UPDATE_X: DO_INTERP is
// fused, u = sum phi d updates the physical displacement and coordinates.
// Each lane issues the loads of up to U list entries before using any of them.

template <class R, int FM, bool UPDATE_X, bool SETS>
__global__ __launch_bounds__(kBlock) void k_gather(Params p, int nblocks)
{
    constexpr bool MIX = std::is_same<R, float>::value;
    constexpr bool DIR = FM == F_DIRECT;
    constexpr int U = SETS ? (MIX ? kUnrollSets : kUnrollSetsF64) : (MIX ? kUnrollMix : kUnrollF64);
    constexpr int GG = gather_group(SETS);
    extern __shared__ __align__(16) unsigned char mega_dyn[];
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int node = blk * gather_npb(SETS) + threadIdx.x / GG;
    const int lane = threadIdx.x & (GG - 1);
    // lists that stay: the block's distinct neighbors' d (and dtot) in LDS,
    // mixed as the FP32 copies (16 B), FP64 as {x, y}, {z, 0} (32 B)
    int m = 0;
    float4 *Ld = reinterpret_cast<float4 *>(mega_dyn);
    double2 *Ld2 = reinterpret_cast<double2 *>(mega_dyn);
    if (SETS) {
        const int s0 = p.gset_ptr[blk];
        m = p.gset_ptr[blk + 1] - s0;
        for (int k = threadIdx.x; k < m; k += kBlock) {
            const int j = p.gset[s0 + k];
            if constexpr (MIX) {
                Ld[k] = p.d4[j];
                if (DIR) Ld[m + k] = p.t4[j];
            } else {
                Ld2[2 * k] = make_double2(p.d[3 * j], p.d[3 * j + 1]);
                Ld2[2 * k + 1] = make_double2(p.d[3 * j + 2], 0.0);
                if (DIR) {
                    Ld2[2 * (m + k)] = make_double2(p.dtot[3 * j], p.dtot[3 * j + 1]);
                    Ld2[2 * (m + k) + 1] = make_double2(p.dtot[3 * j + 2], 0.0);
                }
            }
        }
        if (m > 0) __syncthreads();
    }
    if (node >= p.n) return;
    const int p0 = p.row_ptr[node], p1 = p.row_ptr[node + 1];
    const R *__restrict__ shp = as<R>(p.shp);
    const R *__restrict__ gx = as<R>(p.dshp);
    const R *__restrict__ gy = gx + p.cap;
    const R *__restrict__ gz = gy + p.cap;
    R L[9], F[9], S[3], u[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) { L[k] = R(0); F[k] = R(0); }
    S[0] = S[1] = S[2] = R(0);
    u[0] = u[1] = u[2] = R(0);
    float4 di = make_float4(0.f, 0.f, 0.f, 0.f), ti = di;
    if (MIX) {
        di = p.d4[node];
        if (DIR) ti = p.t4[node];
    }
    // LDS: neighbors by their position in the block's set
    auto sweep = [&](auto lds_tag) {
    constexpr bool LDS = decltype(lds_tag)::value;
    for (int base = p0 + lane; base < p1; base += GG * U) {
        int j[U];
#pragma unroll
        for (int e = 0; e < U; ++e) {
            const int q = base + GG * e;
            if (LDS) j[e] = q < p1 ? (int)p.gloc[q] : 0;
            else j[e] = q < p1 ? p.col[q] : node;
        }
        R g0[U], g1[U], g2[U], f[U], dj[U][3], tj[U][3];
#pragma unroll
        for (int e = 0; e < U; ++e) {
            const int q = base + GG * e;
            const bool ok = q < p1;
            if (LDS) {   // lists that stay: the packed record
                if constexpr (MIX) {
                    const float4 r = ok ? p.fpair[q] : make_float4(0.f, 0.f, 0.f, 0.f);
                    g0[e] = (R)r.x; g1[e] = (R)r.y; g2[e] = (R)r.z;
                    f[e] = UPDATE_X ? (R)r.w : R(0);
                } else {
                    const double2 a = ok ? p.fpair2[2 * (int64_t)q] : make_double2(0.0, 0.0);
                    const double2 b = ok ? p.fpair2[2 * (int64_t)q + 1] : make_double2(0.0, 0.0);
                    g0[e] = a.x; g1[e] = a.y; g2[e] = b.x;
                    f[e] = UPDATE_X ? b.y : R(0);
                }
            } else {
                g0[e] = ok ? gx[q] : R(0);
                g1[e] = ok ? gy[q] : R(0);
                g2[e] = ok ? gz[q] : R(0);
                f[e] = (UPDATE_X && ok) ? shp[q] : R(0);
            }
            if (MIX) {
                const float4 dv = LDS ? Ld[j[e]] : p.d4[j[e]];
                dj[e][0] = dv.x - di.x; dj[e][1] = dv.y - di.y; dj[e][2] = dv.z - di.z;
                if (DIR) {
                    const float4 tv = LDS ? Ld[m + j[e]] : p.t4[j[e]];
                    tj[e][0] = tv.x - ti.x; tj[e][1] = tv.y - ti.y; tj[e][2] = tv.z - ti.z;
                }
            } else if (LDS) {
                const double2 a = Ld2[2 * j[e]], b = Ld2[2 * j[e] + 1];
                dj[e][0] = a.x; dj[e][1] = a.y; dj[e][2] = b.x;
                if (DIR) {
                    const double2 c = Ld2[2 * (m + j[e])], d = Ld2[2 * (m + j[e]) + 1];
                    tj[e][0] = c.x; tj[e][1] = c.y; tj[e][2] = d.x;
                }
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    dj[e][k] = (R)p.d[3 * j[e] + k];
                    if (DIR) tj[e][k] = (R)p.dtot[3 * j[e] + k];
                }
            }
        }
#pragma unroll
        for (int e = 0; e < U; ++e) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                L[k] += g0[e] * dj[e][k]; L[3 + k] += g1[e] * dj[e][k]; L[6 + k] += g2[e] * dj[e][k];
                if (DIR) { F[k] += g0[e] * tj[e][k]; F[3 + k] += g1[e] * tj[e][k]; F[6 + k] += g2[e] * tj[e][k]; }
                if (UPDATE_X) u[k] += f[e] * dj[e][k];
            }
            if (MIX) { S[0] += g0[e]; S[1] += g1[e]; S[2] += g2[e]; }
        }
    }
    };
    if (m > 0) sweep(std::true_type{});
    else sweep(std::false_type{});
#pragma unroll
    for (int k = 0; k < 9; ++k) L[k] = group_sum<GG>(L[k]);
    if (DIR) {
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = group_sum<GG>(F[k]);
    }
    if (MIX) {
#pragma unroll
        for (int k = 0; k < 3; ++k) S[k] = group_sum<GG>(S[k]);
    }
    if (UPDATE_X) {
#pragma unroll
        for (int k = 0; k < 3; ++k) u[k] = group_sum<GG>(u[k]);
    }
    constexpr int W = red_stride<R, DIR>();
    constexpr int FO = MIX ? 12 : 10;
    if (lane == 0) {
        R rec[W];
#pragma unroll
        for (int k = 0; k < W; ++k) rec[k] = R(0);
#pragma unroll
        for (int k = 0; k < 9; ++k) rec[k] = L[k];
        if constexpr (MIX) { rec[9] = S[0]; rec[10] = S[1]; rec[11] = S[2]; }
        if constexpr (DIR) {
#pragma unroll
            for (int k = 0; k < 9; ++k) rec[FO + k] = F[k];
        }
        store_vec<R, W>(as<R>(p.red) + (int64_t)W * node, rec);
    }
    if (UPDATE_X && lane < 3) {
        const int qd = 3 * node + lane;
        const R uk = lane == 0 ? u[0] : (lane == 1 ? u[1] : u[2]);
        // mixed: u = d_i + sum phi (d_j - d_i), with sum phi = 1 by RK consistency
        const double du = MIX ? p.d[qd] + (double)uk : (double)uk;
        const double t = p.uphy[qd] + du;
        p.uphy[qd] = t;
        if (p.out_x) p.x[qd] = p.X0[qd] + t;
    }
}

// ---------------------------------------------------------------------------
// learned return map (mixed precision) and tuple harvest, called by k_node with
// every lane of the wave active

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lanes_below(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// the packed weights (staged once per block that has a candidate) and, per
// wave, the normalised inputs of its candidate lanes, then the outputs,
// feature-major ([feature][slot]: consecutive slots on consecutive banks)
struct SurrLds {
    mlp::bf16x8 tiles[mlp::kTilesFwd * 64];
    float bias[mlp::kBiasPack];
    float norm[mlp::kNorm];
    float stage[kBlock / 64][16][64];
};

// cand: this lane's J2 trial state is plastic and belongs to the trained
// material.  Inside the training envelope the network gives the plastic
// correction; the answer is kept when the residual of the return-map
// equations (mega::return_residual) is within tol and eps does not decrease,
// otherwise the caller runs the exact return map.  A block with a lane inside
// the envelope stages the packed weights (22 KB) in LDS; its waves then run
// the network on their own candidates.
// cnt: this wave's slot of the block's counters (lane 0 writes all four)
__device__ bool surr_return(const Params &p, bool cand, const float *der, const float *sr, const float *pred,
                            float &ep, float *s_out, const Material &mt, unsigned *cnt)
{
    __shared__ SurrLds L;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float x[mlp::kIn];
#pragma unroll
    for (int k = 0; k < 6; ++k) { x[k] = der[k]; x[6 + k] = sr[k]; }
    x[12] = ep;
    bool inenv = cand;
    if (cand)
#pragma unroll
        for (int k = 0; k < mlp::kIn; ++k)
            inenv = inenv && x[k] >= p.surr_env[k] && x[k] <= p.surr_env[mlp::kIn + k];
    const uint64_t mc = __ballot(cand), mu = __ballot(inenv);
    if (lane == 0) {
        cnt[0] = (unsigned)__popcll(mc);
        cnt[1] = (unsigned)(__popcll(mc) - __popcll(mu));
        cnt[2] = (unsigned)__popcll(mu);
        cnt[3] = 0u;
    }
    if (!__syncthreads_or(inenv)) return false;
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(p.surr_tiles);
        uint4 *dst = reinterpret_cast<uint4 *>(L.tiles);
        for (int k = threadIdx.x; k < mlp::kTilesFwd * 64; k += kBlock) dst[k] = src[k];
        for (int k = threadIdx.x; k < mlp::kBiasPack; k += kBlock) L.bias[k] = p.surr_bias[k];
        for (int k = threadIdx.x; k < mlp::kNorm; k += kBlock) L.norm[k] = p.surr_norm[k];
    }
    __syncthreads();
    const float *norm = L.norm;
    const int nc = __popcll(mu);
    const int slot = lanes_below(mu);
    float(*st)[64] = L.stage[wv];
    if (inenv) {
#pragma unroll
        for (int k = 0; k < mlp::kIn; ++k) st[k][slot] = (x[k] - norm[mlp::kInMean + k]) * norm[mlp::kInScale + k];
        st[13][slot] = st[14][slot] = st[15][slot] = 0.0f;
    }
    wave_sync();
    const int g = lane >> 4;
    for (int c0 = 0; c0 < nc; c0 += 16) {
        const int s = c0 + (lane & 15);
        mlp::bf16x8 x0 = {};
        if (g < 2 && s < nc)
            x0 = mlp::pack8(mlp::f32x4{st[8 * g][s], st[8 * g + 1][s], st[8 * g + 2][s], st[8 * g + 3][s]},
                            mlp::f32x4{st[8 * g + 4][s], st[8 * g + 5][s], st[8 * g + 6][s], st[8 * g + 7][s]});
        mlp::f32x4 z1[4], z2[4], z3[4], y;
        mlp::forward16(L.tiles, L.bias, x0, lane, z1, z2, z3, y);
        wave_sync();
        if (g < 2 && s < nc)
#pragma unroll
            for (int t = 0; t < 4; ++t) st[4 * g + t][s] = y[t];
    }
    wave_sync();
    bool ok = false;
    if (inenv) {
        const float *os = norm + mlp::kOutScale;
        float s[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] = pred[k] - os[k] * st[k][slot];
        const float e = ep + os[6] * st[6][slot];
        ok = e >= ep && mega::return_residual(pred, s, ep, e, mt) <= p.surr_tol;
        if (ok) {
#pragma unroll
            for (int k = 0; k < 6; ++k) s_out[k] = s[k];
            ep = e;
        }
    }
    const uint64_t mf = __ballot(inenv && !ok);
    if (lane == 0) cnt[3] = (unsigned)__popcll(mf);
    return ok;
}

// append (de[6], rotated stress[6], eqps_n, stress[6], eqps) of lanes with h set
template <class R>
__device__ void harvest_put(const Params &p, bool h, const R *der, const R *s_in, R ep_n, const R *s_out, R ep)
{
    const uint64_t m = __ballot(h);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(p.hv_count, (unsigned long long)__popcll(m));
    const uint32_t lo = __shfl((uint32_t)base, leader, 64), hi = __shfl((uint32_t)(base >> 32), leader, 64);
    base = ((unsigned long long)hi << 32) | lo;
    if (h) {
        const unsigned long long slot = base + (unsigned long long)lanes_below(m);
        if (slot < (unsigned long long)p.hv_cap) {
            float *o = p.hv + 20 * slot;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                o[k] = (float)der[k];
                o[6 + k] = (float)s_in[k];
                o[13 + k] = (float)s_out[k];
            }
            o[12] = (float)ep_n;
            o[19] = (float)ep;
        }
    }
}

// ---------------------------------------------------------------------------
// per-node update of CONSTRUCT_FINT (FINT.F90:353-391, 758-870), one lane per
// node: F and J = |det F|, push-forward by INV3 (adj F * det F, the
// reference's own INV3), L, Hughes-Winget rotation (FP64), rotated
// stress/strain, elastic trial and J2 return (FP64 or FP32; SURR: the learned
// return map first), state write-back, work, wave speed, and P = V J sigma F^-T
// for the force gather.  Lanes past the last node compute on a copy of it and
// store nothing, so that every lane reaches the wave-wide sections.

// One node's update from its gather sums red (red_stride values): kinematics,
// rotation, trial, return map, state stores and P = V J sigma F^-T (into P,
// when valid); returns the node's work increment.  cnt: the wave's counters
// of the learned return map.
template <class R, int FM, bool SURR>
__device__ __forceinline__ double node_core(const Params &p, int i, bool valid, const R *red, R *P, unsigned *cnt)
{
    constexpr bool MIX = std::is_same<R, float>::value;
    static_assert(!SURR || MIX, "the learned return map runs in mixed precision");
    constexpr int FO = MIX ? 12 : 10;
    double work = 0.0;
    double L0[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) L0[k] = (double)red[k];
    double S[3] = {0, 0, 0}, dI[3] = {0, 0, 0}, tI[3] = {0, 0, 0};
    if constexpr (MIX) {   // add back S (x) own value: L0(k,l) += S_l d_ik
        const float4 dv = p.d4[i];
        S[0] = (double)red[9]; S[1] = (double)red[10]; S[2] = (double)red[11];
        dI[0] = dv.x; dI[1] = dv.y; dI[2] = dv.z;
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int k = 0; k < 3; ++k) L0[k + 3 * l] += S[l] * dI[k];
    }
    const int mi = p.mat[i];
    const Material &mt = p.mats[mi];
    const double vol = p.vol[i];
    double det = 1.0, IF[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Lm[9];
    if (FM != F_NONE) {
        double Fm[9];
        if constexpr (FM == F_DIRECT) {
#pragma unroll
            for (int k = 0; k < 9; ++k) Fm[k] = (double)red[FO + k];
            if (MIX) {
                const float4 tv = p.t4[i];
                tI[0] = tv.x; tI[1] = tv.y; tI[2] = tv.z;
#pragma unroll
                for (int l = 0; l < 3; ++l)
#pragma unroll
                    for (int k = 0; k < 3; ++k) Fm[k + 3 * l] += S[l] * tI[k];
            }
            Fm[0] += 1.0; Fm[4] += 1.0; Fm[8] += 1.0;
        } else {
            {
                double Fo[10];
                load_vec<double, 10>(p.Fst + 10 * (int64_t)i, Fo);
#pragma unroll
                for (int k = 0; k < 9; ++k) Fm[k] = Fo[k] + L0[k];
            }
        }
        if (valid) {
            double Fo[10];
#pragma unroll
            for (int k = 0; k < 9; ++k) Fo[k] = Fm[k];
            Fo[9] = 0.0;
            store_vec<double, 10>(p.Fst + 10 * (int64_t)i, Fo);
        }
        det = mega::det3_abs(Fm);
        mega::inv3_ref(Fm, IF);
        mega::mm3(L0, IF, Lm);   // L = L0 F^-1 (FINT.F90:380-391 then 669-679)
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) Lm[k] = L0[k];
    }
    R *stress = as<R>(p.stress) + 6 * (int64_t)i, *strain = as<R>(p.strain) + 6 * (int64_t)i;
    double sig[6], e6[6], de[6], dd[6], old_s[6], old_e[6];
    {
        R s6[6], e6r[6];
        load6<R>(stress, s6);
        load6<R>(strain, e6r);
#pragma unroll
        for (int k = 0; k < 6; ++k) { sig[k] = old_s[k] = (double)s6[k]; e6[k] = old_e[k] = (double)e6r[k]; }
    }
    if (p.finite_strain) {
        double Rot[9];
        mega::hughes_winget(Lm, Rot, de, dd);
        mega::rotate_voigt(Rot, sig);
        mega::rotate_voigt(Rot, e6);
    } else {   // small strain (FINT.F90:772-783)
        de[0] = Lm[0]; de[1] = Lm[4]; de[2] = Lm[8];
        de[3] = Lm[5] + Lm[7]; de[4] = Lm[2] + Lm[6]; de[5] = Lm[1] + Lm[3];
#pragma unroll
        for (int k = 0; k < 6; ++k) dd[k] = 0.0;
    }
    R sr[6], der[6], pred[6], s_in[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) { e6[k] = e6[k] + de[k]; sr[k] = s_in[k] = (R)sig[k]; der[k] = (R)de[k]; }
    mega::elastic_trial(sr, der, mt, pred);
    R ep = as<R>(p.eps)[i];
    const R ep_n = ep;
    bool done = false;
    if constexpr (SURR) {
        const bool cand = valid && mt.type == 2 && mi == p.surr_mat && mega::trial_excess(pred, ep, mt) > 0.0f;
        float s_new[6];
        done = surr_return(p, cand, der, sr, pred, ep, s_new, mt, cnt);
        if (done)
#pragma unroll
            for (int k = 0; k < 6; ++k) sr[k] = s_new[k];
    }
    if (!done) {
        if (mt.type == 2) {
            if (mega::von_mises(sr, pred, ep, mt) < 0 && valid) atomicAdd(&p.flags[0], 1);
        } else if (mt.type == 1) {
#pragma unroll
            for (int k = 0; k < 6; ++k) sr[k] = pred[k];
        } else if (mt.type == 6 || mt.type == 3) {
            // CONSTITUTION.F90:58-75: the damage models take the rotated stress of the
            // previous step, the total strain and the objective increment
            R *st4 = as<R>(p.st4) + 4 * (int64_t)i;
            R es = st4[0], dam = st4[2], dsv = st4[3], tot[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) tot[k] = (R)e6[k];
            if (mt.type == 6) {
                if (mega::von_mises_dam(sr, tot, der, ep, dam, dsv, mt) < 0 && valid) atomicAdd(&p.flags[0], 1);
            } else {
                mega::druck_prag(sr, tot, der, es, ep, dam, dsv, mt);
            }
            if (valid) { st4[0] = es; st4[2] = dam; st4[3] = dsv; }
        } else if (valid) {
            atomicAdd(&p.flags[1], 1);
        }
    }
    if (p.hv) harvest_put<R>(p, valid && mt.type == 2 && mega::trial_excess(pred, ep_n, mt) > R(0), der, s_in, ep_n,
                             sr, ep);
    if (valid) {
        as<R>(p.eps)[i] = ep;
        double si[6], ei[6];
        R e6r[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) e6r[k] = (R)e6[k];
        store6<R>(stress, sr);
        store6<R>(strain, e6r);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            sig[k] = (double)sr[k];
            si[k] = sig[k] - old_s[k];
            ei[k] = e6[k] - old_e[k];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            if (p.finite_strain) work += 0.5 * dd[k] * (2.0 * sig[k] - si[k]) * vol * det;
            else work += 0.5 * ei[k] * si[k] * vol * det;
        }
        // GMAX_WVEL: output only on the RK path (its time-step estimate is commented
        // out, FINT.F90:1250-1326), so stored with the step's other outputs
        if (p.wave_vel && p.out_dof) p.wave_vel[i] = mega::wave_speed(si, ei, mt);
        // P = V J sigma IF^T  (f_j = P grad phi_ij)
        const double Sg[9] = {sig[0], sig[5], sig[4], sig[5], sig[1], sig[3], sig[4], sig[3], sig[2]};
        const double w = vol * det;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r)
                P[r + 3 * c] = (R)(w * (Sg[r] * IF[c] + Sg[r + 3] * IF[c + 3] + Sg[r + 6] * IF[c + 6]));
        P[9] = (R)(mt.rho * w);
#pragma unroll
        for (int k = 10; k < pstride<R>(); ++k) P[k] = R(0);
    }
    return work;
}

template <class R, int FM, bool SURR>
__device__ __forceinline__ void node_update(const Params &p)
{
    __shared__ double s_red[kBlock / 64];
    __shared__ unsigned s_cnt[kBlock / 64][4];   // learned return map: counters per wave
    constexpr int W = red_stride<R, FM == F_DIRECT>();
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i0 < p.n;
    const int i = valid ? i0 : p.n - 1;
    R red[W];
    load_vec<R, W>(as<R>(p.red) + (int64_t)W * i, red);
    R P[pstride<R>()];
    const double work = node_core<R, FM, SURR>(p, i, valid, red, P, s_cnt[threadIdx.x >> 6]);
    if (valid) store_vec<R, pstride<R>()>(as<R>(p.P) + pstride<R>() * (int64_t)i, P);
    const double ws = block_sum(work, s_red);
    if (threadIdx.x == 0) p.wpart[blockIdx.x] += ws;
    if (SURR && threadIdx.x == 0) {   // the block's counters; no same-address atomics
        unsigned long long *o = p.sstat + 4 * (int64_t)blockIdx.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            unsigned long long t = 0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) t += s_cnt[w][k];
            o[k] += t;
        }
    }
}

template <class R, int FM>
__global__ __launch_bounds__(kBlock) void k_node(Params p)
{
    node_update<R, FM, false>(p);
}

// with the learned return map (mixed precision): registers for three waves
// per SIMD, which the network's tiles fit without spilling (166-168 VGPRs)
template <int FM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_node_surr(Params p)
{
    node_update<float, FM, true>(p);
}

// ---------------------------------------------------------------------------
// force gather over the transposed lists, then (CORRECT) EBC, damping and the
// corrector, energies, and (PREDICT) the next predictor, per degree of freedom.

template <class R>
__device__ __forceinline__ void load_P(const R *Pb, int i, R *P)
{
    if (std::is_same<R, float>::value) {
        const float4 *q = reinterpret_cast<const float4 *>(Pb + 12 * (int64_t)i);
        const float4 a = q[0], b = q[1], c = q[2];
        P[0] = a.x; P[1] = a.y; P[2] = a.z; P[3] = a.w; P[4] = b.x; P[5] = b.y; P[6] = b.z; P[7] = b.w;
        P[8] = c.x; P[9] = c.y;
    } else {
        const double2 *q = reinterpret_cast<const double2 *>(Pb + 10 * (int64_t)i);
#pragma unroll
        for (int k = 0; k < 5; ++k) { const double2 v = q[k]; P[2 * k] = (R)v.x; P[2 * k + 1] = (R)v.y; }
    }
}

// one degree of freedom after its force is known: fint, EBC (fixity 1: v = a = 0,
// reaction into fext), damping and the corrector (MAIN.F90:432-488), the
// energy terms (MAIN.F90:524-553) and, fused, the next step's predictor
// (the arithmetic of predict_dof, without reading back v, a and dtot).
struct DofIn {
    double v, mass, damp, fxo, dtot;
    int ebc;
};

__device__ __forceinline__ DofIn dof_in(const Params &p, int node, int64_t q)
{
    DofIn r;
    r.v = p.v[q];
    r.ebc = p.ebc[q];
    r.mass = p.mass[p.mass_node ? node : q];
    r.damp = p.mats[p.mat[node]].damping;
    r.fxo = p.has_gravity ? p.fext_nmo[q] : 0.0;
    r.dtot = p.dtot[q];
    return r;
}

template <bool CORRECT, bool PREDICT>
__device__ __forceinline__ void dof_out(const Params &p, int64_t q, const DofIn &in, double f, double fx, double &ke,
                                        double &ie, double &ee)
{
    if (!CORRECT) {
        p.fint[q] = f;
        p.fext[q] = fx;
        return;
    }
    if (p.out_dof) p.fint[q] = f;
    const double dt = step_dt(p);
    double v = in.v, acc = 0.0;
    if (in.ebc == 1) {
        v = 0.0;
        fx = -f;
    } else {
        const double c = in.damp * in.mass;
        acc = (fx - f - c * v) / (in.mass + 0.5 * dt * c);
        v = v + 0.5 * dt * acc;
    }
    if (p.out_dof) {
        p.a[q] = acc;
        p.fext[q] = fx;
    }
    if (p.has_gravity) p.fext_nmo[q] = fx;
    ke = 0.5 * in.mass * v * v;
    if (!p.finite_strain) ie = 0.5 * f * in.dtot;
    // external work of the increment: d is read only where a force acts (a
    // zero force contributes +-0, which leaves the block's sum unchanged)
    const double fe = fx + in.fxo;
    ee = fe != 0.0 ? 0.5 * fe * p.d[q] : 0.0;
    if (PREDICT) {
        const double d = predicted_d(dt, v, acc);
        const double t = in.dtot + d;
        p.d[q] = d;
        p.dtot[q] = t;
        p.v[q] = predicted_v(dt, v, acc);
        if (p.pd) {
            const double u = p.uphy[q] + d;
            p.uphy[q] = u;
            p.x[q] = p.X0[q] + u;
        }
        if (p.d4) {   // the FP32 copy of d for the next gather (t4: see predict_dof)
            const int64_t i = q / 3, k = q - 3 * i;
            reinterpret_cast<float *>(p.d4)[4 * i + k] = (float)d;
        }
    } else {
        p.v[q] = v;
    }
}

template <bool CORRECT, bool PREDICT>
__device__ __forceinline__ void dof_update(const Params &p, int node, int64_t q, double f, double fx, double &ke,
                                           double &ie, double &ee)
{
    dof_out<CORRECT, PREDICT>(p, q, dof_in(p, node, q), f, fx, ke, ie, ee);
}

template <class R, bool CORRECT, bool PREDICT>
__global__ __launch_bounds__(kBlock) void k_force(Params p, int nblocks)
{
    __shared__ double s_red[kBlock / 64];
    constexpr bool MIX = std::is_same<R, float>::value;
    constexpr int U = MIX ? kFUnrollMix : kFUnrollF64;
    constexpr int PV = MIX ? 3 : 5;   // 16 B pieces of a P row (pstride values)
    using V = typename std::conditional<MIX, float4, double2>::type;
    extern __shared__ __align__(16) unsigned char mega_dyn[];
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int node = blk * kNodesPerFBlock + threadIdx.x / kFGroup;
    const int lane = threadIdx.x & (kFGroup - 1);
    const bool live = node < p.n;
    // lists that stay: the P rows of the block's distinct sources in LDS, and
    // the src field of each record is the source's position in that set
    int m = 0;
    V *Lp = reinterpret_cast<V *>(mega_dyn);
    if (p.fset_ptr) {
        const int s0 = p.fset_ptr[blk];
        m = p.fset_ptr[blk + 1] - s0;
        const V *Pg = reinterpret_cast<const V *>(p.P);
        if constexpr (MIX) {   // the 10 used values of a row, 40 B in LDS
            float2 *L2 = reinterpret_cast<float2 *>(mega_dyn);
            for (int r = threadIdx.x; r < m; r += kBlock) {
                const float4 *q = reinterpret_cast<const float4 *>(Pg) + 3 * (int64_t)p.fset[s0 + r];
                const float4 a = q[0], b = q[1], c = q[2];
                L2[5 * r] = make_float2(a.x, a.y); L2[5 * r + 1] = make_float2(a.z, a.w);
                L2[5 * r + 2] = make_float2(b.x, b.y); L2[5 * r + 3] = make_float2(b.z, b.w);
                L2[5 * r + 4] = make_float2(c.x, c.y);
            }
        } else {
            for (int k = threadIdx.x; k < PV * m; k += kBlock) {
                const int r = k / PV;
                Lp[k] = Pg[PV * (int64_t)p.fset[s0 + r] + (k - PV * r)];
            }
        }
        if (m > 0) __syncthreads();
    }
    R f0 = R(0), f1 = R(0), f2 = R(0), fe = R(0);
    auto sweep = [&](auto lds_tag) {
        constexpr bool LDS = decltype(lds_tag)::value;
        const int t0 = p.trow[node], t1 = p.trow[node + 1];
        const TPair<R> *__restrict__ tp = as<const TPair<R>>(p.tpair);
        const R *__restrict__ Pb = as<R>(p.P);
        const R *__restrict__ ts = as<R>(p.tshp);
        for (int base = t0 + lane; base < t1; base += kFGroup * U) {
            R g0[U], g1[U], g2[U], P[U][10], sh[U];
            int src[U];
#pragma unroll
            for (int e = 0; e < U; ++e) {
                const int t = base + kFGroup * e;
                const bool ok = t < t1;
                const TPair<R> v = tp[ok ? t : t0];
                g0[e] = ok ? v.gx : R(0);
                g1[e] = ok ? v.gy : R(0);
                g2[e] = ok ? v.gz : R(0);
                src[e] = ok ? v.src : (LDS ? 0 : node);
                sh[e] = (ok && p.has_gravity) ? ts[t] : R(0);
            }
#pragma unroll
            for (int e = 0; e < U; ++e) {
                if constexpr (LDS) {
                    const V *q = Lp + PV * src[e];
                    R *o = P[e];
                    if constexpr (MIX) {
                        const float2 *q2 = reinterpret_cast<const float2 *>(mega_dyn) + 5 * src[e];
#pragma unroll
                        for (int k = 0; k < 5; ++k) { const float2 a = q2[k]; o[2 * k] = a.x; o[2 * k + 1] = a.y; }
                    } else {
#pragma unroll
                        for (int k = 0; k < 5; ++k) { const V v2 = q[k]; o[2 * k] = v2.x; o[2 * k + 1] = v2.y; }
                    }
                } else {
                    load_P<R>(Pb, src[e], P[e]);
                }
            }
#pragma unroll
            for (int e = 0; e < U; ++e) {
                f0 += P[e][0] * g0[e] + P[e][3] * g1[e] + P[e][6] * g2[e];
                f1 += P[e][1] * g0[e] + P[e][4] * g1[e] + P[e][7] * g2[e];
                f2 += P[e][2] * g0[e] + P[e][5] * g1[e] + P[e][8] * g2[e];
                fe += sh[e] * P[e][9];
            }
        }
    };
    if (live) {
        if (m > 0) sweep(std::true_type{});
        else sweep(std::false_type{});
    }
    f0 = group_sum<kFGroup>(f0); f1 = group_sum<kFGroup>(f1); f2 = group_sum<kFGroup>(f2);
    if (p.has_gravity) fe = group_sum<kFGroup>(fe);
    double ke = 0.0, ie = 0.0, ee = 0.0;
    if constexpr (MIX) {
        // the block's degrees of freedom on consecutive lanes: the corrector's
        // node arrays are then read and written in one contiguous run per block
        static_assert(3 * kNodesPerFBlock <= kBlock, "one lane per degree of freedom of the block");
        __shared__ double s_f[3 * kNodesPerFBlock], s_fe[kNodesPerFBlock];
        const int loc = threadIdx.x / kFGroup;
        if (lane == 0) {
            s_f[3 * loc] = (double)f0; s_f[3 * loc + 1] = (double)f1; s_f[3 * loc + 2] = (double)f2;
            s_fe[loc] = (double)fe;
        }
        __syncthreads();
        if (threadIdx.x < 3 * kNodesPerFBlock) {
            const int dn = threadIdx.x / 3, k = threadIdx.x - 3 * dn;
            const int nd = blk * kNodesPerFBlock + dn;
            if (nd < p.n) {
                const double fx = p.has_gravity ? s_fe[dn] * p.grav[k] : 0.0;
                dof_update<CORRECT, PREDICT>(p, nd, 3 * (int64_t)nd + k, s_f[threadIdx.x], fx, ke, ie, ee);
            }
        }
    } else if (live && lane < 3) {
        const double f = (double)(lane == 0 ? f0 : (lane == 1 ? f1 : f2));
        const double fx = p.has_gravity ? (double)fe * p.grav[lane] : 0.0;
        dof_update<CORRECT, PREDICT>(p, node, 3 * (int64_t)node + lane, f, fx, ke, ie, ee);
    }
    if (CORRECT) {   // per-block partials: no same-address atomics, deterministic
        const double sk = block_sum(ke, s_red);
        const double si = block_sum(ie, s_red);
        const double se = block_sum(ee, s_red);
        if (threadIdx.x == 0) {
            p.epart[3 * blk] = sk;
            p.epart[3 * blk + 1] = si;
            p.epart[3 * blk + 2] += se;
        }
    }
}

// ---------------------------------------------------------------------------
// correspondence peridynamics (FINT_PD.F90:14-759, RK.f90:579-807), Lagrangian.
//
//   k_pd_shape   once: influence weights w_ij = prod_k w(|X_j,k - X_i,k| / a_j,k)
//                (list and transposed order), K = sum w V_j xi xi^T and its inverse
//                by the reference's pivoting INVERSE, g_i = sum w V_j (X_i - X_j)
//   k_pd_gather  S = sum w V_j (x_j - x_i) (x) xi, L0 = sum w V_j (d_j - d_i) (x) xi
//   k_pd_node    F = S K^-1, |det F|, F^-1 (pivoting INVERSE), L = L0 K^-1 F^-1,
//                Hughes-Winget, rotation, trial, return map; T = J sigma F^-T K^-1
//   k_pd_force   f_i = V_i (T_i g_i - sum_{k: i in N(k)} w_ki V_k T_k (X_k - X_i)):
//                the reference's +/- bond scatter (FINT_PD.F90:629-655) as one
//                deterministic gather over the transposed bond list, then the
//                corrector as for the RK force

template <class R>
__global__ __launch_bounds__(kBlock) void k_pd_shape(Params p)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int node = (int)(t / kGroup), lane = (int)(t % kGroup);
    const bool live = node < p.n;
    double K[6] = {0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0};   // K: 00 01 02 11 12 22
    if (live) {
        const double xi0 = p.X0[3 * node], xi1 = p.X0[3 * node + 1], xi2 = p.X0[3 * node + 2];
        R *w = as<R>(p.shp), *tw = as<R>(p.tshp);
        for (int q = p.row_ptr[node] + lane; q < p.row_ptr[node + 1]; q += kGroup) {
            const int j = p.col[q];
            const double h0 = -(xi0 - p.X0[3 * j]), h1 = -(xi1 - p.X0[3 * j + 1]), h2 = -(xi2 - p.X0[3 * j + 2]);
            double px, py, pz, dd;
            mega::cubic_kernel(fabs(h0 / p.win[3 * j]), px, dd);
            mega::cubic_kernel(fabs(h1 / p.win[3 * j + 1]), py, dd);
            mega::cubic_kernel(fabs(h2 / p.win[3 * j + 2]), pz, dd);
            const double om = px * py * pz, vj = p.vol[j];
            w[q] = (R)om;
            tw[p.tpos[q]] = (R)om;
            if constexpr (std::is_same<R, float>::value) {
                // mixed: the bond weights of the gather and the force, w V_j in list order
                // (the x block of dshp) and w V_i with the source in the transposed slot;
                // build_pd_sets packs them into 8 B records
                as<float>(p.dshp)[q] = (float)(om * vj);
                as<TPair<float>>(p.tpair)[p.tpos[q]] = TPair<float>{(float)(om * p.vol[node]), 0.f, 0.f,
                                                                    (int32_t)node};
            }
            K[0] += h0 * h0 * om * vj; K[1] += h0 * h1 * om * vj; K[2] += h0 * h2 * om * vj;
            K[3] += h1 * h1 * om * vj; K[4] += h1 * h2 * om * vj; K[5] += h2 * h2 * om * vj;
            g[0] -= om * vj * h0; g[1] -= om * vj * h1; g[2] -= om * vj * h2;
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) K[k] = group_sum<kGroup>(K[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = group_sum<kGroup>(g[k]);
    if (live && lane == 0) {
        const double Km[9] = {K[0], K[1], K[2], K[1], K[3], K[4], K[2], K[4], K[5]};
        double Ki[9];
        if (!mega::inverse_gj3(Km, Ki)) atomicAdd(&p.flags[2], 1);
        double *out = p.Kinv + 9 * (int64_t)node;
#pragma unroll
        for (int k = 0; k < 9; ++k) out[k] = Ki[k];
        double *go = p.gvec + 3 * (int64_t)node;
        go[0] = g[0]; go[1] = g[1]; go[2] = g[2];
    }
}

// Automatic time step (IAUTO_TS = 1; FINT_PD.F90:690-754).  k_pd_gather
// leaves each particle's shortest bond to another particle at the current
// coordinates in char_dist (GCHAR_DIST; a particle with none keeps its
// previous value); here the smallest quotient char_dist / wave speed over all
// particles.  Non-negative doubles order as their bit patterns, so the blocks
// (a grid-stride loop over at most kDtBlocks of them, which keeps the atomics
// on the one address few) meet in one unsigned 64-bit atomicMin; a NaN
// quotient is no candidate.
constexpr int kDtBlocks = 1024;
__global__ __launch_bounds__(kBlock) void k_pd_dt(int n, const double *char_dist, const double *wave_vel,
                                                  unsigned long long *dmin)
{
    __shared__ unsigned long long s_min[kBlock / 64];
    unsigned long long key = ~0ull;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double w = char_dist[i] / wave_vel[i];
        const unsigned long long k = (unsigned long long)__double_as_longlong(w);
        if (w == w && k < key) key = k;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const unsigned long long o = __shfl_xor(key, m, 64);
        key = o < key ? o : key;
    }
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long k = s_min[0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; ++w) k = s_min[w] < k ? s_min[w] : k;
        if (k != ~0ull) atomicMin(dmin, k);
    }
}

// DLT = DLT_FAC * the smallest quotient (FINT_PD.F90:750); the minimum is
// reset for the next step
__global__ void k_pd_dt_fin(unsigned long long *dmin, double dlt_fac, double *dt_dev, double *hist, int s)
{
    if (threadIdx.x != 0) return;
    const double d = __longlong_as_double((long long)*dmin) * dlt_fac;
    dt_dev[0] = d;
    hist[s] = d;
    *dmin = ~0ull;
}

// mixed peridynamics: FP32 u = x - X0 of every particle for the gather
__global__ void k_pd_u4(int n, const double *x, const double *X0, float4 *u4)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u4[i] = make_float4((float)(x[3 * i] - X0[3 * i]), (float)(x[3 * i + 1] - X0[3 * i + 1]),
                        (float)(x[3 * i + 2] - X0[3 * i + 2]), 0.f);
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_pd_gather(Params p, int nblocks)
{
    constexpr int W = std::is_same<R, float>::value ? 21 : 18;
    constexpr bool MIX = std::is_same<R, float>::value;
    constexpr int G = MIX ? kPdG : kGroup;   // lanes per particle
    extern __shared__ __align__(16) unsigned char mega_dyn[];
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int node = blk * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x & (G - 1);
    const bool live = node < p.n;
    double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // automatic time step: the shortest squared bond to another particle at the
    // current coordinates (FINT_PD.F90:710-730)
    const bool cd_on = p.char_dist != nullptr;
    double dmin = INFINITY;
    if constexpr (MIX) {
        // mixed: S - K = sum w V_j (u_j - u_i) (X_j - X_i)^T with u = x - X0 and
        // L0 = sum w V_j (d_j - d_i) (X_j - X_i)^T from FP32 copies of u and d, each
        // lane's sums in FP32 (terms of one sign and size: differences, not
        // positions), the lanes' partials in FP64.  One
        // 8 B record per bond, {w V_j, j's position in the block's source set (or j,
        // for a block without a set)}; the bond X_j - X_i is formed in FP32 from
        // reference positions relative to the block's first particle X_b, staged
        // with u and d (in their fourth lanes, z apart).  k_pd_node adds K back as
        // F = I + (S - K) K^-1.
        // With the automatic time step the set also holds FP32 current positions
        // relative to the block's first particle, x_j - x_b: the bond x_j - x_i is then
        // formed in FP32 from numbers of the bond's own size.
        float4 *Lu = reinterpret_cast<float4 *>(mega_dyn);
        int m = 0;
        const int nb0 = min(blk * (kBlock / G), p.n - 1);
        const double xb[3] = {p.x[3 * (int64_t)nb0], p.x[3 * (int64_t)nb0 + 1], p.x[3 * (int64_t)nb0 + 2]};
        const double Xb[3] = {p.X0[3 * (int64_t)nb0], p.X0[3 * (int64_t)nb0 + 1], p.X0[3 * (int64_t)nb0 + 2]};
        const double iq = p.pd_qb ? 1.0 / p.pd_qb[blk] : 1.0;   // 1 / the block's quantum
        float4 *Ld = Lu, *Lx = Lu;
        float *Lz = reinterpret_cast<float *>(Lu);
        if (p.gset_ptr && p.pd8) {
            const int s0 = p.gset_ptr[blk];
            m = p.gset_ptr[blk + 1] - s0;
            Ld = Lu + m;
            Lz = reinterpret_cast<float *>(Lu + 2 * m);
            Lx = Lu + 2 * m + (m + 3) / 4;
            for (int k = threadIdx.x; k < m; k += kBlock) {
                const int j = p.gset[s0 + k];
                const float4 u = p.u4[j], d = p.d4[j];
                Lu[k] = make_float4(u.x, u.y, u.z, __int_as_float(pd_fix(p.X0[3 * (int64_t)j], Xb[0], iq)));
                Ld[k] = make_float4(d.x, d.y, d.z, __int_as_float(pd_fix(p.X0[3 * (int64_t)j + 1], Xb[1], iq)));
                Lz[k] = __int_as_float(pd_fix(p.X0[3 * (int64_t)j + 2], Xb[2], iq));
                if (cd_on)
                    Lx[k] = make_float4((float)(p.x[3 * (int64_t)j] - xb[0]), (float)(p.x[3 * (int64_t)j + 1] - xb[1]),
                                        (float)(p.x[3 * (int64_t)j + 2] - xb[2]), 0.f);
            }
            if (m > 0) __syncthreads();
        }
        if (live) {
            const float4 ui = p.u4[node], di = p.d4[node];
            const double xi[3] = {p.x[3 * (int64_t)node], p.x[3 * (int64_t)node + 1], p.x[3 * (int64_t)node + 2]};
            const float pi[3] = {(float)(xi[0] - xb[0]), (float)(xi[1] - xb[1]), (float)(xi[2] - xb[2])};
            const int Xi[3] = {pd_fix(p.X0[3 * (int64_t)node], Xb[0], iq),
                               pd_fix(p.X0[3 * (int64_t)node + 1], Xb[1], iq),
                               pd_fix(p.X0[3 * (int64_t)node + 2], Xb[2], iq)};
            const double *Xn = p.X0 + 3 * (int64_t)node;
            const int q0 = p.row_ptr[node], q1 = p.row_ptr[node + 1];
            const float2 *rec8 = p.pd8;
            float Sf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Lf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            const float *rw = as<float>(p.dshp);   // w V_j in list order, before the records are packed
            // LDS: the block's source set (j is a set position), else global memory (j is
            // the particle); one instantiation each, so that no bond chooses between them
            auto sweep = [&](auto lds_tag) {
                constexpr bool LDS = decltype(lds_tag)::value;
                auto load_r = [&](int base, float2 *rec) {
#pragma unroll
                    for (int e = 0; e < kPdUnroll; ++e) {
                        const int q = base + G * e;
                        const int pad = LDS ? 0 : node;   // a zero weight on a valid source
                        if (LDS || rec8) rec[e] = q < q1 ? rec8[q] : make_float2(0.f, __int_as_float(pad));
                        else rec[e] = q < q1 ? make_float2(rw[q], __int_as_float(p.col[q])) : make_float2(0.f, __int_as_float(pad));
                    }
                };
                float2 recn[kPdUnroll];
                if (q0 + lane < q1) load_r(q0 + lane, recn);
                for (int base = q0 + lane; base < q1; base += G * kPdUnroll) {
                    float2 rec[kPdUnroll];
                    float4 uj[kPdUnroll], dj[kPdUnroll];
                    float r[kPdUnroll][3];
#pragma unroll
                    for (int e = 0; e < kPdUnroll; ++e) rec[e] = recn[e];
                    if (base + G * kPdUnroll < q1) load_r(base + G * kPdUnroll, recn);
                    // bond j: (w V_j (X_j - X_i), u_j, d_j)
#pragma unroll
                    for (int e = 0; e < kPdUnroll; ++e) {
                        const int j = __float_as_int(rec[e].y);
                        float xr[3];
                        if constexpr (LDS) {
                            uj[e] = Lu[j];
                            dj[e] = Ld[j];
                            xr[0] = (float)(__float_as_int(uj[e].w) - Xi[0]);
                            xr[1] = (float)(__float_as_int(dj[e].w) - Xi[1]);
                            xr[2] = (float)(__float_as_int(Lz[j]) - Xi[2]);
                        } else {
                            uj[e] = p.u4[j];
                            dj[e] = p.d4[j];
#pragma unroll
                            for (int k = 0; k < 3; ++k) xr[k] = (float)((p.X0[3 * (int64_t)j + k] - Xn[k]) * iq);
                        }
#pragma unroll
                        for (int k = 0; k < 3; ++k) r[e][k] = rec[e].x * xr[k];
                    }
                    if (cd_on) {
#pragma unroll
                        for (int e = 0; e < kPdUnroll; ++e) {
                            const int j = __float_as_int(rec[e].y);
                            if constexpr (LDS) {
                                // the particle itself is the bond of zero length; padding has zero weight
                                if (rec[e].x != 0.f && (r[e][0] != 0.f || r[e][1] != 0.f || r[e][2] != 0.f)) {
                                    const float4 pj = Lx[j];
                                    const float cx = pj.x - pi[0], cy = pj.y - pi[1], cz = pj.z - pi[2];
                                    dmin = fmin(dmin, (double)(cx * cx + cy * cy + cz * cz));
                                }
                            } else if (j != node && rec[e].x != 0.f) {
                                dmin = fmin(dmin, sq_len_unfused(p.x[3 * (int64_t)j] - xi[0], p.x[3 * (int64_t)j + 1] - xi[1],
                                                                 p.x[3 * (int64_t)j + 2] - xi[2]));
                            }
                        }
                    }
#pragma unroll
                    for (int e = 0; e < kPdUnroll; ++e) {
                        const float a[3] = {uj[e].x - ui.x, uj[e].y - ui.y, uj[e].z - ui.z};
                        const float b[3] = {dj[e].x - di.x, dj[e].y - di.y, dj[e].z - di.z};
#pragma unroll
                        for (int bb = 0; bb < 3; ++bb)
#pragma unroll
                            for (int aa = 0; aa < 3; ++aa) {
                                Sf[aa + 3 * bb] += a[aa] * r[e][bb];
                                Lf[aa + 3 * bb] += b[aa] * r[e][bb];
                            }
                    }
                }
            };
            if (m > 0) sweep(std::true_type{});
            else sweep(std::false_type{});
#pragma unroll
            for (int k = 0; k < 9; ++k) { S[k] = (double)Sf[k]; L[k] = (double)Lf[k]; }
        }
    } else {
        // FP64: with the packed records {w V_j, j's set position} a block with a source
        // set stages X0, x and d of its sources in LDS (80 B rows); the same products
        // in the same lanes and order as without them
        double2 *Ls = reinterpret_cast<double2 *>(mega_dyn);
        int m = 0;
        if (p.pd16 && p.gset_ptr) {
            const int s0 = p.gset_ptr[blk];
            m = p.gset_ptr[blk + 1] - s0;
            for (int k = threadIdx.x; k < m; k += kBlock) {
                const int64_t j = p.gset[s0 + k];
                Ls[5 * k] = make_double2(p.X0[3 * j], p.X0[3 * j + 1]);
                Ls[5 * k + 1] = make_double2(p.X0[3 * j + 2], p.x[3 * j]);
                Ls[5 * k + 2] = make_double2(p.x[3 * j + 1], p.x[3 * j + 2]);
                Ls[5 * k + 3] = make_double2(p.d[3 * j], p.d[3 * j + 1]);
                Ls[5 * k + 4] = make_double2(p.d[3 * j + 2], 0.0);
            }
            if (m > 0) __syncthreads();
        }
        if (live) {
            const double X[3] = {p.X0[3 * node], p.X0[3 * node + 1], p.X0[3 * node + 2]};
            const double x[3] = {p.x[3 * node], p.x[3 * node + 1], p.x[3 * node + 2]};
            const double d[3] = {p.d[3 * node], p.d[3 * node + 1], p.d[3 * node + 2]};
            auto term = [&](const double *h, const double *hs, const double *hd, double wv) {
#pragma unroll
                for (int b = 0; b < 3; ++b)
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        S[a + 3 * b] += hs[a] * h[b] * wv;
                        L[a + 3 * b] += hd[a] * h[b] * wv;
                    }
            };
            if (m > 0) {
                constexpr int U = 4;
                const int q0 = p.row_ptr[node], q1 = p.row_ptr[node + 1];
                auto load_r = [&](int base, double2 *rec) {
#pragma unroll
                    for (int e = 0; e < U; ++e) {
                        const int q = base + kGroup * e;
                        rec[e] = q < q1 ? p.pd16[q] : make_double2(0.0, 0.0);
                    }
                };
                double2 recn[U];
                if (q0 + lane < q1) load_r(q0 + lane, recn);
                for (int base = q0 + lane; base < q1; base += kGroup * U) {
                    double2 rec[U];
#pragma unroll
                    for (int e = 0; e < U; ++e) rec[e] = recn[e];
                    if (base + kGroup * U < q1) load_r(base + kGroup * U, recn);
#pragma unroll
                    for (int e = 0; e < U; ++e) {
                        if (base + kGroup * e >= q1) break;
                        const int k = (int)__double_as_longlong(rec[e].y);
                        const double2 v0 = Ls[5 * k], v1 = Ls[5 * k + 1], v2 = Ls[5 * k + 2], v3 = Ls[5 * k + 3],
                                      v4 = Ls[5 * k + 4];
                        const double h[3] = {v0.x - X[0], v0.y - X[1], v1.x - X[2]};
                        const double hs[3] = {v1.y - x[0], v2.x - x[1], v2.y - x[2]};
                        const double hd[3] = {v3.x - d[0], v3.y - d[1], v4.x - d[2]};
                        // the particle itself is the bond of zero reference length
                        if (cd_on && (h[0] != 0.0 || h[1] != 0.0 || h[2] != 0.0))
                            dmin = fmin(dmin, sq_len_unfused(hs[0], hs[1], hs[2]));
                        term(h, hs, hd, rec[e].x);
                    }
                }
            } else {
                const R *w = as<R>(p.shp);
                for (int q = p.row_ptr[node] + lane; q < p.row_ptr[node + 1]; q += kGroup) {
                    const int j = p.col[q];
                    const double wv = (double)w[q] * p.vol[j];
                    const double h[3] = {p.X0[3 * j] - X[0], p.X0[3 * j + 1] - X[1], p.X0[3 * j + 2] - X[2]};
                    const double hs[3] = {p.x[3 * j] - x[0], p.x[3 * j + 1] - x[1], p.x[3 * j + 2] - x[2]};
                    const double hd[3] = {p.d[3 * j] - d[0], p.d[3 * j + 1] - d[1], p.d[3 * j + 2] - d[2]};
                    if (cd_on && j != node) dmin = fmin(dmin, sq_len_unfused(hs[0], hs[1], hs[2]));
                    term(h, hs, hd, wv);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) { S[k] = group_sum<G>(S[k]); L[k] = group_sum<G>(L[k]); }
    if (cd_on) {
        // the square root of the smallest square is the smallest DSQRT (sqrt is
        // monotone); a particle with no other neighbor keeps its distance
        dmin = group_min<G>(dmin);
        if (live && lane == 0 && dmin < INFINITY) p.char_dist[node] = sqrt(dmin);
    }
    if (live && lane == 0) {
        R *red = as<R>(p.red) + (int64_t)W * node;
#pragma unroll
        for (int k = 0; k < 9; ++k) { red[k] = (R)S[k]; red[9 + k] = (R)L[k]; }
    }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_pd_node(Params p)
{
    constexpr int W = std::is_same<R, float>::value ? 21 : 18;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const R *red = as<R>(p.red) + (int64_t)W * i;
    double S[9], L0[9], Ki[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) { S[k] = (double)red[k]; L0[k] = (double)red[9 + k]; Ki[k] = p.Kinv[9 * (int64_t)i + k]; }
    double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, IF[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, det = 1.0;
    if (!p.pd_linit && p.finite_strain) {
        mega::mm3(S, Ki, F);
        if constexpr (std::is_same<R, float>::value) {   // mixed: the gather gave S - K
            F[0] += 1.0; F[4] += 1.0; F[8] += 1.0;
        }
        det = mega::det3_abs(F);
        if (!mega::inverse_gj3(F, IF)) atomicAdd(&p.flags[2], 1);
    }
    double L1[9], Lm[9];
    mega::mm3(L0, Ki, L1);
    if (p.pd_linit) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Lm[k] = L1[k];
    } else {
        mega::mm3(L1, IF, Lm);
    }
    const Material &mt = p.mats[p.mat[i]];
    R *stress = as<R>(p.stress) + 6 * (int64_t)i, *strain = as<R>(p.strain) + 6 * (int64_t)i;
    double sig[6], e6[6], de[6], dd[6], old_s[6], Rot[9];
#pragma unroll
    for (int k = 0; k < 6; ++k) { sig[k] = old_s[k] = (double)stress[k]; e6[k] = (double)strain[k]; }
    mega::hughes_winget(Lm, Rot, de, dd);   // FINT_PD.F90:475 (always, also for small strain)
    mega::rotate_voigt(Rot, sig);
    mega::rotate_voigt(Rot, e6);
    R sr[6], der[6], pred[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) { e6[k] = e6[k] + de[k]; sr[k] = (R)sig[k]; der[k] = (R)de[k]; }
    mega::elastic_trial(sr, der, mt, pred);
    R ep = as<R>(p.eps)[i];
    if (mt.type == 2) {
        if (mega::von_mises(sr, pred, ep, mt) < 0) atomicAdd(&p.flags[0], 1);
    } else if (mt.type == 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) sr[k] = pred[k];
    } else if (mt.type == 6 || mt.type == 3) {
        // CONSTITUTION (FINT_PD.F90:493) as on the RK path: the damage models take
        // the rotated stress, the total strain LSTRAIN and the increment
        R *st4 = as<R>(p.st4) + 4 * (int64_t)i;
        R es = st4[0], dam = st4[2], dsv = st4[3], tot[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) tot[k] = (R)e6[k];
        if (mt.type == 6) {
            if (mega::von_mises_dam(sr, tot, der, ep, dam, dsv, mt) < 0) atomicAdd(&p.flags[0], 1);
        } else {
            mega::druck_prag(sr, tot, der, es, ep, dam, dsv, mt);
        }
        st4[0] = es; st4[2] = dam; st4[3] = dsv;
    } else {
        atomicAdd(&p.flags[1], 1);
    }
    as<R>(p.eps)[i] = ep;
    double si[6], ei[6];
    // mixed: the stress increment of the wave-speed estimate without the FP32
    // roundings of the whole stress (rotated stress to FP32, trial sum): the
    // rotation's change and C:de in FP64 plus the return map's correction,
    // equal in exact arithmetic to new - old.  ESTIMATE_MODULI's secant moduli
    // switch at increments of 1e-8, where those roundings alone would decide.
    double cde[6] = {0, 0, 0, 0, 0, 0};
    if constexpr (std::is_same<R, float>::value) {
        const double z6[6] = {0, 0, 0, 0, 0, 0};
        mega::elastic_trial(z6, de, mt, cde);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        stress[k] = sr[k];
        strain[k] = (R)e6[k];
        if constexpr (std::is_same<R, float>::value)
            si[k] = (sig[k] - old_s[k]) + cde[k] + ((double)sr[k] - (double)pred[k]);
        sig[k] = (double)sr[k];
        if constexpr (!std::is_same<R, float>::value) si[k] = sig[k] - old_s[k];
        ei[k] = e6[k] - old_s[k];   // the reference's STRAIN_INC = LSTRAIN - GSTRESS (FINT_PD.F90:509)
    }
    if (p.wave_vel) {
        double wv = mega::wave_speed(si, ei, mt);
        if (mt.type == 3) wv = wv / 0.15;   // the reference's Drucker-Prager factor (FINT_PD.F90:539-542)
        p.wave_vel[i] = wv;
    }
    // T = det sigma IF^T K^-1 (FINT_PD.F90:620-624)
    const double Sg[9] = {sig[0], sig[5], sig[4], sig[5], sig[1], sig[3], sig[4], sig[3], sig[2]};
    double IFt[9], PK[9], T[9];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) IFt[a + 3 * b] = IF[b + 3 * a];
    mega::mm3(Sg, IFt, PK);
#pragma unroll
    for (int k = 0; k < 9; ++k) PK[k] = det * PK[k];
    mega::mm3(PK, Ki, T);
    R *P = as<R>(p.P) + pstride<R>() * (int64_t)i;
#pragma unroll
    for (int k = 0; k < 9; ++k) P[k] = (R)T[k];
    P[9] = (R)p.vol[i];
}

template <class R, bool CORRECT, bool PREDICT>
__global__ __launch_bounds__(kBlock) void k_pd_force(Params p, int nblocks)
{
    __shared__ double s_red[kBlock / 64];
    constexpr bool MIX = std::is_same<R, float>::value;
    constexpr int G = MIX ? kPdG : kGroup;   // lanes per particle
    extern __shared__ __align__(16) unsigned char mega_dyn[];
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int node = blk * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x & (G - 1);
    const bool live = node < p.n;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const R *Pb = as<R>(p.P);
    if constexpr (MIX) {
        // mixed: one 8 B record per transposed slot, {w V_k, k's position in the
        // block's source set (or k, for a block without a set)}; a block with a set
        // stages per source the 9 used values of its T row and its reference
        // position relative to the block's first particle X_b (FP32, 48 B), and the
        // bond X_k - X_i is formed from those
        float4 *Lt = reinterpret_cast<float4 *>(mega_dyn);
        int m = 0;
        const int nb0 = min(blk * (kBlock / G), p.n - 1);
        const double Xb[3] = {p.X0[3 * (int64_t)nb0], p.X0[3 * (int64_t)nb0 + 1], p.X0[3 * (int64_t)nb0 + 2]};
        const double iq = p.pd_qb ? 1.0 / p.pd_qb[blk] : 1.0;   // 1 / the block's quantum
        if (p.fset_ptr && p.pd8) {
            const int s0 = p.fset_ptr[blk];
            m = p.fset_ptr[blk + 1] - s0;
            for (int r = threadIdx.x; r < m; r += kBlock) {
                const int k = p.fset[s0 + r];
                const float4 *q = reinterpret_cast<const float4 *>(Pb) + 3 * (int64_t)k;
                const float4 a = q[0], b = q[1], c = q[2];
                Lt[3 * r] = a;
                Lt[3 * r + 1] = b;
                Lt[3 * r + 2] = make_float4(c.x, __int_as_float(pd_fix(p.X0[3 * (int64_t)k], Xb[0], iq)),
                                            __int_as_float(pd_fix(p.X0[3 * (int64_t)k + 1], Xb[1], iq)),
                                            __int_as_float(pd_fix(p.X0[3 * (int64_t)k + 2], Xb[2], iq)));
            }
            if (m > 0) __syncthreads();
        }
        if (live) {
            const int Xi[3] = {pd_fix(p.X0[3 * (int64_t)node], Xb[0], iq),
                               pd_fix(p.X0[3 * (int64_t)node + 1], Xb[1], iq),
                               pd_fix(p.X0[3 * (int64_t)node + 2], Xb[2], iq)};
            const double *Xn = p.X0 + 3 * (int64_t)node;
            const float2 *rec8 = p.pd8 ? p.pd8 + p.cap : nullptr;
            const TPair<float> *tp = as<const TPair<float>>(p.tpair);   // {w V_k, -, -, k} before packing
            const int t0 = p.trow[node], t1 = p.trow[node + 1];
            auto load_e = [&](int base, float2 *e) {
#pragma unroll
                for (int u = 0; u < kPdUnroll; ++u) {
                    const int t = base + G * u;
                    if (rec8) e[u] = t < t1 ? rec8[t] : make_float2(0.f, 0.f);
                    else e[u] = t < t1 ? make_float2(tp[t].gx, __int_as_float(tp[t].src)) : make_float2(0.f, 0.f);
                }
            };
            float2 en[kPdUnroll];
            if (t0 + lane < t1) load_e(t0 + lane, en);
            for (int base = t0 + lane; base < t1; base += G * kPdUnroll) {
                float2 e[kPdUnroll];
#pragma unroll
                for (int u = 0; u < kPdUnroll; ++u) e[u] = en[u];
                if (base + G * kPdUnroll < t1) load_e(base + G * kPdUnroll, en);
#pragma unroll
                for (int u = 0; u < kPdUnroll; ++u) {
                    const int k = __float_as_int(e[u].y);
                    float T[10], xr[3];
                    if (m > 0) {
                        const float4 v0 = Lt[3 * k], v1 = Lt[3 * k + 1], v2 = Lt[3 * k + 2];
                        T[0] = v0.x; T[1] = v0.y; T[2] = v0.z; T[3] = v0.w; T[4] = v1.x; T[5] = v1.y;
                        T[6] = v1.z; T[7] = v1.w; T[8] = v2.x;
                        xr[0] = (float)(__float_as_int(v2.y) - Xi[0]);
                        xr[1] = (float)(__float_as_int(v2.z) - Xi[1]);
                        xr[2] = (float)(__float_as_int(v2.w) - Xi[2]);
                    } else {
                        load_P<float>(Pb, k, T);
#pragma unroll
                        for (int c = 0; c < 3; ++c) xr[c] = (float)((p.X0[3 * (int64_t)k + c] - Xn[c]) * iq);
                    }
                    const float gx = e[u].x * xr[0], gy = e[u].x * xr[1], gz = e[u].x * xr[2];
                    a0 += (double)(T[0] * gx + T[3] * gy + T[6] * gz);
                    a1 += (double)(T[1] * gx + T[4] * gy + T[7] * gz);
                    a2 += (double)(T[2] * gx + T[5] * gy + T[8] * gz);
                }
            }
        }
    } else {
        // FP64: with the packed records {w V_k, k's set position} a block with a source
        // set stages the T rows (9 values) and X0 of its sources in LDS (96 B rows);
        // the same products in the same lanes and order as without them
        double2 *Ls = reinterpret_cast<double2 *>(mega_dyn);
        int m = 0;
        if (p.pd16 && p.fset_ptr) {
            const int s0 = p.fset_ptr[blk];
            m = p.fset_ptr[blk + 1] - s0;
            for (int r = threadIdx.x; r < m; r += kBlock) {
                const int64_t k = p.fset[s0 + r];
                const double2 *q = reinterpret_cast<const double2 *>(Pb + 10 * k);
                const double2 t0 = q[0], t1 = q[1], t2 = q[2], t3 = q[3], t4 = q[4];
                Ls[6 * r] = t0; Ls[6 * r + 1] = t1; Ls[6 * r + 2] = t2; Ls[6 * r + 3] = t3;
                Ls[6 * r + 4] = make_double2(t4.x, p.X0[3 * k]);
                Ls[6 * r + 5] = make_double2(p.X0[3 * k + 1], p.X0[3 * k + 2]);
            }
            if (m > 0) __syncthreads();
        }
        if (live) {
            const double X[3] = {p.X0[3 * node], p.X0[3 * node + 1], p.X0[3 * node + 2]};
            auto term = [&](const double *T, const double *dx, double c) {
                a0 += c * (T[0] * dx[0] + T[3] * dx[1] + T[6] * dx[2]);
                a1 += c * (T[1] * dx[0] + T[4] * dx[1] + T[7] * dx[2]);
                a2 += c * (T[2] * dx[0] + T[5] * dx[1] + T[8] * dx[2]);
            };
            if (m > 0) {
                constexpr int U = 4;
                const int s0 = p.trow[node], s1 = p.trow[node + 1];
                auto load_r = [&](int base, double2 *rec) {
#pragma unroll
                    for (int e = 0; e < U; ++e) {
                        const int t = base + kGroup * e;
                        rec[e] = t < s1 ? p.pd16[p.cap + t] : make_double2(0.0, 0.0);
                    }
                };
                double2 recn[U];
                if (s0 + lane < s1) load_r(s0 + lane, recn);
                for (int base = s0 + lane; base < s1; base += kGroup * U) {
                    double2 rec[U];
#pragma unroll
                    for (int e = 0; e < U; ++e) rec[e] = recn[e];
                    if (base + kGroup * U < s1) load_r(base + kGroup * U, recn);
#pragma unroll
                    for (int e = 0; e < U; ++e) {
                        if (base + kGroup * e >= s1) break;
                        const int k = (int)__double_as_longlong(rec[e].y);
                        const double2 v0 = Ls[6 * k], v1 = Ls[6 * k + 1], v2 = Ls[6 * k + 2], v3 = Ls[6 * k + 3],
                                      v4 = Ls[6 * k + 4], v5 = Ls[6 * k + 5];
                        const double T[9] = {v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y, v4.x};
                        const double dx[3] = {v4.y - X[0], v5.x - X[1], v5.y - X[2]};
                        term(T, dx, rec[e].x);
                    }
                }
            } else {
                const R *tw = as<R>(p.tshp);
                for (int t = p.trow[node] + lane; t < p.trow[node + 1]; t += kGroup) {
                    const int k = p.tsrc[t];
                    R T[10];
                    load_P<R>(Pb, k, T);
                    const double c = (double)tw[t] * (double)T[9];
                    const double dx[3] = {p.X0[3 * k] - X[0], p.X0[3 * k + 1] - X[1], p.X0[3 * k + 2] - X[2]};
                    const double Td[9] = {(double)T[0], (double)T[1], (double)T[2], (double)T[3], (double)T[4],
                                          (double)T[5], (double)T[6], (double)T[7], (double)T[8]};
                    term(Td, dx, c);
                }
            }
        }
    }
    a0 = group_sum<G>(a0); a1 = group_sum<G>(a1); a2 = group_sum<G>(a2);
    double ke = 0.0, ie = 0.0, ee = 0.0;
    if (live && lane < 3) {
        R T[10];
        load_P<R>(Pb, node, T);
        const double *g = p.gvec + 3 * (int64_t)node;
        const double own = (double)T[lane] * g[0] + (double)T[lane + 3] * g[1] + (double)T[lane + 6] * g[2];
        const double f = (double)T[9] * (own - (lane == 0 ? a0 : (lane == 1 ? a1 : a2)));
        dof_update<CORRECT, PREDICT>(p, node, 3 * (int64_t)node + lane, f, 0.0, ke, ie, ee);
    }
    if (CORRECT) {
        const double sk = block_sum(ke, s_red);
        const double si = block_sum(ie, s_red);
        const double se = block_sum(ee, s_red);
        if (threadIdx.x == 0) {
            p.epart[3 * blk] = sk;
            p.epart[3 * blk + 1] = si;
            p.epart[3 * blk + 2] += se;
        }
    }
}

// ---------------------------------------------------------------------------
// bins: cell of each node (GET_BINS floor/clamp, 1-based), counting sort of
// nodes into cells, node order restored inside each cell.

struct BinGeom {
    double mn[3], bin[3];
    int nb[3];
};

// the reference's bin capacity: GET_BINS stops once a bin holds MAX_NEIGH
// nodes (HANDELER.F90:245, BIN_SEARCH.F90:47-52)
constexpr int kMaxNodesInBin = 1000;

__global__ void k_bin_assign(int n, const double *x, BinGeom g, int32_t *ijk, int32_t *cell_count, int32_t *full)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // clamped to 1..NBINS before the integer conversion, so that a NaN
        // (to 1) or a far-away node stays in the table (BIN_SEARCH.F90:28-41)
        double s = floor((x[3 * i + k] - g.mn[k]) / g.bin[k]) + 1.0;
        s = s >= 1.0 ? s : 1.0;
        s = s <= (double)g.nb[k] ? s : (double)g.nb[k];
        c[k] = (int)s;
        ijk[3 * i + k] = c[k];
    }
    const int64_t m = (int64_t)(c[2] - 1) * g.nb[0] * g.nb[1] + (int64_t)(c[1] - 1) * g.nb[0] + (c[0] - 1);
    if (atomicAdd(&cell_count[m], 1) + 1 >= kMaxNodesInBin) atomicOr(full, 1);
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

// nodes of a cell in the caller's node order (orig: internal -> caller id), as
// GET_BINS appends them (BIN_SEARCH.F90:45-58)
__global__ void k_bin_sort(int64_t ncell, const int32_t *cell_start, const int32_t *orig, int32_t *cell_nodes)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ncell) return;
    const int s = cell_start[m], e = cell_start[m + 1];
    for (int a = s + 1; a < e; ++a) {   // cells hold a few tens of nodes
        const int v = cell_nodes[a], ov = orig[v];
        int b = a - 1;
        while (b >= s && orig[cell_nodes[b]] > ov) { cell_nodes[b + 1] = cell_nodes[b]; --b; }
        cell_nodes[b + 1] = v;
    }
}

// SOFT_SEARCH: one wavefront per non-empty bin.  All nodes of a bin share the
// same candidates -- the nodes of the 27 surrounding bins, visited in the
// reference's order (ix slowest, then iy, iz; BIN_SEARCH.F90:105-127), each
// bin's nodes in node order -- so the wave loads them into registers, 256 at a
// time (4 per lane), and tests every node of its bin against them; kept
// candidates are compacted with a ballot, so each list comes out in the
// reference's order.  Criterion: x_i inside node j's support,
// |x_j - x_i|_k / a_jk < 1 (box) or the ellipsoidal norm < 1
// (BIN_SEARCH.F90:255-291).  Every node belongs to exactly one bin, so one wave
// owns its list.
//
// The same pass builds the transposed lists.  A candidate j stays in one lane
// while the wave runs through its own nodes, so that lane counts how many of
// them keep j.  The transposed row of j is grouped by the source's bin, in the
// order of the 27 bins around j's bin (bin B is entry 26 - c there when j's bin
// is entry c around B), each group in bin order.  tbin[27 j + e] is that
// group's size after the COUNT pass and its first slot before the FILL pass;
// COUNT also keeps each node's ballot masks of its first 64 kMaskCap candidates,
// so FILL reads them instead of testing again.
// FILL writes col from row_ptr[i] on, tpos (transposed slot of each list
// entry) and, when given, tsrc (source of each transposed slot) and tq (list
// position of each transposed slot, for k_tsort).
#ifndef MEGA_CAND_PER_LANE
#define MEGA_CAND_PER_LANE 4
#endif
constexpr int kCandPerLane = MEGA_CAND_PER_LANE;   // candidates per lane in one pass of a bin's wave
#ifndef MEGA_MASK_CAP
#define MEGA_MASK_CAP 8
#endif
constexpr int kMaskCap = MEGA_MASK_CAP;   // ballot masks kept per node by COUNT for FILL: its first 64 kMaskCap candidates

// Transposed rows place their 27 source-bin groups with dz slowest: group g
// holds source bin tgroup(g) of the 27 neighbor offsets (dz fastest, the order
// of the search's counters).  The nine source bins of one z layer, which the
// bin-ordered pair kernels visit close together in time, then write adjacent
// groups of a row.  tgroup reverses the three base-3 digits, so it is its own
// inverse.
__device__ __forceinline__ int tgroup(int g) { return 9 * (g % 3) + 3 * ((g / 3) % 3) + g / 9; }

// Box support, COUNT: each (node, candidate) pair is first decided in FP32
// from positions relative to the wave's bin corner (mega::box_rel,
// box_limits, box_axis: within 4 bins of the corner, windows above 1/50 bin,
// margin 1e-4 of the window); pairs it leaves undecided (near the window's
// edge, far from the corner, tiny windows, non-finite) take the exact FP64
// test.  FILL, which tests only candidates beyond the kept masks, keeps the
// FP64 test (the prefilter's set-up made it slower).

template <bool FILL, bool SHSUP>
__global__ __launch_bounds__(kBlock) void k_search(int ncells, const int32_t *cells, const double *x,
                                                   const double *win, BinGeom g, const int32_t *cell_start,
                                                   const int32_t *cell_nodes, int32_t *counts,
                                                   const int32_t *row_ptr, int32_t *col, int32_t *tbin,
                                                   int32_t *tsrc, int32_t *tpos, int32_t *tq, uint64_t *masks)
{
    __shared__ int s_beg[kBlock / 64][28], s_off[kBlock / 64][28];
    constexpr bool PRE = !SHSUP && !FILL;   // the FP32 prefilter
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int w = (xcd_block() * blockDim.x + threadIdx.x) >> 6;   // neighboring bins on one XCD
    if (w >= ncells) return;   // whole waves leave together
    const int64_t m0 = cells[w];
    const int nxy = g.nb[0] * g.nb[1];
    const int ck = (int)(m0 / nxy) + 1, cj = (int)((m0 % nxy) / g.nb[0]) + 1, ci = (int)(m0 % g.nb[0]) + 1;
    int cnt = 0;
    if (lane < 27) {
        const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
        const int ix = ci + dx, iy = cj + dy, iz = ck + dz;
        int b = 0;
        if (ix >= 1 && ix <= g.nb[0] && iy >= 1 && iy <= g.nb[1] && iz >= 1 && iz <= g.nb[2]) {
            const int64_t m = (int64_t)(iz - 1) * nxy + (int64_t)(iy - 1) * g.nb[0] + (ix - 1);
            b = cell_start[m];
            cnt = cell_start[m + 1] - b;
        }
        s_beg[wv][lane] = b;
    }
    int incl = cnt;   // inclusive prefix of the 27 counts across the wave
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane < 27) s_off[wv][lane] = incl - cnt;
    const int total = __shfl(incl, 26, 64);
    if (lane == 0) s_off[wv][27] = total;
    __builtin_amdgcn_wave_barrier();
    const int own0 = cell_start[m0], nown = cell_start[m0 + 1] - own0;
    // the bin's corner and the FP32 prefilter's reach (box support)
    const double xr[3] = {g.mn[0] + (ci - 1) * g.bin[0], g.mn[1] + (cj - 1) * g.bin[1], g.mn[2] + (ck - 1) * g.bin[2]};
    const float reach[3] = {(float)(4.0 * g.bin[0]), (float)(4.0 * g.bin[1]), (float)(4.0 * g.bin[2])};
    const float wmin[3] = {(float)(0.02 * g.bin[0]), (float)(0.02 * g.bin[1]), (float)(0.02 * g.bin[2])};
    for (int t0 = 0; t0 < nown; t0 += 64) {   // own nodes, 64 at a time: lane t keeps node t0+t's offset
        const int tn = min(64, nown - t0);
        const bool more = t0 + 64 < nown;
        int myi = 0, mine = 0;
        float myq[3] = {0.f, 0.f, 0.f};   // own node relative to the corner (FP32), NaN when out of reach
        if (lane < tn) {
            myi = cell_nodes[own0 + t0 + lane];
            mine = FILL ? row_ptr[myi] : 0;
            if (PRE) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    myq[k] = mega::box_rel(x[3 * (int64_t)myi + k], xr[k], reach[k]);
                }
            }
        }
        // FILL: the lane's candidates only move forward, so their bins' scan resumes
        // where it stopped (COUNT keeps independent scans, which overlap better there)
        int cell_at = 0;
        for (int base = 0; base < total; base += 64 * kCandPerLane) {
            // does this chunk test candidates here (COUNT, or beyond the kept masks)?
            const bool tests = !FILL || (base >> 6) + kCandPerLane > kMaskCap;
            int cj_[kCandPerLane], tix[kCandPerLane], kc[kCandPerLane], kc0[kCandPerLane];
            double cx[kCandPerLane], cy[kCandPerLane], cz[kCandPerLane];
            double wx[kCandPerLane], wy[kCandPerLane], wz[kCandPerLane];
            float pf[kCandPerLane][3], lo[kCandPerLane][3], hi[kCandPerLane][3];   // box prefilter
            // FILL: this pass's kept masks of own node t0 + lane, loaded together before
            // the walk over the own nodes reads them lane by lane; COUNT: collected there
            uint64_t mym[kCandPerLane];
#pragma unroll
            for (int r = 0; r < kCandPerLane; ++r) {
                const int c64 = (base >> 6) + r;
                mym[r] = (FILL && lane < tn && c64 < kMaskCap) ? masks[(int64_t)myi * kMaskCap + c64] : 0ull;
            }
#pragma unroll
            for (int r = 0; r < kCandPerLane; ++r) {
                const int c = base + 64 * r + lane;
                int j = -1, cell = FILL ? cell_at : 0;
                if (c < total) {
                    while (cell < 26 && s_off[wv][cell + 1] <= c) ++cell;
                    j = cell_nodes[s_beg[wv][cell] + (c - s_off[wv][cell])];
                    if (FILL) cell_at = cell;
                }
                cj_[r] = j;
                const int jj = j < 0 ? 0 : j;
                tix[r] = 27 * jj + (26 - cell);
                // this bin's running count (COUNT) or next slot (FILL) in j's transposed row;
                // an earlier chunk of own nodes left it in tbin, written by this same lane
                kc[r] = (j >= 0 && (FILL || t0 > 0)) ? tbin[tix[r]] : 0;
                kc0[r] = kc[r];
                cx[r] = cy[r] = cz[r] = 0.0;
                wx[r] = wy[r] = wz[r] = 1.0;
                if (!PRE) {
                    if (tests) {
                        cx[r] = x[3 * jj]; cy[r] = x[3 * jj + 1]; cz[r] = x[3 * jj + 2];
                        wx[r] = win[3 * jj]; wy[r] = win[3 * jj + 1]; wz[r] = win[3 * jj + 2];
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        pf[r][k] = 0.f; lo[r][k] = -1.f; hi[r][k] = -1.f;   // no candidate: rejected at once
                    }
                    if (tests && j >= 0) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            pf[r][k] = mega::box_rel(x[3 * (int64_t)jj + k], xr[k], reach[k]);
                            mega::box_limits(win[3 * (int64_t)jj + k], wmin[k], lo[r][k], hi[r][k]);
                        }
                    }
                }
            }
            for (int t = 0; t < tn; ++t) {
                const int i = __builtin_amdgcn_readfirstlane(__shfl(myi, t, 64));
                int out = __shfl(mine, t, 64);
                double xi = 0.0, yi = 0.0, zi = 0.0;
                if (tests && !PRE) { xi = x[3 * i]; yi = x[3 * i + 1]; zi = x[3 * i + 2]; }
                float q[3] = {0.f, 0.f, 0.f};
                if (PRE) {
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        q[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq[k]), t));
                }
#pragma unroll
                for (int r = 0; r < kCandPerLane; ++r) {
                    if (base + 64 * r >= total) break;   // uniform
                    const int c64 = (base >> 6) + r;
                    bool add = false;
                    unsigned long long mask;
                    if (FILL && c64 < kMaskCap) {
                        mask = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(mym[r] >> 32), t) << 32) |
                               (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)mym[r], t);
                        add = (mask >> lane) & 1ull;
                    } else {
                        if (!PRE) {
                            if (cj_[r] >= 0) {
                                const double dx = fabs(cx[r] - xi), dy = fabs(cy[r] - yi), dz = fabs(cz[r] - zi);
                                if (SHSUP) {
                                    const double ax = dx / wx[r], ay = dy / wy[r], az = dz / wz[r];
                                    add = sqrt(ax * ax + ay * ay + az * az) < 1.0;
                                } else {   // the box test, exactly as the quotients decide it
                                    add = mega::quotient_below_one(dx, wx[r]) & mega::quotient_below_one(dy, wy[r]) &
                                          mega::quotient_below_one(dz, wz[r]);
                                }
                            }
                        } else {
                            const int a0 = mega::box_axis(pf[r][0], q[0], lo[r][0], hi[r][0]);
                            const int a1 = mega::box_axis(pf[r][1], q[1], lo[r][1], hi[r][1]);
                            const int a2 = mega::box_axis(pf[r][2], q[2], lo[r][2], hi[r][2]);
                            add = (a0 > 0) & (a1 > 0) & (a2 > 0);
                            const bool rej = (a0 < 0) | (a1 < 0) | (a2 < 0);
                            if (!add && !rej && cj_[r] >= 0) {   // the box test, exactly as the quotients decide it
                                const int jj = cj_[r];
                                const double *xj = x + 3 * (int64_t)jj, *wj = win + 3 * (int64_t)jj,
                                             *xv = x + 3 * (int64_t)i;
                                add = mega::quotient_below_one(fabs(xj[0] - xv[0]), wj[0]) &
                                      mega::quotient_below_one(fabs(xj[1] - xv[1]), wj[1]) &
                                      mega::quotient_below_one(fabs(xj[2] - xv[2]), wj[2]);
                            }
                        }
                        mask = __ballot(add);
                        if (!FILL && lane == t) mym[r] = mask;   // own node t's lane keeps it
                    }
                    if (FILL && add) {
                        const int q = out + __popcll(mask & ((1ull << lane) - 1ull));
                        col[q] = cj_[r];
                        tpos[q] = kc[r];
                        if (tsrc) tsrc[kc[r]] = i;
                        if (tq) tq[kc[r]] = q;
                    }
                    kc[r] += add ? 1 : 0;
                    out += __popcll(mask);
                }
                if (lane == t) mine = out;
            }
            // COUNT: each lane stores its own node's masks of this pass together
            if (!FILL && lane < tn) {
#pragma unroll
                for (int r = 0; r < kCandPerLane; ++r) {
                    const int c64 = (base >> 6) + r;
                    if (c64 < kMaskCap && base + 64 * r < total) masks[(int64_t)myi * kMaskCap + c64] = mym[r];
                }
            }
            // COUNT stores a counter only where this bin's nodes kept the candidate (tbin
            // was cleared, and an earlier chunk's value is already there); FILL stores
            // its cursors for a later chunk of own nodes
#pragma unroll
            for (int r = 0; r < kCandPerLane; ++r)
                if (cj_[r] >= 0 && (FILL ? more : kc[r] != kc0[r])) tbin[tix[r]] = kc[r];
        }
        if (!FILL && lane < tn) counts[myi] = mine;
    }
}

// transposed row lengths: sum of the 27 per-bin group sizes
__global__ void k_trow_len(int n, const int32_t *tbin, int32_t *len)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    int s = 0;
    if (j < n)
        for (int e = 0; e < 27; ++e) s += tbin[27 * (int64_t)j + e];
    len[j] = s;
}

// Lists that stay for the whole run (Lagrangian) get their transposed rows
// ordered by source node: the force gather then reads the sources' P in
// increasing address order.  One wave per row ranks its entries in LDS;
// sources within a row are distinct, so an entry's rank is its slot.  tq holds
// each slot's list position, so tpos follows the entries.
constexpr int kTsortMax = 256;

__global__ __launch_bounds__(kBlock) void k_tsort(int n, const int32_t *trow, int32_t *tsrc, int32_t *tq,
                                                  int32_t *tpos)
{
    __shared__ int2 buf[kBlock / 64][kTsortMax];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (j >= n) return;   // whole waves leave together
    const int s = trow[j], L = trow[j + 1] - s;
    if (L > kTsortMax) {  // long row: one lane, insertion sort in place
        if (lane == 0) {
            for (int a = s + 1; a < s + L; ++a) {
                const int vs = tsrc[a], vq = tq[a];
                int b = a - 1;
                while (b >= s && tsrc[b] > vs) { tsrc[b + 1] = tsrc[b]; tq[b + 1] = tq[b]; --b; }
                tsrc[b + 1] = vs;
                tq[b + 1] = vq;
            }
            for (int a = s; a < s + L; ++a) tpos[tq[a]] = a;
        }
        return;
    }
    int2 *row = buf[w];
    for (int a = lane; a < L; a += 64) row[a] = make_int2(tsrc[s + a], tq[s + a]);
    wave_sync();
    for (int a = lane; a < L; a += 64) {
        const int2 v = row[a];
        int r = 0;
        for (int b = 0; b < L; ++b) r += row[b].x < v.x;
        tsrc[s + r] = v.x;
        tpos[v.y] = s + r;
    }
}

// group sizes -> first slot of each group
__global__ void k_tbin_start(int n, const int32_t *trow, int32_t *tbin)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int run = trow[j];
    for (int g = 0; g < 27; ++g) {
        const int e = tgroup(g);
        const int v = tbin[27 * (int64_t)j + e];
        tbin[27 * (int64_t)j + e] = run;
        run += v;
    }
}

// non-empty bins, in bin order
// Per-block source sets of lists that stay for the run.  The pair kernels give
// block b the npb nodes from b * npb; the distinct
// sources of their rows (in Morton node order a small fraction of the
// entries) are staged once per launch in LDS.  One workgroup per block sorts
// the block's entries (bitonic, in LDS) and keeps the distinct values.
// COUNT: the set size, or 0 when the block has more than cap distinct sources
// or more than kSetSort entries (such blocks gather from global memory), and
// the largest size in *maxm.  FILL: the set in ascending order at set[setp[b]]
// and each entry's position in it: loc[q] for forward lists, the src field of
// the transposed record otherwise.
template <bool FILL, class R>
__global__ __launch_bounds__(kBlock) void k_blk_set(int n, int npb, const int32_t *rowp, const int32_t *vals, int cap,
                                                    int32_t *cnt, int32_t *maxm, const int32_t *setp, int32_t *set,
                                                    uint16_t *loc, TPair<R> *tp)
{
    __shared__ int32_t key[kSetSort];
    __shared__ int32_t uq[kSetSort];
    __shared__ int32_t part[kBlock];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int r0 = b * npb, r1 = min(r0 + npb, n);
    const int q0 = rowp[r0], q1 = rowp[r1], len = q1 - q0;
    if (len > kSetSort) {
        if (!FILL && tid == 0) cnt[b] = 0;
        return;
    }
    int np2 = 2 * kBlock;
    while (np2 < len) np2 <<= 1;
    for (int k = tid; k < np2; k += kBlock) key[k] = k < len ? vals[q0 + k] : INT_MAX;
    __syncthreads();
    for (int size = 2; size <= np2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = tid; k < np2 / 2; k += kBlock) {
                const int i = 2 * stride * (k / stride) + (k & (stride - 1)), j = i + stride;
                const int a = key[i], c = key[j];
                if ((a > c) == ((i & size) == 0)) { key[i] = c; key[j] = a; }
            }
            __syncthreads();
        }
    // thread tid owns the sorted chunk [tid C, tid C + C)
    const int C = np2 / kBlock;
    int own = 0;
    for (int k = tid * C; k < tid * C + C; ++k) own += (k < len && (k == 0 || key[k] != key[k - 1])) ? 1 : 0;
    part[tid] = own;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {   // inclusive scan of the chunk counts
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    const int m = part[kBlock - 1];
    const bool fits = m <= cap;
    if (!FILL) {
        if (tid == 0) {
            cnt[b] = fits ? m : 0;
            if (fits) atomicMax(maxm, m);
        }
        return;
    }
    if (!fits) return;
    const int s0 = setp[b];
    int w = part[tid] - own;
    for (int k = tid * C; k < tid * C + C; ++k)
        if (k < len && (k == 0 || key[k] != key[k - 1])) {
            uq[w] = key[k];
            set[s0 + w] = key[k];
            ++w;
        }
    __syncthreads();
    for (int q = q0 + tid; q < q1; q += kBlock) {
        const int v = vals[q];
        int lo = 0, hi = m - 1;   // v is in uq[0, m)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (uq[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        if (loc) loc[q] = (uint16_t)lo;
        else tp[q].src = lo;
    }
}

// {grad phi, phi} of each forward slot in one 16 B record (mixed, lists that stay)
__global__ void k_fpair(int64_t nnz, int64_t cap, const float *shp, const float *dshp, float4 *out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    out[q] = make_float4(dshp[q], dshp[cap + q], dshp[2 * cap + q], shp[q]);
}

// peridynamics (mixed): the 8 B bond records, out[q] = {w V_j, j's position in the
// block's source set (or j, for a block without a set)} for the gather in list
// order and out[cap + t] = {w V_k, k's set position (or k)} for the force in
// transposed order (the set builder has written the position into the
// transposed slot's src); kPdG lanes per particle
__global__ void k_pd_pack8(int n, const int32_t *row_ptr, const int32_t *col, const int32_t *trow, int64_t cap,
                           const float *wv, const uint16_t *loc, const int32_t *gset_ptr, const TPair<float> *tp,
                           const double *qb, float2 *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(t / kPdG), lane = (int)(t % kPdG);
    if (i >= n) return;
    const int b = i / kPdNpb;
    const bool set = gset_ptr[b + 1] > gset_ptr[b];
    const double quantum = qb[b];
    for (int q = row_ptr[i] + lane; q < row_ptr[i + 1]; q += kPdG)
        out[q] = make_float2((float)((double)wv[q] * quantum), __int_as_float(set ? (int)loc[q] : col[q]));
    for (int s = trow[i] + lane; s < trow[i + 1]; s += kPdG)
        out[cap + s] = make_float2((float)((double)tp[s].gx * quantum), __int_as_float(tp[s].src));
}

// peridynamics (FP64): the 16 B bond records, out[q] = {w_q V_j, j's position in
// the block's source set (or j, for a block without a set)} in list order and
// out[cap + t] = {w_t V_k, k's set position (or k)} in transposed order; the
// products are the ones the gather and the force form without records.  kGroup
// lanes per particle, blocks of npb particles
__global__ void k_pd_pack16(int n, int npb, const int32_t *row_ptr, const int32_t *col, const int32_t *trow,
                            const int32_t *tsrc, int64_t cap, const double *w, const double *tw, const double *vol,
                            const uint16_t *loc, const int32_t *gset_ptr, const int32_t *fset_ptr,
                            const TPair<double> *tp, double2 *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(t / kGroup), lane = (int)(t % kGroup);
    if (i >= n) return;
    const int b = i / npb;
    const bool gs = gset_ptr[b + 1] > gset_ptr[b], fs = fset_ptr[b + 1] > fset_ptr[b];
    for (int q = row_ptr[i] + lane; q < row_ptr[i + 1]; q += kGroup) {
        const int j = col[q];
        out[q] = make_double2(w[q] * vol[j], __longlong_as_double((long long)(gs ? (int)loc[q] : j)));
    }
    for (int s = trow[i] + lane; s < trow[i + 1]; s += kGroup) {
        const int k = tsrc[s];
        out[cap + s] = make_double2(tw[s] * vol[k], __longlong_as_double((long long)(fs ? tp[s].src : k)));
    }
}

// per block b, the quantum q_b = R_b / 2^29 with R_b the largest offset
// |X_j - X_b|_inf of a source-set member or a particle of the block from the
// block's first particle (1 for a block of coincident particles); one thread
// block per set
__global__ __launch_bounds__(kBlock) void k_pd_extent(int n, const double *X0, const int32_t *gptr, const int32_t *gset,
                                                      const int32_t *fptr, const int32_t *fset, double *qb)
{
    __shared__ double s_m[kBlock / 64];
    const int b = blockIdx.x, i0 = b * kPdNpb;
    const double xb[3] = {X0[3 * (int64_t)i0], X0[3 * (int64_t)i0 + 1], X0[3 * (int64_t)i0 + 2]};
    double r = 0.0;
    auto take = [&](int j) {
#pragma unroll
        for (int k = 0; k < 3; ++k) r = fmax(r, fabs(X0[3 * (int64_t)j + k] - xb[k]));
    };
    for (int k = gptr[b] + threadIdx.x; k < gptr[b + 1]; k += kBlock) take(gset[k]);
    for (int k = fptr[b] + threadIdx.x; k < fptr[b + 1]; k += kBlock) take(fset[k]);
    if (threadIdx.x < kPdNpb && i0 + (int)threadIdx.x < n) take(i0 + threadIdx.x);
    for (int m = 32; m >= 1; m >>= 1) r = fmax(r, __shfl_xor(r, m, 64));
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) r = fmax(r, s_m[w]);
        qb[b] = r > 0.0 ? r * 0x1p-29 : 1.0;
    }
}

__global__ void k_fpair2(int64_t nnz, int64_t cap, const double *shp, const double *dshp, double2 *out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    out[2 * q] = make_double2(dshp[q], dshp[cap + q]);
    out[2 * q + 1] = make_double2(dshp[2 * cap + q], shp[q]);
}

__global__ void k_cells_ne(int64_t ncell, const int32_t *cell_start, const int32_t *pos, int32_t *cells)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ncell) return;
    if (cell_start[m + 1] > cell_start[m]) cells[pos[m]] = (int32_t)m;
}

__global__ void k_cell_flag(int64_t ncell, const int32_t *cell_start, int32_t *flag)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m > ncell) return;
    flag[m] = m < ncell ? (cell_start[m + 1] > cell_start[m] ? 1 : 0) : 0;
}

// ---------------------------------------------------------------------------
// per-point constitution (context-free entry point), FP64 or FP32 return map

template <class R>
__global__ void k_constitute(int n, const double *lmat, double *stress, double *strain, double *state,
                             Material mt, int finite, int32_t *fail)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double L[9], sig[6], e6[6], de[6], dd[6];
    for (int k = 0; k < 9; ++k) L[k] = lmat[9 * i + k];
    for (int k = 0; k < 6; ++k) { sig[k] = stress[6 * i + k]; e6[k] = strain[6 * i + k]; }
    if (finite) {
        double Rt[9];
        mega::hughes_winget(L, Rt, de, dd);
        mega::rotate_voigt(Rt, sig);
        mega::rotate_voigt(Rt, e6);
    } else {
        de[0] = L[0]; de[1] = L[4]; de[2] = L[8];
        de[3] = L[5] + L[7]; de[4] = L[2] + L[6]; de[5] = L[1] + L[3];
    }
    R sr[6], der[6], pred[6];
    for (int k = 0; k < 6; ++k) { e6[k] = e6[k] + de[k]; sr[k] = (R)sig[k]; der[k] = (R)de[k]; }
    mega::elastic_trial(sr, der, mt, pred);
    R ep = (R)state[20 * i + 2];
    if (mt.type == 2) {
        if (mega::von_mises(sr, pred, ep, mt) < 0) atomicAdd(fail, 1);
    } else if (mt.type == 6 || mt.type == 3) {
        R es = (R)state[20 * i], dam = (R)state[20 * i + 3], dsv = (R)state[20 * i + 4], tot[6];
        for (int k = 0; k < 6; ++k) tot[k] = (R)e6[k];
        if (mt.type == 6) {
            if (mega::von_mises_dam(sr, tot, der, ep, dam, dsv, mt) < 0) atomicAdd(fail, 1);
        } else {
            mega::druck_prag(sr, tot, der, es, ep, dam, dsv, mt);
        }
        state[20 * i] = (double)es; state[20 * i + 3] = (double)dam; state[20 * i + 4] = (double)dsv;
    } else {
        for (int k = 0; k < 6; ++k) sr[k] = pred[k];
    }
    state[20 * i + 2] = (double)ep;
    for (int k = 0; k < 6; ++k) { stress[6 * i + k] = (double)sr[k]; strain[6 * i + k] = e6[k]; }
}

Material make_material(const double *prop, int type)
{
    Material m{};
    m.nu = prop[0]; m.E = prop[1]; m.rho = prop[2]; m.sig0 = prop[3]; m.B = prop[4]; m.CE = prop[5];
    m.damping = prop[20];
    m.lam = prop[27]; m.mu = prop[28]; m.l2m = prop[29];
    m.bulk = m.E / (3.0 * (1.0 - 2.0 * m.nu));
    m.shear = m.E / (2.0 * (1.0 + m.nu));
    m.k_i = prop[8]; m.k_c = prop[9]; m.dam_max = prop[10];
    m.q_phi = prop[24]; m.k_phi = prop[25]; m.q_psi = prop[26];
    m.type = type;
    return m;
}

}  // namespace

// ===========================================================================
// context

struct mega_ctx {
    mega_control c{};
    int n = 0;
    bool mixed = false;
    size_t rsz = 8;   // bytes of R
    // FP64 node arrays
    double *X0 = nullptr, *x = nullptr, *v = nullptr, *a = nullptr, *d = nullptr, *dtot = nullptr, *uphy = nullptr;
    double *fint = nullptr, *fext = nullptr, *fext_nmo = nullptr, *win = nullptr, *vol = nullptr, *mass = nullptr;
    double *winv = nullptr;   // 1 / win, for the box-support kernel
    double *wave_vel = nullptr, *scratch = nullptr;   // scratch: 6n doubles
    float4 *d4 = nullptr, *t4 = nullptr;             // FP32 gather copies of d, dtot (mixed); d4 kept
                                                     // by every predictor, t4 by sync_d4 before the
                                                     // F_DIRECT launches, its only readers
    float4 *u4 = nullptr;                            // FP32 x - X0 (peridynamics, mixed)
    double *Fst = nullptr;                           // deformation gradient per node
    int F_valid = 0;                                 // Fst consistent with dtot
    int32_t *ebc = nullptr, *ebc_node = nullptr, *mat = nullptr;
    Material *mats = nullptr;
    // R arrays
    void *stress = nullptr, *strain = nullptr, *eps = nullptr, *red = nullptr, *P = nullptr, *st4 = nullptr;
    double *vphy_step = nullptr, *aphy_step = nullptr;   // DO_INTERP of v, a at the last step of a call
    void *shp = nullptr, *dshp = nullptr, *tpair = nullptr, *tshp = nullptr;
    // lists
    int32_t *row_ptr = nullptr, *col = nullptr, *trow = nullptr, *tsrc = nullptr, *tpos = nullptr, *tq = nullptr;
    int32_t *tbin = nullptr;   // [27 n] transposed-row group sizes / first slots
    uint64_t *smask = nullptr; // [kMaskCap n] search ballot masks, COUNT -> FILL
    int32_t *counts = nullptr;
    int64_t nnz = 0, nnz_cap = 0;
    // per-block source sets of lists that stay (k_blk_set); gather sets in mixed precision only
    int32_t *gset_ptr = nullptr, *gset = nullptr, *fset_ptr = nullptr, *fset = nullptr, *set_max = nullptr;
    uint16_t *gloc = nullptr;
    int64_t gset_cap = 0, fset_cap = 0, gloc_cap = 0;
    float4 *fpair = nullptr;   // mixed: {grad phi, phi} per forward slot, beside gloc
    int64_t fpair_cap = 0;
    double2 *fpair2 = nullptr; // FP64: {gx, gy}, {gz, phi} per forward slot, beside gloc
    int64_t fpair2_cap = 0;
    int sets_on = 0, gset_max = 0, fset_max = 0;
    // bins
    BinGeom geom{};
    int64_t ncell = 0;
    int32_t *ijk = nullptr, *cell_start = nullptr, *cell_nodes = nullptr, *cell_cursor = nullptr;
    int32_t *cells_ne = nullptr, *cell_flag = nullptr;   // non-empty bins, in bin order
    int ncell_ne = 0;
    // misc
    double *wpart = nullptr, *epart = nullptr;      // per-block work / energy partial sums
    int nb_node = 0, nb_force = 0;
    int32_t *flags = nullptr;
    void *cub_tmp = nullptr;
    size_t cub_tmp_bytes = 0;
    hipStream_t stream = nullptr;
    int64_t step = 0;
    int search_count = 0;
    int has_lists = 0, has_shapes = 0, predicted = 0;
    int dshp_stale = 0;   // the last shapes formed the gather sums and left the list-order gradients unwritten
    double dt = 0.0;
    // automatic time step (peridynamics): this step's DLT, each step's DLT of the
    // call, the characteristic distances and the running minimum
    // smoothed gradients (ITYPE_INT = 1): per node the smoothing box's half lengths and A_k / V
    int snni = 0;
    double *sml = nullptr, *aov = nullptr;
    int auto_on = 0, auto_step = 0;
    double dlt_fac = 0.0;
    double *dt_dev = nullptr, *dt_hist = nullptr, *char_dist = nullptr;
    unsigned long long *dt_min = nullptr;
    int64_t dt_hist_cap = 0;
    int has_gravity = 0;
    int out_dof = 1, out_x = 1;   // see Params
    int mass_node = 0;            // mass held per node (the caller's three per node are equal)
    std::vector<Material> host_mats;
    // learned return map
    int surr_on = 0, surr_mat = -1;
    float surr_tol = 1e-3f;
    float surr_env[2 * mlp::kIn] = {};   // envelope min / max per input (from the normalisation block)
    uint16_t *surr_tiles = nullptr;
    float *surr_bias = nullptr, *surr_norm = nullptr;
    unsigned long long *sstat = nullptr;
    // tuple harvest
    float *hv = nullptr;
    unsigned long long *hv_count = nullptr;
    int64_t hv_cap = 0;
    int hv_every = 0, hv_active = 0;
    // correspondence peridynamics
    int pd = 0, pd_shapes = 0;
    double *pd_qb = nullptr;   // mixed peridynamics: per block, the fixed-point quantum of the bond vectors
    double *Kinv = nullptr, *gvec = nullptr;
    // node numbering: internal node i is the caller's node perm[i]
    std::vector<int32_t> perm, iperm;
    int32_t *orig = nullptr;   // device copy of perm
    // profiling
    int profile = 0;
    struct Timer { const char *name; hipEvent_t a, b; };
    std::vector<Timer> timers;
    size_t timer_used = 0;
};

namespace {

#define HIPCHK(expr)                                                      \
    do {                                                                  \
        hipError_t e_ = (expr);                                           \
        if (e_ != hipSuccess) {                                           \
            std::fprintf(stderr, "mega: %s failed: %s (%s:%d)\n", #expr, \
                         hipGetErrorString(e_), __FILE__, __LINE__);      \
            return MEGA_ERR_HIP;                                          \
        }                                                                 \
    } while (0)

template <class T>
int dalloc(T **p, size_t count, size_t elem = sizeof(T))
{
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, elem * count) != hipSuccess) { *p = nullptr; return MEGA_ERR_NOMEM; }
    return MEGA_OK;
}

Params params(mega_ctx *c)
{
    Params p{};
    p.n = c->n;
    p.cap = c->nnz_cap;
    p.row_ptr = c->row_ptr; p.col = c->col; p.trow = c->trow; p.tsrc = c->tsrc; p.tpos = c->tpos;
    if (c->sets_on) {
        p.fset_ptr = c->fset_ptr; p.fset = c->fset;
        p.gset_ptr = c->gset_ptr; p.gset = c->gset; p.gloc = c->gloc;
        if (!c->mixed && c->pd) p.pd16 = reinterpret_cast<const double2 *>(c->fpair2);
        else if (!c->mixed) p.fpair2 = c->fpair2;
        else if (c->pd) { p.pd8 = reinterpret_cast<const float2 *>(c->fpair); p.pd_qb = c->pd_qb; }
        else p.fpair = c->fpair;
    }
    p.shp = c->shp; p.dshp = c->dshp; p.tpair = c->tpair; p.tshp = c->tshp;
    p.X0 = c->X0; p.win = c->win; p.winv = c->winv; p.vol = c->vol; p.mass = c->mass;
    p.x = c->x; p.v = c->v; p.a = c->a; p.d = c->d; p.dtot = c->dtot; p.uphy = c->uphy;
    p.fint = c->fint; p.fext = c->fext; p.fext_nmo = c->fext_nmo; p.wave_vel = c->wave_vel;
    p.d4 = c->d4; p.t4 = c->t4; p.u4 = c->u4; p.Fst = c->Fst;
    p.stress = c->stress; p.strain = c->strain; p.eps = c->eps; p.red = c->red; p.P = c->P; p.st4 = c->st4;
    p.ebc = c->ebc; p.ebc_node = c->ebc_node; p.mat = c->mat; p.mats = c->mats;
    p.wpart = c->wpart; p.epart = c->epart; p.flags = c->flags;
    p.ql = c->c.ql; p.ql_coef = c->c.ql_coef; p.ql_len = c->c.ql_len;
    p.shsup = c->c.shsup; p.finite_strain = c->c.finite_strain; p.has_gravity = c->has_gravity;
    for (int k = 0; k < 3; ++k) p.grav[k] = c->c.gravity[k];
    p.dt = c->dt;
    p.dt_dev = c->auto_on ? c->dt_dev : nullptr;
    p.char_dist = c->auto_on ? c->char_dist : nullptr;
    p.surr_tiles = c->surr_tiles; p.surr_bias = c->surr_bias; p.surr_norm = c->surr_norm;
    p.surr_mat = c->surr_mat; p.surr_tol = c->surr_tol; p.sstat = c->sstat;
    for (int k = 0; k < 2 * mlp::kIn; ++k) p.surr_env[k] = c->surr_env[k];
    p.hv = c->hv_active ? c->hv : nullptr; p.hv_count = c->hv_count; p.hv_cap = c->hv_cap;
    p.pd = c->pd; p.pd_linit = 0;
    p.Kinv = c->Kinv; p.gvec = c->gvec;
    p.out_dof = c->out_dof; p.out_x = c->out_x; p.mass_node = c->mass_node;
    return p;
}

// event-bracketed launch for the per-kernel timing of mega_step; keeps an
// index into c->timers, since a nested scope may grow (and move) the vector
struct Scope {
    mega_ctx *c;
    size_t idx = SIZE_MAX;
    Scope(mega_ctx *ctx, const char *name) : c(ctx)
    {
        if (!c->profile) return;
        if (c->profile == 2 && std::strcmp(name, "k_force") != 0 && std::strcmp(name, "k_pd_force") != 0) return;
        if (c->timer_used == c->timers.size()) {
            mega_ctx::Timer nt{name, nullptr, nullptr};
            if (hipEventCreate(&nt.a) != hipSuccess || hipEventCreate(&nt.b) != hipSuccess) return;
            c->timers.push_back(nt);
        }
        idx = c->timer_used++;
        c->timers[idx].name = name;
        (void)hipEventRecord(c->timers[idx].a, c->stream);
    }
    ~Scope()
    {
        if (idx != SIZE_MAX) (void)hipEventRecord(c->timers[idx].b, c->stream);
    }
};

unsigned grid_for(int64_t threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }

// grid for the XCD-mapped pair kernels (kBlock/kGroup nodes per block)
unsigned xcd_grid(int n, int *nblocks, int npb = kNodesPerBlock)
{
    *nblocks = (n + npb - 1) / npb;
    const int per = (*nblocks + kXcd - 1) / kXcd;
    return (unsigned)(per * kXcd);
}

void free_pairs(mega_ctx *c)
{
    void *ptrs[] = {c->col, c->shp, c->dshp, c->tpair, c->tshp, c->tsrc, c->tpos, c->tq};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c->col = c->tsrc = c->tpos = c->tq = nullptr;
    c->shp = c->dshp = c->tpair = c->tshp = nullptr;
    c->nnz_cap = 0;
}

int ensure_pairs(mega_ctx *c, int64_t nnz)
{
    if (nnz > INT32_MAX) return MEGA_ERR_ARG;
    if (nnz <= c->nnz_cap) return MEGA_OK;
    free_pairs(c);
    const int64_t cap = nnz + nnz / 8 + 1024;
    const size_t r = c->rsz;
    int e = 0;
    e |= dalloc(&c->col, cap); e |= dalloc(&c->tpos, cap);
    if (c->c.lagrangian) {   // only lists that stay are sorted and feed the peridynamic force
        e |= dalloc(&c->tsrc, cap);
        e |= dalloc(&c->tq, cap);
    }
    e |= dalloc(&c->shp, cap, r); e |= dalloc(&c->dshp, 3 * cap, r);
    e |= dalloc(&c->tpair, cap, c->mixed ? sizeof(TPair<float>) : sizeof(TPair<double>));
    if (c->has_gravity || c->pd) e |= dalloc(&c->tshp, cap, r);
    if (e) { free_pairs(c); return MEGA_ERR_NOMEM; }
    c->nnz_cap = cap;
    c->has_shapes = 0;
    c->pd_shapes = 0;
    return MEGA_OK;
}

int cub_scan(mega_ctx *c, const int32_t *in, int32_t *out, int64_t count)
{
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)count, c->stream));
    if (bytes > c->cub_tmp_bytes) {
        (void)hipFree(c->cub_tmp);
        c->cub_tmp = nullptr;
        if (hipMalloc(&c->cub_tmp, bytes) != hipSuccess) return MEGA_ERR_NOMEM;
        c->cub_tmp_bytes = bytes;
    }
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp, bytes, in, out, (int)count, c->stream));
    return MEGA_OK;
}

int do_bins(mega_ctx *c, const double *x)
{
    const int n = c->n;
    Scope s(c, "k_bin");
    HIPCHK(hipMemsetAsync(c->cell_cursor, 0, sizeof(int32_t) * (c->ncell + 1), c->stream));
    HIPCHK(hipMemsetAsync(c->flags + 3, 0, sizeof(int32_t), c->stream));
    k_bin_assign<<<grid_for(n), kBlock, 0, c->stream>>>(n, x, c->geom, c->ijk, c->cell_cursor, c->flags + 3);
    int r = cub_scan(c, c->cell_cursor, c->cell_start, c->ncell + 1);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(c->cell_cursor, c->cell_start, sizeof(int32_t) * c->ncell, hipMemcpyDeviceToDevice, c->stream));
    k_bin_fill<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->ijk, c->geom, c->cell_cursor, c->cell_nodes);
    k_bin_sort<<<grid_for(c->ncell), kBlock, 0, c->stream>>>(c->ncell, c->cell_start, c->orig, c->cell_nodes);
    k_cell_flag<<<grid_for(c->ncell + 1), kBlock, 0, c->stream>>>(c->ncell, c->cell_start, c->cell_flag);
    r = cub_scan(c, c->cell_flag, c->cell_cursor, c->ncell + 1);   // positions of the non-empty bins
    if (r) return r;
    k_cells_ne<<<grid_for(c->ncell), kBlock, 0, c->stream>>>(c->ncell, c->cell_start, c->cell_cursor, c->cells_ne);
    int32_t nne = 0, full = 0;
    HIPCHK(hipMemcpyAsync(&nne, c->cell_cursor + c->ncell, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&full, c->flags + 3, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ncell_ne = nne;
    HIPCHK(hipGetLastError());
    // a bin at the reference's capacity: it stops here (BIN_SEARCH.F90:48-51); the
    // one-wave-per-bin search would also grow with the square of such a bin
    if (full) return MEGA_ERR_BINS;
    return MEGA_OK;
}

// one wave per non-empty bin, padded to a multiple of kXcd blocks
unsigned search_grid(const mega_ctx *c)
{
    const unsigned nb = grid_for((int64_t)c->ncell_ne * 64);
    return (nb + kXcd - 1) / kXcd * kXcd;
}

int do_search(mega_ctx *c, const double *x)
{
    const int n = c->n;
    HIPCHK(hipMemsetAsync(c->tbin, 0, sizeof(int32_t) * 27 * (size_t)n, c->stream));
    {
        Scope s(c, "k_search_count");
        auto kern = c->c.shsup ? k_search<false, true> : k_search<false, false>;
        kern<<<search_grid(c), kBlock, 0, c->stream>>>(c->ncell_ne, c->cells_ne, x, c->win, c->geom, c->cell_start,
                                                      c->cell_nodes, c->counts, nullptr, nullptr, c->tbin, nullptr,
                                                      nullptr, nullptr, c->smask);
    }
    HIPCHK(hipMemsetAsync(c->counts + n, 0, sizeof(int32_t), c->stream));
    int r = cub_scan(c, c->counts, c->row_ptr, n + 1);
    if (r) return r;
    int32_t nnz32 = 0;
    HIPCHK(hipMemcpyAsync(&nnz32, c->row_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (nnz32 < 0) return MEGA_ERR_BINS;   // list total beyond the int32 index range of the tables
    r = ensure_pairs(c, nnz32);
    if (r) return r;
    c->nnz = nnz32;
    {
        Scope s(c, "k_transpose");
        k_trow_len<<<grid_for(n + 1), kBlock, 0, c->stream>>>(n, c->tbin, c->counts);
        r = cub_scan(c, c->counts, c->trow, n + 1);
        if (r) return r;
        k_tbin_start<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->trow, c->tbin);
    }
    {
        Scope s(c, "k_search_fill");
        auto kern = c->c.shsup ? k_search<true, true> : k_search<true, false>;
        kern<<<search_grid(c), kBlock, 0, c->stream>>>(c->ncell_ne, c->cells_ne, x, c->win, c->geom, c->cell_start,
                                                      c->cell_nodes, nullptr, c->row_ptr, c->col, c->tbin, c->tsrc,
                                                      c->tpos, c->tq, c->smask);
    }
    if (c->tq) {
        Scope s(c, "k_tsort");
        k_tsort<<<grid_for((int64_t)n * 64), kBlock, 0, c->stream>>>(n, c->trow, c->tsrc, c->tq, c->tpos);
    }
    HIPCHK(hipGetLastError());
    c->has_lists = 1;
    c->has_shapes = 0;
    c->sets_on = 0;
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

// one list's sets: rowp/vals the CSR, loc (forward) or the records (transposed)
template <class R>
int build_set(mega_ctx *c, int npb, const int32_t *rowp, const int32_t *vals, int32_t *setp, int32_t **set,
              int64_t *set_cap, uint16_t *loc, TPair<R> *tp, int *maxm, int cap = kSetCap)
{
    const int nb = (c->n + npb - 1) / npb;
    HIPCHK(hipMemsetAsync(c->set_max, 0, sizeof(int32_t), c->stream));
    HIPCHK(hipMemsetAsync(c->counts + nb, 0, sizeof(int32_t), c->stream));
    k_blk_set<false, R><<<nb, kBlock, 0, c->stream>>>(c->n, npb, rowp, vals, cap, c->counts, c->set_max,
                                                      nullptr, nullptr, nullptr, nullptr);
    int r = cub_scan(c, c->counts, setp, nb + 1);
    if (r) return r;
    int32_t tot = 0, mx = 0;
    HIPCHK(hipMemcpyAsync(&tot, setp + nb, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&mx, c->set_max, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (tot > *set_cap) {
        (void)hipFree(*set);
        *set = nullptr;
        *set_cap = 0;
        if (dalloc(set, tot)) return MEGA_ERR_NOMEM;
        *set_cap = tot;
    }
    k_blk_set<true, R><<<nb, kBlock, 0, c->stream>>>(c->n, npb, rowp, vals, cap, nullptr, nullptr, setp, *set,
                                                     loc, tp);
    HIPCHK(hipGetLastError());
    *maxm = mx;
    return MEGA_OK;
}

// source sets of lists that stay for the run (the Lagrangian RK lists): the
// transposed sets for the force and the forward sets for the gather, with the
// gather's packed {grad phi, phi} records; after the records are written, as
// FILL rewrites their src field
int build_sets(mega_ctx *c)
{
    c->sets_on = 0;
    if (!c->tq || c->pd || c->nnz == 0) return MEGA_OK;
    const int nbf = (c->n + kNodesPerFBlock - 1) / kNodesPerFBlock;
    const int nbg = (c->n + gather_npb(true) - 1) / gather_npb(true);
    int e = 0;
    if (!c->fset_ptr) e |= dalloc(&c->fset_ptr, nbf + 1);
    if (!c->set_max) e |= dalloc(&c->set_max, 1);
    if (!c->gset_ptr) e |= dalloc(&c->gset_ptr, nbg + 1);
    if (c->gloc_cap < c->nnz) {
        (void)hipFree(c->gloc);
        c->gloc = nullptr;
        c->gloc_cap = 0;
        e |= dalloc(&c->gloc, c->nnz_cap);
        if (!e) c->gloc_cap = c->nnz_cap;
    }
    if (c->mixed && c->fpair_cap < c->nnz) {
        (void)hipFree(c->fpair);
        c->fpair = nullptr;
        c->fpair_cap = 0;
        e |= dalloc(&c->fpair, c->nnz_cap);
        if (!e) c->fpair_cap = c->nnz_cap;
    }
    if (!c->mixed && c->fpair2_cap < c->nnz) {
        (void)hipFree(c->fpair2);
        c->fpair2 = nullptr;
        c->fpair2_cap = 0;
        e |= dalloc(&c->fpair2, 2 * (size_t)c->nnz_cap);
        if (!e) c->fpair2_cap = c->nnz_cap;
    }
    if (e) return MEGA_ERR_NOMEM;
    Scope s(c, "k_blk_set");
    int r;
    if (c->mixed) {
        r = build_set<float>(c, kNodesPerFBlock, c->trow, c->tsrc, c->fset_ptr, &c->fset, &c->fset_cap, nullptr,
                             static_cast<TPair<float> *>(c->tpair), &c->fset_max);
        if (!r)
            r = build_set<float>(c, gather_npb(true), c->row_ptr, c->col, c->gset_ptr, &c->gset, &c->gset_cap, c->gloc,
                                 nullptr, &c->gset_max);
        if (!r) {
            k_fpair<<<grid_for(c->nnz), kBlock, 0, c->stream>>>(c->nnz, c->nnz_cap, (const float *)c->shp,
                                                               (const float *)c->dshp, c->fpair);
            HIPCHK(hipGetLastError());
        }
    } else {
        r = build_set<double>(c, kNodesPerFBlock, c->trow, c->tsrc, c->fset_ptr, &c->fset, &c->fset_cap, nullptr,
                              static_cast<TPair<double> *>(c->tpair), &c->fset_max);
        if (!r)
            r = build_set<double>(c, gather_npb(true), c->row_ptr, c->col, c->gset_ptr, &c->gset, &c->gset_cap,
                                  c->gloc, nullptr, &c->gset_max);
        if (!r) {
            k_fpair2<<<grid_for(c->nnz), kBlock, 0, c->stream>>>(c->nnz, c->nnz_cap, (const double *)c->shp,
                                                                (const double *)c->dshp, c->fpair2);
            HIPCHK(hipGetLastError());
        }
    }
    if (r) return r;
    c->sets_on = 1;
    return MEGA_OK;
}

// source sets of the peridynamic bond lists (mixed; the lists stay for the
// run): the force's transposed records get their source's set position (as
// build_sets does for the RK force), the gather's records are packed with
// theirs.  Built after k_pd_shape has written the records.
int build_pd_sets(mega_ctx *c)
{
    c->sets_on = 0;
    if (!c->pd || !c->tsrc || c->nnz == 0) return MEGA_OK;
    const int npb = c->mixed ? kPdNpb : kNodesPerBlock;
    const int nb = (c->n + npb - 1) / npb;
    int e = 0;
    if (!c->fset_ptr) e |= dalloc(&c->fset_ptr, nb + 1);
    if (!c->set_max) e |= dalloc(&c->set_max, 1);
    if (!c->gset_ptr) e |= dalloc(&c->gset_ptr, nb + 1);
    if (c->mixed && !c->pd_qb) e |= dalloc(&c->pd_qb, nb);
    if (c->gloc_cap < c->nnz) {
        (void)hipFree(c->gloc);
        c->gloc = nullptr;
        c->gloc_cap = 0;
        e |= dalloc(&c->gloc, c->nnz_cap);
        if (!e) c->gloc_cap = c->nnz_cap;
    }
    if (c->mixed && c->fpair_cap < c->nnz) {
        (void)hipFree(c->fpair);
        c->fpair = nullptr;
        c->fpair_cap = 0;
        e |= dalloc(&c->fpair, c->nnz_cap);
        if (!e) c->fpair_cap = c->nnz_cap;
    }
    if (!c->mixed && c->fpair2_cap < c->nnz) {
        (void)hipFree(c->fpair2);
        c->fpair2 = nullptr;
        c->fpair2_cap = 0;
        e |= dalloc(&c->fpair2, 2 * (size_t)c->nnz_cap);
        if (!e) c->fpair2_cap = c->nnz_cap;
    }
    if (e) return MEGA_ERR_NOMEM;
    Scope s(c, "k_blk_set");
    if (!c->mixed) {
        int r = build_set<double>(c, npb, c->trow, c->tsrc, c->fset_ptr, &c->fset, &c->fset_cap, nullptr,
                                  static_cast<TPair<double> *>(c->tpair), &c->fset_max, kPdFSetCap64);
        if (!r)
            r = build_set<double>(c, npb, c->row_ptr, c->col, c->gset_ptr, &c->gset, &c->gset_cap, c->gloc, nullptr,
                                  &c->gset_max, kPdSetCap64);
        if (r) return r;
        k_pd_pack16<<<grid_for((int64_t)c->n * kGroup), kBlock, 0, c->stream>>>(
            c->n, npb, c->row_ptr, c->col, c->trow, c->tsrc, c->nnz_cap, (const double *)c->shp,
            (const double *)c->tshp, c->vol, c->gloc, c->gset_ptr, c->fset_ptr,
            static_cast<const TPair<double> *>(c->tpair), reinterpret_cast<double2 *>(c->fpair2));
        HIPCHK(hipGetLastError());
        c->sets_on = 1;
        return MEGA_OK;
    }
    int r = build_set<float>(c, npb, c->trow, c->tsrc, c->fset_ptr, &c->fset, &c->fset_cap, nullptr,
                             static_cast<TPair<float> *>(c->tpair), &c->fset_max, kPdSetCap);
    if (!r)
        r = build_set<float>(c, npb, c->row_ptr, c->col, c->gset_ptr, &c->gset, &c->gset_cap, c->gloc, nullptr,
                             &c->gset_max, kPdSetCap);
    if (r) return r;
    // the fixed-point quantum of each block's bond vectors: offsets below 2^29 units
    k_pd_extent<<<nb, kBlock, 0, c->stream>>>(c->n, c->X0, c->gset_ptr, c->gset, c->fset_ptr, c->fset, c->pd_qb);
    HIPCHK(hipGetLastError());
    k_pd_pack8<<<grid_for((int64_t)c->n * kPdG), kBlock, 0, c->stream>>>(
        c->n, c->row_ptr, c->col, c->trow, c->nnz_cap, (const float *)c->dshp, c->gloc, c->gset_ptr,
        static_cast<const TPair<float> *>(c->tpair), c->pd_qb, reinterpret_cast<float2 *>(c->fpair));
    HIPCHK(hipGetLastError());
    c->sets_on = 1;
    return MEGA_OK;
}

// gather: also form the gather sums of do_fint(F_NONE) (then called with
// gathered = true), without storing the list-order gradients
int do_shapes(mega_ctx *c, const double *xs, bool gather)
{
    if (!c->has_lists) return MEGA_ERR_ARG;
    Params p = params(c);
    {
        Scope s(c, "k_rk");
        const int nb = (int)(((int64_t)c->n * kRkGroup + kBlock - 1) / kBlock);
        const unsigned grid = (unsigned)((nb + kXcd - 1) / kXcd * kXcd);
        if (c->mixed) {
            if (gather) k_rk<float, true><<<grid, kBlock, 0, c->stream>>>(p, xs, c->cell_nodes, nb);
            else k_rk<float, false><<<grid, kBlock, 0, c->stream>>>(p, xs, c->cell_nodes, nb);
        } else {
            if (gather) k_rk<double, true><<<grid, kBlock, 0, c->stream>>>(p, xs, c->cell_nodes, nb);
            else k_rk<double, false><<<grid, kBlock, 0, c->stream>>>(p, xs, c->cell_nodes, nb);
        }
        HIPCHK(hipGetLastError());
    }
    if (c->snni && !gather) {
        Scope s(c, "k_snni");
        const int nb = (int)(((int64_t)c->n * kRkGroup + kBlock - 1) / kBlock);
        const unsigned grid = (unsigned)((nb + kXcd - 1) / kXcd * kXcd);
        if (c->mixed) k_snni<float><<<grid, kBlock, 0, c->stream>>>(p, xs, c->sml, c->aov, nb);
        else k_snni<double><<<grid, kBlock, 0, c->stream>>>(p, xs, c->sml, c->aov, nb);
        HIPCHK(hipGetLastError());
    }
    const int r = build_sets(c);
    if (r) return r;
    c->has_shapes = 1;
    c->dshp_stale = gather ? 1 : 0;
    return MEGA_OK;
}

template <class R, int FM>
int fint_kernels_fm(mega_ctx *c, bool update_x, int force_mode, bool gathered)
{
    Params p = params(c);
    int nb = 0;
    if (!gathered) {
        Scope s(c, "k_gather");
        bool done = false;
        if (p.gset_ptr) {   // lists that stay: source sets, packed records
            const unsigned g = xcd_grid(c->n, &nb, gather_npb(true));
            const size_t row = std::is_same<R, float>::value ? 16 : 32;   // d (and dtot) per source
            const size_t lds = (size_t)c->gset_max * (FM == F_DIRECT ? 2 * row : row);
            if (update_x) k_gather<R, FM, true, true><<<g, kBlock, lds, c->stream>>>(p, nb);
            else k_gather<R, FM, false, true><<<g, kBlock, lds, c->stream>>>(p, nb);
            done = true;
        }
        if (!done) {
            const unsigned g = xcd_grid(c->n, &nb, gather_npb(false));
            if (update_x) k_gather<R, FM, true, false><<<g, kBlock, 0, c->stream>>>(p, nb);
            else k_gather<R, FM, false, false><<<g, kBlock, 0, c->stream>>>(p, nb);
        }
    }
    {
        Scope s(c, "k_node");
        if constexpr (std::is_same<R, float>::value) {
            if (c->surr_on) k_node_surr<FM><<<grid_for(c->n), kBlock, 0, c->stream>>>(p);
            else k_node<R, FM><<<grid_for(c->n), kBlock, 0, c->stream>>>(p);
        } else {
            k_node<R, FM><<<grid_for(c->n), kBlock, 0, c->stream>>>(p);
        }
    }
    if (force_mode >= 0) {
        Scope s(c, "k_force");
        int nbf = 0;
        const unsigned gf = xcd_grid(c->n, &nbf, kNodesPerFBlock);
        const size_t lds = p.fset_ptr ? (size_t)c->fset_max * (std::is_same<R, float>::value ? 40 : 80) : 0;
        if (force_mode == 0) k_force<R, false, false><<<gf, kBlock, lds, c->stream>>>(p, nbf);
        else if (force_mode == 1) k_force<R, true, false><<<gf, kBlock, lds, c->stream>>>(p, nbf);
        else k_force<R, true, true><<<gf, kBlock, lds, c->stream>>>(p, nbf);
    }
    HIPCHK(hipGetLastError());
    if (FM == F_DIRECT) c->F_valid = 1;
    return MEGA_OK;
}

template <class R>
int fint_kernels(mega_ctx *c, int fm, bool update_x, int force_mode, bool gathered)
{
    if (fm == F_DIRECT) return fint_kernels_fm<R, F_DIRECT>(c, update_x, force_mode, false);
    if (fm == F_INCR) return fint_kernels_fm<R, F_INCR>(c, update_x, force_mode, false);
    return fint_kernels_fm<R, F_NONE>(c, update_x, force_mode, gathered);
}

// fm: F_NONE (semi-Lagrangian / small strain), F_DIRECT (F from dtot), F_INCR
// (F carried per node); force_mode: -1 no force, 0 force only, 1 force +
// corrector, 2 + next predictor; gathered: the gather sums are already formed
// (do_shapes with gather, F_NONE without the coordinate update only)
int do_fint(mega_ctx *c, int fm, bool update_x, int force_mode, bool gathered = false)
{
    if (!c->has_shapes) return MEGA_ERR_ARG;
    if (fm == F_INCR && !c->F_valid) fm = F_DIRECT;
    if (gathered && (fm != F_NONE || update_x)) return MEGA_ERR_ARG;
    return c->mixed ? fint_kernels<float>(c, fm, update_x, force_mode, gathered)
                    : fint_kernels<double>(c, fm, update_x, force_mode, gathered);
}

// UDFM_SHAPE_TENSOR for all nodes (FINT_PD.F90:227-290)
int do_pd_shapes(mega_ctx *c)
{
    if (!c->has_lists) return MEGA_ERR_ARG;
    Params p = params(c);
    Scope s(c, "k_pd_shape");
    if (c->mixed) k_pd_shape<float><<<grid_for((int64_t)c->n * kGroup), kBlock, 0, c->stream>>>(p);
    else k_pd_shape<double><<<grid_for((int64_t)c->n * kGroup), kBlock, 0, c->stream>>>(p);
    HIPCHK(hipGetLastError());
    const int r = build_pd_sets(c);
    if (r) return r;
    c->pd_shapes = 1;
    return MEGA_OK;
}

template <class R>
int pd_kernels(mega_ctx *c, int linit, int force_mode)
{
    constexpr bool MIX = std::is_same<R, float>::value;
    Params p = params(c);
    p.pd_linit = linit;
    int nb = 0;
    const unsigned g = xcd_grid(c->n, &nb, MIX ? kPdNpb : kNodesPerBlock);
    // mixed with source sets: u, d and the relative reference position (36 B) per gather
    // source, T rows and the relative reference position (48 B) per force source (+ 16 B
    // per gather source with the automatic time step: FP32 current positions)
    const size_t mg = (size_t)c->gset_max;
    size_t lds_g = (MIX && p.gset_ptr && p.pd8) ? 32 * mg + 16 * ((mg + 3) / 4) + (p.char_dist ? 16 * mg : 0) : 0;
    size_t lds_f = (MIX && p.fset_ptr && p.pd8) ? (size_t)c->fset_max * 48 : 0;
    if (!MIX && p.pd16) {   // FP64: 80 B gather rows, 96 B force rows
        lds_g = 80 * mg;
        lds_f = 96 * (size_t)c->fset_max;
    }
    {
        Scope s(c, "k_pd_gather");
        if (c->mixed) k_pd_u4<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->x, c->X0, c->u4);
        k_pd_gather<R><<<g, kBlock, lds_g, c->stream>>>(p, nb);
    }
    {
        Scope s(c, "k_pd_node");
        k_pd_node<R><<<grid_for(c->n), kBlock, 0, c->stream>>>(p);
    }
    if (c->auto_on && force_mode >= 1) {   // this step's DLT before the corrector uses it
        Scope s(c, "k_pd_dt");
        k_pd_dt<<<std::min<unsigned>(grid_for(c->n), kDtBlocks), kBlock, 0, c->stream>>>(c->n, c->char_dist,
                                                                                           c->wave_vel, c->dt_min);
        k_pd_dt_fin<<<1, 64, 0, c->stream>>>(c->dt_min, c->dlt_fac, c->dt_dev, c->dt_hist, c->auto_step);
    }
    if (force_mode >= 0) {
        Scope s(c, "k_pd_force");
        if (force_mode == 0) k_pd_force<R, false, false><<<g, kBlock, lds_f, c->stream>>>(p, nb);
        else if (force_mode == 1) k_pd_force<R, true, false><<<g, kBlock, lds_f, c->stream>>>(p, nb);
        else k_pd_force<R, true, true><<<g, kBlock, lds_f, c->stream>>>(p, nb);
    }
    HIPCHK(hipGetLastError());
    return MEGA_OK;
}

// CONSTRUCT_FINT_PD on the device state; force_mode as do_fint
int do_pd_fint(mega_ctx *c, int linit, int force_mode)
{
    if (!c->pd_shapes) return MEGA_ERR_ARG;
    return c->mixed ? pd_kernels<float>(c, linit, force_mode) : pd_kernels<double>(c, linit, force_mode);
}

// the learned return map's counters summed over the k_node blocks
int surrogate_totals(mega_ctx *c, int64_t *stats, int reset)
{
    const size_t m = 4 * (size_t)c->nb_node;
    std::vector<unsigned long long> h(m);
    HIPCHK(hipMemcpyAsync(h.data(), c->sstat, sizeof(unsigned long long) * m, hipMemcpyDeviceToHost, c->stream));
    if (reset) HIPCHK(hipMemsetAsync(c->sstat, 0, sizeof(unsigned long long) * m, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; ++k) {
        unsigned long long t = 0;
        for (size_t b = 0; b < (size_t)c->nb_node; ++b) t += h[4 * b + k];
        stats[k] = (int64_t)t;
    }
    return MEGA_OK;
}

double host_sum(mega_ctx *c, const double *dev, int64_t count, int64_t stride, int64_t offset, int *err)
{
    std::vector<double> h((size_t)count * stride);
    if (hipMemcpyAsync(h.data(), dev, sizeof(double) * count * stride, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        *err = MEGA_ERR_HIP;
        return 0.0;
    }
    double s = 0.0;
    for (int64_t b = 0; b < count; ++b) s += h[(size_t)(b * stride + offset)];
    return s;
}

int do_predict(mega_ctx *c)
{
    Params p = params(c);
    Scope s(c, "k_predict");
    k_predict<<<grid_for(3 * (int64_t)c->n), kBlock, 0, c->stream>>>(p);
    HIPCHK(hipGetLastError());
    return MEGA_OK;
}

int sync_d4(mega_ctx *c)
{
    if (!c->mixed) return MEGA_OK;
    k_cast4<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->d, c->d4);
    k_cast4<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->dtot, c->t4);
    HIPCHK(hipGetLastError());
    return MEGA_OK;
}

int check_flags(mega_ctx *c)
{
    int32_t f[3];
    HIPCHK(hipMemcpyAsync(f, c->flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (f[1]) return MEGA_ERR_MATERIAL;
    if (f[0]) return MEGA_ERR_RETURN_MAP;
    if (f[2]) return MEGA_ERR_SINGULAR;   // a warning, as INVERSE.F90:93-95
    return MEGA_OK;
}

// copy an R-typed state array (m values) to/from FP64 host memory
int state_out(mega_ctx *c, const void *src, int64_t m, double *host)
{
    if (!c->mixed) {
        HIPCHK(hipMemcpyAsync(host, src, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    } else {
        k_widen<float><<<grid_for(m), kBlock, 0, c->stream>>>(m, (const float *)src, c->scratch);
        HIPCHK(hipMemcpyAsync(host, c->scratch, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    }
    return MEGA_OK;
}

int state_in(mega_ctx *c, void *dst, int64_t m, const double *host)
{
    if (!c->mixed) {
        HIPCHK(hipMemcpyAsync(dst, host, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(c->scratch, host, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
        k_narrow<float><<<grid_for(m), kBlock, 0, c->stream>>>(m, c->scratch, (float *)dst);
    }
    return MEGA_OK;
}

// Nodes are numbered internally in Morton (Z-curve) order of their initial
// positions on a grid of about one node spacing, so that consecutive internal
// ids -- one block of a pair kernel, one XCD's range -- are compact in space
// and share their neighbors in cache.  Arrays at the C-ABI stay in the
// caller's order; these move them across.
uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

std::vector<int32_t> morton_order(int n, const double *X0, const double *vol)
{
    double vsum = 0.0, mn[3] = {X0[0], X0[1], X0[2]};
    for (int i = 0; i < n; ++i) {
        vsum += vol[i];
        for (int k = 0; k < 3; ++k) mn[k] = std::min(mn[k], X0[3 * (int64_t)i + k]);
    }
    double h = std::cbrt(vsum / n);
    if (!(h > 0.0) || !std::isfinite(h)) h = 1.0;
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (int i = 0; i < n; ++i) {
        uint64_t code = 0;
        for (int k = 0; k < 3; ++k) {
            double q = std::floor((X0[3 * (int64_t)i + k] - mn[k]) / h + 0.5);
            q = std::min(std::max(q, 0.0), 2097151.0);
            code |= spread3((uint64_t)q) << k;
        }
        key[i] = {code, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int32_t> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = key[i].second;
    return perm;
}

// list-order pair values (w per pair) of the internal rows, written out row by
// row in the caller's node order
int rows_to_caller(mega_ctx *c, const double *in, double *out, int w)
{
    std::vector<int32_t> rp(c->n + 1);
    if (hipMemcpyAsync(rp.data(), c->row_ptr, sizeof(int32_t) * (c->n + 1), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return MEGA_ERR_HIP;
    int64_t q = 0;
    for (int o = 0; o < c->n; ++o) {
        const int i = c->iperm[o];
        for (int t = rp[i]; t < rp[i + 1]; ++t, ++q)
            for (int k = 0; k < w; ++k) out[(size_t)w * q + k] = in[(size_t)w * t + k];
    }
    return MEGA_OK;
}

// internal-order copy of a caller-order host array (w values per node)
template <class T>
std::vector<T> to_internal(const mega_ctx *c, const T *h, int w)
{
    std::vector<T> o((size_t)c->n * w);
    for (int i = 0; i < c->n; ++i)
        for (int k = 0; k < w; ++k) o[(size_t)w * i + k] = h[(size_t)w * c->perm[i] + k];
    return o;
}

// in place: h holds internal order on entry, the caller's on return
template <class T>
void to_caller(const mega_ctx *c, T *h, int w)
{
    std::vector<T> t(h, h + (size_t)c->n * w);
    for (int i = 0; i < c->n; ++i)
        for (int k = 0; k < w; ++k) h[(size_t)w * c->perm[i] + k] = t[(size_t)w * i + k];
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
    case MEGA_ERR_BINS: return "a bin holds 1000 nodes, or the bin grid or list total exceeds int32";
    case MEGA_ERR_NOMEM: return "device allocation failed";
    case MEGA_ERR_HIP: return "HIP runtime error";
    case MEGA_ERR_ARG: return "invalid argument or unsupported option";
    case MEGA_ERR_MATERIAL: return "invalid material type on this path";
    case MEGA_ERR_SCRATCH: return "input larger than the fixed scratch of a no-allocation entry point";
    default: return "unknown status";
    }
}

int mega_create(const mega_control *ctrl, int32_t n, const double *X0, const double *win, const double *vol,
                const double *mass, const int32_t *ebc, const int32_t *ebc_node, const int32_t *mat_type,
                const double *prop, const double *vinit, mega_ctx **out)
{
    if (!ctrl || !out || n <= 0 || !X0 || !win || !vol || !mass || !ebc || !ebc_node || !mat_type || !prop)
        return MEGA_ERR_ARG;
    if (ctrl->rk_degree != 1 || ctrl->rk_cont != 3) return MEGA_ERR_ARG;
    // ITYPE_INT: 0 direct nodal integration; 1 the smoothed (stabilized non-conforming)
    // gradients, an extension on RK steps (the reference's is commented out)
    if (ctrl->itype_int != 0 && !(ctrl->itype_int == 1 && !ctrl->peridynamics)) return MEGA_ERR_ARG;
    if (ctrl->precision != MEGA_FP64 && ctrl->precision != MEGA_MIXED) return MEGA_ERR_ARG;
    if (ctrl->surrogate && ctrl->precision != MEGA_MIXED) return MEGA_ERR_ARG;
    // the reference's CONSTRUCT_FINT_PD is Lagrangian only (FINT_PD.F90:408-412)
    if (ctrl->peridynamics && (!ctrl->lagrangian || ctrl->surrogate)) return MEGA_ERR_ARG;
    // the search's per-source-bin counters (27 per node) are indexed in int32
    if ((int64_t)n * 27 > INT32_MAX) return MEGA_ERR_BINS;
    *out = nullptr;
    mega_ctx *c = new mega_ctx();
    c->c = *ctrl;
    if (c->c.pdsearch < 1) c->c.pdsearch = 1;
    c->n = n;
    c->mixed = ctrl->precision == MEGA_MIXED;
    c->rsz = c->mixed ? 4 : 8;
    c->pd = ctrl->peridynamics ? 1 : 0;
    c->has_gravity = (ctrl->gravity[0] != 0.0 || ctrl->gravity[1] != 0.0 || ctrl->gravity[2] != 0.0);
    c->perm = morton_order(n, X0, vol);
    c->iperm.assign(n, 0);
    for (int i = 0; i < n; ++i) c->iperm[c->perm[i]] = i;
    // materials: nodes with identical property rows share a slot
    std::vector<int32_t> mid(n);
    std::map<std::vector<double>, int> seen;
    for (int i = 0; i < n; ++i) {
        std::vector<double> row(prop + 30 * (int64_t)i, prop + 30 * (int64_t)i + 30);
        row.push_back((double)mat_type[i]);
        auto it = seen.find(row);
        if (it == seen.end()) {
            const int k = (int)c->host_mats.size();
            seen.emplace(row, k);
            c->host_mats.push_back(make_material(prop + 30 * (int64_t)i, mat_type[i]));
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
    if (c->ncell <= 0 || c->ncell >= INT32_MAX) { delete c; return MEGA_ERR_BINS; }   // bin grid beyond int32

    const size_t n3 = 3 * (size_t)n, n6 = 6 * (size_t)n, r = c->rsz;
    int e = 0;
    e |= dalloc(&c->X0, n3); e |= dalloc(&c->x, n3); e |= dalloc(&c->v, n3); e |= dalloc(&c->a, n3);
    e |= dalloc(&c->d, n3); e |= dalloc(&c->dtot, n3); e |= dalloc(&c->uphy, n3); e |= dalloc(&c->fint, n3);
    e |= dalloc(&c->fext, n3); e |= dalloc(&c->scratch, n6);
    if (c->has_gravity) e |= dalloc(&c->fext_nmo, n3);
    if (c->mixed) { e |= dalloc(&c->d4, n); e |= dalloc(&c->t4, n); }
    if (c->mixed && c->pd) e |= dalloc(&c->u4, n);
    // wave speeds: FP64, and peridynamics in both precisions (the automatic time step)
    if (!c->mixed || c->pd) e |= dalloc(&c->wave_vel, n);
    e |= dalloc(&c->Fst, 10 * (size_t)n);
    c->snni = ctrl->itype_int == 1;
    if (c->snni) { e |= dalloc(&c->sml, n3); e |= dalloc(&c->aov, n3); }
    if (c->pd) { e |= dalloc(&c->Kinv, 9 * (size_t)n); e |= dalloc(&c->gvec, 3 * (size_t)n); }
    c->mass_node = 1;   // LOCAL_MASS repeats each node's mass over its three dofs (STATE_INIT.f90:111-113)
    for (int i = 0; i < n && c->mass_node; ++i)
        c->mass_node = mass[3 * (int64_t)i] == mass[3 * (int64_t)i + 1] && mass[3 * (int64_t)i] == mass[3 * (int64_t)i + 2];
    e |= dalloc(&c->win, n3); e |= dalloc(&c->winv, n3); e |= dalloc(&c->vol, n);
    e |= dalloc(&c->mass, c->mass_node ? (size_t)n : n3);
    e |= dalloc(&c->ebc, n3); e |= dalloc(&c->ebc_node, n); e |= dalloc(&c->mat, n);
    e |= dalloc(&c->mats, c->host_mats.size());
    e |= dalloc(&c->stress, n6, r); e |= dalloc(&c->strain, n6, r); e |= dalloc(&c->eps, n, r);
    e |= dalloc(&c->st4, 4 * (size_t)n, r);
    e |= dalloc(&c->vphy_step, n3); e |= dalloc(&c->aphy_step, n3);
    e |= dalloc(&c->red, 24 * (size_t)n, r); e |= dalloc(&c->P, (c->mixed ? 12 : 10) * (size_t)n, r);
    e |= dalloc(&c->row_ptr, n + 1); e |= dalloc(&c->trow, n + 1); e |= dalloc(&c->counts, n + 1);
    e |= dalloc(&c->ijk, n3); e |= dalloc(&c->cell_start, c->ncell + 1); e |= dalloc(&c->cell_cursor, c->ncell + 1);
    e |= dalloc(&c->cell_nodes, n);
    e |= dalloc(&c->orig, n);
    e |= dalloc(&c->cells_ne, std::min<int64_t>(c->ncell, n));
    e |= dalloc(&c->tbin, 27 * (int64_t)n);
    e |= dalloc(&c->smask, kMaskCap * (int64_t)n);
    e |= dalloc(&c->cell_flag, c->ncell + 1);
    c->nb_node = (int)grid_for(n);
    {   // energy partials: one per block of the force gather (RK) or of the peridynamic force
        int nbf = 0, nbg = 0;
        xcd_grid(n, &nbf, kNodesPerFBlock);
        xcd_grid(n, &nbg);
        c->nb_force = std::max(nbf, nbg);
    }
    e |= dalloc(&c->wpart, c->nb_node); e |= dalloc(&c->epart, 3 * (size_t)c->nb_force);
    e |= dalloc(&c->flags, 4);
    e |= dalloc(&c->sstat, 4 * (size_t)c->nb_node);
    if (e) { mega_destroy(c); return MEGA_ERR_NOMEM; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { mega_destroy(c); return MEGA_ERR_HIP; }
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    auto zero = [&](void *dst, size_t bytes) { return hipMemset(dst, 0, bytes) == hipSuccess; };
    const auto X0i = to_internal(c, X0, 3), wini = to_internal(c, win, 3), voli = to_internal(c, vol, 1);
    auto massi = to_internal(c, mass, 3);
    if (c->mass_node)
        for (int i = 0; i < n; ++i) massi[i] = massi[3 * (size_t)i];
    const auto ebci = to_internal(c, ebc, 3), ebni = to_internal(c, ebc_node, 1), midi = to_internal(c, mid.data(), 1);
    std::vector<double> winvi(n3);
    for (size_t k = 0; k < n3; ++k) winvi[k] = 1.0 / wini[k];
    bool ok = up(c->X0, X0i.data(), n3 * 8) && up(c->x, X0i.data(), n3 * 8) && up(c->win, wini.data(), n3 * 8) &&
              up(c->winv, winvi.data(), n3 * 8) &&
              up(c->vol, voli.data(), (size_t)n * 8) && up(c->mass, massi.data(), (c->mass_node ? (size_t)n : n3) * 8) &&
              up(c->ebc, ebci.data(), n3 * 4) && up(c->ebc_node, ebni.data(), (size_t)n * 4) &&
              up(c->mat, midi.data(), (size_t)n * 4) && up(c->orig, c->perm.data(), (size_t)n * 4) &&
              up(c->mats, c->host_mats.data(), c->host_mats.size() * sizeof(Material));
    ok = ok && (vinit ? up(c->v, to_internal(c, vinit, 3).data(), n3 * 8) : zero(c->v, n3 * 8));
    ok = ok && zero(c->a, n3 * 8) && zero(c->d, n3 * 8) && zero(c->dtot, n3 * 8) && zero(c->uphy, n3 * 8) &&
         zero(c->fint, n3 * 8) && zero(c->fext, n3 * 8) && zero(c->stress, n6 * r) && zero(c->strain, n6 * r) &&
         zero(c->eps, n * r) && zero(c->st4, 4 * (size_t)n * r) && zero(c->vphy_step, n3 * 8) &&
         zero(c->aphy_step, n3 * 8) && zero(c->wpart, 8 * (size_t)c->nb_node) &&
         zero(c->epart, 24 * (size_t)c->nb_force) && zero(c->flags, 16) && zero(c->Fst, 80 * (size_t)n) &&
         zero(c->sstat, 32 * (size_t)c->nb_node);
    if (ok && c->fext_nmo) ok = zero(c->fext_nmo, n3 * 8);
    if (ok && c->mixed) ok = zero(c->d4, (size_t)n * 16) && zero(c->t4, (size_t)n * 16);
    if (ok && c->wave_vel) ok = zero(c->wave_vel, (size_t)n * 8);
    if (ok && c->snni) {   // GET_SMOOTHING_LENGTHS (PRE_MODEL_SMOOTHING_LENGTHS.F90:49-90)
        const double thrd = 0.333333333333333;   // the reference's THRD
        std::vector<double> l(n3), ao(n3);
        for (size_t i = 0; i < (size_t)n; ++i) {
            const double d1 = wini[3 * i] * 2.0, d2 = wini[3 * i + 1] * 2.0, d3 = wini[3 * i + 2] * 2.0;
            const double ratio = std::pow(voli[i], thrd) / std::pow(d1 * d2 * d3, thrd);
            const double s1 = d1 * ratio / 2.0, s2 = d2 * ratio / 2.0, s3 = d3 * ratio / 2.0;
            const double svol = (s1 + s1) * (s2 + s2) * (s3 + s3);
            l[3 * i] = s1; l[3 * i + 1] = s2; l[3 * i + 2] = s3;
            ao[3 * i] = (s2 + s2) * (s3 + s3) / svol;
            ao[3 * i + 1] = (s1 + s1) * (s3 + s3) / svol;
            ao[3 * i + 2] = (s1 + s1) * (s2 + s2) / svol;
        }
        ok = up(c->sml, l.data(), n3 * 8) && up(c->aov, ao.data(), n3 * 8);
    }
    if (!ok) { mega_destroy(c); return MEGA_ERR_HIP; }
    *out = c;
    return MEGA_OK;
}

int mega_destroy(mega_ctx *c)
{
    if (!c) return MEGA_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_pairs(c);
    void *ptrs[] = {c->X0, c->x, c->v, c->a, c->d, c->dtot, c->uphy, c->fint, c->fext, c->fext_nmo, c->win, c->winv,
                    c->vol, c->mass, c->wave_vel, c->scratch, c->d4, c->t4, c->u4, c->Fst, c->ebc, c->ebc_node, c->mat,
                    c->mats, c->stress, c->strain, c->eps, c->red, c->P, c->row_ptr, c->trow, c->counts,
                    c->ijk, c->cell_start, c->cell_cursor, c->cell_nodes, c->wpart, c->epart, c->flags,
                    c->cub_tmp, c->surr_tiles, c->surr_bias, c->surr_norm, c->sstat, c->hv, c->hv_count,
                    c->Kinv, c->gvec, c->orig, c->cells_ne, c->cell_flag, c->tbin, c->smask, c->gset_ptr,
                    c->gset, c->fset_ptr, c->fset, c->set_max, c->gloc, c->fpair, c->fpair2, c->dt_dev,
                    c->dt_hist, c->char_dist, c->dt_min, c->sml, c->aov, c->pd_qb};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &t : c->timers) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MEGA_OK;
}

int mega_search(mega_ctx *c, const double *x_cur, int32_t rebuild_bins)
{
    if (!c) return MEGA_ERR_ARG;
    const double *xs = c->x;
    std::vector<double> xi;
    if (x_cur) {
        xi = to_internal(c, x_cur, 3);
        HIPCHK(hipMemcpyAsync(c->scratch, xi.data(), sizeof(double) * 3 * c->n, hipMemcpyHostToDevice, c->stream));
        xs = c->scratch;
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
    if (!c || !c->has_lists) return MEGA_ERR_ARG;
    const int n = c->n;
    std::vector<int32_t> rp(n + 1);
    HIPCHK(hipMemcpyAsync(rp.data(), c->row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (nnz) *nnz = rp[n];
    // rows and neighbor ids in the caller's node order, 1-based
    int64_t s = 0;
    for (int o = 0; o < n; ++o) {
        const int i = c->iperm[o], len = rp[i + 1] - rp[i];
        if (gn) gn[o] = len;
        if (gstart) gstart[o] = (int32_t)(s + 1);
        s += len;
    }
    if (gstack) {
        if (cap < rp[n]) return MEGA_ERR_ARG;
        std::vector<int32_t> col((size_t)rp[n]);
        HIPCHK(hipMemcpyAsync(col.data(), c->col, sizeof(int32_t) * rp[n], hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int64_t q = 0;
        for (int o = 0; o < n; ++o) {
            const int i = c->iperm[o];
            for (int t = rp[i]; t < rp[i + 1]; ++t) gstack[q++] = c->perm[col[t]] + 1;
        }
    }
    return MEGA_OK;
}

int mega_shapes(mega_ctx *c, const double *x_eval)
{
    if (!c || !c->has_lists) return MEGA_ERR_ARG;
    if (c->pd) {   // UDFM_SHAPE_TENSOR at the initial positions
        HIPCHK(hipMemsetAsync(c->flags, 0, 16, c->stream));
        int r = do_pd_shapes(c);
        if (r) return r;
        HIPCHK(hipStreamSynchronize(c->stream));
        return check_flags(c);
    }
    const double *xs = c->X0;
    std::vector<double> xi;
    if (x_eval) {
        xi = to_internal(c, x_eval, 3);
        HIPCHK(hipMemcpyAsync(c->scratch, xi.data(), sizeof(double) * 3 * c->n, hipMemcpyHostToDevice, c->stream));
        xs = c->scratch;
    }
    int r = do_shapes(c, xs, false);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    return MEGA_OK;
}

int mega_get_shapes(mega_ctx *c, double *shp, double *dshp, int64_t cap)
{
    if (!c || !c->has_shapes || cap < c->nnz) return MEGA_ERR_ARG;
    if (c->dshp_stale && dshp) {   // form them again at the coordinates the lists were taken at
        const int r = do_shapes(c, c->x, false);
        if (r) return r;
    }
    const int64_t m = c->nnz;
    double *buf = nullptr;
    if (dalloc(&buf, 3 * (size_t)m)) return MEGA_ERR_NOMEM;
    // shp, then the x, y, z gradient blocks (stride nnz_cap on the device)
    for (int part = 0; part < 4; ++part) {
        const size_t off = part == 0 ? 0 : (size_t)(part - 1) * c->nnz_cap;
        const void *src = part == 0 ? c->shp : (const char *)c->dshp + off * c->rsz;
        double *dst = buf + (part == 0 ? 0 : (part - 1) * m);
        if (c->mixed) k_widen<float><<<grid_for(m), kBlock, 0, c->stream>>>(m, (const float *)src, dst);
        else k_widen<double><<<grid_for(m), kBlock, 0, c->stream>>>(m, (const double *)src, dst);
        if (part == 0) {
            std::vector<double> h((size_t)m);
            if (shp) HIPCHK(hipMemcpyAsync(h.data(), buf, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (shp) {
                const int r = rows_to_caller(c, h.data(), shp, 1);
                if (r) { (void)hipFree(buf); return r; }
            }
        }
    }
    int st = MEGA_OK;
    if (dshp) {
        std::vector<double> h(3 * (size_t)m), g(3 * (size_t)m);
        if (hipMemcpyAsync(h.data(), buf, sizeof(double) * 3 * m, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            st = MEGA_ERR_HIP;
        for (int64_t q = 0; q < m && !st; ++q)
            for (int k = 0; k < 3; ++k) g[3 * q + k] = h[(size_t)(k * m + q)];
        if (!st) st = rows_to_caller(c, g.data(), dshp, 3);
    }
    (void)hipFree(buf);
    return st;
}

int mega_get_pd_shape(mega_ctx *c, double *omega, double *invk, int64_t cap)
{
    if (!c || !c->pd || !c->pd_shapes || (omega && cap < c->nnz)) return MEGA_ERR_ARG;
    const int64_t m = c->nnz;
    if (omega) {
        double *buf = nullptr;
        if (dalloc(&buf, (size_t)m)) return MEGA_ERR_NOMEM;
        if (c->mixed) k_widen<float><<<grid_for(m), kBlock, 0, c->stream>>>(m, (const float *)c->shp, buf);
        else k_widen<double><<<grid_for(m), kBlock, 0, c->stream>>>(m, (const double *)c->shp, buf);
        std::vector<double> h((size_t)m);
        const bool ok = hipMemcpyAsync(h.data(), buf, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
                        hipStreamSynchronize(c->stream) == hipSuccess;
        (void)hipFree(buf);
        if (!ok) return MEGA_ERR_HIP;
        const int r = rows_to_caller(c, h.data(), omega, 1);
        if (r) return r;
    }
    if (invk) {
        HIPCHK(hipMemcpyAsync(invk, c->Kinv, sizeof(double) * 9 * c->n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        to_caller(c, invk, 9);
    }
    return MEGA_OK;
}

int mega_interp(mega_ctx *c, const double *coef, double *phys)
{
    if (!c || !coef || !phys) return MEGA_ERR_ARG;
    const size_t b = sizeof(double) * 3 * c->n;
    if (c->pd) {   // physical = generalized (HANDELER.F90:134-137)
        std::memcpy(phys, coef, b);
        return MEGA_OK;
    }
    if (!c->has_shapes) return MEGA_ERR_ARG;
    double *out = nullptr;
    if (dalloc(&out, 3 * (size_t)c->n)) return MEGA_ERR_NOMEM;
    int st = MEGA_OK;
    const auto ci = to_internal(c, coef, 3);
    if (hipMemcpyAsync(c->scratch, ci.data(), b, hipMemcpyHostToDevice, c->stream) != hipSuccess) st = MEGA_ERR_HIP;
    if (!st) {
        Params p = params(c);
        if (c->mixed) k_interp<float, false><<<grid_for((int64_t)c->n * kGroup), kBlock, 0, c->stream>>>(p, c->scratch, out);
        else k_interp<double, false><<<grid_for((int64_t)c->n * kGroup), kBlock, 0, c->stream>>>(p, c->scratch, out);
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(phys, out, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            st = MEGA_ERR_HIP;
    }
    (void)hipFree(out);
    if (!st) to_caller(c, phys, 3);
    return st;
}

int mega_fint(mega_ctx *c, int32_t linit, const double *dinc, const double *dinc_tot, const double *x_cur,
              double *stress, double *strain, double *state, double *fint, double *fext, double *int_work)
{
    if (!c || !dinc || !dinc_tot || !stress || !strain || !state || !fint || !fext || !c->has_lists)
        return MEGA_ERR_ARG;
    const int n = c->n;
    const size_t b3 = sizeof(double) * 3 * n;
    // STATE(3) = EPS, and the damage models' STATE(1), (2), (4), (5) (st4)
    constexpr int kDam[4] = {0, 1, 3, 4};
    std::vector<double> eps0(n), dam0(4 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        eps0[i] = state[20 * (int64_t)i + 2];
        for (int k = 0; k < 4; ++k) dam0[4 * (size_t)i + k] = state[20 * (int64_t)i + kDam[k]];
    }
    const auto di = to_internal(c, dinc, 3), ti = to_internal(c, dinc_tot, 3);
    const auto si = to_internal(c, stress, 6), ei = to_internal(c, strain, 6);
    auto eps = to_internal(c, eps0.data(), 1);
    auto dam = to_internal(c, dam0.data(), 4);
    std::vector<double> xi;
    HIPCHK(hipMemcpyAsync(c->d, di.data(), b3, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dtot, ti.data(), b3, hipMemcpyHostToDevice, c->stream));
    if (x_cur) {
        xi = to_internal(c, x_cur, 3);
        HIPCHK(hipMemcpyAsync(c->x, xi.data(), b3, hipMemcpyHostToDevice, c->stream));
    }
    int r = sync_d4(c);
    if (!r) r = state_in(c, c->stress, 6 * (int64_t)n, si.data());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!r) r = state_in(c, c->strain, 6 * (int64_t)n, ei.data());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!r) r = state_in(c, c->eps, n, eps.data());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!r) r = state_in(c, c->st4, 4 * (int64_t)n, dam.data());
    if (r) return r;
    HIPCHK(hipMemsetAsync(c->flags, 0, 16, c->stream));
    int err = 0;
    const double w0 = host_sum(c, c->wpart, c->nb_node, 1, 0, &err);
    if (err) return err;
    const bool lag = c->c.lagrangian;
    if (c->pd) {   // CONSTRUCT_FINT_PD (FINT_PD.F90:14-759)
        if (linit) r = do_pd_shapes(c);
        if (!r) r = do_pd_fint(c, linit ? 1 : 0, 0);
    } else {
        if (linit || !lag) r = do_shapes(c, lag ? c->X0 : c->x, false);
        else if (!c->has_shapes) r = MEGA_ERR_ARG;
        if (!r) r = do_fint(c, (lag && c->c.finite_strain) ? F_DIRECT : F_NONE, false, 0);
    }
    if (r) return r;
    const int st = check_flags(c);
    const double w = host_sum(c, c->wpart, c->nb_node, 1, 0, &err) - w0;
    if (err) return err;
    r = state_out(c, c->stress, 6 * (int64_t)n, stress);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!r) r = state_out(c, c->strain, 6 * (int64_t)n, strain);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!r) r = state_out(c, c->eps, n, eps.data());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!r) r = state_out(c, c->st4, 4 * (int64_t)n, dam.data());
    if (r) return r;
    HIPCHK(hipMemcpyAsync(fint, c->fint, b3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fext, c->fext, b3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    to_caller(c, stress, 6);
    to_caller(c, strain, 6);
    to_caller(c, eps.data(), 1);
    to_caller(c, dam.data(), 4);
    to_caller(c, fint, 3);
    to_caller(c, fext, 3);
    for (int i = 0; i < n; ++i) {
        state[20 * (int64_t)i + 2] = eps[i];
        for (int k = 0; k < 4; ++k) state[20 * (int64_t)i + kDam[k]] = dam[4 * (size_t)i + k];
    }
    if (int_work) *int_work += w;
    c->predicted = 0;
    return st;
}

int mega_constitute(int32_t n, const double *lmat, double *stress, double *strain, double *state, const double *prop,
                    int32_t mat_type, int32_t finite_strain, int32_t precision)
{
    if (n <= 0 || !lmat || !stress || !strain || !state || !prop) return MEGA_ERR_ARG;
    if (mat_type != 1 && mat_type != 2 && mat_type != 3 && mat_type != 6) return MEGA_ERR_MATERIAL;
    if (precision != MEGA_FP64 && precision != MEGA_MIXED) return MEGA_ERR_ARG;
    const Material mt = make_material(prop, mat_type);
    double *dl = nullptr, *ds = nullptr, *de = nullptr, *dst = nullptr;
    int32_t *fail = nullptr;
    const size_t nn = (size_t)n;
    if (dalloc(&dl, 9 * nn) || dalloc(&ds, 6 * nn) || dalloc(&de, 6 * nn) || dalloc(&dst, 20 * nn) || dalloc(&fail, 1)) {
        (void)hipFree(dl); (void)hipFree(ds); (void)hipFree(de); (void)hipFree(dst); (void)hipFree(fail);
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
        if (precision == MEGA_MIXED) k_constitute<float><<<grid_for(n), kBlock>>>(n, dl, ds, de, dst, mt, finite_strain, fail);
        else k_constitute<double><<<grid_for(n), kBlock>>>(n, dl, ds, de, dst, mt, finite_strain, fail);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(stress, ds, 48 * nn, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(strain, de, 48 * nn, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(state, dst, 160 * nn, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&f, fail, 4, hipMemcpyDeviceToHost) != hipSuccess)
            st = MEGA_ERR_HIP;
        else if (f)
            st = MEGA_ERR_RETURN_MAP;
    }
    (void)hipFree(dl); (void)hipFree(ds); (void)hipFree(de); (void)hipFree(dst); (void)hipFree(fail);
    return st;
}

static int run_steps(mega_ctx *c, int32_t nsteps)
{
    if (c->profile) c->timer_used = 0;
    const int n = c->n;
    const bool lag = c->c.lagrangian;
    const int fm = (lag && c->c.finite_strain) ? F_INCR : F_NONE;
    HIPCHK(hipMemsetAsync(c->flags, 0, 16, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        const bool first = (c->step == 0);
        const bool last = (s == nsteps - 1);
        int r = 0;
        // tuples of this step's return maps (the first call has zero increments)
        c->hv_active = (c->hv && c->hv_every > 0 && c->step > 0 && c->step % c->hv_every == 0) ? 1 : 0;
        if (first) {
            // LINIT call of HANDELER with zero increments (MAIN.F90:320-334)
            HIPCHK(hipMemsetAsync(c->d, 0, sizeof(double) * 3 * n, c->stream));
            r = sync_d4(c);
            if (!r) r = search_step(c, c->x, 1);
            if (c->pd) {
                if (!r) r = do_pd_shapes(c);
                if (!r) r = do_pd_fint(c, 1, -1);
            } else {
                if (!r) r = do_shapes(c, lag ? c->X0 : c->x, false);
                if (!r) r = do_fint(c, fm == F_INCR ? F_DIRECT : F_NONE, false, -1);
            }
            if (r) return r;
        }
        // the predictor is fused into the previous k_force within one call; with
        // the automatic time step the first step has none (MAIN.F90:348-350)
        if (!c->predicted && !(c->auto_on && first)) r = do_predict(c);
        if (r) return r;
        if (last) {   // LOCAL_VEL_PHY, LOCAL_ACL_PHY as DO_INTERP forms them (HANDELER.F90:114-137)
            if (c->pd) {
                HIPCHK(hipMemcpyAsync(c->vphy_step, c->v, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(c->aphy_step, c->a, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->stream));
            } else {
                Params p = params(c);
                const unsigned g = grid_for((int64_t)n * kGroup);
                if (c->mixed) {
                    k_interp<float, false><<<g, kBlock, 0, c->stream>>>(p, c->v, c->vphy_step);
                    k_interp<float, false><<<g, kBlock, 0, c->stream>>>(p, c->a, c->aphy_step);
                } else {
                    k_interp<double, false><<<g, kBlock, 0, c->stream>>>(p, c->v, c->vphy_step);
                    k_interp<double, false><<<g, kBlock, 0, c->stream>>>(p, c->a, c->aphy_step);
                }
            }
        }
        // a is read by the last step's DO_INTERP above, fint, fext and the
        // Lagrangian x by the caller after the call
        c->out_dof = s >= nsteps - 2;
        c->out_x = last || c->pd || !lag;
        if (c->pd) {
            c->auto_step = s;
            r = do_pd_fint(c, 0, last ? 1 : 2);   // DO_INTERP is the identity (HANDELER.F90:134-137)
        } else if (lag) {
            r = do_fint(c, fm, true, last ? 1 : 2);   // DO_INTERP fused into the gather
        } else {
            Params p = params(c);
            {
                Scope sc(c, "k_interp");
                if (c->mixed) k_interp<float, true><<<grid_for((int64_t)n * kGroup), kBlock, 0, c->stream>>>(p, nullptr, nullptr);
                else k_interp<double, true><<<grid_for((int64_t)n * kGroup), kBlock, 0, c->stream>>>(p, nullptr, nullptr);
            }
            r = search_step(c, c->x, first);
            if (c->snni) {   // the smoothed gradients replace the node-point ones before the gather
                if (!r) r = do_shapes(c, c->x, false);
                if (!r) r = do_fint(c, F_NONE, false, last ? 1 : 2, false);
            } else {
                if (!r) r = do_shapes(c, c->x, true);   // with the gather sums
                if (!r) r = do_fint(c, F_NONE, false, last ? 1 : 2, true);
            }
        }
        c->out_dof = c->out_x = 1;
        if (r) return r;
        c->predicted = last ? 0 : 1;
        HIPCHK(hipGetLastError());
        c->step++;
    }
    c->hv_active = 0;
    return check_flags(c);
}

int mega_step(mega_ctx *c, double dt, int32_t nsteps)
{
    if (!c || nsteps < 0 || !(dt > 0.0)) return MEGA_ERR_ARG;
    c->dt = dt;
    c->auto_on = 0;
    return run_steps(c, nsteps);
}

int mega_step_auto(mega_ctx *c, double dlt_fac, int32_t nsteps, double *dlt)
{
    // the RKPM CONSTRUCT_FINT leaves DLT_FINT unset (its estimate is commented
    // out, FINT.F90:1250-1326), so only the peridynamic path has one
    if (!c || !c->pd || !c->wave_vel || nsteps < 0 || !(dlt_fac > 0.0) || !(dlt_fac < INFINITY)) return MEGA_ERR_ARG;
    if (!c->dt_dev) {
        if (dalloc(&c->dt_dev, 1) || dalloc(&c->dt_min, 1) || dalloc(&c->char_dist, c->n)) return MEGA_ERR_NOMEM;
        HIPCHK(hipMemsetAsync(c->dt_min, 0xff, sizeof(unsigned long long), c->stream));
        HIPCHK(hipMemsetAsync(c->char_dist, 0, sizeof(double) * c->n, c->stream));
        HIPCHK(hipMemsetAsync(c->dt_dev, 0, sizeof(double), c->stream));
    }
    if (nsteps > c->dt_hist_cap) {
        if (c->dt_hist) (void)hipFree(c->dt_hist);
        c->dt_hist = nullptr;
        c->dt_hist_cap = 0;
        if (dalloc(&c->dt_hist, (size_t)nsteps)) return MEGA_ERR_NOMEM;
        c->dt_hist_cap = nsteps;
    }
    c->dlt_fac = dlt_fac;
    c->auto_on = 1;
    int r = run_steps(c, nsteps);
    if (r == MEGA_OK || r == MEGA_ERR_RETURN_MAP) {
        std::vector<double> h((size_t)nsteps);
        if (nsteps > 0) {
            HIPCHK(hipMemcpyAsync(h.data(), c->dt_hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            c->dt = h[nsteps - 1];
        }
        if (dlt)
            for (int32_t k = 0; k < nsteps; ++k) dlt[k] = h[k];
    }
    return r;
}

static int field_width(int f)
{
    switch (f) {
    case MEGA_F_STRESS: case MEGA_F_STRAIN: return 6;
    case MEGA_F_EPS: case MEGA_F_STRAIN_EQ: case MEGA_F_WAVE_VEL: return 1;
    case MEGA_F_DAMAGE_STATE: return 4;
    default: return 3;
    }
}

static double *field_ptr(mega_ctx *c, int f, int *width)
{
    *width = 3;
    switch (f) {
    case MEGA_F_X: return c->x;
    case MEGA_F_V: return c->v;
    case MEGA_F_A: return c->a;
    case MEGA_F_D: return c->d;
    case MEGA_F_DTOT: return c->dtot;
    case MEGA_F_UPHY: return c->uphy;
    case MEGA_F_FINT: return c->fint;
    case MEGA_F_FEXT: return c->fext;
    case MEGA_F_WAVE_VEL: *width = 1; return c->wave_vel;
    case MEGA_F_APHY: return c->aphy_step;
    case MEGA_F_VPHY_STEP: return c->vphy_step;
    default: return nullptr;
    }
}

int mega_download(mega_ctx *c, int32_t field, double *host)
{
    if (!c || !host) return MEGA_ERR_ARG;
    const int n = c->n;
    int r = MEGA_OK;
    switch (field) {
    case MEGA_F_STRESS: r = state_out(c, c->stress, 6 * (int64_t)n, host); break;
    case MEGA_F_STRAIN: r = state_out(c, c->strain, 6 * (int64_t)n, host); break;
    case MEGA_F_EPS: r = state_out(c, c->eps, n, host); break;
    case MEGA_F_DAMAGE_STATE: r = state_out(c, c->st4, 4 * (int64_t)n, host); break;
    case MEGA_F_STRAIN_EQ:
        if (c->mixed) k_strain_eq<float><<<grid_for(n), kBlock, 0, c->stream>>>(n, (const float *)c->strain, c->scratch);
        else k_strain_eq<double><<<grid_for(n), kBlock, 0, c->stream>>>(n, (const double *)c->strain, c->scratch);
        HIPCHK(hipMemcpyAsync(host, c->scratch, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        break;
    case MEGA_F_VPHY: {   // sum_j phi_ij v_j of the current velocity (v itself for peridynamics)
        if (c->pd) {
            HIPCHK(hipMemcpyAsync(host, c->v, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
            break;
        }
        if (!c->has_shapes) return MEGA_ERR_ARG;
        Params p = params(c);
        if (c->mixed) k_interp<float, false><<<grid_for((int64_t)n * kGroup), kBlock, 0, c->stream>>>(p, c->v, c->scratch);
        else k_interp<double, false><<<grid_for((int64_t)n * kGroup), kBlock, 0, c->stream>>>(p, c->v, c->scratch);
        HIPCHK(hipMemcpyAsync(host, c->scratch, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        break;
    }
    default: {
        int w = 0;
        double *src = field_ptr(c, field, &w);
        if (!src) return MEGA_ERR_ARG;
        HIPCHK(hipMemcpyAsync(host, src, sizeof(double) * w * n, hipMemcpyDeviceToHost, c->stream));
    }
    }
    if (r) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    to_caller(c, host, field_width(field));
    return MEGA_OK;
}

int mega_upload(mega_ctx *c, int32_t field, const double *host_caller)
{
    if (!c || !host_caller || field < 0 || field >= MEGA_F_NFIELDS) return MEGA_ERR_ARG;
    const int n = c->n;
    int r = MEGA_OK;
    const auto hi = to_internal(c, host_caller, field_width(field));
    const double *host = hi.data();
    switch (field) {
    case MEGA_F_STRESS: r = state_in(c, c->stress, 6 * (int64_t)n, host); break;
    case MEGA_F_STRAIN: r = state_in(c, c->strain, 6 * (int64_t)n, host); break;
    case MEGA_F_EPS: r = state_in(c, c->eps, n, host); break;
    case MEGA_F_DAMAGE_STATE: r = state_in(c, c->st4, 4 * (int64_t)n, host); break;
    default: {
        int w = 0;
        double *dst = field_ptr(c, field, &w);
        if (!dst || field == MEGA_F_WAVE_VEL || field == MEGA_F_APHY || field == MEGA_F_VPHY_STEP) return MEGA_ERR_ARG;
        HIPCHK(hipMemcpyAsync(dst, host, sizeof(double) * w * n, hipMemcpyHostToDevice, c->stream));
        if (field == MEGA_F_D || field == MEGA_F_DTOT) r = sync_d4(c);
        if (field == MEGA_F_DTOT) c->F_valid = 0;
    }
    }
    if (r) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    return MEGA_OK;
}

int mega_energies(mega_ctx *c, double *e_int, double *e_kin, double *e_ext)
{
    if (!c) return MEGA_ERR_ARG;
    int err = 0;
    const double ke = host_sum(c, c->epart, c->nb_force, 3, 0, &err);
    const double ie = host_sum(c, c->epart, c->nb_force, 3, 1, &err);
    const double ee = host_sum(c, c->epart, c->nb_force, 3, 2, &err);
    const double wk = host_sum(c, c->wpart, c->nb_node, 1, 0, &err);
    if (err) return err;
    if (e_int) *e_int = c->c.finite_strain ? wk : ie;
    if (e_kin) *e_kin = ke;
    if (e_ext) *e_ext = ee;
    return MEGA_OK;
}

int mega_stats(mega_ctx *c, int64_t *nnz, int32_t *max_neighbors, int64_t *fallbacks)
{
    if (!c || !c->has_lists) return MEGA_ERR_ARG;
    const int n = c->n;
    std::vector<int32_t> rp(n + 1);
    HIPCHK(hipMemcpyAsync(rp.data(), c->row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int32_t mx = 0;
    for (int i = 0; i < n; ++i) mx = std::max(mx, rp[i + 1] - rp[i]);
    if (nnz) *nnz = rp[n];
    if (max_neighbors) *max_neighbors = mx;
    if (fallbacks) {
        int64_t st[4];
        const int r = surrogate_totals(c, st, 0);
        if (r) return r;
        *fallbacks = st[3];
    }
    return MEGA_OK;
}

int mega_profile(mega_ctx *c, int32_t enable)
{
    if (!c) return MEGA_ERR_ARG;
    if (enable < 0 || enable > 2) return MEGA_ERR_ARG;
    c->profile = enable;
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

int mega_set_surrogate(mega_ctx *c, const float *params, const float *norm, const double *prop, double tol)
{
    if (!c) return MEGA_ERR_ARG;
    if (!params) { c->surr_on = 0; return MEGA_OK; }
    if (!norm || !prop || !c->mixed || c->hv || !(tol > 0.0)) return MEGA_ERR_ARG;
    Material want = make_material(prop, 2);
    int slot = -1;
    for (size_t k = 0; k < c->host_mats.size(); ++k) {
        const Material &m = c->host_mats[k];
        if (m.type == 2 && m.nu == want.nu && m.E == want.E && m.rho == want.rho && m.sig0 == want.sig0 &&
            m.B == want.B && m.CE == want.CE && m.lam == want.lam && m.mu == want.mu && m.l2m == want.l2m)
            slot = (int)k;
    }
    if (slot < 0) return MEGA_ERR_ARG;
    std::vector<uint16_t> tiles(mlp::kTilesFwd * 512);
    for (int t = 0; t < mlp::kTilesFwd; ++t)
        for (int l = 0; l < 64; ++l)
            for (int j = 0; j < 8; ++j) tiles[(t * 64 + l) * 8 + j] = mlp::bf16_bits(mlp::fwd_tile_value(params, t, l, j));
    std::vector<float> bias(mlp::kBiasPack);
    for (int k = 0; k < mlp::kBiasPack; ++k) bias[k] = mlp::bias_pack_value(params, k);
    if (!c->surr_tiles) {
        int e = dalloc(&c->surr_tiles, tiles.size()) | dalloc(&c->surr_bias, bias.size()) |
                dalloc(&c->surr_norm, mlp::kNorm);
        if (e) return MEGA_ERR_NOMEM;
    }
    HIPCHK(hipMemcpyAsync(c->surr_tiles, tiles.data(), tiles.size() * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->surr_bias, bias.data(), bias.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->surr_norm, norm, mlp::kNorm * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < mlp::kIn; ++k) {
        c->surr_env[k] = norm[mlp::kEnvLo + k];
        c->surr_env[mlp::kIn + k] = norm[mlp::kEnvHi + k];
    }
    c->surr_mat = slot;
    c->surr_tol = (float)tol;
    c->surr_on = 1;
    return MEGA_OK;
}

int mega_surrogate_stats(mega_ctx *c, int64_t *stats, int32_t reset)
{
    if (!c || !stats) return MEGA_ERR_ARG;
    return surrogate_totals(c, stats, reset);
}

int mega_harvest(mega_ctx *c, int64_t capacity, int32_t every)
{
    if (!c || capacity < 0 || (capacity > 0 && every < 1)) return MEGA_ERR_ARG;
    if (capacity > 0 && c->surr_on) return MEGA_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->hv) (void)hipFree(c->hv);
    if (c->hv_count) (void)hipFree(c->hv_count);
    c->hv = nullptr;
    c->hv_count = nullptr;
    c->hv_cap = 0;
    c->hv_every = 0;
    if (capacity == 0) return MEGA_OK;
    if (dalloc(&c->hv, 20 * (size_t)capacity) || dalloc(&c->hv_count, 1)) return MEGA_ERR_NOMEM;
    HIPCHK(hipMemsetAsync(c->hv_count, 0, sizeof(unsigned long long), c->stream));
    c->hv_cap = capacity;
    c->hv_every = every;
    return MEGA_OK;
}

int mega_harvest_take(mega_ctx *c, float *host, int64_t cap, int64_t *count)
{
    if (!c || !c->hv || !count) return MEGA_ERR_ARG;
    unsigned long long h = 0;
    HIPCHK(hipMemcpyAsync(&h, c->hv_count, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *count = (int64_t)h;   // may exceed the capacity: only the first capacity tuples were kept
    const int64_t m = std::min<int64_t>((int64_t)h, std::min<int64_t>(cap, c->hv_cap));
    if (host && m > 0) {
        HIPCHK(hipMemcpyAsync(host, c->hv, sizeof(float) * 20 * m, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    HIPCHK(hipMemsetAsync(c->hv_count, 0, sizeof(h), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MEGA_OK;
}

}  // extern "C"
