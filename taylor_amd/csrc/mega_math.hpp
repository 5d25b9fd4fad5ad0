// mega_math.hpp -- per-point math of the explicit meshfree step, as inline
// device functions.  Each follows the reference routine cited beside it,
// including the reference's quirks (they are its behaviour, and parity is
// measured against it):
//   INV3 returns adj(A)*det(A)                       INVERSE.F90:119-152
//   DETERMINANT returns |det A|                      DETERMINANT.F90:18-35
//   y/z kernel derivatives divide by the x window    RK.f90:191-205
//   singular-kernel epsilon is the REAL(4) 1.0E-015  RK.f90:219
//   J2 Newton tolerance uses the REAL(4) 1.0E-05     CONSTITUTE_VONMISES.F90:109
// Small matrices are column-major: a[i + 3*j] = A(i+1, j+1).
// The constitutive per-point functions keep the reference's unfused
// arithmetic (no contraction of products into fused multiply-adds), so that
// in FP64 they give the reference's bits; the kernel and the moment matrix are
// sums over neighbors, whose order already differs.
#pragma once
#include <hip/hip_runtime.h>

#define MEGA_HD __host__ __device__ __forceinline__

namespace mega {

// REAL(4) literals widened to double, as the reference's expressions see them
constexpr double kSingularEps = (double)1.0e-15f;
constexpr double kNewtonTol = (double)1.0e-05f;

struct Material {       // one property row, with derived constants
    double nu, E, rho, sig0, B, CE, damping;
    double lam, mu, l2m;   // props 28-30 (READ.F90:496-507)
    double bulk, shear;    // BULK_MOD, SHEAR_MOD (FINT_FUNCTIONS.F90:197-229)
    double k_i, k_c, dam_max;       // damage props 9-11
    double q_phi, k_phi, q_psi;     // Drucker-Prager props 25-27 (READ.F90:420-483)
    int type;              // 1 elastic, 2 J2, 3 Drucker-Prager + damage, 6 J2 + damage
    int pad;
};

// fl(a / b) < 1 for a >= 0, b > 0 (normal numbers), decided without the
// division.  IEEE division rounds to nearest-even; the doubles next to 1 from
// below are 1 - 2^-53 and 1, with midpoint 1 - 2^-54 (which is never a quotient
// of two doubles: a 2^54 = b (2^54 - 1) has no solution in 53-bit
// significands).  So fl(a/b) < 1 exactly when a/b < 1 - 2^-54, i.e.
// (a - b) + b 2^-54 < 0.  For b/2 <= a < b, a - b is
// exact (Sterbenz) and the fma rounds once, which keeps the sign of the exact
// sum; outside that range the answer is plain.  This is the box test of
// SOFT_SEARCH, |x_J,k - x_k| / a_J,k < 1 (BIN_SEARCH.F90:268-289).
MEGA_HD bool quotient_below_one(double a, double b)
{
    if (a < 0.5 * b) return true;
    if (!(a < b)) return false;
    return fma(b, 0x1p-54, a - b) < 0.0;
}

// FP32 prefilter of the box test.  Positions are taken relative to a corner
// point and kept only within `reach` of it (NaN otherwise, which decides
// nothing); an axis then reads +1 when |dx| < w (1 - margin) for sure, -1
// when |dx| > w (1 + margin) for sure, 0 when the exact test must decide.
// With both points within 4 bins of the corner, the FP32 distance is within
// 2^-24 x 16 bins of the FP64 one; for windows above 1/50 bin that is below
// 5e-5 w, under the margin of 1e-4.
constexpr float kBoxMargin = 1e-4f;
MEGA_HD float box_rel(double x, double corner, float reach)
{
    const float p = (float)(x - corner);
    return fabsf(p) < reach ? p : __builtin_nanf("");
}
// lo = hi = -1 for a window below wmin: every pair then goes to the exact test
MEGA_HD void box_limits(double w, float wmin, float &lo, float &hi)
{
    const float wf = (float)w;
    if (wf > wmin) {
        lo = wf * (1.f - kBoxMargin);
        hi = wf * (1.f + kBoxMargin);
    } else {
        lo = -1.f;
        hi = __builtin_inff();
    }
}
MEGA_HD int box_axis(float p, float q, float lo, float hi)
{
    const float d = fabsf(p - q);
    return d < lo ? 1 : (d > hi ? -1 : 0);
}

// MLS_KERNEL0 cubic B-spline, value and d/d|z| (KERNEL.F90:57-71)
template <class T>
MEGA_HD void cubic_kernel(T x, T &phi, T &dphi)
{
    if (x <= T(0.5)) {
        phi = T(2.0 / 3.0) - T(4) * (x * x) + T(4) * (x * x * x);
        dphi = T(-8) * x + T(12) * (x * x);
    } else if (x <= T(1)) {
        phi = T(4.0 / 3.0) - T(4) * x + T(4) * (x * x) - T(4.0 / 3.0) * (x * x * x);
        dphi = T(-4) + T(8) * x - T(4) * (x * x);
    } else {
        phi = T(0);
        dphi = T(0);
    }
}

// INV3: adj(A) * det(A)  (INVERSE.F90:132-147)
template <class T>
MEGA_HD void inv3_ref(const T *a, T *r)
{
#pragma clang fp contract(off)
#define A_(i, j) a[((i)-1) + 3 * ((j)-1)]
#define R_(i, j) r[((i)-1) + 3 * ((j)-1)]
    T det = A_(1,1)*A_(2,2)*A_(3,3) - A_(1,1)*A_(2,3)*A_(3,2) - A_(1,2)*A_(2,1)*A_(3,3)
          + A_(1,2)*A_(2,3)*A_(3,1) + A_(1,3)*A_(2,1)*A_(3,2) - A_(1,3)*A_(2,2)*A_(3,1);
    R_(1,1) =  (A_(2,2)*A_(3,3)-A_(2,3)*A_(3,2))*det;
    R_(2,1) = -(A_(2,1)*A_(3,3)-A_(2,3)*A_(3,1))*det;
    R_(3,1) =  (A_(2,1)*A_(3,2)-A_(2,2)*A_(3,1))*det;
    R_(1,2) = -(A_(1,2)*A_(3,3)-A_(1,3)*A_(3,2))*det;
    R_(2,2) =  (A_(1,1)*A_(3,3)-A_(1,3)*A_(3,1))*det;
    R_(3,2) = -(A_(1,1)*A_(3,2)-A_(1,2)*A_(3,1))*det;
    R_(1,3) =  (A_(1,2)*A_(2,3)-A_(1,3)*A_(2,2))*det;
    R_(2,3) = -(A_(1,1)*A_(2,3)-A_(1,3)*A_(2,1))*det;
    R_(3,3) =  (A_(1,1)*A_(2,2)-A_(1,2)*A_(2,1))*det;
#undef A_
#undef R_
}

// DETERMINANT (DETERMINANT.F90:28-30)
template <class T>
MEGA_HD T det3_abs(const T *a)
{
#pragma clang fp contract(off)
#define A_(i, j) a[((i)-1) + 3 * ((j)-1)]
    T d = A_(1,1)*(A_(2,2)*A_(3,3) - A_(3,2)*A_(2,3))
        + A_(1,2)*(A_(3,1)*A_(2,3) - A_(2,1)*A_(3,3))
        + A_(1,3)*(A_(2,1)*A_(3,2) - A_(3,1)*A_(2,2));
#undef A_
    return d < T(0) ? -d : d;
}

template <class T>
MEGA_HD void mm3(const T *a, const T *b, T *c)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            c[i + 3 * k] = a[i] * b[3 * k] + a[i + 3] * b[1 + 3 * k] + a[i + 6] * b[2 + 3 * k];
}

// INVERSE (INVERSE.F90:17-99) for a 3x3 matrix: Gauss-Jordan with full
// pivoting.  When the largest remaining pivot d has d + 1 == 1 it stops and
// returns the partly reduced matrix (the reference then only warns).
// Returns 1 on success, 0 on that early stop.
MEGA_HD int inverse_gj3(const double *A, double *Ai)
{
#pragma clang fp contract(off)
#define G_(i, j) a[(i) + 3 * (j)]
    double a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e] = A[e];
    int is[3] = {0, 0, 0}, js[3] = {0, 0, 0};
    int ok = 1;
    for (int k = 0; k < 3 && ok; ++k) {
        double d = 0.0;
        for (int i = k; i < 3; ++i)
            for (int j = k; j < 3; ++j)
                if (fabs(G_(i, j)) > d) { d = fabs(G_(i, j)); is[k] = i; js[k] = j; }
        if (d + 1.0 == 1.0) { ok = 0; break; }
        for (int j = 0; j < 3; ++j) { const double t = G_(k, j); G_(k, j) = G_(is[k], j); G_(is[k], j) = t; }
        for (int i = 0; i < 3; ++i) { const double t = G_(i, k); G_(i, k) = G_(i, js[k]); G_(i, js[k]) = t; }
        G_(k, k) = 1.0 / G_(k, k);
        for (int j = 0; j < 3; ++j)
            if (j != k) G_(k, j) = G_(k, j) * G_(k, k);
        for (int i = 0; i < 3; ++i)
            if (i != k)
                for (int j = 0; j < 3; ++j)
                    if (j != k) G_(i, j) = G_(i, j) - G_(i, k) * G_(k, j);
        for (int i = 0; i < 3; ++i)
            if (i != k) G_(i, k) = -G_(i, k) * G_(k, k);
    }
    if (ok)
        for (int k = 2; k >= 0; --k) {
            for (int j = 0; j < 3; ++j) { const double t = G_(k, j); G_(k, j) = G_(js[k], j); G_(js[k], j) = t; }
            for (int i = 0; i < 3; ++i) { const double t = G_(i, k); G_(i, k) = G_(i, is[k]); G_(i, is[k]) = t; }
        }
#pragma unroll
    for (int e = 0; e < 9; ++e) Ai[e] = a[e];
#undef G_
    return ok;
}

// 4x4 inverse by cofactors (M44INV, INVERSE.F90:185-223) on the symmetric
// moment matrix given by its 10 upper entries m[0..9] =
// (00,01,02,03,11,12,13,22,23,33).  Writes the full inverse (column-major).
MEGA_HD void m44inv_sym(const double *m, double *r)
{
    const double a00 = m[0], a01 = m[1], a02 = m[2], a03 = m[3], a11 = m[4], a12 = m[5], a13 = m[6],
                 a22 = m[7], a23 = m[8], a33 = m[9];
    // 2x2 minors of the lower two rows
    const double s0 = a00 * a11 - a01 * a01, s1 = a00 * a12 - a02 * a01, s2 = a00 * a13 - a03 * a01;
    const double s3 = a01 * a12 - a02 * a11, s4 = a01 * a13 - a03 * a11, s5 = a02 * a13 - a03 * a12;
    const double c5 = a22 * a33 - a23 * a23, c4 = a12 * a33 - a13 * a23, c3 = a12 * a23 - a13 * a22;
    const double c2 = a02 * a33 - a03 * a23, c1 = a02 * a23 - a03 * a22, c0 = a02 * a13 - a03 * a12;
    const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    const double id = 1.0 / det;
    r[0]  = ( a11 * c5 - a12 * c4 + a13 * c3) * id;
    r[4]  = (-a01 * c5 + a02 * c4 - a03 * c3) * id;
    r[8]  = ( a13 * s5 - a23 * s4 + a33 * s3) * id;
    r[12] = (-a12 * s5 + a22 * s4 - a23 * s3) * id;
    r[1]  = (-a01 * c5 + a12 * c2 - a13 * c1) * id;
    r[5]  = ( a00 * c5 - a02 * c2 + a03 * c1) * id;
    r[9]  = (-a03 * s5 + a23 * s2 - a33 * s1) * id;
    r[13] = ( a02 * s5 - a22 * s2 + a23 * s1) * id;
    r[2]  = ( a01 * c4 - a11 * c2 + a13 * c0) * id;
    r[6]  = (-a00 * c4 + a01 * c2 - a03 * c0) * id;
    r[10] = ( a03 * s4 - a13 * s2 + a33 * s0) * id;
    r[14] = (-a02 * s4 + a12 * s2 - a23 * s0) * id;
    r[3]  = (-a01 * c3 + a11 * c1 - a12 * c0) * id;
    r[7]  = ( a00 * c3 - a01 * c1 + a02 * c0) * id;
    r[11] = (-a03 * s3 + a13 * s1 - a23 * s0) * id;
    r[15] = ( a02 * s3 - a12 * s1 + a22 * s0) * id;
}

// HUGHES_WINGET (HUGHES_WINGET.F90:16-81): A = I + L/2, G = L INV3(A),
// W = skew G, R = I + INV3(I - W/2) W, strain = Voigt(sym G) with doubled
// shears, d = Voigt(sym G).
MEGA_HD void hughes_winget(const double *L, double *R, double *de, double *d)
{
#pragma clang fp contract(off)
    double A[9], Ai[9], G[9], W[9], IW[9], IWi[9], T[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) A[i + 3 * j] = (i == j ? 1.0 : 0.0) + 0.5 * L[i + 3 * j];
    inv3_ref(A, Ai);
    mm3(L, Ai, G);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            W[i + 3 * j] = 0.5 * (G[i + 3 * j] - G[j + 3 * i]);
            IW[i + 3 * j] = (i == j ? 1.0 : 0.0) - 0.5 * W[i + 3 * j];
        }
    inv3_ref(IW, IWi);
    mm3(IWi, W, T);
#pragma unroll
    for (int e = 0; e < 9; ++e) R[e] = T[e];
    R[0] += 1.0; R[4] += 1.0; R[8] += 1.0;
    d[0] = G[0]; d[1] = G[4]; d[2] = G[8];
    d[3] = 0.5 * (G[7] + G[5]);
    d[4] = 0.5 * (G[6] + G[2]);
    d[5] = 0.5 * (G[3] + G[1]);
    de[0] = d[0]; de[1] = d[1]; de[2] = d[2];
    de[3] = 2.0 * d[3]; de[4] = 2.0 * d[4]; de[5] = 2.0 * d[5];
}

// ROTATE_TENSOR: v <- R T(v) R^T (ROTATE_TENSOR.F90:21-87), Voigt XX,YY,ZZ,YZ,XZ,XY
MEGA_HD void rotate_voigt(const double *R, double *v)
{
#pragma clang fp contract(off)
    double T[9] = {v[0], v[5], v[4], v[5], v[1], v[3], v[4], v[3], v[2]};
    double RT[9], RTRt[9], Rt[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) Rt[i + 3 * j] = R[j + 3 * i];
    mm3(R, T, RT);
    mm3(RT, Rt, RTRt);
    v[0] = RTRt[0]; v[1] = RTRt[4]; v[2] = RTRt[8];
    v[3] = RTRt[7]; v[4] = RTRt[6]; v[5] = RTRt[3];
}

// (1 + B eps)^CE with the exponents the linear law uses taken exactly
template <class T>
MEGA_HD T hard_pow(T base, T e)
{
#pragma clang fp contract(off)
    if (e == T(1)) return base;
    if (e == T(0)) return T(1);
    return pow(base, e);
}

// VON_MISES radial return with power-law hardening
// (CONSTITUTE_VONMISES.F90:21-158).  Returns Newton iterations, or -1 where the
// reference stops with 'PLASTICITY MATERIAL MODEL DID NOT CONVERGE'.
template <class T>
MEGA_HD int von_mises(T *stress, const T *pred, T &eps, const Material &m)
{
#pragma clang fp contract(off)
    const T mu = T(m.mu), sig0 = T(m.sig0), B = T(m.B), CE = T(m.CE);
    const T kk = pred[0] + pred[1] + pred[2];
    const T p3 = kk / T(3);
    T dev[6] = {pred[0] - p3, pred[1] - p3, pred[2] - p3, pred[3], pred[4], pred[5]};
    const T tn = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                      T(2) * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
    const T r23 = T(sqrt(2.0 / 3.0));
    const T sig23 = r23 * sig0;   // DSQRT(2/3) sigma_0
    T xk = sig23 * hard_pow(T(1) + B * eps, CE);
    T yld = tn - xk;
    int iters = 0;
    if (yld > T(0)) {
        T dg = T(0);
        const T epst = eps;
        const T tol = sig23 * T(kNewtonTol);
        for (int it = 1; it <= 20; ++it) {
            xk = sig23 * hard_pow(T(1) + B * eps, CE);
            yld = tn - xk - (T(2) * mu * dg);
            if (fabs(yld) > tol) {
                const T xkp = B * CE * sig0 * hard_pow(T(1) + B * eps, CE - T(1));
                const T dy = T(-2) * mu * (T(1) + xkp / T(3) / mu);
                dg = dg - yld / dy;
                iters = it;
                if (it == 20) return -1;
            } else {
                eps = epst + r23 * dg;
                break;
            }
            eps = epst + r23 * dg;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) dev[k] = dev[k] - dg * T(2) * mu * dev[k] / tn;   // the reference's order
    }
    stress[0] = p3 + dev[0];
    stress[1] = p3 + dev[1];
    stress[2] = p3 + dev[2];
    stress[3] = dev[3];
    stress[4] = dev[4];
    stress[5] = dev[5];
    return iters;
}

// The damage models (CONSTITUTION.F90:58-75) work in effective stress space:
// eff = (rotated stress of the previous step) / (1 - D) + C : (objective
// strain increment), C from props 28-30 (FORM_CMAT).
template <class T>
MEGA_HD void effective_trial(const T *stress, T dam, const T *inc, const Material &m, T *eff)
{
#pragma clang fp contract(off)
    const T lam = T(m.lam), mu = T(m.mu), l2m = T(m.l2m);
#pragma unroll
    for (int k = 0; k < 6; ++k) eff[k] = stress[k] / (T(1) - dam);
    eff[0] = eff[0] + (l2m * inc[0] + lam * inc[1] + lam * inc[2]);
    eff[1] = eff[1] + (lam * inc[0] + l2m * inc[1] + lam * inc[2]);
    eff[2] = eff[2] + (lam * inc[0] + lam * inc[1] + l2m * inc[2]);
    eff[3] = eff[3] + mu * inc[3];
    eff[4] = eff[4] + mu * inc[4];
    eff[5] = eff[5] + mu * inc[5];
}

// Isotropic damage from the norm of the total strain (TENSOR_NORM, shear
// terms doubled), irreversible through DAM_SV, capped at DAM_MAX
// (CONSTITUTE_VONMISES_DAM.F90:86-120, CONSTITUTE_DRUCKPRAG.F90:88-118).
template <class T>
MEGA_HD void damage_update(const T *total, T &dam, T &dam_sv, const Material &m)
{
#pragma clang fp contract(off)
    const T k_i = T(m.k_i), k_c = T(m.k_c);
    const T eta = sqrt(total[0] * total[0] + total[1] * total[1] + total[2] * total[2] +
                       T(2) * (total[3] * total[3] + total[4] * total[4] + total[5] * total[5]));
    if (eta < dam_sv + k_i) {
        // no new damage
    } else if (eta > k_i && eta < k_c) {
        dam = k_c * (eta - k_i) / (eta * (k_c - k_i));
        dam_sv = eta - k_i;
    }
    if (dam > T(m.dam_max)) dam = T(m.dam_max);
}

// VON_MISES_DAM (CONSTITUTE_VONMISES_DAM.F90:14-196): the radial return of
// VON_MISES in effective stress space, then sigma = (1 - D) sigma_eff.
// stress: in, the rotated stress of the previous step; out, the new stress.
// Returns Newton iterations, or -1 where the reference stops.
template <class T>
MEGA_HD int von_mises_dam(T *stress, const T *total, const T *inc, T &eps, T &dam, T &dam_sv, const Material &m)
{
#pragma clang fp contract(off)
    const T mu = T(m.shear), sig0 = T(m.sig0), B = T(m.B), CE = T(m.CE);
    T eff[6];
    effective_trial(stress, dam, inc, m, eff);
    const T kk = eff[0] + eff[1] + eff[2];
    T dev[6] = {eff[0] - kk / T(3), eff[1] - kk / T(3), eff[2] - kk / T(3), eff[3], eff[4], eff[5]};
    damage_update(total, dam, dam_sv, m);
    const T tn = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                      T(2) * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
    const T r23 = T(sqrt(2.0 / 3.0));
    const T sig23 = r23 * sig0;
    T xk = sig23 * hard_pow(T(1) + B * eps, CE);
    T yld = tn - xk;
    int iters = 0;
    if (yld > T(0)) {
        T dg = T(0);
        const T epst = eps;
        const T tol = sig23 * T(kNewtonTol);
        for (int it = 1; it <= 20; ++it) {
            xk = sig23 * hard_pow(T(1) + B * eps, CE);
            yld = tn - xk - (T(2) * mu * dg);
            if (fabs(yld) > tol) {
                const T xkp = B * CE * sig0 * hard_pow(T(1) + B * eps, CE - T(1));
                const T dy = T(-2) * mu * (T(1) + xkp / T(3) / mu);
                dg = dg - yld / dy;
                iters = it;
                if (it == 20) return -1;
            } else {
                eps = epst + r23 * dg;
                break;
            }
            eps = epst + r23 * dg;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) dev[k] = dev[k] - dg * T(2) * mu * dev[k] / tn;
    }
    const T one_d = T(1) - dam;
    const T p3 = -(-kk / T(3));
#pragma unroll
    for (int k = 0; k < 3; ++k) stress[k] = (p3 + dev[k]) * one_d;
#pragma unroll
    for (int k = 3; k < 6; ++k) stress[k] = (T(-0.0) + dev[k]) * one_d;
    return iters;
}

// DRUCK_PRAG (CONSTITUTE_DRUCKPRAG.F90:14-241): Drucker-Prager shear return in
// effective stress space with isotropic damage; props 25-27 as the reader
// prepares them.  The reference never enters its tension cut-off branch.
// Returns 1 where the shear surface was crossed.
template <class T>
MEGA_HD int druck_prag(T *stress, const T *total, const T *inc, T &eps_shear, T &eps, T &dam, T &dam_sv,
                       const Material &m)
{
#pragma clang fp contract(off)
    const T q_phi = T(m.q_phi), k_phi = T(m.k_phi), q_psi = T(m.q_psi);
    T eff[6];
    effective_trial(stress, dam, inc, m, eff);
    damage_update(total, dam, dam_sv, m);
    const T kk = eff[0] + eff[1] + eff[2];
    const T dev[6] = {eff[0] - kk / T(3), eff[1] - kk / T(3), eff[2] - kk / T(3), eff[3], eff[4], eff[5]};
    const T tau_pre = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                           T(2) * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
    const T sph_pre = (eff[0] + eff[1] + eff[2]) / T(3);
    const T f = tau_pre + q_phi * sph_pre - k_phi;
    const T one_d = T(1) - dam;
    if (f < T(0)) {
#pragma unroll
        for (int k = 0; k < 6; ++k) stress[k] = eff[k] * one_d;
        return 0;
    }
    const T tau = k_phi - q_phi * sph_pre;
    const T bulk = T(m.bulk), shear = T(m.shear);
    const T gamma = f / (T(2) * shear + bulk * q_phi * q_psi);
    const T sph = sph_pre - bulk * q_psi * gamma;
#pragma unroll
    for (int k = 0; k < 6; ++k) stress[k] = dev[k] * tau / tau_pre;
#pragma unroll
    for (int k = 0; k < 3; ++k) stress[k] = stress[k] + sph;
#pragma unroll
    for (int k = 0; k < 6; ++k) stress[k] = stress[k] * one_d;
    eps_shear = eps_shear + gamma * T(sqrt(2.0 / 3.0 + 2.0 / 9.0 * (m.q_psi * m.q_psi)));
    eps = eps + eps_shear;
    return 1;
}

// trial yield value ||dev pred|| - sqrt(2/3) sigma_0 (1 + B eps)^CE, the test
// that opens the return map (CONSTITUTE_VONMISES.F90:87-100)
template <class T>
MEGA_HD T trial_excess(const T *pred, T eps, const Material &m)
{
#pragma clang fp contract(off)
    const T kk = pred[0] + pred[1] + pred[2];
    const T p3 = kk / T(3);
    const T d0 = pred[0] - p3, d1 = pred[1] - p3, d2 = pred[2] - p3;
    const T tn = sqrt(d0 * d0 + d1 * d1 + d2 * d2 + T(2) * (pred[3] * pred[3] + pred[4] * pred[4] + pred[5] * pred[5]));
    return tn - T(sqrt(2.0 / 3.0)) * T(m.sig0) * hard_pow(T(1) + T(m.B) * eps, T(m.CE));
}

// elastic trial: pred = stress + C:de with props 28-30 (FINT.F90:789-794, FORM_CMAT)
template <class T>
MEGA_HD void elastic_trial(const T *stress, const T *de, const Material &m, T *pred)
{
#pragma clang fp contract(off)
    const T lam = T(m.lam), mu = T(m.mu), l2m = T(m.l2m);
    pred[0] = stress[0] + (l2m * de[0] + lam * de[1] + lam * de[2]);
    pred[1] = stress[1] + (lam * de[0] + l2m * de[1] + lam * de[2]);
    pred[2] = stress[2] + (lam * de[0] + lam * de[1] + l2m * de[2]);
    pred[3] = stress[3] + mu * de[3];
    pred[4] = stress[4] + mu * de[4];
    pred[5] = stress[5] + mu * de[5];
}

template <class T>
MEGA_HD T tnorm(const T *v)
{
#pragma clang fp contract(off)
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + T(2) * (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]));
}
template <class T>
MEGA_HD T tnorm_strain(const T *v)
{
#pragma clang fp contract(off)
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]));
}

// Residual of the radial-return equations that VON_MISES solves
// (CONSTITUTE_VONMISES.F90:100-157) at a candidate answer (s, eps) for the
// trial state (pred, eps_n), in units of sigma_y(eps):
//   yield condition     sqrt(3/2) ||dev s|| - sigma_y(eps)
//   flow direction      dev s - (||dev s|| / ||dev pred||) dev pred
//   pressure            tr(s - pred) / 3
//   plastic multiplier  ||dev pred|| - ||dev s|| - 2 mu sqrt(3/2) (eps - eps_n)
// the largest magnitude of the four.  The exact answer has all four zero.
template <class T>
MEGA_HD T return_residual(const T *pred, const T *s, T eps_n, T eps, const Material &m)
{
    const T pp = (pred[0] + pred[1] + pred[2]) / T(3), ps = (s[0] + s[1] + s[2]) / T(3);
    const T a[6] = {pred[0] - pp, pred[1] - pp, pred[2] - pp, pred[3], pred[4], pred[5]};
    const T b[6] = {s[0] - ps, s[1] - ps, s[2] - ps, s[3], s[4], s[5]};
    const T na = tnorm(a), nb = tnorm(b);
    const T sy = T(m.sig0) * hard_pow(T(1) + T(m.B) * eps, T(m.CE));
    const T r = nb / na;
    T c[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = b[k] - r * a[k];
    const T ryield = fabs(sqrt(T(1.5)) * nb - sy);
    const T rdir = tnorm(c);
    const T rpres = fabs(ps - pp);
    const T rmult = fabs(na - nb - T(2) * T(m.mu) * sqrt(T(1.5)) * (eps - eps_n));
    const T res = fmax(fmax(ryield, rdir), fmax(rpres, rmult)) / sy;
    return res == res ? res : T(1e30);   // NaN never passes
}

// GMAX_WVEL from secant moduli (FINT.F90:850-870, ESTIMATE_MODULI.F90:16-86)
MEGA_HD double wave_speed(const double *si, const double *ei, const Material &m)
{
#pragma clang fp contract(off)
    double shear = m.shear, bulk = m.bulk;
    if (m.type > 1) {
        const double sk = (si[0] + si[1] + si[2]) / 3.0, ek = (ei[0] + ei[1] + ei[2]) / 3.0;
        double sd[6] = {si[0] - sk, si[1] - sk, si[2] - sk, si[3], si[4], si[5]};
        double ss[6] = {-sk, -sk, -sk, 0, 0, 0};
        double ed[6] = {ei[0] - ek, ei[1] - ek, ei[2] - ek, ei[3], ei[4], ei[5]};
        double es[6] = {-ek, -ek, -ek, 0, 0, 0};
        const double nsd = tnorm(sd), nss = tnorm(ss), ned = tnorm_strain(ed), nes = tnorm_strain(es);
        double st = (ned > 1.0e-8 && nsd > 1.0e-8) ? nsd / ned / 2.0 : shear;
        st = fmin(fmax(st, shear * 0.5), shear);
        double bt = (nes > 1.0e-8 && nss > 1.0e-8) ? nss / nes / 3.0 : bulk;
        bt = fmin(fmax(bt, bulk * 0.5), bulk);
        shear = st;
        bulk = bt;
    }
    const double pmod = bulk + 4.0 * shear / 3.0;
    return sqrt(fmax(pmod, shear) / m.rho);
}

}  // namespace mega
