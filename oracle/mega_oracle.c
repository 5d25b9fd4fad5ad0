/*
 * mega_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement, in plain C and FP64, of the reference's explicit meshfree
 * step.  Each function cites the reference lines it follows and keeps the
 * reference's operation order, so that with one thread it reproduces the
 * reference bit for bit (x86-64 without FMA contraction, as both compilers
 * build here).  It checks the HIP path (tests/), is timed as the CPU baseline
 * (bench.py) and is never linked by the product library.
 *
 * Reference quirks kept on purpose (they are the reference's behaviour):
 *  - INV3 returns adj(A)*det(A), not adj(A)/det(A)            INVERSE.F90:132-147
 *  - DETERMINANT returns |det|                                  DETERMINANT.F90:250
 *  - y/z kernel derivatives divide by the x-window              RK.f90:191-205
 *  - singular-kernel epsilon is the REAL(4) literal 1.0E-015    RK.f90:219
 *  - QL adds to M and H without its derivative terms            RK.f90:251-303,357-377
 *  - the J2 Newton tolerance uses the REAL(4) literal 1.0E-05   CONSTITUTE_VONMISES.F90:109
 *  - semi-Lagrangian integration weight stays V0 (J = 1)        FINT.F90:230
 */
#include "mega_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CM3(a, i, j) (a)[(i) + 3 * (j)]   /* 0-based column-major 3x3 */
#define CM4(a, i, j) (a)[(i) + 4 * (j)]

/* MLS_KERNEL0, cubic B-spline branch (KERNEL.F90:57-71). */
void or_kernel(double xsa, double *phi, double *phix)
{
    if (xsa <= 0.5) {
        *phi = 2.0 / 3.0 - 4.0 * (xsa * xsa) + 4.0 * (xsa * xsa * xsa);
        *phix = -8.0 * xsa + 12.0 * (xsa * xsa);
    } else if (xsa <= 1.0) {
        *phi = 4.0 / 3.0 - 4.0 * xsa + 4.0 * (xsa * xsa) - 4.0 / 3.0 * (xsa * xsa * xsa);
        *phix = -4.0 + 8.0 * xsa - 4.0 * (xsa * xsa);
    } else {
        *phi = 0.0;
        *phix = 0.0;
    }
}

/* M44INV: 4x4 inverse by cofactors (INVERSE.F90:185-223), same expressions. */
void or_m44inv(const double *a, double *ainv)
{
#define A(i, j) CM4(a, (i) - 1, (j) - 1)
    double det =
        A(1,1)*(A(2,2)*(A(3,3)*A(4,4)-A(3,4)*A(4,3))+A(2,3)*(A(3,4)*A(4,2)-A(3,2)*A(4,4))+A(2,4)*(A(3,2)*A(4,3)-
        A(3,3)*A(4,2)))-A(1,2)*(A(2,1)*(A(3,3)*A(4,4)-A(3,4)*A(4,3))+A(2,3)*(A(3,4)*A(4,1)-A(3,1)*A(4,4))+
        A(2,4)*(A(3,1)*A(4,3)-A(3,3)*A(4,1)))+A(1,3)*(A(2,1)*(A(3,2)*A(4,4)-A(3,4)*A(4,2))+A(2,2)*(A(3,4)*A(4,1)-
        A(3,1)*A(4,4))+A(2,4)*(A(3,1)*A(4,2)-A(3,2)*A(4,1)))-A(1,4)*(A(2,1)*(A(3,2)*A(4,3)-A(3,3)*A(4,2))+
        A(2,2)*(A(3,3)*A(4,1)-A(3,1)*A(4,3))+A(2,3)*(A(3,1)*A(4,2)-A(3,2)*A(4,1)));
    double cf[16];
#define C(i, j) CM4(cf, (i) - 1, (j) - 1)
    C(1,1) = A(2,2)*(A(3,3)*A(4,4)-A(3,4)*A(4,3))+A(2,3)*(A(3,4)*A(4,2)-A(3,2)*A(4,4))+A(2,4)*(A(3,2)*A(4,3)-A(3,3)*A(4,2));
    C(1,2) = A(2,1)*(A(3,4)*A(4,3)-A(3,3)*A(4,4))+A(2,3)*(A(3,1)*A(4,4)-A(3,4)*A(4,1))+A(2,4)*(A(3,3)*A(4,1)-A(3,1)*A(4,3));
    C(1,3) = A(2,1)*(A(3,2)*A(4,4)-A(3,4)*A(4,2))+A(2,2)*(A(3,4)*A(4,1)-A(3,1)*A(4,4))+A(2,4)*(A(3,1)*A(4,2)-A(3,2)*A(4,1));
    C(1,4) = A(2,1)*(A(3,3)*A(4,2)-A(3,2)*A(4,3))+A(2,2)*(A(3,1)*A(4,3)-A(3,3)*A(4,1))+A(2,3)*(A(3,2)*A(4,1)-A(3,1)*A(4,2));
    C(2,1) = A(1,2)*(A(3,4)*A(4,3)-A(3,3)*A(4,4))+A(1,3)*(A(3,2)*A(4,4)-A(3,4)*A(4,2))+A(1,4)*(A(3,3)*A(4,2)-A(3,2)*A(4,3));
    C(2,2) = A(1,1)*(A(3,3)*A(4,4)-A(3,4)*A(4,3))+A(1,3)*(A(3,4)*A(4,1)-A(3,1)*A(4,4))+A(1,4)*(A(3,1)*A(4,3)-A(3,3)*A(4,1));
    C(2,3) = A(1,1)*(A(3,4)*A(4,2)-A(3,2)*A(4,4))+A(1,2)*(A(3,1)*A(4,4)-A(3,4)*A(4,1))+A(1,4)*(A(3,2)*A(4,1)-A(3,1)*A(4,2));
    C(2,4) = A(1,1)*(A(3,2)*A(4,3)-A(3,3)*A(4,2))+A(1,2)*(A(3,3)*A(4,1)-A(3,1)*A(4,3))+A(1,3)*(A(3,1)*A(4,2)-A(3,2)*A(4,1));
    C(3,1) = A(1,2)*(A(2,3)*A(4,4)-A(2,4)*A(4,3))+A(1,3)*(A(2,4)*A(4,2)-A(2,2)*A(4,4))+A(1,4)*(A(2,2)*A(4,3)-A(2,3)*A(4,2));
    C(3,2) = A(1,1)*(A(2,4)*A(4,3)-A(2,3)*A(4,4))+A(1,3)*(A(2,1)*A(4,4)-A(2,4)*A(4,1))+A(1,4)*(A(2,3)*A(4,1)-A(2,1)*A(4,3));
    C(3,3) = A(1,1)*(A(2,2)*A(4,4)-A(2,4)*A(4,2))+A(1,2)*(A(2,4)*A(4,1)-A(2,1)*A(4,4))+A(1,4)*(A(2,1)*A(4,2)-A(2,2)*A(4,1));
    C(3,4) = A(1,1)*(A(2,3)*A(4,2)-A(2,2)*A(4,3))+A(1,2)*(A(2,1)*A(4,3)-A(2,3)*A(4,1))+A(1,3)*(A(2,2)*A(4,1)-A(2,1)*A(4,2));
    C(4,1) = A(1,2)*(A(2,4)*A(3,3)-A(2,3)*A(3,4))+A(1,3)*(A(2,2)*A(3,4)-A(2,4)*A(3,2))+A(1,4)*(A(2,3)*A(3,2)-A(2,2)*A(3,3));
    C(4,2) = A(1,1)*(A(2,3)*A(3,4)-A(2,4)*A(3,3))+A(1,3)*(A(2,4)*A(3,1)-A(2,1)*A(3,4))+A(1,4)*(A(2,1)*A(3,3)-A(2,3)*A(3,1));
    C(4,3) = A(1,1)*(A(2,4)*A(3,2)-A(2,2)*A(3,4))+A(1,2)*(A(2,1)*A(3,4)-A(2,4)*A(3,1))+A(1,4)*(A(2,2)*A(3,1)-A(2,1)*A(3,2));
    C(4,4) = A(1,1)*(A(2,2)*A(3,3)-A(2,3)*A(3,2))+A(1,2)*(A(2,3)*A(3,1)-A(2,1)*A(3,3))+A(1,3)*(A(2,1)*A(3,2)-A(2,2)*A(3,1));
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            CM4(ainv, i, j) = CM4(cf, j, i) / det;
#undef C
#undef A
}

/* INV3 (INVERSE.F90:119-152): adj(A)*det(A), kept as the reference has it. */
void or_inv3(const double *a, double *ainv)
{
#define A(i, j) CM3(a, (i) - 1, (j) - 1)
#define R(i, j) CM3(ainv, (i) - 1, (j) - 1)
    double det = A(1,1)*A(2,2)*A(3,3) - A(1,1)*A(2,3)*A(3,2) - A(1,2)*A(2,1)*A(3,3)
               + A(1,2)*A(2,3)*A(3,1) + A(1,3)*A(2,1)*A(3,2) - A(1,3)*A(2,2)*A(3,1);
    R(1,1) =  (A(2,2)*A(3,3)-A(2,3)*A(3,2))*det;
    R(2,1) = -(A(2,1)*A(3,3)-A(2,3)*A(3,1))*det;
    R(3,1) =  (A(2,1)*A(3,2)-A(2,2)*A(3,1))*det;
    R(1,2) = -(A(1,2)*A(3,3)-A(1,3)*A(3,2))*det;
    R(2,2) =  (A(1,1)*A(3,3)-A(1,3)*A(3,1))*det;
    R(3,2) = -(A(1,1)*A(3,2)-A(1,2)*A(3,1))*det;
    R(1,3) =  (A(1,2)*A(2,3)-A(1,3)*A(2,2))*det;
    R(2,3) = -(A(1,1)*A(2,3)-A(1,3)*A(2,1))*det;
    R(3,3) =  (A(1,1)*A(2,2)-A(1,2)*A(2,1))*det;
#undef R
#undef A
}

/* INVERSE: Gauss-Jordan with full pivoting (INVERSE.F90:17-99).
 * Returns 1 on success, 0 where the reference WARNs 'PROBLEM INVERTING MATRIX'
 * (the partially reduced matrix is left in ainv, as the reference leaves it). */
int or_inverse(const double *a, int n, double *ainv)
{
    int is[16], js[16];
    double t, d;
    memcpy(ainv, a, sizeof(double) * n * n);
#define V(i, j) ainv[((i) - 1) + n * ((j) - 1)]
    for (int k = 1; k <= n; ++k) {
        d = 0.0;
        for (int i = k; i <= n; ++i)
            for (int j = k; j <= n; ++j)
                if (fabs(V(i, j)) > d) { d = fabs(V(i, j)); is[k - 1] = i; js[k - 1] = j; }
        if (d + 1.0 == 1.0) return 0;
        for (int j = 1; j <= n; ++j) { t = V(k, j); V(k, j) = V(is[k - 1], j); V(is[k - 1], j) = t; }
        for (int i = 1; i <= n; ++i) { t = V(i, k); V(i, k) = V(i, js[k - 1]); V(i, js[k - 1]) = t; }
        V(k, k) = 1.0 / V(k, k);
        for (int j = 1; j <= n; ++j) if (j != k) V(k, j) = V(k, j) * V(k, k);
        for (int i = 1; i <= n; ++i)
            if (i != k)
                for (int j = 1; j <= n; ++j)
                    if (j != k) V(i, j) = V(i, j) - V(i, k) * V(k, j);
        for (int i = 1; i <= n; ++i) if (i != k) V(i, k) = -V(i, k) * V(k, k);
    }
    for (int k = n; k >= 1; --k) {
        for (int j = 1; j <= n; ++j) { t = V(k, j); V(k, j) = V(js[k - 1], j); V(js[k - 1], j) = t; }
        for (int i = 1; i <= n; ++i) { t = V(i, k); V(i, k) = V(i, is[k - 1]); V(i, is[k - 1]) = t; }
    }
#undef V
    return 1;
}

/* DETERMINANT: |det(A)| (DETERMINANT.F90:18-35). */
double or_determinant(const double *a)
{
#define A(i, j) CM3(a, (i) - 1, (j) - 1)
    return fabs(A(1,1)*(A(2,2)*A(3,3) - A(3,2)*A(2,3))
              + A(1,2)*(A(3,1)*A(2,3) - A(2,1)*A(3,3))
              + A(1,3)*(A(2,1)*A(3,2) - A(3,1)*A(2,2)));
#undef A
}

static void mm3(const double *a, const double *b, double *c)   /* MATMUL, 3x3 */
{
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += CM3(a, i, j) * CM3(b, j, k);
            CM3(c, i, k) = s;
        }
}

/* HUGHES_WINGET (HUGHES_WINGET.F90:16-81). rot is 3x3 column-major. */
void or_hughes_winget(const double *lmat, double *rot, double *strain, double *d)
{
    double A[9], Ai[9], G[9], W[9], IW[9], IWi[9], T[9];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            CM3(A, i, j) = (i == j ? 1.0 : 0.0) + 0.5 * CM3(lmat, i, j);
    or_inv3(A, Ai);
    mm3(lmat, Ai, G);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) CM3(W, i, j) = 0.5 * (CM3(G, i, j) - CM3(G, j, i));
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) CM3(IW, i, j) = (i == j ? 1.0 : 0.0) - 0.5 * CM3(W, i, j);
    or_inv3(IW, IWi);
    mm3(IWi, W, T);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) CM3(rot, i, j) = (i == j ? 1.0 : 0.0) + CM3(T, i, j);
    /* E = 0.5 (G + G^T); TENSOR_2_VTENSOR order XX,YY,ZZ,YZ,XZ,XY (FINT_FUNCTIONS.F90:56-62) */
    double e11 = 0.5 * (CM3(G, 0, 0) + CM3(G, 0, 0));
    double e22 = 0.5 * (CM3(G, 1, 1) + CM3(G, 1, 1));
    double e33 = 0.5 * (CM3(G, 2, 2) + CM3(G, 2, 2));
    double e23 = 0.5 * (CM3(G, 1, 2) + CM3(G, 2, 1));
    double e13 = 0.5 * (CM3(G, 0, 2) + CM3(G, 2, 0));
    double e12 = 0.5 * (CM3(G, 0, 1) + CM3(G, 1, 0));
    d[0] = e11; d[1] = e22; d[2] = e33; d[3] = e23; d[4] = e13; d[5] = e12;
    strain[0] = e11; strain[1] = e22; strain[2] = e33;
    strain[3] = 2.0 * e23; strain[4] = 2.0 * e13; strain[5] = 2.0 * e12;
}

/* ROTATE_TENSOR: v <- R T(v) R^T in Voigt XX,YY,ZZ,YZ,XZ,XY (ROTATE_TENSOR.F90:21-87). */
void or_rotate_tensor(const double *rot, double *v)
{
    double T[9], RT[9], Rt[9], TR[9];
    CM3(T, 0, 0) = v[0]; CM3(T, 1, 1) = v[1]; CM3(T, 2, 2) = v[2];
    CM3(T, 2, 1) = v[3]; CM3(T, 1, 2) = v[3];
    CM3(T, 0, 2) = v[4]; CM3(T, 2, 0) = v[4];
    CM3(T, 0, 1) = v[5]; CM3(T, 1, 0) = v[5];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) CM3(Rt, i, j) = CM3(rot, j, i);
    mm3(rot, T, RT);
    mm3(RT, Rt, TR);
    v[0] = CM3(TR, 0, 0); v[1] = CM3(TR, 1, 1); v[2] = CM3(TR, 2, 2);
    v[3] = CM3(TR, 1, 2); v[4] = CM3(TR, 0, 2); v[5] = CM3(TR, 0, 1);
}

/* VON_MISES: J2 radial return, power-law hardening (CONSTITUTE_VONMISES.F90:21-158).
 * Returns Newton iterations used (0 if elastic) or -1 where the reference exits
 * with 'PLASTICITY MATERIAL MODEL DID NOT CONVERGE'. */
int or_von_mises(double *stress, const double *pred, double *state, const double *props)
{
    const double poiss = props[0], young = props[1], sig0 = props[3];
    const double mu = young / (2.0 * (1.0 + poiss));
    const double B = props[4], CE = props[5];
    double eps = state[2];
    double kk = pred[0] + pred[1] + pred[2];
    double dev[6], sph[6];
    dev[0] = pred[0] - kk / 3.0; dev[1] = pred[1] - kk / 3.0; dev[2] = pred[2] - kk / 3.0;
    dev[3] = pred[3]; dev[4] = pred[4]; dev[5] = pred[5];
    sph[0] = -kk / 3.0; sph[1] = -kk / 3.0; sph[2] = -kk / 3.0; sph[3] = sph[4] = sph[5] = 0.0;
    double tn = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                     2.0 * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
    double sig23 = sqrt(2.0 / 3.0) * sig0;
    double xk = sig23 * pow(1.0 + B * eps, CE);
    double yld = tn - xk;
    int iters = 0;
    if (yld > 0.0) {
        double dg = 0.0, epst = eps;
        const double tol = sig23 * (double)1.0e-05f;
        for (int i = 1; i <= 20; ++i) {
            xk = sig23 * pow(1.0 + B * eps, CE);
            yld = tn - xk - (2.0 * mu * dg);
            if (fabs(yld) > tol) {
                double xkp = B * CE * sig0 * pow(1.0 + B * eps, CE - 1.0);
                double dyld = -2.0 * mu * (1.0 + xkp / 3.0 / mu);
                dg = dg - yld / dyld;
                iters = i;
                if (i == 20) return -1;
            } else {
                eps = epst + sqrt(2.0 / 3.0) * dg;
                break;
            }
            eps = epst + sqrt(2.0 / 3.0) * dg;
        }
        for (int k = 0; k < 6; ++k) dev[k] = dev[k] - dg * 2.0 * mu * dev[k] / tn;
    }
    state[2] = eps;
    for (int k = 0; k < 6; ++k) stress[k] = -sph[k] + dev[k];
    return iters;
}

/* The effective trial stress of the damage models: the rotated stress of the
 * previous step over (1 - D), plus C : (objective strain increment), with C
 * from props 28-30 (FORM_CMAT, FINT_FUNCTIONS.F90:252-280). */
static void effective_trial(const double *stress, double dam, const double *inc, const double *props, double *eff)
{
    const double lam = props[27], mu = props[28], l2m = props[29];
    for (int k = 0; k < 6; ++k) eff[k] = stress[k] / (1.0 - dam);
    eff[0] = eff[0] + (l2m * inc[0] + lam * inc[1] + lam * inc[2]);
    eff[1] = eff[1] + (lam * inc[0] + l2m * inc[1] + lam * inc[2]);
    eff[2] = eff[2] + (lam * inc[0] + lam * inc[1] + l2m * inc[2]);
    eff[3] = eff[3] + mu * inc[3];
    eff[4] = eff[4] + mu * inc[4];
    eff[5] = eff[5] + mu * inc[5];
}

/* Isotropic damage from the norm of the total strain, irreversible through
 * DAM_SV (CONSTITUTE_VONMISES_DAM.F90:86-120, CONSTITUTE_DRUCKPRAG.F90:88-118):
 * the strain norm is TENSOR_NORM (shear terms doubled). */
static void damage_update(const double *total, double *dam, double *dam_sv, const double *props)
{
    const double k_i = props[8], k_c = props[9], dam_max = props[10];
    const double t = total[0] * total[0] + total[1] * total[1] + total[2] * total[2] +
                     2.0 * (total[3] * total[3] + total[4] * total[4] + total[5] * total[5]);
    const double eta = sqrt(t);
    if (eta < *dam_sv + k_i) {
        /* no new damage */
    } else if (eta > k_i && eta < k_c) {
        *dam = k_c * (eta - k_i) / (eta * (k_c - k_i));
        *dam_sv = eta - k_i;
    }
    if (*dam > dam_max) *dam = dam_max;
}

/* VON_MISES_DAM: J2 radial return with power-law hardening in effective
 * stress space, isotropic damage (CONSTITUTE_VONMISES_DAM.F90:14-196).
 * stress: in, the rotated stress of the previous step; out, the new stress.
 * state[2] = eps, state[3] = D, state[4] = DAM_SV.  Returns Newton
 * iterations (0 if elastic) or -1 where the reference stops. */
int or_von_mises_dam(double *stress, const double *total, const double *inc, double *state, const double *props)
{
    const double poiss = props[0], young = props[1], sig0 = props[3];
    const double mu = young / (2.0 * (1.0 + poiss));
    const double B = props[4], CE = props[5];
    double eps = state[2], dam = state[3], dam_sv = state[4];
    double eff[6], dev[6];
    effective_trial(stress, dam, inc, props, eff);
    double kk = eff[0] + eff[1] + eff[2];
    dev[0] = eff[0] - kk / 3.0; dev[1] = eff[1] - kk / 3.0; dev[2] = eff[2] - kk / 3.0;
    dev[3] = eff[3]; dev[4] = eff[4]; dev[5] = eff[5];
    damage_update(total, &dam, &dam_sv, props);
    const double sph = -kk / 3.0;
    double tn = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                     2.0 * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
    double sig23 = sqrt(2.0 / 3.0) * sig0;
    double xk = sig23 * pow(1.0 + B * eps, CE);
    double yld = tn - xk;
    int iters = 0;
    if (yld > 0.0) {
        double dg = 0.0, epst = eps;
        const double tol = sig23 * (double)1.0e-05f;
        for (int i = 1; i <= 20; ++i) {
            xk = sig23 * pow(1.0 + B * eps, CE);
            yld = tn - xk - (2.0 * mu * dg);
            if (fabs(yld) > tol) {
                double xkp = B * CE * sig0 * pow(1.0 + B * eps, CE - 1.0);
                double dyld = -2.0 * mu * (1.0 + xkp / 3.0 / mu);
                dg = dg - yld / dyld;
                iters = i;
                if (i == 20) return -1;
            } else {
                eps = epst + sqrt(2.0 / 3.0) * dg;
                break;
            }
            eps = epst + sqrt(2.0 / 3.0) * dg;
        }
        for (int k = 0; k < 6; ++k) dev[k] = dev[k] - dg * 2.0 * mu * dev[k] / tn;
    }
    state[2] = eps; state[3] = dam; state[4] = dam_sv;
    for (int k = 0; k < 3; ++k) stress[k] = -sph + dev[k];
    for (int k = 3; k < 6; ++k) stress[k] = -0.0 + dev[k];
    for (int k = 0; k < 6; ++k) stress[k] = stress[k] * (1.0 - dam);
    return iters;
}

/* DRUCK_PRAG: Drucker-Prager shear return with isotropic damage
 * (CONSTITUTE_DRUCKPRAG.F90:14-241).  props 25-27 = Q_PHI, K_PHI, Q_PSI as the
 * reader prepares them; the tension cut-off branch is never entered there.
 * state[0] = EPS_SHEAR, [1] = EPS_VOL, [2] = EPS, [3] = D, [4] = DAM_SV.
 * Returns 1 if the shear surface was crossed, else 0. */
int or_druck_prag(double *stress, const double *total, const double *inc, double *state, const double *props)
{
    const double poiss = props[0], young = props[1];
    const double q_phi = props[24], k_phi = props[25], q_psi = props[26];
    double eps_shear = state[0], eps = state[2], dam = state[3], dam_sv = state[4];
    double eff[6], dev[6];
    effective_trial(stress, dam, inc, props, eff);
    damage_update(total, &dam, &dam_sv, props);
    double kk = eff[0] + eff[1] + eff[2];
    dev[0] = eff[0] - kk / 3.0; dev[1] = eff[1] - kk / 3.0; dev[2] = eff[2] - kk / 3.0;
    dev[3] = eff[3]; dev[4] = eff[4]; dev[5] = eff[5];
    const double tau_pre = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                                2.0 * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
    const double sph_pre = (eff[0] + eff[1] + eff[2]) / 3.0;
    const double f = tau_pre + q_phi * sph_pre - k_phi;
    int sheared = 0;
    if (f < 0.0) {
        for (int k = 0; k < 6; ++k) stress[k] = eff[k];
    } else {
        sheared = 1;
        const double tau = k_phi - q_phi * sph_pre;
        const double bulk = young / (3.0 * (1.0 - 2.0 * poiss));
        const double shear = young / (2.0 * (1.0 + poiss));
        const double gamma = f / (2.0 * shear + bulk * q_phi * q_psi);
        const double sph = sph_pre - bulk * q_psi * gamma;
        for (int k = 0; k < 6; ++k) stress[k] = dev[k] * tau / tau_pre;
        for (int k = 0; k < 3; ++k) stress[k] = stress[k] + sph;
        eps_shear = eps_shear + gamma * sqrt(2.0 / 3.0 + 2.0 / 9.0 * (q_psi * q_psi));
        eps = eps + eps_shear;
    }
    for (int k = 0; k < 6; ++k) stress[k] = stress[k] * (1.0 - dam);
    state[0] = eps_shear; state[2] = eps; state[3] = dam; state[4] = dam_sv;
    return sheared;
}

static double tnorm_strain(const double *v)
{
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]));
}
static double tnorm(const double *v)
{
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 2.0 * (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]));
}
static void dev_sph(const double *v, double *dv, double *sv)
{
    double kk = v[0] + v[1] + v[2];
    dv[0] = v[0] - kk / 3.0; dv[1] = v[1] - kk / 3.0; dv[2] = v[2] - kk / 3.0;
    dv[3] = v[3]; dv[4] = v[4]; dv[5] = v[5];
    sv[0] = sv[1] = sv[2] = -kk / 3.0; sv[3] = sv[4] = sv[5] = 0.0;
}

/* ESTIMATE_MODULI (ESTIMATE_MODULI.F90:16-86), NSNI_FLAG = false. */
void or_estimate_moduli(const double *si, const double *ei, double *shear, double *bulk)
{
    double sd[6], ss[6], ed[6], es[6];
    dev_sph(si, sd, ss);
    dev_sph(ei, ed, es);
    double nsd = tnorm(sd), nss = tnorm(ss), ned = tnorm_strain(ed), nes = tnorm_strain(es);
    double st, bt;
    if (ned > 1.0e-8 && nsd > 1.0e-8) st = nsd / ned / 2.0; else st = *shear;
    st = fmax(st, *shear * 0.5);
    st = fmin(st, *shear);
    *shear = st;
    if (nes > 1.0e-8 && nss > 1.0e-8) bt = nss / nes / 3.0; else bt = *bulk;
    bt = fmax(bt, *bulk * 0.5);
    bt = fmin(bt, *bulk);
    *bulk = bt;
}

/* FINT.F90:758-801 for one integration point. */
int or_constitute_point(const double *lmat, double *stress, double *strain, double *state,
                        const double *props, int mat_type, int finite_strain, double *d_out)
{
    double rot[9], de[6], d[6];
    if (finite_strain) {
        or_hughes_winget(lmat, rot, de, d);
        or_rotate_tensor(rot, stress);
        or_rotate_tensor(rot, strain);
    } else {
        /* small strain: engineering strain increment straight from L (FINT.F90:772-783) */
        de[0] = CM3(lmat, 0, 0); de[1] = CM3(lmat, 1, 1); de[2] = CM3(lmat, 2, 2);
        de[3] = CM3(lmat, 2, 1) + CM3(lmat, 1, 2);
        de[4] = CM3(lmat, 2, 0) + CM3(lmat, 0, 2);
        de[5] = CM3(lmat, 1, 0) + CM3(lmat, 0, 1);
        for (int k = 0; k < 6; ++k) d[k] = 0.0;
    }
    for (int k = 0; k < 6; ++k) strain[k] = strain[k] + de[k];
    const double lam = props[27], mu = props[28], l2m = props[29];
    double pred[6];
    pred[0] = stress[0] + (l2m * de[0] + lam * de[1] + lam * de[2]);
    pred[1] = stress[1] + (lam * de[0] + l2m * de[1] + lam * de[2]);
    pred[2] = stress[2] + (lam * de[0] + lam * de[1] + l2m * de[2]);
    pred[3] = stress[3] + mu * de[3];
    pred[4] = stress[4] + mu * de[4];
    pred[5] = stress[5] + mu * de[5];
    int it = 0;
    if (mat_type == 1) {
        for (int k = 0; k < 6; ++k) stress[k] = pred[k];
    } else if (mat_type == 2) {
        it = or_von_mises(stress, pred, state, props);
    } else if (mat_type == 6) {   /* CONSTITUTION.F90:70-75: stress holds the rotated old stress */
        it = or_von_mises_dam(stress, strain, de, state, props);
    } else if (mat_type == 3) {
        or_druck_prag(stress, strain, de, state, props);
    } else {
        it = -2;
    }
    if (d_out) for (int k = 0; k < 6; ++k) d_out[k] = d[k];
    return it;
}

void or_snni_grad(const double *x, const double *win_x, double vol_x, int n, const double *coo, const double *win,
                  const int32_t *ebc_node, const int32_t *lstack, int ln, const or_control *c, double *shpd)
{
    const double thrd = 0.333333333333333;   /* THRD of PRE_MODEL_SMOOTHING_LENGTHS.F90:41 */
    const double d1 = win_x[0] * 2.0, d2 = win_x[1] * 2.0, d3 = win_x[2] * 2.0;
    const double ratio = pow(vol_x, thrd) / pow(d1 * d2 * d3, thrd);
    const double l[3] = {d1 * ratio / 2.0, d2 * ratio / 2.0, d3 * ratio / 2.0};
    const double svol = (l[0] + l[0]) * (l[1] + l[1]) * (l[2] + l[2]);
    const double aov[3] = {(l[1] + l[1]) * (l[2] + l[2]) / svol, (l[0] + l[0]) * (l[2] + l[2]) / svol,
                           (l[0] + l[0]) * (l[1] + l[1]) / svol};
    double *sp = (double *)malloc(sizeof(double) * (size_t)(ln + 1));
    double *sm = (double *)malloc(sizeof(double) * (size_t)(ln + 1));
    double *dd = (double *)malloc(sizeof(double) * 3 * (size_t)(ln + 1));
    for (int k = 0; k < 3; ++k) {
        double pp[3] = {x[0], x[1], x[2]}, pm[3] = {x[0], x[1], x[2]};
        pp[k] = x[k] + l[k];
        pm[k] = x[k] - l[k];
        or_rk1(pp, n, coo, win, ebc_node, lstack, ln, c, sp, dd);
        or_rk1(pm, n, coo, win, ebc_node, lstack, ln, c, sm, dd);
        for (int J = 0; J < ln; ++J) shpd[3 * J + k] = (sp[J] - sm[J]) * aov[k];
    }
    free(sp); free(sm); free(dd);
}

/* RK1: linear-basis reproducing-kernel shape functions and gradients at x
 * (RK.f90:16-453 with FILL_H :462, DERIV_H :506, MLS_KERNEL0 KERNEL.F90:3). */
void or_rk1(const double *x, int n, const double *coo, const double *win, const int32_t *ebc_node,
            const int32_t *lstack, int ln, const or_control *c, double *shp, double *shpd)
{
    (void)n;
    double M[16] = {0}, Mx[16] = {0}, My[16] = {0}, Mz[16] = {0}, Ms[16] = {0};
    double *phi = (double *)malloc(sizeof(double) * 4 * (size_t)(ln > 0 ? ln : 1));
    double *px = phi + ln, *py = px + ln, *pz = py + ln;
    double phi_sum = 0.0;
    static const double hx[4] = {0, 1, 0, 0}, hy[4] = {0, 0, 1, 0}, hz[4] = {0, 0, 0, 1};
    for (int I = 0; I < ln; ++I) {
        int ii = lstack[I];
        double dx = x[0] - coo[3 * ii], dy = x[1] - coo[3 * ii + 1], dz = x[2] - coo[3 * ii + 2];
        double H[4] = {1.0, dx, dy, dz};
        double w1 = win[3 * ii], w2 = win[3 * ii + 1], w3 = win[3 * ii + 2];
        double zx = dx / w1, zy = dy / w2, zz = dz / w3;
        double p, pxv, pyv, pzv;
        if (c->shsup) {
            double dia = sqrt(zx * zx + zy * zy + zz * zz), k0, k1;
            or_kernel(dia, &k0, &k1);
            double drx, dry, drz;
            if (dia <= 1.0e-13) { drx = dry = drz = 0.0; }
            else { drx = dx / (w1 * w1) / dia; dry = dy / (w2 * w2) / dia; drz = dz / (w3 * w3) / dia; }
            p = k0; pxv = k1 * drx; pyv = k1 * dry; pzv = k1 * drz;
        } else {
            double fx, fxx, fy, fyy, fz, fzz;
            or_kernel(fabs(zx), &fx, &fxx);
            or_kernel(fabs(zy), &fy, &fyy);
            or_kernel(fabs(zz), &fz, &fzz);
            p = fx * fy * fz;
            double drx = (zx == 0.0) ? 0.0 : (zx >= 0.0 ? 1.0 / w1 : -1.0 / w1);
            double dry = (zy == 0.0) ? 0.0 : (zy >= 0.0 ? 1.0 / w1 : -1.0 / w1);
            double drz = (zz == 0.0) ? 0.0 : (zz >= 0.0 ? 1.0 / w1 : -1.0 / w1);
            pxv = fxx * fy * fz * drx;
            pyv = fx * fyy * fz * dry;
            pzv = fx * fy * fzz * drz;
            if (ebc_node[ii]) {
                double wm = 1.0 / (sqrt(zx * zx + zy * zy + zz * zz) + (double)1.0e-15f);
                pxv = pxv * wm - p * wm * wm * wm * zx / w1;
                pyv = pyv * wm - p * wm * wm * wm * zy / w2;
                pzv = pzv * wm - p * wm * wm * wm * zz / w3;
                p = p * wm;
            }
        }
        phi_sum += p;
        phi[I] = p; px[I] = pxv; py[I] = pyv; pz[I] = pzv;
        for (int J = 0; J < 4; ++J)
            for (int K = 0; K < 4; ++K) {
                /* left-to-right sums, as the Fortran statement associates them */
                CM4(M, J, K) = CM4(M, J, K) + H[J] * H[K] * p;
                CM4(Mx, J, K) = CM4(Mx, J, K) + hx[J] * H[K] * p + H[J] * hx[K] * p + H[J] * H[K] * pxv;
                CM4(My, J, K) = CM4(My, J, K) + hy[J] * H[K] * p + H[J] * hy[K] * p + H[J] * H[K] * pyv;
                CM4(Mz, J, K) = CM4(Mz, J, K) + hz[J] * H[K] * p + H[J] * hz[K] * p + H[J] * H[K] * pzv;
            }
    }
    double qx[6][3];
    if (c->ql) {
        const double L = c->ql_len;
        for (int m = 0; m < 6; ++m) { qx[m][0] = x[0]; qx[m][1] = x[1]; qx[m][2] = x[2]; }
        qx[0][0] = x[0] + L; qx[1][0] = x[0] - L;
        qx[2][1] = x[1] + L; qx[3][1] = x[1] - L;
        qx[4][2] = x[2] + L; qx[5][2] = x[2] - L;
        for (int m = 0; m < 6; ++m) {
            double H[4] = {1.0, x[0] - qx[m][0], x[1] - qx[m][1], x[2] - qx[m][2]};
            for (int J = 0; J < 4; ++J)
                for (int K = 0; K < 4; ++K) CM4(Ms, J, K) += H[J] * H[K];
        }
        for (int e = 0; e < 16; ++e) M[e] = M[e] + Ms[e] * phi_sum * c->ql_coef;
    }
    double Mi[16];
    or_m44inv(M, Mi);
    double B[4], BX[4], BY[4], BZ[4];
    for (int k = 0; k < 4; ++k) B[k] = CM4(Mi, 0, k);
    /* BX = first row of (-Mi Mx) Mi */
    const double *Md[3] = {Mx, My, Mz};
    double *Bd[3] = {BX, BY, BZ};
    for (int a = 0; a < 3; ++a) {
        double T1[16], T2[16];
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 4; ++k) {
                double s = 0.0;
                for (int j = 0; j < 4; ++j) s += (-CM4(Mi, i, j)) * CM4(Md[a], j, k);
                CM4(T1, i, k) = s;
            }
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 4; ++k) {
                double s = 0.0;
                for (int j = 0; j < 4; ++j) s += CM4(T1, i, j) * CM4(Mi, j, k);
                CM4(T2, i, k) = s;
            }
        for (int k = 0; k < 4; ++k) Bd[a][k] = CM4(T2, 0, k);
    }
    double Hs[4] = {0, 0, 0, 0};
    if (c->ql)
        for (int m = 0; m < 6; ++m) {
            double Ht[4] = {1.0, x[0] - qx[m][0], x[1] - qx[m][1], x[2] - qx[m][2]};
            for (int k = 0; k < 4; ++k) Hs[k] = Hs[k] + Ht[k];
        }
    for (int I = 0; I < ln; ++I) {
        int ii = lstack[I];
        double H[4] = {1.0, x[0] - coo[3 * ii], x[1] - coo[3 * ii + 1], x[2] - coo[3 * ii + 2]};
        if (c->ql)
            for (int k = 0; k < 4; ++k) H[k] = H[k] + Hs[k] * c->ql_coef;
        double C0 = 0.0, CX = 0.0, CY = 0.0, CZ = 0.0, CX2 = 0.0, CY2 = 0.0, CZ2 = 0.0;
        for (int k = 0; k < 4; ++k) C0 += B[k] * H[k];
        for (int k = 0; k < 4; ++k) { CX += BX[k] * H[k]; CY += BY[k] * H[k]; CZ += BZ[k] * H[k]; }
        for (int k = 0; k < 4; ++k) { CX2 += B[k] * hx[k]; CY2 += B[k] * hy[k]; CZ2 += B[k] * hz[k]; }
        CX = CX + CX2; CY = CY + CY2; CZ = CZ + CZ2;
        shp[I] = C0 * phi[I];
        shpd[3 * I + 0] = CX * phi[I] + C0 * px[I];
        shpd[3 * I + 1] = CY * phi[I] + C0 * py[I];
        shpd[3 * I + 2] = CZ * phi[I] + C0 * pz[I];
    }
    free(phi);
}

/* Bin geometry from initial positions and windows (HANDELER.F90:187-252). */
void or_bin_geometry(int n, const double *coo, const double *win, double xd, double yd, double zd,
                     double *geom, int32_t *nb)
{
    double bin[3] = {0, 0, 0}, mx[3], mn[3], dd[3] = {xd, yd, zd};
    for (int k = 0; k < 3; ++k) mx[k] = mn[k] = coo[k];
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (win[3 * i + k] > bin[k]) bin[k] = win[3 * i + k];
            if (win[3 * i + k] + dd[k] > bin[k]) bin[k] = win[3 * i + k] + dd[k];
            if (coo[3 * i + k] > mx[k]) mx[k] = coo[3 * i + k];
            if (coo[3 * i + k] < mn[k]) mn[k] = coo[3 * i + k];
        }
    for (int k = 0; k < 3; ++k) {
        mn[k] = mn[k] - bin[k] * 10.0;
        mx[k] = mx[k] + bin[k] * 10.0;
        nb[k] = (int32_t)ceil((mx[k] - mn[k]) / bin[k]);
        geom[k] = mn[k];
        geom[3 + k] = bin[k];
    }
}

/* GET_BINS cell assignment (BIN_SEARCH.F90:20-42): 1-based, clamped. */
void or_get_bins(int n, const double *x, const double *geom, const int32_t *nb, int32_t *ijk)
{
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            int s = (int)floor((x[3 * i + k] - geom[k]) / geom[3 + k]) + 1;
            if (s < 1) s = 1;
            if (s > nb[k]) s = nb[k];
            ijk[3 * i + k] = s;
        }
}

/* SOFT_SEARCH + BIN_INODE (BIN_SEARCH.F90:69-133,153-328), direct nodal
 * integration (one point per node).  Candidates are visited in the reference's
 * order: ix, then iy, then iz, each bin in node order; node j is kept if x_i is
 * inside j's support (box: |x_j-x_i|_k/a_jk < 1 for all k; ellipsoid: norm < 1). */
int64_t or_soft_search(int n, const double *x, const double *win, const int32_t *ijk,
                       const int32_t *nb, int shsup, int64_t *row_ptr, int32_t *col)
{
    const int64_t nbins = (int64_t)nb[0] * nb[1] * nb[2];
    int64_t *bstart = (int64_t *)calloc((size_t)nbins + 1, sizeof(int64_t));
    int32_t *blist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbins > 0 ? nbins : 1));
    if (!bstart || !blist || !fill) { free(bstart); free(blist); free(fill); return -1; }
#define BIN_OF(i) ((int64_t)(ijk[3 * (i) + 2] - 1) * nb[0] * nb[1] + (int64_t)(ijk[3 * (i) + 1] - 1) * nb[0] + (ijk[3 * (i)] - 1))
    for (int i = 0; i < n; ++i) bstart[BIN_OF(i) + 1]++;
    for (int64_t b = 0; b < nbins; ++b) bstart[b + 1] += bstart[b];
    for (int64_t b = 0; b < nbins; ++b) fill[b] = bstart[b];
    for (int i = 0; i < n; ++i) blist[fill[BIN_OF(i)]++] = i;
    free(fill);
    for (int pass = 0; pass < (col ? 2 : 1); ++pass) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            int64_t cnt = 0, out = pass ? row_ptr[i] : 0;
            const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
            int i0 = ijk[3 * i] - 1 > 1 ? ijk[3 * i] - 1 : 1, i1 = ijk[3 * i] + 1 < nb[0] ? ijk[3 * i] + 1 : nb[0];
            int j0 = ijk[3 * i + 1] - 1 > 1 ? ijk[3 * i + 1] - 1 : 1, j1 = ijk[3 * i + 1] + 1 < nb[1] ? ijk[3 * i + 1] + 1 : nb[1];
            int k0 = ijk[3 * i + 2] - 1 > 1 ? ijk[3 * i + 2] - 1 : 1, k1 = ijk[3 * i + 2] + 1 < nb[2] ? ijk[3 * i + 2] + 1 : nb[2];
            for (int ix = i0; ix <= i1; ++ix)
                for (int iy = j0; iy <= j1; ++iy)
                    for (int iz = k0; iz <= k1; ++iz) {
                        int64_t m = (int64_t)(iz - 1) * nb[0] * nb[1] + (int64_t)(iy - 1) * nb[0] + (ix - 1);
                        for (int64_t q = bstart[m]; q < bstart[m + 1]; ++q) {
                            int j = blist[q];
                            double ax = fabs(x[3 * j] - xi) / win[3 * j];
                            double ay = fabs(x[3 * j + 1] - yi) / win[3 * j + 1];
                            double az = fabs(x[3 * j + 2] - zi) / win[3 * j + 2];
                            int add;
                            if (shsup) add = sqrt(ax * ax + ay * ay + az * az) < 1.0;
                            else add = (ax < 1.0) && (ay < 1.0) && (az < 1.0);
                            if (add) {
                                if (pass) col[out++] = j;
                                cnt++;
                            }
                        }
                    }
            if (!pass) row_ptr[i + 1] = cnt;
        }
        if (!pass) {
            row_ptr[0] = 0;
            for (int i = 0; i < n; ++i) row_ptr[i + 1] += row_ptr[i];
        }
    }
#undef BIN_OF
    free(bstart);
    free(blist);
    return row_ptr[n];
}

/* DO_INTERP (HANDELER.F90:114-133): phys_i = sum_j phi_ij coef_j, per component. */
void or_interp(int n, const int64_t *row_ptr, const int32_t *col, const double *shp,
               const double *coef, double *phys)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int64_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) s = s + shp[p] * coef[3 * col[p] + k];
            phys[3 * i + k] = s;
        }
}

/* CONSTRUCT_FINT (FINT.F90:16-1247), direct nodal integration, material types
 * 1 (elastic) and 2 (J2).  Lagrangian: shapes stored at the LINIT call and
 * recalled after, pushed forward by F^-1 (bug-compatible INV3).  Semi-Lagrangian:
 * RK1 at the current positions every call, shapes stored for DO_INTERP. */
int or_construct_fint(or_fint_io *io, const or_control *c)
{
    const int n = io->n;
    const int nth = io->nthreads > 0 ? io->nthreads : 1;
    for (int64_t i = 0; i < 3 * (int64_t)n; ++i) { io->fint[i] = 0.0; io->fext[i] = 0.0; }
    double work = 0.0;
    int64_t fails = 0;
    int bad_type = 0;
    int64_t maxln = 0;
    for (int i = 0; i < n; ++i) {
        int64_t l = io->row_ptr[i + 1] - io->row_ptr[i];
        if (l > maxln) maxln = l;
    }
#pragma omp parallel num_threads(nth) reduction(+ : work, fails, bad_type)
    {
        double *shp = (double *)malloc(sizeof(double) * (size_t)(maxln + 1));
        double *shpd = (double *)malloc(sizeof(double) * 3 * (size_t)(maxln + 1));
        double *tmp = (double *)malloc(sizeof(double) * 3 * (size_t)(maxln + 1));
        int32_t *lst = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxln + 1));
        const int atomic = nth > 1;
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            const int64_t p0 = io->row_ptr[i];
            const int ln = (int)(io->row_ptr[i + 1] - p0);
            const int mid = io->mat_id[i];
            const double *lprop = io->props + 30 * mid;
            const int mtype = io->mat_type[mid];
            const double vol = io->vol[i];
            double det = 1.0;
            double lstress[6], lstrain[6], lstate[20];
            for (int k = 0; k < 6; ++k) { lstress[k] = io->stress[6 * i + k]; lstrain[k] = io->strain[6 * i + k]; }
            for (int k = 0; k < 20; ++k) lstate[k] = io->state[20 * i + k];
            for (int J = 0; J < ln; ++J) lst[J] = io->col[p0 + J];
            int push = 0;
            if (c->lagrangian && !io->linit) {
                for (int J = 0; J < ln; ++J) {
                    shp[J] = io->shp[p0 + J];
                    for (int k = 0; k < 3; ++k) shpd[3 * J + k] = io->dshp[3 * (p0 + J) + k];
                }
                push = c->finite_strain;
            } else {
                if (c->lagrangian) {
                    or_rk1(io->coo0 + 3 * i, n, io->coo0, io->win, io->ebc_node, lst, ln, c, shp, shpd);
                    if (c->itype_int == 1)
                        or_snni_grad(io->coo0 + 3 * i, io->win + 3 * i, io->vol[i], n, io->coo0, io->win,
                                     io->ebc_node, lst, ln, c, shpd);
                } else {
                    or_rk1(io->coo_cur + 3 * i, n, io->coo_cur, io->win, io->ebc_node, lst, ln, c, shp, shpd);
                    if (c->itype_int == 1)
                        or_snni_grad(io->coo_cur + 3 * i, io->win + 3 * i, io->vol[i], n, io->coo_cur, io->win,
                                     io->ebc_node, lst, ln, c, shpd);
                }
                for (int J = 0; J < ln; ++J) {
                    io->shp[p0 + J] = shp[J];
                    for (int k = 0; k < 3; ++k) io->dshp[3 * (p0 + J) + k] = shpd[3 * J + k];
                }
                push = c->lagrangian && c->finite_strain;
            }
            if (push) {
                double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, IF[9];
                for (int K = 0; K < 3; ++K)
                    for (int L = 0; L < 3; ++L)
                        for (int J = 0; J < ln; ++J)
                            CM3(F, K, L) = CM3(F, K, L) + shpd[3 * J + L] * io->dinc_tot[3 * lst[J] + K];
                det = or_determinant(F);
                or_inv3(F, IF);
                for (int J = 0; J < ln; ++J) {
                    for (int k = 0; k < 3; ++k) tmp[3 * J + k] = shpd[3 * J + k];
                    for (int L = 0; L < 3; ++L) {
                        double s = 0.0;
                        for (int K = 0; K < 3; ++K) s = s + tmp[3 * J + K] * CM3(IF, K, L);
                        shpd[3 * J + L] = s;
                    }
                }
            }
            double lmat[9] = {0}, d[6] = {0};
            for (int K = 0; K < 3; ++K)
                for (int L = 0; L < 3; ++L)
                    for (int J = 0; J < ln; ++J)
                        CM3(lmat, K, L) = CM3(lmat, K, L) + shpd[3 * J + L] * io->dinc[3 * lst[J] + K];
            if (!c->finite_strain) {
                /* small strain builds the Voigt increment in the reference's own order */
                double s6[6] = {0};
                for (int J = 0; J < ln; ++J) {
                    const double *g = shpd + 3 * J, *u = io->dinc + 3 * lst[J];
                    s6[0] = s6[0] + g[0] * u[0];
                    s6[1] = s6[1] + g[1] * u[1];
                    s6[2] = s6[2] + g[2] * u[2];
                    s6[3] = s6[3] + g[1] * u[2] + g[2] * u[1];
                    s6[4] = s6[4] + g[0] * u[2] + g[2] * u[0];
                    s6[5] = s6[5] + g[0] * u[1] + g[1] * u[0];
                }
                /* encode as L so or_constitute_point rebuilds exactly s6 */
                double L2[9] = {s6[0], 0, 0, s6[5], s6[1], 0, s6[4], s6[3], s6[2]};
                memcpy(lmat, L2, sizeof(L2));
            }
            double old_s[6], old_e[6];
            for (int k = 0; k < 6; ++k) { old_s[k] = io->stress[6 * i + k]; old_e[k] = io->strain[6 * i + k]; }
            int it = or_constitute_point(lmat, lstress, lstrain, lstate, lprop, mtype, c->finite_strain, d);
            if (it == -1) fails++;
            if (it == -2) bad_type++;
            for (int k = 0; k < 20; ++k) io->state[20 * i + k] = lstate[k];
            double si[6], ei[6];
            for (int k = 0; k < 6; ++k) {
                si[k] = lstress[k] - old_s[k];
                ei[k] = lstrain[k] - old_e[k];
                io->stress[6 * i + k] = lstress[k];
                io->strain[6 * i + k] = lstrain[k];
            }
            if (io->strain_eq) {
                double q = lstrain[0] * lstrain[0] + lstrain[1] * lstrain[1] + lstrain[2] * lstrain[2] +
                           2.0 * (lstrain[3] * lstrain[3] + lstrain[4] * lstrain[4] + lstrain[5] * lstrain[5]);
                io->strain_eq[i] = pow(q * 2 / 3, 0.5);
            }
            for (int k = 0; k < 6; ++k) {
                if (c->finite_strain) work += 0.5 * d[k] * (2 * lstress[k] - si[k]) * vol * det;
                else work += 0.5 * ei[k] * si[k] * vol * det;
            }
            if (io->wave_vel) {
                double poiss = lprop[0], young = lprop[1], rho = lprop[2];
                double bulk = young / (3.0 * (1.0 - 2.0 * poiss));
                double shear = young / (2.0 * (1.0 + poiss));
                if (mtype > 1) or_estimate_moduli(si, ei, &shear, &bulk);
                double pmod = bulk + 4.0 * shear / 3.0;
                io->wave_vel[i] = sqrt(fmax(pmod, shear) / rho);
            }
            const double rho = lprop[2];
            const double *lbod = io->bodyforce ? io->bodyforce + 3 * i : NULL;
            for (int J = 0; J < ln; ++J) {
                const double *g = shpd + 3 * J;
                double f[3];
                f[0] = g[0] * lstress[0] + g[2] * lstress[4] + g[1] * lstress[5];
                f[1] = g[1] * lstress[1] + g[2] * lstress[3] + g[0] * lstress[5];
                f[2] = g[2] * lstress[2] + g[1] * lstress[3] + g[0] * lstress[4];
                for (int k = 0; k < 3; ++k) {
                    double t = f[k] * vol * det;
                    if (atomic) {
#pragma omp atomic
                        io->fint[3 * lst[J] + k] += t;
                    } else {
                        io->fint[3 * lst[J] + k] += t;
                    }
                }
                double fe[3];
                for (int k = 0; k < 3; ++k) {
                    double fg = shp[J] * c->gravity[k] * rho;
                    double fb = lbod ? shp[J] * lbod[k] * rho : 0.0;
                    fe[k] = fb + fg;
                }
                for (int k = 0; k < 3; ++k) {
                    double t = fe[k] * vol * det;
                    if (atomic) {
#pragma omp atomic
                        io->fext[3 * lst[J] + k] += t;
                    } else {
                        io->fext[3 * lst[J] + k] += t;
                    }
                }
            }
        }
        free(shp); free(shpd); free(tmp); free(lst);
    }
    io->int_work = io->int_work + work;
    io->ret_fail = fails;
    return bad_type ? -2 : (fails ? -1 : 0);
}

/* ---------------------------------------------------------------------- */
/* Correspondence peridynamics.                                           */

/* UDFM_SHAPE_TENSOR (RK.f90:579-726): w_J = prod_k w(|X_J,k - x_k| / a_J,k),
 * K = sum_J xi xi^T w_J V_J with xi = X_J - x, K^-1 by INVERSE.  The
 * singular-kernel branch there is switched off (IF (.FALSE.)). */
void or_udfm(const double *x, int n, const double *coo, const double *vol, const double *win,
             const int32_t *lstack, int ln, double *shp, double *invk)
{
    (void)n;
    double K[9] = {0};
    for (int I = 0; I < ln; ++I) {
        const int ii = lstack[I];
        const double h[3] = {-(x[0] - coo[3 * ii]), -(x[1] - coo[3 * ii + 1]), -(x[2] - coo[3 * ii + 2])};
        double px, py, pz, d;
        or_kernel(fabs(-(x[0] - coo[3 * ii]) / win[3 * ii]), &px, &d);
        or_kernel(fabs(-(x[1] - coo[3 * ii + 1]) / win[3 * ii + 1]), &py, &d);
        or_kernel(fabs(-(x[2] - coo[3 * ii + 2]) / win[3 * ii + 2]), &pz, &d);
        const double phi = px * py * pz;
        for (int J = 0; J < 3; ++J)
            for (int L = 0; L < 3; ++L) CM3(K, J, L) = CM3(K, J, L) + h[J] * h[L] * phi * vol[ii];
        shp[I] = phi;
    }
    or_inverse(K, 3, invk);
}

/* DFM_SHAPE_TENSOR (RK.f90:730-807) */
void or_dfm(const double *x0, const double *xt, const double *coo, const double *vol, const int32_t *lstack,
            int ln, const double *ycur, const double *shp, double *S)
{
    for (int k = 0; k < 9; ++k) S[k] = 0.0;
    for (int I = 0; I < ln; ++I) {
        const int ii = lstack[I];
        const double h[3] = {coo[3 * ii] - x0[0], coo[3 * ii + 1] - x0[1], coo[3 * ii + 2] - x0[2]};
        const double hs[3] = {ycur[3 * I] - xt[0], ycur[3 * I + 1] - xt[1], ycur[3 * I + 2] - xt[2]};
        for (int J = 0; J < 3; ++J)
            for (int L = 0; L < 3; ++L) CM3(S, J, L) = CM3(S, J, L) + hs[J] * h[L] * shp[I] * vol[ii];
    }
}

/* CONSTRUCT_FINT_PD (FINT_PD.F90:14-759), Lagrangian.  At LINIT the influence
 * weights and K^-1 are formed and stored (:227-290, :418-432), F = I;
 * afterwards F = S K^-1 from the current positions (:383-403) with |det F|
 * and the pivoting INVERSE.  L = (sum w V dd (x) xi) K^-1 F^-1 (:446-470);
 * Hughes-Winget, rotation, trial and CONSTITUTION as in CONSTRUCT_FINT;
 * bond force t = J sigma F^-T K^-1 (X_i - X_j) added to i with V_j w V_i and
 * taken from j with V_i w V_j (:619-655).  The wave-speed estimate keeps the
 * reference's STRAIN_INC = LSTRAIN - GSTRESS (:509). */
int or_construct_fint_pd(or_fint_io *io, const or_control *c)
{
    const int n = io->n;
    const int nth = io->nthreads > 0 ? io->nthreads : 1;
    if (!c->lagrangian) return -3;   /* semi-Lagrangian PD is "not implemented" in the reference */
    for (int64_t i = 0; i < 3 * (int64_t)n; ++i) io->fint[i] = 0.0;
    int64_t maxln = 0;
    for (int i = 0; i < n; ++i) {
        int64_t l = io->row_ptr[i + 1] - io->row_ptr[i];
        if (l > maxln) maxln = l;
    }
    if (io->linit) {
        int32_t *lst = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxln + 1));
        for (int i = 0; i < n; ++i) {
            const int64_t p0 = io->row_ptr[i];
            const int ln = (int)(io->row_ptr[i + 1] - p0);
            for (int J = 0; J < ln; ++J) lst[J] = io->col[p0 + J];
            or_udfm(io->coo0 + 3 * i, n, io->coo0, io->vol, io->win, lst, ln, io->shp + p0, io->invk + 9 * (int64_t)i);
        }
        free(lst);
    }
    int64_t fails = 0;
    int bad_type = 0;
#pragma omp parallel num_threads(nth) reduction(+ : fails, bad_type)
    {
        double *ycur = (double *)malloc(sizeof(double) * 3 * (size_t)(maxln + 1));
        int32_t *lst = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxln + 1));
        const int atomic = nth > 1;
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            const int64_t p0 = io->row_ptr[i];
            const int ln = (int)(io->row_ptr[i + 1] - p0);
            const int mid = io->mat_id[i];
            const double *lprop = io->props + 30 * mid;
            const int mtype = io->mat_type[mid];
            const double *shp = io->shp + p0;
            const double *invk = io->invk + 9 * (int64_t)i;
            double lstress[6], lstrain[6], lstate[20];
            for (int k = 0; k < 6; ++k) { lstress[k] = io->stress[6 * i + k]; lstrain[k] = io->strain[6 * i + k]; }
            for (int k = 0; k < 20; ++k) lstate[k] = io->state[20 * i + k];
            for (int J = 0; J < ln; ++J) lst[J] = io->col[p0 + J];
            double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, IF[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, det = 1.0;
            if (!io->linit && c->finite_strain) {
                double S[9];
                for (int J = 0; J < ln; ++J)
                    for (int k = 0; k < 3; ++k) ycur[3 * J + k] = io->coo_cur[3 * lst[J] + k];
                or_dfm(io->coo0 + 3 * i, io->coo_cur + 3 * i, io->coo0, io->vol, lst, ln, ycur, shp, S);
                mm3(S, invk, F);
                det = or_determinant(F);
                or_inverse(F, 3, IF);
            }
            double L0[9], L1[9], lmat[9];
            for (int J = 0; J < ln; ++J)
                for (int k = 0; k < 3; ++k) ycur[3 * J + k] = io->dinc[3 * lst[J] + k];
            or_dfm(io->coo0 + 3 * i, io->dinc + 3 * i, io->coo0, io->vol, lst, ln, ycur, shp, L0);
            mm3(L0, invk, L1);
            if (!io->linit) mm3(L1, IF, lmat);
            else memcpy(lmat, L1, sizeof(L1));
            double old_s[6];
            for (int k = 0; k < 6; ++k) old_s[k] = io->stress[6 * i + k];
            int it = or_constitute_point(lmat, lstress, lstrain, lstate, lprop, mtype, 1, NULL);
            if (it == -1) fails++;
            if (it == -2) bad_type++;
            for (int k = 0; k < 20; ++k) io->state[20 * i + k] = lstate[k];
            double si[6], ei[6];
            for (int k = 0; k < 6; ++k) {
                si[k] = lstress[k] - old_s[k];
                ei[k] = lstrain[k] - old_s[k];
                io->stress[6 * i + k] = lstress[k];
                io->strain[6 * i + k] = lstrain[k];
            }
            if (io->wave_vel) {
                double poiss = lprop[0], young = lprop[1], rho = lprop[2];
                double bulk = young / (3.0 * (1.0 - 2.0 * poiss));
                double shear = young / (2.0 * (1.0 + poiss));
                if (mtype > 1) or_estimate_moduli(si, ei, &shear, &bulk);
                double pmod = bulk + 4.0 * shear / 3.0;
                io->wave_vel[i] = sqrt(fmax(pmod, shear) / rho);
                if (mtype == 3) io->wave_vel[i] = io->wave_vel[i] / 0.15;
            }
            /* P = det * sigma IF^T, T = P K^-1 */
            const double sg[9] = {lstress[0], lstress[5], lstress[4], lstress[5], lstress[1], lstress[3],
                                  lstress[4], lstress[3], lstress[2]};
            double IFt[9], PK[9], T[9];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) CM3(IFt, a, b) = CM3(IF, b, a);
            mm3(sg, IFt, PK);
            for (int k = 0; k < 9; ++k) PK[k] = det * PK[k];
            mm3(PK, invk, T);
            const double vi = io->vol[i];
            for (int J = 0; J < ln; ++J) {
                const int jj = lst[J];
                const double dx[3] = {io->coo0[3 * i] - io->coo0[3 * jj], io->coo0[3 * i + 1] - io->coo0[3 * jj + 1],
                                      io->coo0[3 * i + 2] - io->coo0[3 * jj + 2]};
                double f3[3];
                for (int a = 0; a < 3; ++a) {
                    double s = 0.0;
                    for (int b = 0; b < 3; ++b) s += CM3(T, a, b) * dx[b];
                    f3[a] = s;
                }
                const double vj = io->vol[jj];
                for (int k = 0; k < 3; ++k) {
                    const double tj = f3[k] * vi * shp[J] * vj, ti = f3[k] * vj * shp[J] * vi;
                    if (atomic) {
#pragma omp atomic
                        io->fint[3 * jj + k] -= tj;
#pragma omp atomic
                        io->fint[3 * i + k] += ti;
                    } else {
                        io->fint[3 * jj + k] = io->fint[3 * jj + k] - tj;
                        io->fint[3 * i + k] = io->fint[3 * i + k] + ti;
                    }
                }
            }
        }
        free(ycur); free(lst);
    }
    io->ret_fail = fails;
    if (io->auto_ts) {
        /* FINT_PD.F90:690-754: GCHAR_DIST(I) the shortest DSQRT((XJ-XI)**2 +
         * (YJ-YI)**2 + (ZJ-ZI)**2) over the neighbors JJ /= I at the current
         * coordinates (left as it was when there is none); DLT_FINT the
         * smallest GCHAR_DIST(I) / GMAX_WVEL(I), in particle order, times DLT_FAC */
        double dlt = 0.0;
        for (int i = 0; i < n; ++i) {
            const int64_t p0 = io->row_ptr[i];
            const int ln = (int)(io->row_ptr[i + 1] - p0);
            const double xi = io->coo_cur[3 * i], yi = io->coo_cur[3 * i + 1], zi = io->coo_cur[3 * i + 2];
            int first = 1;
            for (int J = 0; J < ln; ++J) {
                const int jj = io->col[p0 + J];
                if (jj == i) continue;
                const double xj = io->coo_cur[3 * jj], yj = io->coo_cur[3 * jj + 1], zj = io->coo_cur[3 * jj + 2];
                const double dist = sqrt((xj - xi) * (xj - xi) + (yj - yi) * (yj - yi) + (zj - zi) * (zj - zi));
                if (first) {
                    io->char_dist[i] = dist;
                    first = 0;
                } else {
                    io->char_dist[i] = fmin(io->char_dist[i], dist);
                }
            }
            const double t = io->char_dist[i] / io->wave_vel[i];
            if (i == 0) dlt = t;
            else if (t < dlt) dlt = t;
        }
        io->dlt_fint = dlt * io->dlt_fac;
    }
    return bad_type ? -2 : (fails ? -1 : 0);
}

/* ---------------------------------------------------------------------- */
/* Whole run: MAIN.F90:304-658 with the HANDELER calls it makes.          */

static int model_search(or_model *m, int first)
{
    /* search-count rule of HANDELER.F90:256-268: rebin when the count has
     * reached the period; during the first step the period is not yet set */
    int period = first ? 0 : m->pdsearch;
    if (m->search_count < period) m->search_count++;
    else { or_get_bins(m->n, m->x, m->geom, m->nb, m->ijk); m->search_count = 1; }
    int64_t nnz = or_soft_search(m->n, m->x, m->win, m->ijk, m->nb, m->c.shsup, m->row_ptr, NULL);
    if (nnz < 0) return -1;
    if (nnz > m->nnz_cap) {
        free(m->col); free(m->shp); free(m->dshp);
        m->nnz_cap = nnz + nnz / 4 + 16;
        m->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)m->nnz_cap);
        m->shp = (double *)malloc(sizeof(double) * (size_t)m->nnz_cap);
        m->dshp = (double *)malloc(sizeof(double) * 3 * (size_t)m->nnz_cap);
        if (!m->col || !m->shp || !m->dshp) return -1;
    }
    or_soft_search(m->n, m->x, m->win, m->ijk, m->nb, m->c.shsup, m->row_ptr, m->col);
    m->nnz = nnz;
    return 0;
}

static int model_fint(or_model *m, int linit)
{
    or_fint_io io;
    memset(&io, 0, sizeof(io));
    io.n = m->n; io.win = m->win; io.vol = m->vol; io.coo0 = m->coo0; io.coo_cur = m->x;
    io.props = m->props; io.mat_id = m->mat_id; io.mat_type = m->mat_type; io.ebc_node = m->ebc_node;
    io.dinc = m->d; io.dinc_tot = m->dtot; io.bodyforce = NULL;
    io.row_ptr = m->row_ptr; io.col = m->col; io.shp = m->shp; io.dshp = m->dshp;
    io.stress = m->stress; io.strain = m->strain; io.state = m->state;
    io.fint = m->fint; io.fext = m->fext; io.wave_vel = m->wave_vel; io.strain_eq = m->strain_eq;
    io.int_work = m->int_work; io.linit = linit; io.nthreads = m->nthreads; io.invk = m->invk;
    io.auto_ts = m->auto_ts && m->c.peridynamics; io.dlt_fac = m->dlt_fac; io.char_dist = m->char_dist;
    int r = m->c.peridynamics ? or_construct_fint_pd(&io, &m->c) : or_construct_fint(&io, &m->c);
    m->int_work = io.int_work;
    if (io.auto_ts && !linit) m->dt = io.dlt_fint;   /* IF (AUTO_TS) DLT = LOCAL_DLT (MAIN.F90:428) */
    m->ret_fail += io.ret_fail;
    return r;
}

int or_model_init(or_model *m)
{
    const int n = m->n;
    m->row_ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    m->ijk = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)n);
    m->col = NULL; m->shp = NULL; m->dshp = NULL; m->nnz = 0; m->nnz_cap = 0;
    m->invk = m->c.peridynamics ? (double *)calloc(9 * (size_t)n, sizeof(double)) : NULL;
    if (m->c.peridynamics && !m->invk) return -1;
    m->char_dist = m->auto_ts ? (double *)calloc((size_t)n, sizeof(double)) : NULL;
    if (m->auto_ts && (!m->c.peridynamics || !m->char_dist || !m->wave_vel)) return -1;
    m->step = 0; m->time = 0.0; m->int_work = 0.0; m->ke = m->ie = m->ee = 0.0;
    m->search_count = 0; m->ret_fail = 0;
    if (!m->row_ptr || !m->ijk) return -1;
    for (int64_t q = 0; q < 3 * (int64_t)n; ++q) {
        m->x[q] = m->coo0[q];
        m->a[q] = 0.0; m->d[q] = 0.0; m->dtot[q] = 0.0; m->dphy[q] = 0.0; m->dphy_tot[q] = 0.0;
        m->fint[q] = 0.0; m->fext[q] = 0.0; m->prforce[q] = 0.0;
    }
    for (int64_t q = 0; q < 6 * (int64_t)n; ++q) { m->stress[q] = 0.0; m->strain[q] = 0.0; }
    for (int64_t q = 0; q < 20 * (int64_t)n; ++q) m->state[q] = 0.0;
    or_bin_geometry(n, m->coo0, m->win, m->sm_dist[0], m->sm_dist[1], m->sm_dist[2], m->geom, m->nb);
    return 0;
}

int or_model_step(or_model *m)
{
    const int n = m->n;
    const int first = (m->step == 0);
    double *fext_nmo = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    if (!fext_nmo) return -1;
    int r = 0;
    if (first) {
        for (int64_t q = 0; q < 3 * (int64_t)n; ++q) m->d[q] = 0.0;
        if (model_search(m, 1)) { free(fext_nmo); return -1; }
        r = model_fint(m, 1);
        if (r == -2) { free(fext_nmo); return r; }
    }
    /* predictor (MAIN.F90:348-355) with the step size of the last step; with
     * the automatic time step the first step has none (zero increment) */
    if (first && m->auto_ts) {
        for (int64_t q = 0; q < 3 * (int64_t)n; ++q) m->d[q] = 0.0;
    } else {
        const double dt = m->dt;
        for (int64_t q = 0; q < 3 * (int64_t)n; ++q) {
            m->d[q] = dt * m->v[q] + dt * dt * 0.5 * m->a[q];
            m->dtot[q] = m->dtot[q] + m->d[q];
            m->v[q] = m->v[q] + dt * 0.5 * m->a[q];
        }
    }
    /* DO_INTERP (HANDELER.F90:114-133) and coordinate update (MAIN.F90:377-386) */
    if (m->c.peridynamics) {   /* HANDELER.F90:134-137: physical = generalized */
        for (int64_t q = 0; q < 3 * (int64_t)n; ++q) {
            m->dphy[q] = m->d[q];
            if (m->vphy) m->vphy[q] = m->v[q];
            if (m->aphy) m->aphy[q] = m->a[q];
        }
    } else {
        or_interp(n, m->row_ptr, m->col, m->shp, m->d, m->dphy);
        if (m->vphy) or_interp(n, m->row_ptr, m->col, m->shp, m->v, m->vphy);
        if (m->aphy) or_interp(n, m->row_ptr, m->col, m->shp, m->a, m->aphy);
    }
    for (int64_t q = 0; q < 3 * (int64_t)n; ++q) {
        m->dphy_tot[q] = m->dphy_tot[q] + m->dphy[q];
        m->x[q] = m->coo0[q] + m->dphy_tot[q];
    }
    for (int64_t q = 0; q < 3 * (int64_t)n; ++q) fext_nmo[q] = first ? 0.0 : m->fext[q];
    if (!m->c.lagrangian)
        if (model_search(m, first)) { free(fext_nmo); return -1; }
    r = model_fint(m, 0);
    if (r == -2) { free(fext_nmo); return r; }
    const double dt = m->dt;   /* this step's DLT when it is automatic */
    /* EBC, damping, corrector (MAIN.F90:432-488) */
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            int q = 3 * i + k;
            if (m->ebc[q] == 1) {
                m->v[q] = 0.0; m->a[q] = 0.0;
                m->prforce[q] = -m->fint[q];
                m->fext[q] = -m->fint[q];
            } else {
                double mpdc = m->props[30 * m->mat_id[i] + 20] * m->mass[q];
                m->a[q] = (m->fext[q] - m->fint[q] - mpdc * m->v[q]) / (m->mass[q] + 0.5 * dt * mpdc);
                m->v[q] = m->v[q] + 0.5 * dt * m->a[q];
                m->prforce[q] = 0.0;
            }
        }
    /* energies (MAIN.F90:524-553) */
    double ie = 0.0, ke = 0.0, ee = m->ee;
    for (int64_t q = 0; q < 3 * (int64_t)n; ++q) {
        if (!m->c.finite_strain) ie = ie + 0.5 * m->fint[q] * m->dtot[q];
        ke = ke + 0.5 * m->mass[q] * m->v[q] * m->v[q];
        ee = ee + 0.5 * (m->fext[q] + fext_nmo[q]) * m->d[q];
    }
    if (m->c.finite_strain) ie = ie + m->int_work;
    m->ie = ie; m->ke = ke; m->ee = ee;
    m->time = m->time + dt;
    m->step++;
    free(fext_nmo);
    return r;
}

void or_model_free(or_model *m)
{
    free(m->row_ptr); free(m->ijk); free(m->col); free(m->shp); free(m->dshp); free(m->invk); free(m->char_dist);
    m->row_ptr = NULL; m->ijk = NULL; m->col = NULL; m->shp = NULL; m->dshp = NULL; m->invk = NULL;
    m->char_dist = NULL;
}
