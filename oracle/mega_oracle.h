/*
 * mega_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP64) of the reference's explicit meshfree step
 * (PREDICTOR -> DO_INTERP -> CONSTRUCT_FINT -> CONSTITUTION -> CORRECTOR).
 * It is the checker for the HIP path and the CPU baseline that bench.py times.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it;
 * the product library never links or calls it.
 *
 * Layouts follow the reference: node vectors are interleaved (3*i+k), tensors
 * are Voigt (XX,YY,ZZ,YZ,XZ,XY), small matrices are column-major.  Neighbor
 * lists are 0-based CSR (row_ptr[n+1], col[nnz]); the reference's 1-based
 * GSTART/GSTACK are row_ptr+1 / col+1.
 */
#ifndef MEGA_ORACLE_H
#define MEGA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_control {
    int rk_degree;      /* RK_DEGREE (only 1 is live)            CONTROL.F90 */
    int rk_cont;        /* RK_CONT (3 = cubic B-spline)          */
    int ql;             /* QL on/off                             */
    double ql_coef;     /* QL_COEF                               */
    double ql_len;      /* QL_LEN                                */
    int shsup;          /* SHSUP: ellipsoidal support            */
    int lagrangian;     /* LLAGRANGIAN                           */
    int finite_strain;  /* LFINITE_STRAIN                        */
    double gravity[3];  /* IGRAVITY                              */
    int peridynamics;   /* PERIDYNAMICS: CONSTRUCT_FINT_PD        */
    int itype_int;      /* ITYPE_INT: 0 direct nodal integration; 1 the smoothed
                           gradients (or_snni_grad)                 */
} or_control;

/* ---- small dense math (INVERSE.F90, DETERMINANT.F90, KERNEL.F90) ---- */
void   or_kernel(double xsa, double *phi, double *phix);
void   or_m44inv(const double *a, double *ainv);
void   or_inv3(const double *a, double *ainv);
int    or_inverse(const double *a, int n, double *ainv);
double or_determinant(const double *a);

/* ---- kinematics and constitution ---- */
void or_hughes_winget(const double *lmat, double *rot, double *strain, double *d);
void or_rotate_tensor(const double *rot, double *v);
int  or_von_mises(double *stress, const double *stress_pred, double *state, const double *props);
int  or_von_mises_dam(double *stress, const double *total_strain, const double *inc_strain, double *state,
                      const double *props);
int  or_druck_prag(double *stress, const double *total_strain, const double *inc_strain, double *state,
                   const double *props);
void or_estimate_moduli(const double *stress_inc, const double *strain_inc, double *shear, double *bulk);
/* Per-point update as CONSTRUCT_FINT does it for one node (FINT.F90:758-801):
 * HW(L) -> rotate stress/strain -> strain += dstrain -> trial -> return map.
 * Returns Newton iterations (>=0) or -1 if the return map did not converge. */
int or_constitute_point(const double *lmat, double *stress, double *strain, double *state,
                        const double *props, int mat_type, int finite_strain, double *d_out);

/* ---- RK shape functions (RK.f90:16-453) ---- */
void or_rk1(const double *x, int n, const double *coo, const double *win, const int32_t *ebc_node,
            const int32_t *lstack, int ln, const or_control *c, double *shp, double *shpd);

/* Smoothed gradients at node x (ITYPE_INT = 1): the smoothing box of
 * GET_SMOOTHING_LENGTHS (PRE_MODEL_SMOOTHING_LENGTHS.F90:49-90) from the
 * node's window win_x[3] and volume vol_x, and
 *   shpd[3J+k] = (phi_J(x + l_k e_k) - phi_J(x - l_k e_k)) A_k / V
 * with RK1 at the face centres over the node's list (the smoothing of
 * FINT.F90:441-496, commented out in the reference). */
void or_snni_grad(const double *x, const double *win_x, double vol_x, int n, const double *coo, const double *win,
                  const int32_t *ebc_node, const int32_t *lstack, int ln, const or_control *c, double *shpd);

/* ---- bins and neighbor search (HANDELER.F90:187-252, BIN_SEARCH.F90) ---- */
void or_bin_geometry(int n, const double *coo, const double *win, double xd, double yd, double zd,
                     double *geom /*xmin,ymin,zmin,xbin,ybin,zbin*/, int32_t *nb /*3*/);
/* cell index (1-based i,j,k per node) with the reference's floor/clamp */
void or_get_bins(int n, const double *x, const double *geom, const int32_t *nb, int32_t *ijk);
/* SOFT_SEARCH for all nodes at positions x against bins ijk (possibly stale).
 * If col is NULL only row_ptr is filled.  Returns nnz or -1 on error. */
int64_t or_soft_search(int n, const double *x, const double *win, const int32_t *ijk,
                       const int32_t *nb, int shsup, int64_t *row_ptr, int32_t *col);

/* ---- internal force (FINT.F90:16-1247) ---- */
typedef struct or_fint_io {
    int n;
    const double *win, *vol, *coo0, *coo_cur;   /* 3n, n, 3n, 3n                      */
    const double *props;                         /* 30 per material                    */
    const int32_t *mat_id, *mat_type;            /* n; per material                     */
    const int32_t *ebc_node;                     /* n (LOGICAL as 0/1)                  */
    const double *dinc, *dinc_tot;               /* 3n each                             */
    const double *bodyforce;                     /* 3n or NULL                          */
    const int64_t *row_ptr; const int32_t *col;  /* CSR                                */
    double *shp, *dshp;                          /* nnz, 3nnz (stored shapes, in/out)   */
    double *stress, *strain, *state;             /* 6n, 6n, 20n (in/out)                */
    double *fint, *fext;                         /* 3n out                              */
    double *wave_vel, *strain_eq;                /* n out (may be NULL)                 */
    double *invk;                                /* 9n inverse PD shape tensors (in/out) */
    double int_work;                             /* accumulated (in/out)                */
    int linit;
    int nthreads;                                /* 1 = reference summation order       */
    int64_t ret_fail;                            /* count of non-converged return maps  */
    /* automatic time step (CONSTRUCT_FINT_PD only, FINT_PD.F90:690-754) */
    int auto_ts;                                 /* AUTO_TS                             */
    double dlt_fac;                              /* DLT_FAC                             */
    double *char_dist;                           /* n GCHAR_DIST (in/out; kept where a
                                                    particle has no other neighbor)     */
    double dlt_fint;                             /* DLT_FINT out                        */
} or_fint_io;
int or_construct_fint(or_fint_io *io, const or_control *c);

/* ---- correspondence peridynamics (FINT_PD.F90, RK.f90:579-807) ---- */
/* UDFM_SHAPE_TENSOR: influence weights shp[ln] and K^-1 (3x3) at x */
void or_udfm(const double *x, int n, const double *coo, const double *vol, const double *win,
             const int32_t *lstack, int ln, double *shp, double *invk);
/* DFM_SHAPE_TENSOR: S = sum w V (y_J - x_t) (x) (X_J - x_0) with y given per neighbor */
void or_dfm(const double *x0, const double *xt, const double *coo, const double *vol, const int32_t *lstack,
            int ln, const double *ycur, const double *shp, double *S);
/* CONSTRUCT_FINT_PD: fint only (the reference passes no FEXT and no work) */
int or_construct_fint_pd(or_fint_io *io, const or_control *c);

/* ---- interpolation (HANDELER.F90:114-139) ---- */
void or_interp(int n, const int64_t *row_ptr, const int32_t *col, const double *shp,
               const double *coef, double *phys);

/* ---- whole explicit run (MAIN.F90:304-658) ---- */
typedef struct or_model {
    int n;
    or_control c;
    double dt, time_end;
    int pdsearch;                /* rebin period in steps (semi-Lagrangian)     */
    int nthreads;
    /* model */
    double *coo0, *win, *vol, *mass;   /* 3n,3n,n,3n */
    double *sm_dist;                   /* xdist,ydist,zdist (3)                */
    int32_t *ebc;                      /* 3n fixity                            */
    int32_t *ebc_node;                 /* n                                    */
    int32_t *mat_id; int nmat; double *props; int32_t *mat_type;
    /* state */
    double *x, *v, *a, *d, *dtot, *dphy, *dphy_tot, *fint, *fext, *prforce;
    double *stress, *strain, *state, *strain_eq, *wave_vel;
    double *vphy, *aphy;               /* physical v, a at DO_INTERP (may be NULL) */
    double int_work, ke, ie, ee, time;
    int64_t step, nnz, nnz_cap;
    int64_t *row_ptr; int32_t *col; double *shp, *dshp;
    double *invk;                      /* 9n (peridynamics)                    */
    double geom[6]; int32_t nb[3]; int32_t *ijk; int search_count;
    int64_t ret_fail;
    /* automatic time step (IAUTO_TS = 1, peridynamics): dt is then set by each
     * step's internal force (MAIN.F90:428) and the first step has no predictor
     * (MAIN.F90:348-350) */
    int auto_ts;
    double dlt_fac;
    double *char_dist;                 /* n, allocated by or_model_init        */
} or_model;

/* All arrays in the model struct are caller-owned except the CSR/shape/bin
 * storage, which or_model_init allocates and or_model_free releases. */
int  or_model_init(or_model *m);
int  or_model_step(or_model *m);
void or_model_free(or_model *m);

#ifdef __cplusplus
}
#endif
#endif
