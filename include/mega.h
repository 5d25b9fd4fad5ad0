/*
 * mega.h -- C-ABI of the MI355X explicit meshfree step (libmega_hip.so).
 *
 * The reference program drives one explicit step as
 *   MAIN (predictor, MAIN.F90:348-356)
 *   -> HANDELER(DO_INTERP)            HANDELER.F90:114-139
 *   -> coordinate update              MAIN.F90:377-386
 *   -> HANDELER -> [GET_BINS/SOFT_SEARCH] -> CONSTRUCT_FINT -> CONSTITUTION
 *                                     HANDELER.F90:145-307, FINT.F90:16, CONSTITUTION.F90:18
 *   -> EBC / corrector                MAIN.F90:432-488
 * and passes everything by reference in column-major, 1-based layouts.  This
 * header exposes the same stages as plain C functions over plain pointers:
 * all arrays are in the reference layout (node vectors interleaved 3*i+k,
 * tensors Voigt XX,YY,ZZ,YZ,XZ,XY, LOGICAL as int32 0/1), every call returns
 * an int status (MEGA_OK or an error below), and a context owns all device
 * memory.  A Fortran host binds these with bind(C) / iso_c_binding (see the
 * interface block shipped beside this header).
 *
 * Threading: one host thread per context; all work is issued on the
 * context's own HIP stream; functions that return host data synchronise.
 */
#ifndef MEGA_H
#define MEGA_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference has none: EXIT_PROGRAM / WARN / STOP) */
enum {
    MEGA_OK = 0,
    MEGA_ERR_RETURN_MAP = 1,   /* J2 Newton > 20 iterations       CONSTITUTE_VONMISES.F90:115-119 */
    MEGA_ERR_SINGULAR = 2,     /* singular matrix (warning only)  INVERSE.F90:41-45,93-95         */
    MEGA_ERR_BINS = 3,         /* a bin holds MAX_NEIGH = 1000 nodes, where GET_BINS stops
                                  (BIN_SEARCH.F90:47-52, HANDELER.F90:245): nodes that left the
                                  bar or became non-finite pile into the edge bins; or the bin
                                  grid (NBINSX*NBINSY*NBINSZ) or neighbor-list total is beyond
                                  the int32 index range.  The reference's 10000-candidate stop
                                  (BIN_SEARCH.F90:117-120) has no device counterpart. */
    MEGA_ERR_NOMEM = 4,
    MEGA_ERR_HIP = 5,
    MEGA_ERR_ARG = 6,          /* invalid argument / unsupported option                         */
    MEGA_ERR_MATERIAL = 7,     /* material type not on this path  CONSTITUTION.F90:79-84          */
    MEGA_ERR_SCRATCH = 8       /* input larger than a no-allocation entry point's fixed scratch    */
};

/* precision modes */
enum { MEGA_FP64 = 0, MEGA_MIXED = 1 };

/* What CONSTRUCT_FINT and friends read from the CONTROL module
 * (CONTROL.F90:27-95), passed explicitly. */
typedef struct mega_control {
    int32_t rk_degree;       /* RK_DEGREE, must be 1                          */
    int32_t rk_cont;         /* RK_CONT, must be 3 (cubic B-spline)           */
    int32_t itype_int;       /* ITYPE_INT: 0 direct nodal integration; 1 smoothed
                                (stabilized non-conforming) gradients, Lagrangian RK
                                only -- an extension: the reference's smoothing,
                                FINT.F90:441-496, is commented out */
    int32_t ql;              /* QL switch                                      */
    double  ql_coef;         /* QL_COEF                                        */
    double  ql_len;          /* QL_LEN                                         */
    int32_t shsup;           /* SHSUP: ellipsoidal (1) or box (0) support      */
    int32_t lagrangian;      /* LLAGRANGIAN                                    */
    int32_t finite_strain;   /* LFINITE_STRAIN                                 */
    int32_t pdsearch;        /* bins rebuilt every pdsearch steps (PDSEARCH)   */
    double  gravity[3];      /* IGRAVITY                                       */
    double  dist_max[3];     /* XDIST_MAX, YDIST_MAX, ZDIST_MAX (bin size)     */
    int32_t precision;       /* MEGA_FP64 or MEGA_MIXED                        */
    int32_t surrogate;       /* 1: learned return map with exact fallback      */
    int32_t peridynamics;    /* PERIDYNAMICS: CONSTRUCT_FINT_PD (Lagrangian)   */
    int32_t reserved;        /* 0                                              */
} mega_control;

typedef struct mega_ctx mega_ctx;

/* Device context for one model (replaces the SAVE'd storage HANDELER.F90:72-88
 * and MAIN's LOCAL_* arrays, MAIN.F90:152-214).
 *   X0[3n] initial coordinates (GCOO), win[3n] scaled windows (GWIN),
 *   vol[n] (GVOL), mass[3n] (LOCAL_MASS), ebc[3n] fixities (LOCAL_EBC),
 *   ebc_node[n] (GEBC_NODES), mat_type[n] (GMAT_TYPE), prop[30n] (GPROP),
 *   vinit[3n] initial velocity (LOCAL_VEL after STATE_FEILD_INIT).
 * Nodes with equal property rows share one material slot on the device. */
int mega_create(const mega_control *ctrl, int32_t n, const double *X0, const double *win,
                const double *vol, const double *mass, const int32_t *ebc, const int32_t *ebc_node,
                const int32_t *mat_type, const double *prop, const double *vinit, mega_ctx **out);
int mega_destroy(mega_ctx *ctx);

/* GET_BINS (optional) + SOFT_SEARCH at positions x_cur[3n] (NULL: the
 * context's current positions).  rebuild_bins=0 keeps the previous (stale)
 * bin assignment, as HANDELER does between PDSEARCH rebuilds.
 * BIN_SEARCH.F90:2-64,153-328; HANDELER.F90:187-273. */
int mega_search(mega_ctx *ctx, const double *x_cur, int32_t rebuild_bins);
/* Export the neighbor lists in the reference's 1-based form: gn[n] (GN),
 * gstart[n] (GSTART), gstack[cap] (GSTACK); *nnz receives the total. */
int mega_get_neighbors(mega_ctx *ctx, int32_t *gn, int32_t *gstart, int32_t *gstack, int64_t cap, int64_t *nnz);

/* RK1 for every node at its own point x_eval[3n] with nodal positions x_eval
 * (NULL: initial coordinates), storing shapes on the device (GSTACK_SHP,
 * GSTACK_DSHP).  RK.f90:16-453 called from FINT.F90:406-414. */
int mega_shapes(mega_ctx *ctx, const double *x_eval);
/* Export stored shapes: shp[nnz], dshp[3 nnz], in neighbor-list order.
 * Semi-Lagrangian steps form the gather sums from the gradients without
 * storing them; after such steps the gradients are formed again at the
 * current coordinates (those of the last step's search) before the export. */
int mega_get_shapes(mega_ctx *ctx, double *shp, double *dshp, int64_t cap);

/* With peridynamics on (FINT_PD.F90:14-759): UDFM_SHAPE_TENSOR results,
 * influence weights omega[nnz] in neighbor-list order (GSTACK_SHP) and the
 * inverse shape tensors invk[9n] (GINVK, column-major).  RK.f90:579-726. */
int mega_get_pd_shape(mega_ctx *ctx, double *omega, double *invk, int64_t cap);

/* DO_INTERP: phys[3n] = sum_j phi_ij coef[3j+k].  HANDELER.F90:114-133.
 * With peridynamics, phys = coef (HANDELER.F90:134-137). */
int mega_interp(mega_ctx *ctx, const double *coef, double *phys);

/* One CONSTRUCT_FINT call on host arrays (FINT.F90:16-1247), or with
 * peridynamics one CONSTRUCT_FINT_PD call (FINT_PD.F90:14-759; fext is then
 * zero and int_work unchanged, as the reference passes neither), constitution
 * included: inputs dinc[3n] (GDINC), dinc_tot[3n] (GDINC_TOT), x_cur[3n]
 * (GCOO_CUURENT); in/out stress[6n], strain[6n], state[20n] (GSTATE: of its
 * 20 columns the step reads and writes STATE(3) = EPS and the damage models'
 * STATE(1), (2), (4), (5); the others are left as they are); out fint[3n],
 * fext[3n]; int_work accumulates (GINT_WORK).  linit=1 computes and stores
 * shapes as the first call does. */
int mega_fint(mega_ctx *ctx, int32_t linit, const double *dinc, const double *dinc_tot,
              const double *x_cur, double *stress, double *strain, double *state,
              double *fint, double *fext, double *int_work);

/* Per-point kinematics + constitution without a context (FINT.F90:758-801):
 * for n points, lmat[9n] (velocity-gradient increment, column-major),
 * stress[6n], strain[6n], state[20n] in/out, one property row prop[30] of
 * material type mat_type (1 elastic, 2 J2, 3 Drucker-Prager with damage, 6 J2
 * with damage; CONSTITUTION.F90:43-75).  Status MEGA_ERR_RETURN_MAP if
 * any return map failed to converge. */
int mega_constitute(int32_t n, const double *lmat, double *stress, double *strain, double *state,
                    const double *prop, int32_t mat_type, int32_t finite_strain, int32_t precision);

/* One-time geometry preprocessing of the model on the device, without a
 * context: GET_ADIAL, GET_VOLUMES and GET_SMOOTHING_LENGTHS as PRE_MODEL calls
 * them (PRE_MODEL.F90:248-262; PRE_MODEL_ADIAL.F90:16-158,
 * PRE_MODEL_VOLUMES.F90:16-121, PRE_MODEL_VOL2.F90:36-52,
 * PRE_MODEL_SMOOTHING_LENGTHS.F90:16-102).  In: nump nodes coo[3 nump]
 * (GCOO layout), numel elements elcon[8 numel] (MODEL_ELCON, 1-based node
 * ids), hex[numel] = 1 for a hexahedron (block type 2), 0 for a tetrahedron
 * (its first 4 ids); hex may be NULL (all hexahedra).  Out: win[3 nump]
 * (ADIAL, the unscaled window), vol[nump], sm_len[6 nump] (may be NULL),
 * dist_max[3] (XDIST_MAX, YDIST_MAX, ZDIST_MAX).  MEGA_ERR_ARG if an element
 * names a node outside 1..nump. */
int mega_preprocess(int32_t nump, const double *coo, int32_t numel, const int32_t *elcon, const int32_t *hex,
                    double *win, double *vol, double *sm_len, double *dist_max);

/* No counterpart in the reference: a variant of mega_preprocess for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same work, same results bit for
 * bit, but coo, elcon, hex, win, vol and sm_len are DEVICE pointers the caller
 * owns (hex and sm_len may be NULL), dist_max[3] is host memory, and every
 * intermediate buffer lives in a fixed device array inside the library (about
 * 40 MB per GPU, held until the process ends).  Sizes above
 * MEGA_PRE_NOALLOC_MAX_NODES nodes or MEGA_PRE_NOALLOC_MAX_ELEMS elements
 * return MEGA_ERR_SCRATCH.  Runs on the current device's null stream and
 * synchronises it.  Calls that overlap on one GPU share that array and race;
 * calls on different GPUs do not.  Calls no hipMalloc, hipFree or allocating
 * library routine. */
#define MEGA_PRE_NOALLOC_MAX_NODES 524288
#define MEGA_PRE_NOALLOC_MAX_ELEMS 524288
int mega_preprocess_noalloc(int32_t nump, const double *coo, int32_t numel, const int32_t *elcon,
                            const int32_t *hex, double *win, double *vol, double *sm_len, double *dist_max);

/* nsteps explicit steps of size dt, device resident: predictor, DO_INTERP,
 * coordinate update, (search + shapes when semi-Lagrangian), CONSTRUCT_FINT,
 * EBC and corrector.  MAIN.F90:304-658.  After the call every field
 * mega_download returns is that of the last step; within the call, values no
 * step reads (internal and external force, acceleration, Lagrangian
 * coordinates, the RK wave speed) are stored on the last steps only. */
int mega_step(mega_ctx *ctx, double dt, int32_t nsteps);

/* nsteps explicit steps with the automatic time step (IAUTO_TS = 1,
 * peridynamics only; PRE_CONTROL.F90:522-524, MAIN.F90:348-350,428): each
 * step's DLT is dlt_fac (DLT_FAC) times the smallest, over all particles, of
 * the shortest bond at the current coordinates over the particle's wave speed
 * (FINT_PD.F90:690-754), formed after the internal force and used by that
 * step's corrector and the next step's predictor.  The first step of a
 * context has no predictor.  dlt (host, nsteps entries, may be NULL) receives
 * each step's DLT, so that the caller keeps TIME = TIME + DLT (MAIN.F90:641).
 * MEGA_ERR_ARG for a context without peridynamics: the RKPM CONSTRUCT_FINT
 * leaves DLT_FINT unset (FINT.F90:1250-1326 are commented out). */
int mega_step_auto(mega_ctx *ctx, double dlt_fac, int32_t nsteps, double *dlt);

/* fields for mega_download / mega_upload */
enum {
    MEGA_F_X = 0,        /* current coordinates 3n (LOCAL_COO_CURRENT)     */
    MEGA_F_V = 1,        /* generalized velocity 3n (LOCAL_VEL)            */
    MEGA_F_A = 2,        /* acceleration 3n (LOCAL_ACL)                    */
    MEGA_F_D = 3,        /* displacement increment 3n (LOCAL_DSP)          */
    MEGA_F_DTOT = 4,     /* generalized total displacement 3n              */
    MEGA_F_UPHY = 5,     /* physical total displacement 3n                 */
    MEGA_F_FINT = 6,     /* internal force 3n                              */
    MEGA_F_FEXT = 7,     /* external force 3n                              */
    MEGA_F_STRESS = 8,   /* Cauchy stress 6n                               */
    MEGA_F_STRAIN = 9,   /* strain 6n                                      */
    MEGA_F_EPS = 10,     /* equivalent plastic strain n (STATE(3))         */
    MEGA_F_STRAIN_EQ = 11, /* equivalent total strain n (GSTRAIN_EQ)        */
    MEGA_F_WAVE_VEL = 12,  /* wave speed n (GMAX_WVEL)                     */
    MEGA_F_VPHY = 13,    /* physical velocity 3n = sum_j phi_ij v_j (current v) */
    MEGA_F_DAMAGE_STATE = 14, /* damage models' state 4n: EPS_SHEAR, EPS_VOL, D, DAM_SV
                                 (STATE(1),(2),(4),(5); CONSTITUTE_DRUCKPRAG.F90:63-67) */
    MEGA_F_APHY = 15,    /* LOCAL_ACL_PHY 3n: sum_j phi_ij a_j of the previous acceleration, as
                            DO_INTERP forms it in the last step of mega_step (HANDELER.F90:114-133) */
    MEGA_F_VPHY_STEP = 16, /* LOCAL_VEL_PHY 3n: sum_j phi_ij v_j of the predicted velocity, same place */
    MEGA_F_NFIELDS = 17
};
int mega_download(mega_ctx *ctx, int32_t field, double *host);
int mega_upload(mega_ctx *ctx, int32_t field, const double *host);

/* energies after the last step (MAIN.F90:524-553): internal, kinetic, external */
int mega_energies(mega_ctx *ctx, double *e_int, double *e_kin, double *e_ext);
/* nnz and max neighbor count of the current lists (k-bar of the byte model);
 * return_map_fallbacks: learned-return-map answers rejected so far (exact
 * return map used instead), 0 without the surrogate. */
int mega_stats(mega_ctx *ctx, int64_t *nnz, int32_t *max_neighbors, int64_t *return_map_fallbacks);

/* Per-kernel HIP-event timing of mega_step: 0 off, 1 every kernel, 2 the
 * force kernel only (k_force / k_pd_force; the events of every kernel add a
 * few percent to a step). */
int mega_profile(mega_ctx *ctx, int32_t enable);

/* Average device time per launch (ms) of each kernel over the last
 * mega_step call, measured with HIP events on the context's stream.
 * names[k] receives a static string; returns the number of kernels. */
int mega_kernel_times(mega_ctx *ctx, const char **names, double *ms_avg, int32_t *launches, int32_t cap);

const char *mega_status_string(int status);

/* ---- learned J2 return map ------------------------------------------------
 * No counterpart in the reference.  A 13-64-64-64-7 SiLU network (bf16
 * weights on the matrix cores, fp32 accumulation) stands in for VON_MISES
 * (CONSTITUTE_VONMISES.F90:21-158) at the CONSTITUTION seam (FINT.F90:800).
 * Input x = (strain increment[6], rotated stress[6], eqps_n), normalised as
 * xhat = (x - in_mean) * in_scale; output y maps to the plastic correction
 * t = (sigma_trial - sigma[6], eqps - eqps_n) = y * out_scale, where
 * sigma_trial = rotated stress + C:strain increment (FINT.F90:789-794).
 * params[MEGA_MLP_PARAMS]: per layer W[out][in] row-major then b[out].
 * norm[MEGA_MLP_NORM]: in_mean[13], in_scale[13], out_scale[7], envelope
 * lo[13], hi[13] (min/max of the training inputs). */
#define MEGA_MLP_PARAMS 9671
#define MEGA_MLP_NORM 59
typedef struct mega_mlp mega_mlp;
/* trainer with Glorot-uniform weights from seed (splitmix64), zero biases */
int mega_mlp_create(uint64_t seed, mega_mlp **out);
int mega_mlp_destroy(mega_mlp *m);
int mega_mlp_set_params(mega_mlp *m, const float *params);   /* resets the Adam moments */
int mega_mlp_get_params(mega_mlp *m, float *params);
int mega_mlp_get_norm(mega_mlp *m, float *norm);             /* after mega_mlp_set_data */
/* training set: tuples[20 count] as written by mega_harvest_take, for the
 * material of property row prop[30]; computes norm, normalises, shuffles. */
int mega_mlp_set_data(mega_mlp *m, const float *tuples, int64_t count, const double *prop);
/* iters Adam steps on batches of the (cyclically traversed) training set,
 * learning rate cosine-decayed from lr0 to lr1; loss_history[iters] or NULL
 * receives the mean squared error of each batch (normalised units). */
int mega_mlp_train(mega_mlp *m, int32_t iters, int32_t batch, double lr0, double lr1, float *loss_history);
/* y[7n] = network(xhat[13n]) */
int mega_mlp_forward(mega_mlp *m, const float *xhat, float *y, int64_t n);
/* gradient[MEGA_MLP_PARAMS] and value of (1/n) sum ||network(xhat) - that||^2 */
int mega_mlp_gradient(mega_mlp *m, const float *xhat, const float *that, int64_t n, float *grad, float *loss);

/* Use the network in the context's steps (MEGA_MIXED only) for the J2 nodes
 * whose property row equals prop[30]: a plastic trial state inside the
 * envelope takes the network's answer when the residual of the radial-return
 * equations is <= tol sigma_y(eqps) and eqps does not decrease; otherwise the
 * exact return map runs.  The residual is the largest of four parts, each in
 * units of sigma_y(eqps): the yield condition |sqrt(3/2)||dev sigma|| -
 * sigma_y(eqps)|, the flow direction ||dev sigma - (||dev sigma|| /
 * ||dev sigma_trial||) dev sigma_trial||, the pressure (against the trial
 * pressure) and the plastic multiplier |(||dev sigma_trial|| - ||dev sigma||)
 * - 2 mu sqrt(3/2) (eqps - eqps_n)|.  params NULL disables. */
int mega_set_surrogate(mega_ctx *ctx, const float *params, const float *norm, const double *prop, double tol);
/* counters since the last reset: [0] plastic trial states of the trained
 * material, [1] outside the envelope, [2] evaluated by the network, [3] of
 * those, rejected by the residual test (exact return map used) */
int mega_surrogate_stats(mega_ctx *ctx, int64_t *stats, int32_t reset);
/* Record, every `every` steps of mega_step, one tuple per plastic J2 trial
 * state: (strain increment[6], rotated stress[6], eqps_n, stress[6], eqps)
 * as 20 floats, up to capacity tuples; capacity 0 stops and frees. */
int mega_harvest(mega_ctx *ctx, int64_t capacity, int32_t every);
/* copy up to cap tuples to host[20 cap]; *count receives the number
 * recorded (it can exceed the capacity); the buffer is then emptied */
int mega_harvest_take(mega_ctx *ctx, float *host, int64_t cap, int64_t *count);

#ifdef __cplusplus
}
#endif
#endif
