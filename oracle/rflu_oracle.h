/*
 * rflu_oracle.h -- CPU restatement of RocfluMP's explicit residual/update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the GPU path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and never as
 * the thing measured or shipped.  The product (rocflu_amd) never calls it.
 *
 * Plain C, FP64, built with -ffp-contract=off.  Loops run in the reference's order
 * (face scatter, patch by patch), each function citing the reference lines it follows.
 * Arrays use the reference's component-fastest layout with 0-based indices.
 */
#ifndef RFLU_ORACLE_H
#define RFLU_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n_cells, n_faces, n_patches;
  const int32_t* f2c;        /* (2,F) */
  const double* fn;          /* (4,F) unit normal + |A| */
  const double* fc;          /* (3,F) */
  const double* cofg;        /* (3,N) */
  const double* vol;         /* (N) */
  const int32_t* bpatch_ptr; /* (P+1) boundary faces of patch p: [ptr[p], ptr[p+1]) */
  const int32_t* bf2c;       /* (B) */
  const double* bfn;         /* (4,B) */
  const double* bfc;         /* (3,B) */
  const int32_t* patch_bc;   /* (P) 10 inflow (total conditions and angles), 11 inflow
                                (velocity + temperature), 20 outflow, 60 slip wall,
                                80 farfield, 90 injection */
  const int32_t* patch_opt;  /* (P) inflow/outflow option 0 supersonic, 1 subsonic, 2 mixed */
  const double* patch_pout;  /* (P) outflow static pressure */
  const int32_t* patch_order;/* (P) 1 or 2 (boundary-face order); NULL = all first order */
  const int32_t* c2cs_ptr;   /* (N+1) stencil CSR, order 2 */
  const int32_t* c2cs_idx;
  const double* xyz_moms;    /* (10,N) R-factors */
  const int32_t* c2f_ptr;    /* (N+1) faces of each cell (signed, see rflu_mesh.h) */
  const int32_t* c2f_idx;
  /* gas model and numerics */
  double gamma, cp;          /* global%refGamma, global%refCp */
  int32_t discr;             /* 1 Roe, 3 HLLC, 4 AUSM+ */
  int32_t order;             /* 1 or 2 */
  int32_t reconst;           /* 0 none, 2 WENO-XYZ + clip, 10 Barth-Jespersen (velocity in
                                the frame of its symmetric gradient), 11 Venkatakrishnan */
  int32_t rk;                /* 1 RK4, 2 Wray RK3, 3 Euler, 4 SSP-RK3 */
  double cfl, dt_imposed, dt_min_limit, eps_entr, diss_fact;
  const double* patch_vals;  /* (5,P) BCDAT 1..5 of the patch: inflow 11 u, v, w, T, p
                                (p: non-subsonic options); inflow 10 ptot, ttot, betah,
                                betav, Mach; farfield Mach, attack, sideslip, p, T;
                                injection mass flux, T */
  /* constrained gradients (CRECONSTC): wall faces of each stencil, (patch, face in
   * patch) pairs in CSR c2cs_bptr, and cReconstCellsWeight */
  const int32_t* c2cs_bptr;
  const int32_t* c2cs_bface;
  double creconst_weight;
  int32_t creconst;          /* 0 off, 1 on */
  int32_t pad_;
  const int32_t* patch_opt2; /* (P) BCSWI_INFLOW_FIXED of inflow 10 patches; NULL = all 0 */
  /* symmetry/periodic virtual cells [n_act, n_cells): each takes the state and the final
   * gradients of the actual cell virt_src[k] turned by xforms[9*virt_xf[k]] (row-major);
   * n_act = 0: every cell is actual */
  int32_t n_act;
  int32_t pad2_;
  const int32_t* virt_src;
  const int32_t* virt_xf;
  const double* xforms;
} orc_region;

typedef struct {
  double time;          /* global%currentTime */
  double dt_last;       /* last dtMin */
  int64_t steps;
  int64_t n_bounds_fix; /* EnforceBounds clamps */
  int64_t n_pos_fix;    /* CheckPositivity energy floors */
  double mass_in, mass_out;
} orc_stats;

/* Error codes (mirror of ModError's ERR_INVALID_VALUE / ERR_NEGATIVE_POSDEF). */
#define ORC_OK 0
#define ORC_ERR_INVALID_VALUE 1
#define ORC_ERR_NEGATIVE_POSDEF 2
#define ORC_ERR_ARG 3

void orc_set_dependent_vars(const orc_region* r, const double* cv, double* dv, int beg, int end);
void orc_fluxes(const orc_region* r, const double* cv, const double* dv, const double* grad,
                double* rhs, orc_stats* st, int stage);
void orc_gradients(const orc_region* r, const double* var, double* grad);
void orc_gradients_constr(const orc_region* r, const double* var, double* grad);
void orc_weno_xyz(const orc_region* r, double* grad);
void orc_limit_simple(const orc_region* r, const double* var, double* grad);
void orc_venkat(const orc_region* r, const double* var, double* grad);
/* Boundary state and flux of one face of a patch (test entry points). */
void orc_rind_state(const orc_region* r, int patch, const double* n, const double* q, double pl,
                    double* out);
void orc_sym_eigen3(const double* m, double* vec, double* val);
void orc_barth_jesp(const orc_region* r, const double* var, double* grad);
void orc_interior_flux(const orc_region* r, const double* cv, const double* dv,
                       const double* grad, int f, double* flx);
void orc_patch_flux(const orc_region* r, int patch, const double* n, const double* q, double pl,
                    double* flx);
void orc_cell_gradients(const orc_region* r, double* cv, const double* dv, double* grad);
void orc_rk_update(const orc_region* r, int stage, double dt, double* cv, const double* cvOld,
                   const double* rhs, double* rhsSum);
int orc_after_update(const orc_region* r, double* cv, double* dv, orc_stats* st);
double orc_local_dt(const orc_region* r, const double* cv, const double* dv);
double orc_min_timestep(const orc_region* r, double dt_local_min, double current_time);
int orc_nstages(int rk);
/* Threads of the CPU path (OpenMP; 1 = the sequential loops, the default).  Results
 * are the same bits for any count. */
void orc_set_threads(int n);
int orc_get_threads(void);
/* Full time loop: nsteps of (dt, stages).  cv (5,N) and dv (3,N) are updated in place. */
int orc_run(const orc_region* r, double* cv, double* dv, int64_t nsteps, orc_stats* st);
/* Time loop to end_time with RFLU_TimeStepping's end-time clamp (at most max_steps). */
int orc_run_until(const orc_region* r, double* cv, double* dv, double end_time,
                  int64_t max_steps, orc_stats* st);
/* The virtual cells' state (cv, then dv) and final gradients from the actual cells they
 * mirror (RFLU_MPI_CopyWrapper / RFLU_MPI_ExchangeGradCells with
 * RFLU_RELP_TransformWrapper / RFLU_RELP_TransformGradWrapper); no-ops without virtual
 * cells. */
void orc_fill_virtual_state(const orc_region* r, double* cv, double* dv);
void orc_fill_virtual_grad(const orc_region* r, double* grad);
/* One residual evaluation R(cv) for the current state (gradients included for order 2). */
void orc_residual(const orc_region* r, const double* cv, const double* dv, double* rhs);

#ifdef __cplusplus
}
#endif
#endif
