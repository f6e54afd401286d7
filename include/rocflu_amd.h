/*
 * rocflu_amd.h -- C-ABI of the MI355X explicit residual/update path.
 *
 * The reference has no plugin system: its boundary is the set of Fortran subroutines
 * RungeKuttaMP calls once per stage for each region (modfloflu/ModInterfaces.F90:
 * 85-263, called from libfloflu/RungeKuttaMP.F90:216,220,274,348,432,437) and the
 * time-step pair called from rocflu/RFLU_TimeStepping.F90:429,444.  Each entry point
 * below replaces one of them for one region; a Fortran ISO_C_BINDING module
 * (rocflu_amd/fortran/rflu_gpu.f90) binds them so the reference host keeps its
 * t_region data and input files.
 *
 * Conventions: plain pointers and sizes, column-major arrays with the component index
 * fastest (cv(5,nCells) is double cv[5*ic+iv]), cell/face indices with the base given
 * in the region descriptor (1 for Fortran arrays).  Every call returns 0 or one of the
 * RFLU_ERR_* codes; rflu_gpu_last_error() describes the last failure.  Calls on one
 * region are ordered on that region's HIP stream and return without waiting, except
 * those that hand a value back to the host (documented per call).
 */
#ifndef ROCFLU_AMD_H
#define ROCFLU_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  RFLU_ERR_INVALID_VALUE and RFLU_ERR_NEGATIVE_POSDEF are the
 * reference's ERR_INVALID_VALUE (NaN, libflu/RFLU_CheckValidity.F90:165-222) and
 * ERR_NEGATIVE_POSDEF (libflu/RFLU_CheckPositivity.F90:219-253). */
#define RFLU_OK 0
#define RFLU_ERR_INVALID_VALUE 1
#define RFLU_ERR_NEGATIVE_POSDEF 2
#define RFLU_ERR_ARG 3
#define RFLU_ERR_HIP 4
#define RFLU_ERR_NOMEM 5
#define RFLU_ERR_HANDLE 6
/* Fluid dumps: a file could not be opened, read, written or renamed (the reference's
 * ERR_FILE_OPEN/_READ/_CLOSE), or it is not a dump this region can read
 * (ERR_INVALID_MARKER, ERR_INVALID_NCELLS, ERR_ILLEGAL_VALUE). */
#define RFLU_ERR_FILE 7
#define RFLU_ERR_DUMP 8

/* Flux schemes (DISCR, modfloflu/ModParameters.F90:662-667). */
#define RFLU_DISCR_ROE 1
#define RFLU_DISCR_HLLC 3
#define RFLU_DISCR_AUSMPLUS 4
/* Gradient modifiers (RECONST, modfloflu/ModParameters.F90:678-683). */
#define RFLU_RECONST_NONE 0
#define RFLU_RECONST_WENO_XYZ 2
#define RFLU_RECONST_LIM_BARTHJESP 10
#define RFLU_RECONST_LIM_VENKAT 11
/* Time integrators (RKSCHEME, modfloflu/ModParameters.F90:615-618); 4 is Shu-Osher
 * SSP-RK3, which the reference does not have. */
#define RFLU_RK4_CLASSICAL 1
#define RFLU_RK3_WRAY 2
#define RFLU_RK1_EULER 3
#define RFLU_RK3_SSP 4
/* Arithmetic: FP64 everywhere, or the mixed mode: FP32 cell gradients (stencil
 * weights, accumulation, storage) and FP32 reconstruction increments, with FP64 face
 * fluxes, flux buffer, residual accumulation, state, limiter and positivity checks.
 * (The face-flux functions stay FP64 because HLLC's star pressure responds to the FP32
 * rounding of the normal velocities with rho*a*dq, as large as rounding p itself, which
 * sets the error of the momenta normal to the main flow.) */
#define RFLU_PREC_FP64 0
#define RFLU_PREC_MIXED 1
/* The mixed mode under its older name (FP32 only in the gradients): the same kernels. */
#define RFLU_PREC_FP32_GRAD 2
/* Diagnostic: FP64 gradients, FP32 interior-flux arithmetic with the pressure terms in
 * FP64 and a flux buffer of FP32 mass/energy and FP64 momenta. */
#define RFLU_PREC_FP32_FLUX 3

/*
 * One region's grid and numerics -- the t_grid/t_patch/t_mixt_input fields the hot
 * path reads (modfloflu/ModGrid.F90:95-224, ModBndPatch.F90:113-148,
 * ModMixture.F90:97-105).  Boundary faces of all patches are concatenated patch by
 * patch; bpatch_ptr holds 0-based offsets into them.  The stencil arrays flatten
 * c2cs(icg)%cellMembs into a CSR (c2cs_ptr has n_cells+1 0-based offsets) and
 * xyz_moms is c2cs(icg)%xyzMoms (10 per cell); they are read only for order 2.
 * The library copies everything it needs during rflu_gpu_create_region.
 */
typedef struct {
  int32_t index_base;        /* 0 or 1: base of f2c, bf2c and c2cs_idx values      */
  int32_t n_cells;           /* nCells: actual cells (n_virtual virtual cells follow
                                them in every per-cell array)                        */
  int32_t n_faces;           /* nFaces                                               */
  int32_t n_patches;
  const int32_t* f2c;        /* (2, n_faces)                                         */
  const double* fn;          /* (4, n_faces) unit normal and |A|                     */
  const double* fc;          /* (3, n_faces) face centroid                           */
  const double* cofg;        /* (3, n_cells)                                         */
  const double* vol;         /* (n_cells)                                            */
  const int32_t* bpatch_ptr; /* (n_patches+1)                                        */
  const int32_t* bf2c;       /* (nBFaces)                                            */
  const double* bfn;         /* (4, nBFaces)                                         */
  const double* bfc;         /* (3, nBFaces)                                         */
  const int32_t* patch_bc;   /* (n_patches) 10 inflow (BC_INFLOW_TOTANG), 11 inflow
                                (BC_INFLOW_VELTEMP), 20 outflow, 60 slip wall,
                                80 farfield, 90 injection
                                (modfloflu/ModParameters.F90:397-438)                */
  const int32_t* patch_opt;  /* (n_patches) BCSWI_INFLOW_TYPE / BCSWI_OUTFLOW_TYPE:
                                0 supersonic, 1 subsonic, 2 mixed                    */
  const double* patch_pout;  /* (n_patches) outflow static pressure                  */
  const int32_t* patch_order;/* (n_patches) 1 or 2; NULL = all 1 (reference default) */
  const int32_t* c2cs_ptr;   /* (n_cells+1), order 2 only                            */
  const int32_t* c2cs_idx;
  const double* xyz_moms;    /* (10, n_cells), order 2 only                          */
  double gamma, cp;          /* global%refGamma, global%refCp                        */
  int32_t discr, order, reconst, rk, precision;
  double cfl, dt_imposed, dt_min_limit, eps_entr, diss_fact;
  const double* patch_vals;  /* (5, n_patches) pPatch%mixt%vals(1:5) (constant per
                                patch; modfloflu/ModParameters.F90:452-494):
                                11: BCDAT_INFLOW_U, _V, _W, _T, _P (P used by the
                                    non-subsonic options)
                                10: BCDAT_INFLOW_PTOT, _TTOT, _BETAH, _BETAV, _MACH
                                    (MACH used by the non-subsonic options)
                                80: BCDAT_FARF_MACH, _ATTACK, _SLIP, _PRESS, _TEMP
                                90: BCDAT_INJECT_MFRATE, _TEMP (rest unused)
                                NULL when no patch reads values                      */
  /* Constrained gradients at walls (order 2, mixtInput%cReconstCells = CRECONSTC):
   * c2cs(icg)%bFaceMembs flattened to a CSR (c2cs_bptr, n_cells+1 0-based offsets)
   * of (patch, face within the patch) pairs (index_base applies to both), and
   * mixtInput%cReconstCellsWeight. */
  const int32_t* c2cs_bptr;
  const int32_t* c2cs_bface;
  double creconst_weight;
  int32_t creconst;          /* 0 off (reference default), 1 on                      */
  int32_t pad_;
  const int32_t* patch_opt2; /* (n_patches) second switch: BCSWI_INFLOW_FIXED of
                                inflow 10 patches (0 or 1), BCSWI_FARF_CORR of
                                farfield patches (0 only: no vortex correction);
                                NULL = all 0                                         */
  int32_t n_virtual;         /* nCellsTot - nCells: virtual cells, numbered after the
                                actual ones; they take part in faces (actual-virtual
                                faces) and stencils, while update, checks and time
                                step cover the actual cells; their state comes from
                                rflu_gpu_exchange.  0 for a region of its own and for
                                every case of a batch                                */
  int32_t pad2_;
} rflu_region_desc;

/* Per-region counters of the after-update checks, accumulated since the last reset. */
typedef struct {
  int64_t n_invalid;      /* cells with NaN (ERR_INVALID_VALUE)                     */
  int64_t n_negative;     /* cells with rho <= 0 or p <= 0 after the energy floor   */
  int64_t n_bounds_fix;   /* RFLU_EnforceBounds clamps                              */
  int64_t n_pos_fix;      /* RFLU_CheckPositivity energy floors                     */
  double mass_in, mass_out; /* outflow mass bookkeeping of stage 1 (global%massIn/Out) */
  double time, dt;        /* global%currentTime and the last global%dtMin           */
  int64_t steps;          /* completed time steps                                   */
  int32_t finished;       /* the fused loop reached its end time                    */
  int32_t pad;
} rflu_gpu_stats;

/* Number of visible GPUs (0 when none); never fails. */
int rflu_gpu_device_count(void);

/* Builds the device copy of one region (RFLU_InitFlowSolver's upload point). */
int rflu_gpu_create_region(const rflu_region_desc* desc, int32_t* handle);
int rflu_gpu_destroy_region(int32_t handle);

/* A batch of n_cases independent cases (descs[0..n_cases-1]) in one handle, the
 * batched form of the reference's loop over regions (RungeKuttaMP.F90:148-508 runs
 * each callee per region; here one launch covers every case).  Cases may differ in
 * mesh, gas constants, CFL, dt controls and patches but share discr, order,
 * reconst, rk and precision.  Each case keeps its own time, dt (its own
 * RFLU_MinimumTimeStep), end time and checks: a case that fails a check or reaches
 * its end time stops while the others go on.  State arrays of the handle are the
 * cases' arrays concatenated in case order (cv(5, sum n_cells)).  Advanced by
 * rflu_gpu_time_steps; rflu_gpu_time_step and rflu_gpu_rk_update are single-region
 * calls.  No counterpart in the reference beyond its region loop. */
int rflu_gpu_create_batch(int32_t n_cases, const rflu_region_desc* descs, int32_t* handle);

/* cv (5,n_cells) conservative state; dv (3,n_cells) p,T,c or NULL to compute it
 * (RFLU_SetDependentVars).  Also sets global%currentTime. */
int rflu_gpu_upload_state(int32_t handle, const double* cv, const double* dv, double time);
/* Copies the state back; waits for the region's stream.  dv may be NULL. */
int rflu_gpu_download_state(int32_t handle, double* cv, double* dv);

/* RFLU_TimeStepInviscid + RFLU_MinimumTimeStep (rocflu/RFLU_TimeStepInviscid.F90:
 * 153-334, rocflu/RFLU_MinimumTimeStep.F90:126-245) for a single region at the host's
 * global%currentTime: sets the device dtMin (and the device time to current_time).
 * If dt_out is not NULL, waits and returns it, and returns the error of an after-update
 * check that failed since the last read (the state stopped there). */
int rflu_gpu_time_step(int32_t handle, double current_time, double* dt_out);

/* RKInitMP (libfloflu/RkInitGeneric.F90:128-160). */
int rflu_gpu_rk_init(int32_t handle, int32_t istage);
/* CellGradientsMP (libfloflu/CellGradientsMP.F90:160-296); no-op at order 1. */
int rflu_gpu_cell_gradients(int32_t handle);
/* ConvectiveFluxesMP (libfloflu/ConvectiveFluxes.F90:133-159): fluxes of all faces. */
int rflu_gpu_convective_fluxes(int32_t handle, int32_t istage);
/* RKUpdateMP (libfloflu/RkUpdateGeneric.F90:255-325) with the residual gathered per
 * cell over the cell-to-face list in the reference's summation order; dt_min is the
 * host's global%dtMin (after its end-time clamp).  Queued: launched by the next entry
 * point (see rflu_gpu_after_update). */
int rflu_gpu_rk_update(int32_t handle, int32_t istage, double dt_min);
/* RFLU_SetVarsContWrapper over all cells (libflu/RFLU_SetDependentVars.F90:255-284). */
int rflu_gpu_set_dependent_vars(int32_t handle);
/* AfterUpdateMP (libfloflu/AfterUpdateMP.F90:134-136): RFLU_CheckValidity,
 * RFLU_EnforceBounds, RFLU_CheckPositivity on the device; does not wait.  A failed
 * check stops the region where the reference's ErrorStop stops it: every later launch
 * leaves its state alone, and the next call that reads the device (rflu_gpu_time_step
 * with dt_out, rflu_gpu_check_status, rflu_gpu_time_steps) returns
 * RFLU_ERR_INVALID_VALUE or RFLU_ERR_NEGATIVE_POSDEF.  The update of
 * rflu_gpu_rk_update and the dependent variables of rflu_gpu_set_dependent_vars are
 * launched here with the checks as one pass over the cells when they were the calls
 * just before (otherwise, and before any other entry point, as separate passes; the
 * results are the same bits). */
int rflu_gpu_after_update(int32_t handle, int32_t istage);
/* Waits for the region's work and returns the status of its after-update checks
 * (RFLU_OK, RFLU_ERR_INVALID_VALUE or RFLU_ERR_NEGATIVE_POSDEF).  No counterpart in
 * the reference (whose checks stop the run where they fail): the per-callee host calls
 * it where it needs to know, e.g. after its last step. */
int rflu_gpu_check_status(int32_t handle);

/* Residual R(cv) of the current state (gradients included for order 2), gathered
 * into rhs (5,n_cells); waits.  For inspection and tests. */
int rflu_gpu_residual(int32_t handle, double* rhs);

/*
 * Regions of one problem on one device, coupled through virtual cells (the
 * reference's multi-region runs).  rflu_gpu_set_borders gives a region its borders
 * (t_border, modflu/RFLU_ModCommLists.F90 builds them): border b pairs with border
 * nbr_border[b] (index_base applies) of region nbr_handle[b]; it sends the actual cells
 * send_cells[send_ptr[b] .. send_ptr[b+1]) and receives, in the same order as the
 * neighbour's send list for this region, into the virtual cells
 * recv_cells[recv_ptr[b] .. recv_ptr[b+1]) (cell numbers in the descriptor's numbering
 * and index base).  Replaces earlier borders.
 *
 * rflu_gpu_exchange copies the sending cells' values to the receiving virtual cells
 * of every border of the listed regions (each border's neighbour must be listed):
 * RFLU_XCHG_STATE copies cv and then sets dv (and the primitive copy) of the virtual
 * cells -- RFLU_MPI_CopyWrapper + RFLU_SetVarsContWrapper(nCells+1, nCellsTot)
 * (modflu/RFLU_ModMPI.F90:564-660, libfloflu/RungeKuttaMP.F90:446-453);
 * RFLU_XCHG_GRAD copies the limited (or blended) cell gradients --
 * RFLU_MPI_ExchangeGradCells (modflu/RFLU_ModMPI.F90:1382-1511,
 * RungeKuttaMP.F90:229-236).  The call waits for the listed regions' queued work
 * first.  A step of regions: dt = the minimum of rflu_gpu_time_step over the regions;
 * per stage, rflu_gpu_cell_gradients of each, RFLU_XCHG_GRAD (order 2), then
 * rflu_gpu_convective_fluxes, rflu_gpu_rk_update, rflu_gpu_set_dependent_vars and
 * rflu_gpu_after_update of each, then RFLU_XCHG_STATE.
 *
 * A border may pair a region with itself (nbr_handle[b] = handle, nbr_border[b] = b):
 * the virtual cells of symmetry and periodic patches (rflu_mesh_build_sype) take the
 * values of the actual cells they mirror that way, in the same exchanges.
 */
#define RFLU_XCHG_STATE 1
#define RFLU_XCHG_GRAD 2
int rflu_gpu_set_borders(int32_t handle, int32_t n_borders, const int32_t* nbr_handle,
                         const int32_t* nbr_border, const int32_t* send_ptr,
                         const int32_t* send_cells, const int32_t* recv_ptr,
                         const int32_t* recv_cells);
int rflu_gpu_exchange(int32_t n_regions, const int32_t* handles, int32_t what);
/* Values received through border `border` (index_base applies) are turned by m
 * (row-major 3x3; NULL: copied as they are): the momentum vector m (rho u), the density
 * and pressure gradients m g, the velocity gradient m (G m^T) -- the reflection of a
 * symmetry patch or the turn of a periodic one (RFLU_RELP_TransformWrapper and
 * RFLU_RELP_TransformGradWrapper, modflu/RFLU_ModRelatedPatches.F90:380-435,673-735). */
int rflu_gpu_set_border_transform(int32_t handle, int32_t border, const double* m);

/* RFLU_TimeStepping's loop body (rocflu/RFLU_TimeStepping.F90:353-640) nsteps times:
 * time step, all RungeKuttaMP stages with the update, dependent variables and checks
 * fused, and the time advance.  With use_graph = 1 one step is captured in a HIP
 * graph and replayed.  Stops advancing (state frozen) at the first failed check and
 * returns its code; waits. */
int rflu_gpu_time_steps(int32_t handle, int64_t nsteps, int32_t use_graph);

/* global%maxTime of rflu_gpu_time_steps: the step that would pass it is shortened to
 * end at it and later steps do nothing (rocflu/RFLU_TimeStepping.F90:445-448). */
int rflu_gpu_set_end_time(int32_t handle, double end_time);

/* For a batch: counters and mass summed over the cases, the earliest time, the
 * smallest dt and the fewest steps; finished when every case is. */
int rflu_gpu_get_stats(int32_t handle, rflu_gpu_stats* stats);
/* Statistics of one case of a batch (case 0 of a single region). */
int rflu_gpu_get_case_stats(int32_t handle, int32_t icase, rflu_gpu_stats* stats);
int rflu_gpu_reset_stats(int32_t handle);

/* Waits for the region's stream. */
int rflu_gpu_synchronize(int32_t handle);

/*
 * Differentiation of the residual (verification of the discrete adjoint; no
 * counterpart in the reference).  States live on the device as [5][n_cells] in the
 * region's internal cell order, which rflu_gpu_cell_order maps to the caller's cells
 * (new2old[i] = caller cell of internal cell i).  Device pointers are plain device
 * addresses on the region's GPU; every call waits for its work.  Tangent and adjoint
 * are of the FP64 residual R(cv) of a single region (rflu_gpu_residual's R), first
 * order, or second order with no gradient modifier, the Venkatakrishnan limiter or the
 * WENO-XYZ blend with its positivity clip, with or without wall-constrained gradients,
 * with dv = dv(cv) and the patch values of inflow, farfield and injection patches
 * (their five BCDAT values) as inputs.
 */
int rflu_gpu_cell_order(int32_t handle, int32_t* new2old);
/* vol in the internal cell order (host array of n_cells). */
int rflu_gpu_cell_volumes(int32_t handle, double* vol);
/* Internal cell and |A| of each boundary face of one patch, in patch face order. */
int rflu_gpu_patch_faces(int32_t handle, int32_t patch, int32_t* cells, double* areas);
/* Replaces the 5 values (u, v, w, T, p) of one patch (inflow patches read them). */
int rflu_gpu_set_patch_values(int32_t handle, int32_t patch, const double* vals);
/* cv <- cv_dev, then dv (and the primitive copy) from it. */
int rflu_gpu_set_state_dev(int32_t handle, const double* cv_dev);
int rflu_gpu_get_state_dev(int32_t handle, double* cv_dev);
/* R of the current state into rhs_dev. */
int rflu_gpu_residual_dev(int32_t handle, double* rhs_dev);
/* drhs_dev = dR/dcv . dcv_dev + dR/dvals . dvals (dvals host (5, n_patches) or NULL). */
int rflu_gpu_residual_jvp(int32_t handle, const double* dcv_dev, const double* dvals,
                          double* drhs_dev);
/* rflu_gpu_residual_jvp without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and bits as rflu_gpu_residual_jvp, whose first call allocates the tangent's work
 * arrays; here they live in one static 64 MiB device array per GPU, so a region whose
 * work arrays exceed it (5(F+B) + 5P doubles at first order; plus 35 N, or 65 N with
 * WENO-XYZ, at second order, each rounded up to 256 bytes) gets RFLU_ERR_ARG.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (several host
 * threads or regions on different streams) race and must be serialised by the caller.
 * Returns after the work is done, like rflu_gpu_residual_jvp. */
int rflu_gpu_residual_jvp_static(int32_t handle, const double* dcv_dev, const double* dvals,
                                 double* drhs_dev);
/* lam_dev = (dR/dcv)^T mu_dev; dvals (host (5, n_patches), or NULL) = (dR/dvals)^T mu. */
int rflu_gpu_residual_vjp(int32_t handle, const double* mu_dev, double* lam_dev, double* dvals);

/* Launch tuning: threads per block for the face, cell and gradient kernels (64..1024,
 * multiple of 64; 0 keeps the current value). */
int rflu_gpu_set_launch(int32_t handle, int32_t face_block, int32_t cell_block,
                        int32_t grad_block);

/* Tuning options (set by the autotuner, rocflu_amd/autotune.py): whether the gradient
 * kernel stages each tile's stencil members in LDS (1) or reads them from global
 * memory (0), how many members it handles together (1, 2 or 4; 8 with 256-cell tiles
 * and weights formed in the pass, otherwise read as 4), (read only) the LDS
 * records per tile, over how many XCDs the cell and face kernels spread contiguous
 * block ranges (1 = hardware order; 8 on MI355X), whether the stencil weights are
 * formed once and streamed from HBM (1; 3 weights per stencil entry, 24 B in FP64,
 * 12 B with FP32 gradients; needs LDS tiles), only each member's inverse distance is
 * (2; 8 B per entry in FP64, with 4 members per group; the rest of the weights formed
 * in the pass) or all are re-formed in every gradient pass (0),
 * how many cells a tile holds (128, 256 or 512, one thread each; the default 256;
 * setting it rebuilds the tiles on the host), and whether the second-order interior
 * flux stages tiles of 64, 128 or 256 cells (with the cells their faces reach) in LDS
 * and evaluates each tile's faces from there (FACE_TILE; 0, the default, is one
 * thread per face gathering its two cells from global memory), and whether the
 * update then reads the fluxes of each tile's own faces from an LDS copy
 * (UPDATE_TILE 1; only while face tiles exist; reads as 0 otherwise), and how many
 * threads evaluate a face tile (FACE_THREADS 256, the default, or 512: twice the
 * resident waves where the tiles' LDS holds a CU to two blocks).  None of these
 * changes results.  GRAD_LDS 2 (FP32 gradients only: the mixed mode or FP32_GRAD;
 * groups of 4 or 8 members) stages each union record as eight floats, its centroid and
 * primitive state less those of the tile's first cell rounded once from FP64, and runs
 * the member loop in FP32 with the hardware reciprocal square root: half the LDS of
 * the FP64 records, FP32 rounding in the gradients (within the mixed mode's accuracy,
 * not the bits of GRAD_LDS 1); with GRAD_FAST or stored weights the FP64-layout
 * records are staged and GRAD_LDS reads 1.  GRAD_FAST 1 replaces the gradient's member loop by its moment form
 * (per member one reciprocal of |d|^2 and fused multiply-adds, the cell's inverse
 * moment matrix applied once; LDS tiles only, ahead of stored weights): the same least
 * squares gradient with other rounding, so results differ from the reference's in the
 * last bits (not part of the bit-identical set; 0 is the default).  PERSISTENT 1 (a
 * batch at order 2 with SSP-RK3, no modifier or the Venkatakrishnan limiter, 256-cell
 * gradient tiles and face tiles) advances every case by one time step in one
 * persistent launch whose resident blocks take the gradient tiles, face tiles,
 * boundary faces, per-case time steps and update tiles of all cases phase by phase
 * between grid barriers; the same bits as the separate launches (0, the default).  A
 * barrier wait that times out (5 s) makes rflu_gpu_time_steps return RFLU_ERR_HIP until
 * PERSISTENT is set to 1 again. */
#define RFLU_OPT_GRAD_LDS 1
#define RFLU_OPT_GRAD_UNROLL 2
#define RFLU_OPT_TILE_CAP 3
#define RFLU_OPT_XCD 4
#define RFLU_OPT_GRAD_WEIGHTS 5
#define RFLU_OPT_GRAD_TILE 6
#define RFLU_OPT_FACE_TILE 7
#define RFLU_OPT_UPDATE_TILE 8
#define RFLU_OPT_GRAD_FAST 9
#define RFLU_OPT_PERSISTENT 10
#define RFLU_OPT_FACE_THREADS 11
int rflu_gpu_set_option(int32_t handle, int32_t option, int32_t value);
int rflu_gpu_get_option(int32_t handle, int32_t option, int32_t* value);

/* Times `reps` launches of one kernel of the stage loop on the region's stream with
 * HIP events and returns the average milliseconds per launch (which: 0 face flux,
 * 1 update, 2 gradient, 3 limiter/blend, 4 time step, 5 fused stage).  Waits. */
int rflu_gpu_time_kernel(int32_t handle, int32_t which, int32_t reps, double* ms);

/* Per-kernel timing of rflu_gpu_time_steps: when on, every launch is bracketed by HIP
 * events on the region's stream (graphs are not used) and the milliseconds are
 * accumulated per kernel class (0 face flux, 1 update, 2 gradient, 3 blend, 4 time
 * step, 5 boundary flux).  Turning it on or off clears the totals. */
int rflu_gpu_set_profiling(int32_t handle, int32_t on);
int rflu_gpu_get_kernel_times(int32_t handle, double* ms, int64_t* counts, int32_t n);

/*
 * Fluid dumps (restart files), the reference's RFLUDUMP stream layout, version 5
 * (modflu/RFLU_ModRestartData.F90:7-33): "RFLUDUMP", int32 version, int64 step,
 * double time, int32 region, nCells, nCv, form, withPicl, then cv(iCv,1:nCells) one
 * variable after the other, then int32 number of NSCBC patches (with their face
 * states).  Gas-only regions write the pure form (1), no PICL fields and no NSCBC
 * patches; the reader accepts versions 4 and 5 with the checks of
 * RFLU_ReadRestartData (:693-872): sizes must match the region, the form must be the
 * pure one, PICL fields are read past, and a dump with NSCBC face states is refused.
 */
/* restart_data/fluid_<step:8 digits>_<region:5 digits>.dat under out_dir (which ends
 * in '/' like global%outDir, or is empty), as FluidDumpName (:258-266).  buf gets at
 * most buflen bytes including the terminating NUL. */
int rflu_dump_name(char* buf, int32_t buflen, const char* out_dir, int64_t step, int32_t region);
/* Host arrays: writes cv(n_cv, n_cells) (component fastest) as a dump, through
 * <path>.tmp and a rename (RFLU_WriteRestartData :136-215, RFLU_FinishRestartFile
 * :904-918). */
int rflu_dump_write(const char* path, int64_t step, double time, int32_t region,
                    int32_t n_cells, int32_t n_cv, const double* cv);
/* Host arrays: reads a dump of n_cells cells and n_cv variables into cv(n_cv, n_cells);
 * step, time and region receive the header fields (any may be NULL). */
int rflu_dump_read(const char* path, int32_t n_cells, int32_t n_cv, int64_t* step,
                   double* time, int32_t* region, double* cv);
/* A case of a region handle (0 for a single region; case i of a batch): writes its
 * current state as the dump of `region` at (step, time), the host's
 * global%currentStep/currentTime.  A region with virtual cells writes its actual cells
 * (cv(:,1:nCells), as the reference does).  Waits for the stream. */
int rflu_gpu_write_dump(int32_t handle, int32_t icase, const char* path, int64_t step,
                        double time, int32_t region);
/* Reads a dump into case icase: its state, dependent variables and clock (the dump's
 * time; counters and flags of the case reset as after an upload).  Virtual cells keep
 * their values until the next rflu_gpu_exchange.  step, time and
 * region receive the header fields (any may be NULL).  Waits for the stream. */
int rflu_gpu_read_dump(int32_t handle, int32_t icase, const char* path, int64_t* step,
                       double* time, int32_t* region);

const char* rflu_gpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
