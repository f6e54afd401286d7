/*
 * rflu_input.h -- the reference's user input files, read on the host.
 *
 * A run of the reference is described by <casename>.inp (sections "# NAME" ... "#",
 * one "KEY value" per line, read by libfloflu/ReadInputFile.F90 and the section readers
 * it calls) and <casename>.bc (one "# BC_TYPE" section per patch range, read by
 * modflu/RFLU_ModReadBcInputFile.F90).  These functions read the sections the explicit
 * residual/update path needs with the reference's rules:
 *   - a section starts at a line equal to its header (trailing blanks ignored) and
 *     ends at the next line that starts with '#', which it consumes;
 *   - a key matches a line that starts with the key followed by a blank or a tab
 *     (LineHasKey, libfloflu/ReadInputUtil.F90:686-700) and its value is the first
 *     item after the key; in .inp sections the first occurrence wins (ReadSection),
 *     in .bc sections the last (ReadPatchSection);
 *   - unknown keys and lines outside sections are ignored.
 * Values are converted as the section readers convert them (ABS of reals, NINT of
 * switches, degrees to radians of angles).  Fields a file does not set keep the
 * reference's defaults (libflu/RFLU_InitInputValues.F90:149-205,
 * libflu/RFLU_InitGlobal.F90:229-353).
 */
#ifndef RFLU_INPUT_H
#define RFLU_INPUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sections found (rflu_inp.sections bits). */
#define RFLU_INP_FORMATS   (1u << 0)
#define RFLU_INP_FLOWMODEL (1u << 1)
#define RFLU_INP_NUMERICS  (1u << 2)
#define RFLU_INP_TIMESTEP  (1u << 3)
#define RFLU_INP_REFERENCE (1u << 4)
#define RFLU_INP_MIXTURE   (1u << 5)
#define RFLU_INP_INITFLOW  (1u << 6)

typedef struct rflu_inp {
  /* # FORMATS (libfloflu/ReadFormatsSection.F90): GRID, SOLUTION, GRIDSRC */
  int32_t grid_format, solution_format, grid_source;
  /* # FLOWMODEL (libfloflu/ReadFlowSection.F90): MODEL, MOVEGRID */
  int32_t flow_model, move_grid;
  /* # MIXTURE (libfloflu/ReadMixtureSection.F90): GASMODEL */
  int32_t gas_model;
  /* # NUMERICS (libfloflu/ReadNumericsSection.F90:110-300) */
  double cfl;                     /* CFL (3.0) */
  int32_t discr;                  /* DISCR (1, Roe) */
  int32_t order;                  /* ORDER (1) */
  double eps_entr;                /* ENTROPY (0.05) */
  double diss_fact;               /* DISSFACT (1.0) */
  int32_t reconst;                /* RECONST (2, WENO-XYZ) */
  int32_t dimens;                 /* DIMENS (3) */
  int32_t creconst_cells;         /* CRECONSTC (0) */
  int32_t creconst_faces;         /* CRECONSTF (0) */
  double creconst_cells_weight;   /* CRECONSTCW (1.0) */
  double creconst_faces_weight;   /* CRECONSTFW (1.0) */
  double toler_ict;               /* TOLERICT (1e-11) */
  int32_t stencil_dimens_cells;   /* SDIMENSC (3) */
  int32_t stencil_dimens_faces;   /* SDIMENSF (3) */
  int32_t stencil_dimens_bfaces;  /* SDIMENSBF (3) */
  int32_t order_bfaces;           /* ORDERBF (1): patch order where the .bc has no ORDER */
  int32_t axi_flag;               /* AXIFLAG (0) */
  double sigma, kp, ku, minf;     /* SIGMA, KP, KU, MINF (1.0, 0.25, 0.75, 0.1) */
  /* # TIMESTEP (libfloflu/ReadTimestepSection.F90:95-200) */
  int32_t flow_type;              /* FLOWTYPE (0 steady, 1 unsteady) */
  int32_t solver_type;            /* SOLVERTYPE (1 explicit) */
  int32_t rk_scheme;              /* RKSCHEME (1 classical RK4, 2 Wray RK3, 3 Euler) */
  int32_t start_step;             /* START_STEP (0) */
  int32_t max_iter;               /* MAXITER (10000) */
  int32_t write_iter, print_iter; /* WRIITER, PRNITER */
  double dt_imposed;              /* TIMESTEP (1e-5) */
  double max_time;                /* MAXTIME (1e-4) */
  double write_time;              /* WRITIME (5e-5) */
  double print_time;              /* PRNTIME (WRITIME when the section omits it) */
  double dt_min_limit;            /* DTMINLIMIT (1e-10) */
  double res_tol;                 /* RESTOL */
  /* # REFERENCE (libfloflu/ReadReferenceSection.F90) */
  double ref_velocity, ref_pressure, ref_density, ref_cp, ref_gamma, ref_length;
  /* # INITFLOW (libfloflu/ReadInitFlowSection.F90): FLAG, VELX, VELY, VELZ, PRESS,
   * DENS, TEMP, IVAL1..IVAL19 (truncated to integers), RVAL1..RVAL25 */
  int32_t init_flag;
  double init_vel[3], init_press, init_dens, init_temp;
  int32_t init_ival[19];
  double init_rval[25];
  uint32_t sections;              /* RFLU_INP_* bits of the sections found */
} rflu_inp;

/* Fill *out with the defaults, then read the file.  Returns 0, 7 (cannot open) or 8
 * (a value that does not read, with file:line in rflu_input_last_error). */
int rflu_read_inp(const char* path, rflu_inp* out);

/* Boundary types of the .bc sections (modfloflu/ModParameters.F90:397-441). */
#define RFLU_BCT_INFLOW_TOTANG   10
#define RFLU_BCT_INFLOW_VELTEMP  11
#define RFLU_BCT_OUTFLOW         20
#define RFLU_BCT_SLIPWALL        60
#define RFLU_BCT_NOSLIPWALL_HFLUX 70
#define RFLU_BCT_NOSLIPWALL_TEMP 71
#define RFLU_BCT_FARFIELD        80
#define RFLU_BCT_INJECTION       90
#define RFLU_BCT_SYMMETRY       100
#define RFLU_BCT_PERIODIC       110
#define RFLU_BCT_VIRTUAL        130

typedef struct rflu_bc_patch {
  int32_t bc_type;      /* 0 when no section names the patch */
  int32_t opt;          /* inflow/outflow TYPE (BCSWI_INFLOW_TYPE / BCSWI_OUTFLOW_TYPE) */
  int32_t opt2;         /* inflow 10 FIXED, farfield CORR */
  int32_t order;        /* patch space order: ORDER 2 -> 2, else 1; ORDERBF when absent */
  int32_t creconst;     /* slip wall / injection CRECONST (0 none, 1 weighted) */
  int32_t related;      /* periodic RELPATCH (1-based) */
  int32_t axis;         /* periodic AXIS */
  double angle;         /* periodic ANGLE, radians */
  /* the patch's five values in the order of rflu_region_desc.patch_vals:
   * inflow 11 (VELX, VELY, VELZ, TEMP, PRESS), inflow 10 (PTOT, TTOT, BETAH, BETAV,
   * MACH; angles in radians), farfield (MACH, ATTACK, SLIP, PRESS, TEMP; angles in
   * radians), injection (MFRATE, TEMP, 0, 0, 0); zeros elsewhere */
  double vals[5];
  double pout;          /* outflow PRESS (0 for supersonic outflow) */
  char name[80];        /* NAME */
} rflu_bc_patch;

/* Read a .bc file for a grid of n_patches patches (numbered 1..n_patches as in the
 * grid file).  order_bfaces is the .inp's ORDERBF (the patch order where a section
 * has no ORDER).  A section whose patch range ends beyond n_patches, a missing
 * switch or a missing value is an error (the reference's ErrorStops).
 * Returns 0, 7 (cannot open) or 8. */
int rflu_read_bc(const char* path, int n_patches, int order_bfaces, rflu_bc_patch* out);

const char* rflu_input_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
