!******************************************************************************
!
! rflu_gpu -- ISO_C_BINDING interface of the MI355X explicit residual/update path
! (include/rocflu_amd.h), of the host region builder and grid files
! (include/rflu_mesh.h) and of the user input files (include/rflu_input.h).
!
! A reference host keeps its t_region and replaces the bodies of RungeKuttaMP's
! callees with these calls, one region handle per t_region:
!   RKInitMP            -> rflu_gpu_rk_init              (libfloflu/RkInitMP.F90)
!   CellGradientsMP     -> rflu_gpu_cell_gradients       (libfloflu/CellGradientsMP.F90)
!   ConvectiveFluxesMP  -> rflu_gpu_convective_fluxes    (libfloflu/ConvectiveFluxesMP.F90)
!   RKUpdateMP          -> rflu_gpu_rk_update            (libfloflu/RkUpdateMP.F90)
!   RFLU_SetVarsContWrapper -> rflu_gpu_set_dependent_vars
!   AfterUpdateMP       -> rflu_gpu_after_update         (libfloflu/AfterUpdateMP.F90)
!   RFLU_TimeStepInviscid + RFLU_MinimumTimeStep -> rflu_gpu_time_step
! Arrays are passed as they are stored in t_grid/t_patch (column-major, component
! index fastest) with index_base = 1.
!
!******************************************************************************
MODULE rflu_gpu

  USE ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: RFLU_OK = 0, RFLU_ERR_INVALID_VALUE = 1, &
                               RFLU_ERR_NEGATIVE_POSDEF = 2, RFLU_ERR_ARG = 3, &
                               RFLU_ERR_HIP = 4, RFLU_ERR_NOMEM = 5, &
                               RFLU_ERR_HANDLE = 6, RFLU_ERR_FILE = 7, &
                               RFLU_ERR_DUMP = 8
  INTEGER(C_INT32_T), PARAMETER :: RFLU_XCHG_STATE = 1, RFLU_XCHG_GRAD = 2

  TYPE, BIND(C) :: rflu_region_desc
    INTEGER(C_INT32_T) :: index_base, n_cells, n_faces, n_patches
    TYPE(C_PTR) :: f2c, fn, fc, cofg, vol
    TYPE(C_PTR) :: bpatch_ptr, bf2c, bfn, bfc
    TYPE(C_PTR) :: patch_bc, patch_opt, patch_pout, patch_order
    TYPE(C_PTR) :: c2cs_ptr, c2cs_idx, xyz_moms
    REAL(C_DOUBLE) :: gamma, cp
    INTEGER(C_INT32_T) :: discr, order, reconst, rk, precision
    REAL(C_DOUBLE) :: cfl, dt_imposed, dt_min_limit, eps_entr, diss_fact
    TYPE(C_PTR) :: patch_vals
    TYPE(C_PTR) :: c2cs_bptr, c2cs_bface
    REAL(C_DOUBLE) :: creconst_weight
    INTEGER(C_INT32_T) :: creconst, pad_
    TYPE(C_PTR) :: patch_opt2
    INTEGER(C_INT32_T) :: n_virtual, pad2_
  END TYPE rflu_region_desc

  ! rflu_inp and rflu_bc_patch of include/rflu_input.h
  TYPE, BIND(C) :: rflu_inp
    INTEGER(C_INT32_T) :: grid_format, solution_format, grid_source
    INTEGER(C_INT32_T) :: flow_model, move_grid, gas_model
    REAL(C_DOUBLE) :: cfl
    INTEGER(C_INT32_T) :: discr, order
    REAL(C_DOUBLE) :: eps_entr, diss_fact
    INTEGER(C_INT32_T) :: reconst, dimens, creconst_cells, creconst_faces
    REAL(C_DOUBLE) :: creconst_cells_weight, creconst_faces_weight, toler_ict
    INTEGER(C_INT32_T) :: stencil_dimens_cells, stencil_dimens_faces, stencil_dimens_bfaces
    INTEGER(C_INT32_T) :: order_bfaces, axi_flag
    REAL(C_DOUBLE) :: sigma, kp, ku, minf
    INTEGER(C_INT32_T) :: flow_type, solver_type, rk_scheme, start_step, max_iter
    INTEGER(C_INT32_T) :: write_iter, print_iter
    REAL(C_DOUBLE) :: dt_imposed, max_time, write_time, print_time, dt_min_limit, res_tol
    REAL(C_DOUBLE) :: ref_velocity, ref_pressure, ref_density, ref_cp, ref_gamma, ref_length
    INTEGER(C_INT32_T) :: init_flag
    REAL(C_DOUBLE) :: init_vel(3), init_press, init_dens, init_temp
    INTEGER(C_INT32_T) :: init_ival(19)
    REAL(C_DOUBLE) :: init_rval(25)
    INTEGER(C_INT32_T) :: sections
  END TYPE rflu_inp

  TYPE, BIND(C) :: rflu_bc_patch
    INTEGER(C_INT32_T) :: bc_type, opt, opt2, order, creconst, related, axis
    REAL(C_DOUBLE) :: angle
    REAL(C_DOUBLE) :: vals(5)
    REAL(C_DOUBLE) :: pout
    CHARACTER(KIND=C_CHAR) :: name(80)
  END TYPE rflu_bc_patch

  TYPE, BIND(C) :: rflu_gpu_stats
    INTEGER(C_INT64_T) :: n_invalid, n_negative, n_bounds_fix, n_pos_fix
    REAL(C_DOUBLE) :: mass_in, mass_out, time, dt
    INTEGER(C_INT64_T) :: steps
    INTEGER(C_INT32_T) :: finished, pad
  END TYPE rflu_gpu_stats

  INTERFACE
    INTEGER(C_INT) FUNCTION rflu_gpu_device_count() BIND(C, NAME='rflu_gpu_device_count')
      IMPORT :: C_INT
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_create_region(desc, handle) &
        BIND(C, NAME='rflu_gpu_create_region')
      IMPORT :: C_INT, C_INT32_T, rflu_region_desc
      TYPE(rflu_region_desc), INTENT(IN) :: desc
      INTEGER(C_INT32_T), INTENT(OUT) :: handle
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_destroy_region(handle) &
        BIND(C, NAME='rflu_gpu_destroy_region')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: handle
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_upload_state(handle, cv, dv, time) &
        BIND(C, NAME='rflu_gpu_upload_state')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
      INTEGER(C_INT32_T), VALUE :: handle
      REAL(C_DOUBLE), INTENT(IN) :: cv(*)
      TYPE(C_PTR), VALUE :: dv
      REAL(C_DOUBLE), VALUE :: time
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_download_state(handle, cv, dv) &
        BIND(C, NAME='rflu_gpu_download_state')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: handle
      REAL(C_DOUBLE), INTENT(OUT) :: cv(*), dv(*)
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_time_step(handle, current_time, dt_out) &
        BIND(C, NAME='rflu_gpu_time_step')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: handle
      REAL(C_DOUBLE), VALUE :: current_time
      REAL(C_DOUBLE), INTENT(OUT) :: dt_out
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_rk_init(handle, istage) BIND(C, NAME='rflu_gpu_rk_init')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: handle, istage
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_cell_gradients(handle) &
        BIND(C, NAME='rflu_gpu_cell_gradients')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: handle
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_convective_fluxes(handle, istage) &
        BIND(C, NAME='rflu_gpu_convective_fluxes')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: handle, istage
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_rk_update(handle, istage, dt_min) &
        BIND(C, NAME='rflu_gpu_rk_update')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: handle, istage
      REAL(C_DOUBLE), VALUE :: dt_min
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_set_dependent_vars(handle) &
        BIND(C, NAME='rflu_gpu_set_dependent_vars')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: handle
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_after_update(handle, istage) &
        BIND(C, NAME='rflu_gpu_after_update')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: handle, istage
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_check_status(handle) BIND(C, NAME='rflu_gpu_check_status')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: handle
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_residual(handle, rhs) BIND(C, NAME='rflu_gpu_residual')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: handle
      REAL(C_DOUBLE), INTENT(OUT) :: rhs(*)
    END FUNCTION

    ! Regions coupled through virtual cells: borders (t_border) and the halo copy
    ! (RFLU_MPI_CopyWrapper / RFLU_MPI_ExchangeGradCells for regions on one device).
    INTEGER(C_INT) FUNCTION rflu_gpu_set_borders(handle, n_borders, nbr_handle, nbr_border, &
        send_ptr, send_cells, recv_ptr, recv_cells) BIND(C, NAME='rflu_gpu_set_borders')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: handle, n_borders
      INTEGER(C_INT32_T), INTENT(IN) :: nbr_handle(*), nbr_border(*), send_ptr(*), &
                                        send_cells(*), recv_ptr(*), recv_cells(*)
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_exchange(n_regions, handles, what) &
        BIND(C, NAME='rflu_gpu_exchange')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: n_regions, what
      INTEGER(C_INT32_T), INTENT(IN) :: handles(*)
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_set_border_transform(handle, border, m) &
        BIND(C, NAME='rflu_gpu_set_border_transform')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: handle, border
      REAL(C_DOUBLE), INTENT(IN) :: m(9)
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_time_steps(handle, nsteps, use_graph) &
        BIND(C, NAME='rflu_gpu_time_steps')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T
      INTEGER(C_INT32_T), VALUE :: handle
      INTEGER(C_INT64_T), VALUE :: nsteps
      INTEGER(C_INT32_T), VALUE :: use_graph
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_set_end_time(handle, end_time) &
        BIND(C, NAME='rflu_gpu_set_end_time')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: handle
      REAL(C_DOUBLE), VALUE :: end_time
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_get_stats(handle, stats) BIND(C, NAME='rflu_gpu_get_stats')
      IMPORT :: C_INT, C_INT32_T, rflu_gpu_stats
      INTEGER(C_INT32_T), VALUE :: handle
      TYPE(rflu_gpu_stats), INTENT(OUT) :: stats
    END FUNCTION

    ! fluid dumps: in place of RFLU_WriteRestartData / RFLU_ReadRestartData
    ! (modflu/RFLU_ModRestartData.F90:136-215,693-872); paths end in C_NULL_CHAR
    INTEGER(C_INT) FUNCTION rflu_gpu_write_dump(handle, icase, path, step, time, region) &
        BIND(C, NAME='rflu_gpu_write_dump')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_CHAR
      INTEGER(C_INT32_T), VALUE :: handle, icase, region
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      INTEGER(C_INT64_T), VALUE :: step
      REAL(C_DOUBLE), VALUE :: time
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_gpu_read_dump(handle, icase, path, step, time, region) &
        BIND(C, NAME='rflu_gpu_read_dump')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_CHAR
      INTEGER(C_INT32_T), VALUE :: handle, icase
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      INTEGER(C_INT64_T), INTENT(OUT) :: step
      REAL(C_DOUBLE), INTENT(OUT) :: time
      INTEGER(C_INT32_T), INTENT(OUT) :: region
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_dump_name(buf, buflen, out_dir, step, region) &
        BIND(C, NAME='rflu_dump_name')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_CHAR
      CHARACTER(KIND=C_CHAR), INTENT(OUT) :: buf(*)
      INTEGER(C_INT32_T), VALUE :: buflen, region
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: out_dir(*)
      INTEGER(C_INT64_T), VALUE :: step
    END FUNCTION

    FUNCTION rflu_gpu_last_error() BIND(C, NAME='rflu_gpu_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR) :: rflu_gpu_last_error
    END FUNCTION

    ! host region builder (include/rflu_mesh.h)
    INTEGER(C_INT) FUNCTION rflu_mesh_box(kind, nx, ny, nz, lx, ly, lz, x0, y0, z0, jitter, &
                                          seed, bc_type, bc_opt, bc_pout, build_stencils, out) &
        BIND(C, NAME='rflu_mesh_box')
      IMPORT :: C_INT, C_DOUBLE, C_INT64_T, C_INT32_T, C_PTR
      INTEGER(C_INT), VALUE :: kind, nx, ny, nz, build_stencils
      REAL(C_DOUBLE), VALUE :: lx, ly, lz, x0, y0, z0, jitter
      INTEGER(C_INT64_T), VALUE :: seed
      INTEGER(C_INT32_T), INTENT(IN) :: bc_type(6), bc_opt(6)
      REAL(C_DOUBLE), INTENT(IN) :: bc_pout(6)
      TYPE(C_PTR), INTENT(OUT) :: out
    END FUNCTION

    SUBROUTINE rflu_mesh_destroy(m) BIND(C, NAME='rflu_mesh_destroy')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: m
    END SUBROUTINE

    INTEGER(C_INT) FUNCTION rflu_mesh_sizes(m, out) BIND(C, NAME='rflu_mesh_sizes')
      IMPORT :: C_INT, C_PTR, C_INT64_T
      TYPE(C_PTR), VALUE :: m
      INTEGER(C_INT64_T), INTENT(OUT) :: out(9)
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_mesh_array(m, name, ptr, count, dtype) &
        BIND(C, NAME='rflu_mesh_array')
      IMPORT :: C_INT, C_PTR, C_CHAR, C_INT64_T
      TYPE(C_PTR), VALUE :: m
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: name(*)
      TYPE(C_PTR), INTENT(OUT) :: ptr
      INTEGER(C_INT64_T), INTENT(OUT) :: count
      INTEGER(C_INT), INTENT(OUT) :: dtype
    END FUNCTION

    FUNCTION rflu_mesh_last_error() BIND(C, NAME='rflu_mesh_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR) :: rflu_mesh_last_error
    END FUNCTION

    ! grids (include/rflu_mesh.h)
    INTEGER(C_INT) FUNCTION rflu_grid_read_centaur(path, out) &
        BIND(C, NAME='rflu_grid_read_centaur')
      IMPORT :: C_INT, C_PTR, C_CHAR
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      TYPE(C_PTR), INTENT(OUT) :: out
    END FUNCTION

    SUBROUTINE rflu_grid_destroy(g) BIND(C, NAME='rflu_grid_destroy')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: g
    END SUBROUTINE

    INTEGER(C_INT) FUNCTION rflu_grid_sizes(g, out) BIND(C, NAME='rflu_grid_sizes')
      IMPORT :: C_INT, C_PTR, C_INT64_T
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT64_T), INTENT(OUT) :: out(8)
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_mesh_build_sype(g, bc_type, bc_opt, bc_pout, related, axis, &
                                                 angle, build_stencils, out) &
        BIND(C, NAME='rflu_mesh_build_sype')
      IMPORT :: C_INT, C_PTR, C_INT32_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT32_T), INTENT(IN) :: bc_type(*), bc_opt(*), related(*), axis(*)
      REAL(C_DOUBLE), INTENT(IN) :: bc_pout(*), angle(*)
      INTEGER(C_INT), VALUE :: build_stencils
      TYPE(C_PTR), INTENT(OUT) :: out
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_mesh_build(g, bc_type, bc_opt, bc_pout, build_stencils, out) &
        BIND(C, NAME='rflu_mesh_build')
      IMPORT :: C_INT, C_PTR, C_INT32_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT32_T), INTENT(IN) :: bc_type(*), bc_opt(*)
      REAL(C_DOUBLE), INTENT(IN) :: bc_pout(*)
      INTEGER(C_INT), VALUE :: build_stencils
      TYPE(C_PTR), INTENT(OUT) :: out
    END FUNCTION

    ! user input files (include/rflu_input.h)
    INTEGER(C_INT) FUNCTION rflu_read_inp(path, out) BIND(C, NAME='rflu_read_inp')
      IMPORT :: C_INT, C_CHAR, rflu_inp
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      TYPE(rflu_inp), INTENT(OUT) :: out
    END FUNCTION

    INTEGER(C_INT) FUNCTION rflu_read_bc(path, n_patches, order_bfaces, out) &
        BIND(C, NAME='rflu_read_bc')
      IMPORT :: C_INT, C_CHAR, rflu_bc_patch
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      INTEGER(C_INT), VALUE :: n_patches, order_bfaces
      TYPE(rflu_bc_patch), INTENT(OUT) :: out(*)
    END FUNCTION

    FUNCTION rflu_input_last_error() BIND(C, NAME='rflu_input_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR) :: rflu_input_last_error
    END FUNCTION
  END INTERFACE

CONTAINS

  ! Stops with the library's message, as the reference's ErrorStop would.
  SUBROUTINE rflu_gpu_check(rc, where)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(*), INTENT(IN) :: where
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: n
    IF ( rc /= RFLU_OK ) THEN
      CALL C_F_POINTER(rflu_gpu_last_error(), msg, [512])
      n = 0
      DO WHILE ( n < 512 )
        IF ( msg(n+1) == C_NULL_CHAR ) EXIT
        n = n + 1
      END DO
      WRITE(*,'(A,A,I0,A,512A)') where, ': error ', rc, ': ', msg(1:n)
      STOP 1
    END IF
  END SUBROUTINE rflu_gpu_check

  ! The same for the host libraries: errptr is rflu_mesh_last_error() or
  ! rflu_input_last_error().
  SUBROUTINE rflu_host_check(rc, where, errptr)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(*), INTENT(IN) :: where
    TYPE(C_PTR), INTENT(IN) :: errptr
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: n
    IF ( rc /= 0 ) THEN
      CALL C_F_POINTER(errptr, msg, [512])
      n = 0
      DO WHILE ( n < 512 )
        IF ( msg(n+1) == C_NULL_CHAR ) EXIT
        n = n + 1
      END DO
      WRITE(*,'(A,A,I0,A,512A)') where, ': error ', rc, ': ', msg(1:n)
      STOP 1
    END IF
  END SUBROUTINE rflu_host_check

END MODULE rflu_gpu
