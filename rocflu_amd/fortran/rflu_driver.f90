!******************************************************************************
!
! rflu_driver -- a Fortran host driving the GPU path through rflu_gpu, the way the
! reference's RFLU_TimeStepping / RungeKuttaMP would after the swap: Fortran-owned,
! 1-based t_grid-style arrays, the per-callee entry points in RungeKuttaMP's stage
! order, and the host-side time advance.
!
! Usage: rflu_driver nx ny nz nsteps discr order out.bin
!   Builds the Sod box (nx x ny x nz unit hexahedra, nodes jittered 10%, seed 1,
!   supersonic outflow at both x ends, slip walls elsewhere), runs nsteps Wray RK3
!   steps and writes cv(5,nCells) as raw doubles to out.bin, and the region's fluid
!   dump (step nsteps, region 0) to out.bin.dump.
!
!******************************************************************************
PROGRAM rflu_driver

  USE ISO_C_BINDING
  USE rflu_gpu
  IMPLICIT NONE

  CHARACTER(LEN=256) :: arg, outfile
  INTEGER :: nx, ny, nz, nsteps, discr, order, istep, istage, nStages, icg, k, u
  INTEGER(C_INT32_T) :: handle
  INTEGER(C_INT32_T) :: bct(6), bco(6)
  REAL(C_DOUBLE) :: bcp(6)
  INTEGER(C_INT64_T) :: sizes(9), cnt
  INTEGER(C_INT) :: dtype
  TYPE(C_PTR) :: mesh, p
  TYPE(rflu_region_desc), TARGET :: desc
  INTEGER(C_INT32_T), POINTER :: ip(:)
  REAL(C_DOUBLE), POINTER :: dp(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: f2c(:,:), bf2c(:), bptr(:), pbc(:), popt(:), &
                                             porder(:), sptr(:), sidx(:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: fn(:,:), fc(:,:), cofg(:,:), vol(:), bfn(:,:), &
                                         bfc(:,:), pout(:), moms(:,:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: cv(:,:), dv(:,:)
  REAL(C_DOUBLE) :: time, dt, cp, g, gc, mm, gp, rho, pres
  INTEGER :: nCells, nFaces, nBFaces

  CALL GET_COMMAND_ARGUMENT(1, arg); READ(arg,*) nx
  CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg,*) ny
  CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg,*) nz
  CALL GET_COMMAND_ARGUMENT(4, arg); READ(arg,*) nsteps
  CALL GET_COMMAND_ARGUMENT(5, arg); READ(arg,*) discr
  CALL GET_COMMAND_ARGUMENT(6, arg); READ(arg,*) order
  CALL GET_COMMAND_ARGUMENT(7, outfile)

! ------------------------------------------------------------------------------
! Region: built once on the host (the reference's RFLU_InitFlowSolver)
! ------------------------------------------------------------------------------
  bct = (/ 20, 20, 60, 60, 60, 60 /)
  bco = 0
  bcp = 0.0_C_DOUBLE
  IF ( rflu_mesh_box(0_C_INT, INT(nx,C_INT), INT(ny,C_INT), INT(nz,C_INT), &
                     REAL(nx,C_DOUBLE), REAL(ny,C_DOUBLE), REAL(nz,C_DOUBLE), &
                     0.0_C_DOUBLE, 0.0_C_DOUBLE, 0.0_C_DOUBLE, 0.1_C_DOUBLE, 1_C_INT64_T, &
                     bct, bco, bcp, MERGE(1_C_INT, 0_C_INT, order == 2), mesh) /= 0 ) STOP 2
  IF ( rflu_mesh_sizes(mesh, sizes) /= 0 ) STOP 3
  nCells = INT(sizes(1)); nFaces = INT(sizes(2)); nBFaces = INT(sizes(3))

  ALLOCATE(f2c(2,nFaces), fn(4,nFaces), fc(3,nFaces), cofg(3,nCells), vol(nCells))
  ALLOCATE(bf2c(nBFaces), bfn(4,nBFaces), bfc(3,nBFaces), bptr(7), pbc(6), popt(6), &
           porder(6), pout(6))

  CALL get_int('f2c', ip);   f2c  = RESHAPE(ip, (/2,nFaces/)) + 1      ! 1-based cells
  CALL get_dbl('fn', dp);    fn   = RESHAPE(dp, (/4,nFaces/))
  CALL get_dbl('fc', dp);    fc   = RESHAPE(dp, (/3,nFaces/))
  CALL get_dbl('cofg', dp);  cofg = RESHAPE(dp, (/3,nCells/))
  CALL get_dbl('vol', dp);   vol  = dp
  CALL get_int('bf2c', ip);  bf2c = ip + 1
  CALL get_dbl('bfn', dp);   bfn  = RESHAPE(dp, (/4,nBFaces/))
  CALL get_dbl('bfc', dp);   bfc  = RESHAPE(dp, (/3,nBFaces/))
  CALL get_int('bpatch_ptr', ip); bptr = ip                            ! offsets stay 0-based
  CALL get_int('patch_bc', ip);   pbc  = ip
  CALL get_int('patch_opt', ip);  popt = ip
  CALL get_dbl('patch_pout', dp); pout = dp
  porder = 1
  IF ( order == 2 ) THEN
    CALL get_int('c2cs_ptr', ip)
    ALLOCATE(sptr(nCells+1)); sptr = ip
    CALL get_int('c2cs_idx', ip)
    ALLOCATE(sidx(SIZE(ip))); sidx = ip + 1
    CALL get_dbl('xyz_moms', dp)
    ALLOCATE(moms(10,nCells)); moms = RESHAPE(dp, (/10,nCells/))
  END IF
  CALL rflu_mesh_destroy(mesh)

  cp = 1004.5_C_DOUBLE
  g  = 1.4_C_DOUBLE
  desc%index_base = 1
  desc%n_cells = nCells; desc%n_faces = nFaces; desc%n_patches = 6
  desc%f2c = C_LOC(f2c); desc%fn = C_LOC(fn); desc%fc = C_LOC(fc)
  desc%cofg = C_LOC(cofg); desc%vol = C_LOC(vol)
  desc%bpatch_ptr = C_LOC(bptr); desc%bf2c = C_LOC(bf2c)
  desc%bfn = C_LOC(bfn); desc%bfc = C_LOC(bfc)
  desc%patch_bc = C_LOC(pbc); desc%patch_opt = C_LOC(popt)
  desc%patch_pout = C_LOC(pout); desc%patch_order = C_LOC(porder)
  IF ( order == 2 ) THEN
    desc%c2cs_ptr = C_LOC(sptr); desc%c2cs_idx = C_LOC(sidx); desc%xyz_moms = C_LOC(moms)
  ELSE
    desc%c2cs_ptr = C_NULL_PTR; desc%c2cs_idx = C_NULL_PTR; desc%xyz_moms = C_NULL_PTR
  END IF
  desc%gamma = g; desc%cp = cp
  desc%discr = discr; desc%order = order
  desc%reconst = MERGE(2, 0, order == 2)
  desc%rk = 2; desc%precision = 0
  desc%cfl = 0.5_C_DOUBLE; desc%dt_imposed = 1.0E-3_C_DOUBLE; desc%dt_min_limit = 0.0_C_DOUBLE
  desc%eps_entr = 0.05_C_DOUBLE; desc%diss_fact = 1.0_C_DOUBLE
  desc%patch_vals = C_NULL_PTR
  desc%c2cs_bptr = C_NULL_PTR; desc%c2cs_bface = C_NULL_PTR
  desc%creconst_weight = 1.0_C_DOUBLE; desc%creconst = 0; desc%pad_ = 0
  desc%patch_opt2 = C_NULL_PTR
  desc%n_virtual = 0; desc%pad2_ = 0
  CALL rflu_gpu_check(rflu_gpu_create_region(desc, handle), 'rflu_gpu_create_region')

! ------------------------------------------------------------------------------
! Initial state: Sod along x, E from gamma' (RFLU_InitFlowHardCode.F90:911-915)
! ------------------------------------------------------------------------------
  gc = cp - cp/g
  mm = 8314.3_C_DOUBLE/gc
  gc = 8314.3_C_DOUBLE/mm
  gp = cp/(cp - gc)
  ALLOCATE(cv(5,nCells), dv(3,nCells))
  DO icg = 1,nCells
    IF ( cofg(1,icg) < 0.5_C_DOUBLE*nx ) THEN
      rho = 1.0_C_DOUBLE; pres = 1.0_C_DOUBLE
    ELSE
      rho = 0.125_C_DOUBLE; pres = 0.1_C_DOUBLE
    END IF
    cv(1,icg) = rho
    cv(2:4,icg) = 0.0_C_DOUBLE
    cv(5,icg) = pres/(gp - 1.0_C_DOUBLE)
  END DO
  CALL rflu_gpu_check(rflu_gpu_upload_state(handle, cv, C_NULL_PTR, 0.0_C_DOUBLE), 'upload')

! ------------------------------------------------------------------------------
! Time loop: RFLU_TimeStepping + RungeKuttaMP with the GPU callees
! ------------------------------------------------------------------------------
  nStages = 3
  time = 0.0_C_DOUBLE
  DO istep = 1,nsteps
    CALL rflu_gpu_check(rflu_gpu_time_step(handle, time, dt), 'rflu_gpu_time_step')
    DO istage = 1,nStages
      CALL rflu_gpu_check(rflu_gpu_rk_init(handle, istage), 'RKInitMP')
      CALL rflu_gpu_check(rflu_gpu_cell_gradients(handle), 'CellGradientsMP')
      CALL rflu_gpu_check(rflu_gpu_convective_fluxes(handle, istage), 'ConvectiveFluxesMP')
      CALL rflu_gpu_check(rflu_gpu_rk_update(handle, istage, dt), 'RKUpdateMP')
      CALL rflu_gpu_check(rflu_gpu_set_dependent_vars(handle), 'RFLU_SetVarsContWrapper')
      CALL rflu_gpu_check(rflu_gpu_after_update(handle, istage), 'AfterUpdateMP')
    END DO
    time = time + dt
  END DO
! the after-update checks of the last step (earlier ones surface at rflu_gpu_time_step)
  CALL rflu_gpu_check(rflu_gpu_check_status(handle), 'AfterUpdateMP')

  CALL rflu_gpu_check(rflu_gpu_write_dump(handle, 0, TRIM(outfile)//'.dump'//C_NULL_CHAR, &
                                          INT(nsteps, C_INT64_T), time, 0), 'write_dump')
  CALL rflu_gpu_check(rflu_gpu_download_state(handle, cv, dv), 'download')
  CALL rflu_gpu_check(rflu_gpu_destroy_region(handle), 'destroy')
  OPEN(NEWUNIT=u, FILE=TRIM(outfile), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) cv
  CLOSE(u)
  WRITE(*,'(A,I0,A,I0,A,ES23.16,A,ES23.16)') 'rflu_driver: ', nCells, ' cells, ', nsteps, &
        ' steps, time ', time, ', sum(rho) ', SUM(cv(1,:))

CONTAINS

  SUBROUTINE get_int(name, arr)
    CHARACTER(*), INTENT(IN) :: name
    INTEGER(C_INT32_T), POINTER, INTENT(OUT) :: arr(:)
    IF ( rflu_mesh_array(mesh, TRIM(name)//C_NULL_CHAR, p, cnt, dtype) /= 0 ) STOP 4
    CALL C_F_POINTER(p, arr, [cnt])
  END SUBROUTINE get_int

  SUBROUTINE get_dbl(name, arr)
    CHARACTER(*), INTENT(IN) :: name
    REAL(C_DOUBLE), POINTER, INTENT(OUT) :: arr(:)
    IF ( rflu_mesh_array(mesh, TRIM(name)//C_NULL_CHAR, p, cnt, dtype) /= 0 ) STOP 5
    CALL C_F_POINTER(p, arr, [cnt])
  END SUBROUTINE get_dbl

END PROGRAM rflu_driver
