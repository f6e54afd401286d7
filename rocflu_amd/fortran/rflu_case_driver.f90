!******************************************************************************
!
! rflu_case_driver -- a Fortran host that runs a case from the reference's own input
! files: <dir>/<casename>.inp (NUMERICS, TIMESTEP, REFERENCE, INITFLOW sections),
! <dir>/<casename>.bc (patch boundary conditions) and the CENTAUR ASCII grid
! <dir>/<casename>.hyb.asc, through the readers of include/rflu_input.h and
! include/rflu_mesh.h and the per-callee GPU path of include/rocflu_amd.h, in the
! order of the reference's RFLU_TimeStepping / RungeKuttaMP.
!
! Usage: rflu_case_driver dir casename maxsteps out.bin
!   Runs until MAXTIME (the last step shortened to end there,
!   rocflu/RFLU_TimeStepping.F90:442-447) or maxsteps steps when maxsteps > 0, and
!   writes cv(5,nCells) as raw doubles to out.bin and the fluid dump to out.bin.dump.
!   The initial state is rfluinit's: INITFLOW FLAG 1 (uniform, DENS or TEMP) or
!   FLAG 3 for casename rectshktb (utilities/init/RFLU_InitFlowHardCode.F90:839-901).
!   Symmetry and periodic patches become virtual cells (rflu_mesh_build_sype) that
!   take their actual cells' state and gradients through borders of the region with
!   itself, exchanged where RungeKuttaMP exchanges them.
!
!******************************************************************************
PROGRAM rflu_case_driver

  USE ISO_C_BINDING
  USE rflu_gpu
  IMPLICIT NONE

  CHARACTER(LEN=512) :: dir, casename, arg, outfile, base
  INTEGER :: maxsteps, istep, istage, nStages, icg, ip, u, nPatches, nCells, nFaces, nBFaces, &
             nVirt, nTot, nRegPatches, nXf, ib, k, nb
  LOGICAL :: sype
  INTEGER(C_INT) :: rc
  INTEGER(C_INT32_T) :: handle
  INTEGER(C_INT64_T) :: gsz(8), sizes(9), cnt
  INTEGER(C_INT) :: dtype
  LOGICAL :: finished
  TYPE(C_PTR) :: grid, mesh, p
  TYPE(rflu_inp) :: inp
  TYPE(rflu_bc_patch), ALLOCATABLE :: bcs(:)
  TYPE(rflu_region_desc), TARGET :: desc
  INTEGER(C_INT32_T), POINTER :: ip32(:)
  REAL(C_DOUBLE), POINTER :: dp(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: f2c(:,:), bf2c(:), bptr(:), pbc(:), popt(:), &
                                             popt2(:), porder(:), sptr(:), sidx(:), &
                                             bsptr(:), bsidx(:), prel(:), paxis(:), &
                                             vsrc(:), vxf(:), pmap(:), rbc(:), ropt(:), &
                                             ropt2(:), rorder(:), nbrh(:), nbrb(:), &
                                             sndp(:), sndc(:), rcvp(:), rcvc(:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: fn(:,:), fc(:,:), cofg(:,:), vol(:), bfn(:,:), &
                                         bfc(:,:), pout(:), moms(:,:), pvals(:,:), pang(:), &
                                         xfm(:,:), rpout(:), rvals(:,:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: cv(:,:), dv(:,:)
  REAL(C_DOUBLE) :: time, dt, cp, g, gc, mm, d, uu, vv, ww, pr, x, y, z

  CALL GET_COMMAND_ARGUMENT(1, dir)
  CALL GET_COMMAND_ARGUMENT(2, casename)
  CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg,*) maxsteps
  CALL GET_COMMAND_ARGUMENT(4, outfile)
  base = TRIM(dir)//'/'//TRIM(casename)

! ------------------------------------------------------------------------------
! Input files (ReadInputFile, RFLU_ReadGridWrapper, RFLU_ReadBcInputFileWrapper)
! ------------------------------------------------------------------------------
  rc = rflu_read_inp(TRIM(base)//'.inp'//C_NULL_CHAR, inp)
  CALL rflu_host_check(rc, 'ReadInputFile', rflu_input_last_error())
  IF ( inp%grid_source /= 0 ) STOP 'GRIDSRC: only CENTAUR ASCII grids are read'
  IF ( inp%flow_type /= 1 ) STOP 'FLOWTYPE: the explicit path runs unsteady flows'
  rc = rflu_grid_read_centaur(TRIM(base)//'.hyb.asc'//C_NULL_CHAR, grid)
  CALL rflu_host_check(rc, 'RFLU_ReadGridCENTAURASCII', rflu_mesh_last_error())
  rc = rflu_grid_sizes(grid, gsz)
  nPatches = INT(gsz(6))
  ALLOCATE(bcs(nPatches))
  rc = rflu_read_bc(TRIM(base)//'.bc'//C_NULL_CHAR, INT(nPatches, C_INT), &
                    INT(inp%order_bfaces, C_INT), bcs)
  CALL rflu_host_check(rc, 'RFLU_ReadBcInputFile', rflu_input_last_error())
  ALLOCATE(pbc(nPatches), popt(nPatches), popt2(nPatches), porder(nPatches), &
           pout(nPatches), pvals(5,nPatches), prel(nPatches), paxis(nPatches), pang(nPatches))
  sype = .FALSE.
  DO ip = 1,nPatches
    IF ( bcs(ip)%bc_type == 0 ) THEN
      WRITE(*,'(A,I0,A)') 'patch ', ip, ' has no section in the .bc file'
      STOP 1
    END IF
    pbc(ip) = bcs(ip)%bc_type
    popt(ip) = bcs(ip)%opt
    popt2(ip) = bcs(ip)%opt2
    porder(ip) = bcs(ip)%order
    pout(ip) = bcs(ip)%pout
    pvals(:,ip) = bcs(ip)%vals
    prel(ip) = bcs(ip)%related
    paxis(ip) = bcs(ip)%axis
    pang(ip) = bcs(ip)%angle
    IF ( pbc(ip) == 100 .OR. pbc(ip) == 110 ) sype = .TRUE.
  END DO

! ------------------------------------------------------------------------------
! Region (RFLU_InitFlowSolver): face list, geometry, stencils
! ------------------------------------------------------------------------------
  IF ( sype ) THEN
    ! symmetry and periodic patches become virtual cells (rflupart's RFLU_SYPE_* step)
    rc = rflu_mesh_build_sype(grid, pbc, popt, pout, prel, paxis, pang, &
                              MERGE(1_C_INT, 0_C_INT, inp%order == 2), mesh)
  ELSE
    rc = rflu_mesh_build(grid, pbc, popt, pout, MERGE(1_C_INT, 0_C_INT, inp%order == 2), mesh)
  END IF
  CALL rflu_host_check(rc, 'RFLU_BuildFaceList', rflu_mesh_last_error())
  CALL rflu_grid_destroy(grid)
  IF ( rflu_mesh_sizes(mesh, sizes) /= 0 ) STOP 3
  nTot = INT(sizes(1)); nFaces = INT(sizes(2)); nBFaces = INT(sizes(3))
  nRegPatches = INT(sizes(4))
  nVirt = 0
  IF ( sype ) THEN
    CALL get_int('virt_src', ip32); nVirt = SIZE(ip32)
    ALLOCATE(vsrc(nVirt)); vsrc = ip32 + 1
    CALL get_int('virt_xf', ip32)
    ALLOCATE(vxf(nVirt)); vxf = ip32
    CALL get_dbl('xforms', dp); nXf = SIZE(dp)/9
    ALLOCATE(xfm(9,nXf)); xfm = RESHAPE(dp, (/9,nXf/))
    CALL get_int('patch_map', ip32)
    ALLOCATE(pmap(nRegPatches)); pmap = ip32
  END IF
  nCells = nTot - nVirt
  ! the region's patches: the grid's, or the kept ones and the virtual layers' outer faces
  ALLOCATE(rbc(nRegPatches), ropt(nRegPatches), ropt2(nRegPatches), rorder(nRegPatches), &
           rpout(nRegPatches), rvals(5,nRegPatches))
  CALL get_int('patch_bc', ip32);  rbc = ip32
  CALL get_int('patch_opt', ip32); ropt = ip32
  CALL get_dbl('patch_pout', dp);  rpout = dp
  DO ip = 1,nRegPatches
    k = ip
    IF ( sype ) k = pmap(ip) + 1
    IF ( k >= 1 ) THEN
      ropt2(ip) = popt2(k); rorder(ip) = porder(k); rvals(:,ip) = pvals(:,k)
    ELSE
      ropt2(ip) = 0; rorder(ip) = 1; rvals(:,ip) = 0.0_C_DOUBLE
    END IF
  END DO
  ALLOCATE(f2c(2,nFaces), fn(4,nFaces), fc(3,nFaces), cofg(3,nTot), vol(nTot))
  ALLOCATE(bf2c(nBFaces), bfn(4,nBFaces), bfc(3,nBFaces), bptr(nRegPatches+1))
  CALL get_int('f2c', ip32);   f2c  = RESHAPE(ip32, (/2,nFaces/)) + 1
  CALL get_dbl('fn', dp);      fn   = RESHAPE(dp, (/4,nFaces/))
  CALL get_dbl('fc', dp);      fc   = RESHAPE(dp, (/3,nFaces/))
  CALL get_dbl('cofg', dp);    cofg = RESHAPE(dp, (/3,nTot/))
  CALL get_dbl('vol', dp);     vol  = dp
  CALL get_int('bf2c', ip32);  bf2c = ip32 + 1
  CALL get_dbl('bfn', dp);     bfn  = RESHAPE(dp, (/4,nBFaces/))
  CALL get_dbl('bfc', dp);     bfc  = RESHAPE(dp, (/3,nBFaces/))
  CALL get_int('bpatch_ptr', ip32); bptr = ip32
  IF ( inp%order == 2 ) THEN
    CALL get_int('c2cs_ptr', ip32)
    ALLOCATE(sptr(nTot+1)); sptr = ip32
    CALL get_int('c2cs_idx', ip32)
    ALLOCATE(sidx(SIZE(ip32))); sidx = ip32 + 1
    CALL get_dbl('xyz_moms', dp)
    ALLOCATE(moms(10,nTot)); moms = RESHAPE(dp, (/10,nTot/))
    CALL get_int('c2cs_bptr', ip32)
    ALLOCATE(bsptr(nTot+1)); bsptr = ip32
    CALL get_int('c2cs_bface', ip32)
    ALLOCATE(bsidx(MAX(1,SIZE(ip32)))); bsidx = 0
    IF ( SIZE(ip32) > 0 ) bsidx(1:SIZE(ip32)) = ip32 + 1
  END IF
  CALL rflu_mesh_destroy(mesh)

  cp = inp%ref_cp
  g  = inp%ref_gamma
  desc%index_base = 1
  desc%n_cells = nCells; desc%n_faces = nFaces; desc%n_patches = nRegPatches
  desc%f2c = C_LOC(f2c); desc%fn = C_LOC(fn); desc%fc = C_LOC(fc)
  desc%cofg = C_LOC(cofg); desc%vol = C_LOC(vol)
  desc%bpatch_ptr = C_LOC(bptr); desc%bf2c = C_LOC(bf2c)
  desc%bfn = C_LOC(bfn); desc%bfc = C_LOC(bfc)
  desc%patch_bc = C_LOC(rbc); desc%patch_opt = C_LOC(ropt)
  desc%patch_pout = C_LOC(rpout); desc%patch_order = C_LOC(rorder)
  desc%patch_vals = C_LOC(rvals); desc%patch_opt2 = C_LOC(ropt2)
  IF ( inp%order == 2 ) THEN
    desc%c2cs_ptr = C_LOC(sptr); desc%c2cs_idx = C_LOC(sidx); desc%xyz_moms = C_LOC(moms)
    desc%c2cs_bptr = C_LOC(bsptr); desc%c2cs_bface = C_LOC(bsidx)
  ELSE
    desc%c2cs_ptr = C_NULL_PTR; desc%c2cs_idx = C_NULL_PTR; desc%xyz_moms = C_NULL_PTR
    desc%c2cs_bptr = C_NULL_PTR; desc%c2cs_bface = C_NULL_PTR
  END IF
  desc%gamma = g; desc%cp = cp
  desc%discr = inp%discr; desc%order = inp%order; desc%reconst = inp%reconst
  desc%rk = inp%rk_scheme; desc%precision = 0
  desc%cfl = inp%cfl; desc%dt_imposed = inp%dt_imposed; desc%dt_min_limit = inp%dt_min_limit
  desc%eps_entr = inp%eps_entr; desc%diss_fact = inp%diss_fact
  desc%creconst_weight = inp%creconst_cells_weight
  desc%creconst = inp%creconst_cells; desc%pad_ = 0
  desc%n_virtual = nVirt; desc%pad2_ = 0
  CALL rflu_gpu_check(rflu_gpu_create_region(desc, handle), 'rflu_gpu_create_region')

  ! symmetry/periodic virtual cells: borders of the region with itself, one per
  ! transform, the identity (transform 0) a plain copy
  IF ( nVirt > 0 ) THEN
    nb = 0
    DO k = 0,nXf-1
      IF ( COUNT(vxf == k) > 0 ) nb = nb + 1
    END DO
    ALLOCATE(nbrh(nb), nbrb(nb), sndp(nb+1), rcvp(nb+1), sndc(nVirt), rcvc(nVirt))
    sndp(1) = 0; ib = 0
    DO k = 0,nXf-1
      IF ( COUNT(vxf == k) == 0 ) CYCLE
      ib = ib + 1
      nbrh(ib) = handle; nbrb(ib) = ib
      sndp(ib+1) = sndp(ib)
      DO icg = 1,nVirt
        IF ( vxf(icg) /= k ) CYCLE
        sndp(ib+1) = sndp(ib+1) + 1
        sndc(sndp(ib+1)) = vsrc(icg)
        rcvc(sndp(ib+1)) = nCells + icg
      END DO
    END DO
    rcvp = sndp
    CALL rflu_gpu_check(rflu_gpu_set_borders(handle, nb, nbrh, nbrb, sndp, sndc, rcvp, rcvc), &
                        'RFLU_SYPE borders')
    ib = 0
    DO k = 0,nXf-1
      IF ( COUNT(vxf == k) == 0 ) CYCLE
      ib = ib + 1
      IF ( k > 0 ) CALL rflu_gpu_check(rflu_gpu_set_border_transform(handle, ib, xfm(:,k+1)), &
                                       'RFLU_RELP transform')
    END DO
  END IF

! ------------------------------------------------------------------------------
! Initial state (rfluinit)
! ------------------------------------------------------------------------------
  gc = cp - cp/g
  mm = 8314.3_C_DOUBLE/gc
  gc = 8314.3_C_DOUBLE/mm
  g  = cp/(cp - gc)
  ALLOCATE(cv(5,nTot), dv(3,nTot))
  cv = 0.0_C_DOUBLE
  DO icg = 1,nCells
    x = cofg(1,icg); y = cofg(2,icg); z = cofg(3,icg)
    IF ( inp%init_flag == 1 ) THEN
      uu = inp%init_vel(1); vv = inp%init_vel(2); ww = inp%init_vel(3)
      pr = inp%init_press
      IF ( inp%init_dens > 0.0_C_DOUBLE ) THEN
        d = inp%init_dens
      ELSE
        d = pr/(gc*inp%init_temp)
      END IF
    ELSE IF ( inp%init_flag == 3 .AND. TRIM(casename) == 'rectshktb' ) THEN
      IF ( x < inp%init_rval(1) ) THEN
        d = inp%init_rval(2); uu = inp%init_rval(3); pr = inp%init_rval(4)
      ELSE
        d = inp%init_rval(5); uu = 0.0_C_DOUBLE; pr = inp%init_rval(6)
      END IF
      vv = 0.0_C_DOUBLE; ww = 0.0_C_DOUBLE
      IF ( inp%init_ival(1) == 1 ) THEN
        d = d*(1.0_C_DOUBLE + inp%init_rval(7)*x + inp%init_rval(8)*y + inp%init_rval(9)*z)
      END IF
    ELSE
      STOP 'INITFLOW: FLAG 1, or FLAG 3 for rectshktb'
    END IF
    cv(1,icg) = d
    cv(2,icg) = d*uu
    cv(3,icg) = d*vv
    cv(4,icg) = d*ww
    cv(5,icg) = d*(pr/(d*(g - 1.0_C_DOUBLE)) + 0.5_C_DOUBLE*(uu*uu + vv*vv + ww*ww))
  END DO
  IF ( nVirt > 0 ) cv(:,nCells+1:nTot) = cv(:,vsrc)
  CALL rflu_gpu_check(rflu_gpu_upload_state(handle, cv, C_NULL_PTR, 0.0_C_DOUBLE), 'upload')
  IF ( nVirt > 0 ) CALL rflu_gpu_check(rflu_gpu_exchange(1, (/handle/), RFLU_XCHG_STATE), &
                                       'RFLU_MPI_CopyWrapper')

! ------------------------------------------------------------------------------
! Time loop (RFLU_TimeStepping + RungeKuttaMP with the GPU callees)
! ------------------------------------------------------------------------------
  SELECT CASE ( inp%rk_scheme )
    CASE ( 1 ); nStages = 4
    CASE ( 2 ); nStages = 3
    CASE ( 3 ); nStages = 1
    CASE DEFAULT; STOP 'RKSCHEME: 1, 2 or 3'
  END SELECT
  time = 0.0_C_DOUBLE
  istep = 0
  finished = .FALSE.
  DO WHILE ( .NOT. finished )
    CALL rflu_gpu_check(rflu_gpu_time_step(handle, time, dt), 'RFLU_MinimumTimeStep')
    IF ( time + dt > inp%max_time ) THEN
      dt = inp%max_time - time
      finished = .TRUE.
    END IF
    DO istage = 1,nStages
      CALL rflu_gpu_check(rflu_gpu_rk_init(handle, istage), 'RKInitMP')
      CALL rflu_gpu_check(rflu_gpu_cell_gradients(handle), 'CellGradientsMP')
      IF ( nVirt > 0 .AND. inp%order == 2 ) &
        CALL rflu_gpu_check(rflu_gpu_exchange(1, (/handle/), RFLU_XCHG_GRAD), &
                            'RFLU_MPI_ExchangeGradCells')
      CALL rflu_gpu_check(rflu_gpu_convective_fluxes(handle, istage), 'ConvectiveFluxesMP')
      CALL rflu_gpu_check(rflu_gpu_rk_update(handle, istage, dt), 'RKUpdateMP')
      CALL rflu_gpu_check(rflu_gpu_set_dependent_vars(handle), 'RFLU_SetVarsContWrapper')
      CALL rflu_gpu_check(rflu_gpu_after_update(handle, istage), 'AfterUpdateMP')
      IF ( nVirt > 0 ) CALL rflu_gpu_check(rflu_gpu_exchange(1, (/handle/), RFLU_XCHG_STATE), &
                                           'RFLU_MPI_CopyWrapper')
    END DO
    time = time + dt
    istep = istep + 1
    IF ( maxsteps > 0 .AND. istep >= maxsteps ) finished = .TRUE.
  END DO
  CALL rflu_gpu_check(rflu_gpu_check_status(handle), 'AfterUpdateMP')

  CALL rflu_gpu_check(rflu_gpu_write_dump(handle, 0, TRIM(outfile)//'.dump'//C_NULL_CHAR, &
                                          INT(istep, C_INT64_T), time, 0), 'write_dump')
  CALL rflu_gpu_check(rflu_gpu_download_state(handle, cv, dv), 'download')
  CALL rflu_gpu_check(rflu_gpu_destroy_region(handle), 'destroy')
  OPEN(NEWUNIT=u, FILE=TRIM(outfile), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) cv(:,1:nCells)
  CLOSE(u)
  WRITE(*,'(A,A,A,I0,A,I0,A,ES23.16,A,ES23.16)') 'rflu_case_driver: ', TRIM(casename), ': ', &
        nCells, ' cells, ', istep, ' steps, time ', time, ', sum(rho) ', SUM(cv(1,1:nCells))

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

END PROGRAM rflu_case_driver
