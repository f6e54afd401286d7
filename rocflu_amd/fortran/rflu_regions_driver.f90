! ------------------------------------------------------------------------------
! rflu_regions_driver: a Fortran host of several regions of one problem on one
! device, coupled through virtual cells -- RungeKuttaMP's region loop
! (libfloflu/RungeKuttaMP.F90:148-508) with the MPI halo calls replaced by
! rflu_gpu_exchange.  Each region is read from its own file, as the reference reads
! one grid file per region, with 1-based cell numbers:
!
!   int32 nCells, nVirtual, nFaces, nBFaces, nPatches, nStencil, nBorders, nSend, nRecv
!   int32 f2c(2,nFaces); double fn(4,nFaces), fc(3,nFaces), cofg(3,nTot), vol(nTot)
!   int32 bpatch_ptr(nPatches+1) (0-based offsets), bf2c(nBFaces)
!   double bfn(4,nBFaces), bfc(3,nBFaces)
!   int32 patch_bc(nPatches), patch_opt(nPatches); double patch_pout(nPatches)
!   int32 patch_order(nPatches), c2cs_ptr(nTot+1) (0-based offsets), c2cs_idx(nStencil)
!   double xyz_moms(10,nTot), patch_vals(5,nPatches), cv(5,nTot)
!   int32 nbr_region(nBorders), nbr_border(nBorders), send_ptr(nBorders+1),
!         send(nSend), recv_ptr(nBorders+1), recv(nRecv)
!
! (nTot = nCells + nVirtual).  Scheme: HLLC order 2 + Venkatakrishnan, SSP-RK3.
! Usage: rflu_regions_driver nRegions nSteps in_prefix out_prefix; writes each
! region's actual-cell cv(5,nCells) to out_prefix<iReg>.bin.
! ------------------------------------------------------------------------------
PROGRAM rflu_regions_driver

  USE ISO_C_BINDING
  USE rflu_gpu
  IMPLICIT NONE

  TYPE t_reg
    INTEGER(C_INT32_T) :: hd(9), handle
    INTEGER(C_INT32_T), ALLOCATABLE :: f2c(:,:), bptr(:), bf2c(:), pbc(:), popt(:), porder(:), &
                                       sptr(:), sidx(:), nbr(:), nbrb(:), sendp(:), send(:), &
                                       recvp(:), recv(:), nbrh(:)
    REAL(C_DOUBLE), ALLOCATABLE :: fn(:,:), fc(:,:), cofg(:,:), vol(:), bfn(:,:), bfc(:,:), &
                                   pout(:), moms(:,:), pvals(:,:), cv(:,:), dv(:,:)
  END TYPE t_reg

  TYPE(t_reg), ALLOCATABLE, TARGET :: reg(:)
  TYPE(rflu_region_desc) :: desc
  INTEGER(C_INT32_T), ALLOCATABLE :: handles(:)
  CHARACTER(LEN=512) :: arg, inpre, outpre, fname
  INTEGER :: nReg, nsteps, iReg, istep, istage, u, nTot, ib
  REAL(C_DOUBLE) :: time, dt, dtReg

  CALL GET_COMMAND_ARGUMENT(1, arg); READ(arg,*) nReg
  CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg,*) nsteps
  CALL GET_COMMAND_ARGUMENT(3, inpre)
  CALL GET_COMMAND_ARGUMENT(4, outpre)
  ALLOCATE(reg(nReg), handles(nReg))

! ------------------------------------------------------------------------------
! Read and create the regions
! ------------------------------------------------------------------------------
  DO iReg = 1,nReg
    WRITE(fname,'(A,I0,A)') TRIM(inpre), iReg - 1, '.bin'
    OPEN(NEWUNIT=u, FILE=TRIM(fname), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
    ASSOCIATE ( r => reg(iReg) )
      READ(u) r%hd
      nTot = r%hd(1) + r%hd(2)
      ALLOCATE(r%f2c(2,r%hd(3)), r%fn(4,r%hd(3)), r%fc(3,r%hd(3)), r%cofg(3,nTot), r%vol(nTot))
      READ(u) r%f2c, r%fn, r%fc, r%cofg, r%vol
      ALLOCATE(r%bptr(r%hd(5)+1), r%bf2c(r%hd(4)), r%bfn(4,r%hd(4)), r%bfc(3,r%hd(4)))
      READ(u) r%bptr, r%bf2c, r%bfn, r%bfc
      ALLOCATE(r%pbc(r%hd(5)), r%popt(r%hd(5)), r%pout(r%hd(5)), r%porder(r%hd(5)))
      READ(u) r%pbc, r%popt, r%pout, r%porder
      ALLOCATE(r%sptr(nTot+1), r%sidx(r%hd(6)), r%moms(10,nTot), r%pvals(5,r%hd(5)))
      READ(u) r%sptr, r%sidx, r%moms, r%pvals
      ALLOCATE(r%cv(5,nTot), r%dv(3,nTot))
      READ(u) r%cv
      ALLOCATE(r%nbr(r%hd(7)), r%nbrb(r%hd(7)), r%sendp(r%hd(7)+1), r%send(r%hd(8)), &
               r%recvp(r%hd(7)+1), r%recv(r%hd(9)), r%nbrh(r%hd(7)))
      READ(u) r%nbr, r%nbrb, r%sendp, r%send, r%recvp, r%recv
    END ASSOCIATE
    CLOSE(u)

    desc%index_base = 1
    desc%n_cells = reg(iReg)%hd(1); desc%n_virtual = reg(iReg)%hd(2); desc%pad2_ = 0
    desc%n_faces = reg(iReg)%hd(3); desc%n_patches = reg(iReg)%hd(5)
    desc%f2c = C_LOC(reg(iReg)%f2c); desc%fn = C_LOC(reg(iReg)%fn); desc%fc = C_LOC(reg(iReg)%fc)
    desc%cofg = C_LOC(reg(iReg)%cofg); desc%vol = C_LOC(reg(iReg)%vol)
    desc%bpatch_ptr = C_LOC(reg(iReg)%bptr); desc%bf2c = C_LOC(reg(iReg)%bf2c)
    desc%bfn = C_LOC(reg(iReg)%bfn); desc%bfc = C_LOC(reg(iReg)%bfc)
    desc%patch_bc = C_LOC(reg(iReg)%pbc); desc%patch_opt = C_LOC(reg(iReg)%popt)
    desc%patch_pout = C_LOC(reg(iReg)%pout); desc%patch_order = C_LOC(reg(iReg)%porder)
    desc%c2cs_ptr = C_LOC(reg(iReg)%sptr); desc%c2cs_idx = C_LOC(reg(iReg)%sidx)
    desc%xyz_moms = C_LOC(reg(iReg)%moms)
    desc%gamma = 1.4_C_DOUBLE; desc%cp = 1004.5_C_DOUBLE
    desc%discr = 3; desc%order = 2; desc%reconst = 11; desc%rk = 4; desc%precision = 0
    desc%cfl = 0.5_C_DOUBLE; desc%dt_imposed = 1.0E-3_C_DOUBLE; desc%dt_min_limit = 0.0_C_DOUBLE
    desc%eps_entr = 0.05_C_DOUBLE; desc%diss_fact = 1.0_C_DOUBLE
    desc%patch_vals = C_LOC(reg(iReg)%pvals)
    desc%c2cs_bptr = C_NULL_PTR; desc%c2cs_bface = C_NULL_PTR
    desc%creconst_weight = 1.0_C_DOUBLE; desc%creconst = 0; desc%pad_ = 0
    desc%patch_opt2 = C_NULL_PTR
    CALL rflu_gpu_check(rflu_gpu_create_region(desc, reg(iReg)%handle), 'rflu_gpu_create_region')
    handles(iReg) = reg(iReg)%handle
  END DO

! ------------------------------------------------------------------------------
! Borders (t_border): neighbour handles from the neighbours' region numbers
! ------------------------------------------------------------------------------
  DO iReg = 1,nReg
    ASSOCIATE ( r => reg(iReg) )
      DO ib = 1,r%hd(7)
        r%nbrh(ib) = handles(r%nbr(ib))
      END DO
      CALL rflu_gpu_check(rflu_gpu_set_borders(r%handle, r%hd(7), r%nbrh, r%nbrb, r%sendp, &
                                               r%send, r%recvp, r%recv), 'rflu_gpu_set_borders')
      CALL rflu_gpu_check(rflu_gpu_upload_state(r%handle, r%cv, C_NULL_PTR, 0.0_C_DOUBLE), 'upload')
    END ASSOCIATE
  END DO

! ------------------------------------------------------------------------------
! Time loop: RFLU_TimeStepping (dt = MIN over the regions) + RungeKuttaMP
! ------------------------------------------------------------------------------
  time = 0.0_C_DOUBLE
  DO istep = 1,nsteps
    dt = HUGE(1.0_C_DOUBLE)
    DO iReg = 1,nReg
      CALL rflu_gpu_check(rflu_gpu_time_step(handles(iReg), time, dtReg), 'rflu_gpu_time_step')
      dt = MIN(dt, dtReg)
    END DO
    DO istage = 1,3
      DO iReg = 1,nReg
        CALL rflu_gpu_check(rflu_gpu_rk_init(handles(iReg), istage), 'RKInitMP')
        CALL rflu_gpu_check(rflu_gpu_cell_gradients(handles(iReg)), 'CellGradientsMP')
      END DO
      CALL rflu_gpu_check(rflu_gpu_exchange(nReg, handles, RFLU_XCHG_GRAD), 'RFLU_MPI_ExchangeGradCells')
      DO iReg = 1,nReg
        CALL rflu_gpu_check(rflu_gpu_convective_fluxes(handles(iReg), istage), 'ConvectiveFluxesMP')
        CALL rflu_gpu_check(rflu_gpu_rk_update(handles(iReg), istage, dt), 'RKUpdateMP')
        CALL rflu_gpu_check(rflu_gpu_set_dependent_vars(handles(iReg)), 'RFLU_SetVarsContWrapper')
      END DO
      DO iReg = 1,nReg
        CALL rflu_gpu_check(rflu_gpu_after_update(handles(iReg), istage), 'AfterUpdateMP')
      END DO
      CALL rflu_gpu_check(rflu_gpu_exchange(nReg, handles, RFLU_XCHG_STATE), 'RFLU_MPI_CopyWrapper')
    END DO
    time = time + dt
  END DO
! the after-update checks of the last step (earlier ones surface at rflu_gpu_time_step)
  DO iReg = 1,nReg
    CALL rflu_gpu_check(rflu_gpu_check_status(handles(iReg)), 'AfterUpdateMP')
  END DO

  DO iReg = 1,nReg
    ASSOCIATE ( r => reg(iReg) )
      CALL rflu_gpu_check(rflu_gpu_download_state(r%handle, r%cv, r%dv), 'download')
      CALL rflu_gpu_check(rflu_gpu_destroy_region(r%handle), 'destroy')
      WRITE(fname,'(A,I0,A)') TRIM(outpre), iReg - 1, '.bin'
      OPEN(NEWUNIT=u, FILE=TRIM(fname), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(u) r%cv(:, 1:r%hd(1))
      CLOSE(u)
    END ASSOCIATE
  END DO
  WRITE(*,'(A,I0,A,I0,A,ES23.16)') 'rflu_regions_driver: ', nReg, ' regions, ', nsteps, &
        ' steps, time ', time

END PROGRAM rflu_regions_driver
