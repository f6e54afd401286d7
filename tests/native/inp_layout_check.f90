! Reads a .inp and a .bc file through the Fortran BIND(C) types of rflu_gpu.f90 and
! prints every field, one "name value" per line, for comparison with the C layout.
PROGRAM inp_layout_check
  USE ISO_C_BINDING
  USE rflu_gpu
  IMPLICIT NONE
  CHARACTER(LEN=512) :: inpfile, bcfile, arg
  INTEGER :: npatch, ip, k
  INTEGER(C_INT) :: rc
  TYPE(rflu_inp) :: inp
  TYPE(rflu_bc_patch), ALLOCATABLE :: bcs(:)
  CALL GET_COMMAND_ARGUMENT(1, inpfile)
  CALL GET_COMMAND_ARGUMENT(2, bcfile)
  CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg,*) npatch
  rc = rflu_read_inp(TRIM(inpfile)//C_NULL_CHAR, inp)
  CALL rflu_host_check(rc, 'inp', rflu_input_last_error())
  ALLOCATE(bcs(npatch))
  rc = rflu_read_bc(TRIM(bcfile)//C_NULL_CHAR, INT(npatch, C_INT), INT(inp%order_bfaces, C_INT), bcs)
  CALL rflu_host_check(rc, 'bc', rflu_input_last_error())
  WRITE(*,'(A,I0)') 'grid_source ', inp%grid_source
  WRITE(*,'(A,I0)') 'gas_model ', inp%gas_model
  WRITE(*,'(A,ES25.17)') 'cfl ', inp%cfl
  WRITE(*,'(A,I0)') 'discr ', inp%discr
  WRITE(*,'(A,I0)') 'order ', inp%order
  WRITE(*,'(A,ES25.17)') 'diss_fact ', inp%diss_fact
  WRITE(*,'(A,I0)') 'reconst ', inp%reconst
  WRITE(*,'(A,I0)') 'creconst_cells ', inp%creconst_cells
  WRITE(*,'(A,ES25.17)') 'toler_ict ', inp%toler_ict
  WRITE(*,'(A,I0)') 'order_bfaces ', inp%order_bfaces
  WRITE(*,'(A,ES25.17)') 'minf ', inp%minf
  WRITE(*,'(A,I0)') 'rk_scheme ', inp%rk_scheme
  WRITE(*,'(A,I0)') 'print_iter ', inp%print_iter
  WRITE(*,'(A,ES25.17)') 'dt_imposed ', inp%dt_imposed
  WRITE(*,'(A,ES25.17)') 'max_time ', inp%max_time
  WRITE(*,'(A,ES25.17)') 'res_tol ', inp%res_tol
  WRITE(*,'(A,ES25.17)') 'ref_cp ', inp%ref_cp
  WRITE(*,'(A,ES25.17)') 'ref_length ', inp%ref_length
  WRITE(*,'(A,I0)') 'init_flag ', inp%init_flag
  WRITE(*,'(A,ES25.17)') 'init_temp ', inp%init_temp
  WRITE(*,'(A,I0)') 'init_ival19 ', inp%init_ival(19)
  DO k = 1,25
    WRITE(*,'(A,I0,A,ES25.17)') 'init_rval', k, ' ', inp%init_rval(k)
  END DO
  WRITE(*,'(A,I0)') 'sections ', inp%sections
  DO ip = 1,npatch
    WRITE(*,'(A,I0,A,7(1X,I0))') 'patch', ip, ' ', bcs(ip)%bc_type, bcs(ip)%opt, bcs(ip)%opt2, &
      bcs(ip)%order, bcs(ip)%creconst, bcs(ip)%related, bcs(ip)%axis
    WRITE(*,'(A,I0,A,7(1X,ES25.17))') 'patchval', ip, ' ', bcs(ip)%angle, bcs(ip)%vals, bcs(ip)%pout
    WRITE(*,'(A,I0,A,80A)') 'patchname', ip, ' ', (bcs(ip)%name(k), k=1,80)
  END DO
END PROGRAM inp_layout_check
