! Writes small fluid dumps with Fortran stream I/O in the statement forms of the
! reference's writer (modflu/RFLU_ModRestartData.F90:136-215: one WRITE of the header
! items, one WRITE per variable of cv(iCv,1:nCells), then the NSCBC count), so the
! tests can check the C codec against the bytes a Fortran compiler produces for them.
! Build and run from tests/golden:  amdflang -O0 make_dump_fixtures.f90 -o /tmp/mkd && /tmp/mkd
PROGRAM make_dump_fixtures
  IMPLICIT NONE
  INTEGER, PARAMETER :: nCells = 3, nCv = 5
  REAL(8) :: cv(nCv,nCells)
  INTEGER :: iCv, ic, iFile

  DO ic = 1,nCells
    DO iCv = 1,nCv
      cv(iCv,ic) = 10.0D0*ic + iCv + 0.125D0
    END DO
  END DO

! version 5, gas only (pure form, no PICL fields, no NSCBC patches)
  OPEN(NEWUNIT=iFile,FILE='dumps/v5_gas.dat',FORM='UNFORMATTED',ACCESS='STREAM', &
       STATUS='REPLACE')
  WRITE(iFile) 'RFLUDUMP',INT(5,4),INT(1234,8),REAL(1.5D-4,8),INT(7,4),INT(nCells,4), &
               INT(nCv,4),INT(1,4),INT(0,4)
  DO iCv = 1,nCv
    WRITE(iFile) REAL(cv(iCv,1:nCells),8)
  END DO
  WRITE(iFile) INT(0,4)
  CLOSE(iFile)

! version 5 with PICL fields: two nonzero cells, then the gradient fields
  OPEN(NEWUNIT=iFile,FILE='dumps/v5_picl.dat',FORM='UNFORMATTED',ACCESS='STREAM', &
       STATUS='REPLACE')
  WRITE(iFile) 'RFLUDUMP',INT(5,4),INT(20,8),REAL(2.0D-5,8),INT(1,4),INT(nCells,4), &
               INT(nCv,4),INT(1,4),INT(1,4)
  DO iCv = 1,nCv
    WRITE(iFile) REAL(cv(iCv,1:nCells),8)
  END DO
  WRITE(iFile) INT(2,4),(/ INT(1,4),INT(3,4) /),(/ 0.25D0,0.5D0 /)
  WRITE(iFile) 1_4
  DO ic = 1,3
    WRITE(iFile) REAL((/ 1.0D0,2.0D0,3.0D0 /)*ic,8)
  END DO
  WRITE(iFile) INT(0,4)
  CLOSE(iFile)

! version 4 with PICL fields: piclVF of every cell, then three gradient fields
  OPEN(NEWUNIT=iFile,FILE='dumps/v4_picl.dat',FORM='UNFORMATTED',ACCESS='STREAM', &
       STATUS='REPLACE')
  WRITE(iFile) 'RFLUDUMP',INT(4,4),INT(5,8),REAL(1.0D-6,8),INT(2,4),INT(nCells,4), &
               INT(nCv,4),INT(1,4),INT(1,4)
  DO iCv = 1,nCv
    WRITE(iFile) REAL(cv(iCv,1:nCells),8)
  END DO
  WRITE(iFile) REAL((/ 0.1D0,0.0D0,0.3D0 /),8)
  DO ic = 1,3
    WRITE(iFile) REAL((/ 4.0D0,5.0D0,6.0D0 /)*ic,8)
  END DO
  WRITE(iFile) INT(0,4)
  CLOSE(iFile)

! version 5 with one NSCBC patch (face states): a gas-only region has none to match
  OPEN(NEWUNIT=iFile,FILE='dumps/v5_nscbc.dat',FORM='UNFORMATTED',ACCESS='STREAM', &
       STATUS='REPLACE')
  WRITE(iFile) 'RFLUDUMP',INT(5,4),INT(3,8),REAL(3.0D-6,8),INT(0,4),INT(nCells,4), &
               INT(nCv,4),INT(1,4),INT(0,4)
  DO iCv = 1,nCv
    WRITE(iFile) REAL(cv(iCv,1:nCells),8)
  END DO
  WRITE(iFile) INT(1,4)
  WRITE(iFile) INT(4,4),INT(2,4)
  DO iCv = 1,nCv
    WRITE(iFile) REAL((/ 1.0D0,2.0D0 /)*iCv,8)
  END DO
  CLOSE(iFile)

! volume-weighted form (a particle run's state after step 0)
  OPEN(NEWUNIT=iFile,FILE='dumps/v5_weighted.dat',FORM='UNFORMATTED',ACCESS='STREAM', &
       STATUS='REPLACE')
  WRITE(iFile) 'RFLUDUMP',INT(5,4),INT(9,8),REAL(9.0D-6,8),INT(0,4),INT(nCells,4), &
               INT(nCv,4),INT(2,4),INT(0,4)
  DO iCv = 1,nCv
    WRITE(iFile) REAL(cv(iCv,1:nCells),8)
  END DO
  WRITE(iFile) INT(0,4)
  CLOSE(iFile)
END PROGRAM make_dump_fixtures
