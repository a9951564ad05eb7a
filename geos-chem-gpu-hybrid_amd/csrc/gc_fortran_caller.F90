!------------------------------------------------------------------------------
! gc_fortran_caller.F90 -- the drop-in library called from Fortran through
! ISO_C_BINDING, the way GEOS-Chem's fullchem driver calls it
! (src/geos-chem/GeosCore/fullchem_mod.F90:153-212 declares the same five
! interfaces).  Built with flang -fopenmp; run by the GPU test suite.
!
! Checks, one line each, then OK or FAIL:
!   * gc_gpu_info / gc_gpu_set_device see the device;
!   * gc_gpu_selftest gives the reference's known answer (rc 1, 12 steps,
!     gpulib/src/gc_gpu_batch_api.F90:349-376);
!   * the dummy chemistry of the reference's multi-device driver
!     (gpulib/src/mgpu_test.F90:46-56: C = 1e6 + mod(i,97), RCONST(1:50) =
!     1e-4*(1 + mod(i,13)/13), [0, 1200] s) converges in every cell;
!   * OpenMP threads solving contiguous slices through
!     gc_gpu_integrate_batch_dev (slice base addresses, the full identity
!     Perm, as fullchem_mod.F90:1253-1257 passes them) reproduce the single
!     call bit for bit;
!   * gc_amd_hybrid_plan through its Fortran binding puts the longest prefix
!     within half of the work on the device.
!
! usage: gc_fortran_caller [nCell [threads]]
!------------------------------------------------------------------------------
PROGRAM gc_fortran_caller
  USE ISO_C_BINDING
  IMPLICIT NONE
  INTERFACE
     SUBROUTINE gc_gpu_integrate_batch_dev( dev, nCell, Tstart, Tend, C, RCONST, ATOL,  &
                                            RTOL, ICNTRL, RCNTRL, Perm, ISTAT, RSTAT,   &
                                            IERR, tTotal, tKernel )                     &
                BIND(C, name='gc_gpu_integrate_batch_dev')
       IMPORT :: c_int, c_double
       INTEGER(c_int), VALUE         :: dev, nCell
       REAL(c_double), VALUE         :: Tstart, Tend
       REAL(c_double), INTENT(INOUT) :: C(*)
       REAL(c_double), INTENT(IN)    :: RCONST(*), ATOL(*), RTOL(*), RCNTRL(*)
       INTEGER(c_int), INTENT(IN)    :: ICNTRL(*), Perm(*)
       INTEGER(c_int), INTENT(OUT)   :: ISTAT(*), IERR(*)
       REAL(c_double), INTENT(OUT)   :: RSTAT(*), tTotal, tKernel
     END SUBROUTINE gc_gpu_integrate_batch_dev
     SUBROUTINE gc_gpu_integrate_batch( nCell, Tstart, Tend, C, RCONST, ATOL, RTOL,     &
                                        ICNTRL, RCNTRL, Perm, ISTAT, RSTAT, IERR,       &
                                        tTotal, tKernel )                               &
                BIND(C, name='gc_gpu_integrate_batch')
       IMPORT :: c_int, c_double
       INTEGER(c_int), VALUE         :: nCell
       REAL(c_double), VALUE         :: Tstart, Tend
       REAL(c_double), INTENT(INOUT) :: C(*)
       REAL(c_double), INTENT(IN)    :: RCONST(*), ATOL(*), RTOL(*), RCNTRL(*)
       INTEGER(c_int), INTENT(IN)    :: ICNTRL(*), Perm(*)
       INTEGER(c_int), INTENT(OUT)   :: ISTAT(*), IERR(*)
       REAL(c_double), INTENT(OUT)   :: RSTAT(*), tTotal, tKernel
     END SUBROUTINE gc_gpu_integrate_batch
     SUBROUTINE gc_gpu_info( nDev, devName, nameLen ) BIND(C, name='gc_gpu_info')
       IMPORT :: c_int, c_char
       INTEGER(c_int),         INTENT(OUT) :: nDev
       INTEGER(c_int),         VALUE       :: nameLen
       CHARACTER(KIND=c_char), INTENT(OUT) :: devName(*)
     END SUBROUTINE gc_gpu_info
     INTEGER(c_int) FUNCTION gc_gpu_set_device( d ) BIND(C, name='gc_gpu_set_device')
       IMPORT :: c_int
       INTEGER(c_int), VALUE :: d
     END FUNCTION gc_gpu_set_device
     SUBROUTINE gc_gpu_selftest( rc, nsteps, tTotal ) BIND(C, name='gc_gpu_selftest')
       IMPORT :: c_int, c_double
       INTEGER(c_int), INTENT(OUT) :: rc, nsteps
       REAL(c_double), INTENT(OUT) :: tTotal
     END SUBROUTINE gc_gpu_selftest
     SUBROUTINE gc_amd_mech_info( nvar, nfix, nreact, nnz ) BIND(C, name='gc_amd_mech_info')
       IMPORT :: c_int
       INTEGER(c_int), INTENT(OUT) :: nvar, nfix, nreact, nnz
     END SUBROUTINE gc_amd_mech_info
     INTEGER(c_int) FUNCTION gc_amd_hybrid_plan( nCell, slotNsteps, fixedFrac, workFrac, &
                                rateGpu, rateCpu, nDev, devRate, work, nGpu, nDevUse,    &
                                devLo, devHi ) BIND(C, name='gc_amd_hybrid_plan')
       IMPORT :: c_int, c_double, c_long_long
       INTEGER(c_int),       VALUE         :: nCell, fixedFrac, nDev
       INTEGER(c_int),       INTENT(IN)    :: slotNsteps(*)
       REAL(c_double),       INTENT(INOUT) :: workFrac
       REAL(c_double),       VALUE         :: rateGpu, rateCpu
       REAL(c_double),       INTENT(IN)    :: devRate(*)
       INTEGER(c_long_long), INTENT(OUT)   :: work(0:*)
       INTEGER(c_int),       INTENT(OUT)   :: nGpu, nDevUse, devLo(*), devHi(*)
     END FUNCTION gc_amd_hybrid_plan
  END INTERFACE

  INTEGER(c_int) :: nvar, nfix, nreact, nnz, nspec, nDev, rc, nsteps, nGpu, nDevUse
  INTEGER(c_int) :: nCell, nThr, i, t, lo, hi, n, nFail, nDiff, devLo(1), devHi(1)
  CHARACTER(KIND=c_char) :: devName(128)
  CHARACTER(LEN=32) :: arg
  REAL(c_double), ALLOCATABLE :: C0(:,:), C(:,:), C1(:,:), RCONST(:,:), RCNTRL(:,:)
  REAL(c_double), ALLOCATABLE :: RSTAT(:,:), RSTAT1(:,:), ATOL(:), RTOL(:)
  INTEGER(c_int), ALLOCATABLE :: ICNTRL(:,:), ISTAT(:,:), ISTAT1(:,:), IERR(:), IERR1(:)
  INTEGER(c_int), ALLOCATABLE :: Perm(:)
  INTEGER(c_long_long), ALLOCATABLE :: work(:)
  INTEGER(c_long_long) :: wsum, wtot
  REAL(c_double) :: tT, tK, frac, devRate(1)
  INTEGER :: fails, brute

  nCell = 4096
  nThr = 4
  CALL GET_COMMAND_ARGUMENT( 1, arg ); IF ( LEN_TRIM(arg) > 0 ) READ(arg,*) nCell
  CALL GET_COMMAND_ARGUMENT( 2, arg ); IF ( LEN_TRIM(arg) > 0 ) READ(arg,*) nThr
  fails = 0

  CALL gc_amd_mech_info( nvar, nfix, nreact, nnz )
  nspec = nvar + nfix
  WRITE(*,'(a,4(1x,i0))') 'mechanism: nvar nfix nreact nnz', nvar, nfix, nreact, nnz
  devName = c_null_char
  CALL gc_gpu_info( nDev, devName, 128_c_int )
  WRITE(*,'(a,i0,a)') 'info: ', nDev, ' device(s)'
  IF ( nDev < 1 .or. gc_gpu_set_device( 0_c_int ) /= 0 ) THEN
     WRITE(*,'(a)') 'FAIL: no usable device'
     STOP 1
  ENDIF

  CALL gc_gpu_selftest( rc, nsteps, tT )
  WRITE(*,'(a,i0,a,i0,a)') 'selftest: rc ', rc, ', ', nsteps, ' steps'
  IF ( rc /= 1 .or. nsteps /= 12 ) fails = fails + 1

  ALLOCATE( C0(nspec,nCell), C(nspec,nCell), C1(nspec,nCell), RCONST(nreact,nCell),     &
            RCNTRL(20,nCell), RSTAT(3,nCell), RSTAT1(3,nCell), ATOL(nvar), RTOL(nvar),   &
            ICNTRL(20,nCell), ISTAT(8,nCell), ISTAT1(8,nCell), IERR(nCell), IERR1(nCell), &
            Perm(nCell), work(0:nCell) )
  ATOL = 1.0d-2
  RTOL = 0.5d-2
  ICNTRL = 0
  ICNTRL(1,:) = 1
  ICNTRL(3,:) = 4
  ICNTRL(7,:) = 1
  ICNTRL(15,:) = -1
  RCNTRL = 0.0d0
  RCONST = 0.0d0
  DO i = 1, nCell
     C0(:,i) = 1.0d6 + REAL( MOD(i,97), c_double )
     RCONST(1:MIN(50,nreact),i) = 1.0d-4 * ( 1.0d0 + REAL( MOD(i,13), c_double ) / 13.0d0 )
     Perm(i) = i
  ENDDO

  ! the whole batch in one call on the current device
  C1 = C0
  CALL gc_gpu_integrate_batch( nCell, 0.0d0, 1200.0d0, C1, RCONST, ATOL, RTOL, ICNTRL,     &
                               RCNTRL, Perm, ISTAT1, RSTAT1, IERR1, tT, tK )
  nFail = COUNT( IERR1 /= 1 )
  WRITE(*,'(a,i0,a,i0,a,f8.3,a)') 'single call: ', nCell, ' cells, ', nFail,                &
       ' failed, ', tT, ' s'
  IF ( nFail /= 0 ) fails = fails + 1

  ! contiguous slices from OpenMP threads, slice base addresses, full Perm
  C = C0
  IERR = 0
  !$OMP PARALLEL DO NUM_THREADS(nThr) PRIVATE(t, lo, hi, n, tT, tK) SCHEDULE(STATIC, 1)
  DO t = 0, nThr - 1
     lo = 1 + INT( (INT(nCell, 8) * t) / nThr )
     hi = INT( (INT(nCell, 8) * (t + 1)) / nThr )
     n = hi - lo + 1
     IF ( n > 0 ) THEN
        CALL gc_gpu_integrate_batch_dev( 0_c_int, n, 0.0d0, 1200.0d0, C(1,lo), RCONST(1,lo), &
                                         ATOL, RTOL, ICNTRL(1,lo), RCNTRL(1,lo), Perm,     &
                                         ISTAT(1,lo), RSTAT(1,lo), IERR(lo), tT, tK )
     ENDIF
  ENDDO
  !$OMP END PARALLEL DO
  nDiff = 0
  DO i = 1, nCell
     IF ( ANY( C(1:nvar,i) /= C1(1:nvar,i) ) .or. ANY( ISTAT(:,i) /= ISTAT1(:,i) ) .or.     &
          ANY( RSTAT(:,i) /= RSTAT1(:,i) ) .or. IERR(i) /= IERR1(i) ) nDiff = nDiff + 1
  ENDDO
  WRITE(*,'(a,i0,a,i0,a)') 'slices: ', nThr, ' OpenMP threads, ', nDiff,                    &
       ' cells differ from the single call'
  IF ( nDiff /= 0 ) fails = fails + 1

  ! the split of the batch by its step counts: half of the work to the device
  frac = 0.5d0
  devRate = -1.0d0
  rc = gc_amd_hybrid_plan( nCell, ISTAT1(3,:), 0_c_int, frac, -1.0d0, -1.0d0, 1_c_int,    &
                           devRate, work, nGpu, nDevUse, devLo, devHi )
  wtot = SUM( INT( MAX( ISTAT1(3,:), 1 ), c_long_long ) )
  wsum = 0
  brute = 0
  DO i = 1, nCell
     wsum = wsum + MAX( ISTAT1(3,i), 1 )
     IF ( REAL(wsum, c_double) <= 0.5d0 * REAL(wtot, c_double) ) brute = i
  ENDDO
  WRITE(*,'(a,i0,a,i0,a,i0,a,i0,a,i0)') 'hybrid plan: rc ', rc, ', nGpu ', nGpu,            &
       ' (longest prefix within half: ', brute, '), device slice ', devLo(1), '-', devHi(1)
  IF ( rc /= 0 .or. nGpu /= brute .or. nDevUse /= 1 .or. devLo(1) /= 1 .or.               &
       devHi(1) /= nGpu .or. work(nCell) /= wsum ) fails = fails + 1

  IF ( fails == 0 ) THEN
     WRITE(*,'(a)') 'OK'
  ELSE
     WRITE(*,'(a)') 'FAIL'
     STOP 1
  ENDIF
END PROGRAM gc_fortran_caller
