MODULE operators
  !---------------------------------------------------------------------------!
  ! The time step seen from the host: advance(itime) keeps the reference's
  ! interface (src/operators.f90:62-188) and output files, and hands the
  ! step itself to the GPU library.
  !---------------------------------------------------------------------------!
  USE ISO_C_BINDING
  USE parameters
  USE grid
  USE ibfs_c
  USE variables
  IMPLICIT NONE

  ! force.dat, circ.dat and slip.dat stay open for the whole run on units of
  ! their own (the reference reopens them every step on units 100 and 106,
  ! which it also uses for the restart and Cholesky files)
  INTEGER, PARAMETER :: uforce = 201, ucirc = 202, uslip = 206
  LOGICAL :: files_open = .FALSE.
  LOGICAL :: have_factor = .FALSE.
  REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE :: motion, hist

CONTAINS

  SUBROUTINE preprocess
    !-------------------------------------------------------------------------!
    ! src/operators.f90:29-59: reuse input/ibfs.chd if it is this case's
    ! factor (chd_is_current: record size and the key written beside it),
    ! else build the force-system matrix and its Cholesky factor (on the
    ! GPU) and cache it with its key.
    !-------------------------------------------------------------------------!
    LOGICAL :: readchol

    IF (.NOT. stationary) THEN
       ! the library solves each step's force system with cjgr
       ! (src/operators.f90:169-171); there is no factor to build or cache
       have_factor = .TRUE.
       RETURN
    END IF
    INQUIRE(file='input/ibfs.chd',exist=readchol)
    IF (readchol) readchol = chd_is_current()
    IF (readchol) THEN
       CALL read_cholesky
       CALL ibfs_check(ibfs_set_cholesky(handle, cholmat, cholvec), 'ibfs_set_cholesky')
    ELSE
       WRITE(*,*) 'Precomputing body matrix on the GPU...'
       WRITE(111,*) 'Precomputing body matrix...'
       CALL ibfs_check(ibfs_build_cholesky(handle, cholmat, cholvec), 'ibfs_build_cholesky')
       CALL write_cholesky
    END IF
    have_factor = .TRUE.
  END SUBROUTINE preprocess

  SUBROUTINE open_outputs(first_it)
    INTEGER :: first_it
    CALL open_one(uforce, 'output/force.dat', first_it)
    CALL open_one(ucirc, 'output/circ.dat', first_it)
    CALL open_one(uslip, 'output/slip.dat', first_it)
    files_open = .TRUE.
  END SUBROUTINE open_outputs

  SUBROUTINE open_one(u, fname, first_it)
    ! same status choice as the per-step open of the reference
    ! (src/operators.f90:566-576): replace at it==1, else append or create
    INTEGER :: u, first_it
    CHARACTER(*) :: fname
    LOGICAL :: ex
    INQUIRE(file=fname, exist=ex)
    IF (first_it == 1) THEN
       OPEN(unit=u,file=fname,form="formatted",status="replace")
    ELSE IF (ex) THEN
       OPEN(unit=u,file=fname,form="formatted",status="old",position="append")
    ELSE
       OPEN(unit=u,file=fname,form="formatted",status="new")
    END IF
  END SUBROUTINE open_one

  SUBROUTINE close_outputs
    IF (files_open) THEN
       CLOSE(uforce); CLOSE(ucirc); CLOSE(uslip)
       files_open = .FALSE.
    END IF
  END SUBROUTINE close_outputs

  SUBROUTINE advance(itime)
    !-------------------------------------------------------------------------!
    ! Advances by one step, or by up to nblock steps in one library call when
    ! requested (never past istop or a restart save).  Writes the lines of
    ! force.dat (it, fx(bmin..bmax), fy(bmin..bmax)), circ.dat (it,
    ! circ(1..L)) and slip.dat (it, slip) for every step, and the restart file
    ! every isave steps and at istop (src/operators.f90:180-186).
    !-------------------------------------------------------------------------!
    INTEGER :: itime
    INTEGER(C_INT) :: it_c, nsteps
    INTEGER :: k, s, ncol, t, b
    INTEGER(C_INT), DIMENSION(:), ALLOCATABLE :: cgit
    LOGICAL :: moving

    moving = ANY(bdyno > 0)
    nsteps = MIN(nblock, istop - itime)
    IF (moving) nsteps = 1  ! the body moves before every step
    DO t = itime + 1, itime + nsteps
       IF (MOD(t, isave) == 0) THEN
          nsteps = t - itime
          EXIT
       END IF
    END DO
    IF (nsteps < 1) nsteps = 1

    IF (.NOT. have_factor) CALL preprocess
    IF (.NOT. files_open) CALL open_outputs(itime + 1)

    ncol = 2*nbody + mgridlev + 1
    IF (.NOT. ALLOCATED(motion)) THEN
       ALLOCATE(motion(MAX(nf,1)))
       motion = 0.d0
    END IF
    IF (ALLOCATED(hist)) THEN
       IF (SIZE(hist) < ncol*nsteps) DEALLOCATE(hist)
    END IF
    IF (.NOT. ALLOCATED(hist)) ALLOCATE(hist(ncol*nsteps))

    IF (moving) THEN
       ! BOUNDARY MOTION (src/operators.f90:142-144): the prescribed motion of
       ! this step, then setup_reg on the device for the new positions
       motion = move_body(itime + 1)
       CALL ibfs_check(ibfs_move_body(handle, xb, yb), 'ibfs_move_body')
    END IF
    it_c = itime
    CALL ibfs_check(ibfs_advance(handle, it_c, nsteps, motion, hist), 'ibfs_advance')
    IF (.NOT. stationary) THEN  ! what cjgr reports (src/operators.f90:435-440)
       ALLOCATE(cgit(nsteps))
       CALL ibfs_check(ibfs_cg_iterations(handle, nsteps, cgit), 'ibfs_cg_iterations')
       DO s = 1, nsteps
          IF (cgit(s) >= cg_max_iter) WRITE(111,*) "WARNING!! CJGR used max iterations!"
          WRITE(111,*) "CJGR used ", cgit(s), " iterations."
       END DO
       WRITE(*,*) "* CJGR iterations = ", cgit(nsteps)
       DEALLOCATE(cgit)
    END IF

    DO s = 1, nsteps
       k = (s-1)*ncol
       WRITE(uforce,*) itime+s, (hist(k+b), b=1,nbody), (hist(k+nbody+b), b=1,nbody)
       WRITE(ucirc,*) itime+s, (hist(k+2*nbody+b), b=1,mgridlev)
       WRITE(uslip,*) itime+s, hist(k+ncol)
    END DO
    itime = it_c
    WRITE(*,*) "* itime =", itime, " SLIP =", hist((nsteps-1)*ncol+ncol)
    WRITE(111,*) "...Advanced to itime =", itime

    IF ((MOD(itime,isave).eq.0).or.(itime==istop)) THEN
       CALL write_variables(itime)
    END IF
  END SUBROUTINE advance

END MODULE operators
