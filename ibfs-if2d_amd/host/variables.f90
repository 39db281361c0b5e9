MODULE variables
  !---------------------------------------------------------------------------!
  ! Host copies of the solver state, used only for restart and Cholesky
  ! files; the live state is on the GPU.  File layouts are the reference's:
  ! output/ibfsNNNNNNN.var holds one unformatted record
  ! (omega, fb, rhs_old, q, q0) and input/ibfs.chd one record
  ! (cholmat, cholvec) (src/variables.f90:110-175).
  !---------------------------------------------------------------------------!
  USE ISO_C_BINDING
  USE grid
  USE ibfs_c
  IMPLICIT NONE

  REAL(C_DOUBLE), DIMENSION(:,:,:), ALLOCATABLE :: omega, rhs_old
  REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE :: fb
  REAL(C_DOUBLE), DIMENSION(:,:), ALLOCATABLE :: q, q0
  REAL(C_DOUBLE), DIMENSION(:,:), ALLOCATABLE :: cholmat
  REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE :: cholvec
  INTEGER :: nbody, bmin, bmax

CONTAINS

  SUBROUTINE setup_variables
    !-------------------------------------------------------------------------!
    ! Creates the GPU solver (tables, initial condition: zero flow or uniform
    ! flow as src/variables.f90:54-67), or restarts from a saved state.
    !-------------------------------------------------------------------------!
    INTEGER(C_INT) :: dims(12)
    INTEGER(C_INT), ALLOCATABLE :: bd(:)
    INTEGER :: uf
    INTEGER(C_INT) :: st
    LOGICAL :: readchol

    uf = 0
    IF (unif_flow) uf = 1
    ALLOCATE(bd(MAX(nb,1)))
    bd = 0
    IF (nb > 0) bd(1:nb) = bdyno
    CALL ibfs_check(ibfs_create(m, n, mgridlev, nb, dt, Re, len, offsetx, offsety, &
                                xb, yb, bd, uf, precision, handle), 'ibfs_create')
    CALL ibfs_check(ibfs_dims(handle, dims), 'ibfs_dims')
    IF (dims(3) /= mgridlev) STOP 'level count mismatch with the library'
    ! the force solve: the factor for stationary bodies, else cjgr
    ! (src/operators.f90:166-171)
    st = 0
    IF (stationary) st = 1
    CALL ibfs_check(ibfs_set_solver(handle, st, cgtol, cg_max_iter), 'ibfs_set_solver')
    ! the jets of the actuator slots (actuation, src/operators.f90:150-152)
    IF (n_act > 0) CALL ibfs_check(ibfs_set_actuators(handle, n_act, act_npt, act_par), 'ibfs_set_actuators')
    nbody = dims(9); bmin = dims(10); bmax = dims(11)

    ALLOCATE( omega(2:m,2:n,mgridlev), fb(MAX(nf,1)), rhs_old(2:m,2:n,mgridlev) )
    ALLOCATE( q(nq,mgridlev), q0(nq,mgridlev) )
    IF (stationary) ALLOCATE( cholmat(MAX(nf,1),MAX(nf,1)), cholvec(MAX(nf,1)) )

    IF (istart /= 0) THEN
       CALL read_variables(istart)
       CALL ibfs_check(ibfs_set_state(handle, omega, rhs_old, q, q0, fb), 'ibfs_set_state')
       IF (stationary) THEN
          readchol = .FALSE.
          INQUIRE(file='input/ibfs.chd', exist=readchol)
          IF (readchol) readchol = chd_is_current()
          IF (readchol) THEN  ! else preprocess builds it before the first step
             CALL read_cholesky
             CALL ibfs_check(ibfs_set_cholesky(handle, cholmat, cholvec), 'ibfs_set_cholesky')
          END IF
       END IF
    END IF
  END SUBROUTINE setup_variables

  SUBROUTINE destroy_variables
    INTEGER(C_INT) :: rc
    rc = ibfs_destroy(handle)
    handle = C_NULL_PTR
    DEALLOCATE( omega, fb, rhs_old, q, q0 )
    IF (ALLOCATED(cholmat)) DEALLOCATE( cholmat, cholvec )
  END SUBROUTINE destroy_variables

  SUBROUTINE write_cholesky
    OPEN(unit=100,file="input/ibfs.chd",form="unformatted",status="unknown")
    WRITE(100) cholmat, cholvec
    CLOSE(100)
    ! what the factor was built for, checked before the file is reused
    OPEN(unit=100,file="input/ibfs.chd.key",form="formatted",status="replace")
    WRITE(100,"(*(ES25.17))") chd_key()
    CLOSE(100)
    WRITE(111,*) "Saving cholesky decomposition..."
  END SUBROUTINE write_cholesky

  FUNCTION chd_key() RESULT(key)
    !-------------------------------------------------------------------------!
    ! The inputs the force-system factor depends on: grid, levels, domain,
    ! dt and Re (ainv of level 1 is in A), the point count and moments of
    ! the point list that change with any point moved or reordered.
    !-------------------------------------------------------------------------!
    REAL(KIND(0.D0)), DIMENSION(14) :: key
    INTEGER :: k
    key = 0.d0
    key(1) = m; key(2) = n; key(3) = mgridlev; key(4) = nb
    key(5) = len; key(6) = offsetx; key(7) = offsety; key(8) = dt; key(9) = Re
    DO k = 1, nb
       key(10) = key(10) + xb(k)
       key(11) = key(11) + yb(k)
       key(12) = key(12) + REAL(k,KIND(0.D0))*xb(k)
       key(13) = key(13) + REAL(k,KIND(0.D0))*yb(k)
       key(14) = key(14) + xb(k)*xb(k) + yb(k)*yb(k)
    END DO
  END FUNCTION chd_key

  LOGICAL FUNCTION chd_is_current()
    !-------------------------------------------------------------------------!
    ! The reference reuses input/ibfs.chd whatever run wrote it
    ! (src/operators.f90:41-43).  Here a cached factor is used only if its
    ! record has the size of this run's factor and, when input/ibfs.chd.key
    ! is present (every factor this program writes has one), that key
    ! matches this run.  A file without a key (written by the reference) is
    ! trusted on its size alone.
    !-------------------------------------------------------------------------!
    INTEGER(KIND=8) :: sz, want
    LOGICAL :: haskey
    REAL(KIND(0.D0)), DIMENSION(14) :: key, cur
    INTEGER :: k, ios

    chd_is_current = .FALSE.
    INQUIRE(file="input/ibfs.chd", size=sz)
    want = 8_8 * (INT(nf,8)*INT(nf,8) + INT(nf,8)) + 8_8
    IF (sz /= want) THEN
       WRITE(111,*) "input/ibfs.chd has", sz, " bytes, this run's factor needs", want
       RETURN
    END IF
    INQUIRE(file="input/ibfs.chd.key", exist=haskey)
    IF (haskey) THEN
       OPEN(unit=100,file="input/ibfs.chd.key",form="formatted",status="old")
       READ(100,*,iostat=ios) key
       CLOSE(100)
       IF (ios /= 0) RETURN
       cur = chd_key()
       DO k = 1, 14
          IF (ABS(key(k) - cur(k)) > 1.d-12*MAX(1.d0, ABS(cur(k)))) THEN
             WRITE(111,*) "input/ibfs.chd was built for another case (key entry", k, ")"
             RETURN
          END IF
       END DO
    END IF
    chd_is_current = .TRUE.
  END FUNCTION chd_is_current

  SUBROUTINE read_cholesky
    OPEN(unit=100,file="input/ibfs.chd",form="unformatted",status="unknown")
    READ(100) cholmat, cholvec
    CLOSE(100)
    WRITE(111,*) "Reading Cholesky decompopsition..."
    WRITE(*,*) "Reading Cholesky premade decomposition..."
  END SUBROUTINE read_cholesky

  SUBROUTINE write_variables(it)
    CHARACTER(7) :: charit
    INTEGER :: it
    CALL ibfs_check(ibfs_get_state(handle, omega, rhs_old, q, q0, fb), 'ibfs_get_state')
    WRITE(charit,"(I7.7)") it
    OPEN(unit=100,file="output/ibfs"//charit//".var",form="unformatted",status="unknown")
    WRITE(100) omega,fb,rhs_old, q, q0
    CLOSE(100)
    WRITE(111,*) "Writing restart file at itime=",it
  END SUBROUTINE write_variables

  SUBROUTINE read_variables(it)
    CHARACTER(7) :: charit
    INTEGER :: it
    LOGICAL :: ex
    WRITE(charit,"(I7.7)") it
    INQUIRE(file="output/ibfs"//charit//".var", exist=ex)
    IF (.NOT. ex) THEN
       WRITE(*,*) "Restart file output/ibfs"//charit//".var not found (istart =", it, ")"
       WRITE(111,*) "Restart file output/ibfs"//charit//".var not found (istart =", it, ")"
       STOP 'restart file not found'
    END IF
    OPEN(unit=100,file="output/ibfs"//charit//".var",form="unformatted",status="unknown")
    READ(100) omega, fb, rhs_old, q, q0
    CLOSE(100)
    WRITE(111,*) "Reading restart file at itime=",it
    WRITE(*,*) "Reading restart file at itime =",it
  END SUBROUTINE read_variables

END MODULE variables
