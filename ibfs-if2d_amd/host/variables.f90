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
    nbody = dims(9); bmin = dims(10); bmax = dims(11)

    ALLOCATE( omega(2:m,2:n,mgridlev), fb(MAX(nf,1)), rhs_old(2:m,2:n,mgridlev) )
    ALLOCATE( q(nq,mgridlev), q0(nq,mgridlev) )
    IF (stationary) ALLOCATE( cholmat(MAX(nf,1),MAX(nf,1)), cholvec(MAX(nf,1)) )

    IF (istart /= 0) THEN
       CALL read_variables(istart)
       CALL ibfs_check(ibfs_set_state(handle, omega, rhs_old, q, q0, fb), 'ibfs_set_state')
       IF (stationary) THEN
          CALL read_cholesky
          CALL ibfs_check(ibfs_set_cholesky(handle, cholmat, cholvec), 'ibfs_set_cholesky')
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
    WRITE(111,*) "Saving cholesky decomposition..."
  END SUBROUTINE write_cholesky

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
    WRITE(charit,"(I7.7)") it
    OPEN(unit=100,file="output/ibfs"//charit//".var",form="unformatted",status="unknown")
    READ(100) omega, fb, rhs_old, q, q0
    CLOSE(100)
    WRITE(111,*) "Reading restart file at itime=",it
    WRITE(*,*) "Reading restart file at itime =",it
  END SUBROUTINE read_variables

END MODULE variables
