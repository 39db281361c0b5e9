MODULE ibfs_c
  !---------------------------------------------------------------------------!
  ! ISO_C_BINDING interfaces of the gfx950 library (include/ibfs.h).  Every
  ! entry returns 0 on success; ibfs_check turns a failure into STOP with the
  ! library's message, the way the reference stops on its own errors.
  !---------------------------------------------------------------------------!
  USE ISO_C_BINDING
  IMPLICIT NONE

  TYPE(C_PTR) :: handle = C_NULL_PTR

  INTERFACE
     INTEGER(C_INT) FUNCTION ibfs_create(m, n, mgridlev, nb, dt, re, len, offsetx, offsety, &
          xb, yb, bdyno, unif_flow, precision, out) BIND(C, name='ibfs_create')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       INTEGER(C_INT), VALUE :: m, n, mgridlev, nb, unif_flow, precision
       REAL(C_DOUBLE), VALUE :: dt, re, len, offsetx, offsety
       REAL(C_DOUBLE), DIMENSION(*) :: xb, yb
       INTEGER(C_INT), DIMENSION(*) :: bdyno
       TYPE(C_PTR) :: out
     END FUNCTION ibfs_create

     INTEGER(C_INT) FUNCTION ibfs_dims(h, dims) BIND(C, name='ibfs_dims')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: h
       INTEGER(C_INT), DIMENSION(*) :: dims
     END FUNCTION ibfs_dims

     INTEGER(C_INT) FUNCTION ibfs_set_state(h, omega, rhs_old, q, q0, fb) BIND(C, name='ibfs_set_state')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), DIMENSION(*) :: omega, rhs_old, q, q0, fb
     END FUNCTION ibfs_set_state

     INTEGER(C_INT) FUNCTION ibfs_get_state(h, omega, rhs_old, q, q0, fb) BIND(C, name='ibfs_get_state')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), DIMENSION(*) :: omega, rhs_old, q, q0, fb
     END FUNCTION ibfs_get_state

     INTEGER(C_INT) FUNCTION ibfs_set_cholesky(h, cholmat, cholvec) BIND(C, name='ibfs_set_cholesky')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), DIMENSION(*) :: cholmat, cholvec
     END FUNCTION ibfs_set_cholesky

     INTEGER(C_INT) FUNCTION ibfs_build_cholesky(h, cholmat, cholvec) BIND(C, name='ibfs_build_cholesky')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), DIMENSION(*) :: cholmat, cholvec
     END FUNCTION ibfs_build_cholesky

     INTEGER(C_INT) FUNCTION ibfs_advance(h, itime, nsteps, motion, hist) BIND(C, name='ibfs_advance')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       INTEGER(C_INT) :: itime
       INTEGER(C_INT), VALUE :: nsteps
       REAL(C_DOUBLE), DIMENSION(*) :: motion, hist
     END FUNCTION ibfs_advance

     INTEGER(C_INT) FUNCTION ibfs_set_solver(h, stationary, cgtol, cg_max_iter) BIND(C, name='ibfs_set_solver')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       INTEGER(C_INT), VALUE :: stationary, cg_max_iter
       REAL(C_DOUBLE), VALUE :: cgtol
     END FUNCTION ibfs_set_solver

     INTEGER(C_INT) FUNCTION ibfs_move_body(h, xb, yb) BIND(C, name='ibfs_move_body')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), DIMENSION(*) :: xb, yb
     END FUNCTION ibfs_move_body

     INTEGER(C_INT) FUNCTION ibfs_set_actuators(h, n_act, npt, par) BIND(C, name='ibfs_set_actuators')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       INTEGER(C_INT), VALUE :: n_act
       INTEGER(C_INT), DIMENSION(*) :: npt
       REAL(C_DOUBLE), DIMENSION(*) :: par
     END FUNCTION ibfs_set_actuators

     INTEGER(C_INT) FUNCTION ibfs_cg_iterations(h, nsteps, iters) BIND(C, name='ibfs_cg_iterations')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: h
       INTEGER(C_INT), VALUE :: nsteps
       INTEGER(C_INT), DIMENSION(*) :: iters
     END FUNCTION ibfs_cg_iterations

     INTEGER(C_INT) FUNCTION ibfs_op_reg(h, h0, hq) BIND(C, name='ibfs_op_reg')
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: h
       REAL(C_DOUBLE), DIMENSION(*) :: h0, hq
     END FUNCTION ibfs_op_reg

     FUNCTION ibfs_last_error(h) BIND(C, name='ibfs_last_error') RESULT(msg)
       IMPORT :: C_PTR
       TYPE(C_PTR), VALUE :: h
       TYPE(C_PTR) :: msg
     END FUNCTION ibfs_last_error

     INTEGER(C_INT) FUNCTION ibfs_destroy(h) BIND(C, name='ibfs_destroy')
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: h
     END FUNCTION ibfs_destroy
  END INTERFACE

CONTAINS

  SUBROUTINE ibfs_check(rc, what)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(*), INTENT(IN) :: what
    CHARACTER(KIND=C_CHAR), DIMENSION(:), POINTER :: cmsg
    CHARACTER(512) :: msg
    INTEGER :: i
    IF (rc == 0) RETURN
    msg = ' '
    IF (C_ASSOCIATED(handle)) THEN
       CALL C_F_POINTER(ibfs_last_error(handle), cmsg, [512])
       DO i = 1, 512
          IF (cmsg(i) == C_NULL_CHAR) EXIT
          msg(i:i) = cmsg(i)
       END DO
    END IF
    WRITE(*,*) 'IBFS library error in ', what, ' (code ', rc, '): ', TRIM(msg)
    WRITE(111,*) 'IBFS library error in ', what, ' (code ', rc, '): ', TRIM(msg)
    STOP 'ibfs library call failed'
  END SUBROUTINE ibfs_check

END MODULE ibfs_c
