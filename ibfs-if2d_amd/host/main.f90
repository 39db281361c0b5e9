PROGRAM main
  !---------------------------------------------------------------------------!
  ! IBFS multi-domain solver, GPU host.  Same flow as the reference driver
  ! (src/main.f90:1-63): log file, input, setup, variables, the advance loop,
  ! teardown; the step runs in the gfx950 library.
  !---------------------------------------------------------------------------!
  USE parameters
  USE grid
  USE variables
  USE operators
  IMPLICIT NONE
  INTEGER :: it
  INTEGER(KIND=8) :: c0, c1, rate
  CHARACTER(10) :: date, time

  OPEN(unit=111,file='output/ibfs.inf',status='unknown',position='append', form='formatted')
  WRITE(111,*) "==================================="
  WRITE(111,*) "|   Welcome to IBFS MGRID (GPU)   |"
  WRITE(111,*) "==================================="
  CALL DATE_AND_TIME(date,time)
  WRITE(111,*) "Executed at time:",time," on date:",date
  WRITE(*,*) "IBFS MGRID, gfx950 step library"

  CALL input
  CALL setup
  CALL setup_variables

  it = istart
  CALL SYSTEM_CLOCK(c0, rate)
  DO WHILE (it < istop)
     CALL advance(it)
  END DO
  CALL SYSTEM_CLOCK(c1)

  CALL close_outputs
  CALL destroy_variables

  WRITE(*,*) "Total wall time: ", REAL(c1-c0)/REAL(rate), " s"
  WRITE(111,*) "Total wall time: ", REAL(c1-c0)/REAL(rate), " s"
  CLOSE(111)
END PROGRAM main
