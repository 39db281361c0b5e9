PROGRAM post
  !---------------------------------------------------------------------------!
  ! Post-processing driver with the flow of src/post.f90:1-43: for every
  ! saved step istart..istop (multiples of isave, and istop) it reads the
  ! restart file output/ibfsNNNNNNN.var and writes one Tecplot ASCII zone per
  ! level (write_tecplot, src/operators.f90:702-763) to
  ! post/ibfs_g<level>t<step>_vel.dat.  reg(fb) comes from the library.
  !---------------------------------------------------------------------------!
  USE ISO_C_BINDING
  USE parameters
  USE grid
  USE ibfs_c
  USE variables
  IMPLICIT NONE
  INTEGER :: it, k
  CHARACTER(10) :: date, time

  OPEN(unit=111,file='output/ibfs.inf',status='unknown',position='append', form='formatted')
  WRITE(111,*) "==================================="
  WRITE(111,*) "|   IBFS post-processing (GPU)    |"
  WRITE(111,*) "==================================="
  CALL DATE_AND_TIME(date,time)
  WRITE(111,*) "Executed at time:",time," on date:",date

  CALL input
  CALL setup

  it = istart
  DO WHILE (it <= istop)
     IF ((MOD(it,isave) == 0) .OR. (it == istop)) THEN
        istart = it
        WRITE(*,*) "Post-processing: itime = ", it
        CALL setup_variables
        DO k = 1, mgridlev
           CALL write_tecplot(it, k)
        END DO
        CALL destroy_variables
     END IF
     it = it + 1
  END DO
  CLOSE(111)

CONTAINS

  INTEGER FUNCTION iu(i, j)  ! u face (i,j), i = 1..m+1, j = 1..n (src/grid.f90:119-125)
    INTEGER :: i, j
    iu = i + (j-1)*(m+1)
  END FUNCTION iu

  INTEGER FUNCTION iv(i, j)  ! v face (i,j), i = 1..m, j = 1..n+1 (src/grid.f90:126-131)
    INTEGER :: i, j
    iv = (m+1)*n + i + (j-1)*m
  END FUNCTION iv

  SUBROUTINE write_tecplot(it, igrid)
    !-------------------------------------------------------------------------!
    ! Level igrid at the vertices 2..m x 2..n: velocity (q + q0 averaged to
    ! the vertex, over the level's spacing), the smeared force density of
    ! the finest level's reg(fb) (delta * face sum / dt), and vorticity over
    ! the spacing squared, in the layout and point format of the
    ! reference's Tecplot files.
    !-------------------------------------------------------------------------!
    INTEGER :: it, igrid
    CHARACTER(7) :: charit
    CHARACTER(1) :: chargrid
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE :: etf
    REAL(KIND(0.D0)) :: fac, del, offx, offy, u1, v1, fx, fy, w
    INTEGER :: i, j

    fac = REAL(2**(igrid-1), KIND(0.D0))
    offx = (fac-1.d0)/2.d0 * len
    offy = (fac-1.d0)/2.d0 * delta*REAL(n, KIND(0.D0))
    del = fac*delta
    ALLOCATE(etf(nq))
    etf = 0.d0
    IF (nf > 0) CALL ibfs_check(ibfs_op_reg(handle, fb, etf), 'ibfs_op_reg')

    WRITE(charit,"(I7.7)") it
    WRITE(chargrid,"(I1.1)") igrid
    OPEN(unit=100,file="post/ibfs_g"//chargrid//"t"//charit//"_vel.dat",form="formatted",status="unknown")
    WRITE(100,*) 'TITLE     = "IBFS V1.0 AT ITIME=',it,'"'
    WRITE(100,*) 'VARIABLES = "x"'
    WRITE(100,*) '"y"'
    WRITE(100,*) '"u"'
    WRITE(100,*) '"v"'
    WRITE(100,*) '"etfx"'
    WRITE(100,*) '"etfy"'
    WRITE(100,*) '"vort"'
    WRITE(100,*) 'ZONE T="Rectangular zone"'
    WRITE(100,*) 'I=',m-1,', J=',n-1,', K=1, ZONETYPE=Ordered'
    WRITE(100,*) 'DATAPACKING=POINT'
    WRITE(100,*) 'DT=(SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE)'
    DO j = 2, n
       DO i = 2, m
          u1 = 0.5d0*(q(iu(i,j),igrid) + q(iu(i,j-1),igrid) + q0(iu(i,j),igrid) + q0(iu(i,j-1),igrid))/del
          v1 = 0.5d0*(q(iv(i,j),igrid) + q(iv(i-1,j),igrid) + q0(iv(i,j),igrid) + q0(iv(i-1,j),igrid))/del
          fx = delta*(etf(iu(i,j)) + etf(iu(i,j-1)))/dt
          fy = delta*(etf(iv(i,j)) + etf(iv(i-1,j)))/dt
          w = omega(i,j,igrid)/del/del
          WRITE(100,"(8E14.5)") fac*delta*REAL(i-1, KIND(0.D0)) - offx - offsetx, &
                                fac*delta*REAL(j-1, KIND(0.D0)) - offy - offsety, u1, v1, fx, fy, w
       END DO
    END DO
    CLOSE(100)
    DEALLOCATE(etf)
  END SUBROUTINE write_tecplot

END PROGRAM post
