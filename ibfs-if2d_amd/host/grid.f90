MODULE grid
  !---------------------------------------------------------------------------!
  ! Grid sizes and the immersed-boundary point set.  Follows the set-up of
  ! the reference (src/grid.f90:44-139): spacing, the divisible-by-4 check,
  ! the level clamp, the geometry (input/geom.inp, or a default circle of
  ! diameter 1 written to it) and the index counts.  The smearing tables
  ! (src/grid.f90:142-230) are built by the GPU library from xb, yb.
  !---------------------------------------------------------------------------!
  USE parameters
  IMPLICIT NONE

  REAL(KIND(0.D0)) :: delta
  REAL(KIND(0.D0)), DIMENSION(:), ALLOCATABLE :: xb, yb
  INTEGER, DIMENSION(:), ALLOCATABLE :: bdyno
  INTEGER :: nq, nb, nf, ns, nw

  INTEGER, PARAMETER :: maxpts = 100000
  TYPE bodylist
     INTEGER :: ntot = 0
     REAL(KIND(0.D0)), DIMENSION(maxpts) :: xbody, ybody
     INTEGER, DIMENSION(maxpts) :: bodyno
  END TYPE bodylist

CONTAINS

  SUBROUTINE setup
    INTEGER :: i, nlevmax
    TYPE(bodylist), ALLOCATABLE :: geom

    delta = len/REAL(m)
    IF ((MOD(m,4).ne.0).or.(MOD(n,4).ne.0)) THEN
       STOP 'HEY! Grid nodes must be divisible by 4!!'
    END IF

    nlevmax = FLOOR(  LOG( REAL( MIN(m,n) /2)  ) / LOG( 2.d0) ) - 1
    WRITE(*,*) 'Maximum:',nlevmax,'levels in MGRID'
    mgridlev = MIN(mgridlev,nlevmax)
    WRITE(*,*) 'Using:',mgridlev,'levels in MGRID'

    ALLOCATE(geom)
    CALL setup_geometry( delta, geom )
    nb = geom%ntot
    ALLOCATE( xb(nb), yb(nb), bdyno(nb) )
    DO i=1,nb
       xb(i) = geom%xbody(i)
       yb(i) = geom%ybody(i)
       bdyno(i) = geom%bodyno(i)
    END DO
    DEALLOCATE(geom)

    ns = (m-1)*(n-1)
    nw = ns
    nf = 2*nb
    nq = (m+1)*n + (n+1)*m
    IF (ANY(bdyno>0) .AND. stationary) THEN
       ! a factor of the initial geometry would be wrong after the first move
       WRITE(*,*) 'Moving bodies (bdyno > 0) need stationary = F (the CG force solve)'
       STOP 'bdyno > 0 with stationary = T'
    END IF
    WRITE(111,*) "Set up grid and body..."
  END SUBROUTINE setup

  SUBROUTINE setup_geometry( ds, geom )
    !-------------------------------------------------------------------------!
    ! input/geom.inp (Nb, then "x y bodyno" lines, src/grid.f90:270-276), or
    ! the reference's default body: a circle of diameter 1 at the origin
    ! (src/grid.f90:353), written back to input/geom.inp.
    !-------------------------------------------------------------------------!
    TYPE(bodylist) :: geom
    REAL(KIND(0.D0)) :: ds
    LOGICAL :: readinput
    INTEGER :: i

    INQUIRE(file='input/geom.inp',exist=readinput)
    IF (readinput) THEN
       OPEN(unit=8,file='input/geom.inp',form='formatted',status='old')
       READ(8,*) geom%ntot
       IF (geom%ntot > maxpts) STOP 'too many points in geom.inp'
       DO i=1,geom%ntot
          READ(8,*) geom%xbody(i), geom%ybody(i), geom%bodyno(i)
       END DO
       CLOSE(8)
       WRITE(111,*) "Reading geometry from geom.inp"
       WRITE(*,*) "Reading geometry from geom.inp"
    ELSE
       CALL circle( ds, 1.d0, 0.0d0, 0.d0, geom, 0)
       OPEN(unit=8,file='input/geom.inp',form='formatted',status='new')
       WRITE(8,*) geom%ntot
       DO i=1,geom%ntot
          WRITE(8,*) geom%xbody(i), geom%ybody(i), geom%bodyno(i)
       END DO
       CLOSE(8)
       WRITE(111,*) "Generated geometry from primatives."
       WRITE(*,*) "Generated geometry from primatives."
    END IF
  END SUBROUTINE setup_geometry

  SUBROUTINE circle( ds, dia, x, y, geom, num )
    ! npt = INT(pi*dia/ds) points on the circle, first at angle 0
    ! (src/grid.f90:496-520)
    TYPE(bodylist) :: geom
    REAL(KIND(0.D0)) :: ds, dia, x, y
    INTEGER :: num, npt, i

    npt = INT( pi*dia/ds )
    IF (geom%ntot + npt > maxpts) STOP 'too many body points'
    DO i=1,npt
       geom%bodyno(geom%ntot+i) = num
       geom%xbody(geom%ntot+i) = x + 0.5D0*dia*COS(2.*pi*REAL(i-1)/REAL(npt))
       geom%ybody(geom%ntot+i) = y + 0.5D0*dia*SIN(2.*pi*REAL(i-1)/REAL(npt))
    END DO
    geom%ntot = geom%ntot+npt
  END SUBROUTINE circle

  FUNCTION move_body(itime) RESULT(motion)
    !-------------------------------------------------------------------------!
    ! Prescribed motion of the moving bodies (bdyno > 0) at step itime, with
    ! the interface of src/grid.f90:523-576: the point velocities in the
    ! force layout (all x components, then all y components), scaled by
    ! delta (src/grid.f90:573-574), and xb, yb updated to the positions of
    ! this step.  No motion is prescribed here, as in the reference as
    ! shipped: the points stay put and their velocities are zero.  A case
    ! with a moving body sets, for every point i of it, motion(i) and
    ! motion(nb+i) to the velocity at t = dt*itime and moves xb(i), yb(i) by
    ! the displacement since the previous step (itime is the step number).
    !-------------------------------------------------------------------------!
    INTEGER :: itime
    REAL(KIND(0.D0)), DIMENSION(MAX(nf,1)) :: motion
    INTEGER :: i

    motion = 0.d0
    DO i = 1, nb
       IF (bdyno(i) > 0) THEN
          motion(i) = 0.d0
          motion(nb+i) = 0.d0
       END IF
    END DO
    motion = motion * delta
  END FUNCTION move_body

END MODULE grid
