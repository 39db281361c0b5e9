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

  ! actuator slots: points per slot and each slot's read_act_parameters
  ! (x_j, y_j, h_j, u_j, u_j_osc, theta_j, omg_j), handed to the library,
  ! which evaluates the jets on the device every step
  INTEGER, PARAMETER :: max_act = 100
  INTEGER, DIMENSION(max_act) :: act_npt = 0
  REAL(KIND(0.D0)), DIMENSION(7, max_act) :: act_par = 0.d0

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
    ! (src/grid.f90:353), written back to input/geom.inp.  Actuators
    ! (input/actuator.NNN.inp, NNN = 001..100) are handled as in
    ! src/grid.f90:286-316,383-392: a geom.inp with negative body numbers
    ! already holds their slots (their point counts are read back from
    ! input/slot.NNN.inp); otherwise each actuator's slot is appended with
    ! body number -NNN and geom.inp is rewritten.
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
       IF (ANY(geom%bodyno(1:geom%ntot) < 0)) THEN
          CALL count_actuators
       ELSE
          CALL append_actuators( ds, geom )
          CALL write_geometry( geom, 'replace' )
          WRITE(111,*) "...appended actuator info to geom.inp"
       END IF
    ELSE
       CALL circle( ds, 1.d0, 0.0d0, 0.d0, geom, 0)
       CALL append_actuators( ds, geom )
       CALL write_geometry( geom, 'new' )
       WRITE(111,*) "Generated geometry from primatives."
       WRITE(*,*) "Generated geometry from primatives."
    END IF
  END SUBROUTINE setup_geometry

  SUBROUTINE write_geometry( geom, stat )
    TYPE(bodylist) :: geom
    CHARACTER(*) :: stat
    INTEGER :: i
    OPEN(unit=8,file='input/geom.inp',form='formatted',status=stat)
    WRITE(8,*) geom%ntot
    DO i=1,geom%ntot
       WRITE(8,*) geom%xbody(i), geom%ybody(i), geom%bodyno(i)
    END DO
    CLOSE(8)
  END SUBROUTINE write_geometry

  SUBROUTINE read_actuator( i_act, par )
    ! the namelist of input/actuator.NNN.inp (src/grid.f90:262-263,309-311)
    INTEGER :: i_act
    REAL(KIND(0.D0)), DIMENSION(7) :: par
    CHARACTER(3) :: act_num
    REAL(KIND(0.D0)) :: x_j, y_j, h_j, u_j, u_j_osc, theta_j, omg_j
    NAMELIST /read_act_parameters/ x_j,y_j,h_j,u_j,u_j_osc,theta_j,omg_j

    x_j = 0.d0; y_j = 0.d0; h_j = 0.d0; u_j = 0.d0; u_j_osc = 0.d0; theta_j = 0.d0; omg_j = 0.d0
    WRITE(act_num,"(I3.3)") i_act
    OPEN(unit=7,file="input/actuator."//act_num//".inp",form='formatted',status='old')
    READ(unit=7,nml=read_act_parameters)
    CLOSE(7)
    par = (/ x_j, y_j, h_j, u_j, u_j_osc, theta_j, omg_j /)
  END SUBROUTINE read_actuator

  SUBROUTINE append_actuators( ds, geom )
    ! every input/actuator.NNN.inp adds its slot with body number -NNN
    ! (src/grid.f90:303-316 and 383-392)
    TYPE(bodylist) :: geom
    REAL(KIND(0.D0)) :: ds
    INTEGER :: i_act, ntot0
    CHARACTER(3) :: act_num
    LOGICAL :: readact

    DO i_act = 1, max_act
       WRITE(act_num,"(I3.3)") i_act
       INQUIRE(file="input/actuator."//act_num//".inp",exist=readact)
       IF (readact) THEN
          n_act = n_act + 1
          IF (n_act /= i_act) STOP 'actuator files must be numbered 001, 002, ... without gaps'
          CALL read_actuator( i_act, act_par(:,n_act) )
          ntot0 = geom%ntot
          CALL actuator_slot( act_par(1,n_act), act_par(2,n_act), act_par(3,n_act), ds, &
                              act_par(6,n_act), geom, -i_act )
          act_npt(n_act) = geom%ntot - ntot0
       END IF
    END DO
  END SUBROUTINE append_actuators

  SUBROUTINE count_actuators
    ! slots already in geom.inp: their point counts from input/slot.NNN.inp
    ! (src/grid.f90:286-300)
    INTEGER :: i_act, n_pt
    CHARACTER(3) :: act_num
    LOGICAL :: readact

    DO i_act = 1, max_act
       WRITE(act_num,"(I3.3)") i_act
       INQUIRE(file="input/actuator."//act_num//".inp",exist=readact)
       IF (readact) THEN
          OPEN(unit=9,file="input/slot."//act_num//".inp",form='formatted',status='old')
          READ(9,*) n_pt
          CLOSE(9)
          n_act = n_act + 1
          IF (n_act /= i_act) STOP 'actuator files must be numbered 001, 002, ... without gaps'
          n_act_pt = n_act_pt + n_pt
          act_npt(n_act) = n_pt
          CALL read_actuator( i_act, act_par(:,n_act) )
       END IF
    END DO
  END SUBROUTINE count_actuators

  SUBROUTINE actuator_slot( x, y, hs, ds, angle, geom, num )
    !-------------------------------------------------------------------------!
    ! src/grid.f90:636-718: INT(hs/ds)+1 points on the segment of width hs
    ! through (x, y), normal to the jet direction angle (degrees), or the
    ! single point (x, y) when hs < ds; body number num.  Logs the slot in
    ! input/slot.dat and its point count in input/slot.NNN.inp.
    !-------------------------------------------------------------------------!
    TYPE(bodylist) :: geom
    REAL(KIND(0.D0)) :: x, y, hs, ds, angle, seg
    INTEGER :: num, npt, i
    CHARACTER(3) :: act_num
    LOGICAL :: ex

    npt = INT(hs/ds)
    IF (geom%ntot + npt + 1 > maxpts) STOP 'too many body points'
    IF (npt == 0) THEN
       geom%xbody(geom%ntot+1) = x
       geom%ybody(geom%ntot+1) = y
       geom%bodyno(geom%ntot+1) = num
       geom%ntot = geom%ntot + 1
    ELSE
       DO i = 1, npt+1
          seg = REAL(i-1)*hs/REAL(npt)
          geom%xbody(geom%ntot+i) = x - hs/2.0d0*COS(pi*(angle+90.)/180.) + seg*COS(pi*(angle+90.)/180.)
          geom%ybody(geom%ntot+i) = y - hs/2.0d0*SIN(pi*(angle+90.)/180.) + seg*SIN(pi*(angle+90.)/180.)
          geom%bodyno(geom%ntot+i) = num
       END DO
       geom%ntot = geom%ntot + npt + 1
    END IF

    WRITE(act_num,"(I3.3)") -num
    IF (num == -1) THEN
       OPEN(unit=116,file="input/slot.dat",form="formatted",status="replace")
       OPEN(unit=118,file="input/slot."//act_num//".inp",form="formatted",status="replace")
    ELSE
       INQUIRE(file="input/slot.dat",exist=ex)
       IF (ex) THEN
          OPEN(unit=116,file="input/slot.dat",form="formatted",status="old",position="append")
       ELSE
          OPEN(unit=116,file="input/slot.dat",form="formatted",status="new")
       END IF
       OPEN(unit=118,file="input/slot."//act_num//".inp",form="formatted",status="replace")
    END IF
    WRITE(116,*) " ----------- ACTUATOR No. ",-num
    WRITE(116,*) "   o Desired slot width               = ",hs
    IF (npt == 0) THEN
       WRITE(116,*) "   o Actual slot width in simulation  = ",ds
    ELSE
       WRITE(116,*) "   o Actual slot width in simulation  = ",hs
    END IF
    WRITE(116,*) "   o Number of pts used to model slot = ",(npt+1)
    CLOSE(116)
    WRITE(118,*) (npt+1)
    CLOSE(118)
    n_act_pt = n_act_pt + (npt + 1)
  END SUBROUTINE actuator_slot

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
