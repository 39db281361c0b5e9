MODULE parameters
  !---------------------------------------------------------------------------!
  ! Run parameters of the GPU host.  The &READ_PARAMETERS namelist in
  ! input/ibfs.inp has the same sixteen names, defaults and meaning as the
  ! reference (src/parameters.f90:12-31, 48-50); if the file is absent the
  ! defaults are written to it, as there.
  !
  ! Library options that the reference does not have live in a separate,
  ! optional namelist &IBFS_GPU in input/ibfs_gpu.inp so that ibfs.inp stays
  ! unchanged:
  !   precision  0 = plain FP64 (default), 1 = mixed FP32/FP64
  !   nblock     steps per library call (outputs are unchanged; default 1)
  !---------------------------------------------------------------------------!
  IMPLICIT NONE

  INTEGER :: istart = 0
  INTEGER :: istop = 400
  INTEGER :: isave = 25
  INTEGER :: m  = 100
  INTEGER :: n  = 100
  REAL(KIND(0.0D0)) :: dt = 0.005d0
  REAL(KIND(0.0D0)) :: Re = 300.d0
  REAL(KIND(0.0D0)) :: cgtol = 1.D-7
  INTEGER :: cg_max_iter = 3000
  INTEGER :: n_act = 0
  INTEGER :: n_act_pt = 0
  REAL(KIND(0.D0)) :: len = 2.d0
  REAL(KIND(0.D0)) :: offsetx = 1.d0
  REAL(KIND(0.D0)) :: offsety = 1.d0
  INTEGER :: mgridlev = 4
  LOGICAL :: unif_flow = .TRUE.
  LOGICAL :: make_vortex = .FALSE.
  LOGICAL :: stationary = .FALSE.

  INTEGER :: precision = 0
  INTEGER :: nblock = 1

  REAL(KIND(0.D0)) :: pi = 3.1415926535897932384d0

CONTAINS

  SUBROUTINE input
    LOGICAL :: readinput, readgpu
    NAMELIST /read_parameters/ istart,istop,isave,m,n,dt,Re,cgtol, &
                                cg_max_iter,unif_flow,len, offsetx, &
                                offsety,mgridlev,make_vortex,stationary
    NAMELIST /ibfs_gpu/ precision, nblock

    INQUIRE(file='input/ibfs.inp',exist=readinput)
    IF (readinput) THEN
      OPEN(unit=3,file='input/ibfs.inp',form='formatted',status='old')
      READ(unit=3,nml=read_parameters)
      CLOSE(3)
      WRITE(111,*) "... Read parameters from ibfs.inp"
      WRITE(111,*) "... Input parameters follow"
      WRITE(unit=111,nml=read_parameters)
    ELSE
      OPEN(unit=3,file='input/ibfs.inp',form='formatted',status='new')
      WRITE(unit=3,nml=read_parameters)
      CLOSE(3)
      WRITE(111,*) "... loading default run"
      WRITE(unit=111,nml=read_parameters)
    END IF

    INQUIRE(file='input/ibfs_gpu.inp',exist=readgpu)
    IF (readgpu) THEN
      OPEN(unit=3,file='input/ibfs_gpu.inp',form='formatted',status='old')
      READ(unit=3,nml=ibfs_gpu)
      CLOSE(3)
    END IF
    IF (nblock < 1) nblock = 1
    WRITE(unit=111,nml=ibfs_gpu)
  END SUBROUTINE input

END MODULE parameters
