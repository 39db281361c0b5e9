"""Host-side Python mirror of the gfx950 IBFS library (ctypes over the C-ABI
in include/ibfs.h).  The compute path is the HIP library only: importing
``ibfs_amd.hip`` raises if ``libibfs_hip.so`` has not been built."""
