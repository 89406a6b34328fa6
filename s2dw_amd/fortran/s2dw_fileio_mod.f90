!------------------------------------------------------------------------------
! s2dw_fileio_mod -- present so that host programs which `use` the module
! compile against this library.  The FITS and MATLAB wavelet-file readers and
! writers of the reference (src/mod/s2dw_fileio_mod.f90) are outside the
! transform path this library accelerates and are not provided.
!------------------------------------------------------------------------------
module s2dw_fileio_mod

  use s2dw_types_mod
  use s2dw_core_mod

  implicit none

  private

end module s2dw_fileio_mod
