!------------------------------------------------------------------------------
! s2dw_types_mod -- kinds, tolerances and constants of the S2DW Fortran API.
! Same public names and values as the reference module
! (src/mod/s2dw_types_mod.f90:25-78) so that host programs compile unchanged.
!------------------------------------------------------------------------------
module s2dw_types_mod

  implicit none

  private

  integer, public, parameter :: sp  = selected_real_kind(5, 30)
  integer, public, parameter :: dp_p = 12
  integer, public, parameter :: dp_r = 200
  integer, public, parameter :: dp  = selected_real_kind(dp_p, dp_r)
  integer, public, parameter :: spc = kind((1.0_sp, 1.0_sp))
  integer, public, parameter :: dpc = kind((1.0_dp, 1.0_dp))

  ! Tolerances.  TOL_ADMISS and TOL_ZERO are default-real literals in the
  ! reference, kept as such.
  real(dp), public, parameter :: TOL_ADMISS = 1e-12
  real(dp), public, parameter :: TOL_ZERO = 1e-12
  real(dp), public, parameter :: TOL_QUAD = 1d-12
  real(dp), public, parameter :: TOL_LIMIT = 1d-12
  real(dp), public, parameter :: TOL_CEIL = 1d-5

  integer, public, parameter :: STRING_LEN = 256
  real(dp), public, parameter :: PI = 3.141592653589793238462643383279502884197_dp
  real(dp), public, parameter :: PION2 = 1.570796326794896619231321691639751442099_dp
  complex(dpc), public, parameter :: I = (0.0_dp, 1.0_dp)

end module s2dw_types_mod
