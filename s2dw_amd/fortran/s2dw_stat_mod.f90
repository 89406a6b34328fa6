!------------------------------------------------------------------------------
! s2dw_stat_mod -- Fortran drop-in for the moments of wavelet coefficients.
!
! s2dw_stat_moments keeps the interface of the reference routine
! (src/mod/s2dw_stat_mod.f90:81-275): same argument order, kinds, bounds and
! optional arguments (wavdyn_mask, five, six), and the mask is applied to
! wavdyn in place.  The moments are computed by the GPU library
! (s2dw_hip_stat_moments in include/s2dw_hip.h).  Errors follow the
! reference: nj < 2 on a scale halts (S2DW_ERROR_SIZE_INVALID), a variance
! below TOL_ZERO reports the non-halting S2DW_ERROR_ARTH_WARNING with zero
! skewness and higher moments for that scale.  s2dw_stat_moments_write writes
! the moments in the reference's text format (one line per set of
! coefficients, reference 315-367).  The reference's s2dw_stat_histogram is
! not provided (no program calls it; its last bin reads past hist_bins).
!------------------------------------------------------------------------------
module s2dw_stat_mod

  use, intrinsic :: iso_c_binding
  use, intrinsic :: ieee_arithmetic, only: ieee_is_nan
  use s2dw_types_mod
  use s2dw_error_mod
  use s2dw_core_mod

  implicit none

  private

  public :: &
       s2dw_stat_moments, &
       s2dw_stat_moments_write

  interface
     function c_stat_moments(ptrs, J, B, N, alpha, mean, var, skew, kur, mptrs, five, six) &
          bind(C, name='s2dw_hip_stat_moments')
       import :: c_int, c_double, c_ptr
       type(c_ptr), intent(in) :: ptrs(*)
       integer(c_int), value :: J, B, N
       real(c_double), value :: alpha
       real(c_double), intent(out) :: mean(*), var(*), skew(*), kur(*)
       type(c_ptr), value :: mptrs, five, six
       integer(c_int) :: c_stat_moments
     end function c_stat_moments
  end interface

contains

  ! s2dw_stat_moments, reference 81-275
  subroutine s2dw_stat_moments(wavdyn, J, B, N, alpha, mean, var, skew, kur, wavdyn_mask, five, six)
    type(s2dw_wav_abg), intent(inout), target :: wavdyn(0:J)
    integer, intent(in) :: J
    integer, intent(in) :: B
    integer, intent(in) :: N
    real(dp), intent(in) :: alpha
    real(dp), intent(out) :: mean(0:J)
    real(dp), intent(out) :: var(0:J)
    real(dp), intent(out) :: skew(0:J)
    real(dp), intent(out) :: kur(0:J)
    real(dp), intent(out), optional, target :: five(0:J)
    real(dp), intent(out), optional, target :: six(0:J)
    type(s2dw_wav_abg), intent(in), optional, target :: wavdyn_mask(0:J)
    type(c_ptr), target :: ptrs(0:J), mptrs(0:J)
    type(c_ptr) :: mp, fp, sp
    integer(c_int) :: code
    integer :: jj
    do jj = 0, J
       ptrs(jj) = c_loc(wavdyn(jj)%coeff)
    end do
    mp = c_null_ptr
    if (present(wavdyn_mask)) then
       do jj = 0, J
          mptrs(jj) = c_loc(wavdyn_mask(jj)%coeff)
       end do
       mp = c_loc(mptrs)
    end if
    fp = c_null_ptr
    if (present(five)) fp = c_loc(five)
    sp = c_null_ptr
    if (present(six)) sp = c_loc(six)
    code = c_stat_moments(ptrs, int(J, c_int), int(B, c_int), int(N, c_int), real(alpha, c_double), &
         mean, var, skew, kur, mp, fp, sp)
    select case (code)
    case (0, S2DW_ERROR_ARTH_WARNING, S2DW_ERROR_SIZE_INVALID)
       ! as the reference's scale loop: a warning for each zero-variance scale
       ! in scale order, and the halt at the first scale with fewer than two
       ! samples (its variance comes back NaN; later scales are not computed)
       do jj = 0, J
          if (ieee_is_nan(var(jj))) exit
          if (var(jj) < TOL_ZERO) call s2dw_error(S2DW_ERROR_ARTH_WARNING, 's2dw_stat_moments', &
               'Skewness/kurtosis & higher undefined when variance zero')
       end do
       if (code == S2DW_ERROR_SIZE_INVALID) call s2dw_error(S2DW_ERROR_SIZE_INVALID, 's2dw_stat_moments', &
            comment_add='Require at least two samples to compute moments')
    case default
       call s2dw_error(int(code), 's2dw_stat_moments')
    end select
  end subroutine s2dw_stat_moments

  ! s2dw_stat_moments_write, reference 315-367: per set isim one record
  ! "description, J, " followed by mu(isim,0:J), var, skew, kur and, when
  ! present, five and six, each value as e27.20 followed by ", ".
  subroutine s2dw_stat_moments_write(filename, nsim, description, J, mu, var, skew, kur, five, six)
    character(len=*), intent(in) :: filename
    integer, intent(in) :: nsim
    character(len=STRING_LEN), intent(in) :: description(0:nsim-1)
    integer, intent(in) :: J
    real(dp), intent(in) :: mu(0:nsim-1,0:J)
    real(dp), intent(in) :: var(0:nsim-1,0:J)
    real(dp), intent(in) :: skew(0:nsim-1,0:J)
    real(dp), intent(in) :: kur(0:nsim-1,0:J)
    real(dp), optional :: five(0:nsim-1,0:J)
    real(dp), optional :: six(0:nsim-1,0:J)
    integer :: u, isim
    open(newunit=u, file=filename, form='formatted', status='replace', action='write')
    do isim = 0, nsim-1
       write(u, '(a,a)', advance='no') trim(description(isim)), ', '
       write(u, '(i4,a)', advance='no') J, ', '
       call put(mu(isim,:))
       call put(var(isim,:))
       call put(skew(isim,:))
       call put(kur(isim,:))
       if (present(five)) then
          call put(five(isim,:))
          if (present(six)) call put(six(isim,:))
       end if
       write(u, *)
    end do
    close(u)
  contains
    subroutine put(v)
      real(dp), intent(in) :: v(0:J)
      integer :: jj
      do jj = 0, J
         write(u, '(e27.20,a)', advance='no') v(jj), ', '
      end do
    end subroutine put
  end subroutine s2dw_stat_moments_write

end module s2dw_stat_mod
