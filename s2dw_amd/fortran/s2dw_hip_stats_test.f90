!------------------------------------------------------------------------------
! s2dw_hip_stats_test -- the moments of s2dw_stat_mod through the GPU library.
!
! Usage: s2dw_hip_stats_test B N alpha mode
!   mode 0: no mask; 1: a mask of zeros and ones; 2: an all-zero mask on
!   scale 0 (the routine halts as the reference does); 3: no mask, scales 1
!   and 2 constant (zero variance: one warning each, as the reference).
! Usage: s2dw_hip_stats_test write nsim J nopt out
!   s2dw_stat_moments_write into the text file out, for the values
!   v(isim, jj, k) = (-1)**k * (isim + 1.5) * 10**(jj - 3) / (k + 3)
!   (k = 0 mean .. 5 six), descriptions "set<isim>", and nopt = 0, 1, 2 of
!   the optional five, six passed (no GPU work).
!
! The coefficients of scale jj are x(k) = sin(0.37 k + jj) + 0.1 cos(1.3 k)**3
! over the column-major index k = 0, 1, ...; the mask is 1 where
! mod(k, 7) /= 3.  Prints one line per scale: jj, mean, var, skew, kur, five,
! six, and the sum of the coefficients after the call.
!------------------------------------------------------------------------------
program s2dw_hip_stats_test

  use s2dw_types_mod
  use s2dw_error_mod
  use s2dw_core_mod
  use s2dw_stat_mod

  implicit none

  character(len=64) :: arg
  integer :: B, N, J, mode, jj, k, a, bb, g, hi(0:63)
  real(dp) :: alpha
  real(dp), allocatable :: mean(:), var(:), skew(:), kur(:), five(:), six(:)
  type(s2dw_wav_abg), allocatable :: wavdyn(:), mask(:)
  logical :: ok
  real(dp), allocatable :: mv(:,:,:)
  character(len=STRING_LEN), allocatable :: desc(:)
  character(len=STRING_LEN) :: outf
  integer :: nsim, isim, nopt

  call get_command_argument(1, arg)
  if (trim(arg) == 'write') then
     call get_command_argument(2, arg); read(arg,*) nsim
     call get_command_argument(3, arg); read(arg,*) J
     call get_command_argument(4, arg); read(arg,*) nopt
     call get_command_argument(5, outf)
     allocate(mv(0:nsim-1, 0:J, 0:5), desc(0:nsim-1))
     do isim = 0, nsim-1
        write(desc(isim), '(a,i0)') 'set', isim
        do jj = 0, J
           do k = 0, 5
              mv(isim, jj, k) = (-1)**k * (isim + 1.5d0) * 10d0**(jj - 3) / (k + 3)
           end do
        end do
     end do
     if (nopt == 0) then
        call s2dw_stat_moments_write(trim(outf), nsim, desc, J, mv(:,:,0), mv(:,:,1), mv(:,:,2), mv(:,:,3))
     else if (nopt == 1) then
        call s2dw_stat_moments_write(trim(outf), nsim, desc, J, mv(:,:,0), mv(:,:,1), mv(:,:,2), mv(:,:,3), &
             mv(:,:,4))
     else
        call s2dw_stat_moments_write(trim(outf), nsim, desc, J, mv(:,:,0), mv(:,:,1), mv(:,:,2), mv(:,:,3), &
             mv(:,:,4), mv(:,:,5))
     end if
     stop
  end if
  read(arg,*) B
  call get_command_argument(2, arg); read(arg,*) N
  call get_command_argument(3, arg); read(arg,*) alpha
  call get_command_argument(4, arg); read(arg,*) mode
  J = s2dw_core_comp_Jmax(B, alpha)
  allocate(wavdyn(0:J), mask(0:J), mean(0:J), var(0:J), skew(0:J), kur(0:J), five(0:J), six(0:J))
  do jj = 0, J
     ok = s2dw_core_assimilate_int(B / (alpha**(jj-1)), hi(jj))
     if (ok) then
        hi(jj) = min(hi(jj), B)
     else
        hi(jj) = min(ceiling(B / (alpha**(jj-1))), B)
     end if
     allocate(wavdyn(jj)%coeff(0:2*hi(jj)-2, 0:2*hi(jj)-1, 0:N-1))
     allocate(mask(jj)%coeff(0:2*hi(jj)-2, 0:2*hi(jj)-1, 0:N-1))
     k = 0
     do g = 0, N-1
        do bb = 0, 2*hi(jj)-1
           do a = 0, 2*hi(jj)-2
              wavdyn(jj)%coeff(a, bb, g) = sin(0.37d0 * k + jj) + 0.1d0 * cos(1.3d0 * k)**3
              mask(jj)%coeff(a, bb, g) = merge(0d0, 1d0, mod(k, 7) == 3)
              k = k + 1
           end do
        end do
     end do
  end do
  if (mode == 2) mask(0)%coeff = 0d0
  if (mode == 3) then
     wavdyn(1)%coeff = 0.25d0
     wavdyn(2)%coeff = -1.5d0
  end if
  if (mode == 0 .or. mode == 3) then
     call s2dw_stat_moments(wavdyn, J, B, N, alpha, mean, var, skew, kur, five=five, six=six)
  else
     call s2dw_stat_moments(wavdyn, J, B, N, alpha, mean, var, skew, kur, mask, five, six)
  end if
  do jj = 0, J
     write(*,'(i4,7es26.17)') jj, mean(jj), var(jj), skew(jj), kur(jj), five(jj), six(jj), sum(wavdyn(jj)%coeff)
  end do

end program s2dw_hip_stats_test
