!------------------------------------------------------------------------------
! s2dw_hip_test -- round-trip check of the Fortran API on the GPU library.
!
! Usage: s2dw_hip_test B [N] [alpha] [J_min] [tol]
!
! tol (default 1e-12) is the reconstruction error the run must stay below.
!
! Follows the steps of a typical S2DW host program: kernels, directionality,
! admissibility, dynamic analysis, dynamic synthesis, then reports the
! maximum reconstruction error of f_lm and the wall time of each transform.
! The test signal is a seeded real-signal f_lm (f_l0 real, Re/Im uniform in
! (-1,1) for 1 <= m <= l) from a 64-bit xorshift generator.
!------------------------------------------------------------------------------
program s2dw_hip_test

  use s2dw_types_mod
  use s2dw_error_mod
  use s2dw_core_mod

  implicit none

  character(len=64) :: arg
  integer :: B, N, J, J_min, bl_scoeff, el, m, nargs
  real(dp) :: alpha, err, tol
  integer(8) :: state, t0, t1, t2, rate
  complex(dpc), allocatable :: flm(:,:), flm_syn(:,:), Slm(:,:), scoeff(:,:)
  real(dp), allocatable :: K_gamma(:,:), Phi2(:), admiss(:)
  type(s2dw_wav_abg), allocatable :: wavdyn(:)

  nargs = command_argument_count()
  B = 64; N = 3; alpha = 2d0; J_min = 0; tol = 1d-12
  if (nargs >= 1) then
     call get_command_argument(1, arg); read(arg,*) B
  end if
  if (nargs >= 2) then
     call get_command_argument(2, arg); read(arg,*) N
  end if
  if (nargs >= 3) then
     call get_command_argument(3, arg); read(arg,*) alpha
  end if
  if (nargs >= 4) then
     call get_command_argument(4, arg); read(arg,*) J_min
  end if
  if (nargs >= 5) then
     call get_command_argument(5, arg); read(arg,*) tol
  end if
  J = s2dw_core_comp_Jmax(B, alpha) - J_min

  allocate(flm(0:B-1,0:B-1), flm_syn(0:B-1,0:B-1), Slm(0:B-1,0:N-1), K_gamma(0:J,0:B-1), &
       Phi2(0:B-1), admiss(0:B-1))
  state = 88172645463325252_8
  flm = cmplx(0d0, 0d0, dpc)
  do el = 0, B-1
     flm(el,0) = cmplx(2d0*uniform(state)-1d0, 0d0, dpc)
     do m = 1, el
        flm(el,m) = cmplx(2d0*uniform(state)-1d0, 2d0*uniform(state)-1d0, dpc)
     end do
  end do

  call s2dw_core_init_kernels(K_gamma, Phi2, bl_scoeff, J, B, alpha)
  call s2dw_core_init_directionality(Slm, B, N)
  if (.not. s2dw_core_admiss(admiss, K_gamma, Phi2, B, J)) then
     call s2dw_error(S2DW_ERROR_ADMISS_FAIL, 's2dw_hip_test')
  end if
  allocate(scoeff(0:bl_scoeff-1,0:bl_scoeff-1))

  call system_clock(t0, rate)
  call s2dw_core_analysis_flm2wav_dynamic(wavdyn, scoeff, flm, K_gamma, Slm, J, B, N, bl_scoeff, alpha)
  call system_clock(t1)
  call s2dw_core_synthesis_wav2flm_dynamic(flm_syn, wavdyn, scoeff, K_gamma, Phi2, Slm, J, B, N, bl_scoeff, alpha)
  call system_clock(t2)

  err = maxval(abs(flm - flm_syn))
  write(*,'(a,i6,a,i3,a,f5.2,a,i3,a,i4)') 'B=', B, ' N=', N, ' alpha=', alpha, ' J=', J, ' bl_scoeff=', bl_scoeff
  write(*,'(a,f12.6,a,f12.6,a)') 'analysis ', real(t1-t0,dp)/rate, ' s, synthesis ', real(t2-t1,dp)/rate, ' s'
  write(*,'(a,es12.4)') 'max reconstruction error = ', err
  if (allocated(wavdyn)) then
     write(*,'(a)') 'wavdyn not freed by synthesis'
     stop 2
  end if
  if (err < tol) then
     write(*,'(a)') 'Tests passed!'
  else
     write(*,'(a)') 'Tests failed!'
     stop 1
  end if

contains

  function uniform(s) result(u)
    integer(8), intent(inout) :: s
    real(dp) :: u
    s = ieor(s, ishft(s, 13))
    s = ieor(s, ishft(s, -7))
    s = ieor(s, ishft(s, 17))
    u = real(ishft(s, -11), dp) * (1d0 / 9007199254740992d0)
  end function uniform

end program s2dw_hip_test
