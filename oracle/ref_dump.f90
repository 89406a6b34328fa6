!------------------------------------------------------------------------------
! ref_dump -- driver (project code) around the reference's own Fortran modules
! s2dw_types_mod, s2dw_error_mod, s2dw_dl_mod, s2dw_core_mod and s2dw_stat_mod,
! compiled from the reference sources by oracle/Makefile into oracle/_ref/.
! TEST INFRASTRUCTURE ONLY: it writes golden vectors of the reference's setup
! routines, Wigner planes and wavelet-coefficient moments for tests/golden/.
! It never calls the reference's analysis/synthesis routines (they need FFTW,
! which is absent here).
!
! usage (all output is little-endian stream binary to <out>):
!   ref_dump kernels B alpha J out    -> J_max, bl_scoeff, pass, K_gamma, Phi2, admiss
!   ref_dump direc   B N out          -> Slm (complex, B*N)
!   ref_dump dl      l out            -> dl(-l:l,-l:l) at beta = pi/2
!   ref_dump dlbeta  l beta out       -> dl(-l:l,-l:l) at beta
!   ref_dump bands   B alpha J out    -> J_max, then (hi, lo) per scale
!   ref_dump moments B N alpha J masked in out
!       in: the coefficients of scales 0..J ((2hi-1)*2hi*N each, in order),
!       then, if masked = 1, the masks alike; out: mean, var, skew, kur, five,
!       six (J+1 each) from s2dw_stat_moments, then the coefficients after the
!       call (masked in place when masked = 1)
!   ref_dump momwrite nsim J nopt out
!       the reference's s2dw_stat_moments_write into the text file out, for
!       the values v(isim, jj, k) = (-1)**k * (isim + 1.5) * 10**(jj - 3) / (k + 3)
!       (k = 0 mean .. 5 six), descriptions "set<isim>", and nopt = 0, 1, 2
!       of the optional five, six passed
!------------------------------------------------------------------------------
program ref_dump

  use s2dw_types_mod
  use s2dw_core_mod
  use s2dw_dl_mod
  use s2dw_stat_mod

  implicit none

  character(len=256) :: mode, arg, outf
  integer :: B, J, N, l, bl_scoeff, J_max, u, jj, bl_hi, bl_lo
  real(dp) :: alpha, beta
  real(dp), allocatable :: K_gamma(:,:), Phi2(:), admiss(:), dl(:,:)
  complex(dpc), allocatable :: Slm(:,:)
  logical :: pass
  integer :: ipass, masked
  type(s2dw_wav_abg), allocatable :: wavdyn(:), wavmask(:)
  real(dp), allocatable :: mean(:), var(:), skew(:), kur(:), five(:), six(:)
  real(dp), allocatable :: mv(:,:,:)
  character(len=STRING_LEN), allocatable :: desc(:)
  integer :: nsim, isim, k, nopt

  call get_command_argument(1, mode)

  select case (trim(mode))

  case ('kernels')
     call get_command_argument(2, arg); read(arg,*) B
     call get_command_argument(3, arg); read(arg,*) alpha
     call get_command_argument(4, arg); read(arg,*) J
     call get_command_argument(5, outf)
     J_max = s2dw_core_comp_Jmax(B, alpha)
     allocate(K_gamma(0:J,0:B-1), Phi2(0:B-1), admiss(0:B-1))
     call s2dw_core_init_kernels(K_gamma, Phi2, bl_scoeff, J, B, alpha)
     pass = s2dw_core_admiss(admiss, K_gamma, Phi2, B, J)
     ipass = 0
     if (pass) ipass = 1
     open(newunit=u, file=trim(outf), access='stream', form='unformatted', status='replace')
     write(u) J_max, bl_scoeff, ipass
     write(u) K_gamma, Phi2, admiss
     close(u)

  case ('direc')
     call get_command_argument(2, arg); read(arg,*) B
     call get_command_argument(3, arg); read(arg,*) N
     call get_command_argument(4, outf)
     allocate(Slm(0:B-1,0:N-1))
     call s2dw_core_init_directionality(Slm, B, N)
     open(newunit=u, file=trim(outf), access='stream', form='unformatted', status='replace')
     write(u) Slm
     close(u)

  case ('dl', 'dlbeta')
     call get_command_argument(2, arg); read(arg,*) l
     beta = PION2
     if (trim(mode) == 'dlbeta') then
        call get_command_argument(3, arg); read(arg,*) beta
        call get_command_argument(4, outf)
     else
        call get_command_argument(3, outf)
     end if
     allocate(dl(-l:l,-l:l))
     call s2dw_dl_beta_operator(dl, beta, l)
     open(newunit=u, file=trim(outf), access='stream', form='unformatted', status='replace')
     write(u) dl
     close(u)

  case ('bands')
     call get_command_argument(2, arg); read(arg,*) B
     call get_command_argument(3, arg); read(arg,*) alpha
     call get_command_argument(4, arg); read(arg,*) J
     call get_command_argument(5, outf)
     J_max = s2dw_core_comp_Jmax(B, alpha)
     open(newunit=u, file=trim(outf), access='stream', form='unformatted', status='replace')
     write(u) J_max
     ! Band limits exactly as written in s2dw_core_analysis_flm2wav_dynamic
     ! (s2dw_core_mod.f90:606-616), through the public assimilate_int.
     do jj = 0, J
        if (s2dw_core_assimilate_int(B / (alpha**(jj-1)), bl_hi)) then
           bl_hi = min(bl_hi, B)
        else
           bl_hi = min(ceiling(B / (alpha**(jj-1))), B)
        end if
        if (s2dw_core_assimilate_int(B / (alpha**(jj+1)), bl_lo)) then
           bl_lo = max(bl_lo, 0)
        else
           bl_lo = max(floor(B / (alpha**(jj+1))), 0)
        end if
        write(u) bl_hi, bl_lo
     end do
     close(u)

  case ('moments')
     call get_command_argument(2, arg); read(arg,*) B
     call get_command_argument(3, arg); read(arg,*) N
     call get_command_argument(4, arg); read(arg,*) alpha
     call get_command_argument(5, arg); read(arg,*) J
     call get_command_argument(6, arg); read(arg,*) masked
     call get_command_argument(7, arg)
     call get_command_argument(8, outf)
     allocate(wavdyn(0:J), wavmask(0:J))
     allocate(mean(0:J), var(0:J), skew(0:J), kur(0:J), five(0:J), six(0:J))
     open(newunit=u, file=trim(arg), access='stream', form='unformatted', status='old')
     do jj = 0, J
        if (s2dw_core_assimilate_int(B / (alpha**(jj-1)), bl_hi)) then
           bl_hi = min(bl_hi, B)
        else
           bl_hi = min(ceiling(B / (alpha**(jj-1))), B)
        end if
        allocate(wavdyn(jj)%coeff(0:2*bl_hi-2, 0:2*bl_hi-1, 0:N-1))
        allocate(wavmask(jj)%coeff(0:2*bl_hi-2, 0:2*bl_hi-1, 0:N-1))
        read(u) wavdyn(jj)%coeff
     end do
     if (masked == 1) then
        do jj = 0, J
           read(u) wavmask(jj)%coeff
        end do
     end if
     close(u)
     if (masked == 1) then
        call s2dw_stat_moments(wavdyn, J, B, N, alpha, mean, var, skew, kur, wavmask, five, six)
     else
        call s2dw_stat_moments(wavdyn, J, B, N, alpha, mean, var, skew, kur, five=five, six=six)
     end if
     open(newunit=u, file=trim(outf), access='stream', form='unformatted', status='replace')
     write(u) mean, var, skew, kur, five, six
     do jj = 0, J
        write(u) wavdyn(jj)%coeff
     end do
     close(u)

  case ('momwrite')
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
     select case (nopt)
     case (0)
        call s2dw_stat_moments_write(trim(outf), nsim, desc, J, mv(:,:,0), mv(:,:,1), mv(:,:,2), mv(:,:,3))
     case (1)
        call s2dw_stat_moments_write(trim(outf), nsim, desc, J, mv(:,:,0), mv(:,:,1), mv(:,:,2), mv(:,:,3), &
             mv(:,:,4))
     case default
        call s2dw_stat_moments_write(trim(outf), nsim, desc, J, mv(:,:,0), mv(:,:,1), mv(:,:,2), mv(:,:,3), &
             mv(:,:,4), mv(:,:,5))
     end select

  case default
     write(*,'(a)') 'ref_dump: unknown mode'
     stop 1

  end select

end program ref_dump
