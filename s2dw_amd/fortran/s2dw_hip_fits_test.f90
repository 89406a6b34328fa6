!------------------------------------------------------------------------------
! s2dw_hip_fits_test -- the wavelet-file formats of s2dw_fileio_mod (FITS,
! and with an m prefix on the mode the MATLAB .m/.dat form).
!
! Usage: s2dw_hip_fits_test [m]write_static  file B N alpha [J]
!        s2dw_hip_fits_test [m]write_dynamic file B N alpha [J]
!        s2dw_hip_fits_test [m]read_static   file
!        s2dw_hip_fits_test [m]read_dynamic  file
! (MATLAB files are name.m with name_scoeff.dat and name_wcoeff.dat beside it.)
!
! Writing stores, for J = J_max (or the J given), the coefficients
!   wav(jj, a, b, g) = sin(0.1 a + 0.2 b + 0.3 g + jj) (inside each scale's
!   (2hi-1) x 2hi block; zero elsewhere in the static array)
!   scoeff(l, m) = (l + 0.5 m) - 0.25 l i
! with the comment 'fits test'.  Reading prints J, B, N, bl_scoeff, alpha,
! then per scale its cube sizes and the largest difference from the formula,
! and the largest difference of the scaling coefficients.
!------------------------------------------------------------------------------
program s2dw_hip_fits_test

  use s2dw_types_mod
  use s2dw_error_mod
  use s2dw_core_mod
  use s2dw_fileio_mod

  implicit none

  character(len=256) :: mode, fname, arg
  logical :: matlab
  integer :: B, N, J, bl_scoeff, jj, a, bb, g, hi
  real(dp) :: alpha, d
  real(dp), allocatable :: wav(:,:,:,:)
  complex(dpc), allocatable :: scoeff(:,:)
  type(s2dw_wav_abg), allocatable :: wavdyn(:)

  call get_command_argument(1, mode)
  call get_command_argument(2, fname)
  matlab = mode(1:1) == 'm'
  if (matlab) mode = mode(2:)
  if (mode(1:5) == 'write') then
     call get_command_argument(3, arg); read(arg,*) B
     call get_command_argument(4, arg); read(arg,*) N
     call get_command_argument(5, arg); read(arg,*) alpha
     J = s2dw_core_comp_Jmax(B, alpha)
     if (command_argument_count() >= 6) then  ! a smaller J: more scaling coefficients
        call get_command_argument(6, arg)
        read(arg,*) J
     end if
     bl_scoeff = bl_of(J + 1)
     allocate(scoeff(0:bl_scoeff-1, 0:bl_scoeff-1))
     do a = 0, bl_scoeff-1
        do bb = 0, bl_scoeff-1
           scoeff(a, bb) = cmplx(a + 0.5d0 * bb, -0.25d0 * a, dpc)
        end do
     end do
     if (trim(mode) == 'write_static') then
        allocate(wav(0:J, 0:2*B-2, 0:2*B-1, 0:N-1))
        wav = 0d0
        do jj = 0, J
           hi = bl_of(jj)
           do g = 0, N-1
              do bb = 0, 2*hi-1
                 do a = 0, 2*hi-2
                    wav(jj, a, bb, g) = value(jj, a, bb, g)
                 end do
              end do
           end do
        end do
        if (matlab) then
           call s2dw_fileio_matlab_wav_write(wav, scoeff, J, B, N, bl_scoeff, alpha, trim(fname), 'fits test')
        else
           call s2dw_fileio_fits_wav_write(wav, scoeff, J, B, N, bl_scoeff, alpha, trim(fname), 'fits test')
        end if
     else
        allocate(wavdyn(0:J))
        do jj = 0, J
           hi = bl_of(jj)
           allocate(wavdyn(jj)%coeff(0:2*hi-2, 0:2*hi-1, 0:N-1))
           do g = 0, N-1
              do bb = 0, 2*hi-1
                 do a = 0, 2*hi-2
                    wavdyn(jj)%coeff(a, bb, g) = value(jj, a, bb, g)
                 end do
              end do
           end do
        end do
        if (matlab) then
           call s2dw_fileio_matlab_wav_write(wavdyn, scoeff, J, B, N, bl_scoeff, alpha, trim(fname), 'fits test')
        else
           call s2dw_fileio_fits_wav_write(wavdyn, scoeff, J, B, N, bl_scoeff, alpha, trim(fname), 'fits test')
        end if
     end if
  else
     if (trim(mode) == 'read_static') then
        if (matlab) then
           call s2dw_fileio_matlab_wav_read(wav, scoeff, J, B, N, bl_scoeff, alpha, trim(fname))
        else
           call s2dw_fileio_fits_wav_read(wav, scoeff, J, B, N, bl_scoeff, alpha, trim(fname))
        end if
     else
        if (matlab) then
           call s2dw_fileio_matlab_wav_read(wavdyn, scoeff, J, B, N, bl_scoeff, alpha, trim(fname))
        else
           call s2dw_fileio_fits_wav_read(wavdyn, scoeff, J, B, N, bl_scoeff, alpha, trim(fname))
        end if
     end if
     write(*,'(a,4i6,f14.8)') 'header', J, B, N, bl_scoeff, alpha
     do jj = 0, J
        hi = bl_of(jj)
        d = 0d0
        do g = 0, N-1
           do bb = 0, 2*hi-1
              do a = 0, 2*hi-2
                 if (trim(mode) == 'read_static') then
                    d = max(d, abs(wav(jj, a, bb, g) - value(jj, a, bb, g)))
                 else
                    d = max(d, abs(wavdyn(jj)%coeff(a, bb, g) - value(jj, a, bb, g)))
                 end if
              end do
           end do
        end do
        if (trim(mode) == 'read_static') then
           write(*,'(a,i4,3i6,es12.3)') 'scale', jj, 2*hi-1, 2*hi, N, d
        else
           write(*,'(a,i4,3i6,es12.3)') 'scale', jj, size(wavdyn(jj)%coeff, 1), size(wavdyn(jj)%coeff, 2), &
                size(wavdyn(jj)%coeff, 3), d
        end if
     end do
     ! the MATLAB form stores scoeff(l, m) for m <= l only; the rest reads as 0
     d = 0d0
     do a = 0, bl_scoeff-1
        do bb = 0, bl_scoeff-1
           if (matlab .and. bb > a) then
              d = max(d, abs(scoeff(a, bb)))
           else
              d = max(d, abs(scoeff(a, bb) - cmplx(a + 0.5d0 * bb, -0.25d0 * a, dpc)))
           end if
        end do
     end do
     write(*,'(a,es12.3)') 'scoeff', d
  end if

contains

  real(dp) function value(jj, a, bb, g)
    integer, intent(in) :: jj, a, bb, g
    value = sin(0.1d0 * a + 0.2d0 * bb + 0.3d0 * g + jj)
  end function value

  integer function bl_of(jj) result(h)
    integer, intent(in) :: jj
    if (s2dw_core_assimilate_int(B / (alpha**(jj-1)), h)) then
       h = min(h, B)
    else
       h = min(ceiling(B / (alpha**(jj-1))), B)
    end if
  end function bl_of

end program s2dw_hip_fits_test
