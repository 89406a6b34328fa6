!------------------------------------------------------------------------------
! s2dw_fileio_mod -- the S2DW wavelet-coefficient file format (FITS) for
! programs that use this library in place of the reference.
!
! Public interface as the reference's module (src/mod/s2dw_fileio_mod.f90):
! the generics s2dw_fileio_fits_wav_write / _read and
! s2dw_fileio_matlab_wav_write / _read, each for statically allocated
! coefficients wav(0:J, 0:2B-2, 0:2B-1, 0:N-1) and for the per-scale type
! s2dw_wav_abg.  The file layout is the reference's, so files
! are interchangeable both ways:
!   HDU 1  empty primary image; comment cards, DATE, the optional comment and
!          the keys J, B, N, BLSCOEFF (integers) and ALPHA (8 decimals)
!          (reference 845-862)
!   HDU 2  SCOEFF_REAL, bl_scoeff x bl_scoeff FP64 image (871-875)
!   HDU 3  SCOEFF_IMAG, likewise (876-879)
!   HDU 4+jj  WAV, the (2hi-1) x 2hi x N FP64 cube of scale jj (900-915)
! Writing refuses an existing file (S2DW_ERROR_FILEIO, then stop), reading
! checks the number of HDUs, the scaling-coefficient sizes and for null
! values, as the reference does (1140-1240).  I/O goes through CFITSIO's
! Fortran interface (link with -lcfitsio).
!
! The MATLAB form (reference 102-763): a text file name.m, MATLAB code that
! loads name_scoeff.dat (Re, then Im, of scoeff(l, m) for m <= l, one e27.20
! value per line) and name_wcoeff.dat (every scale's (2hi-1) x 2hi x N cube,
! first index fastest), preceded by a 19-line header comment and the lines
! 'J = ...', 'N = ...', 'B = ...', 'bl_scoeff = ...', 'alpha = ...' and the
! cube sizes that the readers parse.  Writing opens all three files as new
! (an existing one is a Fortran I/O error); reading needs the layout exactly
! as written.
!------------------------------------------------------------------------------
module s2dw_fileio_mod

  use s2dw_types_mod
  use s2dw_error_mod
  use s2dw_core_mod

  implicit none

  private

  public :: &
       s2dw_fileio_matlab_wav_write, &
       s2dw_fileio_matlab_wav_read, &
       s2dw_fileio_fits_wav_write, &
       s2dw_fileio_fits_wav_read

  interface s2dw_fileio_matlab_wav_write
     module procedure matlab_write_static, matlab_write_dynamic
  end interface s2dw_fileio_matlab_wav_write

  interface s2dw_fileio_matlab_wav_read
     module procedure matlab_read_static, matlab_read_dynamic
  end interface s2dw_fileio_matlab_wav_read

  interface s2dw_fileio_fits_wav_write
     module procedure fits_write_static, fits_write_dynamic
  end interface s2dw_fileio_fits_wav_write

  interface s2dw_fileio_fits_wav_read
     module procedure fits_read_static, fits_read_dynamic
  end interface s2dw_fileio_fits_wav_read

  integer, parameter :: BITPIX_DP = -64     ! FP64 images
  integer, parameter :: ALPHA_DECIMALS = 8  ! digits of ALPHA in the header
  integer, parameter :: FITS_NO_FILE = 104  ! CFITSIO: could not open the named file

contains

  ! Band limit bl_hi of scale jj, as the transform computes it
  ! (s2dw_core_mod.f90:606-611).
  integer function scale_hi(jj, B, alpha) result(hi)
    integer, intent(in) :: jj, B
    real(dp), intent(in) :: alpha
    if (s2dw_core_assimilate_int(B / (alpha**(jj-1)), hi)) then
       hi = min(hi, B)
    else
       hi = min(ceiling(B / (alpha**(jj-1))), B)
    end if
  end function scale_hi

  ! Reports a CFITSIO failure with its messages and halts when asked.
  subroutine fits_fail(status, halt)
    integer, intent(inout) :: status
    logical, intent(in) :: halt
    character(len=30) :: text
    character(len=80) :: msg
    if (status <= 0) return
    call ftgerr(status, text)
    print *, 'FITSIO Error Status =', status, ': ', text
    do
       call ftgmsg(msg)
       if (msg == ' ') exit
       write(*,*) trim(msg)
    end do
    if (halt) stop
  end subroutine fits_fail

  ! True when filename can be opened (or exists but fails to open).
  logical function fits_present(filename) result(exists)
    character(len=*), intent(in) :: filename
    integer :: unit, status, blocksize
    status = 0
    call ftgiou(unit, status)
    call ftopen(unit, filename, 0, blocksize, status)
    if (status == 0) then
       exists = .true.
       call ftclos(unit, status)
    else if (status == FITS_NO_FILE) then
       exists = .false.
       status = 0
       call ftcmsg()
    else
       exists = .true.
       call fits_fail(status, .false.)
       status = 0
       call ftclos(unit, status)
       status = 0
    end if
    call ftfiou(unit, status)
  end function fits_present

  ! Creates filename (which must not exist) with the primary header and the
  ! scaling coefficients; returns the open unit.
  subroutine fits_create(filename, proc, J, B, N, bl_scoeff, alpha, scoeff, comment, unit, status)
    character(len=*), intent(in) :: filename, proc
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    complex(dpc), intent(in) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    character(len=*), intent(in), optional :: comment
    integer, intent(out) :: unit
    integer, intent(inout) :: status
    integer :: naxes(2)
    real(dp), allocatable :: part(:,:)
    if (fits_present(filename)) then
       call s2dw_error(S2DW_ERROR_FILEIO, proc, comment_add='File already exists')
       stop
    end if
    call ftgiou(unit, status)
    call ftinit(unit, filename, 1, status)
    naxes = 0
    call ftphpr(unit, .true., BITPIX_DP, 0, naxes, 0, 1, .true., status)
    call ftpcom(unit, '  Scale discretised wavelet coefficients created by S2DW code', status)
    call ftpcom(unit, '  Written by Jason McEwen (mcewen@mrao.cam.ac.uk)', status)
    call ftpcom(unit, '  Primary extension empty.', status)
    call ftpdat(unit, status)
    if (present(comment)) call ftpcom(unit, comment, status)
    call ftpkyj(unit, 'J', J, 'No. of scales', status)
    call ftpkyj(unit, 'B', B, 'Band limit', status)
    call ftpkyj(unit, 'N', N, 'Azimuthal band limit', status)
    call ftpkyj(unit, 'BLSCOEFF', bl_scoeff, 'Scaling coefficient band limit', status)
    call ftpkyd(unit, 'ALPHA', alpha, ALPHA_DECIMALS, 'Alpha', status)
    allocate(part(0:bl_scoeff-1, 0:bl_scoeff-1))
    naxes = bl_scoeff
    part = real(scoeff, dp)
    call ftiimg(unit, BITPIX_DP, 2, naxes, status)
    call ftpkys(unit, 'EXTNAME', 'SCOEFF_REAL', 'entension name', status)
    call ftp2dd(unit, 1, bl_scoeff, bl_scoeff, bl_scoeff, part, status)
    part = aimag(scoeff)
    call ftiimg(unit, BITPIX_DP, 2, naxes, status)
    call ftpkys(unit, 'EXTNAME', 'SCOEFF_IMAG', 'entension name', status)
    call ftp2dd(unit, 1, bl_scoeff, bl_scoeff, bl_scoeff, part, status)
  end subroutine fits_create

  ! Appends one WAV cube (n1 x n2 x n3, contiguous).
  subroutine fits_put_cube(unit, cube, n1, n2, n3, status)
    integer, intent(in) :: unit, n1, n2, n3
    real(dp), intent(in) :: cube(n1, n2, n3)
    integer, intent(inout) :: status
    integer :: naxes(3)
    naxes = [n1, n2, n3]
    call ftiimg(unit, BITPIX_DP, 3, naxes, status)
    call ftpkys(unit, 'EXTNAME', 'WAV', 'entension name', status)
    call ftp3dd(unit, 1, n1, n2, n1, n2, n3, cube, status)
  end subroutine fits_put_cube

  subroutine fits_finish(unit, status)
    integer, intent(in) :: unit
    integer, intent(inout) :: status
    call ftclos(unit, status)
    call ftfiou(unit, status)
    if (status > 0) call fits_fail(status, .true.)
  end subroutine fits_finish

  ! s2dw_fileio_fits_wav_write_static, reference 796-926.
  subroutine fits_write_static(wav, scoeff, J, B, N, bl_scoeff, alpha, filename, comment)
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    real(dp), intent(in) :: wav(0:J, 0:2*B-2, 0:2*B-1, 0:N-1)
    complex(dpc), intent(in) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    character(len=*), intent(in) :: filename
    character(len=*), intent(in), optional :: comment
    integer :: unit, status, jj, hi
    real(dp), allocatable :: cube(:,:,:)
    status = 0
    call fits_create(filename, 's2dw_fileio_fits_wav_write', J, B, N, bl_scoeff, alpha, scoeff, comment, unit, status)
    do jj = 0, J
       hi = scale_hi(jj, B, alpha)
       cube = wav(jj, 0:2*hi-2, 0:2*hi-1, 0:N-1)
       call fits_put_cube(unit, cube, 2*hi-1, 2*hi, N, status)
    end do
    call fits_finish(unit, status)
  end subroutine fits_write_static

  ! s2dw_fileio_fits_wav_write_dynamic, reference 956-1077.
  subroutine fits_write_dynamic(wavdyn, scoeff, J, B, N, bl_scoeff, alpha, filename, comment)
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    type(s2dw_wav_abg), intent(in), allocatable :: wavdyn(:)
    complex(dpc), intent(in) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    character(len=*), intent(in) :: filename
    character(len=*), intent(in), optional :: comment
    integer :: unit, status, jj
    status = 0
    call fits_create(filename, 's2dw_fileio_fits_wav_write_dynamic', J, B, N, bl_scoeff, alpha, scoeff, comment, &
         unit, status)
    do jj = 0, J
       call fits_put_cube(unit, wavdyn(jj)%coeff, size(wavdyn(jj)%coeff, 1), size(wavdyn(jj)%coeff, 2), &
            size(wavdyn(jj)%coeff, 3), status)
    end do
    call fits_finish(unit, status)
  end subroutine fits_write_dynamic

  ! Opens filename, reads the primary keys and the scaling coefficients and
  ! checks the HDU count; leaves the unit at HDU 3.
  subroutine fits_open_read(filename, proc, J, B, N, bl_scoeff, alpha, scoeff, unit, status)
    character(len=*), intent(in) :: filename, proc
    integer, intent(out) :: J, B, N, bl_scoeff
    real(dp), intent(out) :: alpha
    complex(dpc), allocatable, intent(out) :: scoeff(:,:)
    integer, intent(out) :: unit
    integer, intent(inout) :: status
    character(len=20) :: note
    integer :: blocksize, nhdu_file, hdutype, n1, n2, part, fail
    logical :: anynull
    real(dp), allocatable :: re(:,:)
    if (.not. fits_present(filename)) &
         call s2dw_error(S2DW_ERROR_FILEIO, proc, comment_add='File does not exist')
    call ftgiou(unit, status)
    call ftopen(unit, filename, 0, blocksize, status)
    call ftgkyj(unit, 'J', J, note, status)
    call ftgkyj(unit, 'B', B, note, status)
    call ftgkyj(unit, 'N', N, note, status)
    call ftgkyj(unit, 'BLSCOEFF', bl_scoeff, note, status)
    call ftgkyd(unit, 'ALPHA', alpha, note, status)
    ! primary, two scaling-coefficient images, one cube per scale
    call ftthdu(unit, nhdu_file, status)
    if (nhdu_file /= 3 + (J + 1)) call s2dw_error(S2DW_ERROR_FILEIO, proc, comment_add='Invalid number of headers')
    fail = 0
    allocate(scoeff(0:bl_scoeff-1, 0:bl_scoeff-1), re(0:bl_scoeff-1, 0:bl_scoeff-1), stat=fail)
    if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, proc)
    scoeff = cmplx(0d0, 0d0, dpc)
    do part = 0, 1
       call ftmahd(unit, 2 + part, hdutype, status)
       call ftgkyj(unit, 'NAXIS1', n1, note, status)
       call ftgkyj(unit, 'NAXIS2', n2, note, status)
       if (n1 /= bl_scoeff .or. n2 /= bl_scoeff) &
            call s2dw_error(S2DW_ERROR_FILEIO, proc, comment_add='Inconsistent scaling coefficients sizes in file')
       call ftg2dd(unit, 1, -999d0, bl_scoeff, bl_scoeff, bl_scoeff, re, anynull, status)
       if (part == 0) then
          scoeff = cmplx(re, 0d0, dpc)
       else
          scoeff = scoeff + I * re
       end if
    end do
  end subroutine fits_open_read

  ! Reads the cube of HDU 4 + jj into cube (allocated to its NAXIS sizes).
  subroutine fits_get_cube(unit, jj, proc, cube, status)
    integer, intent(in) :: unit, jj
    character(len=*), intent(in) :: proc
    real(dp), allocatable, intent(out) :: cube(:,:,:)
    integer, intent(inout) :: status
    character(len=20) :: note
    integer :: hdutype, n1, n2, n3, fail
    logical :: anynull
    call ftmahd(unit, 4 + jj, hdutype, status)
    call ftgkyj(unit, 'NAXIS1', n1, note, status)
    call ftgkyj(unit, 'NAXIS2', n2, note, status)
    call ftgkyj(unit, 'NAXIS3', n3, note, status)
    fail = 0
    allocate(cube(0:n1-1, 0:n2-1, 0:n3-1), stat=fail)
    if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, proc)
    cube = 0d0
    anynull = .false.
    call ftg3dd(unit, 1, -999d0, n1, n2, n1, n2, n3, cube, anynull, status)
    if (anynull) call s2dw_error(S2DW_ERROR_FILEIO, proc, comment_add='Null wavelet coefficients read from file')
  end subroutine fits_get_cube

  ! s2dw_fileio_fits_wav_read_static, reference 1109-1253: wav allocated
  ! (0:J, 0:2B-2, 0:2B-1, 0:N-1), zero outside each scale's block.
  subroutine fits_read_static(wav, scoeff, J, B, N, bl_scoeff, alpha, filename)
    integer, intent(out) :: J, B, N, bl_scoeff
    real(dp), intent(out) :: alpha
    real(dp), allocatable, intent(out) :: wav(:,:,:,:)
    complex(dpc), allocatable, intent(out) :: scoeff(:,:)
    character(len=*), intent(in) :: filename
    integer :: unit, status, jj, fail
    real(dp), allocatable :: cube(:,:,:)
    status = 0
    call fits_open_read(filename, 's2dw_fileio_fits_wav_read', J, B, N, bl_scoeff, alpha, scoeff, unit, status)
    fail = 0
    allocate(wav(0:J, 0:2*B-2, 0:2*B-1, 0:N-1), stat=fail)
    if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, 's2dw_fileio_fits_wav_read')
    wav = 0d0
    do jj = 0, J
       call fits_get_cube(unit, jj, 's2dw_fileio_fits_wav_read', cube, status)
       wav(jj, 0:size(cube, 1)-1, 0:size(cube, 2)-1, 0:size(cube, 3)-1) = cube
    end do
    call fits_finish(unit, status)
  end subroutine fits_read_static

  ! s2dw_fileio_fits_wav_read_dynamic, reference 1286-1437.
  subroutine fits_read_dynamic(wavdyn, scoeff, J, B, N, bl_scoeff, alpha, filename)
    integer, intent(out) :: J, B, N, bl_scoeff
    real(dp), intent(out) :: alpha
    type(s2dw_wav_abg), intent(out), allocatable :: wavdyn(:)
    complex(dpc), allocatable, intent(out) :: scoeff(:,:)
    character(len=*), intent(in) :: filename
    integer :: unit, status, jj, fail
    status = 0
    call fits_open_read(filename, 's2dw_fileio_fits_wav_read', J, B, N, bl_scoeff, alpha, scoeff, unit, status)
    fail = 0
    allocate(wavdyn(0:J), stat=fail)
    if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, 's2dw_fileio_fits_wav_read_dynamic')
    do jj = 0, J
       call fits_get_cube(unit, jj, 's2dw_fileio_fits_wav_read_dynamic', wavdyn(jj)%coeff, status)
    end do
    call fits_finish(unit, status)
  end subroutine fits_read_dynamic

  !----------------------------------------------------------------------------
  ! MATLAB .m / .dat form
  !----------------------------------------------------------------------------

  ! name.m without its ".m"
  function mat_base(filename) result(base)
    character(len=*), intent(in) :: filename
    character(len=STRING_LEN) :: base
    base = filename(1:len_trim(filename) - 2)
  end function mat_base

  ! Header, parameters and cube sizes of name.m (the lines the readers count
  ! and parse); sa, sb, sg: per-scale sizes.
  subroutine mat_write_head(u, filename, J, N, B, bl_scoeff, alpha, sa, sb, sg, comment)
    integer, intent(in) :: u, J, N, B, bl_scoeff
    character(len=*), intent(in) :: filename
    real(dp), intent(in) :: alpha
    integer, intent(in) :: sa(0:J), sb(0:J), sg(0:J)
    character(len=*), intent(in), optional :: comment
    character(len=*), parameter :: sig = 'function [J, N, B, bl_scoeff, alpha, scoeff, wcoeff] = '
    integer :: jj
    write(u, '(a,a)') sig, trim(mat_base(filename))
    write(u, '(a,a)') '% ' // sig, trim(mat_base(filename))
    write(u, '(a)') '%'
    write(u, '(a)') '%  Scale discretised wavelet coefficients created by S2DW code'
    write(u, '(a)') '%  S2DW code written by Jason McEwen (mcewen@mrao.cam.ac.uk)'
    write(u, '(a)') '%'
    write(u, '(a)') '%  This file written by: Fortran'
    if (present(comment)) then
       write(u, '(a,a)') '%  Comment: ', trim(comment)
    else
       write(u, '(a)') '%  Comment: '
    end if
    write(u, '(a)') '%'
    write(u, '(a)') '%  Variables:'
    write(u, '(a)') '%   J: Maximum analysis scale depth.'
    write(u, '(a)') '%   B: Harmonic band limit.'
    write(u, '(a)') '%   N: Azimuthal band limit.'
    write(u, '(a)') '%   bl_scoeff: Upper band limit for scaling coefficients.'
    write(u, '(a)') '%   alpha: Basis dilation factor.'
    write(u, '(a)') '%   scoeff(1:bl_scoeff,1:bl_scoeff): Scaling coefficients.'
    write(u, '(a)') '%   wcoeff{1:J+1}(aa,bb,gg): Wavelet coefficients (ranges of aa, bb and gg, depend on scale).'
    write(u, '(a)')
    write(u, '(a)') '% Size parameters'
    write(u, '(a,i27,a)') 'J         = ', J, ';'
    write(u, '(a,i27,a)') 'N         = ', N, ';'
    write(u, '(a,i27,a)') 'B         = ', B, ';'
    write(u, '(a,i27,a)') 'bl_scoeff = ', bl_scoeff, ';'
    write(u, '(a,e27.20,a)') 'alpha     = ', alpha, ';'
    write(u, '(a)')
    write(u, '(a)') '% Wavelet coefficient array sizes'
    do jj = 0, J
       write(u, '(a,i5,a,i15,a)') 'wcoeff_size_aa(', jj + 1, ') = ', sa(jj), ';'
       write(u, '(a,i5,a,i15,a)') 'wcoeff_size_bb(', jj + 1, ') = ', sb(jj), ';'
       write(u, '(a,i5,a,i15,a)') 'wcoeff_size_gg(', jj + 1, ') = ', sg(jj), ';'
    end do
    write(u, '(a)')
  end subroutine mat_write_head

  ! name_scoeff.dat and the .m lines that load it.
  subroutine mat_write_scoeff(u, filename, scoeff, bl_scoeff)
    integer, intent(in) :: u, bl_scoeff
    character(len=*), intent(in) :: filename
    complex(dpc), intent(in) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    character(len=STRING_LEN) :: fdat
    integer :: ud, el, m, part
    write(fdat, '(a,a)') trim(mat_base(filename)), '_scoeff.dat'
    open(newunit=ud, file=fdat, status='new', action='write', form='formatted')
    do part = 0, 1
       do el = 0, bl_scoeff - 1
          do m = 0, el
             if (part == 0) then
                write(ud, '(e27.20)') real(scoeff(el, m), dp)
             else
                write(ud, '(e27.20)') real(aimag(scoeff(el, m)), dp)
             end if
          end do
       end do
    end do
    close(ud)
    write(u, '(a)') '% Scaling coefficients'
    write(u, '(a,a,a)') 'scoeff_array = load(''', trim(fdat), ''');'
    write(u, '(a)') 'n = length(scoeff_array) / 2;'
    write(u, '(a)') 'scoeff_array = scoeff_array(1:n) + i*scoeff_array(n+1:end);'
    write(u, '(a)') 'scoeff = zeros(bl_scoeff,bl_scoeff);'
    write(u, '(a)') 'ind_start = 1;'
    write(u, '(a)') 'for el = 0:bl_scoeff-1'
    write(u, '(a)') '    ind_end = ind_start - 1 + el + 1;'
    write(u, '(a)') '    scoeff(el+1, 1:el+1) = scoeff_array(ind_start:ind_end).'';'
    write(u, '(a)') '    ind_start = ind_end + 1;'
    write(u, '(a)') 'end'
    write(u, '(a)')
  end subroutine mat_write_scoeff

  ! One scale's cube into the open name_wcoeff.dat (first index fastest).
  subroutine mat_put_cube(ud, cube, sa, sb, sg)
    integer, intent(in) :: ud, sa, sb, sg
    real(dp), intent(in) :: cube(0:sa-1, 0:sb-1, 0:sg-1)
    integer :: aa, bb, gg
    do gg = 0, sg - 1
       do bb = 0, sb - 1
          do aa = 0, sa - 1
             write(ud, '(e27.20)') cube(aa, bb, gg)
          end do
       end do
    end do
  end subroutine mat_put_cube

  ! The .m lines that load name_wcoeff.dat.
  subroutine mat_write_wload(u, fdat)
    integer, intent(in) :: u
    character(len=*), intent(in) :: fdat
    write(u, '(a)') '% Wavelet coefficients'
    write(u, '(a,a,a)') 'wcoeff_array = load(''', trim(fdat), ''');'
    write(u, '(a)') 'ind_start = 1;'
    write(u, '(a)') 'for jj = 1:J+1'
    write(u, '(a)') '    ind_end = ind_start - 1 + wcoeff_size_aa(jj) * wcoeff_size_bb(jj) * wcoeff_size_gg(jj);'
    write(u, '(a,a)') '    wcoeff{jj} = reshape(wcoeff_array(ind_start:ind_end),', &
         'wcoeff_size_aa(jj), wcoeff_size_bb(jj), wcoeff_size_gg(jj));'
    write(u, '(a)') '    ind_start = ind_end+1;'
    write(u, '(a)') 'end'
    write(u, '(a)')
  end subroutine mat_write_wload

  subroutine matlab_write_static(wav, scoeff, J, B, N, bl_scoeff, alpha, filename, comment)
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    real(dp), intent(in) :: wav(0:J, 0:2*B-2, 0:2*B-1, 0:N-1)
    complex(dpc), intent(in) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    character(len=*), intent(in) :: filename
    character(len=*), intent(in), optional :: comment
    integer :: u, ud, jj, hi, sa(0:J), sb(0:J), sg(0:J)
    character(len=STRING_LEN) :: fdat
    do jj = 0, J
       hi = scale_hi(jj, B, alpha)
       sa(jj) = 2 * hi - 1
       sb(jj) = 2 * hi
       sg(jj) = N
    end do
    open(newunit=u, file=trim(filename), status='new', action='write', form='formatted')
    call mat_write_head(u, filename, J, N, B, bl_scoeff, alpha, sa, sb, sg, comment)
    call mat_write_scoeff(u, filename, scoeff, bl_scoeff)
    write(fdat, '(a,a)') trim(mat_base(filename)), '_wcoeff.dat'
    open(newunit=ud, file=fdat, status='new', action='write', form='formatted')
    do jj = 0, J
       call mat_put_cube(ud, wav(jj, 0:sa(jj)-1, 0:sb(jj)-1, 0:N-1), sa(jj), sb(jj), N)
    end do
    close(ud)
    call mat_write_wload(u, fdat)
    close(u)
  end subroutine matlab_write_static

  subroutine matlab_write_dynamic(wavdyn, scoeff, J, B, N, bl_scoeff, alpha, filename, comment)
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    type(s2dw_wav_abg), intent(in) :: wavdyn(0:J)
    complex(dpc), intent(in) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    character(len=*), intent(in) :: filename
    character(len=*), intent(in), optional :: comment
    integer :: u, ud, jj, sa(0:J), sb(0:J), sg(0:J)
    character(len=STRING_LEN) :: fdat
    do jj = 0, J  ! sizes from the arrays themselves, as the reference does
       sa(jj) = size(wavdyn(jj)%coeff, 1)
       sb(jj) = size(wavdyn(jj)%coeff, 2)
       sg(jj) = size(wavdyn(jj)%coeff, 3)
    end do
    open(newunit=u, file=trim(filename), status='new', action='write', form='formatted')
    call mat_write_head(u, filename, J, N, B, bl_scoeff, alpha, sa, sb, sg, comment)
    call mat_write_scoeff(u, filename, scoeff, bl_scoeff)
    write(fdat, '(a,a)') trim(mat_base(filename)), '_wcoeff.dat'
    open(newunit=ud, file=fdat, status='new', action='write', form='formatted')
    do jj = 0, J
       call mat_put_cube(ud, wavdyn(jj)%coeff, sa(jj), sb(jj), sg(jj))
    end do
    close(ud)
    call mat_write_wload(u, fdat)
    close(u)
  end subroutine matlab_write_dynamic

  ! Value after '=' and before the final ';' of a header line.
  subroutine mat_value_line(u, text)
    integer, intent(in) :: u
    character(len=STRING_LEN), intent(out) :: text
    character(len=STRING_LEN) :: line
    read(u, '(a)') line
    text = line(index(line, '=') + 1:len_trim(line) - 1)
  end subroutine mat_value_line

  ! Parameters and cube sizes from name.m.
  subroutine mat_read_head(filename, J, N, B, bl_scoeff, alpha, sa, sb, sg)
    character(len=*), intent(in) :: filename
    integer, intent(out) :: J, N, B, bl_scoeff
    real(dp), intent(out) :: alpha
    integer, allocatable, intent(out) :: sa(:), sb(:), sg(:)
    character(len=STRING_LEN) :: line
    integer :: u, i, jj
    open(newunit=u, file=trim(filename), status='old', action='read', form='formatted')
    do i = 1, 19
       read(u, '(a)') line
    end do
    call mat_value_line(u, line)
    read(line, *) J
    call mat_value_line(u, line)
    read(line, *) N
    call mat_value_line(u, line)
    read(line, *) B
    call mat_value_line(u, line)
    read(line, *) bl_scoeff
    call mat_value_line(u, line)
    read(line, *) alpha
    read(u, '(a)') line
    read(u, '(a)') line
    allocate(sa(0:J), sb(0:J), sg(0:J))
    do jj = 0, J
       call mat_value_line(u, line)
       read(line, *) sa(jj)
       call mat_value_line(u, line)
       read(line, *) sb(jj)
       call mat_value_line(u, line)
       read(line, *) sg(jj)
    end do
    close(u)
  end subroutine mat_read_head

  ! scoeff(l, m), m <= l, from name_scoeff.dat (the rest zero).
  subroutine mat_read_scoeff(filename, bl_scoeff, scoeff, proc)
    character(len=*), intent(in) :: filename, proc
    integer, intent(in) :: bl_scoeff
    complex(dpc), allocatable, intent(out) :: scoeff(:,:)
    real(dp) :: x
    character(len=STRING_LEN) :: fdat
    integer :: ud, el, m, fail
    allocate(scoeff(0:bl_scoeff-1, 0:bl_scoeff-1), stat=fail)
    if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, proc)
    scoeff = (0d0, 0d0)
    write(fdat, '(a,a)') trim(mat_base(filename)), '_scoeff.dat'
    open(newunit=ud, file=fdat, status='old', action='read', form='formatted')
    do el = 0, bl_scoeff - 1
       do m = 0, el
          read(ud, '(e27.20)') x
          scoeff(el, m) = cmplx(x, 0d0, dpc)
       end do
    end do
    do el = 0, bl_scoeff - 1
       do m = 0, el
          read(ud, '(e27.20)') x
          scoeff(el, m) = cmplx(real(scoeff(el, m), dp), x, dpc)
       end do
    end do
    close(ud)
  end subroutine mat_read_scoeff

  subroutine mat_get_cube(ud, cube, sa, sb, sg)
    integer, intent(in) :: ud, sa, sb, sg
    real(dp), intent(out) :: cube(0:sa-1, 0:sb-1, 0:sg-1)
    integer :: aa, bb, gg
    do gg = 0, sg - 1
       do bb = 0, sb - 1
          do aa = 0, sa - 1
             read(ud, '(e27.20)') cube(aa, bb, gg)
          end do
       end do
    end do
  end subroutine mat_get_cube

  subroutine matlab_read_static(wav, scoeff, J, B, N, bl_scoeff, alpha, filename)
    integer, intent(out) :: J, B, N, bl_scoeff
    real(dp), intent(out) :: alpha
    real(dp), allocatable, intent(out) :: wav(:,:,:,:)
    complex(dpc), allocatable, intent(out) :: scoeff(:,:)
    character(len=*), intent(in) :: filename
    integer, allocatable :: sa(:), sb(:), sg(:)
    real(dp), allocatable :: cube(:,:,:)
    character(len=STRING_LEN) :: fdat
    integer :: ud, jj, fail
    call mat_read_head(filename, J, N, B, bl_scoeff, alpha, sa, sb, sg)
    allocate(wav(0:J, 0:2*B-2, 0:2*B-1, 0:N-1), stat=fail)
    if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, 's2dw_fileio_matlab_wav_read_static')
    wav = 0d0
    call mat_read_scoeff(filename, bl_scoeff, scoeff, 's2dw_fileio_matlab_wav_read_static')
    write(fdat, '(a,a)') trim(mat_base(filename)), '_wcoeff.dat'
    open(newunit=ud, file=fdat, status='old', action='read', form='formatted')
    do jj = 0, J
       allocate(cube(0:sa(jj)-1, 0:sb(jj)-1, 0:sg(jj)-1))
       call mat_get_cube(ud, cube, sa(jj), sb(jj), sg(jj))
       wav(jj, 0:sa(jj)-1, 0:sb(jj)-1, 0:sg(jj)-1) = cube
       deallocate(cube)
    end do
    close(ud)
  end subroutine matlab_read_static

  subroutine matlab_read_dynamic(wavdyn, scoeff, J, B, N, bl_scoeff, alpha, filename)
    integer, intent(out) :: J, B, N, bl_scoeff
    real(dp), intent(out) :: alpha
    type(s2dw_wav_abg), allocatable, intent(out) :: wavdyn(:)
    complex(dpc), allocatable, intent(out) :: scoeff(:,:)
    character(len=*), intent(in) :: filename
    integer, allocatable :: sa(:), sb(:), sg(:)
    character(len=STRING_LEN) :: fdat
    integer :: ud, jj, fail
    call mat_read_head(filename, J, N, B, bl_scoeff, alpha, sa, sb, sg)
    allocate(wavdyn(0:J), stat=fail)
    if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, 's2dw_fileio_matlab_wav_read_dynamic')
    do jj = 0, J
       allocate(wavdyn(jj)%coeff(0:sa(jj)-1, 0:sb(jj)-1, 0:sg(jj)-1), stat=fail)
       if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, 's2dw_fileio_matlab_wav_read_dynamic')
    end do
    call mat_read_scoeff(filename, bl_scoeff, scoeff, 's2dw_fileio_matlab_wav_read_dynamic')
    write(fdat, '(a,a)') trim(mat_base(filename)), '_wcoeff.dat'
    open(newunit=ud, file=fdat, status='old', action='read', form='formatted')
    do jj = 0, J
       call mat_get_cube(ud, wavdyn(jj)%coeff, sa(jj), sb(jj), sg(jj))
    end do
    close(ud)
  end subroutine matlab_read_dynamic

end module s2dw_fileio_mod
