!------------------------------------------------------------------------------
! s2dw_core_mod -- Fortran drop-in for the S2DW library core on MI355X.
!
! Same public procedures, argument order, kinds and explicit-shape bounds as
! the reference module (src/mod/s2dw_core_mod.f90, public list 29-39, type
! s2dw_wav_abg 56-58).  Each procedure forwards to the C ABI of
! libs2dw_hip.so (include/s2dw_hip.h) through ISO_C_BINDING; the transforms
! run as HIP kernels on the GPU.  Non-zero return codes are re-raised through
! s2dw_error, which prints the reference's message and halts as it does.
!
! Ownership follows the reference: the _dynamic analysis allocates
! wavdyn(0:J) and each wavdyn(jj)%coeff(0:2hi-2,0:2hi-1,0:N-1); the _dynamic
! synthesis and s2dw_core_free_wavdyn deallocate them.
!------------------------------------------------------------------------------
module s2dw_core_mod

  use, intrinsic :: iso_c_binding
  use s2dw_types_mod
  use s2dw_error_mod

  implicit none

  private

  public :: &
       s2dw_core_init_kernels, &
       s2dw_core_init_directionality, &
       s2dw_core_comp_Jmax, &
       s2dw_core_admiss, &
       s2dw_core_analysis_flm2wav, &
       s2dw_core_analysis_flm2wav_dynamic, &
       s2dw_core_synthesis_wav2flm, &
       s2dw_core_synthesis_wav2flm_dynamic, &
       s2dw_core_free_wavdyn, &
       s2dw_core_assimilate_int

  type, public :: s2dw_wav_abg
     real(dp), allocatable :: coeff(:,:,:)
  end type s2dw_wav_abg

  interface
     function c_init_kernels(K, Phi2, bl, J, B, alpha) bind(C, name='s2dw_hip_init_kernels')
       import :: c_int, c_double
       real(c_double), intent(out) :: K(*), Phi2(*)
       integer(c_int), intent(out) :: bl
       integer(c_int), value :: J, B
       real(c_double), value :: alpha
       integer(c_int) :: c_init_kernels
     end function c_init_kernels
     function c_init_directionality(S, B, N) bind(C, name='s2dw_hip_init_directionality')
       import :: c_int, c_double_complex
       complex(c_double_complex), intent(out) :: S(*)
       integer(c_int), value :: B, N
       integer(c_int) :: c_init_directionality
     end function c_init_directionality
     function c_comp_Jmax(B, alpha) bind(C, name='s2dw_hip_comp_Jmax')
       import :: c_int, c_double
       integer(c_int), value :: B
       real(c_double), value :: alpha
       integer(c_int) :: c_comp_Jmax
     end function c_comp_Jmax
     function c_admiss(a, K, Phi2, B, J, pass) bind(C, name='s2dw_hip_admiss')
       import :: c_int, c_double
       real(c_double), intent(out) :: a(*)
       real(c_double), intent(in) :: K(*), Phi2(*)
       integer(c_int), value :: B, J
       integer(c_int), intent(out) :: pass
       integer(c_int) :: c_admiss
     end function c_admiss
     function c_assimilate_int(x, y) bind(C, name='s2dw_hip_assimilate_int')
       import :: c_int, c_double
       real(c_double), value :: x
       integer(c_int), intent(out) :: y
       integer(c_int) :: c_assimilate_int
     end function c_assimilate_int
     function c_band_limits(J, B, alpha, hi, lo) bind(C, name='s2dw_hip_band_limits')
       import :: c_int, c_double
       integer(c_int), value :: J, B
       real(c_double), value :: alpha
       integer(c_int), intent(out) :: hi(*), lo(*)
       integer(c_int) :: c_band_limits
     end function c_band_limits
     function c_analysis(wav, scoeff, flm, K, S, J, B, N, bl, alpha) bind(C, name='s2dw_hip_analysis_flm2wav')
       import :: c_int, c_double, c_double_complex
       real(c_double), intent(out) :: wav(*)
       complex(c_double_complex), intent(out) :: scoeff(*)
       complex(c_double_complex), intent(in) :: flm(*), S(*)
       real(c_double), intent(in) :: K(*)
       integer(c_int), value :: J, B, N, bl
       real(c_double), value :: alpha
       integer(c_int) :: c_analysis
     end function c_analysis
     function c_analysis_dyn(ptrs, scoeff, flm, K, S, J, B, N, bl, alpha) &
          bind(C, name='s2dw_hip_analysis_flm2wav_dynamic')
       import :: c_int, c_double, c_double_complex, c_ptr
       type(c_ptr), intent(in) :: ptrs(*)
       complex(c_double_complex), intent(out) :: scoeff(*)
       complex(c_double_complex), intent(in) :: flm(*), S(*)
       real(c_double), intent(in) :: K(*)
       integer(c_int), value :: J, B, N, bl
       real(c_double), value :: alpha
       integer(c_int) :: c_analysis_dyn
     end function c_analysis_dyn
     function c_synthesis(flm, wav, scoeff, K, Phi2, S, J, B, N, bl, alpha) &
          bind(C, name='s2dw_hip_synthesis_wav2flm')
       import :: c_int, c_double, c_double_complex
       complex(c_double_complex), intent(out) :: flm(*)
       real(c_double), intent(in) :: wav(*), K(*), Phi2(*)
       complex(c_double_complex), intent(in) :: scoeff(*), S(*)
       integer(c_int), value :: J, B, N, bl
       real(c_double), value :: alpha
       integer(c_int) :: c_synthesis
     end function c_synthesis
     function c_synthesis_dyn(flm, ptrs, scoeff, K, Phi2, S, J, B, N, bl, alpha) &
          bind(C, name='s2dw_hip_synthesis_wav2flm_dynamic')
       import :: c_int, c_double, c_double_complex, c_ptr
       complex(c_double_complex), intent(out) :: flm(*)
       type(c_ptr), intent(in) :: ptrs(*)
       real(c_double), intent(in) :: K(*), Phi2(*)
       complex(c_double_complex), intent(in) :: scoeff(*), S(*)
       integer(c_int), value :: J, B, N, bl
       real(c_double), value :: alpha
       integer(c_int) :: c_synthesis_dyn
     end function c_synthesis_dyn
     function c_last_error() bind(C, name='s2dw_hip_last_error')
       import :: c_ptr
       type(c_ptr) :: c_last_error
     end function c_last_error
  end interface

contains

  ! Re-raise a code returned by the C ABI through s2dw_error.
  subroutine raise(code, proc)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: proc
    character(kind=c_char), pointer :: msg(:)
    character(len=STRING_LEN) :: detail
    type(c_ptr) :: p
    integer :: k
    if (code == 0) return
    detail = ''
    p = c_last_error()
    if (c_associated(p)) then
       call c_f_pointer(p, msg, [STRING_LEN])
       do k = 1, STRING_LEN
          if (msg(k) == c_null_char) exit
          detail(k:k) = msg(k)
       end do
    end if
    if (len_trim(detail) > 0) then
       call s2dw_error(int(code), proc, comment_add=trim(detail))
    else
       call s2dw_error(int(code), proc)
    end if
  end subroutine raise

  ! s2dw_core_init_kernels, reference s2dw_core_mod.f90:97-157
  subroutine s2dw_core_init_kernels(K_gamma, Phi2, bl_scoeff, J, B, alpha)
    integer, intent(in) :: J, B
    real(dp), intent(in) :: alpha
    real(dp), intent(out) :: K_gamma(0:J,0:B-1)
    real(dp), intent(out) :: Phi2(0:B-1)
    integer, intent(out) :: bl_scoeff
    integer(c_int) :: bl
    call raise(c_init_kernels(K_gamma, Phi2, bl, int(J, c_int), int(B, c_int), real(alpha, c_double)), &
         's2dw_core_init_kernels')
    bl_scoeff = bl
  end subroutine s2dw_core_init_kernels

  ! s2dw_core_init_directionality, reference 219-250
  subroutine s2dw_core_init_directionality(Slm, B, N)
    integer, intent(in) :: B, N
    complex(dpc), intent(out) :: Slm(0:B-1,0:N-1)
    call raise(c_init_directionality(Slm, int(B, c_int), int(N, c_int)), 's2dw_core_init_directionality')
  end subroutine s2dw_core_init_directionality

  ! s2dw_core_comp_Jmax, reference 1365-1376
  function s2dw_core_comp_Jmax(B, alpha) result(J_max)
    integer, intent(in) :: B
    real(dp), intent(in) :: alpha
    integer :: J_max
    J_max = c_comp_Jmax(int(B, c_int), real(alpha, c_double))
  end function s2dw_core_comp_Jmax

  ! s2dw_core_admiss, reference 285-304
  function s2dw_core_admiss(admiss, K_gamma, Phi2, B, J) result(pass)
    integer, intent(in) :: J, B
    real(dp), intent(in) :: K_gamma(0:J,0:B-1)
    real(dp), intent(in) :: Phi2(0:B-1)
    real(dp), intent(out) :: admiss(0:B-1)
    logical :: pass
    integer(c_int) :: ok
    call raise(c_admiss(admiss, K_gamma, Phi2, int(B, c_int), int(J, c_int), ok), 's2dw_core_admiss')
    pass = ok /= 0
  end function s2dw_core_admiss

  ! s2dw_core_assimilate_int, reference 1398-1415
  function s2dw_core_assimilate_int(x, y) result(assimilated)
    real(dp), intent(in) :: x
    integer, intent(out) :: y
    logical :: assimilated
    integer(c_int) :: yy
    assimilated = c_assimilate_int(real(x, c_double), yy) /= 0
    y = yy
  end function s2dw_core_assimilate_int

  ! s2dw_core_analysis_flm2wav, reference 342-534
  subroutine s2dw_core_analysis_flm2wav(wav, scoeff, flm, K_gamma, Slm, J, B, N, bl_scoeff, alpha)
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    real(dp), intent(out) :: wav(0:J, 0:2*B-2, 0:2*B-1, 0:N-1)
    complex(dpc), intent(out) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    complex(dpc), intent(in) :: flm(0:B-1, 0:B-1)
    real(dp), intent(in) :: K_gamma(0:J, 0:B-1)
    complex(dpc), intent(in) :: Slm(0:B-1, 0:N-1)
    call raise(c_analysis(wav, scoeff, flm, K_gamma, Slm, int(J, c_int), int(B, c_int), int(N, c_int), &
         int(bl_scoeff, c_int), real(alpha, c_double)), 's2dw_core_analysis_flm2wav')
  end subroutine s2dw_core_analysis_flm2wav

  ! s2dw_core_analysis_flm2wav_dynamic, reference 572-775
  subroutine s2dw_core_analysis_flm2wav_dynamic(wavdyn, scoeff, flm, K_gamma, Slm, J, B, N, bl_scoeff, alpha)
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    type(s2dw_wav_abg), intent(out), allocatable, target :: wavdyn(:)
    complex(dpc), intent(out) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    complex(dpc), intent(in) :: flm(0:B-1, 0:B-1)
    real(dp), intent(in) :: K_gamma(0:J, 0:B-1)
    complex(dpc), intent(in) :: Slm(0:B-1, 0:N-1)
    integer(c_int) :: hi(0:J), lo(0:J)
    type(c_ptr) :: ptrs(0:J)
    integer :: jj, fail
    call raise(c_band_limits(int(J, c_int), int(B, c_int), real(alpha, c_double), hi, lo), &
         's2dw_core_analysis_flm2wav_dynamic')
    allocate(wavdyn(0:J), stat=fail)
    if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, 's2dw_core_analysis_flm2wav_dynamic')
    do jj = 0, J
       allocate(wavdyn(jj)%coeff(0:2*hi(jj)-2, 0:2*hi(jj)-1, 0:N-1), stat=fail)
       if (fail /= 0) call s2dw_error(S2DW_ERROR_MEM_ALLOC_FAIL, 's2dw_core_analysis_flm2wav_dynamic')
       ptrs(jj) = c_loc(wavdyn(jj)%coeff(0, 0, 0))
    end do
    call raise(c_analysis_dyn(ptrs, scoeff, flm, K_gamma, Slm, int(J, c_int), int(B, c_int), int(N, c_int), &
         int(bl_scoeff, c_int), real(alpha, c_double)), 's2dw_core_analysis_flm2wav_dynamic')
  end subroutine s2dw_core_analysis_flm2wav_dynamic

  ! s2dw_core_synthesis_wav2flm, reference 816-1019
  subroutine s2dw_core_synthesis_wav2flm(flm, wav, scoeff, K_gamma, Phi2, Slm, J, B, N, bl_scoeff, alpha)
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    real(dp), intent(in) :: wav(0:J, 0:2*B-2, 0:2*B-1, 0:N-1)
    complex(dpc), intent(in) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    complex(dpc), intent(out) :: flm(0:B-1, 0:B-1)
    real(dp), intent(in) :: K_gamma(0:J, 0:B-1)
    real(dp), intent(in) :: Phi2(0:B-1)
    complex(dpc), intent(in) :: Slm(0:B-1, 0:N-1)
    call raise(c_synthesis(flm, wav, scoeff, K_gamma, Phi2, Slm, int(J, c_int), int(B, c_int), int(N, c_int), &
         int(bl_scoeff, c_int), real(alpha, c_double)), 's2dw_core_synthesis_wav2flm')
  end subroutine s2dw_core_synthesis_wav2flm

  ! s2dw_core_synthesis_wav2flm_dynamic, reference 1060-1266 (consumes wavdyn)
  subroutine s2dw_core_synthesis_wav2flm_dynamic(flm, wavdyn, scoeff, K_gamma, Phi2, Slm, J, B, N, bl_scoeff, alpha)
    integer, intent(in) :: J, B, N, bl_scoeff
    real(dp), intent(in) :: alpha
    type(s2dw_wav_abg), intent(inout), allocatable, target :: wavdyn(:)
    complex(dpc), intent(in) :: scoeff(0:bl_scoeff-1, 0:bl_scoeff-1)
    complex(dpc), intent(out) :: flm(0:B-1, 0:B-1)
    real(dp), intent(in) :: K_gamma(0:J, 0:B-1)
    real(dp), intent(in) :: Phi2(0:B-1)
    complex(dpc), intent(in) :: Slm(0:B-1, 0:N-1)
    type(c_ptr) :: ptrs(0:J)
    integer :: jj
    do jj = 0, J
       ptrs(jj) = c_loc(wavdyn(jj)%coeff(0, 0, 0))
    end do
    call raise(c_synthesis_dyn(flm, ptrs, scoeff, K_gamma, Phi2, Slm, int(J, c_int), int(B, c_int), &
         int(N, c_int), int(bl_scoeff, c_int), real(alpha, c_double)), 's2dw_core_synthesis_wav2flm_dynamic')
    call s2dw_core_free_wavdyn(wavdyn)
  end subroutine s2dw_core_synthesis_wav2flm_dynamic

  ! s2dw_core_free_wavdyn, reference 1289-1301
  subroutine s2dw_core_free_wavdyn(wavdyn)
    type(s2dw_wav_abg), intent(inout), allocatable :: wavdyn(:)
    integer :: jj
    do jj = lbound(wavdyn, 1), ubound(wavdyn, 1)
       if (allocated(wavdyn(jj)%coeff)) deallocate(wavdyn(jj)%coeff)
    end do
    deallocate(wavdyn)
  end subroutine s2dw_core_free_wavdyn

end module s2dw_core_mod
