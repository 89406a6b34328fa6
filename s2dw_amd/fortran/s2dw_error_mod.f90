!------------------------------------------------------------------------------
! s2dw_error_mod -- error codes and the halting error routine of the S2DW
! Fortran API.  Codes, messages and default halt policy follow the reference
! (src/mod/s2dw_error_mod.f90:37-96, routine s2dw_error at 127-191); the
! library behind s2dw_core_mod returns these codes and the module re-raises
! them here.
!------------------------------------------------------------------------------
module s2dw_error_mod

  use s2dw_types_mod, only: STRING_LEN

  implicit none

  private

  public :: s2dw_error

  integer, parameter :: S2DW_ERROR_NUM = 14

  integer, public, parameter :: &
       S2DW_ERROR_NONE = 0, S2DW_ERROR_INIT = 1, S2DW_ERROR_NOT_INIT = 2, &
       S2DW_ERROR_INIT_FAIL = 3, S2DW_ERROR_MEM_ALLOC_FAIL = 4, S2DW_ERROR_ARTH = 5, &
       S2DW_ERROR_ARTH_WARNING = 6, S2DW_ERROR_SIZE_WARNING = 7, S2DW_ERROR_SIZE_INVALID = 8, &
       S2DW_ERROR_SIZE_NOT_DEF = 9, S2DW_ERROR_ARG_INVALID = 10, S2DW_ERROR_ARG_WARNING = 11, &
       S2DW_ERROR_QUAD_STEP_EXCEED = 12, S2DW_ERROR_ADMISS_FAIL = 13, S2DW_ERROR_FILEIO = 14

  character(len=STRING_LEN), parameter :: error_text(0:S2DW_ERROR_NUM) = (/ character(len=STRING_LEN) :: &
       'No error', &
       'Attempt to initialise object that has already been initialised', &
       'Object not initialised', &
       'Object initialisation failed', &
       'Memory allocation failed', &
       'Arithmetic exception', &
       'Arithmetic warning', &
       'Warning: Sizes not in recommended range', &
       'Invalid sizes', &
       'Sizes not defined', &
       'Arguments invalid', &
       'Argument warning', &
       'Exceeded number of steps limit whem computing quadrature', &
       'Admissibility test failed (resolution of identity not satisfied)', &
       'File IO error' /)

  ! Codes 0, 6, 7 and 11 are warnings; every other code halts by default.
  logical, parameter :: halts(0:S2DW_ERROR_NUM) = (/ .false., .true., .true., .true., .true., &
       .true., .false., .false., .true., .true., .true., .false., .true., .true., .true. /)

contains

  subroutine s2dw_error(error_code, procedure, comment_add, comment_out, halt_in)

    integer, intent(in) :: error_code
    character(len=*), intent(in), optional :: procedure, comment_add
    character(len=*), intent(inout), optional :: comment_out
    logical, intent(in), optional :: halt_in

    character(len=*), parameter :: prefix = 'S2DW_ERROR: '
    integer :: code
    logical :: halt

    code = error_code
    if (code < 0 .or. code > S2DW_ERROR_NUM) code = S2DW_ERROR_INIT_FAIL

    if (present(procedure)) then
       if (present(comment_add)) then
          write(*,'(a,a,a,a,a,a,a,a)') prefix, 'Error ''', trim(error_text(code)), &
               ''' occured in procedure ''', trim(procedure), '''', ' - ', trim(comment_add)
       else
          write(*,'(a,a,a,a,a,a)') prefix, 'Error ''', trim(error_text(code)), &
               ''' occured in procedure ''', trim(procedure), ''''
       end if
    else
       if (present(comment_add)) then
          write(*,'(a,a,a,a)') prefix, trim(error_text(code)), ' - ', trim(comment_add)
       else
          write(*,'(a,a)') prefix, trim(error_text(code))
       end if
    end if

    if (present(comment_out)) comment_out = error_text(code)

    halt = halts(code)
    if (present(halt_in)) halt = halt_in
    if (halt) then
       write(*,'(a,a,a,a,a)') prefix, '  Halting program execution ', &
            'due to error ''', trim(error_text(code)), ''''
       stop
    end if

  end subroutine s2dw_error

end module s2dw_error_mod
