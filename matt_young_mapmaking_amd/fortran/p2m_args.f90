! p2m_args.f90 -- positional command-line arguments with the semantics of the
! reference's textlib (src/modules/External/textlib.f90:33-91, 171-179):
! argument `argno` converted to the requested kind; a lone '.' or a parse
! failure or a missing argument yields the default.
module p2m_args
  implicit none
contains

  integer function i4arg(argno, defval)
    integer, intent(in) :: argno, defval
    character(len=256) :: a
    integer :: ios, v
    i4arg = defval
    if (argno > command_argument_count()) return
    call get_command_argument(argno, a)
    if (trim(a) == '.') return
    read(a, *, iostat=ios) v
    if (ios == 0) i4arg = v
  end function i4arg

  ! parsed as real*8 then stored as default real, as r4arg does
  real function r4arg(argno, defval)
    integer, intent(in) :: argno
    real, intent(in) :: defval
    character(len=256) :: a
    integer :: ios
    double precision :: v
    r4arg = defval
    if (argno > command_argument_count()) return
    call get_command_argument(argno, a)
    if (trim(a) == '.') return
    read(a, *, iostat=ios) v
    if (ios == 0) r4arg = real(v)
  end function r4arg

end module p2m_args
