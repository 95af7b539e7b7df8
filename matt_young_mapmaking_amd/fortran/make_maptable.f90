! make_maptable.f90 -- drop-in make_maptable over the GPU table kernel.
!
! Same command line and output file as the reference program
! src/pks2map_src/make_maptable.f90:
!   make_maptable <fileout> <model> [<frequency> <alpha_CMB>]
! model 1..3 selects the Battaglia et al. 2012 fit (bbps_profile.f90:9-21);
! frequency > 0 multiplies y by G(x(z)) (maptable.f90:105-123).  The table is
! computed on the GPU (p2m_make_maptable) and written with the layout of
! maptable.f90:94,142: int32 nchit,nmht,nrt; float32 bounds; float32 table.
program make_maptable
  use iso_c_binding
  use p2m_iface
  use p2m_args
  implicit none
  character(len=1024) :: fileout
  integer :: model, u, ios
  real :: nu, alpha_cmb
  integer(c_int32_t) :: hdr_i(3)
  real(c_float) :: hdr_f(6)
  real(c_float), allocatable :: table(:)
  integer(c_int32_t), parameter :: nchit = 100, nmht = 100, nrt = 200   ! maptable.f90:69-87
  integer(c_int64_t) :: nt
  integer(c_int) :: rc

  if (command_argument_count() < 3) then
     write(*, '(a)') 'Usage: make_maptable <fileout> <model> [<frequency> <alpha_CMB>]'
     stop
  end if
  call get_command_argument(1, fileout)
  model = i4arg(2, 1)
  nu = r4arg(3, 0.0)
  alpha_cmb = r4arg(4, 0.0)
  if (model < 1 .or. model > 3) then
     write(0, '(a)') 'make_maptable: model must be 1, 2 or 3'
     error stop 2
  end if

  write(*, *) ''
  write(*, *) 'Making Maptable on the GPU'
  write(*, *) ''
  nt = 2_8 * nchit * nmht * nrt
  allocate(table(nt))
  rc = p2m_make_maptable(0_c_int32_t, int(model, c_int32_t), dble(nu), dble(alpha_cmb), nchit, nmht, nrt, &
       hdr_i, hdr_f, table, nt)
  if (rc /= 0) then
     write(0, '(a)') 'make_maptable: ' // p2m_strerror(rc)
     error stop 4
  end if
  open(newunit=u, file=trim(fileout), access='stream', form='unformatted', status='replace', action='write', &
       iostat=ios)
  if (ios /= 0) then
     write(0, '(a)') 'make_maptable: cannot write ' // trim(fileout)
     error stop 3
  end if
  write(u) hdr_i, hdr_f
  write(u) table
  close(u)
end program make_maptable
