! p2m_iface.f90 -- iso_c_binding interface of libp2m.so (include/pks2map_gpu.h)
! for the Fortran host programs.  The derived types mirror the C structs
! field by field; every routine is a C function returning 0 or a P2M_E_* code.
module p2m_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: P2M_PREC_PLAIN_FP64 = 0
  integer(c_int), parameter :: P2M_PREC_FP64       = 1
  integer(c_int), parameter :: P2M_PREC_FP32_COMP  = 2
  integer(c_int), parameter :: P2M_FLAG_MEMBERSHIP = 1
  integer(c_int), parameter :: P2M_RUN_PRECUT      = 1

  type, bind(C) :: p2m_config
     integer(c_int32_t) :: nside
     integer(c_int32_t) :: npix_fs
     real(c_double)     :: fov_deg
     integer(c_int32_t) :: model
     integer(c_int32_t) :: profile
     integer(c_int32_t) :: n_maps
     integer(c_int32_t) :: precision
     integer(c_int64_t) :: chunk_halos
     integer(c_int32_t) :: device
     integer(c_int32_t) :: flags
  end type p2m_config

  type, bind(C) :: p2m_run
     real(c_double)     :: omegam, omegal, h
     real(c_double)     :: zmin, zmax
     real(c_double)     :: chihview
     real(c_double)     :: mmin
     integer(c_int32_t) :: flags
  end type p2m_run

  type, bind(C) :: p2m_stats
     integer(c_int64_t) :: n_in, n_cut, n_painted
     integer(c_int64_t) :: n_listed_hp, n_upd_hp, n_upd_fs
     integer(c_int64_t) :: n_fallback, n_segments, n_ring_gaps, n_radius_err
     real(c_double)     :: cut_low, cut_high
     real(c_double)     :: ms_total, ms_decode, ms_bin, ms_paint, ms_finish
     integer(c_int32_t) :: tile_w, tile_wg
     real(c_double)     :: ms_tile
     integer(c_int64_t) :: h2d_bytes
     real(c_double)     :: ms_h2d
     integer(c_int32_t) :: h2d_pinned, n_redo_tiles
  end type p2m_stats

  interface
     integer(c_int) function p2m_create(ctx, cfg) bind(C, name="p2m_create")
       import :: c_int, c_ptr, p2m_config
       type(c_ptr), intent(out) :: ctx
       type(p2m_config), intent(in) :: cfg
     end function p2m_create

     subroutine p2m_destroy(ctx) bind(C, name="p2m_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine p2m_destroy

     type(c_ptr) function p2m_strerror_c(code) bind(C, name="p2m_strerror")
       import :: c_ptr, c_int
       integer(c_int), value :: code
     end function p2m_strerror_c

     integer(c_int) function p2m_set_run(ctx, run) bind(C, name="p2m_set_run")
       import :: c_int, c_ptr, p2m_run
       type(c_ptr), value :: ctx
       type(p2m_run), intent(in) :: run
     end function p2m_set_run

     integer(c_int) function p2m_rofz(ctx, z, chi) bind(C, name="p2m_rofz")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: z
       real(c_double), intent(out) :: chi
     end function p2m_rofz

     integer(c_int) function p2m_load_table(ctx, slot, hdr_i, hdr_f, table, n_table) bind(C, name="p2m_load_table")
       import :: c_int, c_ptr, c_int32_t, c_float, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: slot
       integer(c_int32_t), intent(in) :: hdr_i(3)
       real(c_float), intent(in) :: hdr_f(6)
       real(c_float), intent(in) :: table(*)
       integer(c_int64_t), value :: n_table
     end function p2m_load_table

     integer(c_int) function p2m_paint(ctx, halos, n, index_base) bind(C, name="p2m_paint")
       import :: c_int, c_ptr, c_float, c_int64_t
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: halos(*)
       integer(c_int64_t), value :: n, index_base
     end function p2m_paint

     integer(c_int) function p2m_host_alloc(ctx, bytes, ptr) bind(C, name="p2m_host_alloc")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: bytes
       type(c_ptr), intent(out) :: ptr
     end function p2m_host_alloc

     integer(c_int) function p2m_host_free(ctx, ptr) bind(C, name="p2m_host_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, ptr
     end function p2m_host_free

     integer(c_int) function p2m_stream_begin(ctx, n_total) bind(C, name="p2m_stream_begin")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n_total
     end function p2m_stream_begin

     integer(c_int) function p2m_stream_push(ctx, halos, n, index_base) bind(C, name="p2m_stream_push")
       import :: c_int, c_ptr, c_float, c_int64_t
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: halos(*)
       integer(c_int64_t), value :: n, index_base
     end function p2m_stream_push

     integer(c_int) function p2m_stream_finish(ctx) bind(C, name="p2m_stream_finish")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function p2m_stream_finish

     integer(c_int) function p2m_get_stats(ctx, st) bind(C, name="p2m_get_stats")
       import :: c_int, c_ptr, p2m_stats
       type(c_ptr), value :: ctx
       type(p2m_stats), intent(out) :: st
     end function p2m_get_stats

     integer(c_int) function p2m_products(ctx, slot, moments, rms, dpdy) bind(C, name="p2m_products")
       import :: c_int, c_ptr, c_double, c_float
       type(c_ptr), value :: ctx
       integer(c_int), value :: slot
       real(c_double), intent(out) :: moments(6), rms(2)
       real(c_float), intent(out) :: dpdy(10000)
     end function p2m_products

     integer(c_int) function p2m_write_outputs(ctx, slot, base, wfs, whp, wdp) bind(C, name="p2m_write_outputs")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       integer(c_int), value :: slot
       character(kind=c_char), intent(in) :: base(*)
       integer(c_int), value :: wfs, whp, wdp
     end function p2m_write_outputs

     integer(c_int) function p2m_make_maptable(device, model, nu, alpha, nchit, nmht, nrt, hdr_i, hdr_f, table, &
          n_table) bind(C, name="p2m_make_maptable")
       import :: c_int, c_int32_t, c_double, c_float, c_int64_t
       integer(c_int32_t), value :: device, model
       real(c_double), value :: nu, alpha
       integer(c_int32_t), value :: nchit, nmht, nrt
       integer(c_int32_t), intent(out) :: hdr_i(3)
       real(c_float), intent(out) :: hdr_f(6)
       real(c_float), intent(out) :: table(*)
       integer(c_int64_t), value :: n_table
     end function p2m_make_maptable
  end interface

contains

  ! message of a P2M_E_* code as a Fortran string
  function p2m_strerror(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = p2m_strerror_c(code)
    call c_f_pointer(p, s, [512])
    n = 0
    do while (n < 512)
       if (s(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function p2m_strerror

end module p2m_iface
