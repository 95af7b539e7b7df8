! pks2map.f90 -- drop-in pks2map driver over the MI355X painter (libp2m.so).
!
! Same command line, stdout lines and output files as the reference program
! src/pks2map_src/pks2map.f90:
!   pks2map <filein> <fileout> <tablefile> [<zmin> <nside> <scramble> <center>
!           <npix> <fov> <zmax> <chihview> <model> <profile>]
! The host keeps the CLI (:62-82), the hard-coded cosmology (:118-128), the
! catalogue read (pksc.f90:22-90 for one rank), the table read
! (maptable.f90:150-172), the optional scramble/centre branches (:162-262)
! and the reporting; the halo loop (:276-303), the products (:321-357) and
! the writers (:359-381) run through the C ABI.
!
! Extensions (environment, so the CLI stays the reference's):
!   P2M_PRECISION = fp32 (default) | fp64 | plain
!   P2M_TILE      = "W,WG" tile shape of the tiled painter (default: autotuned)
! <tablefile> may be a comma-separated list: all tables are painted in one
! catalogue pass; map k is written as <fileout>_<table basename>_{fs.map,hp.fits,dp.bin}.
! Inputs are validated before the GPU is touched; any failure exits with a
! nonzero status and writes no output file.
!
! The catalogue goes to the GPU in blocks of 2^22 halos read into two
! page-locked buffers in turn (p2m_stream_push): the copy of one block runs
! while the next is read.  The scramble and centre branches need every halo
! on the host first, so with those the whole catalogue is read into one
! page-locked array and painted with p2m_paint.
program pks2map
  use iso_c_binding
  use p2m_iface
  use p2m_args
  implicit none

  character(len=1024) :: filein, fileout, tablearg, base, env
  character(len=1024), allocatable :: tables(:)
  integer :: ntab, k, ios, u, i, m, p0, p1
  real :: zmin, fov, zmax, chihview, mypi, dpix
  integer :: hpnside, scramble, centre, npix, model, profile
  integer(c_int) :: rc
  type(p2m_config) :: cfg
  type(p2m_run) :: run
  type(p2m_stats) :: st
  type(c_ptr) :: ctx
  integer(c_int32_t) :: nhalo_hdr
  integer(c_int32_t) :: hdr_i(3), hdr_i0(3)
  real(c_float) :: rthmax, hdr_f(6), hdr_f0(6)
  real(c_float), pointer :: halos(:,:)
  real(c_float), allocatable :: tabs(:,:)
  integer(c_int64_t) :: nh, ntv, fsize
  integer :: ucat
  logical :: whole
  real(c_double) :: moments(6), rms(2), chi_lo, chi_hi
  real(c_float) :: dpdy(10000)
  integer(c_int64_t) :: hpnpix

  if (command_argument_count() < 3) then
     write(*, '(a)') 'Usage: pks2map <filein> <fileout> <tablefile> [<zmin> <nside> <scramble> <center> ' // &
          '<npix> <fov> <zmax> <chihview> <model> <profile>]'
     stop
  end if

  call get_command_argument(1, filein)
  call get_command_argument(2, fileout)
  call get_command_argument(3, tablearg)
  zmin     = r4arg(4, 0.0)
  hpnside  = i4arg(5, 4096)
  scramble = i4arg(6, 0)
  centre   = i4arg(7, 0)
  npix     = i4arg(8, 16)
  fov      = r4arg(9, 10.0)
  zmax     = r4arg(10, 1.25)
  chihview = r4arg(11, 0.0)
  model    = i4arg(12, 1)
  profile  = i4arg(13, 1)

  mypi = 4. * atan(1.0)
  fov  = fov / 360 * 2 * mypi
  dpix = fov / npix
  write(*, 12) npix, npix, dpix, fov / 2 / mypi * 360.
12 format(/, /, 'Resolution:         ', i0, ' x ', i0, ' pixels (', 1pe8.2, ' radians)', /, &
          'Field of view:      ', 1pe8.2, ' degrees')
  hpnpix = -1
  if (hpnside >= 1 .and. iand(hpnside - 1, hpnside) == 0) hpnpix = 12_8 * hpnside * hpnside
  write(*, 13) hpnside, hpnpix
13 format('Healpix Nside:      ', i0, /, 'Number of pixels:   ', i0)
  if (hpnpix < 0) call fail('nside must be a power of 2', 2)

  ! ---------------------------------------------------------------- catalogue header (pksc.f90:46-48)
  open(newunit=ucat, file=trim(filein), access='stream', form='unformatted', status='old', action='read', iostat=ios)
  if (ios /= 0) call fail('cannot open catalogue '//trim(filein), 3)
  inquire(unit=ucat, size=fsize)
  read(ucat, iostat=ios) nhalo_hdr, rthmax
  if (ios /= 0 .or. nhalo_hdr < 0) call fail('bad catalogue header in '//trim(filein), 3)
  if (fsize /= 8_8 + 28_8 * int(nhalo_hdr, 8)) call fail('catalogue size does not match its header (truncated?)', 3)
  nh = nhalo_hdr

  ! ---------------------------------------------------------------- tables
  ntab = 1
  do i = 1, len_trim(tablearg)
     if (tablearg(i:i) == ',') ntab = ntab + 1
  end do
  if (ntab > 6) call fail('at most 6 tables per pass', 2)
  allocate(tables(ntab))
  p0 = 1
  k = 0
  do i = 1, len_trim(tablearg) + 1
     if (i > len_trim(tablearg) .or. tablearg(min(i, len(tablearg)):min(i, len(tablearg))) == ',') then
        k = k + 1
        tables(k) = tablearg(p0:i - 1)
        p0 = i + 1
     end if
  end do
  do k = 1, ntab
     open(newunit=u, file=trim(tables(k)), access='stream', form='unformatted', status='old', action='read', iostat=ios)
     if (ios /= 0) call fail('cannot open table '//trim(tables(k)), 3)
     inquire(unit=u, size=fsize)
     read(u, iostat=ios) hdr_i, hdr_f
     if (ios /= 0 .or. any(hdr_i < 2)) call fail('bad table header in '//trim(tables(k)), 3)
     ntv = 2_8 * hdr_i(1) * hdr_i(2) * hdr_i(3)
     if (fsize /= 36_8 + 4_8 * ntv) call fail('table size does not match its header: '//trim(tables(k)), 3)
     if (k == 1) then
        hdr_i0 = hdr_i
        hdr_f0 = hdr_f
        allocate(tabs(ntv, ntab))
     else if (any(hdr_i /= hdr_i0) .or. any(hdr_f /= hdr_f0)) then
        call fail('all tables of one pass must share the grid', 3)
     end if
     read(u, pos=37, iostat=ios) tabs(:, k)
     if (ios /= 0) call fail('table read failed', 3)
     close(u)
  end do

  ! ---------------------------------------------------------------- GPU context
  cfg%nside = hpnside
  cfg%npix_fs = npix
  cfg%fov_deg = dble(fov / 2 / mypi * 360.)
  cfg%fov_deg = dble(r4arg(9, 10.0))
  cfg%model = model
  cfg%profile = profile
  cfg%n_maps = ntab
  cfg%precision = P2M_PREC_FP32_COMP
  call get_environment_variable('P2M_PRECISION', env)
  if (trim(env) == 'fp64') cfg%precision = P2M_PREC_FP64
  if (trim(env) == 'plain') cfg%precision = P2M_PREC_PLAIN_FP64
  cfg%chunk_halos = 0
  cfg%device = 0
  cfg%flags = 0
  rc = p2m_create(ctx, cfg)
  call check(rc, 'p2m_create')

  ! pks2map.f90:118-128 hard-coded cosmology; mmin of :84
  run%omegam = 0.25d0
  run%omegal = 0.75d0
  run%h = 0.7d0
  run%zmin = dble(zmin)
  run%zmax = dble(zmax)
  run%chihview = dble(chihview)
  run%mmin = 1d13
  run%flags = 0
  call check(p2m_set_run(ctx, run), 'p2m_set_run')
  call check(p2m_rofz(ctx, dble(zmin), chi_lo), 'p2m_rofz')
  call check(p2m_rofz(ctx, dble(zmax), chi_hi), 'p2m_rofz')
  write(*, 14) chi_lo / 1e3, chi_hi / 1e3
14 format('cut_low, cut_high:  ', f4.2, 1x, f4.2, ' Gpc')

  ! the whole catalogue on the host: for the scramble and centre branches,
  ! and for the optional dump of the records as painted
  call get_environment_variable('P2M_DUMP_CATALOGUE', env)
  whole = scramble == 1 .or. centre /= 0 .or. len_trim(env) > 0
  if (whole) then
     call pinned_block(max(nh, 1_8), halos)
     if (nh > 0) then
        read(ucat, pos=9, iostat=ios) halos(:, 1:nh)
        if (ios /= 0) call fail('catalogue read failed', 3)
     end if
     close(ucat)
  end if

  if (scramble == 1 .or. centre /= 0) then
     call host_branches(halos, nh, real(chi_lo), real(chi_hi), chihview, scramble, centre)
     run%flags = P2M_RUN_PRECUT
     call check(p2m_set_run(ctx, run), 'p2m_set_run')
  end if

  ! optional: the records as painted (after cut, scramble, centring), in .pksc layout
  if (len_trim(env) > 0) then
     open(newunit=u, file=trim(env), access='stream', form='unformatted', status='replace', action='write', &
          iostat=ios)
     if (ios /= 0) call fail('cannot write '//trim(env), 3)
     if (nh > 0) then
        write(u) int(nh, 4), maxval(halos(7, 1:nh)), halos(:, 1:nh)
     else
        write(u) 0_4, 0.0
     end if
     close(u)
  end if

  call get_environment_variable('P2M_TILE', env)
  if (len_trim(env) > 0) then
     block
       integer :: tw, twg
       integer(c_int) :: p2m_set_tile_rc
       interface
          integer(c_int) function p2m_set_tile(ctx, w, wg) bind(C, name="p2m_set_tile")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: w, wg
          end function p2m_set_tile
       end interface
       read(env, *, iostat=ios) tw, twg
       if (ios == 0) then
          p2m_set_tile_rc = p2m_set_tile(ctx, int(tw, c_int32_t), int(twg, c_int32_t))
          call check(p2m_set_tile_rc, 'p2m_set_tile')
       end if
     end block
  end if

  do k = 1, ntab
     call check(p2m_load_table(ctx, int(k - 1, c_int), hdr_i0, hdr_f0, tabs(:, k), ntv), 'p2m_load_table')
  end do

  ! ---------------------------------------------------------------- halo loop on the GPU
  if (whole) then
     call check(p2m_paint(ctx, halos, nh, 0_c_int64_t), 'p2m_paint')
  else
     call stream_catalogue()
  end if
  call check(p2m_get_stats(ctx, st), 'p2m_get_stats')
  if (st%n_radius_err > 0) write(*, '(a,i0,a)') 'Warning: ', st%n_radius_err, &
       ' halos with disc radius outside [0,pi] skipped'

  write(*, 15) st%n_painted
15 format('Number of halos in map:    ', i0)

  do k = 1, ntab
     call check(p2m_products(ctx, int(k - 1, c_int), moments, rms, dpdy), 'p2m_products')
     write(*, 16) moments(1:3), moments(4:6)
16   format('Flatsky moments:    ', 3(1pe9.3, 1x), /, 'Healpix moments:    ', 3(1pe9.3, 1x))
     write(*, 17) rms(1), moments(1), rms(2), moments(4)
17   format('Flatsky RMS:                ', 1pe9.3, /, 'Flatsky Mean:               ', 1pe8.2, /, &
            'Healpix RMS:                ', 1pe9.3, /, 'Healpix Mean:               ', 1pe8.2, /)
     if (ntab == 1) then
        base = fileout
     else
        m = index(tables(k), '/', back=.true.)
        p1 = len_trim(tables(k))
        base = trim(fileout) // '_' // tables(k)(m + 1:p1)
     end if
     call check(p2m_write_outputs(ctx, int(k - 1, c_int), trim(base) // c_null_char, 1, 1, 1), 'p2m_write_outputs')
  end do
  write(*, *)

  call p2m_destroy(ctx)

contains

  ! an (7, n) array in page-locked host memory of the context
  subroutine pinned_block(n, a)
    integer(c_int64_t), intent(in) :: n
    real(c_float), pointer, intent(out) :: a(:,:)
    type(c_ptr) :: p
    call check(p2m_host_alloc(ctx, 28_c_int64_t * n, p), 'p2m_host_alloc')
    call c_f_pointer(p, a, [7_c_int64_t, n])
  end subroutine pinned_block

  ! The block read of loadpksc (pksc.f90:60-66) for one rank holding the
  ! whole catalogue, block by block into two page-locked buffers in turn:
  ! a buffer is refilled only after the next push has returned, when its
  ! copy to the device is complete.
  subroutine stream_catalogue()
    integer(c_int64_t), parameter :: blk = 4194304_c_int64_t
    real(c_float), pointer :: b1(:,:), b2(:,:), bk(:,:)
    integer(c_int64_t) :: i0, nb, nblk
    integer :: k
    nblk = min(max(nh, 1_c_int64_t), blk)
    call pinned_block(nblk, b1)
    call pinned_block(nblk, b2)
    call check(p2m_stream_begin(ctx, nh), 'p2m_stream_begin')
    i0 = 0
    k = 0
    do while (i0 < nh)
       nb = min(blk, nh - i0)
       if (mod(k, 2) == 0) then
          bk => b1
       else
          bk => b2
       end if
       read(ucat, pos=9_c_int64_t + 28_c_int64_t * i0, iostat=ios) bk(:, 1:nb)
       if (ios /= 0) call fail('catalogue read failed', 3)
       call check(p2m_stream_push(ctx, bk, nb, i0), 'p2m_stream_push')
       i0 = i0 + nb
       k = k + 1
    end do
    close(ucat)
    call check(p2m_stream_finish(ctx), 'p2m_stream_finish')
    call check(p2m_host_free(ctx, c_loc(b1)), 'p2m_host_free')
    call check(p2m_host_free(ctx, c_loc(b2)), 'p2m_host_free')
  end subroutine stream_catalogue

  subroutine fail(msg, code)
    character(len=*), intent(in) :: msg
    integer, intent(in) :: code
    write(0, '(a)') 'pks2map: ' // msg
    select case (code)
    case (2)
       error stop 2
    case (3)
       error stop 3
    case default
       error stop 4
    end select
  end subroutine fail

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
       write(0, '(a)') 'pks2map: ' // what // ': ' // p2m_strerror(code)
       error stop 4
    end if
  end subroutine check

  ! Radial cut + compaction (pks2map.f90:144-160), scramble (:162-187) and
  ! centring on the largest halo (:189-262), in default real as the reference.
  subroutine host_branches(h, n, cut_low, cut_high, chv, scr, cen)
    real(c_float), intent(inout) :: h(:, :)
    integer(c_int64_t), intent(inout) :: n
    real, intent(in) :: cut_low, cut_high, chv
    integer, intent(in) :: scr, cen
    integer(c_int64_t) :: i, j
    real :: xh, yh, zh, chih, rh, muran, phiran
    real :: rthm, xm, ym, zm, rmaxtheta, rmaxphi, xml
    integer :: seed(4)
    double precision :: randnum
    j = 0
    do i = 1, n
       xh = h(3, i) - chv
       yh = h(2, i)
       zh = h(1, i)
       chih = sqrt(xh**2 + yh**2 + zh**2)
       if (chih > cut_high .or. chih < cut_low) cycle
       j = j + 1
       h(:, j) = h(:, i)
    end do
    n = j
    if (scr == 1) then
       seed = 13579            ! every limb set to the same value (quirk Q10)
       do i = 1, n
          xh = h(1, i)
          yh = h(2, i)
          zh = h(3, i)
          chih = sqrt(xh**2 + yh**2 + zh**2)
          call ranf(seed, randnum)
          muran = 2 * randnum - 1
          call ranf(seed, randnum)
          phiran = 2 * 3.14159 * randnum
          rh = sqrt(1 - muran**2) * chih
          h(1, i) = rh * cos(phiran)
          h(2, i) = rh * sin(phiran)
          h(3, i) = muran * chih
       end do
    end if
    if (cen /= 0 .and. n > 0) then
       rthm = maxval(h(7, 1:n))
       xm = -1e5
       ym = -1e5
       zm = -1e5
       do i = 1, n
          if (h(7, i) == rthm) then
             xm = h(3, i)
             ym = h(2, i)
             zm = h(1, i)
          end if
       end do
       if (xm < 0.0) h(3, 1:n) = -h(3, 1:n)
       if (ym < 0.0) h(2, 1:n) = -h(2, 1:n)
       if (zm < 0.0) h(1, 1:n) = -h(1, 1:n)
       xm = abs(xm)
       ym = abs(ym)
       zm = abs(zm)
       rmaxtheta = asin(zm / sqrt(xm**2 + zm**2))
       xml = xm / abs(xm) * sqrt(xm**2 + zm**2)
       rmaxphi = asin(ym / sqrt(ym**2 + xml**2))
       do i = 1, n
          xh = h(3, i) * cos(rmaxtheta) + h(1, i) * sin(rmaxtheta)
          zh = -h(3, i) * sin(rmaxtheta) + h(1, i) * cos(rmaxtheta)
          yh = -xh * sin(rmaxphi) + h(2, i) * cos(rmaxphi)
          xh = xh * cos(rmaxphi) + h(2, i) * sin(rmaxphi)
          h(3, i) = zh
          h(2, i) = yh
          h(1, i) = xh
       end do
    end if
  end subroutine host_branches

  ! The 48-bit multiplicative congruential generator of ranf
  ! (RandomField/random.f90:40-57, 264-286) in closed form: the draw is the
  ! state X / 2^48, X read from the four 12-bit limbs of seed (which the
  ! scramble seed holds unreduced on the first draw, quirk Q10), and the next
  ! state is (A X) mod 2^48 for the multiplier A with limbs 373, 3707, 1442,
  ! 647, formed from 24-bit halves so every product fits 64 bits.
  subroutine ranf(seed, r)
    integer, intent(inout) :: seed(4)
    double precision, intent(out) :: r
    integer(8), parameter :: b12 = 4096_8, b24 = 16777216_8, b48 = 281474976710656_8
    integer(8), parameter :: a = 373_8 + b12 * (3707_8 + b12 * (1442_8 + b12 * 647_8))
    integer(8) :: x, xl, xh, p
    x = int(seed(1), 8) + b12 * (int(seed(2), 8) + b12 * (int(seed(3), 8) + b12 * int(seed(4), 8)))
    r = dble(x) / dble(b48)
    xl = mod(x, b24)
    xh = x / b24
    p = mod(mod(a, b24) * xl + b24 * mod(mod(a, b24) * xh + (a / b24) * xl, b24), b48)
    seed(1) = int(mod(p, b12))
    seed(2) = int(mod(p / b12, b12))
    seed(3) = int(mod(p / b24, b12))
    seed(4) = int(p / (b24 * b12))
  end subroutine ranf

end program pks2map
