! A Fortran host that runs a whole case through the tandem_hip module, the
! way the reference's time loop (src/Main3D.f90:552-965) would call the
! library: it reads the case (layout, deck parameters, grids) from a stream
! file written by the test, creates a context, sets the grids up, advances
! nsteps time steps on the GPU and writes qa of every sub-block.
!
!   run_case <case file> <result file>
! case file (stream, native byte order): nb, njct, dims(3 nb), nbc(6 nb),
! ncd(6 nb), nsz(6 nb), jct(12 njct), jvel(2 njct), iparams(16) as int32;
! dparams(48) as real64; per sub-block x, y, z (real64); nsteps, exact lines
! (int32).  Result file: qa(0:lmx, 1:5) of each sub-block in order (real64).
program run_case
  use iso_c_binding, only: c_ptr, c_int32_t, c_int64_t, c_double
  use tandem_hip
  implicit none
  character(len=1024) :: fin, fout
  type(c_ptr) :: ctx
  integer(c_int32_t) :: nb, njct, cid, nsteps, exact, ip(16)
  integer(c_int32_t), allocatable :: dims(:), nbc(:), ncd(:), nsz(:), jct(:), jvel(:)
  real(c_double) :: dp(48)
  real(c_double), allocatable :: xyz(:), qa(:)
  integer :: b, n, u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) nb, njct
  allocate(dims(3 * nb), nbc(6 * nb), ncd(6 * nb), nsz(6 * nb), jct(max(1, 12 * njct)), jvel(max(1, 2 * njct)))
  read(u) dims, nbc, ncd, nsz
  if (njct > 0) read(u) jct(1:12 * njct), jvel(1:2 * njct)
  read(u) ip, dp

  call tp_check(tp_create(ctx), 'tp_create')
  call tp_check(tp_add_case(ctx, nb, dims, nbc, ncd, nsz, njct, jct, jvel, ip, dp, cid), 'tp_add_case')
  call tp_check(tp_finalize(ctx), 'tp_finalize')
  do b = 0, nb - 1
    n = (dims(3 * b + 1) + 1) * (dims(3 * b + 2) + 1) * (dims(3 * b + 3) + 1)
    allocate(xyz(3 * n))
    read(u) xyz
    call tp_check(tp_set_grid(ctx, cid, int(b, c_int32_t), xyz), 'tp_set_grid')
    deallocate(xyz)
  end do
  read(u) nsteps, exact
  close(u)

  call tp_check(tp_set_option(ctx, TP_OPT_EXACT_LINES, int(exact, c_int64_t)), 'tp_set_option')
  call tp_check(tp_setup_geometry(ctx), 'tp_setup_geometry')
  call tp_check(tp_initialo(ctx, cid), 'tp_initialo')
  call tp_check(tp_step(ctx, nsteps), 'tp_step')
  call tp_check(tp_sync(ctx), 'tp_sync')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  do b = 0, nb - 1
    n = (dims(3 * b + 1) + 1) * (dims(3 * b + 2) + 1) * (dims(3 * b + 3) + 1)
    allocate(qa(5 * n))
    call tp_check(tp_download(ctx, cid, int(b, c_int32_t), TP_QA, qa, int(5 * n, c_int64_t)), 'tp_download')
    write(u) qa
    deallocate(qa)
  end do
  close(u)
  call tp_check(tp_destroy(ctx), 'tp_destroy')
  write(*, '(a,i0,a,i0,a)') 'fortran host: ', nsteps, ' steps on ', nb, ' sub-blocks'
end program run_case
