!> edlat_gpu: minimal Fortran host for one sector.  Reads the hopping list in
!! the Hij_read format (GRAPH_MATRIX.f90:280-301), solves the ground state of
!! sector (Nup, Ndw) on the GPU and prints E0 and the site densities.
!! Usage: edlat_gpu HIJ_FILE NS NUP NDW ULOC [NITER] [PREC]
program edlat_gpu
  use iso_c_binding
  use ED_GPU
  implicit none
  type(c_ptr) :: ctx
  character(len=512) :: arg, hfile
  integer :: Ns, nup, ndw, niter, prec, unit, ios, i, j, a, b, spin
  real(8) :: re, im, U(1), e0
  real(8), allocatable :: Hij(:, :, :), alpha(:), beta(:), du(:), dd(:), docc(:)
  integer(c_int) :: Nsites(1), isector, nsteps
  integer(c_int64_t) :: dim
  integer(c_int32_t) :: dimup, dimdw

  call get_command_argument(1, hfile)
  call get_command_argument(2, arg); read(arg, *) Ns
  call get_command_argument(3, arg); read(arg, *) nup
  call get_command_argument(4, arg); read(arg, *) ndw
  call get_command_argument(5, arg); read(arg, *) U(1)
  niter = 200
  if (command_argument_count() >= 6) then
     call get_command_argument(6, arg); read(arg, *) niter
  endif
  prec = EDL_FP64
  if (command_argument_count() >= 7) then
     call get_command_argument(7, arg); read(arg, *) prec
  endif
  allocate(Hij(1, Ns, Ns)); Hij = 0d0
  open(newunit=unit, file=trim(hfile), status="old", action="read")
  do
     read(unit, *, iostat=ios) i, j, a, b, spin, re, im
     if (ios /= 0) exit
     Hij(spin, i, j) = re
  enddo
  close(unit)
  Nsites(1) = Ns
  call edl_check(c_null_ptr, edl_ctx_create(ctx, 0), "edl_ctx_create")
  call edl_check(ctx, edl_set_hamiltonian(ctx, Ns, 1, 1, Nsites, Hij, c_null_ptr, U, 0d0, 0d0, 0d0, 0d0, 0d0, &
       1, 0), "edl_set_hamiltonian")
  isector = edl_sector_index(Ns, nup, ndw)
  call edl_check(ctx, edl_sector_build(ctx, isector, dim, dimup, dimdw), "build_sector")
  allocate(alpha(niter), beta(niter), du(Ns), dd(Ns), docc(Ns))
  call edl_check(ctx, edl_lanczos_gs(ctx, isector, prec, niter, 0d0, int(z'5EED', c_int64_t), 1d-13, e0, &
       nsteps, alpha, beta, c_null_ptr), "sp_lanc_eigh")
  call edl_check(ctx, edl_observables(ctx, du, dd, docc), "lanc_observables")
  write(*, "(A,I6,A,I4,A,I4,A,I12)") "sector", isector, " nup", nup, " ndw", ndw, " dim", dim
  write(*, "(A,F20.12,A,I5)") "Egs =", e0, "  Lanczos steps", nsteps
  write(*, "(A,100F12.7)") "dens=", (du(i) + dd(i), i = 1, Ns)
  write(*, "(A,100F12.7)") "docc=", (docc(i), i = 1, Ns)
  call edl_check(ctx, edl_ctx_destroy(ctx), "edl_ctx_destroy")
end program edlat_gpu
