!> ED_GPU: iso_c_binding interfaces of libedlat_hip.so (include/edlat_gpu.h)
!! and thin wrappers that keep the reference's error semantics (a non-zero
!! status becomes `stop` with the library's message, as the reference stops
!! in HAMILTONIAN_DIRECT_HxV.f90:30,33 or GRAPH_MATRIX.f90:325).
module ED_GPU
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: EDL_FP64 = 0, EDL_FP32 = 1, EDL_MIXED = 2
  integer(c_int), parameter, public :: EDL_OP_C = 1, EDL_OP_CDG = 2, EDL_OP_SZ = 3, EDL_OP_J = 4

  type, bind(C), public :: edl_seed_desc
     integer(c_int) :: op, ispin, io, jo
  end type edl_seed_desc

  public :: edl_ctx_create, edl_ctx_destroy, edl_last_error
  public :: edl_set_hamiltonian, edl_sector_index, edl_sector_qn
  public :: edl_sector_build, edl_sector_maps, edl_sector_hops, edl_sector_rank
  public :: edl_hxv, edl_hxv_host, edl_hxv_host_complex
  public :: edl_device_alloc, edl_copy_h2d, edl_copy_d2h, edl_vec_random, edl_vec_dot
  public :: edl_lanczos_gs, edl_lanczos_stats, edl_set_state, edl_get_state
  public :: edl_tridiag, edl_tridiag_batched, edl_lanczos_poles, edl_observables, edl_local_energy
  public :: edl_correlations
  public :: edl_lanczos_states
  public :: edl_last_timing, edl_bench_hxv, edl_bench_passes
  public :: edl_check, edl_error_message

  interface
     integer(c_int) function edl_ctx_create(ctx, device) bind(C, name="edl_ctx_create")
       import :: c_int, c_ptr
       type(c_ptr) :: ctx
       integer(c_int), value :: device
     end function edl_ctx_create

     integer(c_int) function edl_ctx_destroy(ctx) bind(C, name="edl_ctx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function edl_ctx_destroy

     type(c_ptr) function edl_last_error(ctx) bind(C, name="edl_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function edl_last_error

     integer(c_int) function edl_set_hamiltonian(ctx, Ns, Nspin, Norb, Nsites, Hij_re, Hij_im, Uloc, Ust, Jh, Jx, &
          Jp, xmu, hfmode, ed_filling) bind(C, name="edl_set_hamiltonian")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: Ns, Nspin, Norb
       integer(c_int) :: Nsites(*)
       real(c_double) :: Hij_re(*)
       type(c_ptr), value :: Hij_im
       real(c_double) :: Uloc(*)
       real(c_double), value :: Ust, Jh, Jx, Jp, xmu
       integer(c_int), value :: hfmode, ed_filling
     end function edl_set_hamiltonian

     integer(c_int) function edl_sector_index(Ns, nup, ndw) bind(C, name="edl_sector_index")
       import :: c_int
       integer(c_int), value :: Ns, nup, ndw
     end function edl_sector_index

     integer(c_int) function edl_sector_qn(Ns, isector, nup, ndw) bind(C, name="edl_sector_qn")
       import :: c_int
       integer(c_int), value :: Ns, isector
       integer(c_int) :: nup, ndw
     end function edl_sector_qn

     integer(c_int) function edl_sector_build(ctx, isector, dim, dimup, dimdw) bind(C, name="edl_sector_build")
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector
       integer(c_int64_t) :: dim
       integer(c_int32_t) :: dimup, dimdw
     end function edl_sector_build

     integer(c_int) function edl_sector_maps(ctx, isector, map_up, map_dw) bind(C, name="edl_sector_maps")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector
       integer(c_int32_t) :: map_up(*), map_dw(*)
     end function edl_sector_maps

     integer(c_int) function edl_sector_hops(ctx, isector, ispin, kmax, nbr, coef) bind(C, name="edl_sector_hops")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, ispin
       integer(c_int32_t) :: kmax
       type(c_ptr), value :: nbr, coef
     end function edl_sector_hops

     integer(c_int) function edl_sector_rank(ctx, isector, ispin, n, states, pos) bind(C, name="edl_sector_rank")
       import :: c_int, c_ptr, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, ispin
       integer(c_int64_t), value :: n
       integer(c_int32_t) :: states(*), pos(*)
     end function edl_sector_rank

     integer(c_int) function edl_hxv(ctx, isector, prec, v_dev, hv_dev) bind(C, name="edl_hxv")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, v_dev, hv_dev
       integer(c_int), value :: isector, prec
     end function edl_hxv

     integer(c_int) function edl_hxv_host(ctx, isector, prec, v, hv) bind(C, name="edl_hxv_host")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, prec
       real(c_double) :: v(*), hv(*)
     end function edl_hxv_host

     integer(c_int) function edl_hxv_host_complex(ctx, isector, prec, v, hv) bind(C, name="edl_hxv_host_complex")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, prec
       complex(c_double_complex) :: v(*), hv(*)
     end function edl_hxv_host_complex

     integer(c_int) function edl_vec_random(ctx, prec, dev, n, seed) bind(C, name="edl_vec_random")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, dev
       integer(c_int), value :: prec
       integer(c_int64_t), value :: n, seed
     end function edl_vec_random

     integer(c_int) function edl_vec_dot(ctx, prec, a, b, n, res) bind(C, name="edl_vec_dot")
       import :: c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: ctx, a, b
       integer(c_int), value :: prec
       integer(c_int64_t), value :: n
       real(c_double) :: res
     end function edl_vec_dot

     integer(c_int) function edl_device_alloc(ctx, bytes, dev_ptr) bind(C, name="edl_device_alloc")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: bytes
       type(c_ptr) :: dev_ptr
     end function edl_device_alloc

     integer(c_int) function edl_copy_h2d(ctx, dst, src, bytes) bind(C, name="edl_copy_h2d")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, dst, src
       integer(c_int64_t), value :: bytes
     end function edl_copy_h2d

     integer(c_int) function edl_copy_d2h(ctx, dst, src, bytes) bind(C, name="edl_copy_d2h")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, dst, src
       integer(c_int64_t), value :: bytes
     end function edl_copy_d2h

     integer(c_int) function edl_lanczos_gs(ctx, isector, prec, niter, tol, seed, refine_tol, e0, nsteps, alpha, &
          beta, gs_host) bind(C, name="edl_lanczos_gs")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, prec, niter
       real(c_double), value :: tol, refine_tol
       integer(c_int64_t), value :: seed
       real(c_double) :: e0
       integer(c_int) :: nsteps
       real(c_double) :: alpha(*), beta(*)
       type(c_ptr), value :: gs_host
     end function edl_lanczos_gs

     integer(c_int) function edl_lanczos_stats(ctx, steps_lo, steps_hi, e0_lo) bind(C, name="edl_lanczos_stats")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int) :: steps_lo, steps_hi
       real(c_double) :: e0_lo
     end function edl_lanczos_stats

     integer(c_int) function edl_set_state(ctx, isector, energy, vec) bind(C, name="edl_set_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector
       real(c_double), value :: energy
       real(c_double) :: vec(*)
     end function edl_set_state

     integer(c_int) function edl_get_state(ctx, vec) bind(C, name="edl_get_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: vec(*)
     end function edl_get_state

     integer(c_int) function edl_tridiag(ctx, isector, prec, seed, nlanc, alpha, beta, norm2, nsteps) &
          bind(C, name="edl_tridiag")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, prec, nlanc
       real(c_double) :: seed(*), alpha(*), beta(*), norm2
       integer(c_int) :: nsteps
     end function edl_tridiag

     integer(c_int) function edl_tridiag_batched(ctx, nseeds, seeds, nlanc, prec, group, alpha, beta, norm2, &
          nsteps, isector_out) bind(C, name="edl_tridiag_batched")
       import :: c_int, c_ptr, c_double, edl_seed_desc
       type(c_ptr), value :: ctx
       integer(c_int), value :: nseeds, nlanc, prec, group
       type(edl_seed_desc) :: seeds(*)
       real(c_double) :: alpha(*), beta(*), norm2(*)
       integer(c_int) :: nsteps(*), isector_out(*)
     end function edl_tridiag_batched

     integer(c_int) function edl_lanczos_poles(nlanc, alpha, beta, norm2, Ei, isign, poles, weights) &
          bind(C, name="edl_lanczos_poles")
       import :: c_int, c_double
       integer(c_int), value :: nlanc, isign
       real(c_double) :: alpha(*), beta(*), poles(*), weights(*)
       real(c_double), value :: norm2, Ei
     end function edl_lanczos_poles

     integer(c_int) function edl_observables(ctx, dens_up, dens_dw, docc) bind(C, name="edl_observables")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: dens_up(*), dens_dw(*), docc(*)
     end function edl_observables

     integer(c_int) function edl_lanczos_states(ctx, isector, prec, nstates, niter, tol, seed, energies, vecs) &
          bind(C, name="edl_lanczos_states")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, prec, nstates, niter
       real(c_double), value :: tol
       integer(c_int64_t), value :: seed
       real(c_double) :: energies(*)
       type(c_ptr), value :: vecs
     end function edl_lanczos_states

     integer(c_int) function edl_correlations(ctx, sz2, n2) bind(C, name="edl_correlations")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: sz2(*), n2(*)
     end function edl_correlations

     integer(c_int) function edl_local_energy(ctx, e) bind(C, name="edl_local_energy")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: e(*)
     end function edl_local_energy

     integer(c_int) function edl_last_timing(ctx, ms, hxv_count) bind(C, name="edl_last_timing")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double) :: ms
       integer(c_int64_t) :: hxv_count
     end function edl_last_timing

     integer(c_int) function edl_bench_hxv(ctx, isector, prec, reps, ms_per_hxv) bind(C, name="edl_bench_hxv")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, prec, reps
       real(c_double) :: ms_per_hxv
     end function edl_bench_hxv

     integer(c_int) function edl_bench_passes(ctx, isector, prec, reps, ms_pass, cfg) bind(C, name="edl_bench_passes")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: isector, prec, reps
       real(c_double) :: ms_pass(2)
       integer(c_int32_t) :: cfg(14)
     end function edl_bench_passes

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  !> Message of the last failure of `ctx` (or of a context-free call).
  function edl_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = edl_last_error(ctx)
    if (.not. c_associated(p)) then
       msg = "unknown error"
       return
    endif
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    enddo
  end function edl_error_message

  !> Reference error semantics: stop with the library's message on failure.
  subroutine edl_check(ctx, status, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= 0) then
       write(*, "(A)") where // " ERROR: " // edl_error_message(ctx)
       stop 1
    endif
  end subroutine edl_check

end module ED_GPU
