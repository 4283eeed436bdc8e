!> edlat_gpu: Fortran host of the GPU engine for the DIAG / GF / CHI path.
!!
!! Follows the reference driver flow ed_init_solver + ed_solve (MAIN.f90:34-108)
!! for Norb = 1 at T = 0 with lanc_method = lanczos:
!!   diagonalize_lattice  -> edl_lanczos_gs per sector, state list kept within
!!                           gs_threshold of the minimum (DIAG.f90:264-280)
!!   observables_lattice  -> edl_observables, dens.ed / docc.ed / egs.ed
!!                           (OBSERVABLES.f90:318-352)
!!   build/eval_gf        -> edl_tridiag_batched on all c, c^dg (and mixed)
!!                           seeds, poles/weights (GF_ELECTRONS.f90:417-512),
!!                           G_*_iw.ed / G_*_realw.ed, gfmatrix.restart
!!   build/eval_chi       -> S^z seeds, spinChi_*_tau.ed, ChiSpinMatrix.restart
!!                           (CHI_SPIN_ELECTRONS.f90:305-468)
!! Input: `edlat_gpu [input file]` (default inputED.conf), NAME=value lines
!! with the reference's variable names and defaults (INPUT_VARS.f90:106-166),
!! plus HIJ_FILE (the Hij_read file, GRAPH_MATRIX.f90:280-301),
!! SECTOR_NUP / SECTOR_NDW (one sector, -1 = every sector allowed by
!! ED_FILLING), ED_PRECISION (fp64 | fp32 | mixed) and LANC_SEED.
program edlat_gpu
  use iso_c_binding
  use ED_GPU
  implicit none

  ! ---- input variables (reference names and defaults)
  integer :: Norb = 1, Ns = 0, Nspin = 1, ed_filling = 0, Lmats = 4096, Lreal = 5000, Ltau = 1024
  integer :: lanc_niter = 512, lanc_ngfiter = 200, sector_nup = -1, sector_ndw = -1, ed_verbose = 3
  real(8) :: Uloc = 2d0, Ust = 0d0, Jh = 0d0, Jx = 0d0, Jp = 0d0, temp = 0.001d0, xmu = 0d0, eps = 0.01d0
  real(8) :: wini = -5d0, wfin = 5d0, gs_threshold = 1d-9, lanc_tolerance = 1d-18, cutoff = 1d-9
  logical :: hfmode = .true., gf_flag = .false., offdiag_gf_flag = .false., chispin_flag = .false.
  ! ED_SPARSE_H (INPUT_VARS.f90:151, default T): the stored-H path's operator
  ! keeps the -xmu*N and hfmode -U/2*N diagonal terms at any ED_FILLING
  ! (stored/H_local.f90:61-90); both paths run on the on-the-fly kernels
  logical :: ed_sparse_h = .true.
  ! OC_FLAG (INPUT_VARS.f90:138): optical conductivity and Drude weight
  logical :: oc_flag = .false.
  ! ED_FINITE_TEMP with LANC_NSTATES_SECTOR / LANC_NSTATES_TOTAL
  ! (SETUP.f90:87-103): several states per sector, Boltzmann weights
  logical :: ed_finite_temp = .false.
  integer :: lanc_nstates_sector = 2, lanc_nstates_total = 10
  character(len=256) :: hij_file = "hij.in", ed_precision = "fp64", input_file = "inputED.conf"
  integer(c_int64_t) :: lanc_seed = int(z'5EED', c_int64_t)

  type(c_ptr) :: ctx
  real(8), allocatable :: Hij(:, :, :), wm(:), wr(:), tau(:)
  integer :: prec, nsel, k, isel, ist, nstates
  integer, allocatable :: sel(:)
  ! state list (EIGENSPACE.f90): sector, energy, Boltzmann factor
  ! exp(-(E - Emin)/T) at finite temperature (1 at T = 0), vector
  integer, parameter :: maxstates = 1024
  integer :: st_sector(maxstates)
  real(8) :: st_energy(maxstates), st_w(maxstates), emin
  type vecbox
     real(8), allocatable :: v(:)
  end type vecbox
  type(vecbox) :: st_vec(maxstates)
  real(8), allocatable :: dens_up(:), dens_dw(:), docc(:), du(:), dd(:), dc(:)
  complex(8), allocatable :: Gmats(:, :, :, :), Greal(:, :, :, :)
  real(8), allocatable :: chitau(:, :)
  real(8) :: zeta
  real(8) :: energy(11)   ! lanc_energy columns of energy_last.ed

  call get_command_argument(1, input_file)
  if (len_trim(input_file) == 0) input_file = "inputED.conf"
  call read_input(trim(input_file))
  call read_hij(trim(hij_file))
  if (ed_finite_temp) then   ! SETUP.f90:91-103
     if (mod(lanc_nstates_sector, 2) /= 0) lanc_nstates_sector = lanc_nstates_sector + 1
     lanc_nstates_total = lanc_nstates_sector * (Ns + 1)**2 + 10
     if (mod(lanc_nstates_total, 2) /= 0) lanc_nstates_total = lanc_nstates_total + 1
  endif

  select case (trim(ed_precision))
  case ("fp64"); prec = EDL_FP64
  case ("fp32"); prec = EDL_FP32
  case ("mixed"); prec = EDL_MIXED
  case default; stop "ED_PRECISION must be fp64, fp32 or mixed"
  end select

  call edl_check(c_null_ptr, edl_ctx_create(ctx, 0), "edl_ctx_create")
  call set_model()
  call select_sectors()
  call diagonalize()
  ! get_Z_function (DIAG.f90:56-76): the number of kept states at T = 0,
  ! sum of the Boltzmann factors of the trimmed list at finite temperature
  zeta = sum(st_w(1:nstates))
  call observables()
  if (gf_flag) call greens_functions()
  if (chispin_flag) call spin_chi()
  if (oc_flag) call optical_conductivity()
  call edl_check(ctx, edl_ctx_destroy(ctx), "edl_ctx_destroy")

contains

  ! ------------------------------------------------------------- input
  subroutine read_input(fname)
    character(len=*), intent(in) :: fname
    character(len=512) :: line, name, val
    integer :: unit, ios, p
    logical :: ex
    inquire(file=fname, exist=ex)
    if (.not. ex) then
       write(*, "(A)") "edlat_gpu: input file " // fname // " not found"
       stop 1
    endif
    open(newunit=unit, file=fname, status="old", action="read")
    do
       read(unit, "(A)", iostat=ios) line
       if (ios /= 0) exit
       p = index(line, "!")
       if (p > 0) line = line(:p - 1)
       p = index(line, "=")
       if (p == 0) cycle
       name = upper(adjustl(line(:p - 1)))
       val = adjustl(line(p + 1:))
       select case (trim(name))
       case ("NORB"); read(val, *) Norb
       case ("NSITES"); read(val, *) Ns
       case ("NSPIN"); read(val, *) Nspin
       case ("ED_FILLING"); read(val, *) ed_filling
       case ("ULOC"); read(val, *) Uloc
       case ("UST"); read(val, *) Ust
       case ("JH"); read(val, *) Jh
       case ("JX"); read(val, *) Jx
       case ("JP"); read(val, *) Jp
       case ("TEMP"); read(val, *) temp
       case ("XMU"); read(val, *) xmu
       case ("HFMODE"); read(val, *) hfmode
       case ("ED_SPARSE_H"); read(val, *) ed_sparse_h
       case ("EPS"); read(val, *) eps
       case ("WINI"); read(val, *) wini
       case ("WFIN"); read(val, *) wfin
       case ("LMATS"); read(val, *) Lmats
       case ("LREAL"); read(val, *) Lreal
       case ("LTAU"); read(val, *) Ltau
       case ("GF_FLAG"); read(val, *) gf_flag
       case ("OFFDIAG_GF_FLAG"); read(val, *) offdiag_gf_flag
       case ("CHISPIN_FLAG"); read(val, *) chispin_flag
       case ("OC_FLAG"); read(val, *) oc_flag
       case ("CUTOFF"); read(val, *) cutoff
       case ("ED_FINITE_TEMP"); read(val, *) ed_finite_temp
       case ("LANC_NSTATES_SECTOR"); read(val, *) lanc_nstates_sector
       case ("LANC_NSTATES_TOTAL"); read(val, *) lanc_nstates_total
       case ("LANC_NITER"); read(val, *) lanc_niter
       case ("LANC_NGFITER"); read(val, *) lanc_ngfiter
       case ("LANC_TOLERANCE"); read(val, *) lanc_tolerance
       case ("GS_THRESHOLD"); read(val, *) gs_threshold
       case ("ED_VERBOSE"); read(val, *) ed_verbose
       case ("HIJ_FILE"); hij_file = trim(val)
       case ("SECTOR_NUP"); read(val, *) sector_nup
       case ("SECTOR_NDW"); read(val, *) sector_ndw
       case ("ED_PRECISION"); ed_precision = trim(val)
       case ("LANC_SEED"); read(val, *) lanc_seed
       case default
          ! other reference variables (ED_METHOD, LANC_METHOD, ...) are accepted and ignored
       end select
    enddo
    close(unit)
    if (Norb /= 1) stop "edlat_gpu: only NORB=1 is driven by this host"
    if (Ns < 1) stop "edlat_gpu: NSITES must be given"
  end subroutine read_input

  function upper(s) result(u)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: i
    u = s
    do i = 1, len(s)
       if (u(i:i) >= "a" .and. u(i:i) <= "z") u(i:i) = achar(iachar(u(i:i)) - 32)
    enddo
  end function upper

  !> Hij_read (GRAPH_MATRIX.f90:280-301): lines `i j a b spin re im`.
  subroutine read_hij(fname)
    character(len=*), intent(in) :: fname
    integer :: unit, ios, i, j, a, b, spin
    real(8) :: re, im
    allocate(Hij(Nspin, Ns, Ns)); Hij = 0d0
    open(newunit=unit, file=fname, status="old", action="read")
    do
       read(unit, *, iostat=ios) i, j, a, b, spin, re, im
       if (ios /= 0) exit
       if (spin < 1 .or. spin > Nspin .or. i < 1 .or. i > Ns .or. j < 1 .or. j > Ns) &
            stop "Hij_add_link ERROR: index out of range"
       if (im /= 0d0) stop "edlat_gpu: complex Hij is not supported"
       Hij(spin, i, j) = re
    enddo
    close(unit)
  end subroutine read_hij

  subroutine set_model()
    integer(c_int) :: nsites(1)
    real(8) :: U(1)
    nsites(1) = Ns
    U(1) = Uloc
    call edl_check(ctx, edl_set_hamiltonian(ctx, Ns, Nspin, 1, nsites, Hij, c_null_ptr, U, Ust, Jh, Jx, Jp, xmu, &
         merge(1, 0, hfmode), merge(0, ed_filling, ed_sparse_h)), "edl_set_hamiltonian")
  end subroutine set_model

  subroutine select_sectors()
    integer :: nup, ndw
    allocate(sel((Ns + 1)**2))
    nsel = 0
    if (sector_nup >= 0 .and. sector_ndw >= 0) then
       nsel = 1
       sel(1) = edl_sector_index(Ns, sector_nup, sector_ndw)
       return
    endif
    do nup = 0, Ns
       do ndw = 0, Ns
          if (ed_filling /= 0 .and. nup + ndw /= ed_filling) cycle
          nsel = nsel + 1
          sel(nsel) = edl_sector_index(Ns, nup, ndw)
       enddo
    enddo
  end subroutine select_sectors

  ! ---------------------------------------------------------------- DIAG
  subroutine diagonalize()
    integer :: unit, isector, nup, ndw, nsteps, niter, ierr
    integer(c_int64_t) :: dim
    integer(c_int32_t) :: dimup, dimdw
    real(8) :: e0, oldzero
    real(8), allocatable :: alpha(:), beta(:)
    real(8), allocatable, target :: vec(:)
    nstates = 0
    oldzero = 1000d0
    open(newunit=unit, file="eigenvalues_list.ed", position="append", action="write")
    do isel = 1, nsel
       isector = sel(isel)
       ierr = edl_sector_qn(Ns, isector, nup, ndw)
       call edl_check(ctx, edl_sector_build(ctx, isector, dim, dimup, dimdw), "build_sector")
       niter = int(min(int(lanc_niter, c_int64_t), dim))
       if (ed_finite_temp) then
          call sector_states(isector, nup, ndw, dim, niter, unit)
          cycle
       endif
       allocate(alpha(niter), beta(niter), vec(dim))
       call edl_check(ctx, edl_lanczos_gs(ctx, isector, prec, niter, 0d0, lanc_seed, 1d-13, e0, nsteps, alpha, &
            beta, c_loc(vec)), "sp_lanc_eigh")
       if (ed_verbose > 2) write(*, "(A,I6,A,I3,A,I3,A,I12,A,F20.12,A,I5)") "sector", isector, " nup", nup, &
            " ndw", ndw, " dim", dim, " E0", e0, " steps", nsteps
       write(unit, "(I6,2I4,F24.15)") isector, nup, ndw, e0
       ! state list at T = 0 (DIAG.f90:268-279)
       if (e0 < oldzero - 10d0 * gs_threshold) then
          oldzero = e0
          call free_states()
          call add_state(isector, e0, vec)
       elseif (abs(e0 - oldzero) <= gs_threshold) then
          oldzero = min(oldzero, e0)
          call add_state(isector, e0, vec)
       endif
       deallocate(alpha, beta, vec)
    enddo
    close(unit)
    emin = minval(st_energy(1:nstates))
    st_w(1:nstates) = 1d0
    if (ed_finite_temp) call trim_states()
    open(newunit=unit, file="state_list.ed", action="write")
    do ist = 1, nstates
       ierr = edl_sector_qn(Ns, st_sector(ist), nup, ndw)
       write(unit, "(I6,F24.15,2I4)") st_sector(ist), st_energy(ist), nup, ndw
       write(*, "(A,F20.12,2I4)") "Egs =", st_energy(ist), nup, ndw
    enddo
    close(unit)
  end subroutine diagonalize

  !> ed_diag_d at finite temperature (DIAG.f90:142-151, 262-265): the lowest
  !> Neigen = min(lanc_nstates_sector, Dim) states of the sector by deflated
  !> Lanczos, added to the state list.
  subroutine sector_states(isector, nup, ndw, dim, niter, unit)
    integer, intent(in) :: isector, nup, ndw, niter, unit
    integer(c_int64_t), intent(in) :: dim
    integer :: neigen, i
    real(8), allocatable :: e(:)
    real(8), allocatable, target :: vecs(:)
    neigen = int(min(int(lanc_nstates_sector, c_int64_t), dim))
    allocate(e(neigen), vecs(dim * neigen))
    call edl_check(ctx, edl_lanczos_states(ctx, isector, merge(EDL_FP32, EDL_FP64, prec == EDL_FP32), neigen, &
         niter, 1d-13, lanc_seed, e, c_loc(vecs)), "sp_eigh")
    do i = 1, neigen
       write(unit, "(I6,2I4,F24.15)") isector, nup, ndw, e(i)
       call add_state(isector, e(i), vecs((i - 1) * dim + 1:i * dim))
    enddo
    if (ed_verbose > 2) write(*, "(A,I6,A,I3,A,I3,A,I12,A,20F16.10)") "sector", isector, " nup", nup, " ndw", &
         ndw, " dim", dim, " E", e
  end subroutine sector_states

  !> es_add_state with size = lanc_nstates_total keeps the lowest states of
  !> all sectors; es_trim_size (EIGENSPACE.f90:440-475) then drops the
  !> states from the first one with exp(-(E - Emin)/T) <= CUTOFF on.
  subroutine trim_states()
    integer :: i, j, keep
    integer :: isec_t
    real(8) :: e_t
    type(vecbox) :: v_t
    do i = 2, nstates   ! insertion sort by energy (stable)
       j = i
       do while (j > 1)
          if (st_energy(j - 1) <= st_energy(j)) exit
          isec_t = st_sector(j); st_sector(j) = st_sector(j - 1); st_sector(j - 1) = isec_t
          e_t = st_energy(j); st_energy(j) = st_energy(j - 1); st_energy(j - 1) = e_t
          call move_alloc(st_vec(j)%v, v_t%v)
          call move_alloc(st_vec(j - 1)%v, st_vec(j)%v)
          call move_alloc(v_t%v, st_vec(j - 1)%v)
          j = j - 1
       enddo
    enddo
    keep = min(nstates, lanc_nstates_total)
    do i = 1, keep
       if (exp(-(st_energy(i) - emin) / temp) <= cutoff) then
          keep = i - 1
          exit
       endif
    enddo
    if (keep == 0) stop "es_return_size: list does not contain any node"
    do i = keep + 1, nstates
       if (allocated(st_vec(i)%v)) deallocate(st_vec(i)%v)
    enddo
    nstates = keep
    do i = 1, nstates
       st_w(i) = exp(-(st_energy(i) - emin) / temp)
    enddo
  end subroutine trim_states

  subroutine free_states()
    do ist = 1, nstates
       if (allocated(st_vec(ist)%v)) deallocate(st_vec(ist)%v)
    enddo
    nstates = 0
  end subroutine free_states

  subroutine add_state(isector, e, vec)
    integer, intent(in) :: isector
    real(8), intent(in) :: e, vec(:)
    if (nstates >= maxstates) stop "edlat_gpu: state list full"
    nstates = nstates + 1
    st_sector(nstates) = isector
    st_energy(nstates) = e
    allocate(st_vec(nstates)%v(size(vec)))
    st_vec(nstates)%v = vec
  end subroutine add_state

  subroutine install(ist_)
    integer, intent(in) :: ist_
    call edl_check(ctx, edl_set_state(ctx, st_sector(ist_), st_energy(ist_), st_vec(ist_)%v), "es_return_cvector")
  end subroutine install

  ! --------------------------------------------------------- OBSERVABLES
  subroutine observables()
    integer :: unit
    real(8) :: e(11)
    allocate(dens_up(Ns), dens_dw(Ns), docc(Ns), du(Ns), dd(Ns), dc(Ns))
    dens_up = 0d0; dens_dw = 0d0; docc = 0d0; energy = 0d0
    do ist = 1, nstates
       call install(ist)
       call edl_check(ctx, edl_observables(ctx, du, dd, dc), "lanc_observables")
       dens_up = dens_up + du * st_w(ist) / zeta
       dens_dw = dens_dw + dd * st_w(ist) / zeta
       docc = docc + dc * st_w(ist) / zeta
       call edl_check(ctx, edl_local_energy(ctx, e), "lanc_energy")
       energy = energy + e * st_w(ist) / zeta
    enddo
    ! write_energy (ENERGY.f90:635-659)
    open(newunit=unit, file="energy_info.ed", action="write")
    write(unit, "(A1,90(A14,1X))") "#", "1<K>", "2<Hi>", "3<V>=<Hi-Ehf>", "4<Eloc>", "5<Ehf>", "6<Dst>", &
         "7<Dnd>", "8<Dse>", "9<Dph>", "10<Dkxy>", "11<Dkz>"
    close(unit)
    open(newunit=unit, file="energy_last.ed", action="write", position="append")
    write(unit, "(90F15.9)") energy
    close(unit)
    open(newunit=unit, file="dens.ed", action="write")
    write(unit, "(90(F20.12,1X))") (dens_up(k) + dens_dw(k), k = 1, Ns)
    close(unit)
    open(newunit=unit, file="docc.ed", action="write")
    write(unit, "(90(F20.12,1X))") (docc(k), k = 1, Ns)
    close(unit)
    open(newunit=unit, file="magz.ed", action="write")
    write(unit, "(90(F20.12,1X))") (dens_up(k) - dens_dw(k), k = 1, Ns)
    close(unit)
    open(newunit=unit, file="egs.ed", action="write")
    write(unit, *) emin
    close(unit)
    write(*, "(A,100F15.8)") "dens=", (dens_up(k) + dens_dw(k), k = 1, Ns), sum(dens_up + dens_dw)
  end subroutine observables

  ! ------------------------------------------------------------- grids
  subroutine grids()
    integer :: i
    Ltau = max(int(1d0 / temp), Ltau)
    allocate(wm(Lmats), wr(Lreal), tau(0:Ltau))
    do i = 1, Lmats
       wm(i) = acos(-1d0) * temp * (2 * i - 1)
    enddo
    do i = 1, Lreal
       wr(i) = wini + (wfin - wini) * dble(i - 1) / dble(max(Lreal - 1, 1))
    enddo
    do i = 0, Ltau
       tau(i) = (1d0 / temp) * dble(i) / dble(Ltau)
    enddo
  end subroutine grids

  ! ------------------------------------------------------------------ GF
  subroutine greens_functions()
    integer :: nseed, s, io, jo, ispin, ich, j, nl, isign, unit, ninfo
    type(edl_seed_desc), allocatable :: seeds(:)
    real(8), allocatable :: alpha(:), beta(:), norm2(:), poles(:), weights(:)
    integer(c_int), allocatable :: nsteps(:), jsec(:)
    integer(c_int64_t) :: dimj
    integer(c_int32_t) :: du_, dd_
    character(len=64) :: fname
    type chan
       real(8), allocatable :: p(:), w(:)
    end type chan
    type(chan), allocatable :: store(:, :, :, :, :)   ! (ich, istate, ispin, io, jo)
    call grids()
    allocate(Gmats(Nspin, Ns, Ns, Lmats), Greal(Nspin, Ns, Ns, Lreal))
    Gmats = (0d0, 0d0); Greal = (0d0, 0d0)
    allocate(store(2, nstates, Nspin, Ns, Ns))
    nseed = 2 * Nspin * Ns
    if (offdiag_gf_flag) nseed = nseed + 2 * Nspin * Ns * (Ns - 1)
    allocate(seeds(nseed), norm2(nseed), nsteps(nseed), jsec(nseed))
    allocate(alpha(nseed * lanc_ngfiter), beta(nseed * lanc_ngfiter))
    do ist = 1, nstates
       call install(ist)
       s = 0
       do ispin = 1, Nspin
          do io = 1, Ns
             do jo = 1, Ns
                if (io /= jo .and. .not. offdiag_gf_flag) cycle
                s = s + 1; seeds(s) = edl_seed_desc(EDL_OP_CDG, ispin, io, merge(0, jo, io == jo))
                s = s + 1; seeds(s) = edl_seed_desc(EDL_OP_C, ispin, io, merge(0, jo, io == jo))
             enddo
          enddo
       enddo
       call edl_check(ctx, edl_tridiag_batched(ctx, nseed, seeds, lanc_ngfiter, merge(EDL_FP32, EDL_FP64, &
            prec == EDL_FP32), 0, alpha, beta, norm2, nsteps, jsec), "tridiag_Hv_sector")
       do s = 1, nseed
          ich = merge(1, 2, seeds(s)%op == EDL_OP_CDG)
          isign = merge(1, -1, ich == 1)
          ispin = seeds(s)%ispin; io = seeds(s)%io; jo = merge(io, seeds(s)%jo, seeds(s)%jo == 0)
          if (jsec(s) == 0) then
             allocate(store(ich, ist, ispin, io, jo)%p(0), store(ich, ist, ispin, io, jo)%w(0))
             cycle
          endif
          call edl_check(ctx, edl_sector_build(ctx, jsec(s), dimj, du_, dd_), "build_sector")
          nl = int(min(int(lanc_ngfiter, c_int64_t), dimj))
          allocate(poles(nl), weights(nl))
          if (norm2(s) /= 0d0) then
             call edl_check(c_null_ptr, edl_lanczos_poles(nl, alpha((s - 1) * lanc_ngfiter + 1), &
                  beta((s - 1) * lanc_ngfiter + 1), norm2(s), st_energy(ist), isign, poles, weights), &
                  "add_to_lanczos_gf_normal")
          else
             poles = 0d0; weights = 0d0
          endif
          do j = 1, nl
             Gmats(ispin, io, jo, :) = Gmats(ispin, io, jo, :) + weights(j) * st_w(ist) / zeta / &
                  (dcmplx(0d0, wm) - poles(j))
             Greal(ispin, io, jo, :) = Greal(ispin, io, jo, :) + weights(j) * st_w(ist) / zeta / &
                  (dcmplx(wr, eps) - poles(j))
          enddo
          call move_alloc(poles, store(ich, ist, ispin, io, jo)%p)
          call move_alloc(weights, store(ich, ist, ispin, io, jo)%w)
       enddo
    enddo
    ! off-diagonal from the mixed seeds (GF_ELECTRONS.f90:172-175)
    if (offdiag_gf_flag) then
       do ispin = 1, Nspin
          do io = 1, Ns
             do jo = 1, Ns
                if (io == jo) cycle
                Gmats(ispin, io, jo, :) = 0.5d0 * (Gmats(ispin, io, jo, :) - Gmats(ispin, io, io, :) - Gmats(ispin, jo, jo, :))
                Greal(ispin, io, jo, :) = 0.5d0 * (Greal(ispin, io, jo, :) - Greal(ispin, io, io, :) - Greal(ispin, jo, jo, :))
             enddo
          enddo
       enddo
    endif
    ! G files (ed_write_func, IO.f90:159-261): x, Im G, Re G
    do ispin = 1, Nspin
       do io = 1, Ns
          do jo = 1, Ns
             if (io /= jo .and. .not. offdiag_gf_flag) cycle
             if (offdiag_gf_flag) then
                write(fname, "(A,I4.4,A,I4.4,A,I0)") "G_i", io, "j", jo, "_l1m1_s", ispin
             else
                write(fname, "(A,I4.4,A,I0)") "G_i", io, "_l1_s", ispin
             endif
             call splot(trim(fname) // "_iw.ed", wm, Gmats(ispin, io, jo, :))
             call splot(trim(fname) // "_realw.ed", wr, Greal(ispin, io, jo, :))
          enddo
       enddo
    enddo
    ! gfmatrix.restart (write_GFmatrix_all3 / write_formatted_gfmatrix, VARS_GLOBAL.f90:456-471,587-606)
    open(newunit=unit, file="gfmatrix.restart", action="write")
    do ispin = 1, Nspin
       do io = 1, Ns
          do jo = 1, Ns
             if (io /= jo .and. .not. offdiag_gf_flag) then
                write(unit, *) .false.
                cycle
             endif
             write(unit, *) .true.
             write(unit, *) nstates
             do ist = 1, nstates
                write(unit, *) 2
                do ich = 1, 2
                   ninfo = size(store(ich, ist, ispin, io, jo)%p)
                   write(unit, *) ninfo
                   write(unit, *) store(ich, ist, ispin, io, jo)%p
                   write(unit, *) store(ich, ist, ispin, io, jo)%w
                enddo
             enddo
             write(unit, *) ""
          enddo
       enddo
    enddo
    close(unit)
  end subroutine greens_functions

  ! ----------------------------------------------------------------- OC
  !> build_oc_electrons / eval_oc_electrons (OC_ELECTRONS.f90:103-319) for
  !> the one orbital of this host, then eval_oc_lattice
  !> (OC_FUNCTIONS.f90:45-89): Drude weight -pi/Nsites*(<K> + 2 D) with D the
  !> current-correlation sum of peso/dE, resistivity -1/Drude.
  subroutine optical_conductivity()
    integer :: j, nl, unit
    type(edl_seed_desc) :: seeds(1)
    real(8), allocatable :: alpha(:), beta(:), poles(:), weights(:)
    real(8) :: norm2(1), d_jj, drude, res, peso
    integer(c_int) :: nsteps(1), jsec(1)
    allocate(alpha(lanc_ngfiter), beta(lanc_ngfiter), poles(lanc_ngfiter), weights(lanc_ngfiter))
    d_jj = 0d0
    do ist = 1, nstates
       call install(ist)
       seeds(1) = edl_seed_desc(EDL_OP_J, 0, 1, 0)
       call edl_check(ctx, edl_tridiag_batched(ctx, 1, seeds, lanc_ngfiter, merge(EDL_FP32, EDL_FP64, &
            prec == EDL_FP32), 0, alpha, beta, norm2, nsteps, jsec), "tridiag_Hv_sector")
       nl = nsteps(1)
       if (nl <= 0 .or. norm2(1) == 0d0) cycle
       call edl_check(c_null_ptr, edl_lanczos_poles(nl, alpha, beta, norm2(1), st_energy(ist), 1, poles, weights), &
            "add_to_lanczos_oc")
       do j = 1, nl
          peso = weights(j) * st_w(ist) / zeta
          if (abs(poles(j)) < cutoff) cycle
          if (abs(peso / poles(j)) < cutoff) cycle
          d_jj = d_jj + peso / poles(j)
       enddo
    enddo
    drude = -acos(-1d0) / dble(Ns) * (energy(1) + 2d0 * d_jj)
    res = -1d0 / drude
    open(newunit=unit, file="drude.ed", action="write", position="append")
    write(unit, *) drude
    close(unit)
    open(newunit=unit, file="resistivity.ed", action="write", position="append")
    write(unit, *) res
    close(unit)
    open(newunit=unit, file="jjcorr.ed", action="write", position="append")
    write(unit, *) d_jj
    close(unit)
    write(*, "(A,3F18.10)") "Drude, resistivity, D_jj =", drude, res, d_jj
  end subroutine optical_conductivity

  ! ---------------------------------------------------------------- CHI
  subroutine spin_chi()
    integer :: s, io, j, nl, unit, i
    type(edl_seed_desc), allocatable :: seeds(:)
    real(8), allocatable :: alpha(:), beta(:), norm2(:), poles(:), weights(:)
    integer(c_int), allocatable :: nsteps(:), jsec(:)
    integer(c_int64_t) :: dimj
    integer(c_int32_t) :: du_, dd_
    character(len=64) :: fname
    real(8) :: beta_t, peso, de
    if (.not. allocated(tau)) call grids()
    beta_t = 1d0 / temp
    allocate(chitau(Ns, 0:Ltau)); chitau = 0d0
    allocate(seeds(Ns), norm2(Ns), nsteps(Ns), jsec(Ns), alpha(Ns * lanc_ngfiter), beta(Ns * lanc_ngfiter))
    open(newunit=unit, file="ChiSpinMatrix.restart", action="write")
    do ist = 1, nstates
       call install(ist)
       do io = 1, Ns
          seeds(io) = edl_seed_desc(EDL_OP_SZ, 1, io, 0)
       enddo
       call edl_check(ctx, edl_tridiag_batched(ctx, Ns, seeds, lanc_ngfiter, merge(EDL_FP32, EDL_FP64, &
            prec == EDL_FP32), 0, alpha, beta, norm2, nsteps, jsec), "tridiag_Hv_sector")
       do s = 1, Ns
          call edl_check(ctx, edl_sector_build(ctx, jsec(s), dimj, du_, dd_), "build_sector")
          nl = int(min(int(lanc_ngfiter, c_int64_t), dimj))
          allocate(poles(nl), weights(nl))
          poles = 0d0; weights = 0d0
          if (norm2(s) /= 0d0) call edl_check(c_null_ptr, edl_lanczos_poles(nl, alpha((s - 1) * lanc_ngfiter + 1), &
               beta((s - 1) * lanc_ngfiter + 1), norm2(s), st_energy(ist), 1, poles, weights), "add_to_lanczos_spinChi")
          ! chi(tau) with the Ltau/2 symmetrisation (CHI_SPIN_ELECTRONS.f90:435-441)
          do j = 1, nl
             peso = weights(j) * st_w(ist) / zeta
             de = poles(j)
             do i = 0, Ltau / 2 - 1
                chitau(s, i) = chitau(s, i) + peso * exp(-tau(i) * de)
             enddo
             chitau(s, Ltau / 2) = chitau(s, Ltau / 2) + peso * 0.5d0 * (exp(-tau(Ltau / 2) * de) + &
                  exp(-(beta_t - tau(Ltau / 2)) * de))
             do i = Ltau / 2 + 1, Ltau
                chitau(s, i) = chitau(s, i) + peso * exp(-(beta_t - tau(i)) * de)
             enddo
          enddo
          write(unit, *) .true.
          write(unit, *) 1
          write(unit, *) 1
          write(unit, *) nl
          write(unit, *) poles
          write(unit, *) weights
          write(unit, *) ""
          deallocate(poles, weights)
       enddo
    enddo
    close(unit)
    do io = 1, Ns
       write(fname, "(A,I4.4,A)") "spinChi_i", io, "_l1_tau.ed"
       open(newunit=unit, file=trim(fname), action="write")
       do i = 0, Ltau
          write(unit, "(2ES24.15)") tau(i), chitau(io, i)
       enddo
       close(unit)
    enddo
  end subroutine spin_chi

  subroutine splot(fname, x, f)
    character(len=*), intent(in) :: fname
    real(8), intent(in) :: x(:)
    complex(8), intent(in) :: f(:)
    integer :: unit, i
    open(newunit=unit, file=fname, action="write")
    do i = 1, size(x)
       write(unit, "(3ES24.15)") x(i), dimag(f(i)), dreal(f(i))
    enddo
    close(unit)
  end subroutine splot

end program edlat_gpu
