!> EDLAT: the solver module a user driver `use`s, with the reference's entry
!! points, input variables and outputs (EDlattice.f90:1-51), over the GPU
!! engine libedlat_hip.so (module ED_GPU).
!!
!!   ed_read_input                        INPUT_VARS.f90:88-202
!!   ed_Hij_init / add_link / read / get  GRAPH_MATRIX.f90:79-407
!!   ed_Hij_local / info / write          GRAPH_MATRIX.f90:140-155,373-442
!!   ed_Hij_get_g0inv / g0func            GRAPH_MATRIX.f90:451-551
!!   ed_init_solver / ed_solve            MAIN.f90:34-108
!!   ed_get_gimp_matsubara / realaxis,    IO.f90:40-79
!!   ed_get_dens / mag / docc / energy /
!!   doubles
!!
!! ed_solve follows the reference flow: diagonalize_lattice (the sector
!! loop of ed_diag_d, DIAG.f90:103-332, with the twin mask and the state
!! list of EIGENSPACE.f90), build_gf_lattice / build_chi_lattice /
!! build_oc_lattice (every seed of a state in one batched call of the
!! engine; poles and weights stored per state), then for every temperature
!! of the list: partition_function_lattice, observables_lattice,
!! energy_lattice, eval_gf_lattice, eval_chi_lattice, eval_oc_lattice.
!!
!! Engine input variables beyond the reference's: ED_PRECISION (fp64 |
!! fp32 | mixed: the vectors and H x v of the Lanczos runs), LANC_SEED (the
!! SplitMix64 start vector), SECTOR_NUP / SECTOR_NDW (solve one sector
!! only; -1 = every sector the reference visits) and HIJ_FILE (read by the
!! edlat_gpu driver).
module EDLAT
  use iso_c_binding
  use ED_GPU
  implicit none
  private

  ! ------------------------------------------------------------ input variables
  ! (INPUT_VARS.f90:10-78; defaults of INPUT_VARS.f90:106-166)
  integer, public :: Norb = 1
  integer, allocatable, public :: Nsites(:)
  integer, public :: Nspin = 1
  real(8), allocatable, public :: Uloc(:)
  real(8), public :: Ust = 0d0, Jh = 0d0, Jx = 0d0, Jp = 0d0
  real(8), public :: xmu = 0d0, temp = 0.001d0, eps = 0.01d0, wini = -5d0, wfin = 5d0
  logical, public :: HFmode = .true.
  real(8), public :: cutoff = 1d-9, gs_threshold = 1d-9, sb_field = 0.1d0
  logical, allocatable, public :: gf_flag(:), chispin_flag(:), oc_flag(:)
  logical, public :: dm_flag = .false., rdm_flag = .false., offdiag_gf_flag = .false., offdiag_chispin_flag = .false.
  integer, public :: ed_filling = 0
  logical, public :: ed_finite_temp = .false., ed_sparse_H = .true., ed_print_G = .true., ed_twin = .false.
  character(len=12), public :: ed_method = "lanczos", lanc_method = "arpack"
  integer, public :: ed_verbose = 3
  real(8), public :: lanc_tolerance = 1d-18
  integer, public :: lanc_niter = 512, lanc_ngfiter = 200, lanc_ncv_factor = 10, lanc_ncv_add = 0
  integer, public :: lanc_nstates_sector = 2, lanc_dim_threshold = 1024, lanc_nstates_total = 1
  integer, public :: Lmats = 4096, Lreal = 5000, Ltau = 1024
  character(len=100), public :: Tfile = "temperature"
  integer, public :: LOGfile = 6
  character(len=200), public :: ed_input_file = ""
  ! engine variables
  character(len=16), public :: ed_precision = "fp64"
  integer(c_int64_t), public :: lanc_seed = int(z'5EED', c_int64_t)
  integer, public :: sector_nup = -1, sector_ndw = -1
  character(len=256), public :: hij_file = "hij.in"

  public :: ed_read_input
  public :: ed_Hij_init, ed_Hij_delete, ed_Hij_add_link, ed_Hij_read, ed_Hij_get, ed_Hij_local, ed_Hij_info
  public :: ed_Hij_write, ed_Hij_get_g0inv, ed_Hij_get_g0func
  public :: ed_init_solver, ed_solve, ed_finalize_solver
  public :: ed_get_gimp_matsubara, ed_get_gimp_realaxis, ed_get_dens, ed_get_mag, ed_get_docc
  public :: ed_get_energy, ed_get_doubles

  interface ed_Hij_get
     module procedure :: Hij_get_main, Hij_get_ispin, Hij_get_whole
  end interface ed_Hij_get
  interface ed_Hij_local
     module procedure :: Hij_local_main, Hij_local_ispin, Hij_local_whole
  end interface ed_Hij_local

  ! ------------------------------------------------------------ Hij structure
  ! (GRAPH_MATRIX.f90:17-71): a list of links (site, site, orbital,
  ! orbital, spin, value) and the matrices built from it
  type link_t
     integer :: siteI = 0, siteJ = 0, orbI = 0, orbJ = 0, spin = 0
     complex(8) :: Hval = (0d0, 0d0)
  end type link_t
  type(link_t), allocatable :: links(:)
  integer :: nlinks = 0
  logical :: hm_status = .false., hm_built = .false.
  integer :: hm_Ns = 0, hm_Norb = 0, hm_Nspin = 0
  integer, allocatable :: hm_Nsites(:)
  complex(8), allocatable :: hm_Hij(:, :, :), hm_Hloc(:, :, :)
  character(len=512) :: hm_file = ""

  ! ------------------------------------------------------------ state list
  ! (EIGENSPACE.f90): ascending energies; a twin entry (itwin) carries no
  ! vector, it is the spin flip of the entry just before it
  type estate_t
     real(8) :: e = 0d0
     integer :: sector = 0
     logical :: itwin = .false.
     integer :: uid = 0
     real(8), allocatable :: v(:)
  end type estate_t
  type(estate_t), allocatable :: slist(:)
  integer :: ssize = 0, strimd = 0, next_uid = 0, installed_uid = -1
  real(8) :: s_emin = 0d0, s_emax = 0d0, zeta_function = 0d0

  ! ------------------------------------------------------------ GFmatrix (VARS_GLOBAL.f90:47-61)
  type gfchannel_t
     real(8), allocatable :: poles(:), weight(:)
  end type gfchannel_t
  type gfstate_t
     type(gfchannel_t), allocatable :: channel(:)
  end type gfstate_t
  type gfmatrix_t
     logical :: status = .false.
     type(gfstate_t), allocatable :: state(:)
  end type gfmatrix_t
  type(gfmatrix_t), allocatable :: impGmatrix(:, :, :), SpinChiMatrix(:, :), OcMatrix(:)

  ! ------------------------------------------------------------ solver state and results
  type(c_ptr) :: ctx = c_null_ptr
  integer :: Ns = 0, Nsectors = 0, prec = 0
  logical :: finiteT = .false., solver_ready = .false.
  real(8), allocatable :: temperature_list(:)
  character(len=32) :: ed_file_suffix = ""
  integer, parameter :: site_indx_padding = 4
  real(8), allocatable :: wm(:), wr(:), vr(:), tau(:)
  complex(8), allocatable :: impGmats(:, :, :, :), impGreal(:, :, :, :)
  real(8), allocatable :: spinChi_tau(:, :, :)
  complex(8), allocatable :: spinChi_w(:, :, :)
  real(8), allocatable :: ed_dens(:), ed_dens_up(:), ed_dens_dw(:), ed_docc(:), ed_mag(:)
  ! <K>, <Hi>, <V>, <Eloc>, <Ehf>, Dust, Dund, Dse, Dph, Dkxy, Dkz (energy_last.ed)
  real(8) :: ed_energy(11) = 0d0
  real(8), allocatable :: Drude_weight(:), OptCond_w(:, :)

contains

  ! ======================================================================
  ! INPUT (INPUT_VARS.f90:88-202)
  ! ======================================================================

  !> Reads NAME=value lines (names case-insensitive, `!` and `#` start a
  !> comment, arrays as comma-separated values).  A missing file leaves the
  !> defaults, as the reference's parser does.
  subroutine ed_read_input(INPUTunit)
    character(len=*), intent(in) :: INPUTunit
    logical :: ex
    ed_input_file = INPUTunit
    inquire(file=INPUTunit, exist=ex)
    if (.not. ex) write(LOGfile, "(A)") "ed_read_input: " // trim(INPUTunit) // " not found, defaults used"
    Norb = 1
    if (ex) call parse_file(INPUTunit, .true.)
    if (Norb < 1) stop "ed_read_input ERROR: NORB < 1"
    if (allocated(Nsites)) deallocate(Nsites)
    if (allocated(Uloc)) deallocate(Uloc)
    if (allocated(gf_flag)) deallocate(gf_flag)
    if (allocated(chispin_flag)) deallocate(chispin_flag)
    if (allocated(oc_flag)) deallocate(oc_flag)
    allocate(Nsites(Norb), Uloc(Norb), gf_flag(Norb), chispin_flag(Norb), oc_flag(Norb))
    Nsites = 1
    Uloc = 2d0
    gf_flag = .false.
    chispin_flag = .false.
    oc_flag = .false.
    if (ex) call parse_file(INPUTunit, .false.)
    ! INPUT_VARS.f90:199-200
    call substring_delete(Tfile, ".restart")
    call substring_delete(Tfile, ".ed")
  end subroutine ed_read_input

  subroutine parse_file(fname, norb_only)
    character(len=*), intent(in) :: fname
    logical, intent(in) :: norb_only
    character(len=1024) :: line, name, val
    integer :: unit, ios, p
    open(newunit=unit, file=fname, status="old", action="read")
    do
       read(unit, "(A)", iostat=ios) line
       if (ios /= 0) exit
       p = index(line, "!")
       if (p > 0) line = line(:p - 1)
       p = index(line, "#")
       if (p > 0) line = line(:p - 1)
       p = index(line, "=")
       if (p == 0) cycle
       name = upper(adjustl(line(:p - 1)))
       val = adjustl(line(p + 1:))
       if (norb_only) then
          if (trim(name) == "NORB") read(val, *) Norb
          cycle
       endif
       select case (trim(name))
       case ("NORB"); read(val, *) Norb
       case ("NSITES"); read(val, *) Nsites
       case ("NSPIN"); read(val, *) Nspin
       case ("ED_FILLING"); read(val, *) ed_filling
       case ("ULOC"); read(val, *) Uloc
       case ("UST"); read(val, *) Ust
       case ("JH"); read(val, *) Jh
       case ("JX"); read(val, *) Jx
       case ("JP"); read(val, *) Jp
       case ("TEMP"); read(val, *) temp
       case ("ED_FINITE_TEMP"); read(val, *) ed_finite_temp
       case ("XMU"); read(val, *) xmu
       case ("SB_FIELD"); read(val, *) sb_field
       case ("WINI"); read(val, *) wini
       case ("WFIN"); read(val, *) wfin
       case ("LMATS"); read(val, *) Lmats
       case ("LREAL"); read(val, *) Lreal
       case ("LTAU"); read(val, *) Ltau
       case ("GF_FLAG"); read(val, *) gf_flag
       case ("CHISPIN_FLAG"); read(val, *) chispin_flag
       case ("OC_FLAG"); read(val, *) oc_flag
       case ("DM_FLAG"); read(val, *) dm_flag
       case ("RDM_FLAG"); read(val, *) rdm_flag
       case ("OFFDIAG_GF_FLAG"); read(val, *) offdiag_gf_flag
       case ("OFFDIAG_CHISPIN_FLAG"); read(val, *) offdiag_chispin_flag
       case ("HFMODE"); read(val, *) HFmode
       case ("EPS"); read(val, *) eps
       case ("CUTOFF"); read(val, *) cutoff
       case ("GS_THRESHOLD"); read(val, *) gs_threshold
       case ("ED_METHOD"); ed_method = lower(trim(val))
       case ("ED_TWIN"); read(val, *) ed_twin
       case ("ED_SPARSE_H"); read(val, *) ed_sparse_H
       case ("ED_PRINT_G"); read(val, *) ed_print_G
       case ("ED_VERBOSE"); read(val, *) ed_verbose
       case ("LANC_METHOD"); lanc_method = lower(trim(val))
       case ("LANC_NSTATES_SECTOR"); read(val, *) lanc_nstates_sector
       case ("LANC_NSTATES_TOTAL"); read(val, *) lanc_nstates_total
       case ("LANC_NCV_FACTOR"); read(val, *) lanc_ncv_factor
       case ("LANC_NCV_ADD"); read(val, *) lanc_ncv_add
       case ("LANC_NITER"); read(val, *) lanc_niter
       case ("LANC_NGFITER"); read(val, *) lanc_ngfiter
       case ("LANC_TOLERANCE"); read(val, *) lanc_tolerance
       case ("LANC_DIM_THRESHOLD"); read(val, *) lanc_dim_threshold
       case ("TFILE"); Tfile = trim(val)
       case ("LOGFILE"); read(val, *) LOGfile
       case ("ED_PRECISION"); ed_precision = lower(trim(val))
       case ("LANC_SEED"); read(val, *) lanc_seed
       case ("SECTOR_NUP"); read(val, *) sector_nup
       case ("SECTOR_NDW"); read(val, *) sector_ndw
       case ("HIJ_FILE"); hij_file = trim(val)
       case default
          ! other reference variables (ED_PRINT_DM, ED_ALL_G, ...) are accepted
       end select
    enddo
    close(unit)
  end subroutine parse_file

  function upper(s) result(u)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: i
    u = s
    do i = 1, len(s)
       if (u(i:i) >= "a" .and. u(i:i) <= "z") u(i:i) = achar(iachar(u(i:i)) - 32)
    enddo
  end function upper

  function lower(s) result(u)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: i
    u = s
    do i = 1, len(s)
       if (u(i:i) >= "A" .and. u(i:i) <= "Z") u(i:i) = achar(iachar(u(i:i)) + 32)
    enddo
  end function lower

  !> Removes every occurrence of `sub` from `s`.
  subroutine substring_delete(s, sub)
    character(len=*), intent(inout) :: s
    character(len=*), intent(in) :: sub
    integer :: loc
    do
       loc = index(s, sub)
       if (loc == 0) exit
       s = s(:loc - 1) // s(loc + len(sub):)
    enddo
  end subroutine substring_delete

  ! ======================================================================
  ! Hij (GRAPH_MATRIX.f90)
  ! ======================================================================

  subroutine ed_Hij_init(Nsites_, Nspin_)
    integer, intent(in) :: Nsites_(:)
    integer, intent(in), optional :: Nspin_
    if (hm_status) call ed_Hij_delete()
    hm_Nspin = 1
    if (present(Nspin_)) hm_Nspin = Nspin_
    hm_Norb = size(Nsites_)
    hm_Ns = sum(Nsites_)
    allocate(hm_Nsites(hm_Norb))
    hm_Nsites = Nsites_
    allocate(links(16))
    nlinks = 0
    hm_file = ""
    allocate(hm_Hij(hm_Nspin, hm_Ns, hm_Ns), hm_Hloc(hm_Nspin, hm_Ns, hm_Ns))
    hm_Hij = (0d0, 0d0)
    hm_Hloc = (0d0, 0d0)
    hm_built = .false.
    hm_status = .true.
  end subroutine ed_Hij_init

  subroutine ed_Hij_delete()
    if (.not. hm_status) return
    if (allocated(links)) deallocate(links)
    if (allocated(hm_Nsites)) deallocate(hm_Nsites)
    if (allocated(hm_Hij)) deallocate(hm_Hij)
    if (allocated(hm_Hloc)) deallocate(hm_Hloc)
    nlinks = 0
    hm_Ns = 0
    hm_Norb = 0
    hm_Nspin = 0
    hm_built = .false.
    hm_status = .false.
  end subroutine ed_Hij_delete

  subroutine ed_Hij_info()
    if (.not. hm_status) then
       write(*, "(A)") "Hij_info WARNING: Hmatrix not allocated"
       return
    endif
    write(*, "(A,10I8)") "Nsites  =", hm_Nsites
    write(*, "(A,I8)") "Ns      =", hm_Ns
    write(*, "(A,I8)") "Nspin   =", hm_Nspin
    write(*, "(A,I8)") "Norb    =", hm_Norb
    write(*, "(A,I8)") "# link  =", nlinks
    if (hm_file /= "") write(*, "(A,1x,A)") "file    =", trim(hm_file)
    write(*, "(A,L8)") "built H =", hm_built
  end subroutine ed_Hij_info

  subroutine ed_Hij_add_link(siteI, siteJ, orbI, orbJ, spin, Hval)
    integer, intent(in) :: siteI, siteJ, orbI, orbJ, spin
    complex(8), intent(in) :: Hval
    type(link_t), allocatable :: tmp(:)
    if (.not. hm_status) stop "Hij_add_link ERROR: Hmatrix not allocated"
    if (spin > hm_Nspin .or. spin <= 0) stop "Hij_add_link ERROR: spinI or spinJ either > Nspin OR < 0"
    if (orbI > hm_Norb .or. orbJ > hm_Norb .or. orbI <= 0 .or. orbJ <= 0) &
         stop "Hij_add_link ERROR: orbI or orbJ either > Norb OR < 0"
    if (siteI > hm_Nsites(orbI) .or. siteJ > hm_Nsites(orbJ) .or. siteI <= 0 .or. siteJ <= 0) &
         stop "Hij_add_link ERROR: siteI or siteJ either > Nlat OR < 0"
    if (nlinks == size(links)) then
       allocate(tmp(2 * nlinks))
       tmp(1:nlinks) = links(1:nlinks)
       call move_alloc(tmp, links)
    endif
    nlinks = nlinks + 1
    links(nlinks) = link_t(siteI, siteJ, orbI, orbJ, spin, Hval)
    hm_built = .false.
  end subroutine ed_Hij_add_link

  !> Hij_read (GRAPH_MATRIX.f90:280-301): lines `i j a b spin re im`.
  subroutine ed_Hij_read(file)
    character(len=*), intent(in) :: file
    integer :: unit, ios, i, j, a, b, spin
    real(8) :: re, im
    character(len=512) :: line
    if (.not. hm_status) stop "Hij_read ERROR: Hmatrix not allocated"
    open(newunit=unit, file=file, status="old", action="read")
    hm_file = file
    do
       read(unit, "(A)", iostat=ios) line
       if (ios /= 0) exit
       if (len_trim(line) == 0) cycle
       read(line, *) i, j, a, b, spin, re, im
       call ed_Hij_add_link(i, j, a, b, spin, dcmplx(re, im))
    enddo
    close(unit)
  end subroutine ed_Hij_read

  ! indices_link (GRAPH_MATRIX.f90:258-269), pack_indices (AUX_FUNX.f90:42-53)
  integer function pack_indices(site, orb) result(indx)
    integer, intent(in) :: site, orb
    if (orb == 1) then
       indx = site
    else
       indx = site + (orb - 1) * Nsites(orb - 1)
    endif
  end function pack_indices

  integer function link_index(site, orb) result(indx)
    integer, intent(in) :: site, orb
    if (orb == 1) then
       indx = site
    else
       indx = site + (orb - 1) * hm_Nsites(orb - 1)
    endif
  end function link_index

  !> Hij_build (GRAPH_MATRIX.f90:306-329)
  subroutine Hij_build()
    integer :: k, io, jo, spin
    if (.not. hm_status) stop "Hij_build ERROR: Hmatrix not allocated"
    hm_Hij = (0d0, 0d0)
    hm_Hloc = (0d0, 0d0)
    do k = 1, nlinks
       io = link_index(links(k)%siteI, links(k)%orbI)
       jo = link_index(links(k)%siteJ, links(k)%orbJ)
       hm_Hij(links(k)%spin, io, jo) = links(k)%Hval
       if (links(k)%siteI == links(k)%siteJ) hm_Hloc(links(k)%spin, io, jo) = links(k)%Hval
    enddo
    do spin = 1, hm_Nspin
       if (any(abs(hm_Hij(spin, :, :) - conjg(transpose(hm_Hij(spin, :, :)))) > 1d-6)) &
            stop "Hij_build ERROR: Hij is not Hermitian"
    enddo
    hm_built = .true.
  end subroutine Hij_build

  subroutine Hij_get_main(Hij)
    complex(8), intent(inout) :: Hij(:, :, :)
    if (.not. hm_status) stop "Hij_get ERROR: Hmatrix not allocated"
    if (any(shape(Hij) /= [hm_Nspin, hm_Ns, hm_Ns])) stop "Hij_get ERROR: wrong shape of Hij"
    if (.not. hm_built) call Hij_build()
    Hij = hm_Hij
  end subroutine Hij_get_main

  subroutine Hij_get_ispin(Hij, spin)
    complex(8), intent(inout) :: Hij(:, :)
    integer, intent(in) :: spin
    if (.not. hm_status) stop "Hij_get ERROR: Hmatrix not allocated"
    if (any(shape(Hij) /= [hm_Ns, hm_Ns])) stop "Hij_get ERROR: wrong shape of Hij"
    if (.not. hm_built) call Hij_build()
    Hij = hm_Hij(spin, :, :)
  end subroutine Hij_get_ispin

  subroutine Hij_get_whole(Hij)
    complex(8), intent(inout) :: Hij(:, :)
    if (.not. hm_status) stop "Hij_get ERROR: Hmatrix not allocated"
    if (any(shape(Hij) /= [hm_Nspin * hm_Ns, hm_Nspin * hm_Ns])) stop "Hij_get ERROR: wrong shape of Hij"
    if (.not. hm_built) call Hij_build()
    Hij = pack_spin(hm_Hij)
  end subroutine Hij_get_whole

  subroutine Hij_local_main(Hloc)
    complex(8), intent(inout) :: Hloc(:, :, :)
    if (.not. hm_status) stop "Hij_local ERROR: Hmatrix not allocated"
    if (any(shape(Hloc) /= [hm_Nspin, hm_Ns, hm_Ns])) stop "Hij_local ERROR: wrong shape of Hloc"
    if (.not. hm_built) call Hij_build()
    Hloc = hm_Hloc
  end subroutine Hij_local_main

  subroutine Hij_local_ispin(Hloc, spin)
    complex(8), intent(inout) :: Hloc(:, :)
    integer, intent(in) :: spin
    if (.not. hm_status) stop "Hij_local ERROR: Hmatrix not allocated"
    if (any(shape(Hloc) /= [hm_Ns, hm_Ns])) stop "Hij_local ERROR: wrong shape of Hloc"
    if (.not. hm_built) call Hij_build()
    Hloc = hm_Hloc(spin, :, :)
  end subroutine Hij_local_ispin

  subroutine Hij_local_whole(Hloc)
    complex(8), intent(inout) :: Hloc(:, :)
    if (.not. hm_status) stop "Hij_local ERROR: Hmatrix not allocated"
    if (any(shape(Hloc) /= [hm_Nspin * hm_Ns, hm_Nspin * hm_Ns])) stop "Hij_local ERROR: wrong shape of Hloc"
    if (.not. hm_built) call Hij_build()
    Hloc = pack_spin(hm_Hloc)
  end subroutine Hij_local_whole

  ! pack_matrix (GRAPH_MATRIX.f90:636-651): spin blocks on the diagonal
  function pack_spin(Hin) result(Hout)
    complex(8), intent(in) :: Hin(:, :, :)
    complex(8) :: Hout(size(Hin, 1) * size(Hin, 2), size(Hin, 1) * size(Hin, 2))
    integer :: ispin, n
    n = size(Hin, 2)
    Hout = (0d0, 0d0)
    do ispin = 1, size(Hin, 1)
       Hout((ispin - 1) * n + 1:ispin * n, (ispin - 1) * n + 1:ispin * n) = Hin(ispin, :, :)
    enddo
  end function pack_spin

  !> Hij_write (GRAPH_MATRIX.f90:414-442): the links, one per line.
  subroutine ed_Hij_write(unit, file)
    integer, intent(in), optional :: unit
    character(len=*), intent(in), optional :: file
    integer :: u, k
    if (.not. hm_status) stop "Hij_write ERROR: Hmatrix not allocated"
    u = 6
    if (present(unit)) u = unit
    if (present(file)) open(newunit=u, file=file, action="write")
    do k = 1, nlinks
       write(u, "(5I6,2F12.4)") links(k)%siteI, links(k)%siteJ, links(k)%orbI, links(k)%orbJ, links(k)%spin, &
            dreal(links(k)%Hval), dimag(links(k)%Hval)
    enddo
    if (present(file)) close(u)
  end subroutine ed_Hij_write

  !> Hij_get_g0inv (GRAPH_MATRIX.f90:451-497): G0^-1(z) = z - Hij.
  subroutine ed_Hij_get_g0inv(zeta, Gloc)
    complex(8), intent(in) :: zeta(:)
    complex(8), intent(inout) :: Gloc(:, :, :, :)
    integer :: ispin, i, k
    if (.not. hm_status) stop "Hij_get_gfunc ERROR: Hmatrix not allocated"
    if (any(shape(Gloc) /= [hm_Nspin, hm_Ns, hm_Ns, size(zeta)])) stop "Hij_get_g0inv ERROR: wrong shape of Gloc"
    if (.not. hm_built) call Hij_build()
    do ispin = 1, hm_Nspin
       do i = 1, size(zeta)
          Gloc(ispin, :, :, i) = -hm_Hij(ispin, :, :)
          do k = 1, hm_Ns
             Gloc(ispin, k, k, i) = Gloc(ispin, k, k, i) + zeta(i)
          enddo
       enddo
    enddo
  end subroutine ed_Hij_get_g0inv

  !> Hij_get_g0func (GRAPH_MATRIX.f90:501-551): G0(z) = (z - Hij)^-1.
  subroutine ed_Hij_get_g0func(zeta, Gloc)
    complex(8), intent(in) :: zeta(:)
    complex(8), intent(inout) :: Gloc(:, :, :, :)
    complex(8), allocatable :: A(:, :)
    integer :: ispin, i
    call ed_Hij_get_g0inv(zeta, Gloc)
    allocate(A(hm_Ns, hm_Ns))
    do ispin = 1, hm_Nspin
       do i = 1, size(zeta)
          A = Gloc(ispin, :, :, i)
          call cinverse(A)
          Gloc(ispin, :, :, i) = A
       enddo
    enddo
  end subroutine ed_Hij_get_g0func

  !> In-place inverse of a complex matrix: Gauss-Jordan elimination with
  !> partial pivoting on [A | I].
  subroutine cinverse(A)
    complex(8), intent(inout) :: A(:, :)
    complex(8), allocatable :: B(:, :), row(:)
    complex(8) :: piv, f
    integer :: n, i, k, p
    n = size(A, 1)
    allocate(B(n, n), row(n))
    B = (0d0, 0d0)
    do i = 1, n
       B(i, i) = (1d0, 0d0)
    enddo
    do k = 1, n
       p = k - 1 + maxloc(abs(A(k:n, k)), 1)
       if (abs(A(p, k)) == 0d0) stop "ed_Hij_get_g0func ERROR: z - Hij is singular"
       if (p /= k) then
          row = A(k, :); A(k, :) = A(p, :); A(p, :) = row
          row = B(k, :); B(k, :) = B(p, :); B(p, :) = row
       endif
       piv = A(k, k)
       A(k, :) = A(k, :) / piv
       B(k, :) = B(k, :) / piv
       do i = 1, n
          if (i == k) cycle
          f = A(i, k)
          if (f == (0d0, 0d0)) cycle
          A(i, :) = A(i, :) - f * A(k, :)
          B(i, :) = B(i, :) - f * B(k, :)
       enddo
    enddo
    A = B
  end subroutine cinverse

  ! ======================================================================
  ! SOLVER (MAIN.f90:34-108)
  ! ======================================================================

  !> ed_init_solver (MAIN.f90:34-55): checks (ed_checks_global,
  !> SETUP.f90:209-225), the structure set-up (init_ed_structure,
  !> SETUP.f90:25-200) and the engine context with the model.
  subroutine ed_init_solver()
    integer :: i, unit, nt, ios
    real(8) :: t
    real(8), allocatable :: tl(:)
    logical :: ex
    if (.not. hm_status) stop "ED_INIT_SOLVER ERROR: Hmatrix is not allocated"
    if (.not. allocated(Nsites)) stop "ED_INIT_SOLVER ERROR: call ed_read_input first"
    Ns = sum(Nsites(1:Norb))
    if (Ns /= hm_Ns .or. Nspin /= hm_Nspin) stop "ED_INIT_SOLVER ERROR: Hmatrix does not match Nsites / Nspin"
    Nsectors = (Ns + 1) * (Ns + 1)
    ! ed_checks_global
    if (Nspin > 2) stop "ED ERROR: Nspin > 2 is currently not supported"
    if (Norb > 5) stop "ED ERROR: Norb > 5 is currently not supported"
    if (Nspin > 1 .and. ed_twin) write(LOGfile, "(A)") "WARNING: using twin_sector with Nspin>1"
    if (trim(ed_method) == "lapack" .or. trim(ed_method) == "full") &
         stop "ED ERROR: ed_method=lapack/full is not provided by this engine (use lanczos)"
    if (trim(lanc_method) == "lanczos") then
       if (ed_finite_temp) stop "ED ERROR: lanc_method==lanczos available only for T=0"
       if (lanc_nstates_sector > 1) stop "ED ERROR: lanc_method==lanczos available only for lanc_nstates_sector==1, T=0"
    endif
    lanc_nstates_total = 1
    finiteT = ed_finite_temp
    if (finiteT) then
       if (mod(lanc_nstates_sector, 2) /= 0) then
          lanc_nstates_sector = lanc_nstates_sector + 1
          write(LOGfile, "(A,I10)") "Increased Lanc_nstates_sector:", lanc_nstates_sector
       endif
       lanc_nstates_total = lanc_nstates_sector * Nsectors + 10
       if (mod(lanc_nstates_total, 2) /= 0) then
          lanc_nstates_total = lanc_nstates_total + 1
          write(LOGfile, "(A,I10)") "Increased Lanc_nstates_total:", lanc_nstates_total
       endif
       write(LOGfile, "(A)") "Lanczos FINITE temperature calculation:"
    else
       write(LOGfile, "(A)") "Lanczos ZERO temperature calculation:"
    endif
    ! temperature list: Tfile.restart, largest first (SETUP.f90:123-142)
    if (allocated(temperature_list)) deallocate(temperature_list)
    allocate(temperature_list(1))
    temperature_list = temp
    inquire(file=trim(Tfile) // ".restart", exist=ex)
    if (finiteT .and. ex) then
       write(LOGfile, "(A)") "Reading temperature list from file " // trim(Tfile) // ".restart"
       allocate(tl(4096))
       nt = 0
       open(newunit=unit, file=trim(Tfile) // ".restart", status="old", action="read")
       do
          read(unit, *, iostat=ios) t
          if (ios /= 0) exit
          nt = nt + 1
          if (nt > size(tl)) stop "ED_INIT_SOLVER ERROR: temperature list longer than 4096"
          tl(nt) = t
       enddo
       close(unit)
       if (nt == 0) stop "ED_INIT_SOLVER ERROR: empty temperature list"
       call sort_desc(tl(1:nt))
       deallocate(temperature_list)
       allocate(temperature_list(nt))
       temperature_list = tl(1:nt)
       open(newunit=unit, file=trim(Tfile) // ".used", action="write")
       do i = 1, nt
          write(unit, *) temperature_list(i)
       enddo
       close(unit)
       temp = temperature_list(1)
    endif
    select case (trim(ed_precision))
    case ("fp64"); prec = EDL_FP64
    case ("fp32"); prec = EDL_FP32
    case ("mixed"); prec = EDL_MIXED
    case default; stop "ED_INIT_SOLVER ERROR: ED_PRECISION must be fp64, fp32 or mixed"
    end select
    if (.not. c_associated(ctx)) call edl_check(c_null_ptr, edl_ctx_create(ctx, 0), "edl_ctx_create")
    call set_model()
    if (allocated(impGmatrix)) deallocate(impGmatrix)
    if (allocated(SpinChiMatrix)) deallocate(SpinChiMatrix)
    if (allocated(OcMatrix)) deallocate(OcMatrix)
    allocate(impGmatrix(Nspin, Ns, Ns), SpinChiMatrix(Ns, Ns), OcMatrix(Ns))
    if (allocated(ed_dens)) deallocate(ed_dens, ed_dens_up, ed_dens_dw, ed_docc, ed_mag)
    allocate(ed_dens(Ns), ed_dens_up(Ns), ed_dens_dw(Ns), ed_docc(Ns), ed_mag(Ns))
    ed_dens = 0d0; ed_dens_up = 0d0; ed_dens_dw = 0d0; ed_docc = 0d0; ed_mag = 0d0
    if (allocated(Drude_weight)) deallocate(Drude_weight, OptCond_w)
    allocate(Drude_weight(Norb), OptCond_w(Norb, Lreal))
    Drude_weight = 0d0
    OptCond_w = (0d0, 0d0)
    if (allocated(impGmats)) deallocate(impGmats, impGreal)
    allocate(impGmats(Nspin, Ns, Ns, Lmats), impGreal(Nspin, Ns, Ns, Lreal))
    impGmats = (0d0, 0d0)
    impGreal = (0d0, 0d0)
    solver_ready = .true.
  end subroutine ed_init_solver

  subroutine sort_desc(a)
    real(8), intent(inout) :: a(:)
    integer :: i, j
    real(8) :: t
    do i = 2, size(a)
       t = a(i)
       j = i - 1
       do while (j >= 1)
          if (a(j) >= t) exit
          a(j + 1) = a(j)
          j = j - 1
       enddo
       a(j + 1) = t
    enddo
  end subroutine sort_desc

  subroutine set_model()
    real(8), allocatable :: hre(:, :, :)
    if (.not. hm_built) call Hij_build()
    if (any(dimag(hm_Hij) /= 0d0)) stop "ED_INIT_SOLVER ERROR: complex Hij is not supported by the engine"
    allocate(hre(Nspin, Ns, Ns))
    hre = dreal(hm_Hij)
    ! ED_SPARSE_H=T: the stored path's diagonal keeps -xmu*N and the
    ! hfmode -U/2*N terms at any ED_FILLING (stored/H_local.f90:61-90)
    call edl_check(ctx, edl_set_hamiltonian(ctx, Ns, Nspin, Norb, Nsites, hre, c_null_ptr, Uloc, Ust, Jh, Jx, Jp, &
         xmu, merge(1, 0, HFmode), merge(0, ed_filling, ed_sparse_H)), "edl_set_hamiltonian")
    installed_uid = -1
  end subroutine set_model

  !> Releases the engine context (no reference counterpart: the reference
  !> keeps its state in module globals until the process ends).
  subroutine ed_finalize_solver()
    call es_free()
    if (c_associated(ctx)) call edl_check(ctx, edl_ctx_destroy(ctx), "edl_ctx_destroy")
    ctx = c_null_ptr
    solver_ready = .false.
  end subroutine ed_finalize_solver

  !> ed_solve (MAIN.f90:61-108)
  subroutine ed_solve()
    integer :: itemp
    logical :: gf_any, chi_any, oc_any
    if (.not. solver_ready) stop "ED_SOLVE ERROR: call ed_init_solver first"
    if (.not. hm_status) stop "ED_INIT_SOLVER ERROR: Hmatrix is not allocated"
    call ed_Hij_write(unit=LOGfile)
    gf_any = any(gf_flag)
    chi_any = any(chispin_flag)
    oc_any = any(oc_flag)
    call diagonalize_lattice()
    if (gf_any .or. chi_any .or. oc_any) call build_functions(gf_any, chi_any, oc_any)
    do itemp = 1, size(temperature_list)
       temp = temperature_list(itemp)
       if (size(temperature_list) > 1) then
          write(ed_file_suffix, "(A,F0.6)") "_T", temp
          write(LOGfile, "(A)") ""
          write(LOGfile, "(A,F12.6)") "Solving for temperature T=", temp
       endif
       call partition_function_lattice()
       call observables_lattice()
       if (gf_any) call eval_gf_lattice()
       if (chi_any) call eval_chi_lattice()
       if (oc_any) call eval_oc_lattice()
       ed_file_suffix = ""
    enddo
    call es_free()
  end subroutine ed_solve

  ! ======================================================================
  ! DIAG (DIAG.f90:103-332) and the state list (EIGENSPACE.f90)
  ! ======================================================================

  subroutine diagonalize_lattice()
    integer :: isector, nup, ndw, ierr, neigen, niter, nsteps, i, unit, iter, ist
    integer(c_int64_t) :: dim
    integer(c_int32_t) :: dimup, dimdw
    real(8) :: oldzero, enemin, ec, ei
    real(8), allocatable :: ev(:), alpha(:), beta(:)
    real(8), allocatable, target :: vecs(:)
    logical :: tflag, lanc_solve
    call es_free()
    oldzero = 1000d0
    iter = 0
    write(LOGfile, "(A)") "Diagonalize H:"
    do isector = 1, Nsectors
       ierr = edl_sector_qn(Ns, isector, nup, ndw)
       if (sector_nup >= 0 .and. sector_ndw >= 0) then
          if (nup /= sector_nup .or. ndw /= sector_ndw) cycle
       endif
       if (ed_filling /= 0 .and. nup + ndw /= ed_filling) cycle
       if (ed_twin .and. nup < ndw) cycle            ! twin_mask (SETUP.f90:254-262)
       iter = iter + 1
       tflag = ed_twin .and. nup /= ndw
       call edl_check(ctx, edl_sector_build(ctx, isector, dim, dimup, dimdw), "build_sector")
       if (trim(lanc_method) == "lanczos") then
          neigen = 1
       else
          neigen = int(min(int(lanc_nstates_sector, c_int64_t), dim))
       endif
       niter = int(min(int(lanc_niter, c_int64_t), dim))
       lanc_solve = neigen /= dim .and. dim > lanc_dim_threshold
       allocate(ev(neigen), vecs(dim * neigen))
       if (neigen == 1) then
          allocate(alpha(niter), beta(niter))
          call edl_check(ctx, edl_lanczos_gs(ctx, isector, prec, niter, 0d0, lanc_seed, 1d-13, ev(1), nsteps, &
               alpha, beta, c_loc(vecs)), "sp_lanc_eigh")
          deallocate(alpha, beta)
       else
          call edl_check(ctx, edl_lanczos_states(ctx, isector, merge(EDL_FP32, EDL_FP64, prec == EDL_FP32), &
               neigen, niter, 1d-13, lanc_seed, ev, c_loc(vecs)), "sp_eigh")
       endif
       installed_uid = -1
       if (ed_verbose > 2) write(LOGfile, "(1X,I9,A,I9,A6,I3,A6,I3,A7,I20,A,20F16.10)") iter, "-Solving sector:", &
            isector, ", nup:", nup, ", ndw:", ndw, ", dim=", dim, "  E:", ev
       do i = 1, neigen
          if (finiteT) then
             call es_add_state(ev(i), vecs((i - 1) * dim + 1:i * dim), isector, tflag, lanc_nstates_total)
          else
             enemin = ev(i)
             if (enemin < oldzero - 10d0 * gs_threshold) then
                oldzero = enemin
                call es_free()
                call es_add_state(enemin, vecs((i - 1) * dim + 1:i * dim), isector, tflag)
             elseif (abs(enemin - oldzero) <= gs_threshold) then
                oldzero = min(oldzero, enemin)
                call es_add_state(enemin, vecs((i - 1) * dim + 1:i * dim), isector, tflag)
             endif
          endif
       enddo
       ! print_eigenvalues_list (DIAG.f90:538-556)
       open(newunit=unit, file="eigenvalues_list.ed", position="append", action="write")
       if (lanc_solve) then
          write(unit, "(A9,A15)") " # Sector", "Indices"
       else
          write(unit, "(A10,A15)") " #X Sector", "Indices"
       endif
       write(unit, "(I9,2I6)") isector, nup, ndw
       do i = 1, neigen
          write(unit, *) ev(i)
       enddo
       write(unit, *) ""
       close(unit)
       deallocate(ev, vecs)
    enddo
    if (ssize == 0) stop "ED_DIAG ERROR: no sector was solved"
    call save_state_list()
    if (ed_verbose >= 2) call print_state_list(LOGfile)
    do ist = 1, es_gs_degeneracy()
       ierr = edl_sector_qn(Ns, slist(ist)%sector, nup, ndw)
       write(LOGfile, "(A,F20.12,2I4)") "Egs =", slist(ist)%e, nup, ndw
    enddo
    if (finiteT) then
       ! DIAG.f90:306-330: enough states for the cutoff at the largest T,
       ! then drop those beyond it
       ec = s_emax
       if (exp(-(ec - s_emin) / temp) > cutoff) &
            stop "ED_DIAG:  exp(-(Ei-Egs)/T)>Cutoff condition not met, try increasing lanc_nstates_sector"
       ei = slist(ssize)%e
       do while (exp(-(ei - s_emin) / temp) <= cutoff)
          call es_pop_state()
          ei = slist(ssize)%e
       enddo
       write(LOGfile, "(A,I4)") "Adjusting lanc_nstates_total to:", ssize
    endif
  end subroutine diagonalize_lattice

  !> es_add_state / es_insert_state (EIGENSPACE.f90:167-247): sorted
  !> insertion before the first entry of equal or higher energy; with
  !> `twin` the spin-flipped copy follows; with `size` the list keeps its
  !> lowest entries.
  subroutine es_add_state(e, vec, isector, twin, size)
    real(8), intent(in) :: e, vec(:)
    integer, intent(in) :: isector
    logical, intent(in) :: twin
    integer, intent(in), optional :: size
    if (present(size)) then
       if (ssize < size) then
          call es_insert(e, vec, isector, twin)
       elseif (e < slist(ssize)%e) then
          call es_pop_state()
          call es_insert(e, vec, isector, twin)
       endif
    else
       call es_insert(e, vec, isector, twin)
    endif
  end subroutine es_add_state

  subroutine es_insert(e, vec, isector, twin)
    real(8), intent(in) :: e, vec(:)
    integer, intent(in) :: isector
    logical, intent(in) :: twin
    type(estate_t), allocatable :: tmp(:)
    integer :: pos, n, k, nup, ndw, ierr
    n = merge(2, 1, twin)
    if (.not. allocated(slist)) allocate(slist(64))
    if (ssize + n > size(slist)) then
       allocate(tmp(2 * (ssize + n)))
       do k = 1, ssize
          call move_estate(slist(k), tmp(k))
       enddo
       call move_alloc(tmp, slist)
    endif
    pos = ssize + 1
    do k = 1, ssize
       if (e <= slist(k)%e) then
          pos = k
          exit
       endif
    enddo
    do k = ssize, pos, -1
       call move_estate(slist(k), slist(k + n))
    enddo
    slist(pos)%e = e
    slist(pos)%sector = isector
    slist(pos)%itwin = .false.
    next_uid = next_uid + 1
    slist(pos)%uid = next_uid
    if (allocated(slist(pos)%v)) deallocate(slist(pos)%v)
    allocate(slist(pos)%v(size(vec)))
    slist(pos)%v = vec
    if (twin) then
       ierr = edl_sector_qn(Ns, isector, nup, ndw)
       slist(pos + 1)%e = e
       slist(pos + 1)%sector = edl_sector_index(Ns, ndw, nup)   ! get_twin_sector
       slist(pos + 1)%itwin = .true.
       next_uid = next_uid + 1
       slist(pos + 1)%uid = next_uid
       if (allocated(slist(pos + 1)%v)) deallocate(slist(pos + 1)%v)
    endif
    ssize = ssize + n
    s_emin = slist(1)%e
    s_emax = max(s_emax, e)
    if (ssize == n) s_emax = e
  end subroutine es_insert

  subroutine move_estate(src, dst)
    type(estate_t), intent(inout) :: src, dst
    dst%e = src%e
    dst%sector = src%sector
    dst%itwin = src%itwin
    dst%uid = src%uid
    if (allocated(dst%v)) deallocate(dst%v)
    if (allocated(src%v)) call move_alloc(src%v, dst%v)
  end subroutine move_estate

  !> es_pop_state without n (EIGENSPACE.f90:256-330): drop the last entry,
  !> and its partner when the last is a twin.
  subroutine es_pop_state()
    integer :: n
    if (ssize == 0) stop "es_pop_state: empty list"
    n = 1
    if (slist(ssize)%itwin) n = 2
    if (allocated(slist(ssize)%v)) deallocate(slist(ssize)%v)
    if (n == 2) then
       if (allocated(slist(ssize - 1)%v)) deallocate(slist(ssize - 1)%v)
    endif
    ssize = ssize - n
    if (ssize > 0) then
       s_emax = slist(ssize)%e
       s_emin = slist(1)%e
    endif
  end subroutine es_pop_state

  subroutine es_free()
    integer :: k
    if (allocated(slist)) then
       do k = 1, size(slist)
          if (allocated(slist(k)%v)) deallocate(slist(k)%v)
       enddo
    endif
    ssize = 0
    strimd = 0
    s_emin = 0d0
    s_emax = 0d0
  end subroutine es_free

  !> es_return_gs_degeneracy (EIGENSPACE.f90:340-369)
  integer function es_gs_degeneracy() result(numzero)
    real(8) :: oldzero, enemin
    integer :: k
    oldzero = 1000d0
    numzero = 0
    do k = 1, ssize
       enemin = slist(k)%e
       if (enemin < oldzero - 10d0 * gs_threshold) then
          numzero = 1
          oldzero = enemin
       elseif (abs(enemin - oldzero) <= gs_threshold) then
          numzero = numzero + 1
          oldzero = min(oldzero, enemin)
       endif
    enddo
  end function es_gs_degeneracy

  !> es_trim_size (EIGENSPACE.f90:440-475)
  subroutine es_trim_size()
    integer :: k
    strimd = ssize
    if (.not. finiteT) return
    do k = 1, ssize
       if (exp(-(slist(k)%e - s_emin) / temp) <= cutoff) then
          strimd = k - 1
          exit
       endif
    enddo
    if (strimd == 0) stop "es_return_size: list does not contain any node"
  end subroutine es_trim_size

  !> Sets entry k as the engine's installed state (es_return_cvector,
  !> EIGENSPACE.f90:479-512): a twin entry is the spin flip of its partner,
  !> V'(idw, iup) = V(iup, idw), the transposed matrix of the partner.
  subroutine install(k)
    integer, intent(in) :: k
    integer(c_int64_t) :: dim
    integer(c_int32_t) :: dimup, dimdw
    real(8), allocatable :: w(:)
    if (slist(k)%uid == installed_uid) return
    if (.not. slist(k)%itwin) then
       call edl_check(ctx, edl_sector_build(ctx, slist(k)%sector, dim, dimup, dimdw), "build_sector")
       call edl_check(ctx, edl_set_state(ctx, slist(k)%sector, slist(k)%e, slist(k)%v), "es_return_cvector")
    else
       call edl_check(ctx, edl_sector_build(ctx, slist(k - 1)%sector, dim, dimup, dimdw), "build_sector")
       allocate(w(dim))
       w = reshape(transpose(reshape(slist(k - 1)%v, [int(dimup), int(dimdw)])), [dim])
       call edl_check(ctx, edl_sector_build(ctx, slist(k)%sector, dim, dimup, dimdw), "build_sector")
       call edl_check(ctx, edl_set_state(ctx, slist(k)%sector, slist(k)%e, w), "es_return_cvector")
    endif
    installed_uid = slist(k)%uid
  end subroutine install

  !> save_state_list (DIAG.f90:495-509)
  subroutine save_state_list()
    integer :: unit, k, nup, ndw, ierr
    open(newunit=unit, file="state_list.ed", action="write")
    do k = 1, ssize
       ierr = edl_sector_qn(Ns, slist(k)%sector, nup, ndw)
       write(unit, "(i8,i12,2i8)") k, slist(k)%sector, nup, ndw
    enddo
    close(unit)
  end subroutine save_state_list

  !> print_state_list (DIAG.f90:512-530)
  subroutine print_state_list(unit)
    integer, intent(in) :: unit
    integer :: k, nup, ndw, ierr
    integer(c_int64_t) :: dim
    integer(c_int32_t) :: dimup, dimdw
    write(unit, "(A1,A6,A18,2x,A19,1x,2A10,A12)") "#", "i", "E_i", "exp(-(E-E0)/T)", "Sect", "Dim", "Indices:"
    do k = 1, ssize
       ierr = edl_sector_qn(Ns, slist(k)%sector, nup, ndw)
       call edl_check(ctx, edl_sector_build(ctx, slist(k)%sector, dim, dimup, dimdw), "build_sector")
       write(unit, "(i6,f18.12,2x,ES19.12,1x,2I10,2I4)") k, slist(k)%e, exp(-(slist(k)%e - s_emin) / temp), &
            slist(k)%sector, int(dim), nup, ndw
    enddo
  end subroutine print_state_list

  !> partition_function_lattice (DIAG.f90:54-91)
  subroutine partition_function_lattice()
    integer :: k
    zeta_function = 0d0
    if (finiteT) then
       call es_trim_size()
       do k = 1, strimd
          zeta_function = zeta_function + exp(-(slist(k)%e - s_emin) / temp)
       enddo
    else
       zeta_function = dble(ssize)
    endif
    if (ed_verbose >= 2) write(LOGfile, "(A,F20.12)") "Z   =", zeta_function
  end subroutine partition_function_lattice

  real(8) function boltzmann(k) result(w)
    integer, intent(in) :: k
    w = 1d0
    if (finiteT) w = exp(-(slist(k)%e - s_emin) / temp)
    w = w / zeta_function
  end function boltzmann

  ! ======================================================================
  ! OBSERVABLES (OBSERVABLES.f90:78-192,318-352) and ENERGY (ENERGY.f90:70-353)
  ! ======================================================================

  subroutine observables_lattice()
    integer :: k, unit, io, jo
    real(8) :: du(Ns), dd(Ns), dc(Ns), sz2(Ns, Ns), n2(Ns, Ns), csz(Ns, Ns), cn(Ns, Ns), e(11), w
    ed_dens_up = 0d0; ed_dens_dw = 0d0; ed_docc = 0d0
    sz2 = 0d0; n2 = 0d0; ed_energy = 0d0
    call es_trim_size()
    do k = 1, strimd
       call install(k)
       w = boltzmann(k)
       call edl_check(ctx, edl_observables(ctx, du, dd, dc), "lanc_observables")
       call edl_check(ctx, edl_correlations(ctx, csz, cn), "lanc_observables")
       call edl_check(ctx, edl_local_energy(ctx, e), "lanc_energy")
       ed_dens_up = ed_dens_up + w * du
       ed_dens_dw = ed_dens_dw + w * dd
       ed_docc = ed_docc + w * dc
       sz2 = sz2 + w * csz
       n2 = n2 + w * cn
       ed_energy = ed_energy + w * e
    enddo
    ed_dens = ed_dens_up + ed_dens_dw
    ed_mag = ed_dens_up - ed_dens_dw
    ! write_observables (OBSERVABLES.f90:318-352): every file appended
    open(newunit=unit, file="parameters_last.ed", position="append", action="write")
    write(unit, "(90F15.9)") xmu, temp, Uloc(1:Norb), Ust, Jh, Jx, Jp
    close(unit)
    open(newunit=unit, file="dens.ed", position="append", action="write")
    write(unit, "(90(F20.12,1X))") (ed_dens(io), io = 1, Ns)
    close(unit)
    open(newunit=unit, file="docc.ed", position="append", action="write")
    write(unit, "(90(F20.12,1X))") (ed_docc(io), io = 1, Ns)
    close(unit)
    open(newunit=unit, file="magz.ed", position="append", action="write")
    write(unit, "(90(F20.12,1X))") (ed_mag(io), io = 1, Ns)
    close(unit)
    open(newunit=unit, file="egs.ed", position="append", action="write")
    write(unit, *) s_emin
    close(unit)
    open(newunit=unit, file="Sz_corr.ed", position="append", action="write")
    do io = 1, Ns
       write(unit, "(90(F15.9,1X))") (sz2(io, jo), jo = 1, io)
    enddo
    close(unit)
    open(newunit=unit, file="N_corr.ed", position="append", action="write")
    do io = 1, Ns
       write(unit, "(90(F15.9,1X))") (n2(io, jo), jo = 1, io)
    enddo
    close(unit)
    write(LOGfile, "(A,100f15.8,f15.8)") "dens=", (ed_dens(io), io = 1, Ns), sum(ed_dens)
    if (Nspin == 2) write(LOGfile, "(A,10f18.12,A)") "magZ=", (ed_mag(io), io = 1, Ns)
    ! write_energy (ENERGY.f90:635-659)
    open(newunit=unit, file="energy_info.ed", action="write")
    write(unit, "(A1,90(A14,1X))") "#", "1<K>", "2<Hi>", "3<V>=<Hi-Ehf>", "4<Eloc>", "5<Ehf>", "6<Dst>", &
         "7<Dnd>", "8<Dse>", "9<Dph>", "10<Dkxy>", "11<Dkz>"
    close(unit)
    open(newunit=unit, file="energy_last.ed", position="append", action="write")
    write(unit, "(90F15.9)") ed_energy
    close(unit)
  end subroutine observables_lattice

  ! ======================================================================
  ! GF / CHI / OC build (GF_ELECTRONS.f90:35-94,202-457;
  ! CHI_SPIN_ELECTRONS.f90:45-95,189-366; OC_ELECTRONS.f90:103-259)
  ! ======================================================================

  !> For every state of the list, all its seeds (c^dg and c of every
  !> flagged site and spin, the mixed pairs, S^z, the currents) run as one
  !> batched call; poles and weights go to impGmatrix, SpinChiMatrix and
  !> OcMatrix (the reference's build_*_lattice), whose restart files are
  !> then written.
  subroutine build_functions(gf_any, chi_any, oc_any)
    logical, intent(in) :: gf_any, chi_any, oc_any
    integer :: ist, nseed, s, ispin, iorb, jorb, isite, jsite, io, jo, kind_s, nl, ierr
    type(edl_seed_desc), allocatable :: seeds(:)
    integer, allocatable :: role(:, :)       ! (kind, a, b, c) per seed: 1 GF, 2 chi, 3 OC
    real(8), allocatable :: alpha(:), beta(:), norm2(:), poles(:), weights(:)
    integer(c_int), allocatable :: nsteps(:), jsec(:)
    integer(c_int64_t) :: dimj
    integer(c_int32_t) :: du_, dd_
    integer :: ich, isign, maxseed
    logical :: gfsel(Ns, Ns)
    write(LOGfile, "(A)") "Build lattice functions:"
    maxseed = 2 * Nspin * Ns * Ns + Ns * Ns + Norb
    allocate(seeds(maxseed), role(4, maxseed))
    nseed = 0
    gfsel = .false.
    if (gf_any) then
       do ispin = 1, Nspin
          do iorb = 1, Norb
             if (.not. gf_flag(iorb)) cycle
             do isite = 1, Nsites(iorb)
                io = pack_indices(isite, iorb)
                gfsel(io, io) = .true.
                call add_seed(EDL_OP_CDG, ispin, io, 0, 1, 1)
                call add_seed(EDL_OP_C, ispin, io, 0, 1, 2)
             enddo
          enddo
       enddo
       if (offdiag_gf_flag) then
          do ispin = 1, Nspin
             do iorb = 1, Norb
                if (.not. gf_flag(iorb)) cycle
                do jorb = 1, Norb
                   if (.not. gf_flag(jorb)) cycle
                   do isite = 1, Nsites(iorb)
                      do jsite = 1, Nsites(jorb)
                         io = pack_indices(isite, iorb)
                         jo = pack_indices(jsite, jorb)
                         ! the mixed seed is symmetric in (io, jo): run io < jo, copy to (jo, io)
                         if (io == jo) cycle
                         gfsel(io, jo) = .true.
                         if (io > jo) cycle
                         call add_seed(EDL_OP_CDG, ispin, io, jo, 1, 1)
                         call add_seed(EDL_OP_C, ispin, io, jo, 1, 2)
                      enddo
                   enddo
                enddo
             enddo
          enddo
       endif
       do ispin = 1, Nspin
          do io = 1, Ns
             do jo = 1, Ns
                if (gfsel(io, jo)) call alloc_gfm(impGmatrix(ispin, io, jo), 2)
             enddo
          enddo
       enddo
    endif
    if (chi_any) then
       do iorb = 1, Norb
          if (.not. chispin_flag(iorb)) cycle
          do isite = 1, Nsites(iorb)
             io = pack_indices(isite, iorb)
             call add_seed(EDL_OP_SZ, 1, io, 0, 2, 1)
             call alloc_gfm(SpinChiMatrix(io, io), 1)
          enddo
       enddo
       if (offdiag_chispin_flag .and. Norb > 1) then
          do iorb = 1, Norb
             do jorb = 1, Norb
                if (.not. chispin_flag(iorb) .and. .not. chispin_flag(jorb)) cycle
                do isite = 1, Nsites(iorb)
                   do jsite = 1, Nsites(jorb)
                      io = pack_indices(isite, iorb)
                      jo = pack_indices(jsite, jorb)
                      if (io == jo) cycle
                      call alloc_gfm(SpinChiMatrix(io, jo), 1)
                      if (io > jo) cycle
                      call add_seed(EDL_OP_SZ, 1, io, jo, 2, 1)
                   enddo
                enddo
             enddo
          enddo
       endif
    endif
    if (oc_any) then
       do iorb = 1, Norb
          if (.not. oc_flag(iorb)) cycle
          call add_seed(EDL_OP_J, 0, iorb, 0, 3, 1)
          call alloc_gfm(OcMatrix(iorb), 1)
       enddo
    endif
    allocate(norm2(nseed), nsteps(nseed), jsec(nseed), alpha(nseed * lanc_ngfiter), beta(nseed * lanc_ngfiter))
    do ist = 1, ssize
       call install(ist)
       call edl_check(ctx, edl_tridiag_batched(ctx, nseed, seeds, lanc_ngfiter, merge(EDL_FP32, EDL_FP64, &
            prec == EDL_FP32), 0, alpha, beta, norm2, nsteps, jsec), "tridiag_Hv_sector")
       do s = 1, nseed
          kind_s = role(1, s)
          ich = role(2, s)
          isign = merge(1, -1, kind_s /= 1 .or. ich == 1)
          if (jsec(s) == 0) then
             ! no target sector: an empty channel (GF_ELECTRONS.f90:254,280)
             allocate(poles(0), weights(0))
          else
             call edl_check(ctx, edl_sector_build(ctx, jsec(s), dimj, du_, dd_), "build_sector")
             nl = int(min(int(lanc_ngfiter, c_int64_t), dimj))
             allocate(poles(nl), weights(nl))
             if (norm2(s) /= 0d0) then
                call edl_check(c_null_ptr, edl_lanczos_poles(nl, alpha((s - 1) * lanc_ngfiter + 1), &
                     beta((s - 1) * lanc_ngfiter + 1), norm2(s), slist(ist)%e, isign, poles, weights), &
                     "add_to_lanczos")
             else
                ! zero seed: alpha = beta = 0, every Ritz value 0, weight 0
                poles = isign * (0d0 - slist(ist)%e)
                weights = 0d0
             endif
          endif
          ispin = seeds(s)%ispin
          io = seeds(s)%io
          jo = merge(io, seeds(s)%jo, seeds(s)%jo == 0)
          select case (kind_s)
          case (1)
             call put_channel(impGmatrix(ispin, io, jo), ist, ich, poles, weights)
             if (jo /= io) call put_channel(impGmatrix(ispin, jo, io), ist, ich, poles, weights)
          case (2)
             call put_channel(SpinChiMatrix(io, jo), ist, ich, poles, weights)
             if (jo /= io) call put_channel(SpinChiMatrix(jo, io), ist, ich, poles, weights)
          case (3)
             call put_channel(OcMatrix(io), ist, ich, poles, weights)
          end select
          deallocate(poles, weights)
       enddo
    enddo
    installed_uid = -1
    ierr = 0
    if (gf_any) call write_gfmatrix3(impGmatrix, "gfmatrix" // trim(ed_file_suffix) // ".restart")
    if (chi_any) call write_gfmatrix2(SpinChiMatrix, "ChiSpinMatrix" // trim(ed_file_suffix) // ".restart")
    if (oc_any) call write_gfmatrix1(OcMatrix, "OcMatrix" // trim(ed_file_suffix) // ".restart")
  contains
    subroutine add_seed(op, ispin_, io_, jo_, kind_, ich_)
      integer, intent(in) :: op, ispin_, io_, jo_, kind_, ich_
      nseed = nseed + 1
      seeds(nseed) = edl_seed_desc(op, ispin_, io_, jo_)
      role(:, nseed) = [kind_, ich_, 0, 0]
    end subroutine add_seed
  end subroutine build_functions

  subroutine alloc_gfm(g, nchan)
    type(gfmatrix_t), intent(inout) :: g
    integer, intent(in) :: nchan
    integer :: ist
    if (allocated(g%state)) deallocate(g%state)
    allocate(g%state(ssize))
    do ist = 1, ssize
       allocate(g%state(ist)%channel(nchan))
    enddo
    g%status = .true.
  end subroutine alloc_gfm

  subroutine put_channel(g, ist, ich, poles, weights)
    type(gfmatrix_t), intent(inout) :: g
    integer, intent(in) :: ist, ich
    real(8), intent(in) :: poles(:), weights(:)
    if (allocated(g%state(ist)%channel(ich)%poles)) deallocate(g%state(ist)%channel(ich)%poles)
    if (allocated(g%state(ist)%channel(ich)%weight)) deallocate(g%state(ist)%channel(ich)%weight)
    allocate(g%state(ist)%channel(ich)%poles(size(poles)), g%state(ist)%channel(ich)%weight(size(weights)))
    g%state(ist)%channel(ich)%poles = poles
    g%state(ist)%channel(ich)%weight = weights
  end subroutine put_channel

  ! write_formatted_gfmatrix (VARS_GLOBAL.f90:587-606) and the writers of
  ! rank 1, 2, 3 arrays (write_gfmatrix_all1/2/3, VARS_GLOBAL.f90:425-471)
  subroutine write_formatted_gfmatrix(g, unit)
    type(gfmatrix_t), intent(in) :: g
    integer, intent(in) :: unit
    integer :: ist, ich
    write(unit, *) g%status
    if (.not. g%status) return
    write(unit, *) size(g%state)
    do ist = 1, size(g%state)
       write(unit, *) size(g%state(ist)%channel)
       do ich = 1, size(g%state(ist)%channel)
          write(unit, *) size(g%state(ist)%channel(ich)%poles)
          write(unit, *) g%state(ist)%channel(ich)%poles
          write(unit, *) g%state(ist)%channel(ich)%weight
       enddo
    enddo
    write(unit, *) ""
  end subroutine write_formatted_gfmatrix

  subroutine write_gfmatrix1(g, file)
    type(gfmatrix_t), intent(in) :: g(:)
    character(len=*), intent(in) :: file
    integer :: unit, i1
    open(newunit=unit, file=file, action="write")
    do i1 = 1, size(g)
       call write_formatted_gfmatrix(g(i1), unit)
    enddo
    close(unit)
  end subroutine write_gfmatrix1

  subroutine write_gfmatrix2(g, file)
    type(gfmatrix_t), intent(in) :: g(:, :)
    character(len=*), intent(in) :: file
    integer :: unit, i1, i2
    open(newunit=unit, file=file, action="write")
    do i1 = 1, size(g, 1)
       do i2 = 1, size(g, 2)
          call write_formatted_gfmatrix(g(i1, i2), unit)
       enddo
    enddo
    close(unit)
  end subroutine write_gfmatrix2

  subroutine write_gfmatrix3(g, file)
    type(gfmatrix_t), intent(in) :: g(:, :, :)
    character(len=*), intent(in) :: file
    integer :: unit, i1, i2, i3
    open(newunit=unit, file=file, action="write")
    do i1 = 1, size(g, 1)
       do i2 = 1, size(g, 2)
          do i3 = 1, size(g, 3)
             call write_formatted_gfmatrix(g(i1, i2, i3), unit)
          enddo
       enddo
    enddo
    close(unit)
  end subroutine write_gfmatrix3

  ! ======================================================================
  ! GF / CHI / OC evaluation at the current temperature
  ! ======================================================================

  !> allocate_grids (AUX_FUNX.f90:357-373); linspace end points included.
  subroutine allocate_grids()
    integer :: i
    Ltau = max(int(1d0 / temp), Ltau)
    if (allocated(wm)) deallocate(wm, wr, vr, tau)
    allocate(wm(Lmats), wr(Lreal), vr(Lreal), tau(0:Ltau))
    do i = 1, Lmats
       wm(i) = acos(-1d0) * temp * (2 * i - 1)
    enddo
    do i = 1, Lreal
       wr(i) = wini + (wfin - wini) * dble(i - 1) / dble(max(Lreal - 1, 1))
       vr(i) = wfin * dble(i - 1) / dble(max(Lreal - 1, 1))
    enddo
    do i = 0, Ltau
       tau(i) = (1d0 / temp) * dble(i) / dble(Ltau)
    enddo
  end subroutine allocate_grids

  !> eval_gf_lattice (GREENS_FUNCTIONS.f90:38-51), lanc_eval_gf_electrons
  !> (GF_ELECTRONS.f90:472-512) and the off-diagonal combination
  !> G_ij = (G_mix - G_ii - G_jj)/2 (GF_ELECTRONS.f90:160-180).
  subroutine eval_gf_lattice()
    integer :: ispin, io, jo, k, ich, j
    real(8) :: pesobz
    call allocate_grids()
    impGmats = (0d0, 0d0)
    impGreal = (0d0, 0d0)
    call es_trim_size()
    do ispin = 1, Nspin
       do io = 1, Ns
          do jo = 1, Ns
             if (.not. impGmatrix(ispin, io, jo)%status) cycle
             do k = 1, strimd
                pesobz = boltzmann(k)
                do ich = 1, size(impGmatrix(ispin, io, jo)%state(k)%channel)
                   associate(ch => impGmatrix(ispin, io, jo)%state(k)%channel(ich))
                     do j = 1, size(ch%poles)
                        impGmats(ispin, io, jo, :) = impGmats(ispin, io, jo, :) + pesobz * ch%weight(j) / &
                             (dcmplx(0d0, wm) - ch%poles(j))
                        impGreal(ispin, io, jo, :) = impGreal(ispin, io, jo, :) + pesobz * ch%weight(j) / &
                             (dcmplx(wr, eps) - ch%poles(j))
                     enddo
                   end associate
                enddo
             enddo
          enddo
       enddo
    enddo
    if (offdiag_gf_flag) then
       do ispin = 1, Nspin
          do io = 1, Ns
             do jo = 1, Ns
                if (io == jo .or. .not. impGmatrix(ispin, io, jo)%status) cycle
                impGmats(ispin, io, jo, :) = 0.5d0 * (impGmats(ispin, io, jo, :) - impGmats(ispin, io, io, :) - &
                     impGmats(ispin, jo, jo, :))
                impGreal(ispin, io, jo, :) = 0.5d0 * (impGreal(ispin, io, jo, :) - impGreal(ispin, io, io, :) - &
                     impGreal(ispin, jo, jo, :))
             enddo
          enddo
       enddo
    endif
    if (ed_print_G) call print_impG()
  end subroutine eval_gf_lattice

  !> ed_write_func (IO.f90:159-261): iprint 1 (diagonal) or 3 (all pairs).
  subroutine print_impG()
    integer :: ispin, iorb, jorb, ilat, jlat, io, jo
    character(len=256) :: fname
    do ispin = 1, Nspin
       do iorb = 1, Norb
          if (.not. gf_flag(iorb)) cycle
          if (.not. offdiag_gf_flag) then
             do ilat = 1, Nsites(iorb)
                io = pack_indices(ilat, iorb)
                write(fname, "(A,I4.4,A,I0,A,I0)") "G_i", ilat, "_l", iorb, "_s", ispin
                call splot(trim(fname) // "_iw" // trim(ed_file_suffix) // ".ed", wm, impGmats(ispin, io, io, :))
                call splot(trim(fname) // "_realw" // trim(ed_file_suffix) // ".ed", wr, impGreal(ispin, io, io, :))
             enddo
          else
             do jorb = 1, Norb
                if (.not. gf_flag(jorb)) cycle
                do ilat = 1, Nsites(iorb)
                   do jlat = 1, Nsites(jorb)
                      io = pack_indices(ilat, iorb)
                      jo = pack_indices(jlat, jorb)
                      write(fname, "(A,I4.4,A,I4.4,A,I0,A,I0,A,I0)") "G_i", ilat, "j", jlat, "_l", iorb, "m", jorb, &
                           "_s", ispin
                      call splot(trim(fname) // "_iw" // trim(ed_file_suffix) // ".ed", wm, impGmats(ispin, io, jo, :))
                      call splot(trim(fname) // "_realw" // trim(ed_file_suffix) // ".ed", wr, &
                           impGreal(ispin, io, jo, :))
                   enddo
                enddo
             enddo
          endif
       enddo
    enddo
  end subroutine print_impG

  !> eval_chi_lattice (CHI_FUNCTIONS.f90:45-73) with
  !> lanc_eval_spinChi_electrons (CHI_SPIN_ELECTRONS.f90:382-468): chi(tau)
  !> with the Ltau/2 symmetrisation and chi(w) on vr = [0, wfin]; chiT.ed
  !> gets T and trapz(chi_ii(tau), 0, beta) per site.  Every (io, jo) keeps
  !> its own chi (the reference assigns its whole temporary arrays per pair,
  !> CHI_SPIN_ELECTRONS.f90:455-465, so that only the last pair survives).
  subroutine eval_chi_lattice()
    integer :: io, jo, k, j, i, unit, iorb, ilat, jorb, jlat
    real(8) :: pesobz, peso, de, beta, h
    real(8) :: chiT(Ns)
    character(len=256) :: fname
    call allocate_grids()
    if (allocated(spinChi_tau)) deallocate(spinChi_tau, spinChi_w)
    allocate(spinChi_tau(Ns, Ns, 0:Ltau), spinChi_w(Ns, Ns, Lreal))
    spinChi_tau = 0d0
    spinChi_w = (0d0, 0d0)
    beta = 1d0 / temp
    call es_trim_size()
    do io = 1, Ns
       do jo = 1, Ns
          if (.not. SpinChiMatrix(io, jo)%status) cycle
          do k = 1, strimd
             pesobz = boltzmann(k)
             associate(ch => SpinChiMatrix(io, jo)%state(k)%channel(1))
               do j = 1, size(ch%poles)
                  peso = ch%weight(j) * pesobz
                  de = ch%poles(j)
                  do i = 0, Ltau / 2 - 1
                     spinChi_tau(io, jo, i) = spinChi_tau(io, jo, i) + peso * exp(-tau(i) * de)
                  enddo
                  spinChi_tau(io, jo, Ltau / 2) = spinChi_tau(io, jo, Ltau / 2) + peso * 0.5d0 * &
                       (exp(-tau(Ltau / 2) * de) + exp(-(beta - tau(Ltau / 2)) * de))
                  do i = Ltau / 2 + 1, Ltau
                     spinChi_tau(io, jo, i) = spinChi_tau(io, jo, i) + peso * exp(-(beta - tau(i)) * de)
                  enddo
                  spinChi_w(io, jo, :) = spinChi_w(io, jo, :) - peso * (1d0 - exp(-beta * de)) * &
                       (1d0 / (dcmplx(vr, eps) - de) - 1d0 / (dcmplx(vr, eps) + de))
               enddo
             end associate
          enddo
       enddo
    enddo
    if (offdiag_chispin_flag .and. Norb > 1) then
       do io = 1, Ns
          do jo = 1, Ns
             if (io == jo .or. .not. SpinChiMatrix(io, jo)%status) cycle
             if (io > jo) cycle
             spinChi_w(io, jo, :) = 0.5d0 * (spinChi_w(io, jo, :) - spinChi_w(io, io, :) - spinChi_w(jo, jo, :))
             spinChi_tau(io, jo, :) = 0.5d0 * (spinChi_tau(io, jo, :) - spinChi_tau(io, io, :) - spinChi_tau(jo, jo, :))
             spinChi_w(jo, io, :) = spinChi_w(io, jo, :)
             spinChi_tau(jo, io, :) = spinChi_tau(io, jo, :)
          enddo
       enddo
    endif
    ! print_chi_spin (IO.f90:112-149)
    do iorb = 1, Norb
       if (.not. chispin_flag(iorb)) cycle
       do ilat = 1, Nsites(iorb)
          io = pack_indices(ilat, iorb)
          write(fname, "(A,I4.4,A,I0)") "spinChi_i", ilat, "_l", iorb
          call splot_r(trim(fname) // "_tau" // trim(ed_file_suffix) // ".ed", tau, spinChi_tau(io, io, 0:))
       enddo
    enddo
    if (offdiag_chispin_flag .and. Norb > 1) then
       do iorb = 1, Norb
          do jorb = 1, Norb
             if (.not. chispin_flag(iorb) .and. .not. chispin_flag(jorb)) cycle
             do ilat = 1, Nsites(iorb)
                do jlat = 1, Nsites(jorb)
                   io = pack_indices(ilat, iorb)
                   jo = pack_indices(jlat, jorb)
                   if (io == jo) cycle
                   write(fname, "(A,I4.4,A,I4.4,A,I0,A,I0)") "spinChi_i", ilat, "j", jlat, "_l", iorb, "m", jorb
                   call splot_r(trim(fname) // "_tau" // trim(ed_file_suffix) // ".ed", tau, spinChi_tau(io, jo, 0:))
                enddo
             enddo
          enddo
       enddo
    endif
    ! chiT.ed: trapz on the uniform tau grid (CHI_FUNCTIONS.f90:57-67)
    chiT = 0d0
    h = beta / dble(Ltau)
    do iorb = 1, Norb
       if (.not. chispin_flag(iorb)) cycle
       do ilat = 1, Nsites(iorb)
          io = pack_indices(ilat, iorb)
          chiT(io) = h * (sum(spinChi_tau(io, io, :)) - 0.5d0 * (spinChi_tau(io, io, 0) + spinChi_tau(io, io, Ltau)))
       enddo
    enddo
    open(newunit=unit, file="chiT.ed", position="append", action="write")
    write(unit, *) temp, (chiT(io), io = 1, Ns)
    close(unit)
  end subroutine eval_chi_lattice

  !> eval_oc_lattice (OC_FUNCTIONS.f90:45-89) with lanc_ed_eval_oc
  !> (OC_ELECTRONS.f90:265-305): D = sum peso/dE over the poles past the
  !> cutoffs, Drude = -pi/Nsites (<K> + 2 D), resistivity -1/Drude.
  subroutine eval_oc_lattice()
    integer :: iorb, k, j, unit
    real(8) :: pesobz, peso, de, d(Norb), res(Norb)
    call allocate_grids()
    Drude_weight = 0d0
    OptCond_w = (0d0, 0d0)
    res = 0d0
    call es_trim_size()
    do iorb = 1, Norb
       if (.not. OcMatrix(iorb)%status) cycle
       do k = 1, strimd
          pesobz = boltzmann(k)
          associate(ch => OcMatrix(iorb)%state(k)%channel(1))
            do j = 1, size(ch%poles)
               peso = ch%weight(j) * pesobz
               de = ch%poles(j)
               if (de == 0d0) cycle
               if (abs(de) < cutoff) cycle
               if (abs(peso / de) < cutoff) cycle
               Drude_weight(iorb) = Drude_weight(iorb) + peso / de
               OptCond_w(iorb, :) = OptCond_w(iorb, :) + peso / de * eps / ((vr - de)**2 + eps**2)
            enddo
          end associate
       enddo
    enddo
    d = Drude_weight
    do iorb = 1, Norb
       if (.not. oc_flag(iorb)) cycle
       Drude_weight(iorb) = -acos(-1d0) / Nsites(iorb) * (ed_energy(1) + 2d0 * d(iorb))
       res(iorb) = -1d0 / Drude_weight(iorb)
    enddo
    open(newunit=unit, file="drude.ed", position="append", action="write")
    write(unit, *) (Drude_weight(iorb), iorb = 1, Norb)
    close(unit)
    open(newunit=unit, file="resistivity.ed", position="append", action="write")
    write(unit, *) (res(iorb), iorb = 1, Norb)
    close(unit)
    open(newunit=unit, file="jjcorr.ed", position="append", action="write")
    write(unit, *) (d(iorb), iorb = 1, Norb)
    close(unit)
    write(LOGfile, "(A,3F18.10)") "Drude, resistivity, D_jj =", Drude_weight(1), res(1), d(1)
  end subroutine eval_oc_lattice

  ! splot (SciFortran): x, Im f, Re f for complex data; x, f for real data
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

  subroutine splot_r(fname, x, f)
    character(len=*), intent(in) :: fname
    real(8), intent(in) :: x(0:)
    real(8), intent(in) :: f(0:)
    integer :: unit, i
    open(newunit=unit, file=fname, action="write")
    do i = 0, size(x) - 1
       write(unit, "(2ES24.15)") x(i), f(i)
    enddo
    close(unit)
  end subroutine splot_r

  ! ======================================================================
  ! IO accessors (IO.f90:40-79)
  ! ======================================================================

  subroutine ed_get_gimp_matsubara(Func)
    complex(8), intent(inout) :: Func(:, :, :, :)
    Func = impGmats
  end subroutine ed_get_gimp_matsubara

  subroutine ed_get_gimp_realaxis(Func)
    complex(8), intent(inout) :: Func(:, :, :, :)
    Func = impGreal
  end subroutine ed_get_gimp_realaxis

  subroutine ed_get_dens(dens)
    real(8), intent(inout) :: dens(:)
    dens = ed_dens
  end subroutine ed_get_dens

  subroutine ed_get_mag(mag)
    real(8), intent(inout) :: mag(:)
    mag = ed_dens_up - ed_dens_dw
  end subroutine ed_get_mag

  subroutine ed_get_docc(docc)
    real(8), intent(inout) :: docc(:)
    docc = ed_docc
  end subroutine ed_get_docc

  !> [Dust, Dund, Dse, Dph, Dkxy, Dkz]
  subroutine ed_get_doubles(docc)
    real(8), intent(inout) :: docc(6)
    docc = ed_energy(6:11)
  end subroutine ed_get_doubles

  !> [Ekin, Epot, Eint, Ehartree, Eknot]; Eint is never assigned by the
  !> reference's solver (VARS_GLOBAL.f90:183) and is 0 here.
  subroutine ed_get_energy(eimp)
    real(8), intent(inout) :: eimp(5)
    eimp = [ed_energy(1), ed_energy(2), 0d0, ed_energy(5), ed_energy(4)]
  end subroutine ed_get_energy

end module EDLAT
