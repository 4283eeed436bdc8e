!> edlat_gpu: a user driver of the EDLAT module, the shape of the reference's
!! drivers (README.md usage: ed_read_input, ed_Hij_init, ed_Hij_read,
!! ed_init_solver, ed_solve; EDlattice.f90:1-51, MAIN.f90:34-108).
!!
!! Usage: `edlat_gpu [input file]` (default inputED.conf).  The Hij links
!! are read from HIJ_FILE (default hij.in) in the Hij_read format
!! `i j a b spin re im` (GRAPH_MATRIX.f90:280-301).
program edlat_gpu
  use EDLAT
  implicit none
  character(len=256) :: input_file

  call get_command_argument(1, input_file)
  if (len_trim(input_file) == 0) input_file = "inputED.conf"
  call ed_read_input(trim(input_file))
  call ed_Hij_init(Nsites(1:Norb), Nspin)
  call ed_Hij_read(trim(hij_file))
  call ed_init_solver()
  call ed_solve()
  call ed_finalize_solver()
end program edlat_gpu
