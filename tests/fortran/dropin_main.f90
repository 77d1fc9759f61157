! dropin_main.f90 -- the reference main program's hot path (main.f90:29-66)
! with `use assembly_mod` / `use solver_mod` replaced by `use quadesth_gpu`.
! Test program: built against the reference's own data_mod / stress_mod /
! io_mod by `make -C oracle dropin`.
!
! usage: dropin DECK OUTFILE RESULTS [fp64|fp32|fp16]
! RESULTS receives, as raw stream data: n, ne, l (int32), FI (3n), DFI (27, ne),
! u_int (3, l), sigma_int (6, l), sigma_int_prin (3, l), sigma_int_vm (l).
program dropin_main
  use data_mod
  use quadesth_gpu
  use stress_mod, only: u_int, sigma_int, sigma_int_prin, sigma_int_vm, deallocate_stress_data
  use io_mod, only: input_data, reorganize_solution
  implicit none
  character(len=256) :: arg, resfile
  integer :: u

  call get_command_argument(1, arg)
  filein = trim(arg)
  call get_command_argument(2, arg)
  fileout = trim(arg)
  call get_command_argument(3, resfile)
  call get_command_argument(4, arg)
  select case (trim(arg))
  case ('fp32')
    qd_solver_precision = QD_FP32_IR
  case ('fp16')
    qd_solver_precision = QD_FP16_IR
  case default
    qd_solver_precision = QD_FP64
  end select

  call input_data()
  close(ipr)
  call precompute_element_geometry()
  call ghmat8()
  call solve_linear_system(A, FI, 3*n)
  call reorganize_solution()
  if (l > 0) call compute_internal_points()

  open(newunit=u, file=trim(resfile), access='stream', form='unformatted', status='replace')
  write(u) n, ne, l
  write(u) FI
  write(u) DFI
  if (l > 0) write(u) u_int, sigma_int, sigma_int_prin, sigma_int_vm
  close(u)

  call quadesth_gpu_finalize()
  call deallocate_stress_data()
  call deallocate_bem_data()
end program dropin_main
