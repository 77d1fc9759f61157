! ref_driver.f90 -- C-callable driver around the reference's own compiled
! modules (data_mod, assembly_mod, stress_mod, io_mod), used ONLY to generate
! and check golden vectors.  It reads a deck through the reference's reader
! (io_mod.f90:36-146) and exposes each hot-path phase so that the test
! harness can supply the linear solve (the reference calls LAPACK dgesv,
! solver_mod.f90:28, which is not vendored).
module qref_driver
  use iso_c_binding
  use data_mod
  use assembly_mod, only: precompute_element_geometry, ghmat8
  use stress_mod, only: compute_internal_points, compute_boundary_stresses, &
                        u_int, sigma_int, sigma_int_prin, sigma_int_vm, &
                        sigma_node, sigma_principal, sigma_vm, deallocate_stress_data
  use io_mod, only: input_data, reorganize_solution
  implicit none
contains

  ! Load a deck with the reference reader.  path and outpath are
  ! NUL-terminated and at most 63 characters.
  function qref_load(path, outpath) result(rc) bind(C, name="qref_load")
    character(kind=c_char), intent(in) :: path(*), outpath(*)
    integer(c_int) :: rc
    integer :: i
    filein = ' '
    fileout = ' '
    i = 1
    do while (path(i) /= c_null_char .and. i <= 63)
      filein(i:i) = path(i)
      i = i + 1
    end do
    i = 1
    do while (outpath(i) /= c_null_char .and. i <= 63)
      fileout(i:i) = outpath(i)
      i = i + 1
    end do
    young = 0.0_dp
    rge = 0.0_dp
    xnu = 0.0_dp
    call input_data()
    close(ipr)
    rc = 0
  end function qref_load

  subroutine qref_sizes(nn, nne, nl, mu, nu_) bind(C, name="qref_sizes")
    integer(c_int), intent(out) :: nn, nne, nl
    real(c_double), intent(out) :: mu, nu_
    nn = n
    nne = ne
    nl = l
    mu = rge
    nu_ = xnu
  end subroutine qref_sizes

  ! assembly_mod.f90:28-51 then 53-223; copies A (3n x 3n, column-major) and FI.
  subroutine qref_assemble(a_out, fi_out) bind(C, name="qref_assemble")
    real(c_double), intent(out) :: a_out(3*n, 3*n), fi_out(3*n)
    call precompute_element_geometry()
    call ghmat8()
    a_out = A
    fi_out = FI
  end subroutine qref_assemble

  ! Replace FI by the solution of A X = FI, then io_mod.f90:148-183.
  subroutine qref_reorganize(x_in, fi_out, dfi_out) bind(C, name="qref_reorganize")
    real(c_double), intent(in) :: x_in(3*n)
    real(c_double), intent(out) :: fi_out(3*n), dfi_out(27, ne)
    FI = x_in
    call reorganize_solution()
    fi_out = FI
    dfi_out = DFI
  end subroutine qref_reorganize

  ! stress_mod.f90:511-781
  subroutine qref_internal(u_out, s_out, p_out, vm_out) bind(C, name="qref_internal")
    real(c_double), intent(out) :: u_out(3, l), s_out(6, l), p_out(3, l), vm_out(l)
    call compute_internal_points()
    u_out = u_int
    s_out = sigma_int
    p_out = sigma_int_prin
    vm_out = sigma_int_vm
  end subroutine qref_internal

  ! stress_mod.f90:366-406
  subroutine qref_boundary(s_out, p_out, vm_out) bind(C, name="qref_boundary")
    real(c_double), intent(out) :: s_out(6, n), p_out(3, n), vm_out(n)
    call compute_boundary_stresses()
    s_out = sigma_node
    p_out = sigma_principal
    vm_out = sigma_vm
  end subroutine qref_boundary

  ! The reference's writers into directory dir (NUL-terminated, at most 40
  ! characters), in main.f90:75-80's order: output_data (io_mod.f90:191-228)
  ! continues the file qref_load's input_data echoed the deck into (fileout,
  ! reopened for appending), then output_boundary_stresses (stress_mod.f90:
  ! 412-468, also writing stresses.dat), output_internal_points
  ! (stress_mod.f90:892-936, internal_points.dat) and save_disp_trac
  ! (io_mod.f90:230-255, disp_trac.dat).  Needs qref_reorganize,
  ! qref_internal and qref_boundary first.
  subroutine qref_outputs(dir) bind(C, name="qref_outputs")
    use stress_mod, only: output_boundary_stresses, output_internal_points, filestress, fileintpts
    use io_mod, only: output_data, save_disp_trac
    character(kind=c_char), intent(in) :: dir(*)
    character(len=64) :: d
    integer :: i
    d = ' '
    i = 1
    do while (dir(i) /= c_null_char .and. i <= 40)
      d(i:i) = dir(i)
      i = i + 1
    end do
    filestress = trim(d) // '/stresses.dat'
    fileintpts = trim(d) // '/internal_points.dat'
    filetrack = trim(d) // '/disp_trac.dat'
    open(unit=ipr, file=trim(fileout), status='old', position='append', action='write')
    open(unit=iout_track, file=trim(filetrack), status='replace', action='write')
    call output_data()
    call output_boundary_stresses()
    if (l > 0) call output_internal_points()
    call save_disp_trac()
    close(ipr)
    close(iout_track)
  end subroutine qref_outputs

  subroutine qref_free() bind(C, name="qref_free")
    call deallocate_stress_data()
    call deallocate_bem_data()
  end subroutine qref_free

end module qref_driver
