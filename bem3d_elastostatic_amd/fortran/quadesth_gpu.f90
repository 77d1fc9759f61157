! quadesth_gpu.f90 -- Fortran side of libquadesth.
!
! Two modules:
!
!   quadesth_capi  ISO_C_BINDING interfaces to the C ABI of include/quadesth.h
!                  (one interface per extern "C" function, same argument order).
!
!   quadesth_gpu   Drop-in replacements for the reference's four hot-path
!                  routines, with the same names and arguments, working on the
!                  data_mod globals:
!                    precompute_element_geometry  (assembly_mod.f90:28-51)
!                    ghmat8                       (assembly_mod.f90:53-223)
!                    solve_linear_system          (solver_mod.f90:17-36)
!                    compute_internal_points      (stress_mod.f90:511-781)
!                  A main program switches by replacing `use assembly_mod` and
!                  `use solver_mod` with `use quadesth_gpu`, and importing only
!                  the non-hot-path routines of stress_mod.
!
! Differences from the reference routines:
!   * ghmat8 keeps A in GPU memory and returns only FI on the host; the host A of
!     data_mod is left untouched unless qd_copy_matrix_to_host is .true.
!   * solve_linear_system solves on the device-resident A when its order matches
!     the last ghmat8; any other matrix goes through the LAPACK-ABI dgesv_ of
!     libquadesth (host A copied to the GPU).
!   * The solve precision is qd_solver_precision (QD_FP64 by default, as DGESV).
!   * Failures stop with the library's message, where the reference stops only
!     on a DGESV info /= 0 (solver_mod.f90:30-33, same text here).
module quadesth_capi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: QD_OK = 0
  integer(c_int), parameter :: QD_FP64 = 0, QD_FP32_IR = 1, QD_FP16_IR = 2
  integer(c_int), parameter :: QD_SOMIGLIANA_REFERENCE = 0, QD_SOMIGLIANA_CORRECTED = 1
  integer(c_int), parameter :: QD_ASSEMBLY_FP64 = 0, QD_ASSEMBLY_FAR_FP32 = 1, QD_ASSEMBLY_FAR_FP32_FACTOR = 2

  type, bind(C) :: qd_solve_report
    integer(c_int) :: precision = 0, info = 0, iterations = 0, converged = 0, gmres = 0, pivoting = 0
    real(c_double) :: residual = 0.0_c_double, t_factor_ms = 0.0_c_double, t_solve_ms = 0.0_c_double
  end type qd_solve_report

  interface
    integer(c_int) function qd_create(h, device) bind(C, name="qd_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: device
    end function qd_create

    subroutine qd_destroy(h) bind(C, name="qd_destroy")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine qd_destroy

    type(c_ptr) function qd_last_error(h) bind(C, name="qd_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function qd_last_error

    integer(c_int) function qd_synchronize(h) bind(C, name="qd_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function qd_synchronize

    integer(c_int) function qd_set_mesh(h, n, ne, x, y, z, nconec, kconec) bind(C, name="qd_set_mesh")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n, ne
      real(c_double), intent(in) :: x(*), y(*), z(*)
      integer(c_int), intent(in) :: nconec(*), kconec(*)
    end function qd_set_mesh

    integer(c_int) function qd_set_material(h, mu, nu) bind(C, name="qd_set_material")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: mu, nu
    end function qd_set_material

    integer(c_int) function qd_set_bc(h, kode, bc) bind(C, name="qd_set_bc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), intent(in) :: kode(*)
      real(c_double), intent(in) :: bc(*)
    end function qd_set_bc

    ! Each table pointer may be c_null_ptr.
    integer(c_int) function qd_precompute_element_geometry(h, eta1, eta2, eta3, xja, xg1, xg2, xg3, ff) &
        bind(C, name="qd_precompute_element_geometry")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, eta1, eta2, eta3, xja, xg1, xg2, xg3, ff
    end function qd_precompute_element_geometry

    integer(c_int) function qd_ghmat8(h) bind(C, name="qd_ghmat8")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function qd_ghmat8

    integer(c_int) function qd_set_assembly_precision(h, mode) bind(C, name="qd_set_assembly_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mode
    end function qd_set_assembly_precision

    integer(c_int) function qd_ghmat8_rhs(h, bc, fi) bind(C, name="qd_ghmat8_rhs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: bc(*)
      real(c_double), intent(out) :: fi(*)
    end function qd_ghmat8_rhs

    ! a and fi may be c_null_ptr.
    integer(c_int) function qd_get_system(h, a, lda, fi) bind(C, name="qd_get_system")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, a, fi
      integer(c_int), value :: lda
    end function qd_get_system

    integer(c_int) function qd_solve(h, precision, nrhs, b, ldb, tol, max_iter, rep) bind(C, name="qd_solve")
      import :: c_int, c_ptr, c_double, qd_solve_report
      type(c_ptr), value :: h
      integer(c_int), value :: precision, nrhs, ldb, max_iter
      real(c_double), intent(inout) :: b(*)
      real(c_double), value :: tol
      type(qd_solve_report), intent(out) :: rep
    end function qd_solve

    integer(c_int) function qd_reorganize_solution(h, x, fi, dfi) bind(C, name="qd_reorganize_solution")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: fi(*), dfi(*)
    end function qd_reorganize_solution

    ! u, sigma, prin, vm, counts may be c_null_ptr.
    integer(c_int) function qd_internal_points(h, l, cx, cy, cz, fi, dfi, variant, u, sigma, prin, vm, counts) &
        bind(C, name="qd_internal_points")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: l, variant
      real(c_double), intent(in) :: cx(*), cy(*), cz(*), fi(*), dfi(*)
      type(c_ptr), value :: u, sigma, prin, vm, counts
    end function qd_internal_points

    ! sigma_elem may be c_null_ptr.
    integer(c_int) function qd_boundary_stresses(h, fi, dfi, sigma_elem, sigma_node, prin, vm) &
        bind(C, name="qd_boundary_stresses")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h, sigma_elem
      real(c_double), intent(in) :: fi(*), dfi(*)
      real(c_double), intent(out) :: sigma_node(*), prin(*), vm(*)
    end function qd_boundary_stresses

    ! LAPACK ABI (Fortran calling convention, all arguments by reference).
    subroutine qd_dgesv_lapack(n, nrhs, a, lda, ipiv, b, ldb, info) bind(C, name="dgesv_")
      import :: c_int, c_double
      integer(c_int), intent(in) :: n, nrhs, lda, ldb
      real(c_double), intent(inout) :: a(lda, *), b(ldb, *)
      integer(c_int), intent(out) :: ipiv(*), info
    end subroutine qd_dgesv_lapack

    integer(c_size_t) function qd_c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function qd_c_strlen
  end interface

contains

  ! Message of the last failed call on h, as a Fortran string.
  function qd_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, nlen
    p = qd_last_error(h)
    if (.not. c_associated(p)) then
      msg = 'unknown error'
      return
    end if
    nlen = int(qd_c_strlen(p))
    call c_f_pointer(p, chars, [nlen])
    allocate(character(len=nlen) :: msg)
    do i = 1, nlen
      msg(i:i) = chars(i)
    end do
  end function qd_error_message

end module quadesth_capi


module quadesth_gpu
  use, intrinsic :: iso_c_binding
  use data_mod
  use quadesth_capi
  use stress_mod, only: u_int, sigma_int, sigma_int_prin, sigma_int_vm
  implicit none
  private

  public :: precompute_element_geometry, ghmat8, solve_linear_system, compute_internal_points
  public :: quadesth_gpu_finalize, quadesth_gpu_handle
  public :: qd_solver_precision, qd_somigliana_variant, qd_copy_matrix_to_host, qd_last_report
  public :: qd_assembly_precision
  public :: QD_FP64, QD_FP32_IR, QD_FP16_IR, QD_SOMIGLIANA_REFERENCE, QD_SOMIGLIANA_CORRECTED
  public :: QD_ASSEMBLY_FP64, QD_ASSEMBLY_FAR_FP32, QD_ASSEMBLY_FAR_FP32_FACTOR

  ! Precision of solve_linear_system (QD_FP64 reproduces DGESV).
  integer(c_int), save :: qd_solver_precision = QD_FP64
  ! Sign convention of the Somigliana displacement (reference by default).
  integer(c_int), save :: qd_somigliana_variant = QD_SOMIGLIANA_REFERENCE
  ! Assembly precision of ghmat8 (QD_ASSEMBLY_FP64 integrates as the reference).
  integer(c_int), save :: qd_assembly_precision = QD_ASSEMBLY_FP64
  ! Also fill data_mod's host A in ghmat8 (3n x 3n doubles over PCIe).
  logical, save :: qd_copy_matrix_to_host = .false.
  ! Report of the last solve.
  type(qd_solve_report), save :: qd_last_report

  type(c_ptr), save :: handle = c_null_ptr
  integer, save :: assembled_order = -1

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= QD_OK) then
      write(*, '(A, A, A, I4, A, A)') ' *** libquadesth ', what, ' failed (', rc, '): ', qd_error_message(handle)
      error stop 'libquadesth failure'
    end if
  end subroutine check

  ! The library handle of this module (created on first use, device 0).
  function quadesth_gpu_handle() result(h)
    type(c_ptr) :: h
    if (.not. c_associated(handle)) call check(qd_create(handle, 0_c_int), 'qd_create')
    h = handle
  end function quadesth_gpu_handle

  subroutine quadesth_gpu_finalize()
    if (c_associated(handle)) call qd_destroy(handle)
    handle = c_null_ptr
    assembled_order = -1
  end subroutine quadesth_gpu_finalize

  ! assembly_mod.f90:28-51.  Uploads the mesh, material and boundary codes of
  ! data_mod and computes the Gauss-point geometry on the GPU.
  subroutine precompute_element_geometry()
    type(c_ptr) :: h
    h = quadesth_gpu_handle()
    call check(qd_set_mesh(h, int(n, c_int), int(ne, c_int), x, y, z, nconec, kconec), 'qd_set_mesh')
    call check(qd_set_material(h, rge, xnu), 'qd_set_material')
    call check(qd_set_bc(h, kode, bc), 'qd_set_bc')
    call check(qd_precompute_element_geometry(h, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                              c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), &
               'qd_precompute_element_geometry')
  end subroutine precompute_element_geometry

  ! assembly_mod.f90:53-223.  A stays on the GPU; FI (3n) comes back.
  subroutine ghmat8()
    type(c_ptr) :: h
    h = quadesth_gpu_handle()
    call check(qd_set_assembly_precision(h, qd_assembly_precision), 'qd_set_assembly_precision')
    call check(qd_ghmat8(h), 'qd_ghmat8')
    if (qd_copy_matrix_to_host) then
      call system_into(h, 3 * n, FI, A)
    else
      call system_into(h, 3 * n, FI)
    end if
    assembled_order = 3 * n
  end subroutine ghmat8

  ! qd_get_system into fi(nsize) and, when present, a(nsize, nsize).
  subroutine system_into(h, nsize, fi, a)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: nsize
    real(c_double), intent(out), target :: fi(*)
    real(c_double), intent(out), target, optional :: a(*)
    type(c_ptr) :: pa
    pa = c_null_ptr
    if (present(a)) pa = c_loc(a(1))
    call check(qd_get_system(h, pa, int(nsize, c_int), c_loc(fi(1))), 'qd_get_system')
  end subroutine system_into

  ! solver_mod.f90:17-36.  B_vec is overwritten by the solution.
  subroutine solve_linear_system(A_mat, B_vec, n_size)
    integer, intent(in) :: n_size
    real(dp), intent(inout) :: A_mat(n_size, n_size)
    real(dp), intent(inout) :: B_vec(n_size)
    integer(c_int), allocatable :: ipiv(:)
    integer(c_int) :: info, rc
    if (c_associated(handle) .and. n_size == assembled_order) then
      rc = qd_solve(handle, qd_solver_precision, 1_c_int, B_vec, int(n_size, c_int), 0.0_c_double, &
                    0_c_int, qd_last_report)
      info = qd_last_report%info
      if (rc < 0) call check(rc, 'qd_solve')
    else
      allocate(ipiv(n_size))
      call qd_dgesv_lapack(int(n_size, c_int), 1_c_int, A_mat, int(n_size, c_int), ipiv, B_vec, &
                           int(n_size, c_int), info)
      deallocate(ipiv)
    end if
    if (info /= 0) then
      write(*, '(A, I6)') ' *** ERROR in LAPACK DGESV: info = ', info
      stop 'Solver failure: Singular matrix or bad arguments in LAPACK DGESV.'
    end if
  end subroutine solve_linear_system

  ! stress_mod.f90:511-781.  Fills stress_mod's u_int (3,l), sigma_int (6,l),
  ! sigma_int_prin (3,l) and sigma_int_vm (l) from FI and DFI.
  subroutine compute_internal_points()
    type(c_ptr) :: h
    h = quadesth_gpu_handle()
    if (allocated(u_int)) deallocate(u_int)
    if (allocated(sigma_int)) deallocate(sigma_int)
    if (allocated(sigma_int_prin)) deallocate(sigma_int_prin)
    if (allocated(sigma_int_vm)) deallocate(sigma_int_vm)
    allocate(u_int(3, l), sigma_int(6, l), sigma_int_prin(3, l), sigma_int_vm(l))
    if (l <= 0) return
    call internal_points_into(h, u_int, sigma_int, sigma_int_prin, sigma_int_vm)
  end subroutine compute_internal_points

  subroutine internal_points_into(h, u, sig, prin, vm)
    type(c_ptr), intent(in) :: h
    real(c_double), intent(out), target :: u(*), sig(*), prin(*), vm(*)
    call check(qd_internal_points(h, int(l, c_int), cx, cy, cz, FI, DFI, qd_somigliana_variant, &
                                  c_loc(u(1)), c_loc(sig(1)), c_loc(prin(1)), c_loc(vm(1)), c_null_ptr), &
               'qd_internal_points')
  end subroutine internal_points_into

end module quadesth_gpu
