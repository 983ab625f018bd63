! libmicropp.f95 -- Fortran binding of the micropp3_* C interface
! (include/micropp_c.h of micropp_amd), for macro codes that
! `use libmicropp`.
!
! Module name, the micropp3 handle type and every interface (names,
! argument order, value / reference passing, logical(c_bool) results) are
! those of the reference module (src/micropp.f95:24-182), so a Fortran
! caller of the reference compiles and links against this library
! unchanged.  Beyond the reference: micropp3_get_non_linear_gps (declared in
! the reference C header), and the micropp3_params extension with
! micropp3_params_default / micropp3_new_ext (tolerances, lin_stress, A0,
! log, mixed precision).
module libmicropp
  use, intrinsic :: iso_c_binding, only: c_bool, c_char, c_double, c_int, c_ptr
  use libmaterial
  implicit none
  private
  public :: material_base, material_set
  public :: micropp3, micropp3_params
  public :: micropp3_new, micropp3_free, micropp3_set_strain, micropp3_get_stress, micropp3_get_ctan
  public :: micropp3_homogenize, micropp3_homogenize_linear, micropp3_update_vars
  public :: micropp3_is_non_linear, micropp3_get_cost, micropp3_has_converged, micropp3_has_subiterated
  public :: micropp3_output, micropp3_output2, micropp3_print_info, micropp3_get_non_linear_gps
  public :: micropp3_write_restart, micropp3_read_restart
  public :: micropp3_params_default, micropp3_new_ext

  ! struct micropp3 { void *ptr; }
  type, bind(C) :: micropp3
    type(c_ptr) :: ptr
  end type micropp3

  ! struct micropp3_params (extension; field order of include/micropp_c.h)
  type, bind(C) :: micropp3_params
    integer(c_int) :: ngp
    integer(c_int) :: size(3)
    integer(c_int) :: type
    real(c_double) :: geo_params(4)
    type(material_base) :: materials(3)
    type(c_ptr) :: coupling
    integer(c_int) :: subiterations, nsubiterations, mpi_rank, nr_max_its
    real(c_double) :: nr_max_tol, nr_rel_tol
    integer(c_int) :: cg_max_its
    real(c_double) :: cg_abs_tol, cg_rel_tol
    integer(c_int) :: calc_ctan_lin, use_A0, its_with_A0, lin_stress, write_log, mixed_precision
  end type micropp3_params

  interface
    subroutine micropp3_new(self, ngp, size, micro_type, micro_params, materials, coupling, &
                            nsubiterations, mpi_rank) bind(C, name="micropp3_new")
      import :: micropp3, material_base, c_int, c_double
      type(micropp3), intent(out) :: self
      integer(c_int), value :: ngp
      integer(c_int), intent(in) :: size(3)
      integer(c_int), value :: micro_type
      real(c_double), intent(in) :: micro_params(*)
      type(material_base), intent(in) :: materials(*)
      integer(c_int), intent(in) :: coupling(*)
      integer(c_int), value :: nsubiterations
      integer(c_int), value :: mpi_rank
    end subroutine micropp3_new

    subroutine micropp3_free(this) bind(C, name="micropp3_free")
      import :: micropp3
      type(micropp3), intent(inout) :: this
    end subroutine micropp3_free

    subroutine micropp3_set_strain(this, gp_id, strain) bind(C, name="micropp3_set_strain")
      import :: micropp3, c_int, c_double
      type(micropp3), intent(inout) :: this
      integer(c_int), value, intent(in) :: gp_id
      real(c_double), intent(in) :: strain(*)
    end subroutine micropp3_set_strain

    subroutine micropp3_get_stress(this, gp_id, stress) bind(C, name="micropp3_get_stress")
      import :: micropp3, c_int, c_double
      type(micropp3), intent(inout) :: this
      integer(c_int), value, intent(in) :: gp_id
      real(c_double), intent(out) :: stress(*)
    end subroutine micropp3_get_stress

    subroutine micropp3_get_ctan(this, gp_id, ctan) bind(C, name="micropp3_get_ctan")
      import :: micropp3, c_int, c_double
      type(micropp3), intent(inout) :: this
      integer(c_int), value, intent(in) :: gp_id
      real(c_double), intent(out) :: ctan(*)
    end subroutine micropp3_get_ctan

    subroutine micropp3_homogenize(this) bind(C, name="micropp3_homogenize")
      import :: micropp3
      type(micropp3), intent(inout) :: this
    end subroutine micropp3_homogenize

    subroutine micropp3_homogenize_linear(this) bind(C, name="micropp3_homogenize_linear")
      import :: micropp3
      type(micropp3), intent(inout) :: this
    end subroutine micropp3_homogenize_linear

    logical(c_bool) function micropp3_is_non_linear(this, gp_id) bind(C, name="micropp3_is_non_linear")
      import :: micropp3, c_bool, c_int
      type(micropp3), intent(in) :: this
      integer(c_int), value, intent(in) :: gp_id
    end function micropp3_is_non_linear

    integer(c_int) function micropp3_get_cost(this, gp_id) bind(C, name="micropp3_get_cost")
      import :: micropp3, c_int
      type(micropp3), intent(in) :: this
      integer(c_int), value, intent(in) :: gp_id
    end function micropp3_get_cost

    logical(c_bool) function micropp3_has_converged(this, gp_id) bind(C, name="micropp3_has_converged")
      import :: micropp3, c_bool, c_int
      type(micropp3), intent(in) :: this
      integer(c_int), value, intent(in) :: gp_id
    end function micropp3_has_converged

    logical(c_bool) function micropp3_has_subiterated(this, gp_id) bind(C, name="micropp3_has_subiterated")
      import :: micropp3, c_bool, c_int
      type(micropp3), intent(in) :: this
      integer(c_int), value, intent(in) :: gp_id
    end function micropp3_has_subiterated

    subroutine micropp3_update_vars(this) bind(C, name="micropp3_update_vars")
      import :: micropp3
      type(micropp3), intent(inout) :: this
    end subroutine micropp3_update_vars

    ! filename: a NUL-terminated character array (c_char), as in the reference
    subroutine micropp3_output(this, gp_id, filename) bind(C, name="micropp3_output")
      import :: micropp3, c_int, c_char
      type(micropp3), intent(inout) :: this
      integer(c_int), value, intent(in) :: gp_id
      character(kind=c_char), intent(in) :: filename(*)
    end subroutine micropp3_output

    subroutine micropp3_output2(this, gp_id, elem_global, time_step) bind(C, name="micropp3_output2")
      import :: micropp3, c_int
      type(micropp3), intent(inout) :: this
      integer(c_int), value, intent(in) :: gp_id, elem_global, time_step
    end subroutine micropp3_output2

    subroutine micropp3_print_info(this) bind(C, name="micropp3_print_info")
      import :: micropp3
      type(micropp3), intent(in) :: this
    end subroutine micropp3_print_info

    integer(c_int) function micropp3_get_non_linear_gps(this) bind(C, name="micropp3_get_non_linear_gps")
      import :: micropp3, c_int
      type(micropp3), intent(in) :: this
    end function micropp3_get_non_linear_gps

    subroutine micropp3_write_restart(this, restart_id) bind(C, name="micropp3_write_restart")
      import :: micropp3, c_int
      type(micropp3), intent(in) :: this
      integer(c_int), value, intent(in) :: restart_id
    end subroutine micropp3_write_restart

    subroutine micropp3_read_restart(this, restart_id) bind(C, name="micropp3_read_restart")
      import :: micropp3, c_int
      type(micropp3), intent(in) :: this
      integer(c_int), value, intent(in) :: restart_id
    end subroutine micropp3_read_restart

    subroutine micropp3_params_default(p) bind(C, name="micropp3_params_default")
      import :: micropp3_params
      type(micropp3_params), intent(out) :: p
    end subroutine micropp3_params_default

    subroutine micropp3_new_ext(self, p) bind(C, name="micropp3_new_ext")
      import :: micropp3, micropp3_params
      type(micropp3), intent(out) :: self
      type(micropp3_params), intent(in) :: p
    end subroutine micropp3_new_ext
  end interface
end module libmicropp
