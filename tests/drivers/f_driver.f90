! f_driver.f90 -- a Fortran caller of the micropp3_* C interface through
! iso_c_binding, the way a Fortran macro code binds it (the struct layouts and
! signatures of include/micropp_c.h and include/material_base.h).  Same
! problem and output as c_driver.c.
module micropp3_binding
  use iso_c_binding
  implicit none
  type, bind(C) :: material_base
    real(c_double) :: E, nu, Ka, Sy, k, mu, lambda, Xt
    integer(c_int) :: mtype
  end type
  type, bind(C) :: micropp3
    type(c_ptr) :: ptr
  end type
  interface
    subroutine material_set(self, mtype, E, nu, Ka, Sy, Xt) bind(C, name="material_set")
      import :: material_base, c_int, c_double
      type(material_base), intent(out) :: self
      integer(c_int), value :: mtype
      real(c_double), value :: E, nu, Ka, Sy, Xt
    end subroutine
    subroutine micropp3_new(self, ngp, nsize, micro_type, micro_params, materials, coupling, nsub, rank) &
        bind(C, name="micropp3_new")
      import :: micropp3, material_base, c_int, c_double
      type(micropp3), intent(out) :: self
      integer(c_int), value :: ngp, micro_type, nsub, rank
      integer(c_int), intent(in) :: nsize(3), coupling(*)
      real(c_double), intent(in) :: micro_params(4)
      type(material_base), intent(in) :: materials(3)
    end subroutine
    subroutine micropp3_free(self) bind(C, name="micropp3_free")
      import :: micropp3
      type(micropp3), intent(inout) :: self
    end subroutine
    subroutine micropp3_set_strain(self, gp, strain) bind(C, name="micropp3_set_strain")
      import :: micropp3, c_int, c_double
      type(micropp3), intent(inout) :: self
      integer(c_int), value :: gp
      real(c_double), intent(in) :: strain(6)
    end subroutine
    subroutine micropp3_get_stress(self, gp, stress) bind(C, name="micropp3_get_stress")
      import :: micropp3, c_int, c_double
      type(micropp3), intent(in) :: self
      integer(c_int), value :: gp
      real(c_double), intent(out) :: stress(6)
    end subroutine
    subroutine micropp3_get_ctan(self, gp, ctan) bind(C, name="micropp3_get_ctan")
      import :: micropp3, c_int, c_double
      type(micropp3), intent(in) :: self
      integer(c_int), value :: gp
      real(c_double), intent(out) :: ctan(36)
    end subroutine
    subroutine micropp3_homogenize(self) bind(C, name="micropp3_homogenize")
      import :: micropp3
      type(micropp3), intent(inout) :: self
    end subroutine
    subroutine micropp3_update_vars(self) bind(C, name="micropp3_update_vars")
      import :: micropp3
      type(micropp3), intent(inout) :: self
    end subroutine
    logical(c_bool) function micropp3_is_non_linear(self, gp) bind(C, name="micropp3_is_non_linear")
      import :: micropp3, c_int, c_bool
      type(micropp3), intent(in) :: self
      integer(c_int), value :: gp
    end function
    integer(c_int) function micropp3_get_cost(self, gp) bind(C, name="micropp3_get_cost")
      import :: micropp3, c_int
      type(micropp3), intent(in) :: self
      integer(c_int), value :: gp
    end function
    logical(c_bool) function micropp3_has_converged(self, gp) bind(C, name="micropp3_has_converged")
      import :: micropp3, c_int, c_bool
      type(micropp3), intent(in) :: self
      integer(c_int), value :: gp
    end function
    logical(c_bool) function micropp3_has_subiterated(self, gp) bind(C, name="micropp3_has_subiterated")
      import :: micropp3, c_int, c_bool
      type(micropp3), intent(in) :: self
      integer(c_int), value :: gp
    end function
  end interface
end module

program f_driver
  use micropp3_binding
  implicit none
  integer(c_int), parameter :: ngp = 2, steps = 3
  integer(c_int) :: nsize(3) = [6, 6, 6], coupling(2) = [2, 1]
  real(c_double) :: geo(4) = [0.25d0, 0d0, 0d0, 0d0]
  type(material_base) :: mats(3)
  type(micropp3) :: m
  real(c_double) :: eps(6), sig(6), ctan(36)
  integer(c_int) :: t, g
  call material_set(mats(1), 1_c_int, 1.0d6, 0.3d0, 1.0d5, 1.0d3, 0d0)
  call material_set(mats(2), 0_c_int, 1.0d7, 0.25d0, 0d0, 0d0, 0d0)
  call material_set(mats(3), 0_c_int, 1.0d7, 0.25d0, 0d0, 0d0, 0d0)
  call micropp3_new(m, ngp, nsize, 1_c_int, geo, mats, coupling, 4_c_int, 0_c_int)
  if (.not. c_associated(m%ptr)) stop 1
  do t = 0, steps - 1
    do g = 0, ngp - 1
      eps = 0d0
      eps(1) = 1.0d-3 * (t + 1) * (g + 1)
      eps(4) = -4.0d-4 * (t + 1)
      call micropp3_set_strain(m, g, eps)
    end do
    call micropp3_homogenize(m)
    do g = 0, ngp - 1
      call micropp3_get_stress(m, g, sig)
      call micropp3_get_ctan(m, g, ctan)
      write (*, '(A,6I6)') 'flags', t, g, merge(1, 0, logical(micropp3_is_non_linear(m, g))), &
        micropp3_get_cost(m, g), merge(1, 0, logical(micropp3_has_converged(m, g))), &
        merge(1, 0, logical(micropp3_has_subiterated(m, g)))
      write (*, '(A,6ES25.17)') 'sig', sig
      write (*, '(A,36ES25.17)') 'ctan', ctan
    end do
    call micropp3_update_vars(m)
  end do
  call micropp3_free(m)
end program
