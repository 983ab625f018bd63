! f_driver.f90 -- a Fortran caller of micropp_amd through the drop-in
! modules `libmicropp` / `libmaterial` (micropp_amd/fortran/), the way a
! Fortran macro code uses the reference's modules.  Same problem and output
! as c_driver.c: one instance with two Gauss points (FE_FULL and FE_ONE_WAY)
! of a 6x6x6-node sphere RVE with a J2-plastic matrix, three load steps.
program f_driver
  use, intrinsic :: iso_c_binding, only: c_int, c_double, c_associated
  use libmicropp
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
end program f_driver
