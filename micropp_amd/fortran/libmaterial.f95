! libmaterial.f95 -- Fortran binding of material_base / material_set
! (include/material_base.h of micropp_amd), for macro codes that
! `use libmaterial`.
!
! Module name, derived type, component order and the material_set interface
! are those of the reference module (src/material.f95:24-48), so a Fortran
! caller of the reference compiles against this one unchanged.  The type
! mirrors the C struct field for field; the C side checks nothing, so the
! two must not drift apart.
module libmaterial
  use, intrinsic :: iso_c_binding, only: c_double, c_int
  implicit none
  private
  public :: material_base, material_set

  ! struct material_base { double E, nu, Ka, Sy, k, mu, lambda, Xt; int type; }
  type, bind(C) :: material_base
    real(c_double) :: E, nu, Ka, Sy
    real(c_double) :: k, mu, lambda
    real(c_double) :: Xt
    integer(c_int) :: type
  end type material_base

  interface
    ! void material_set(struct material_base *, int type, double E, double nu,
    !                   double Ka, double Sy, double Xt)
    subroutine material_set(this, mtype, E, nu, Ka, Sy, Xt) bind(C, name="material_set")
      import :: material_base, c_int, c_double
      type(material_base) :: this
      integer(c_int), value, intent(in) :: mtype
      real(c_double), value, intent(in) :: E, nu, Ka, Sy, Xt
    end subroutine material_set
  end interface
end module libmaterial
