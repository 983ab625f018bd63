﻿!mod$ v1 sum:63a488294f10313a
!need$ 0bde2ac47243ead2 i iso_c_binding
module micropp3_binding
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
type,bind(c)::material_base
real(8)::e
real(8)::nu
real(8)::ka
real(8)::sy
real(8)::k
real(8)::mu
real(8)::lambda
real(8)::xt
integer(4)::mtype
end type
type,bind(c)::micropp3
type(c_ptr)::ptr
end type
interface
subroutine material_set(self,mtype,e,nu,ka,sy,xt) bind(c,name="material_set")
import::material_base
type(material_base),intent(out)::self
integer(4),value::mtype
real(8),value::e
real(8),value::nu
real(8),value::ka
real(8),value::sy
real(8),value::xt
end
end interface
interface
subroutine micropp3_new(self,ngp,nsize,micro_type,micro_params,materials,coupling,nsub,rank) bind(c,name="micropp3_new")
import::material_base
import::micropp3
type(micropp3),intent(out)::self
integer(4),value::ngp
integer(4),intent(in)::nsize(1_8:3_8)
integer(4),value::micro_type
real(8),intent(in)::micro_params(1_8:4_8)
type(material_base),intent(in)::materials(1_8:3_8)
integer(4),intent(in)::coupling(1_8:*)
integer(4),value::nsub
integer(4),value::rank
end
end interface
interface
subroutine micropp3_free(self) bind(c,name="micropp3_free")
import::micropp3
type(micropp3),intent(inout)::self
end
end interface
interface
subroutine micropp3_set_strain(self,gp,strain) bind(c,name="micropp3_set_strain")
import::micropp3
type(micropp3),intent(inout)::self
integer(4),value::gp
real(8),intent(in)::strain(1_8:6_8)
end
end interface
interface
subroutine micropp3_get_stress(self,gp,stress) bind(c,name="micropp3_get_stress")
import::micropp3
type(micropp3),intent(in)::self
integer(4),value::gp
real(8),intent(out)::stress(1_8:6_8)
end
end interface
interface
subroutine micropp3_get_ctan(self,gp,ctan) bind(c,name="micropp3_get_ctan")
import::micropp3
type(micropp3),intent(in)::self
integer(4),value::gp
real(8),intent(out)::ctan(1_8:36_8)
end
end interface
interface
subroutine micropp3_homogenize(self) bind(c,name="micropp3_homogenize")
import::micropp3
type(micropp3),intent(inout)::self
end
end interface
interface
subroutine micropp3_update_vars(self) bind(c,name="micropp3_update_vars")
import::micropp3
type(micropp3),intent(inout)::self
end
end interface
interface
function micropp3_is_non_linear(self,gp) bind(c,name="micropp3_is_non_linear")
import::micropp3
type(micropp3),intent(in)::self
integer(4),value::gp
logical(1)::micropp3_is_non_linear
end
end interface
interface
function micropp3_get_cost(self,gp) bind(c,name="micropp3_get_cost")
import::micropp3
type(micropp3),intent(in)::self
integer(4),value::gp
integer(4)::micropp3_get_cost
end
end interface
interface
function micropp3_has_converged(self,gp) bind(c,name="micropp3_has_converged")
import::micropp3
type(micropp3),intent(in)::self
integer(4),value::gp
logical(1)::micropp3_has_converged
end
end interface
interface
function micropp3_has_subiterated(self,gp) bind(c,name="micropp3_has_subiterated")
import::micropp3
type(micropp3),intent(in)::self
integer(4),value::gp
logical(1)::micropp3_has_subiterated
end
end interface
end
