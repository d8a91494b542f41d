! C-callable entry points over the reference's own spline module, compiled
! with the reference's real(8) promotion.  Test infrastructure only: it lets
! tests/test_oracle_pins.py evaluate the reference's spline_value itself.
module ref_shim
  use iso_c_binding, only: c_int, c_double
  use spline, only: spline_init, spline_value, spline_cleanup
  implicit none
contains
  subroutine ref_spline_init(order) bind(C, name="ref_spline_init")
    integer(c_int), value, intent(in) :: order
    call spline_init(int(order))
  end subroutine ref_spline_init

  real(c_double) function ref_spline_value(x) bind(C, name="ref_spline_value")
    real(c_double), value, intent(in) :: x
    ref_spline_value = spline_value(x)
  end function ref_spline_value

  subroutine ref_spline_cleanup() bind(C, name="ref_spline_cleanup")
    call spline_cleanup()
  end subroutine ref_spline_cleanup
end module ref_shim
