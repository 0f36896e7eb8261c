! TEST INFRASTRUCTURE.  C-callable entries to the reference's own advect_scalar
! (dyn_em/module_advect_em.F:3029-4359) and advect_scalar_pd (6069-7885), compiled in the
! reference's standalone
! ADVECT_KERNEL mode (module_advect_em.F:4-113, which supplies its own
! grid_config_rec_type).  The oracle Makefile compiles the reference file in place;
! this wrapper only copies plain C arguments into that derived type and forwards.
! Used by the tests to pin the oracle's C restatements of the two routines.
module ref_advect_bindc
  use iso_c_binding
  use advection_kernel
  implicit none
contains
  ! flags(1:13) = periodic_x, periodic_y, symmetric_xs, symmetric_xe, symmetric_ys,
  !               symmetric_ye, open_xs, open_xe, open_ys, open_ye, specified, nested, polar
  ! flags(14:15) = h_sca_adv_order, v_sca_adv_order
  ! d(1:18) = ids,ide,jds,jde,kds,kde, ims,ime,jms,jme,kms,kme, its,ite,jts,jte,kts,kte
  subroutine ref_advect_scalar(field, field_old, tendency, ru, rv, rom, c1, c2, mut, time_step, &
                               flags, msfux, msfuy, msfvx, msfvy, msftx, msfty, fzm, fzp,      &
                               rdx, rdy, rdzw, d) bind(C, name="ref_advect_scalar")
    integer(c_int), intent(in) :: d(18)
    integer(c_int), intent(in) :: flags(15)
    integer(c_int), value :: time_step
    real(c_float), value :: rdx, rdy
    real(c_float), intent(in) :: field(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: field_old(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(inout) :: tendency(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: ru(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: rv(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: rom(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: mut(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: msfux(d(7):d(8), d(9):d(10)), msfuy(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: msfvx(d(7):d(8), d(9):d(10)), msfvy(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: msftx(d(7):d(8), d(9):d(10)), msfty(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: c1(d(11):d(12)), c2(d(11):d(12))
    real(c_float), intent(in) :: fzm(d(11):d(12)), fzp(d(11):d(12)), rdzw(d(11):d(12))
    type(grid_config_rec_type) :: cf
    integer :: itime
    cf%scalar_adv_opt = 0
    cf%periodic_x = flags(1) /= 0
    cf%periodic_y = flags(2) /= 0
    cf%symmetric_xs = flags(3) /= 0
    cf%symmetric_xe = flags(4) /= 0
    cf%symmetric_ys = flags(5) /= 0
    cf%symmetric_ye = flags(6) /= 0
    cf%open_xs = flags(7) /= 0
    cf%open_xe = flags(8) /= 0
    cf%open_ys = flags(9) /= 0
    cf%open_ye = flags(10) /= 0
    cf%specified = flags(11) /= 0
    cf%nested = flags(12) /= 0
    cf%polar = flags(13) /= 0
    cf%h_sca_adv_order = flags(14)
    cf%v_sca_adv_order = flags(15)
    itime = time_step
    call advect_scalar(field, field_old, tendency, ru, rv, rom, c1, c2, mut, itime, cf, &
                       msfux, msfuy, msfvx, msfvy, msftx, msfty, fzm, fzp, rdx, rdy, rdzw, &
                       d(1), d(2), d(3), d(4), d(5), d(6), d(7), d(8), d(9), d(10), d(11), d(12), &
                       d(13), d(14), d(15), d(16), d(17), d(18))
  end subroutine ref_advect_scalar
  ! the positive-definite form: field_old, mub and mu_old enter the first-order fluxes and
  ! the limiter; h_tendency and z_tendency are written when tenddec /= 0
  subroutine ref_advect_scalar_pd(field, field_old, tendency, h_tendency, z_tendency, ru, rv, rom, c1, c2, &
                                  mut, mub, mu_old, time_step, flags, tenddec, msfux, msfuy, msfvx, msfvy, &
                                  msftx, msfty, fzm, fzp, rdx, rdy, rdzw, dt, d) bind(C, name="ref_advect_scalar_pd")
    integer(c_int), intent(in) :: d(18)
    integer(c_int), intent(in) :: flags(15)
    integer(c_int), value :: time_step, tenddec
    real(c_float), value :: rdx, rdy, dt
    real(c_float), intent(in) :: field(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: field_old(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(inout) :: tendency(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(inout) :: h_tendency(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(inout) :: z_tendency(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: ru(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: rv(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: rom(d(7):d(8), d(11):d(12), d(9):d(10))
    real(c_float), intent(in) :: mut(d(7):d(8), d(9):d(10)), mub(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: mu_old(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: msfux(d(7):d(8), d(9):d(10)), msfuy(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: msfvx(d(7):d(8), d(9):d(10)), msfvy(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: msftx(d(7):d(8), d(9):d(10)), msfty(d(7):d(8), d(9):d(10))
    real(c_float), intent(in) :: c1(d(11):d(12)), c2(d(11):d(12))
    real(c_float), intent(in) :: fzm(d(11):d(12)), fzp(d(11):d(12)), rdzw(d(11):d(12))
    type(grid_config_rec_type) :: cf
    integer :: itime
    logical :: tdec
    cf%scalar_adv_opt = 1
    cf%periodic_x = flags(1) /= 0
    cf%periodic_y = flags(2) /= 0
    cf%symmetric_xs = flags(3) /= 0
    cf%symmetric_xe = flags(4) /= 0
    cf%symmetric_ys = flags(5) /= 0
    cf%symmetric_ye = flags(6) /= 0
    cf%open_xs = flags(7) /= 0
    cf%open_xe = flags(8) /= 0
    cf%open_ys = flags(9) /= 0
    cf%open_ye = flags(10) /= 0
    cf%specified = flags(11) /= 0
    cf%nested = flags(12) /= 0
    cf%polar = flags(13) /= 0
    cf%h_sca_adv_order = flags(14)
    cf%v_sca_adv_order = flags(15)
    itime = time_step
    tdec = tenddec /= 0
    call advect_scalar_pd(field, field_old, tendency, h_tendency, z_tendency, ru, rv, rom, c1, c2, &
                          mut, mub, mu_old, itime, cf, tdec, msfux, msfuy, msfvx, msfvy, msftx, msfty, &
                          fzm, fzp, rdx, rdy, rdzw, dt, &
                          d(1), d(2), d(3), d(4), d(5), d(6), d(7), d(8), d(9), d(10), d(11), d(12), &
                          d(13), d(14), d(15), d(16), d(17), d(18))
  end subroutine ref_advect_scalar_pd
end module ref_advect_bindc
