! C entry points into reference routines compiled from the reference's own
! sources.  TEST INFRASTRUCTURE ONLY: the tests use these to check the oracle's
! restatements against the reference code itself; nothing in the product
! loads this library.
!
! Routines reached (all FP64, reference module state set by ref_set_*):
!   utilities/box_test.f95, lbgrid/utilities/tri_volume.f95,
!   linking/modules/linkwise.f95:45-57 (crossing_point_and_dist, with
!   link_dist set as general/background_calcs/link_info.f95:23-28 does),
!   utilities/modules/quick_calcs.f95:24-215 (Sutherland viscosity and heat
!   conductivity), looping/timestepping/lagrangian_multipliers/lu_decomp.f95
!   (one 13x13 LU-solve Newton update of the Lagrange multipliers).
module uclbs_ref_entry
  use iso_c_binding
  use precise
  use constants
  use linkwise
  use quick_calcs
  implicit none
  external :: tri_volume, box_test, lu_decomp
contains

  function ref_tri_volume(xc, e, t) bind(c, name='ref_tri_volume') result(r)
    real(c_double), intent(in) :: xc(3), e(3), t(9)
    integer(c_int) :: r
    real(dp) :: a(3), b(3)
    integer :: cross
    a = xc
    b = e
    cross = 0
    call tri_volume(a, b, t(1), t(2), t(3), t(4), t(5), t(6), t(7), t(8), t(9), cross)
    r = cross
  end function

  ! bb = max_x, min_x, max_y, min_y, max_z, min_z
  function ref_box_test(p1, p2, bb) bind(c, name='ref_box_test') result(r)
    real(c_double), intent(in) :: p1(3), p2(3), bb(6)
    integer(c_int) :: r
    integer :: outside
    outside = 0
    call box_test(p1, p2, bb(1), bb(2), bb(3), bb(4), bb(5), bb(6), outside)
    r = outside
  end function

  ! wall-link code of a crossing at xc on the link from node e1 along a
  ! lattice vector of squared length cmag2, grid spacing dx
  function ref_link_code(xc, e1, cmag2, dx) bind(c, name='ref_link_code') result(r)
    real(c_double), intent(in) :: xc(3), e1(3)
    integer(c_int), value :: cmag2
    real(c_double), value :: dx
    integer(c_int) :: r
    real(dp) :: a(3), b(3), d
    if (allocated(link_dist)) deallocate(link_dist)
    allocate(link_dist(1, 0:0))
    link_dist(1, 0) = sqrt(real(cmag2, dp))*dx/2.0D0**0
    a = xc
    b = e1
    d = dx
    r = crossing_point_and_dist(a, b, 1, d, 0)
  end function

  subroutine ref_set_constants(g, tr, mr, rr, kr, rgas) bind(c, name='ref_set_constants')
    real(c_double), value :: g, tr, mr, rr, kr, rgas
    gama = g
    t_ref = tr
    mu_ref = mr
    rho_ref = rr
    kappa_ref = kr
    gas_const_R = rgas
  end subroutine

  function ref_sutherland(temp) bind(c, name='ref_sutherland') result(mu)
    real(c_double), value :: temp
    real(c_double) :: mu
    call quick_sutherlands(mu, temp)
  end function

  function ref_therm_cond(rho, temp) bind(c, name='ref_therm_cond') result(kappa)
    real(c_double), value :: rho, temp
    real(c_double) :: kappa
    call quick_therm_cond(kappa, rho, temp)
  end function

  ! jac: 13x13 column-major; tgt: rho, u, v, w, Pxx, Pyy, Pzz, Pxy, Pxz,
  ! Pyz, qx, qy, qz (velocities, not momenta); lm: chi, zeta(3), pi_xx, pi_yy,
  ! pi_zz, pi_xy, pi_xz, pi_yz, lambda(3), updated in place.
  function ref_lu_decomp(jac, tgt, lm) bind(c, name='ref_lu_decomp') result(max_diff)
    real(c_double), intent(in) :: jac(13, 13), tgt(13)
    real(c_double), intent(inout) :: lm(13)
    real(c_double) :: max_diff
    real(dp) :: j(13, 13)
    j = jac
    max_diff = 0.0D0
    call lu_decomp(j, 13, tgt(1), tgt(2), tgt(3), tgt(4), tgt(5), tgt(6), tgt(7), tgt(8), &
                   tgt(9), tgt(10), tgt(11), tgt(12), tgt(13), lm(1), lm(2), lm(3), lm(4), &
                   lm(5), lm(6), lm(7), lm(8), lm(9), lm(10), lm(11), lm(12), lm(13), max_diff)
  end function

  ! nrep calls of lu_decomp on the same inputs (timing): returns the sum of
  ! each call's max_diff and first updated multiplier
  function ref_lu_decomp_batch(nrep, jac, tgt, lm) bind(c, name='ref_lu_decomp_batch') result(s)
    integer(c_int), value :: nrep
    real(c_double), intent(in) :: jac(13, 13), tgt(13), lm(13)
    real(c_double) :: s
    real(dp) :: j(13, 13), l(13), max_diff
    integer :: k
    s = 0.0D0
    do k = 1, nrep
      j = jac
      l = lm
      max_diff = 0.0D0
      call lu_decomp(j, 13, tgt(1), tgt(2), tgt(3), tgt(4), tgt(5), tgt(6), tgt(7), tgt(8), &
                     tgt(9), tgt(10), tgt(11), tgt(12), tgt(13), l(1), l(2), l(3), l(4), &
                     l(5), l(6), l(7), l(8), l(9), l(10), l(11), l(12), l(13), max_diff)
      s = s + max_diff + l(1)
    end do
  end function

  ! quick_sutherlands and quick_therm_cond at n nodes (timing): returns the
  ! sum of mu + kappa
  function ref_quick_batch(n, rho, temp) bind(c, name='ref_quick_batch') result(s)
    integer(c_int), value :: n
    real(c_double), intent(in) :: rho(n), temp(n)
    real(c_double) :: s, mu, kappa
    integer :: k
    s = 0.0D0
    do k = 1, n
      call quick_sutherlands(mu, temp(k))
      call quick_therm_cond(kappa, rho(k), temp(k))
      s = s + mu + kappa
    end do
  end function

end module
