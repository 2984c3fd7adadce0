! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! bind(C) entry points over the reference's own Fortran modules, compiled from
! the reference sources by `make ref` (oracle/Makefile) into oracle/_ref/.
! The tests use them to pin the C oracle bit for bit:
!   ref_ran2_seq   -> ran2.f:1-33
!   ref_mc         -> gridset.f90:6-46 + sourceph.f90:7-49 + inttau2.f90:7-72,
!                     driven like the photon loop of mcpolar.f90:151-170
!   ref_stokes     -> stokes.f90:6-153
!   ref_fits       -> thermalConst_mod.f90:11-87
!   ref_mk_cdf, ref_sample, ref_search, ref_lin_inter
!                  -> ch_opt.f90:28-152 (spectral helpers)
! The reference grid is fixed at compile time (constants.f90:12, 80^3).
module ref_harness
    use iso_c_binding
    implicit none
contains

    subroutine ref_ran2_seq(iseed, n, out) bind(C, name="ref_ran2_seq")
        integer(c_int), intent(inout) :: iseed
        integer(c_int), value         :: n
        real(c_double), intent(out)   :: out(n)
        real    :: ran2
        integer :: i, s
        s = iseed
        do i = 1, n
            out(i) = ran2(s)
        end do
        iseed = s
    end subroutine ref_ran2_seq

    subroutine ref_grid_n(n) bind(C, name="ref_grid_n")
        use constants, only : nxg
        integer(c_int), intent(out) :: n
        n = nxg
    end subroutine ref_grid_n

    ! Runs nphotons packets; rhokap_in (halo array, (nxg+2)^3) replaces the
    ! homogeneous grid when use_rhokap /= 0.  jmean_out is nxg^3, col-major.
    subroutine ref_mc(nphotons, xmax_c, ymax_c, zmax_c, iseed_c, use_rhokap, rhokap_in, &
                      jmean_out, nesc) bind(C, name="ref_mc")
        use constants,   only : nxg, nyg, nzg
        use iarray,      only : xface, yface, zface, rhokap, jmean
        use gridset_mod, only : gridset
        use sourceph_mod, only : sourcephCO2
        use inttau2,     only : tauint1
        integer(c_long), value        :: nphotons
        real(c_double),  value        :: xmax_c, ymax_c, zmax_c
        integer(c_int),  intent(inout):: iseed_c
        integer(c_int),  value        :: use_rhokap
        real(c_double),  intent(in)   :: rhokap_in(0:nxg+1, 0:nyg+1, 0:nzg+1)
        real(c_double),  intent(out)  :: jmean_out(nxg, nyg, nzg)
        integer(c_long), intent(out)  :: nesc
        real    :: xmax, ymax, zmax, delta
        integer :: iseed, xcell, ycell, zcell
        integer(c_long) :: j
        logical :: tflag

        xmax = xmax_c; ymax = ymax_c; zmax = zmax_c
        if (allocated(xface)) deallocate(xface, yface, zface, rhokap, jmean)
        allocate(xface(nxg+1), yface(nyg+1), zface(nzg+1))
        allocate(rhokap(0:nxg+1, 0:nyg+1, 0:nzg+1), jmean(nxg, nyg, nzg))
        xface = 0.; yface = 0.; zface = 0.; rhokap = 0.; jmean = 0.
        call gridset(xmax, ymax, zmax, 1)
        if (use_rhokap /= 0) rhokap = rhokap_in
        delta = 1.e-8*(2.*zmax/nzg)
        iseed = iseed_c
        nesc = 0
        do j = 1, nphotons
            tflag = .false.
            call sourcephCO2(xmax, ymax, zmax, xcell, ycell, zcell, iseed)
            call tauint1(xmax, ymax, zmax, xcell, ycell, zcell, tflag, iseed, delta)
            if (tflag) nesc = nesc + 1
        end do
        iseed_c = iseed
        jmean_out = jmean
    end subroutine ref_mc

    ! n successive stokes calls from the direction (cost,sint,cosp,sinp,phi);
    ! dirs(1:3,i) = (nxp,nyp,nzp) after call i, angles updated in place.
    subroutine ref_stokes(n, hgg_c, iseed_c, ang, dirs) bind(C, name="ref_stokes")
        use photon_vars, only : nxp, nyp, nzp, cost, sint, cosp, sinp, phi
        use opt_prop,    only : hgg, g2
        use stokes_mod,  only : stokes
        integer(c_int), value         :: n
        real(c_double), value         :: hgg_c
        integer(c_int), intent(inout) :: iseed_c
        real(c_double), intent(inout) :: ang(5)
        real(c_double), intent(out)   :: dirs(3, n)
        integer :: i, iseed
        hgg = hgg_c
        g2 = hgg**2.
        cost = ang(1); sint = ang(2); cosp = ang(3); sinp = ang(4); phi = ang(5)
        nxp = sint*cosp; nyp = sint*sinp; nzp = cost
        iseed = iseed_c
        do i = 1, n
            call stokes(iseed)
            dirs(1, i) = nxp; dirs(2, i) = nyp; dirs(3, i) = nzp
        end do
        iseed_c = iseed
        ang(1) = cost; ang(2) = sint; ang(3) = cosp; ang(4) = sinp; ang(5) = phi
    end subroutine ref_stokes

    ! out(1..6, i) = airThermalCond(T), airDensity(T), getWaterContent(Q,W),
    !                getSkinDensity(W), getSkinHeatCap(W), getSkinThermalCond(W, rho)
    subroutine ref_fits(n, T, W, Q, qvap, rho, out) bind(C, name="ref_fits")
        use thermalConstants
        integer(c_int), value       :: n
        real(c_double), intent(in)  :: T(n), W(n), Q(n), rho(n)
        real(c_double), value       :: qvap
        real(c_double), intent(out) :: out(6, n)
        integer :: i
        real    :: tt, ww, qq, rr
        QVapor = qvap
        do i = 1, n
            tt = T(i); ww = W(i); qq = Q(i); rr = rho(i)
            out(1, i) = airThermalCond(tt, 0)
            out(2, i) = airDensity(tt)
            out(3, i) = getWaterContent(qq, ww)
            out(4, i) = getSkinDensity(ww)
            out(5, i) = getSkinHeatCap(ww)
            out(6, i) = getSkinThermalCond(ww, rr)
        end do
    end subroutine ref_fits

    ! mk_cdf on a (n, 2) table; cdf(n) comes in as the caller's value
    subroutine ref_mk_cdf(n, arr, cdf) bind(C, name="ref_mk_cdf")
        use ch_opt, only : mk_cdf
        integer(c_int), value         :: n
        real(c_double), intent(in)    :: arr(n, 2)
        real(c_double), intent(inout) :: cdf(n)
        real :: a(n, 2), c(n)
        a = arr
        c = cdf
        call mk_cdf(a, c, n)
        cdf = c
    end subroutine ref_mk_cdf

    ! ns successive sample calls (one ran2 draw each) from seed iseed_c
    subroutine ref_sample(n, ns, arr, cdf, iseed_c, waves) bind(C, name="ref_sample")
        use ch_opt, only : sample
        integer(c_int), value         :: n, ns
        real(c_double), intent(in)    :: arr(n, 2), cdf(n)
        integer(c_int), intent(inout) :: iseed_c
        real(c_double), intent(out)   :: waves(ns)
        real    :: a(n, 2), c(n), w
        integer :: i, iseed
        a = arr
        c = cdf
        iseed = iseed_c
        do i = 1, ns
            call sample(a, n, c, w, iseed)
            waves(i) = w
        end do
        iseed_c = iseed
    end subroutine ref_sample

    ! nlow(i, 1) = search_1D over cdf, nlow(i, 2) = search_2D over arr's
    ! first column, for each value(i)
    subroutine ref_search(n, arr, cdf, nv, values, nlow) bind(C, name="ref_search")
        use ch_opt, only : search_1D, search_2D
        integer(c_int), value        :: n, nv
        real(c_double), intent(in)   :: arr(n, 2), cdf(n), values(nv)
        integer(c_int), intent(out)  :: nlow(nv, 2)
        real    :: a(n, 2), c(n), v
        integer :: i, k
        a = arr
        c = cdf
        do i = 1, nv
            v = values(i)
            call search_1D(n, c, k, v)
            nlow(i, 1) = k
            call search_2D(n, a, k, v)
            nlow(i, 2) = k
        end do
    end subroutine ref_search

    ! out(i, 1) = lin_inter_1D(arr, cdf, values(i), n, nlow1(i)),
    ! out(i, 2) = lin_inter_2D(arr, values(i), n, nlow2(i))
    subroutine ref_lin_inter(n, arr, cdf, nv, values, nlow1, nlow2, out) bind(C, name="ref_lin_inter")
        use ch_opt, only : lin_inter_1D, lin_inter_2D
        integer(c_int), value       :: n, nv
        real(c_double), intent(in)  :: arr(n, 2), cdf(n), values(nv)
        integer(c_int), intent(in)  :: nlow1(nv), nlow2(nv)
        real(c_double), intent(out) :: out(nv, 2)
        real    :: a(n, 2), c(n), v, y
        integer :: i, k
        a = arr
        c = cdf
        do i = 1, nv
            v = values(i)
            k = nlow1(i)
            call lin_inter_1D(a, c, v, n, k, y)
            out(i, 1) = y
            k = nlow2(i)
            call lin_inter_2D(a, v, n, k, y)
            out(i, 2) = y
        end do
    end subroutine ref_lin_inter

end module ref_harness
