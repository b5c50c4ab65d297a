! iso_c_binding interfaces of libraishin_hip.so (include/raishin_hip.h,
! include/raishin_ref.h, include/raishin_setup.h, include/raishin_io.h).  Integers and arrays are
! passed by address, as Fortran passes them, so the reference-named wrappers
! in raishin_wrappers.f90 forward their dummy arguments unchanged.
module raishin_c
  use iso_c_binding
  implicit none

  ! module pram of the reference (pram.f90:11-80)
  type, bind(C) :: raishin_pram
    integer(c_int) :: imax, jmax, kmax
    integer(c_int) :: irec, irkt, ict, ihll, icha, iwvec, ieos, model, metric, iter
    real(c_double) :: cfl, gam, akm, dmin, pmin, pmax
    integer(c_int) :: iboux1in(9), iboux1ot(9), iboux2in(9), iboux2ot(9), iboux3in(9), iboux3ot(9)
  end type raishin_pram

  type, bind(C) :: raishin_counters
    integer(c_long_long) :: recov_fallback, recov_fail, recov_neg_rho, recov_neg_p, recov_vsq_gt1
    integer(c_long_long) :: recov_det0, recov_wg_carry, atm_floor, pmm_scale, pmm_floor, guard
  end type raishin_counters

  type, bind(C) :: raishin_grid_ext
    integer(c_int) :: ix1
    real(c_double) :: xmin, xmax, ymin, ymax, zmin, zmax
  end type raishin_grid_ext

  type, bind(C) :: raishin_torus
    real(c_double) :: rin, rmax, beta, acut, noise
    integer(c_long_long) :: seed
  end type raishin_torus

  interface
    ! ---- errors and residency
    function raishin_last_error(buf, len) bind(C, name="raishin_last_error") result(r)
      import :: c_char, c_int
      character(kind=c_char) :: buf(*)
      integer(c_int), value :: len
      integer(c_int) :: r
    end function
    subroutine raishin_init(p) bind(C, name="raishin_init")
      import :: raishin_pram
      type(raishin_pram), intent(in) :: p
    end subroutine
    subroutine raishin_geometry(x1, x2, x3, dx1b, dx2b, dx3b, gcov, gcon, gcovi, gconi, gcovj, gconj, &
                                gcovk, gconk, detg, detgi, detgj, detgk, christ) bind(C, name="raishin_geometry")
      import :: c_double
      real(c_double) :: x1(*), x2(*), x3(*), dx1b(*), dx2b(*), dx3b(*), gcov(*), gcon(*), gcovi(*), gconi(*)
      real(c_double) :: gcovj(*), gconj(*), gcovk(*), gconk(*), detg(*), detgi(*), detgj(*), detgk(*), christ(*)
    end subroutine
    subroutine raishin_finalize() bind(C, name="raishin_finalize")
    end subroutine
    subroutine raishin_push(a) bind(C, name="raishin_push")
      import :: c_double
      real(c_double) :: a(*)
    end subroutine
    subroutine raishin_pull(a) bind(C, name="raishin_pull")
      import :: c_double
      real(c_double) :: a(*)
    end subroutine
    subroutine raishin_sync() bind(C, name="raishin_sync")
    end subroutine
    ! ---- restart and analysis files (include/raishin_io.h); path is NUL-terminated
    subroutine raishin_store1(path, uu, u0, uri, uri0, it0, ih0, time, nd, npe, myrank, &
                              is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_store1")
      import :: c_char, c_double, c_int
      character(kind=c_char) :: path(*)
      real(c_double) :: uu(*), u0(*), uri(*), uri0(*), time
      integer(c_int) :: it0, ih0, nd, npe, myrank, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_restar2(path, uu, u0, uri, uri0, it0, ih0, time, nd, npe, myrank, &
                               is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_restar2")
      import :: c_char, c_double, c_int
      character(kind=c_char) :: path(*)
      real(c_double) :: uu(*), u0(*), uri(*), uri0(*), time
      integer(c_int) :: it0, ih0, nd, npe, myrank, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_output1(path, uri, detg, x1, x2, x3, time, nm1, nd, npe, myrank, myranki, myrankj, &
                               myrankk, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_output1")
      import :: c_char, c_double, c_int
      character(kind=c_char) :: path(*)
      real(c_double) :: uri(*), detg(*), x1(*), x2(*), x3(*), time
      integer(c_int) :: nm1, nd, npe, myrank, myranki, myrankj, myrankk, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_reset_counters() bind(C, name="raishin_reset_counters")
    end subroutine
    subroutine raishin_get_counters(c) bind(C, name="raishin_get_counters")
      import :: raishin_counters
      type(raishin_counters) :: c
    end subroutine

    ! ---- set-up (grid.f90:2, coord.f90:5/1089, mdgrmhd.f90:879)
    subroutine raishin_grid(p, ext, x1, x2, x3, dx1, dx2, dx3, dx1b, dx2b, dx3b, x1a, x2a, x3a, x1b, x2b, x3b) &
        bind(C, name="raishin_grid")
      import :: raishin_pram, raishin_grid_ext, c_double
      type(raishin_pram), intent(in) :: p
      type(raishin_grid_ext), intent(in) :: ext
      real(c_double) :: x1(*), x2(*), x3(*), dx1(*), dx2(*), dx3(*), dx1b(*), dx2b(*), dx3b(*)
      real(c_double) :: x1a(*), x2a(*), x3a(*), x1b(*), x2b(*), x3b(*)
    end subroutine
    subroutine raishin_coord1(p, gcov, gcovi, gcovj, gcovk, gcon, gconi, gconj, gconk, detg, detgi, detgj, detgk, &
                              x1, x2, x3, x1a, x2a, x3a) bind(C, name="raishin_coord1")
      import :: raishin_pram, c_double
      type(raishin_pram), intent(in) :: p
      real(c_double) :: gcov(*), gcovi(*), gcovj(*), gcovk(*), gcon(*), gconi(*), gconj(*), gconk(*)
      real(c_double) :: detg(*), detgi(*), detgj(*), detgk(*), x1(*), x2(*), x3(*), x1a(*), x2a(*), x3a(*)
    end subroutine
    subroutine raishin_calchrist(p, christ, x1, x2, x3) bind(C, name="raishin_calchrist")
      import :: raishin_pram, c_double
      type(raishin_pram), intent(in) :: p
      real(c_double) :: christ(*), x1(*), x2(*), x3(*)
    end subroutine
    function raishin_mdtorus1(p, t, uri, gcov, gcon, detg, x1, x2, x3) bind(C, name="raishin_mdtorus1") result(r)
      import :: raishin_pram, raishin_torus, c_double, c_int
      type(raishin_pram), intent(in) :: p
      type(raishin_torus), intent(in) :: t
      real(c_double) :: uri(*), gcov(*), gcon(*), detg(*), x1(*), x2(*), x3(*)
      integer(c_int) :: r
    end function

    ! ---- reference-named entry points (include/raishin_ref.h)
    subroutine raishin_ident(uu, un, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_ident")
      import :: c_double, c_int
      real(c_double) :: uu(*), un(*)
      integer(c_int) :: is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_cdtcfl4(uri, gcov, gcon, dtcfl, dx1, dx2, dx3, dt1, dt2, dt3, nm1, &
                               is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_cdtcfl4")
      import :: c_double, c_int
      real(c_double) :: uri(*), gcov(*), gcon(*), dtcfl, dx1(*), dx2(*), dx3(*), dt1, dt2, dt3
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rec2(uri, x1, x2, x3, x1a, x2a, x3a, dx1, dx2, dx3, dx1b, dx2b, dx3b, &
                            x1b, x2b, x3b, x1ab, x2ab, x3ab, uriir, urijr, urikr, uriil, urijl, urikl, &
                            uuir, uujr, uukr, uuil, uujl, uukl, wwir, wwjr, wwkr, wwil, wwjl, wwkl, &
                            gcovi, gcovj, gcovk, gconi, gconj, gconk, detgi, detgj, detgk, nm0, &
                            is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_rec2")
      import :: c_double, c_int
      real(c_double) :: uri(*), x1(*), x2(*), x3(*), x1a(*), x2a(*), x3a(*), dx1(*), dx2(*), dx3(*)
      real(c_double) :: dx1b(*), dx2b(*), dx3b(*), x1b(*), x2b(*), x3b(*), x1ab(*), x2ab(*), x3ab(*)
      real(c_double) :: uriir(*), urijr(*), urikr(*), uriil(*), urijl(*), urikl(*)
      real(c_double) :: uuir(*), uujr(*), uukr(*), uuil(*), uujl(*), uukl(*)
      real(c_double) :: wwir(*), wwjr(*), wwkr(*), wwil(*), wwjl(*), wwkl(*)
      real(c_double) :: gcovi(*), gcovj(*), gcovk(*), gconi(*), gconj(*), gconk(*), detgi(*), detgj(*), detgk(*)
      integer(c_int) :: nm0, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_calcha(uriir, urijr, urikr, uriil, urijl, urikl, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
                              cmaxi, cmini, cmaxj, cminj, cmaxk, cmink, nm0, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_calcha")
      import :: c_double, c_int
      real(c_double) :: uriir(*), urijr(*), urikr(*), uriil(*), urijl(*), urikl(*)
      real(c_double) :: gcovi(*), gcovj(*), gcovk(*), gconi(*), gconj(*), gconk(*)
      real(c_double) :: cmaxi(*), cmini(*), cmaxj(*), cminj(*), cmaxk(*), cmink(*)
      integer(c_int) :: nm0, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_hll(ww, uuir, uujr, uukr, uuil, uujl, uukl, wwir, wwjr, wwkr, wwil, wwjl, wwkl, &
                           cmaxi, cmini, cmaxj, cminj, cmaxk, cmink, nm0, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_hll")
      import :: c_double, c_int
      real(c_double) :: ww(*), uuir(*), uujr(*), uukr(*), uuil(*), uujl(*), uukl(*)
      real(c_double) :: wwir(*), wwjr(*), wwkr(*), wwil(*), wwjl(*), wwkl(*)
      real(c_double) :: cmaxi(*), cmini(*), cmaxj(*), cminj(*), cmaxk(*), cmink(*)
      integer(c_int) :: nm0, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_ct(ww, wwo, uri, nm0, nm1, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_ct")
      import :: c_double, c_int
      real(c_double) :: ww(*), wwo(*), uri(*)
      integer(c_int) :: nm0, nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_caltenr(uri, tenr, gcov, gcon, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_caltenr")
      import :: c_double, c_int
      real(c_double) :: uri(*), tenr(*), gcov(*), gcon(*)
      integer(c_int) :: is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_calsf(sf, tenr, christ, detg, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_calsf")
      import :: c_double, c_int
      real(c_double) :: sf(*), tenr(*), christ(*), detg(*)
      integer(c_int) :: is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rk2fst(uh, ww, uu, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_rk2fst")
      import :: c_double, c_int
      real(c_double) :: uh(*), ww(*), uu(*), akap1b(*), akap2b(*), akap3b(*)
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rk2snd(uu, ww, uo, uh, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_rk2snd")
      import :: c_double, c_int
      real(c_double) :: uu(*), ww(*), uo(*), uh(*), akap1b(*), akap2b(*), akap3b(*)
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rk3snd(uu, ww, uo, uh, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_rk3snd")
      import :: c_double, c_int
      real(c_double) :: uu(*), ww(*), uo(*), uh(*), akap1b(*), akap2b(*), akap3b(*)
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rk3trd(uu, ww, uo, us, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_rk3trd")
      import :: c_double, c_int
      real(c_double) :: uu(*), ww(*), uo(*), us(*), akap1b(*), akap2b(*), akap3b(*)
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rk2adsff(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_rk2adsff")
      import :: c_double, c_int
      real(c_double) :: uu(*), sf(*), dt
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rk2adsfs(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_rk2adsfs")
      import :: c_double, c_int
      real(c_double) :: uu(*), sf(*), dt
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rk3adsfs(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_rk3adsfs")
      import :: c_double, c_int
      real(c_double) :: uu(*), sf(*), dt
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_rk3adsft(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1) bind(C, name="raishin_rk3adsft")
      import :: c_double, c_int
      real(c_double) :: uu(*), sf(*), dt
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_recov(uu, uri, gcov, gcon, detg, x1, x2, x3, nm1, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_recov")
      import :: c_double, c_int
      real(c_double) :: uu(*), uri(*), gcov(*), gcon(*), detg(*), x1(*), x2(*), x3(*)
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_bnd4(uu, uo, x1, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable) &
        bind(C, name="raishin_bnd4")
      import :: c_double, c_int
      real(c_double) :: uu(*), uo(*), x1(*)
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable(*)
    end subroutine
    subroutine raishin_infchk(uri, gcov, gcon, nm1, is1, ie1, js1, je1, ks1, ke1, myranki) &
        bind(C, name="raishin_infchk")
      import :: c_double, c_int
      real(c_double) :: uri(*), gcov(*), gcon(*)
      integer(c_int) :: nm1, is1, ie1, js1, je1, ks1, ke1, myranki
    end subroutine
    subroutine raishin_atmosphere1a(uu, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_atmosphere1a")
      import :: c_double, c_int
      real(c_double) :: uu(*), uri(*), gcov(*), gcon(*), detg(*), x1(*)
      integer(c_int) :: is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_pminmax2b(uu, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_pminmax2b")
      import :: c_double, c_int
      real(c_double) :: uu(*), uri(*), gcov(*), gcon(*), detg(*), x1(*)
      integer(c_int) :: is1, ie1, js1, je1, ks1, ke1
    end subroutine
    subroutine raishin_caluu1(uri, uu, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1) &
        bind(C, name="raishin_caluu1")
      import :: c_double, c_int
      real(c_double) :: uri(*), uu(*), gcov(*), gcon(*), detg(*), x1(*)
      integer(c_int) :: is1, ie1, js1, je1, ks1, ke1
    end subroutine
  end interface

contains

  ! Stop with the library's message if a call failed.
  subroutine raishin_check(where)
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: n
    if (raishin_last_error(buf, 512_c_int) /= 0) then
      msg = ' '
      do n = 1, 512
        if (buf(n) == c_null_char) exit
        msg(n:n) = buf(n)
      end do
      write(0, '(a, ": ", a)') where, trim(msg)
      error stop 2
    end if
  end subroutine raishin_check

end module raishin_c
