! External subroutines with the reference's names and argument lists, each
! forwarding to the device library.  A Fortran main loop written against the
! reference (main.f90:743-1149) links these in place of rec.f90, calcha.f90,
! hll.f90, ct.f90, calsf.f90, rkt.f90, recov.f90, bnd.f90, correction.f90,
! caluu.f90, cdtcfl.f90 and func.f90.  Arrays are assumed-size: only their
! base addresses cross the boundary.

subroutine ident(uu, un, is1, ie1, js1, je1, ks1, ke1)          ! func.f90:2
  use raishin_c
  real(8) :: uu(*), un(*)
  integer :: is1, ie1, js1, je1, ks1, ke1
  call raishin_ident(uu, un, is1, ie1, js1, je1, ks1, ke1)
end subroutine ident

subroutine cdtcfl4(uri, gcov, gcon, dtcfl, dx1, dx2, dx3, dt1, dt2, dt3, nm1, is1, ie1, js1, je1, ks1, ke1)
  use raishin_c                                                  ! cdtcfl.f90:2
  real(8) :: uri(*), gcov(*), gcon(*), dtcfl, dx1(*), dx2(*), dx3(*), dt1, dt2, dt3
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_cdtcfl4(uri, gcov, gcon, dtcfl, dx1, dx2, dx3, dt1, dt2, dt3, nm1, is1, ie1, js1, je1, ks1, ke1)
end subroutine cdtcfl4

subroutine rec2(uri, x1, x2, x3, x1a, x2a, x3a, dx1, dx2, dx3, dx1b, dx2b, dx3b, &
                x1b, x2b, x3b, x1ab, x2ab, x3ab, uriir, urijr, urikr, uriil, urijl, urikl, &
                uuir, uujr, uukr, uuil, uujl, uukl, wwir, wwjr, wwkr, wwil, wwjl, wwkl, &
                gcovi, gcovj, gcovk, gconi, gconj, gconk, detgi, detgj, detgk, nm0, &
                is1, ie1, js1, je1, ks1, ke1)                    ! rec.f90:2
  use raishin_c
  real(8) :: uri(*), x1(*), x2(*), x3(*), x1a(*), x2a(*), x3a(*), dx1(*), dx2(*), dx3(*)
  real(8) :: dx1b(*), dx2b(*), dx3b(*), x1b(*), x2b(*), x3b(*), x1ab(*), x2ab(*), x3ab(*)
  real(8) :: uriir(*), urijr(*), urikr(*), uriil(*), urijl(*), urikl(*)
  real(8) :: uuir(*), uujr(*), uukr(*), uuil(*), uujl(*), uukl(*)
  real(8) :: wwir(*), wwjr(*), wwkr(*), wwil(*), wwjl(*), wwkl(*)
  real(8) :: gcovi(*), gcovj(*), gcovk(*), gconi(*), gconj(*), gconk(*), detgi(*), detgj(*), detgk(*)
  integer :: nm0, is1, ie1, js1, je1, ks1, ke1
  call raishin_rec2(uri, x1, x2, x3, x1a, x2a, x3a, dx1, dx2, dx3, dx1b, dx2b, dx3b, &
                    x1b, x2b, x3b, x1ab, x2ab, x3ab, uriir, urijr, urikr, uriil, urijl, urikl, &
                    uuir, uujr, uukr, uuil, uujl, uukl, wwir, wwjr, wwkr, wwil, wwjl, wwkl, &
                    gcovi, gcovj, gcovk, gconi, gconj, gconk, detgi, detgj, detgk, nm0, &
                    is1, ie1, js1, je1, ks1, ke1)
end subroutine rec2

subroutine calcha(uriir, urijr, urikr, uriil, urijl, urikl, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
                  cmaxi, cmini, cmaxj, cminj, cmaxk, cmink, nm0, is1, ie1, js1, je1, ks1, ke1)   ! calcha.f90:2
  use raishin_c
  real(8) :: uriir(*), urijr(*), urikr(*), uriil(*), urijl(*), urikl(*)
  real(8) :: gcovi(*), gcovj(*), gcovk(*), gconi(*), gconj(*), gconk(*)
  real(8) :: cmaxi(*), cmini(*), cmaxj(*), cminj(*), cmaxk(*), cmink(*)
  integer :: nm0, is1, ie1, js1, je1, ks1, ke1
  call raishin_calcha(uriir, urijr, urikr, uriil, urijl, urikl, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
                      cmaxi, cmini, cmaxj, cminj, cmaxk, cmink, nm0, is1, ie1, js1, je1, ks1, ke1)
end subroutine calcha

subroutine hll(ww, uuir, uujr, uukr, uuil, uujl, uukl, wwir, wwjr, wwkr, wwil, wwjl, wwkl, &
               cmaxi, cmini, cmaxj, cminj, cmaxk, cmink, nm0, is1, ie1, js1, je1, ks1, ke1)      ! hll.f90:2
  use raishin_c
  real(8) :: ww(*), uuir(*), uujr(*), uukr(*), uuil(*), uujl(*), uukl(*)
  real(8) :: wwir(*), wwjr(*), wwkr(*), wwil(*), wwjl(*), wwkl(*)
  real(8) :: cmaxi(*), cmini(*), cmaxj(*), cminj(*), cmaxk(*), cmink(*)
  integer :: nm0, is1, ie1, js1, je1, ks1, ke1
  call raishin_hll(ww, uuir, uujr, uukr, uuil, uujl, uukl, wwir, wwjr, wwkr, wwil, wwjl, wwkl, &
                   cmaxi, cmini, cmaxj, cminj, cmaxk, cmink, nm0, is1, ie1, js1, je1, ks1, ke1)
end subroutine hll

subroutine ct(ww, wwo, uri, nm0, nm1, is1, ie1, js1, je1, ks1, ke1)                          ! ct.f90:4
  use raishin_c
  real(8) :: ww(*), wwo(*), uri(*)
  integer :: nm0, nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_ct(ww, wwo, uri, nm0, nm1, is1, ie1, js1, je1, ks1, ke1)
end subroutine ct

subroutine caltenr(uri, tenr, gcov, gcon, is1, ie1, js1, je1, ks1, ke1)                      ! calsf.f90:6
  use raishin_c
  real(8) :: uri(*), tenr(*), gcov(*), gcon(*)
  integer :: is1, ie1, js1, je1, ks1, ke1
  call raishin_caltenr(uri, tenr, gcov, gcon, is1, ie1, js1, je1, ks1, ke1)
end subroutine caltenr

subroutine calsf(sf, tenr, christ, detg, is1, ie1, js1, je1, ks1, ke1)                      ! calsf.f90:107
  use raishin_c
  real(8) :: sf(*), tenr(*), christ(*), detg(*)
  integer :: is1, ie1, js1, je1, ks1, ke1
  call raishin_calsf(sf, tenr, christ, detg, is1, ie1, js1, je1, ks1, ke1)
end subroutine calsf

subroutine rk2fst(uh, ww, uu, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)     ! rkt.f90:2
  use raishin_c
  real(8) :: uh(*), ww(*), uu(*), akap1b(*), akap2b(*), akap3b(*)
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_rk2fst(uh, ww, uu, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)
end subroutine rk2fst

subroutine rk2snd(uu, ww, uo, uh, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1) ! rkt.f90:45
  use raishin_c
  real(8) :: uu(*), ww(*), uo(*), uh(*), akap1b(*), akap2b(*), akap3b(*)
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_rk2snd(uu, ww, uo, uh, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)
end subroutine rk2snd

subroutine rk3snd(uu, ww, uo, uh, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1) ! rkt.f90:91
  use raishin_c
  real(8) :: uu(*), ww(*), uo(*), uh(*), akap1b(*), akap2b(*), akap3b(*)
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_rk3snd(uu, ww, uo, uh, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)
end subroutine rk3snd

subroutine rk3trd(uu, ww, uo, us, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1) ! rkt.f90:137
  use raishin_c
  real(8) :: uu(*), ww(*), uo(*), us(*), akap1b(*), akap2b(*), akap3b(*)
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_rk3trd(uu, ww, uo, us, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)
end subroutine rk3trd

subroutine rk2adsff(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)                          ! rkt.f90:183
  use raishin_c
  real(8) :: uu(*), sf(*), dt
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_rk2adsff(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)
end subroutine rk2adsff

subroutine rk2adsfs(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)                          ! rkt.f90:218
  use raishin_c
  real(8) :: uu(*), sf(*), dt
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_rk2adsfs(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)
end subroutine rk2adsfs

subroutine rk3adsfs(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)                          ! rkt.f90:253
  use raishin_c
  real(8) :: uu(*), sf(*), dt
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_rk3adsfs(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)
end subroutine rk3adsfs

subroutine rk3adsft(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)                          ! rkt.f90:288
  use raishin_c
  real(8) :: uu(*), sf(*), dt
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_rk3adsft(uu, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)
end subroutine rk3adsft

subroutine recov(uu, uri, gcov, gcon, detg, x1, x2, x3, nm1, is1, ie1, js1, je1, ks1, ke1)   ! recov.f90:2
  use raishin_c
  real(8) :: uu(*), uri(*), gcov(*), gcon(*), detg(*), x1(*), x2(*), x3(*)
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1
  call raishin_recov(uu, uri, gcov, gcon, detg, x1, x2, x3, nm1, is1, ie1, js1, je1, ks1, ke1)
end subroutine recov

subroutine bnd4(uu, uo, x1, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
  use raishin_c                                                  ! bnd.f90:4
  real(8) :: uu(*), uo(*), x1(*)
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable(*)
  call raishin_bnd4(uu, uo, x1, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
end subroutine bnd4

subroutine infchk(uri, gcov, gcon, nm1, is1, ie1, js1, je1, ks1, ke1, myranki)               ! bnd.f90:1636
  use raishin_c
  real(8) :: uri(*), gcov(*), gcon(*)
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1, myranki
  call raishin_infchk(uri, gcov, gcon, nm1, is1, ie1, js1, je1, ks1, ke1, myranki)
end subroutine infchk

subroutine atmosphere1a(uu, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)          ! correction.f90:226
  use raishin_c
  real(8) :: uu(*), uri(*), gcov(*), gcon(*), detg(*), x1(*)
  integer :: is1, ie1, js1, je1, ks1, ke1
  call raishin_atmosphere1a(uu, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
end subroutine atmosphere1a

subroutine pminmax2b(uu, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)             ! correction.f90:744
  use raishin_c
  real(8) :: uu(*), uri(*), gcov(*), gcon(*), detg(*), x1(*)
  integer :: is1, ie1, js1, je1, ks1, ke1
  call raishin_pminmax2b(uu, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
end subroutine pminmax2b

subroutine caluu1(uri, uu, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)                ! caluu.f90:2
  use raishin_c
  real(8) :: uri(*), uu(*), gcov(*), gcon(*), detg(*), x1(*)
  integer :: is1, ie1, js1, je1, ks1, ke1
  call raishin_caluu1(uri, uu, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
end subroutine caluu1

! mpisub.f90:62 mpiex: the halo exchange has nothing to exchange with one rank.
subroutine mpiex(uri, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
  real(8) :: uri(*)
  integer :: nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable(*)
end subroutine mpiex

! ---- restart and analysis files (restar.f90) ------------------------------
! main.f90 keeps unit 9 open on restartRRR.outdat (main.f90:339-343); the
! wrappers write or read that file through the library and leave the unit
! open on it again.  The state arrays are pulled from the device first
! (store1, output1) or pushed to it after reading (restar2).

subroutine unit9_name(myrank, fname, was_open)
  integer :: myrank
  character(len=*) :: fname
  logical :: was_open
  inquire(unit=9, opened=was_open, name=fname)
  if (was_open) then
    close(9)
  else
    write(fname, '(a, i3.3, a)') 'restart', myrank, '.outdat'
  end if
end subroutine unit9_name

subroutine store1(uu, u0, uri, uri0, it0, ih0, time, nd, npe, myrank, is1, ie1, js1, je1, ks1, ke1)
  use raishin_c                                                  ! restar.f90:5
  real(8) :: uu(*), u0(*), uri(*), uri0(*), time
  integer :: it0, ih0, nd, npe, myrank, is1, ie1, js1, je1, ks1, ke1
  character(len=1024) :: fname
  logical :: was_open
  call raishin_pull(uu)
  call raishin_pull(u0)
  call raishin_pull(uri)
  call raishin_pull(uri0)
  call unit9_name(myrank, fname, was_open)
  call raishin_store1(trim(fname) // c_null_char, uu, u0, uri, uri0, it0, ih0, time, nd, npe, myrank, &
                      is1, ie1, js1, je1, ks1, ke1)
  if (was_open) open(9, file=trim(fname), form='unformatted', status='old')
end subroutine store1

subroutine restar2(uu, u0, uri, uri0, it0, ih0, time, nd, npe, myrank, is1, ie1, js1, je1, ks1, ke1)
  use raishin_c                                                  ! restar.f90:273
  real(8) :: uu(*), u0(*), uri(*), uri0(*), time
  integer :: it0, ih0, nd, npe, myrank, is1, ie1, js1, je1, ks1, ke1
  character(len=1024) :: fname
  logical :: was_open
  call unit9_name(myrank, fname, was_open)
  call raishin_restar2(trim(fname) // c_null_char, uu, u0, uri, uri0, it0, ih0, time, nd, npe, myrank, &
                       is1, ie1, js1, je1, ks1, ke1)
  call raishin_push(uu)
  call raishin_push(u0)
  call raishin_push(uri)
  call raishin_push(uri0)
  if (was_open) open(9, file=trim(fname), form='unformatted', status='old')
end subroutine restar2

subroutine output1(uri, detg, x1, x2, x3, time, nm1, nd, npe, myrank, myranki, myrankj, myrankk, &
                   is1, ie1, js1, je1, ks1, ke1)
  use raishin_c                                                  ! restar.f90:91
  real(8) :: uri(*), detg(*), x1(*), x2(*), x3(*), time
  integer :: nm1, nd, npe, myrank, myranki, myrankj, myrankk, is1, ie1, js1, je1, ks1, ke1
  character(len=64) :: fname
  call raishin_pull(uri)
  write(fname, '(a, i4.4, a, i4.4, a)') 'structr', myrank, '-', nd, '.outdat'
  call raishin_output1(trim(fname) // c_null_char, uri, detg, x1, x2, x3, time, nm1, nd, npe, myrank, &
                       myranki, myrankj, myrankk, is1, ie1, js1, je1, ks1, ke1)
end subroutine output1
