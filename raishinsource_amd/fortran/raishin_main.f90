! A Fortran main loop in the shape of the reference's (main.f90:586-1151),
! calling the reference-named routines, which raishin_wrappers.f90 binds to
! the device library.  Set-up uses the library's host routines (grid, Kerr-
! Schild metric, Christoffel symbols, magnetised torus).
!
! usage: raishin_main imax jmax kmax irec irkt cfl nsteps out.bin [noise seed [files]]
! out.bin (stream, native endian): int32 imax,jmax,kmax,nsteps; real64
! dt(nsteps), time; uri(9,imax,jmax,kmax); uu(9,imax,jmax,kmax);
! int64 counters(11).
! files (bits): 2 = start from restart000.outdat through restar2 (icres=1,
! main.f90:560-571) instead of the torus; 1 = after the last step write
! restart000.outdat (store1) and structr0000-0001.outdat (output1), as the
! output branch of main.f90:642-716 does.
program raishin_main
  use iso_c_binding
  use raishin_c
  implicit none
  integer, parameter :: nv = 9
  integer :: imax, jmax, kmax, irec, irkt, nsteps, nm0, nm1, it, i, j, k, files, it0, ihi, nd, npe, myrank
  integer :: is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable(1)
  real(8) :: cfl, noise, dtcfl, dtcflg, dt1, dt2, dt3, dt, time
  integer(c_long_long) :: seed
  character(len=256) :: arg, outfile
  type(raishin_pram) :: p
  type(raishin_grid_ext) :: ext
  type(raishin_torus) :: tor
  type(raishin_counters) :: cnt
  real(8), allocatable :: x1(:), x2(:), x3(:), dx1(:), dx2(:), dx3(:), dx1b(:), dx2b(:), dx3b(:)
  real(8), allocatable :: x1a(:), x2a(:), x3a(:), x1b(:), x2b(:), x3b(:), x1ab(:), x2ab(:), x3ab(:)
  real(8), allocatable :: akap1b(:), akap2b(:), akap3b(:)
  real(8), allocatable :: gcov(:,:,:,:,:), gcon(:,:,:,:,:), gcovi(:,:,:,:,:), gconi(:,:,:,:,:)
  real(8), allocatable :: gcovj(:,:,:,:,:), gconj(:,:,:,:,:), gcovk(:,:,:,:,:), gconk(:,:,:,:,:)
  real(8), allocatable :: detg(:,:,:), detgi(:,:,:), detgj(:,:,:), detgk(:,:,:), christ(:,:,:,:,:,:)
  real(8), allocatable :: uri(:,:,:,:), urio(:,:,:,:), uu(:,:,:,:), uo(:,:,:,:), uh(:,:,:,:)
  real(8), allocatable :: us(:,:,:,:), un(:,:,:,:), ww(:,:,:,:,:), wwo(:), u0(:,:,:,:), uri0(:,:,:,:)
  ! face states, speeds, tenr and sf live inside the library: never written
  real(8) :: face(1), cmx(1), tenr(1), sf(1)
  real(8), allocatable :: dts(:)
  real(8), parameter :: pi = 3.14159265358979323846d0

  if (command_argument_count() < 8) then
    write(0, *) 'usage: raishin_main imax jmax kmax irec irkt cfl nsteps out.bin [noise seed]'
    error stop 1
  end if
  call get_command_argument(1, arg); read(arg, *) imax
  call get_command_argument(2, arg); read(arg, *) jmax
  call get_command_argument(3, arg); read(arg, *) kmax
  call get_command_argument(4, arg); read(arg, *) irec
  call get_command_argument(5, arg); read(arg, *) irkt
  call get_command_argument(6, arg); read(arg, *) cfl
  call get_command_argument(7, arg); read(arg, *) nsteps
  call get_command_argument(8, outfile)
  noise = 0.02d0
  seed = 0
  files = 0
  if (command_argument_count() >= 10) then
    call get_command_argument(9, arg); read(arg, *) noise
    call get_command_argument(10, arg); read(arg, *) seed
  end if
  if (command_argument_count() >= 11) then
    call get_command_argument(11, arg); read(arg, *) files
  end if
  npe = 1; myrank = 0; nd = 1; it0 = 1; ihi = 0

  ! ghost widths, main.f90:179-211 (flux-CT adds one)
  if (irec == 5) then
    nm0 = 4
  else if (irec <= 3) then
    nm0 = 2
  else
    nm0 = 3
  end if
  nm1 = nm0 + 1
  is1 = 1; ie1 = imax; js1 = 1; je1 = jmax; ks1 = 1; ke1 = kmax
  myranki = 0; myrankj = 0; myrankk = 0; icputable = 0

  ! pram.f90 values of the Kerr torus runs
  p%imax = imax; p%jmax = jmax; p%kmax = kmax
  p%irec = irec; p%irkt = irkt; p%ict = 1; p%ihll = 1; p%icha = 2; p%iwvec = 1; p%ieos = 0
  p%model = 7; p%metric = 303; p%iter = 5000
  p%cfl = cfl; p%gam = 4.d0 / 3.d0; p%akm = 0.9375d0; p%dmin = 1.d-4; p%pmin = 1.d-6; p%pmax = 1.d6
  p%iboux1in = 5
  p%iboux1ot = (/5, 18, 18, 18, 5, 5, 5, 5, 5/)
  p%iboux2in = 1
  p%iboux2ot = 1
  p%iboux3in = (/14, 14, 14, 12, 14, 14, 14, 14, 12/)
  p%iboux3ot = (/14, 14, 14, 12, 14, 14, 14, 14, 12/)
  ext%ix1 = 2
  ext%xmin = 0.98d0 * (1.d0 + sqrt(1.d0 - p%akm**2))
  ext%xmax = 50.d0
  ext%ymin = 0.d0
  ext%ymax = 2.d0 * pi / dble(jmax - 2 * nm1) * dble(jmax - 1)
  ext%zmin = 0.d0
  ext%zmax = pi
  tor%rin = 6.d0; tor%rmax = 12.d0; tor%beta = 100.d0; tor%acut = 0.2d0; tor%noise = noise; tor%seed = seed

  allocate(x1(imax), x2(jmax), x3(kmax), dx1(imax), dx2(jmax), dx3(kmax), dx1b(imax), dx2b(jmax), dx3b(kmax))
  allocate(x1a(imax), x2a(jmax), x3a(kmax), x1b(imax), x2b(jmax), x3b(kmax), x1ab(imax), x2ab(jmax), x3ab(kmax))
  allocate(akap1b(imax), akap2b(jmax), akap3b(kmax))
  allocate(gcov(4,4,imax,jmax,kmax), gcon(4,4,imax,jmax,kmax), gcovi(4,4,imax,jmax,kmax), gconi(4,4,imax,jmax,kmax))
  allocate(gcovj(4,4,imax,jmax,kmax), gconj(4,4,imax,jmax,kmax), gcovk(4,4,imax,jmax,kmax), gconk(4,4,imax,jmax,kmax))
  allocate(detg(imax,jmax,kmax), detgi(imax,jmax,kmax), detgj(imax,jmax,kmax), detgk(imax,jmax,kmax))
  allocate(christ(4,4,4,imax,jmax,kmax))
  allocate(uri(nv,imax,jmax,kmax), urio(nv,imax,jmax,kmax), uu(nv,imax,jmax,kmax), uo(nv,imax,jmax,kmax))
  allocate(uh(nv,imax,jmax,kmax), us(nv,imax,jmax,kmax), un(nv,imax,jmax,kmax), ww(3,nv,imax,jmax,kmax), wwo(1))
  allocate(dts(nsteps), u0(nv,imax,jmax,kmax), uri0(nv,imax,jmax,kmax))
  x1ab = 0.d0; x2ab = 0.d0; x3ab = 0.d0
  ww = 0.d0                                       ! main.f90:372-381
  uh = 0.d0; us = 0.d0; un = 0.d0; uo = 0.d0; urio = 0.d0

  call raishin_grid(p, ext, x1, x2, x3, dx1, dx2, dx3, dx1b, dx2b, dx3b, x1a, x2a, x3a, x1b, x2b, x3b)
  call raishin_coord1(p, gcov, gcovi, gcovj, gcovk, gcon, gconi, gconj, gconk, detg, detgi, detgj, detgk, &
                      x1, x2, x3, x1a, x2a, x3a)
  call raishin_calchrist(p, christ, x1, x2, x3)
  if (raishin_mdtorus1(p, tor, uri, gcov, gcon, detg, x1, x2, x3) /= 0) error stop 'no torus cells'
  call raishin_init(p)
  call raishin_check('raishin_init')
  call raishin_geometry(x1, x2, x3, dx1b, dx2b, dx3b, gcov, gcon, gcovi, gconi, gcovj, gconj, gcovk, gconk, &
                        detg, detgi, detgj, detgk, christ)
  call raishin_check('raishin_geometry')

  call caluu1(uri, uu, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)     ! main.f90:530
  call ident(uu, u0, is1, ie1, js1, je1, ks1, ke1)                             ! main.f90:535-538
  call ident(uri, uri0, is1, ie1, js1, je1, ks1, ke1)
  time = 0.d0
  if (files /= 0) open(9, file='restart000.outdat', status='unknown', form='unformatted')  ! main.f90:339-343
  if (iand(files, 2) /= 0) then
    call restar2(uu, u0, uri, uri0, it0, ihi, time, nd, npe, myrank, is1, ie1, js1, je1, ks1, ke1)
    call raishin_check('restar2')
  end if

  do it = 1, nsteps
    call ident(uu, uo, is1, ie1, js1, je1, ks1, ke1)
    call ident(uri, urio, is1, ie1, js1, je1, ks1, ke1)
    call cdtcfl4(uri, gcov, gcon, dtcfl, dx1b, dx2b, dx3b, dt1, dt2, dt3, nm1, is1, ie1, js1, je1, ks1, ke1)
    call raishin_check('cdtcfl4')
    dtcflg = dtcfl                                ! MIN over one rank
    dt = cfl * dtcflg
    do i = 1, imax - 1
      akap1b(i) = dt / dx1b(i)
    end do
    do j = 1, jmax - 1
      akap2b(j) = dt / dx2b(j)
    end do
    do k = 1, kmax - 1
      akap3b(k) = dt / dx3b(k)
    end do
    dts(it) = dt
    ! first stage
    call rec2(uri, x1, x2, x3, x1a, x2a, x3a, dx1, dx2, dx3, dx1b, dx2b, dx3b, x1b, x2b, x3b, x1ab, x2ab, x3ab, &
              face, face, face, face, face, face, face, face, face, face, face, face, &
              face, face, face, face, face, face, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
              detgi, detgj, detgk, nm0, is1, ie1, js1, je1, ks1, ke1)
    call calcha(face, face, face, face, face, face, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
                cmx, cmx, cmx, cmx, cmx, cmx, nm0, is1, ie1, js1, je1, ks1, ke1)
    call hll(ww, face, face, face, face, face, face, face, face, face, face, face, face, &
             cmx, cmx, cmx, cmx, cmx, cmx, nm0, is1, ie1, js1, je1, ks1, ke1)
    call ct(ww, wwo, uri, nm0, nm1, is1, ie1, js1, je1, ks1, ke1)
    call caltenr(uri, tenr, gcov, gcon, is1, ie1, js1, je1, ks1, ke1)
    call calsf(sf, tenr, christ, detg, is1, ie1, js1, je1, ks1, ke1)
    call rk2fst(uh, ww, uu, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)
    call rk2adsff(uh, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)
    call recov(uh, uri, gcov, gcon, detg, x1, x2, x3, nm1, is1, ie1, js1, je1, ks1, ke1)
    call bnd4(uri, urio, x1, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
    call infchk(uri, gcov, gcon, nm1, is1, ie1, js1, je1, ks1, ke1, myranki)
    call atmosphere1a(uh, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
    call pminmax2b(uh, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
    call caluu1(uri, uh, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
    call mpiex(uri, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
    call ident(uh, uu, is1, ie1, js1, je1, ks1, ke1)
    call raishin_check('first stage')
    ! second stage
    call rec2(uri, x1, x2, x3, x1a, x2a, x3a, dx1, dx2, dx3, dx1b, dx2b, dx3b, x1b, x2b, x3b, x1ab, x2ab, x3ab, &
              face, face, face, face, face, face, face, face, face, face, face, face, &
              face, face, face, face, face, face, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
              detgi, detgj, detgk, nm0, is1, ie1, js1, je1, ks1, ke1)
    call calcha(face, face, face, face, face, face, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
                cmx, cmx, cmx, cmx, cmx, cmx, nm0, is1, ie1, js1, je1, ks1, ke1)
    call hll(ww, face, face, face, face, face, face, face, face, face, face, face, face, &
             cmx, cmx, cmx, cmx, cmx, cmx, nm0, is1, ie1, js1, je1, ks1, ke1)
    call ct(ww, wwo, uri, nm0, nm1, is1, ie1, js1, je1, ks1, ke1)
    call caltenr(uri, tenr, gcov, gcon, is1, ie1, js1, je1, ks1, ke1)
    call calsf(sf, tenr, christ, detg, is1, ie1, js1, je1, ks1, ke1)
    if (irkt == 2) then
      call rk2snd(us, ww, uo, uh, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)
      call rk2adsfs(us, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)
    else
      call rk3snd(us, ww, uo, uh, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)
      call rk3adsfs(us, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)
    end if
    call recov(us, uri, gcov, gcon, detg, x1, x2, x3, nm1, is1, ie1, js1, je1, ks1, ke1)
    call bnd4(uri, urio, x1, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
    call infchk(uri, gcov, gcon, nm1, is1, ie1, js1, je1, ks1, ke1, myranki)
    call atmosphere1a(us, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
    call pminmax2b(us, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
    call caluu1(uri, us, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
    call mpiex(uri, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
    call ident(us, uu, is1, ie1, js1, je1, ks1, ke1)
    call raishin_check('second stage')
    if (irkt == 3) then
      call rec2(uri, x1, x2, x3, x1a, x2a, x3a, dx1, dx2, dx3, dx1b, dx2b, dx3b, x1b, x2b, x3b, x1ab, x2ab, x3ab, &
                face, face, face, face, face, face, face, face, face, face, face, face, &
                face, face, face, face, face, face, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
                detgi, detgj, detgk, nm0, is1, ie1, js1, je1, ks1, ke1)
      call calcha(face, face, face, face, face, face, gcovi, gcovj, gcovk, gconi, gconj, gconk, &
                  cmx, cmx, cmx, cmx, cmx, cmx, nm0, is1, ie1, js1, je1, ks1, ke1)
      call hll(ww, face, face, face, face, face, face, face, face, face, face, face, face, &
               cmx, cmx, cmx, cmx, cmx, cmx, nm0, is1, ie1, js1, je1, ks1, ke1)
      call ct(ww, wwo, uri, nm0, nm1, is1, ie1, js1, je1, ks1, ke1)
      call caltenr(uri, tenr, gcov, gcon, is1, ie1, js1, je1, ks1, ke1)
      call calsf(sf, tenr, christ, detg, is1, ie1, js1, je1, ks1, ke1)
      call rk3trd(un, ww, uo, us, nm1, akap1b, akap2b, akap3b, is1, ie1, js1, je1, ks1, ke1)
      call rk3adsft(un, sf, dt, nm1, is1, ie1, js1, je1, ks1, ke1)
      call recov(un, uri, gcov, gcon, detg, x1, x2, x3, nm1, is1, ie1, js1, je1, ks1, ke1)
      call bnd4(uri, urio, x1, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
      call infchk(uri, gcov, gcon, nm1, is1, ie1, js1, je1, ks1, ke1, myranki)
      call atmosphere1a(un, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
      call pminmax2b(un, uri, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
      call caluu1(uri, un, gcov, gcon, detg, x1, is1, ie1, js1, je1, ks1, ke1)
      call mpiex(uri, nm1, is1, ie1, js1, je1, ks1, ke1, myranki, myrankj, myrankk, icputable)
      call ident(un, uu, is1, ie1, js1, je1, ks1, ke1)
      call raishin_check('third stage')
    end if
    time = time + dt
  end do

  if (iand(files, 1) /= 0) then
    ihi = ihi + nsteps
    call store1(uu, u0, uri, uri0, it0 + nsteps, ihi, time, nd, npe, myrank, is1, ie1, js1, je1, ks1, ke1)
    call output1(uri, detg, x1, x2, x3, time, nm1, nd, npe, myrank, myranki, myrankj, myrankk, &
                 is1, ie1, js1, je1, ks1, ke1)
    call raishin_check('store1/output1')
  end if
  if (files /= 0) close(9)
  call raishin_pull(uri)
  call raishin_pull(uu)
  call raishin_get_counters(cnt)
  call raishin_check('pull')
  open(10, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(10) int(imax, 4), int(jmax, 4), int(kmax, 4), int(nsteps, 4)
  write(10) dts, time
  write(10) uri
  write(10) uu
  write(10) cnt
  close(10)
  call raishin_finalize()
  write(6, '(a, i6, a, es24.16)') 'raishin_main: ', nsteps, ' steps, time =', time
end program raishin_main
