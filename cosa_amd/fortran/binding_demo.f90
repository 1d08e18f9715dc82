!-----------------------------------------------------------------------
! binding_demo.f90 -- drives libcosahip from Fortran through
! cosahip_binding, the way COSA's driver does: one 9x9x9-node Cartesian
! block with farfield (30) on every face, reference-layout arrays built in
! Fortran (q(-1:imax+1,...,npde), si(lmet,0:imax+1,...)), uploaded, one
! smooth() iteration, then the state is read back.  A uniform freestream is
! a steady solution, so the residual norm must vanish and q be unchanged.
! Exit status 0 on success.
!-----------------------------------------------------------------------
program binding_demo
  use iso_c_binding
  use cosahip_binding
  implicit none
  integer, parameter :: im = 9, jm = 9, km = 9, npde = 5, lmet = 4
  type(cosahip_config) :: cfg
  type(c_ptr) :: s
  real(c_double), allocatable :: q(:,:,:,:), q2(:,:,:,:), si(:,:,:,:), sj(:,:,:,:), sk(:,:,:,:), vol(:,:,:)
  integer(c_int32_t) :: imax(1), jmax(1), kmax(1), nbcs(1), bctopo(10,6), cutopo(21), done
  real(c_double) :: resrms(8), gam, mach, al, h
  integer :: rc, i, ib, ext(3)
  integer(c_int64_t) :: n

  gam = 1.4d0; mach = 0.5d0; al = 2.d0 * acos(-1.d0) / 180.d0; h = 1.d0 / (im - 1)
  cfg%npde = npde; cfg%nharms = 0; cfg%hbmove = 0; cfg%lmet = lmet; cfg%nlevel = 1; cfg%nl_crs = 4
  cfg%nstage = 4; cfg%irsop = 1; cfg%cutcirs = 0; cfg%ilimf = 4; cfg%ilimt = 4; cfg%iupdt = 1
  cfg%iprint = 1; cfg%ncycle = 10
  cfg%kom = 0; cfg%kom_bsl = 0; cfg%kom_sst = 0; cfg%viscous = 0; cfg%moving = 0; cfg%harbal = 0
  cfg%foturb = 0; cfg%efx_setup_f = 0; cfg%efx_all_f = 0; cfg%eig_limit_f = 0; cfg%eig_lim_all_f = 0
  cfg%efx_setup_t = 0; cfg%efx_all_t = 0; cfg%eig_limit_t = 0; cfg%eig_lim_all_t = 0
  cfg%entfx_f_w = 0; cfg%entfx_t_w = 0; cfg%bln_prol = 1; cfg%ho_rest = 0; cfg%rkimp = 0
  cfg%gamma = gam; cfg%machfs = mach; cfg%alpha = al; cfg%beta = 0.d0
  cfg%rhoinf = 1.d0; cfg%pinf = 1.d0 / gam; cfg%ainf = 1.d0
  cfg%cfl = 4.d0; cfg%cflt = 4.d0; cfg%rat = 2.d0; cfg%ratt = 2.d0; cfg%psirs = 0.d0; cfg%rkap = 0.333333d0
  cfg%epslimf = 1.d-6; cfg%epslimt = 1.d-6; cfg%cntrpy_f = 0.d0; cfg%cntrpy_t = 0.d0
  cfg%entfxctf_f = 0.d0; cfg%entfxctf_t = 0.d0; cfg%eig_cutoff_f = 0.d0; cfg%eig_cutoff_t = 0.d0
  cfg%omega = 0.d0; cfg%toler = 0.d0
  cfg%qmin = (/1.d-8, 1.d0, 1.d0, 1.d0, 1.d-8, 1.d-12, 1.d-8/)
  cfg%alphas = (/0.25d0, 0.333333d0, 0.5d0, 1.d0, 0.d0/)
  cfg%tkefar = 0.d0; cfg%omefar = 0.d0
  cfg%turcmp = 0; cfg%lim_prodtke = 0; cfg%lim_prodomega = 0; cfg%wallwilc = 1; cfg%ifixq = 0
  cfg%prd_all_lvls = 0; cfg%lim_corr = 0; cfg%pad1 = 0
  cfg%reyno = 1.d6; cfg%pranl = 0.72d0; cfg%prant = 0.9d0; cfg%stemp = 0.383d0; cfg%cdff = 1.d0
  cfg%epst = 1.d-12; cfg%cmu = 1.d0; cfg%sigk = 0.5d0; cfg%sigw = 0.5d0; cfg%bkw = 0.075d0
  cfg%bstrkw = 0.09d0; cfg%gkw = 0.5527d0; cfg%prdlim = 20.d0; cfg%roughk = 0.d0
  cfg%ramping = 0; cfg%ncycr = 0; cfg%pad2 = 0; cfg%vcfli = 0.d0; cfg%vcflit = 0.d0

  rc = cosahip_create(cfg, 0, s)
  if (rc /= COSAHIP_OK) stop 'cosahip_create failed'

  ! one block, farfield on all six faces (bctopo rows of setdim_bbc)
  imax = im; jmax = jm; kmax = km; nbcs = 6
  ext = (/im, jm, km/)
  do ib = 1, 6
    bctopo(:, ib) = 0
    bctopo(1, ib) = 30
    bctopo(2, ib) = (ib + 1) / 2
    bctopo(3, ib) = merge(1, ext((ib + 1) / 2), mod(ib, 2) == 1)
    bctopo(4:9, ib) = (/1, im - 1, 1, jm - 1, 1, km - 1/)
  end do
  cutopo = 0
  rc = cosahip_define_level(s, 1, 1, imax, jmax, kmax, nbcs, bctopo, 0, cutopo)
  if (rc /= COSAHIP_OK) then
    print *, trim(cosahip_last_error(s))
    stop 'define_level failed'
  end if

  allocate(q(-1:im+1, -1:jm+1, -1:km+1, npde), q2(-1:im+1, -1:jm+1, -1:km+1, npde))
  allocate(si(lmet, 0:im+1, 0:jm+1, 0:km+1), sj(lmet, 0:im+1, 0:jm+1, 0:km+1), sk(lmet, 0:im+1, 0:jm+1, 0:km+1))
  allocate(vol(0:im, 0:jm, 0:km))
  si = 0; sj = 0; sk = 0
  si(1, :, :, :) = 1.d0; si(4, :, :, :) = h * h
  sj(2, :, :, :) = 1.d0; sj(4, :, :, :) = h * h
  sk(3, :, :, :) = 1.d0; sk(4, :, :, :) = h * h
  vol = h * h * h
  q(:, :, :, 1) = 1.d0
  q(:, :, :, 2) = mach * cos(al)
  q(:, :, :, 3) = mach * sin(al)
  q(:, :, :, 4) = 0.d0
  q(:, :, :, 5) = 1.d0 / gam / (gam - 1.d0) + 0.5d0 * mach**2

  n = size(si, kind=c_int64_t)
  if (cosahip_upload(s, 1, COSAHIP_SI, si, n) /= 0) stop 'upload si'
  if (cosahip_upload(s, 1, COSAHIP_SJ, sj, n) /= 0) stop 'upload sj'
  if (cosahip_upload(s, 1, COSAHIP_SK, sk, n) /= 0) stop 'upload sk'
  if (cosahip_upload(s, 1, COSAHIP_VOL, vol, size(vol, kind=c_int64_t)) /= 0) stop 'upload vol'
  if (cosahip_upload(s, 1, COSAHIP_Q, q, size(q, kind=c_int64_t)) /= 0) stop 'upload q'

  rc = cosahip_smooth(s, 1, 0, 0, 1, 1, done, resrms)
  if (rc /= COSAHIP_OK) then
    print *, trim(cosahip_last_error(s))
    stop 'smooth failed'
  end if
  if (cosahip_download(s, 1, COSAHIP_Q, q2, size(q2, kind=c_int64_t)) /= 0) stop 'download q'
  rc = cosahip_destroy(s)
  print '(a, 6es12.4)', ' resrms(1:6) =', resrms(1:6)
  print '(a, es12.4)', ' max |q - q0| (interior) =', &
      maxval(abs(q2(1:im-1, 1:jm-1, 1:km-1, :) - q(1:im-1, 1:jm-1, 1:km-1, :)))
  if (maxval(resrms(1:6)) > 1.d-14) stop 1
  if (maxval(abs(q2(1:im-1, 1:jm-1, 1:km-1, :) - q(1:im-1, 1:jm-1, 1:km-1, :))) > 1.d-13) stop 2
  print *, 'binding_demo: ok'
end program binding_demo
