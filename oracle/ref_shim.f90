!-----------------------------------------------------------------------
! ref_shim.f90 -- bind(C) entry points that drive COSA's OWN compiled
! routines (src/flux.f90, src/update.f90, src/src.f90, src/util.f90 and the
! self-contained subroutines of src/bc.f90 and src/cosa.f90), so tests can
! feed the reference and the C restatement (cosa_oracle.c) the same inputs.
!
! TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/
! (git-ignored) from the sources where they lie under /root/reference; no
! reference source is copied into this repository.
!-----------------------------------------------------------------------
module ref_shim
  use iso_c_binding
  use common_variables
  use cosa_variables
  use cosa_precision
  implicit none

  ! layout of cosahip_config (include/cosahip.h)
  type, bind(C) :: cfg_t
    integer(c_int32_t) :: npde, nharms, hbmove, lmet, nlevel, nl_crs, nstage
    integer(c_int32_t) :: irsop, cutcirs, ilimf, ilimt, iupdt, iprint, ncycle
    integer(c_int32_t) :: kom, kom_bsl, kom_sst, viscous
    integer(c_int32_t) :: moving, harbal, foturb
    integer(c_int32_t) :: efx_setup_f, efx_all_f, eig_limit_f, eig_lim_all_f
    integer(c_int32_t) :: efx_setup_t, efx_all_t, eig_limit_t, eig_lim_all_t
    integer(c_int32_t) :: entfx_f_w(3), entfx_t_w(3)
    integer(c_int32_t) :: bln_prol, ho_rest, rkimp
    real(c_double) :: gamma, machfs, alpha, beta
    real(c_double) :: rhoinf, pinf, ainf
    real(c_double) :: cfl, cflt, rat, ratt, psirs, rkap
    real(c_double) :: epslimf, epslimt
    real(c_double) :: cntrpy_f, cntrpy_t, entfxctf_f, entfxctf_t
    real(c_double) :: eig_cutoff_f, eig_cutoff_t
    real(c_double) :: omega, toler
    real(c_double) :: qmin(7), alphas(5)
    real(c_double) :: tkefar, omefar
    integer(c_int32_t) :: turcmp, lim_prodtke, lim_prodomega, wallwilc, ifixq, prd_all_lvls
    integer(c_int32_t) :: lim_corr(3), pad1
    real(c_double) :: reyno, pranl, prant, stemp, cdff, epst, cmu, sigk, sigw, bkw, bstrkw, gkw
    real(c_double) :: prdlim, roughk
    integer(c_int32_t) :: ramping, ncycr(4), pad2
    real(c_double) :: vcfli(3), vcflit(3)
  end type cfg_t

contains

  integer(c_int) function ref_cfg_size() bind(C, name='ref_cfg_size')
    type(cfg_t) :: c
    ref_cfg_size = int(c_sizeof(c), c_int)
  end function

  subroutine ref_set_config(c) bind(C, name='ref_set_config')
    type(cfg_t), intent(in) :: c
    cyc = reshape((/1,2,3,2,3,1,3,1,2/),(/3,3/))
    krd = reshape((/1,0,0,0,1,0,0,0,1/),(/3,3/))
    npde = c%npde; nharms = c%nharms; hbmove = c%hbmove; lmet = c%lmet
    dim5 = 2*nharms+1; dim5h = 2*nharms*hbmove+1
    nlevel = c%nlevel; nl_crs = c%nl_crs; nstage = c%nstage
    irsop = c%irsop; cutcirs = c%cutcirs; ilimf = c%ilimf; ilimt = c%ilimt
    iupdt = c%iupdt; iprint = c%iprint; ncycle = c%ncycle
    kom = c%kom /= 0; kom_bsl = c%kom_bsl /= 0; kom_sst = c%kom_sst /= 0
    viscous = c%viscous /= 0; moving = c%moving /= 0; harbal = c%harbal /= 0
    foturb = c%foturb /= 0
    efx_setup_f = c%efx_setup_f /= 0; efx_all_f = c%efx_all_f /= 0
    eig_limit_f = c%eig_limit_f /= 0; eig_lim_all_f = c%eig_lim_all_f /= 0
    efx_setup_t = c%efx_setup_t /= 0; efx_all_t = c%efx_all_t /= 0
    eig_limit_t = c%eig_limit_t /= 0; eig_lim_all_t = c%eig_lim_all_t /= 0
    entfx_f_w = c%entfx_f_w; entfx_t_w = c%entfx_t_w
    bln_prol = c%bln_prol /= 0; ho_rest = c%ho_rest /= 0
    gamma = c%gamma; machfs = c%machfs; alpha = c%alpha; beta = c%beta
    rhoinf = c%rhoinf; pinf = c%pinf; ainf = c%ainf
    cfl = c%cfl; cflt = c%cflt; rat = c%rat; ratt = c%ratt; psirs = c%psirs
    rkap = c%rkap; epslimf = c%epslimf; epslimt = c%epslimt
    cntrpy_f = c%cntrpy_f; cntrpy_t = c%cntrpy_t
    entfxctf_f = c%entfxctf_f; entfxctf_t = c%entfxctf_t
    eig_cutoff_f = c%eig_cutoff_f; eig_cutoff_t = c%eig_cutoff_t
    omega = c%omega; toler = c%toler; qmin = c%qmin
    alphas(1:5) = c%alphas; tkefar = c%tkefar; omefar = c%omefar
    ! options of the configurations exercised here
    aircraft = .true.; vawt = .false.; hawt = .false.; rotating = .false.
    tpitch = .false.; lomach = .false.; rkimp = .false.; rgkuns = .false.
    dualt = .false.; relframe = .false.; unsteady = harbal
    mgit = 1
    lim_corr = c%lim_corr /= 0
    turcmp = c%turcmp; lim_prodtke = c%lim_prodtke /= 0; lim_prodomega = c%lim_prodomega /= 0
    wallwilc = c%wallwilc /= 0; wallment = .not. wallwilc; ifixq = c%ifixq
    prd_all_lvls = c%prd_all_lvls /= 0
    reyno = c%reyno; pranl = c%pranl; prant = c%prant; stemp = c%stemp; cdff = c%cdff
    epst = c%epst; cmu = c%cmu; sigk = c%sigk; sigw = c%sigw; bkw = c%bkw; bstrkw = c%bstrkw
    gkw = c%gkw; prdlim = c%prdlim; roughk = c%roughk
    ramping1 = c%ramping == 1; ramping2 = c%ramping == 2
    ncycr(0:3) = c%ncycr; vcfli(1:3) = c%vcfli; vcflit(1:3) = c%vcflit
    ! BSL / SST constants as input sets them
    bstr = 0.09d0; a1 = 0.31d0
    sigk1_bsl = 0.5d0; sigk1_sst = 0.85d0; sigk2 = 1.0d0
    sigw1 = 0.5d0; b1 = 3.d0 / 40.d0; g1 = b1/bstr - sigw1*0.41d0**2/dsqrt(bstr)
    sigw2 = 0.856d0; b2 = 0.0828d0; g2 = b2/bstr - sigw2*0.41d0**2/dsqrt(bstr)
  end subroutine

  ! the multigrid cycle counter of common_variables (CFL ramping reads it)
  subroutine ref_set_mgit(m) bind(C, name='ref_set_mgit')
    integer(c_int), value :: m
    mgit = m
  end subroutine

  ! dual time stepping: dualt, rkimp, dt, dual_prop_start, itime
  subroutine ref_set_dual(on, rkim, dtp, prop, it) bind(C, name='ref_set_dual')
    integer(c_int), value :: on, rkim, prop, it
    real(c_double), value :: dtp
    dualt = on /= 0; rkimp = rkim /= 0; dt = dtp; dual_prop_start = prop /= 0; itime = it
    unsteady = unsteady .or. dualt
  end subroutine

  subroutine ref_src_uns_b(q, vol, res, imax, jmax, kmax) bind(C, name='ref_src_uns_b')
    real(c_double) :: q(*), vol(*), res(*)
    integer(c_int), value :: imax, jmax, kmax
    call src_uns_b(q, vol, res, imax, jmax, kmax, npde)
  end subroutine

  subroutine ref_rhs_uns_b(rhs, q1, q2, vol, imax, jmax, kmax) bind(C, name='ref_rhs_uns_b')
    real(c_double) :: rhs(*), q1(*), q2(*), vol(*)
    integer(c_int), value :: imax, jmax, kmax
    call rhs_uns_b(rhs, q1, q2, vol, imax, jmax, kmax, npde)
  end subroutine

  subroutine ref_set_dhb(nsnap, d) bind(C, name='ref_set_dhb')
    integer(c_int), value :: nsnap
    real(c_double), intent(in) :: d(nsnap, nsnap)
    dhb(0:nsnap-1, 0:nsnap-1) = d
  end subroutine

  subroutine ref_cons2prim_b(q, work, imax, jmax, kmax, xtnd) bind(C, name='ref_cons2prim_b')
    real(c_double) :: q(*), work(*)
    integer(c_int), value :: imax, jmax, kmax, xtnd
    call cons2prim_b(q, work, xtnd, npde, imax, jmax, kmax, nharms)
  end subroutine

  subroutine ref_prim2cons_b(q, work, imax, jmax, kmax) bind(C, name='ref_prim2cons_b')
    real(c_double) :: q(*), work(*)
    integer(c_int), value :: imax, jmax, kmax
    call prim2cons_b(q, work, npde, imax, jmax, kmax, nharms)
  end subroutine

  subroutine ref_bbc(q, si, sj, sk, dist, nl, bctopo, nbcs, imax, jmax, kmax) bind(C, name='ref_bbc')
    real(c_double) :: q(*), si(*), sj(*), sk(*), dist(*)
    integer(c_int) :: bctopo(*)
    integer(c_int), value :: nl, nbcs, imax, jmax, kmax
    integer(c_int) :: ijkmax
    ijkmax = max(imax, jmax, kmax)
    call bbc(q, si, sj, sk, dist, nl, bctopo, nbcs, imax, jmax, kmax, ijkmax, npde, nharms, lmet)
  end subroutine

  subroutine ref_lambdab(q, si, sj, sk, xdot, ydot, zdot, vol, mut, work, lambdas, cut1, cut2, &
                         imax, jmax, kmax) bind(C, name='ref_lambdab')
    real(c_double) :: q(*), si(*), sj(*), sk(*), xdot(*), ydot(*), zdot(*), vol(*), mut(*)
    real(c_double) :: work(*), lambdas(*), cut1(*), cut2(*)
    integer(c_int), value :: imax, jmax, kmax
    call lambdab(1, 1, q, si, sj, sk, xdot, ydot, zdot, vol, mut, work, lambdas, cut1, cut2, &
                 imax, jmax, kmax, npde, nharms, lmet)
  end subroutine

  subroutine ref_bdeltat(dtvol, dtvolt, lambdas, lambdasv, bx, be, bz, tbx, tbe, tbz, &
                         imax, jmax, kmax) bind(C, name='ref_bdeltat')
    real(c_double) :: dtvol(*), dtvolt(*), lambdas(*), lambdasv(*), bx(*), be(*), bz(*)
    real(c_double) :: tbx(*), tbe(*), tbz(*)
    integer(c_int), value :: imax, jmax, kmax
    call bdeltat(1, 1, dtvol, dtvolt, lambdas, lambdasv, bx, be, bz, tbx, tbe, tbz, &
                 imax, jmax, kmax, npde, nharms, lmet)
  end subroutine

  subroutine ref_muscl_b(idir, nl, bctopo, nbcs, q, qp, qm, dq, dqp, dqm, imax, jmax, kmax) &
      bind(C, name='ref_muscl_b')
    integer(c_int), value :: idir, nl, nbcs, imax, jmax, kmax
    integer(c_int) :: bctopo(*)
    real(c_double) :: q(*), qp(*), qm(*), dq(*), dqp(*), dqm(*)
    if (idir == 1) then
      call muscl_bi(nl, bctopo, q, qp, qm, dq, dqp, dqm, nbcs, imax, jmax, kmax, npde, nharms)
    else if (idir == 2) then
      call muscl_bj(nl, bctopo, q, qp, qm, dq, dqp, dqm, nbcs, imax, jmax, kmax, npde, nharms)
    else
      call muscl_bk(nl, bctopo, q, qp, qm, dq, dqp, dqm, nbcs, imax, jmax, kmax, npde, nharms)
    end if
  end subroutine

  subroutine ref_roflux_b(idir, qr, ql, alphar, work, flux, sijk, imax, jmax, kmax) &
      bind(C, name='ref_roflux_b')
    integer(c_int), value :: idir, imax, jmax, kmax
    real(c_double) :: qr(*), ql(*), alphar(*), work(*), flux(*), sijk(*)
    integer(c_int) :: ic, jc, kc
    ic = imax - 1; jc = jmax - 1; kc = kmax - 1
    if (idir == 1) ic = imax
    if (idir == 2) jc = jmax
    if (idir == 3) kc = kmax
    call roflux_b(qr, ql, alphar, work, flux, sijk, imax, jmax, kmax, npde, nharms, lmet, ic, jc, kc)
  end subroutine

  subroutine ref_bresid(idir, flux, res, imax, jmax, kmax) bind(C, name='ref_bresid')
    integer(c_int), value :: idir, imax, jmax, kmax
    real(c_double) :: flux(*), res(*)
    call bresid(flux, res, idir, imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_src_bhb(q, vol, res, qhb, imax, jmax, kmax) bind(C, name='ref_src_bhb')
    real(c_double) :: q(*), vol(*), res(*), qhb(*)
    integer(c_int), value :: imax, jmax, kmax
    call src_bhb(q, vol, res, qhb, imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_resg_brms(q, res, rhs, vol, rb, imax, jmax, kmax) bind(C, name='ref_resg_brms')
    real(c_double) :: q(*), res(*), rhs(*), vol(*), rb(*)
    integer(c_int), value :: imax, jmax, kmax
    call resg_brms(q, res, rhs, vol, rb, imax, jmax, kmax, npde)
  end subroutine

  subroutine ref_resg_brms_hb(q, work1, vol, rb, imax, jmax, kmax) bind(C, name='ref_resg_brms_hb')
    real(c_double) :: q(*), work1(*), vol(*), rb(*)
    integer(c_int), value :: imax, jmax, kmax
    call resg_brms_hb(q, work1, vol, rb, imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_cirs_b(res, dtvol, bx, be, bz, dtvolt, tbx, tbe, tbz, work, imax, jmax, kmax) &
      bind(C, name='ref_cirs_b')
    real(c_double) :: res(*), dtvol(*), bx(*), be(*), bz(*), dtvolt(*), tbx(*), tbe(*), tbz(*), work(*)
    integer(c_int), value :: imax, jmax, kmax
    call cirs_b(res, dtvol, bx, be, bz, dtvolt, tbx, tbe, tbz, work, imax, jmax, kmax, npde, nharms, &
                max(imax, jmax, kmax))
  end subroutine

  subroutine ref_ccirs_b(res, ores, bx, be, bz, tbx, tbe, tbz, imax, jmax, kmax) bind(C, name='ref_ccirs_b')
    real(c_double) :: res(*), ores(*), bx(*), be(*), bz(*), tbx(*), tbe(*), tbz(*)
    integer(c_int), value :: imax, jmax, kmax
    real(c_double), allocatable :: work(:)
    integer :: ijkmax
    ijkmax = max(imax, jmax, kmax)
    allocate(work((ijkmax - 1) * max(npde, 4)))
    call ccirs_b(res, ores, bx, be, bz, tbx, tbe, tbz, work, imax, jmax, kmax, npde, nharms, ijkmax)
    deallocate(work)
  end subroutine

  subroutine ref_scale_b(res, dtvol, dtvolt, back, imax, jmax, kmax) bind(C, name='ref_scale_b')
    real(c_double) :: res(*), dtvol(*), dtvolt(*)
    integer(c_int), value :: back, imax, jmax, kmax
    if (back == 0) then
      call scalef_b(res, dtvol, dtvolt, imax, jmax, kmax, npde, nharms)
    else
      call scaleb_b(res, dtvol, dtvolt, imax, jmax, kmax, npde, nharms)
    end if
  end subroutine

  ! cut_q (src/bc.f90:1894-2141), the branch where both blocks belong to
  ! this process (lowernblock..uppernblock cover them): q1's two halo
  ! layers from q2's two interior layers; the message arrays of the other
  ! branches are sized but unused
  subroutine ref_cut_q(imax1, jmax1, kmax1, q1, imax2, jmax2, kmax2, q2, cutopo) bind(C, name='ref_cut_q')
    integer(c_int32_t), value :: imax1, jmax1, kmax1, imax2, jmax2, kmax2
    real(c_double) :: q1(*), q2(*)
    integer(c_int32_t) :: cutopo(21)
    integer(kind=cosa_int) :: rreq(1), sreq(1), nrecv(1), nsend(1), myr(0:0), mys(0:0)
    integer(kind=cosa_int), allocatable :: rind(:,:,:)
    real(kind=cosa_real), allocatable :: sarr(:,:,:), rarr(:,:,:)
    allocate(rind(1,8,1), sarr(npde*(2*nharms+1),1,1), rarr(npde*(2*nharms+1),1,1))
    lowernblock = 1
    uppernblock = max(cutopo(1), cutopo(10))
    nrecv = 0; nsend = 0; myr = 1; mys = 1
    call cut_q(imax1, jmax1, kmax1, q1, imax2, jmax2, kmax2, q2, npde, nharms, cutopo, rarr, rreq, nrecv, &
         rind, sarr, sreq, nsend, 1, 1, 1, 1, myr, mys)
    deallocate(rind, sarr, rarr)
  end subroutine

  subroutine ref_tridi(a, b, c, x, dimx, n) bind(C, name='ref_tridi')
    real(c_double) :: a(*), b(*), c(*), x(*)
    integer(c_int), value :: dimx, n
    call tridi(a, b, c, x, dimx, n)
  end subroutine

  subroutine ref_update(res, q, qold, vol, dtvol, dtvolt, work1, qhb0, delpls, prd, ttrm, ar, &
                        imax, jmax, kmax) bind(C, name='ref_update')
    real(c_double) :: res(*), q(*), qold(*), vol(*), dtvol(*), dtvolt(*), work1(*), qhb0(*)
    real(c_double) :: delpls(*), prd(*), ttrm(*)
    real(c_double), value :: ar
    integer(c_int), value :: imax, jmax, kmax
    if (harbal) then
      call update_bhb(res, q, qold, vol, dtvol, dtvolt, work1, qhb0, delpls, prd, ttrm, ar, &
                      imax, jmax, kmax, npde, nharms)
    else
      call update_b(res, q, qold, vol, dtvol, dtvolt, delpls, prd, ttrm, ar, imax, jmax, kmax, npde, nharms)
    end if
  end subroutine

  subroutine ref_rest1_b(i1, j1, k1, i2, j2, k2, q1, q2, vol1, vol2) bind(C, name='ref_rest1_b')
    integer(c_int), value :: i1, j1, k1, i2, j2, k2
    real(c_double) :: q1(*), q2(*), vol1(*), vol2(*)
    call rest1_b(i1, j1, k1, i2, j2, k2, npde, nharms, q1, q2, vol1, vol2)
  end subroutine

  subroutine ref_rest2_b(ho, i1, j1, k1, i2, j2, k2, rhs1, rhs2, res1) bind(C, name='ref_rest2_b')
    integer(c_int), value :: ho, i1, j1, k1, i2, j2, k2
    real(c_double) :: rhs1(*), rhs2(*), res1(*)
    if (ho /= 0) then
      call rest2_ho_b(i1, j1, k1, i2, j2, k2, npde, nharms, rhs1, rhs2, res1)
    else
      call rest2_lo_b(i1, j1, k1, i2, j2, k2, npde, nharms, rhs1, rhs2, res1)
    end if
  end subroutine

  subroutine ref_prol_b(bln, i1, j1, k1, i2, j2, k2, q1, q2, vol1) bind(C, name='ref_prol_b')
    integer(c_int), value :: bln, i1, j1, k1, i2, j2, k2
    real(c_double) :: q1(*), q2(*), vol1(*)
    logical :: lc(3)
    lc = .false.
    if (bln /= 0) then
      call prol_bln_b(i1, j1, k1, i2, j2, k2, npde, nharms, q1, q2, vol1, lc)
    else
      call prol_lin_b(i1, j1, k1, i2, j2, k2, npde, nharms, q1, q2, vol1)
    end if
  end subroutine

  subroutine ref_q_bedges(q, mut, imax, jmax, kmax) bind(C, name='ref_q_bedges')
    real(c_double) :: q(*), mut(*)
    integer(c_int), value :: imax, jmax, kmax
    calmut = .false.
    call q_bedges(q, mut, imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_q_bcorns(q, mut, imax, jmax, kmax) bind(C, name='ref_q_bcorns')
    real(c_double) :: q(*), mut(*)
    integer(c_int), value :: imax, jmax, kmax
    calmut = .false.
    call q_bcorns(q, mut, imax, jmax, kmax, npde, nharms)
  end subroutine

  ! ---- viscous fluxes and Wilcox k-omega (src/flux.f90, src/src.f90, src/update.f90)

  subroutine ref_q_face_b(idir, bctopo, nbcs, q, qf, mut, mutf, cm, imax, jmax, kmax) &
      bind(C, name='ref_q_face_b')
    integer(c_int), value :: idir, nbcs, cm, imax, jmax, kmax
    integer(c_int) :: bctopo(*)
    real(c_double) :: q(*), qf(*), mut(*), mutf(*)
    calmut = cm /= 0
    if (idir == 1) then
      call bq_face_i(bctopo, nbcs, q, qf, mut, mutf, imax, jmax, kmax, npde, nharms)
    else if (idir == 2) then
      call bq_face_j(bctopo, nbcs, q, qf, mut, mutf, imax, jmax, kmax, npde, nharms)
    else
      call bq_face_k(bctopo, nbcs, q, qf, mut, mutf, imax, jmax, kmax, npde, nharms)
    end if
    calmut = .false.
  end subroutine

  subroutine ref_q_der_b(idir, bctopo, nbcs, q, dqxi, dqeta, dqzeta, imax, jmax, kmax) &
      bind(C, name='ref_q_der_b')
    integer(c_int), value :: idir, nbcs, imax, jmax, kmax
    integer(c_int) :: bctopo(*)
    real(c_double) :: q(*), dqxi(*), dqeta(*), dqzeta(*)
    if (idir == 1) then
      call bq_der_i(bctopo, nbcs, q, dqxi, dqeta, dqzeta, imax, jmax, kmax, npde, nharms)
    else if (idir == 2) then
      call bq_der_j(bctopo, nbcs, q, dqxi, dqeta, dqzeta, imax, jmax, kmax, npde, nharms)
    else
      call bq_der_k(bctopo, nbcs, q, dqxi, dqeta, dqzeta, imax, jmax, kmax, npde, nharms)
    end if
  end subroutine

  subroutine ref_rtst_b(idir, qf, dqxi, dqeta, dqzeta, xider, etader, zetader, tau1, mutf, &
                        imax, jmax, kmax) bind(C, name='ref_rtst_b')
    integer(c_int), value :: idir, imax, jmax, kmax
    real(c_double) :: qf(*), dqxi(*), dqeta(*), dqzeta(*), xider(*), etader(*), zetader(*)
    real(c_double) :: tau1(*), mutf(*)
    integer(c_int) :: ic, jc, kc
    ic = imax - 1; jc = jmax - 1; kc = kmax - 1
    if (idir == 1) ic = imax
    if (idir == 2) jc = jmax
    if (idir == 3) kc = kmax
    call b_rtst(qf, dqxi, dqeta, dqzeta, xider, etader, zetader, tau1, mutf, imax, jmax, kmax, &
                npde, nharms, ic, jc, kc)
  end subroutine

  subroutine ref_vflux_b(idir, qf, dqxi, dqeta, dqzeta, tau1, flux, fdist, xider, etader, zetader, &
                         sijk, imax, jmax, kmax) bind(C, name='ref_vflux_b')
    integer(c_int), value :: idir, imax, jmax, kmax
    real(c_double) :: qf(*), dqxi(*), dqeta(*), dqzeta(*), tau1(*), flux(*), fdist(*)
    real(c_double) :: xider(*), etader(*), zetader(*), sijk(*)
    integer(c_int) :: ic, jc, kc
    ic = imax - 1; jc = jmax - 1; kc = kmax - 1
    if (idir == 1) ic = imax
    if (idir == 2) jc = jmax
    if (idir == 3) kc = kmax
    call vflux_b(qf, dqxi, dqeta, dqzeta, tau1, flux, fdist, xider, etader, zetader, sijk, &
                 imax, jmax, kmax, npde, nharms, lmet, ic, jc, kc)
  end subroutine

  subroutine ref_q_der_c_b(idir, qf, dqc, imax, jmax, kmax) bind(C, name='ref_q_der_c_b')
    integer(c_int), value :: idir, imax, jmax, kmax
    real(c_double) :: qf(*), dqc(*)
    if (idir == 1) then
      call bq_der_ci(qf, dqc, imax, jmax, kmax, npde, nharms)
    else if (idir == 2) then
      call bq_der_cj(qf, dqc, imax, jmax, kmax, npde, nharms)
    else
      call bq_der_ck(qf, dqc, imax, jmax, kmax, npde, nharms)
    end if
  end subroutine

  subroutine ref_turprod_b(q, dqxi, dqeta, dqzeta, prd, divvel, delpls, xdi, edi, zdi, xdj, edj, zdj, &
                           xdk, edk, zdk, imax, jmax, kmax) bind(C, name='ref_turprod_b')
    real(c_double) :: q(*), dqxi(*), dqeta(*), dqzeta(*), prd(*), divvel(*), delpls(*)
    real(c_double) :: xdi(*), edi(*), zdi(*), xdj(*), edj(*), zdj(*), xdk(*), edk(*), zdk(*)
    integer(c_int), value :: imax, jmax, kmax
    call turprod_b(q, dqxi, dqeta, dqzeta, prd, divvel, delpls, xdi, edi, zdi, xdj, edj, zdj, &
                   xdk, edk, zdk, imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_srckw_b(q, mut, prd, divvel, res, vol, nl, imax, jmax, kmax) bind(C, name='ref_srckw_b')
    real(c_double) :: q(*), mut(*), prd(*), divvel(*), res(*), vol(*)
    integer(c_int), value :: nl, imax, jmax, kmax
    call srckw_b(q, mut, prd, divvel, res, vol, nl, imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_lim_prod_mg_b(q, prd, mut, imax, jmax, kmax) bind(C, name='ref_lim_prod_mg_b')
    real(c_double) :: q(*), prd(*), mut(*)
    integer(c_int), value :: imax, jmax, kmax
    call lim_prod_mg_b(q, prd, mut, imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_fixq_b(q, imax, jmax, kmax) bind(C, name='ref_fixq_b')
    real(c_double) :: q(*)
    integer(c_int), value :: imax, jmax, kmax
    call fixq_b(q, imax, jmax, kmax, npde, nharms, 1, 1)
  end subroutine

  subroutine ref_tur_vis_b(q, mut, imax, jmax, kmax) bind(C, name='ref_tur_vis_b')
    real(c_double) :: q(*), mut(*)
    integer(c_int), value :: imax, jmax, kmax
    call tur_vis_b(q, mut, npde, imax, jmax, kmax, nharms)
  end subroutine

  subroutine ref_update_ttrm_b(q, dqxi, dqeta, dqzeta, xdi, edi, zdi, xdj, edj, zdj, xdk, edk, zdk, &
                               ttrm, dist, imax, jmax, kmax) bind(C, name='ref_update_ttrm_b')
    real(c_double) :: q(*), dqxi(*), dqeta(*), dqzeta(*), ttrm(*), dist(*)
    real(c_double) :: xdi(*), edi(*), zdi(*), xdj(*), edj(*), zdj(*), xdk(*), edk(*), zdk(*)
    integer(c_int), value :: imax, jmax, kmax
    call update_ttrm_b(q, dqxi, dqeta, dqzeta, xdi, edi, zdi, xdj, edj, zdj, xdk, edk, zdk, ttrm, dist, &
                       npde, imax, jmax, kmax, nharms)
  end subroutine

  subroutine ref_lim_res_ome_b(res, q, vol, dtvolt, delpls, prd, ttrm, ar, imax, jmax, kmax) &
      bind(C, name='ref_lim_res_ome_b')
    real(c_double) :: res(*), q(*), vol(*), dtvolt(*), delpls(*), prd(*), ttrm(*)
    real(c_double), value :: ar
    integer(c_int), value :: imax, jmax, kmax
    real(c_double) :: dummy(1)
    call lim_res_ome_b(res, q, vol, dtvolt, dummy, dummy, dummy, delpls, prd, ttrm, dummy, ar, &
                       imax, jmax, kmax, npde)
  end subroutine

  subroutine ref_srckw_sst_b(q, prd, divvel, res, mut, ttrm, vol, nl, imax, jmax, kmax) &
      bind(C, name='ref_srckw_sst_b')
    real(c_double) :: q(*), prd(*), divvel(*), res(*), mut(*), ttrm(*), vol(*)
    integer(c_int), value :: nl, imax, jmax, kmax
    call srckw_sst_b(q, prd, divvel, res, mut, ttrm, vol, nl, imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_tur_vis_sst_b(q, mut, dqxi, dqeta, dqzeta, xdi, edi, zdi, xdj, edj, zdj, xdk, edk, zdk, &
                               dist, imax, jmax, kmax) bind(C, name='ref_tur_vis_sst_b')
    real(c_double) :: q(*), mut(*), dqxi(*), dqeta(*), dqzeta(*), dist(*)
    real(c_double) :: xdi(*), edi(*), zdi(*), xdj(*), edj(*), zdj(*), xdk(*), edk(*), zdk(*)
    integer(c_int), value :: imax, jmax, kmax
    call tur_vis_sst_b(q, mut, dqxi, dqeta, dqzeta, xdi, edi, zdi, xdj, edj, zdj, xdk, edk, zdk, dist, &
                       npde, imax, jmax, kmax, nharms)
  end subroutine

  ! explicit branch only: rkimp stays .false. here (its branch calls LAPACK,
  ! which is not linked into this library)
  subroutine ref_lim_res_ome_bhb(res, q, qold, vol, dtvolt, delpls, prd, ttrm, imax, jmax, kmax) &
      bind(C, name='ref_lim_res_ome_bhb')
    real(c_double) :: res(*), q(*), qold(*), vol(*), dtvolt(*), delpls(*), prd(*), ttrm(*)
    integer(c_int), value :: imax, jmax, kmax
    real(c_double) :: dummy(1)
    call lim_res_ome_bhb(res, q, qold, vol, dtvolt, dummy, dummy, dummy, delpls, prd, ttrm, dummy, &
                         imax, jmax, kmax, npde, nharms)
  end subroutine

  subroutine ref_rest4_b(i1, j1, k1, i2, j2, k2, prd1, prd2, vol1, vol2) bind(C, name='ref_rest4_b')
    integer(c_int), value :: i1, j1, k1, i2, j2, k2
    real(c_double) :: prd1(*), prd2(*), vol1(*), vol2(*)
    call rest4_b(i1, j1, k1, i2, j2, k2, nharms, prd1, prd2, vol1, vol2)
  end subroutine

  ! force_b (src/output.f90:4388-4886) for aircraft; pitching about (xh, yh)
  ! with angles dth(0:2N) when ipitch /= 0.  coef(12, 0:2N, nsurf).
  subroutine ref_force_b(q, x, y, z, xdot, ydot, zdot, si, sj, sk, xideri, etaderj, zetaderk, nbcs, bctopo, &
                         nsurf, lng, xyzmom, ipitch, xhh, yhh, dth, coef, imax, jmax, kmax) bind(C, name='ref_force_b')
    real(c_double) :: q(*), x(*), y(*), z(*), xdot(*), ydot(*), zdot(*), si(*), sj(*), sk(*)
    real(c_double) :: xideri(*), etaderj(*), zetaderk(*)
    integer(c_int) :: bctopo(*)
    integer(c_int), value :: nbcs, nsurf, ipitch, imax, jmax, kmax
    real(c_double), value :: xhh, yhh
    real(c_double) :: lng(3), xyzmom(3), dth(*)
    real(c_double) :: coef(12, 0:2*nharms, nsurf)
    real(c_double), allocatable :: c(:, :, :)
    integer :: m, n, is
    allocate(c(0:2*nharms, nsurf, 12))
    nsurface = nsurf
    lngth = lng
    xmom = xyzmom(1); ymom = xyzmom(2); zmom = xyzmom(3)
    pitching = ipitch /= 0; plunging = .false.; plupitching = .false.; rotating = .false.
    xh = xhh; yh = yhh
    dtheta(0:2*nharms) = dth(1:2*nharms+1)
    call force_b(q, x, y, z, xdot, ydot, zdot, si, sj, sk, xideri, etaderj, zetaderk, nbcs, bctopo, &
                 c(:, :, 1), c(:, :, 2), c(:, :, 3), c(:, :, 4), c(:, :, 5), c(:, :, 6), &
                 c(:, :, 7), c(:, :, 8), c(:, :, 9), c(:, :, 10), c(:, :, 11), c(:, :, 12), &
                 imax, jmax, kmax, npde, nharms, lmet)
    do is = 1, nsurf
      do n = 0, 2*nharms
        do m = 1, 12
          coef(m, n, is) = c(n, is, m)
        end do
      end do
    end do
    deallocate(c)
  end subroutine

  ! writerest / readrest (src/output.f90:3481-3529, src/input.f90:4620-4663)
  ! over nblocks blocks stored one after another (q: npde*(2N+1) P3 arrays per
  ! block, mut: 2N+1), calling write_brest / read_brest per block
  subroutine ref_restart(mode, path, plen, q, mut, nblocks, imax, jmax, kmax) bind(C, name='ref_restart')
    integer(c_int), value :: mode, plen, nblocks
    character(kind=c_char) :: path(plen)
    real(c_double) :: q(*), mut(*)
    integer(c_int) :: imax(nblocks), jmax(nblocks), kmax(nblocks)
    character(len=plen) :: fname
    real(c_double) :: q1(1)
    integer :: i, ib, iq, imt, n3, fid
    do i = 1, plen
      fname(i:i) = path(i)
    end do
    dualt = .false.; dual_prop_start = .false.
    fid = 61
    if (mode == 0) then
      open(fid, file=fname, form='unformatted', status='replace')
    else
      open(fid, file=fname, form='unformatted', status='old')
    end if
    iq = 1
    imt = 1
    do ib = 1, nblocks
      if (mode == 0) then
        call write_brest(q(iq), q1, mut(imt), imax(ib), jmax(ib), kmax(ib), npde, nharms, fid)
      else
        call read_brest(q(iq), q1, mut(imt), imax(ib), jmax(ib), kmax(ib), npde, nharms, fid)
      end if
      n3 = (imax(ib) + 3) * (jmax(ib) + 3) * (kmax(ib) + 3)
      iq = iq + n3 * npde * dim5
      imt = imt + n3 * dim5
    end do
    close(fid)
  end subroutine

  ! ---- grid motion and metric set-up (src/grid.f90; coarsen_b of src/cosa.f90)

  subroutine ref_imetrix_b(x, y, z, xdot, ydot, zdot, si, sj, sk, imax, jmax, kmax) bind(C, name='ref_imetrix_b')
    real(c_double) :: x(*), y(*), z(*), xdot(*), ydot(*), zdot(*), si(*), sj(*), sk(*)
    integer(c_int32_t), value :: imax, jmax, kmax
    integer(kind=cosa_int) :: ib, im, jm, km, lm, nh
    dgn_cell = .false.
    ib = 1; im = imax; jm = jmax; km = kmax; lm = lmet; nh = nharms
    call imetrix_b(x, y, z, xdot, ydot, zdot, si, sj, sk, ib, im, jm, km, lm, nh)
  end subroutine

  subroutine ref_bvols(x, y, z, vol, imax, jmax, kmax) bind(C, name='ref_bvols')
    real(c_double) :: x(*), y(*), z(*), vol(*)
    integer(c_int32_t), value :: imax, jmax, kmax
    integer(kind=cosa_int) :: ib, im, jm, km, nh, lnb
    real(kind=cosa_real), allocatable :: rad(:)
    debug = .false.
    ib = 1; im = imax; jm = jmax; km = kmax; nh = nharms; lnb = 1
    allocate(rad(max(1, (imax - 1) * (jmax - 1) * (kmax - 1) * dim5h)))
    call bvols(ib, x, y, z, x, y, z, vol, rad, im, jm, km, nh, lnb)
    deallocate(rad)
  end subroutine

  subroutine ref_volex_b(vol, imax, jmax, kmax) bind(C, name='ref_volex_b')
    real(c_double) :: vol(*)
    integer(c_int32_t), value :: imax, jmax, kmax
    integer(kind=cosa_int) :: im, jm, km
    im = imax; jm = jmax; km = kmax
    call volex_b(vol, im, jm, km)
  end subroutine

  subroutine ref_vol_bedges(vol, imax, jmax, kmax) bind(C, name='ref_vol_bedges')
    real(c_double) :: vol(*)
    integer(c_int32_t), value :: imax, jmax, kmax
    integer(kind=cosa_int) :: im, jm, km
    im = imax; jm = jmax; km = kmax
    call vol_bedges(vol, im, jm, km)
  end subroutine

  subroutine ref_cordex_b(x, y, z, imax, jmax, kmax) bind(C, name='ref_cordex_b')
    real(c_double) :: x(*), y(*), z(*)
    integer(c_int32_t), value :: imax, jmax, kmax
    integer(kind=cosa_int) :: im, jm, km, nh
    im = imax; jm = jmax; km = kmax; nh = nharms
    call cordex_b(x, y, z, im, jm, km, nh)
  end subroutine

  ! d(1:9): xideri, etaderi, zetaderi, xiderj, etaderj, zetaderj, xiderk,
  ! etaderk, zetaderk (vmetrix_b takes them grouped by face family first)
  subroutine ref_vmetrix_b(x, y, z, d1, d2, d3, d4, d5, d6, d7, d8, d9, vol, imax, jmax, kmax) &
      bind(C, name='ref_vmetrix_b')
    real(c_double) :: x(*), y(*), z(*), d1(*), d2(*), d3(*), d4(*), d5(*), d6(*), d7(*), d8(*), d9(*), vol(*)
    integer(c_int32_t), value :: imax, jmax, kmax
    integer(kind=cosa_int) :: ib, im, jm, km, nh
    ib = 1; im = imax; jm = jmax; km = kmax; nh = nharms
    call vmetrix_b(x, y, z, d1, d4, d7, d2, d5, d8, d3, d6, d9, vol, ib, im, jm, km, nh)
  end subroutine

  subroutine ref_coarsen_b(x, y, z, imax, jmax, kmax, x2, y2, z2, imax2, jmax2, kmax2) bind(C, name='ref_coarsen_b')
    real(c_double) :: x(*), y(*), z(*), x2(*), y2(*), z2(*)
    integer(c_int32_t), value :: imax, jmax, kmax, imax2, jmax2, kmax2
    integer(kind=cosa_int) :: im, jm, km, im2, jm2, km2
    real(kind=cosa_real) :: dd(1)
    im = imax; jm = jmax; km = kmax; im2 = imax2; jm2 = jmax2; km2 = kmax2
    pitching = .false.
    call coarsen_b(x, y, z, dd, dd, dd, im, jm, km, x2, y2, z2, dd, dd, dd, im2, jm2, km2)
  end subroutine

  subroutine ref_pitch(zeit, dtheta, dtheta0, omegas, xh, yh, x0, y0, z0, dx, dy, dz, x, y, z, xdot, ydot, zdot, &
                       imax, jmax, kmax) bind(C, name='ref_pitch')
    real(c_double), value :: zeit, dtheta, dtheta0, omegas, xh, yh
    real(c_double) :: x0(*), y0(*), z0(*), dx(*), dy(*), dz(*), x(*), y(*), z(*), xdot(*), ydot(*), zdot(*)
    integer(c_int32_t), value :: imax, jmax, kmax
    real(kind=cosa_real) :: a1, a2, a3, a4, a5, a6
    integer(kind=cosa_int) :: im, jm, km
    a1 = zeit; a2 = dtheta; a3 = dtheta0; a4 = omegas; a5 = xh; a6 = yh
    im = imax; jm = jmax; km = kmax
    call pitch(a1, a2, a3, a4, a5, a6, x0, y0, z0, dx, dy, dz, x, y, z, xdot, ydot, zdot, im, jm, km)
  end subroutine

  ! cut_v / cut_x (src/grid.f90:612-838, 242-524), the branch where both
  ! blocks belong to this process, as ref_cut_q
  subroutine ref_cut_v(imax1, jmax1, kmax1, v1, imax2, jmax2, kmax2, v2, cutopo) bind(C, name='ref_cut_v')
    integer(c_int32_t), value :: imax1, jmax1, kmax1, imax2, jmax2, kmax2
    real(c_double) :: v1(*), v2(*)
    integer(c_int32_t) :: cutopo(21)
    integer(kind=cosa_int) :: rreq(1), sreq(1), nrecv(1), nsend(1), myr(0:0), mys(0:0), rind(1, 8, 1)
    real(kind=cosa_real) :: sarr(1, 1), rarr(1, 1)
    integer(kind=cosa_int) :: i1, j1, k1, i2, j2, k2
    lowernblock = 1
    uppernblock = max(cutopo(1), cutopo(10))
    nrecv = 0; nsend = 0; myr = 1; mys = 1
    i1 = imax1; j1 = jmax1; k1 = kmax1; i2 = imax2; j2 = jmax2; k2 = kmax2
    call cut_v(i1, j1, k1, v1, i2, j2, k2, v2, cutopo, rarr, rreq, nrecv, rind, sarr, sreq, nsend, 1, 1, 1, 1, &
               myr, mys)
  end subroutine

  subroutine ref_cut_x(imax1, jmax1, kmax1, x1, y1, z1, imax2, jmax2, kmax2, x2, y2, z2, cutopo) &
      bind(C, name='ref_cut_x')
    integer(c_int32_t), value :: imax1, jmax1, kmax1, imax2, jmax2, kmax2
    real(c_double) :: x1(*), y1(*), z1(*), x2(*), y2(*), z2(*)
    integer(c_int32_t) :: cutopo(21)
    integer(kind=cosa_int) :: rreq(1), sreq(1), nrecv(1), nsend(1), myr(0:0), mys(0:0), rind(1, 8, 1)
    real(kind=cosa_real), allocatable :: sarr(:,:,:), rarr(:,:,:)
    integer(kind=cosa_int) :: i1, j1, k1, i2, j2, k2, nh, d5h
    allocate(sarr(dim5h, 1, 1), rarr(dim5h, 1, 1))
    lowernblock = 1
    uppernblock = max(cutopo(1), cutopo(10))
    nrecv = 0; nsend = 0; myr = 1; mys = 1
    i1 = imax1; j1 = jmax1; k1 = kmax1; i2 = imax2; j2 = jmax2; k2 = kmax2; nh = nharms; d5h = dim5h
    call cut_x(i1, j1, k1, x1, y1, z1, i2, j2, k2, x2, y2, z2, nh, d5h, cutopo, rarr, rreq, nrecv, rind, sarr, &
               sreq, nsend, 1, 1, 1, 1, myr, mys)
    deallocate(sarr, rarr)
  end subroutine

end module ref_shim
