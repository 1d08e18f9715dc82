!-----------------------------------------------------------------------
! cosahip_binding.f90 -- iso_c_binding interface to libcosahip
! (include/cosahip.h) for COSA's Fortran driver.
!
! The driver keeps its input, grid, metrics, output and the mg() cycle
! (src/cosa.f90:806-930); with this module it creates one device solve
! after alloc/metrics, uploads the level arrays it already holds in the
! reference layout (Cray pointers p_q, p_si, ... of cosa_variables), and
! replaces the calls to smooth/restrict/prolong by cosahip_smooth /
! cosahip_restrict / cosahip_prolong.  Array arguments are the flat
! whole-level arrays (one copy each way).
!-----------------------------------------------------------------------
module cosahip_binding
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: COSAHIP_OK = 0, COSAHIP_ERR_POSITIVITY = 5
  integer(c_int), parameter :: COSAHIP_Q = 0, COSAHIP_QOLD = 1, COSAHIP_RES = 2, COSAHIP_RHS = 3
  integer(c_int), parameter :: COSAHIP_SI = 4, COSAHIP_SJ = 5, COSAHIP_SK = 6, COSAHIP_VOL = 7
  integer(c_int), parameter :: COSAHIP_DTVOL = 8, COSAHIP_XDOT = 12, COSAHIP_YDOT = 13, COSAHIP_ZDOT = 14
  integer(c_int), parameter :: COSAHIP_MUT = 15, COSAHIP_DIST = 16, COSAHIP_XIDERI = 17, COSAHIP_ETADERI = 18
  integer(c_int), parameter :: COSAHIP_ZETADERI = 19, COSAHIP_XIDERJ = 20, COSAHIP_ETADERJ = 21
  integer(c_int), parameter :: COSAHIP_ZETADERJ = 22, COSAHIP_XIDERK = 23, COSAHIP_ETADERK = 24
  integer(c_int), parameter :: COSAHIP_ZETADERK = 25, COSAHIP_PRD = 26, COSAHIP_DIVVEL = 27
  integer(c_int), parameter :: COSAHIP_DELPLS = 28, COSAHIP_DTVOLT = 29
  integer(c_int), parameter :: COSAHIP_X = 30, COSAHIP_Y = 31, COSAHIP_Z = 32, COSAHIP_TTRM = 33
  integer(c_int), parameter :: COSAHIP_Q1 = 34, COSAHIP_Q2 = 35

  ! cosahip_config, member for member
  type, bind(C) :: cosahip_config
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
  end type cosahip_config

  interface
    integer(c_int) function cosahip_create(cfg, device, solve) bind(C, name='cosahip_create')
      import :: c_int, c_ptr, cosahip_config
      type(cosahip_config), intent(in) :: cfg
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: solve
    end function
    integer(c_int) function cosahip_destroy(solve) bind(C, name='cosahip_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
    end function
    integer(c_int) function cosahip_error_info(solve, info) bind(C, name='cosahip_error_info')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: solve
      integer(c_int32_t), intent(out) :: info(8)
    end function
    integer(c_int) function cosahip_set_hb_operator(solve, nsnap, dhb) bind(C, name='cosahip_set_hb_operator')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solve
      integer(c_int), value :: nsnap
      real(c_double), intent(in) :: dhb(*)
    end function
    integer(c_int) function cosahip_define_level(solve, nl, nblocks, imax, jmax, kmax, nbcs, bctopo, ncuts, &
                                                 cutopo) bind(C, name='cosahip_define_level')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: solve
      integer(c_int), value :: nl, nblocks, ncuts
      integer(c_int32_t), intent(in) :: imax(*), jmax(*), kmax(*), nbcs(*), bctopo(*), cutopo(*)
    end function
    integer(c_int64_t) function cosahip_field_size(solve, nl, field) bind(C, name='cosahip_field_size')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: solve
      integer(c_int), value :: nl, field
    end function
    integer(c_int) function cosahip_upload(solve, nl, field, host, count) bind(C, name='cosahip_upload')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: solve
      integer(c_int), value :: nl, field
      real(c_double), intent(in) :: host(*)
      integer(c_int64_t), value :: count
    end function
    integer(c_int) function cosahip_download(solve, nl, field, host, count) bind(C, name='cosahip_download')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: solve
      integer(c_int), value :: nl, field
      real(c_double), intent(out) :: host(*)
      integer(c_int64_t), value :: count
    end function
    integer(c_int) function cosahip_smooth(solve, nl, from_fine, to_coarse, niter, mgit, done, resrms) &
        bind(C, name='cosahip_smooth')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: solve
      integer(c_int), value :: nl, from_fine, to_coarse, niter, mgit
      integer(c_int32_t), intent(out) :: done
      real(c_double), intent(out) :: resrms(8)
    end function
    ! dual time stepping: set up once, then begin / mg() / end per physical step
    integer(c_int) function cosahip_set_dual_time(solve, dt, dual_prop_start) bind(C, name='cosahip_set_dual_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solve
      real(c_double), value :: dt
      integer(c_int), value :: dual_prop_start
    end function
    integer(c_int) function cosahip_dual_begin(solve, itime) bind(C, name='cosahip_dual_begin')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
      integer(c_int), value :: itime
    end function
    integer(c_int) function cosahip_dual_end(solve) bind(C, name='cosahip_dual_end')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
    end function
    integer(c_int) function cosahip_src_uns(solve, nl) bind(C, name='cosahip_src_uns')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
      integer(c_int), value :: nl
    end function
    integer(c_int) function cosahip_restrict(solve, nl) bind(C, name='cosahip_restrict')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
      integer(c_int), value :: nl
    end function
    integer(c_int) function cosahip_prolong(solve, nl) bind(C, name='cosahip_prolong')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
      integer(c_int), value :: nl
    end function
    integer(c_int) function cosahip_force(solve, nl, nsurface, include_inviscid, lngth, momr, coef) &
        bind(C, name='cosahip_force')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solve
      integer(c_int), value :: nl, nsurface, include_inviscid
      real(c_double), intent(in) :: lngth(3), momr(*)
      real(c_double), intent(out) :: coef(*)
    end function
    integer(c_int) function cosahip_force_noalloc(solve, nl, nsurface, include_inviscid, lngth, momr, coef) &
        bind(C, name='cosahip_force_noalloc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solve
      integer(c_int), value :: nl, nsurface, include_inviscid
      real(c_double), intent(in) :: lngth(3), momr(*)
      real(c_double), intent(out) :: coef(*)
    end function
    integer(c_int) function cosahip_grid_metrics(solve, nl) bind(C, name='cosahip_grid_metrics')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
      integer(c_int), value :: nl
    end function
    integer(c_int) function cosahip_grid_coarsen(solve, nl) bind(C, name='cosahip_grid_coarsen')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
      integer(c_int), value :: nl
    end function
    integer(c_int) function cosahip_grid_pitch(solve, nl, dtheta0, omegas, xh, yh, pose) &
        bind(C, name='cosahip_grid_pitch')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solve
      integer(c_int), value :: nl
      real(c_double), value :: dtheta0, omegas, xh, yh
      real(c_double), intent(in) :: pose(*)
    end function
    integer(c_int) function cosahip_set_precision(solve, mode, tail_rms) bind(C, name='cosahip_set_precision')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solve
      integer(c_int), value :: mode
      real(c_double), value :: tail_rms
    end function
    integer(c_int) function cosahip_precision_in_use(solve) bind(C, name='cosahip_precision_in_use')
      import :: c_int, c_ptr
      type(c_ptr), value :: solve
    end function
    type(c_ptr) function cosahip_last_error_c(solve) bind(C, name='cosahip_last_error')
      import :: c_ptr
      type(c_ptr), value :: solve
    end function
  end interface

contains

  ! Fortran string of cosahip_last_error
  function cosahip_last_error(solve) result(msg)
    type(c_ptr), intent(in) :: solve
    character(len=256) :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: i
    msg = ' '
    call c_f_pointer(cosahip_last_error_c(solve), cs, [256])
    do i = 1, 256
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function

end module cosahip_binding
