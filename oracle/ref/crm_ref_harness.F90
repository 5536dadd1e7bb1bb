! crm_ref_harness.F90 -- test harness around the reference CRM phase routines.
!
! ORACLE / TEST INFRASTRUCTURE ONLY.  Built by oracle/ref/build_ref.sh together with the
! reference's own phase sources (compiled unmodified from the reference tree) into
! oracle/_ref/libcrmref_<class>.so.  Only tests, smoke() and bench.py's CPU baseline use it.
!
! The reference's driver, crm() in components/cam/src/physics/crm/crm_module.F90, cannot be
! compiled here: it uses CAM modules (ppgrid, phys_grid, time_manager, cam_abortutils,
! perf_mod, and through accelerate_crm.F90 phys_control, cam_logfile, spmd_utils) whose own
! sources need MPI/ESMF/PIO.  This file therefore drives the reference's phase subroutines
! (abcoefs, zero, buoyancy, forcing, damping, ice_fall, boundaries, sgs_proc, advect_mom,
! sgs_mom, pressure, adams, advect_all_scalars, uvw, sgs_scalars, micro_proc, diagnose,
! kurant, precip_init, micro_init, sgs_init, setperturb) in the order crm() calls them
! (crm_module.F90:728-888), and restates crm()'s set-up (crm_module.F90:286-713) and
! end-of-call output code (crm_module.F90:896-1659) for the 1-moment microphysics without
! ECPP/CLUBB/MAML/SP_ESMT/mean-state acceleration.
!
! Interface: plain bind(C) structs of pointers, identical to include/crm_hip.h.
module crm_ref_harness
  use iso_c_binding
  implicit none

  type, bind(C) :: crm_dims_t
    integer(c_int) :: ncrms, ld, plev, nx, ny, nz, nx_rad, ny_rad
    real(c_double) :: dx, dt
  end type

  type, bind(C) :: crm_column_info_t
    integer(c_int) :: lchnk
    type(c_ptr)    :: gcol, lat, lon
    integer(c_int) :: igstep
  end type

  type, bind(C) :: crm_input_view_t
    type(c_ptr) :: zmid, zint, tl, ql, qccl, qiil, ps, pmid, pint, pdel, phis, &
                   ul, vl, ocnfrac, tau00, wndls, bflxls, fluxu00, fluxv00, fluxt00, fluxq00
  end type

  type, bind(C) :: crm_state_view_t
    type(c_ptr) :: u_wind, v_wind, w_wind, temperature, qt, qp, qn
  end type

  type, bind(C) :: crm_rad_view_t
    type(c_ptr) :: qrad, temperature, qv, qc, qi, cld
  end type

  type, bind(C) :: crm_output_view_t
    type(c_ptr) :: qcl, qci, qpl, qpi, tk, tkh
    type(c_ptr) :: prec_crm
    type(c_ptr) :: cltot, clhgh, clmed, cllow
    type(c_ptr) :: precc, precl, precsc, precsl
    type(c_ptr) :: jt_crm, mx_crm, prectend, precstend, taux, tauy, z0m, timing_factor
    type(c_ptr) :: cldtop, qc_mean, qi_mean, qs_mean, qg_mean, qr_mean
    type(c_ptr) :: sltend, qltend, qcltend, qiltend
    type(c_ptr) :: cld, gicewp, gliqwp, mctot, mcup, mcdn, mcuup, mcudn
    type(c_ptr) :: mu_crm, md_crm, du_crm, eu_crm, ed_crm
    type(c_ptr) :: flux_qt, fluxsgs_qt, tkez, tkesgsz, tkz, flux_u, flux_v, flux_qp
    type(c_ptr) :: precflux, qt_ls, qt_trans, qp_trans, qp_fall, qp_src, qp_evp, t_ls
  end type

  logical, save :: live = .false.   ! module arrays of the reference still allocated
  integer, save :: live_ncrms = 0

contains

  ! Sizes this library was compiled for (the reference fixes them with CPP constants).
  subroutine ref_compiled_dims(d) bind(C, name="ref_compiled_dims")
    use grid, only: nx, ny, nzm
    use crmdims, only: crm_nx_rad, crm_ny_rad, crm_dx, crm_dt
    type(crm_dims_t), intent(out) :: d
    d%ncrms = 0; d%ld = 0; d%plev = 0
    d%nx = nx; d%ny = ny; d%nz = nzm
    d%nx_rad = crm_nx_rad; d%ny_rad = crm_ny_rad
    d%dx = crm_dx; d%dt = crm_dt
  end subroutine

  subroutine free_reference_state()
    use params, only: deallocate_params
    use vars, only: deallocate_vars
    use grid, only: deallocate_grid
    use crmtracers, only: deallocate_tracers
    use sgs, only: deallocate_sgs
    use microphysics, only: deallocate_micro
    use micro_params, only: deallocate_micro_params
    if (.not. live) return
    call deallocate_params()
    call deallocate_grid()
    call deallocate_tracers()
    call deallocate_sgs()
    call deallocate_vars()
    call deallocate_micro()
    call deallocate_micro_params()
    live = .false.
  end subroutine

  subroutine ref_release() bind(C, name="ref_release")
    call free_reference_state()
  end subroutine

  ! The reference's Coriolis switch (params.F90:107, docoriolis, off by default) for later calls.
  subroutine ref_set_coriolis(flag) bind(C, name="ref_set_coriolis")
    use params, only: docoriolis
    integer(c_int), value :: flag
    docoriolis = flag /= 0
  end subroutine

  ! Returns 0 on success, CRM_ERR_* otherwise.  phase_stop > 0 returns right after that phase
  ! of the last substep (module state left allocated for ref_get); 0 runs the whole call.
  function ref_crm(d, col, dt_gl, inp, st, rad, out, phase_stop) result(status) &
      bind(C, name="ref_crm")
    use params
    use vars
    use grid
    use microphysics
    use micro_params, only: allocate_micro_params, tgrmin, a_gr
    use sgs
    use crmtracers, only: allocate_tracers
    use crmdims, only: crm_nx_rad, crm_ny_rad, crm_dx, crm_dt
    use setparm_mod, only: setparm
    use task_init_mod, only: task_init
    use abcoefs_mod, only: abcoefs
    use kurant_mod, only: kurant
    use setperturb_mod, only: setperturb
    use boundaries_mod, only: boundaries
    use forcing_mod, only: forcing
    use advect_mom_mod, only: advect_mom
    use adams_mod, only: adams
    use advect_all_scalars_mod, only: advect_all_scalars
    use sat_mod, only: qsatw_crm
    use crmsurface_mod, only: crmsurface, z0_est
    use precip_init_mod, only: precip_init
    use zero_mod, only: zero
    use buoyancy_mod, only: buoyancy
    use pressure_mod, only: pressure
    use uvw_mod, only: uvw
    use diagnose_mod, only: diagnose
    use damping_mod, only: damping
    use ice_fall_mod, only: ice_fall
    use coriolis_mod, only: coriolis
    type(crm_dims_t), intent(in) :: d
    type(crm_column_info_t), intent(in) :: col
    real(c_double), value :: dt_gl
    type(crm_input_view_t), intent(in) :: inp
    type(crm_state_view_t), intent(in) :: st
    type(crm_rad_view_t), intent(in) :: rad
    type(crm_output_view_t), intent(in) :: out
    integer(c_int), value :: phase_stop
    integer(c_int) :: status

    ! caller arrays
    real(c_double), pointer :: zmid(:,:), zint(:,:), tl(:,:), ql(:,:), qccl(:,:), qiil(:,:)
    real(c_double), pointer :: pmid(:,:), pint(:,:), pdel(:,:), ul(:,:), vl(:,:)
    real(c_double), pointer :: ocnfrac(:), tau00(:), wndls(:), bflxls(:)
    real(c_double), pointer :: s_u(:,:,:,:), s_v(:,:,:,:), s_w(:,:,:,:), s_t(:,:,:,:)
    real(c_double), pointer :: s_qt(:,:,:,:), s_qp(:,:,:,:), s_qn(:,:,:,:)
    real(c_double), pointer :: r_qrad(:,:,:,:), r_t(:,:,:,:), r_qv(:,:,:,:), r_qc(:,:,:,:)
    real(c_double), pointer :: r_qi(:,:,:,:), r_cld(:,:,:,:)
    real(c_double), pointer :: o_qcl(:,:,:,:), o_qci(:,:,:,:), o_qpl(:,:,:,:), o_qpi(:,:,:,:)
    real(c_double), pointer :: o_tk(:,:,:,:), o_tkh(:,:,:,:), o_prec_crm(:,:,:)
    real(c_double), pointer :: o_cltot(:), o_clhgh(:), o_clmed(:), o_cllow(:)
    real(c_double), pointer :: o_precc(:), o_precl(:), o_precsc(:), o_precsl(:)
    real(c_double), pointer :: o_jt(:), o_mx(:), o_prectend(:), o_precstend(:)
    real(c_double), pointer :: o_taux(:), o_tauy(:), o_z0m(:), o_timing(:)
    real(c_double), pointer :: o_cldtop(:,:), o_qc_mean(:,:), o_qi_mean(:,:), o_qs_mean(:,:)
    real(c_double), pointer :: o_qg_mean(:,:), o_qr_mean(:,:), o_sltend(:,:), o_qltend(:,:)
    real(c_double), pointer :: o_qcltend(:,:), o_qiltend(:,:), o_cld(:,:), o_gicewp(:,:)
    real(c_double), pointer :: o_gliqwp(:,:), o_mctot(:,:), o_mcup(:,:), o_mcdn(:,:)
    real(c_double), pointer :: o_mcuup(:,:), o_mcudn(:,:), o_mu(:,:), o_md(:,:), o_du(:,:)
    real(c_double), pointer :: o_eu(:,:), o_ed(:,:), o_flux_qt(:,:), o_fluxsgs_qt(:,:)
    real(c_double), pointer :: o_tkez(:,:), o_tkesgsz(:,:), o_tkz(:,:), o_flux_u(:,:)
    real(c_double), pointer :: o_flux_v(:,:), o_flux_qp(:,:), o_precflux(:,:), o_qt_ls(:,:)
    real(c_double), pointer :: o_qt_trans(:,:), o_qp_trans(:,:), o_qp_fall(:,:), o_qp_src(:,:)
    real(c_double), pointer :: o_qp_evp(:,:), o_t_ls(:,:)
    integer(c_int), pointer :: gcol(:)
    real(c_double), pointer :: lat(:), lon(:)

    real(8), parameter :: umax = 0.5*crm_dx/crm_dt
    real(8), parameter :: wmin = 2.
    real(crm_rknd), parameter :: cwp_threshold = 0.001
    integer, parameter :: perturb_seed_scale = 1000
    integer :: ncrms, plev, ld, ptop, i, j, k, l, icrm, icyc, nn, kx, i_rad, j_rad, igstep
    real(8) :: factor_xy, factor_xyt, idt_gl, crm_run_time, icrm_run_time
    real(crm_rknd) :: tmp, tmp1, tmp2, u2z, v2z, w2z, qsat, omg
    real(crm_rknd) :: crm_nx_rad_fac, crm_ny_rad_fac
    real(crm_rknd), allocatable :: t00(:,:), tln(:,:), qln(:,:), qccln(:,:), qiiln(:,:)
    real(crm_rknd), allocatable :: uln(:,:), vln(:,:), colprec(:), colprecs(:)
    real(crm_rknd), allocatable :: cwp(:,:,:), cwph(:,:,:), cwpm(:,:,:), cwpl(:,:,:)
    real(crm_rknd), allocatable :: cltemp(:,:,:), cmtemp(:,:,:), chtemp(:,:,:), cttemp(:,:,:)
    logical, allocatable :: flag_top(:,:,:)
    real(8), allocatable :: dd_crm(:,:), mui_crm(:,:), mdi_crm(:,:)
    real(crm_rknd), allocatable :: ustar(:), bflx(:), wnd(:)
    integer :: iseed

    status = 0
    ncrms = d%ncrms; plev = d%plev; ld = d%ld
    if (d%nx /= nx .or. d%ny /= ny .or. d%nz /= nzm .or. d%nx_rad /= crm_nx_rad .or. &
        d%ny_rad /= crm_ny_rad .or. d%dx /= crm_dx .or. d%dt /= crm_dt .or. ncrms < 1 .or. &
        ld < ncrms .or. plev < nzm + 1) then
      status = 1
      return
    end if
    if (mod(nx, crm_nx_rad) /= 0 .or. mod(ny, crm_ny_rad) /= 0) then
      status = 3
      return
    end if

    call c_f_pointer(inp%zmid, zmid, [ld, plev]);   call c_f_pointer(inp%zint, zint, [ld, plev+1])
    call c_f_pointer(inp%tl, tl, [ld, plev]);       call c_f_pointer(inp%ql, ql, [ld, plev])
    call c_f_pointer(inp%qccl, qccl, [ld, plev]);   call c_f_pointer(inp%qiil, qiil, [ld, plev])
    call c_f_pointer(inp%pmid, pmid, [ld, plev]);   call c_f_pointer(inp%pint, pint, [ld, plev+1])
    call c_f_pointer(inp%pdel, pdel, [ld, plev]);   call c_f_pointer(inp%ul, ul, [ld, plev])
    call c_f_pointer(inp%vl, vl, [ld, plev]);       call c_f_pointer(inp%ocnfrac, ocnfrac, [ld])
    call c_f_pointer(inp%tau00, tau00, [ld]);       call c_f_pointer(inp%wndls, wndls, [ld])
    call c_f_pointer(inp%bflxls, bflxls, [ld])
    call c_f_pointer(st%u_wind, s_u, [ld, nx, ny, nzm]); call c_f_pointer(st%v_wind, s_v, [ld, nx, ny, nzm])
    call c_f_pointer(st%w_wind, s_w, [ld, nx, ny, nzm]); call c_f_pointer(st%temperature, s_t, [ld, nx, ny, nzm])
    call c_f_pointer(st%qt, s_qt, [ld, nx, ny, nzm]);    call c_f_pointer(st%qp, s_qp, [ld, nx, ny, nzm])
    call c_f_pointer(st%qn, s_qn, [ld, nx, ny, nzm])
    call c_f_pointer(rad%qrad, r_qrad, [ld, crm_nx_rad, crm_ny_rad, nzm])
    call c_f_pointer(rad%temperature, r_t, [ld, crm_nx_rad, crm_ny_rad, nzm])
    call c_f_pointer(rad%qv, r_qv, [ld, crm_nx_rad, crm_ny_rad, nzm])
    call c_f_pointer(rad%qc, r_qc, [ld, crm_nx_rad, crm_ny_rad, nzm])
    call c_f_pointer(rad%qi, r_qi, [ld, crm_nx_rad, crm_ny_rad, nzm])
    call c_f_pointer(rad%cld, r_cld, [ld, crm_nx_rad, crm_ny_rad, nzm])
    call c_f_pointer(out%qcl, o_qcl, [ld, nx, ny, nzm]); call c_f_pointer(out%qci, o_qci, [ld, nx, ny, nzm])
    call c_f_pointer(out%qpl, o_qpl, [ld, nx, ny, nzm]); call c_f_pointer(out%qpi, o_qpi, [ld, nx, ny, nzm])
    call c_f_pointer(out%tk, o_tk, [ld, nx, ny, nzm]);   call c_f_pointer(out%tkh, o_tkh, [ld, nx, ny, nzm])
    call c_f_pointer(out%prec_crm, o_prec_crm, [ld, nx, ny])
    call c_f_pointer(out%cltot, o_cltot, [ld]); call c_f_pointer(out%clhgh, o_clhgh, [ld])
    call c_f_pointer(out%clmed, o_clmed, [ld]); call c_f_pointer(out%cllow, o_cllow, [ld])
    call c_f_pointer(out%precc, o_precc, [ld]); call c_f_pointer(out%precl, o_precl, [ld])
    call c_f_pointer(out%precsc, o_precsc, [ld]); call c_f_pointer(out%precsl, o_precsl, [ld])
    call c_f_pointer(out%jt_crm, o_jt, [ld]); call c_f_pointer(out%mx_crm, o_mx, [ld])
    call c_f_pointer(out%prectend, o_prectend, [ld]); call c_f_pointer(out%precstend, o_precstend, [ld])
    call c_f_pointer(out%taux, o_taux, [ld]); call c_f_pointer(out%tauy, o_tauy, [ld])
    call c_f_pointer(out%z0m, o_z0m, [ld]); call c_f_pointer(out%timing_factor, o_timing, [ld])
    call c_f_pointer(out%cldtop, o_cldtop, [ld, plev]); call c_f_pointer(out%qc_mean, o_qc_mean, [ld, plev])
    call c_f_pointer(out%qi_mean, o_qi_mean, [ld, plev]); call c_f_pointer(out%qs_mean, o_qs_mean, [ld, plev])
    call c_f_pointer(out%qg_mean, o_qg_mean, [ld, plev]); call c_f_pointer(out%qr_mean, o_qr_mean, [ld, plev])
    call c_f_pointer(out%sltend, o_sltend, [ld, plev]); call c_f_pointer(out%qltend, o_qltend, [ld, plev])
    call c_f_pointer(out%qcltend, o_qcltend, [ld, plev]); call c_f_pointer(out%qiltend, o_qiltend, [ld, plev])
    call c_f_pointer(out%cld, o_cld, [ld, plev]); call c_f_pointer(out%gicewp, o_gicewp, [ld, plev])
    call c_f_pointer(out%gliqwp, o_gliqwp, [ld, plev]); call c_f_pointer(out%mctot, o_mctot, [ld, plev])
    call c_f_pointer(out%mcup, o_mcup, [ld, plev]); call c_f_pointer(out%mcdn, o_mcdn, [ld, plev])
    call c_f_pointer(out%mcuup, o_mcuup, [ld, plev]); call c_f_pointer(out%mcudn, o_mcudn, [ld, plev])
    call c_f_pointer(out%mu_crm, o_mu, [ld, plev]); call c_f_pointer(out%md_crm, o_md, [ld, plev])
    call c_f_pointer(out%du_crm, o_du, [ld, plev]); call c_f_pointer(out%eu_crm, o_eu, [ld, plev])
    call c_f_pointer(out%ed_crm, o_ed, [ld, plev]); call c_f_pointer(out%flux_qt, o_flux_qt, [ld, plev])
    call c_f_pointer(out%fluxsgs_qt, o_fluxsgs_qt, [ld, plev]); call c_f_pointer(out%tkez, o_tkez, [ld, plev])
    call c_f_pointer(out%tkesgsz, o_tkesgsz, [ld, plev]); call c_f_pointer(out%tkz, o_tkz, [ld, plev])
    call c_f_pointer(out%flux_u, o_flux_u, [ld, plev]); call c_f_pointer(out%flux_v, o_flux_v, [ld, plev])
    call c_f_pointer(out%flux_qp, o_flux_qp, [ld, plev]); call c_f_pointer(out%precflux, o_precflux, [ld, plev])
    call c_f_pointer(out%qt_ls, o_qt_ls, [ld, plev]); call c_f_pointer(out%qt_trans, o_qt_trans, [ld, plev])
    call c_f_pointer(out%qp_trans, o_qp_trans, [ld, plev]); call c_f_pointer(out%qp_fall, o_qp_fall, [ld, plev])
    call c_f_pointer(out%qp_src, o_qp_src, [ld, plev]); call c_f_pointer(out%qp_evp, o_qp_evp, [ld, plev])
    call c_f_pointer(out%t_ls, o_t_ls, [ld, plev])
    call c_f_pointer(col%gcol, gcol, [ncrms])
    call c_f_pointer(col%lat, lat, [ncrms]); call c_f_pointer(col%lon, lon, [ncrms])

    call free_reference_state()
    allocate(t00(ncrms,nz), tln(ncrms,plev), qln(ncrms,plev), qccln(ncrms,plev))
    allocate(qiiln(ncrms,plev), uln(ncrms,plev), vln(ncrms,plev))
    allocate(cwp(ncrms,nx,ny), cwph(ncrms,nx,ny), cwpm(ncrms,nx,ny), cwpl(ncrms,nx,ny))
    allocate(flag_top(ncrms,nx,ny), cltemp(ncrms,nx,ny), cmtemp(ncrms,nx,ny))
    allocate(chtemp(ncrms,nx,ny), cttemp(ncrms,nx,ny))
    allocate(dd_crm(ncrms,plev), mui_crm(ncrms,plev+1), mdi_crm(ncrms,plev+1))
    allocate(ustar(ncrms), bflx(ncrms), wnd(ncrms), colprec(ncrms), colprecs(ncrms))

    call allocate_params(ncrms)
    call allocate_vars(ncrms)
    call allocate_grid(ncrms)
    call allocate_tracers(ncrms)
    call allocate_sgs(ncrms)
    call allocate_micro(ncrms)
    call allocate_micro_params(ncrms)
    live = .true.
    live_ncrms = ncrms
    nstep = 0

    ! set-up: crm_module.F90:286-375
    do icrm = 1, ncrms
      latitude0(icrm) = lat(icrm) * 57.296_8
      longitude0(icrm) = lon(icrm) * 57.296_8
    end do
    igstep = col%igstep
    dostatis = .false.
    idt_gl = 1._8/dt_gl
    ptop = plev - nzm + 1
    factor_xy = 1._8/dble(nx*ny)
    r_t(1:ncrms,:,:,:) = 0.; r_qv(1:ncrms,:,:,:) = 0.; r_qc(1:ncrms,:,:,:) = 0.
    r_qi(1:ncrms,:,:,:) = 0.; r_cld(1:ncrms,:,:,:) = 0.
    do icrm = 1, ncrms
      bflx(icrm) = bflxls(icrm)
      wnd(icrm) = wndls(icrm)
    end do
    call task_init()
    call setparm()
    if (2*dt > 60.) then      ! damping.F90:39-42 would abort (tau_min = 60 s)
      status = 4
      return
    end if
    do icrm = 1, ncrms
      fcor(icrm) = 4*pi/86400.*sin(latitude0(icrm)*pi/180.)
      fcorz(icrm) = sqrt(4.*(2*pi/(3600.*24.))**2 - fcor(icrm)**2)
      fcory(icrm,:) = fcor(icrm)
      fcorzy(icrm,:) = fcorz(icrm)
      latitude(icrm,:,:) = latitude0(icrm)
      longitude(icrm,:,:) = longitude0(icrm)
      if (ocnfrac(icrm) > 0.5) then
        ocean(icrm) = .true.
      else
        land(icrm) = .true.
      end if
      do k = 1, nzm
        z(icrm,k) = zmid(icrm,plev-k+1) - zint(icrm,plev+1)
        zi(icrm,k) = zint(icrm,plev-k+2) - zint(icrm,plev+1)
        pres(icrm,k) = pmid(icrm,plev-k+1)/100.
        presi(icrm,k) = pint(icrm,plev-k+2)/100.
        prespot(icrm,k) = (1000./pres(icrm,k))**(rgas/cp)
        bet(icrm,k) = ggr/tl(icrm,plev-k+1)
        gamaz(icrm,k) = ggr/cp*z(icrm,k)
      end do
      zi(icrm,nz) = zint(icrm,plev-nz+2) - zint(icrm,plev+1)
      presi(icrm,nz) = pint(icrm,plev-nz+2)/100.
      dz(icrm) = 0.5*(z(icrm,1) + z(icrm,2))
      do k = 2, nzm
        adzw(icrm,k) = (z(icrm,k) - z(icrm,k-1))/dz(icrm)
      end do
      adzw(icrm,1) = 1.
      adzw(icrm,nz) = adzw(icrm,nzm)
      do k = 1, nzm
        adz(icrm,k) = (zi(icrm,k+1) - zi(icrm,k))/dz(icrm)
      end do
      do k = 1, nzm
        rho(icrm,k) = pdel(icrm,plev-k+1)/ggr/(adz(icrm,k)*dz(icrm))
      end do
      do k = 2, nzm
        rhow(icrm,k) = (pmid(icrm,plev-k+2) - pmid(icrm,plev-k+1))/ggr/(adzw(icrm,k)*dz(icrm))
      end do
      rhow(icrm,1) = 2.*rhow(icrm,2) - rhow(icrm,3)
      rhow(icrm,nz) = 2.*rhow(icrm,nzm) - rhow(icrm,nzm-1)
    end do

    ! initial fields: crm_module.F90:379-452
    do k = 1, nzm
      do j = 1, ny
        do i = 1, nx
          do icrm = 1, ncrms
            u(icrm,i,j,k) = s_u(icrm,i,j,k)
            v(icrm,i,j,k) = s_v(icrm,i,j,k)*YES3D
            w(icrm,i,j,k) = s_w(icrm,i,j,k)
            tabs(icrm,i,j,k) = s_t(icrm,i,j,k)
          end do
        end do
      end do
    end do
    ! first-step wind clamp, decided by instance 1 (crm_module.F90:399-417)
    if (u(1,1,1,1) == u(1,2,1,1) .and. u(1,3,1,2) == u(1,4,1,2)) then
      do k = 1, nzm
        do j = 1, ny
          do i = 1, nx
            do icrm = 1, ncrms
              u(icrm,i,j,k) = min(umax, max(-umax, u(icrm,i,j,k)))
              v(icrm,i,j,k) = min(umax, max(-umax, v(icrm,i,j,k)))*YES3D
            end do
          end do
        end do
      end do
    end if
    do k = 1, nzm
      do j = 1, ny
        do i = 1, nx
          do icrm = 1, ncrms
            micro_field(icrm,i,j,k,1) = s_qt(icrm,i,j,k)
            micro_field(icrm,i,j,k,2) = s_qp(icrm,i,j,k)
            qn(icrm,i,j,k) = s_qn(icrm,i,j,k)
          end do
        end do
      end do
    end do
    call micro_init(ncrms)
    call sgs_init(ncrms)

    ! domain means and large-scale tendencies: crm_module.F90:482-594
    colprec = 0; colprecs = 0
    u0 = 0.; v0 = 0.; t0 = 0.; t00 = 0.; tabs0 = 0.; q0 = 0.; qv0 = 0.; qn0 = 0.; qp0 = 0.; tke0 = 0.
    do k = 1, nzm
      do j = 1, ny
        do i = 1, nx
          do icrm = 1, ncrms
            t(icrm,i,j,k) = tabs(icrm,i,j,k) + gamaz(icrm,k) - fac_cond*qcl(icrm,i,j,k) - fac_sub*qci(icrm,i,j,k) &
                            - fac_cond*qpl(icrm,i,j,k) - fac_sub*qpi(icrm,i,j,k)
            colprec(icrm) = colprec(icrm) + (qpl(icrm,i,j,k) + qpi(icrm,i,j,k))*pdel(icrm,plev-k+1)
            colprecs(icrm) = colprecs(icrm) + qpi(icrm,i,j,k)*pdel(icrm,plev-k+1)
            u0(icrm,k) = u0(icrm,k) + u(icrm,i,j,k)
            v0(icrm,k) = v0(icrm,k) + v(icrm,i,j,k)
            t0(icrm,k) = t0(icrm,k) + t(icrm,i,j,k)
            t00(icrm,k) = t00(icrm,k) + (t(icrm,i,j,k) + fac_cond*qpl(icrm,i,j,k) + fac_sub*qpi(icrm,i,j,k))
            tabs0(icrm,k) = tabs0(icrm,k) + tabs(icrm,i,j,k)
            q0(icrm,k) = q0(icrm,k) + (qv(icrm,i,j,k) + qcl(icrm,i,j,k) + qci(icrm,i,j,k))
            qv0(icrm,k) = qv0(icrm,k) + qv(icrm,i,j,k)
            qn0(icrm,k) = qn0(icrm,k) + (qcl(icrm,i,j,k) + qci(icrm,i,j,k))
            qp0(icrm,k) = qp0(icrm,k) + (qpl(icrm,i,j,k) + qpi(icrm,i,j,k))
            tke0(icrm,k) = tke0(icrm,k) + sgs_field(icrm,i,j,k,1)
          end do
        end do
      end do
    end do
    do k = 1, nzm
      do icrm = 1, ncrms
        u0(icrm,k) = u0(icrm,k)*factor_xy
        v0(icrm,k) = v0(icrm,k)*factor_xy
        t0(icrm,k) = t0(icrm,k)*factor_xy
        t00(icrm,k) = t00(icrm,k)*factor_xy
        tabs0(icrm,k) = tabs0(icrm,k)*factor_xy
        q0(icrm,k) = q0(icrm,k)*factor_xy
        qv0(icrm,k) = qv0(icrm,k)*factor_xy
        qn0(icrm,k) = qn0(icrm,k)*factor_xy
        qp0(icrm,k) = qp0(icrm,k)*factor_xy
        tke0(icrm,k) = tke0(icrm,k)*factor_xy
        l = plev - k + 1
        uln(icrm,l) = min(umax, max(-umax, ul(icrm,l)))
        vln(icrm,l) = min(umax, max(-umax, vl(icrm,l)))*YES3D
        ttend(icrm,k) = (tl(icrm,l) + gamaz(icrm,k) - fac_cond*(qccl(icrm,l) + qiil(icrm,l)) &
                         - fac_fus*qiil(icrm,l) - t00(icrm,k))*idt_gl
        qtend(icrm,k) = (ql(icrm,l) + qccl(icrm,l) + qiil(icrm,l) - q0(icrm,k))*idt_gl
        utend(icrm,k) = (uln(icrm,l) - u0(icrm,k))*idt_gl
        vtend(icrm,k) = (vln(icrm,l) - v0(icrm,k))*idt_gl
        ug0(icrm,k) = uln(icrm,l)
        vg0(icrm,k) = vln(icrm,l)
        tg0(icrm,k) = tl(icrm,l) + gamaz(icrm,k) - fac_cond*qccl(icrm,l) - fac_sub*qiil(icrm,l)
        qg0(icrm,k) = ql(icrm,l) + qccl(icrm,l) + qiil(icrm,l)
      end do
    end do
    do icrm = 1, ncrms
      uhl(icrm) = u0(icrm,1)
      vhl(icrm) = v0(icrm,1)
      ustar(icrm) = sqrt(tau00(icrm)/rho(icrm,1))
      z0(icrm) = z0_est(z(icrm,1), bflx(icrm), wnd(icrm), ustar(icrm))
      z0(icrm) = max(real(0.00001,crm_rknd), min(real(1.,crm_rknd), z0(icrm)))
      o_timing(icrm) = 0.
      o_prectend(icrm) = colprec(icrm)
      o_precstend(icrm) = colprecs(icrm)
    end do
    o_cld(1:ncrms,:) = 0.; o_cldtop(1:ncrms,:) = 0.; o_gicewp(1:ncrms,:) = 0.; o_gliqwp(1:ncrms,:) = 0.
    o_mctot(1:ncrms,:) = 0.; o_mcup(1:ncrms,:) = 0.; o_mcdn(1:ncrms,:) = 0.; o_mcuup(1:ncrms,:) = 0.
    o_mcudn(1:ncrms,:) = 0.; o_qc_mean(1:ncrms,:) = 0.; o_qi_mean(1:ncrms,:) = 0.
    o_qs_mean(1:ncrms,:) = 0.; o_qg_mean(1:ncrms,:) = 0.; o_qr_mean(1:ncrms,:) = 0.
    o_mu(1:ncrms,:) = 0.; o_md(1:ncrms,:) = 0.; o_eu(1:ncrms,:) = 0.; o_du(1:ncrms,:) = 0.
    o_ed(1:ncrms,:) = 0.; o_flux_qt(1:ncrms,:) = 0.; o_flux_u(1:ncrms,:) = 0.; o_flux_v(1:ncrms,:) = 0.
    o_fluxsgs_qt(1:ncrms,:) = 0.; o_tkez(1:ncrms,:) = 0.; o_tkesgsz(1:ncrms,:) = 0.; o_tkz(1:ncrms,:) = 0.
    o_flux_qp(1:ncrms,:) = 0.; o_precflux(1:ncrms,:) = 0.; o_qt_trans(1:ncrms,:) = 0.
    o_qp_trans(1:ncrms,:) = 0.; o_qp_fall(1:ncrms,:) = 0.; o_qp_evp(1:ncrms,:) = 0.
    o_qp_src(1:ncrms,:) = 0.; o_qt_ls(1:ncrms,:) = 0.; o_t_ls(1:ncrms,:) = 0.
    dd_crm = 0.; mui_crm = 0.; mdi_crm = 0.
    o_jt(1:ncrms) = 0.; o_mx(1:ncrms) = 0.

    if (doprecip) call precip_init(ncrms)
    do icrm = 1, ncrms
      if (igstep <= 1) then
        iseed = gcol(icrm)*perturb_seed_scale
        call setperturb(ncrms, icrm, iseed)
      end if
    end do
    crm_nx_rad_fac = real(crm_nx_rad,crm_rknd)/real(nx,crm_rknd)
    crm_ny_rad_fac = real(crm_ny_rad,crm_rknd)/real(ny,crm_rknd)

    nstop = dt_gl/dt
    dt = dt_gl/nstop
    crm_run_time = dt_gl
    icrm_run_time = 1._8/crm_run_time

    ! main loop: crm_module.F90:727-1075
    nstep = 0
    do while (nstep < nstop)
      nstep = nstep + 1
      o_timing(1:ncrms) = o_timing(1:ncrms) + 1
      call kurant(ncrms)
      if (ncycle > 16) then
        status = 2
        return
      end if
      do icyc = 1, ncycle
        icycle = icyc
        dtn = dt/ncycle
        dt3(na) = dtn
        dtfactor = dtn/dt
        call abcoefs(ncrms)
        call zero(ncrms)
        if (stop_here(1)) return
        call buoyancy(ncrms)
        if (stop_here(2)) return
        call forcing(ncrms)
        if (stop_here(3)) return
        do k = 1, nzm
          do j = 1, ny
            do i = 1, nx
              do icrm = 1, ncrms
                i_rad = (i-1)/(nx/crm_nx_rad) + 1
                j_rad = (j-1)/(ny/crm_ny_rad) + 1
                t(icrm,i,j,k) = t(icrm,i,j,k) + r_qrad(icrm,i_rad,j_rad,k)*dtn
              end do
            end do
          end do
        end do
        if (stop_here(4)) return
        if (dodamping) call damping(ncrms)
        if (stop_here(5)) return
        if (docloud) call ice_fall(ncrms)
        if (stop_here(6)) return
        call boundaries(ncrms, 3)
        call boundaries(ncrms, 0)
        if (dosurface) call crmsurface(ncrms, bflx)
        if (stop_here(7)) return
        if (dosgs) call sgs_proc(ncrms)
        if (stop_here(8)) return
        call boundaries(ncrms, 4)
        if (stop_here(9)) return
        call advect_mom(ncrms)
        if (stop_here(10)) return
        if (dosgs) call sgs_mom(ncrms)
        if (stop_here(11)) return
        if (docoriolis) call coriolis(ncrms)
        call pressure(ncrms)
        if (stop_here(12)) return
        call adams(ncrms)
        if (stop_here(13)) return
        call boundaries(ncrms, 2)
        if (stop_here(14)) return
        call advect_all_scalars(ncrms)
        if (stop_here(15)) return
        call uvw(ncrms)
        if (stop_here(16)) return
        call boundaries(ncrms, 3)
        if (stop_here(17)) return
        if (dosgs) call sgs_scalars(ncrms)
        if (stop_here(18)) return
        if (docloud .or. dosmoke) call micro_proc(ncrms)
        if (stop_here(19)) return
        call diagnose(ncrms)
        if (stop_here(20)) return
        nn = na
        na = nc
        nc = nb
        nb = nn
      end do

      ! per-step statistics: crm_module.F90:896-1073
      cwp = 0.; cwph = 0.; cwpm = 0.; cwpl = 0.; flag_top = .true.
      cltemp = 0.; cmtemp = 0.; chtemp = 0.; cttemp = 0.
      do j = 1, ny
        do i = 1, nx
          do icrm = 1, ncrms
            do k = 1, nzm
              l = plev - k + 1
              tmp1 = rho(icrm,nz-k)*adz(icrm,nz-k)*dz(icrm)*(qcl(icrm,i,j,nz-k) + qci(icrm,i,j,nz-k))
              cwp(icrm,i,j) = cwp(icrm,i,j) + tmp1
              cttemp(icrm,i,j) = max(cf3d(icrm,i,j,nz-k), cttemp(icrm,i,j))
              if (cwp(icrm,i,j) > cwp_threshold .and. flag_top(icrm,i,j)) then
                o_cldtop(icrm,l) = o_cldtop(icrm,l) + 1
                flag_top(icrm,i,j) = .false.
              end if
              if (pres(icrm,nz-k) >= 700.) then
                cwpl(icrm,i,j) = cwpl(icrm,i,j) + tmp1
                cltemp(icrm,i,j) = max(cf3d(icrm,i,j,nz-k), cltemp(icrm,i,j))
              else if (pres(icrm,nz-k) < 400.) then
                cwph(icrm,i,j) = cwph(icrm,i,j) + tmp1
                chtemp(icrm,i,j) = max(cf3d(icrm,i,j,nz-k), chtemp(icrm,i,j))
              else
                cwpm(icrm,i,j) = cwpm(icrm,i,j) + tmp1
                cmtemp(icrm,i,j) = max(cf3d(icrm,i,j,nz-k), cmtemp(icrm,i,j))
              end if
              tmp1 = rho(icrm,k)*adz(icrm,k)*dz(icrm)
              if (tmp1*(qcl(icrm,i,j,k) + qci(icrm,i,j,k)) > cwp_threshold) then
                o_cld(icrm,l) = o_cld(icrm,l) + cf3d(icrm,i,j,k)
                if (w(icrm,i,j,k+1) + w(icrm,i,j,k) > 2*wmin) then
                  o_mcup(icrm,l) = o_mcup(icrm,l) + rho(icrm,k)*0.5*(w(icrm,i,j,k+1) + w(icrm,i,j,k))*cf3d(icrm,i,j,k)
                  o_mcuup(icrm,l) = o_mcuup(icrm,l) + rho(icrm,k)*0.5*(w(icrm,i,j,k+1) + w(icrm,i,j,k))*(1.0 - cf3d(icrm,i,j,k))
                end if
                if (w(icrm,i,j,k+1) + w(icrm,i,j,k) < -2*wmin) then
                  o_mcdn(icrm,l) = o_mcdn(icrm,l) + rho(icrm,k)*0.5*(w(icrm,i,j,k+1) + w(icrm,i,j,k))*cf3d(icrm,i,j,k)
                  o_mcudn(icrm,l) = o_mcudn(icrm,l) + rho(icrm,k)*0.5*(w(icrm,i,j,k+1) + w(icrm,i,j,k))*(1. - cf3d(icrm,i,j,k))
                end if
              else
                if (w(icrm,i,j,k+1) + w(icrm,i,j,k) > 2*wmin) then
                  o_mcuup(icrm,l) = o_mcuup(icrm,l) + rho(icrm,k)*0.5*(w(icrm,i,j,k+1) + w(icrm,i,j,k))
                end if
                if (w(icrm,i,j,k+1) + w(icrm,i,j,k) < -2*wmin) then
                  o_mcudn(icrm,l) = o_mcudn(icrm,l) + rho(icrm,k)*0.5*(w(icrm,i,j,k+1) + w(icrm,i,j,k))
                end if
              end if
              o_gliqwp(icrm,l) = o_gliqwp(icrm,l) + qcl(icrm,i,j,k)
              o_gicewp(icrm,l) = o_gicewp(icrm,l) + qci(icrm,i,j,k)
            end do
          end do
        end do
      end do
      do k = 1, nzm
        do j = 1, ny
          do i = 1, nx
            do icrm = 1, ncrms
              i_rad = (i-1)/(nx/crm_nx_rad) + 1
              j_rad = (j-1)/(ny/crm_ny_rad) + 1
              r_t(icrm,i_rad,j_rad,k) = r_t(icrm,i_rad,j_rad,k) + tabs(icrm,i,j,k)
              r_qv(icrm,i_rad,j_rad,k) = r_qv(icrm,i_rad,j_rad,k) + max(real(0.,crm_rknd), qv(icrm,i,j,k))
              r_qc(icrm,i_rad,j_rad,k) = r_qc(icrm,i_rad,j_rad,k) + qcl(icrm,i,j,k)
              r_qi(icrm,i_rad,j_rad,k) = r_qi(icrm,i_rad,j_rad,k) + qci(icrm,i,j,k)
              if (qcl(icrm,i,j,k) + qci(icrm,i,j,k) > 0) then
                r_cld(icrm,i_rad,j_rad,k) = r_cld(icrm,i_rad,j_rad,k) + cf3d(icrm,i,j,k)
              end if
            end do
          end do
        end do
      end do
      do j = 1, ny
        do i = 1, nx
          do icrm = 1, ncrms
            do k = 1, nzm+1
              l = plev + 1 - k + 1
              if (w(icrm,i,j,k) > 0.) then
                kx = max(1, k-1)
                qsat = qsatw_crm(tabs(icrm,i,j,kx), pres(icrm,kx))
                if (qcl(icrm,i,j,kx) + qci(icrm,i,j,kx) > min(real(1.e-5,crm_rknd), 0.01*qsat)) then
                  mui_crm(icrm,l) = mui_crm(icrm,l) + rhow(icrm,k)*w(icrm,i,j,k)
                end if
              else if (w(icrm,i,j,k) < 0.) then
                kx = min(k+1, nzm)
                qsat = qsatw_crm(tabs(icrm,i,j,kx), pres(icrm,kx))
                if (qcl(icrm,i,j,kx) + qci(icrm,i,j,kx) > min(real(1.e-5,crm_rknd), 0.01*qsat)) then
                  mdi_crm(icrm,l) = mdi_crm(icrm,l) + rhow(icrm,k)*w(icrm,i,j,k)
                else if (qpl(icrm,i,j,kx) + qpi(icrm,i,j,kx) > 1.0e-4) then
                  mdi_crm(icrm,l) = mdi_crm(icrm,l) + rhow(icrm,k)*w(icrm,i,j,k)
                end if
              end if
            end do
          end do
        end do
      end do
      do j = 1, ny
        do i = 1, nx
          do icrm = 1, ncrms
            if (cwp(icrm,i,j) > cwp_threshold) o_cltot(icrm) = o_cltot(icrm) + cttemp(icrm,i,j)
            if (cwph(icrm,i,j) > cwp_threshold) o_clhgh(icrm) = o_clhgh(icrm) + chtemp(icrm,i,j)
            if (cwpm(icrm,i,j) > cwp_threshold) o_clmed(icrm) = o_clmed(icrm) + cmtemp(icrm,i,j)
            if (cwpl(icrm,i,j) > cwp_threshold) o_cllow(icrm) = o_cllow(icrm) + cltemp(icrm,i,j)
          end do
        end do
      end do
    end do

    ! end-of-call outputs: crm_module.F90:1078-1659
    factor_xyt = factor_xy/real(nstop,crm_rknd)
    tmp1 = crm_nx_rad_fac*crm_ny_rad_fac/real(nstop,crm_rknd)
    r_t(1:ncrms,:,:,:) = r_t(1:ncrms,:,:,:)*tmp1
    r_qv(1:ncrms,:,:,:) = r_qv(1:ncrms,:,:,:)*tmp1
    r_qc(1:ncrms,:,:,:) = r_qc(1:ncrms,:,:,:)*tmp1
    r_qi(1:ncrms,:,:,:) = r_qi(1:ncrms,:,:,:)*tmp1
    r_cld(1:ncrms,:,:,:) = r_cld(1:ncrms,:,:,:)*tmp1
    do k = 1, ptop-1
      do icrm = 1, ncrms
        tln(icrm,k) = tl(icrm,k); qln(icrm,k) = ql(icrm,k); qccln(icrm,k) = qccl(icrm,k)
        qiiln(icrm,k) = qiil(icrm,k); uln(icrm,k) = ul(icrm,k); vln(icrm,k) = vl(icrm,k)
      end do
    end do
    do k = ptop, plev
      do icrm = 1, ncrms
        tln(icrm,k) = 0.; qln(icrm,k) = 0.; qccln(icrm,k) = 0.
        qiiln(icrm,k) = 0.; uln(icrm,k) = 0.; vln(icrm,k) = 0.
      end do
    end do
    colprec = 0; colprecs = 0
    do k = 1, nzm
      do i = 1, nx
        do j = 1, ny
          do icrm = 1, ncrms
            l = plev - k + 1
            colprec(icrm) = colprec(icrm) + (qpl(icrm,i,j,k) + qpi(icrm,i,j,k))*pdel(icrm,plev-k+1)
            colprecs(icrm) = colprecs(icrm) + qpi(icrm,i,j,k)*pdel(icrm,plev-k+1)
            tln(icrm,l) = tln(icrm,l) + tabs(icrm,i,j,k)
            qln(icrm,l) = qln(icrm,l) + qv(icrm,i,j,k)
            qccln(icrm,l) = qccln(icrm,l) + qcl(icrm,i,j,k)
            qiiln(icrm,l) = qiiln(icrm,l) + qci(icrm,i,j,k)
            uln(icrm,l) = uln(icrm,l) + u(icrm,i,j,k)
            vln(icrm,l) = vln(icrm,l) + v(icrm,i,j,k)
          end do
        end do
      end do
    end do
    do k = ptop, plev
      do icrm = 1, ncrms
        tln(icrm,k) = tln(icrm,k)*factor_xy; qln(icrm,k) = qln(icrm,k)*factor_xy
        qccln(icrm,k) = qccln(icrm,k)*factor_xy; qiiln(icrm,k) = qiiln(icrm,k)*factor_xy
        uln(icrm,k) = uln(icrm,k)*factor_xy; vln(icrm,k) = vln(icrm,k)*factor_xy
      end do
    end do
    do k = 1, plev
      do icrm = 1, ncrms
        o_sltend(icrm,k) = cp*(tln(icrm,k) - tl(icrm,k))*icrm_run_time
        o_qltend(icrm,k) = (qln(icrm,k) - ql(icrm,k))*icrm_run_time
        o_qcltend(icrm,k) = (qccln(icrm,k) - qccl(icrm,k))*icrm_run_time
        o_qiltend(icrm,k) = (qiiln(icrm,k) - qiil(icrm,k))*icrm_run_time
      end do
    end do
    do icrm = 1, ncrms
      o_prectend(icrm) = (colprec(icrm) - o_prectend(icrm))/ggr*factor_xy*icrm_run_time
      o_precstend(icrm) = (colprecs(icrm) - o_precstend(icrm))/ggr*factor_xy*icrm_run_time
    end do
    do k = ptop, ptop+1
      do icrm = 1, ncrms
        o_sltend(icrm,k) = 0.; o_qltend(icrm,k) = 0.; o_qcltend(icrm,k) = 0.; o_qiltend(icrm,k) = 0.
      end do
    end do
    do k = 1, nzm
      do j = 1, ny
        do i = 1, nx
          do icrm = 1, ncrms
            s_u(icrm,i,j,k) = u(icrm,i,j,k)
            s_v(icrm,i,j,k) = v(icrm,i,j,k)
            s_w(icrm,i,j,k) = w(icrm,i,j,k)
            s_t(icrm,i,j,k) = tabs(icrm,i,j,k)
            s_qt(icrm,i,j,k) = micro_field(icrm,i,j,k,1)
            s_qp(icrm,i,j,k) = micro_field(icrm,i,j,k,2)
            s_qn(icrm,i,j,k) = qn(icrm,i,j,k)
            o_tk(icrm,i,j,k) = sgs_field_diag(icrm,i,j,k,1)
            o_tkh(icrm,i,j,k) = sgs_field_diag(icrm,i,j,k,2)
            o_qcl(icrm,i,j,k) = qcl(icrm,i,j,k)
            o_qci(icrm,i,j,k) = qci(icrm,i,j,k)
            o_qpl(icrm,i,j,k) = qpl(icrm,i,j,k)
            o_qpi(icrm,i,j,k) = qpi(icrm,i,j,k)
          end do
        end do
      end do
    end do
    do icrm = 1, ncrms
      o_z0m(icrm) = z0(icrm)
      o_taux(icrm) = taux0(icrm)/dble(nstop)
      o_tauy(icrm) = tauy0(icrm)/dble(nstop)
    end do
    do k = 1, nzm
      do j = 1, ny
        do i = 1, nx
          do icrm = 1, ncrms
            l = plev - k + 1
            o_qc_mean(icrm,l) = o_qc_mean(icrm,l) + qcl(icrm,i,j,k)
            o_qi_mean(icrm,l) = o_qi_mean(icrm,l) + qci(icrm,i,j,k)
            o_qr_mean(icrm,l) = o_qr_mean(icrm,l) + qpl(icrm,i,j,k)
            omg = max(real(0.,crm_rknd), min(real(1.,crm_rknd), (tabs(icrm,i,j,k) - tgrmin)*a_gr))
            o_qg_mean(icrm,l) = o_qg_mean(icrm,l) + qpi(icrm,i,j,k)*omg
            o_qs_mean(icrm,l) = o_qs_mean(icrm,l) + qpi(icrm,i,j,k)*(1. - omg)
          end do
        end do
      end do
    end do
    do k = 1, plev
      do icrm = 1, ncrms
        o_cld(icrm,k) = min(1._8, o_cld(icrm,k)*factor_xyt)
        o_cldtop(icrm,k) = min(1._8, o_cldtop(icrm,k)*factor_xyt)
        o_gicewp(icrm,k) = o_gicewp(icrm,k)*pdel(icrm,k)*1000./ggr*factor_xyt
        o_gliqwp(icrm,k) = o_gliqwp(icrm,k)*pdel(icrm,k)*1000./ggr*factor_xyt
        o_mcup(icrm,k) = o_mcup(icrm,k)*factor_xyt
        o_mcdn(icrm,k) = o_mcdn(icrm,k)*factor_xyt
        o_mcuup(icrm,k) = o_mcuup(icrm,k)*factor_xyt
        o_mcudn(icrm,k) = o_mcudn(icrm,k)*factor_xyt
        o_mctot(icrm,k) = o_mcup(icrm,k) + o_mcdn(icrm,k) + o_mcuup(icrm,k) + o_mcudn(icrm,k)
        o_qc_mean(icrm,k) = o_qc_mean(icrm,k)*factor_xy
        o_qi_mean(icrm,k) = o_qi_mean(icrm,k)*factor_xy
        o_qs_mean(icrm,k) = o_qs_mean(icrm,k)*factor_xy
        o_qg_mean(icrm,k) = o_qg_mean(icrm,k)*factor_xy
        o_qr_mean(icrm,k) = o_qr_mean(icrm,k)*factor_xy
      end do
    end do
    do icrm = 1, ncrms
      o_precc(icrm) = 0.; o_precl(icrm) = 0.; o_precsc(icrm) = 0.; o_precsl(icrm) = 0.
    end do
    do j = 1, ny
      do i = 1, nx
        do icrm = 1, ncrms
          precsfc(icrm,i,j) = precsfc(icrm,i,j)*dz(icrm)/dt/dble(nstop)
          precssfc(icrm,i,j) = precssfc(icrm,i,j)*dz(icrm)/dt/dble(nstop)
          if (precsfc(icrm,i,j) > 10./86400.) then
            o_precc(icrm) = o_precc(icrm) + precsfc(icrm,i,j)
            o_precsc(icrm) = o_precsc(icrm) + precssfc(icrm,i,j)
          else
            o_precl(icrm) = o_precl(icrm) + precsfc(icrm,i,j)
            o_precsl(icrm) = o_precsl(icrm) + precssfc(icrm,i,j)
          end if
          o_prec_crm(icrm,i,j) = precsfc(icrm,i,j)/1000.
        end do
      end do
    end do
    do icrm = 1, ncrms
      o_precc(icrm) = o_precc(icrm)*factor_xy/1000.
      o_precl(icrm) = o_precl(icrm)*factor_xy/1000.
      o_precsc(icrm) = o_precsc(icrm)*factor_xy/1000.
      o_precsl(icrm) = o_precsl(icrm)*factor_xy/1000.
      o_cltot(icrm) = o_cltot(icrm)*factor_xyt
      o_clhgh(icrm) = o_clhgh(icrm)*factor_xyt
      o_clmed(icrm) = o_clmed(icrm)*factor_xyt
      o_cllow(icrm) = o_cllow(icrm)*factor_xyt
      o_jt(icrm) = plev*1.0
      o_mx(icrm) = 1.0
    end do
    do k = 1, plev
      do icrm = 1, ncrms
        o_mu(icrm,k) = 0.5*(mui_crm(icrm,k) + mui_crm(icrm,k+1))
        o_md(icrm,k) = 0.5*(mdi_crm(icrm,k) + mdi_crm(icrm,k+1))
        o_mu(icrm,k) = o_mu(icrm,k)*ggr/100.
        o_md(icrm,k) = o_md(icrm,k)*ggr/100.
        o_eu(icrm,k) = 0.
        if (mui_crm(icrm,k) - mui_crm(icrm,k+1) > 0) then
          o_eu(icrm,k) = (mui_crm(icrm,k) - mui_crm(icrm,k+1))*ggr/pdel(icrm,k)
        else
          o_du(icrm,k) = -1.0*(mui_crm(icrm,k) - mui_crm(icrm,k+1))*ggr/pdel(icrm,k)
        end if
        if (mdi_crm(icrm,k+1) - mdi_crm(icrm,k) < 0) then
          o_ed(icrm,k) = (mdi_crm(icrm,k) - mdi_crm(icrm,k+1))*ggr/pdel(icrm,k)
        else
          dd_crm(icrm,k) = -1.*(mdi_crm(icrm,k) - mdi_crm(icrm,k+1))*ggr/pdel(icrm,k)
        end if
        if (abs(o_mu(icrm,k)) > 1.0e-15 .or. abs(o_md(icrm,k)) > 1.0e-15) then
          tmp = k
          o_jt(icrm) = min(tmp, o_jt(icrm))
          o_mx(icrm) = max(tmp, o_mx(icrm))
        end if
      end do
    end do
    do k = 1, nzm
      do icrm = 1, ncrms
        u2z = 0.; v2z = 0.; w2z = 0.
        do j = 1, ny
          do i = 1, nx
            u2z = u2z + (u(icrm,i,j,k) - u0(icrm,k))**2
            v2z = v2z + (v(icrm,i,j,k) - v0(icrm,k))**2
            w2z = w2z + 0.5*(w(icrm,i,j,k+1)**2 + w(icrm,i,j,k)**2)
          end do
        end do
        tmp1 = dz(icrm)/rhow(icrm,k)
        tmp2 = tmp1/dtn
        mkwsb(icrm,k,:) = mkwsb(icrm,k,:)*tmp1*rhow(icrm,k)*factor_xy/nstop
        mkwle(icrm,k,:) = mkwle(icrm,k,:)*tmp2*rhow(icrm,k)*factor_xy/nstop
        mkadv(icrm,k,:) = mkadv(icrm,k,:)*factor_xy*icrm_run_time
        mkdiff(icrm,k,:) = mkdiff(icrm,k,:)*factor_xy*icrm_run_time
        qpsrc(icrm,k) = qpsrc(icrm,k)*factor_xy*icrm_run_time
        qpevp(icrm,k) = qpevp(icrm,k)*factor_xy*icrm_run_time
        qpfall(icrm,k) = qpfall(icrm,k)*factor_xy*icrm_run_time
        precflux(icrm,k) = precflux(icrm,k)*factor_xy*dz(icrm)/dt/nstop
        l = plev - k + 1
        o_flux_u(icrm,l) = (uwle(icrm,k) + uwsb(icrm,k))*tmp1*factor_xy/nstop
        o_flux_v(icrm,l) = (vwle(icrm,k) + vwsb(icrm,k))*tmp1*factor_xy/nstop
        o_flux_qt(icrm,l) = mkwle(icrm,k,1) + mkwsb(icrm,k,1)
        o_fluxsgs_qt(icrm,l) = mkwsb(icrm,k,1)
        o_flux_qp(icrm,l) = mkwle(icrm,k,2) + mkwsb(icrm,k,2)
        o_qt_trans(icrm,l) = mkadv(icrm,k,1) + mkdiff(icrm,k,1)
        o_qp_trans(icrm,l) = mkadv(icrm,k,2) + mkdiff(icrm,k,2)
        tmp = 0
        do j = 1, ny
          do i = 1, nx
            tmp = tmp + sgs_field(icrm,i,j,k,1)
          end do
        end do
        o_tkesgsz(icrm,l) = rho(icrm,k)*tmp*factor_xy
        o_tkez(icrm,l) = rho(icrm,k)*0.5*(u2z + v2z*YES3D + w2z)*factor_xy + o_tkesgsz(icrm,l)
        tmp = 0
        do j = 1, ny
          do i = 1, nx
            tmp = tmp + sgs_field_diag(icrm,i,j,k,1)
          end do
        end do
        o_tkz(icrm,l) = tmp*factor_xy
        o_precflux(icrm,l) = precflux(icrm,k)/1000.
        o_qp_fall(icrm,l) = qpfall(icrm,k)
        o_qp_evp(icrm,l) = qpevp(icrm,k)
        o_qp_src(icrm,l) = qpsrc(icrm,k)
        o_qt_ls(icrm,l) = qtend(icrm,k)
        o_t_ls(icrm,l) = ttend(icrm,k)
      end do
    end do
    o_timing(1:ncrms) = o_timing(1:ncrms)/nstop

  contains
    logical function stop_here(id)
      integer, intent(in) :: id
      stop_here = (phase_stop == id .and. nstep == nstop)
    end function
  end function ref_crm

  ! Copy one reference module field (interior points) into buf, laid out [icrm][k][j][i]
  ! (i fastest) for grid fields and [icrm][k] for profiles.  Returns the number of values
  ! written, or -1 for an unknown name.  For debugging phase-by-phase parity.
  function ref_get(cname, buf) result(n) bind(C, name="ref_get")
    use vars
    use grid
    use microphysics, only: micro_field, qn, mkwle, mkwsb, mkadv, mkdiff, qpsrc, qpevp
    use sgs, only: sgs_field, sgs_field_diag
    character(kind=c_char), intent(in) :: cname(*)
    real(c_double), intent(out) :: buf(*)
    integer(c_int) :: n
    character(len=32) :: name
    integer :: i, j, k, icrm, m
    name = ' '
    do i = 1, 32
      if (cname(i) == c_null_char) exit
      name(i:i) = cname(i)
    end do
    n = -1
    if (.not. live) return
    m = 0
    select case (trim(name))
    case ('u');    call grid3(u(:,1:nx,1:ny,1:nzm))
    case ('v');    call grid3(v(:,1:nx,1:ny,1:nzm))
    case ('w');    call grid3(w(:,1:nx,1:ny,1:nz))
    case ('t');    call grid3(t(:,1:nx,1:ny,1:nzm))
    case ('tabs'); call grid3(tabs(:,1:nx,1:ny,1:nzm))
    case ('qv');   call grid3(qv(:,1:nx,1:ny,1:nzm))
    case ('qcl');  call grid3(qcl(:,1:nx,1:ny,1:nzm))
    case ('qci');  call grid3(qci(:,1:nx,1:ny,1:nzm))
    case ('qpl');  call grid3(qpl(:,1:nx,1:ny,1:nzm))
    case ('qpi');  call grid3(qpi(:,1:nx,1:ny,1:nzm))
    case ('p');    call grid3(p(:,1:nx,1:ny,1:nzm))
    case ('qt');   call grid3(micro_field(:,1:nx,1:ny,1:nzm,1))
    case ('qp');   call grid3(micro_field(:,1:nx,1:ny,1:nzm,2))
    case ('qn');   call grid3(qn(:,1:nx,1:ny,1:nzm))
    case ('tke');  call grid3(sgs_field(:,1:nx,1:ny,1:nzm,1))
    case ('tk');   call grid3(sgs_field_diag(:,1:nx,1:ny,1:nzm,1))
    case ('tkh');  call grid3(sgs_field_diag(:,1:nx,1:ny,1:nzm,2))
    case ('misc'); call grid3(misc(:,1:nx,1:ny,1:nz))
    case ('dudt_na'); call grid3(dudt(:,1:nx,1:ny,1:nzm,na))
    case ('dudt_nb'); call grid3(dudt(:,1:nx,1:ny,1:nzm,nb))
    case ('dudt_nc'); call grid3(dudt(:,1:nx,1:ny,1:nzm,nc))
    case ('dwdt_na'); call grid3(dwdt(:,1:nx,1:ny,1:nz,na))
    case ('dwdt_nb'); call grid3(dwdt(:,1:nx,1:ny,1:nz,nb))
    case ('dwdt_nc'); call grid3(dwdt(:,1:nx,1:ny,1:nz,nc))
    case ('dvdt_na'); call grid3(dvdt(:,1:nx,1:ny,1:nzm,na))
    case ('dvdt_nb'); call grid3(dvdt(:,1:nx,1:ny,1:nzm,nb))
    case ('dvdt_nc'); call grid3(dvdt(:,1:nx,1:ny,1:nzm,nc))
    case ('precsfc');  call grid2(precsfc(:,1:nx,1:ny))
    case ('precssfc'); call grid2(precssfc(:,1:nx,1:ny))
    case ('t0');    call prof(t0)
    case ('q0');    call prof(q0)
    case ('qv0');   call prof(qv0)
    case ('tabs0'); call prof(tabs0)
    case ('u0');    call prof(u0)
    case ('v0');    call prof(v0)
    case ('qn0');   call prof(qn0)
    case ('qp0');   call prof(qp0)
    case ('p0');    call prof(p0)
    case ('rho');   call prof(rho)
    case ('rhow');  call prof(rhow)
    case ('adz');   call prof(adz)
    case ('adzw');  call prof(adzw)
    case ('pres');  call prof(pres)
    case ('presi'); call prof(presi)
    case ('z');     call prof(z)
    case ('zi');    call prof(zi)
    case ('bet');   call prof(bet)
    case ('gamaz'); call prof(gamaz)
    case ('ttend'); call prof(ttend)
    case ('qtend'); call prof(qtend)
    case ('utend'); call prof(utend)
    case ('vtend'); call prof(vtend)
    case ('vwle');  call prof(vwle)
    case ('vwsb');  call prof(vwsb)
    case ('uwle');  call prof(uwle)
    case ('uwsb');  call prof(uwsb)
    case ('mkwle1'); call prof(mkwle(:,:,1))
    case ('mkwsb1'); call prof(mkwsb(:,:,1))
    case ('mkadv1'); call prof(mkadv(:,:,1))
    case ('mkdiff1'); call prof(mkdiff(:,:,1))
    case ('qpsrc'); call prof(qpsrc)
    case ('qpevp'); call prof(qpevp)
    case ('qpfall'); call prof(qpfall)
    case ('precflux'); call prof(precflux)
    case ('dz')
      do icrm = 1, size(dz)
        buf(icrm) = dz(icrm)
      end do
      m = size(dz)
    case default
      return
    end select
    n = m
  contains
    subroutine grid3(a)
      real(crm_rknd), intent(in) :: a(:,:,:,:)
      integer :: ii, jj, kk, cc
      do cc = 1, size(a,1)
        do kk = 1, size(a,4)
          do jj = 1, size(a,3)
            do ii = 1, size(a,2)
              m = m + 1
              buf(m) = a(cc,ii,jj,kk)
            end do
          end do
        end do
      end do
    end subroutine
    subroutine grid2(a)
      real(crm_rknd), intent(in) :: a(:,:,:)
      integer :: ii, jj, cc
      do cc = 1, size(a,1)
        do jj = 1, size(a,3)
          do ii = 1, size(a,2)
            m = m + 1
            buf(m) = a(cc,ii,jj)
          end do
        end do
      end do
    end subroutine
    subroutine prof(a)
      real(crm_rknd), intent(in) :: a(:,:)
      integer :: kk, cc
      do cc = 1, size(a,1)
        do kk = 1, size(a,2)
          m = m + 1
          buf(m) = a(cc,kk)
        end do
      end do
    end subroutine
  end function ref_get

  ! The reference's own x-transform (fft.F90:5-153) on one row of n+2 values, for pinning the
  ! restated transform and the dense operator matrices built from it.
  subroutine ref_fft991(a, n, isign) bind(C, name="ref_fft991")
    use fft_mod, only: fft991_crm, fftfax_crm
    integer(c_int), value :: n, isign
    real(c_double), intent(inout) :: a(n+2)
    real(8) :: work(n+2), trigs(3*n/2+1+8)
    integer :: ifax(100)
    call fftfax_crm(n, ifax, trigs)
    call fft991_crm(a, work, trigs, ifax, 1, n+2, n, 1, isign)
  end subroutine

end module crm_ref_harness
