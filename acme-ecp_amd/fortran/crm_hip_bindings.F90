! crm_hip_bindings.F90 -- Fortran ISO_C_BINDING interface to libcrm_hip.so (include/crm_hip.h).
!
! The bind(C) derived types mirror the C structs member for member; every array member is a
! type(c_ptr) to the caller's column-major array (pcols leading), exactly as the Fortran
! crm_input_type / crm_state_type / crm_rad_type / crm_output_type components are laid out
! (components/cam/src/physics/crm/crm_input_module.F90:10-47, crm_state_module.F90:12-56,
! crm_rad_module.F90:17-34, crm_output_module.F90:7-133).
module crm_hip_bindings
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: CRM_OK = 0, CRM_ERR_ARG = 1, CRM_ERR_CFL = 2, &
       CRM_ERR_NXRAD = 3, CRM_ERR_DAMPING = 4, CRM_ERR_NAN = 5, CRM_ERR_DEVICE = 6, &
       CRM_ERR_NOMEM = 7, CRM_ERR_UNSUPPORTED = 8, CRM_ERR_ACCEL = 9, CRM_ERR_SCRATCH = 10
  integer(c_int), parameter, public :: CRM_PREC_FP64 = 0, CRM_PREC_MIXED = 1
  integer(c_int), parameter, public :: CRM_GRAPHS_OFF = 0, CRM_GRAPHS_SUBSTEP = 1, &
                                       CRM_GRAPHS_CALL = 2, CRM_GRAPHS_JOINT = 3

  type, bind(C), public :: crm_dims
    integer(c_int) :: ncrms, ld, plev, nx, ny, nz, nx_rad, ny_rad
    real(c_double) :: dx, dt
  end type

  type, bind(C), public :: crm_column_info
    integer(c_int) :: lchnk
    type(c_ptr)    :: gcol, lat, lon
    integer(c_int) :: igstep
  end type

  type, bind(C), public :: crm_input_view
    type(c_ptr) :: zmid, zint, tl, ql, qccl, qiil, ps, pmid, pint, pdel, phis, ul, vl, &
                   ocnfrac, tau00, wndls, bflxls, fluxu00, fluxv00, fluxt00, fluxq00
  end type

  type, bind(C), public :: crm_state_view
    type(c_ptr) :: u_wind, v_wind, w_wind, temperature, qt, qp, qn
  end type

  type, bind(C), public :: crm_rad_view
    type(c_ptr) :: qrad, temperature, qv, qc, qi, cld
  end type

  type, bind(C), public :: crm_output_view
    type(c_ptr) :: qcl, qci, qpl, qpi, tk, tkh
    type(c_ptr) :: prec_crm
    type(c_ptr) :: cltot, clhgh, clmed, cllow, precc, precl, precsc, precsl
    type(c_ptr) :: jt_crm, mx_crm, prectend, precstend, taux, tauy, z0m, timing_factor
    type(c_ptr) :: cldtop, qc_mean, qi_mean, qs_mean, qg_mean, qr_mean
    type(c_ptr) :: sltend, qltend, qcltend, qiltend
    type(c_ptr) :: cld, gicewp, gliqwp, mctot, mcup, mcdn, mcuup, mcudn
    type(c_ptr) :: mu_crm, md_crm, du_crm, eu_crm, ed_crm
    type(c_ptr) :: flux_qt, fluxsgs_qt, tkez, tkesgsz, tkz, flux_u, flux_v, flux_qp
    type(c_ptr) :: precflux, qt_ls, qt_trans, qp_trans, qp_fall, qp_src, qp_evp, t_ls
  end type

  ! GCM-side glue (include/crm_glue.h): the physics_state members crm_physics_tend reads (q as
  ! its vapour, cloud-liquid and cloud-ice slices) and the cam_in_t members
  type, bind(C), public :: crm_gcm_state_view
    type(c_ptr) :: zm, zi, t, qv, qcl, qci, ps, pmid, pint, pdel, phis, u, v
  end type

  type, bind(C), public :: crm_gcm_surface_view
    type(c_ptr) :: ocnfrac, wsx, wsy, shf, lhf
  end type

  type, bind(C), public :: crm_hip_config
    integer(c_int) :: device, precision, max_ncrms
  end type

  type, bind(C), public :: crm_hip_stats
    integer(c_int) :: nstop, ncycle_max, nprec_max
    real(c_double) :: ms_device
    integer(c_int) :: accel_ceased
  end type

  public :: crm_hip_init, crm_hip_run, crm_hip_run_device, crm_hip_run_device_noalloc, &
            crm_hip_accel_init, crm_hip_get_stats, &
            crm_hip_finalize, crm_hip_errstr, crm_hip_message
  public :: crm_hip_gcm_input, crm_hip_gcm_flux_bypass, crm_hip_gcm_tend, crm_hip_check_energy_chng

  interface
    integer(c_int) function crm_hip_init(ctx, cfg) bind(C, name="crm_hip_init")
      import :: c_int, c_ptr, crm_hip_config
      type(c_ptr), intent(out) :: ctx
      type(crm_hip_config), intent(in) :: cfg
    end function

    integer(c_int) function crm_hip_run(ctx, dims, col, dt_gl, inp, state, rad, out) &
        bind(C, name="crm_hip_run")
      import :: c_int, c_ptr, c_double, crm_dims, crm_column_info, crm_input_view, &
                crm_state_view, crm_rad_view, crm_output_view
      type(c_ptr), value :: ctx
      type(crm_dims), intent(in) :: dims
      type(crm_column_info), intent(in) :: col
      real(c_double), value :: dt_gl
      type(crm_input_view), intent(in) :: inp
      type(crm_state_view), intent(inout) :: state
      type(crm_rad_view), intent(inout) :: rad
      type(crm_output_view), intent(inout) :: out
    end function

    integer(c_int) function crm_hip_run_device(ctx, dims, col, dt_gl, inp, state, rad, out) &
        bind(C, name="crm_hip_run_device")
      import :: c_int, c_ptr, c_double, crm_dims, crm_column_info, crm_input_view, &
                crm_state_view, crm_rad_view, crm_output_view
      type(c_ptr), value :: ctx
      type(crm_dims), intent(in) :: dims
      type(crm_column_info), intent(in) :: col
      real(c_double), value :: dt_gl
      type(crm_input_view), intent(in) :: inp
      type(crm_state_view), intent(inout) :: state
      type(crm_rad_view), intent(inout) :: rad
      type(crm_output_view), intent(inout) :: out
    end function

    ! crm_hip_run_device for callers that cannot allocate device memory: every device buffer of
    ! the call comes from the library's fixed per-GPU array (no counterpart in the reference)
    integer(c_int) function crm_hip_run_device_noalloc(ctx, dims, col, dt_gl, inp, state, rad, out) &
        bind(C, name="crm_hip_run_device_noalloc")
      import :: c_int, c_ptr, c_double, crm_dims, crm_column_info, crm_input_view, &
                crm_state_view, crm_rad_view, crm_output_view
      type(c_ptr), value :: ctx
      type(crm_dims), intent(in) :: dims
      type(crm_column_info), intent(in) :: col
      real(c_double), value :: dt_gl
      type(crm_input_view), intent(in) :: inp
      type(crm_state_view), intent(inout) :: state
      type(crm_rad_view), intent(inout) :: rad
      type(crm_output_view), intent(inout) :: out
    end function

    ! crm_accel_init replacement (accelerate_crm.F90:30-38,45-75)
    integer(c_int) function crm_hip_accel_init(ctx, use_crm_accel, crm_accel_factor, crm_accel_uv) &
        bind(C, name="crm_hip_accel_init")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: use_crm_accel, crm_accel_uv
      real(c_double), value :: crm_accel_factor
    end function

    ! the reference's docoriolis switch (params.F90:107); 3D CRMs only
    integer(c_int) function crm_hip_set_coriolis(ctx, docoriolis) bind(C, name="crm_hip_set_coriolis")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: docoriolis
    end function

    integer(c_int) function crm_hip_set_graphs(ctx, mode) bind(C, name="crm_hip_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function

    integer(c_int) function crm_hip_get_stats(ctx, stats) bind(C, name="crm_hip_get_stats")
      import :: c_int, c_ptr, crm_hip_stats
      type(c_ptr), value :: ctx
      type(crm_hip_stats), intent(out) :: stats
    end function

    integer(c_int) function crm_hip_finalize(ctx) bind(C, name="crm_hip_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function crm_hip_errstr(code) bind(C, name="crm_hip_errstr")
      import :: c_int, c_ptr
      integer(c_int), value :: code
    end function

    ! GCM-side glue on device arrays (crm_physics.F90:1005-1010,1050-1125; :1652-1703;
    ! :1158-1169,1203-1209,1310-1330; check_energy.F90:358-561); pass c_loc of device buffers
    integer(c_int) function crm_hip_gcm_input(ctx, ncol, pcols, pver, crm_nz, nrad, state, srf, crm_angle, &
                                              inp, rad) bind(C, name="crm_hip_gcm_input")
      import :: c_int, c_ptr, crm_gcm_state_view, crm_gcm_surface_view, crm_input_view, crm_rad_view
      type(c_ptr), value :: ctx, crm_angle
      integer(c_int), value :: ncol, pcols, pver, crm_nz, nrad
      type(crm_gcm_state_view), intent(in) :: state
      type(crm_gcm_surface_view), intent(in) :: srf
      type(crm_input_view), intent(in) :: inp
      type(crm_rad_view), intent(in) :: rad
    end function

    integer(c_int) function crm_hip_gcm_flux_bypass(ctx, ncol, pcols, pver, pdel, shf, cflx_vap, ptend_s, &
                                                    ptend_qv) bind(C, name="crm_hip_gcm_flux_bypass")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pdel, shf, cflx_vap, ptend_s, ptend_qv
      integer(c_int), value :: ncol, pcols, pver
    end function

    integer(c_int) function crm_hip_gcm_tend(ctx, ncol, pcols, pver, crm_nz, nrad, pdel, out, rad, qrs, qrl, &
                                             ptend_s, ptend_qv, ptend_qcl, ptend_qci, sp_rad_flux) &
        bind(C, name="crm_hip_gcm_tend")
      import :: c_int, c_ptr, crm_output_view, crm_rad_view
      type(c_ptr), value :: ctx, pdel, qrs, qrl, ptend_s, ptend_qv, ptend_qcl, ptend_qci, sp_rad_flux
      integer(c_int), value :: ncol, pcols, pver, crm_nz, nrad
      type(crm_output_view), intent(inout) :: out
      type(crm_rad_view), intent(in) :: rad
    end function

    integer(c_int) function crm_hip_check_energy_chng(ctx, ncol, pcols, pver, state, ztodt, flx_vap, flx_cnd, &
                                                      flx_ice, flx_sen, te_cur, tw_cur, te_tnd, tw_tnd, &
                                                      te_rer, tw_rer) bind(C, name="crm_hip_check_energy_chng")
      import :: c_int, c_ptr, c_double, crm_gcm_state_view
      type(c_ptr), value :: ctx, flx_vap, flx_cnd, flx_ice, flx_sen, te_cur, tw_cur, te_tnd, tw_tnd, te_rer, tw_rer
      integer(c_int), value :: ncol, pcols, pver
      type(crm_gcm_state_view), intent(in) :: state
      real(c_double), value :: ztodt
    end function
  end interface

contains

  ! The status text of code as a Fortran string (for endrun messages).
  function crm_hip_message(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    interface
      integer(c_size_t) function c_strlen(p) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: p
      end function
    end interface
    type(c_ptr) :: p
    p = crm_hip_errstr(code)
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

end module crm_hip_bindings
