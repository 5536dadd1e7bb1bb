! crm_module_hip.F90 -- drop-in replacement of the reference's crm_module for SP-CAM builds
! that run the cloud-resolving model on an MI355X through libcrm_hip.so.
!
! It keeps the interface of
!   subroutine crm(lchnk, icol, ncrms, dt_gl, plev, crm_input, crm_state, crm_rad,
!                  crm_ecpp_output, crm_output)     (crm/crm_module.F90:45-56, 102-124)
! so crm_physics_tend (cam/crm_physics.F90:1145-1156) calls it unchanged.  The CAM module
! queries crm() makes (get_gcol_p, get_rlat_p, get_rlon_p, get_nstep; crm_module.F90:286-291,
! 677-681) are resolved here on the host and passed as plain arrays; a nonzero status from the
! library becomes endrun with the library's message, where the reference calls endrun /
! task_abort itself (kurant.F90:64-75, crm_module.F90:686-692, damping.F90:39-42).
!
! Built inside an E3SM tree (it uses CAM's ppgrid, phys_grid, time_manager, cam_abortutils
! and the CRM type modules); crm_hip_bindings.F90 holds the ISO_C_BINDING layer and builds
! on its own.  Not covered: CLUBB_CRM, MAML, ECPP statistics (crm_ecpp_output is accepted
! and left untouched).  Mean-state acceleration takes its settings from the physics namelist
! as crm_accel_init does (accelerate_crm.F90:30-38,45-75).
module crm_module
  use iso_c_binding
  use shr_kind_mod,          only: r8 => shr_kind_r8
  use crm_hip_bindings
  use crm_input_module,      only: crm_input_type
  use crm_state_module,      only: crm_state_type
  use crm_rad_module,        only: crm_rad_type
  use crm_output_module,     only: crm_output_type
  use crm_ecpp_output_module, only: crm_ecpp_output_type
  implicit none
  private
  public :: crm

  ! one library context (HIP stream + device buffers) per OpenMP thread of the chunk loop
  type(c_ptr), save :: ctx = c_null_ptr
  !$omp threadprivate(ctx)

contains

  subroutine crm(lchnk, icol, ncrms, dt_gl, plev, crm_input, crm_state, crm_rad, &
                 crm_ecpp_output, crm_output)
    use ppgrid,         only: pcols
    use phys_grid,      only: get_rlat_p, get_rlon_p, get_gcol_p
    use time_manager,   only: get_nstep
    use cam_abortutils, only: endrun
    use crmdims,        only: crm_nx, crm_ny, crm_nz, crm_nx_rad, crm_ny_rad, crm_dx, crm_dt
    use phys_control,   only: phys_getopts

    integer,  intent(in) :: lchnk, ncrms, plev
    integer,  intent(in) :: icol(ncrms)
    real(r8), intent(in) :: dt_gl
    type(crm_input_type),       target, intent(in)    :: crm_input
    type(crm_state_type),       target, intent(inout) :: crm_state
    type(crm_rad_type),         target, intent(inout) :: crm_rad
    type(crm_ecpp_output_type),         intent(inout) :: crm_ecpp_output
    type(crm_output_type),      target, intent(inout) :: crm_output

    type(crm_dims)        :: dims
    type(crm_column_info) :: col
    type(crm_input_view)  :: vin
    type(crm_state_view)  :: vst
    type(crm_rad_view)    :: vrad
    type(crm_output_view) :: vout
    type(crm_hip_config)  :: cfg
    integer(c_int), target :: gcol(ncrms)
    real(c_double), target :: lat(ncrms), lon(ncrms)
    integer(c_int) :: rc
    integer :: i
    logical :: use_crm_accel, crm_accel_uv
    real(r8) :: crm_accel_factor

    if (.not. c_associated(ctx)) then
      cfg = crm_hip_config(0, CRM_PREC_FP64, pcols)
      rc = crm_hip_init(ctx, cfg)
      if (rc /= CRM_OK) call endrun('crm_hip_init: '//crm_hip_message(rc))
      use_crm_accel = .false.
      crm_accel_factor = 0._r8
      crm_accel_uv = .false.
      call phys_getopts(use_crm_accel_out = use_crm_accel, crm_accel_factor_out = crm_accel_factor, &
                        crm_accel_uv_out = crm_accel_uv)
      rc = crm_hip_accel_init(ctx, merge(1_c_int, 0_c_int, use_crm_accel), &
                              real(crm_accel_factor, c_double), merge(1_c_int, 0_c_int, crm_accel_uv))
      if (rc /= CRM_OK) call endrun('crm_hip_accel_init: '//crm_hip_message(rc))
    end if

    do i = 1, ncrms
      gcol(i) = get_gcol_p(lchnk, icol(i))
      lat(i)  = get_rlat_p(lchnk, icol(i))
      lon(i)  = get_rlon_p(lchnk, icol(i))
    end do
    dims = crm_dims(ncrms, pcols, plev, crm_nx, crm_ny, crm_nz, crm_nx_rad, crm_ny_rad, &
                    real(crm_dx, c_double), real(crm_dt, c_double))
    col = crm_column_info(lchnk, c_loc(gcol), c_loc(lat), c_loc(lon), get_nstep())

    vin = crm_input_view(c_loc(crm_input%zmid), c_loc(crm_input%zint), c_loc(crm_input%tl), &
         c_loc(crm_input%ql), c_loc(crm_input%qccl), c_loc(crm_input%qiil), &
         c_loc(crm_input%ps), c_loc(crm_input%pmid), c_loc(crm_input%pint), &
         c_loc(crm_input%pdel), c_loc(crm_input%phis), c_loc(crm_input%ul), &
         c_loc(crm_input%vl), c_loc(crm_input%ocnfrac), c_loc(crm_input%tau00), &
         c_loc(crm_input%wndls), c_loc(crm_input%bflxls), c_loc(crm_input%fluxu00), &
         c_loc(crm_input%fluxv00), c_loc(crm_input%fluxt00), c_loc(crm_input%fluxq00))
    vst = crm_state_view(c_loc(crm_state%u_wind), c_loc(crm_state%v_wind), &
         c_loc(crm_state%w_wind), c_loc(crm_state%temperature), c_loc(crm_state%qt), &
         c_loc(crm_state%qp), c_loc(crm_state%qn))
    vrad = crm_rad_view(c_loc(crm_rad%qrad), c_loc(crm_rad%temperature), c_loc(crm_rad%qv), &
         c_loc(crm_rad%qc), c_loc(crm_rad%qi), c_loc(crm_rad%cld))
    vout = crm_output_view(c_loc(crm_output%qcl), c_loc(crm_output%qci), &
         c_loc(crm_output%qpl), c_loc(crm_output%qpi), c_loc(crm_output%tk), &
         c_loc(crm_output%tkh), c_loc(crm_output%prec_crm), &
         c_loc(crm_output%cltot), c_loc(crm_output%clhgh), c_loc(crm_output%clmed), &
         c_loc(crm_output%cllow), c_loc(crm_output%precc), c_loc(crm_output%precl), &
         c_loc(crm_output%precsc), c_loc(crm_output%precsl), c_loc(crm_output%jt_crm), &
         c_loc(crm_output%mx_crm), c_loc(crm_output%prectend), c_loc(crm_output%precstend), &
         c_loc(crm_output%taux), c_loc(crm_output%tauy), c_loc(crm_output%z0m), &
         c_loc(crm_output%timing_factor), c_loc(crm_output%cldtop), &
         c_loc(crm_output%qc_mean), c_loc(crm_output%qi_mean), c_loc(crm_output%qs_mean), &
         c_loc(crm_output%qg_mean), c_loc(crm_output%qr_mean), c_loc(crm_output%sltend), &
         c_loc(crm_output%qltend), c_loc(crm_output%qcltend), c_loc(crm_output%qiltend), &
         c_loc(crm_output%cld), c_loc(crm_output%gicewp), c_loc(crm_output%gliqwp), &
         c_loc(crm_output%mctot), c_loc(crm_output%mcup), c_loc(crm_output%mcdn), &
         c_loc(crm_output%mcuup), c_loc(crm_output%mcudn), c_loc(crm_output%mu_crm), &
         c_loc(crm_output%md_crm), c_loc(crm_output%du_crm), c_loc(crm_output%eu_crm), &
         c_loc(crm_output%ed_crm), c_loc(crm_output%flux_qt), c_loc(crm_output%fluxsgs_qt), &
         c_loc(crm_output%tkez), c_loc(crm_output%tkesgsz), c_loc(crm_output%tkz), &
         c_loc(crm_output%flux_u), c_loc(crm_output%flux_v), c_loc(crm_output%flux_qp), &
         c_loc(crm_output%precflux), c_loc(crm_output%qt_ls), c_loc(crm_output%qt_trans), &
         c_loc(crm_output%qp_trans), c_loc(crm_output%qp_fall), c_loc(crm_output%qp_src), &
         c_loc(crm_output%qp_evp), c_loc(crm_output%t_ls))

    rc = crm_hip_run(ctx, dims, col, real(dt_gl, c_double), vin, vst, vrad, vout)
    if (rc /= CRM_OK) call endrun('crm (HIP): '//crm_hip_message(rc))
  end subroutine crm

end module crm_module
