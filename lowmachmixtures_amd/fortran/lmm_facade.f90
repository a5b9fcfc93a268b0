! Fortran side of the drop-in boundary: iso_c_binding interfaces of the
! C ABI (include/lmm.h) and wrappers that keep the reference's routine names
! and argument order for the hot path of BoxLib-Codes/LowMachMixtures:
!
!   gmres                  src_gmres/gmres.f90:30-45
!   apply_matrix           src_gmres/apply_matrix.f90:23-39
!   apply_precon           src_gmres/apply_precon.f90:29-43
!   stag_mg_solver         src_gmres/stag_mg_solver.f90:37-50
!   macproject             src_gmres/macproject.f90:34-45
!   stochastic_m_fluxdiv   src_common/stochastic_m_fluxdiv.f90:41-51
!   diffusive_mass_fluxdiv src_multiSpec/diffusive_mass_fluxdiv.f90:26-42
!   diffusive_m_fluxdiv    src_common/diffusive_m_fluxdiv.f90:16-17
!
! A multifab of the reference becomes a device-resident field handle
! (type(c_ptr)); ml_layout, bc_tower, dx and the probin_* globals become the
! context handle.  lmm_fab_in / lmm_fab_out move one component of one box
! between the reference's fab (dataptr(mf, box, comp, 1): lower bound lo-ng,
! column-major, nodal hi+1 faces and ghosts included) and a field.
!
! Errors: the reference aborts through bl_error; every wrapper stops the
! program with the library's message when the C call returns a nonzero
! status.  Non-convergence of GMRES is a printed warning, as in gmres.f90.
module lmm_facade

  use iso_c_binding
  implicit none

  private

  ! status codes (enum lmm_status)
  integer(c_int), parameter, public :: LMM_OK = 0, LMM_ERR_ARG = 1, LMM_ERR_UNSUPPORTED = 2, &
       LMM_ERR_BREAKDOWN = 3, LMM_ERR_DEVICE = 4, LMM_ERR_NOMEM = 5, LMM_NOT_CONVERGED = 6
  ! field kinds (enum lmm_kind)
  integer(c_int), parameter, public :: LMM_CELL = 0, LMM_FACE_X = 1, LMM_FACE_Y = 2, LMM_FACE_Z = 3, &
       LMM_NODE = 4, LMM_EDGE_XY = 5, LMM_EDGE_XZ = 6, LMM_EDGE_YZ = 7
  ! precision forms (enum lmm_precision)
  integer(c_int), parameter, public :: LMM_FP64 = 0, LMM_FP32_MG = 1, LMM_LEARNED = 2

  ! lmm_params: the probin_gmres knobs (probin_gmres.f90:119-157) and the
  ! probin_common ones the path reads
  type, bind(c), public :: lmm_params
     integer(c_int) :: visc_type, precon_type
     real(c_double) :: p_norm_weight, scale_factor
     integer(c_int) :: mg_max_vcycles, mg_minwidth
     integer(c_int) :: mg_nsmooths_down, mg_nsmooths_up, mg_nsmooths_bottom
     real(c_double) :: mg_rel_tol
     integer(c_int) :: stag_mg_max_vcycles, stag_mg_minwidth
     integer(c_int) :: stag_mg_nsmooths_down, stag_mg_nsmooths_up, stag_mg_nsmooths_bottom
     integer(c_int) :: stag_mg_smoother
     real(c_double) :: stag_mg_omega, stag_mg_rel_tol
     real(c_double) :: gmres_rel_tol, gmres_abs_tol
     integer(c_int) :: gmres_max_outer, gmres_max_inner, gmres_max_iter, gmres_min_iter
     integer(c_int) :: gmres_verbose
     integer(c_int) :: precision
     integer(c_int) :: stoch_stress_form, filtering_width
  end type lmm_params

  type, bind(c), public :: lmm_gmres_info
     integer(c_int) :: total_iter, outer_iter
     real(c_double) :: norm_pre_rhs, norm_resid, norm_b, norm_resid_true
     integer(c_int) :: status, vcycles
     real(c_double) :: seconds
     integer(c_int) :: learned_applied, learned_fallback
  end type lmm_gmres_info

  ! lmm_species_params: the probin_multispecies knobs (probin_multispecies.f90:
  ! 11-90) and probin_common's molmass, rhobar, k_B, variance_coef_mass,
  ! barodiffusion_type
  type, bind(c), public :: lmm_species_params
     integer(c_int) :: nspecies
     real(c_double) :: molmass(8), rhobar(8)
     real(c_double) :: dbar(28), dtherm(8), h_offdiag(28), h_diag(8)
     real(c_double) :: k_b, variance_coef_mass, fraction_tolerance
     integer(c_int) :: is_ideal_mixture, is_nonisothermal, use_lapack, inverse_type, chi_iterations
     integer(c_int) :: correct_flux, barodiffusion_type
  end type lmm_species_params

  ! lmm_inertial_params: probin_common's grav, variance_coef_mom, visc_coef,
  ! advection_type (probin_common.f90:209-227), the eta model of
  ! compute_eta_kappa and the noise stream of the time step
  type, bind(c), public :: lmm_inertial_params
     real(c_double) :: grav(3), variance_coef_mom, visc_coef
     integer(c_int) :: eta_model, advection_type, noise, algorithm_type
     integer(c_int64_t) :: noise_seed
     integer(c_int) :: restart, reserved
  end type lmm_inertial_params

  public :: lmm_params_default, lmm_ctx_create, lmm_ctx_destroy, lmm_ctx_set_params, lmm_ctx_last_error, &
       lmm_field_create, lmm_field_destroy, lmm_field_upload_fab, lmm_field_download_fab, lmm_fill_normal_host, &
       lmm_gmres, lmm_apply_matrix, lmm_apply_precon, lmm_stag_mg_solver, lmm_macproject, &
       lmm_init_m_stochastic, lmm_fill_m_stochastic, lmm_stochastic_m_fluxdiv, lmm_diffusive_mass_fluxdiv, &
       lmm_diffusive_m_fluxdiv, lmm_species_params_default, lmm_ctx_set_species, lmm_init_mass_stochastic, &
       lmm_fill_mass_stochastic, lmm_stochastic_mass_fluxdiv, lmm_compute_mass_fluxdiv, lmm_compute_div_comps, &
       lmm_inertial_params_default, lmm_advance_timestep_inertial, lmm_mk_advective_s_fluxdiv, &
       lmm_mk_advective_m_fluxdiv, lmm_mk_grav_force, lmm_advance_timestep_overdamped
  public :: lmm_check, lmm_error_message, lmm_fab_in, lmm_fab_out
  public :: gmres, apply_matrix, apply_precon, stag_mg_solver, macproject, stochastic_m_fluxdiv, &
       diffusive_mass_fluxdiv, diffusive_m_fluxdiv, compute_mass_fluxdiv, stochastic_mass_fluxdiv, &
       init_mass_stochastic, fill_mass_stochastic, compute_div, advance_timestep_inertial, &
       mk_advective_s_fluxdiv, mk_advective_m_fluxdiv, mk_grav_force, advance_timestep_overdamped

  interface
     subroutine lmm_params_default(p) bind(c, name='lmm_params_default')
       import :: lmm_params
       type(lmm_params), intent(out) :: p
     end subroutine lmm_params_default

     integer(c_int) function lmm_ctx_create(dim, n_cells, dx, p, ctx) bind(c, name='lmm_ctx_create')
       import :: c_int, c_double, c_ptr, lmm_params
       integer(c_int), value :: dim
       integer(c_int), intent(in) :: n_cells(3)
       real(c_double), value :: dx
       type(lmm_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function lmm_ctx_create

     integer(c_int) function lmm_ctx_destroy(ctx) bind(c, name='lmm_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function lmm_ctx_destroy

     integer(c_int) function lmm_ctx_set_params(ctx, p) bind(c, name='lmm_ctx_set_params')
       import :: c_int, c_ptr, lmm_params
       type(c_ptr), value :: ctx
       type(lmm_params), intent(in) :: p
     end function lmm_ctx_set_params

     type(c_ptr) function lmm_ctx_last_error(ctx) bind(c, name='lmm_ctx_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function lmm_ctx_last_error

     integer(c_int) function lmm_field_create(ctx, kind, ncomp, f) bind(c, name='lmm_field_create')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: kind, ncomp
       type(c_ptr), intent(out) :: f
     end function lmm_field_create

     integer(c_int) function lmm_field_destroy(f) bind(c, name='lmm_field_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: f
     end function lmm_field_destroy

     integer(c_int) function lmm_field_upload_fab(f, comp, host, lo, hi, ng) bind(c, name='lmm_field_upload_fab')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: f
       integer(c_int), value :: comp, ng
       real(c_double), intent(in) :: host(*)
       integer(c_int), intent(in) :: lo(3), hi(3)
     end function lmm_field_upload_fab

     integer(c_int) function lmm_field_download_fab(f, comp, host, lo, hi, ng) bind(c, name='lmm_field_download_fab')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: f
       integer(c_int), value :: comp, ng
       real(c_double), intent(inout) :: host(*)
       integer(c_int), intent(in) :: lo(3), hi(3)
     end function lmm_field_download_fab

     integer(c_int) function lmm_fill_normal_host(seed, count, out) bind(c, name='lmm_fill_normal_host')
       import :: c_int, c_int32_t, c_int64_t, c_double
       integer(c_int32_t), value :: seed
       integer(c_int64_t), value :: count
       real(c_double), intent(out) :: out(*)
     end function lmm_fill_normal_host

     integer(c_int) function lmm_gmres(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha, &
          norm_pre_rhs, info) bind(c, name='lmm_gmres')
       import :: c_int, c_double, c_ptr, lmm_gmres_info
       type(c_ptr), value :: ctx, b_p, x_p, beta, gamma
       type(c_ptr), intent(in) :: b_u(3), x_u(3), alpha_fc(3), beta_ed(3)
       real(c_double), value :: theta_alpha
       real(c_double), intent(out) :: norm_pre_rhs
       type(lmm_gmres_info), intent(out) :: info
     end function lmm_gmres

     integer(c_int) function lmm_apply_matrix(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, &
          theta_alpha) bind(c, name='lmm_apply_matrix')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, b_p, x_p, beta, gamma
       type(c_ptr), intent(in) :: b_u(3), x_u(3), alpha_fc(3), beta_ed(3)
       real(c_double), value :: theta_alpha
     end function lmm_apply_matrix

     integer(c_int) function lmm_apply_precon(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, &
          theta_alpha) bind(c, name='lmm_apply_precon')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, b_p, x_p, beta, gamma
       type(c_ptr), intent(in) :: b_u(3), x_u(3), alpha_fc(3), beta_ed(3)
       real(c_double), value :: theta_alpha
     end function lmm_apply_precon

     integer(c_int) function lmm_stag_mg_solver(ctx, alpha_fc, beta_cc, beta_ed, gamma_cc, theta_alpha, phi_fc, &
          rhs_fc) bind(c, name='lmm_stag_mg_solver')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, beta_cc, gamma_cc
       type(c_ptr), intent(in) :: alpha_fc(3), beta_ed(3), phi_fc(3), rhs_fc(3)
       real(c_double), value :: theta_alpha
     end function lmm_stag_mg_solver

     integer(c_int) function lmm_macproject(ctx, phi, umac, alphainv_fc, mac_rhs) bind(c, name='lmm_macproject')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, phi, mac_rhs
       type(c_ptr), intent(in) :: umac(3), alphainv_fc(3)
     end function lmm_macproject

     integer(c_int) function lmm_init_m_stochastic(ctx, n_rngs) bind(c, name='lmm_init_m_stochastic')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_rngs
     end function lmm_init_m_stochastic

     integer(c_int) function lmm_fill_m_stochastic(ctx, seed) bind(c, name='lmm_fill_m_stochastic')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: seed
     end function lmm_fill_m_stochastic

     integer(c_int) function lmm_stochastic_m_fluxdiv(ctx, stoch_m_force, eta, eta_ed, temperature, &
          temperature_ed, dt, weights, variance_coef_mom, k_b) bind(c, name='lmm_stochastic_m_fluxdiv')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, eta, temperature
       type(c_ptr), intent(in) :: stoch_m_force(3), eta_ed(3), temperature_ed(3)
       real(c_double), value :: dt, variance_coef_mom, k_b
       real(c_double), intent(in) :: weights(*)
     end function lmm_stochastic_m_fluxdiv

     integer(c_int) function lmm_diffusive_mass_fluxdiv(ctx, nspecies, rho, rhotot, molarconc, rhowchi, gama, &
          diff_fluxdiv, temp, zeta_by_temp, gradp_baro, flux_total, correct_flux) &
          bind(c, name='lmm_diffusive_mass_fluxdiv')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, rho, rhotot, molarconc, rhowchi, gama, diff_fluxdiv, temp, zeta_by_temp
       type(c_ptr), value :: gradp_baro
       type(c_ptr), intent(in) :: flux_total(3)
       integer(c_int), value :: nspecies, correct_flux
     end function lmm_diffusive_mass_fluxdiv

     integer(c_int) function lmm_diffusive_m_fluxdiv(ctx, m_update, umac, eta, eta_ed, kappa) &
          bind(c, name='lmm_diffusive_m_fluxdiv')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, eta, kappa
       type(c_ptr), intent(in) :: m_update(3), umac(3), eta_ed(3)
     end function lmm_diffusive_m_fluxdiv

     subroutine lmm_species_params_default(p, nspecies) bind(c, name='lmm_species_params_default')
       import :: lmm_species_params, c_int
       type(lmm_species_params), intent(out) :: p
       integer(c_int), value :: nspecies
     end subroutine lmm_species_params_default

     integer(c_int) function lmm_ctx_set_species(ctx, p) bind(c, name='lmm_ctx_set_species')
       import :: c_int, c_ptr, lmm_species_params
       type(c_ptr), value :: ctx
       type(lmm_species_params), intent(in) :: p
     end function lmm_ctx_set_species

     integer(c_int) function lmm_init_mass_stochastic(ctx, n_rngs) bind(c, name='lmm_init_mass_stochastic')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_rngs
     end function lmm_init_mass_stochastic

     integer(c_int) function lmm_fill_mass_stochastic(ctx, seed) bind(c, name='lmm_fill_mass_stochastic')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: seed
     end function lmm_fill_mass_stochastic

     integer(c_int) function lmm_stochastic_mass_fluxdiv(ctx, rho, rhotot, molarconc, molmtot, chi, gama, &
          stoch_fluxdiv, flux_total, dt, weights, increment) bind(c, name='lmm_stochastic_mass_fluxdiv')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, rho, rhotot, molarconc, molmtot, chi, gama, stoch_fluxdiv
       type(c_ptr), intent(in) :: flux_total(3)
       real(c_double), value :: dt
       real(c_double), intent(in) :: weights(*)
       integer(c_int), value :: increment
     end function lmm_stochastic_mass_fluxdiv

     integer(c_int) function lmm_compute_mass_fluxdiv(ctx, rho, gradp_baro, diff_fluxdiv, stoch_fluxdiv, temp, &
          flux_total, dt, stage_time, weights) bind(c, name='lmm_compute_mass_fluxdiv')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, rho, diff_fluxdiv, stoch_fluxdiv, temp
       type(c_ptr), value :: gradp_baro
       type(c_ptr), intent(in) :: flux_total(3)
       real(c_double), value :: dt, stage_time
       real(c_double), intent(in) :: weights(*)
     end function lmm_compute_mass_fluxdiv

     integer(c_int) function lmm_compute_div_comps(ctx, phi_fc, div, start_incomp, start_outcomp, num_comp, &
          increment) bind(c, name='lmm_compute_div_comps')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, div
       type(c_ptr), intent(in) :: phi_fc(3)
       integer(c_int), value :: start_incomp, start_outcomp, num_comp, increment
     end function lmm_compute_div_comps

     subroutine lmm_inertial_params_default(p) bind(c, name='lmm_inertial_params_default')
       import :: lmm_inertial_params
       type(lmm_inertial_params), intent(out) :: p
     end subroutine lmm_inertial_params_default

     integer(c_int) function lmm_advance_timestep_inertial(ctx, umac, rho_old, rho_new, rhotot_old, rhotot_new, &
          gradp_baro, pi, eta, eta_ed, kappa, temp, temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, time, &
          istep, p, info) bind(c, name='lmm_advance_timestep_inertial')
       import :: c_int, c_int64_t, c_double, c_ptr, lmm_inertial_params, lmm_gmres_info
       type(c_ptr), value :: ctx, rho_old, rho_new, rhotot_old, rhotot_new, pi, eta, kappa, temp
       type(c_ptr), value :: diff_mass_fluxdiv, stoch_mass_fluxdiv
       type(c_ptr), intent(in) :: umac(3), gradp_baro(3), eta_ed(3), temp_ed(3)
       real(c_double), value :: dt, time
       integer(c_int64_t), value :: istep
       type(lmm_inertial_params), intent(in) :: p
       type(lmm_gmres_info), intent(out) :: info(2)
     end function lmm_advance_timestep_inertial

     integer(c_int) function lmm_advance_timestep_overdamped(ctx, umac, rho_old, rho_new, rhotot_old, rhotot_new, &
          gradp_baro, pi, eta, eta_ed, kappa, temp, temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, time, &
          istep, p, info) bind(c, name='lmm_advance_timestep_overdamped')
       import :: c_int, c_int64_t, c_double, c_ptr, lmm_inertial_params, lmm_gmres_info
       type(c_ptr), value :: ctx, rho_old, rho_new, rhotot_old, rhotot_new, pi, eta, kappa, temp
       type(c_ptr), value :: diff_mass_fluxdiv, stoch_mass_fluxdiv
       type(c_ptr), intent(in) :: umac(3), gradp_baro(3), eta_ed(3), temp_ed(3)
       real(c_double), value :: dt, time
       integer(c_int64_t), value :: istep
       type(lmm_inertial_params), intent(in) :: p
       type(lmm_gmres_info), intent(out) :: info(2)
     end function lmm_advance_timestep_overdamped

     integer(c_int) function lmm_mk_advective_s_fluxdiv(ctx, umac, s_fc, s_update, start_comp, num_comp) &
          bind(c, name='lmm_mk_advective_s_fluxdiv')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, s_update
       type(c_ptr), intent(in) :: umac(3), s_fc(3)
       integer(c_int), value :: start_comp, num_comp
     end function lmm_mk_advective_s_fluxdiv

     integer(c_int) function lmm_mk_advective_m_fluxdiv(ctx, umac, m, m_update) &
          bind(c, name='lmm_mk_advective_m_fluxdiv')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: umac(3), m(3), m_update(3)
     end function lmm_mk_advective_m_fluxdiv

     integer(c_int) function lmm_mk_grav_force(ctx, m_force, s_fc_old, s_fc_new, grav) &
          bind(c, name='lmm_mk_grav_force')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: m_force(3), s_fc_old(3), s_fc_new(3)
       real(c_double), intent(in) :: grav(3)
     end function lmm_mk_grav_force

     integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! the context's last error message as a Fortran string
  function lmm_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = lmm_ctx_last_error(ctx)
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    end do
  end function lmm_error_message

  ! bl_error semantics: a nonzero status stops the program with the message
  subroutine lmm_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= LMM_OK) then
       print '(a,a,i0,a,a)', what, ': status ', rc, ': ', lmm_error_message(ctx)
       error stop 1
    end if
  end subroutine lmm_check

  ! one component of one box of a multifab (dataptr(mf, box, comp, 1)) into a field
  subroutine lmm_fab_in(ctx, f, comp, fp, lo, hi, ng)
    type(c_ptr), intent(in) :: ctx, f
    integer, intent(in) :: comp, lo(3), hi(3), ng
    real(c_double), intent(in) :: fp(*)
    call lmm_check(ctx, lmm_field_upload_fab(f, int(comp - 1, c_int), fp, int(lo, c_int), int(hi, c_int), &
         int(ng, c_int)), 'lmm_fab_in')
  end subroutine lmm_fab_in

  subroutine lmm_fab_out(ctx, f, comp, fp, lo, hi, ng)
    type(c_ptr), intent(in) :: ctx, f
    integer, intent(in) :: comp, lo(3), hi(3), ng
    real(c_double), intent(inout) :: fp(*)
    call lmm_check(ctx, lmm_field_download_fab(f, int(comp - 1, c_int), fp, int(lo, c_int), int(hi, c_int), &
         int(ng, c_int)), 'lmm_fab_out')
  end subroutine lmm_fab_out

  ! gmres.f90:30-45 (theta_alpha and the coefficients are read only here)
  subroutine gmres(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha, norm_pre_rhs)
    type(c_ptr), intent(in) :: ctx, b_u(3), b_p, x_u(3), x_p, alpha_fc(3), beta, beta_ed(3), gamma
    real(c_double), intent(in) :: theta_alpha
    real(c_double), intent(out), optional :: norm_pre_rhs
    type(lmm_gmres_info) :: info
    real(c_double) :: npr
    call lmm_check(ctx, lmm_gmres(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha, npr, info), &
         'gmres')
    if (info%status /= LMM_OK) print '(a,i0,a)', 'GMRES did not converge in ', info%total_iter, ' iterations'
    if (present(norm_pre_rhs)) norm_pre_rhs = npr
  end subroutine gmres

  ! apply_matrix.f90:23-39
  subroutine apply_matrix(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha)
    type(c_ptr), intent(in) :: ctx, b_u(3), b_p, x_u(3), x_p, alpha_fc(3), beta, beta_ed(3), gamma
    real(c_double), intent(in) :: theta_alpha
    call lmm_check(ctx, lmm_apply_matrix(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha), &
         'apply_matrix')
  end subroutine apply_matrix

  ! apply_precon.f90:29-43
  subroutine apply_precon(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha)
    type(c_ptr), intent(in) :: ctx, b_u(3), b_p, x_u(3), x_p, alpha_fc(3), beta, beta_ed(3), gamma
    real(c_double), intent(in) :: theta_alpha
    call lmm_check(ctx, lmm_apply_precon(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha), &
         'apply_precon')
  end subroutine apply_precon

  ! stag_mg_solver.f90:37-50
  subroutine stag_mg_solver(ctx, alpha_fc, beta_cc, beta_ed, gamma_cc, theta_alpha, phi_fc, rhs_fc)
    type(c_ptr), intent(in) :: ctx, alpha_fc(3), beta_cc, beta_ed(3), gamma_cc, phi_fc(3), rhs_fc(3)
    real(c_double), intent(in) :: theta_alpha
    call lmm_check(ctx, lmm_stag_mg_solver(ctx, alpha_fc, beta_cc, beta_ed, gamma_cc, theta_alpha, phi_fc, rhs_fc), &
         'stag_mg_solver')
  end subroutine stag_mg_solver

  ! macproject.f90:34-45
  subroutine macproject(ctx, phi, umac, alphainv_fc, mac_rhs)
    type(c_ptr), intent(in) :: ctx, phi, umac(3), alphainv_fc(3), mac_rhs
    call lmm_check(ctx, lmm_macproject(ctx, phi, umac, alphainv_fc, mac_rhs), 'macproject')
  end subroutine macproject

  ! stochastic_m_fluxdiv.f90:41-51 (dx is the context's; k_B and
  ! variance_coef_mom are probin_common's)
  subroutine stochastic_m_fluxdiv(ctx, stoch_m_force, eta, eta_ed, temperature, temperature_ed, dt, weights, &
       variance_coef_mom, k_b)
    type(c_ptr), intent(in) :: ctx, stoch_m_force(3), eta, eta_ed(3), temperature, temperature_ed(3)
    real(c_double), intent(in) :: dt, weights(:), variance_coef_mom, k_b
    call lmm_check(ctx, lmm_stochastic_m_fluxdiv(ctx, stoch_m_force, eta, eta_ed, temperature, temperature_ed, dt, &
         weights, variance_coef_mom, k_b), 'stochastic_m_fluxdiv')
  end subroutine stochastic_m_fluxdiv

  ! diffusive_mass_fluxdiv.f90:26-42; temp and zeta_by_temp give the
  ! temperature piece (non-isothermal knobs), gradp_baro (a c_loc of three
  ! face handles) the barodiffusion piece; correct_flux is the context's
  subroutine diffusive_mass_fluxdiv(ctx, nspecies, rho, rhotot, molarconc, rhowchi, gama, diff_fluxdiv, flux_total, &
       temp, zeta_by_temp, gradp_baro)
    type(c_ptr), intent(in) :: ctx, rho, rhotot, molarconc, rhowchi, gama, diff_fluxdiv, flux_total(3)
    integer, intent(in) :: nspecies
    type(c_ptr), intent(in), optional :: temp, zeta_by_temp, gradp_baro
    type(c_ptr) :: t, z, g
    t = c_null_ptr
    z = c_null_ptr
    g = c_null_ptr
    if (present(temp)) t = temp
    if (present(zeta_by_temp)) z = zeta_by_temp
    if (present(gradp_baro)) g = gradp_baro
    call lmm_check(ctx, lmm_diffusive_mass_fluxdiv(ctx, int(nspecies, c_int), rho, rhotot, molarconc, rhowchi, gama, &
         diff_fluxdiv, t, z, g, flux_total, 1_c_int), 'diffusive_mass_fluxdiv')
  end subroutine diffusive_mass_fluxdiv

  ! stochastic_mass_fluxdiv.f90:158-188 / :219-246 (periodic: no wall conditions)
  subroutine init_mass_stochastic(ctx, n_rngs)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n_rngs
    call lmm_check(ctx, lmm_init_mass_stochastic(ctx, int(n_rngs, c_int)), 'init_mass_stochastic')
  end subroutine init_mass_stochastic

  subroutine fill_mass_stochastic(ctx, seed)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: seed
    call lmm_check(ctx, lmm_fill_mass_stochastic(ctx, int(seed, c_int32_t)), 'fill_mass_stochastic')
  end subroutine fill_mass_stochastic

  ! stochastic_mass_fluxdiv.f90:32-154 (dx is the context's)
  subroutine stochastic_mass_fluxdiv(ctx, rho, rhotot, molarconc, molmtot, chi, gama, stoch_fluxdiv, flux_total, dt, &
       weights, increment_in)
    type(c_ptr), intent(in) :: ctx, rho, rhotot, molarconc, molmtot, chi, gama, stoch_fluxdiv, flux_total(3)
    real(c_double), intent(in) :: dt, weights(:)
    logical, intent(in), optional :: increment_in
    integer(c_int) :: inc
    inc = 0
    if (present(increment_in)) then
       if (increment_in) inc = 1
    end if
    call lmm_check(ctx, lmm_stochastic_mass_fluxdiv(ctx, rho, rhotot, molarconc, molmtot, chi, gama, stoch_fluxdiv, &
         flux_total, dt, weights, inc), 'stochastic_mass_fluxdiv')
  end subroutine stochastic_mass_fluxdiv

  ! compute_mass_fluxdiv.f90:24-38 (dx and the bc tower are the context's)
  subroutine compute_mass_fluxdiv(ctx, rho, gradp_baro, diff_fluxdiv, stoch_fluxdiv, temp, flux_total, dt, &
       stage_time, weights)
    type(c_ptr), intent(in) :: ctx, rho, gradp_baro, diff_fluxdiv, stoch_fluxdiv, temp, flux_total(3)
    real(c_double), intent(in) :: dt, stage_time, weights(:)
    call lmm_check(ctx, lmm_compute_mass_fluxdiv(ctx, rho, gradp_baro, diff_fluxdiv, stoch_fluxdiv, temp, flux_total, &
         dt, stage_time, weights), 'compute_mass_fluxdiv')
  end subroutine compute_mass_fluxdiv

  ! div_and_grad.f90:304 compute_div(mla,phi_fc,div,dx,start_incomp,start_outcomp,num_comp,increment_in),
  ! components 1-based as in the reference
  subroutine compute_div(ctx, phi_fc, div, start_incomp, start_outcomp, num_comp, increment_in)
    type(c_ptr), intent(in) :: ctx, phi_fc(3), div
    integer, intent(in) :: start_incomp, start_outcomp, num_comp
    logical, intent(in), optional :: increment_in
    integer(c_int) :: inc
    inc = 0
    if (present(increment_in)) then
       if (increment_in) inc = 1
    end if
    call lmm_check(ctx, lmm_compute_div_comps(ctx, phi_fc, div, int(start_incomp - 1, c_int), &
         int(start_outcomp - 1, c_int), int(num_comp, c_int), inc), 'compute_div')
  end subroutine compute_div

  ! advance_timestep_inertial.f90:58-79 advance_timestep_inertial(mla,umac,rho_old,rho_new,
  ! rhotot_old,rhotot_new,gradp_baro,pi,eta,eta_ed,kappa,Temp,Temp_ed,diff_mass_fluxdiv,
  ! stoch_mass_fluxdiv,dx,dt,time,the_bc_tower,istep); the probin_common knobs in p
  subroutine advance_timestep_inertial(ctx, umac, rho_old, rho_new, rhotot_old, rhotot_new, gradp_baro, pi, eta, &
       eta_ed, kappa, temp, temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, time, istep, p, info)
    type(c_ptr), intent(in) :: ctx, umac(3), rho_old, rho_new, rhotot_old, rhotot_new, gradp_baro(3), pi, eta, &
         eta_ed(3), kappa, temp, temp_ed(3), diff_mass_fluxdiv, stoch_mass_fluxdiv
    real(c_double), intent(in) :: dt, time
    integer, intent(in) :: istep
    type(lmm_inertial_params), intent(in) :: p
    type(lmm_gmres_info), intent(out), optional :: info(2)
    type(lmm_gmres_info) :: loc(2)
    call lmm_check(ctx, lmm_advance_timestep_inertial(ctx, umac, rho_old, rho_new, rhotot_old, rhotot_new, &
         gradp_baro, pi, eta, eta_ed, kappa, temp, temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, time, &
         int(istep, c_int64_t), p, loc), 'advance_timestep_inertial')
    if (present(info)) info = loc
  end subroutine advance_timestep_inertial

  ! advance_timestep_overdamped.f90:60-66 advance_timestep_overdamped(mla,umac,rho_old,rho_new,
  ! rhotot_old,rhotot_new,gradp_baro,pi,eta,eta_ed,kappa,Temp,Temp_ed,diff_mass_fluxdiv,
  ! stoch_mass_fluxdiv,dx,dt,time,the_bc_tower,istep); p%algorithm_type is 1 or 2
  subroutine advance_timestep_overdamped(ctx, umac, rho_old, rho_new, rhotot_old, rhotot_new, gradp_baro, pi, eta, &
       eta_ed, kappa, temp, temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, time, istep, p, info)
    type(c_ptr), intent(in) :: ctx, umac(3), rho_old, rho_new, rhotot_old, rhotot_new, gradp_baro(3), pi, eta, &
         eta_ed(3), kappa, temp, temp_ed(3), diff_mass_fluxdiv, stoch_mass_fluxdiv
    real(c_double), intent(in) :: dt, time
    integer, intent(in) :: istep
    type(lmm_inertial_params), intent(in) :: p
    type(lmm_gmres_info), intent(out), optional :: info(2)
    type(lmm_gmres_info) :: loc(2)
    call lmm_check(ctx, lmm_advance_timestep_overdamped(ctx, umac, rho_old, rho_new, rhotot_old, rhotot_new, &
         gradp_baro, pi, eta, eta_ed, kappa, temp, temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, time, &
         int(istep, c_int64_t), p, loc), 'advance_timestep_overdamped')
    if (present(info)) info = loc
  end subroutine advance_timestep_overdamped

  ! mk_advective_s_fluxdiv.f90:14 mk_advective_s_fluxdiv(mla,umac,s_fc,s_update,dx,start_comp,num_comp),
  ! start_comp 1-based as in the reference
  subroutine mk_advective_s_fluxdiv(ctx, umac, s_fc, s_update, start_comp, num_comp)
    type(c_ptr), intent(in) :: ctx, umac(3), s_fc(3), s_update
    integer, intent(in) :: start_comp, num_comp
    call lmm_check(ctx, lmm_mk_advective_s_fluxdiv(ctx, umac, s_fc, s_update, int(start_comp - 1, c_int), &
         int(num_comp, c_int)), 'mk_advective_s_fluxdiv')
  end subroutine mk_advective_s_fluxdiv

  ! mk_advective_m_fluxdiv.f90:16 mk_advective_m_fluxdiv(mla,umac,m,m_update,dx,the_bc_level)
  subroutine mk_advective_m_fluxdiv(ctx, umac, m, m_update)
    type(c_ptr), intent(in) :: ctx, umac(3), m(3), m_update(3)
    call lmm_check(ctx, lmm_mk_advective_m_fluxdiv(ctx, umac, m, m_update), 'mk_advective_m_fluxdiv')
  end subroutine mk_advective_m_fluxdiv

  ! mk_grav_force.f90:17 mk_grav_force(mla,m_force,s_fc_old,s_fc_new,the_bc_tower)
  subroutine mk_grav_force(ctx, m_force, s_fc_old, s_fc_new, grav)
    type(c_ptr), intent(in) :: ctx, m_force(3), s_fc_old(3), s_fc_new(3)
    real(c_double), intent(in) :: grav(3)
    call lmm_check(ctx, lmm_mk_grav_force(ctx, m_force, s_fc_old, s_fc_new, grav), 'mk_grav_force')
  end subroutine mk_grav_force

  ! diffusive_m_fluxdiv.f90:16-17
  subroutine diffusive_m_fluxdiv(ctx, m_update, umac, eta, eta_ed, kappa)
    type(c_ptr), intent(in) :: ctx, m_update(3), umac(3), eta, eta_ed(3), kappa
    call lmm_check(ctx, lmm_diffusive_m_fluxdiv(ctx, m_update, umac, eta, eta_ed, kappa), 'diffusive_m_fluxdiv')
  end subroutine diffusive_m_fluxdiv

end module lmm_facade
