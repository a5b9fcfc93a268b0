! Calls the C ABI through the Fortran facade without a GPU: the reference
! defaults, the argument checks that return before the device is touched,
! the host Gaussian generator and the sizes of the bind(c) structs.  Prints
! one "key value" line per check for the pytest driver to compare.
program facade_check
  use iso_c_binding
  use lmm_facade
  implicit none
  type(lmm_params) :: p, bad
  type(lmm_species_params) :: sp
  type(lmm_inertial_params) :: ip
  type(lmm_gmres_info) :: info2(2)
  type(c_ptr) :: f3(3)
  type(lmm_gmres_info) :: info
  type(c_ptr) :: ctx, b_u(3), x_u(3), a3(3), e3(3)
  integer(c_int) :: rc
  real(c_double) :: z(4), npr

  call lmm_params_default(p)
  print '(a,1x,i0)', 'visc_type', p%visc_type
  print '(a,1x,i0)', 'precon_type', p%precon_type
  print '(a,1x,i0)', 'gmres_max_inner', p%gmres_max_inner
  print '(a,1x,i0)', 'gmres_max_outer', p%gmres_max_outer
  print '(a,1x,i0)', 'stag_mg_nsmooths_bottom', p%stag_mg_nsmooths_bottom
  print '(a,1x,es23.16)', 'gmres_rel_tol', p%gmres_rel_tol
  print '(a,1x,i0)', 'stoch_stress_form', p%stoch_stress_form
  print '(a,1x,i0)', 'sizeof_params', c_sizeof(p)
  print '(a,1x,i0)', 'sizeof_info', c_sizeof(info)

  ! dim 4
  rc = lmm_ctx_create(4_c_int, [8_c_int, 8_c_int, 8_c_int], 1.0_c_double, p, ctx)
  print '(a,1x,i0)', 'create_dim4', rc
  ! a level that coarsens to an odd extent
  rc = lmm_ctx_create(2_c_int, [6_c_int, 8_c_int, 1_c_int], 1.0_c_double, p, ctx)
  print '(a,1x,i0)', 'create_odd', rc
  ! visc_type outside +-1, +-2, +-3
  bad = p
  bad%visc_type = 4
  rc = lmm_ctx_create(2_c_int, [8_c_int, 8_c_int, 1_c_int], 1.0_c_double, bad, ctx)
  print '(a,1x,i0)', 'create_visc4', rc
  ! filtering width 3
  bad = p
  bad%filtering_width = 3
  rc = lmm_ctx_create(3_c_int, [8_c_int, 8_c_int, 8_c_int], 1.0_c_double, bad, ctx)
  print '(a,1x,i0)', 'create_filter3', rc
  ! gmres without a context
  b_u = c_null_ptr
  x_u = c_null_ptr
  a3 = c_null_ptr
  e3 = c_null_ptr
  rc = lmm_gmres(c_null_ptr, b_u, c_null_ptr, x_u, c_null_ptr, a3, c_null_ptr, e3, c_null_ptr, 1.0_c_double, &
       npr, info)
  print '(a,1x,i0)', 'gmres_null_ctx', rc
  ! species knobs: the reference defaults and the null-context rejections
  call lmm_species_params_default(sp, 4_c_int)
  print '(a,1x,i0)', 'species_n', sp%nspecies
  print '(a,1x,i0)', 'chi_iterations', sp%chi_iterations
  print '(a,1x,es23.16)', 'fraction_tolerance', sp%fraction_tolerance
  print '(a,1x,i0,1x,i0,1x,i0)', 'species_flags', sp%is_ideal_mixture, sp%is_nonisothermal, sp%use_lapack
  print '(a,1x,i0)', 'sizeof_species', c_sizeof(sp)
  rc = lmm_ctx_set_species(c_null_ptr, sp)
  print '(a,1x,i0)', 'set_species_null_ctx', rc
  rc = lmm_compute_mass_fluxdiv(c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, a3, &
       1.0_c_double, 0.0_c_double, z)
  print '(a,1x,i0)', 'mass_fluxdiv_null_ctx', rc
  ! inertial step knobs (probin_common defaults) and the null-context rejection
  call lmm_inertial_params_default(ip)
  print '(a,1x,es23.16,1x,es23.16)', 'inertial_coefs', ip%variance_coef_mom, ip%visc_coef
  print '(a,1x,i0,1x,i0,1x,i0)', 'inertial_flags', ip%eta_model, ip%advection_type, ip%noise
  print '(a,1x,i0)', 'sizeof_inertial', c_sizeof(ip)
  f3 = c_null_ptr
  rc = lmm_advance_timestep_inertial(c_null_ptr, f3, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, f3, &
       c_null_ptr, c_null_ptr, f3, c_null_ptr, c_null_ptr, f3, c_null_ptr, c_null_ptr, 1.0_c_double, &
       0.0_c_double, 0_c_int64_t, ip, info2)
  print '(a,1x,i0)', 'inertial_null_ctx', rc
  print '(a,1x,i0,1x,i0)', 'overdamped_knobs', ip%algorithm_type, ip%restart
  rc = lmm_advance_timestep_overdamped(c_null_ptr, f3, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, f3, &
       c_null_ptr, c_null_ptr, f3, c_null_ptr, c_null_ptr, f3, c_null_ptr, c_null_ptr, 1.0_c_double, &
       0.0_c_double, 0_c_int64_t, ip, info2)
  print '(a,1x,i0)', 'overdamped_null_ctx', rc
  ! the host Gaussian stream (MT19937 + Acklam)
  rc = lmm_fill_normal_host(99_c_int32_t, 4_c_int64_t, z)
  print '(a,1x,i0)', 'fill_rc', rc
  print '(a,4(1x,es25.17))', 'normals', z
end program facade_check
