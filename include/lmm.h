/*
 * lmm.h - C ABI of the MI355X (gfx950) GMRES + staggered-multigrid Stokes
 * solver and the fluctuating-flux stencils that feed it.
 *
 * The entry points keep the argument order and meaning of the reference's
 * Fortran routines (BoxLib-Codes/LowMachMixtures); each declaration cites
 * the routine it replaces.  A multifab argument of the reference becomes a
 * device-resident field handle (lmm_field); the ml_layout, bc_tower, dx and
 * the probin_* module globals become the context (lmm_ctx).
 *
 * Data model: one periodic box of n[0] x n[1] x n[2] cells (n[2] = 1 in 2D)
 * per context.  Fields live in HBM, x fastest, without ghost cells; a face
 * component stores the lower face of each cell, the periodic hi+1 face being
 * the same unknown.  lmm_field_upload_fab / lmm_field_download_fab convert
 * from / to the reference's own per-box Fortran arrays (lower bound lo-ng,
 * column-major, hi+1 faces and ghost cells included).
 *
 * Every call returns an lmm_status.  Calls on one context are stream
 * ordered on the context's HIP stream and are not re-entrant; a call that
 * returns data to the host synchronises that stream.
 */
#ifndef LMM_H
#define LMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lmm_ctx lmm_ctx;
typedef struct lmm_field lmm_field;

enum lmm_status {
  LMM_OK = 0,
  LMM_ERR_ARG = 1,          /* bad argument (null, wrong kind, size mismatch) */
  LMM_ERR_UNSUPPORTED = 2,  /* option outside the implemented set */
  LMM_ERR_BREAKDOWN = 3,    /* GMRES orthogonalisation breakdown (gmres.f90:377) */
  LMM_ERR_DEVICE = 4,       /* HIP runtime error */
  LMM_ERR_NOMEM = 5,
  LMM_NOT_CONVERGED = 6     /* reported in lmm_gmres_info only; the call returns LMM_OK */
};

enum lmm_kind {
  LMM_CELL = 0,
  LMM_FACE_X = 1,
  LMM_FACE_Y = 2,
  LMM_FACE_Z = 3,
  LMM_NODE = 4,     /* 2D nodes: node (i,j) is the lower-left corner of cell (i,j) */
  LMM_EDGE_XY = 5,  /* 3D edges, nodal in the two named directions */
  LMM_EDGE_XZ = 6,
  LMM_EDGE_YZ = 7
};

enum lmm_precision {
  LMM_FP64 = 0,      /* the reference algorithm: FP64 left-preconditioned GMRES */
  LMM_FP32_MG = 1,   /* FP32 V-cycles inside an FP64 right-preconditioned flexible GMRES */
  LMM_LEARNED = 2    /* LMM_FP32_MG plus the learned coarse-grid correction stage */
};

/* Solver knobs.  Names, meaning and defaults follow
 * src_gmres/probin_gmres.f90:119-157 and probin_common.f90 (visc_type). */
typedef struct {
  int visc_type;               /* +-1, +-2, +-3 (stag_applyop.f90) */
  int precon_type;             /* 1 = projection preconditioner (apply_precon.f90:103) */
  double p_norm_weight;
  double scale_factor;
  int mg_max_vcycles, mg_minwidth;
  int mg_nsmooths_down, mg_nsmooths_up, mg_nsmooths_bottom;
  double mg_rel_tol;
  int stag_mg_max_vcycles, stag_mg_minwidth;
  int stag_mg_nsmooths_down, stag_mg_nsmooths_up, stag_mg_nsmooths_bottom;
  int stag_mg_smoother;        /* 0 = Jacobi, 1 = 2*dim-colour Gauss-Seidel */
  double stag_mg_omega, stag_mg_rel_tol;
  double gmres_rel_tol, gmres_abs_tol;
  int gmres_max_outer, gmres_max_inner, gmres_max_iter, gmres_min_iter;
  int gmres_verbose;
  int precision;               /* enum lmm_precision */
  /* probin_common.f90 knobs of the stochastic momentum flux */
  int stoch_stress_form;       /* 0 non-symmetric, else symmetric (default 1, probin_common.f90:230) */
  int filtering_width;         /* 0, 1, 2 or 4 (multifab_filter.f90) */
} lmm_params;

typedef struct {
  int total_iter;              /* inner iterations (gmres.f90 total_iter) */
  int outer_iter;
  double norm_pre_rhs;         /* |M^-1 b| (gmres.f90:125-134) */
  double norm_resid;           /* final (preconditioned for LMM_FP64) residual */
  double norm_b;
  double norm_resid_true;      /* last un-preconditioned residual |b - Ax| */
  int status;                  /* LMM_OK or LMM_NOT_CONVERGED */
  int vcycles;                 /* V-cycles run (stag + cell-centred) */
  double seconds;              /* wall time of the call */
  int learned_applied;         /* LMM_LEARNED: preconditioner applications with the learned stage */
  int learned_fallback;        /* LMM_LEARNED: 1 if the solve switched to the plain V-cycle */
} lmm_gmres_info;

/* Kernel forms (no counterpart in the reference).  Every choice but
 * flex_cgs gives the same bits; they exist for measurement, tuning and the
 * tests that compare kernel forms.  flex_cgs changes the rounding of the
 * flexible GMRES (LMM_FP32_MG, LMM_LEARNED) only, never of the LMM_FP64
 * reference algorithm.  A new context takes the defaults, overridden by the
 * LMM_* environment switches present at its creation. */
typedef struct {
  int fused_sweeps;        /* 1: fused red-black sweeps where a level allows; 0: colour kernels */
  int fused_min_tiles;     /* a level below this many workgroups per launch uses the colour kernels */
  int fused_kernel;        /* fused-sweep form: 0 plane marching, 1 block, 2 paired plane marching */
  int fused_target;        /* workgroups the z-slab choice aims for (FP32; FP64 twice) */
  int fused_x3;            /* 1: the three face components in one launch */
  int resid_restrict;      /* 1: residual and restriction in one pass */
  int coarse_one_launch;   /* 1: levels of at most 16^3 cells in one single-workgroup launch */
  int am_brick;            /* z depth of the operator's brick order (0: row order) */
  int z32_direct;          /* 1: FGMRES reads its FP32 preconditioned vectors directly */
  int krylov_fused;        /* 1: fused Gram-Schmidt and basis-scaling passes */
  int flex_cgs;            /* flexible GMRES: 1 classical Gram-Schmidt (two passes per
                              step over the basis, unnormalised stored vectors),
                              0 modified Gram-Schmidt */
  int sweep_prefetch;      /* 1: the FP32 fused sweep loads the next plane's coefficients
                              one step ahead */
  int precon_graph;        /* 1: each FP32 preconditioner application of the flexible
                              GMRES is captured once per output slot into a HIP graph and
                              replayed (same kernels, same bits; fewer host launches);
                              default 0: replays ran no faster than eager launches */
} lmm_kernel_options;

/* No counterpart in the reference: times the context's fine-level fused
 * sweeps (staggered and cell-centred, in the context's precision) for a set
 * of z-slab workgroup targets, 256 .. 2048, on this GPU, and keeps the
 * fastest as the context's fused_target (returned in *fused_target; may be
 * null).  The slab shape never changes results.  The level buffers it runs
 * on are scratch between calls; the call synchronises the context. */
int lmm_ctx_autotune(lmm_ctx *ctx, int *fused_target);

/* ---------------- context ---------------- */

/* Fills p with the reference defaults (probin_gmres.f90:119-157, visc_type 1). */
void lmm_params_default(lmm_params *p);

/* Replaces ml_layout + bc_tower + dx for one periodic box of n_cells cells.
 * Every multigrid level must have an even number of cells per direction
 * (periodic hi+1 faces then coincide with lo faces under coloured sweeps). */
int lmm_ctx_create(int dim, const int n_cells[3], double dx, const lmm_params *p, lmm_ctx **out);
int lmm_ctx_destroy(lmm_ctx *ctx);
int lmm_ctx_set_params(lmm_ctx *ctx, const lmm_params *p);
int lmm_ctx_get_params(const lmm_ctx *ctx, lmm_params *p);
int lmm_ctx_synchronize(lmm_ctx *ctx);
void lmm_kernel_options_default(lmm_kernel_options *o);
int lmm_ctx_get_kernel_options(const lmm_ctx *ctx, lmm_kernel_options *o);
int lmm_ctx_set_kernel_options(lmm_ctx *ctx, const lmm_kernel_options *o);
/* Number of staggered multigrid levels (compute_nlevs_mg, stag_mg_layout.f90:179-222). */
int lmm_ctx_nlevels(const lmm_ctx *ctx, int *nlevs_stag, int *nlevs_cc);
/* Last HIP or argument error message of this context. */
const char *lmm_ctx_last_error(const lmm_ctx *ctx);
/* Kernel timing (no counterpart in the reference).  on = 1: HIP events
 * bracket every launch of the named kernel families on the context stream;
 * on = 2: only the fused red-black sweeps are timed, with events their own
 * dispatch timestamps (hipExtLaunchKernelGGL), which costs the solve almost
 * nothing; on = 0: off.  family: "stag_smooth_l0" / "cc_smooth_l0" (finest
 * level sweeps), "stag_smooth" / "cc_smooth" (coarser levels),
 * "stag_rb_fused", "stag_restrict_l0", "stag_restrict", "stag_rr_fused",
 * "stag_prolong", "stag_coarse", "cc_coarse", "apply_matrix", "krylov". */
int lmm_ctx_set_profiling(lmm_ctx *ctx, int on);
int lmm_ctx_kernel_stats(lmm_ctx *ctx, const char *family, int64_t *launches, double *total_ms,
                         double *algorithmic_bytes);
int lmm_ctx_reset_stats(lmm_ctx *ctx);

/* ---------------- fields (multifab replacement) ---------------- */

int lmm_field_create(lmm_ctx *ctx, int kind, int ncomp, lmm_field **out);
int lmm_field_destroy(lmm_field *f);
int lmm_field_info(const lmm_field *f, int *kind, int *ncomp, int64_t *count_per_comp);
/* Deduplicated host layout: ncomp consecutive blocks of nx*ny*nz doubles. */
int lmm_field_upload(lmm_field *f, const double *host);
int lmm_field_download(const lmm_field *f, double *host);
/* Reference fab layout for component comp (0-based): host array with lower
 * bound lo-ng and upper bound hi+ng (hi includes the nodal +1), column-major.
 * Upload reads the valid region (the hi+1 face must equal the lo face);
 * download writes the valid region, both copies of periodic faces, and fills
 * the ng ghost layers from the periodic images. */
int lmm_field_upload_fab(lmm_field *f, int comp, const double *host, const int lo[3],
                         const int hi[3], int ng);
int lmm_field_download_fab(const lmm_field *f, int comp, double *host, const int lo[3],
                           const int hi[3], int ng);
int lmm_field_setval(lmm_field *f, double v);
int lmm_field_copy(lmm_field *dst, const lmm_field *src);
/* Device address of the field (ncomp blocks of nx*ny*nz doubles). */
double *lmm_field_device_ptr(lmm_field *f);

/* ---------------- operators and solvers ---------------- */

/* gmres.f90:30-45  gmres(mla,the_bc_tower,dx,b_u,b_p,x_u,x_p,alpha_fc,beta,beta_ed,
 *                        gamma,theta_alpha,norm_pre_rhs)
 * x_u, x_p: initial guess in, solution out.  b_u, coefficients are read only
 * (the reference scales and restores them when scale_factor != 1).
 * norm_pre_rhs may be NULL.  info may be NULL. */
int lmm_gmres(lmm_ctx *ctx, lmm_field *const b_u[3], lmm_field *b_p, lmm_field *x_u[3],
              lmm_field *x_p, lmm_field *const alpha_fc[3], lmm_field *beta,
              lmm_field *const beta_ed[3], lmm_field *gamma, double theta_alpha,
              double *norm_pre_rhs, lmm_gmres_info *info);

/* One member of a batched solve: the arguments of lmm_gmres for one
 * context, and its results. */
typedef struct {
  lmm_ctx *ctx;
  lmm_field *b_u[3], *b_p, *x_u[3], *x_p;
  lmm_field *alpha_fc[3], *beta, *beta_ed[3], *gamma;
  double theta_alpha;
  double norm_pre_rhs;   /* out */
  lmm_gmres_info info;   /* out */
  int status;            /* out: what lmm_gmres returned for this member */
} lmm_gmres_member;

/* Ensembles of independent solves (the reference runs one gmres call per
 * ensemble member, each on its own MPI ranks): n members, each with its own
 * context (contexts must be distinct), solved concurrently by up to
 * nthreads host threads (0: one per member, at most 16), each driving its
 * context's stream, so the members' kernels and host round trips overlap on
 * the GPU.  Every member's result equals lmm_gmres on it alone.  Returns
 * LMM_OK when every member did, else the first member's nonzero status. */
int lmm_gmres_batched(int n, lmm_gmres_member *members, int nthreads);

/* apply_matrix.f90:23-39  apply_matrix(mla,b_u,b_p,x_u,x_p,alpha_fc,beta,beta_ed,gamma,
 *                                       theta_alpha,dx,the_bc_tower)
 * [b_u; b_p] = [A G; -D 0][x_u; x_p], A = theta_alpha*alpha - L_beta. */
int lmm_apply_matrix(lmm_ctx *ctx, lmm_field *b_u[3], lmm_field *b_p, lmm_field *const x_u[3],
                     lmm_field *x_p, lmm_field *const alpha_fc[3], lmm_field *beta,
                     lmm_field *const beta_ed[3], lmm_field *gamma, double theta_alpha);

/* apply_precon.f90:29-43  apply_precon(mla,b_u,b_p,x_u,x_p,alpha_fc,beta,beta_ed,gamma,
 *                                       theta_alpha,dx,the_bc_tower)
 * x = M^{-1} b with the projection preconditioner (precon_type 1). */
int lmm_apply_precon(lmm_ctx *ctx, lmm_field *const b_u[3], lmm_field *b_p, lmm_field *x_u[3],
                     lmm_field *x_p, lmm_field *const alpha_fc[3], lmm_field *beta,
                     lmm_field *const beta_ed[3], lmm_field *gamma, double theta_alpha);

/* stag_mg_solver.f90:37-50  stag_mg_solver(mla,la_mg,alpha_fc,beta_cc,beta_ed,gamma_cc,
 *                                          theta_alpha,phi_fc,rhs_fc,dx,the_bc_tower)
 * Solves (theta_alpha*alpha - L)phi = rhs; phi_fc is the initial guess. */
int lmm_stag_mg_solver(lmm_ctx *ctx, lmm_field *const alpha_fc[3], lmm_field *beta_cc,
                       lmm_field *const beta_ed[3], lmm_field *gamma_cc, double theta_alpha,
                       lmm_field *phi_fc[3], lmm_field *const rhs_fc[3]);

/* macproject.f90:34-45  macproject(mla,phi,umac,alphainv_fc,mac_rhs,dx,the_bc_tower)
 * Solves div(alphainv grad phi) = mac_rhs with one cell-centred V-cycle
 * (mg_* knobs); phi is zeroed on entry.  umac feeds only an unused norm in
 * the reference and may be NULL. */
int lmm_macproject(lmm_ctx *ctx, lmm_field *phi, lmm_field *const umac[3],
                   lmm_field *const alphainv_fc[3], lmm_field *mac_rhs);

/* stag_applyop.f90:70-81  stag_applyop_level(la,the_bc_level,phi_fc,Lphi_fc,alpha_fc,beta_cc,
 *                                            beta_ed,gamma_cc,dx,color)
 * alpha_fc is used as given (stag_applyop multiplies by theta_alpha first; pass
 * theta_alpha != 1 to get stag_applyop.f90:20-67). color 0 = every face. */
int lmm_stag_applyop(lmm_ctx *ctx, lmm_field *const phi_fc[3], lmm_field *Lphi_fc[3],
                     lmm_field *const alpha_fc[3], lmm_field *beta_cc,
                     lmm_field *const beta_ed[3], lmm_field *gamma_cc, double theta_alpha,
                     int color);

/* diffusive_m_fluxdiv.f90:16-17  diffusive_m_fluxdiv(mla,m_update,umac,eta,eta_ed,kappa,dx,
 *                                                  the_bc_level)
 * m_update = m_update + L_eta(umac): stag_applyop_level with alpha_fc = 0
 * (visc_type of the context, kappa in the role of gamma), subtracted from
 * m_update (the explicit viscous term of a Crank-Nicolson step).  m_update
 * and umac must be distinct fields. */
int lmm_diffusive_m_fluxdiv(lmm_ctx *ctx, lmm_field *m_update[3], lmm_field *const umac[3], lmm_field *eta,
                            lmm_field *const eta_ed[3], lmm_field *kappa);

/* norm_inner_product.f90:16-21 stag_inner_prod / :266-271 cc_inner_prod
 * prod_val[d] per face component; box-boundary faces weighted 1.5 as in the reference. */
int lmm_stag_inner_prod(lmm_ctx *ctx, lmm_field *const m1[3], lmm_field *const m2[3],
                        double prod_val[3]);
int lmm_cc_inner_prod(lmm_ctx *ctx, lmm_field *m1, lmm_field *m2, double *prod_val);

/* div_and_grad.f90:304-310 compute_div (comp 0 only, increment flag) and
 * div_and_grad.f90:16-19 compute_grad (periodic box). */
int lmm_compute_div(lmm_ctx *ctx, lmm_field *const phi_fc[3], lmm_field *div, int increment);
int lmm_compute_grad(lmm_ctx *ctx, lmm_field *phi, lmm_field *gradp[3]);
/* The same with the reference's component arguments (0-based here):
 * div_and_grad.f90:304 compute_div(mla,phi_fc,div,dx,start_incomp,start_outcomp,num_comp,
 * increment_in) and :16-18 compute_grad(mla,phi,gradp,dx,start_incomp,start_bccomp,
 * start_outcomp,num_comp,the_bc_level) -- on the periodic box the boundary
 * component is unused; the wall-face 2/dx stencils (:139-155) need walls,
 * which this data model does not carry. */
int lmm_compute_div_comps(lmm_ctx *ctx, lmm_field *const phi_fc[3], lmm_field *div, int start_incomp,
                          int start_outcomp, int num_comp, int increment);
int lmm_compute_grad_comps(lmm_ctx *ctx, lmm_field *phi, lmm_field *gradp[3], int start_incomp,
                           int start_outcomp, int num_comp);

/* convert_stag.f90:132 average_cc_to_face, :395 average_cc_to_node (2D), :612
 * average_cc_to_edge (3D) for component 0 of a cell field. */
int lmm_average_cc_to_face(lmm_ctx *ctx, lmm_field *cc, lmm_field *face[3]);
/* convert_stag.f90:132 average_cc_to_face(nlevs,cc,face,start_scomp,start_bccomp,num_comp,...):
 * components [start_comp, start_comp + num_comp) of cc to the same components
 * of the faces (0-based). */
int lmm_average_cc_to_face_comps(lmm_ctx *ctx, lmm_field *cc, lmm_field *face[3], int start_comp,
                                 int num_comp);
int lmm_average_cc_to_node(lmm_ctx *ctx, lmm_field *cc, lmm_field *node);
int lmm_average_cc_to_edge(lmm_ctx *ctx, lmm_field *cc, lmm_field *edge[3]);

/* stochastic_m_fluxdiv.f90:41-51  stochastic_m_fluxdiv(mla,the_bc_level,stoch_m_force,eta,
 *        eta_ed,temperature,temperature_ed,dx,dt,weights)
 * Increments stoch_m_force by div(sqrt(variance_coef_mom*2*k_B/(dV*dt)) sqrt(eta*T) W)
 * where W = sum_r weights[r] * W_r is the noise held by the context
 * (lmm_init_m_stochastic / lmm_fill_m_stochastic).  With filtering_width > 0
 * the scaled cell and node/edge fluxes pass through multifab_filter
 * (multifab_filter.f90:13-200: x, then y, then z) before the divergence
 * (stochastic_m_fluxdiv.f90:166-174, :224-230).  The stress form enters only
 * through the fill. */
int lmm_stochastic_m_fluxdiv(lmm_ctx *ctx, lmm_field *stoch_m_force[3], lmm_field *eta,
                             lmm_field *const eta_ed[3], lmm_field *temperature,
                             lmm_field *const temperature_ed[3], double dt,
                             const double *weights, double variance_coef_mom, double k_B);
/* stochastic_m_fluxdiv.f90:613 init_m_stochastic(mla,n_rngs) */
int lmm_init_m_stochastic(lmm_ctx *ctx, int n_rngs);
/* stochastic_m_fluxdiv.f90:556 fill_m_stochastic(mla): fills every stage with
 * N(0,1) numbers from the host generator (MT19937 seeded with seed, Acklam
 * inverse normal), in the order: stage, W_cc components, then node/edge
 * arrays and their 2 components, x fastest.  With the symmetric stress form
 * (stoch_stress_form != 0, :569-604) W_cc is scaled by sqrt(2) (variance 2)
 * and only component 1 of each node/edge array is drawn; component 2 is a
 * copy of it. */
int lmm_fill_m_stochastic(lmm_ctx *ctx, uint32_t seed);
/* No counterpart in the reference, whose noise comes from the host MT19937
 * stream above: fills every stage of the context's noise on the device, in
 * the same value order, with N(0,1) values of the counter-based stream
 * (seed, step) -- Philox4x32-10 (Salmon et al., SC'11) of the counter
 * (value/2, step) under the key seed, Box-Muller pairs.  For callers that
 * draw noise every time step without a host round trip; each (seed, step)
 * gives an independent, reproducible field. */
int lmm_fill_m_stochastic_device(lmm_ctx *ctx, uint64_t seed, uint64_t step);
/* The same stream into a field: values v0, v0+1, ... in storage order
 * (component, then x fastest).  No counterpart in the reference. */
int lmm_field_fill_normal(lmm_field *f, uint64_t seed, uint64_t step, int64_t v0);
/* Same divergence with caller-provided noise: W_cc has dim components, W_ed
 * has 2 components per node (2D, W_ed[0]) or edge array (3D); filtered as
 * above when the context's filtering_width > 0. */
int lmm_stoch_m_force(lmm_ctx *ctx, lmm_field *stoch_m_force[3], lmm_field *W_cc,
                      lmm_field *const W_ed[3], lmm_field *eta, lmm_field *const eta_ed[3],
                      lmm_field *temperature, lmm_field *const temperature_ed[3], double variance);

/* ---------------- species mass fluxes ----------------
 * The reference keeps these knobs in module globals (probin_multispecies.f90:
 * 11-90, probin_common.f90:31-32,214-220); a context carries one set.  Dbar
 * and H_offdiag list the strict lower triangle by rows, (2,1), (3,1), (3,2),
 * (4,1), ... (compute_mixture_properties.f90:209-230).  Species matrices are
 * fields of nspecies^2 components in Fortran (ns,ns) order (row r, column c
 * in component r + ns*c).  Defaults are the reference's. */
typedef struct {
  int nspecies;               /* 2..8 */
  double molmass[8];          /* 1 */
  double rhobar[8];           /* 1 */
  double Dbar[28];            /* 1 */
  double Dtherm[8];           /* 0 */
  double H_offdiag[28];       /* 0 */
  double H_diag[8];           /* 0 */
  double k_B;                 /* 1 */
  double variance_coef_mass;  /* 1 */
  double fraction_tolerance;  /* 1e-13 */
  int is_ideal_mixture;       /* 1 */
  int is_nonisothermal;       /* 1 */
  int use_lapack;             /* 0: Dbar2chi_iterative and choldc */
  int inverse_type;           /* 1 inverse, 2 pseudo-inverse (use_lapack only) */
  int chi_iterations;         /* 10 */
  int correct_flux;           /* 1 */
  int barodiffusion_type;     /* 0 */
} lmm_species_params;

void lmm_species_params_default(lmm_species_params *p, int nspecies);
int lmm_ctx_set_species(lmm_ctx *ctx, const lmm_species_params *p);
int lmm_ctx_get_species(const lmm_ctx *ctx, lmm_species_params *p);

/* mass_flux_utilities.f90:26-150 correct_rho_with_drho(mla,rho,drho):
 * rho = max(rho, fraction_tolerance*sum(rho)) in place, drho = the change. */
int lmm_correct_rho_with_drho(lmm_ctx *ctx, lmm_field *rho, lmm_field *drho);
/* :152-313 convert_cons_to_prim(mla,rho,rhotot,molarconc,molmtot) */
int lmm_convert_cons_to_prim(lmm_ctx *ctx, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc,
                             lmm_field *molmtot);
/* compute_mixture_properties.f90:24-233 compute_mixture_properties(mla,rho,rhotot,
 * D_bar,D_therm,Hessian,Temp): the constant-coefficient model (prob_type 9's
 * water-glycerol fit is not provided); rho, rhotot, Temp may be NULL. */
int lmm_compute_mixture_properties(lmm_ctx *ctx, lmm_field *rho, lmm_field *rhotot, lmm_field *D_bar,
                                   lmm_field *D_therm, lmm_field *Hessian, lmm_field *Temp);
/* mass_flux_utilities.f90:428-599 compute_Gama(mla,rho,rhotot,molarconc,molmtot,Hessian,Gama);
 * rho, rhotot, molmtot may be NULL (unused by the reference). */
int lmm_compute_Gama(lmm_ctx *ctx, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc,
                     lmm_field *molmtot, lmm_field *Hessian, lmm_field *Gama);
/* :601-927 compute_chi(mla,rho,rhotot,molarconc,chi,D_bar,D_therm,Temp,zeta_by_Temp);
 * zeta_by_Temp is written when is_nonisothermal (then D_therm and Temp are
 * required).  use_lapack computes the inverse (inverse_type 1) or the
 * pseudo-inverse (2) that LAPACK dgetrf/dgetri or gesvd would, by Gauss-Jordan
 * elimination and a one-sided Jacobi SVD: equal to LAPACK to rounding. */
int lmm_compute_chi(lmm_ctx *ctx, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc, lmm_field *chi,
                    lmm_field *D_bar, lmm_field *D_therm, lmm_field *Temp, lmm_field *zeta_by_Temp);
/* :1120-1273 compute_rhoWchi(mla,rho,rhotot,chi,rhoWchi): rhoWchi = -rho_r chi_rc;
 * rhotot may be NULL. */
int lmm_compute_rhoWchi(lmm_ctx *ctx, lmm_field *rho, lmm_field *rhotot, lmm_field *chi, lmm_field *rhoWchi);
/* :929-1118 compute_Lonsager(mla,rho,rhotot,molarconc,molmtot,chi,Gama,Lonsager):
 * Cholesky factor (lower triangle, upper zeroed) of molmtot rho W chi W / k_B,
 * by choldc (matrix_utilities.f90:160-239) or, with use_lapack, the dpotrf
 * factor with its (ns,ns) entry zeroed.  molarconc and Gama may be NULL. */
int lmm_compute_Lonsager(lmm_ctx *ctx, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc,
                         lmm_field *molmtot, lmm_field *chi, lmm_field *Gama, lmm_field *Lonsager);

/* diffusive_mass_fluxdiv.f90:26-42  diffusive_mass_fluxdiv(mla,rho,rhotot,molarconc,rhoWchi,
 *        Gama,diff_fluxdiv,Temp,zeta_by_Temp,gradp_baro,flux_total,dx,the_bc_tower)
 * F = rhoWchi_f (Gama_f grad x [+ zeta_f grad T] [+ baro_f gradp_baro]) per
 * face (:92-262), corrected so the species fluxes sum to zero when
 * correct_flux is nonzero (correction_flux.f90:71-251); flux_total[d] is
 * incremented, diff_fluxdiv overwritten.  The temperature piece enters when
 * the context's species knobs say is_nonisothermal and Temp and
 * zeta_by_Temp are given (both or neither); the barodiffusion piece when
 * barodiffusion_type > 0 (2D only, as compute_baro_coef,
 * mass_flux_utilities.f90:1291-1340; rho,
 * Temp, gradp_baro required).  Without lmm_ctx_set_species the context uses
 * the default knobs for nspecies species.  rhoWchi and Gama hold
 * nspecies^2 components. */
int lmm_diffusive_mass_fluxdiv(lmm_ctx *ctx, int nspecies, lmm_field *rho, lmm_field *rhotot,
                               lmm_field *molarconc, lmm_field *rhoWchi, lmm_field *Gama,
                               lmm_field *diff_fluxdiv, lmm_field *Temp, lmm_field *zeta_by_Temp,
                               lmm_field *const gradp_baro[3], lmm_field *flux_total[3],
                               int correct_flux);

/* stochastic_mass_fluxdiv.f90:158-188 init_mass_stochastic(mla,n_rngs): n_rngs
 * stages of face noise, nspecies components per face direction. */
int lmm_init_mass_stochastic(lmm_ctx *ctx, int n_rngs);
/* :219-246 fill_mass_stochastic(mla,the_bc_level): every stage filled with
 * N(0,1) numbers of the host generator (as lmm_fill_m_stochastic), in the
 * order stage, face direction, species, x fastest (periodic: no wall
 * boundary conditions to apply). */
int lmm_fill_mass_stochastic(lmm_ctx *ctx, uint32_t seed);
/* No counterpart in the reference: the same fill on the device from the
 * counter-based stream of lmm_fill_m_stochastic_device, same value order. */
int lmm_fill_mass_stochastic_device(lmm_ctx *ctx, uint64_t seed, uint64_t step);
/* stochastic_mass_fluxdiv.f90:32-154  stochastic_mass_fluxdiv(mla,rho,rhotot,molarconc,
 *        molmtot,chi,Gama,stoch_fluxdiv,flux_total,dx,dt,weights,the_bc_level,increment)
 * F = sqrt(2 k_B variance_coef_mass/(dV dt)) * L_f * (sum_r weights[r] W_r)
 * with L_f the face average of the cell Cholesky factors of L-Onsager,
 * corrected when correct_flux; flux_total[d] incremented; stoch_fluxdiv
 * overwritten, or incremented when increment is nonzero. */
int lmm_stochastic_mass_fluxdiv(lmm_ctx *ctx, lmm_field *rho, lmm_field *rhotot, lmm_field *molarconc,
                                lmm_field *molmtot, lmm_field *chi, lmm_field *Gama,
                                lmm_field *stoch_fluxdiv, lmm_field *flux_total[3], double dt,
                                const double *weights, int increment);
/* compute_mass_fluxdiv.f90:24-147  compute_mass_fluxdiv(mla,rho,gradp_baro,diff_fluxdiv,
 *        stoch_fluxdiv,Temp,flux_total,dt,stage_time,dx,weights,the_bc_tower)
 * The whole chain in two device passes per flux: rho corrected
 * (correct_rho_with_drho), primitive variables, mixture properties, Gama,
 * chi, rhoWchi, the diffusive flux, the stochastic flux from the context's
 * mass noise (when variance_coef_mass != 0; else stoch_fluxdiv = 0), rho
 * restored as rho_c - drho.  flux_total is overwritten with the sum of both
 * face fluxes.  Temp is required when is_nonisothermal or
 * barodiffusion_type > 0.  stage_time only feeds the reference's
 * manufactured-solution sources (external_source, prob_types 4/5), which
 * this library does not provide. */
int lmm_compute_mass_fluxdiv(lmm_ctx *ctx, lmm_field *rho, lmm_field *const gradp_baro[3],
                             lmm_field *diff_fluxdiv, lmm_field *stoch_fluxdiv, lmm_field *Temp,
                             lmm_field *flux_total[3], double dt, double stage_time,
                             const double *weights);

/* ---------------- inertial time step ----------------
 * advance_timestep_inertial.f90:58-990 (algorithm_type 0) on the periodic
 * box: the predictor Euler step for rho, the Crank-Nicolson predictor solve,
 * the trapezoidal scalar corrector and the corrector solve, each gmres on
 * theta_alpha = 1/dt, alpha_fc = rhotot on faces, beta = eta/2, gamma =
 * kappa/2 with the corrector stopping at max(gmres_abs_tol,
 * norm_pre_rhs * gmres_rel_tol) of the predictor (:480-486).  Centred
 * advection only (advection_type 0: mk_advective_s_fluxdiv and
 * mk_advective_m_fluxdiv; BDS is not built); barodiffusion_type 0 or 2 (3
 * needs compute_HSE_pres, which needs walls).  The knobs the reference reads
 * from probin_common (grav, variance_coef_mom, visc_coef, advection_type)
 * travel here; the species knobs come from lmm_ctx_set_species. */
typedef struct {
  double grav[3];           /* probin_common.f90:209 */
  double variance_coef_mom; /* :213 */
  double visc_coef;         /* :225 */
  int eta_model;            /* compute_eta_3d (compute_mixture_properties.f90:352-406):
                               0 = visc_coef (the default prob_type branch),
                               1 = visc_coef*(-5 + 6 rhotot) (prob_type -12) */
  int advection_type;       /* :227; only 0 */
  int noise;                /* the step's noise fills: 0 = device counter-based stream
                               (seed, 2*istep) for the momentum noise and (seed,
                               2*istep+1) for the mass noise; 1 = host MT19937 seeded
                               with seed + 2*istep and seed + 2*istep + 1 */
  int algorithm_type;       /* probin_common.f90:218; the overdamped step takes 1 or 2 */
  uint64_t noise_seed;
  int restart;              /* probin_common.f90:205; the overdamped step draws no new
                               noise at istep 1 and restart+1 */
  int reserved;
} lmm_inertial_params;

/* probin_common defaults: no gravity, variance_coef_mom 1, visc_coef 1,
 * centred advection, algorithm_type 0, restart -1, device noise with seed 1.
 * The same knobs drive lmm_advance_timestep_overdamped. */
void lmm_inertial_params_default(lmm_inertial_params *p);

/* advance_timestep_inertial.f90:58-79 advance_timestep_inertial(mla,umac,
 * rho_old,rho_new,rhotot_old,rhotot_new,gradp_baro,pi,eta,eta_ed,kappa,Temp,
 * Temp_ed,diff_mass_fluxdiv,stoch_mass_fluxdiv,dx,dt,time,the_bc_tower,istep).
 * rho fields have nspecies components.  On entry diff/stoch_mass_fluxdiv
 * hold the step's starting mass fluxes F^n (as the reference's step 0 or
 * previous step leaves them) and the context holds the noise the step starts
 * with (lmm_init_m_stochastic(ctx, 1) / lmm_init_mass_stochastic(ctx, 1) and
 * a first fill, when the variance coefficients are nonzero).  On return
 * umac, pi, rho_new, rhotot_new, eta, eta_ed, kappa, the mass fluxes and the
 * noise are those of t + dt; gradp_baro is written when barodiffusion_type
 * is 2 (null otherwise).  info (may be null) receives the predictor's and
 * the corrector's gmres_info. */
int lmm_advance_timestep_inertial(lmm_ctx *ctx, lmm_field *umac[3], lmm_field *rho_old,
                                  lmm_field *rho_new, lmm_field *rhotot_old, lmm_field *rhotot_new,
                                  lmm_field *gradp_baro[3], lmm_field *pi, lmm_field *eta,
                                  lmm_field *eta_ed[3], lmm_field *kappa, lmm_field *Temp,
                                  lmm_field *Temp_ed[3], lmm_field *diff_mass_fluxdiv,
                                  lmm_field *stoch_mass_fluxdiv, double dt, double time,
                                  int64_t istep, const lmm_inertial_params *params,
                                  lmm_gmres_info info[2]);
/* advance_timestep_overdamped.f90:60-66 advance_timestep_overdamped(mla,umac,rho_old,
 * rho_new,rhotot_old,rhotot_new,gradp_baro,pi,eta,eta_ed,kappa,Temp,Temp_ed,
 * diff_mass_fluxdiv,stoch_mass_fluxdiv,dx,dt,time,the_bc_tower,istep): the overdamped
 * (theta_alpha = 0) midpoint scheme, algorithm_type 1 or 2 (params->algorithm_type):
 * new noise at the start of every step but the first (istep 1 or restart+1), a
 * predictor Stokes solve with the fluxes at rho^n (dt/2 for algorithm_type 2, noise
 * weights (1, 0)), a half step for rho, a corrector solve with the fluxes at the
 * midpoint state (dt, weights 1 or (1/sqrt 2, 1/sqrt 2)) and the full step for rho.
 * The context's noise needs algorithm_type stages (lmm_init_m_stochastic /
 * lmm_init_mass_stochastic).  Same field arguments and limits as the inertial
 * step; diff/stoch_mass_fluxdiv are outputs. */
int lmm_advance_timestep_overdamped(lmm_ctx *ctx, lmm_field *umac[3], lmm_field *rho_old,
                                    lmm_field *rho_new, lmm_field *rhotot_old, lmm_field *rhotot_new,
                                    lmm_field *gradp_baro[3], lmm_field *pi, lmm_field *eta,
                                    lmm_field *eta_ed[3], lmm_field *kappa, lmm_field *Temp,
                                    lmm_field *Temp_ed[3], lmm_field *diff_mass_fluxdiv,
                                    lmm_field *stoch_mass_fluxdiv, double dt, double time,
                                    int64_t istep, const lmm_inertial_params *params,
                                    lmm_gmres_info info[2]);
/* mk_advective_s_fluxdiv.f90:14-21 mk_advective_s_fluxdiv(mla,umac,s_fc,s_update,dx,
 * start_comp,num_comp): s_update -= div(umac s_fc) for components
 * start_comp .. start_comp+num_comp-1 (0-based). */
int lmm_mk_advective_s_fluxdiv(lmm_ctx *ctx, lmm_field *const umac[3], lmm_field *const s_fc[3],
                               lmm_field *s_update, int start_comp, int num_comp);
/* mk_advective_m_fluxdiv.f90:16-22 mk_advective_m_fluxdiv(mla,umac,m,m_update,dx,
 * the_bc_level): m_update += -div(m u), centred. */
int lmm_mk_advective_m_fluxdiv(lmm_ctx *ctx, lmm_field *const umac[3], lmm_field *const m[3],
                               lmm_field *m_update[3]);
/* mk_grav_force.f90:17-22 mk_grav_force(mla,m_force,s_fc_old,s_fc_new,the_bc_tower):
 * m_force += 0.5 grav (s_fc_old + s_fc_new); grav is probin_common's. */
int lmm_mk_grav_force(lmm_ctx *ctx, lmm_field *m_force[3], lmm_field *const s_fc_old[3],
                      lmm_field *const s_fc_new[3], const double grav[3]);

/* ---------------- learned coarse-grid correction ----------------
 * No counterpart in the reference.  A small bias-free 3-level U-Net
 * (3x3x3 periodic convolutions, channels c1/c2/c3 = 16/32/64 by default,
 * BF16 weights and activations, FP32 accumulation on the gfx950 matrix
 * cores) predicts the error that the projection preconditioner leaves.
 * Precision LMM_LEARNED inserts it into the flexible GMRES as an extra
 * stage after the FP32 preconditioner (apply_precon.f90:103-195, the stage
 * gmres.f90:335 reaches): the Stokes residual v - A z of the preconditioned
 * vector z is restricted to multigrid level `level` (velocity with the
 * 12-point weights, pressure with the 8-cell mean); the network maps
 * [r_u dx^2/s, r_p dx/s, log beta, log theta*alpha] there (s the residual's
 * scale in velocity units) to the error [e_u/s, e_p dx/s]; the prediction,
 * its pressure made mean-free, is prolongated to the finest level and added
 * to z.  The plain preconditioner is the fallback: a step whose Hessenberg
 * column is not finite is redone without the stage, and an inner step that
 * cuts the residual estimate by less than 10 % ends the stage for the rest
 * of the solve (lmm_gmres_info.learned_fallback).  The FGMRES stopping test
 * on the true residual is unchanged, so the solution meets the same
 * tolerance.
 *
 * Training pairs come from the plain FP32 preconditioner of the context on
 * given coefficients: phi* = M(counter-based white noise) (seed, pair
 * index), then sample_steps times phi* <- phi* - M(A phi*), which leaves the
 * modes the preconditioner handles worst (those the flexible GMRES iterates
 * on); b = A phi*, z = M(b); the input is the residual b - A z, the target
 * the error phi* - z, both restricted to `level` (the error with face and
 * cell means).  Pairs stay on the device. */
typedef struct lmm_cnn lmm_cnn;
typedef struct {
  int cin, c1, c2, c3, cout;  /* 8, 16, 32, 64, 4; c1..c3 multiples of 16, cin of 4, cout <= 4 */
  int level;                  /* multigrid level of the correction (default 3) */
  int sample_steps;           /* training samples: applications of the preconditioned error
                                 propagation I - M A to the noise before a sample (default 3) */
} lmm_cnn_config;
void lmm_cnn_config_default(lmm_cnn_config *cfg);
/* He-initialised network (seeded) on the context's stream; cfg may be NULL. */
int lmm_cnn_create(lmm_ctx *ctx, const lmm_cnn_config *cfg, uint64_t seed, lmm_cnn **out);
int lmm_cnn_destroy(lmm_cnn *net);
/* FP32 master weights, layers L0..L10 in order, each W[co][tap][ci] with
 * tap = (dz+1)*9 + (dy+1)*3 + (dx+1). */
int lmm_cnn_num_params(const lmm_cnn *net, int64_t *n);
int lmm_cnn_get_params(lmm_cnn *net, float *host);
int lmm_cnn_set_params(lmm_cnn *net, const float *host);
/* One forward pass on host data: x [nb][nz][ny][nx][cin] (rounded to BF16),
 * y [..][cout] FP32; n = (nx, ny, nz), multiples of 4, nb*(n/4)^3 a multiple of 16. */
int lmm_cnn_forward_host(lmm_cnn *net, int nb, const int n[3], const float *x, float *y);
/* loss = sum((y - t)^2) / (2 * nb * voxels * cout) and its gradient (the
 * flat parameter layout) by the device's backward kernels. */
int lmm_cnn_loss_grad_host(lmm_cnn *net, int nb, const int n[3], const float *x, const float *t, double *loss,
                           float *grad);
/* Appends npairs training pairs generated on the coefficients given
 * (the arguments of lmm_gmres; 3D, the network's level must exist). */
int lmm_cnn_generate_pairs(lmm_cnn *net, lmm_field *const alpha_fc[3], lmm_field *beta,
                           lmm_field *const beta_ed[3], lmm_field *gamma, double theta_alpha, int npairs,
                           uint64_t seed);
int lmm_cnn_num_pairs(const lmm_cnn *net, int64_t *count);
/* steps of Adam (lr) on random batches of `batch` pairs, the last held_out
 * pairs left out; loss_history (steps values, the loss of the targets divided
 * by their RMS over the training pairs) may be NULL.  Training runs on those
 * unit-scale targets; the RMS is folded into the linear output layer after. */
int lmm_cnn_train(lmm_cnn *net, int steps, int batch, double lr, uint64_t seed, int64_t held_out,
                  float *loss_history);
/* Mean over pairs [first, last) of |t - y| / |t| (1 = no better than a zero
 * prediction). */
int lmm_cnn_evaluate(lmm_cnn *net, int64_t first, int64_t last, double *rel_err);
/* The ablation of the learned stage on fresh samples of the plain
 * preconditioner: for npairs samples made as lmm_cnn_generate_pairs makes
 * them (phi* = M(white noise), b = A phi*), the mean over samples of
 * |phi* - z| / |phi*| (velocity and pressure) with z = M(b) (err_plain) and
 * with z corrected by the learned stage (err_learned). */
int lmm_cnn_error_reduction(lmm_cnn *net, lmm_field *const alpha_fc[3], lmm_field *beta,
                            lmm_field *const beta_ed[3], lmm_field *gamma, double theta_alpha, int npairs,
                            uint64_t seed, double *err_plain, double *err_learned);
/* Network used by precision LMM_LEARNED on this context (NULL detaches). */
int lmm_ctx_set_cnn(lmm_ctx *ctx, lmm_cnn *net);

/* No counterpart in the reference.  lmm_diffusive_mass_fluxdiv for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): same arguments and
 * results, bit for bit, but the face fluxes go to a file-scope device array
 * of the library instead of a per-call allocation, so the call makes no
 * hipMalloc / hipFree.  Limit: dim * nspecies * cells <= 4194304 (32 MiB of
 * FP64), else LMM_ERR_ARG.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two contexts' streams, two threads) race. */
int lmm_diffusive_mass_fluxdiv_noalloc(lmm_ctx *ctx, int nspecies, lmm_field *rho, lmm_field *rhotot,
                                       lmm_field *molarconc, lmm_field *rhoWchi, lmm_field *Gama,
                                       lmm_field *diff_fluxdiv, lmm_field *Temp, lmm_field *zeta_by_Temp,
                                       lmm_field *const gradp_baro[3], lmm_field *flux_total[3],
                                       int correct_flux);

/* Host Gaussian generator of the noise fields (RNGs.c:247 genrandn stream:
 * MT19937 init_genrand(seed) + Acklam inverse normal). */
int lmm_fill_normal_host(uint32_t seed, int64_t count, double *out);

#ifdef __cplusplus
}
#endif
#endif
