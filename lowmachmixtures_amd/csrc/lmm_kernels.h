// Host-side launchers of the gfx950 kernels (lmm_kernels.hip).
#pragma once

#include "lmm_common.h"

namespace lmm {

// Device scalars used by Krylov kernels live in one small array per context.
struct DotOut {
  double *partial;   // [nblk_red * 4]
  double *scalars;   // result slots
};

constexpr int RED_BLOCKS = 1024;

// Launch bookkeeping: every launch is followed by post_launch(), which checks
// hipGetLastError() and, when debug_sync is set (environment LMM_DEBUG_SYNC=1),
// synchronises the stream so a failing kernel is named in first_error().
void set_debug_sync(int on);
int debug_sync_on();

// Kernel-choice settings of one context.  Every setting gives the same bits;
// they exist for measurement and for the tests that compare kernel forms.
struct Tuning {
  int rb_variant = 2;   // fused sweep: 0 plane marching, 1 block, 2 paired plane marching
  int rb_target = 768;  // workgroups the z-slab choice aims for (FP32; FP64 twice)
  int am_dz = 32;       // z depth of the operator's brick order (0: row order)
  int rb_prefetch = 0;  // FP32 fused sweep: load the next plane's coefficients one step ahead
  int rb_pairs = 2;     // FP32 three-component sweep: point pairs per thread (2: 64 x 16 tiles)
  int prolong_tiled = 1;  // 3D staggered prolongation from coarse values staged in LDS
  int coarse_split = 1;   // coarse staggered V-cycle: one workgroup per component (visc_type -1 / 1)
  int rr_three = 1;       // fused residual + restriction: the three components in one launch
};
// The settings the launchers on this thread use: those of the context whose
// library call is running on it (TuningScope), else the defaults.
const Tuning &tuning();
struct TuningScope {
  const Tuning *prev;
  explicit TuningScope(const Tuning *t);
  ~TuningScope();
  TuningScope(const TuningScope &) = delete;
  TuningScope &operator=(const TuningScope &) = delete;
};
void post_launch(hipStream_t s, const char *name);
// Kernel-exact timing: a profiling scope may leave a start / stop event pair
// here; the next fused-sweep launch then has the dispatch itself timestamp
// them (hipExtLaunchKernelGGL) and clears the pair.
struct LaunchEvents {
  hipEvent_t start, stop;
};
LaunchEvents *launch_events();
const char *first_error();
void clear_first_error();
// records a host-side launch error (first one wins) for the next status check
void report_error(const char *msg);
constexpr int RED_TPB = 256;

template <typename T>
struct Faces {
  T *c[3];
};

// ---- staggered operator ----
// Fused red-black sweep of one component (3D, visc_type -1, 1, 2, 3): reads
// the old field from pin, writes the new one to pout (lmm_smooth.hip); the
// form and the z-slab target come from tuning().  visc_type 2 and 3 read the
// other components from cur (their state at this component's turn in the
// colour order).
// z-slab thickness for element size es (4 or 8 bytes) when one launch
// sweeps ncomp components, 0: not eligible
int fused_rb_tz(const Grid &g, int min_tiles, int es, int ncomp);
bool fused_rb_ok(const Grid &g, int visc, int min_tiles);
template <typename T>
void launch_stag_rb_fused(hipStream_t s, const Grid &g, int visc, int a, T omega, const T *pin, T *pout,
                          const T *rhs, const StagCoef<T> &K, int min_tiles, const T *const cur[3] = nullptr);
// All three components of one fused sweep in one launch (3D, visc -1 / 1,
// the paired plane-marching form); reads pin[a] (zero_in: a field of zeros,
// pin not read), writes pout[a].
bool fused_rb3_ok(const Grid &g, int visc, int min_tiles);
template <typename T>
void launch_stag_rb_fused3(hipStream_t s, const Grid &g, int visc, T omega, T *const pin[3], T *const pout[3],
                           T *const rhs[3], const StagCoef<T> &K, int min_tiles, bool zero_in);
// Both components of one fused red-black sweep in one launch (2D, visc -1 /
// 1, x a multiple of 64, y of 8, at least min_tiles workgroups); reads pin[a]
// (zero_in: a field of zeros, pin not read), writes pout[a].
bool fused_rb2d_ok(const Grid &g, int visc, int min_tiles);
template <typename T>
void launch_stag_rb2d(hipStream_t s, const Grid &g, int visc, T omega, T *const pin[3], T *const pout[3],
                      T *const rhs[3], const StagCoef<T> &K, bool zero_in);
// One face component of a 2D fused red-black sweep of a visc_type 2 / 3
// level (the components couple, so they go one after the other): reads the
// component from pin and the other component's current state from other.
bool fused_rb2d_comp_ok(const Grid &g, int visc, int min_tiles);
template <typename T>
void launch_stag_rb2d_comp(hipStream_t s, const Grid &g, int visc, int a, T omega, const T *pin, T *pout,
                           const T *rhs, const StagCoef<T> &K, const T *other);
// The cell-centred red-black sweep of one 2D level in one launch (x a
// multiple of 64, y of 8, at least min_tiles workgroups); a null pin is a
// field of zeros.
bool cc_rb2d_ok(const Grid &g, int min_tiles);
template <typename T>
void launch_cc_rb2d(hipStream_t s, const Grid &g, const T *pin, T *pout, const T *rhs, T *const a[3]);
// Residual of every face component and its full-weighting restriction into
// the coarse right-hand side crhs in one pass per component (3D, visc_type
// -1, 1, 2, 3, x a multiple of 64, y of 8); false (nothing launched) elsewhere.
bool stag_resid_restrict_ok(const Grid &g, const Grid &gc, int visc);
template <typename T>
bool launch_stag_resid_restrict(hipStream_t s, const Grid &g, const Grid &gc, int visc, T *const phi[3],
                                T *const rhs[3], T *const crhs[3], const StagCoef<T> &K);
// whether the fused sweeps take a null input as a field of zeros (the
// paired plane-marching form does)
bool fused_rb_takes_zeros();
// The same fused sweep for the cell-centred red-black Gauss-Seidel (3D).
template <typename T>
void launch_cc_rb_fused(hipStream_t s, const Grid &g, const T *pin, T *pout, const T *rhs, T *const a[3],
                        int min_tiles);
template <typename T>
void launch_stag_gs_colour(hipStream_t s, const Grid &g, int visc, int a, int par, T omega,
                           T *const phi[3], const StagCoef<T> &K, const T *rhs);
// Lphi = (alpha - L)phi; sub: Lphi = Lphi - (alpha - L)phi instead
template <typename T>
void launch_stag_applyop(hipStream_t s, const Grid &g, int visc, int color, T *const phi[3],
                         T *const Lphi[3], const StagCoef<T> &K, bool sub = false);
template <typename T>
void launch_stag_jacobi(hipStream_t s, const Grid &g, int visc, T omega, T *const phi[3],
                        T *const rhs[3], T *const tmp[3], const StagCoef<T> &K);
template <typename T>
void launch_stag_residual(hipStream_t s, const Grid &g, int visc, T *const phi[3], T *const rhs[3],
                          T *const res[3], const StagCoef<T> &K);
template <typename T>
void launch_stag_restrict(hipStream_t s, const Grid &gc, T *const fine[3], T *const coarse[3],
                          int simple);
template <typename T>
void launch_stag_prolong(hipStream_t s, const Grid &gf, T *const fine[3], T *const coarse[3]);
template <typename T>
void launch_cc_restrict(hipStream_t s, const Grid &gc, const T *fine, T *coarse);
template <typename T>
void launch_edge_restrict(hipStream_t s, const Grid &gc, T *const fine[3], T *const coarse[3]);
template <typename T>
void launch_nodal_restrict(hipStream_t s, const Grid &gc, const T *fine, T *coarse);

// ---- coarse levels of a V-cycle in one single-workgroup launch (lmm_coarse.hip) ----
constexpr int COARSE_MAX_LEVELS = 12;
// levels with at most this many cells run inside the coarse-level launch
// (one workgroup: above 16^3 the separate wide launches are faster)
constexpr long COARSE_MAX_CELLS = 4096;
template <typename T>
struct CoarseStagLevel {
  Grid g;
  T *phi[3], *rhs[3], *res[3];
  StagCoef<T> K;
};
template <typename T>
struct CoarseCCLevel {
  Grid g;
  T *phi, *rhs, *res;
  T *a[3];
};
// Levels lv[0..m-1] (finest first): zero phi on all of them, then the
// down-leg (nd sweeps, residual, 12-point restriction), nb bottom sweeps on
// lv[m-1] and the up-leg (prolongation, nu sweeps), as stag_mg_solver's
// level loop does; 2*dim-colour Gauss-Seidel.
template <typename T>
void launch_stag_coarse_vcycle(hipStream_t s, int m, const CoarseStagLevel<T> *lv, int visc, T omega, int nd,
                               int nu, int nb);
template <typename T>
void launch_cc_coarse_vcycle(hipStream_t s, int m, const CoarseCCLevel<T> *lv, int nd, int nu, int nb);

// ---- cell-centred multigrid ----
template <typename T>
void launch_cc_rb_colour(hipStream_t s, const Grid &g, int colour, T *phi, const T *rhs,
                         T *const a[3]);
template <typename T>
void launch_cc_residual(hipStream_t s, const Grid &g, const T *phi, const T *rhs, T *res,
                        T *const a[3]);
template <typename T>
void launch_cc_resid_restrict(hipStream_t s, const Grid &gf, const T *phi, const T *rhs, T *const a[3],
                              const Grid &gc, T *coarse_rhs);
template <typename T>
void launch_cc_prolong_add(hipStream_t s, const Grid &gf, T *fine, const T *coarse);

// ---- elementwise / conversions ----
template <typename T>
void launch_fill(hipStream_t s, T *x, long n, T v);
template <typename T>
void launch_scale(hipStream_t s, T *x, long n, T v);
template <typename TD, typename TS>
void launch_convert(hipStream_t s, TD *dst, const TS *src, long n);
void launch_reciprocal(hipStream_t s, double *dst, const double *src, long n);
void launch_scale_copy(hipStream_t s, double *dst, const double *src, long n, double v);

// ---- Stokes pieces (FP64 unless templated) ----
template <typename T>
void launch_apply_matrix(hipStream_t s, const Grid &g, int visc, T *const x_u[3], const T *x_p,
                         T *const b_u[3], T *b_p, const StagCoef<T> &K);
template <typename T>
void launch_div_plus(hipStream_t s, const Grid &g, T *const u[3], const T *b_p, T *div,
                     int mode);  // mode 0: div; 1: div + b_p; 2: div*(-1); 3: div_in + div
template <typename T>
void launch_precon_final(hipStream_t s, const Grid &g, int visc, T *const x_u[3], T *x_p,
                         const T *phi, const T *mac_rhs, T *const ainv[3], const T *beta,
                         const T *gamma, double theta_alpha, double *partial);
void launch_grad(hipStream_t s, const Grid &g, const double *p, double *const gp[3]);
void launch_avg_cc_to_face(hipStream_t s, const Grid &g, const double *cc, double *const face[3]);
void launch_avg_cc_to_node(hipStream_t s, const Grid &g, const double *cc, double *node);
void launch_avg_cc_to_edge(hipStream_t s, const Grid &g, const double *cc, double *const edge[3]);

// ---- reductions (deterministic two-stage) ----
// Segmented weighted dot over the Stokes vector layout [u0|u1|(u2)|p]:
// face segments weight face index 0 along their direction by 3 (the lo and
// hi+1 copies of the reference, each 1 + 0.5).
void launch_stokes_dot(hipStream_t s, const Grid &g, const double *a, const double *b,
                       double *partial);
// Finalize: op 0 -> out[0] = (s0 + s1 (+ s2)) + pnw2*s3
//           op 1 -> out[0] = sqrt(sqrt(su)^2 + (pnw*sqrt(s3))^2), out[1] = sqrt(su), out[2] = pnw*sqrt(s3)
//           op 2 -> out[0..3] = raw segment sums
//           op 3 -> out[0] = s3 / total_volume (pressure mean)
//           op 4 -> out[d] = face-sum_d / total_volume (velocity means)
void launch_finalize(hipStream_t s, const Grid &g, const double *partial, double *out, int op,
                     double pnw);
void launch_face_maxabs(hipStream_t s, const Grid &g, const double *const r[3], double *partial);
void launch_finalize_max(hipStream_t s, const double *partial, double *out);
template <typename T>
void launch_sum_faces_weighted(hipStream_t s, const Grid &g, T *const u[3], double *partial);

// ---- Krylov vector kernels on [u0|u1|(u2)|p] of length (dim+1)*N ----
void launch_bmax_neg(hipStream_t s, double *tmp, const double *b, long n);       // tmp = -(tmp - b)
void launch_div_by(hipStream_t s, double *v, const double *src, long n, const double *scal);  // v = src / *scal
void launch_axmy(hipStream_t s, double *w, const double *v, long n, const double *scal);     // w = w - v * *scal
void launch_update_sol(hipStream_t s, double *x, double *const V[], const double *y, int i, long n);
// w -= h[0]*v, then the Stokes dot of the new w with vn (vn = nullptr: with
// itself) into partial, bit-identical to launch_axmy + launch_stokes_dot
void launch_axmy_dot(hipStream_t s, const Grid &g, double *w, const double *v, const double *h, const double *vn,
                     double *partial);
// v = src / scal[0] and v32 = (float)v in one pass
void launch_div_by_f32(hipStream_t s, double *v, float *v32, const double *src, long n, const double *scal);
// values v0 .. v0+count-1 of the counter-based Gaussian stream (seed, step)
// into out (lmm_rng.hip)
void launch_fill_normal_philox(hipStream_t s, double *out, long count, long v0, uint64_t seed, uint64_t step);
// x += sum_k y_k Z_k, Z in FP32 (at most 32 vectors)
void launch_update_sol_z32(hipStream_t s, double *x, float *const Z[], const double *y, int i, long n);
// Classical Gram-Schmidt passes of the flexible GMRES (1..CGS_MAXV basis
// vectors per pass).  launch_cgs_dots: segment sums of <w, V_k> into
// partial[blk*CGS_STRIDE + 4k + seg]; launch_cgs_finalize: d_k = <w, V_k>
// (Stokes product, box-face weights, pressure weighted by pnw^2) and
// c_k = sig2[k]*d_k; launch_cgs_update: vn = src - sum_k c_k V_k, its FP32
// copy (v32 may be null) and the segment sums of <vn, vn> for
// launch_finalize.  Every level extent must be even (cell pairs).
constexpr int CGS_MAXV = 8;
constexpr int CGS_STRIDE = 4 * CGS_MAXV;
void launch_cgs_dots(hipStream_t s, const Grid &g, const double *w, const double *const V[], int nv,
                     double *partial);
void launch_cgs_finalize(hipStream_t s, const Grid &g, const double *partial, int nv, const double *sig2,
                         double pnw, double *d_out, double *c_out);
void launch_cgs_update(hipStream_t s, const Grid &g, const double *src, const double *const V[], int nv,
                       const double *c, double *vn, float *v32, double *partial);
// r = b - A x in one pass (FP64, 3D visc_type -1 / 1); false when not available
bool launch_apply_matrix_resid(hipStream_t s, const Grid &g, int visc, double *const x_u[3], const double *x_p,
                               double *const b_u[3], const double *b_p, double *const r_u[3], double *r_p,
                               const StagCoef<double> &K);
// FP64 A x of an FP32 x (3D visc_type -1 / 1); false when not available
bool launch_apply_matrix_x32(hipStream_t s, const Grid &g, int visc, float *const x_u[3], const float *x_p,
                             double *const b_u[3], double *b_p, const StagCoef<double> &K);
template <typename T>
void launch_sub_scalar(hipStream_t s, T *x, long n, const double *scal);
template <typename T>
void launch_sub_scalar3(hipStream_t s, const Grid &g, T *const u[3], const double *scal);

// ---- inertial time step (lmm_inertial.hip) ----
struct AdvArgs {
  const double *u[3];  // umac
  const double *s[3];  // face values of the advected components
  double dxinv;
};
struct RhoArgs {
  double *upd;          // mode 0: the updated cell components
  const double *diff;   // modes 1, 2: diffusive mass fluxdiv
  const double *stoch;  // modes 1, 2: stochastic mass fluxdiv, or null
  const double *rho_old;
  double *rho_new;
  double dt;            // mode 1: dt; mode 2: dt/2
};
struct RhsVArgs {
  double *v[3], *mold[3], *gradpi[3], *m_a[3], *m_d[3], *m_s[3], *rfo[3], *rfn[3];
  double g[3];
  double inv_dt;
  int half_ma, half_md, half_ms, grav;
};
struct RhobarInv {
  double v[8];  // -1/rhobar per species
};
// s_update (mode 0) or the rho update of the predictor (1) / corrector (2)
void launch_adv_s(hipStream_t s, const Grid &g, int ncomp, const AdvArgs &a, const RhoArgs &r, int mode);
void launch_rhotot_conc(hipStream_t s, const Grid &g, int ns, double *rho, double *rhotot);
void launch_adv_m(hipStream_t s, const Grid &g, double *const u[3], double *const m[3], double *const upd[3]);
void launch_grav(hipStream_t s, const Grid &g, const double grav[3], double *const f[3], double *const so[3],
                 double *const sn[3]);
void launch_mul_faces(hipStream_t s, const Grid &g, double *const out[3], double *const a[3], double *const b[3]);
void launch_rhs_v(hipStream_t s, const Grid &g, const RhsVArgs &r);
void launch_v_sub_mt(hipStream_t s, const Grid &g, double *const v[3], double *const u[3], double *const rf[3],
                     double inv_dt);
void launch_v_add_half(hipStream_t s, const Grid &g, double *const v[3], double *const w[3]);
void launch_zero_minus(hipStream_t s, const Grid &g, double *const v[3], double *const w[3]);
void launch_add(hipStream_t s, double *a, const double *b, long n);
void launch_rhs_p(hipStream_t s, const Grid &g, int ns, double *bp, double *diff, double *stoch, const double *divu,
                  const RhobarInv &a);
void launch_eta(hipStream_t s, const Grid &g, double *eta, double *kappa, const double *rhotot, double visc_coef,
                int model);

// ---- fluctuating and diffusive fluxes ----
// width > 0: the scaled fluxes are written to work (2*(dim + 2*ne)*N doubles,
// ne = 1 in 2D, 3 in 3D) and filtered (multifab_filter) before the divergence
void launch_stoch_m_force(hipStream_t s, const Grid &g, double *const force[3], const double *Wcc,
                          double *const Wed[3], int nstages, const double *weights, long stage_stride_cc,
                          long stage_stride_ed, const double *eta, double *const eta_ed[3],
                          const double *temp, double *const temp_ed[3], double variance, int width = 0,
                          double *work = nullptr);

}  // namespace lmm
