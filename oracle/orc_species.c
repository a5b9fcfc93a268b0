/*
 * CPU oracle of the species mass-flux chain (test infrastructure only):
 * the per-cell algebra of src_multiSpec/mass_flux_utilities.f90,
 * compute_mixture_properties.f90 and matrix_utilities.f90, and the face
 * fluxes of diffusive_mass_fluxdiv.f90 / stochastic_mass_fluxdiv.f90 as
 * composed by compute_mass_fluxdiv.f90:24-147.
 *
 *   orc_sp_correct_rho      <- mass_flux_utilities.f90:127-150 correct_rho_with_drho_local
 *   orc_sp_cons_to_prim     <- :279-313 compute_molconc_rhotot_local
 *   orc_sp_mixture          <- compute_mixture_properties.f90:165-233
 *                               mixture_properties_mass_local (constant D_bar,
 *                               D_therm, H; prob_type 9 is not restated)
 *   orc_sp_gama             <- mass_flux_utilities.f90:562-599 compute_Gama_local
 *   orc_sp_chi              <- :767-831 compute_chi_local, with
 *                               matrix_utilities.f90:21-156 Dbar2chi_iterative
 *                               (use_lapack = F) and :833-927 compute_chi_lapack
 *                               (use_lapack = T)
 *   orc_sp_rhowchi          <- :1250-1273 compute_rhoWchi_local
 *   orc_sp_lonsager         <- :1068-1118 compute_Lonsager_local with
 *                               matrix_utilities.f90:160-239 choldc (use_lapack = F)
 *   orc_sp_baro_coef        <- :1342-1369 compute_baro_coef_2d
 *   orc_diffusive_mass_flux_full <- diffusive_mass_fluxdiv.f90:26-262 with the
 *                               temperature and barodiffusion pieces
 *   orc_stochastic_mass_fluxdiv <- stochastic_mass_fluxdiv.f90:32-154
 *   orc_compute_mass_fluxdiv <- compute_mass_fluxdiv.f90:24-147 (external_source
 *                               is a manufactured-solution source of prob_types
 *                               4/5 and is not part of the chain here)
 *
 * Matrices are column-major (ns, ns): entry (r, c) at r + ns * c, as the
 * reference reshapes its nspecies^2 components.  Fields are deduplicated
 * periodic arrays, component-major.  Sums accumulate from 0.0 in ascending
 * index order, the order of the reference's explicit loops; the reference's
 * matmul intrinsic (Gama, matvec_mul) is taken in the same ascending order.
 *
 * use_lapack = T replaces LAPACK dgetrf/dgetri (inverse_type 1), la_gesvd
 * (inverse_type 2) and dpotrf by textbook algorithms of the same
 * mathematical result (Gauss-Jordan with partial pivoting, one-sided Jacobi
 * SVD, row-oriented Cholesky), so that path agrees with LAPACK to rounding,
 * not bit for bit.
 */
#include "orc_fab.h"
#include "lmm_oracle.h"
#include <float.h>

#define MS ORC_MAX_SPECIES

/* ---------------- per-cell algebra ---------------- */

void orc_sp_correct_rho(const orc_species *sp, double *rho, double *drho) {
  int ns = sp->nspecies;
  double old[MS], tot = 0.0;
  for (int n = 0; n < ns; n++) {
    old[n] = rho[n];
    tot = tot + rho[n];
  }
  double floor_v = sp->fraction_tolerance * tot;
  for (int n = 0; n < ns; n++) rho[n] = rho[n] > floor_v ? rho[n] : floor_v;
  for (int n = 0; n < ns; n++) drho[n] = rho[n] - old[n];
}

void orc_sp_cons_to_prim(const orc_species *sp, const double *rho, double *rhotot, double *molarconc,
                         double *molmtot) {
  int ns = sp->nspecies;
  double tot = 0.0, W[MS], s = 0.0;
  for (int n = 0; n < ns; n++) tot = tot + rho[n];
  *rhotot = tot;
  for (int n = 0; n < ns; n++) {
    W[n] = rho[n] / tot;
    s = s + W[n] / sp->molmass[n];
  }
  double mt = 1.0 / s;
  *molmtot = mt;
  for (int n = 0; n < ns; n++) molarconc[n] = mt * W[n] / sp->molmass[n];
}

void orc_sp_mixture(const orc_species *sp, double *D_bar, double *D_therm, double *H) {
  int ns = sp->nspecies, q = 0;
  for (int r = 0; r < ns; r++) {
    for (int c = 0; c < r; c++, q++) {
      D_bar[r + ns * c] = sp->Dbar[q];
      D_bar[c + ns * r] = sp->Dbar[q];
      if (!sp->is_ideal_mixture) {
        H[r + ns * c] = sp->H_offdiag[q];
        H[c + ns * r] = sp->H_offdiag[q];
      }
    }
    D_bar[r + ns * r] = 0.0;
    D_therm[r] = sp->Dtherm[r];
    if (!sp->is_ideal_mixture)
      H[r + ns * r] = sp->H_diag[r];
    else
      for (int k = 0; k < ns * ns; k++) H[k] = 0.0;
  }
}

void orc_sp_gama(const orc_species *sp, const double *molarconc, const double *H, double *Gama) {
  int ns = sp->nspecies;
  double X[MS * MS];
  for (int k = 0; k < ns * ns; k++) X[k] = 0.0;
  if (!sp->is_ideal_mixture)
    for (int r = 0; r < ns; r++) {
      for (int c = 0; c < r; c++) {
        X[r + ns * c] = -(molarconc[r] * molarconc[c]);
        X[c + ns * r] = X[r + ns * c];
      }
      X[r + ns * r] = molarconc[r] - molarconc[r] * molarconc[r];
    }
  for (int r = 0; r < ns; r++)
    for (int c = 0; c < ns; c++) {
      double acc = 0.0;
      for (int k = 0; k < ns; k++) acc = acc + X[r + ns * k] * H[k + ns * c];
      Gama[r + ns * c] = (r == c ? 1.0 : 0.0) + acc;
    }
}

/* matrix_utilities.f90:21-156 */
static void dbar2chi_iterative(const orc_species *sp, const double *D_bar, const double *Xk, double *chi) {
  int ns = sp->nspecies;
  double MW = 0.0, Y[MS], Di[MS], Mm[MS], Mi[MS];
  double P[MS * MS], Z[MS * MS], J[MS * MS], PJ[MS * MS], M2[MS * MS], M1[MS * MS];
  for (int i = 0; i < ns; i++) MW = MW + Xk[i] * sp->molmass[i];
  for (int i = 0; i < ns; i++) Y[i] = sp->molmass[i] / MW * Xk[i];
  for (int i = 0; i < ns; i++) {
    double t2 = 0.0;
    for (int j = 0; j < ns; j++)
      if (j != i) t2 = t2 + Xk[j] / D_bar[i + ns * j];
    Di[i] = (1.0 - Y[i]) / t2;
  }
  for (int i = 0; i < ns; i++) {
    Mm[i] = Xk[i] / Di[i];
    Mi[i] = Di[i] / Xk[i];
  }
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) P[i + ns * j] = (i == j) ? -Y[j] + 1.0 : -Y[j];
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) {
      double delta;
      if (i == j) {
        double t1 = 0.0;
        for (int k = 0; k < ns; k++)
          if (k != i) t1 = t1 + Xk[i] * Xk[k] / D_bar[i + ns * k];
        delta = t1;
      } else {
        delta = -(Xk[i] * Xk[j] / D_bar[i + ns * j]);
      }
      Z[i + ns * j] = -delta;
    }
  for (int i = 0; i < ns; i++) Z[i + ns * i] = Z[i + ns * i] + Mm[i];
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) J[i + ns * j] = Mi[i] * Z[i + ns * j];
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) {
      double acc = 0.0;
      for (int k = 0; k < ns; k++) acc = acc + P[i + ns * k] * J[k + ns * j];
      PJ[i + ns * j] = acc;
    }
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) {
      double scr = 0.0;
      for (int k = 0; k < ns; k++) scr = scr + P[i + ns * k] * Mi[k] * P[j + ns * k];
      M2[i + ns * j] = scr;
      chi[i + ns * j] = scr;
    }
  for (int it = 0; it < sp->chi_iterations; it++) {
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) {
        double scr = 0.0;
        for (int k = 0; k < ns; k++) scr = scr + PJ[i + ns * k] * chi[k + ns * j];
        M1[i + ns * j] = scr + M2[i + ns * j];
      }
    for (int k = 0; k < ns * ns; k++) chi[k] = M1[k];
  }
}

/* Inverse by Gauss-Jordan elimination with partial pivoting (the matrix
 * inverse LAPACK dgetrf + dgetri compute).  Returns nonzero if singular. */
int orc_sp_invert(int ns, double *A) {
  double B[MS * MS];
  for (int r = 0; r < ns; r++)
    for (int c = 0; c < ns; c++) B[r + ns * c] = (r == c) ? 1.0 : 0.0;
  for (int c = 0; c < ns; c++) {
    int p = c;
    for (int r = c + 1; r < ns; r++)
      if (fabs(A[r + ns * c]) > fabs(A[p + ns * c])) p = r;
    if (A[p + ns * c] == 0.0) return 1;
    if (p != c)
      for (int k = 0; k < ns; k++) {
        double t = A[c + ns * k];
        A[c + ns * k] = A[p + ns * k];
        A[p + ns * k] = t;
        t = B[c + ns * k];
        B[c + ns * k] = B[p + ns * k];
        B[p + ns * k] = t;
      }
    double inv = 1.0 / A[c + ns * c];
    for (int k = 0; k < ns; k++) {
      A[c + ns * k] = A[c + ns * k] * inv;
      B[c + ns * k] = B[c + ns * k] * inv;
    }
    for (int r = 0; r < ns; r++) {
      if (r == c) continue;
      double f = A[r + ns * c];
      if (f == 0.0) continue;
      for (int k = 0; k < ns; k++) {
        A[r + ns * k] = A[r + ns * k] - f * A[c + ns * k];
        B[r + ns * k] = B[r + ns * k] - f * B[c + ns * k];
      }
    }
  }
  for (int k = 0; k < ns * ns; k++) A[k] = B[k];
  return 0;
}

/* Pseudo-inverse V diag(1/S) U^T through a one-sided Jacobi SVD (Hestenes),
 * dropping singular values at or below tol * sum(S), the rule of
 * mass_flux_utilities.f90:905-915. */
void orc_sp_pinv(int ns, const double *A0, double tol, double *out) {
  double U[MS * MS], V[MS * MS], S[MS];
  for (int k = 0; k < ns * ns; k++) U[k] = A0[k];
  for (int r = 0; r < ns; r++)
    for (int c = 0; c < ns; c++) V[r + ns * c] = (r == c) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0;
    for (int p = 0; p < ns - 1; p++)
      for (int q = p + 1; q < ns; q++) {
        double a = 0.0, b = 0.0, g = 0.0;
        for (int k = 0; k < ns; k++) {
          a = a + U[k + ns * p] * U[k + ns * p];
          b = b + U[k + ns * q] * U[k + ns * q];
          g = g + U[k + ns * p] * U[k + ns * q];
        }
        if (g == 0.0) continue;
        double rel = fabs(g) / sqrt(a * b);
        if (rel > off) off = rel;
        if (rel < 1e-15) continue;
        double zeta = (b - a) / (2.0 * g);
        double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for (int k = 0; k < ns; k++) {
          double up = U[k + ns * p], uq = U[k + ns * q];
          U[k + ns * p] = cs * up - sn * uq;
          U[k + ns * q] = sn * up + cs * uq;
          double vp = V[k + ns * p], vq = V[k + ns * q];
          V[k + ns * p] = cs * vp - sn * vq;
          V[k + ns * q] = sn * vp + cs * vq;
        }
      }
    if (off < 1e-15) break;
  }
  double ssum = 0.0;
  for (int c = 0; c < ns; c++) {
    double s = 0.0;
    for (int k = 0; k < ns; k++) s = s + U[k + ns * c] * U[k + ns * c];
    S[c] = sqrt(s);
    ssum = ssum + S[c];
  }
  for (int k = 0; k < ns * ns; k++) out[k] = 0.0;
  for (int c = 0; c < ns; c++) {
    if (!(S[c] > tol * ssum)) continue;
    /* column c of U holds S_c u_c: V_c (S_c u_c)^T / S_c^2 = V_c u_c^T / S_c */
    double w = 1.0 / (S[c] * S[c]);
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) out[i + ns * j] = out[i + ns * j] + V[i + ns * c] * w * U[j + ns * c];
  }
}

/* mass_flux_utilities.f90:833-927 compute_chi_lapack */
static void chi_lapack(const orc_species *sp, const double *Lambda, const double *W, double *chi) {
  int ns = sp->nspecies;
  double alpha = 0.0, A[MS * MS];
  for (int r = 0; r < ns; r++) alpha = alpha + Lambda[r + ns * r];
  for (int r = 0; r < ns; r++)
    for (int c = 0; c < ns; c++) A[r + ns * c] = alpha * W[r] * W[c] + Lambda[r + ns * c];
  if (sp->inverse_type == 2) {
    orc_sp_pinv(ns, A, sp->fraction_tolerance, chi);
  } else {
    orc_sp_invert(ns, A);
    for (int k = 0; k < ns * ns; k++) chi[k] = A[k];
  }
  for (int r = 0; r < ns; r++)
    for (int c = 0; c < ns; c++) chi[r + ns * c] = chi[r + ns * c] - 1.0 / alpha;
}

void orc_sp_chi(const orc_species *sp, const double *rho, double rhotot, const double *molarconc,
                const double *D_bar, double Temp, const double *D_therm, double *chi, double *zeta_by_Temp) {
  int ns = sp->nspecies;
  double Lambda[MS * MS], W[MS];
  for (int k = 0; k < ns * ns; k++) Lambda[k] = 0.0;
  for (int r = 0; r < ns; r++) {
    for (int c = 0; c < r; c++) {
      Lambda[r + ns * c] = -(molarconc[r] * molarconc[c] / D_bar[r + ns * c]);
      Lambda[c + ns * r] = Lambda[r + ns * c];
    }
    W[r] = rho[r] / rhotot;
  }
  for (int r = 0; r < ns; r++) {
    double s = 0.0;
    for (int c = 0; c < ns; c++) {
      if (c != r) s = s - Lambda[r + ns * c];
      Lambda[r + ns * r] = s;
    }
  }
  if (sp->is_nonisothermal)
    for (int r = 0; r < ns; r++) {
      double s = 0.0;
      for (int c = 0; c < ns; c++)
        if (c != r) s = s + Lambda[r + ns * c] * (D_therm[r] - D_therm[c]);
      zeta_by_Temp[r] = s / Temp;
    }
  if (sp->use_lapack)
    chi_lapack(sp, Lambda, W, chi);
  else
    dbar2chi_iterative(sp, D_bar, molarconc, chi);
}

void orc_sp_rhowchi(const orc_species *sp, const double *rho, const double *chi, double *rhoWchi) {
  int ns = sp->nspecies;
  for (int r = 0; r < ns; r++)
    for (int c = 0; c < ns; c++) rhoWchi[r + ns * c] = -(rho[r] * chi[r + ns * c]);
}

/* matrix_utilities.f90:160-239 choldc: lower triangle and diagonal hold the
 * factor, the upper triangle is zeroed; a non-positive pivot zeroes its
 * column below the diagonal. */
static void choldc(int np, double *a) {
  double p[MS];
  for (int i = 0; i < np; i++) {
    int ising = 0;
    for (int j = i; j < np; j++) {
      double s = a[i + np * j];
      for (int k = i - 1; k >= 0; k--) s = s - a[i + np * k] * a[j + np * k];
      if (i == j) {
        if (s <= 0.0) {
          p[i] = 0.0;
          ising = 1;
        } else {
          p[i] = sqrt(s);
        }
      } else {
        a[j + np * i] = ising ? 0.0 : s / p[i];
      }
    }
  }
  for (int i = 0; i < np; i++) {
    for (int j = i + 1; j < np; j++) a[i + np * j] = 0.0;
    a[i + np * i] = p[i];
  }
}

/* The factor LAPACK dpotrf('L') computes, with the reference's clean-up
 * (upper triangle and the (ns, ns) entry zeroed, :1101-1109): column j's
 * pivot is a(j,j) - sum_k a(j,k)^2, the column below it is scaled by its
 * reciprocal; a non-positive pivot stops the factorisation as dpotf2 does. */
static void chol_lapack(int np, double *a) {
  for (int j = 0; j < np; j++) {
    double s = a[j + np * j];
    for (int k = 0; k < j; k++) s = s - a[j + np * k] * a[j + np * k];
    if (!(s > 0.0)) {
      a[j + np * j] = s;
      break;
    }
    double d = sqrt(s);
    a[j + np * j] = d;
    double rinv = 1.0 / d;
    for (int i = j + 1; i < np; i++) {
      double t = a[i + np * j];
      for (int k = 0; k < j; k++) t = t - a[i + np * k] * a[j + np * k];
      a[i + np * j] = t * rinv;
    }
  }
  for (int r = 0; r < np; r++)
    for (int c = r + 1; c < np; c++) a[r + np * c] = 0.0;
  a[(np - 1) + np * (np - 1)] = 0.0;
}

void orc_sp_lonsager(const orc_species *sp, const double *rho, double rhotot, double molmtot, const double *chi,
                     double *L) {
  int ns = sp->nspecies;
  double W[MS];
  for (int r = 0; r < ns; r++) W[r] = rho[r] / rhotot;
  for (int c = 0; c < ns; c++)
    for (int r = 0; r < ns; r++) L[r + ns * c] = molmtot * rhotot * W[r] * chi[r + ns * c] * W[c] / sp->k_B;
  if (sp->use_lapack)
    chol_lapack(ns, L);
  else
    choldc(ns, L);
}

void orc_sp_baro_coef(const orc_species *sp, const double *rho, double Temp, double *baro) {
  int ns = sp->nspecies;
  double n = 0.0;
  for (int c = 0; c < ns; c++) n = n + rho[c] / sp->molmass[c];
  for (int c = 0; c < ns; c++) baro[c] = rho[c] / sp->rhobar[c] / (n * sp->k_B * Temp);
}

/* ---------------- whole-field restatements ---------------- */

static long ncells(const orc_geom *g) { return (long)g->n[0] * g->n[1] * (g->dim == 3 ? g->n[2] : 1); }

static int sp_ok(const orc_species *sp) { return sp && sp->nspecies >= 2 && sp->nspecies <= MS; }

/* Gather the ncomp components of cell t from a component-major field. */
static void gather(const double *f, int ncomp, long N, long t, double *v) {
  for (int q = 0; q < ncomp; q++) v[q] = f[q * N + t];
}
static void scatter(double *f, int ncomp, long N, long t, const double *v) {
  for (int q = 0; q < ncomp; q++) f[q * N + t] = v[q];
}

int orc_correct_rho_with_drho(const orc_geom *g, const orc_species *sp, double *rho, double *drho) {
  if (!sp_ok(sp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies;
  for (long t = 0; t < N; t++) {
    double r[MS], d[MS];
    gather(rho, ns, N, t, r);
    orc_sp_correct_rho(sp, r, d);
    scatter(rho, ns, N, t, r);
    scatter(drho, ns, N, t, d);
  }
  return 0;
}

int orc_convert_cons_to_prim(const orc_geom *g, const orc_species *sp, const double *rho, double *rhotot,
                             double *molarconc, double *molmtot) {
  if (!sp_ok(sp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies;
  for (long t = 0; t < N; t++) {
    double r[MS], x[MS];
    gather(rho, ns, N, t, r);
    orc_sp_cons_to_prim(sp, r, &rhotot[t], x, &molmtot[t]);
    scatter(molarconc, ns, N, t, x);
  }
  return 0;
}

int orc_compute_mixture_properties(const orc_geom *g, const orc_species *sp, double *D_bar, double *D_therm,
                                   double *Hessian) {
  if (!sp_ok(sp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies;
  double Db[MS * MS], Dt[MS], H[MS * MS];
  for (int k = 0; k < ns * ns; k++) H[k] = 0.0; /* unset entries of an ideal mixture's loop */
  orc_sp_mixture(sp, Db, Dt, H);
  for (long t = 0; t < N; t++) {
    scatter(D_bar, ns * ns, N, t, Db);
    scatter(D_therm, ns, N, t, Dt);
    scatter(Hessian, ns * ns, N, t, H);
  }
  return 0;
}

int orc_compute_Gama(const orc_geom *g, const orc_species *sp, const double *molarconc, const double *Hessian,
                     double *Gama) {
  if (!sp_ok(sp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies;
  for (long t = 0; t < N; t++) {
    double x[MS], H[MS * MS], G[MS * MS];
    gather(molarconc, ns, N, t, x);
    gather(Hessian, ns * ns, N, t, H);
    orc_sp_gama(sp, x, H, G);
    scatter(Gama, ns * ns, N, t, G);
  }
  return 0;
}

int orc_compute_chi(const orc_geom *g, const orc_species *sp, const double *rho, const double *rhotot,
                    const double *molarconc, double *chi, const double *D_bar, const double *D_therm,
                    const double *Temp, double *zeta_by_Temp) {
  if (!sp_ok(sp)) return 2;
  if (sp->is_nonisothermal && (!Temp || !D_therm || !zeta_by_Temp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies;
  for (long t = 0; t < N; t++) {
    double r[MS], x[MS], Db[MS * MS], Dt[MS], ch[MS * MS], z[MS];
    gather(rho, ns, N, t, r);
    gather(molarconc, ns, N, t, x);
    gather(D_bar, ns * ns, N, t, Db);
    if (D_therm) gather(D_therm, ns, N, t, Dt);
    orc_sp_chi(sp, r, rhotot[t], x, Db, Temp ? Temp[t] : 1.0, Dt, ch, z);
    scatter(chi, ns * ns, N, t, ch);
    if (sp->is_nonisothermal && zeta_by_Temp) scatter(zeta_by_Temp, ns, N, t, z);
  }
  return 0;
}

int orc_compute_rhoWchi(const orc_geom *g, const orc_species *sp, const double *rho, const double *chi,
                        double *rhoWchi) {
  if (!sp_ok(sp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies;
  for (long t = 0; t < N; t++) {
    double r[MS], ch[MS * MS], rw[MS * MS];
    gather(rho, ns, N, t, r);
    gather(chi, ns * ns, N, t, ch);
    orc_sp_rhowchi(sp, r, ch, rw);
    scatter(rhoWchi, ns * ns, N, t, rw);
  }
  return 0;
}

int orc_compute_Lonsager(const orc_geom *g, const orc_species *sp, const double *rho, const double *rhotot,
                         const double *molmtot, const double *chi, double *Lonsager) {
  if (!sp_ok(sp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies;
  for (long t = 0; t < N; t++) {
    double r[MS], ch[MS * MS], L[MS * MS];
    gather(rho, ns, N, t, r);
    gather(chi, ns * ns, N, t, ch);
    orc_sp_lonsager(sp, r, rhotot[t], molmtot[t], ch, L);
    scatter(Lonsager, ns * ns, N, t, L);
  }
  return 0;
}

/* Periodic neighbour index of cell (i,j,k) shifted by -1 (s = -1) or +1 in dir d. */
static long nb(const orc_geom *g, long t, int d, int s) {
  int n0 = g->n[0], n1 = g->n[1], n2 = g->dim == 3 ? g->n[2] : 1;
  int i = (int)(t % n0), j = (int)((t / n0) % n1), k = (int)(t / ((long)n0 * n1));
  int c[3] = {i, j, k}, n[3] = {n0, n1, n2};
  c[d] = (c[d] + s + n[d]) % n[d];
  return c[0] + (long)n0 * (c[1] + (long)n1 * c[2]);
}

/* diffusive_mass_fluxdiv.f90:92-262 on one face: grad x (compute_grad),
 * Gama_face * grad x, + zeta_face * grad T, + baro_face * gradp_baro,
 * rhoWchi_face * (...), then correction_flux.f90:71-251.  Fields are
 * component-major with N cells; (t, m) are the cells above and below the face. */
static void diff_face(const orc_species *sp, long N, double dxinv, long t, long m, const double *molarconc,
                      const double *rhoWchi, const double *Gama, const double *Temp, const double *zeta,
                      const double *baro, double gpb, double *out) {
  int ns = sp->nspecies;
  double x[MS], y[MS];
  for (int s = 0; s < ns; s++) x[s] = (molarconc[s * N + t] - molarconc[s * N + m]) * dxinv;
  for (int r = 0; r < ns; r++) {
    double acc = 0.0;
    for (int c = 0; c < ns; c++) {
      long q = r + ns * c;
      acc = acc + 0.5 * (Gama[q * N + t] + Gama[q * N + m]) * x[c];
    }
    y[r] = acc;
  }
  if (zeta) {
    double gT = (Temp[t] - Temp[m]) * dxinv;
    for (int s = 0; s < ns; s++) y[s] = y[s] + 0.5 * (zeta[s * N + t] + zeta[s * N + m]) * gT;
  }
  if (baro)
    for (int s = 0; s < ns; s++) y[s] = y[s] + 0.5 * (baro[s * N + t] + baro[s * N + m]) * gpb;
  for (int r = 0; r < ns; r++) {
    double acc = 0.0;
    for (int c = 0; c < ns; c++) {
      long q = r + ns * c;
      acc = acc + 0.5 * (rhoWchi[q * N + t] + rhoWchi[q * N + m]) * y[c];
    }
    out[r] = acc;
  }
  if (sp->correct_flux && ns > 1) {
    double sum = 0.0;
    for (int s = 0; s < ns - 1; s++) sum = sum + out[s];
    out[ns - 1] = -sum;
  }
}

/* compute_div (div_and_grad.f90:304-455) of species face fluxes flux[d][s*N+t]. */
static void div_species(const orc_geom *g, int ns, long N, double dxinv, double *const flux[3], double *div,
                        int increment) {
  for (int s = 0; s < ns; s++)
    for (long t = 0; t < N; t++) {
      double v = (flux[0][s * N + nb(g, t, 0, 1)] - flux[0][s * N + t]) * dxinv;
      v = v + (flux[1][s * N + nb(g, t, 1, 1)] - flux[1][s * N + t]) * dxinv;
      if (g->dim == 3) v = v + (flux[2][s * N + nb(g, t, 2, 1)] - flux[2][s * N + t]) * dxinv;
      div[s * N + t] = increment ? div[s * N + t] + v : v;
    }
}

int orc_diffusive_mass_flux_full(const orc_geom *g, const orc_species *sp, const double *rho,
                                 const double *rhotot, const double *molarconc, const double *rhoWchi,
                                 const double *Gama, double *diff_fluxdiv, const double *Temp,
                                 const double *zeta_by_Temp, double *const gradp_baro[3],
                                 double *flux_total[3]) {
  if (!sp_ok(sp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies, dm = g->dim;
  double dxinv = 1.0 / g->dx;
  /* the temperature piece enters when the mixture is non-isothermal and
   * both Temp and zeta_by_Temp are given (one without the other is an error) */
  if (sp->is_nonisothermal && (!Temp) != (!zeta_by_Temp)) return 2;
  const double *zeta = (sp->is_nonisothermal && Temp) ? zeta_by_Temp : NULL;
  double *baro = NULL;
  if (sp->barodiffusion_type > 0) {
    if (dm != 2 || !gradp_baro || !Temp) return 2; /* compute_baro_coef is 2D only (mass_flux_utilities.f90:1291-1340) */
    baro = malloc(sizeof(double) * ns * N);
    for (long t = 0; t < N; t++) {
      double r[MS], b[MS];
      gather(rho, ns, N, t, r);
      orc_sp_baro_coef(sp, r, Temp[t], b);
      scatter(baro, ns, N, t, b);
    }
  }
  double *flux[3] = {NULL, NULL, NULL};
  for (int d = 0; d < dm; d++) {
    flux[d] = malloc(sizeof(double) * ns * N);
    for (long t = 0; t < N; t++) {
      long m = nb(g, t, d, -1);
      double out[MS];
      diff_face(sp, N, dxinv, t, m, molarconc, rhoWchi, Gama, Temp, zeta, baro, baro ? gradp_baro[d][t] : 0.0, out);
      for (int s = 0; s < ns; s++) {
        flux[d][s * N + t] = out[s];
        flux_total[d][s * N + t] = flux_total[d][s * N + t] + out[s];
      }
    }
  }
  (void)rhotot;
  div_species(g, ns, N, dxinv, flux, diff_fluxdiv, 0);
  for (int d = 0; d < dm; d++) free(flux[d]);
  free(baro);
  return 0;
}

/* stochastic_mass_fluxdiv.f90:32-154.  W holds n_rngs stages of dim face
 * arrays of ns components: stage r, direction d, species s at
 * ((r * dim + d) * ns + s) * N. */
int orc_stochastic_mass_fluxdiv(const orc_geom *g, const orc_species *sp, const double *rho,
                                const double *rhotot, const double *molarconc, const double *molmtot,
                                const double *chi, const double *Gama, double *stoch_fluxdiv,
                                double *flux_total[3], double dt, const double *W, int n_rngs,
                                const double *weights, int increment) {
  if (!sp_ok(sp) || n_rngs < 1) return 2;
  (void)molarconc;
  (void)Gama;
  long N = ncells(g);
  int ns = sp->nspecies, dm = g->dim, nn = ns * ns;
  double dv = 1.0;
  for (int d = 0; d < dm; d++) dv = dv * g->dx;
  double variance = sqrt(2.0 * sp->k_B * sp->variance_coef_mass / (dv * dt));
  double *L = malloc(sizeof(double) * nn * N);
  for (long t = 0; t < N; t++) {
    double r[MS], ch[MS * MS], Lc[MS * MS];
    gather(rho, ns, N, t, r);
    gather(chi, nn, N, t, ch);
    orc_sp_lonsager(sp, r, rhotot[t], molmtot[t], ch, Lc);
    scatter(L, nn, N, t, Lc);
  }
  double *flux[3] = {NULL, NULL, NULL};
  double dxinv = 1.0 / g->dx;
  for (int d = 0; d < dm; d++) {
    flux[d] = malloc(sizeof(double) * ns * N);
    for (long t = 0; t < N; t++) {
      long m = nb(g, t, d, -1);
      double w[MS], out[MS];
      for (int s = 0; s < ns; s++) {
        double acc = 0.0;
        for (int r = 0; r < n_rngs; r++) acc = acc + weights[r] * W[(((long)r * dm + d) * ns + s) * N + t];
        w[s] = acc;
      }
      for (int r = 0; r < ns; r++) {
        double acc = 0.0;
        for (int c = 0; c < ns; c++) {
          long q = r + ns * c;
          acc = acc + 0.5 * (L[q * N + t] + L[q * N + m]) * w[c];
        }
        out[r] = acc * variance;
      }
      if (sp->correct_flux && ns > 1) {
        double sum = 0.0;
        for (int s = 0; s < ns - 1; s++) sum = sum + out[s];
        out[ns - 1] = -sum;
      }
      for (int s = 0; s < ns; s++) {
        flux[d][s * N + t] = out[s];
        flux_total[d][s * N + t] = flux_total[d][s * N + t] + out[s];
      }
    }
  }
  div_species(g, ns, N, dxinv, flux, stoch_fluxdiv, increment);
  for (int d = 0; d < dm; d++) free(flux[d]);
  free(L);
  return 0;
}

/* compute_mass_fluxdiv.f90:24-147: rho is corrected in place for the
 * duration of the call and restored by rho - drho at its end. */
int orc_compute_mass_fluxdiv(const orc_geom *g, const orc_species *sp, double *rho, double *const gradp_baro[3],
                             double *diff_fluxdiv, double *stoch_fluxdiv, const double *Temp,
                             double *flux_total[3], double dt, const double *W, int n_rngs,
                             const double *weights) {
  if (!sp_ok(sp)) return 2;
  long N = ncells(g);
  int ns = sp->nspecies, nn = ns * ns, dm = g->dim;
  double *drho = malloc(sizeof(double) * ns * N), *rhotot = malloc(sizeof(double) * N);
  double *xm = malloc(sizeof(double) * ns * N), *mt = malloc(sizeof(double) * N);
  double *Db = malloc(sizeof(double) * nn * N), *Dt = malloc(sizeof(double) * ns * N);
  double *H = malloc(sizeof(double) * nn * N), *G = malloc(sizeof(double) * nn * N);
  double *ch = malloc(sizeof(double) * nn * N), *z = malloc(sizeof(double) * ns * N);
  double *rw = malloc(sizeof(double) * nn * N);
  for (long k = 0; k < (long)ns * N; k++) z[k] = 0.0;
  int rc = orc_correct_rho_with_drho(g, sp, rho, drho);
  if (!rc) rc = orc_convert_cons_to_prim(g, sp, rho, rhotot, xm, mt);
  if (!rc) rc = orc_compute_mixture_properties(g, sp, Db, Dt, H);
  if (!rc) rc = orc_compute_Gama(g, sp, xm, H, G);
  if (!rc) rc = orc_compute_chi(g, sp, rho, rhotot, xm, ch, Db, Dt, Temp, z);
  if (!rc) rc = orc_compute_rhoWchi(g, sp, rho, ch, rw);
  if (!rc)
    for (int d = 0; d < dm; d++)
      for (long k = 0; k < (long)ns * N; k++) flux_total[d][k] = 0.0;
  if (!rc)
    rc = orc_diffusive_mass_flux_full(g, sp, rho, rhotot, xm, rw, G, diff_fluxdiv, Temp, z, gradp_baro,
                                      flux_total);
  if (!rc) {
    if (sp->variance_coef_mass != 0.0)
      rc = orc_stochastic_mass_fluxdiv(g, sp, rho, rhotot, xm, mt, ch, G, stoch_fluxdiv, flux_total, dt, W,
                                       n_rngs, weights, 0);
    else
      for (long k = 0; k < (long)ns * N; k++) stoch_fluxdiv[k] = 0.0;
  }
  for (long k = 0; k < (long)ns * N; k++) rho[k] = rho[k] + -1.0 * drho[k];
  free(drho);
  free(rhotot);
  free(xm);
  free(mt);
  free(Db);
  free(Dt);
  free(H);
  free(G);
  free(ch);
  free(z);
  free(rw);
  return rc;
}
