/*
 * Advection and gravity terms of the inertial time step, CPU oracle (test
 * infrastructure only).
 *   orc_mk_advective_s_fluxdiv <- src_common/mk_advective_s_fluxdiv.f90:14-166
 *   orc_mk_advective_m_fluxdiv <- src_common/mk_advective_m_fluxdiv.f90:16-331
 *   orc_mk_grav_force          <- src_common/mk_grav_force.f90:17-161
 * Each runs the reference's per-box loops on ghost-filled fabs of the single
 * periodic box (the physical-boundary calls of the reference,
 * multifab_physbc_domainvel and zero_edgeval_physical, do nothing on a
 * periodic box).  Every flux is the product the reference forms, in its
 * order: 0.25d0*(a+b)*(c+d) is ((0.25*(a+b))*(c+d)).
 */
#include "orc_fab.h"

static orc_geom gnorm(const orc_geom *g) {
  orc_geom h = *g;
  if (h.dim == 2) h.n[2] = 1;
  return h;
}

static void faces_load(ofab *u, const orc_geom *g, int ng, double *const src[3], long off) {
  for (int d = 0; d < g->dim; d++) {
    face_build(&u[d], g, d, ng);
    if (src && src[d]) fab_from_dedup(&u[d], src[d] + off);
  }
}
static void faces_drop(ofab *u, const orc_geom *g) {
  for (int d = 0; d < g->dim; d++) fab_free(&u[d]);
}

/* mk_advective_s_fluxdiv_2d/3d (mk_advective_s_fluxdiv.f90:68-166):
 * s_update -= div(umac * s_fc), per component; s_fc holds ncomp components
 * per face direction, s_update ncomp cell components (component-major). */
int orc_mk_advective_s_fluxdiv(const orc_geom *g0, int ncomp, double *const umac[3], double *const s_fc[3],
                               double *s_update) {
  orc_geom g = gnorm(g0);
  const long N = (long)g.n[0] * g.n[1] * g.n[2];
  const double dxinv = 1.0 / g.dx;
  ofab u[3], s[3], a;
  faces_load(u, &g, 1, umac, 0);
  for (int c = 0; c < ncomp; c++) {
    faces_load(s, &g, 0, s_fc, c * N);
    cell_build(&a, &g, 0);
    fab_from_dedup(&a, s_update + c * N);
    for (int k = 0; k < g.n[2]; k++)
      for (int j = 0; j < g.n[1]; j++)
        for (int i = 0; i < g.n[0]; i++) {
          /* fluxx(i) = umac(i)*sx(i) on the faces of cell (i,j,k) */
          double fxl = F(&u[0], i, j, k) * F(&s[0], i, j, k);
          double fxh = F(&u[0], i + 1, j, k) * F(&s[0], i + 1, j, k);
          double fyl = F(&u[1], i, j, k) * F(&s[1], i, j, k);
          double fyh = F(&u[1], i, j + 1, k) * F(&s[1], i, j + 1, k);
          double div = (fxh - fxl) * dxinv + (fyh - fyl) * dxinv;
          if (g.dim == 3) {
            double fzl = F(&u[2], i, j, k) * F(&s[2], i, j, k);
            double fzh = F(&u[2], i, j, k + 1) * F(&s[2], i, j, k + 1);
            div = div + (fzh - fzl) * dxinv;
          }
          F(&a, i, j, k) = F(&a, i, j, k) - div;
        }
    fab_to_dedup(&a, s_update + c * N);
    fab_free(&a);
    faces_drop(s, &g);
  }
  faces_drop(u, &g);
  return 0;
}

/* 0.25d0*(a+b)*(c+d) */
static inline double q4(double a, double b, double c, double d) { return 0.25 * (a + b) * (c + d); }

/* mk_advective_m_fluxdiv_2d/3d (mk_advective_m_fluxdiv.f90:87-329):
 * m_update += -div(m u) on every face component, the momentum fluxes formed
 * from centred averages of m and umac (ng 1 on both). */
int orc_mk_advective_m_fluxdiv(const orc_geom *g0, double *const umac[3], double *const m[3],
                               double *m_update[3]) {
  orc_geom g = gnorm(g0);
  const int dm = g.dim;
  const double dxinv = 1.0 / g.dx;
  ofab u[3], mm[3], up[3];
  faces_load(u, &g, 1, umac, 0);
  faces_load(mm, &g, 1, m, 0);
  faces_load(up, &g, 0, m_update, 0);
  ofab *U = &u[0], *V = &u[1], *W = dm == 3 ? &u[2] : NULL;
  for (int a = 0; a < dm; a++) {
    ofab *M = &mm[a];
    for (int k = 0; k <= vhi(&up[a], 2); k++)
      for (int j = 0; j <= vhi(&up[a], 1); j++)
        for (int i = 0; i <= vhi(&up[a], 0); i++) {
          double fd;
          if (a == 0) {
            /* mx_fluxx(i) at cell i-1; mx_fluxy(i,j) on the x-y node; mx_fluxz likewise */
            double xl = q4(F(M, i - 1, j, k), F(M, i, j, k), F(U, i - 1, j, k), F(U, i, j, k));
            double xh = q4(F(M, i, j, k), F(M, i + 1, j, k), F(U, i, j, k), F(U, i + 1, j, k));
            double yl = q4(F(M, i, j - 1, k), F(M, i, j, k), F(V, i - 1, j, k), F(V, i, j, k));
            double yh = q4(F(M, i, j, k), F(M, i, j + 1, k), F(V, i - 1, j + 1, k), F(V, i, j + 1, k));
            double s = (xh - xl) * dxinv + (yh - yl) * dxinv;
            if (dm == 3) {
              double zl = q4(F(M, i, j, k - 1), F(M, i, j, k), F(W, i - 1, j, k), F(W, i, j, k));
              double zh = q4(F(M, i, j, k), F(M, i, j, k + 1), F(W, i - 1, j, k + 1), F(W, i, j, k + 1));
              s = s + (zh - zl) * dxinv;
            }
            fd = -s;
          } else if (a == 1) {
            double xl = q4(F(M, i - 1, j, k), F(M, i, j, k), F(U, i, j - 1, k), F(U, i, j, k));
            double xh = q4(F(M, i, j, k), F(M, i + 1, j, k), F(U, i + 1, j - 1, k), F(U, i + 1, j, k));
            double yl = q4(F(M, i, j - 1, k), F(M, i, j, k), F(V, i, j - 1, k), F(V, i, j, k));
            double yh = q4(F(M, i, j, k), F(M, i, j + 1, k), F(V, i, j, k), F(V, i, j + 1, k));
            double s = (xh - xl) * dxinv + (yh - yl) * dxinv;
            if (dm == 3) {
              double zl = q4(F(M, i, j, k - 1), F(M, i, j, k), F(W, i, j - 1, k), F(W, i, j, k));
              double zh = q4(F(M, i, j, k), F(M, i, j, k + 1), F(W, i, j - 1, k + 1), F(W, i, j, k + 1));
              s = s + (zh - zl) * dxinv;
            }
            fd = -s;
          } else {
            double xl = q4(F(M, i - 1, j, k), F(M, i, j, k), F(U, i, j, k - 1), F(U, i, j, k));
            double xh = q4(F(M, i, j, k), F(M, i + 1, j, k), F(U, i + 1, j, k - 1), F(U, i + 1, j, k));
            double yl = q4(F(M, i, j - 1, k), F(M, i, j, k), F(V, i, j, k - 1), F(V, i, j, k));
            double yh = q4(F(M, i, j, k), F(M, i, j + 1, k), F(V, i, j + 1, k - 1), F(V, i, j + 1, k));
            double zl = q4(F(M, i, j, k - 1), F(M, i, j, k), F(W, i, j, k - 1), F(W, i, j, k));
            double zh = q4(F(M, i, j, k), F(M, i, j, k + 1), F(W, i, j, k), F(W, i, j, k + 1));
            fd = -((xh - xl) * dxinv + (yh - yl) * dxinv + (zh - zl) * dxinv);
          }
          F(&up[a], i, j, k) = F(&up[a], i, j, k) + fd;
        }
  }
  int rc = 0;
  for (int a = 0; a < dm; a++)
    if (fab_to_dedup(&up[a], m_update[a]) != 0.0) rc = 3;
  faces_drop(u, &g);
  faces_drop(mm, &g);
  faces_drop(up, &g);
  return rc;
}

/* mk_grav_force_2d/3d (mk_grav_force.f90:96-159):
 * m_force += 0.5d0*grav(d)*(rho_old + rho_new) on face component d. */
int orc_mk_grav_force(const orc_geom *g0, const double grav[3], double *const s_old[3], double *const s_new[3],
                      double *m_force[3]) {
  orc_geom g = gnorm(g0);
  const long N = (long)g.n[0] * g.n[1] * g.n[2];
  for (int d = 0; d < g.dim; d++)
    for (long q = 0; q < N; q++) m_force[d][q] = m_force[d][q] + 0.5 * grav[d] * (s_old[d][q] + s_new[d][q]);
  return 0;
}
