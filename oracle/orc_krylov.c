/*
 * Stokes operator, preconditioner and GMRES of the CPU oracle (test
 * infrastructure only).
 *   o_apply_matrix  <- src_gmres/apply_matrix.f90:23-122
 *   o_apply_precon  <- src_gmres/apply_precon.f90:29-195 (precon_type 1,
 *                      visc_schur_approx 0) and :434-460 (null spaces)
 *   o_gmres         <- src_gmres/gmres.f90:30-607
 * plus the deduplicated-array entry points declared in lmm_oracle.h.
 */
#include "orc_fab.h"
#include <stdio.h>

static int nedges(int dim) { return dim == 2 ? 1 : 3; }

static void faces_build(ofab *u, const orc_geom *g, int ng) {
  for (int d = 0; d < g->dim; d++) face_build(&u[d], g, d, ng);
}
static void faces_free(ofab *u, const orc_geom *g) {
  for (int d = 0; d < g->dim; d++) fab_free(&u[d]);
}

void o_apply_matrix(const orc_geom *g, const orc_params *p, ocoef *c, double theta_alpha,
                    ofab *b_u, ofab *b_p, ofab *x_u, ofab *x_p) {
  int dm = g->dim;
  fab_fill_boundary(x_p);
  for (int d = 0; d < dm; d++) fab_fill_boundary(&x_u[d]);
  ofab gx[3], at[3];
  faces_build(gx, g, 0);
  faces_build(at, g, 0);
  /* stag_applyop: alpha_fc_temp = alpha_fc * theta_alpha */
  for (int d = 0; d < dm; d++) {
    fab_copy_valid(&at[d], &c->alpha[d]);
    ofab *a = &at[d];
    for (int k = 0; k <= vhi(a, 2); k++)
      for (int j = 0; j <= vhi(a, 1); j++)
        for (int i = 0; i <= vhi(a, 0); i++) F(a, i, j, k) = F(a, i, j, k) * theta_alpha;
  }
  o_stag_applyop_level(g, p->visc_type, x_u, b_u, at, &c->beta, c->bed, &c->gamma, 0);
  o_compute_grad(g, x_p, gx);
  for (int d = 0; d < dm; d++) {
    ofab *b = &b_u[d], *q = &gx[d];
    for (int k = 0; k <= vhi(b, 2); k++)
      for (int j = 0; j <= vhi(b, 1); j++)
        for (int i = 0; i <= vhi(b, 0); i++) F(b, i, j, k) = F(b, i, j, k) + F(q, i, j, k);
  }
  o_compute_div(g, x_u, b_p, 0);
  int nk = (dm == 3) ? g->n[2] : 1;
  for (int k = 0; k < nk; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) F(b_p, i, j, k) = F(b_p, i, j, k) * -1.0;
  faces_free(gx, g);
  faces_free(at, g);
}

void o_apply_precon(const orc_geom *g, const orc_params *p, ocoef *c, double theta_alpha,
                    ofab *b_u, ofab *b_p, ofab *x_u, ofab *x_p) {
  int dm = g->dim;
  int nk = (dm == 3) ? g->n[2] : 1;
  ofab phi, mac_rhs, ainv[3];
  cell_build(&phi, g, 1);
  cell_build(&mac_rhs, g, 0);
  faces_build(ainv, g, 0);
  for (int d = 0; d < dm; d++) {
    ofab *a = &ainv[d];
    for (int k = 0; k <= vhi(a, 2); k++)
      for (int j = 0; j <= vhi(a, 1); j++)
        for (int i = 0; i <= vhi(a, 0); i++) F(a, i, j, k) = 1.0 / F(&c->alpha[d], i, j, k);
  }
  fab_setval(&phi, 0.0);
  for (int d = 0; d < dm; d++) fab_setval(&x_u[d], 0.0);
  /* STEP 1: x_u* = A^{-1} b_u */
  o_stag_mg_solver(g, p, c, theta_alpha, x_u, b_u);
  /* STEP 2: mac_rhs = D x_u* + b_p */
  o_compute_div(g, x_u, &mac_rhs, 0);
  for (int k = 0; k < nk; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) F(&mac_rhs, i, j, k) = F(&mac_rhs, i, j, k) + F(b_p, i, j, k);
  /* STEP 3: solve div(alpha^-1 grad phi) = mac_rhs; x_u = x_u* - alpha^-1 grad phi */
  o_macproject(g, p, &phi, ainv, &mac_rhs);
  {
    ofab gp[3];
    faces_build(gp, g, 1);
    o_compute_grad(g, &phi, gp);
    for (int d = 0; d < dm; d++) {
      ofab *u = &x_u[d], *q = &gp[d], *a = &ainv[d];
      for (int k = 0; k <= vhi(u, 2); k++)
        for (int j = 0; j <= vhi(u, 1); j++)
          for (int i = 0; i <= vhi(u, 0); i++) {
            double gq = F(q, i, j, k) * F(a, i, j, k);
            F(u, i, j, k) = F(u, i, j, k) + -1.0 * gq;
          }
    }
    faces_free(gp, g);
  }
  /* STEP 4: x_p = theta_alpha*phi - c*beta*mac_rhs */
  int av = abs(p->visc_type);
  for (int k = 0; k < nk; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        double xp = F(&mac_rhs, i, j, k) * -1.0;
        if (av == 1 || av == 2) {
          xp = xp * F(&c->beta, i, j, k);
          if (av == 2) xp = xp * 2.0;
        } else {
          double t = xp * F(&c->gamma, i, j, k);
          xp = xp * F(&c->beta, i, j, k);
          xp = xp * (4.0 / 3.0);
          xp = xp + t;
        }
        double ph = F(&phi, i, j, k) * theta_alpha;
        F(&phi, i, j, k) = ph;
        F(x_p, i, j, k) = xp + ph;
      }
  /* null spaces */
  double mean_p, mean_u[3];
  o_sum_umac_press(g, x_p, x_u, &mean_p, mean_u);
  for (int k = 0; k < nk; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) F(x_p, i, j, k) = F(x_p, i, j, k) - mean_p;
  if (theta_alpha == 0.0) {
    for (int d = 0; d < dm; d++) {
      ofab *u = &x_u[d];
      for (int k = 0; k <= vhi(u, 2); k++)
        for (int j = 0; j <= vhi(u, 1); j++)
          for (int i = 0; i <= vhi(u, 0); i++) F(u, i, j, k) = F(u, i, j, k) - mean_u[d];
    }
  }
  fab_free(&phi);
  fab_free(&mac_rhs);
  faces_free(ainv, g);
}

/* ---------------- GMRES host-scalar kernels (gmres.f90:517-605) ---------------- */

#define H_(r, cidx) H[((r)-1) + (long)(m + 1) * ((cidx)-1)]

static void rotmat(double a, double b, double *cs, double *sn) {
  if (b == 0.0) {
    *cs = 1.0;
    *sn = 0.0;
  } else if (fabs(b) > fabs(a)) {
    double temp = a / b;
    *sn = 1.0 / sqrt(1.0 + temp * temp);
    *cs = temp * *sn;
  } else {
    double temp = b / a;
    *cs = 1.0 / sqrt(1.0 + temp * temp);
    *sn = temp * *cs;
  }
}

void orc_least_squares(int i, int m, double *H, double *cs, double *sn, double *s) {
  for (int k = 1; k <= i - 1; k++) {
    double temp = cs[k - 1] * H_(k, i) + sn[k - 1] * H_(k + 1, i);
    H_(k + 1, i) = -sn[k - 1] * H_(k, i) + cs[k - 1] * H_(k + 1, i);
    H_(k, i) = temp;
  }
  rotmat(H_(i, i), H_(i + 1, i), &cs[i - 1], &sn[i - 1]);
  double temp = cs[i - 1] * s[i - 1];
  s[i] = -sn[i - 1] * s[i - 1];
  s[i - 1] = temp;
  H_(i, i) = cs[i - 1] * H_(i, i) + sn[i - 1] * H_(i + 1, i);
  H_(i + 1, i) = 0.0;
}

void orc_solve_utriangular(int k, int m, const double *H, const double *s, double *y) {
  y[k] = s[k] / H_(k + 1, k + 1);
  for (int i = k; i >= 1; i--) {
    double dp = 0.0;
    for (int q = i + 1; q <= k + 1; q++) dp = dp + H_(i, q) * y[q - 1];
    y[i - 1] = (s[i - 1] - dp) / H_(i, i);
  }
}

/* ---------------- GMRES (gmres.f90:30-607) ---------------- */

typedef struct {
  ofab u[3];
  ofab p;
} svec;

static void svec_build(svec *v, const orc_geom *g, int ngu, int ngp) {
  faces_build(v->u, g, ngu);
  cell_build(&v->p, g, ngp);
}
static void svec_free(svec *v, const orc_geom *g) {
  faces_free(v->u, g);
  fab_free(&v->p);
}

static void norms(const orc_geom *g, const orc_params *p, svec *v, double *nu, double *np,
                  double *nt) {
  double ip[3] = {0, 0, 0};
  o_stag_inner_prod(g, v->u, v->u, ip);
  double su = 0.0;
  for (int d = 0; d < g->dim; d++) su = su + ip[d];
  *nu = sqrt(su);
  *np = sqrt(o_cc_inner_prod(g, &v->p, &v->p));
  *np = p->p_norm_weight * *np;
  *nt = sqrt((*nu) * (*nu) + (*np) * (*np));
}

/* dst = op(dst, src) over the valid region of every component */
static void sv_copy(const orc_geom *g, svec *dst, svec *src) {
  for (int d = 0; d < g->dim; d++) fab_copy_valid(&dst->u[d], &src->u[d]);
  fab_copy_valid(&dst->p, &src->p);
}

static void fab_map(ofab *a, ofab *b, double s, int op) {
  for (int k = 0; k <= vhi(a, 2); k++)
    for (int j = 0; j <= vhi(a, 1); j++)
      for (int i = 0; i <= vhi(a, 0); i++) {
        double x = F(a, i, j, k);
        switch (op) {
          case 0: x = (x - F(b, i, j, k)) * -1.0; break;   /* a = -(a - b)   */
          case 1: x = x / s; break;                         /* a = a / s      */
          case 2: x = x - F(b, i, j, k) * s; break;         /* a = a - s*b    */
          case 3: x = x * s; break;                         /* a = a * s      */
        }
        F(a, i, j, k) = x;
      }
}

static void sv_map(const orc_geom *g, svec *a, svec *b, double s, int op) {
  for (int d = 0; d < g->dim; d++) fab_map(&a->u[d], b ? &b->u[d] : NULL, s, op);
  fab_map(&a->p, b ? &b->p : NULL, s, op);
}

static void scale_coefs(const orc_geom *g, ocoef *c, double s) {
  fab_scale_all(&c->beta, s);
  fab_scale_all(&c->gamma, s);
  for (int e = 0; e < nedges(g->dim); e++) fab_scale_all(&c->bed[e], s);
}

static int o_gmres(const orc_geom *g, const orc_params *p, ocoef *c, double *theta_alpha_io,
                   svec *b, svec *x, orc_gmres_info *info) {
  int dm = g->dim, m = p->gmres_max_inner;
  double theta_alpha = *theta_alpha_io;
  svec r, w, tmp;
  svec *V = (svec *)calloc((size_t)m + 1, sizeof(svec));
  svec_build(&r, g, 1, 1);
  svec_build(&w, g, 0, 0);
  svec_build(&tmp, g, 0, 0);
  for (int q = 0; q <= m; q++) svec_build(&V[q], g, 0, 0);
  double *cs = calloc((size_t)m, sizeof(double)), *sn = calloc((size_t)m, sizeof(double));
  double *y = calloc((size_t)m, sizeof(double)), *s = calloc((size_t)m + 1, sizeof(double));
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double nu, np, norm_pre_b, norm_b, norm_resid = 0.0, norm_init_resid = 0.0;
  double norm_resid_Stokes = 0.0, norm_init_Stokes = 0.0, norm_resid_est = 0.0;
  int status = 0;

  if (p->scale_factor != 1.0) {
    theta_alpha = theta_alpha * p->scale_factor;
    fab_scale_all(&x->p, p->scale_factor);
    for (int d = 0; d < dm; d++) fab_scale_all(&b->u[d], p->scale_factor);
    scale_coefs(g, c, p->scale_factor);
  }
  o_apply_precon(g, p, c, theta_alpha, b->u, &b->p, tmp.u, &tmp.p);
  norms(g, p, &tmp, &nu, &np, &norm_pre_b);
  info->norm_pre_rhs = norm_pre_b;
  norms(g, p, b, &nu, &np, &norm_b);
  info->norm_b = norm_b;
  int total_iter = 0, iter = 0, i = 0, i_copy = 0;
  if (norm_b <= 0.0) {
    for (int d = 0; d < dm; d++) fab_setval(&x->u[d], 0.0);
    fab_setval(&x->p, 0.0);
  } else {
    for (;;) {
      o_apply_matrix(g, p, c, theta_alpha, tmp.u, &tmp.p, x->u, &x->p);
      sv_map(g, &tmp, b, 0.0, 0);
      norms(g, p, &tmp, &nu, &np, &norm_resid_Stokes);
      if (iter == 0) norm_init_Stokes = norm_resid_Stokes;
      o_apply_precon(g, p, c, theta_alpha, tmp.u, &tmp.p, r.u, &r.p);
      norms(g, p, &r, &nu, &np, &norm_resid);
      if (iter == 0) {
        norm_init_resid = norm_resid;
        norm_resid_est = norm_resid;
      }
      if (p->verbose >= 1)
        printf("oracle gmres outer %d total %d |Pr| %.10e |r| %.10e\n", iter, total_iter, norm_resid,
               norm_resid_Stokes);
      if (total_iter >= p->gmres_max_iter) break;
      if (total_iter >= p->gmres_min_iter) {
        double mn = norm_pre_b < norm_init_resid ? norm_pre_b : norm_init_resid;
        if (norm_resid <= p->gmres_rel_tol * mn) break;
        if (norm_resid <= p->gmres_abs_tol) break;
      }
      if (iter >= p->gmres_max_outer) break;
      iter = iter + 1;
      sv_copy(g, &V[0], &r);
      sv_map(g, &V[0], NULL, norm_resid, 1);
      for (int q = 0; q <= m; q++) s[q] = 0.0;
      s[0] = norm_resid;
      for (i = 1; i <= m; i++) {
        total_iter = total_iter + 1;
        i_copy = i;
        sv_copy(g, &r, &V[i - 1]);
        o_apply_matrix(g, p, c, theta_alpha, tmp.u, &tmp.p, r.u, &r.p);
        o_apply_precon(g, p, c, theta_alpha, tmp.u, &tmp.p, w.u, &w.p);
        for (int k = 1; k <= i; k++) {
          double ipv[3] = {0, 0, 0};
          o_stag_inner_prod(g, w.u, V[k - 1].u, ipv);
          double ipp = o_cc_inner_prod(g, &w.p, &V[k - 1].p);
          double sv = 0.0;
          for (int d = 0; d < dm; d++) sv = sv + ipv[d];
          H_(k, i) = sv + (p->p_norm_weight * p->p_norm_weight) * ipp;
          sv_map(g, &w, &V[k - 1], H_(k, i), 2);
        }
        norms(g, p, &w, &nu, &np, &H_(i + 1, i));
        if (H_(i + 1, i) != 0.0) {
          sv_copy(g, &V[i], &w);
          sv_map(g, &V[i], NULL, H_(i + 1, i), 1);
        } else {
          status = 1;
          break;
        }
        orc_least_squares(i, m, H, cs, sn, s);
        norm_resid_est = fabs(s[i]);
        if (p->verbose >= 2) printf("oracle gmres inner %d est %.10e\n", total_iter, norm_resid_est);
        if (total_iter >= p->gmres_max_iter) break;
        if (total_iter >= p->gmres_min_iter) {
          double mn = norm_pre_b < norm_init_resid ? norm_pre_b : norm_init_resid;
          if (norm_resid_est <= p->gmres_rel_tol * mn || norm_resid_est <= p->gmres_abs_tol) break;
        }
      }
      if (status) break;
      orc_solve_utriangular(i_copy - 1, m, H, s, y);
      /* update_sol: V(k) *= y(k); x += V(k), pressure first then velocity */
      for (int k = 1; k <= i_copy; k++) {
        fab_map(&V[k - 1].p, NULL, y[k - 1], 3);
        fab_map(&x->p, &V[k - 1].p, -1.0, 2);
        for (int d = 0; d < dm; d++) {
          fab_map(&V[k - 1].u[d], NULL, y[k - 1], 3);
          fab_map(&x->u[d], &V[k - 1].u[d], -1.0, 2);
        }
      }
    }
  }
  for (int d = 0; d < dm; d++) fab_internal_sync(&x->u[d]);
  if (p->scale_factor != 1.0) {
    theta_alpha = theta_alpha / p->scale_factor;
    fab_scale_all(&x->p, 1.0 / p->scale_factor);
    for (int d = 0; d < dm; d++) fab_scale_all(&b->u[d], 1.0 / p->scale_factor);
    scale_coefs(g, c, 1.0 / p->scale_factor);
  }
  *theta_alpha_io = theta_alpha;
  info->total_iter = total_iter;
  info->outer_iter = iter;
  info->norm_resid = norm_resid;
  info->norm_resid_true = norm_resid_Stokes;
  info->status = status;
  (void)norm_init_Stokes;
  svec_free(&r, g);
  svec_free(&w, g);
  svec_free(&tmp, g);
  for (int q = 0; q <= m; q++) svec_free(&V[q], g);
  free(V);
  free(cs);
  free(sn);
  free(y);
  free(s);
  free(H);
  return status;
}

/* ---------------- deduplicated-array entry points ---------------- */

static int geom_ok(const orc_geom *g) {
  if (g->dim != 2 && g->dim != 3) return 0;
  for (int d = 0; d < g->dim; d++)
    if (g->n[d] < 2) return 0;
  return 1;
}

static orc_geom geom_norm(const orc_geom *g) {
  orc_geom h = *g;
  if (h.dim == 2) h.n[2] = 1;
  return h;
}

int orc_compute_nlevs_mg(const orc_geom *g, int minwidth) { return o_compute_nlevs(g, minwidth); }

static void load_faces(ofab *u, const orc_geom *g, int ng, double *const src[3]) {
  for (int d = 0; d < g->dim; d++) {
    face_build(&u[d], g, d, ng);
    if (src && src[d]) fab_from_dedup(&u[d], src[d]);
  }
}
static double store_faces(ofab *u, const orc_geom *g, double *dst[3]) {
  double mm = 0.0;
  for (int d = 0; d < g->dim; d++) {
    double e = fab_to_dedup(&u[d], dst[d]);
    if (e > mm) mm = e;
  }
  return mm;
}

int orc_stag_applyop(const orc_geom *g0, int visc, const orc_coef *c, double theta_alpha,
                     double *const phi[3], double *Lphi[3], int color) {
  if (!geom_ok(g0)) return 2;
  orc_geom g = geom_norm(g0);
  ocoef oc;
  ocoef_build(&oc, &g, c);
  ofab ph[3], lp[3];
  load_faces(ph, &g, 1, phi);
  load_faces(lp, &g, 1, (double *const *)Lphi);
  for (int d = 0; d < g.dim; d++) {
    ofab *a = &oc.alpha[d];
    for (int k = 0; k <= vhi(a, 2); k++)
      for (int j = 0; j <= vhi(a, 1); j++)
        for (int i = 0; i <= vhi(a, 0); i++) F(a, i, j, k) = F(a, i, j, k) * theta_alpha;
  }
  o_stag_applyop_level(&g, visc, ph, lp, oc.alpha, &oc.beta, oc.bed, &oc.gamma, color);
  store_faces(lp, &g, Lphi);
  for (int d = 0; d < g.dim; d++) {
    fab_free(&ph[d]);
    fab_free(&lp[d]);
  }
  ocoef_free(&oc, &g);
  return 0;
}

int orc_stag_mg_smooth(const orc_geom *g0, const orc_params *p, const orc_coef *c,
                       double theta_alpha, double *phi[3], double *const rhs[3], int nsweeps) {
  if (!geom_ok(g0)) return 2;
  orc_geom g = geom_norm(g0);
  ocoef oc;
  ocoef_build(&oc, &g, c);
  ofab ph[3], rh[3], lp[3];
  load_faces(ph, &g, 1, (double *const *)phi);
  load_faces(rh, &g, 0, rhs);
  load_faces(lp, &g, 1, NULL);
  for (int d = 0; d < g.dim; d++) {
    ofab *a = &oc.alpha[d];
    for (int k = 0; k <= vhi(a, 2); k++)
      for (int j = 0; j <= vhi(a, 1); j++)
        for (int i = 0; i <= vhi(a, 0); i++) F(a, i, j, k) = F(a, i, j, k) * theta_alpha;
  }
  int cs = (p->stag_mg_smoother == 0) ? 0 : 1, ce = (p->stag_mg_smoother == 0) ? 0 : 2 * g.dim;
  for (int m = 0; m < nsweeps; m++)
    for (int color = cs; color <= ce; color++) {
      o_stag_applyop_level(&g, p->visc_type, ph, lp, oc.alpha, &oc.beta, oc.bed, &oc.gamma, color);
      o_stag_mg_update(&g, p->visc_type, p->stag_mg_omega, ph, rh, lp, oc.alpha, &oc.beta, oc.bed,
                       &oc.gamma, color);
      for (int d = 0; d < g.dim; d++) fab_fill_boundary(&ph[d]);
    }
  int rc = store_faces(ph, &g, phi) == 0.0 ? 0 : 3;
  for (int d = 0; d < g.dim; d++) {
    fab_free(&ph[d]);
    fab_free(&rh[d]);
    fab_free(&lp[d]);
  }
  ocoef_free(&oc, &g);
  return rc;
}

static orc_geom coarse_of(const orc_geom *gf) {
  orc_geom gc = geom_norm(gf);
  for (int d = 0; d < gc.dim; d++) gc.n[d] /= 2;
  gc.dx = gf->dx * 2.0;
  return gc;
}

int orc_stag_restriction(const orc_geom *gf0, double *const fine[3], double *coarse[3], int simple) {
  if (!geom_ok(gf0)) return 2;
  orc_geom gf = geom_norm(gf0), gc = coarse_of(&gf);
  ofab f[3], c[3];
  load_faces(f, &gf, 1, fine);
  load_faces(c, &gc, 0, NULL);
  o_stag_restriction(&gc, f, c, simple);
  int rc = store_faces(c, &gc, coarse) == 0.0 ? 0 : 3;
  faces_free(f, &gf);
  faces_free(c, &gc);
  return rc;
}

int orc_stag_prolongation(const orc_geom *gf0, double *fine[3], double *const coarse[3]) {
  if (!geom_ok(gf0)) return 2;
  orc_geom gf = geom_norm(gf0), gc = coarse_of(&gf);
  ofab f[3], c[3];
  load_faces(f, &gf, 1, (double *const *)fine);
  load_faces(c, &gc, 1, coarse);
  o_stag_prolongation(&gf, f, c);
  int rc = store_faces(f, &gf, fine) == 0.0 ? 0 : 3;
  faces_free(f, &gf);
  faces_free(c, &gc);
  return rc;
}

int orc_cc_restriction(const orc_geom *gf0, const double *fine, double *coarse) {
  orc_geom gf = geom_norm(gf0), gc = coarse_of(&gf);
  ofab f, c;
  cell_build(&f, &gf, 1);
  fab_from_dedup(&f, fine);
  cell_build(&c, &gc, 1);
  o_cc_restriction(&gc, &c, &f);
  fab_to_dedup(&c, coarse);
  fab_free(&f);
  fab_free(&c);
  return 0;
}

int orc_edge_restriction(const orc_geom *gf0, double *const fine[3], double *coarse[3]) {
  orc_geom gf = geom_norm(gf0), gc = coarse_of(&gf);
  if (gf.dim != 3) return 2;
  ofab f[3], c[3];
  for (int e = 0; e < 3; e++) {
    edge_build(&f[e], &gf, e, 0);
    fab_from_dedup(&f[e], fine[e]);
    edge_build(&c[e], &gc, e, 0);
  }
  o_edge_restriction(&gc, f, c);
  double mm = 0.0;
  for (int e = 0; e < 3; e++) {
    double x = fab_to_dedup(&c[e], coarse[e]);
    if (x > mm) mm = x;
    fab_free(&f[e]);
    fab_free(&c[e]);
  }
  return mm == 0.0 ? 0 : 3;
}

int orc_nodal_restriction(const orc_geom *gf0, const double *fine, double *coarse) {
  orc_geom gf = geom_norm(gf0), gc = coarse_of(&gf);
  if (gf.dim != 2) return 2;
  ofab f, c;
  edge_build(&f, &gf, 0, 0);
  fab_from_dedup(&f, fine);
  edge_build(&c, &gc, 0, 0);
  o_nodal_restriction(&gc, &f, &c);
  double mm = fab_to_dedup(&c, coarse);
  fab_free(&f);
  fab_free(&c);
  return mm == 0.0 ? 0 : 3;
}

int orc_compute_div(const orc_geom *g0, double *const u[3], double *div, int increment) {
  orc_geom g = geom_norm(g0);
  ofab uf[3], dv;
  load_faces(uf, &g, 1, u);
  cell_build(&dv, &g, 0);
  if (increment) fab_from_dedup(&dv, div);
  o_compute_div(&g, uf, &dv, increment);
  fab_to_dedup(&dv, div);
  faces_free(uf, &g);
  fab_free(&dv);
  return 0;
}

int orc_compute_grad(const orc_geom *g0, const double *p, double *gp[3]) {
  orc_geom g = geom_norm(g0);
  ofab pc, gf[3];
  cell_build(&pc, &g, 1);
  fab_from_dedup(&pc, p);
  load_faces(gf, &g, 0, NULL);
  o_compute_grad(&g, &pc, gf);
  int rc = store_faces(gf, &g, gp) == 0.0 ? 0 : 3;
  fab_free(&pc);
  faces_free(gf, &g);
  return rc;
}

int orc_stag_inner_prod(const orc_geom *g0, double *const m1[3], double *const m2[3], double out[3]) {
  orc_geom g = geom_norm(g0);
  ofab a[3], b[3];
  load_faces(a, &g, 0, m1);
  load_faces(b, &g, 0, m2);
  o_stag_inner_prod(&g, a, b, out);
  faces_free(a, &g);
  faces_free(b, &g);
  return 0;
}

double orc_cc_inner_prod(const orc_geom *g0, const double *m1, const double *m2) {
  orc_geom g = geom_norm(g0);
  ofab a, b;
  cell_build(&a, &g, 0);
  cell_build(&b, &g, 0);
  fab_from_dedup(&a, m1);
  fab_from_dedup(&b, m2);
  double r = o_cc_inner_prod(&g, &a, &b);
  fab_free(&a);
  fab_free(&b);
  return r;
}

int orc_sum_umac_press(const orc_geom *g0, const double *p, double *const u[3], double *av_p,
                       double av_u[3]) {
  orc_geom g = geom_norm(g0);
  ofab pc, uf[3];
  cell_build(&pc, &g, 0);
  fab_from_dedup(&pc, p);
  load_faces(uf, &g, 0, u);
  o_sum_umac_press(&g, &pc, uf, av_p, av_u);
  fab_free(&pc);
  faces_free(uf, &g);
  return 0;
}

int orc_stag_mg_solver(const orc_geom *g0, const orc_params *p, const orc_coef *c,
                       double theta_alpha, double *phi[3], double *const rhs[3]) {
  if (!geom_ok(g0)) return 2;
  orc_geom g = geom_norm(g0);
  ocoef oc;
  ocoef_build(&oc, &g, c);
  ofab ph[3], rh[3];
  load_faces(ph, &g, 1, (double *const *)phi);
  load_faces(rh, &g, 0, rhs);
  o_stag_mg_solver(&g, p, &oc, theta_alpha, ph, rh);
  int rc = store_faces(ph, &g, phi) == 0.0 ? 0 : 3;
  faces_free(ph, &g);
  faces_free(rh, &g);
  ocoef_free(&oc, &g);
  return rc;
}

int orc_macproject(const orc_geom *g0, const orc_params *p, double *const alphainv[3],
                   const double *mac_rhs, double *phi) {
  if (!geom_ok(g0)) return 2;
  orc_geom g = geom_norm(g0);
  ofab ai[3], rhs, ph;
  load_faces(ai, &g, 0, alphainv);
  cell_build(&rhs, &g, 0);
  fab_from_dedup(&rhs, mac_rhs);
  cell_build(&ph, &g, 1);
  o_macproject(&g, p, &ph, ai, &rhs);
  fab_to_dedup(&ph, phi);
  faces_free(ai, &g);
  fab_free(&rhs);
  fab_free(&ph);
  return 0;
}

int orc_apply_matrix(const orc_geom *g0, const orc_params *p, const orc_coef *c, double theta_alpha,
                     double *const x_u[3], const double *x_p, double *b_u[3], double *b_p) {
  if (!geom_ok(g0)) return 2;
  orc_geom g = geom_norm(g0);
  ocoef oc;
  ocoef_build(&oc, &g, c);
  ofab xu[3], bu[3], xp, bp;
  load_faces(xu, &g, 1, x_u);
  load_faces(bu, &g, 0, NULL);
  cell_build(&xp, &g, 1);
  fab_from_dedup(&xp, x_p);
  cell_build(&bp, &g, 0);
  o_apply_matrix(&g, p, &oc, theta_alpha, bu, &bp, xu, &xp);
  int rc = store_faces(bu, &g, b_u) == 0.0 ? 0 : 3;
  fab_to_dedup(&bp, b_p);
  faces_free(xu, &g);
  faces_free(bu, &g);
  fab_free(&xp);
  fab_free(&bp);
  ocoef_free(&oc, &g);
  return rc;
}

int orc_apply_precon(const orc_geom *g0, const orc_params *p, const orc_coef *c, double theta_alpha,
                     double *const b_u[3], const double *b_p, double *x_u[3], double *x_p) {
  if (!geom_ok(g0)) return 2;
  if (p->precon_type != 1) return 2;
  orc_geom g = geom_norm(g0);
  ocoef oc;
  ocoef_build(&oc, &g, c);
  ofab bu[3], xu[3], bp, xp;
  load_faces(bu, &g, 0, b_u);
  load_faces(xu, &g, 1, NULL);
  cell_build(&bp, &g, 0);
  fab_from_dedup(&bp, b_p);
  cell_build(&xp, &g, 1);
  o_apply_precon(&g, p, &oc, theta_alpha, bu, &bp, xu, &xp);
  int rc = store_faces(xu, &g, x_u) == 0.0 ? 0 : 3;
  fab_to_dedup(&xp, x_p);
  faces_free(xu, &g);
  faces_free(bu, &g);
  fab_free(&xp);
  fab_free(&bp);
  ocoef_free(&oc, &g);
  return rc;
}

int orc_gmres(const orc_geom *g0, const orc_params *p, const orc_coef *c, double theta_alpha,
              double *b_u[3], const double *b_p, double *x_u[3], double *x_p, orc_gmres_info *info) {
  memset(info, 0, sizeof(*info));
  if (!geom_ok(g0) || p->precon_type != 1 || p->gmres_max_inner < 1) {
    info->status = 2;
    return 2;
  }
  orc_geom g = geom_norm(g0);
  ocoef oc;
  ocoef_build(&oc, &g, c);
  svec b, x;
  svec_build(&b, &g, 0, 0);
  svec_build(&x, &g, 1, 1);
  for (int d = 0; d < g.dim; d++) {
    fab_from_dedup(&b.u[d], b_u[d]);
    fab_from_dedup(&x.u[d], x_u[d]);
  }
  fab_from_dedup(&b.p, b_p);
  fab_from_dedup(&x.p, x_p);
  double ta = theta_alpha;
  int rc = o_gmres(&g, p, &oc, &ta, &b, &x, info);
  info->dup_mismatch = store_faces(x.u, &g, x_u);
  fab_to_dedup(&x.p, x_p);
  svec_free(&b, &g);
  svec_free(&x, &g);
  ocoef_free(&oc, &g);
  return rc;
}
