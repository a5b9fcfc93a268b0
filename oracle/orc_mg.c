/*
 * Multigrid solvers of the CPU oracle (test infrastructure only).
 *
 * o_stag_mg_solver restates src_gmres/stag_mg_solver.f90:37-868 (single box,
 * periodic, no "fancy bottom": that path needs n_cells/max_grid_size >= 2,
 * i.e. more than one box, stag_mg_solver.f90:450).
 * o_compute_nlevs restates compute_nlevs_mg (stag_mg_layout.f90:179-222).
 *
 * o_macproject restates src_gmres/macproject.f90:34-318.  Its inner
 * cell-centred multigrid is BoxLib F_MG, which is not in the reference tree;
 * it is restated from the method in doc/StokesPreconditioners.pdf, Appendix
 * C (pp. 23-24): V-cycle in residual-correction form, coarsening by 2 down to
 * mg_minwidth cells, red-black Gauss-Seidel with omega = 1 (red = even i+j+k
 * first), restriction by averaging the 2^d fine cells, prolongation by
 * injection, mg_nsmooths_bottom relaxations on the coarsest level, face
 * coefficients coarsened by averaging the overlying fine faces, and the
 * tower settings of mgt_macproj_precon_build (macproject.f90:366-430):
 * nu1 = mg_nsmooths_down, nu2 = mg_nsmooths_up, at most mg_max_vcycles cycles
 * with F_MG's stopping test (max|r| <= mg_rel_tol*max|rhs| or <= 1e-16).
 * Bit-level agreement with F_MG itself is not pinned by anything in the tree.
 */
#include "orc_fab.h"
#include <stdio.h>

int o_compute_nlevs(const orc_geom *g, int minwidth) {
  int nl = -1;
  for (int d = 0; d < g->dim; d++) {
    int temp = g->n[d], rdir = 1;
    while (temp % 2 == 0 && temp / minwidth != 1) {
      temp /= 2;
      rdir++;
    }
    nl = (nl == -1) ? rdir : (rdir < nl ? rdir : nl);
  }
  return nl;
}

typedef struct {
  orc_geom g;
  ofab alpha[3], rhs[3], phi[3], Lphi[3], resid[3];
  ofab beta, gamma, bed[3];
} slevel;

static int nedges(int dim) { return dim == 2 ? 1 : 3; }

static void smooth_level(slevel *L, const orc_params *p, int nsmooths) {
  int dm = L->g.dim;
  int cs = (p->stag_mg_smoother == 0) ? 0 : 1;
  int ce = (p->stag_mg_smoother == 0) ? 0 : 2 * dm;
  for (int m = 0; m < nsmooths; m++)
    for (int color = cs; color <= ce; color++) {
      o_stag_applyop_level(&L->g, p->visc_type, L->phi, L->Lphi, L->alpha, &L->beta, L->bed,
                           &L->gamma, color);
      o_stag_mg_update(&L->g, p->visc_type, p->stag_mg_omega, L->phi, L->rhs, L->Lphi, L->alpha,
                       &L->beta, L->bed, &L->gamma, color);
      for (int j = 0; j < dm; j++) fab_fill_boundary(&L->phi[j]);
    }
}

/* Lphi <- -(Lphi - rhs) = rhs - Lphi, then ghost fill (stag_mg_solver.f90:404-434) */
static void residual_level(slevel *L, const orc_params *p) {
  int dm = L->g.dim;
  o_stag_applyop_level(&L->g, p->visc_type, L->phi, L->Lphi, L->alpha, &L->beta, L->bed,
                       &L->gamma, 0);
  for (int j = 0; j < dm; j++) {
    ofab *l = &L->Lphi[j], *r = &L->rhs[j];
    for (int k = 0; k <= vhi(l, 2); k++)
      for (int jj = 0; jj <= vhi(l, 1); jj++)
        for (int i = 0; i <= vhi(l, 0); i++) {
          double v = F(l, i, jj, k) - F(r, i, jj, k);
          F(l, i, jj, k) = v * -1.0;
        }
    fab_fill_boundary(l);
  }
}

void o_stag_mg_solver(const orc_geom *g, const orc_params *p, ocoef *c, double theta_alpha,
                      ofab *phi_fc, ofab *rhs_fc) {
  int dm = g->dim;
  int nl = o_compute_nlevs(g, p->stag_mg_minwidth);
  slevel *lev = (slevel *)calloc((size_t)nl, sizeof(slevel));
  for (int n = 0; n < nl; n++) {
    slevel *L = &lev[n];
    L->g = *g;
    for (int d = 0; d < dm; d++) L->g.n[d] = g->n[d] >> n;
    L->g.dx = g->dx * (double)(1 << n);
    cell_build(&L->beta, &L->g, 1);
    cell_build(&L->gamma, &L->g, 1);
    for (int d = 0; d < dm; d++) {
      face_build(&L->alpha[d], &L->g, d, 0);
      face_build(&L->rhs[d], &L->g, d, 0);
      face_build(&L->phi[d], &L->g, d, 1);
      face_build(&L->Lphi[d], &L->g, d, 1);
      face_build(&L->resid[d], &L->g, d, 0);
    }
    for (int e = 0; e < nedges(dm); e++) edge_build(&L->bed[e], &L->g, e, 0);
  }
  /* level-1 coefficients: beta/gamma with one ghost cell, alpha*theta */
  {
    slevel *L = &lev[0];
    memcpy(L->beta.p, c->beta.p, sizeof(double) * (size_t)c->beta.sz * c->beta.ext[2]);
    memcpy(L->gamma.p, c->gamma.p, sizeof(double) * (size_t)c->gamma.sz * c->gamma.ext[2]);
    for (int e = 0; e < nedges(dm); e++) fab_copy_valid(&L->bed[e], &c->bed[e]);
    for (int d = 0; d < dm; d++) {
      fab_copy_valid(&L->alpha[d], &c->alpha[d]);
      ofab *a = &L->alpha[d];
      for (int k = 0; k <= vhi(a, 2); k++)
        for (int j = 0; j <= vhi(a, 1); j++)
          for (int i = 0; i <= vhi(a, 0); i++) F(a, i, j, k) = F(a, i, j, k) * theta_alpha;
    }
  }
  for (int n = 1; n < nl; n++) {
    fab_setval(&lev[n].beta, 0.0);
    fab_setval(&lev[n].gamma, 0.0);
    o_cc_restriction(&lev[n].g, &lev[n].beta, &lev[n - 1].beta);
    o_cc_restriction(&lev[n].g, &lev[n].gamma, &lev[n - 1].gamma);
    o_stag_restriction(&lev[n].g, lev[n - 1].alpha, lev[n].alpha, 1);
    if (dm == 2)
      o_nodal_restriction(&lev[n].g, &lev[n - 1].bed[0], &lev[n].bed[0]);
    else
      o_edge_restriction(&lev[n].g, lev[n - 1].bed, lev[n].bed);
  }
  for (int d = 0; d < dm; d++) {
    fab_copy_valid(&lev[0].phi[d], &phi_fc[d]);
    fab_fill_boundary(&lev[0].phi[d]);
    fab_copy_valid(&lev[0].rhs[d], &rhs_fc[d]);
  }
  /* initial residual */
  double resid0[3] = {0, 0, 0}, resid[3];
  o_stag_applyop_level(&lev[0].g, p->visc_type, lev[0].phi, lev[0].Lphi, lev[0].alpha, &lev[0].beta,
                       lev[0].bed, &lev[0].gamma, 0);
  for (int j = 0; j < dm; j++) {
    ofab *l = &lev[0].Lphi[j], *r = &lev[0].rhs[j];
    for (int k = 0; k <= vhi(l, 2); k++)
      for (int jj = 0; jj <= vhi(l, 1); jj++)
        for (int i = 0; i <= vhi(l, 0); i++) F(l, i, jj, k) = F(l, i, jj, k) - F(r, i, jj, k);
    resid0[j] = fab_norm_inf(l);
  }
  int allzero = 1, anyzero = 0;
  double mx = 0.0;
  for (int j = 0; j < dm; j++) {
    if (resid0[j] != 0.0) allzero = 0; else anyzero = 1;
    if (resid0[j] > mx) mx = resid0[j];
  }
  if (!allzero) {
    if (anyzero)
      for (int j = 0; j < dm; j++) resid0[j] = mx;
    for (int vcycle = 1; vcycle <= p->stag_mg_max_vcycles; vcycle++) {
      for (int j = 0; j < dm; j++)
        for (int n = 1; n < nl; n++) fab_setval(&lev[n].phi[j], 0.0);
      for (int n = 0; n < nl - 1; n++) {
        smooth_level(&lev[n], p, p->stag_mg_nsmooths_down);
        residual_level(&lev[n], p);
        o_stag_restriction(&lev[n + 1].g, lev[n].Lphi, lev[n + 1].rhs, 0);
      }
      {
        slevel *L = &lev[nl - 1];
        smooth_level(L, p, p->stag_mg_nsmooths_bottom);
        residual_level(L, p); /* computed by the reference, result unused */
      }
      for (int n = nl - 2; n >= 0; n--) {
        o_stag_prolongation(&lev[n].g, lev[n].phi, lev[n + 1].phi);
        for (int j = 0; j < dm; j++) fab_fill_boundary(&lev[n].phi[j]);
        smooth_level(&lev[n], p, p->stag_mg_nsmooths_up);
      }
      /* convergence test on the finest level */
      o_stag_applyop_level(&lev[0].g, p->visc_type, lev[0].phi, lev[0].Lphi, lev[0].alpha,
                           &lev[0].beta, lev[0].bed, &lev[0].gamma, 0);
      int done = 1;
      for (int j = 0; j < dm; j++) {
        ofab *l = &lev[0].Lphi[j], *r = &lev[0].rhs[j];
        for (int k = 0; k <= vhi(l, 2); k++)
          for (int jj = 0; jj <= vhi(l, 1); jj++)
            for (int i = 0; i <= vhi(l, 0); i++) F(l, i, jj, k) = F(l, i, jj, k) - F(r, i, jj, k);
        resid[j] = fab_norm_inf(l) / resid0[j];
        if (!(resid[j] < p->stag_mg_rel_tol)) done = 0;
      }
      if (p->verbose >= 2) printf("oracle stag_mg vcycle %d resid %g\n", vcycle, resid[0]);
      if (done) break;
    }
    for (int j = 0; j < dm; j++) {
      fab_copy_valid(&phi_fc[j], &lev[0].phi[j]);
      fab_fill_boundary(&phi_fc[j]);
    }
  }
  for (int n = 0; n < nl; n++) {
    slevel *L = &lev[n];
    fab_free(&L->beta);
    fab_free(&L->gamma);
    for (int d = 0; d < dm; d++) {
      fab_free(&L->alpha[d]);
      fab_free(&L->rhs[d]);
      fab_free(&L->phi[d]);
      fab_free(&L->Lphi[d]);
      fab_free(&L->resid[d]);
    }
    for (int e = 0; e < nedges(dm); e++) fab_free(&L->bed[e]);
  }
  free(lev);
}

/* ------------------------------------------------------------------ */
/* cell-centred multigrid for div(a grad phi) = rhs                     */
/* ------------------------------------------------------------------ */

typedef struct {
  orc_geom g;
  ofab a[3];     /* face coefficients, ng 0 */
  ofab phi;      /* ng 1 */
  ofab rhs, res; /* ng 0 */
} clevel;

/* (L phi)(i) = ( -(sum of the 2*dim face coefficients) phi(i)
 *                + a_lo,x phi(i-1) + a_hi,x phi(i+1) + ... ) / dx^2          */
static inline double cc_apply_pt(const clevel *L, int i, int j, int k, double *diag) {
  double inv = 1.0 / (L->g.dx * L->g.dx);
  const ofab *ax = &L->a[0], *ay = &L->a[1], *az = &L->a[2];
  const ofab *p = &L->phi;
  double axl = F(ax, i, j, k), axh = F(ax, i + 1, j, k);
  double ayl = F(ay, i, j, k), ayh = F(ay, i, j + 1, k);
  double s = axl + axh + ayl + ayh;
  double off = axl * F(p, i - 1, j, k) + axh * F(p, i + 1, j, k) + ayl * F(p, i, j - 1, k) +
               ayh * F(p, i, j + 1, k);
  if (L->g.dim == 3) {
    double azl = F(az, i, j, k), azh = F(az, i, j, k + 1);
    s = s + azl + azh;
    off = off + azl * F(p, i, j, k - 1) + azh * F(p, i, j, k + 1);
  }
  *diag = -s * inv;
  return (-s * F(p, i, j, k) + off) * inv;
}

static void cc_smooth(clevel *L, int nsweeps) {
  int nk = (L->g.dim == 3) ? L->g.n[2] : 1;
  for (int m = 0; m < nsweeps; m++)
    for (int colour = 0; colour < 2; colour++) {
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 0; k < nk; k++)
        for (int j = 0; j < L->g.n[1]; j++)
          for (int i = (colour + j + k) & 1; i < L->g.n[0]; i += 2) {
            double diag;
            double lp = cc_apply_pt(L, i, j, k, &diag);
            F(&L->phi, i, j, k) = F(&L->phi, i, j, k) + (F(&L->rhs, i, j, k) - lp) / diag;
          }
      fab_fill_boundary(&L->phi);
    }
}

static void cc_residual(clevel *L) {
  int nk = (L->g.dim == 3) ? L->g.n[2] : 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nk; k++)
    for (int j = 0; j < L->g.n[1]; j++)
      for (int i = 0; i < L->g.n[0]; i++) {
        double diag;
        F(&L->res, i, j, k) = F(&L->rhs, i, j, k) - cc_apply_pt(L, i, j, k, &diag);
      }
}

void o_macproject(const orc_geom *g, const orc_params *p, ofab *phi, ofab *alphainv,
                  ofab *mac_rhs) {
  int dm = g->dim;
  fab_setval(phi, 0.0);
  int nl = o_compute_nlevs(g, p->mg_minwidth);
  clevel *lev = (clevel *)calloc((size_t)nl, sizeof(clevel));
  for (int n = 0; n < nl; n++) {
    clevel *L = &lev[n];
    L->g = *g;
    for (int d = 0; d < dm; d++) L->g.n[d] = g->n[d] >> n;
    L->g.dx = g->dx * (double)(1 << n);
    for (int d = 0; d < dm; d++) face_build(&L->a[d], &L->g, d, 0);
    cell_build(&L->phi, &L->g, 1);
    cell_build(&L->rhs, &L->g, 0);
    cell_build(&L->res, &L->g, 0);
  }
  for (int d = 0; d < dm; d++) fab_copy_valid(&lev[0].a[d], &alphainv[d]);
  for (int n = 1; n < nl; n++) o_stag_restriction(&lev[n].g, lev[n - 1].a, lev[n].a, 1);
  fab_copy_valid(&lev[0].rhs, mac_rhs);
  int nk0 = (dm == 3) ? g->n[2] : 1;
  /* F_MG's stopping test with the tower's eps = mg_rel_tol and abs_eps =
   * 1e-16 (macproject.f90:402-425): max|r| <= eps*max|rhs| or <= abs_eps,
   * before the first cycle (phi = 0, so r = rhs) and after each cycle but the
   * last; evaluated when more than one cycle may run. */
  int check = p->mg_max_vcycles > 1;
  double bnorm = check ? fab_norm_inf(&lev[0].rhs) : 0.0;
  int ncyc = p->mg_max_vcycles;
  if (check && (bnorm <= p->mg_rel_tol * bnorm || bnorm <= 1e-16)) ncyc = 0;
  for (int cyc = 0; cyc < ncyc; cyc++) {
    for (int n = 1; n < nl; n++) fab_setval(&lev[n].phi, 0.0);
    for (int n = 0; n < nl - 1; n++) {
      cc_smooth(&lev[n], p->mg_nsmooths_down);
      cc_residual(&lev[n]);
      o_cc_restriction(&lev[n + 1].g, &lev[n + 1].rhs, &lev[n].res);
    }
    cc_smooth(&lev[nl - 1], p->mg_nsmooths_bottom);
    for (int n = nl - 2; n >= 0; n--) {
      clevel *L = &lev[n], *C = &lev[n + 1];
      int nk = (dm == 3) ? L->g.n[2] : 1;
      for (int k = 0; k < nk; k++)
        for (int j = 0; j < L->g.n[1]; j++)
          for (int i = 0; i < L->g.n[0]; i++)
            F(&L->phi, i, j, k) = F(&L->phi, i, j, k) + F(&C->phi, i / 2, j / 2, (dm == 3) ? k / 2 : 0);
      fab_fill_boundary(&L->phi);
      cc_smooth(L, p->mg_nsmooths_up);
    }
    if (check && cyc + 1 < ncyc) {
      cc_residual(&lev[0]);
      double r = fab_norm_inf(&lev[0].res);
      if (r <= p->mg_rel_tol * bnorm || r <= 1e-16) break;
    }
  }
  for (int k = 0; k < nk0; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) F(phi, i, j, k) = F(&lev[0].phi, i, j, k);
  fab_fill_boundary(phi);
  for (int n = 0; n < nl; n++) {
    clevel *L = &lev[n];
    for (int d = 0; d < dm; d++) fab_free(&L->a[d]);
    fab_free(&L->phi);
    fab_free(&L->rhs);
    fab_free(&L->res);
  }
  free(lev);
}
