/*
 * Internal storage of the CPU oracle (test infrastructure only).
 *
 * ofab restates one BoxLib fab of a single periodic box: valid region
 * 0..n-1 (cell direction) or 0..n (nodal direction: the lo face and the
 * hi+1 face are both stored), plus ng ghost cells in every direction of the
 * problem (none in z for 2D).  Layout is Fortran column-major with i fastest,
 * like dataptr() in F_BaseLib.
 */
#ifndef ORC_FAB_H
#define ORC_FAB_H

#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "lmm_oracle.h"

typedef struct {
  int dim;
  int n[3];      /* cells per direction (n[2] = 1 in 2D) */
  int nodal[3];
  int ng, ngz;
  int ext[3];
  long sy, sz;
  double *p;     /* allocation */
  double *o;     /* address of (0,0,0) */
} ofab;

#define F(f, i, j, k) ((f)->o[(long)(i) + (f)->sy * (long)(j) + (f)->sz * (long)(k)])

static inline int vhi(const ofab *f, int d) { return f->n[d] - 1 + f->nodal[d]; }

void fab_build(ofab *f, const orc_geom *g, int nx, int ny, int nz, int ng);
void fab_free(ofab *f);
void fab_setval(ofab *f, double v);                 /* all=.true. */
void fab_copy_valid(ofab *dst, const ofab *src);    /* valid region only */
void fab_from_dedup(ofab *f, const double *src);    /* + periodic ghost fill */
double fab_to_dedup(const ofab *f, double *dst);    /* returns dup mismatch */
void fab_fill_boundary(ofab *f);                    /* multifab_fill_boundary */
void fab_internal_sync(ofab *f);                    /* multifab_internal_sync */
void fab_scale_all(ofab *f, double s);              /* mult_mult_s incl. ghosts */
double fab_norm_inf(const ofab *f);                 /* valid region */
double fab_norm_l2(const ofab *f);

/* face component d of a level with n cells */
static inline void face_build(ofab *f, const orc_geom *g, int d, int ng) {
  fab_build(f, g, d == 0, d == 1, d == 2, ng);
}
static inline void cell_build(ofab *f, const orc_geom *g, int ng) { fab_build(f, g, 0, 0, 0, ng); }
/* edge e of 3D: 0 = xy (nodal x,y), 1 = xz, 2 = yz; 2D node: e = 0, nodal x,y */
static inline void edge_build(ofab *f, const orc_geom *g, int e, int ng) {
  if (g->dim == 2) fab_build(f, g, 1, 1, 0, ng);
  else if (e == 0) fab_build(f, g, 1, 1, 0, ng);
  else if (e == 1) fab_build(f, g, 1, 0, 1, ng);
  else fab_build(f, g, 0, 1, 1, ng);
}

/* level-1 coefficient fabs staged from deduplicated inputs */
typedef struct {
  ofab alpha[3];   /* ng 0 */
  ofab beta;       /* ng 1 */
  ofab bed[3];     /* ng 0 */
  ofab gamma;      /* ng 1 */
} ocoef;

void ocoef_build(ocoef *c, const orc_geom *g, const orc_coef *src);
void ocoef_free(ocoef *c, const orc_geom *g);

/* stencil kernels on fabs (orc_stencil.c) */
void o_stag_applyop_level(const orc_geom *g, int visc, ofab *phi, ofab *Lphi, ofab *alpha,
                          ofab *beta, ofab *bed, ofab *gamma, int color);
void o_stag_mg_update(const orc_geom *g, int visc, double omega, ofab *phi, ofab *rhs,
                      ofab *Lphi, ofab *alpha, ofab *beta, ofab *bed, ofab *gamma, int color);
void o_cc_restriction(const orc_geom *gc, ofab *phi_c, ofab *phi_f);
void o_stag_restriction(const orc_geom *gc, ofab *phi_f, ofab *phi_c, int simple);
void o_nodal_restriction(const orc_geom *gc, ofab *phi_f, ofab *phi_c);
void o_edge_restriction(const orc_geom *gc, ofab *phi_f, ofab *phi_c);
void o_stag_prolongation(const orc_geom *gf, ofab *phi_f, ofab *phi_c);
void o_compute_div(const orc_geom *g, ofab *u, ofab *div, int increment);
void o_compute_grad(const orc_geom *g, ofab *p, ofab *gp);
void o_stag_inner_prod(const orc_geom *g, ofab *m1, ofab *m2, double out[3]);
double o_cc_inner_prod(const orc_geom *g, ofab *m1, ofab *m2);
void o_sum_umac_press(const orc_geom *g, ofab *p, ofab *u, double *av_p, double av_u[3]);

/* solvers on fabs (orc_mg.c, orc_krylov.c) */
int o_compute_nlevs(const orc_geom *g, int minwidth);
void o_stag_mg_solver(const orc_geom *g, const orc_params *p, ocoef *c, double theta_alpha,
                      ofab *phi, ofab *rhs);
void o_macproject(const orc_geom *g, const orc_params *p, ofab *phi, ofab *alphainv,
                  ofab *mac_rhs);
void o_apply_matrix(const orc_geom *g, const orc_params *p, ocoef *c, double theta_alpha,
                    ofab *b_u, ofab *b_p, ofab *x_u, ofab *x_p);
void o_apply_precon(const orc_geom *g, const orc_params *p, ocoef *c, double theta_alpha,
                    ofab *b_u, ofab *b_p, ofab *x_u, ofab *x_p);

#endif
