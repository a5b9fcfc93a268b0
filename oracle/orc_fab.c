/*
 * Fab storage and BoxLib-style periodic ghost exchange for the CPU oracle
 * (test infrastructure only).
 *
 * fab_fill_boundary restates F_BaseLib's multifab_fill_boundary for one
 * periodic box: each ghost cell takes the value of the valid point it is the
 * periodic image of (ghost -g <- n-g, ghost lastvalid+g <- lastvalid+g-n),
 * direction by direction so edges and corners are filled too.
 * fab_internal_sync restates multifab_internal_sync for one periodic box:
 * the duplicated hi+1 point of a nodal direction takes the lo point's value.
 */
#include "orc_fab.h"
#include <stdio.h>

void fab_build(ofab *f, const orc_geom *g, int nx, int ny, int nz, int ng) {
  f->dim = g->dim;
  for (int d = 0; d < 3; d++) f->n[d] = g->n[d];
  if (g->dim == 2) f->n[2] = 1;
  f->nodal[0] = nx;
  f->nodal[1] = ny;
  f->nodal[2] = (g->dim == 3) ? nz : 0;
  f->ng = ng;
  f->ngz = (g->dim == 3) ? ng : 0;
  f->ext[0] = f->n[0] + f->nodal[0] + 2 * ng;
  f->ext[1] = f->n[1] + f->nodal[1] + 2 * ng;
  f->ext[2] = f->n[2] + f->nodal[2] + 2 * f->ngz;
  f->sy = f->ext[0];
  f->sz = (long)f->ext[0] * f->ext[1];
  size_t cnt = (size_t)f->sz * f->ext[2];
  f->p = (double *)calloc(cnt, sizeof(double));
  if (!f->p) {
    fprintf(stderr, "oracle: out of memory (%zu doubles)\n", cnt);
    abort();
  }
  f->o = f->p + ng + f->sy * ng + f->sz * f->ngz;
}

void fab_free(ofab *f) {
  free(f->p);
  f->p = f->o = NULL;
}

void fab_setval(ofab *f, double v) {
  size_t cnt = (size_t)f->sz * f->ext[2];
  for (size_t q = 0; q < cnt; q++) f->p[q] = v;
}

void fab_scale_all(ofab *f, double s) {
  size_t cnt = (size_t)f->sz * f->ext[2];
  for (size_t q = 0; q < cnt; q++) f->p[q] = f->p[q] * s;
}

void fab_copy_valid(ofab *dst, const ofab *src) {
  for (int k = 0; k <= vhi(src, 2); k++)
    for (int j = 0; j <= vhi(src, 1); j++)
      for (int i = 0; i <= vhi(src, 0); i++) F(dst, i, j, k) = F(src, i, j, k);
}

void fab_from_dedup(ofab *f, const double *src) {
  int nx = f->n[0], ny = f->n[1], nz = f->n[2];
  for (int k = 0; k <= vhi(f, 2); k++)
    for (int j = 0; j <= vhi(f, 1); j++)
      for (int i = 0; i <= vhi(f, 0); i++)
        F(f, i, j, k) = src[(i % nx) + (long)nx * ((j % ny) + (long)ny * (k % nz))];
  fab_fill_boundary(f);
}

double fab_to_dedup(const ofab *f, double *dst) {
  int nx = f->n[0], ny = f->n[1], nz = f->n[2];
  double mism = 0.0;
  for (int k = 0; k <= vhi(f, 2); k++)
    for (int j = 0; j <= vhi(f, 1); j++)
      for (int i = 0; i <= vhi(f, 0); i++) {
        double v = F(f, i, j, k);
        if (i < nx && j < ny && k < nz) {
          dst[i + (long)nx * (j + (long)ny * k)] = v;
        } else {
          double w = F(f, i % nx, j % ny, k % nz);
          double e = fabs(v - w);
          if (e > mism || e != e) mism = (e != e) ? INFINITY : e;
        }
      }
  return mism;
}

void fab_fill_boundary(ofab *f) {
  int ngs[3] = {f->ng, f->ng, f->ngz};
  for (int d = 0; d < f->dim; d++) {
    int ng = ngs[d];
    if (ng == 0) continue;
    int n = f->n[d];
    int last = vhi(f, d);
    /* iterate over the whole allocated range of the other directions */
    int lo[3], hi[3];
    for (int e = 0; e < 3; e++) {
      lo[e] = -ngs[e];
      hi[e] = vhi(f, e) + ngs[e];
    }
    for (int g = 1; g <= ng; g++) {
      int ghost_lo = -g, img_lo = n - g;
      int ghost_hi = last + g, img_hi = last + g - n;
      if (d == 0) {
        for (int k = lo[2]; k <= hi[2]; k++)
          for (int j = lo[1]; j <= hi[1]; j++) {
            F(f, ghost_lo, j, k) = F(f, img_lo, j, k);
            F(f, ghost_hi, j, k) = F(f, img_hi, j, k);
          }
      } else if (d == 1) {
        for (int k = lo[2]; k <= hi[2]; k++)
          for (int i = lo[0]; i <= hi[0]; i++) {
            F(f, i, ghost_lo, k) = F(f, i, img_lo, k);
            F(f, i, ghost_hi, k) = F(f, i, img_hi, k);
          }
      } else {
        for (int j = lo[1]; j <= hi[1]; j++)
          for (int i = lo[0]; i <= hi[0]; i++) {
            F(f, i, j, ghost_lo) = F(f, i, j, img_lo);
            F(f, i, j, ghost_hi) = F(f, i, j, img_hi);
          }
      }
    }
  }
}

void fab_internal_sync(ofab *f) {
  for (int d = 0; d < f->dim; d++) {
    if (!f->nodal[d]) continue;
    int n = f->n[d];
    for (int k = 0; k <= vhi(f, 2); k++)
      for (int j = 0; j <= vhi(f, 1); j++)
        for (int i = 0; i <= vhi(f, 0); i++) {
          int idx[3] = {i, j, k};
          if (idx[d] != n) continue;
          int s[3] = {i, j, k};
          s[d] = 0;
          F(f, i, j, k) = F(f, s[0], s[1], s[2]);
        }
  }
}

double fab_norm_inf(const ofab *f) {
  double r = 0.0;
  for (int k = 0; k <= vhi(f, 2); k++)
    for (int j = 0; j <= vhi(f, 1); j++)
      for (int i = 0; i <= vhi(f, 0); i++) {
        double a = fabs(F(f, i, j, k));
        if (a > r) r = a;
      }
  return r;
}

double fab_norm_l2(const ofab *f) {
  double r = 0.0;
  for (int k = 0; k <= vhi(f, 2); k++)
    for (int j = 0; j <= vhi(f, 1); j++)
      for (int i = 0; i <= vhi(f, 0); i++) r += F(f, i, j, k) * F(f, i, j, k);
  return sqrt(r);
}

void ocoef_build(ocoef *c, const orc_geom *g, const orc_coef *src) {
  for (int d = 0; d < g->dim; d++) {
    face_build(&c->alpha[d], g, d, 0);
    fab_from_dedup(&c->alpha[d], src->alpha[d]);
  }
  cell_build(&c->beta, g, 1);
  fab_from_dedup(&c->beta, src->beta);
  cell_build(&c->gamma, g, 1);
  fab_from_dedup(&c->gamma, src->gamma);
  int ne = (g->dim == 2) ? 1 : 3;
  for (int e = 0; e < ne; e++) {
    edge_build(&c->bed[e], g, e, 0);
    fab_from_dedup(&c->bed[e], src->beta_ed[e]);
  }
}

void ocoef_free(ocoef *c, const orc_geom *g) {
  for (int d = 0; d < g->dim; d++) fab_free(&c->alpha[d]);
  fab_free(&c->beta);
  fab_free(&c->gamma);
  int ne = (g->dim == 2) ? 1 : 3;
  for (int e = 0; e < ne; e++) fab_free(&c->bed[e]);
}

/* OpenMP threads the oracle's parallel loops use (the CPU baseline reports it) */
#include <omp.h>
int orc_max_threads(void) { return omp_get_max_threads(); }
