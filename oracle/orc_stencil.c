/*
 * Stencil kernels of the CPU oracle (test infrastructure only), restated
 * expression by expression, left to right, from the reference:
 *   o_stag_applyop_level  <- src_common/stag_applyop.f90:70-1067
 *   o_stag_mg_update      <- src_gmres/stag_mg_solver.f90:1637-2463
 *   o_*_restriction       <- src_gmres/stag_mg_solver.f90:873-1367
 *   o_stag_prolongation   <- src_gmres/stag_mg_solver.f90:1370-1633
 *   o_compute_div/grad    <- src_common/div_and_grad.f90:16-455 (periodic)
 *   o_stag_inner_prod etc <- src_gmres/norm_inner_product.f90:16-794
 * Colour c (1..2*dim) touches component ceil(c/2) and the points with
 * (i+j+k) mod 2 == (c+1) mod 2; colour 0 touches every point.
 * Built with -ffp-contract=off so no multiply-add is fused.
 */
#include "orc_fab.h"

static void colour_setup(int dim, int color, int dos[3], int *off, int *par) {
  dos[0] = dos[1] = dos[2] = 1;
  if (dim == 2) dos[2] = 0;
  *off = 1;
  if (color == 1 || color == 2) { dos[1] = dos[2] = 0; *off = 2; }
  else if (color == 3 || color == 4) { dos[0] = dos[2] = 0; *off = 2; }
  else if (color == 5 || color == 6) { dos[0] = dos[1] = 0; *off = 2; }
  *par = (color + 1) % 2;
}

#define I0 ((off == 2) ? ((par + j + k) & 1) : 0)

/* ------------------------------------------------------------------ */
/* stag_applyop_2d / stag_applyop_3d                                   */
/* ------------------------------------------------------------------ */

static void applyop_2d(const orc_geom *g, int visc, ofab *phi, ofab *Lp, ofab *al, ofab *be,
                       ofab *bn, ofab *ga, int color) {
  int dos[3], off, par;
  colour_setup(2, color, dos, &off, &par);
  double dx = g->dx;
  double dxsq = dx * dx, dxsqinv = 1.0 / dxsq;
  double onethird = 1.0 / 3.0, twothirds = 2.0 / 3.0, fourthirds = 4.0 / 3.0;
  ofab *px = &phi[0], *py = &phi[1], *lx = &Lp[0], *ly = &Lp[1], *ax = &al[0], *ay = &al[1];
  const int k = 0;
  double b = F(be, 0, 0, 0), c = F(ga, 0, 0, 0);
#define X(a, bb) F(px, a, bb, 0)
#define Y(a, bb) F(py, a, bb, 0)
#define B(a, bb) F(be, a, bb, 0)
#define N(a, bb) F(bn, a, bb, 0)
#define G(a, bb) F(ga, a, bb, 0)
  if (dos[0]) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= vhi(px, 1); j++)
      for (int i = I0; i <= vhi(px, 0); i += off) {
        double r;
        if (visc == -1) {
          r = X(i, j) * (F(ax, i, j, 0) + (B(i, j) + B(i - 1, j) + N(i, j + 1) + N(i, j)) * dxsqinv) -
              (X(i + 1, j) * B(i, j) + X(i - 1, j) * B(i - 1, j) + X(i, j + 1) * N(i, j + 1) +
               X(i, j - 1) * N(i, j)) * dxsqinv;
        } else if (visc == 1) {
          r = X(i, j) * (F(ax, i, j, 0) + 4.0 * b * dxsqinv) -
              (X(i + 1, j) + X(i - 1, j) + X(i, j + 1) + X(i, j - 1)) * b * dxsqinv;
        } else if (visc == -2) {
          r = X(i, j) * (F(ax, i, j, 0) +
                         (2.0 * B(i, j) + 2.0 * B(i - 1, j) + N(i, j + 1) + N(i, j)) * dxsqinv) -
              (2.0 * X(i + 1, j) * B(i, j) + 2.0 * X(i - 1, j) * B(i - 1, j) + X(i, j + 1) * N(i, j + 1) +
               X(i, j - 1) * N(i, j) + Y(i, j + 1) * N(i, j + 1) - Y(i, j) * N(i, j) -
               Y(i - 1, j + 1) * N(i, j + 1) + Y(i - 1, j) * N(i, j)) * dxsqinv;
        } else if (visc == 2) {
          r = X(i, j) * (F(ax, i, j, 0) + 6.0 * b * dxsqinv) -
              (2.0 * X(i + 1, j) + 2.0 * X(i - 1, j) + X(i, j + 1) + X(i, j - 1) + Y(i, j + 1) - Y(i, j) -
               Y(i - 1, j + 1) + Y(i - 1, j)) * b * dxsqinv;
        } else if (visc == -3) {
          r = X(i, j) * (F(ax, i, j, 0) +
                         (fourthirds * B(i, j) + G(i, j) + fourthirds * B(i - 1, j) + G(i - 1, j) +
                          N(i, j + 1) + N(i, j)) * dxsqinv) -
              (X(i + 1, j) * (fourthirds * B(i, j) + G(i, j)) +
               X(i - 1, j) * (fourthirds * B(i - 1, j) + G(i - 1, j)) + X(i, j + 1) * N(i, j + 1) +
               X(i, j - 1) * N(i, j) + Y(i, j + 1) * (N(i, j + 1) - twothirds * B(i, j) + G(i, j)) -
               Y(i, j) * (N(i, j) - twothirds * B(i, j) + G(i, j)) -
               Y(i - 1, j + 1) * (N(i, j + 1) - twothirds * B(i - 1, j) + G(i - 1, j)) +
               Y(i - 1, j) * (N(i, j) - twothirds * B(i - 1, j) + G(i - 1, j))) * dxsqinv;
        } else { /* 3 */
          r = X(i, j) * (F(ax, i, j, 0) + (14.0 * b * onethird + 2.0 * c) * dxsqinv) -
              ((X(i + 1, j) + X(i - 1, j)) * (fourthirds * b + c) + (X(i, j + 1) + X(i, j - 1)) * b +
               (Y(i, j + 1) - Y(i, j) - Y(i - 1, j + 1) + Y(i - 1, j)) * (onethird * b + c)) * dxsqinv;
        }
        F(lx, i, j, 0) = r;
      }
  }
  if (dos[1]) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= vhi(py, 1); j++)
      for (int i = I0; i <= vhi(py, 0); i += off) {
        double r;
        if (visc == -1) {
          r = Y(i, j) * (F(ay, i, j, 0) + (B(i, j) + B(i, j - 1) + N(i + 1, j) + N(i, j)) * dxsqinv) -
              (Y(i, j + 1) * B(i, j) + Y(i, j - 1) * B(i, j - 1) + Y(i + 1, j) * N(i + 1, j) +
               Y(i - 1, j) * N(i, j)) * dxsqinv;
        } else if (visc == 1) {
          r = Y(i, j) * (F(ay, i, j, 0) + 4.0 * b * dxsqinv) -
              (Y(i, j + 1) + Y(i, j - 1) + Y(i + 1, j) + Y(i - 1, j)) * b * dxsqinv;
        } else if (visc == -2) {
          r = Y(i, j) * (F(ay, i, j, 0) +
                         (2.0 * B(i, j) + 2.0 * B(i, j - 1) + N(i + 1, j) + N(i, j)) * dxsqinv) -
              (2.0 * Y(i, j + 1) * B(i, j) + 2.0 * Y(i, j - 1) * B(i, j - 1) + Y(i + 1, j) * N(i + 1, j) +
               Y(i - 1, j) * N(i, j) + X(i + 1, j) * N(i + 1, j) - X(i, j) * N(i, j) -
               X(i + 1, j - 1) * N(i + 1, j) + X(i, j - 1) * N(i, j)) * dxsqinv;
        } else if (visc == 2) {
          r = Y(i, j) * (F(ay, i, j, 0) + 6.0 * b * dxsqinv) -
              (2.0 * Y(i, j + 1) + 2.0 * Y(i, j - 1) + Y(i + 1, j) + Y(i - 1, j) + X(i + 1, j) - X(i, j) -
               X(i + 1, j - 1) + X(i, j - 1)) * b * dxsqinv;
        } else if (visc == -3) {
          r = Y(i, j) * (F(ay, i, j, 0) +
                         (fourthirds * B(i, j) + G(i, j) + fourthirds * B(i, j - 1) + G(i, j - 1) +
                          N(i + 1, j) + N(i, j)) * dxsqinv) -
              (Y(i, j + 1) * (fourthirds * B(i, j) + G(i, j)) +
               Y(i, j - 1) * (fourthirds * B(i, j - 1) + G(i, j - 1)) + Y(i + 1, j) * N(i + 1, j) +
               Y(i - 1, j) * N(i, j) + X(i + 1, j) * (N(i + 1, j) - twothirds * B(i, j) + G(i, j)) -
               X(i, j) * (N(i, j) - twothirds * B(i, j) + G(i, j)) -
               X(i + 1, j - 1) * (N(i + 1, j) - twothirds * B(i, j - 1) + G(i, j - 1)) +
               X(i, j - 1) * (N(i, j) - twothirds * B(i, j - 1) + G(i, j - 1))) * dxsqinv;
        } else {
          r = Y(i, j) * (F(ay, i, j, 0) + (14.0 * b * onethird + 2.0 * c) * dxsqinv) -
              ((Y(i, j + 1) + Y(i, j - 1)) * (fourthirds * b + c) + (Y(i + 1, j) + Y(i - 1, j)) * b +
               (X(i + 1, j) - X(i, j) - X(i + 1, j - 1) + X(i, j - 1)) * (onethird * b + c)) * dxsqinv;
        }
        F(ly, i, j, 0) = r;
      }
  }
#undef X
#undef Y
#undef B
#undef N
#undef G
  (void)k;
}

static void applyop_3d(const orc_geom *g, int visc, ofab *phi, ofab *Lp, ofab *al, ofab *be,
                       ofab *bed, ofab *ga, int color) {
  int dos[3], off, par;
  colour_setup(3, color, dos, &off, &par);
  double dx = g->dx;
  double dxsq = dx * dx, dxsqinv = 1.0 / dxsq;
  double onethird = 1.0 / 3.0, twothirds = 2.0 / 3.0, fourthirds = 4.0 / 3.0;
  ofab *px = &phi[0], *py = &phi[1], *pz = &phi[2];
  ofab *bxy = &bed[0], *bxz = &bed[1], *byz = &bed[2];
  double b = F(be, 0, 0, 0), c = F(ga, 0, 0, 0);
#define X(a, bb, cc) F(px, a, bb, cc)
#define Y(a, bb, cc) F(py, a, bb, cc)
#define Z(a, bb, cc) F(pz, a, bb, cc)
#define B(a, bb, cc) F(be, a, bb, cc)
#define G(a, bb, cc) F(ga, a, bb, cc)
#define BXY(a, bb, cc) F(bxy, a, bb, cc)
#define BXZ(a, bb, cc) F(bxz, a, bb, cc)
#define BYZ(a, bb, cc) F(byz, a, bb, cc)
  if (dos[0]) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= vhi(px, 2); k++)
      for (int j = 0; j <= vhi(px, 1); j++)
        for (int i = I0; i <= vhi(px, 0); i += off) {
          double r, a = F(&al[0], i, j, k);
          if (visc == -1) {
            r = X(i, j, k) * (a + (B(i, j, k) + B(i - 1, j, k) + BXY(i, j, k) + BXY(i, j + 1, k) +
                                   BXZ(i, j, k) + BXZ(i, j, k + 1)) * dxsqinv) -
                (X(i + 1, j, k) * B(i, j, k) + X(i - 1, j, k) * B(i - 1, j, k) +
                 X(i, j + 1, k) * BXY(i, j + 1, k) + X(i, j - 1, k) * BXY(i, j, k) +
                 X(i, j, k + 1) * BXZ(i, j, k + 1) + X(i, j, k - 1) * BXZ(i, j, k)) * dxsqinv;
          } else if (visc == 1) {
            r = X(i, j, k) * (a + 6.0 * b * dxsqinv) -
                (X(i + 1, j, k) + X(i - 1, j, k) + X(i, j + 1, k) + X(i, j - 1, k) + X(i, j, k + 1) +
                 X(i, j, k - 1)) * b * dxsqinv;
          } else if (visc == -2) {
            r = X(i, j, k) * (a + (2.0 * B(i, j, k) + 2.0 * B(i - 1, j, k) + BXY(i, j, k) +
                                   BXY(i, j + 1, k) + BXZ(i, j, k) + BXZ(i, j, k + 1)) * dxsqinv) -
                (2.0 * X(i + 1, j, k) * B(i, j, k) + 2.0 * X(i - 1, j, k) * B(i - 1, j, k) +
                 X(i, j + 1, k) * BXY(i, j + 1, k) + X(i, j - 1, k) * BXY(i, j, k) +
                 X(i, j, k + 1) * BXZ(i, j, k + 1) + X(i, j, k - 1) * BXZ(i, j, k) +
                 Y(i, j + 1, k) * BXY(i, j + 1, k) - Y(i, j, k) * BXY(i, j, k) -
                 Y(i - 1, j + 1, k) * BXY(i, j + 1, k) + Y(i - 1, j, k) * BXY(i, j, k) +
                 Z(i, j, k + 1) * BXZ(i, j, k + 1) - Z(i, j, k) * BXZ(i, j, k) -
                 Z(i - 1, j, k + 1) * BXZ(i, j, k + 1) + Z(i - 1, j, k) * BXZ(i, j, k)) * dxsqinv;
          } else if (visc == 2) {
            r = X(i, j, k) * (a + 8.0 * b * dxsqinv) -
                (2.0 * X(i + 1, j, k) + 2.0 * X(i - 1, j, k) + X(i, j + 1, k) + X(i, j - 1, k) +
                 X(i, j, k + 1) + X(i, j, k - 1) + Y(i, j + 1, k) - Y(i, j, k) - Y(i - 1, j + 1, k) +
                 Y(i - 1, j, k) + Z(i, j, k + 1) - Z(i, j, k) - Z(i - 1, j, k + 1) + Z(i - 1, j, k)) *
                    b * dxsqinv;
          } else if (visc == -3) {
            r = X(i, j, k) * (a + (fourthirds * B(i, j, k) + G(i, j, k) + fourthirds * B(i - 1, j, k) +
                                   G(i - 1, j, k) + BXY(i, j, k) + BXY(i, j + 1, k) + BXZ(i, j, k) +
                                   BXZ(i, j, k + 1)) * dxsqinv) -
                (X(i + 1, j, k) * (fourthirds * B(i, j, k) + G(i, j, k)) +
                 X(i - 1, j, k) * (fourthirds * B(i - 1, j, k) + G(i - 1, j, k)) +
                 X(i, j + 1, k) * BXY(i, j + 1, k) + X(i, j - 1, k) * BXY(i, j, k) +
                 X(i, j, k + 1) * BXZ(i, j, k + 1) + X(i, j, k - 1) * BXZ(i, j, k) +
                 Y(i, j + 1, k) * (BXY(i, j + 1, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 Y(i, j, k) * (BXY(i, j, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 Y(i - 1, j + 1, k) * (BXY(i, j + 1, k) - twothirds * B(i - 1, j, k) + G(i - 1, j, k)) +
                 Y(i - 1, j, k) * (BXY(i, j, k) - twothirds * B(i - 1, j, k) + G(i - 1, j, k)) +
                 Z(i, j, k + 1) * (BXZ(i, j, k + 1) - twothirds * B(i, j, k) + G(i, j, k)) -
                 Z(i, j, k) * (BXZ(i, j, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 Z(i - 1, j, k + 1) * (BXZ(i, j, k + 1) - twothirds * B(i - 1, j, k) + G(i - 1, j, k)) +
                 Z(i - 1, j, k) * (BXZ(i, j, k) - twothirds * B(i - 1, j, k) + G(i - 1, j, k))) *
                    dxsqinv;
          } else {
            r = X(i, j, k) * (a + (20.0 * b * onethird + 2.0 * c) * dxsqinv) -
                ((X(i + 1, j, k) + X(i - 1, j, k)) * (fourthirds * b + c) +
                 (X(i, j + 1, k) + X(i, j - 1, k) + X(i, j, k + 1) + X(i, j, k - 1)) * b +
                 (Y(i, j + 1, k) - Y(i, j, k) - Y(i - 1, j + 1, k) + Y(i - 1, j, k) + Z(i, j, k + 1) -
                  Z(i, j, k) - Z(i - 1, j, k + 1) + Z(i - 1, j, k)) * (onethird * b + c)) * dxsqinv;
          }
          F(&Lp[0], i, j, k) = r;
        }
  }
  if (dos[1]) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= vhi(py, 2); k++)
      for (int j = 0; j <= vhi(py, 1); j++)
        for (int i = I0; i <= vhi(py, 0); i += off) {
          double r, a = F(&al[1], i, j, k);
          if (visc == -1) {
            r = Y(i, j, k) * (a + (B(i, j, k) + B(i, j - 1, k) + BXY(i, j, k) + BXY(i + 1, j, k) +
                                   BYZ(i, j, k) + BYZ(i, j, k + 1)) * dxsqinv) -
                (Y(i, j + 1, k) * B(i, j, k) + Y(i, j - 1, k) * B(i, j - 1, k) +
                 Y(i + 1, j, k) * BXY(i + 1, j, k) + Y(i - 1, j, k) * BXY(i, j, k) +
                 Y(i, j, k + 1) * BYZ(i, j, k + 1) + Y(i, j, k - 1) * BYZ(i, j, k)) * dxsqinv;
          } else if (visc == 1) {
            r = Y(i, j, k) * (a + 6.0 * b * dxsqinv) -
                (Y(i + 1, j, k) + Y(i - 1, j, k) + Y(i, j + 1, k) + Y(i, j - 1, k) + Y(i, j, k + 1) +
                 Y(i, j, k - 1)) * b * dxsqinv;
          } else if (visc == -2) {
            r = Y(i, j, k) * (a + (2.0 * B(i, j, k) + 2.0 * B(i, j - 1, k) + BXY(i, j, k) +
                                   BXY(i + 1, j, k) + BYZ(i, j, k) + BYZ(i, j, k + 1)) * dxsqinv) -
                (2.0 * Y(i, j + 1, k) * B(i, j, k) + 2.0 * Y(i, j - 1, k) * B(i, j - 1, k) +
                 Y(i + 1, j, k) * BXY(i + 1, j, k) + Y(i - 1, j, k) * BXY(i, j, k) +
                 Y(i, j, k + 1) * BYZ(i, j, k + 1) + Y(i, j, k - 1) * BYZ(i, j, k) +
                 X(i + 1, j, k) * BXY(i + 1, j, k) - X(i, j, k) * BXY(i, j, k) -
                 X(i + 1, j - 1, k) * BXY(i + 1, j, k) + X(i, j - 1, k) * BXY(i, j, k) +
                 Z(i, j, k + 1) * BYZ(i, j, k + 1) - Z(i, j, k) * BYZ(i, j, k) -
                 Z(i, j - 1, k + 1) * BYZ(i, j, k + 1) + Z(i, j - 1, k) * BYZ(i, j, k)) * dxsqinv;
          } else if (visc == 2) {
            r = Y(i, j, k) * (a + 8.0 * b * dxsqinv) -
                (2.0 * Y(i, j + 1, k) + 2.0 * Y(i, j - 1, k) + Y(i + 1, j, k) + Y(i - 1, j, k) +
                 Y(i, j, k + 1) + Y(i, j, k - 1) + X(i + 1, j, k) - X(i, j, k) - X(i + 1, j - 1, k) +
                 X(i, j - 1, k) + Z(i, j, k + 1) - Z(i, j, k) - Z(i, j - 1, k + 1) + Z(i, j - 1, k)) *
                    b * dxsqinv;
          } else if (visc == -3) {
            r = Y(i, j, k) * (a + (fourthirds * B(i, j, k) + G(i, j, k) + fourthirds * B(i, j - 1, k) +
                                   G(i, j - 1, k) + BXY(i, j, k) + BXY(i + 1, j, k) + BYZ(i, j, k) +
                                   BYZ(i, j, k + 1)) * dxsqinv) -
                (Y(i, j + 1, k) * (fourthirds * B(i, j, k) + G(i, j, k)) +
                 Y(i, j - 1, k) * (fourthirds * B(i, j - 1, k) + G(i, j - 1, k)) +
                 Y(i + 1, j, k) * BXY(i + 1, j, k) + Y(i - 1, j, k) * BXY(i, j, k) +
                 Y(i, j, k + 1) * BYZ(i, j, k + 1) + Y(i, j, k - 1) * BYZ(i, j, k) +
                 X(i + 1, j, k) * (BXY(i + 1, j, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 X(i, j, k) * (BXY(i, j, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 X(i + 1, j - 1, k) * (BXY(i + 1, j, k) - twothirds * B(i, j - 1, k) + G(i, j - 1, k)) +
                 X(i, j - 1, k) * (BXY(i, j, k) - twothirds * B(i, j - 1, k) + G(i, j - 1, k)) +
                 Z(i, j, k + 1) * (BYZ(i, j, k + 1) - twothirds * B(i, j, k) + G(i, j, k)) -
                 Z(i, j, k) * (BYZ(i, j, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 Z(i, j - 1, k + 1) * (BYZ(i, j, k + 1) - twothirds * B(i, j - 1, k) + G(i, j - 1, k)) +
                 Z(i, j - 1, k) * (BYZ(i, j, k) - twothirds * B(i, j - 1, k) + G(i, j - 1, k))) *
                    dxsqinv;
          } else {
            r = Y(i, j, k) * (a + (20.0 * b * onethird + 2.0 * c) * dxsqinv) -
                ((Y(i, j + 1, k) + Y(i, j - 1, k)) * (fourthirds * b + c) +
                 (Y(i + 1, j, k) + Y(i - 1, j, k) + Y(i, j, k + 1) + Y(i, j, k - 1)) * b +
                 (X(i + 1, j, k) - X(i, j, k) - X(i + 1, j - 1, k) + X(i, j - 1, k) + Z(i, j, k + 1) -
                  Z(i, j, k) - Z(i, j - 1, k + 1) + Z(i, j - 1, k)) * (onethird * b + c)) * dxsqinv;
          }
          F(&Lp[1], i, j, k) = r;
        }
  }
  if (dos[2]) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= vhi(pz, 2); k++)
      for (int j = 0; j <= vhi(pz, 1); j++)
        for (int i = I0; i <= vhi(pz, 0); i += off) {
          double r, a = F(&al[2], i, j, k);
          if (visc == -1) {
            r = Z(i, j, k) * (a + (B(i, j, k) + B(i, j, k - 1) + BXZ(i, j, k) + BXZ(i + 1, j, k) +
                                   BYZ(i, j, k) + BYZ(i, j + 1, k)) * dxsqinv) -
                (Z(i, j, k + 1) * B(i, j, k) + Z(i, j, k - 1) * B(i, j, k - 1) +
                 Z(i + 1, j, k) * BXZ(i + 1, j, k) + Z(i - 1, j, k) * BXZ(i, j, k) +
                 Z(i, j + 1, k) * BYZ(i, j + 1, k) + Z(i, j - 1, k) * BYZ(i, j, k)) * dxsqinv;
          } else if (visc == 1) {
            r = Z(i, j, k) * (a + 6.0 * b * dxsqinv) -
                (Z(i + 1, j, k) + Z(i - 1, j, k) + Z(i, j + 1, k) + Z(i, j - 1, k) + Z(i, j, k + 1) +
                 Z(i, j, k - 1)) * b * dxsqinv;
          } else if (visc == -2) {
            r = Z(i, j, k) * (a + (2.0 * B(i, j, k) + 2.0 * B(i, j, k - 1) + BXZ(i, j, k) +
                                   BXZ(i + 1, j, k) + BYZ(i, j, k) + BYZ(i, j + 1, k)) * dxsqinv) -
                (2.0 * Z(i, j, k + 1) * B(i, j, k) + 2.0 * Z(i, j, k - 1) * B(i, j, k - 1) +
                 Z(i + 1, j, k) * BXZ(i + 1, j, k) + Z(i - 1, j, k) * BXZ(i, j, k) +
                 Z(i, j + 1, k) * BYZ(i, j + 1, k) + Z(i, j - 1, k) * BYZ(i, j, k) +
                 X(i + 1, j, k) * BXZ(i + 1, j, k) - X(i, j, k) * BXZ(i, j, k) -
                 X(i + 1, j, k - 1) * BXZ(i + 1, j, k) + X(i, j, k - 1) * BXZ(i, j, k) +
                 Y(i, j + 1, k) * BYZ(i, j + 1, k) - Y(i, j, k) * BYZ(i, j, k) -
                 Y(i, j + 1, k - 1) * BYZ(i, j + 1, k) + Y(i, j, k - 1) * BYZ(i, j, k)) * dxsqinv;
          } else if (visc == 2) {
            r = Z(i, j, k) * (a + 8.0 * b * dxsqinv) -
                (2.0 * Z(i, j, k + 1) + 2.0 * Z(i, j, k - 1) + Z(i + 1, j, k) + Z(i - 1, j, k) +
                 Z(i, j + 1, k) + Z(i, j - 1, k) + X(i + 1, j, k) - X(i, j, k) - X(i + 1, j, k - 1) +
                 X(i, j, k - 1) + Y(i, j + 1, k) - Y(i, j, k) - Y(i, j + 1, k - 1) + Y(i, j, k - 1)) *
                    b * dxsqinv;
          } else if (visc == -3) {
            r = Z(i, j, k) * (a + (fourthirds * B(i, j, k) + G(i, j, k) + fourthirds * B(i, j, k - 1) +
                                   G(i, j, k - 1) + BXZ(i, j, k) + BXZ(i + 1, j, k) + BYZ(i, j, k) +
                                   BYZ(i, j + 1, k)) * dxsqinv) -
                (Z(i, j, k + 1) * (fourthirds * B(i, j, k) + G(i, j, k)) +
                 Z(i, j, k - 1) * (fourthirds * B(i, j, k - 1) + G(i, j, k - 1)) +
                 Z(i + 1, j, k) * BXZ(i + 1, j, k) + Z(i - 1, j, k) * BXZ(i, j, k) +
                 Z(i, j + 1, k) * BYZ(i, j + 1, k) + Z(i, j - 1, k) * BYZ(i, j, k) +
                 X(i + 1, j, k) * (BXZ(i + 1, j, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 X(i, j, k) * (BXZ(i, j, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 X(i + 1, j, k - 1) * (BXZ(i + 1, j, k) - twothirds * B(i, j, k - 1) + G(i, j, k - 1)) +
                 X(i, j, k - 1) * (BXZ(i, j, k) - twothirds * B(i, j, k - 1) + G(i, j, k - 1)) +
                 Y(i, j + 1, k) * (BYZ(i, j + 1, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 Y(i, j, k) * (BYZ(i, j, k) - twothirds * B(i, j, k) + G(i, j, k)) -
                 Y(i, j + 1, k - 1) * (BYZ(i, j + 1, k) - twothirds * B(i, j, k - 1) + G(i, j, k - 1)) +
                 Y(i, j, k - 1) * (BYZ(i, j, k) - twothirds * B(i, j, k - 1) + G(i, j, k - 1))) *
                    dxsqinv;
          } else {
            r = Z(i, j, k) * (a + (20.0 * b * onethird + 2.0 * c) * dxsqinv) -
                ((Z(i, j, k + 1) + Z(i, j, k - 1)) * (fourthirds * b + c) +
                 (Z(i + 1, j, k) + Z(i - 1, j, k) + Z(i, j + 1, k) + Z(i, j - 1, k)) * b +
                 (X(i + 1, j, k) - X(i, j, k) - X(i + 1, j, k - 1) + X(i, j, k - 1) + Y(i, j + 1, k) -
                  Y(i, j, k) - Y(i, j + 1, k - 1) + Y(i, j, k - 1)) * (onethird * b + c)) * dxsqinv;
          }
          F(&Lp[2], i, j, k) = r;
        }
  }
#undef X
#undef Y
#undef Z
#undef B
#undef G
#undef BXY
#undef BXZ
#undef BYZ
}

void o_stag_applyop_level(const orc_geom *g, int visc, ofab *phi, ofab *Lphi, ofab *alpha,
                          ofab *beta, ofab *bed, ofab *gamma, int color) {
  if (g->dim == 2)
    applyop_2d(g, visc, phi, Lphi, alpha, beta, bed, gamma, color);
  else
    applyop_3d(g, visc, phi, Lphi, alpha, beta, bed, gamma, color);
  /* multifab_physbc_domainvel: no-op on a periodic domain */
}

/* ------------------------------------------------------------------ */
/* stag_mg_update_2d / _3d: phi += omega*(rhs - Lphi)/fac               */
/* ------------------------------------------------------------------ */

static void update_2d(const orc_geom *g, int visc, double omega, ofab *phi, ofab *rhs, ofab *Lp,
                      ofab *al, ofab *be, ofab *bn, ofab *ga, int color) {
  int dos[3], off, par;
  colour_setup(2, color, dos, &off, &par);
  double dxsq = g->dx * g->dx, dxsqinv = 1.0 / dxsq;
  double fourthirds = 4.0 / 3.0, fourteenthirds = 14.0 / 3.0;
  double b = F(be, 0, 0, 0), c = F(ga, 0, 0, 0);
  const int k = 0;
#define B(a, bb) F(be, a, bb, 0)
#define N(a, bb) F(bn, a, bb, 0)
#define G(a, bb) F(ga, a, bb, 0)
  for (int d = 0; d < 2; d++) {
    if (!dos[d]) continue;
    ofab *p = &phi[d], *r = &rhs[d], *l = &Lp[d], *a = &al[d];
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= vhi(p, 1); j++)
      for (int i = I0; i <= vhi(p, 0); i += off) {
        double fac;
        if (d == 0) {
          if (visc == -1) fac = F(a, i, j, 0) + (B(i, j) + B(i - 1, j) + N(i, j) + N(i, j + 1)) * dxsqinv;
          else if (visc == 1) fac = F(a, i, j, 0) + 4.0 * b * dxsqinv;
          else if (visc == -2)
            fac = F(a, i, j, 0) + (2.0 * B(i, j) + 2.0 * B(i - 1, j) + N(i, j) + N(i, j + 1)) * dxsqinv;
          else if (visc == 2) fac = F(a, i, j, 0) + 6.0 * b * dxsqinv;
          else if (visc == -3)
            fac = F(a, i, j, 0) + (fourthirds * B(i, j) + G(i, j) + fourthirds * B(i - 1, j) + G(i - 1, j) +
                                   N(i, j) + N(i, j + 1)) * dxsqinv;
          else fac = F(a, i, j, 0) + (fourteenthirds * b + 2.0 * c) * dxsqinv;
        } else {
          if (visc == -1) fac = F(a, i, j, 0) + (B(i, j) + B(i, j - 1) + N(i, j) + N(i + 1, j)) * dxsqinv;
          else if (visc == 1) fac = F(a, i, j, 0) + 4.0 * b * dxsqinv;
          else if (visc == -2)
            fac = F(a, i, j, 0) + (2.0 * B(i, j) + 2.0 * B(i, j - 1) + N(i, j) + N(i + 1, j)) * dxsqinv;
          else if (visc == 2) fac = F(a, i, j, 0) + 6.0 * b * dxsqinv;
          else if (visc == -3)
            fac = F(a, i, j, 0) + (fourthirds * B(i, j) + G(i, j) + fourthirds * B(i, j - 1) + G(i, j - 1) +
                                   N(i, j) + N(i + 1, j)) * dxsqinv;
          else fac = F(a, i, j, 0) + (fourteenthirds * b + 2.0 * c) * dxsqinv;
        }
        F(p, i, j, 0) = F(p, i, j, 0) + omega * (F(r, i, j, 0) - F(l, i, j, 0)) / fac;
      }
  }
#undef B
#undef N
#undef G
  (void)k;
}

static void update_3d(const orc_geom *g, int visc, double omega, ofab *phi, ofab *rhs, ofab *Lp,
                      ofab *al, ofab *be, ofab *bed, ofab *ga, int color) {
  int dos[3], off, par;
  colour_setup(3, color, dos, &off, &par);
  double dxsq = g->dx * g->dx, dxsqinv = 1.0 / dxsq;
  double fourthirds = 4.0 / 3.0, twentythirds = 20.0 / 3.0;
  double b = F(be, 0, 0, 0), c = F(ga, 0, 0, 0);
  ofab *bxy = &bed[0], *bxz = &bed[1], *byz = &bed[2];
#define B(a, bb, cc) F(be, a, bb, cc)
#define G(a, bb, cc) F(ga, a, bb, cc)
#define BXY(a, bb, cc) F(bxy, a, bb, cc)
#define BXZ(a, bb, cc) F(bxz, a, bb, cc)
#define BYZ(a, bb, cc) F(byz, a, bb, cc)
  for (int d = 0; d < 3; d++) {
    if (!dos[d]) continue;
    ofab *p = &phi[d], *r = &rhs[d], *l = &Lp[d], *a = &al[d];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= vhi(p, 2); k++)
      for (int j = 0; j <= vhi(p, 1); j++)
        for (int i = I0; i <= vhi(p, 0); i += off) {
          double fac, av = F(a, i, j, k);
          if (visc == 1) fac = av + 6.0 * b * dxsqinv;
          else if (visc == 2) fac = av + 8.0 * b * dxsqinv;
          else if (visc == 3) fac = av + (twentythirds * b + 2.0 * c) * dxsqinv;
          else if (d == 0) {
            if (visc == -1)
              fac = av + (B(i, j, k) + B(i - 1, j, k) + BXY(i, j, k) + BXY(i, j + 1, k) + BXZ(i, j, k) +
                          BXZ(i, j, k + 1)) * dxsqinv;
            else if (visc == -2)
              fac = av + (2.0 * B(i, j, k) + 2.0 * B(i - 1, j, k) + BXY(i, j, k) + BXY(i, j + 1, k) +
                          BXZ(i, j, k) + BXZ(i, j, k + 1)) * dxsqinv;
            else
              fac = av + (fourthirds * B(i, j, k) + G(i, j, k) + fourthirds * B(i - 1, j, k) + G(i - 1, j, k) +
                          BXY(i, j, k) + BXY(i, j + 1, k) + BXZ(i, j, k) + BXZ(i, j, k + 1)) * dxsqinv;
          } else if (d == 1) {
            if (visc == -1)
              fac = av + (B(i, j, k) + B(i, j - 1, k) + BXY(i, j, k) + BXY(i + 1, j, k) + BYZ(i, j, k) +
                          BYZ(i, j, k + 1)) * dxsqinv;
            else if (visc == -2)
              fac = av + (2.0 * B(i, j, k) + 2.0 * B(i, j - 1, k) + BXY(i, j, k) + BXY(i + 1, j, k) +
                          BYZ(i, j, k) + BYZ(i, j, k + 1)) * dxsqinv;
            else
              fac = av + (fourthirds * B(i, j, k) + G(i, j, k) + fourthirds * B(i, j - 1, k) + G(i, j - 1, k) +
                          BXY(i, j, k) + BXY(i + 1, j, k) + BYZ(i, j, k) + BYZ(i, j, k + 1)) * dxsqinv;
          } else {
            if (visc == -1)
              fac = av + (B(i, j, k) + B(i, j, k - 1) + BXZ(i, j, k) + BXZ(i + 1, j, k) + BYZ(i, j, k) +
                          BYZ(i, j + 1, k)) * dxsqinv;
            else if (visc == -2)
              fac = av + (2.0 * B(i, j, k) + 2.0 * B(i, j, k - 1) + BXZ(i, j, k) + BXZ(i + 1, j, k) +
                          BYZ(i, j, k) + BYZ(i, j + 1, k)) * dxsqinv;
            else
              fac = av + (fourthirds * B(i, j, k) + G(i, j, k) + fourthirds * B(i, j, k - 1) + G(i, j, k - 1) +
                          BXZ(i, j, k) + BXZ(i + 1, j, k) + BYZ(i, j, k) + BYZ(i, j + 1, k)) * dxsqinv;
          }
          F(p, i, j, k) = F(p, i, j, k) + omega * (F(r, i, j, k) - F(l, i, j, k)) / fac;
        }
  }
#undef B
#undef G
#undef BXY
#undef BXZ
#undef BYZ
}

void o_stag_mg_update(const orc_geom *g, int visc, double omega, ofab *phi, ofab *rhs,
                      ofab *Lphi, ofab *alpha, ofab *beta, ofab *bed, ofab *gamma, int color) {
  if (g->dim == 2)
    update_2d(g, visc, omega, phi, rhs, Lphi, alpha, beta, bed, gamma, color);
  else
    update_3d(g, visc, omega, phi, rhs, Lphi, alpha, beta, bed, gamma, color);
}

/* ------------------------------------------------------------------ */
/* restriction / prolongation (gc = coarse geometry)                    */
/* ------------------------------------------------------------------ */

void o_cc_restriction(const orc_geom *gc, ofab *c, ofab *f) {
  if (gc->dim == 2) {
    for (int j = 0; j < gc->n[1]; j++)
      for (int i = 0; i < gc->n[0]; i++)
        F(c, i, j, 0) = 0.25 * (F(f, 2 * i, 2 * j, 0) + F(f, 2 * i + 1, 2 * j, 0) + F(f, 2 * i, 2 * j + 1, 0) +
                                F(f, 2 * i + 1, 2 * j + 1, 0));
  } else {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < gc->n[2]; k++)
      for (int j = 0; j < gc->n[1]; j++)
        for (int i = 0; i < gc->n[0]; i++)
          F(c, i, j, k) = 0.125 * (F(f, 2 * i, 2 * j, 2 * k) + F(f, 2 * i + 1, 2 * j, 2 * k) +
                                   F(f, 2 * i, 2 * j + 1, 2 * k) + F(f, 2 * i + 1, 2 * j + 1, 2 * k) +
                                   F(f, 2 * i, 2 * j, 2 * k + 1) + F(f, 2 * i + 1, 2 * j, 2 * k + 1) +
                                   F(f, 2 * i, 2 * j + 1, 2 * k + 1) + F(f, 2 * i + 1, 2 * j + 1, 2 * k + 1));
  }
  fab_fill_boundary(c);
}

void o_stag_restriction(const orc_geom *gc, ofab *pf, ofab *pc, int simple) {
  int nx = gc->n[0], ny = gc->n[1], nz = gc->n[2];
  if (gc->dim == 2) {
    ofab *fx = &pf[0], *fy = &pf[1], *cx = &pc[0], *cy = &pc[1];
    if (simple) {
      for (int j = 0; j < ny; j++)
        for (int i = 0; i <= nx; i++) F(cx, i, j, 0) = 0.5 * (F(fx, 2 * i, 2 * j, 0) + F(fx, 2 * i, 2 * j + 1, 0));
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i < nx; i++) F(cy, i, j, 0) = 0.5 * (F(fy, 2 * i, 2 * j, 0) + F(fy, 2 * i + 1, 2 * j, 0));
    } else {
      for (int j = 0; j < ny; j++)
        for (int i = 0; i <= nx; i++)
          F(cx, i, j, 0) = 0.25 * (F(fx, 2 * i, 2 * j, 0) + F(fx, 2 * i, 2 * j + 1, 0)) +
                           0.125 * (F(fx, 2 * i + 1, 2 * j, 0) + F(fx, 2 * i + 1, 2 * j + 1, 0) +
                                    F(fx, 2 * i - 1, 2 * j, 0) + F(fx, 2 * i - 1, 2 * j + 1, 0));
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i < nx; i++)
          F(cy, i, j, 0) = 0.25 * (F(fy, 2 * i, 2 * j, 0) + F(fy, 2 * i + 1, 2 * j, 0)) +
                           0.125 * (F(fy, 2 * i, 2 * j + 1, 0) + F(fy, 2 * i + 1, 2 * j + 1, 0) +
                                    F(fy, 2 * i, 2 * j - 1, 0) + F(fy, 2 * i + 1, 2 * j - 1, 0));
    }
    return;
  }
  ofab *fx = &pf[0], *fy = &pf[1], *fz = &pf[2], *cx = &pc[0], *cy = &pc[1], *cz = &pc[2];
  if (simple) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i <= nx; i++)
          F(cx, i, j, k) = 0.25 * (F(fx, 2 * i, 2 * j, 2 * k) + F(fx, 2 * i, 2 * j + 1, 2 * k) +
                                   F(fx, 2 * i, 2 * j, 2 * k + 1) + F(fx, 2 * i, 2 * j + 1, 2 * k + 1));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i < nx; i++)
          F(cy, i, j, k) = 0.25 * (F(fy, 2 * i, 2 * j, 2 * k) + F(fy, 2 * i + 1, 2 * j, 2 * k) +
                                   F(fy, 2 * i, 2 * j, 2 * k + 1) + F(fy, 2 * i + 1, 2 * j, 2 * k + 1));
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          F(cz, i, j, k) = 0.25 * (F(fz, 2 * i, 2 * j, 2 * k) + F(fz, 2 * i + 1, 2 * j, 2 * k) +
                                   F(fz, 2 * i, 2 * j + 1, 2 * k) + F(fz, 2 * i + 1, 2 * j + 1, 2 * k));
  } else {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i <= nx; i++)
          F(cx, i, j, k) = 0.125 * (F(fx, 2 * i, 2 * j, 2 * k) + F(fx, 2 * i, 2 * j + 1, 2 * k) +
                                    F(fx, 2 * i, 2 * j, 2 * k + 1) + F(fx, 2 * i, 2 * j + 1, 2 * k + 1)) +
                           0.0625 * (F(fx, 2 * i + 1, 2 * j, 2 * k) + F(fx, 2 * i + 1, 2 * j + 1, 2 * k) +
                                     F(fx, 2 * i + 1, 2 * j, 2 * k + 1) + F(fx, 2 * i + 1, 2 * j + 1, 2 * k + 1)) +
                           0.0625 * (F(fx, 2 * i - 1, 2 * j, 2 * k) + F(fx, 2 * i - 1, 2 * j + 1, 2 * k) +
                                     F(fx, 2 * i - 1, 2 * j, 2 * k + 1) + F(fx, 2 * i - 1, 2 * j + 1, 2 * k + 1));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i < nx; i++)
          F(cy, i, j, k) = 0.125 * (F(fy, 2 * i, 2 * j, 2 * k) + F(fy, 2 * i + 1, 2 * j, 2 * k) +
                                    F(fy, 2 * i, 2 * j, 2 * k + 1) + F(fy, 2 * i + 1, 2 * j, 2 * k + 1)) +
                           0.0625 * (F(fy, 2 * i, 2 * j + 1, 2 * k) + F(fy, 2 * i + 1, 2 * j + 1, 2 * k) +
                                     F(fy, 2 * i, 2 * j + 1, 2 * k + 1) + F(fy, 2 * i + 1, 2 * j + 1, 2 * k + 1)) +
                           0.0625 * (F(fy, 2 * i, 2 * j - 1, 2 * k) + F(fy, 2 * i + 1, 2 * j - 1, 2 * k) +
                                     F(fy, 2 * i, 2 * j - 1, 2 * k + 1) + F(fy, 2 * i + 1, 2 * j - 1, 2 * k + 1));
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          F(cz, i, j, k) = 0.125 * (F(fz, 2 * i, 2 * j, 2 * k) + F(fz, 2 * i + 1, 2 * j, 2 * k) +
                                    F(fz, 2 * i, 2 * j + 1, 2 * k) + F(fz, 2 * i + 1, 2 * j + 1, 2 * k)) +
                           0.0625 * (F(fz, 2 * i, 2 * j, 2 * k + 1) + F(fz, 2 * i + 1, 2 * j, 2 * k + 1) +
                                     F(fz, 2 * i, 2 * j + 1, 2 * k + 1) + F(fz, 2 * i + 1, 2 * j + 1, 2 * k + 1)) +
                           0.0625 * (F(fz, 2 * i, 2 * j, 2 * k - 1) + F(fz, 2 * i + 1, 2 * j, 2 * k - 1) +
                                     F(fz, 2 * i, 2 * j + 1, 2 * k - 1) + F(fz, 2 * i + 1, 2 * j + 1, 2 * k - 1));
  }
}

void o_nodal_restriction(const orc_geom *gc, ofab *f, ofab *c) {
  for (int j = 0; j <= gc->n[1]; j++)
    for (int i = 0; i <= gc->n[0]; i++) F(c, i, j, 0) = F(f, 2 * i, 2 * j, 0);
  fab_internal_sync(c);
}

void o_edge_restriction(const orc_geom *gc, ofab *f, ofab *c) {
  int nx = gc->n[0], ny = gc->n[1], nz = gc->n[2];
  for (int k = 0; k < nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++)
        F(&c[0], i, j, k) = 0.5 * (F(&f[0], 2 * i, 2 * j, 2 * k) + F(&f[0], 2 * i, 2 * j, 2 * k + 1));
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i <= nx; i++)
        F(&c[1], i, j, k) = 0.5 * (F(&f[1], 2 * i, 2 * j, 2 * k) + F(&f[1], 2 * i, 2 * j + 1, 2 * k));
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i < nx; i++)
        F(&c[2], i, j, k) = 0.5 * (F(&f[2], 2 * i, 2 * j, 2 * k) + F(&f[2], 2 * i + 1, 2 * j, 2 * k));
  for (int e = 0; e < 3; e++) fab_internal_sync(&c[e]);
}

void o_stag_prolongation(const orc_geom *gf, ofab *pf, ofab *pc) {
  int nx = gf->n[0], ny = gf->n[1], nz = gf->n[2];
  if (gf->dim == 2) {
    ofab *fx = &pf[0], *fy = &pf[1], *cx = &pc[0], *cy = &pc[1];
    for (int j = 0; j < ny; j++)
      for (int i = 0; i <= nx; i++) {
        int joff = (j % 2 == 0) ? -1 : 1;
        if (i % 2 == 0)
          F(fx, i, j, 0) = F(fx, i, j, 0) + 0.75 * F(cx, i / 2, j / 2, 0) + 0.25 * F(cx, i / 2, j / 2 + joff, 0);
        else
          F(fx, i, j, 0) = F(fx, i, j, 0) + 0.375 * F(cx, i / 2, j / 2, 0) + 0.125 * F(cx, i / 2, j / 2 + joff, 0) +
                           0.375 * F(cx, i / 2 + 1, j / 2, 0) + 0.125 * F(cx, i / 2 + 1, j / 2 + joff, 0);
      }
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i < nx; i++) {
        int ioff = (i % 2 == 0) ? -1 : 1;
        if (j % 2 == 0)
          F(fy, i, j, 0) = F(fy, i, j, 0) + 0.75 * F(cy, i / 2, j / 2, 0) + 0.25 * F(cy, i / 2 + ioff, j / 2, 0);
        else
          F(fy, i, j, 0) = F(fy, i, j, 0) + 0.375 * F(cy, i / 2, j / 2, 0) + 0.125 * F(cy, i / 2 + ioff, j / 2, 0) +
                           0.375 * F(cy, i / 2, j / 2 + 1, 0) + 0.125 * F(cy, i / 2 + ioff, j / 2 + 1, 0);
      }
    return;
  }
  const double nine16 = 9.0 / 16.0, three16 = 3.0 / 16.0, one16 = 1.0 / 16.0;
  const double nine32 = 9.0 / 32.0, three32 = 3.0 / 32.0, one32 = 1.0 / 32.0;
  ofab *fx = &pf[0], *fy = &pf[1], *fz = &pf[2], *cx = &pc[0], *cy = &pc[1], *cz = &pc[2];
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i <= nx; i++) {
        int joff = (j % 2 == 0) ? -1 : 1, koff = (k % 2 == 0) ? -1 : 1;
        int I = i / 2, J = j / 2, K = k / 2;
        if (i % 2 == 0)
          F(fx, i, j, k) = F(fx, i, j, k) + nine16 * F(cx, I, J, K) + three16 * F(cx, I, J + joff, K) +
                           three16 * F(cx, I, J, K + koff) + one16 * F(cx, I, J + joff, K + koff);
        else
          F(fx, i, j, k) = F(fx, i, j, k) + nine32 * F(cx, I, J, K) + three32 * F(cx, I, J + joff, K) +
                           three32 * F(cx, I, J, K + koff) + one32 * F(cx, I, J + joff, K + koff) +
                           nine32 * F(cx, I + 1, J, K) + three32 * F(cx, I + 1, J + joff, K) +
                           three32 * F(cx, I + 1, J, K + koff) + one32 * F(cx, I + 1, J + joff, K + koff);
      }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i < nx; i++) {
        int ioff = (i % 2 == 0) ? -1 : 1, koff = (k % 2 == 0) ? -1 : 1;
        int I = i / 2, J = j / 2, K = k / 2;
        if (j % 2 == 0)
          F(fy, i, j, k) = F(fy, i, j, k) + nine16 * F(cy, I, J, K) + three16 * F(cy, I + ioff, J, K) +
                           three16 * F(cy, I, J, K + koff) + one16 * F(cy, I + ioff, J, K + koff);
        else
          F(fy, i, j, k) = F(fy, i, j, k) + nine32 * F(cy, I, J, K) + three32 * F(cy, I + ioff, J, K) +
                           three32 * F(cy, I, J, K + koff) + one32 * F(cy, I + ioff, J, K + koff) +
                           nine32 * F(cy, I, J + 1, K) + three32 * F(cy, I + ioff, J + 1, K) +
                           three32 * F(cy, I, J + 1, K + koff) + one32 * F(cy, I + ioff, J + 1, K + koff);
      }
#pragma omp parallel for schedule(static)
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int ioff = (i % 2 == 0) ? -1 : 1, joff = (j % 2 == 0) ? -1 : 1;
        int I = i / 2, J = j / 2, K = k / 2;
        if (k % 2 == 0)
          F(fz, i, j, k) = F(fz, i, j, k) + nine16 * F(cz, I, J, K) + three16 * F(cz, I + ioff, J, K) +
                           three16 * F(cz, I, J + joff, K) + one16 * F(cz, I + ioff, J + joff, K);
        else
          F(fz, i, j, k) = F(fz, i, j, k) + nine32 * F(cz, I, J, K) + three32 * F(cz, I + ioff, J, K) +
                           three32 * F(cz, I, J + joff, K) + one32 * F(cz, I + ioff, J + joff, K) +
                           nine32 * F(cz, I, J, K + 1) + three32 * F(cz, I + ioff, J, K + 1) +
                           three32 * F(cz, I, J + joff, K + 1) + one32 * F(cz, I + ioff, J + joff, K + 1);
      }
}

/* ------------------------------------------------------------------ */
/* compute_div / compute_grad (periodic: no boundary stencil changes)   */
/* ------------------------------------------------------------------ */

void o_compute_div(const orc_geom *g, ofab *u, ofab *div, int increment) {
  double dxinv = 1.0 / g->dx;
  int nz = (g->dim == 3) ? g->n[2] : 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        double v;
        if (g->dim == 2) {
          v = (F(&u[0], i + 1, j, 0) - F(&u[0], i, j, 0)) * dxinv;
          if (increment) v = F(div, i, j, 0) + v;
          v = v + (F(&u[1], i, j + 1, 0) - F(&u[1], i, j, 0)) * dxinv;
        } else {
          v = (F(&u[0], i + 1, j, k) - F(&u[0], i, j, k)) * dxinv;
          if (increment) v = F(div, i, j, k) + v;
          v = v + (F(&u[1], i, j + 1, k) - F(&u[1], i, j, k)) * dxinv;
          v = v + (F(&u[2], i, j, k + 1) - F(&u[2], i, j, k)) * dxinv;
        }
        F(div, i, j, k) = v;
      }
}

void o_compute_grad(const orc_geom *g, ofab *p, ofab *gp) {
  double dxinv = 1.0 / g->dx;
  int nz = (g->dim == 3) ? g->n[2] : 1;
  for (int d = 0; d < g->dim; d++) {
    ofab *o = &gp[d];
    int s[3] = {0, 0, 0};
    s[d] = 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= vhi(o, 2); k++)
      for (int j = 0; j <= vhi(o, 1); j++)
        for (int i = 0; i <= vhi(o, 0); i++)
          F(o, i, j, k) = (F(p, i, j, k) - F(p, i - s[0], j - s[1], k - s[2])) * dxinv;
  }
  (void)nz;
}

/* ------------------------------------------------------------------ */
/* inner products: box-boundary normal faces (lo and hi+1) get an extra */
/* 0.5*m1*m2 on top of the plain sum, i.e. weight 1.5                   */
/* ------------------------------------------------------------------ */

void o_stag_inner_prod(const orc_geom *g, ofab *m1, ofab *m2, double out[3]) {
  for (int d = 0; d < g->dim; d++) {
    ofab *a = &m1[d], *b = &m2[d];
    int nk = vhi(a, 2) + 1;
    double *part = (double *)calloc((size_t)nk, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nk; k++) {
      double s = 0.0;
      for (int j = 0; j <= vhi(a, 1); j++)
        for (int i = 0; i <= vhi(a, 0); i++) s += F(a, i, j, k) * F(b, i, j, k);
      part[k] = s;
    }
    double s = 0.0;
    for (int k = 0; k < nk; k++) s += part[k];
    free(part);
    /* boundary faces */
    int n = g->n[d];
    for (int side = 0; side < 2; side++) {
      int idx = side ? n : 0;
      for (int k = 0; k <= vhi(a, 2); k++)
        for (int j = 0; j <= vhi(a, 1); j++)
          for (int i = 0; i <= vhi(a, 0); i++) {
            int c[3] = {i, j, k};
            if (c[d] != idx) continue;
            s += 0.5 * F(a, i, j, k) * F(b, i, j, k);
          }
    }
    out[d] = s;
  }
}

double o_cc_inner_prod(const orc_geom *g, ofab *a, ofab *b) {
  int nk = (g->dim == 3) ? g->n[2] : 1;
  double *part = (double *)calloc((size_t)nk, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nk; k++) {
    double s = 0.0;
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) s += F(a, i, j, k) * F(b, i, j, k);
    part[k] = s;
  }
  double s = 0.0;
  for (int k = 0; k < nk; k++) s += part[k];
  free(part);
  return s;
}

/* sum_umac_press (norm_inner_product.f90:579-794) for one tile spanning the
 * whole box: pressure sum over valid cells; each velocity component sums
 * all valid faces and adds 0.5*u on the lo and hi+1 faces. */
/* sum_umac_2d/3d add the 0.5 boundary correction in direction d only on a
 * tile whose nodal box spans the whole box in d (norm_inner_product.f90:
 * 701-705 and the 3D analogue).  sum_umac_press iterates tiles of BoxLib's
 * default tile size, 1024000 x 8 x 8 (F_BaseLib layout; the BoxLib version
 * is not pinned): the x tiles always span the box, y and z tiles only when
 * the box is at most 8 cells long in that direction. */
static const int o_tile_size[3] = {1024000, 8, 8};

void o_sum_umac_press(const orc_geom *g, ofab *p, ofab *u, double *av_p, double av_u[3]) {
  double total_volume = (double)g->n[0] * g->n[1] * ((g->dim == 3) ? g->n[2] : 1);
  double sp = 0.0;
  int nk = (g->dim == 3) ? g->n[2] : 1;
  for (int k = 0; k < nk; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) sp += F(p, i, j, k);
  *av_p = sp / total_volume;
  if (!u) return;
  for (int d = 0; d < g->dim; d++) {
    ofab *a = &u[d];
    double s = 0.0;
    for (int k = 0; k <= vhi(a, 2); k++)
      for (int j = 0; j <= vhi(a, 1); j++)
        for (int i = 0; i <= vhi(a, 0); i++) s += F(a, i, j, k);
    int n = g->n[d];
    if (n > o_tile_size[d]) {
      av_u[d] = s / total_volume;
      continue;
    }
    for (int k = 0; k <= vhi(a, 2); k++)
      for (int j = 0; j <= vhi(a, 1); j++)
        for (int i = 0; i <= vhi(a, 0); i++) {
          int c[3] = {i, j, k};
          if (c[d] != 0) continue;
          int h[3] = {i, j, k};
          h[d] = n;
          s += 0.5 * F(a, i, j, k);
          s += 0.5 * F(a, h[0], h[1], h[2]);
        }
    av_u[d] = s / total_volume;
  }
}
