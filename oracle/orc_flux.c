/*
 * Fluctuating and diffusive fluxes, coefficient averaging and the Gaussian
 * generator of the CPU oracle (test infrastructure only).
 *   orc_stochastic_m_fluxdiv  <- src_common/stochastic_m_fluxdiv.f90:41-553
 *        (periodic box: mflux_bc_* and physbc are no-ops; the caller passes
 *         the already weighted noise sum of :114-123), with fab_filter
 *        <- src_common/multifab_filter.f90:13-200
 *   orc_diffusive_mass_fluxdiv <- src_multiSpec/diffusive_mass_fluxdiv.f90:26-262
 *        (isothermal, barodiffusion_type 0) with matvec_mul.f90 and
 *        correction_flux.f90:71-143
 *   orc_average_cc_to_*        <- src_common/convert_stag.f90:229-863 (interior
 *        formulas; periodic box has no boundary overrides)
 *   orc_mt_* / orc_inv_norm_dist <- src_random/RNGs.c:62-260, restated from the
 *        published methods: MT19937 (Matsumoto & Nishimura, ACM TOMACS 8(1),
 *        1998, with the 2002 initialisation) and P. J. Acklam's rational
 *        approximation of the inverse normal CDF.
 */
#include "orc_fab.h"

/* ---------------- coefficient averaging ---------------- */

int orc_average_cc_to_face(const orc_geom *g0, const double *cc, double *face[3]) {
  orc_geom g = *g0;
  if (g.dim == 2) g.n[2] = 1;
  ofab c, f;
  cell_build(&c, &g, 1);
  fab_from_dedup(&c, cc);
  for (int d = 0; d < g.dim; d++) {
    face_build(&f, &g, d, 0);
    int s[3] = {0, 0, 0};
    s[d] = 1;
    for (int k = 0; k <= vhi(&f, 2); k++)
      for (int j = 0; j <= vhi(&f, 1); j++)
        for (int i = 0; i <= vhi(&f, 0); i++)
          F(&f, i, j, k) = 0.5 * (F(&c, i, j, k) + F(&c, i - s[0], j - s[1], k - s[2]));
    fab_to_dedup(&f, face[d]);
    fab_free(&f);
  }
  fab_free(&c);
  return 0;
}

int orc_average_cc_to_node(const orc_geom *g0, const double *cc, double *node) {
  orc_geom g = *g0;
  if (g.dim != 2) return 2;
  g.n[2] = 1;
  ofab c, nd;
  cell_build(&c, &g, 1);
  fab_from_dedup(&c, cc);
  edge_build(&nd, &g, 0, 0);
  for (int j = 0; j <= g.n[1]; j++)
    for (int i = 0; i <= g.n[0]; i++)
      F(&nd, i, j, 0) = 0.25 * (F(&c, i, j, 0) + F(&c, i - 1, j, 0) + F(&c, i, j - 1, 0) + F(&c, i - 1, j - 1, 0));
  fab_to_dedup(&nd, node);
  fab_free(&c);
  fab_free(&nd);
  return 0;
}

int orc_average_cc_to_edge(const orc_geom *g0, const double *cc, double *edge[3]) {
  orc_geom g = *g0;
  if (g.dim != 3) return 2;
  ofab c, e;
  cell_build(&c, &g, 1);
  fab_from_dedup(&c, cc);
  edge_build(&e, &g, 0, 0);
  for (int k = 0; k < g.n[2]; k++)
    for (int j = 0; j <= g.n[1]; j++)
      for (int i = 0; i <= g.n[0]; i++)
        F(&e, i, j, k) = 0.25 * (F(&c, i, j, k) + F(&c, i - 1, j, k) + F(&c, i, j - 1, k) + F(&c, i - 1, j - 1, k));
  fab_to_dedup(&e, edge[0]);
  fab_free(&e);
  edge_build(&e, &g, 1, 0);
  for (int k = 0; k <= g.n[2]; k++)
    for (int j = 0; j < g.n[1]; j++)
      for (int i = 0; i <= g.n[0]; i++)
        F(&e, i, j, k) = 0.25 * (F(&c, i, j, k) + F(&c, i - 1, j, k) + F(&c, i, j, k - 1) + F(&c, i - 1, j, k - 1));
  fab_to_dedup(&e, edge[1]);
  fab_free(&e);
  edge_build(&e, &g, 2, 0);
  for (int k = 0; k <= g.n[2]; k++)
    for (int j = 0; j <= g.n[1]; j++)
      for (int i = 0; i < g.n[0]; i++)
        F(&e, i, j, k) = 0.25 * (F(&c, i, j, k) + F(&c, i, j - 1, k) + F(&c, i, j, k - 1) + F(&c, i, j - 1, k - 1));
  fab_to_dedup(&e, edge[2]);
  fab_free(&e);
  fab_free(&c);
  return 0;
}

/* ---------------- stochastic momentum flux divergence ---------------- */

/* multifab_filter(mfab, width): along x, then y, then z, each pass over the
 * valid region (nodal hi+1 included) from the current ghosts into a
 * temporary, copied back, ghosts refilled before the next pass (not after the
 * last).  Terms in the reference's order. */
static void fab_filter(ofab *f, int width) {
  const int dm = f->dim;
  const int lo0 = 0, hi0 = vhi(f, 0), hi1 = vhi(f, 1), hi2 = (dm == 3) ? vhi(f, 2) : 0;
  const long cnt = (long)(hi0 + 1) * (hi1 + 1) * (hi2 + 1);
  double *tmp = (double *)malloc(sizeof(double) * cnt);
  for (int dir = 0; dir < dm; dir++) {
    if (dir > 0) fab_fill_boundary(f);
    const int sx = dir == 0, sy = dir == 1, sz = dir == 2;
    long q = 0;
    for (int k = 0; k <= hi2; k++)
      for (int j = 0; j <= hi1; j++)
        for (int i = lo0; i <= hi0; i++) {
#define FP(o) F(f, i + (o) * sx, j + (o) * sy, k + (o) * sz)
          double v;
          if (width == 1)
            v = FP(0) / 2 + FP(-1) / 4 + FP(1) / 4;
          else if (width == 2)
            v = 5 * FP(0) / 8 - FP(-2) / 16 + FP(-1) / 4 - FP(2) / 16 + FP(1) / 4;
          else
            v = 93 * FP(0) / 128 + 7 * FP(-1) / 32 - 7 * FP(-2) / 64 + FP(-3) / 32 - FP(-4) / 256 +
                7 * FP(1) / 32 - 7 * FP(2) / 64 + FP(3) / 32 - FP(4) / 256;
#undef FP
          tmp[q++] = v;
        }
    q = 0;
    for (int k = 0; k <= hi2; k++)
      for (int j = 0; j <= hi1; j++)
        for (int i = lo0; i <= hi0; i++) F(f, i, j, k) = tmp[q++];
  }
  free(tmp);
}

int orc_stochastic_m_fluxdiv(const orc_geom *g0, const double *W_cc, double *const W_ed[3],
                             const double *eta, double *const eta_ed[3], const double *temp,
                             double *const temp_ed[3], double variance, int width, double *force[3]) {
  if (!(width == 0 || width == 1 || width == 2 || width == 4)) return 1;
  const int ngw = width > 1 ? width : 1;
  orc_geom g = *g0;
  int dm = g.dim;
  if (dm == 2) g.n[2] = 1;
  long N = (long)g.n[0] * g.n[1] * g.n[2];
  ofab wc[3], we[3][2], et, tt, ee[3], te[3], fo[3];
  cell_build(&et, &g, 1);
  fab_from_dedup(&et, eta);
  cell_build(&tt, &g, 1);
  fab_from_dedup(&tt, temp);
  int ne = (dm == 2) ? 1 : 3;
  for (int e = 0; e < ne; e++) {
    edge_build(&ee[e], &g, e, 0);
    fab_from_dedup(&ee[e], eta_ed[e]);
    edge_build(&te[e], &g, e, 0);
    fab_from_dedup(&te[e], temp_ed[e]);
  }
  /* cell noise: variance * W * sqrt(eta*T) on valid cells, then periodic ghosts */
  for (int d = 0; d < dm; d++) {
    cell_build(&wc[d], &g, ngw);
    fab_from_dedup(&wc[d], W_cc + d * N);
    for (int k = 0; k < (dm == 3 ? g.n[2] : 1); k++)
      for (int j = 0; j < g.n[1]; j++)
        for (int i = 0; i < g.n[0]; i++) {
          double v = variance * F(&wc[d], i, j, k);
          F(&wc[d], i, j, k) = v * sqrt(F(&et, i, j, k) * F(&tt, i, j, k));
        }
    fab_fill_boundary(&wc[d]);
  }
  /* nodal/edge noise: variance * W * sqrt(eta_ed*T_ed) over the valid range */
  for (int e = 0; e < ne; e++)
    for (int q = 0; q < 2; q++) {
      ofab *w = &we[e][q];
      edge_build(w, &g, e, ngw);
      fab_from_dedup(w, W_ed[e] + q * N);
      for (int k = 0; k <= vhi(w, 2); k++)
        for (int j = 0; j <= vhi(w, 1); j++)
          for (int i = 0; i <= vhi(w, 0); i++) {
            double v = variance * F(w, i, j, k);
            F(w, i, j, k) = v * sqrt(F(&ee[e], i, j, k) * F(&te[e], i, j, k));
          }
      fab_internal_sync(w);
      fab_fill_boundary(w);
    }
  if (width > 0) {  /* stochastic_m_fluxdiv.f90:166-174 (2D), :224-230 (3D) */
    for (int e = 0; e < ne; e++)
      for (int q = 0; q < 2; q++) fab_filter(&we[e][q], width);
    for (int d = 0; d < dm; d++) {
      fab_filter(&wc[d], width);
      fab_fill_boundary(&wc[d]);
    }
  }
  double dxinv = 1.0 / g.dx;
  for (int d = 0; d < dm; d++) {
    face_build(&fo[d], &g, d, 0);
    fab_from_dedup(&fo[d], force[d]);
  }
  if (dm == 2) {
    ofab *nd1 = &we[0][0], *nd2 = &we[0][1];
    for (int j = 0; j < g.n[1]; j++)
      for (int i = 0; i <= g.n[0]; i++)
        F(&fo[0], i, j, 0) = F(&fo[0], i, j, 0) + (F(&wc[0], i, j, 0) - F(&wc[0], i - 1, j, 0)) * dxinv +
                             (F(nd1, i, j + 1, 0) - F(nd1, i, j, 0)) * dxinv;
    for (int j = 0; j <= g.n[1]; j++)
      for (int i = 0; i < g.n[0]; i++)
        F(&fo[1], i, j, 0) = F(&fo[1], i, j, 0) + (F(nd2, i + 1, j, 0) - F(nd2, i, j, 0)) * dxinv +
                             (F(&wc[1], i, j, 0) - F(&wc[1], i, j - 1, 0)) * dxinv;
  } else {
    ofab *xy1 = &we[0][0], *xy2 = &we[0][1], *xz1 = &we[1][0], *xz2 = &we[1][1];
    ofab *yz1 = &we[2][0], *yz2 = &we[2][1];
    for (int k = 0; k < g.n[2]; k++)
      for (int j = 0; j < g.n[1]; j++)
        for (int i = 0; i <= g.n[0]; i++)
          F(&fo[0], i, j, k) = F(&fo[0], i, j, k) + (F(&wc[0], i, j, k) - F(&wc[0], i - 1, j, k)) * dxinv +
                               (F(xy1, i, j + 1, k) - F(xy1, i, j, k)) * dxinv +
                               (F(xz1, i, j, k + 1) - F(xz1, i, j, k)) * dxinv;
    for (int k = 0; k < g.n[2]; k++)
      for (int j = 0; j <= g.n[1]; j++)
        for (int i = 0; i < g.n[0]; i++)
          F(&fo[1], i, j, k) = F(&fo[1], i, j, k) + (F(xy2, i + 1, j, k) - F(xy2, i, j, k)) * dxinv +
                               (F(&wc[1], i, j, k) - F(&wc[1], i, j - 1, k)) * dxinv +
                               (F(yz1, i, j, k + 1) - F(yz1, i, j, k)) * dxinv;
    for (int k = 0; k <= g.n[2]; k++)
      for (int j = 0; j < g.n[1]; j++)
        for (int i = 0; i < g.n[0]; i++)
          F(&fo[2], i, j, k) = F(&fo[2], i, j, k) + (F(xz2, i + 1, j, k) - F(xz2, i, j, k)) * dxinv +
                               (F(yz2, i, j + 1, k) - F(yz2, i, j, k)) * dxinv +
                               (F(&wc[2], i, j, k) - F(&wc[2], i, j, k - 1)) * dxinv;
  }
  double mm = 0.0;
  for (int d = 0; d < dm; d++) {
    double x = fab_to_dedup(&fo[d], force[d]);
    if (x > mm) mm = x;
    fab_free(&fo[d]);
    fab_free(&wc[d]);
  }
  for (int e = 0; e < ne; e++) {
    fab_free(&ee[e]);
    fab_free(&te[e]);
    fab_free(&we[e][0]);
    fab_free(&we[e][1]);
  }
  fab_free(&et);
  fab_free(&tt);
  return mm == 0.0 ? 0 : 3;
}

/* ---------------- species diffusion flux divergence ---------------- */

int orc_diffusive_mass_fluxdiv(const orc_geom *g0, int ns, const double *rho, const double *rhotot,
                               const double *molarconc, const double *rhoWchi, const double *Gama,
                               double *diff_fluxdiv, double *flux_total[3], int correct_flux) {
  orc_geom g = *g0;
  int dm = g.dim;
  if (dm == 2) g.n[2] = 1;
  long N = (long)g.n[0] * g.n[1] * g.n[2];
  int nn = ns * ns;
  (void)rho;
  (void)rhotot;
  ofab *xc = calloc((size_t)ns, sizeof(ofab));
  ofab *rw = calloc((size_t)nn, sizeof(ofab)), *gm = calloc((size_t)nn, sizeof(ofab));
  for (int s = 0; s < ns; s++) {
    cell_build(&xc[s], &g, 1);
    fab_from_dedup(&xc[s], molarconc + s * N);
  }
  for (int q = 0; q < nn; q++) {
    cell_build(&rw[q], &g, 1);
    fab_from_dedup(&rw[q], rhoWchi + q * N);
    cell_build(&gm[q], &g, 1);
    fab_from_dedup(&gm[q], Gama + q * N);
  }
  double dxinv = 1.0 / g.dx;
  double *x = malloc(sizeof(double) * ns), *y = malloc(sizeof(double) * ns);
  double *A = malloc(sizeof(double) * nn), *Bm = malloc(sizeof(double) * nn);
  double *flux = malloc(sizeof(double) * (size_t)ns * N * 3);
  for (int d = 0; d < dm; d++) {
    int sft[3] = {0, 0, 0};
    sft[d] = 1;
    for (int k = 0; k < g.n[2]; k++)
      for (int j = 0; j < g.n[1]; j++)
        for (int i = 0; i < g.n[0]; i++) {
          int im = i - sft[0], jm = j - sft[1], km = k - sft[2];
          for (int s = 0; s < ns; s++) x[s] = (F(&xc[s], i, j, k) - F(&xc[s], im, jm, km)) * dxinv;
          for (int q = 0; q < nn; q++) {
            A[q] = 0.5 * (F(&gm[q], i, j, k) + F(&gm[q], im, jm, km));
            Bm[q] = 0.5 * (F(&rw[q], i, j, k) + F(&rw[q], im, jm, km));
          }
          /* x = Gama_face * x */
          for (int r = 0; r < ns; r++) {
            double acc = 0.0;
            for (int c = 0; c < ns; c++) acc = acc + A[r + ns * c] * x[c];
            y[r] = acc;
          }
          /* x = rhoWchi_face * x */
          for (int r = 0; r < ns; r++) {
            double acc = 0.0;
            for (int c = 0; c < ns; c++) acc = acc + Bm[r + ns * c] * y[c];
            x[r] = acc;
          }
          if (correct_flux && ns > 1) {
            double sum = 0.0;
            for (int s = 0; s < ns - 1; s++) sum = sum + x[s];
            x[ns - 1] = -sum;
          }
          long idx = i + (long)g.n[0] * (j + (long)g.n[1] * k);
          for (int s = 0; s < ns; s++) {
            flux[((long)d * ns + s) * N + idx] = x[s];
            flux_total[d][s * N + idx] = flux_total[d][s * N + idx] + x[s];
          }
        }
  }
  for (int s = 0; s < ns; s++)
    for (int k = 0; k < g.n[2]; k++)
      for (int j = 0; j < g.n[1]; j++)
        for (int i = 0; i < g.n[0]; i++) {
          long idx = i + (long)g.n[0] * (j + (long)g.n[1] * k);
          long ip = ((i + 1) % g.n[0]) + (long)g.n[0] * (j + (long)g.n[1] * k);
          long jp = i + (long)g.n[0] * (((j + 1) % g.n[1]) + (long)g.n[1] * k);
          long kp = i + (long)g.n[0] * (j + (long)g.n[1] * ((k + 1) % g.n[2]));
          double v = (flux[(0L * ns + s) * N + ip] - flux[(0L * ns + s) * N + idx]) * dxinv;
          v = v + (flux[(1L * ns + s) * N + jp] - flux[(1L * ns + s) * N + idx]) * dxinv;
          if (dm == 3) v = v + (flux[(2L * ns + s) * N + kp] - flux[(2L * ns + s) * N + idx]) * dxinv;
          diff_fluxdiv[s * N + idx] = v;
        }
  for (int s = 0; s < ns; s++) fab_free(&xc[s]);
  for (int q = 0; q < nn; q++) {
    fab_free(&rw[q]);
    fab_free(&gm[q]);
  }
  free(xc);
  free(rw);
  free(gm);
  free(x);
  free(y);
  free(A);
  free(Bm);
  free(flux);
  return 0;
}

/* ---------------- MT19937 + inverse normal ---------------- */

void orc_mt_seed(orc_mt *s, uint32_t seed) {
  s->mt[0] = seed;
  for (int q = 1; q < 624; q++) s->mt[q] = 1812433253u * (s->mt[q - 1] ^ (s->mt[q - 1] >> 30)) + (uint32_t)q;
  s->mti = 624;
}

static void mt_twist(orc_mt *s) {
  static const uint32_t A = 0x9908b0dfu;
  for (int q = 0; q < 624; q++) {
    uint32_t y = (s->mt[q] & 0x80000000u) | (s->mt[(q + 1) % 624] & 0x7fffffffu);
    uint32_t v = s->mt[(q + 397) % 624] ^ (y >> 1);
    if (y & 1u) v ^= A;
    s->mt[q] = v;
  }
  s->mti = 0;
}

uint32_t orc_mt_next(orc_mt *s) {
  if (s->mti >= 624) mt_twist(s);
  uint32_t y = s->mt[s->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

double orc_mt_uniform(orc_mt *s) { return (double)orc_mt_next(s) * (1.0 / 4294967296.0); }

double orc_inv_norm_dist(double p) {
  /* Acklam's coefficients (published with the algorithm) */
  static const double a1 = -3.969683028665376e+01, a2 = 2.209460984245205e+02, a3 = -2.759285104469687e+02,
                      a4 = 1.383577518672690e+02, a5 = -3.066479806614716e+01, a6 = 2.506628277459239e+00;
  static const double b1 = -5.447609879822406e+01, b2 = 1.615858368580409e+02, b3 = -1.556989798598866e+02,
                      b4 = 6.680131188771972e+01, b5 = -1.328068155288572e+01;
  static const double c1 = -7.784894002430293e-03, c2 = -3.223964580411365e-01, c3 = -2.400758277161838e+00,
                      c4 = -2.549732539343734e+00, c5 = 4.374664141464968e+00, c6 = 2.938163982698783e+00;
  static const double d1 = 7.784695709041462e-03, d2 = 3.224671290700398e-01, d3 = 2.445134137142996e+00,
                      d4 = 3.754408661907416e+00;
  const double plow = 0.02425, phigh = 1.0 - plow;
  if (!(p > 0.0 && p < 1.0)) return NAN;
  if (p < plow) {
    double q = sqrt(-2.0 * log(p));
    return (((((c1 * q + c2) * q + c3) * q + c4) * q + c5) * q + c6) / ((((d1 * q + d2) * q + d3) * q + d4) * q + 1.0);
  }
  if (p > phigh) {
    double q = sqrt(-2.0 * log(1.0 - p));
    return -(((((c1 * q + c2) * q + c3) * q + c4) * q + c5) * q + c6) /
           ((((d1 * q + d2) * q + d3) * q + d4) * q + 1.0);
  }
  double q = p - 0.5, r = q * q;
  return (((((a1 * r + a2) * r + a3) * r + a4) * r + a5) * r + a6) * q /
         (((((b1 * r + b2) * r + b3) * r + b4) * r + b5) * r + 1.0);
}

double orc_mt_normal(orc_mt *s) {
  double u;
  do {
    u = orc_mt_uniform(s);
  } while (u == 0.0);
  return orc_inv_norm_dist(u);
}

void orc_fill_normal(uint32_t seed, int64_t count, double *out) {
  orc_mt s;
  orc_mt_seed(&s, seed);
  for (int64_t q = 0; q < count; q++) out[q] = orc_mt_normal(&s);
}
