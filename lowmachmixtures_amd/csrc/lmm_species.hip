// Species mass-flux chain on gfx950 (declarations and layout in
// lmm_species.h).
//
// Every per-cell routine is a template on the species count NS (2..8; the
// face fluxes also take 1) so that the small dense matrices live in VGPRs
// with static indices.  The arithmetic follows the reference routine by
// routine, operation by operation (sums from 0.0 in ascending index order,
// the same association of products and quotients), so the FP64 results
// equal the CPU restatement bit for bit:
//   correct_rho_with_drho_local   mass_flux_utilities.f90:127-150
//   compute_molconc_rhotot_local  :279-313
//   mixture_properties_mass_local compute_mixture_properties.f90:165-233
//   compute_Gama_local            mass_flux_utilities.f90:562-599
//   compute_chi_local             :767-831 (+ Dbar2chi_iterative,
//                                 matrix_utilities.f90:21-156, or the
//                                 use_lapack form :833-927)
//   compute_rhoWchi_local         :1250-1273
//   compute_Lonsager_local        :1068-1118 (+ choldc, matrix_utilities.f90:160-239)
//   compute_baro_coef_2d          :1342-1369
//   diffusive_mass_flux           diffusive_mass_fluxdiv.f90:92-262
//   stochastic_mass_fluxdiv       stochastic_mass_fluxdiv.f90:32-154
//   correction_flux               correction_flux.f90:71-251
//
// Dbar2chi_iterative updates chi column by column: column j of iteration
// k+1 depends only on column j of iteration k, so running all iterations on
// one column before the next gives the same operations on every entry while
// holding one column instead of three matrices.
//
// compute_mass_fluxdiv.f90 runs as one cell pass (sp_chain_cell), which
// writes only what the faces need (mole fractions, rho W chi, Gama unless
// the mixture is ideal, zeta/T, the packed lower triangle of the Cholesky
// factor of L-Onsager) and restores rho, then (up to 4 species) one face
// pass computing both fluxes of every face (flux_total written once) and one
// pass for both divergences, or (more species) a pass and a divergence per
// flux (sp_chain_face).  Giving each cell
// its upper faces too (recomputed, so the face arrays need not be stored)
// was tried: the two matrix-vector chains per face kept so many loads live
// that the kernel ran at one wave per SIMD with spills from 5 species up.
#include "lmm.h"
#include "lmm_kernels.h"
#include "lmm_points.h"
#include "lmm_species.h"

namespace lmm {
namespace {

constexpr int STPB = 256;

#define SLAUNCH(K, G, B, SHM, S, ...)                   \
  do {                                                  \
    hipLaunchKernelGGL(K, G, B, SHM, S, __VA_ARGS__);   \
    post_launch(S, #K);                                 \
  } while (0)

inline int sblk(long n) { return (int)((n + STPB - 1) / STPB); }

// strict lower triangle index of (r, c), r > c, in the Dbar / H_offdiag order
__host__ __device__ __forceinline__ int tri(int r, int c) { return r * (r - 1) / 2 + c; }
// packed lower triangle by columns: (r, c), r >= c
template <int NS>
__host__ __device__ __forceinline__ int lpk_index(int r, int c) {
  return c * NS - c * (c - 1) / 2 + (r - c);
}

// D_bar(i, j) from the constant knobs (mixture_properties_mass_local)
struct DParam {
  const SpParams &sp;
  __device__ double operator()(int i, int j) const {
    return i == j ? 0.0 : sp.Dbar[i > j ? tri(i, j) : tri(j, i)];
  }
};
// D_bar(i, j) from a field of ns*ns components
template <int NS>
struct DField {
  const double *p;
  long N, t;
  __device__ double operator()(int i, int j) const { return p[(long)(i + NS * j) * N + t]; }
};

template <int NS>
__device__ __forceinline__ void load(const double *f, long N, long t, double (&v)[NS]) {
#pragma unroll
  for (int q = 0; q < NS; q++) v[q] = f[(long)q * N + t];
}

template <int NS>
__device__ __forceinline__ void cell_correct(const SpParams &sp, double (&rho)[NS], double (&drho)[NS]) {
  double old[NS], tot = 0.0;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    old[n] = rho[n];
    tot = tot + rho[n];
  }
  const double f = sp.ftol * tot;
#pragma unroll
  for (int n = 0; n < NS; n++) rho[n] = rho[n] > f ? rho[n] : f;
#pragma unroll
  for (int n = 0; n < NS; n++) drho[n] = rho[n] - old[n];
}

template <int NS>
__device__ __forceinline__ void cell_prim(const SpParams &sp, const double (&rho)[NS], double &tot,
                                          double (&x)[NS], double &mt) {
  tot = 0.0;
#pragma unroll
  for (int n = 0; n < NS; n++) tot = tot + rho[n];
  double W[NS], s = 0.0;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    W[n] = rho[n] / tot;
    s = s + W[n] / sp.molmass[n];
  }
  mt = 1.0 / s;
#pragma unroll
  for (int n = 0; n < NS; n++) x[n] = mt * W[n] / sp.molmass[n];
}

// Gama(r, c) = delta_rc + sum_k X(r, k) H(k, c), X = diag(x) - x x^T
// (zero for an ideal mixture); H(k, c) from the callable.
template <int NS, class HF>
__device__ __forceinline__ double gama_entry(const SpParams &sp, const double (&x)[NS], int r, int c, HF H) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    double X = 0.0;
    if (!sp.ideal) X = (k == r) ? x[r] - x[r] * x[r] : -(x[r] * x[k]);
    acc = acc + X * H(k, c);
  }
  return (r == c ? 1.0 : 0.0) + acc;
}

// Lambda(r, c), r != c, of compute_chi_local: the value set at (max, min)
template <int NS, class DF>
__device__ __forceinline__ double lambda_off(const double (&x)[NS], int r, int c, DF D) {
  const int hi = r > c ? r : c, lo = r > c ? c : r;
  return -(x[hi] * x[lo] / D(hi, lo));
}

// zeta_by_Temp of compute_chi_local (non-isothermal)
template <int NS, class DF>
__device__ __forceinline__ void cell_zeta(const double (&x)[NS], DF D, const double (&Dt)[NS], double T,
                                          double (&z)[NS]) {
#pragma unroll
  for (int r = 0; r < NS; r++) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < NS; c++)
      if (c != r) s = s + lambda_off<NS>(x, r, c, D) * (Dt[r] - Dt[c]);
    z[r] = s / T;
  }
}

// Dbar2chi_iterative: hands each finished column j of chi to col(j, v).
template <int NS, class DF, class COL>
__device__ __forceinline__ void chi_iterative(const SpParams &sp, const double (&X)[NS], DF D, COL col) {
  double MW = 0.0, Y[NS], Mm[NS], Mi[NS], t1[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) MW = MW + X[i] * sp.molmass[i];
#pragma unroll
  for (int i = 0; i < NS; i++) Y[i] = sp.molmass[i] / MW * X[i];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    double t2 = 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++)
      if (j != i) t2 = t2 + X[j] / D(i, j);
    const double Di = (1.0 - Y[i]) / t2;
    Mm[i] = X[i] / Di;
    Mi[i] = Di / X[i];
  }
#pragma unroll
  for (int i = 0; i < NS; i++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NS; k++)
      if (k != i) s = s + X[i] * X[k] / D(i, k);
    t1[i] = s;
  }
  auto P = [&](int i, int j) { return (i == j) ? -Y[j] + 1.0 : -Y[j]; };
  // Z = -Delta + diag(Mm); J = diag(Mi) Z; PJ = P J
  double PJ[NS * NS];
#pragma unroll
  for (int i = 0; i < NS; i++)
#pragma unroll
    for (int j = 0; j < NS; j++) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < NS; k++) {
        const double Z = (k == j) ? -t1[k] + Mm[k] : -(-(X[k] * X[j] / D(k, j)));
        acc = acc + P(i, k) * (Mi[k] * Z);
      }
      PJ[i + NS * j] = acc;
    }
#pragma unroll
  for (int j = 0; j < NS; j++) {
    double m2[NS], c[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      double scr = 0.0;
#pragma unroll
      for (int k = 0; k < NS; k++) scr = scr + P(i, k) * Mi[k] * P(j, k);
      m2[i] = scr;
      c[i] = scr;
    }
    for (int it = 0; it < sp.chi_iter; it++) {
      double nc[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) {
        double scr = 0.0;
#pragma unroll
        for (int k = 0; k < NS; k++) scr = scr + PJ[i + NS * k] * c[k];
        nc[i] = scr + m2[i];
      }
#pragma unroll
      for (int i = 0; i < NS; i++) c[i] = nc[i];
    }
    col(j, c);
  }
}

// Gauss-Jordan inverse with partial pivoting (the result of LAPACK
// dgetrf + dgetri); row exchanges by static-index selects.
template <int NS>
__device__ __forceinline__ void invert(double (&A)[NS * NS]) {
  double B[NS * NS];
#pragma unroll
  for (int r = 0; r < NS; r++)
#pragma unroll
    for (int c = 0; c < NS; c++) B[r + NS * c] = (r == c) ? 1.0 : 0.0;
#pragma unroll
  for (int c = 0; c < NS; c++) {
    int p = c;
    double best = fabs(A[c + NS * c]);
#pragma unroll
    for (int r = c + 1; r < NS; r++)
      if (fabs(A[r + NS * c]) > best) {
        best = fabs(A[r + NS * c]);
        p = r;
      }
#pragma unroll
    for (int r = c + 1; r < NS; r++)
      if (r == p)
#pragma unroll
        for (int k = 0; k < NS; k++) {
          double t = A[c + NS * k];
          A[c + NS * k] = A[r + NS * k];
          A[r + NS * k] = t;
          t = B[c + NS * k];
          B[c + NS * k] = B[r + NS * k];
          B[r + NS * k] = t;
        }
    const double inv = 1.0 / A[c + NS * c];
#pragma unroll
    for (int k = 0; k < NS; k++) {
      A[c + NS * k] = A[c + NS * k] * inv;
      B[c + NS * k] = B[c + NS * k] * inv;
    }
#pragma unroll
    for (int r = 0; r < NS; r++) {
      if (r == c) continue;
      const double f = A[r + NS * c];
      if (f == 0.0) continue;
#pragma unroll
      for (int k = 0; k < NS; k++) {
        A[r + NS * k] = A[r + NS * k] - f * A[c + NS * k];
        B[r + NS * k] = B[r + NS * k] - f * B[c + NS * k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NS * NS; k++) A[k] = B[k];
}

// Pseudo-inverse V diag(1/S) U^T by one-sided Jacobi (Hestenes), singular
// values at or below tol * sum(S) dropped (mass_flux_utilities.f90:905-915).
template <int NS>
__device__ __forceinline__ void pinv(const double (&A0)[NS * NS], double tol, double (&out)[NS * NS]) {
  double U[NS * NS], V[NS * NS];
#pragma unroll
  for (int k = 0; k < NS * NS; k++) U[k] = A0[k];
#pragma unroll
  for (int r = 0; r < NS; r++)
#pragma unroll
    for (int c = 0; c < NS; c++) V[r + NS * c] = (r == c) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0;
#pragma unroll
    for (int p = 0; p < NS - 1; p++)
#pragma unroll
      for (int q = p + 1; q < NS; q++) {
        double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll
        for (int k = 0; k < NS; k++) {
          a = a + U[k + NS * p] * U[k + NS * p];
          b = b + U[k + NS * q] * U[k + NS * q];
          g = g + U[k + NS * p] * U[k + NS * q];
        }
        if (g == 0.0) continue;
        const double rel = fabs(g) / sqrt(a * b);
        if (rel > off) off = rel;
        if (rel < 1e-15) continue;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll
        for (int k = 0; k < NS; k++) {
          const double up = U[k + NS * p], uq = U[k + NS * q];
          U[k + NS * p] = cs * up - sn * uq;
          U[k + NS * q] = sn * up + cs * uq;
          const double vp = V[k + NS * p], vq = V[k + NS * q];
          V[k + NS * p] = cs * vp - sn * vq;
          V[k + NS * q] = sn * vp + cs * vq;
        }
      }
    if (off < 1e-15) break;
  }
  double S[NS], ssum = 0.0;
#pragma unroll
  for (int c = 0; c < NS; c++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NS; k++) s = s + U[k + NS * c] * U[k + NS * c];
    S[c] = sqrt(s);
    ssum = ssum + S[c];
  }
#pragma unroll
  for (int k = 0; k < NS * NS; k++) out[k] = 0.0;
#pragma unroll
  for (int c = 0; c < NS; c++) {
    if (!(S[c] > tol * ssum)) continue;
    const double w = 1.0 / (S[c] * S[c]);
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int j = 0; j < NS; j++) out[i + NS * j] = out[i + NS * j] + V[i + NS * c] * w * U[j + NS * c];
  }
}

// compute_chi_lapack: hands the columns of chi to col(j, v).
template <int NS, class DF, class COL>
__device__ __forceinline__ void chi_lapack(const SpParams &sp, const double (&x)[NS], const double (&W)[NS], DF D,
                                           COL col) {
  double Lam[NS * NS];
#pragma unroll
  for (int r = 0; r < NS; r++)
#pragma unroll
    for (int c = 0; c < NS; c++) Lam[r + NS * c] = (r == c) ? 0.0 : lambda_off<NS>(x, r, c, D);
#pragma unroll
  for (int r = 0; r < NS; r++) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < NS; c++)
      if (c != r) s = s - Lam[r + NS * c];
    Lam[r + NS * r] = s;
  }
  double alpha = 0.0;
#pragma unroll
  for (int r = 0; r < NS; r++) alpha = alpha + Lam[r + NS * r];
  double A[NS * NS];
#pragma unroll
  for (int r = 0; r < NS; r++)
#pragma unroll
    for (int c = 0; c < NS; c++) A[r + NS * c] = alpha * W[r] * W[c] + Lam[r + NS * c];
  double C[NS * NS];
  if (sp.inverse_type == 2) {
    pinv<NS>(A, sp.ftol, C);
  } else {
    invert<NS>(A);
#pragma unroll
    for (int k = 0; k < NS * NS; k++) C[k] = A[k];
  }
#pragma unroll
  for (int j = 0; j < NS; j++) {
    double v[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) v[i] = C[i + NS * j] - 1.0 / alpha;
    col(j, v);
  }
}

template <int NS, bool LP, class DF, class COL>
__device__ __forceinline__ void cell_chi(const SpParams &sp, const double (&rho)[NS], double rhotot,
                                         const double (&x)[NS], DF D, COL col) {
  if (LP) {
    double W[NS];
#pragma unroll
    for (int r = 0; r < NS; r++) W[r] = rho[r] / rhotot;
    chi_lapack<NS>(sp, x, W, D, col);
  } else {
    chi_iterative<NS>(sp, x, D, col);
  }
}

// choldc (use_lapack = F) or the dpotrf factor with the reference's clean-up
// (use_lapack = T), in place on a full (NS, NS) matrix in registers.
template <int NS, bool LP>
__device__ __forceinline__ void cholesky(double (&a)[NS * NS]) {
  if (!LP) {
    double p[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      bool sing = false;
#pragma unroll
      for (int j = i; j < NS; j++) {
        double s = a[i + NS * j];
#pragma unroll
        for (int k = i - 1; k >= 0; k--) s = s - a[i + NS * k] * a[j + NS * k];
        if (i == j) {
          if (s <= 0.0) {
            p[i] = 0.0;
            sing = true;
          } else {
            p[i] = sqrt(s);
          }
        } else {
          a[j + NS * i] = sing ? 0.0 : s / p[i];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
#pragma unroll
      for (int j = i + 1; j < NS; j++) a[i + NS * j] = 0.0;
      a[i + NS * i] = p[i];
    }
    return;
  }
  bool stop = false;
#pragma unroll
  for (int j = 0; j < NS; j++) {
    if (stop) break;
    double s = a[j + NS * j];
#pragma unroll
    for (int k = 0; k < j; k++) s = s - a[j + NS * k] * a[j + NS * k];
    if (!(s > 0.0)) {
      a[j + NS * j] = s;
      stop = true;
      continue;
    }
    const double d = sqrt(s);
    a[j + NS * j] = d;
    const double rinv = 1.0 / d;
#pragma unroll
    for (int i = j + 1; i < NS; i++) {
      double t = a[i + NS * j];
#pragma unroll
      for (int k = 0; k < j; k++) t = t - a[i + NS * k] * a[j + NS * k];
      a[i + NS * j] = t * rinv;
    }
  }
#pragma unroll
  for (int r = 0; r < NS; r++)
#pragma unroll
    for (int c = r + 1; c < NS; c++) a[r + NS * c] = 0.0;
  a[(NS - 1) + NS * (NS - 1)] = 0.0;
}

// ---------------- kernels: the reference routines one by one ----------------

template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_correct(Grid g, SpParams sp, double *rho, double *drho) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  double r[NS], d[NS];
  load<NS>(rho, N, t, r);
  cell_correct<NS>(sp, r, d);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    rho[(long)n * N + t] = r[n];
    drho[(long)n * N + t] = d[n];
  }
}

template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_prim(Grid g, SpParams sp, const double *rho, double *rhotot, double *xm,
                                                   double *mt) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  double r[NS], x[NS], tot, m;
  load<NS>(rho, N, t, r);
  cell_prim<NS>(sp, r, tot, x, m);
  rhotot[t] = tot;
  mt[t] = m;
#pragma unroll
  for (int n = 0; n < NS; n++) xm[(long)n * N + t] = x[n];
}

template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_mixture(Grid g, SpParams sp, double *D_bar, double *D_therm, double *H) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  const DParam D{sp};
#pragma unroll
  for (int r = 0; r < NS; r++) {
    D_therm[(long)r * N + t] = sp.Dtherm[r];
#pragma unroll
    for (int c = 0; c < NS; c++) {
      D_bar[(long)(r + NS * c) * N + t] = D(r, c);
      double h = 0.0;
      if (!sp.ideal) h = (r == c) ? sp.Hdiag[r] : sp.Hoff[r > c ? tri(r, c) : tri(c, r)];
      H[(long)(r + NS * c) * N + t] = h;
    }
  }
}

template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_gama(Grid g, SpParams sp, const double *xm, const double *H,
                                                   double *Gama) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  double x[NS];
  load<NS>(xm, N, t, x);
  auto Hf = [&](int k, int c) { return H[(long)(k + NS * c) * N + t]; };
#pragma unroll
  for (int r = 0; r < NS; r++)
#pragma unroll
    for (int c = 0; c < NS; c++) Gama[(long)(r + NS * c) * N + t] = gama_entry<NS>(sp, x, r, c, Hf);
}

template <int NS, bool LP>
__global__ void __launch_bounds__(STPB) k_sp_chi(Grid g, SpParams sp, const double *rho, const double *rhotot,
                                                  const double *xm, const double *Dbar, const double *Dtherm,
                                                  const double *Temp, double *chi, double *zeta) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  double r[NS], x[NS];
  load<NS>(rho, N, t, r);
  load<NS>(xm, N, t, x);
  const DField<NS> D{Dbar, N, t};
  if (sp.noniso) {
    double Dt[NS], z[NS];
    load<NS>(Dtherm, N, t, Dt);
    cell_zeta<NS>(x, D, Dt, Temp[t], z);
#pragma unroll
    for (int n = 0; n < NS; n++) zeta[(long)n * N + t] = z[n];
  }
  cell_chi<NS, LP>(sp, r, rhotot[t], x, D, [&](int j, const double (&v)[NS]) {
#pragma unroll
    for (int i = 0; i < NS; i++) chi[(long)(i + NS * j) * N + t] = v[i];
  });
}

template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_rhowchi(Grid g, const double *rho, const double *chi, double *rw) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  double r[NS];
  load<NS>(rho, N, t, r);
#pragma unroll
  for (int i = 0; i < NS; i++)
#pragma unroll
    for (int j = 0; j < NS; j++) {
      const long q = (long)(i + NS * j) * N + t;
      rw[q] = -(r[i] * chi[q]);
    }
}

template <int NS>
__device__ __forceinline__ double lonsager_entry(const SpParams &sp, double mt, double rt, const double (&W)[NS],
                                                 int r, int c, double chi_rc) {
  return mt * rt * W[r] * chi_rc * W[c] / sp.k_B;
}

template <int NS, bool LP>
__global__ void __launch_bounds__(STPB) k_sp_lonsager(Grid g, SpParams sp, const double *rho, const double *rhotot,
                                                       const double *molmtot, const double *chi, double *L) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  double r[NS], W[NS], a[NS * NS];
  load<NS>(rho, N, t, r);
  const double rt = rhotot[t], mt = molmtot[t];
#pragma unroll
  for (int n = 0; n < NS; n++) W[n] = r[n] / rt;
#pragma unroll
  for (int c = 0; c < NS; c++)
#pragma unroll
    for (int i = 0; i < NS; i++) a[i + NS * c] = lonsager_entry<NS>(sp, mt, rt, W, i, c, chi[(long)(i + NS * c) * N + t]);
  cholesky<NS, LP>(a);
#pragma unroll
  for (int k = 0; k < NS * NS; k++) L[(long)k * N + t] = a[k];
}

// ---------------- face fluxes ----------------

// diffusive_mass_flux on the face between cells t (above) and m (below):
// grad x, Gama_face grad x (identity when gama is null), + zeta_face grad T,
// + baro_face gradp_baro, rhoWchi_face (...), correction_flux.
template <int NS>
__device__ __forceinline__ void diff_face(const SpParams &sp, long N, double dxinv, long t, long m, const double *xm,
                                          const double *rw, const double *gama, const double *temp, const double *zeta,
                                          const double *baro, double gpb, double (&out)[NS]) {
  // opaque copies of the cell indices: keeps the compiler from hoisting the
  // 2*NS^2 per-lane addresses of this face out of a caller's loop (which
  // would hold them all in VGPRs); no instruction is emitted
  asm volatile("" : "+v"(t), "+v"(m));
  double x[NS], y[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) x[s] = (xm[(long)s * N + t] - xm[(long)s * N + m]) * dxinv;
  if (gama) {
    // column-outer order: each y[r] still sums c = 0, 1, ... in turn, and the
    // scheduling fence after each column keeps only one column's loads live
#pragma unroll
    for (int r = 0; r < NS; r++) y[r] = 0.0;
#pragma unroll
    for (int c = 0; c < NS; c++) {
#pragma unroll
      for (int r = 0; r < NS; r++) {
        const long q = (long)(r + NS * c) * N;
        y[r] = y[r] + 0.5 * (gama[q + t] + gama[q + m]) * x[c];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    // Gama_face = I: the sum 0 + 0*x_0 + ... + 1*x_r + ... is 0 + x_r
#pragma unroll
    for (int r = 0; r < NS; r++) y[r] = 0.0 + x[r];
  }
  if (zeta) {
    const double gT = (temp[t] - temp[m]) * dxinv;
#pragma unroll
    for (int s = 0; s < NS; s++) y[s] = y[s] + 0.5 * (zeta[(long)s * N + t] + zeta[(long)s * N + m]) * gT;
  }
  if (baro) {
#pragma unroll
    for (int s = 0; s < NS; s++) y[s] = y[s] + 0.5 * (baro[(long)s * N + t] + baro[(long)s * N + m]) * gpb;
  }
#pragma unroll
  for (int r = 0; r < NS; r++) out[r] = 0.0;
#pragma unroll
  for (int c = 0; c < NS; c++) {
#pragma unroll
    for (int r = 0; r < NS; r++) {
      const long q = (long)(r + NS * c) * N;
      out[r] = out[r] + 0.5 * (rw[q + t] + rw[q + m]) * y[c];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  if (sp.correct && NS > 1) {
    double sum = 0.0;
#pragma unroll
    for (int s = 0; s < NS - 1; s++) sum = sum + out[s];
    out[NS - 1] = -sum;
  }
}

// stochastic flux on the face of cell t in direction d: variance *
// L_face (sum_r w_r W_r), L_face the face average of the cell factors
// (upper triangle zero, so only c <= r enters: adding 0*w leaves the sum's
// bits unchanged), then correction_flux.  Packed (lpk) or full (Lf) factors.
template <int NS, bool PACKED>
__device__ __forceinline__ void stoch_face(const SpParams &sp, long N, long t, long m, int d, int dim,
                                           const double *L, const double *W, int n_rngs, const double *weights,
                                           double variance, double (&out)[NS]) {
  asm volatile("" : "+v"(t), "+v"(m));
  double w[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    double acc = 0.0;
    for (int r = 0; r < n_rngs; r++) acc = acc + weights[r] * W[((long)(r * dim + d) * NS + s) * N + t];
    w[s] = acc;
  }
#pragma unroll
  for (int r = 0; r < NS; r++) out[r] = 0.0;
#pragma unroll
  for (int c = 0; c < NS; c++) {
#pragma unroll
    for (int r = c; r < NS; r++) {
      const long q = (long)(PACKED ? lpk_index<NS>(r, c) : r + NS * c) * N;
      out[r] = out[r] + 0.5 * (L[q + t] + L[q + m]) * w[c];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int r = 0; r < NS; r++) out[r] = out[r] * variance;
  if (sp.correct && NS > 1) {
    double sum = 0.0;
#pragma unroll
    for (int s = 0; s < NS - 1; s++) sum = sum + out[s];
    out[NS - 1] = -sum;
  }
}

// flux_total: ft_zero set -> 0.0 + F (setval then add), else ft + F
template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_diff_flux(Grid g, SpParams sp, const double *xm, const double *rw,
                                                        const double *gama, const double *temp, const double *zeta,
                                                        const double *baro, Ptr3<const double> gpb, double *flux,
                                                        Ptr3<double> ft, int ft_zero) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  const Pt P = make_pt(g, i, j, k);
  const long N = g.N;
  const double dxinv = 1.0 / g.dx;
  for (int d = 0; d < g.dim; d++) {
    const long m = at_sh(g, P, d, -1, -1, 0);
    double out[NS];
    diff_face<NS>(sp, N, dxinv, t, m, xm, rw, gama, temp, zeta, baro, baro ? gpb.c[d][t] : 0.0, out);
#pragma unroll
    for (int s = 0; s < NS; s++) {
      flux[((long)d * NS + s) * N + t] = out[s];
      const long q = (long)s * N + t;
      ft.c[d][q] = (ft_zero ? 0.0 : ft.c[d][q]) + out[s];
    }
  }
}

template <int NS, bool PACKED>
__global__ void __launch_bounds__(STPB) k_sp_stoch_flux(Grid g, SpParams sp, const double *L, const double *W,
                                                         int n_rngs, const double *weights, double variance,
                                                         double *flux, Ptr3<double> ft) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  const Pt P = make_pt(g, i, j, k);
  const long N = g.N;
  for (int d = 0; d < g.dim; d++) {
    const long m = at_sh(g, P, d, -1, -1, 0);
    double out[NS];
    stoch_face<NS, PACKED>(sp, N, t, m, d, g.dim, L, W, n_rngs, weights, variance, out);
#pragma unroll
    for (int s = 0; s < NS; s++) {
      flux[((long)d * NS + s) * N + t] = out[s];
      ft.c[d][(long)s * N + t] = ft.c[d][(long)s * N + t] + out[s];
    }
  }
}

// compute_div of species face fluxes (div_and_grad.f90:304-455)
template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_div(Grid g, const double *flux, double *div, int increment) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  const Pt P = make_pt(g, i, j, k);
  const long N = g.N;
  const double dxinv = 1.0 / g.dx;
  const long ip = at3(g, P, 1, 0, 0), jp = at3(g, P, 0, 1, 0), kp = at3(g, P, 0, 0, 1);
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const double *fx = flux + (long)s * N, *fy = flux + (long)(NS + s) * N;
    double v = (fx[ip] - fx[t]) * dxinv;
    v = v + (fy[jp] - fy[t]) * dxinv;
    if (g.dim == 3) {
      const double *fz = flux + (long)(2 * NS + s) * N;
      v = v + (fz[kp] - fz[t]) * dxinv;
    }
    div[(long)s * N + t] = increment ? div[(long)s * N + t] + v : v;
  }
}

// compute_mass_fluxdiv's face pass: both fluxes of each lower face in one
// thread (the diffusive then the stochastic, fenced so their loads do not
// pile up), flux_total = (0 + F_diff) + F_stoch written once.
template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_chain_flux(Grid g, SpParams sp, SpFaceIn in, double *fd, double *fs,
                                                         Ptr3<double> ft) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  const Pt P = make_pt(g, i, j, k);
  const long N = g.N;
  const double dxinv = 1.0 / g.dx;
  for (int d = 0; d < g.dim; d++) {
    const long m = at_sh(g, P, d, -1, -1, 0);
    double od[NS], os[NS];
    diff_face<NS>(sp, N, dxinv, t, m, in.xm, in.rw, in.gama, in.temp, in.zeta, in.baro,
                  in.baro ? in.gpb[d][t] : 0.0, od);
    __builtin_amdgcn_sched_barrier(0);
    stoch_face<NS, true>(sp, N, t, m, d, g.dim, in.lpk, in.W, in.n_rngs, in.weights, in.variance, os);
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const long q = ((long)d * NS + s) * N + t;
      fd[q] = od[s];
      fs[q] = os[s];
      ft.c[d][(long)s * N + t] = (0.0 + od[s]) + os[s];
    }
  }
}

// compute_div of two species face-flux sets in one pass
template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_div2(Grid g, const double *fa, double *da, const double *fb,
                                                   double *db) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  const Pt P = make_pt(g, i, j, k);
  const long N = g.N;
  const double dxinv = 1.0 / g.dx;
  const long ip = at3(g, P, 1, 0, 0), jp = at3(g, P, 0, 1, 0), kp = at3(g, P, 0, 0, 1);
  for (int w = 0; w < 2; w++) {
    const double *flux = w ? fb : fa;
    double *div = w ? db : da;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const double *fx = flux + (long)s * N, *fy = flux + (long)(NS + s) * N;
      double v = (fx[ip] - fx[t]) * dxinv;
      v = v + (fy[jp] - fy[t]) * dxinv;
      if (g.dim == 3) {
        const double *fz = flux + (long)(2 * NS + s) * N;
        v = v + (fz[kp] - fz[t]) * dxinv;
      }
      div[(long)s * N + t] = v;
    }
  }
}

// ---------------- compute_mass_fluxdiv, fused ----------------

template <int NS, bool LP>
__global__ void __launch_bounds__(STPB) k_sp_chain_cell(Grid g, SpParams sp, const double *temp, SpCellOut o) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  double r[NS], dr[NS], x[NS], rt, mt;
  load<NS>(o.rho, N, t, r);
  cell_correct<NS>(sp, r, dr);
  cell_prim<NS>(sp, r, rt, x, mt);
#pragma unroll
  for (int n = 0; n < NS; n++) o.xm[(long)n * N + t] = x[n];
  const DParam D{sp};
  if (o.gama) {
    auto Hf = [&](int kk, int c) {
      return (kk == c) ? sp.Hdiag[kk] : sp.Hoff[kk > c ? tri(kk, c) : tri(c, kk)];
    };
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int c = 0; c < NS; c++) o.gama[(long)(i + NS * c) * N + t] = gama_entry<NS>(sp, x, i, c, Hf);
  }
  if (o.zeta) {
    double Dt[NS], z[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) Dt[n] = sp.Dtherm[n];
    cell_zeta<NS>(x, D, Dt, temp[t], z);
#pragma unroll
    for (int n = 0; n < NS; n++) o.zeta[(long)n * N + t] = z[n];
  }
  if (o.baro) {
    double nn = 0.0;
#pragma unroll
    for (int c = 0; c < NS; c++) nn = nn + r[c] / sp.molmass[c];
#pragma unroll
    for (int c = 0; c < NS; c++) o.baro[(long)c * N + t] = r[c] / sp.rhobar[c] / (nn * sp.k_B * temp[t]);
  }
  double W[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) W[n] = r[n] / rt;
  const bool want_l = o.lpk != nullptr;
  double a[NS * NS];
  cell_chi<NS, LP>(sp, r, rt, x, D, [&](int jc, const double (&v)[NS]) {
#pragma unroll
    for (int i = 0; i < NS; i++) {
      o.rw[(long)(i + NS * jc) * N + t] = -(r[i] * v[i]);
      a[i + NS * jc] = lonsager_entry<NS>(sp, mt, rt, W, i, jc, v[i]);
    }
  });
  if (want_l) {
    cholesky<NS, LP>(a);
#pragma unroll
    for (int c = 0; c < NS; c++)
#pragma unroll
      for (int i = c; i < NS; i++) o.lpk[(long)lpk_index<NS>(i, c) * N + t] = a[i + NS * c];
  }
#pragma unroll
  for (int n = 0; n < NS; n++) o.rho[(long)n * N + t] = r[n] + -1.0 * dr[n];
}

template <typename T>
Ptr3<const T> cp3(const T *const a[3]) {
  Ptr3<const T> r;
  for (int d = 0; d < 3; d++) r.c[d] = a ? a[d] : nullptr;
  return r;
}

// Dispatch on the species count.
template <template <int> class F, class... A>
int dispatch(int ns, A &&...args) {
  switch (ns) {
    case 1: return F<1>::run(args...);
    case 2: return F<2>::run(args...);
    case 3: return F<3>::run(args...);
    case 4: return F<4>::run(args...);
    case 5: return F<5>::run(args...);
    case 6: return F<6>::run(args...);
    case 7: return F<7>::run(args...);
    case 8: return F<8>::run(args...);
    default: return LMM_ERR_UNSUPPORTED;
  }
}

// NS = 1 has no chi (Dbar2chi divides by an empty sum): those routines
// reject it; the face fluxes take it.
#define SP_NEEDS2 \
  if (NS < 2) return LMM_ERR_UNSUPPORTED

template <int NS>
struct RunCorrect {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, double *rho, double *drho) {
    SP_NEEDS2;
    SLAUNCH(k_sp_correct<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, rho, drho);
    return 0;
  }
};
template <int NS>
struct RunPrim {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, double *rt, double *xm,
                 double *mt) {
    SP_NEEDS2;
    SLAUNCH(k_sp_prim<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, rho, rt, xm, mt);
    return 0;
  }
};
template <int NS>
struct RunMixture {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, double *Db, double *Dt, double *H) {
    SP_NEEDS2;
    SLAUNCH(k_sp_mixture<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, Db, Dt, H);
    return 0;
  }
};
template <int NS>
struct RunGama {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const double *xm, const double *H, double *G) {
    SP_NEEDS2;
    SLAUNCH(k_sp_gama<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, xm, H, G);
    return 0;
  }
};
template <int NS>
struct RunChi {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *rt,
                 const double *xm, const double *Db, const double *Dt, const double *T, double *chi, double *z) {
    SP_NEEDS2;
    if (sp.lapack)
      SLAUNCH((k_sp_chi<NS, true>), dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, rho, rt, xm, Db, Dt, T, chi, z);
    else
      SLAUNCH((k_sp_chi<NS, false>), dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, rho, rt, xm, Db, Dt, T, chi, z);
    return 0;
  }
};
template <int NS>
struct RunRhoWchi {
  static int run(hipStream_t s, const Grid &g, const double *rho, const double *chi, double *rw) {
    SP_NEEDS2;
    SLAUNCH(k_sp_rhowchi<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, rho, chi, rw);
    return 0;
  }
};
template <int NS>
struct RunLonsager {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *rt,
                 const double *mt, const double *chi, double *L) {
    SP_NEEDS2;
    if (sp.lapack)
      SLAUNCH((k_sp_lonsager<NS, true>), dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, rho, rt, mt, chi, L);
    else
      SLAUNCH((k_sp_lonsager<NS, false>), dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, rho, rt, mt, chi, L);
    return 0;
  }
};
template <int NS>
struct RunDiffusive {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const double *xm, const double *rw,
                 const double *G, const double *T, const double *z, const double *baro, const double *const gpb[3],
                 double *flux, double *const ft[3], double *div) {
    SLAUNCH(k_sp_diff_flux<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, xm, rw, G, T, z, baro, cp3(gpb), flux,
            p3(ft), 0);
    SLAUNCH(k_sp_div<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, flux, div, 0);
    return 0;
  }
};
template <int NS>
struct RunBaro {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *T, double *baro);
};
template <int NS>
struct RunStochastic {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const double *L, const double *W, int nr,
                 const double *w, double var, double *flux, double *const ft[3], double *div, int inc) {
    SLAUNCH((k_sp_stoch_flux<NS, false>), dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, L, W, nr, w, var, flux, p3(ft));
    SLAUNCH(k_sp_div<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, flux, div, inc);
    return 0;
  }
};
template <int NS>
struct RunChainCell {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const double *T, const SpCellOut &o) {
    SP_NEEDS2;
    if (sp.lapack)
      SLAUNCH((k_sp_chain_cell<NS, true>), dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, T, o);
    else
      SLAUNCH((k_sp_chain_cell<NS, false>), dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, T, o);
    return 0;
  }
};
template <int NS>
struct RunChainFace {
  static int run(hipStream_t s, const Grid &g, const SpParams &sp, const SpFaceIn &in, double *flux, double *dd,
                 double *sd, double *const ft[3]) {
    SP_NEEDS2;
    // up to 4 species both fluxes share one face pass (at 5 and more its
    // registers drop it to 1-3 waves per SIMD: separate passes there)
    if (in.W && NS <= 4) {  // flux holds 2 * dim * ns * N: the diffusive then the stochastic face fluxes
      double *fs = flux + (long)g.dim * NS * g.N;
      SLAUNCH(k_sp_chain_flux<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, in, flux, fs, p3(ft));
      SLAUNCH(k_sp_div2<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, flux, dd, fs, sd);
      return 0;
    }
    SLAUNCH(k_sp_diff_flux<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, in.xm, in.rw, in.gama, in.temp, in.zeta,
            in.baro, cp3(in.baro ? in.gpb : nullptr), flux, p3(ft), 1);
    SLAUNCH(k_sp_div<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, flux, dd, 0);
    if (in.W) {
      SLAUNCH((k_sp_stoch_flux<NS, true>), dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, in.lpk, in.W, in.n_rngs,
              in.weights, in.variance, flux, p3(ft));
      SLAUNCH(k_sp_div<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, flux, sd, 0);
    }
    return 0;
  }
};

// compute_baro_coef on cells (the diffusive entry computes it from rho)
template <int NS>
__global__ void __launch_bounds__(STPB) k_sp_baro(Grid g, SpParams sp, const double *rho, const double *temp,
                                                   double *baro) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  double r[NS];
  load<NS>(rho, N, t, r);
  double nn = 0.0;
#pragma unroll
  for (int c = 0; c < NS; c++) nn = nn + r[c] / sp.molmass[c];
#pragma unroll
  for (int c = 0; c < NS; c++) baro[(long)c * N + t] = r[c] / sp.rhobar[c] / (nn * sp.k_B * temp[t]);
}

template <int NS>
int RunBaro<NS>::run(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *T,
                     double *baro) {
  SLAUNCH(k_sp_baro<NS>, dim3(sblk(g.N)), dim3(STPB), 0, s, g, sp, rho, T, baro);
  return 0;
}

}  // namespace

int sp_correct_rho(hipStream_t s, const Grid &g, const SpParams &sp, double *rho, double *drho) {
  return dispatch<RunCorrect>(sp.ns, s, g, sp, rho, drho);
}
int sp_cons_to_prim(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, double *rhotot,
                    double *molarconc, double *molmtot) {
  return dispatch<RunPrim>(sp.ns, s, g, sp, rho, rhotot, molarconc, molmtot);
}
int sp_mixture(hipStream_t s, const Grid &g, const SpParams &sp, double *D_bar, double *D_therm, double *H) {
  return dispatch<RunMixture>(sp.ns, s, g, sp, D_bar, D_therm, H);
}
int sp_gama(hipStream_t s, const Grid &g, const SpParams &sp, const double *molarconc, const double *H,
            double *Gama) {
  return dispatch<RunGama>(sp.ns, s, g, sp, molarconc, H, Gama);
}
int sp_chi(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *rhotot,
           const double *molarconc, const double *D_bar, const double *D_therm, const double *Temp, double *chi,
           double *zeta) {
  return dispatch<RunChi>(sp.ns, s, g, sp, rho, rhotot, molarconc, D_bar, D_therm, Temp, chi, zeta);
}
int sp_rhowchi(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *chi,
               double *rhoWchi) {
  return dispatch<RunRhoWchi>(sp.ns, s, g, rho, chi, rhoWchi);
}
int sp_lonsager(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *rhotot,
                const double *molmtot, const double *chi, double *L) {
  return dispatch<RunLonsager>(sp.ns, s, g, sp, rho, rhotot, molmtot, chi, L);
}
int sp_diffusive(hipStream_t s, const Grid &g, const SpParams &sp, const double *rho, const double *molarconc,
                 const double *rhoWchi, const double *Gama, const double *Temp, const double *zeta,
                 const double *const gpb[3], double *baro, double *flux, double *const flux_total[3], double *div) {
  if (gpb) {
    if (int rc = dispatch<RunBaro>(sp.ns, s, g, sp, rho, Temp, baro)) return rc;
  } else {
    baro = nullptr;
  }
  return dispatch<RunDiffusive>(sp.ns, s, g, sp, molarconc, rhoWchi, Gama, Temp, zeta, baro, gpb, flux, flux_total,
                                div);
}
int sp_stochastic(hipStream_t s, const Grid &g, const SpParams &sp, const double *L, const double *W, int n_rngs,
                  const double *weights, double variance, double *flux, double *const flux_total[3], double *div,
                  int increment) {
  return dispatch<RunStochastic>(sp.ns, s, g, sp, L, W, n_rngs, weights, variance, flux, flux_total, div, increment);
}
int sp_chain_cell(hipStream_t s, const Grid &g, const SpParams &sp, const double *Temp, const SpCellOut &o) {
  return dispatch<RunChainCell>(sp.ns, s, g, sp, Temp, o);
}
int sp_chain_face(hipStream_t s, const Grid &g, const SpParams &sp, const SpFaceIn &in, double *flux,
                  double *diff_div, double *stoch_div, double *const flux_total[3]) {
  return dispatch<RunChainFace>(sp.ns, s, g, sp, in, flux, diff_div, stoch_div, flux_total);
}

}  // namespace lmm
