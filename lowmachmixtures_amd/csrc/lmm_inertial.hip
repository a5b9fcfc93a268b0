// gfx950 kernels of the inertial time step (advance_timestep_inertial.f90):
// centred scalar and momentum advection, gravity, and the element-wise
// assembly of the right-hand sides, each fused into one pass over the
// periodic box where the reference makes several multifab calls in a row.
// Every fused pass performs the reference's operations on each value in the
// reference's order (-ffp-contract=off), so the FP64 results equal the
// reference's sequence of multifab operations bit for bit:
//   mk_advective_s_fluxdiv   mk_advective_s_fluxdiv.f90:68-166
//   mk_advective_m_fluxdiv   mk_advective_m_fluxdiv.f90:87-329
//   mk_grav_force            mk_grav_force.f90:96-159
//   compute_rhotot + convert_rhoc_to_c  mass_flux_utilities.f90:376-420,
//                            convert_rhoc_to_c.f90:21-58
//   the rho updates, gmres_rhs_v and gmres_rhs_p assembly of
//                            advance_timestep_inertial.f90:169-257,
//                            :264-496, :585-606, :641-957
#include "lmm_kernels.h"
#include "lmm_points.h"

namespace lmm {

constexpr int ITPB = 256;
static inline int iblk(long n) { return (int)((n + ITPB - 1) / ITPB); }

#define ILAUNCH(K, G, S, ...)                                   \
  do {                                                          \
    hipLaunchKernelGGL(K, dim3(G), dim3(ITPB), 0, S, __VA_ARGS__); \
    post_launch(S, #K);                                         \
  } while (0)

namespace {

// div(u s) of component c at cell (i,j,k): ((fxh - fxl)*dxinv + (fyh -
// fyl)*dxinv) [+ (fzh - fzl)*dxinv], fluxes umac*s_fc on the cell's faces
__device__ __forceinline__ double adv_div(const Grid &g, int i, int j, int k, long p, const AdvArgs &a, long off) {
  const long sy = g.s[1], sz = g.s[2];
  const long xp = wp(i, g.n[0]) + sy * j + sz * k;
  const long yp = i + sy * wp(j, g.n[1]) + sz * k;
  const double fxl = a.u[0][p] * a.s[0][off + p], fxh = a.u[0][xp] * a.s[0][off + xp];
  const double fyl = a.u[1][p] * a.s[1][off + p], fyh = a.u[1][yp] * a.s[1][off + yp];
  double div = (fxh - fxl) * a.dxinv + (fyh - fyl) * a.dxinv;
  if (g.dim == 3) {
    const long zp = i + sy * j + sz * wp(k, g.n[2]);
    const double fzl = a.u[2][p] * a.s[2][off + p], fzh = a.u[2][zp] * a.s[2][off + zp];
    div = div + (fzh - fzl) * a.dxinv;
  }
  return div;
}

// mode 0: upd -= div (mk_advective_s_fluxdiv)
// mode 1: rho_new = rho_old + (((0 + diff [+ stoch]) - div) * dt)   (step 2)
// mode 2: rho_new = ((rho_new + rho_old) * 0.5) + (((0 + diff [+ stoch]) - div) * (dt/2))   (step 5)
__global__ void __launch_bounds__(ITPB) k_adv_s(Grid g, int ncomp, AdvArgs a, RhoArgs r, int mode) {
  const long N = g.N;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  for (int c = 0; c < ncomp; c++) {
    const long off = (long)c * N;
    const double div = adv_div(g, i, j, k, t, a, off);
    if (mode == 0) {
      r.upd[off + t] = r.upd[off + t] - div;
      continue;
    }
    double ru = 0.0 + r.diff[off + t];
    if (r.stoch) ru = ru + r.stoch[off + t];
    ru = ru - div;
    if (mode == 1) {
      ru = ru * r.dt;
      r.rho_new[off + t] = r.rho_old[off + t] + ru;
    } else {
      double rn = r.rho_new[off + t] + r.rho_old[off + t];
      rn = rn * 0.5;
      ru = ru * r.dt;  // dt/2 from the host
      r.rho_new[off + t] = rn + ru;
    }
  }
}

// rhotot = sum over species from 0; rho = (rho / rhotot) * rhotot
__global__ void __launch_bounds__(ITPB) k_rhotot_conc(long N, int ns, double *rho, double *rhotot) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  double s = 0.0;
  for (int c = 0; c < ns; c++) s = s + rho[(long)c * N + t];
  rhotot[t] = s;
  for (int c = 0; c < ns; c++) {
    const long q = (long)c * N + t;
    rho[q] = (rho[q] / s) * s;
  }
}

// 0.25d0*(a+b)*(c+d)
__device__ __forceinline__ double q4(double a, double b, double c, double d) { return 0.25 * (a + b) * (c + d); }

// m_update[a] += -div(m_a u) at the face (a; i,j,k)
__global__ void __launch_bounds__(ITPB) k_adv_m(Grid g, Ptr3<double> u, Ptr3<double> m, Ptr3<double> upd,
                                                double dxinv) {
  const long N = g.N;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.dim * N) return;
  const int a = (int)(t / N);
  const long p = t - (long)a * N;
  int i, j, k;
  decode(g, p, i, j, k);
  const long sy = g.s[1], sz = g.s[2];
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  auto at = [&](int ii, int jj, int kk) { return (long)ii + sy * jj + sz * kk; };
  const int im = wm(i, nx), ip = wp(i, nx), jm = wm(j, ny), jp = wp(j, ny);
  const int km = g.dim == 3 ? wm(k, nz) : 0, kp = g.dim == 3 ? wp(k, nz) : 0;
  const double *M = m.c[a], *U = u.c[0], *V = u.c[1], *W = u.c[2];
  const double m0 = M[p];
  double fd;
  if (a == 0) {
    const double xl = q4(M[at(im, j, k)], m0, U[at(im, j, k)], U[p]);
    const double xh = q4(m0, M[at(ip, j, k)], U[p], U[at(ip, j, k)]);
    const double yl = q4(M[at(i, jm, k)], m0, V[at(im, j, k)], V[p]);
    const double yh = q4(m0, M[at(i, jp, k)], V[at(im, jp, k)], V[at(i, jp, k)]);
    double s = (xh - xl) * dxinv + (yh - yl) * dxinv;
    if (g.dim == 3) {
      const double zl = q4(M[at(i, j, km)], m0, W[at(im, j, k)], W[p]);
      const double zh = q4(m0, M[at(i, j, kp)], W[at(im, j, kp)], W[at(i, j, kp)]);
      s = s + (zh - zl) * dxinv;
    }
    fd = -s;
  } else if (a == 1) {
    const double xl = q4(M[at(im, j, k)], m0, U[at(i, jm, k)], U[p]);
    const double xh = q4(m0, M[at(ip, j, k)], U[at(ip, jm, k)], U[at(ip, j, k)]);
    const double yl = q4(M[at(i, jm, k)], m0, V[at(i, jm, k)], V[p]);
    const double yh = q4(m0, M[at(i, jp, k)], V[p], V[at(i, jp, k)]);
    double s = (xh - xl) * dxinv + (yh - yl) * dxinv;
    if (g.dim == 3) {
      const double zl = q4(M[at(i, j, km)], m0, W[at(i, jm, k)], W[p]);
      const double zh = q4(m0, M[at(i, j, kp)], W[at(i, jm, kp)], W[at(i, j, kp)]);
      s = s + (zh - zl) * dxinv;
    }
    fd = -s;
  } else {
    const double xl = q4(M[at(im, j, k)], m0, U[at(i, j, km)], U[p]);
    const double xh = q4(m0, M[at(ip, j, k)], U[at(ip, j, km)], U[at(ip, j, k)]);
    const double yl = q4(M[at(i, jm, k)], m0, V[at(i, j, km)], V[p]);
    const double yh = q4(m0, M[at(i, jp, k)], V[at(i, jp, km)], V[at(i, jp, k)]);
    const double zl = q4(M[at(i, j, km)], m0, W[at(i, j, km)], W[p]);
    const double zh = q4(m0, M[at(i, j, kp)], W[p], W[at(i, j, kp)]);
    fd = -((xh - xl) * dxinv + (yh - yl) * dxinv + (zh - zl) * dxinv);
  }
  upd.c[a][p] = upd.c[a][p] + fd;
}

// m_force[d] += 0.5*grav(d)*(s_old + s_new)
__global__ void __launch_bounds__(ITPB) k_grav(long N, int dim, Ptr3<double> f, Ptr3<double> so, Ptr3<double> sn,
                                               double g0, double g1, double g2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)dim * N) return;
  const int d = (int)(t / N);
  const long p = t - (long)d * N;
  const double gd = d == 0 ? g0 : (d == 1 ? g1 : g2);
  f.c[d][p] = f.c[d][p] + 0.5 * gd * (so.c[d][p] + sn.c[d][p]);
}

// out = a * b on every face component
__global__ void __launch_bounds__(ITPB) k_mul_faces(long N, int dim, Ptr3<double> out, Ptr3<double> a,
                                                    Ptr3<double> b) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)dim * N) return;
  const int d = (int)(t / N);
  const long p = t - (long)d * N;
  out.c[d][p] = a.c[d][p] * b.c[d][p];
}

// gmres_rhs_v = ((((mold*(1/dt) - gradpi) + m_a) + m_d) + m_s) [+ 0.5 g (rfo + rfn)],
// halving m_a / m_d / m_s in place first where the reference does
__global__ void __launch_bounds__(ITPB) k_rhs_v(long N, int dim, RhsVArgs r) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)dim * N) return;
  const int d = (int)(t / N);
  const long p = t - (long)d * N;
  double ma = r.m_a[d][p], md = r.m_d[d][p], ms = r.m_s[d][p];
  if (r.half_ma) {
    ma = ma * 0.5;
    r.m_a[d][p] = ma;
  }
  if (r.half_md) {
    md = md * 0.5;
    r.m_d[d][p] = md;
  }
  if (r.half_ms) {
    ms = ms * 0.5;
    r.m_s[d][p] = ms;
  }
  double v = r.mold[d][p] * r.inv_dt;
  v = v - r.gradpi[d][p];
  v = v + ma;
  v = v + md;
  v = v + ms;
  if (r.grav) v = v + 0.5 * r.g[d] * (r.rfo[d][p] + r.rfn[d][p]);
  r.v[d][p] = v;
}

// v = v - ((umac * rf) * (1/dt))
__global__ void __launch_bounds__(ITPB) k_v_sub_mt(long N, int dim, Ptr3<double> v, Ptr3<double> u,
                                                   Ptr3<double> rf, double inv_dt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)dim * N) return;
  const int d = (int)(t / N);
  const long p = t - (long)d * N;
  double mt = u.c[d][p] * rf.c[d][p];
  mt = mt * inv_dt;
  v.c[d][p] = v.c[d][p] - mt;
}

// v = v + (w * 0.5)
__global__ void __launch_bounds__(ITPB) k_v_add_half(long N, int dim, Ptr3<double> v, Ptr3<double> w) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)dim * N) return;
  const int d = (int)(t / N);
  const long p = t - (long)d * N;
  v.c[d][p] = v.c[d][p] + w.c[d][p] * 0.5;
}

// v = 0 - w on every face (a zeroed right-hand side minus grad pi)
__global__ void __launch_bounds__(ITPB) k_zero_minus(long N, int dim, Ptr3<double> v, Ptr3<double> w) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)dim * N) return;
  const int d = (int)(t / N);
  const long p = t - (long)d * N;
  v.c[d][p] = 0.0 - w.c[d][p];
}

// a = a + b
__global__ void __launch_bounds__(ITPB) k_add(long n, double *a, const double *__restrict__ b) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = a[t] + b[t];
}

// diff = diff*(-1), stoch = stoch*(-1) in place; gmres_rhs_p = (0 + sum_c
// (a_c diff_c [+ a_c stoch_c])) + divu with a_c = -1/rhobar_c
__global__ void __launch_bounds__(ITPB) k_rhs_p(long N, int ns, double *bp, double *diff, double *stoch,
                                                const double *__restrict__ divu, RhobarInv a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  double s = 0.0;
  for (int c = 0; c < ns; c++) {
    const long q = (long)c * N + t;
    const double dq = diff[q] * -1.0;
    diff[q] = dq;
    s = s + a.v[c] * dq;
    if (stoch) {
      const double sq = stoch[q] * -1.0;
      stoch[q] = sq;
      s = s + a.v[c] * sq;
    }
  }
  bp[t] = s + divu[t];
}

// compute_eta_2d/3d: eta = visc_coef, or visc_coef*(-5 + 6 rhotot); kappa = 1
__global__ void __launch_bounds__(ITPB) k_eta(long N, double *eta, double *kappa, const double *__restrict__ rhotot,
                                              double visc_coef, int model) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  eta[t] = model == 0 ? visc_coef : visc_coef * (-5.0 + 6.0 * rhotot[t]);
  kappa[t] = 1.0;
}

}  // namespace

void launch_adv_s(hipStream_t s, const Grid &g, int ncomp, const AdvArgs &a, const RhoArgs &r, int mode) {
  if (g.N > 0 && ncomp > 0) ILAUNCH(k_adv_s, iblk(g.N), s, g, ncomp, a, r, mode);
}
void launch_rhotot_conc(hipStream_t s, const Grid &g, int ns, double *rho, double *rhotot) {
  ILAUNCH(k_rhotot_conc, iblk(g.N), s, g.N, ns, rho, rhotot);
}
void launch_adv_m(hipStream_t s, const Grid &g, double *const u[3], double *const m[3], double *const upd[3]) {
  Ptr3<double> U, M, P;
  for (int d = 0; d < 3; d++) {
    U.c[d] = d < g.dim ? u[d] : nullptr;
    M.c[d] = d < g.dim ? m[d] : nullptr;
    P.c[d] = d < g.dim ? upd[d] : nullptr;
  }
  ILAUNCH(k_adv_m, iblk((long)g.dim * g.N), s, g, U, M, P, 1.0 / g.dx);
}
void launch_grav(hipStream_t s, const Grid &g, const double grav[3], double *const f[3], double *const so[3],
                 double *const sn[3]) {
  Ptr3<double> F, O, W;
  for (int d = 0; d < 3; d++) {
    F.c[d] = d < g.dim ? f[d] : nullptr;
    O.c[d] = d < g.dim ? so[d] : nullptr;
    W.c[d] = d < g.dim ? sn[d] : nullptr;
  }
  ILAUNCH(k_grav, iblk((long)g.dim * g.N), s, g.N, g.dim, F, O, W, grav[0], grav[1], grav[2]);
}
void launch_mul_faces(hipStream_t s, const Grid &g, double *const out[3], double *const a[3], double *const b[3]) {
  Ptr3<double> X, A, B;
  for (int d = 0; d < 3; d++) {
    X.c[d] = d < g.dim ? out[d] : nullptr;
    A.c[d] = d < g.dim ? a[d] : nullptr;
    B.c[d] = d < g.dim ? b[d] : nullptr;
  }
  ILAUNCH(k_mul_faces, iblk((long)g.dim * g.N), s, g.N, g.dim, X, A, B);
}
void launch_rhs_v(hipStream_t s, const Grid &g, const RhsVArgs &r) {
  ILAUNCH(k_rhs_v, iblk((long)g.dim * g.N), s, g.N, g.dim, r);
}
void launch_v_sub_mt(hipStream_t s, const Grid &g, double *const v[3], double *const u[3], double *const rf[3],
                     double inv_dt) {
  Ptr3<double> V, U, R;
  for (int d = 0; d < 3; d++) {
    V.c[d] = d < g.dim ? v[d] : nullptr;
    U.c[d] = d < g.dim ? u[d] : nullptr;
    R.c[d] = d < g.dim ? rf[d] : nullptr;
  }
  ILAUNCH(k_v_sub_mt, iblk((long)g.dim * g.N), s, g.N, g.dim, V, U, R, inv_dt);
}
void launch_v_add_half(hipStream_t s, const Grid &g, double *const v[3], double *const w[3]) {
  Ptr3<double> V, W;
  for (int d = 0; d < 3; d++) {
    V.c[d] = d < g.dim ? v[d] : nullptr;
    W.c[d] = d < g.dim ? w[d] : nullptr;
  }
  ILAUNCH(k_v_add_half, iblk((long)g.dim * g.N), s, g.N, g.dim, V, W);
}
void launch_zero_minus(hipStream_t s, const Grid &g, double *const v[3], double *const w[3]) {
  Ptr3<double> V, W;
  for (int d = 0; d < 3; d++) {
    V.c[d] = d < g.dim ? v[d] : nullptr;
    W.c[d] = d < g.dim ? w[d] : nullptr;
  }
  ILAUNCH(k_zero_minus, iblk((long)g.dim * g.N), s, g.N, g.dim, V, W);
}
void launch_add(hipStream_t s, double *a, const double *b, long n) {
  if (n > 0) ILAUNCH(k_add, iblk(n), s, n, a, b);
}
void launch_rhs_p(hipStream_t s, const Grid &g, int ns, double *bp, double *diff, double *stoch, const double *divu,
                  const RhobarInv &a) {
  ILAUNCH(k_rhs_p, iblk(g.N), s, g.N, ns, bp, diff, stoch, divu, a);
}
void launch_eta(hipStream_t s, const Grid &g, double *eta, double *kappa, const double *rhotot, double visc_coef,
                int model) {
  ILAUNCH(k_eta, iblk(g.N), s, g.N, eta, kappa, rhotot, visc_coef, model);
}

}  // namespace lmm
