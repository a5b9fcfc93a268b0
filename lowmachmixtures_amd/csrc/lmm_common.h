// Shared device/host definitions of the MI355X Stokes library.
//
// Storage: one periodic box, x fastest.  A face component stores the lower
// face of every cell; the periodic hi+1 face is the same unknown, so every
// array of a level holds exactly nx*ny*nz values.  Neighbour indices wrap
// periodically, which replaces BoxLib's ghost cells and fill_boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

namespace lmm {

struct Grid {
  int dim;
  int n[3];      // cells per direction (n[2] = 1 in 2D)
  long s[3];     // strides: 1, nx, nx*ny
  long N;        // nx*ny*nz
  double dx;
};

inline Grid make_grid(int dim, const int n[3], double dx) {
  Grid g;
  g.dim = dim;
  g.n[0] = n[0];
  g.n[1] = n[1];
  g.n[2] = (dim == 3) ? n[2] : 1;
  g.s[0] = 1;
  g.s[1] = g.n[0];
  g.s[2] = (long)g.n[0] * g.n[1];
  g.N = g.s[2] * g.n[2];
  g.dx = dx;
  return g;
}

inline Grid coarsen(const Grid &f) {
  int n[3] = {f.n[0] / 2, f.n[1] / 2, f.dim == 3 ? f.n[2] / 2 : 1};
  return make_grid(f.dim, n, f.dx * 2.0);
}

// The division of a Gauss-Seidel / Jacobi update, x / diag.  FP64 keeps the
// correctly rounded quotient (the reference's rounding, bit-exact parity).
// FP32 is only the multigrid preconditioner inside the FP64 flexible GMRES,
// which has no reference rounding to follow: there the hardware reciprocal
// (v_rcp_f32, 1 ulp) and a multiply replace the ten-instruction IEEE
// division sequence.
template <typename T>
__device__ __forceinline__ T gs_div(T x, T diag) {
  return x / diag;
}
template <>
__device__ __forceinline__ float gs_div<float>(float x, float diag) {
  return x * __builtin_amdgcn_rcpf(diag);
}

__device__ __forceinline__ int wp(int i, int n) { return (i == n - 1) ? 0 : i + 1; }
__device__ __forceinline__ int wm(int i, int n) { return (i == 0) ? n - 1 : i - 1; }

// Edge (3D) / node (2D) array shared by face directions a and b.
__host__ __device__ __forceinline__ int edge_of(int a, int b) {
  int lo = a < b ? a : b, hi = a < b ? b : a;
  return (lo == 0 && hi == 1) ? 0 : (lo == 0 ? 1 : 2);
}

// Coefficients of the staggered viscous operator on one level.
template <typename T>
struct StagCoef {
  const T *alpha[3];  // theta_alpha * alpha on faces
  const T *beta;      // cells
  const T *gamma;     // cells
  const T *bed[3];    // 2D: bed[0] = nodes; 3D: xy, xz, yz edges
};

// Point coordinates and wrapped neighbours.
struct Pt {
  int c[3], cp[3], cm[3];
};

__device__ __forceinline__ Pt make_pt(const Grid &g, int i, int j, int k) {
  Pt p;
  p.c[0] = i;
  p.c[1] = j;
  p.c[2] = k;
  for (int d = 0; d < 3; d++) {
    if (d < g.dim) {
      p.cp[d] = wp(p.c[d], g.n[d]);
      p.cm[d] = wm(p.c[d], g.n[d]);
    } else {
      p.cp[d] = p.cm[d] = 0;
    }
  }
  return p;
}

// index of point p with selected directions shifted: sh[d] in {-1,0,1}
__device__ __forceinline__ long at3(const Grid &g, const Pt &p, int sx, int sy, int sz) {
  int i = sx > 0 ? p.cp[0] : (sx < 0 ? p.cm[0] : p.c[0]);
  int j = sy > 0 ? p.cp[1] : (sy < 0 ? p.cm[1] : p.c[1]);
  int k = sz > 0 ? p.cp[2] : (sz < 0 ? p.cm[2] : p.c[2]);
  return i + g.s[1] * j + g.s[2] * k;
}

__device__ __forceinline__ long at_sh(const Grid &g, const Pt &p, int d1, int s1, int d2, int s2) {
  int sh[3] = {0, 0, 0};
  if (d1 >= 0) sh[d1] += s1;
  if (d2 >= 0) sh[d2] += s2;
  return at3(g, p, sh[0], sh[1], sh[2]);
}

// Staggered operator at one face point of component a, written term by term
// in the order of stag_applyop_2d/3d (stag_applyop.f90:204-1067); fac is the
// diagonal of stag_mg_update_2d/3d (stag_mg_solver.f90:1769-2463) in its own
// order.  Lout = (theta*alpha - L)phi at the point.
template <typename T>
__device__ __forceinline__ void stag_point(int visc, const Grid &g, int a, const Pt &P,
                                           const T *const phi[3], const StagCoef<T> &K,
                                           T dxsqinv, T &Lout, T &fac) {
  const int dim = g.dim;
  int t1, t2 = -1;
  if (dim == 2) {
    t1 = 1 - a;
  } else {
    t1 = (a == 0) ? 1 : 0;
    t2 = (a == 2) ? 1 : 2;
  }
  const long p = at3(g, P, 0, 0, 0);
  const long pa = at_sh(g, P, a, 1, -1, 0), pma = at_sh(g, P, a, -1, -1, 0);
  const long p1p = at_sh(g, P, t1, 1, -1, 0), p1m = at_sh(g, P, t1, -1, -1, 0);
  const T *F = phi[a];
  const T av = K.alpha[a][p];
  const T fourthirds = (T)(4.0 / 3.0), twothirds = (T)(2.0 / 3.0), onethird = (T)(1.0 / 3.0);
  if (visc > 0) {
    const T b = K.beta[0], c = K.gamma[0];
    if (dim == 3) {
      const long p2p = at_sh(g, P, t2, 1, -1, 0), p2m = at_sh(g, P, t2, -1, -1, 0);
      if (visc == 1) {
        // absolute x, y, z order in 3D
        const long xp = at3(g, P, 1, 0, 0), xm = at3(g, P, -1, 0, 0);
        const long yp = at3(g, P, 0, 1, 0), ym = at3(g, P, 0, -1, 0);
        const long zp = at3(g, P, 0, 0, 1), zm = at3(g, P, 0, 0, -1);
        T s = F[xp] + F[xm] + F[yp] + F[ym] + F[zp] + F[zm];
        Lout = F[p] * (av + (T)6.0 * b * dxsqinv) - s * b * dxsqinv;
        fac = av + (T)6.0 * b * dxsqinv;
        return;
      }
      const T *G1 = phi[t1], *G2 = phi[t2];
      const long q1 = at_sh(g, P, a, -1, t1, 1), q2 = at_sh(g, P, a, -1, t2, 1);
      T cr = G1[p1p] - G1[p] - G1[q1] + G1[pma] + G2[p2p] - G2[p] - G2[q2] + G2[pma];
      if (visc == 2) {
        T s = (T)2.0 * F[pa] + (T)2.0 * F[pma] + F[p1p] + F[p1m] + F[p2p] + F[p2m];
        s = s + G1[p1p] - G1[p] - G1[q1] + G1[pma] + G2[p2p] - G2[p] - G2[q2] + G2[pma];
        Lout = F[p] * (av + (T)8.0 * b * dxsqinv) - s * b * dxsqinv;
        fac = av + (T)8.0 * b * dxsqinv;
      } else {
        Lout = F[p] * (av + ((T)20.0 * b * onethird + (T)2.0 * c) * dxsqinv) -
               ((F[pa] + F[pma]) * (fourthirds * b + c) + (F[p1p] + F[p1m] + F[p2p] + F[p2m]) * b +
                cr * (onethird * b + c)) * dxsqinv;
        fac = av + ((T)(20.0 / 3.0) * b + (T)2.0 * c) * dxsqinv;
      }
      return;
    }
    // 2D constant coefficients: along direction first
    const T *G1 = phi[t1];
    const long q1 = at_sh(g, P, a, -1, t1, 1);
    if (visc == 1) {
      T s = F[pa] + F[pma] + F[p1p] + F[p1m];
      Lout = F[p] * (av + (T)4.0 * b * dxsqinv) - s * b * dxsqinv;
      fac = av + (T)4.0 * b * dxsqinv;
    } else if (visc == 2) {
      T s = (T)2.0 * F[pa] + (T)2.0 * F[pma] + F[p1p] + F[p1m] + G1[p1p] - G1[p] - G1[q1] + G1[pma];
      Lout = F[p] * (av + (T)6.0 * b * dxsqinv) - s * b * dxsqinv;
      fac = av + (T)6.0 * b * dxsqinv;
    } else {
      T cr = G1[p1p] - G1[p] - G1[q1] + G1[pma];
      Lout = F[p] * (av + ((T)14.0 * b * onethird + (T)2.0 * c) * dxsqinv) -
             ((F[pa] + F[pma]) * (fourthirds * b + c) + (F[p1p] + F[p1m]) * b + cr * (onethird * b + c)) *
                 dxsqinv;
      fac = av + ((T)(14.0 / 3.0) * b + (T)2.0 * c) * dxsqinv;
    }
    return;
  }
  // variable coefficients
  const T *B = K.beta;
  const T B0 = B[p], B1 = B[pma];
  const T *E1 = K.bed[edge_of(a, t1)];
  const T E10 = E1[p], E11 = E1[p1p];
  if (dim == 2) {
    const T *G1 = phi[t1];
    const long q1 = at_sh(g, P, a, -1, t1, 1);
    if (visc == -1) {
      T dg = B0 + B1 + E11 + E10;
      T off = F[pa] * B0 + F[pma] * B1 + F[p1p] * E11 + F[p1m] * E10;
      Lout = F[p] * (av + dg * dxsqinv) - off * dxsqinv;
      fac = av + (B0 + B1 + E10 + E11) * dxsqinv;
    } else if (visc == -2) {
      T dg = (T)2.0 * B0 + (T)2.0 * B1 + E11 + E10;
      T off = (T)2.0 * F[pa] * B0 + (T)2.0 * F[pma] * B1 + F[p1p] * E11 + F[p1m] * E10 + G1[p1p] * E11 -
              G1[p] * E10 - G1[q1] * E11 + G1[pma] * E10;
      Lout = F[p] * (av + dg * dxsqinv) - off * dxsqinv;
      fac = av + ((T)2.0 * B0 + (T)2.0 * B1 + E10 + E11) * dxsqinv;
    } else {
      const T G0 = K.gamma[p], Gm = K.gamma[pma];
      T dg = fourthirds * B0 + G0 + fourthirds * B1 + Gm + E11 + E10;
      T off = F[pa] * (fourthirds * B0 + G0) + F[pma] * (fourthirds * B1 + Gm) + F[p1p] * E11 +
              F[p1m] * E10 + G1[p1p] * (E11 - twothirds * B0 + G0) - G1[p] * (E10 - twothirds * B0 + G0) -
              G1[q1] * (E11 - twothirds * B1 + Gm) + G1[pma] * (E10 - twothirds * B1 + Gm);
      Lout = F[p] * (av + dg * dxsqinv) - off * dxsqinv;
      fac = av + (fourthirds * B0 + G0 + fourthirds * B1 + Gm + E10 + E11) * dxsqinv;
    }
    return;
  }
  const long p2p = at_sh(g, P, t2, 1, -1, 0), p2m = at_sh(g, P, t2, -1, -1, 0);
  const T *E2 = K.bed[edge_of(a, t2)];
  const T E20 = E2[p], E21 = E2[p2p];
  if (visc == -1) {
    T dg = B0 + B1 + E10 + E11 + E20 + E21;
    T off = F[pa] * B0 + F[pma] * B1 + F[p1p] * E11 + F[p1m] * E10 + F[p2p] * E21 + F[p2m] * E20;
    fac = av + dg * dxsqinv;
    Lout = F[p] * fac - off * dxsqinv;
    return;
  }
  const T *G1 = phi[t1], *G2 = phi[t2];
  const long q1 = at_sh(g, P, a, -1, t1, 1), q2 = at_sh(g, P, a, -1, t2, 1);
  if (visc == -2) {
    T dg = (T)2.0 * B0 + (T)2.0 * B1 + E10 + E11 + E20 + E21;
    T off = (T)2.0 * F[pa] * B0 + (T)2.0 * F[pma] * B1 + F[p1p] * E11 + F[p1m] * E10 + F[p2p] * E21 +
            F[p2m] * E20 + G1[p1p] * E11 - G1[p] * E10 - G1[q1] * E11 + G1[pma] * E10 + G2[p2p] * E21 -
            G2[p] * E20 - G2[q2] * E21 + G2[pma] * E20;
    fac = av + dg * dxsqinv;
    Lout = F[p] * fac - off * dxsqinv;
    return;
  }
  const T G0 = K.gamma[p], Gm = K.gamma[pma];
  T dg = fourthirds * B0 + G0 + fourthirds * B1 + Gm + E10 + E11 + E20 + E21;
  T off = F[pa] * (fourthirds * B0 + G0) + F[pma] * (fourthirds * B1 + Gm) + F[p1p] * E11 + F[p1m] * E10 +
          F[p2p] * E21 + F[p2m] * E20 + G1[p1p] * (E11 - twothirds * B0 + G0) -
          G1[p] * (E10 - twothirds * B0 + G0) - G1[q1] * (E11 - twothirds * B1 + Gm) +
          G1[pma] * (E10 - twothirds * B1 + Gm) + G2[p2p] * (E21 - twothirds * B0 + G0) -
          G2[p] * (E20 - twothirds * B0 + G0) - G2[q2] * (E21 - twothirds * B1 + Gm) +
          G2[pma] * (E20 - twothirds * B1 + Gm);
  fac = av + dg * dxsqinv;
  Lout = F[p] * fac - off * dxsqinv;
}

}  // namespace lmm
