// Point-level device functions of the multigrid kernels, shared by the
// one-launch-per-operation kernels (lmm_kernels.hip) and the single-workgroup
// coarse-level V-cycle (lmm_coarse.hip).  Each computes one point of its
// operation, given the point's linear index t.
#pragma once

#include "lmm_common.h"

namespace lmm {

template <typename T>
struct Ptr3 {
  T *c[3];
};

template <typename T>
static Ptr3<T> p3(T *const a[3]) {
  Ptr3<T> r;
  for (int d = 0; d < 3; d++) r.c[d] = a[d];
  return r;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup w runs on
// XCD w % 8), each with its own L2.  Renumbering them so that each XCD takes
// one contiguous eighth of the index range keeps the neighbouring rows and
// planes that a stencil reads in the L2 of the XCD that reads them.
__device__ __forceinline__ long xcd_block() {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  if (nb % 8 != 0) return b;
  return (long)(b % 8) * (nb / 8) + b / 8;
}

// Linear point index -> (i, j, k).  Point counts stay below 2^31 (checked
// when a context is created), so 32-bit unsigned division suffices; the
// 64-bit form costs several times the instructions of the stencil itself.
__device__ __forceinline__ void decode(const Grid &g, long t, int &i, int &j, int &k) {
  const unsigned u = (unsigned)t, nx = (unsigned)g.n[0], ny = (unsigned)g.n[1];
  const unsigned r = u / nx;
  i = (int)(u - r * nx);
  const unsigned q = r / ny;
  j = (int)(r - q * ny);
  k = (int)q;
}

// Index of the colour points (one parity) -> (i, j, k): the x extent is even.
__device__ __forceinline__ void decode_half(const Grid &g, long t, int par, int &i, int &j, int &k) {
  const unsigned u = (unsigned)t, half = (unsigned)(g.n[0] >> 1), ny = (unsigned)g.n[1];
  const unsigned r = u / half;
  const int ii = (int)(u - r * half);
  const unsigned q = r / ny;
  j = (int)(r - q * ny);
  k = (int)q;
  i = 2 * ii + ((par + j + k) & 1);
}

// One colour point t of the 2*dim-colour Gauss-Seidel: applyop at the
// point followed by phi += omega*(rhs - Lphi)/fac, fused.  Points of one
// colour read only points of other colours, so the in-place update equals
// the reference's two passes (stag_mg_solver.f90:373-382).
template <typename T>
__device__ __forceinline__ void gs_colour_at(const Grid &g, int visc, int a, int par, T omega, T dxsqinv,
                                             const Ptr3<T> &ph, const StagCoef<T> &K, const T *rhs, long t) {
  int i, j, k;
  decode_half(g, t, par, i, j, k);
  Pt P = make_pt(g, i, j, k);
  const T *phr[3] = {ph.c[0], ph.c[1], ph.c[2]};
  T L, fac;
  stag_point<T>(visc, g, a, P, phr, K, dxsqinv, L, fac);
  long p = i + g.s[1] * j + g.s[2] * k;
  ph.c[a][p] = ph.c[a][p] + gs_div(omega * (rhs[p] - L), fac);
}

// res = rhs - L phi at point t of component a (stag_mg_solver.f90:404-416: -(Lphi - rhs))
template <typename T>
__device__ __forceinline__ void residual_at(const Grid &g, int visc, T dxsqinv, const Ptr3<T> &ph,
                                            const Ptr3<T> &rh, const Ptr3<T> &res, const StagCoef<T> &K, int a,
                                            long t) {
  int i, j, k;
  decode(g, t, i, j, k);
  Pt P = make_pt(g, i, j, k);
  const T *phr[3] = {ph.c[0], ph.c[1], ph.c[2]};
  T L, fac;
  stag_point<T>(visc, g, a, P, phr, K, dxsqinv, L, fac);
  long p = i + g.s[1] * j + g.s[2] * k;
  res.c[a][p] = rh.c[a][p] - L;
}

// fine index of coarse point with fine offsets (along a: oa, transverse t1: o1, t2: o2)
__device__ __forceinline__ long fidx(const Grid &gf, int a, int t1, int t2, const int C[3], int oa, int o1,
                                     int o2) {
  int f[3] = {2 * C[0], 2 * C[1], 2 * C[2]};
  if (gf.dim == 2) f[2] = 0;
  f[a] += oa;
  f[t1] += o1;
  if (t2 >= 0) f[t2] += o2;
  for (int d = 0; d < gf.dim; d++) {
    if (f[d] < 0) f[d] += gf.n[d];
    if (f[d] >= gf.n[d]) f[d] -= gf.n[d];
  }
  return f[0] + gf.s[1] * f[1] + gf.s[2] * f[2];
}

// Coarse point t of component a of the staggered restriction
// (stag_mg_solver.f90:971-1175): simple = face average, else 6-point (2D) /
// 12-point (3D) full weighting.
template <typename T>
__device__ __forceinline__ void restrict_at(const Grid &gc, const Grid &gf, int simple, const Ptr3<T> &fi,
                                            const Ptr3<T> &co, int a, long t) {
  int C[3];
  decode(gc, t, C[0], C[1], C[2]);
  int t1, t2 = -1;
  if (gc.dim == 2) {
    t1 = 1 - a;
  } else {
    t1 = (a == 0) ? 1 : 0;
    t2 = (a == 2) ? 1 : 2;
  }
  const T *f = fi.c[a];
  T v;
  if (gc.dim == 2) {
    if (simple) {
      v = (T)0.5 * (f[fidx(gf, a, t1, t2, C, 0, 0, 0)] + f[fidx(gf, a, t1, t2, C, 0, 1, 0)]);
    } else {
      v = (T)0.25 * (f[fidx(gf, a, t1, t2, C, 0, 0, 0)] + f[fidx(gf, a, t1, t2, C, 0, 1, 0)]) +
          (T)0.125 * (f[fidx(gf, a, t1, t2, C, 1, 0, 0)] + f[fidx(gf, a, t1, t2, C, 1, 1, 0)] +
                      f[fidx(gf, a, t1, t2, C, -1, 0, 0)] + f[fidx(gf, a, t1, t2, C, -1, 1, 0)]);
    }
  } else {
    T g0 = f[fidx(gf, a, t1, t2, C, 0, 0, 0)] + f[fidx(gf, a, t1, t2, C, 0, 1, 0)] +
           f[fidx(gf, a, t1, t2, C, 0, 0, 1)] + f[fidx(gf, a, t1, t2, C, 0, 1, 1)];
    if (simple) {
      v = (T)0.25 * g0;
    } else {
      T gp = f[fidx(gf, a, t1, t2, C, 1, 0, 0)] + f[fidx(gf, a, t1, t2, C, 1, 1, 0)] +
             f[fidx(gf, a, t1, t2, C, 1, 0, 1)] + f[fidx(gf, a, t1, t2, C, 1, 1, 1)];
      T gm = f[fidx(gf, a, t1, t2, C, -1, 0, 0)] + f[fidx(gf, a, t1, t2, C, -1, 1, 0)] +
             f[fidx(gf, a, t1, t2, C, -1, 0, 1)] + f[fidx(gf, a, t1, t2, C, -1, 1, 1)];
      v = (T)0.125 * g0 + (T)0.0625 * gp + (T)0.0625 * gm;
    }
  }
  co.c[a][C[0] + gc.s[1] * C[1] + gc.s[2] * C[2]] = v;
}

// phi_f += P phi_c for face component A in dimension D (stag_prolongation_2d/3d,
// stag_mg_solver.f90:1370-1633): bilinear in the transverse directions,
// linear along A.  The component is a template argument so every index is static.
// The prolongation arithmetic of one fine point c of component A
// (stag_prolongation_2d/3d, stag_mg_solver.f90:1370-1633): at(sa, s1, s2) is
// the coarse value at Q + sa along A, + s1 along the first transverse
// direction and + s2 along the second (s1, s2 in {0, o1 / o2}).
template <typename T, int D, int A, typename F>
__device__ __forceinline__ T prolong_combine(const int c[3], T v, F &&at) {
  constexpr int T1 = (D == 2) ? 1 - A : (A == 0 ? 1 : 0);
  constexpr int T2 = (D == 2) ? -1 : (A == 2 ? 1 : 2);
  const int o1 = (c[T1] & 1) ? 1 : -1;
  const int o2 = (T2 >= 0 && (c[T2 >= 0 ? T2 : 0] & 1)) ? 1 : -1;
  const bool even = (c[A] & 1) == 0;
  if (D == 2) {
    if (even) {
      v = v + (T)0.75 * at(0, 0, 0) + (T)0.25 * at(0, o1, 0);
    } else {
      v = v + (T)0.375 * at(0, 0, 0) + (T)0.125 * at(0, o1, 0) + (T)0.375 * at(1, 0, 0) +
          (T)0.125 * at(1, o1, 0);
    }
  } else {
    const T nine16 = (T)0.5625, three16 = (T)0.1875, one16 = (T)0.0625;
    const T nine32 = (T)0.28125, three32 = (T)0.09375, one32 = (T)0.03125;
    if (even) {
      v = v + nine16 * at(0, 0, 0) + three16 * at(0, o1, 0) + three16 * at(0, 0, o2) + one16 * at(0, o1, o2);
    } else {
      v = v + nine32 * at(0, 0, 0) + three32 * at(0, o1, 0) + three32 * at(0, 0, o2) + one32 * at(0, o1, o2) +
          nine32 * at(1, 0, 0) + three32 * at(1, o1, 0) + three32 * at(1, 0, o2) + one32 * at(1, o1, o2);
    }
  }
  return v;
}

template <typename T, int D, int A>
__device__ __forceinline__ T prolong_value(const Grid &gf, const Grid &gc, const T *C, const int c[3], T v) {
  constexpr int T1 = (D == 2) ? 1 - A : (A == 0 ? 1 : 0);
  constexpr int T2 = (D == 2) ? -1 : (A == 2 ? 1 : 2);
  const int Q[3] = {c[0] >> 1, c[1] >> 1, c[2] >> 1};
  return prolong_combine<T, D, A>(c, v, [&](int sa, int s1, int s2) -> T {
    int q[3] = {Q[0], Q[1], Q[2]};
    q[A] += sa;
    q[T1] += s1;
    if (T2 >= 0) q[T2 >= 0 ? T2 : 0] += s2;
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (q[d] < 0) q[d] += gc.n[d];
      if (q[d] >= gc.n[d]) q[d] -= gc.n[d];
    }
    return C[q[0] + gc.s[1] * q[1] + gc.s[2] * q[2]];
  });
}

template <typename T, int D, int A>
__device__ __forceinline__ void prolong_at(const Grid &gf, const Grid &gc, T *fine, const T *C, long t) {
  int c[3];
  decode(gf, t, c[0], c[1], c[2]);
  const long p = c[0] + gf.s[1] * c[1] + gf.s[2] * c[2];
  fine[p] = prolong_value<T, D, A>(gf, gc, C, c, fine[p]);
}

template <typename T>
__device__ __forceinline__ T cc_apply(const Grid &g, int i, int j, int k, const T *phi, const Ptr3<T> &A,
                                      T inv, T &diag) {
  long p = i + g.s[1] * j + g.s[2] * k;
  long xm = wm(i, g.n[0]) + g.s[1] * j + g.s[2] * k, xp = wp(i, g.n[0]) + g.s[1] * j + g.s[2] * k;
  long ym = i + g.s[1] * wm(j, g.n[1]) + g.s[2] * k, yp = i + g.s[1] * wp(j, g.n[1]) + g.s[2] * k;
  T axl = A.c[0][p], axh = A.c[0][xp], ayl = A.c[1][p], ayh = A.c[1][yp];
  T s = axl + axh + ayl + ayh;
  T off = axl * phi[xm] + axh * phi[xp] + ayl * phi[ym] + ayh * phi[yp];
  if (g.dim == 3) {
    long zm = i + g.s[1] * j + g.s[2] * wm(k, g.n[2]), zp = i + g.s[1] * j + g.s[2] * wp(k, g.n[2]);
    T azl = A.c[2][p], azh = A.c[2][zp];
    s = s + azl + azh;
    off = off + azl * phi[zm] + azh * phi[zp];
  }
  diag = -s * inv;
  return (-s * phi[p] + off) * inv;
}

// One red/black point t of the cell-centred Gauss-Seidel.
template <typename T>
__device__ __forceinline__ void cc_rb_at(const Grid &g, int colour, T inv, T *phi, const T *rhs, const Ptr3<T> &A,
                                         long t) {
  int i, j, k;
  decode_half(g, t, colour, i, j, k);
  T diag;
  T lp = cc_apply<T>(g, i, j, k, phi, A, inv, diag);
  long p = i + g.s[1] * j + g.s[2] * k;
  phi[p] = phi[p] + gs_div(rhs[p] - lp, diag);
}

template <typename T>
__device__ __forceinline__ void cc_residual_at(const Grid &g, T inv, const T *phi, const T *rhs, T *res,
                                               const Ptr3<T> &A, long t) {
  int i, j, k;
  decode(g, t, i, j, k);
  T diag;
  res[t] = rhs[t] - cc_apply<T>(g, i, j, k, phi, A, inv, diag);
}

// average of the 2^dim fine cells (cc_restriction, stag_mg_solver.f90:873-968)
template <typename T>
__device__ __forceinline__ void cc_restrict_at(const Grid &gc, const Grid &gf, const T *f, T *c, long t) {
  int i, j, k;
  decode(gc, t, i, j, k);
  long b = 2 * i + gf.s[1] * (2 * j) + gf.s[2] * (2 * k);
  long sy = gf.s[1], sz = gf.s[2];
  T v;
  if (gc.dim == 2) {
    v = (T)0.25 * (f[b] + f[b + 1] + f[b + sy] + f[b + 1 + sy]);
  } else {
    v = (T)0.125 * (f[b] + f[b + 1] + f[b + sy] + f[b + 1 + sy] + f[b + sz] + f[b + 1 + sz] + f[b + sy + sz] +
                    f[b + 1 + sy + sz]);
  }
  c[t] = v;
}

// injection prolongation, added (the cell-centred V-cycle's correction)
template <typename T>
__device__ __forceinline__ void cc_prolong_add_at(const Grid &gf, const Grid &gc, T *f, const T *c, long t) {
  int i, j, k;
  decode(gf, t, i, j, k);
  f[t] = f[t] + c[i / 2 + gc.s[1] * (j / 2) + gc.s[2] * (k / 2)];
}

__device__ __forceinline__ int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Two consecutive elements a[p], a[p+1] with one vector load (p even, so the
// pair is aligned to its size).
template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};
template <typename T>
__device__ __forceinline__ void ld2(const T *a, long p, T &x0, T &x1) {
  const typename Vec2<T>::type v = *reinterpret_cast<const typename Vec2<T>::type *>(a + p);
  x0 = v.x;
  x1 = v.y;
}
template <typename T>
__device__ __forceinline__ void st2(T *a, long p, T x0, T x1) {
  typename Vec2<T>::type v;
  v.x = x0;
  v.y = x1;
  *reinterpret_cast<typename Vec2<T>::type *>(a + p) = v;
}

// A red-black stencil is described by a policy: the per-point coefficients
// (Coef, loaded by load) and the update of one point from its old value and
// its six neighbours in the same array (update), with the kernel's scalar
// arguments in Args.

// Component A of the staggered operator, visc_type -1, 1, 2 or 3.  The
// terms are those of stag_applyop_3d (off-diagonal sum) and
// stag_mg_update_3d, in their order.  visc_type 2 and 3 couple the
// components: the point reads the other two components (e1, e2 then hold
// the arrays of t1 and t2, in the state the colour order gives them: already
// swept components new, later ones old) and the constant coefficients
// beta[0], gamma[0] (the tile corner).
template <typename T, int A, int VISC>
struct StagRB {
  struct Args {
    T omega, dxsqinv;
    const T *rhs, *alpha, *beta, *e1, *e2;
    const T *gamma;  // visc_type 3
  };
  struct Coef {
    T av, rhs, B0, B1, E10, E11, E20, E21;
    T G[8];  // visc_type 2: the eight cross-term values; 3: their sum in G[0]
  };
  // the array pointers in a fixed order, and back (the kernels receive them
  // as separate __restrict__ parameters)
  static __host__ __device__ void ptrs(const Args &a, const T *q[5]) {
    q[0] = a.rhs;
    q[1] = a.alpha;
    q[2] = a.beta;
    q[3] = a.e1;
    q[4] = a.e2;
  }
  static __device__ __forceinline__ Args bind(Args a, const T *q0, const T *q1, const T *q2, const T *q3,
                                              const T *q4) {
    a.rhs = q0;
    a.alpha = q1;
    a.beta = q2;
    a.e1 = q3;
    a.e2 = q4;
    return a;
  }
  static __device__ __forceinline__ Coef load(const Grid &g, int i, int j, int k, const Args &a) {
    Coef c;
    const long sy = g.s[1], sz = g.s[2];
    const long p = i + sy * j + sz * k;
    c.av = a.alpha[p];
    c.rhs = a.rhs[p];
    if (VISC == -1) {
      // along a = A; t1, t2 = the other two directions in increasing order
      const int im = (A == 0) ? wrapc(i - 1, g.n[0]) : i;
      const int jm = (A == 1) ? wrapc(j - 1, g.n[1]) : j;
      const int km = (A == 2) ? wrapc(k - 1, g.n[2]) : k;
      c.B0 = a.beta[p];
      c.B1 = a.beta[im + sy * jm + sz * km];
      // t1: y for A = 0, x otherwise; t2: z for A < 2, y for A = 2
      const long p1 = (A == 0) ? (i + sy * wrapc(j + 1, g.n[1]) + sz * k) : (wrapc(i + 1, g.n[0]) + sy * j + sz * k);
      const long p2 = (A == 2) ? (i + sy * wrapc(j + 1, g.n[1]) + sz * k) : (i + sy * j + sz * wrapc(k + 1, g.n[2]));
      c.E10 = a.e1[p];
      c.E11 = a.e1[p1];
      c.E20 = a.e2[p];
      c.E21 = a.e2[p2];
    } else if (VISC >= 2) {
      c.B0 = a.beta[0];
      c.B1 = (VISC == 3) ? a.gamma[0] : (T)0;
      c.E10 = c.E11 = c.E20 = c.E21 = (T)0;
      const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
      auto at = [&](int ii, int jj, int kk) { return (long)ii + sy * jj + sz * kk; };
      // pma: -1 along A; p1p / p2p: +1 along t1 / t2; q1 / q2: both shifts
      const int ia = (A == 0) ? wrapc(i - 1, nx) : i, ja = (A == 1) ? wrapc(j - 1, ny) : j;
      const int ka = (A == 2) ? wrapc(k - 1, nz) : k;
      const long pma = at(ia, ja, ka);
      long p1p, q1, p2p, q2;
      if (A == 0) {
        p1p = at(i, wrapc(j + 1, ny), k);
        q1 = at(ia, wrapc(j + 1, ny), k);
        p2p = at(i, j, wrapc(k + 1, nz));
        q2 = at(ia, j, wrapc(k + 1, nz));
      } else if (A == 1) {
        p1p = at(wrapc(i + 1, nx), j, k);
        q1 = at(wrapc(i + 1, nx), ja, k);
        p2p = at(i, j, wrapc(k + 1, nz));
        q2 = at(i, ja, wrapc(k + 1, nz));
      } else {
        p1p = at(wrapc(i + 1, nx), j, k);
        q1 = at(wrapc(i + 1, nx), j, ka);
        p2p = at(i, wrapc(j + 1, ny), k);
        q2 = at(i, wrapc(j + 1, ny), ka);
      }
      const T *G1 = a.e1, *G2 = a.e2;
      if (VISC == 2) {
        c.G[0] = G1[p1p];
        c.G[1] = G1[p];
        c.G[2] = G1[q1];
        c.G[3] = G1[pma];
        c.G[4] = G2[p2p];
        c.G[5] = G2[p];
        c.G[6] = G2[q2];
        c.G[7] = G2[pma];
      } else {
        c.G[0] = G1[p1p] - G1[p] - G1[q1] + G1[pma] + G2[p2p] - G2[p] - G2[q2] + G2[pma];
      }
    } else {
      c.B0 = a.beta[0];  // constant coefficients: beta at the tile corner
      c.B1 = c.E10 = c.E11 = c.E20 = c.E21 = (T)0;
    }
    return c;
  }
  // Coefficients of the points (i, j, k) and (i+1, j, k), i even, with one
  // vector load per array where the two points' values are adjacent.
  static __device__ __forceinline__ void load_pair(const Grid &g, int i, int j, int k, const Args &a, Coef &c0,
                                                   Coef &c1) {
    if (VISC >= 2) {
      c0 = load(g, i, j, k, a);
      c1 = load(g, i + 1, j, k, a);
      return;
    }
    const long sy = g.s[1], sz = g.s[2];
    const long p = i + sy * j + sz * k;
    ld2(a.alpha, p, c0.av, c1.av);
    ld2(a.rhs, p, c0.rhs, c1.rhs);
    if (VISC == -1) {
      ld2(a.beta, p, c0.B0, c1.B0);
      if (A == 0) {
        c0.B1 = a.beta[wrapc(i - 1, g.n[0]) + sy * j + sz * k];
        c1.B1 = c0.B0;
        ld2(a.e1, p, c0.E10, c1.E10);  // xy edges
        ld2(a.e1, i + sy * wrapc(j + 1, g.n[1]) + sz * k, c0.E11, c1.E11);
        ld2(a.e2, p, c0.E20, c1.E20);  // xz edges
        ld2(a.e2, i + sy * j + sz * wrapc(k + 1, g.n[2]), c0.E21, c1.E21);
      } else {
        if (A == 1)
          ld2(a.beta, i + sy * wrapc(j - 1, g.n[1]) + sz * k, c0.B1, c1.B1);
        else
          ld2(a.beta, i + sy * j + sz * wrapc(k - 1, g.n[2]), c0.B1, c1.B1);
        ld2(a.e1, p, c0.E10, c1.E10);  // xy (A = 1) or xz (A = 2) edges, neighbour along x
        c0.E11 = c1.E10;
        c1.E11 = a.e1[wrapc(i + 2, g.n[0]) + sy * j + sz * k];
        ld2(a.e2, p, c0.E20, c1.E20);  // yz edges
        if (A == 1)
          ld2(a.e2, i + sy * j + sz * wrapc(k + 1, g.n[2]), c0.E21, c1.E21);
        else
          ld2(a.e2, i + sy * wrapc(j + 1, g.n[1]) + sz * k, c0.E21, c1.E21);
      }
    } else {
      c0.B0 = c1.B0 = a.beta[0];
      c0.B1 = c0.E10 = c0.E11 = c0.E20 = c0.E21 = (T)0;
      c1.B1 = c1.E10 = c1.E11 = c1.E20 = c1.E21 = (T)0;
    }
  }
  // (theta*alpha - L) phi at the point, as stag_applyop_3d computes it
  static __device__ __forceinline__ T apply(const Coef &c, T F, T xp, T xm, T yp, T ym, T zp, T zm, T dxsqinv) {
    if (VISC == 1) {
      const T b = c.B0;
      T s = xp + xm + yp + ym + zp + zm;
      return F * (c.av + (T)6.0 * b * dxsqinv) - s * b * dxsqinv;
    } else if (VISC >= 2) {
      const T Fa = (A == 0) ? xp : (A == 1 ? yp : zp);
      const T Fma = (A == 0) ? xm : (A == 1 ? ym : zm);
      const T F1p = (A == 0) ? yp : xp;
      const T F1m = (A == 0) ? ym : xm;
      const T F2p = (A == 2) ? yp : zp;
      const T F2m = (A == 2) ? ym : zm;
      const T b = c.B0;
      if (VISC == 2) {
        T s = (T)2.0 * Fa + (T)2.0 * Fma + F1p + F1m + F2p + F2m;
        s = s + c.G[0] - c.G[1] - c.G[2] + c.G[3] + c.G[4] - c.G[5] - c.G[6] + c.G[7];
        return F * (c.av + (T)8.0 * b * dxsqinv) - s * b * dxsqinv;
      }
      const T cg = c.B1, fourthirds = (T)(4.0 / 3.0), onethird = (T)(1.0 / 3.0);
      return F * (c.av + ((T)20.0 * b * onethird + (T)2.0 * cg) * dxsqinv) -
             ((Fa + Fma) * (fourthirds * b + cg) + (F1p + F1m + F2p + F2m) * b + c.G[0] * (onethird * b + cg)) *
                 dxsqinv;
    } else {
      const T Fa = (A == 0) ? xp : (A == 1 ? yp : zp);
      const T Fma = (A == 0) ? xm : (A == 1 ? ym : zm);
      const T F1p = (A == 0) ? yp : xp;
      const T F1m = (A == 0) ? ym : xm;
      const T F2p = (A == 2) ? yp : zp;
      const T F2m = (A == 2) ? ym : zm;
      T dg = c.B0 + c.B1 + c.E10 + c.E11 + c.E20 + c.E21;
      T off = Fa * c.B0 + Fma * c.B1 + F1p * c.E11 + F1m * c.E10 + F2p * c.E21 + F2m * c.E20;
      T fac = c.av + dg * dxsqinv;
      return F * fac - off * dxsqinv;
    }
  }
  static __device__ __forceinline__ T update(const Coef &c, T F, T xp, T xm, T yp, T ym, T zp, T zm,
                                             const Args &a) {
    const T omega = a.omega, dxsqinv = a.dxsqinv;
    if (VISC == 1) {
      const T b = c.B0;
      T s = xp + xm + yp + ym + zp + zm;
      T L = F * (c.av + (T)6.0 * b * dxsqinv) - s * b * dxsqinv;
      T fac = c.av + (T)6.0 * b * dxsqinv;
      return F + gs_div(omega * (c.rhs - L), fac);
    } else if (VISC >= 2) {
      const T Fa = (A == 0) ? xp : (A == 1 ? yp : zp);
      const T Fma = (A == 0) ? xm : (A == 1 ? ym : zm);
      const T F1p = (A == 0) ? yp : xp;
      const T F1m = (A == 0) ? ym : xm;
      const T F2p = (A == 2) ? yp : zp;
      const T F2m = (A == 2) ? ym : zm;
      const T b = c.B0;
      T L, fac;
      if (VISC == 2) {
        T s = (T)2.0 * Fa + (T)2.0 * Fma + F1p + F1m + F2p + F2m;
        s = s + c.G[0] - c.G[1] - c.G[2] + c.G[3] + c.G[4] - c.G[5] - c.G[6] + c.G[7];
        L = F * (c.av + (T)8.0 * b * dxsqinv) - s * b * dxsqinv;
        fac = c.av + (T)8.0 * b * dxsqinv;
      } else {
        const T cg = c.B1, fourthirds = (T)(4.0 / 3.0), onethird = (T)(1.0 / 3.0);
        L = F * (c.av + ((T)20.0 * b * onethird + (T)2.0 * cg) * dxsqinv) -
            ((Fa + Fma) * (fourthirds * b + cg) + (F1p + F1m + F2p + F2m) * b + c.G[0] * (onethird * b + cg)) *
                dxsqinv;
        fac = c.av + ((T)(20.0 / 3.0) * b + (T)2.0 * cg) * dxsqinv;
      }
      return F + gs_div(omega * (c.rhs - L), fac);
    } else {
      // along / transverse neighbours of component A
      const T Fa = (A == 0) ? xp : (A == 1 ? yp : zp);
      const T Fma = (A == 0) ? xm : (A == 1 ? ym : zm);
      const T F1p = (A == 0) ? yp : xp;
      const T F1m = (A == 0) ? ym : xm;
      const T F2p = (A == 2) ? yp : zp;
      const T F2m = (A == 2) ? ym : zm;
      T dg = c.B0 + c.B1 + c.E10 + c.E11 + c.E20 + c.E21;
      T off = Fa * c.B0 + Fma * c.B1 + F1p * c.E11 + F1m * c.E10 + F2p * c.E21 + F2m * c.E20;
      T fac = c.av + dg * dxsqinv;
      T L = F * fac - off * dxsqinv;
      return F + gs_div(omega * (c.rhs - L), fac);
    }
  }
};

// Cell-centred div(a grad phi) of the MAC-projection multigrid: the point
// arithmetic of cc_apply / the red-black update of lmm_points.h.
template <typename T>
struct CCRB {
  struct Args {
    T inv;
    const T *rhs, *ax, *ay, *az;
  };
  struct Coef {
    T rhs, axl, axh, ayl, ayh, azl, azh;
  };
  static __host__ __device__ void ptrs(const Args &a, const T *q[5]) {
    q[0] = a.rhs;
    q[1] = a.ax;
    q[2] = a.ay;
    q[3] = a.az;
    q[4] = nullptr;
  }
  static __device__ __forceinline__ Args bind(Args a, const T *q0, const T *q1, const T *q2, const T *q3,
                                              const T *) {
    a.rhs = q0;
    a.ax = q1;
    a.ay = q2;
    a.az = q3;
    return a;
  }
  static __device__ __forceinline__ Coef load(const Grid &g, int i, int j, int k, const Args &a) {
    Coef c;
    const long sy = g.s[1], sz = g.s[2];
    const long p = i + sy * j + sz * k;
    c.rhs = a.rhs[p];
    c.axl = a.ax[p];
    c.axh = a.ax[wrapc(i + 1, g.n[0]) + sy * j + sz * k];
    c.ayl = a.ay[p];
    c.ayh = a.ay[i + sy * wrapc(j + 1, g.n[1]) + sz * k];
    c.azl = a.az[p];
    c.azh = a.az[i + sy * j + sz * wrapc(k + 1, g.n[2])];
    return c;
  }
  static __device__ __forceinline__ void load_pair(const Grid &g, int i, int j, int k, const Args &a, Coef &c0,
                                                   Coef &c1) {
    const long sy = g.s[1], sz = g.s[2];
    const long p = i + sy * j + sz * k;
    ld2(a.rhs, p, c0.rhs, c1.rhs);
    ld2(a.ax, p, c0.axl, c1.axl);
    c0.axh = c1.axl;
    c1.axh = a.ax[wrapc(i + 2, g.n[0]) + sy * j + sz * k];
    ld2(a.ay, p, c0.ayl, c1.ayl);
    ld2(a.ay, i + sy * wrapc(j + 1, g.n[1]) + sz * k, c0.ayh, c1.ayh);
    ld2(a.az, p, c0.azl, c1.azl);
    ld2(a.az, i + sy * j + sz * wrapc(k + 1, g.n[2]), c0.azh, c1.azh);
  }
  static __device__ __forceinline__ T update(const Coef &c, T F, T xp, T xm, T yp, T ym, T zp, T zm,
                                             const Args &a) {
    T s = c.axl + c.axh + c.ayl + c.ayh;
    T off = c.axl * xm + c.axh * xp + c.ayl * ym + c.ayh * yp;
    s = s + c.azl + c.azh;
    off = off + c.azl * zm + c.azh * zp;
    const T diag = -s * a.inv;
    const T lp = (-s * F + off) * a.inv;
    return F + gs_div(c.rhs - lp, diag);
  }
};

}  // namespace lmm
