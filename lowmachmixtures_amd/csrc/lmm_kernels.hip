// gfx950 kernels of the Stokes library: staggered and cell-centred multigrid
// pieces, the fused Stokes operator, preconditioner assembly, deterministic
// reductions, Krylov vector kernels and the flux stencils.
//
// Arithmetic follows the reference expression order (see lmm_common.h) and
// the library is built with -ffp-contract=off, so the FP64 kernels reproduce
// the reference formulas bit for bit.
#include <utility>

#include "lmm_kernels.h"
#include "lmm_points.h"

namespace lmm {

constexpr int TPB = 256;

static int g_debug_sync = 0;
// per host thread: each thread drives its own contexts (lmm_gmres_batched)
static thread_local char g_first_error[256] = {0};
void set_debug_sync(int on) { g_debug_sync = on; }
int debug_sync_on() { return g_debug_sync; }
// the context settings in force on this thread (null: the defaults)
static thread_local const Tuning *t_tuning = nullptr;
const Tuning &tuning() {
  static const Tuning defaults;
  return t_tuning ? *t_tuning : defaults;
}
TuningScope::TuningScope(const Tuning *t) : prev(t_tuning) { t_tuning = t; }
TuningScope::~TuningScope() { t_tuning = prev; }
const char *first_error() { return g_first_error; }
void report_error(const char *msg) {
  if (g_first_error[0] == 0) snprintf(g_first_error, sizeof(g_first_error), "%s", msg);
}
void clear_first_error() { g_first_error[0] = 0; }
void post_launch(hipStream_t s, const char *name) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && g_debug_sync) e = hipStreamSynchronize(s);
  if (e != hipSuccess && g_first_error[0] == 0)
    snprintf(g_first_error, sizeof(g_first_error), "%s: %s", name, hipGetErrorString(e));
}
#define LAUNCH(K, G, B, SHM, S, ...)                      \
  do {                                                   \
    hipLaunchKernelGGL(K, G, B, SHM, S, __VA_ARGS__);    \
    post_launch(S, #K);                                  \
  } while (0)
static inline int nblk(long n, int tpb = TPB) { return (int)((n + tpb - 1) / tpb); }

// ------------------------------------------------------------------------
// staggered smoother / operator
// ------------------------------------------------------------------------

// One colour of the 2*dim-colour Gauss-Seidel: applyop at the colour's points
// followed by phi += omega*(rhs - Lphi)/fac, fused.  Points of one colour
// read only points of other colours, so the in-place update equals the
// reference's two passes (stag_mg_solver.f90:373-382).
template <typename T>
__global__ void __launch_bounds__(256) k_stag_gs_colour(Grid g, int visc, int a, int par, T omega,
                                                        T dxsqinv, Ptr3<T> ph, StagCoef<T> K,
                                                        const T *__restrict__ rhs) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= (long)(g.n[0] >> 1) * g.n[1] * g.n[2]) return;
  gs_colour_at<T>(g, visc, a, par, omega, dxsqinv, ph, K, rhs, t);
}

template <typename T>
void launch_stag_gs_colour(hipStream_t s, const Grid &g, int visc, int a, int par, T omega,
                           T *const phi[3], const StagCoef<T> &K, const T *rhs) {
  long total = (long)(g.n[0] / 2) * g.n[1] * g.n[2];
  T dxsqinv = (T)(1.0 / (g.dx * g.dx));
  LAUNCH(k_stag_gs_colour<T>, dim3(nblk(total)), dim3(TPB), 0, s, g, visc, a, par, omega,
                     dxsqinv, p3(phi), K, rhs);
}

// Lphi at the points of one colour (or every point for colour 0).  The
// component is blockIdx.y, uniform across a workgroup.
template <typename T>
__global__ void __launch_bounds__(256) k_stag_applyop(Grid g, int visc, int color, T dxsqinv, Ptr3<T> ph,
                                                      Ptr3<T> lp, StagCoef<T> K, int sub) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  int i, j, k, par = (color + 1) % 2;
  const int a = (color == 0) ? (int)blockIdx.y : (color - 1) / 2;
  const long per = (color == 0) ? g.N : g.N / 2;
  if (t >= per) return;
  if (color == 0)
    decode(g, t, i, j, k);
  else
    decode_half(g, t, par, i, j, k);
  Pt P = make_pt(g, i, j, k);
  const T *phr[3] = {ph.c[0], ph.c[1], ph.c[2]};
  T L, fac;
  stag_point<T>(visc, g, a, P, phr, K, dxsqinv, L, fac);
  const long p = i + g.s[1] * j + g.s[2] * k;
  if (sub) lp.c[a][p] = lp.c[a][p] - L;  // multifab_sub_sub_c(m_update, Lphi)
  else lp.c[a][p] = L;
}

template <typename T>
void launch_stag_applyop(hipStream_t s, const Grid &g, int visc, int color, T *const phi[3],
                         T *const Lphi[3], const StagCoef<T> &K, bool sub) {
  long per = (color == 0) ? g.N : g.N / 2;
  dim3 grid(nblk(per), color == 0 ? g.dim : 1);
  T dxsqinv = (T)(1.0 / (g.dx * g.dx));
  LAUNCH(k_stag_applyop<T>, grid, dim3(TPB), 0, s, g, visc, color, dxsqinv, p3(phi), p3(Lphi), K, sub ? 1 : 0);
}

// Jacobi sweep (stag_mg_smoother = 0): tmp = phi + omega*(rhs - L phi)/fac with
// L phi from the old phi at every point; the caller copies tmp back.
template <typename T>
__global__ void __launch_bounds__(256) k_stag_jacobi(Grid g, int visc, T omega, T dxsqinv, Ptr3<T> ph,
                                                     Ptr3<T> rh, Ptr3<T> tmp, StagCoef<T> K) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const int a = (int)blockIdx.y;
  int i, j, k;
  decode(g, t, i, j, k);
  Pt P = make_pt(g, i, j, k);
  const T *phr[3] = {ph.c[0], ph.c[1], ph.c[2]};
  T L, fac;
  stag_point<T>(visc, g, a, P, phr, K, dxsqinv, L, fac);
  long p = i + g.s[1] * j + g.s[2] * k;
  tmp.c[a][p] = ph.c[a][p] + gs_div(omega * (rh.c[a][p] - L), fac);
}

template <typename T>
void launch_stag_jacobi(hipStream_t s, const Grid &g, int visc, T omega, T *const phi[3],
                        T *const rhs[3], T *const tmp[3], const StagCoef<T> &K) {
  T dxsqinv = (T)(1.0 / (g.dx * g.dx));
  LAUNCH(k_stag_jacobi<T>, dim3(nblk(g.N), g.dim), dim3(TPB), 0, s, g, visc, omega, dxsqinv,
                     p3(phi), p3(rhs), p3(tmp), K);
}

// res = rhs - L phi  (stag_mg_solver.f90:404-416: -(Lphi - rhs))
template <typename T>
__global__ void __launch_bounds__(256) k_stag_residual(Grid g, int visc, T dxsqinv, Ptr3<T> ph, Ptr3<T> rh,
                                                       Ptr3<T> res, StagCoef<T> K) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  residual_at<T>(g, visc, dxsqinv, ph, rh, res, K, (int)blockIdx.y, t);
}

template <typename T>
void launch_stag_residual(hipStream_t s, const Grid &g, int visc, T *const phi[3], T *const rhs[3],
                          T *const res[3], const StagCoef<T> &K) {
  T dxsqinv = (T)(1.0 / (g.dx * g.dx));
  LAUNCH(k_stag_residual<T>, dim3(nblk(g.N), g.dim), dim3(TPB), 0, s, g, visc, dxsqinv, p3(phi),
                     p3(rhs), p3(res), K);
}

// ------------------------------------------------------------------------
// restriction / prolongation (stag_mg_solver.f90:873-1633)
// ------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256) k_stag_restrict(Grid gc, Grid gf, int simple, Ptr3<T> fi, Ptr3<T> co) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= gc.N) return;
  restrict_at<T>(gc, gf, simple, fi, co, (int)blockIdx.y, t);
}

template <typename T>
void launch_stag_restrict(hipStream_t s, const Grid &gc, T *const fine[3], T *const coarse[3], int simple) {
  Grid gf = gc;
  for (int d = 0; d < gc.dim; d++) gf.n[d] = gc.n[d] * 2;
  gf = make_grid(gc.dim, gf.n, gc.dx * 0.5);
  LAUNCH(k_stag_restrict<T>, dim3(nblk(gc.N), gc.dim), dim3(TPB), 0, s, gc, gf, simple, p3(fine),
                     p3(coarse));
}

template <typename T, int D, int A>
__global__ void __launch_bounds__(256) k_stag_prolong(Grid gf, Grid gc, T *fine, const T *__restrict__ C) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= gf.N) return;
  prolong_at<T, D, A>(gf, gc, fine, C, t);
}

// Two x-neighbours (2m, 2m+1) per thread: one vector load and store of the
// fine pair; each point gets prolong_value as in the one-point kernel, and
// for the x-face component the even and odd point no longer sit in
// alternate lanes of a wave.
template <typename T, int A>
__global__ void __launch_bounds__(256) k_stag_prolong2(Grid gf, Grid gc, T *fine, const T *__restrict__ C) {
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (2 * t >= gf.N) return;
  int c0[3];
  decode(gf, 2 * t, c0[0], c0[1], c0[2]);
  const int c1[3] = {c0[0] + 1, c0[1], c0[2]};
  const long p = 2 * t;
  T f0, f1;
  ld2(fine, p, f0, f1);
  f0 = prolong_value<T, 3, A>(gf, gc, C, c0, f0);
  f1 = prolong_value<T, 3, A>(gf, gc, C, c1, f1);
  st2(fine, p, f0, f1);
}

// Tiled prolongation of the three face components (3D; fine level with x a
// multiple of 64, y and z of 4): a workgroup owns a 64 x 4 x 4 block of fine
// points of one component, stages the coarse values the block reads -- its
// 32 x 2 x 2 coarse points, one more along the component's direction and one
// on each side across it -- in LDS once, and each thread prolongs two pairs
// of fine points from there with prolong_combine, the arithmetic of the
// one-point kernel, so the same bits with one coarse load per staged value
// instead of 4 or 8 per fine point.
constexpr int PTX = 64, PTY = 4, PTZ = 4;

template <typename T, int A>
__device__ __forceinline__ void prolong_tile(const Grid &gf, const Grid &gc, T *fine, const T *__restrict__ C,
                                             long tile, int ntx, int nty, T (*S)[PTY][PTX / 2 + 2]) {
  constexpr int T1 = (A == 0 ? 1 : 0), T2 = (A == 2 ? 1 : 2);
  // the box: along A from the tile's first coarse point, one extra above;
  // across A one extra on each side
  constexpr int LX0 = (A == 0) ? 0 : -1, LY0 = (A == 1) ? 0 : -1, LZ0 = (A == 2) ? 0 : -1;
  constexpr int SX = PTX / 2 + (A == 0 ? 1 : 2), SY = PTY / 2 + (A == 1 ? 1 : 2), SZ = PTZ / 2 + (A == 2 ? 1 : 2);
  const int bx = (int)(tile % ntx), by = (int)((tile / ntx) % nty), bz = (int)(tile / ((long)ntx * nty));
  const int i0 = bx * PTX, j0 = by * PTY, k0 = bz * PTZ;
  const int ox = (i0 >> 1) + LX0, oy = (j0 >> 1) + LY0, oz = (k0 >> 1) + LZ0;
  const int tid = threadIdx.x;
  for (int q = tid; q < SX * SY * SZ; q += 256) {
    const int x = q % SX, y = (q / SX) % SY, z = q / (SX * SY);
    S[z][y][x] = C[wrapc(ox + x, gc.n[0]) + gc.s[1] * wrapc(oy + y, gc.n[1]) + gc.s[2] * wrapc(oz + z, gc.n[2])];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 2; r++) {
    // 512 pairs per block: 32 per row, 4 rows, 4 planes
    const int pidx = tid + 256 * r;
    const int m = pidx & 31, yy = (pidx >> 5) & 3, zz = pidx >> 7;
    const long p = (i0 + 2 * m) + gf.s[1] * (j0 + yy) + gf.s[2] * (k0 + zz);
    T f[2];
    ld2(fine, p, f[0], f[1]);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int c[3] = {i0 + 2 * m + h, j0 + yy, k0 + zz};
      const int lq[3] = {(c[0] >> 1) - ox, (c[1] >> 1) - oy, (c[2] >> 1) - oz};
      f[h] = prolong_combine<T, 3, A>(c, f[h], [&](int sa, int s1, int s2) -> T {
        int q[3] = {lq[0], lq[1], lq[2]};
        q[A] += sa;
        q[T1] += s1;
        q[T2] += s2;
        return S[q[2]][q[1]][q[0]];
      });
    }
    st2(fine, p, f[0], f[1]);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_stag_prolong_tiled(Grid gf, Grid gc, Ptr3<T> fine, Ptr3<T> coarse, int ntx,
                                                            int nty) {
  __shared__ T S[PTZ][PTY][PTX / 2 + 2];
  const long tile = xcd_block();
  const int a = (int)blockIdx.y;
  if (a == 0) prolong_tile<T, 0>(gf, gc, fine.c[0], coarse.c[0], tile, ntx, nty, S);
  else if (a == 1) prolong_tile<T, 1>(gf, gc, fine.c[1], coarse.c[1], tile, ntx, nty, S);
  else prolong_tile<T, 2>(gf, gc, fine.c[2], coarse.c[2], tile, ntx, nty, S);
}

template <typename T>
void launch_stag_prolong(hipStream_t s, const Grid &gf, T *const fine[3], T *const coarse[3]) {
  Grid gc = coarsen(gf);
  dim3 grid(nblk(gf.N));
  if (gf.dim == 3 && tuning().prolong_tiled && gf.n[0] % PTX == 0 && gf.n[1] % PTY == 0 && gf.n[2] % PTZ == 0) {
    const int ntx = gf.n[0] / PTX, nty = gf.n[1] / PTY, ntz = gf.n[2] / PTZ;
    Ptr3<T> F, Cc;
    for (int d = 0; d < 3; d++) {
      F.c[d] = fine[d];
      Cc.c[d] = coarse[d];
    }
    LAUNCH(k_stag_prolong_tiled<T>, dim3(ntx * nty * ntz, 3), dim3(256), 0, s, gf, gc, F, Cc, ntx, nty);
    return;
  }
  if (gf.dim == 3 && gf.n[0] % 2 == 0) {
    dim3 grid2(nblk(gf.N / 2));
    LAUNCH((k_stag_prolong2<T, 0>), grid2, dim3(TPB), 0, s, gf, gc, fine[0], coarse[0]);
    LAUNCH((k_stag_prolong2<T, 1>), grid2, dim3(TPB), 0, s, gf, gc, fine[1], coarse[1]);
    LAUNCH((k_stag_prolong2<T, 2>), grid2, dim3(TPB), 0, s, gf, gc, fine[2], coarse[2]);
    return;
  }
  if (gf.dim == 2) {
    LAUNCH((k_stag_prolong<T, 2, 0>), grid, dim3(TPB), 0, s, gf, gc, fine[0], coarse[0]);
    LAUNCH((k_stag_prolong<T, 2, 1>), grid, dim3(TPB), 0, s, gf, gc, fine[1], coarse[1]);
  } else {
    LAUNCH((k_stag_prolong<T, 3, 0>), grid, dim3(TPB), 0, s, gf, gc, fine[0], coarse[0]);
    LAUNCH((k_stag_prolong<T, 3, 1>), grid, dim3(TPB), 0, s, gf, gc, fine[1], coarse[1]);
    LAUNCH((k_stag_prolong<T, 3, 2>), grid, dim3(TPB), 0, s, gf, gc, fine[2], coarse[2]);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_cc_restrict(Grid gc, Grid gf, const T *__restrict__ f, T *c) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= gc.N) return;
  cc_restrict_at<T>(gc, gf, f, c, t);
}

template <typename T>
void launch_cc_restrict(hipStream_t s, const Grid &gc, const T *fine, T *coarse) {
  int n[3] = {gc.n[0] * 2, gc.n[1] * 2, gc.n[2] * 2};
  Grid gf = make_grid(gc.dim, n, gc.dx * 0.5);
  LAUNCH(k_cc_restrict<T>, dim3(nblk(gc.N)), dim3(TPB), 0, s, gc, gf, fine, coarse);
}

// Edge e is averaged along its own direction: xy edges along z, xz along y,
// yz along x.  The fine-grid stride of that direction comes in as data
// (along.o[e]) so the kernel has no branch on the edge type.
struct Along3 {
  long o[3];
};

template <typename T>
__global__ void __launch_bounds__(256) k_edge_restrict(Grid gc, Grid gf, Along3 along, Ptr3<T> fi, Ptr3<T> co) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= gc.N) return;
  const int e = (int)blockIdx.y;
  int i, j, k;
  decode(gc, t, i, j, k);
  long b = 2 * i + gf.s[1] * (2 * j) + gf.s[2] * (2 * k);
  const T *f = fi.c[e];
  co.c[e][t] = (T)0.5 * (f[b] + f[b + along.o[e]]);
}

template <typename T>
void launch_edge_restrict(hipStream_t s, const Grid &gc, T *const fine[3], T *const coarse[3]) {
  int n[3] = {gc.n[0] * 2, gc.n[1] * 2, gc.n[2] * 2};
  Grid gf = make_grid(3, n, gc.dx * 0.5);
  Along3 along = {{gf.s[2], gf.s[1], 1}};
  LAUNCH(k_edge_restrict<T>, dim3(nblk(gc.N), 3), dim3(TPB), 0, s, gc, gf, along, p3(fine), p3(coarse));
}

template <typename T>
__global__ void __launch_bounds__(256) k_nodal_restrict(Grid gc, Grid gf, const T *__restrict__ f, T *c) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= gc.N) return;
  int i, j, k;
  decode(gc, t, i, j, k);
  c[t] = f[2 * i + gf.s[1] * (2 * j)];
}

template <typename T>
void launch_nodal_restrict(hipStream_t s, const Grid &gc, const T *fine, T *coarse) {
  int n[3] = {gc.n[0] * 2, gc.n[1] * 2, 1};
  Grid gf = make_grid(2, n, gc.dx * 0.5);
  LAUNCH(k_nodal_restrict<T>, dim3(nblk(gc.N)), dim3(TPB), 0, s, gc, gf, fine, coarse);
}

// ------------------------------------------------------------------------
// cell-centred multigrid: div(a grad phi) = rhs
//   (L phi)(p) = ( -(sum of the 2*dim face coefficients) phi(p)
//                  + a_lo,x phi(p-x) + a_hi,x phi(p+x) + ... ) / dx^2
// ------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256) k_cc_rb(Grid g, int colour, T inv, T *phi, const T *__restrict__ rhs,
                                               Ptr3<T> A) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= (long)(g.n[0] >> 1) * g.n[1] * g.n[2]) return;
  cc_rb_at<T>(g, colour, inv, phi, rhs, A, t);
}

template <typename T>
void launch_cc_rb_colour(hipStream_t s, const Grid &g, int colour, T *phi, const T *rhs, T *const a[3]) {
  long total = (long)(g.n[0] / 2) * g.n[1] * g.n[2];
  T inv = (T)(1.0 / (g.dx * g.dx));
  LAUNCH(k_cc_rb<T>, dim3(nblk(total)), dim3(TPB), 0, s, g, colour, inv, phi, rhs, p3(a));
}

template <typename T>
__global__ void __launch_bounds__(256) k_cc_residual(Grid g, T inv, const T *phi, const T *__restrict__ rhs,
                                                     T *res, Ptr3<T> A) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  cc_residual_at<T>(g, inv, phi, rhs, res, A, t);
}

template <typename T>
void launch_cc_residual(hipStream_t s, const Grid &g, const T *phi, const T *rhs, T *res, T *const a[3]) {
  T inv = (T)(1.0 / (g.dx * g.dx));
  LAUNCH(k_cc_residual<T>, dim3(nblk(g.N)), dim3(TPB), 0, s, g, inv, phi, rhs, res, p3(a));
}

// Residual of the cell-centred level and its average onto the next coarser
// level in one pass; the fine residual is never stored.  A workgroup covers
// a run of 64 fine cells in x on the 2 (2D) or 2x2 (3D) fine rows of one
// coarse row: each thread forms one fine residual (64-wide coalesced rows, as
// the residual pass alone reads them) into LDS, then one thread per coarse
// cell sums its 2^dim children.  The fine residuals and the sum are the ones
// cc_residual_at / cc_restrict_at form, in the same order, so the coarse
// right-hand side has the same bits as the two-pass form
// (stag_mg_solver.f90:873-968 restriction of the residual).
template <typename T, int DIM>
__global__ void __launch_bounds__(256) k_cc_resid_restrict(Grid gc, Grid gf, T inv, const T *phi,
                                                           const T *__restrict__ rhs, Ptr3<T> A, T *c) {
  __shared__ T sr[DIM == 3 ? 2 : 1][2][64];
  const int tiles_x = (gf.n[0] + 63) >> 6;
  const long w = xcd_block();
  const int tx = (int)(w % tiles_x);
  const long row = w / tiles_x;  // coarse row: cj + ny_c * ck
  const int cj = (int)(row % gc.n[1]), ck = (int)(row / gc.n[1]);
  const int x = threadIdx.x & 63, dy = (threadIdx.x >> 6) & 1, dz = (DIM == 3) ? (threadIdx.x >> 7) : 0;
  const int fi = tx * 64 + x, fj = 2 * cj + dy, fk = (DIM == 3) ? 2 * ck + dz : 0;
  if (fi < gf.n[0]) {
    const long p = fi + gf.s[1] * fj + gf.s[2] * fk;
    T diag;
    sr[dz][dy][x] = rhs[p] - cc_apply<T>(gf, fi, fj, fk, phi, A, inv, diag);
  }
  __syncthreads();
  const int ci = tx * 32 + (int)threadIdx.x;
  if (threadIdx.x >= 32 || ci >= gc.n[0]) return;
  const int l = 2 * (int)threadIdx.x;
  const long t = ci + gc.s[1] * cj + gc.s[2] * ck;
  if (DIM == 2) {
    c[t] = (T)0.25 * (sr[0][0][l] + sr[0][0][l + 1] + sr[0][1][l] + sr[0][1][l + 1]);
  } else {
    c[t] = (T)0.125 * (sr[0][0][l] + sr[0][0][l + 1] + sr[0][1][l] + sr[0][1][l + 1] + sr[1][0][l] +
                       sr[1][0][l + 1] + sr[1][1][l] + sr[1][1][l + 1]);
  }
}

template <typename T>
void launch_cc_resid_restrict(hipStream_t s, const Grid &gf, const T *phi, const T *rhs, T *const a[3],
                              const Grid &gc, T *coarse_rhs) {
  for (int d = 0; d < gf.dim; d++)
    if (gf.n[d] != 2 * gc.n[d]) {
      report_error("cc residual-restriction: fine extent is not twice the coarse one");
      return;
    }
  T inv = (T)(1.0 / (gf.dx * gf.dx));
  // one workgroup per 64-cell run of each coarse row
  const long nwg = (long)((gf.n[0] + 63) / 64) * gc.n[1] * (gf.dim == 3 ? gc.n[2] : 1);
  if (gf.dim == 2)
    LAUNCH((k_cc_resid_restrict<T, 2>), dim3((unsigned)nwg), dim3(128), 0, s, gc, gf, inv, phi, rhs, p3(a),
           coarse_rhs);
  else
    LAUNCH((k_cc_resid_restrict<T, 3>), dim3((unsigned)nwg), dim3(256), 0, s, gc, gf, inv, phi, rhs, p3(a),
           coarse_rhs);
}

template <typename T>
__global__ void __launch_bounds__(256) k_cc_prolong_add(Grid gf, Grid gc, T *f, const T *__restrict__ c) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= gf.N) return;
  cc_prolong_add_at<T>(gf, gc, f, c, t);
}

template <typename T>
void launch_cc_prolong_add(hipStream_t s, const Grid &gf, T *fine, const T *coarse) {
  Grid gc = coarsen(gf);
  LAUNCH(k_cc_prolong_add<T>, dim3(nblk(gf.N)), dim3(TPB), 0, s, gf, gc, fine, coarse);
}

// ------------------------------------------------------------------------
// elementwise
// ------------------------------------------------------------------------

template <typename T>
__global__ void k_fill(T *x, long n, T v) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}
template <typename T>
void launch_fill(hipStream_t s, T *x, long n, T v) {
  if (n > 0) LAUNCH(k_fill<T>, dim3(nblk(n)), dim3(TPB), 0, s, x, n, v);
}

template <typename T>
__global__ void k_scale(T *x, long n, T v) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = x[t] * v;
}
template <typename T>
void launch_scale(hipStream_t s, T *x, long n, T v) {
  if (n > 0) LAUNCH(k_scale<T>, dim3(nblk(n)), dim3(TPB), 0, s, x, n, v);
}

template <typename TD, typename TS>
__global__ void k_convert(TD *d, const TS *__restrict__ src, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) d[t] = (TD)src[t];
}
template <typename TD, typename TS>
void launch_convert(hipStream_t s, TD *dst, const TS *src, long n) {
  if (n > 0) LAUNCH((k_convert<TD, TS>), dim3(nblk(n)), dim3(TPB), 0, s, dst, src, n);
}

__global__ void k_recip(double *d, const double *__restrict__ src, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) d[t] = 1.0 / src[t];
}
void launch_reciprocal(hipStream_t s, double *dst, const double *src, long n) {
  LAUNCH(k_recip, dim3(nblk(n)), dim3(TPB), 0, s, dst, src, n);
}

__global__ void k_scale_copy(double *d, const double *__restrict__ src, long n, double v) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) d[t] = src[t] * v;
}
void launch_scale_copy(hipStream_t s, double *dst, const double *src, long n, double v) {
  LAUNCH(k_scale_copy, dim3(nblk(n)), dim3(TPB), 0, s, dst, src, n, v);
}

// ------------------------------------------------------------------------
// Stokes operator (apply_matrix.f90:60-112), fused per cell:
//   b_u[a] = (theta alpha - L) x_u + (p - p(-e_a))/dx ;  b_p = -(D x_u)
// ------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256) k_apply_matrix(Grid g, int visc, T dxsqinv, T dxinv, Ptr3<T> xu,
                                                      const T *__restrict__ xp, Ptr3<T> bu, T *bp,
                                                      StagCoef<T> K) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  Pt P = make_pt(g, i, j, k);
  const T *phr[3] = {xu.c[0], xu.c[1], xu.c[2]};
  for (int a = 0; a < g.dim; a++) {
    T L, fac;
    stag_point<T>(visc, g, a, P, phr, K, dxsqinv, L, fac);
    long pm = at_sh(g, P, a, -1, -1, 0);
    T gx = (xp[t] - xp[pm]) * dxinv;
    bu.c[a][t] = L + gx;
  }
  T dv = (xu.c[0][at3(g, P, 1, 0, 0)] - xu.c[0][t]) * dxinv;
  dv = dv + (xu.c[1][at3(g, P, 0, 1, 0)] - xu.c[1][t]) * dxinv;
  if (g.dim == 3) dv = dv + (xu.c[2][at3(g, P, 0, 0, 1)] - xu.c[2][t]) * dxinv;
  bp[t] = dv * (T)-1.0;
}

// 3D visc_type -1 / 1: the same arithmetic with every index static (the
// stencil policy of the fused smoother supplies L at a point).
template <typename T, int VISC, typename TX>
__device__ __forceinline__ T am_component(const Grid &g, int i, int j, int k, const Ptr3<TX> &xu,
                                          const typename StagRB<T, 0, VISC>::Args &a0,
                                          const typename StagRB<T, 1, VISC>::Args &a1,
                                          const typename StagRB<T, 2, VISC>::Args &a2, int comp, T dxsqinv) {
  const long sy = g.s[1], sz = g.s[2];
  const long p = i + sy * j + sz * k;
  const int ip = wrapc(i + 1, g.n[0]), im = wrapc(i - 1, g.n[0]);
  const int jp = wrapc(j + 1, g.n[1]), jm = wrapc(j - 1, g.n[1]);
  const int kp = wrapc(k + 1, g.n[2]), km = wrapc(k - 1, g.n[2]);
  const long ro = sy * j + sz * k;
  // TX: storage type of the operand (FP32 preconditioned vectors of the
  // flexible GMRES are widened exactly on load)
  auto nb = [&](const TX *F, T &xp, T &xm, T &yp, T &ym, T &zp, T &zm) {
    xp = (T)F[ip + ro];
    xm = (T)F[im + ro];
    yp = (T)F[i + sy * jp + sz * k];
    ym = (T)F[i + sy * jm + sz * k];
    zp = (T)F[i + sy * j + sz * kp];
    zm = (T)F[i + sy * j + sz * km];
  };
  T xp, xm, yp, ym, zp, zm;
  if (comp == 0) {
    nb(xu.c[0], xp, xm, yp, ym, zp, zm);
    return StagRB<T, 0, VISC>::apply(StagRB<T, 0, VISC>::load(g, i, j, k, a0), (T)xu.c[0][p], xp, xm, yp, ym, zp,
                                     zm, dxsqinv);
  } else if (comp == 1) {
    nb(xu.c[1], xp, xm, yp, ym, zp, zm);
    return StagRB<T, 1, VISC>::apply(StagRB<T, 1, VISC>::load(g, i, j, k, a1), (T)xu.c[1][p], xp, xm, yp, ym, zp,
                                     zm, dxsqinv);
  }
  nb(xu.c[2], xp, xm, yp, ym, zp, zm);
  return StagRB<T, 2, VISC>::apply(StagRB<T, 2, VISC>::load(g, i, j, k, a2), (T)xu.c[2][p], xp, xm, yp, ym, zp, zm,
                                   dxsqinv);
}

// Slab depth (z planes) of the operator's brick order: the largest of the
// configured depth and its halvings that divides nz; 0 (row order) when x is not a
// multiple of 64 or y of 4, or when brick order is switched off.
static int am_brick_depth(const Grid &g) {
  if (tuning().am_dz <= 0 || g.dim != 3 || g.n[0] % 64 != 0 || g.n[1] % 4 != 0) return 0;
  for (int d = tuning().am_dz; d >= 1; d /= 2)
    if (g.n[2] % d == 0) return d;
  return 1;
}

template <typename T, int VISC, typename TX, bool RES>
__global__ void __launch_bounds__(256) k_apply_matrix3(Grid g, T dxsqinv, T dxinv, Ptr3<TX> xu,
                                                       const TX *__restrict__ xp, Ptr3<T> rbu, const T *rbp,
                                                       Ptr3<T> bu, T *bp, typename StagRB<T, 0, VISC>::Args a0,
                                                       typename StagRB<T, 1, VISC>::Args a1,
                                                       typename StagRB<T, 2, VISC>::Args a2,
                                                       const T *__restrict__ al0, const T *__restrict__ al1,
                                                       const T *__restrict__ al2, const T *__restrict__ be,
                                                       const T *__restrict__ ed0, const T *__restrict__ ed1,
                                                       const T *__restrict__ ed2, int dz) {
  // rhs is not read by apply(); it gets a valid address of the right type
  a0 = StagRB<T, 0, VISC>::bind(a0, be, al0, be, ed0, ed1);
  a1 = StagRB<T, 1, VISC>::bind(a1, be, al1, be, ed0, ed2);
  a2 = StagRB<T, 2, VISC>::bind(a2, be, al2, be, ed1, ed2);
  const long sy = g.s[1], sz = g.s[2];
  long t;
  int i, j, k;
  if (dz > 0) {
    // bricks of 64 x 4 x 1 cells; consecutive workgroups go x-brick, then z
    // within a slab of dz planes, then y, then slab, so the workgroups in
    // flight on one XCD cover a compact box whose neighbours are in flight
    // too (row order keeps a whole plane of every array in flight, more than
    // that XCD's L2 holds, and re-reads the z-neighbour planes from HBM)
    const unsigned w = (unsigned)xcd_block(), bx = (unsigned)g.n[0] >> 6, by = (unsigned)g.n[1] >> 2;
    const unsigned r = w / bx, r2 = r / (unsigned)dz, zs = r2 / by;
    i = (int)((w - r * bx) << 6) + (int)(threadIdx.x & 63);
    j = (int)((r2 - zs * by) << 2) + (int)(threadIdx.x >> 6);
    k = (int)(zs * (unsigned)dz + (r - r2 * (unsigned)dz));
    t = i + sy * j + sz * k;
  } else {
    t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    decode(g, t, i, j, k);
  }
  const long pmx = wrapc(i - 1, g.n[0]) + sy * j + sz * k;
  const long pmy = i + sy * wrapc(j - 1, g.n[1]) + sz * k;
  const long pmz = i + sy * j + sz * wrapc(k - 1, g.n[2]);
  const T pc = (T)xp[t];
  // RES: write (Ax - b) * -1 = b - Ax instead of Ax, as k_bmax_neg does to a
  // stored Ax, so the same bits
  auto out = [&](T v, T bv) { return RES ? (v - bv) * (T)-1.0 : v; };
  bu.c[0][t] = out(am_component<T, VISC, TX>(g, i, j, k, xu, a0, a1, a2, 0, dxsqinv) + (pc - (T)xp[pmx]) * dxinv,
                   RES ? rbu.c[0][t] : (T)0);
  bu.c[1][t] = out(am_component<T, VISC, TX>(g, i, j, k, xu, a0, a1, a2, 1, dxsqinv) + (pc - (T)xp[pmy]) * dxinv,
                   RES ? rbu.c[1][t] : (T)0);
  bu.c[2][t] = out(am_component<T, VISC, TX>(g, i, j, k, xu, a0, a1, a2, 2, dxsqinv) + (pc - (T)xp[pmz]) * dxinv,
                   RES ? rbu.c[2][t] : (T)0);
  const long xq = wrapc(i + 1, g.n[0]) + sy * j + sz * k;
  const long yq = i + sy * wrapc(j + 1, g.n[1]) + sz * k;
  const long zq = i + sy * j + sz * wrapc(k + 1, g.n[2]);
  T dv = ((T)xu.c[0][xq] - (T)xu.c[0][t]) * dxinv;
  dv = dv + ((T)xu.c[1][yq] - (T)xu.c[1][t]) * dxinv;
  dv = dv + ((T)xu.c[2][zq] - (T)xu.c[2][t]) * dxinv;
  bp[t] = out(dv * (T)-1.0, RES ? rbp[t] : (T)0);
}

template <typename T, int VISC, typename TX>
static void launch_am3(hipStream_t s, const Grid &g, T dxsqinv, T dxinv, TX *const x_u[3], const TX *x_p,
                       T *const b_u[3], T *b_p, const StagCoef<T> &K, T *const rb_u[3] = nullptr,
                       const T *rb_p = nullptr) {
  typename StagRB<T, 0, VISC>::Args a0;
  typename StagRB<T, 1, VISC>::Args a1;
  typename StagRB<T, 2, VISC>::Args a2;
  a0.omega = a1.omega = a2.omega = (T)0;
  a0.dxsqinv = a1.dxsqinv = a2.dxsqinv = dxsqinv;
  a0.rhs = a1.rhs = a2.rhs = K.beta;  // unused by apply(); any valid address
  a0.alpha = K.alpha[0];
  a1.alpha = K.alpha[1];
  a2.alpha = K.alpha[2];
  a0.beta = a1.beta = a2.beta = K.beta;
  // edges shared with (A, t1) and (A, t2): xy = 0, xz = 1, yz = 2
  a0.e1 = K.bed[0];
  a0.e2 = K.bed[1];
  a1.e1 = K.bed[0];
  a1.e2 = K.bed[2];
  a2.e1 = K.bed[1];
  a2.e2 = K.bed[2];
  const int dz = am_brick_depth(g);
  if (rb_u) {
    LAUNCH((k_apply_matrix3<T, VISC, TX, true>), dim3(nblk(g.N)), dim3(TPB), 0, s, g, dxsqinv, dxinv, p3(x_u), x_p,
           p3(rb_u), rb_p, p3(b_u), b_p, a0, a1, a2, K.alpha[0], K.alpha[1], K.alpha[2], K.beta, K.bed[0], K.bed[1],
           K.bed[2], dz);
    return;
  }
  LAUNCH((k_apply_matrix3<T, VISC, TX, false>), dim3(nblk(g.N)), dim3(TPB), 0, s, g, dxsqinv, dxinv, p3(x_u), x_p,
         p3(b_u), (const T *)nullptr, p3(b_u), b_p, a0, a1, a2, K.alpha[0], K.alpha[1], K.alpha[2], K.beta, K.bed[0],
         K.bed[1], K.bed[2], dz);
}

// r = b - A x (FP64, 3D visc_type -1 / 1) in one pass: the operator's output
// goes through k_bmax_neg's arithmetic before it is stored; false elsewhere
bool launch_apply_matrix_resid(hipStream_t s, const Grid &g, int visc, double *const x_u[3], const double *x_p,
                               double *const b_u[3], const double *b_p, double *const r_u[3], double *r_p,
                               const StagCoef<double> &K) {
  const double dxsqinv = 1.0 / (g.dx * g.dx), dxinv = 1.0 / g.dx;
  if (g.dim == 3 && visc == -1) {
    launch_am3<double, -1, double>(s, g, dxsqinv, dxinv, x_u, x_p, r_u, r_p, K, b_u, b_p);
    return true;
  }
  if (g.dim == 3 && visc == 1) {
    launch_am3<double, 1, double>(s, g, dxsqinv, dxinv, x_u, x_p, r_u, r_p, K, b_u, b_p);
    return true;
  }
  return false;
}

// A x with x held in FP32 (the flexible GMRES's preconditioned vectors) and
// the operator applied in FP64: every value is widened exactly on load, so
// the result equals converting x to FP64 first.  3D visc_type -1 and 1 only;
// returns false (nothing launched) elsewhere.
bool launch_apply_matrix_x32(hipStream_t s, const Grid &g, int visc, float *const x_u[3], const float *x_p,
                             double *const b_u[3], double *b_p, const StagCoef<double> &K) {
  const double dxsqinv = 1.0 / (g.dx * g.dx), dxinv = 1.0 / g.dx;
  if (g.dim == 3 && visc == -1) {
    launch_am3<double, -1, float>(s, g, dxsqinv, dxinv, x_u, x_p, b_u, b_p, K);
    return true;
  }
  if (g.dim == 3 && visc == 1) {
    launch_am3<double, 1, float>(s, g, dxsqinv, dxinv, x_u, x_p, b_u, b_p, K);
    return true;
  }
  return false;
}

template <typename T>
void launch_apply_matrix(hipStream_t s, const Grid &g, int visc, T *const x_u[3], const T *x_p,
                         T *const b_u[3], T *b_p, const StagCoef<T> &K) {
  T dxsqinv = (T)(1.0 / (g.dx * g.dx)), dxinv = (T)(1.0 / g.dx);
  if (g.dim == 3 && visc == -1) {
    launch_am3<T, -1, T>(s, g, dxsqinv, dxinv, x_u, x_p, b_u, b_p, K);
    return;
  }
  if (g.dim == 3 && visc == 1) {
    launch_am3<T, 1, T>(s, g, dxsqinv, dxinv, x_u, x_p, b_u, b_p, K);
    return;
  }
  LAUNCH(k_apply_matrix<T>, dim3(nblk(g.N)), dim3(TPB), 0, s, g, visc, dxsqinv, dxinv, p3(x_u), x_p,
                     p3(b_u), b_p, K);
}

// div kernel (div_and_grad.f90:400-451) with combination modes
template <typename T>
__global__ void __launch_bounds__(256) k_div(Grid g, T dxinv, Ptr3<T> u, const T *b_p, T *div, int mode) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  long xp = wp(i, g.n[0]) + g.s[1] * j + g.s[2] * k;
  long yp = i + g.s[1] * wp(j, g.n[1]) + g.s[2] * k;
  T v = (u.c[0][xp] - u.c[0][t]) * dxinv;
  if (mode == 3) v = div[t] + v;
  v = v + (u.c[1][yp] - u.c[1][t]) * dxinv;
  if (g.dim == 3) {
    long zp = i + g.s[1] * j + g.s[2] * wp(k, g.n[2]);
    v = v + (u.c[2][zp] - u.c[2][t]) * dxinv;
  }
  if (mode == 1) v = v + b_p[t];
  if (mode == 2) v = v * (T)-1.0;
  div[t] = v;
}

// k_div on an x-pair of cells per thread (x extents are even on every level
// the library builds, so a pair never leaves its row): the pair's own face
// values, the y / z neighbour rows, b_p and div move as 2-wide loads and
// stores. Each cell's value is the same expression in the same order as
// k_div's, so the same bits.
template <typename T>
__global__ void __launch_bounds__(256) k_div2(Grid g, T dxinv, Ptr3<T> u, const T *b_p, T *div, int mode) {
  const long t = 2 * (xcd_block() * blockDim.x + threadIdx.x);
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  // x neighbour of cell i is i + 1 (i is even); of cell i + 1 it wraps
  const long x1 = wp(i + 1, g.n[0]) + g.s[1] * j + g.s[2] * k;
  const long yr = i + g.s[1] * wp(j, g.n[1]) + g.s[2] * k;
  T a0, a1, n0, n1;
  ld2(u.c[0], t, a0, a1);
  T v0 = (a1 - a0) * dxinv;
  T v1 = (u.c[0][x1] - a1) * dxinv;
  if (mode == 3) {
    T d0, d1;
    ld2((const T *)div, t, d0, d1);
    v0 = d0 + v0;
    v1 = d1 + v1;
  }
  ld2(u.c[1], t, a0, a1);
  ld2(u.c[1], yr, n0, n1);
  v0 = v0 + (n0 - a0) * dxinv;
  v1 = v1 + (n1 - a1) * dxinv;
  if (g.dim == 3) {
    const long zr = i + g.s[1] * j + g.s[2] * wp(k, g.n[2]);
    ld2(u.c[2], t, a0, a1);
    ld2(u.c[2], zr, n0, n1);
    v0 = v0 + (n0 - a0) * dxinv;
    v1 = v1 + (n1 - a1) * dxinv;
  }
  if (mode == 1) {
    T b0, b1;
    ld2(b_p, t, b0, b1);
    v0 = v0 + b0;
    v1 = v1 + b1;
  }
  if (mode == 2) {
    v0 = v0 * (T)-1.0;
    v1 = v1 * (T)-1.0;
  }
  st2(div, t, v0, v1);
}

template <typename T>
void launch_div_plus(hipStream_t s, const Grid &g, T *const u[3], const T *b_p, T *div, int mode) {
  T dxinv = (T)(1.0 / g.dx);
  if (g.n[0] % 2 == 0) {
    LAUNCH(k_div2<T>, dim3(nblk(g.N / 2)), dim3(TPB), 0, s, g, dxinv, p3(u), b_p, div, mode);
    return;
  }
  LAUNCH(k_div<T>, dim3(nblk(g.N)), dim3(TPB), 0, s, g, dxinv, p3(u), b_p, div, mode);
}

// Block sum into partial[blockIdx * 4 + seg] (seg fixed per call site).
__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; w++) r += sh[w];
  }
  __syncthreads();
  return r;
}

// P1 steps 3-4 (apply_precon.f90:131-188, macproject.f90:320-364), fused per cell:
//   x_u[a] = x_u[a] + (-1)*(((phi - phi(-e_a))/dx) * ainv[a])
//   x_p    = -(mac_rhs)*beta*c + theta*phi      (c per |visc_type|)
// and the per-block partial sums of x_p for the pressure mean.
template <typename T>
__global__ void __launch_bounds__(256) k_precon_final(Grid g, int visc, T dxinv, Ptr3<T> xu, T *xp,
                                                      const T *__restrict__ phi, const T *__restrict__ mrhs,
                                                      Ptr3<T> ainv, const T *__restrict__ beta,
                                                      const T *__restrict__ gamma, T theta, double *partial) {
  __shared__ double sh[8];
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double mine = 0.0;
  if (t < g.N) {
    int i, j, k;
    decode(g, t, i, j, k);
    Pt P = make_pt(g, i, j, k);
    for (int a = 0; a < g.dim; a++) {
      long pm = at_sh(g, P, a, -1, -1, 0);
      T gp = (phi[t] - phi[pm]) * dxinv;
      T gq = gp * ainv.c[a][t];
      xu.c[a][t] = xu.c[a][t] + (T)-1.0 * gq;
    }
    int av = visc < 0 ? -visc : visc;
    T x = mrhs[t] * (T)-1.0;
    if (av == 1 || av == 2) {
      x = x * beta[t];
      if (av == 2) x = x * (T)2.0;
    } else {
      T tt = x * gamma[t];
      x = x * beta[t];
      x = x * (T)(4.0 / 3.0);
      x = x + tt;
    }
    T ph = phi[t] * theta;
    x = x + ph;
    xp[t] = x;
    mine = (double)x;
  }
  if (partial) {
    double r = block_sum(mine, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
  }
}

template <typename T>
void launch_precon_final(hipStream_t s, const Grid &g, int visc, T *const x_u[3], T *x_p, const T *phi,
                         const T *mac_rhs, T *const ainv[3], const T *beta, const T *gamma,
                         double theta_alpha, double *partial) {
  T dxinv = (T)(1.0 / g.dx);
  LAUNCH(k_precon_final<T>, dim3(nblk(g.N)), dim3(TPB), 0, s, g, visc, dxinv, p3(x_u), x_p, phi,
                     mac_rhs, p3(ainv), beta, gamma, (T)theta_alpha, partial);
}

__global__ void k_grad(Grid g, double dxinv, const double *__restrict__ p, Ptr3<double> gp) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  Pt P = make_pt(g, i, j, k);
  for (int a = 0; a < g.dim; a++) gp.c[a][t] = (p[t] - p[at_sh(g, P, a, -1, -1, 0)]) * dxinv;
}
void launch_grad(hipStream_t s, const Grid &g, const double *p, double *const gp[3]) {
  LAUNCH(k_grad, dim3(nblk(g.N)), dim3(TPB), 0, s, g, 1.0 / g.dx, p, p3(gp));
}

// convert_stag.f90 interior formulas
__global__ void k_avg_face(Grid g, const double *__restrict__ cc, Ptr3<double> f) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  Pt P = make_pt(g, i, j, k);
  for (int a = 0; a < g.dim; a++) f.c[a][t] = 0.5 * (cc[t] + cc[at_sh(g, P, a, -1, -1, 0)]);
}
void launch_avg_cc_to_face(hipStream_t s, const Grid &g, const double *cc, double *const face[3]) {
  LAUNCH(k_avg_face, dim3(nblk(g.N)), dim3(TPB), 0, s, g, cc, p3(face));
}
__global__ void k_avg_node(Grid g, const double *__restrict__ cc, double *nd) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  Pt P = make_pt(g, i, j, k);
  nd[t] = 0.25 * (cc[t] + cc[at3(g, P, -1, 0, 0)] + cc[at3(g, P, 0, -1, 0)] + cc[at3(g, P, -1, -1, 0)]);
}
void launch_avg_cc_to_node(hipStream_t s, const Grid &g, const double *cc, double *node) {
  LAUNCH(k_avg_node, dim3(nblk(g.N)), dim3(TPB), 0, s, g, cc, node);
}
__global__ void k_avg_edge(Grid g, const double *__restrict__ cc, Ptr3<double> e) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  Pt P = make_pt(g, i, j, k);
  e.c[0][t] = 0.25 * (cc[t] + cc[at3(g, P, -1, 0, 0)] + cc[at3(g, P, 0, -1, 0)] + cc[at3(g, P, -1, -1, 0)]);
  e.c[1][t] = 0.25 * (cc[t] + cc[at3(g, P, -1, 0, 0)] + cc[at3(g, P, 0, 0, -1)] + cc[at3(g, P, -1, 0, -1)]);
  e.c[2][t] = 0.25 * (cc[t] + cc[at3(g, P, 0, -1, 0)] + cc[at3(g, P, 0, 0, -1)] + cc[at3(g, P, 0, -1, -1)]);
}
void launch_avg_cc_to_edge(hipStream_t s, const Grid &g, const double *cc, double *const edge[3]) {
  LAUNCH(k_avg_edge, dim3(nblk(g.N)), dim3(TPB), 0, s, g, cc, p3(edge));
}

// ------------------------------------------------------------------------
// deterministic reductions
// ------------------------------------------------------------------------

// partial[blk*4 + seg]: seg d < dim face components (face index 0 along d
// weighted 3), seg 3 = cells.  a, b: Stokes vectors [u0|u1|(u2)|p].
__global__ void __launch_bounds__(RED_TPB) k_stokes_dot(Grid g, const double *__restrict__ a,
                                                        const double *__restrict__ b, double *partial) {
  __shared__ double sh[8];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < g.N; q += stride) {
    int i, j, k;
    decode(g, q, i, j, k);
    int c[3] = {i, j, k};
    for (int d = 0; d < g.dim; d++) {
      double w = (c[d] == 0) ? 3.0 : 1.0;
      acc[d] += w * (a[d * g.N + q] * b[d * g.N + q]);
    }
    acc[3] += a[g.dim * g.N + q] * b[g.dim * g.N + q];
  }
  for (int seg = 0; seg < 4; seg++) {
    double r = block_sum(acc[seg], sh);
    if (threadIdx.x == 0) partial[blockIdx.x * 4 + seg] = r;
  }
}

void launch_stokes_dot(hipStream_t s, const Grid &g, const double *a, const double *b, double *partial) {
  LAUNCH(k_stokes_dot, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, a, b, partial);
}

__global__ void __launch_bounds__(256) k_finalize(Grid g, const double *__restrict__ partial, int nparts,
                                                  double *out, int op, double pnw) {
  __shared__ double sh[8];
  double seg[4];
  for (int q = 0; q < 4; q++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) v += partial[b * 4 + q];
    seg[q] = block_sum(v, sh);
  }
  if (threadIdx.x != 0) return;
  double su = seg[0] + seg[1];
  if (g.dim == 3) su = su + seg[2];
  double tv = (double)g.N;
  if (op == 0) {
    out[0] = su + (pnw * pnw) * seg[3];
  } else if (op == 1) {
    double nu = sqrt(su);
    double np = pnw * sqrt(seg[3]);
    out[0] = sqrt(nu * nu + np * np);
    out[1] = nu;
    out[2] = np;
  } else if (op == 2) {
    for (int q = 0; q < 4; q++) out[q] = seg[q];
  } else if (op == 3) {
    out[0] = seg[3] / tv;
  } else if (op == 4) {
    for (int q = 0; q < 3; q++) out[q] = seg[q] / tv;
  } else if (op == 5) {
    out[0] = seg[0];
  }
}

void launch_finalize(hipStream_t s, const Grid &g, const double *partial, double *out, int op, double pnw) {
  LAUNCH(k_finalize, dim3(1), dim3(256), 0, s, g, partial, RED_BLOCKS, out, op, pnw);
}

// sum of a cell field into partial[blk*4+3] (pressure mean); reads partial
// from k_precon_final (one value per block of 256 cells).
__global__ void k_fold_cells(const double *__restrict__ src, int nsrc, double *partial) {
  __shared__ double sh[8];
  // RED_BLOCKS blocks: block b folds src[b], src[b+RED_BLOCKS], ...
  double v = 0.0;
  for (int q = blockIdx.x + RED_BLOCKS * threadIdx.x; q < nsrc; q += RED_BLOCKS * blockDim.x) v += src[q];
  double r = block_sum(v, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * 4 + 0] = 0.0;
    partial[blockIdx.x * 4 + 1] = 0.0;
    partial[blockIdx.x * 4 + 2] = 0.0;
    partial[blockIdx.x * 4 + 3] = r;
  }
}

__global__ void __launch_bounds__(RED_TPB) k_face_maxabs(Grid g, Ptr3<double> r, double *partial) {
  __shared__ double sh[8];
  double m[3] = {0.0, 0.0, 0.0};
  long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < g.N; q += stride)
    for (int d = 0; d < g.dim; d++) m[d] = fmax(m[d], fabs(r.c[d][q]));
  for (int d = 0; d < 3; d++) {
    double v = m[d];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double x = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) x = fmax(x, sh[w]);
      partial[blockIdx.x * 4 + d] = x;
    }
    __syncthreads();
  }
}
void launch_face_maxabs(hipStream_t s, const Grid &g, const double *const r[3], double *partial) {
  Ptr3<double> p;
  for (int d = 0; d < 3; d++) p.c[d] = const_cast<double *>(r[d]);
  LAUNCH(k_face_maxabs, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, p, partial);
}
__global__ void k_finalize_max(const double *__restrict__ partial, double *out) {
  if (threadIdx.x >= 3) return;
  double m = 0.0;
  for (int b = 0; b < RED_BLOCKS; b++) m = fmax(m, partial[b * 4 + threadIdx.x]);
  out[threadIdx.x] = m;
}
void launch_finalize_max(hipStream_t s, const double *partial, double *out) {
  LAUNCH(k_finalize_max, dim3(1), dim3(64), 0, s, partial, out);
}

// sum_umac_press velocity part (norm_inner_product.f90:579-721): every face
// of the reference layout once (face 0 and its periodic copy hi+1: weight 2),
// plus 0.5*u on the lo and hi+1 copies in a direction where the reference's
// tile spans the whole box (weight 3).  Its tiles are BoxLib's default
// 1024000 x 8 x 8: x always spans, y and z only when the box is at most 8
// cells long there.
template <typename T>
__global__ void __launch_bounds__(RED_TPB) k_sum_faces_w(Grid g, Ptr3<T> u, double *partial) {
  __shared__ double sh[8];
  double acc[3] = {0.0, 0.0, 0.0};
  long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < g.N; q += stride) {
    int i, j, k;
    decode(g, q, i, j, k);
    int c[3] = {i, j, k};
    for (int d = 0; d < g.dim; d++) {
      const double w0 = (d == 0 || g.n[d] <= 8) ? 3.0 : 2.0;
      acc[d] += (c[d] == 0 ? w0 : 1.0) * (double)u.c[d][q];
    }
  }
  for (int seg = 0; seg < 4; seg++) {
    double r = block_sum(seg < 3 ? acc[seg] : 0.0, sh);
    if (threadIdx.x == 0) partial[blockIdx.x * 4 + seg] = r;
  }
}
template <typename T>
void launch_sum_faces_weighted(hipStream_t s, const Grid &g, T *const u[3], double *partial) {
  LAUNCH(k_sum_faces_w<T>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, p3(u), partial);
}

// ------------------------------------------------------------------------
// Krylov vector kernels
// ------------------------------------------------------------------------

__global__ void k_bmax_neg(double *tmp, const double *__restrict__ b, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) tmp[t] = (tmp[t] - b[t]) * -1.0;
}
void launch_bmax_neg(hipStream_t s, double *tmp, const double *b, long n) {
  LAUNCH(k_bmax_neg, dim3(nblk(n)), dim3(TPB), 0, s, tmp, b, n);
}

__global__ void k_div_by(double *v, const double *__restrict__ src, long n, const double *scal) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = src[t] / scal[0];
}
void launch_div_by(hipStream_t s, double *v, const double *src, long n, const double *scal) {
  LAUNCH(k_div_by, dim3(nblk(n)), dim3(TPB), 0, s, v, src, n, scal);
}

__global__ void k_axmy(double *w, const double *__restrict__ v, long n, const double *scal) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) w[t] = w[t] - v[t] * scal[0];
}
void launch_axmy(hipStream_t s, double *w, const double *v, long n, const double *scal) {
  LAUNCH(k_axmy, dim3(nblk(n)), dim3(TPB), 0, s, w, v, n, scal);
}

// One modified Gram-Schmidt step fused with the next reduction: w -= h*v
// (k_axmy's arithmetic), then the weighted dot of the new w with vn (the next
// H entry) or, NORM, with itself (|w|^2 for H(i+1,i)).  Cells are visited in
// k_stokes_dot's grid-stride order and summed the same way, so the partial
// sums are bit-identical to an axmy pass followed by a dot pass, while w is
// read and written once.
template <bool NORM>
__global__ void __launch_bounds__(RED_TPB) k_axmy_dot(Grid g, double *w, const double *__restrict__ v,
                                                      const double *__restrict__ h, const double *__restrict__ vn,
                                                      double *partial) {
  __shared__ double sh[8];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double hv = h[0];
  long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < g.N; q += stride) {
    int i, j, k;
    decode(g, q, i, j, k);
    int c[3] = {i, j, k};
    for (int d = 0; d < g.dim; d++) {
      const long p = d * g.N + q;
      const double wn = w[p] - v[p] * hv;
      w[p] = wn;
      const double wt = (c[d] == 0) ? 3.0 : 1.0;
      acc[d] += wt * (wn * (NORM ? wn : vn[p]));
    }
    const long p = g.dim * g.N + q;
    const double wn = w[p] - v[p] * hv;
    w[p] = wn;
    acc[3] += wn * (NORM ? wn : vn[p]);
  }
  for (int seg = 0; seg < 4; seg++) {
    double r = block_sum(acc[seg], sh);
    if (threadIdx.x == 0) partial[blockIdx.x * 4 + seg] = r;
  }
}
void launch_axmy_dot(hipStream_t s, const Grid &g, double *w, const double *v, const double *h, const double *vn,
                     double *partial) {
  if (vn) LAUNCH(k_axmy_dot<false>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, v, h, vn, partial);
  else LAUNCH(k_axmy_dot<true>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, v, h, vn, partial);
}

// v = src / scal (k_div_by) and its FP32 copy v32 (k_convert's rounding) in
// one pass: the next Krylov vector and the FP32 preconditioner input.
__global__ void k_div_by_f32(double *v, float *v32, const double *__restrict__ src, long n, const double *scal) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const double x = src[t] / scal[0];
    v[t] = x;
    v32[t] = (float)x;
  }
}
void launch_div_by_f32(hipStream_t s, double *v, float *v32, const double *src, long n, const double *scal) {
  LAUNCH(k_div_by_f32, dim3(nblk(n)), dim3(TPB), 0, s, v, v32, src, n, scal);
}

struct VList {
  const double *v[32];
};
// update_sol (gmres.f90:481-515): x = x + V(k)*y(k), k = 1..i in order.
__global__ void k_update_sol(double *x, VList V, const double *__restrict__ y, int i, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double xv = x[t];
  for (int k = 0; k < i; k++) xv = xv + V.v[k][t] * y[k];
  x[t] = xv;
}
void launch_update_sol(hipStream_t s, double *x, double *const V[], const double *y, int i, long n) {
  VList L = {};
  for (int k = 0; k < i && k < 32; k++) L.v[k] = V[k];
  LAUNCH(k_update_sol, dim3(nblk(n)), dim3(TPB), 0, s, x, L, y, i, n);
}

// x += sum_k y_k Z_k with Z in FP32, widened exactly, added in order k = 0..i-1
struct ZList {
  const float *v[32];
};
__global__ void k_update_sol_z32(double *x, ZList Z, const double *__restrict__ y, int i, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double xv = x[t];
  for (int k = 0; k < i; k++) xv = xv + (double)Z.v[k][t] * y[k];
  x[t] = xv;
}
void launch_update_sol_z32(hipStream_t s, double *x, float *const Z[], const double *y, int i, long n) {
  ZList L = {};
  for (int k = 0; k < i && k < 32; k++) L.v[k] = Z[k];
  LAUNCH(k_update_sol_z32, dim3(nblk(n)), dim3(TPB), 0, s, x, L, y, i, n);
}

// ------------------------------------------------------------------------
// Classical Gram-Schmidt of the flexible GMRES (no counterpart in the
// reference, whose gmres.f90:338-358 is modified Gram-Schmidt)
// ------------------------------------------------------------------------

// Weights of the Stokes inner product for the cell pair (q, q+1) of one
// x-row (nx even): face component d counts 3x where the pair's coordinate
// along d is 0 (only the first cell of the pair for d = 0).
struct PairW {
  double w0a, w0b, w1, w2;
};
__device__ __forceinline__ PairW pair_weights(const Grid &g, long q) {
  const long row = q / g.n[0];
  const int i = (int)(q - row * g.n[0]);
  const int j = (int)(row % g.n[1]);
  const int k = (int)(row / g.n[1]);
  PairW w;
  w.w0a = (i == 0) ? 3.0 : 1.0;
  w.w0b = 1.0;
  w.w1 = (j == 0) ? 3.0 : 1.0;
  w.w2 = (k == 0) ? 3.0 : 1.0;
  return w;
}

// Sums acc[0..n-1] over the workgroup (RED_TPB threads) into out[0..n-1]
// (thread 0 writes).
template <int NACC>
__device__ __forceinline__ void block_sum_multi(double (&acc)[NACC], double *sh, double *out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NACC; a++) {
    double v = acc[a];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) sh[wid * NACC + a] = v;
  }
  __syncthreads();
  if (threadIdx.x < NACC) {
    double r = 0.0;
    for (int w = 0; w < RED_TPB / 64; w++) r += sh[w * NACC + threadIdx.x];
    out[threadIdx.x] = r;
  }
}

// Segment sums of <w, V_k> for k < NV: partial[blk*CGS_STRIDE + 4k + seg]
// (seg d < dim: face component d with the box-face weights, seg 3: cells).
template <int NV>
__global__ void __launch_bounds__(RED_TPB) k_cgs_dots(Grid g, const double *__restrict__ w, VList V, double *partial) {
  __shared__ double sh[(RED_TPB / 64) * 4 * NV];
  double acc[4 * NV];
#pragma unroll
  for (int a = 0; a < 4 * NV; a++) acc[a] = 0.0;
  const long npair = g.N / 2;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < npair; t += stride) {
    const long q = 2 * t;
    const PairW pw = pair_weights(g, q);
#pragma unroll
    for (int seg = 0; seg < 4; seg++) {
      if (seg == 2 && g.dim == 2) continue;
      const long p = (long)(seg == 3 ? g.dim : seg) * g.N + q;
      const double2 wv = *reinterpret_cast<const double2 *>(w + p);
      const double wa = (seg == 0) ? pw.w0a : (seg == 1) ? pw.w1 : (seg == 2) ? pw.w2 : 1.0;
      const double wb = (seg == 0) ? pw.w0b : wa;
#pragma unroll
      for (int k = 0; k < NV; k++) {
        const double2 vv = *reinterpret_cast<const double2 *>(V.v[k] + p);
        acc[4 * k + seg] += wa * (wv.x * vv.x) + wb * (wv.y * vv.y);
      }
    }
  }
  block_sum_multi<4 * NV>(acc, sh, partial + (long)blockIdx.x * CGS_STRIDE);
}

// d_k = <w, V_k> (Stokes product, pressure weighted by pnw^2) and
// c_k = sig2[k] * d_k for k < nv.
struct Sig2 {
  double s[CGS_MAXV];
};
__global__ void __launch_bounds__(256) k_cgs_finalize(const double *__restrict__ partial, int nparts, int nv, int dim,
                                                      Sig2 sig2, double pnw, double *d_out, double *c_out) {
  __shared__ double sh[4 * 4 * CGS_MAXV];
  __shared__ double tot[4 * CGS_MAXV];
  double acc[4 * CGS_MAXV];
#pragma unroll
  for (int a = 0; a < 4 * CGS_MAXV; a++) acc[a] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
#pragma unroll
    for (int a = 0; a < 4 * CGS_MAXV; a++)
      if (a < 4 * nv) acc[a] += partial[(long)b * CGS_STRIDE + a];
  }
  block_sum_multi<4 * CGS_MAXV>(acc, sh, tot);
  __syncthreads();
  if (threadIdx.x < nv) {
    const double *s = tot + 4 * threadIdx.x;
    double su = s[0] + s[1];
    if (dim == 3) su = su + s[2];
    const double dk = su + (pnw * pnw) * s[3];
    d_out[threadIdx.x] = dk;
    c_out[threadIdx.x] = sig2.s[threadIdx.x] * dk;
  }
}

// vn = src - sum_k c_k V_k (k < NV, in order), its FP32 copy v32 (when not
// null) and the segment sums of <vn, vn> into partial[blk*4 + seg].
template <int NV>
__global__ void __launch_bounds__(RED_TPB) k_cgs_update(Grid g, const double *src, VList V,
                                                        const double *__restrict__ cdev, double *vn, float *v32,
                                                        double *partial) {
  __shared__ double sh[(RED_TPB / 64) * 4];
  double c[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) c[k] = cdev[k];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const long npair = g.N / 2;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < npair; t += stride) {
    const long q = 2 * t;
    const PairW pw = pair_weights(g, q);
#pragma unroll
    for (int seg = 0; seg < 4; seg++) {
      if (seg == 2 && g.dim == 2) continue;
      const long p = (long)(seg == 3 ? g.dim : seg) * g.N + q;
      double2 x = *reinterpret_cast<const double2 *>(src + p);
#pragma unroll
      for (int k = 0; k < NV; k++) {
        const double2 vv = *reinterpret_cast<const double2 *>(V.v[k] + p);
        x.x = x.x - c[k] * vv.x;
        x.y = x.y - c[k] * vv.y;
      }
      *reinterpret_cast<double2 *>(vn + p) = x;
      if (v32) *reinterpret_cast<float2 *>(v32 + p) = make_float2((float)x.x, (float)x.y);
      const double wa = (seg == 0) ? pw.w0a : (seg == 1) ? pw.w1 : (seg == 2) ? pw.w2 : 1.0;
      const double wb = (seg == 0) ? pw.w0b : wa;
      acc[seg] += wa * (x.x * x.x) + wb * (x.y * x.y);
    }
  }
  block_sum_multi<4>(acc, sh, partial + (long)blockIdx.x * 4);
}

void launch_cgs_dots(hipStream_t s, const Grid &g, const double *w, const double *const V[], int nv,
                     double *partial) {
  VList L = {};
  for (int k = 0; k < nv && k < CGS_MAXV; k++) L.v[k] = V[k];
  switch (nv) {
    case 1: LAUNCH(k_cgs_dots<1>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, L, partial); break;
    case 2: LAUNCH(k_cgs_dots<2>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, L, partial); break;
    case 3: LAUNCH(k_cgs_dots<3>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, L, partial); break;
    case 4: LAUNCH(k_cgs_dots<4>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, L, partial); break;
    case 5: LAUNCH(k_cgs_dots<5>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, L, partial); break;
    case 6: LAUNCH(k_cgs_dots<6>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, L, partial); break;
    case 7: LAUNCH(k_cgs_dots<7>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, L, partial); break;
    case 8: LAUNCH(k_cgs_dots<8>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, w, L, partial); break;
    default: report_error("launch_cgs_dots: 1..8 vectors per pass");
  }
}

void launch_cgs_finalize(hipStream_t s, const Grid &g, const double *partial, int nv, const double *sig2,
                         double pnw, double *d_out, double *c_out) {
  if (nv < 1 || nv > CGS_MAXV) {
    report_error("launch_cgs_finalize: 1..8 vectors per pass");
    return;
  }
  Sig2 q = {};
  for (int k = 0; k < nv; k++) q.s[k] = sig2[k];
  LAUNCH(k_cgs_finalize, dim3(1), dim3(256), 0, s, partial, RED_BLOCKS, nv, g.dim, q, pnw, d_out, c_out);
}

void launch_cgs_update(hipStream_t s, const Grid &g, const double *src, const double *const V[], int nv,
                       const double *c, double *vn, float *v32, double *partial) {
  VList L = {};
  for (int k = 0; k < nv && k < CGS_MAXV; k++) L.v[k] = V[k];
  switch (nv) {
    case 1: LAUNCH(k_cgs_update<1>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, src, L, c, vn, v32, partial); break;
    case 2: LAUNCH(k_cgs_update<2>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, src, L, c, vn, v32, partial); break;
    case 3: LAUNCH(k_cgs_update<3>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, src, L, c, vn, v32, partial); break;
    case 4: LAUNCH(k_cgs_update<4>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, src, L, c, vn, v32, partial); break;
    case 5: LAUNCH(k_cgs_update<5>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, src, L, c, vn, v32, partial); break;
    case 6: LAUNCH(k_cgs_update<6>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, src, L, c, vn, v32, partial); break;
    case 7: LAUNCH(k_cgs_update<7>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, src, L, c, vn, v32, partial); break;
    case 8: LAUNCH(k_cgs_update<8>, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, g, src, L, c, vn, v32, partial); break;
    default: report_error("launch_cgs_update: 1..8 vectors per pass");
  }
}

template <typename T>
__global__ void k_sub_scalar(T *x, long n, const double *scal) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = (T)((double)x[t] - scal[0]);
}
template <typename T>
void launch_sub_scalar(hipStream_t s, T *x, long n, const double *scal) {
  LAUNCH(k_sub_scalar<T>, dim3(nblk(n)), dim3(TPB), 0, s, x, n, scal);
}
template <typename T>
__global__ void k_sub_scalar3(Grid g, Ptr3<T> u, const double *scal) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  for (int d = 0; d < g.dim; d++) u.c[d][t] = (T)((double)u.c[d][t] - scal[d]);
}
template <typename T>
void launch_sub_scalar3(hipStream_t s, const Grid &g, T *const u[3], const double *scal) {
  LAUNCH(k_sub_scalar3<T>, dim3(nblk(g.N)), dim3(TPB), 0, s, g, p3(u), scal);
}

// ------------------------------------------------------------------------
// fluctuating momentum flux (stochastic_m_fluxdiv.f90:41-553, periodic)
// ------------------------------------------------------------------------

struct Noise {
  const double *cc;
  const double *ed[3];
  int nstages;
  long scc, sed;  // stage strides
  const double *w;
  // materialized scaled fluxes (filtered form): cell components [dim][N],
  // node/edge arrays [ne][2][N]; when set, cc/ed/w/variance/eta/T are not read
  const double *mcc, *med;
};

// sum_r w_r W_r (saxpy from zero), times variance, times sqrt(eta*T)
__device__ __forceinline__ double wsum(const Noise &nz, const double *base, long stride, long idx) {
  double v = 0.0;
  for (int r = 0; r < nz.nstages; r++) v = v + nz.w[r] * base[r * stride + idx];
  return v;
}

__global__ void __launch_bounds__(256) k_stoch_force(Grid g, Ptr3<double> force, Noise nz,
                                                     const double *__restrict__ eta, Ptr3<double> eta_ed,
                                                     const double *__restrict__ temp, Ptr3<double> temp_ed,
                                                     double variance) {
  long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int i, j, k;
  decode(g, t, i, j, k);
  Pt P = make_pt(g, i, j, k);
  const long N = g.N;
  double dxinv = 1.0 / g.dx;
  auto WC = [&](int comp, long q) -> double {
    if (nz.mcc) return nz.mcc[comp * N + q];
    double v = variance * wsum(nz, nz.cc + comp * N, nz.scc, q);
    return v * sqrt(eta[q] * temp[q]);
  };
  auto WE = [&](int e, int comp, long q) -> double {
    if (nz.med) return nz.med[((long)e * 2 + comp) * N + q];
    double v = variance * wsum(nz, nz.ed[e] + comp * N, nz.sed, q);
    return v * sqrt(eta_ed.c[e][q] * temp_ed.c[e][q]);
  };
  long xm = at3(g, P, -1, 0, 0), yp = at3(g, P, 0, 1, 0), xp = at3(g, P, 1, 0, 0), ym = at3(g, P, 0, -1, 0);
  if (g.dim == 2) {
    force.c[0][t] = force.c[0][t] + (WC(0, t) - WC(0, xm)) * dxinv + (WE(0, 0, yp) - WE(0, 0, t)) * dxinv;
    force.c[1][t] = force.c[1][t] + (WE(0, 1, xp) - WE(0, 1, t)) * dxinv + (WC(1, t) - WC(1, ym)) * dxinv;
    return;
  }
  long zp = at3(g, P, 0, 0, 1), zm = at3(g, P, 0, 0, -1);
  force.c[0][t] = force.c[0][t] + (WC(0, t) - WC(0, xm)) * dxinv + (WE(0, 0, yp) - WE(0, 0, t)) * dxinv +
                  (WE(1, 0, zp) - WE(1, 0, t)) * dxinv;
  force.c[1][t] = force.c[1][t] + (WE(0, 1, xp) - WE(0, 1, t)) * dxinv + (WC(1, t) - WC(1, ym)) * dxinv +
                  (WE(2, 0, zp) - WE(2, 0, t)) * dxinv;
  force.c[2][t] = force.c[2][t] + (WE(1, 1, xp) - WE(1, 1, t)) * dxinv + (WE(2, 1, yp) - WE(2, 1, t)) * dxinv +
                  (WC(2, t) - WC(2, zm)) * dxinv;
}

// The scaled fluxes of k_stoch_force's WC / WE (same arithmetic) written out:
// Scc[comp][N] (dim components), Sed[e][comp][N] (2 components per node / edge array).
__global__ void __launch_bounds__(256) k_stoch_scale(Grid g, Noise nz, const double *__restrict__ eta,
                                                     Ptr3<double> eta_ed, const double *__restrict__ temp,
                                                     Ptr3<double> temp_ed, double variance, double *Scc,
                                                     double *Sed) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const long N = g.N;
  const int ne = (g.dim == 2) ? 1 : 3;
  const double se = sqrt(eta[t] * temp[t]);
  for (int comp = 0; comp < g.dim; comp++) {
    const double v = variance * wsum(nz, nz.cc + comp * N, nz.scc, t);
    Scc[comp * N + t] = v * se;
  }
  for (int e = 0; e < ne; e++) {
    const double sq = sqrt(eta_ed.c[e][t] * temp_ed.c[e][t]);
    for (int comp = 0; comp < 2; comp++) {
      const double v = variance * wsum(nz, nz.ed[e] + comp * N, nz.sed, t);
      Sed[((long)e * 2 + comp) * N + t] = v * sq;
    }
  }
}

// One direction of multifab_filter (multifab_filter.f90:13-200) on narr
// arrays of N values each: out = the width-W filter of in along dir, terms
// in the reference's order; periodic neighbours.
template <int W>
__global__ void __launch_bounds__(256) k_filter_dir(Grid g, int dir, const double *__restrict__ in, double *out,
                                                    int narr) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N * narr) return;
  const long arr = t / g.N, q = t - arr * g.N;
  int i, j, k;
  decode(g, q, i, j, k);
  const double *f = in + arr * g.N;
  const int cd = (dir == 0) ? i : (dir == 1) ? j : k, nd = g.n[dir];
  const long base = q - (long)cd * g.s[dir];
  // periodic neighbour at offset o (|o| <= 4) along dir
  auto at = [&](int o) -> double {
    int c = (cd + o) % nd;
    if (c < 0) c += nd;
    return f[base + (long)c * g.s[dir]];
  };
  const double f0 = f[q];
  double v;
  if (W == 1) {
    v = f0 / 2 + at(-1) / 4 + at(1) / 4;
  } else if (W == 2) {
    v = 5 * f0 / 8 - at(-2) / 16 + at(-1) / 4 - at(2) / 16 + at(1) / 4;
  } else {
    v = 93 * f0 / 128 + 7 * at(-1) / 32 - 7 * at(-2) / 64 + at(-3) / 32 - at(-4) / 256 + 7 * at(1) / 32 -
        7 * at(2) / 64 + at(3) / 32 - at(4) / 256;
  }
  out[t] = v;
}

void launch_stoch_m_force(hipStream_t s, const Grid &g, double *const force[3], const double *Wcc,
                          double *const Wed[3], int nstages, const double *weights, long scc, long sed,
                          const double *eta, double *const eta_ed[3], const double *temp,
                          double *const temp_ed[3], double variance, int width, double *work) {
  Noise nz = {};
  nz.cc = Wcc;
  for (int e = 0; e < 3; e++) nz.ed[e] = Wed[e];
  nz.nstages = nstages;
  nz.scc = scc;
  nz.sed = sed;
  nz.w = weights;
  if (width > 0) {
    // scaled fluxes, then x, y (, z) filter passes between two work halves
    const int ne = (g.dim == 2) ? 1 : 3;
    const int narr = g.dim + 2 * ne;
    double *a = work, *b = work + (long)narr * g.N;
    LAUNCH(k_stoch_scale, dim3(nblk(g.N)), dim3(TPB), 0, s, g, nz, eta, p3(eta_ed), temp, p3(temp_ed), variance,
           a, a + (long)g.dim * g.N);
    for (int dir = 0; dir < g.dim; dir++) {
      const long n = g.N * narr;
      if (width == 1) LAUNCH(k_filter_dir<1>, dim3(nblk(n)), dim3(TPB), 0, s, g, dir, a, b, narr);
      else if (width == 2) LAUNCH(k_filter_dir<2>, dim3(nblk(n)), dim3(TPB), 0, s, g, dir, a, b, narr);
      else LAUNCH(k_filter_dir<4>, dim3(nblk(n)), dim3(TPB), 0, s, g, dir, a, b, narr);
      std::swap(a, b);
    }
    nz.mcc = a;
    nz.med = a + (long)g.dim * g.N;
  }
  LAUNCH(k_stoch_force, dim3(nblk(g.N)), dim3(TPB), 0, s, g, p3(force), nz, eta, p3(eta_ed), temp,
                     p3(temp_ed), variance);
}


// ------------------------------------------------------------------------
// explicit instantiations
// ------------------------------------------------------------------------
#define INST(T)                                                                                              \
  template void launch_stag_gs_colour<T>(hipStream_t, const Grid &, int, int, int, T, T *const[3],          \
                                         const StagCoef<T> &, const T *);                                    \
  template void launch_stag_applyop<T>(hipStream_t, const Grid &, int, int, T *const[3], T *const[3],        \
                                       const StagCoef<T> &, bool);                                           \
  template void launch_stag_jacobi<T>(hipStream_t, const Grid &, int, T, T *const[3], T *const[3],           \
                                      T *const[3], const StagCoef<T> &);                                     \
  template void launch_stag_residual<T>(hipStream_t, const Grid &, int, T *const[3], T *const[3],            \
                                        T *const[3], const StagCoef<T> &);                                   \
  template void launch_stag_restrict<T>(hipStream_t, const Grid &, T *const[3], T *const[3], int);           \
  template void launch_stag_prolong<T>(hipStream_t, const Grid &, T *const[3], T *const[3]);                 \
  template void launch_cc_restrict<T>(hipStream_t, const Grid &, const T *, T *);                            \
  template void launch_edge_restrict<T>(hipStream_t, const Grid &, T *const[3], T *const[3]);                \
  template void launch_nodal_restrict<T>(hipStream_t, const Grid &, const T *, T *);                         \
  template void launch_cc_rb_colour<T>(hipStream_t, const Grid &, int, T *, const T *, T *const[3]);         \
  template void launch_cc_residual<T>(hipStream_t, const Grid &, const T *, const T *, T *, T *const[3]);    \
  template void launch_cc_resid_restrict<T>(hipStream_t, const Grid &, const T *, const T *, T *const[3],   \
                                            const Grid &, T *);                                            \
  template void launch_cc_prolong_add<T>(hipStream_t, const Grid &, T *, const T *);                         \
  template void launch_fill<T>(hipStream_t, T *, long, T);                                                   \
  template void launch_scale<T>(hipStream_t, T *, long, T);                                                  \
  template void launch_apply_matrix<T>(hipStream_t, const Grid &, int, T *const[3], const T *, T *const[3], \
                                       T *, const StagCoef<T> &);                                            \
  template void launch_div_plus<T>(hipStream_t, const Grid &, T *const[3], const T *, T *, int);             \
  template void launch_precon_final<T>(hipStream_t, const Grid &, int, T *const[3], T *, const T *,          \
                                       const T *, T *const[3], const T *, const T *, double, double *);  \
  template void launch_sum_faces_weighted<T>(hipStream_t, const Grid &, T *const[3], double *);              \
  template void launch_sub_scalar<T>(hipStream_t, T *, long, const double *);                                \
  template void launch_sub_scalar3<T>(hipStream_t, const Grid &, T *const[3], const double *);
INST(double)
INST(float)
template void launch_convert<float, double>(hipStream_t, float *, const double *, long);
template void launch_convert<double, float>(hipStream_t, double *, const float *, long);

// fold per-block pressure sums (k_precon_final partials) into the 4-segment layout
void launch_fold_cells(hipStream_t s, const double *src, int nsrc, double *partial) {
  LAUNCH(k_fold_cells, dim3(RED_BLOCKS), dim3(RED_TPB), 0, s, src, nsrc, partial);
}

}  // namespace lmm
