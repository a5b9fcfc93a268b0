// Coarse levels of a V-cycle in one launch.
//
// Below a few tens of thousands of points a multigrid level is far too small
// to fill the GPU: each colour sweep, residual, restriction and prolongation
// becomes a launch of a few microseconds that does almost no work, and a
// V-cycle over eight levels spends most of its time in such launches.  Here
// one workgroup of 1024 threads runs the whole coarse part of the V-cycle --
// down-leg smoothing, residual and restriction on every coarse level, the
// bottom smoothing, then prolongation and up-leg smoothing -- with a
// workgroup barrier between dependent passes.  Every point is computed by the
// same device function as in the one-launch-per-pass kernels, in the same
// order of passes, so the result is bit for bit that of the separate launches
// (and of the reference's stag_mg_solver / F_MG level loops).  For
// visc_type -1 and 1 a face component's operator reads only that component,
// so the staggered V-cycle runs one workgroup per component side by side
// (the serial chain of barriers a third as long); the coupled viscosities
// keep all components in one workgroup.
#include "lmm_kernels.h"
#include "lmm_points.h"

namespace lmm {

namespace {

constexpr int CTPB = 1024;
const char kStagCoarseName[] = "k_stag_coarse_vcycle";
const char kCCCoarseName[] = "k_cc_coarse_vcycle";

template <typename T>
struct SLevelDesc {
  Grid g;
  Ptr3<T> phi, rhs, res;
  StagCoef<T> K;
};

template <typename T>
struct SCoarseArgs {
  SLevelDesc<T> L[COARSE_MAX_LEVELS];
  int m;                 // coarse levels held in L[0..m-1], finest first
  int visc, nd, nu, nb;  // visc_type, sweeps down / up / bottom
  int split;             // 1: workgroup a handles component a only
  T omega;
};

template <typename T>
__device__ void blk_fill(T *x, long n, T v) {
  for (long t = threadIdx.x; t < n; t += CTPB) x[t] = v;
}

// the colour loop is unrolled for each dimension so that the component a of
// every colour is a constant after inlining (no per-lane stack for the
// per-component pointer tables)
template <typename T, int DIM>
__device__ void blk_stag_smooth_d(const SLevelDesc<T> &L, int visc, T omega, int nsweeps, int alo, int ahi) {
  const long half = (long)(L.g.n[0] >> 1) * L.g.n[1] * L.g.n[2];
  const T dxsqinv = (T)(1.0 / (L.g.dx * L.g.dx));
  for (int m = 0; m < nsweeps; m++) {
#pragma unroll
    for (int color = 1; color <= 2 * DIM; color++) {
      const int a = (color - 1) / 2, par = (color + 1) % 2;
      if (a < alo || a >= ahi) continue;  // uniform over the workgroup
      for (long t = threadIdx.x; t < half; t += CTPB)
        gs_colour_at<T>(L.g, visc, a, par, omega, dxsqinv, L.phi, L.K, L.rhs.c[a], t);
      __syncthreads();
    }
  }
}

template <typename T>
__device__ void blk_stag_smooth(const SLevelDesc<T> &L, int visc, T omega, int nsweeps, int alo, int ahi) {
  if (L.g.dim == 3) blk_stag_smooth_d<T, 3>(L, visc, omega, nsweeps, alo, ahi);
  else blk_stag_smooth_d<T, 2>(L, visc, omega, nsweeps, alo, ahi);
}

template <typename T>
__device__ void blk_prolong(const SLevelDesc<T> &F, const SLevelDesc<T> &C, int alo, int ahi) {
  const long N = F.g.N;
  if (F.g.dim == 2) {
    if (alo <= 0 && 0 < ahi)
      for (long t = threadIdx.x; t < N; t += CTPB) prolong_at<T, 2, 0>(F.g, C.g, F.phi.c[0], C.phi.c[0], t);
    if (alo <= 1 && 1 < ahi)
      for (long t = threadIdx.x; t < N; t += CTPB) prolong_at<T, 2, 1>(F.g, C.g, F.phi.c[1], C.phi.c[1], t);
  } else {
    if (alo <= 0 && 0 < ahi)
      for (long t = threadIdx.x; t < N; t += CTPB) prolong_at<T, 3, 0>(F.g, C.g, F.phi.c[0], C.phi.c[0], t);
    if (alo <= 1 && 1 < ahi)
      for (long t = threadIdx.x; t < N; t += CTPB) prolong_at<T, 3, 1>(F.g, C.g, F.phi.c[1], C.phi.c[1], t);
    if (alo <= 2 && 2 < ahi)
      for (long t = threadIdx.x; t < N; t += CTPB) prolong_at<T, 3, 2>(F.g, C.g, F.phi.c[2], C.phi.c[2], t);
  }
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(CTPB) k_stag_coarse_vcycle(SCoarseArgs<T> A) {
  const int m = A.m, dim = A.L[0].g.dim;
  // the components this workgroup handles
  const int alo = A.split ? (int)blockIdx.x : 0, ahi = A.split ? alo + 1 : dim;
  for (int n = 0; n < m; n++)
    for (int d = alo; d < ahi; d++) blk_fill<T>(A.L[n].phi.c[d], A.L[n].g.N, (T)0);
  __syncthreads();
  for (int n = 0; n < m - 1; n++) {
    const SLevelDesc<T> &L = A.L[n];
    blk_stag_smooth<T>(L, A.visc, A.omega, A.nd, alo, ahi);
    const T dxsqinv = (T)(1.0 / (L.g.dx * L.g.dx));
    for (int a = alo; a < ahi; a++)
      for (long t = threadIdx.x; t < L.g.N; t += CTPB) residual_at<T>(L.g, A.visc, dxsqinv, L.phi, L.rhs, L.res, L.K, a, t);
    __syncthreads();
    const SLevelDesc<T> &Cn = A.L[n + 1];
    for (int a = alo; a < ahi; a++)
      for (long t = threadIdx.x; t < Cn.g.N; t += CTPB) restrict_at<T>(Cn.g, L.g, 0, L.res, Cn.rhs, a, t);
    __syncthreads();
  }
  blk_stag_smooth<T>(A.L[m - 1], A.visc, A.omega, A.nb, alo, ahi);
  for (int n = m - 2; n >= 0; n--) {
    blk_prolong<T>(A.L[n], A.L[n + 1], alo, ahi);
    blk_stag_smooth<T>(A.L[n], A.visc, A.omega, A.nu, alo, ahi);
  }
}

template <typename T>
struct CLevelDesc {
  Grid g;
  T *phi, *rhs, *res;
  Ptr3<T> a;
};

template <typename T>
struct CCoarseArgs {
  CLevelDesc<T> L[COARSE_MAX_LEVELS];
  int m, nd, nu, nb;
};

template <typename T>
__device__ void blk_cc_rb(const CLevelDesc<T> &L, int nsweeps) {
  const long half = (long)(L.g.n[0] >> 1) * L.g.n[1] * L.g.n[2];
  const T inv = (T)(1.0 / (L.g.dx * L.g.dx));
  for (int s = 0; s < nsweeps; s++)
    for (int colour = 0; colour < 2; colour++) {
      for (long t = threadIdx.x; t < half; t += CTPB) cc_rb_at<T>(L.g, colour, inv, L.phi, L.rhs, L.a, t);
      __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(CTPB) k_cc_coarse_vcycle(CCoarseArgs<T> A) {
  const int m = A.m;
  for (int n = 0; n < m; n++) blk_fill<T>(A.L[n].phi, A.L[n].g.N, (T)0);
  __syncthreads();
  for (int n = 0; n < m - 1; n++) {
    const CLevelDesc<T> &L = A.L[n];
    blk_cc_rb<T>(L, A.nd);
    const T inv = (T)(1.0 / (L.g.dx * L.g.dx));
    for (long t = threadIdx.x; t < L.g.N; t += CTPB) cc_residual_at<T>(L.g, inv, L.phi, L.rhs, L.res, L.a, t);
    __syncthreads();
    const CLevelDesc<T> &Cn = A.L[n + 1];
    for (long t = threadIdx.x; t < Cn.g.N; t += CTPB) cc_restrict_at<T>(Cn.g, L.g, L.res, Cn.rhs, t);
    __syncthreads();
  }
  blk_cc_rb<T>(A.L[m - 1], A.nb);
  for (int n = m - 2; n >= 0; n--) {
    const CLevelDesc<T> &F = A.L[n], &Cn = A.L[n + 1];
    for (long t = threadIdx.x; t < F.g.N; t += CTPB) cc_prolong_add_at<T>(F.g, Cn.g, F.phi, Cn.phi, t);
    __syncthreads();
    blk_cc_rb<T>(F, A.nu);
  }
}

}  // namespace

template <typename T>
void launch_stag_coarse_vcycle(hipStream_t s, int m, const CoarseStagLevel<T> *lv, int visc, T omega, int nd,
                               int nu, int nb) {
  SCoarseArgs<T> A;
  A.m = m;
  A.visc = visc;
  A.nd = nd;
  A.nu = nu;
  A.nb = nb;
  A.omega = omega;
  for (int n = 0; n < m; n++) {
    A.L[n].g = lv[n].g;
    A.L[n].phi = p3(lv[n].phi);
    A.L[n].rhs = p3(lv[n].rhs);
    A.L[n].res = p3(lv[n].res);
    A.L[n].K = lv[n].K;
  }
  // visc_type -1 / 1: independent components, one workgroup each
  const int dim = lv[0].g.dim;
  A.split = (visc == -1 || visc == 1) && tuning().coarse_split ? 1 : 0;
  hipLaunchKernelGGL(k_stag_coarse_vcycle<T>, dim3(A.split ? dim : 1), dim3(CTPB), 0, s, A);
  post_launch(s, kStagCoarseName);
}

template <typename T>
void launch_cc_coarse_vcycle(hipStream_t s, int m, const CoarseCCLevel<T> *lv, int nd, int nu, int nb) {
  CCoarseArgs<T> A;
  A.m = m;
  A.nd = nd;
  A.nu = nu;
  A.nb = nb;
  for (int n = 0; n < m; n++) {
    A.L[n].g = lv[n].g;
    A.L[n].phi = lv[n].phi;
    A.L[n].rhs = lv[n].rhs;
    A.L[n].res = lv[n].res;
    A.L[n].a = p3(lv[n].a);
  }
  hipLaunchKernelGGL(k_cc_coarse_vcycle<T>, dim3(1), dim3(CTPB), 0, s, A);
  post_launch(s, kCCCoarseName);
}

template void launch_stag_coarse_vcycle<double>(hipStream_t, int, const CoarseStagLevel<double> *, int, double,
                                                int, int, int);
template void launch_stag_coarse_vcycle<float>(hipStream_t, int, const CoarseStagLevel<float> *, int, float, int,
                                               int, int);
template void launch_cc_coarse_vcycle<double>(hipStream_t, int, const CoarseCCLevel<double> *, int, int, int);
template void launch_cc_coarse_vcycle<float>(hipStream_t, int, const CoarseCCLevel<float> *, int, int, int);

}  // namespace lmm
