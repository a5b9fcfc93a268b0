// Fused red-black Gauss-Seidel sweep of one face component (3D, visc_type
// -1 and 1), the finest-level hot loop of the staggered multigrid.
//
// The reference runs the 2*dim colours of stag_mg_smoother = 1 as separate
// passes (stag_mg_solver.f90:373-382): colour 2a+1 updates the even points
// (i+j+k even) of component a, colour 2a+2 the odd ones.  For visc_type -1
// and 1 a component's update reads only that component, so the sweep is
// component by component, and within a component
//   red   = phi_old + omega*(rhs - L(phi_old))/diag    at even points,
//   black = phi_old + omega*(rhs - L(red))/diag        at odd points,
// where L at an even point reads only odd neighbours (old values) and L at an
// odd point reads only even neighbours (red values).  Every red value is thus
// a pure function of the old field.  This kernel reads the old field from
// one buffer and writes the new field to another, so a workgroup can
// recompute the red values of its 1-point halo from the old field instead of
// waiting for its neighbours.  Each array is then streamed once per
// component instead of once per colour.  The arithmetic of every point is
// that of stag_applyop_3d + stag_mg_update_3d term for term, so the result
// equals the reference's two colour passes bit for bit.
//
// Tiling: a workgroup owns a 64 x 8 column of points in (x, y) and marches
// through tz planes in z.  Per plane it stages the old field with a 2-point
// halo in LDS, computes red on the 66 x 10 region (interior + 1-point halo)
// into a second LDS ring, then black on the interior of the previous plane
// and writes that plane of the new field.  The global loads of the next
// plane (old field and the coefficients of its points) are issued before
// the current plane is computed.
// Workgroups are numbered so that tiles sharing halo rows are dispatched to
// the same XCD (consecutive workgroup ids are spread over the 8 XCDs).
#include "lmm_kernels.h"
#include "lmm_points.h"

namespace lmm {

namespace {

constexpr int FX = 64, FY = 8;          // interior tile
constexpr int HX = FX + 4, HY = FY + 4; // old field, 2-point halo
constexpr int RX = FX + 2, RY = FY + 2; // red values, 1-point halo
constexpr int NRED = RY * (RX / 2);     // red points per plane of the red region

const char kFusedName[] = "k_rb_fused";

template <typename T, typename P>
__global__ void __launch_bounds__(256) k_rb_fused(Grid g, int tz, int ntx, int nty, int ntiles,
                                                  const T *__restrict__ pin, T *__restrict__ pout,
                                                  typename P::Args args0, const T *__restrict__ q0,
                                                  const T *__restrict__ q1, const T *__restrict__ q2,
                                                  const T *__restrict__ q3, const T *__restrict__ q4) {
  using Coef = typename P::Coef;
  const typename P::Args args = P::bind(args0, q0, q1, q2, q3, q4);
  // Four-plane rings: the plane stored at step kk+1 and the red plane
  // written at step kk never alias a plane that a slower thread may still
  // read for the black update of step kk-1, so two barriers per plane do.
  __shared__ T sp[4][HY][HX];
  __shared__ T sr[4][RY][RX];
  // XCD-aware tile order: workgroup w runs on XCD w % 8; give each XCD a
  // contiguous range of tiles so halo rows are shared in that XCD's L2.
  int w = blockIdx.x, tile = w;
  if (ntiles % 8 == 0) tile = (w % 8) * (ntiles / 8) + w / 8;
  const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
  const int i0 = bx * FX, j0 = by * FY, k0 = bz * tz;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const long sy = g.s[1], sz = g.s[2];
  const int tid = threadIdx.x;
  constexpr int NLD = (HY * HX + 255) / 256;   // old-field values staged per thread
  constexpr int NRI = (NRED + 255) / 256;      // red points per thread

  // positions of this thread's red items and black item for plane kk
  auto red_at = [&](int q, int kk, int &row, int &rr, int &ig, int &jg) {
    row = q / (RX / 2);
    const int cc = q - row * (RX / 2);
    const int jl = row - 1;                          // tile-local j, -1 .. FY
    jg = wrapc(j0 + jl, ny);
    rr = 2 * cc + ((j0 + jl + kk + 1) & 1);          // red: i + j + k even
    ig = wrapc(i0 + rr - 1, nx);
  };
  auto black_at = [&](int kk, int &row, int &il) {
    row = tid / (FX / 2);
    const int cc = tid - row * (FX / 2);
    il = 2 * cc + ((j0 + row + kk + 1) & 1);         // black: i + j + k odd
  };

  auto fetch_plane = [&](int kk, T *dst) {
    const int kg = wrapc(kk, nz);
#pragma unroll
    for (int n = 0; n < NLD; n++) {
      const int q = tid + 256 * n;
      if (q < HY * HX) {
        const int r = q / HX, c = q - r * HX;
        dst[n] = pin[wrapc(i0 - 2 + c, nx) + sy * wrapc(j0 - 2 + r, ny) + sz * kg];
      }
    }
  };
  auto store_plane = [&](int kk, const T *src) {
    T(*dst)[HX] = sp[kk & 3];
#pragma unroll
    for (int n = 0; n < NLD; n++) {
      const int q = tid + 256 * n;
      if (q < HY * HX) {
        const int r = q / HX, c = q - r * HX;
        dst[r][c] = src[n];
      }
    }
  };
  auto fetch_red = [&](int kk, Coef *dst) {
    const int kg = wrapc(kk, nz);
#pragma unroll
    for (int n = 0; n < NRI; n++) {
      const int q = tid + 256 * n;
      if (q < NRED) {
        int row, rr, ig, jg;
        red_at(q, kk, row, rr, ig, jg);
        dst[n] = P::load(g, ig, jg, kg, args);
      }
    }
  };
  auto fetch_black = [&](int kk, Coef &dst) {
    int row, il;
    black_at(kk, row, il);
    dst = P::load(g, i0 + il, j0 + row, kk, args);
  };
  auto red_plane = [&](int kk, const Coef *rc) {
    const int s = kk & 3, sm = (kk - 1) & 3, spl = (kk + 1) & 3;
#pragma unroll
    for (int n = 0; n < NRI; n++) {
      const int q = tid + 256 * n;
      if (q < NRED) {
        int row, rr, ig, jg;
        red_at(q, kk, row, rr, ig, jg);
        const int hr = row + 1, hc = rr + 1;         // position in sp
        sr[s][row][rr] = P::update(rc[n], sp[s][hr][hc], sp[s][hr][hc + 1], sp[s][hr][hc - 1], sp[s][hr + 1][hc],
                                   sp[s][hr - 1][hc], sp[spl][hr][hc], sp[sm][hr][hc], args);
      }
    }
  };
  auto black_plane = [&](int kk, const Coef &bc) {
    const int s = kk & 3, sm = (kk - 1) & 3, spl = (kk + 1) & 3;
    int row, il;
    black_at(kk, row, il);
    const int rrow = row + 1, rc0 = il + 1;          // position in sr
    const T v = P::update(bc, sp[s][row + 2][il + 2], sr[s][rrow][rc0 + 1], sr[s][rrow][rc0 - 1], sr[s][rrow + 1][rc0],
                          sr[s][rrow - 1][rc0], sr[spl][rrow][rc0], sr[sm][rrow][rc0], args);
    const long base = sy * (j0 + row) + sz * (long)kk;
    pout[base + i0 + il] = v;
    const int ir = il ^ 1;                           // the red point of the pair
    pout[base + i0 + ir] = sr[s][rrow][ir + 1];
  };

  // Operand registers of a step: old-field values of plane kk+1, red
  // coefficients of plane kk, black coefficients of plane kk-1.  Two sets
  // alternate so that the loads issued for the next step are first waited on
  // where they are used, not at the end of the current step.
  struct Regs {
    T pv[NLD];
    Coef rc[NRI];
    Coef bc;
  };
  auto step = [&](int kk, const Regs &cur, Regs &nxt) {
    store_plane(kk + 1, cur.pv);
    __syncthreads();
    if (kk + 2 <= k0 + tz + 1) fetch_plane(kk + 2, nxt.pv);
    if (kk + 1 <= k0 + tz) fetch_red(kk + 1, nxt.rc);
    if (kk >= k0 && kk < k0 + tz) fetch_black(kk, nxt.bc);
    red_plane(kk, cur.rc);
    __syncthreads();
    if (kk - 1 >= k0) black_plane(kk - 1, cur.bc);
  };

  // prologue: old-field planes k0-2, k0-1 staged, k0 in registers, red
  // coefficients of plane k0-1 in registers
  Regs ra = {}, rb = {};
  fetch_plane(k0 - 2, ra.pv);
  store_plane(k0 - 2, ra.pv);
  fetch_plane(k0 - 1, ra.pv);
  store_plane(k0 - 1, ra.pv);
  fetch_plane(k0, ra.pv);
  fetch_red(k0 - 1, ra.rc);
  // tz + 2 steps (k0-1 .. k0+tz), an even count since tz is even
  for (int kk = k0 - 1; kk <= k0 + tz; kk += 2) {
    step(kk, ra, rb);
    step(kk + 1, rb, ra);
  }
}

template <typename T, typename P>
void launch_rb(hipStream_t s, const Grid &g, int tz, const T *pin, T *pout, const typename P::Args &args) {
  const int ntx = g.n[0] / FX, nty = g.n[1] / FY, ntz = g.n[2] / tz;
  const int ntiles = ntx * nty * ntz;
  const T *q[5];
  P::ptrs(args, q);
  hipLaunchKernelGGL((k_rb_fused<T, P>), dim3(ntiles), dim3(256), 0, s, g, tz, ntx, nty, ntiles, pin, pout, args,
                     q[0], q[1], q[2], q[3], q[4]);
  post_launch(s, kFusedName);
}

template <typename T, int A, int VISC>
void launch_stag_one(hipStream_t s, const Grid &g, T omega, T dxsqinv, int tz, const T *pin, T *pout,
                     const T *rhs, const StagCoef<T> &K) {
  typename StagRB<T, A, VISC>::Args a;
  a.omega = omega;
  a.dxsqinv = dxsqinv;
  a.rhs = rhs;
  a.alpha = K.alpha[A];
  a.beta = K.beta;
  // xy = 0, xz = 1, yz = 2; edges shared with (A, t1) and (A, t2)
  a.e1 = (VISC == -1) ? K.bed[A == 2 ? 1 : 0] : nullptr;
  a.e2 = (VISC == -1) ? K.bed[A == 0 ? 1 : 2] : nullptr;
  launch_rb<T, StagRB<T, A, VISC>>(s, g, tz, pin, pout, a);
}

}  // namespace

// Plane count per workgroup, or 0 when the level cannot use the fused
// sweep.  By default a level needs enough workgroups (min_tiles) to fill the
// 256 CUs several times over; smaller levels are faster with the colour
// kernels.
int fused_rb_tz(const Grid &g, int min_tiles) {
  if (g.dim != 3 || g.n[0] % FX != 0 || g.n[1] % FY != 0) return 0;
  const int per_plane_set = (g.n[0] / FX) * (g.n[1] / FY);
  int tz = 32;
  while (tz > 8 && (g.n[2] % tz != 0 || per_plane_set * (g.n[2] / tz) < min_tiles)) tz /= 2;
  if (g.n[2] % tz != 0 || per_plane_set * (g.n[2] / tz) < min_tiles) return 0;
  return tz;
}

bool fused_rb_ok(const Grid &g, int visc, int min_tiles) {
  return (visc == -1 || visc == 1) && fused_rb_tz(g, min_tiles) > 0;
}

template <typename T>
void launch_stag_rb_fused(hipStream_t s, const Grid &g, int visc, int a, T omega, const T *pin, T *pout,
                          const T *rhs, const StagCoef<T> &K, int min_tiles) {
  const int tz = fused_rb_tz(g, min_tiles);
  const T dxsqinv = (T)(1.0 / (g.dx * g.dx));
  if (visc == -1) {
    if (a == 0) launch_stag_one<T, 0, -1>(s, g, omega, dxsqinv, tz, pin, pout, rhs, K);
    if (a == 1) launch_stag_one<T, 1, -1>(s, g, omega, dxsqinv, tz, pin, pout, rhs, K);
    if (a == 2) launch_stag_one<T, 2, -1>(s, g, omega, dxsqinv, tz, pin, pout, rhs, K);
  } else {
    if (a == 0) launch_stag_one<T, 0, 1>(s, g, omega, dxsqinv, tz, pin, pout, rhs, K);
    if (a == 1) launch_stag_one<T, 1, 1>(s, g, omega, dxsqinv, tz, pin, pout, rhs, K);
    if (a == 2) launch_stag_one<T, 2, 1>(s, g, omega, dxsqinv, tz, pin, pout, rhs, K);
  }
}

template <typename T>
void launch_cc_rb_fused(hipStream_t s, const Grid &g, const T *pin, T *pout, const T *rhs, T *const a[3],
                        int min_tiles) {
  typename CCRB<T>::Args args;
  args.inv = (T)(1.0 / (g.dx * g.dx));
  args.rhs = rhs;
  args.ax = a[0];
  args.ay = a[1];
  args.az = a[2];
  launch_rb<T, CCRB<T>>(s, g, fused_rb_tz(g, min_tiles), pin, pout, args);
}

template void launch_stag_rb_fused<double>(hipStream_t, const Grid &, int, int, double, const double *, double *,
                                           const double *, const StagCoef<double> &, int);
template void launch_stag_rb_fused<float>(hipStream_t, const Grid &, int, int, float, const float *, float *,
                                          const float *, const StagCoef<float> &, int);
template void launch_cc_rb_fused<double>(hipStream_t, const Grid &, const double *, double *, const double *,
                                          double *const[3], int);
template void launch_cc_rb_fused<float>(hipStream_t, const Grid &, const float *, float *, const float *,
                                         float *const[3], int);

}  // namespace lmm
