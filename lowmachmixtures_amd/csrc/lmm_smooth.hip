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
// Tiling: a workgroup owns a 64 x 8 column of points in (x, y) (64 x 16 in
// the FP32 three-component sweep, two point pairs per thread) and marches
// through tz planes in z.  Per plane it stages the old field with a 2-point
// halo in LDS, computes red on the 66 x 10 region (interior + 1-point halo)
// into a second LDS ring, then black on the interior of the previous plane
// and writes that plane of the new field.  The global loads of the next
// plane (old field and the coefficients of its points) are issued before
// the current plane is computed.
// Workgroups are numbered so that tiles sharing halo rows are dispatched to
// the same XCD (consecutive workgroup ids are spread over the 8 XCDs).
#include <hip/hip_ext.h>

#include "lmm_kernels.h"
#include "lmm_points.h"

namespace lmm {

namespace {

constexpr int FX = 64, FY = 8;          // interior tile
constexpr int HX = FX + 4, HY = FY + 4; // old field, 2-point halo
constexpr int RX = FX + 2, RY = FY + 2; // red values, 1-point halo
constexpr int NRED = RY * (RX / 2);     // red points per plane of the red region
constexpr int FX2 = 32, FY2 = 16;       // narrow tile of the three-component sweep (x = 32 levels)

const char kFusedName[] = "k_rb_fused";
const char kBlockName[] = "k_rb_block";
const char kMarchName[] = "k_rb_march2";
const char kMarch3Name[] = "k_rb_march2x3";
const char kResidRestrictName[] = "k_resid_restrict";

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

// Plane-marching form with paired coefficient loads: thread t owns the
// point pair (2m, 2m+1) of row r of the 64 x 8 tile in every plane.  At the
// step that computes red on plane kk it loads the coefficients of both
// points of its pair with one vector load per array, uses the red half at
// once and keeps the black half in registers for the next step, which
// computes black on plane kk.  Every coefficient line of the tile is thus
// read once, at one time; the red points of the 1-point halo ring (and of
// the two halo planes) use scalar loads of lines that the neighbouring
// tiles read in the same step.
//
// PF (prefetch): the coefficients of the next interior plane's red step
// (the thread's pair and its halo-ring point) are loaded one step ahead, so
// a step's global loads are consumed by the following step instead of
// stalling the one that issued them.  Same arithmetic, same bits.
template <typename T, typename P, int TX, bool PF, int NP = 1>
__device__ __forceinline__ void march2_tile(const Grid &g, int tz, int ntx, int nty, int tile,
                                            const T *__restrict__ pin, T *__restrict__ pout,
                                            const typename P::Args &args, T (*sp)[512 * NP / TX + 4][TX + 4],
                                            T (*sr)[512 * NP / TX + 2][TX + 2]) {
  using Coef = typename P::Coef;
  // tile TX x (512 NP / TX): NP point pairs per thread of the 256, rows
  // 512 / TX apart
  constexpr int FX = TX, FY = 512 * NP / TX, HX = FX + 4, HY = FY + 4, RX = FX + 2, RY = FY + 2;
  constexpr int RPP = 512 / TX;
  constexpr int NRED = RY * (RX / 2);
  const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
  const int i0 = bx * FX, j0 = by * FY, k0 = bz * tz;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const long sy = g.s[1], sz = g.s[2];
  const int tid = threadIdx.x;
  constexpr int NLD = (HY * HX + 255) / 256;
  // this thread's pairs: columns 2m, 2m+1 of tile rows r + RPP q, q < NP
  const int pr = tid / (FX / 2), pm = tid - pr * (FX / 2);
  const int pj = j0 + pr;

  auto fetch_plane = [&](int kk, T *dst) {
    const int kg = wrapc(kk, nz);
#pragma unroll
    for (int n = 0; n < NLD; n++) {
      const int q = tid + 256 * n;
      if (q < HY * HX) {
        const int r = q / HX, c = q - r * HX;
        // a null input stands for a field of zeros (the first sweep from a zero guess)
        dst[n] = pin ? pin[wrapc(i0 - 2 + c, nx) + sy * wrapc(j0 - 2 + r, ny) + sz * kg] : (T)0;
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
  // red point of plane kk at red-region position (row, rr) with coefficients c
  auto red_with = [&](int kk, int row, int rr, const Coef &c) {
    const int s = kk & 3, sm = (kk - 1) & 3, spl = (kk + 1) & 3;
    const int hr = row + 1, hc = rr + 1;
    sr[s][row][rr] = P::update(c, sp[s][hr][hc], sp[s][hr][hc + 1], sp[s][hr][hc - 1], sp[s][hr + 1][hc],
                               sp[s][hr - 1][hc], sp[spl][hr][hc], sp[sm][hr][hc], args);
  };
  auto point_coef = [&](int kk, int row, int rr) -> Coef {
    return P::load(g, wrapc(i0 + rr - 1, nx), wrapc(j0 + row - 1, ny), wrapc(kk, nz), args);
  };
  // one red point of plane kk at red-region position (row, rr), scalar loads
  auto red_point = [&](int kk, int row, int rr) { red_with(kk, row, rr, point_coef(kk, row, rr)); };
  // halo ring of an interior plane: rows -1 and FY (RX / 2 red points each),
  // then one end point of each interior row; NHALO <= 256, one per thread
  constexpr int NHALO = 2 * (RX / 2) + FY;
  static_assert(NHALO <= 256, "one halo-ring point per thread");
  auto halo_pos = [&](int kk, int h, int &row2, int &rr2) {
    if (h < 2 * (RX / 2)) {
      row2 = (h < RX / 2) ? 0 : RY - 1;
      const int jj = j0 + row2 - 1;
      rr2 = 2 * (h % (RX / 2)) + ((jj + kk + 1) & 1);
    } else {
      row2 = 1 + (h - 2 * (RX / 2));
      const int jj = j0 + row2 - 1;
      rr2 = ((jj + kk + 1) & 1) ? RX - 1 : 0;  // red: i + j + k even, i = i0 + rr - 1
    }
  };
  auto interior = [&](int kk) { return kk >= k0 && kk < k0 + tz; };
  // the coefficients an interior plane's red step needs: the thread's pair
  // and, for tid < NHALO, its halo-ring point
  auto load_red = [&](int kk, Coef (&c0)[NP], Coef (&c1)[NP], Coef &ch) {
#pragma unroll
    for (int q = 0; q < NP; q++) P::load_pair(g, i0 + 2 * pm, pj + RPP * q, kk, args, c0[q], c1[q]);
    if (tid < NHALO) {
      int row2, rr2;
      halo_pos(kk, tid, row2, rr2);
      ch = point_coef(kk, row2, rr2);
    }
  };
  // red on plane kk; for interior planes returns the black half of the pair.
  // PF: c0, c1, ch hold the plane's coefficients, loaded a step earlier.
  auto red_plane = [&](int kk, Coef (&cb)[NP], const Coef (&c0)[NP], const Coef (&c1)[NP], const Coef &ch) {
    if (interior(kk)) {
      Coef x0[NP], x1[NP], xh;
      if constexpr (PF) {
#pragma unroll
        for (int q = 0; q < NP; q++) {
          x0[q] = c0[q];
          x1[q] = c1[q];
        }
        xh = ch;
      } else {
        load_red(kk, x0, x1, xh);
      }
      // rows RPP apart have the same parity
      const bool first_red = ((pj + kk) & 1) == 0;  // column 2m is red when j + k is even
#pragma unroll
      for (int q = 0; q < NP; q++) {
        const Coef cr = first_red ? x0[q] : x1[q];
        cb[q] = first_red ? x1[q] : x0[q];
        red_with(kk, pr + RPP * q + 1, 2 * pm + (first_red ? 0 : 1) + 1, cr);
      }
      if (tid < NHALO) {
        int row2, rr2;
        halo_pos(kk, tid, row2, rr2);
        red_with(kk, row2, rr2, xh);
      }
    } else {
      for (int q = tid; q < NRED; q += 256) {
        const int row = q / (RX / 2), cc = q - row * (RX / 2);
        const int jj = j0 + row - 1;
        red_point(kk, row, 2 * cc + ((jj + kk + 1) & 1));
      }
    }
  };
  auto black_plane = [&](int kk, const Coef (&cb)[NP]) {
    const int s = kk & 3, sm = (kk - 1) & 3, spl = (kk + 1) & 3;
    const bool first_red = ((pj + kk) & 1) == 0;
    const int il = 2 * pm + (first_red ? 1 : 0);
    const int rc0 = il + 1;
#pragma unroll
    for (int q = 0; q < NP; q++) {
      const int rrow = pr + RPP * q + 1;
      const T v = P::update(cb[q], sp[s][rrow + 1][il + 2], sr[s][rrow][rc0 + 1], sr[s][rrow][rc0 - 1],
                            sr[s][rrow + 1][rc0], sr[s][rrow - 1][rc0], sr[spl][rrow][rc0], sr[sm][rrow][rc0], args);
      const T r = sr[s][rrow][first_red ? rc0 - 1 : rc0 + 1];
      st2(pout, i0 + 2 * pm + sy * (pj + RPP * q) + sz * (long)kk, first_red ? r : v, first_red ? v : r);
    }
  };

  T pva[NLD] = {}, pvb[NLD] = {};
  Coef cba[NP] = {}, cbb[NP] = {};
  // red coefficients of the current plane (a) and the next one (b)
  Coef ra0[NP] = {}, ra1[NP] = {}, rah = {}, rb0[NP] = {}, rb1[NP] = {}, rbh = {};
  auto step = [&](int kk, const T *pv, T *pvn, const Coef (&cb_prev)[NP], Coef (&cb_cur)[NP], Coef (&c0)[NP],
                  Coef (&c1)[NP], Coef &ch, Coef (&n0)[NP], Coef (&n1)[NP], Coef &nh) {
    store_plane(kk + 1, pv);
    __syncthreads();
    if (kk + 2 <= k0 + tz + 1) fetch_plane(kk + 2, pvn);
    red_plane(kk, cb_cur, c0, c1, ch);
    // the next plane's red coefficients, issued once this plane's are used
    if constexpr (PF) {
      if (interior(kk + 1)) load_red(kk + 1, n0, n1, nh);
    }
    __syncthreads();
    if (kk - 1 >= k0) black_plane(kk - 1, cb_prev);
  };
  fetch_plane(k0 - 2, pva);
  store_plane(k0 - 2, pva);
  fetch_plane(k0 - 1, pva);
  store_plane(k0 - 1, pva);
  fetch_plane(k0, pva);
  // tz + 2 steps, an even count
  for (int kk = k0 - 1; kk <= k0 + tz; kk += 2) {
    step(kk, pva, pvb, cba, cbb, ra0, ra1, rah, rb0, rb1, rbh);
    step(kk + 1, pvb, pva, cbb, cba, rb0, rb1, rbh, ra0, ra1, rah);
  }
}

template <typename T, typename P, bool PF, int TX>
__global__ void __launch_bounds__(256) k_rb_march2(Grid g, int tz, int ntx, int nty, int ntiles,
                                                   const T *__restrict__ pin, T *__restrict__ pout,
                                                   typename P::Args args0, const T *__restrict__ q0,
                                                   const T *__restrict__ q1, const T *__restrict__ q2,
                                                   const T *__restrict__ q3, const T *__restrict__ q4) {
  const typename P::Args args = P::bind(args0, q0, q1, q2, q3, q4);
  __shared__ T sp[4][512 / TX + 4][TX + 4];
  __shared__ T sr[4][512 / TX + 2][TX + 2];
  int w = blockIdx.x, tile = w;
  if (ntiles % 8 == 0) tile = (w % 8) * (ntiles / 8) + w / 8;
  march2_tile<T, P, TX, PF>(g, tz, ntx, nty, tile, pin, pout, args, sp, sr);
}

// The three face components of one sweep in one launch (visc_type -1 / 1:
// the components are independent).  Workgroup slots (tile, component) are
// numbered so that the three components of a tile go to the same XCD one
// after another: they run side by side and share the lines of beta_cc and
// of the edge arrays each of them reads in that XCD's L2, where three
// separate launches stream those arrays from memory once per component.
// Each workgroup runs the single-component tile code, so every point is
// computed exactly as in k_rb_march2.
template <typename T, typename P0, typename P1, typename P2, int TX, bool PF, int NP = 1>
__global__ void __launch_bounds__(256)
    k_rb_march2x3(Grid g, int tz, int ntx, int nty, int ntiles, const T *__restrict__ pin0, T *__restrict__ pout0,
                  const T *__restrict__ pin1, T *__restrict__ pout1, const T *__restrict__ pin2,
                  T *__restrict__ pout2, typename P0::Args a0, typename P1::Args a1, typename P2::Args a2,
                  const T *__restrict__ q00, const T *__restrict__ q01, const T *__restrict__ q02,
                  const T *__restrict__ q03, const T *__restrict__ q04, const T *__restrict__ q10,
                  const T *__restrict__ q11, const T *__restrict__ q12, const T *__restrict__ q13,
                  const T *__restrict__ q14, const T *__restrict__ q20, const T *__restrict__ q21,
                  const T *__restrict__ q22, const T *__restrict__ q23, const T *__restrict__ q24) {
  __shared__ T sp[4][512 * NP / TX + 4][TX + 4];
  __shared__ T sr[4][512 * NP / TX + 2][TX + 2];
  const int nslots = 3 * ntiles, w = blockIdx.x;
  int slot = w;
  if (nslots % 8 == 0) slot = (w % 8) * (nslots / 8) + w / 8;
  const int tile = slot / 3, comp = slot - 3 * tile;
  if (comp == 0) {
    march2_tile<T, P0, TX, PF, NP>(g, tz, ntx, nty, tile, pin0, pout0, P0::bind(a0, q00, q01, q02, q03, q04), sp, sr);
  } else if (comp == 1) {
    march2_tile<T, P1, TX, PF, NP>(g, tz, ntx, nty, tile, pin1, pout1, P1::bind(a1, q10, q11, q12, q13, q14), sp, sr);
  } else {
    march2_tile<T, P2, TX, PF, NP>(g, tz, ntx, nty, tile, pin2, pout2, P2::bind(a2, q20, q21, q22, q23, q24), sp, sr);
  }
}

// Block form of the same sweep: a workgroup owns a 32 x 8 x 8 block and
// updates it in three phases with no plane-to-plane dependence -- stage the
// old field of the block and a 2-point halo in LDS, update the red points of
// the block and a 1-point halo in place in LDS (red points read only black,
// i.e. old, neighbours), then update the black points of the block from the
// red values and write the block.  There is no serial chain of plane steps,
// so many workgroups per CU overlap their memory latency.
constexpr int BX = 32, BY = 8, BZ = 8;
constexpr int LX = BX + 4, LY = BY + 4, LZ = BZ + 4;      // staged region
constexpr int RBX = BX + 2, RBY = BY + 2, RBZ = BZ + 2;   // red region
constexpr int NBB = BZ * BY * (BX / 2);                  // black (and red) points of the block
// red points of the 1-point halo ring: end planes, end rows, one end of each inner row
constexpr int NRING = 2 * RBY * (RBX / 2) + 2 * BZ * (RBX / 2) + BZ * BY;

template <typename T, typename P>
__global__ void __launch_bounds__(256) k_rb_block(Grid g, int nbx, int nby, int nblocks, const T *__restrict__ pin,
                                                  T *__restrict__ pout, typename P::Args args0,
                                                  const T *__restrict__ q0, const T *__restrict__ q1,
                                                  const T *__restrict__ q2, const T *__restrict__ q3,
                                                  const T *__restrict__ q4) {
  using Coef = typename P::Coef;
  __shared__ T sp[LZ][LY][LX];
  const typename P::Args args = P::bind(args0, q0, q1, q2, q3, q4);
  int w = blockIdx.x, b = w;
  if (nblocks % 8 == 0) b = (w % 8) * (nblocks / 8) + w / 8;
  const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
  const int i0 = bx * BX, j0 = by * BY, k0 = bz * BZ;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const long sy = g.s[1], sz = g.s[2];
  const int tid = threadIdx.x;
  for (int q = tid; q < LZ * LY * LX; q += 256) {
    const int x = q % LX, y = (q / LX) % LY, z = q / (LX * LY);
    sp[z][y][x] = pin[wrapc(i0 - 2 + x, nx) + sy * wrapc(j0 - 2 + y, ny) + sz * wrapc(k0 - 2 + z, nz)];
  }
  // Each thread owns NPT point pairs (x, x+1) of the block; one point of a
  // pair is red and the other black.  Their coefficients come with one
  // vector load per array; the black half is kept for the black phase.
  constexpr int NPT = NBB / 256;
  Coef cb[NPT];
  __syncthreads();
#pragma unroll
  for (int n = 0; n < NPT; n++) {
    const int q = tid + 256 * n;
    const int m = q % (BX / 2), yy = (q / (BX / 2)) % BY, zz = q / ((BX / 2) * BY);
    const int j = j0 + yy, k = k0 + zz;
    Coef c0, c1;
    P::load_pair(g, i0 + 2 * m, j, k, args, c0, c1);
    const bool first_red = ((j + k) & 1) == 0;  // x = 2m is red when j + k is even
    const Coef cr = first_red ? c0 : c1;
    cb[n] = first_red ? c1 : c0;
    const int x = 2 * m + (first_red ? 0 : 1) + 2, y = yy + 2, z = zz + 2;
    sp[z][y][x] = P::update(cr, sp[z][y][x], sp[z][y][x + 1], sp[z][y][x - 1], sp[z][y + 1][x], sp[z][y - 1][x],
                            sp[z + 1][y][x], sp[z - 1][y][x], args);
  }
  // red points of the 1-point halo ring: the two end planes, the two end rows
  // of the inner planes, and one end point of each inner row
  for (int h = tid; h < NRING; h += 256) {
    int rz, ry, rx;
    if (h < 2 * RBY * (RBX / 2)) {
      rz = (h < RBY * (RBX / 2)) ? 0 : RBZ - 1;
      const int hh = h % (RBY * (RBX / 2));
      ry = hh / (RBX / 2);
      rx = 2 * (hh % (RBX / 2));
    } else if (h < 2 * RBY * (RBX / 2) + 2 * BZ * (RBX / 2)) {
      const int hh = h - 2 * RBY * (RBX / 2);
      rz = 1 + hh / RBX;
      const int rem = hh % RBX;
      ry = (rem < RBX / 2) ? 0 : RBY - 1;
      rx = 2 * (rem % (RBX / 2));
    } else {
      const int hh = h - 2 * RBY * (RBX / 2) - 2 * BZ * (RBX / 2);
      rz = 1 + hh / BY;
      ry = 1 + hh % BY;
      rx = 0;
    }
    const int j = j0 - 1 + ry, k = k0 - 1 + rz;
    const int par = (j + k + 1) & 1;  // red: i + j + k even, i = i0 - 1 + rx
    if (rx == 0 && ry >= 1 && ry <= BY && rz >= 1 && rz <= BZ)
      rx = par ? RBX - 1 : 0;  // inner row: the red one of its two end points
    else
      rx += par;
    const int i = i0 - 1 + rx;
    const int x = rx + 1, y = ry + 1, z = rz + 1;
    const Coef c = P::load(g, wrapc(i, nx), wrapc(j, ny), wrapc(k, nz), args);
    sp[z][y][x] = P::update(c, sp[z][y][x], sp[z][y][x + 1], sp[z][y][x - 1], sp[z][y + 1][x], sp[z][y - 1][x],
                            sp[z + 1][y][x], sp[z - 1][y][x], args);
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < NPT; n++) {
    const int q = tid + 256 * n;
    const int m = q % (BX / 2), yy = (q / (BX / 2)) % BY, zz = q / ((BX / 2) * BY);
    const int j = j0 + yy, k = k0 + zz;
    const bool first_red = ((j + k) & 1) == 0;
    const int xb = 2 * m + (first_red ? 1 : 0);
    const int x = xb + 2, y = yy + 2, z = zz + 2;
    const T v = P::update(cb[n], sp[z][y][x], sp[z][y][x + 1], sp[z][y][x - 1], sp[z][y + 1][x], sp[z][y - 1][x],
                          sp[z + 1][y][x], sp[z - 1][y][x], args);
    const T r = sp[z][y][first_red ? x - 1 : x + 1];
    st2(pout, i0 + 2 * m + sy * j + sz * (long)k, first_red ? r : v, first_red ? v : r);
  }
}

// tx: the march's tile width (64, or 32 with 32 x 16 tiles)
template <typename T, typename P>
void launch_rb(hipStream_t s, const Grid &g, int tz, int tx, const T *pin, T *pout, const typename P::Args &args) {
  const T *q[5];
  P::ptrs(args, q);
  if (tuning().rb_variant == 1) {
    const int nbx = g.n[0] / BX, nby = g.n[1] / BY, nbz = g.n[2] / BZ;
    const int nb = nbx * nby * nbz;
    hipLaunchKernelGGL((k_rb_block<T, P>), dim3(nb), dim3(256), 0, s, g, nbx, nby, nb, pin, pout, args, q[0], q[1],
                       q[2], q[3], q[4]);
    post_launch(s, kBlockName);
    return;
  }
  const int fy = (tx == FX2) ? FY2 : FY;
  const int ntx = g.n[0] / tx, nty = g.n[1] / fy, ntz = g.n[2] / tz;
  const int ntiles = ntx * nty * ntz;
  if (tuning().rb_variant == 2) {
    LaunchEvents *ev = launch_events();
    // FP32 prefetches the next plane's coefficients (FP64 would lose occupancy)
    const bool pf = sizeof(T) == 4 && tuning().rb_prefetch;
    auto kern = (tx == FX2) ? (pf ? k_rb_march2<T, P, true, FX2> : k_rb_march2<T, P, false, FX2>)
                            : (pf ? k_rb_march2<T, P, true, FX> : k_rb_march2<T, P, false, FX>);
    if (ev->start) {
      hipExtLaunchKernelGGL(kern, dim3(ntiles), dim3(256), 0, s, ev->start, ev->stop, 0, g, tz, ntx, nty, ntiles, pin,
                            pout, args, q[0], q[1], q[2], q[3], q[4]);
      ev->start = ev->stop = nullptr;
    } else {
      hipLaunchKernelGGL(kern, dim3(ntiles), dim3(256), 0, s, g, tz, ntx, nty, ntiles, pin, pout, args, q[0], q[1],
                         q[2], q[3], q[4]);
    }
    post_launch(s, kMarchName);
    return;
  }
  hipLaunchKernelGGL((k_rb_fused<T, P>), dim3(ntiles), dim3(256), 0, s, g, tz, ntx, nty, ntiles, pin, pout, args,
                     q[0], q[1], q[2], q[3], q[4]);
  post_launch(s, kFusedName);
}

// Residual r = rhs - L phi of one face component and its full-weighting
// restriction to the next coarser level in one pass: the down-leg step
// "residual, then restrict" of stag_mg_solver.f90:384-404 (the residual of
// stag_applyop, the restriction of stag_restriction, :971-1175).  The fine
// residual never goes to memory.  A workgroup owns a 64 x 8 column of fine
// points and marches through its z-slab; per plane it stages phi with a halo
// in LDS, computes the residual of its points (and of the one-point halo on
// the low side of the component's own direction, which the restriction's
// i-1 / j-1 terms read) into a ring of residual planes, and after every odd
// plane writes the 32 x 4 coarse points of the plane pair.  Component 2
// starts one plane early for its k-1 term.  Each residual is c.rhs - L with
// the operator of stag_applyop_3d and each coarse value sums its 12 fine
// residuals in the order of the reference, so the coarse right-hand side is
// bit for bit that of the two separate passes.
template <typename T, typename P, int A, int TX = 64>
__device__ __forceinline__ void rr_tile(const Grid &g, const Grid &gc, int tz, int ntx, int nty, int tile,
                                        const T *__restrict__ pin, T *__restrict__ cout, const typename P::Args &args,
                                        T (*sp)[512 / TX + 4][TX + 4], T (*sres)[512 / TX + 1][TX + 1]) {
  using Coef = typename P::Coef;
  // tile TX x (512 / TX): 64 x 8, or 32 x 16 on levels 32 points wide
  constexpr int FX = TX, FY = 512 / TX, HX = FX + 4, HY = FY + 4;
  const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
  const int i0 = bx * FX, j0 = by * FY, k0 = bz * tz;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const long sy = g.s[1], sz = g.s[2];
  const int tid = threadIdx.x;
  constexpr int NLD = (HY * HX + 255) / 256;
  const int pr = tid / (FX / 2), pm = tid - pr * (FX / 2);
  const T dxsqinv = args.dxsqinv;

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
  // residual at tile-local (il, jl) of plane kk from a loaded coefficient set
  auto resid = [&](const Coef &c, int kk, int jl, int il) -> T {
    const int s = kk & 3, sm = (kk - 1) & 3, spl = (kk + 1) & 3;
    const int hr = jl + 2, hc = il + 2;
    return c.rhs - P::apply(c, sp[s][hr][hc], sp[s][hr][hc + 1], sp[s][hr][hc - 1], sp[s][hr + 1][hc],
                            sp[s][hr - 1][hc], sp[spl][hr][hc], sp[sm][hr][hc], dxsqinv);
  };
  auto resid_plane = [&](int kk) {
    const int s = kk & 3, kg = wrapc(kk, nz);
    Coef c0, c1;
    P::load_pair(g, i0 + 2 * pm, j0 + pr, kg, args, c0, c1);
    sres[s][pr + 1][2 * pm + 1] = resid(c0, kk, pr, 2 * pm);
    sres[s][pr + 1][2 * pm + 2] = resid(c1, kk, pr, 2 * pm + 1);
    if (A == 0 && tid < FY) {
      const Coef c = P::load(g, wrapc(i0 - 1, nx), j0 + tid, kg, args);
      sres[s][tid + 1][0] = resid(c, kk, tid, -1);
    }
    if (A == 1 && tid < FX) {
      const Coef c = P::load(g, i0 + tid, wrapc(j0 - 1, ny), kg, args);
      sres[s][0][tid + 1] = resid(c, kk, -1, tid);
    }
  };
  // coarse plane K = (kk - 1) / 2 from fine planes kk - 2 (A = 2), kk - 1, kk
  auto restrict_plane = [&](int kk) {
    if (tid >= (FX / 2) * (FY / 2)) return;
    const int cj = tid / (FX / 2), ci = tid - cj * (FX / 2);
    const int s0 = (kk - 1) & 3, s1 = kk & 3, sm = (kk - 2) & 3;
    const int il = 2 * ci + 1, jl = 2 * cj + 1;  // position of fine (2I, 2J) in sres
    T g0, gp, gm;
    if (A == 0) {
      // transverse t1 = y, t2 = z
      g0 = sres[s0][jl][il] + sres[s0][jl + 1][il] + sres[s1][jl][il] + sres[s1][jl + 1][il];
      gp = sres[s0][jl][il + 1] + sres[s0][jl + 1][il + 1] + sres[s1][jl][il + 1] + sres[s1][jl + 1][il + 1];
      gm = sres[s0][jl][il - 1] + sres[s0][jl + 1][il - 1] + sres[s1][jl][il - 1] + sres[s1][jl + 1][il - 1];
    } else if (A == 1) {
      // t1 = x, t2 = z
      g0 = sres[s0][jl][il] + sres[s0][jl][il + 1] + sres[s1][jl][il] + sres[s1][jl][il + 1];
      gp = sres[s0][jl + 1][il] + sres[s0][jl + 1][il + 1] + sres[s1][jl + 1][il] + sres[s1][jl + 1][il + 1];
      gm = sres[s0][jl - 1][il] + sres[s0][jl - 1][il + 1] + sres[s1][jl - 1][il] + sres[s1][jl - 1][il + 1];
    } else {
      // t1 = x, t2 = y
      g0 = sres[s0][jl][il] + sres[s0][jl][il + 1] + sres[s0][jl + 1][il] + sres[s0][jl + 1][il + 1];
      gp = sres[s1][jl][il] + sres[s1][jl][il + 1] + sres[s1][jl + 1][il] + sres[s1][jl + 1][il + 1];
      gm = sres[sm][jl][il] + sres[sm][jl][il + 1] + sres[sm][jl + 1][il] + sres[sm][jl + 1][il + 1];
    }
    const T v = (T)0.125 * g0 + (T)0.0625 * gp + (T)0.0625 * gm;
    cout[(i0 / 2 + ci) + gc.s[1] * (j0 / 2 + cj) + gc.s[2] * (long)((kk - 1) / 2)] = v;
  };

  const int ks = (A == 2) ? k0 - 1 : k0, ke = k0 + tz;  // fine planes ks .. ke-1
  T pva[NLD] = {}, pvb[NLD] = {};
  auto step = [&](int kk, const T *pv, T *pvn) {
    store_plane(kk + 1, pv);
    __syncthreads();
    if (kk + 2 <= ke) fetch_plane(kk + 2, pvn);
    resid_plane(kk);
    // after fine planes k0+1, k0+3, ...; not after the extra plane k0-1 of
    // component 2
    if (kk > k0 && ((kk - k0) & 1) == 1) {
      __syncthreads();
      restrict_plane(kk);
    }
  };
  fetch_plane(ks - 1, pva);
  store_plane(ks - 1, pva);
  fetch_plane(ks, pva);
  store_plane(ks, pva);
  fetch_plane(ks + 1, pva);
  for (int kk = ks; kk < ke; kk += 2) {
    step(kk, pva, pvb);
    if (kk + 1 < ke) step(kk + 1, pvb, pva);
  }
}

template <typename T, typename P, int A>
__global__ void __launch_bounds__(256) k_resid_restrict(Grid g, Grid gc, int tz, int ntx, int nty, int ntiles,
                                                        const T *__restrict__ pin, T *__restrict__ cout,
                                                        typename P::Args args0, const T *__restrict__ q0,
                                                        const T *__restrict__ q1, const T *__restrict__ q2,
                                                        const T *__restrict__ q3, const T *__restrict__ q4) {
  __shared__ T sp[4][HY][HX];
  __shared__ T sres[4][FY + 1][FX + 1];  // residuals, row / column 0 = the low-side halo
  int w = blockIdx.x, tile = w;
  if (ntiles % 8 == 0) tile = (w % 8) * (ntiles / 8) + w / 8;
  rr_tile<T, P, A>(g, gc, tz, ntx, nty, tile, pin, cout, P::bind(args0, q0, q1, q2, q3, q4), sp, sres);
}

// The three components' residual + restriction in one launch: slots (tile,
// component) numbered as in the three-component sweep, so the components of
// a tile run side by side on one XCD and share the lines of beta_cc and the
// edge arrays in its L2.  The residual reads phi only, so the components are
// independent for every visc_type.
template <typename T, int V, int TX>
__global__ void __launch_bounds__(256)
    k_resid_restrict3(Grid g, Grid gc, int tz, int ntx, int nty, int ntiles, const T *__restrict__ pin0,
                      T *__restrict__ cout0, const T *__restrict__ pin1, T *__restrict__ cout1,
                      const T *__restrict__ pin2, T *__restrict__ cout2, typename StagRB<T, 0, V>::Args a0,
                      typename StagRB<T, 1, V>::Args a1, typename StagRB<T, 2, V>::Args a2,
                      const T *__restrict__ q00, const T *__restrict__ q01, const T *__restrict__ q02,
                      const T *__restrict__ q03, const T *__restrict__ q04, const T *__restrict__ q10,
                      const T *__restrict__ q11, const T *__restrict__ q12, const T *__restrict__ q13,
                      const T *__restrict__ q14, const T *__restrict__ q20, const T *__restrict__ q21,
                      const T *__restrict__ q22, const T *__restrict__ q23, const T *__restrict__ q24) {
  using P0 = StagRB<T, 0, V>;
  using P1 = StagRB<T, 1, V>;
  using P2 = StagRB<T, 2, V>;
  __shared__ T sp[4][512 / TX + 4][TX + 4];
  __shared__ T sres[4][512 / TX + 1][TX + 1];
  const int nslots = 3 * ntiles, w = blockIdx.x;
  int slot = w;
  if (nslots % 8 == 0) slot = (w % 8) * (nslots / 8) + w / 8;
  const int tile = slot / 3, comp = slot - 3 * tile;
  if (comp == 0)
    rr_tile<T, P0, 0, TX>(g, gc, tz, ntx, nty, tile, pin0, cout0, P0::bind(a0, q00, q01, q02, q03, q04), sp, sres);
  else if (comp == 1)
    rr_tile<T, P1, 1, TX>(g, gc, tz, ntx, nty, tile, pin1, cout1, P1::bind(a1, q10, q11, q12, q13, q14), sp, sres);
  else
    rr_tile<T, P2, 2, TX>(g, gc, tz, ntx, nty, tile, pin2, cout2, P2::bind(a2, q20, q21, q22, q23, q24), sp, sres);
}

template <typename T, int A, int VISC>
typename StagRB<T, A, VISC>::Args stag_args(T omega, T dxsqinv, const T *rhs, const StagCoef<T> &K) {
  typename StagRB<T, A, VISC>::Args a = {};
  a.omega = omega;
  a.dxsqinv = dxsqinv;
  a.rhs = rhs;
  a.alpha = K.alpha[A];
  a.beta = K.beta;
  // xy = 0, xz = 1, yz = 2; edges shared with (A, t1) and (A, t2)
  a.e1 = (VISC == -1) ? K.bed[A == 2 ? 1 : 0] : nullptr;
  a.e2 = (VISC == -1) ? K.bed[A == 0 ? 1 : 2] : nullptr;
  return a;
}

template <typename T, int VISC>
void launch_stag_three(hipStream_t s, const Grid &g, T omega, T dxsqinv, int tz, int tx, int np, T *const pin_[3],
                       T *const pout[3], T *const rhs[3], const StagCoef<T> &K, bool zero_in) {
  T *const zeros[3] = {nullptr, nullptr, nullptr};
  T *const *pin = zero_in ? zeros : pin_;
  using P0 = StagRB<T, 0, VISC>;
  using P1 = StagRB<T, 1, VISC>;
  using P2 = StagRB<T, 2, VISC>;
  const auto a0 = stag_args<T, 0, VISC>(omega, dxsqinv, rhs[0], K);
  const auto a1 = stag_args<T, 1, VISC>(omega, dxsqinv, rhs[1], K);
  const auto a2 = stag_args<T, 2, VISC>(omega, dxsqinv, rhs[2], K);
  const T *q0[5], *q1[5], *q2[5];
  P0::ptrs(a0, q0);
  P1::ptrs(a1, q1);
  P2::ptrs(a2, q2);
  LaunchEvents *ev = launch_events();
  // FP32 prefetches the next plane's coefficients (FP64 would lose occupancy)
  const bool pf = sizeof(T) == 4 && tuning().rb_prefetch;
  if (tx == FX && np == 2) {
    // 64 x 16 tiles, two point pairs per thread
    const int ntx = g.n[0] / FX, nty = g.n[1] / (2 * FY), ntz = g.n[2] / tz;
    const int ntiles = ntx * nty * ntz;
    auto kern = pf ? k_rb_march2x3<T, P0, P1, P2, FX, true, 2> : k_rb_march2x3<T, P0, P1, P2, FX, false, 2>;
    hipExtLaunchKernelGGL(kern, dim3(3 * ntiles), dim3(256), 0, s, ev->start, ev->stop, 0, g, tz, ntx, nty, ntiles,
                          pin[0], pout[0], pin[1], pout[1], pin[2], pout[2], a0, a1, a2, q0[0], q0[1], q0[2], q0[3],
                          q0[4], q1[0], q1[1], q1[2], q1[3], q1[4], q2[0], q2[1], q2[2], q2[3], q2[4]);
  } else if (tx == FX) {
    const int ntx = g.n[0] / FX, nty = g.n[1] / FY, ntz = g.n[2] / tz;
    const int ntiles = ntx * nty * ntz;
    auto kern = pf ? k_rb_march2x3<T, P0, P1, P2, FX, true> : k_rb_march2x3<T, P0, P1, P2, FX, false>;
    hipExtLaunchKernelGGL(kern, dim3(3 * ntiles), dim3(256), 0, s, ev->start, ev->stop, 0, g, tz, ntx, nty, ntiles,
                          pin[0], pout[0], pin[1], pout[1], pin[2], pout[2], a0, a1, a2, q0[0], q0[1], q0[2], q0[3],
                          q0[4], q1[0], q1[1], q1[2], q1[3], q1[4], q2[0], q2[1], q2[2], q2[3], q2[4]);
  } else {
    const int ntx = g.n[0] / FX2, nty = g.n[1] / FY2, ntz = g.n[2] / tz;
    const int ntiles = ntx * nty * ntz;
    auto kern = pf ? k_rb_march2x3<T, P0, P1, P2, FX2, true> : k_rb_march2x3<T, P0, P1, P2, FX2, false>;
    hipExtLaunchKernelGGL(kern, dim3(3 * ntiles), dim3(256), 0, s, ev->start, ev->stop, 0, g, tz, ntx, nty, ntiles,
                          pin[0], pout[0], pin[1], pout[1], pin[2], pout[2], a0, a1, a2, q0[0], q0[1], q0[2], q0[3],
                          q0[4], q1[0], q1[1], q1[2], q1[3], q1[4], q2[0], q2[1], q2[2], q2[3], q2[4]);
  }
  ev->start = ev->stop = nullptr;
  post_launch(s, kMarch3Name);
}

// ---------------------------------------------------------------------------
// 2D fused red-black sweep (visc_type -1 / 1): both face components in one
// launch, a workgroup per (tile, component).  In 2D a component's colours
// read only that component, red (i + j even) reads only black (old)
// neighbours, so a workgroup stages the old field of its 64 x 8 tile with a
// 2-point halo in LDS, computes the red values of the tile and its 1-point
// ring, then the black values of the tile from them, and writes the tile:
// the reference's colours 2a+1, 2a+2 (stag_mg_solver.f90:373-382, the 2D
// stag_applyop / stag_mg_update arithmetic term by term), in one pass over
// memory instead of four colour passes.
template <typename T, int A, int VISC>
struct Stag2RB {
  struct Args {
    T omega, dxsqinv;
    const T *rhs, *alpha, *beta, *bnd;  // bnd: node beta (2D edges)
    const T *g1, *gamma;                // visc_type 2, 3: the other component; gamma
  };
  struct Coef {
    T av, rhs, B0, B1, E10, E11;
    T G[4];  // visc_type 2: the four cross-term values; 3: their sum in G[0]
  };
  static __device__ __forceinline__ Coef load(const Grid &g, int i, int j, const Args &a) {
    Coef c;
    const long sy = g.s[1];
    const long p = i + sy * j;
    c.av = a.alpha[p];
    c.rhs = a.rhs[p];
    if (VISC == -1) {
      const long pma = (A == 0) ? wrapc(i - 1, g.n[0]) + sy * j : i + sy * wrapc(j - 1, g.n[1]);
      const long p1p = (A == 0) ? i + sy * wrapc(j + 1, g.n[1]) : wrapc(i + 1, g.n[0]) + sy * j;
      c.B0 = a.beta[p];
      c.B1 = a.beta[pma];
      c.E10 = a.bnd[p];
      c.E11 = a.bnd[p1p];
    } else if (VISC >= 2) {
      // pma: -1 along A; p1p: +1 along the other direction; q1: both shifts
      const int nx = g.n[0], ny = g.n[1];
      const int ia = (A == 0) ? wrapc(i - 1, nx) : i, ja = (A == 1) ? wrapc(j - 1, ny) : j;
      const long pma = ia + sy * ja;
      const long p1p = (A == 0) ? i + sy * wrapc(j + 1, ny) : wrapc(i + 1, nx) + sy * j;
      const long q1 = (A == 0) ? ia + sy * wrapc(j + 1, ny) : wrapc(i + 1, nx) + sy * ja;
      c.B0 = a.beta[0];
      c.B1 = (VISC == 3) ? a.gamma[0] : (T)0;
      c.E10 = c.E11 = (T)0;
      const T *G1 = a.g1;
      if (VISC == 2) {
        c.G[0] = G1[p1p];
        c.G[1] = G1[p];
        c.G[2] = G1[q1];
        c.G[3] = G1[pma];
      } else {
        c.G[0] = G1[p1p] - G1[p] - G1[q1] + G1[pma];
      }
    } else {
      c.B0 = a.beta[0];
      c.B1 = c.E10 = c.E11 = (T)0;
    }
    return c;
  }
  // F: the point and its x / y neighbours
  static __device__ __forceinline__ T update(const Coef &c, T F, T xp, T xm, T yp, T ym, const Args &a) {
    // (fpa, fpma) along A; (fp1p, fp1m) along the other direction
    const T fpa = (A == 0) ? xp : yp, fpma = (A == 0) ? xm : ym;
    const T fp1p = (A == 0) ? yp : xp, fp1m = (A == 0) ? ym : xm;
    const T dxsqinv = a.dxsqinv;
    T L, fac;
    if (VISC == 1) {
      const T b = c.B0;
      T s = fpa + fpma + fp1p + fp1m;
      L = F * (c.av + (T)4.0 * b * dxsqinv) - s * b * dxsqinv;
      fac = c.av + (T)4.0 * b * dxsqinv;
    } else if (VISC == 2) {
      const T b = c.B0;
      T s = (T)2.0 * fpa + (T)2.0 * fpma + fp1p + fp1m + c.G[0] - c.G[1] - c.G[2] + c.G[3];
      L = F * (c.av + (T)6.0 * b * dxsqinv) - s * b * dxsqinv;
      fac = c.av + (T)6.0 * b * dxsqinv;
    } else if (VISC == 3) {
      const T b = c.B0, cg = c.B1, fourthirds = (T)(4.0 / 3.0), onethird = (T)(1.0 / 3.0);
      L = F * (c.av + ((T)14.0 * b * onethird + (T)2.0 * cg) * dxsqinv) -
          ((fpa + fpma) * (fourthirds * b + cg) + (fp1p + fp1m) * b + c.G[0] * (onethird * b + cg)) * dxsqinv;
      fac = c.av + ((T)(14.0 / 3.0) * b + (T)2.0 * cg) * dxsqinv;
    } else {
      T dg = c.B0 + c.B1 + c.E11 + c.E10;
      T off = fpa * c.B0 + fpma * c.B1 + fp1p * c.E11 + fp1m * c.E10;
      L = F * (c.av + dg * dxsqinv) - off * dxsqinv;
      fac = c.av + (c.B0 + c.B1 + c.E10 + c.E11) * dxsqinv;
    }
    return F + gs_div(a.omega * (c.rhs - L), fac);
  }
};

// The 2D cell-centred red-black update of the MAC-projection multigrid:
// cc_apply's arithmetic for dim 2 (x terms, then y).
template <typename T>
struct CC2RB {
  struct Args {
    T inv;
    const T *rhs, *ax, *ay;
  };
  struct Coef {
    T rhs, axl, axh, ayl, ayh;
  };
  static __device__ __forceinline__ Coef load(const Grid &g, int i, int j, const Args &a) {
    Coef c;
    const long sy = g.s[1];
    const long p = i + sy * j;
    c.rhs = a.rhs[p];
    c.axl = a.ax[p];
    c.axh = a.ax[wrapc(i + 1, g.n[0]) + sy * j];
    c.ayl = a.ay[p];
    c.ayh = a.ay[i + sy * wrapc(j + 1, g.n[1])];
    return c;
  }
  static __device__ __forceinline__ T update(const Coef &c, T F, T xp, T xm, T yp, T ym, const Args &a) {
    const T s = c.axl + c.axh + c.ayl + c.ayh;
    const T off = c.axl * xm + c.axh * xp + c.ayl * ym + c.ayh * yp;
    const T diag = -s * a.inv;
    const T lp = (-s * F + off) * a.inv;
    return F + gs_div(c.rhs - lp, diag);
  }
};

constexpr int TX2D = 64, TY2D = 8;

template <typename T, typename P>
__device__ __forceinline__ void tile2d(const Grid &g, int ntx, int tile, const T *__restrict__ pin,
                                       T *__restrict__ pout, const typename P::Args &args,
                                       T (*sp)[TX2D + 4], T (*sr)[TX2D + 2]) {
  constexpr int HX2 = TX2D + 4, HY2 = TY2D + 4, RX2 = TX2D + 2, RY2 = TY2D + 2;
  const int bx = tile % ntx, by = tile / ntx;
  const int i0 = bx * TX2D, j0 = by * TY2D;
  const int nx = g.n[0], ny = g.n[1];
  const long sy = g.s[1];
  const int tid = threadIdx.x;
  // the old field of the tile and a 2-point halo (a null input: zeros)
  for (int q = tid; q < HY2 * HX2; q += 256) {
    const int r = q / HX2, c = q - r * HX2;
    sp[r][c] = pin ? pin[wrapc(i0 - 2 + c, nx) + sy * wrapc(j0 - 2 + r, ny)] : (T)0;
  }
  __syncthreads();
  // red points of the tile and its 1-point ring: region (row, rr) is point
  // (i0 + rr - 1, j0 + row - 1); red where i + j is even (i0 is even)
  constexpr int NRED2 = RY2 * (RX2 / 2);
  for (int q = tid; q < NRED2; q += 256) {
    const int row = q / (RX2 / 2), cc = q - row * (RX2 / 2);
    const int jj = j0 + row - 1;
    const int rr = 2 * cc + ((jj + 1) & 1);
    const int hr = row + 1, hc = rr + 1;
    const typename P::Coef c = P::load(g, wrapc(i0 + rr - 1, nx), wrapc(jj, ny), args);
    const T xp = sp[hr][hc + 1], xm = sp[hr][hc - 1], yp = sp[hr + 1][hc], ym = sp[hr - 1][hc];
    sr[row][rr] = P::update(c, sp[hr][hc], xp, xm, yp, ym, args);
  }
  __syncthreads();
  // black: thread t owns the pair (2m, 2m+1) of tile row r
  const int pr = tid / (TX2D / 2), pm = tid - pr * (TX2D / 2);
  const int j = j0 + pr;
  const bool first_red = (j & 1) == 0;  // column 2m is red when j is even
  const int bc = 2 * pm + (first_red ? 1 : 0), rc = 2 * pm + (first_red ? 0 : 1);
  const typename P::Coef cb = P::load(g, i0 + bc, j, args);
  const int rrow = pr + 1, rcol = bc + 1;
  const T xp = sr[rrow][rcol + 1], xm = sr[rrow][rcol - 1], yp = sr[rrow + 1][rcol], ym = sr[rrow - 1][rcol];
  const T F = sp[pr + 2][bc + 2];
  const T v = P::update(cb, F, xp, xm, yp, ym, args);
  const T r = sr[rrow][rc + 1];
  st2(pout, i0 + 2 * pm + sy * j, first_red ? r : v, first_red ? v : r);
}

template <typename T, typename P0, typename P1>
__global__ void __launch_bounds__(256) k_rb_tile2d(Grid g, int ntx, int ntiles, const T *__restrict__ pin0,
                                                   T *__restrict__ pout0, const T *__restrict__ pin1,
                                                   T *__restrict__ pout1, typename P0::Args a0,
                                                   typename P1::Args a1) {
  __shared__ T sp[TY2D + 4][TX2D + 4];
  __shared__ T sr[TY2D + 2][TX2D + 2];
  const int nslots = 2 * ntiles, w = blockIdx.x;
  int slot = w;
  if (nslots % 8 == 0) slot = (w % 8) * (nslots / 8) + w / 8;
  const int tile = slot / 2, comp = slot - 2 * tile;
  if (comp == 0)
    tile2d<T, P0>(g, ntx, tile, pin0, pout0, a0, sp, sr);
  else
    tile2d<T, P1>(g, ntx, tile, pin1, pout1, a1, sp, sr);
}

template <typename T, typename P>
__global__ void __launch_bounds__(256) k_rb_tile2d1(Grid g, int ntx, int ntiles, const T *__restrict__ pin,
                                                    T *__restrict__ pout, typename P::Args a) {
  __shared__ T sp[TY2D + 4][TX2D + 4];
  __shared__ T sr[TY2D + 2][TX2D + 2];
  const int w = blockIdx.x;
  const int tile = (ntiles % 8 == 0) ? (w % 8) * (ntiles / 8) + w / 8 : w;
  tile2d<T, P>(g, ntx, tile, pin, pout, a, sp, sr);
}

}  // namespace

// 2D levels are fused from 16 workgroups on (a lower threshold than 3D's:
// a 2D level has no z-slabs to spread, and its colour passes are as short)
static int min_tiles_2d(int min_tiles) { return min_tiles < 16 ? min_tiles : 16; }

bool fused_rb2d_ok(const Grid &g, int visc, int min_tiles) {
  return g.dim == 2 && (visc == -1 || visc == 1) && g.n[0] % TX2D == 0 && g.n[1] % TY2D == 0 &&
         2 * (g.n[0] / TX2D) * (g.n[1] / TY2D) >= min_tiles_2d(min_tiles);
}

template <typename T, int VISC>
static void launch_rb2d(hipStream_t s, const Grid &g, T omega, T dxsqinv, T *const pin[3], T *const pout[3],
                        T *const rhs[3], const StagCoef<T> &K) {
  using P0 = Stag2RB<T, 0, VISC>;
  using P1 = Stag2RB<T, 1, VISC>;
  typename P0::Args a0 = {omega, dxsqinv, rhs[0], K.alpha[0], K.beta, K.bed[0]};
  typename P1::Args a1 = {omega, dxsqinv, rhs[1], K.alpha[1], K.beta, K.bed[0]};
  const int ntx = g.n[0] / TX2D, ntiles = ntx * (g.n[1] / TY2D);
  LaunchEvents *ev = launch_events();
  hipExtLaunchKernelGGL((k_rb_tile2d<T, P0, P1>), dim3(2 * ntiles), dim3(256), 0, s, ev->start, ev->stop, 0, g, ntx,
                        ntiles, pin[0], pout[0], pin[1], pout[1], a0, a1);
  ev->start = ev->stop = nullptr;
  post_launch(s, "k_rb_tile2d");
}

template <typename T>
void launch_stag_rb2d(hipStream_t s, const Grid &g, int visc, T omega, T *const pin_[3], T *const pout[3],
                      T *const rhs[3], const StagCoef<T> &K, bool zero_in) {
  if (!fused_rb2d_ok(g, visc, 1)) {
    report_error("launch_stag_rb2d: level not eligible");
    return;
  }
  T *const zeros[3] = {nullptr, nullptr, nullptr};
  T *const *pin = zero_in ? zeros : pin_;
  const T dxsqinv = (T)(1.0 / (g.dx * g.dx));
  if (visc == -1) launch_rb2d<T, -1>(s, g, omega, dxsqinv, pin, pout, rhs, K);
  else launch_rb2d<T, 1>(s, g, omega, dxsqinv, pin, pout, rhs, K);
}
template void launch_stag_rb2d<double>(hipStream_t, const Grid &, int, double, double *const[3], double *const[3],
                                       double *const[3], const StagCoef<double> &, bool);
template void launch_stag_rb2d<float>(hipStream_t, const Grid &, int, float, float *const[3], float *const[3],
                                      float *const[3], const StagCoef<float> &, bool);

bool fused_rb2d_comp_ok(const Grid &g, int visc, int min_tiles) {
  return g.dim == 2 && (visc == 2 || visc == 3) && g.n[0] % TX2D == 0 && g.n[1] % TY2D == 0 &&
         2 * (g.n[0] / TX2D) * (g.n[1] / TY2D) >= min_tiles_2d(min_tiles);
}

template <typename T, int A, int VISC>
static void launch_rb2d_one(hipStream_t s, const Grid &g, T omega, const T *pin, T *pout, const T *rhs,
                            const StagCoef<T> &K, const T *g1) {
  using P = Stag2RB<T, A, VISC>;
  typename P::Args a = {};
  a.omega = omega;
  a.dxsqinv = (T)(1.0 / (g.dx * g.dx));
  a.rhs = rhs;
  a.alpha = K.alpha[A];
  a.beta = K.beta;
  a.bnd = K.bed[0];
  a.g1 = g1;
  a.gamma = K.gamma;
  const int ntx = g.n[0] / TX2D, ntiles = ntx * (g.n[1] / TY2D);
  LaunchEvents *ev = launch_events();
  hipExtLaunchKernelGGL((k_rb_tile2d1<T, P>), dim3(ntiles), dim3(256), 0, s, ev->start, ev->stop, 0, g, ntx, ntiles,
                        pin, pout, a);
  ev->start = ev->stop = nullptr;
  post_launch(s, "k_rb_tile2d1");
}

template <typename T>
void launch_stag_rb2d_comp(hipStream_t s, const Grid &g, int visc, int a, T omega, const T *pin, T *pout,
                           const T *rhs, const StagCoef<T> &K, const T *other) {
  if (!fused_rb2d_comp_ok(g, visc, 1) || !pin || !other) {
    report_error("launch_stag_rb2d_comp: level not eligible");
    return;
  }
  if (visc == 2) {
    if (a == 0) launch_rb2d_one<T, 0, 2>(s, g, omega, pin, pout, rhs, K, other);
    else launch_rb2d_one<T, 1, 2>(s, g, omega, pin, pout, rhs, K, other);
  } else {
    if (a == 0) launch_rb2d_one<T, 0, 3>(s, g, omega, pin, pout, rhs, K, other);
    else launch_rb2d_one<T, 1, 3>(s, g, omega, pin, pout, rhs, K, other);
  }
}
template void launch_stag_rb2d_comp<double>(hipStream_t, const Grid &, int, int, double, const double *, double *,
                                            const double *, const StagCoef<double> &, const double *);
template void launch_stag_rb2d_comp<float>(hipStream_t, const Grid &, int, int, float, const float *, float *,
                                           const float *, const StagCoef<float> &, const float *);

bool cc_rb2d_ok(const Grid &g, int min_tiles) {
  return g.dim == 2 && g.n[0] % TX2D == 0 && g.n[1] % TY2D == 0 &&
         (g.n[0] / TX2D) * (g.n[1] / TY2D) >= min_tiles_2d(min_tiles);
}

template <typename T>
void launch_cc_rb2d(hipStream_t s, const Grid &g, const T *pin, T *pout, const T *rhs, T *const a[3]) {
  if (!cc_rb2d_ok(g, 1)) {
    report_error("launch_cc_rb2d: level not eligible");
    return;
  }
  typename CC2RB<T>::Args args = {};  // zeroed padding: kernel arguments are copied bytewise
  args.inv = (T)(1.0 / (g.dx * g.dx));
  args.rhs = rhs;
  args.ax = a[0];
  args.ay = a[1];
  const int ntx = g.n[0] / TX2D, ntiles = ntx * (g.n[1] / TY2D);
  LaunchEvents *ev = launch_events();
  hipExtLaunchKernelGGL((k_rb_tile2d1<T, CC2RB<T>>), dim3(ntiles), dim3(256), 0, s, ev->start, ev->stop, 0, g, ntx,
                        ntiles, pin, pout, args);
  ev->start = ev->stop = nullptr;
  post_launch(s, "k_rb_tile2d1");
}
template void launch_cc_rb2d<double>(hipStream_t, const Grid &, const double *, double *, const double *,
                                     double *const[3]);
template void launch_cc_rb2d<float>(hipStream_t, const Grid &, const float *, float *, const float *,
                                    float *const[3]);

namespace {

template <typename T, int A, int VISC>
void launch_stag_one(hipStream_t s, const Grid &g, T omega, T dxsqinv, int tz, int tx, const T *pin, T *pout,
                     const T *rhs, const StagCoef<T> &K, const T *const cur[3]) {
  typename StagRB<T, A, VISC>::Args a = {};
  a.omega = omega;
  a.dxsqinv = dxsqinv;
  a.rhs = rhs;
  a.alpha = K.alpha[A];
  a.beta = K.beta;
  a.gamma = K.gamma;
  if (VISC >= 2) {
    // the other components: t1 = y for A = 0, x otherwise; t2 = z for A < 2, y for A = 2
    a.e1 = cur[A == 0 ? 1 : 0];
    a.e2 = cur[A == 2 ? 1 : 2];
  } else {
    // xy = 0, xz = 1, yz = 2; edges shared with (A, t1) and (A, t2)
    a.e1 = (VISC == -1) ? K.bed[A == 2 ? 1 : 0] : nullptr;
    a.e2 = (VISC == -1) ? K.bed[A == 0 ? 1 : 2] : nullptr;
  }
  launch_rb<T, StagRB<T, A, VISC>>(s, g, tz, tx, pin, pout, a);
}

}  // namespace

LaunchEvents *launch_events() {
  static thread_local LaunchEvents ev = {nullptr, nullptr};
  return &ev;
}

bool fused_rb_takes_zeros() { return tuning().rb_variant == 2; }

// Workgroups aimed for by FP32 (es = 4) and FP64 (es = 8) launches: an FP64
// workgroup's LDS rings allow 3 resident per CU instead of 4, and twice the
// count keeps the last round of workgroups from running on a third of the
// machine.
static int rb_target(int es) {
  const int t = tuning().rb_target > 0 ? tuning().rb_target : 1;
  return es == 8 ? 2 * t : t;
}

// Plane count per workgroup, or 0 when the level cannot use the fused
// sweep.  The plane-marching forms take the thickest z-slab (at most 32
// planes) that still gives rb_target(es) workgroups over the ncomp
// components one launch sweeps, down to 2 planes for the paired form (8 for
// the first form); a launch with fewer than min_tiles workgroups is left to
// the colour kernels.  rb_tz is the rule for fx x fy tiles.
// nsweep: the launches one sweep makes of this shape (the components swept
// one after another), counted against min_tiles with the launch's own tiles
static int rb_tz(const Grid &g, int min_tiles, int es, int ncomp, int fx, int fy, int nsweep = 1) {
  if (g.dim != 3 || g.n[0] % fx != 0 || g.n[1] % fy != 0) return 0;
  const int per_plane_set = (g.n[0] / fx) * (g.n[1] / fy);
  const int tz_min = (tuning().rb_variant == 2) ? 2 : 8;
  const int t = rb_target(es);
  const int target = t > min_tiles ? t : min_tiles;
  int tz = 32;
  auto fits = [&](int z) { return g.n[2] % z == 0 && ncomp * per_plane_set * (g.n[2] / z) >= target; };
  while (tz > tz_min && !fits(tz)) tz /= 2;
  // at the thinnest slab only the minimum (and divisibility) must hold
  if (g.n[2] % tz != 0 || nsweep * ncomp * per_plane_set * (g.n[2] / tz) < min_tiles) return 0;
  return tz;
}

// z-slab and tile width of a one-component fused sweep: 64 x 8 tiles, or (the
// plane-marching form only) 32 x 16 tiles on levels 32 points wide.  nsweep
// launches of this shape make one sweep (the staggered components one after
// another, 1 for the cell-centred field); together they must reach min_tiles.
static int rb1_shape(const Grid &g, int min_tiles, int es, int &tx, int nsweep = 1) {
  tx = FX;
  if (tuning().rb_variant == 1) {
    if (g.dim != 3 || g.n[0] % BX != 0 || g.n[1] % BY != 0 || g.n[2] % BZ != 0) return 0;
    const int nb = (g.n[0] / BX) * (g.n[1] / BY) * (g.n[2] / BZ);
    return nsweep * nb >= min_tiles ? BZ : 0;
  }
  const int tz = rb_tz(g, min_tiles, es, 1, FX, FY, nsweep);
  if (tz > 0 || tuning().rb_variant != 2) return tz;
  tx = FX2;
  return rb_tz(g, min_tiles, es, 1, FX2, FY2, nsweep);
}

int fused_rb_tz(const Grid &g, int min_tiles, int es, int ncomp) {
  int tx;
  if (ncomp == 1) return rb1_shape(g, min_tiles, es, tx);
  if (tuning().rb_variant == 1) return rb1_shape(g, min_tiles, es, tx);
  return rb_tz(g, min_tiles, es, ncomp, FX, FY);
}

bool fused_rb_ok(const Grid &g, int visc, int min_tiles) {
  int tx;
  return (visc == -1 || visc == 1 || visc == 2 || visc == 3) && rb1_shape(g, min_tiles, 4, tx, g.dim) > 0;
}

template <typename T>
void launch_stag_rb_fused(hipStream_t s, const Grid &g, int visc, int a, T omega, const T *pin, T *pout,
                          const T *rhs, const StagCoef<T> &K, int min_tiles, const T *const cur[3]) {
  int tx = FX;
  const int tz = rb1_shape(g, min_tiles, sizeof(T), tx, g.dim);
  if (tz <= 0 || !fused_rb_ok(g, visc, min_tiles) || ((visc == 2 || visc == 3) && !cur)) {
    report_error("launch_stag_rb_fused: level not eligible");
    return;
  }
  const T dxsqinv = (T)(1.0 / (g.dx * g.dx));
#define ONE(V)                                                                             \
  do {                                                                                     \
    if (a == 0) launch_stag_one<T, 0, V>(s, g, omega, dxsqinv, tz, tx, pin, pout, rhs, K, cur); \
    if (a == 1) launch_stag_one<T, 1, V>(s, g, omega, dxsqinv, tz, tx, pin, pout, rhs, K, cur); \
    if (a == 2) launch_stag_one<T, 2, V>(s, g, omega, dxsqinv, tz, tx, pin, pout, rhs, K, cur); \
  } while (0)
  if (visc == -1) ONE(-1);
  else if (visc == 1) ONE(1);
  else if (visc == 2) ONE(2);
  else ONE(3);
#undef ONE
}

// tile width (64, or 32 on levels 32 points wide), pairs per thread (2:
// 64 x 16 tiles, where tuning asks for them) and z-slab of the
// three-component sweep; 0 when the level cannot use it
static int rb3_shape(const Grid &g, int min_tiles, int es, int &tx, int &np) {
  tx = FX;
  np = 2;
  if (tuning().rb_pairs == 2 && es == 4) {  // FP64 would need 80 KB of LDS per workgroup
    const int tz2 = rb_tz(g, min_tiles, es, 3, FX, 2 * FY);
    if (tz2 > 0) return tz2;
  }
  np = 1;
  const int tz = rb_tz(g, min_tiles, es, 3, FX, FY);
  if (tz > 0) return tz;
  tx = FX2;
  return rb_tz(g, min_tiles, es, 3, FX2, FY2);
}

bool fused_rb3_ok(const Grid &g, int visc, int min_tiles) {
  int tx, np;
  return tuning().rb_variant == 2 && (visc == -1 || visc == 1) && rb3_shape(g, min_tiles, 4, tx, np) > 0;
}

template <typename T>
void launch_stag_rb_fused3(hipStream_t s, const Grid &g, int visc, T omega, T *const pin[3], T *const pout[3],
                           T *const rhs[3], const StagCoef<T> &K, int min_tiles, bool zero_in) {
  int tx = FX, np = 1;
  const int tz = rb3_shape(g, min_tiles, sizeof(T), tx, np);
  if (tz <= 0) {
    report_error("launch_stag_rb_fused3: level not eligible");
    return;
  }
  const T dxsqinv = (T)(1.0 / (g.dx * g.dx));
  if (visc == -1) launch_stag_three<T, -1>(s, g, omega, dxsqinv, tz, tx, np, pin, pout, rhs, K, zero_in);
  else launch_stag_three<T, 1>(s, g, omega, dxsqinv, tz, tx, np, pin, pout, rhs, K, zero_in);
}

template <typename T>
void launch_cc_rb_fused(hipStream_t s, const Grid &g, const T *pin, T *pout, const T *rhs, T *const a[3],
                        int min_tiles) {
  typename CCRB<T>::Args args = {};  // zeroed padding: kernel arguments are copied bytewise
  args.inv = (T)(1.0 / (g.dx * g.dx));
  args.rhs = rhs;
  args.ax = a[0];
  args.ay = a[1];
  args.az = a[2];
  int tx = FX;
  const int tz = rb1_shape(g, min_tiles, sizeof(T), tx);
  if (tz <= 0) {
    report_error("launch_cc_rb_fused: level not eligible");
    return;
  }
  launch_rb<T, CCRB<T>>(s, g, tz, tx, pin, pout, args);
}

// z-slab thickness and tile width of the fused residual + restriction, or 0
// when the level cannot use it (3D, x a multiple of 64 and y of 8, or -- the
// three-component launch only -- x of 32 and y of 16; z of 2; coarse = fine / 2)
static int resid_restrict_tz(const Grid &g, const Grid &gc, int es, int &tx) {
  tx = FX;
  if (g.dim != 3 || g.n[2] % 2 != 0) return 0;
  for (int d = 0; d < 3; d++)
    if (gc.n[d] * 2 != g.n[d]) return 0;
  if (g.n[0] % FX != 0 || g.n[1] % FY != 0) {
    if (!tuning().rr_three || g.n[0] % FX2 != 0 || g.n[1] % FY2 != 0) return 0;
    tx = FX2;
  }
  const int per_plane_set = (g.n[0] / tx) * (g.n[1] / (512 / tx));
  int tz = 32;
  while (tz > 2 && (g.n[2] % tz != 0 || per_plane_set * (g.n[2] / tz) < rb_target(es))) tz /= 2;
  return (g.n[2] % tz == 0) ? tz : 0;
}

namespace {
template <typename T, int A, int VISC>
typename StagRB<T, A, VISC>::Args rr_args(const Grid &g, T *const phi_all[3], const T *rhs, const StagCoef<T> &K) {
  typename StagRB<T, A, VISC>::Args a = {};
  a.omega = (T)0;
  a.dxsqinv = (T)(1.0 / (g.dx * g.dx));
  a.rhs = rhs;
  a.alpha = K.alpha[A];
  a.beta = K.beta;
  a.gamma = K.gamma;
  if (VISC >= 2) {
    // the other components, all in the state the residual sees
    a.e1 = phi_all[A == 0 ? 1 : 0];
    a.e2 = phi_all[A == 2 ? 1 : 2];
  } else {
    a.e1 = (VISC == -1) ? K.bed[A == 2 ? 1 : 0] : nullptr;
    a.e2 = (VISC == -1) ? K.bed[A == 0 ? 1 : 2] : nullptr;
  }
  return a;
}

template <typename T, int VISC>
void launch_rr_three(hipStream_t s, const Grid &g, const Grid &gc, int tz, int tx, T *const phi[3], T *const rhs[3],
                     T *const crhs[3], const StagCoef<T> &K) {
  const auto a0 = rr_args<T, 0, VISC>(g, phi, rhs[0], K);
  const auto a1 = rr_args<T, 1, VISC>(g, phi, rhs[1], K);
  const auto a2 = rr_args<T, 2, VISC>(g, phi, rhs[2], K);
  const T *q0[5], *q1[5], *q2[5];
  StagRB<T, 0, VISC>::ptrs(a0, q0);
  StagRB<T, 1, VISC>::ptrs(a1, q1);
  StagRB<T, 2, VISC>::ptrs(a2, q2);
  const int fy = 512 / tx;
  const int ntx = g.n[0] / tx, nty = g.n[1] / fy, ntz = g.n[2] / tz;
  const int ntiles = ntx * nty * ntz;
  auto kern = (tx == FX2) ? k_resid_restrict3<T, VISC, FX2> : k_resid_restrict3<T, VISC, FX>;
  hipLaunchKernelGGL(kern, dim3(3 * ntiles), dim3(256), 0, s, g, gc, tz, ntx, nty, ntiles, phi[0], crhs[0], phi[1],
                     crhs[1], phi[2], crhs[2], a0, a1, a2, q0[0], q0[1], q0[2], q0[3], q0[4], q1[0], q1[1], q1[2],
                     q1[3], q1[4], q2[0], q2[1], q2[2], q2[3], q2[4]);
  post_launch(s, kResidRestrictName);
}

template <typename T, int A, int VISC>
void launch_rr_one(hipStream_t s, const Grid &g, const Grid &gc, int tz, T *const phi_all[3], const T *rhs, T *crhs,
                   const StagCoef<T> &K) {
  using P = StagRB<T, A, VISC>;
  const T *phi = phi_all[A];
  const typename P::Args a = rr_args<T, A, VISC>(g, phi_all, rhs, K);
  const T *q[5];
  P::ptrs(a, q);
  const int ntx = g.n[0] / FX, nty = g.n[1] / FY, ntz = g.n[2] / tz;
  const int ntiles = ntx * nty * ntz;
  hipLaunchKernelGGL((k_resid_restrict<T, P, A>), dim3(ntiles), dim3(256), 0, s, g, gc, tz, ntx, nty, ntiles, phi,
                     crhs, a, q[0], q[1], q[2], q[3], q[4]);
  post_launch(s, kResidRestrictName);
}
}  // namespace

bool stag_resid_restrict_ok(const Grid &g, const Grid &gc, int visc) {
  int tx;
  return (visc == -1 || visc == 1 || visc == 2 || visc == 3) && resid_restrict_tz(g, gc, 4, tx) > 0;
}

template <typename T>
bool launch_stag_resid_restrict(hipStream_t s, const Grid &g, const Grid &gc, int visc, T *const phi[3],
                                T *const rhs[3], T *const crhs[3], const StagCoef<T> &K) {
  if (!stag_resid_restrict_ok(g, gc, visc)) return false;
  int tx = FX;
  const int tz = resid_restrict_tz(g, gc, sizeof(T), tx);
#define RR(V)                                                         \
  do {                                                                \
    if (tuning().rr_three) {                                          \
      launch_rr_three<T, V>(s, g, gc, tz, tx, phi, rhs, crhs, K);     \
      break;                                                          \
    }                                                                 \
    launch_rr_one<T, 0, V>(s, g, gc, tz, phi, rhs[0], crhs[0], K);    \
    launch_rr_one<T, 1, V>(s, g, gc, tz, phi, rhs[1], crhs[1], K);    \
    launch_rr_one<T, 2, V>(s, g, gc, tz, phi, rhs[2], crhs[2], K);    \
  } while (0)
  if (visc == -1) RR(-1);
  else if (visc == 1) RR(1);
  else if (visc == 2) RR(2);
  else RR(3);
#undef RR
  return true;
}

template void launch_stag_rb_fused3<double>(hipStream_t, const Grid &, int, double, double *const[3],
                                            double *const[3], double *const[3], const StagCoef<double> &, int,
                                            bool);
template void launch_stag_rb_fused3<float>(hipStream_t, const Grid &, int, float, float *const[3], float *const[3],
                                           float *const[3], const StagCoef<float> &, int, bool);
template void launch_stag_rb_fused<double>(hipStream_t, const Grid &, int, int, double, const double *, double *,
                                           const double *, const StagCoef<double> &, int, const double *const[3]);
template void launch_stag_rb_fused<float>(hipStream_t, const Grid &, int, int, float, const float *, float *,
                                          const float *, const StagCoef<float> &, int, const float *const[3]);
template void launch_cc_rb_fused<double>(hipStream_t, const Grid &, const double *, double *, const double *,
                                          double *const[3], int);
template bool launch_stag_resid_restrict<double>(hipStream_t, const Grid &, const Grid &, int, double *const[3],
                                                 double *const[3], double *const[3], const StagCoef<double> &);
template bool launch_stag_resid_restrict<float>(hipStream_t, const Grid &, const Grid &, int, float *const[3],
                                                float *const[3], float *const[3], const StagCoef<float> &);
template void launch_cc_rb_fused<float>(hipStream_t, const Grid &, const float *, float *, const float *,
                                         float *const[3], int);

}  // namespace lmm
