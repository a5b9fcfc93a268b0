// The learned coarse-grid correction network on gfx950 (no counterpart in
// the reference).  Periodic 3x3x3 convolutions as implicit GEMMs on the
// matrix cores: v_mfma_f32_16x16x16_bf16 with BF16 operands and FP32
// accumulation.  One wave computes 16 output voxels x (16*MT) output
// channels; the K dimension runs over (tap, input channel).  The weight
// gradient is the GEMM dZ^T * X_shift with K over voxels, split over waves
// and summed by a second pass.  Everything else (pooling, upsampling, ReLU
// masks, the loss, Adam) is elementwise.
//
// Lane layout of v_mfma_f32_16x16x16_bf16 (wave64): lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16] and D[4(l/16)..+3][l%16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/lmm.h"
#include "lmm_cnn.h"
#include "lmm_kernels.h"

namespace lmm {

namespace {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int CTPB = 256;

#define CLAUNCH(K, G, B, SHM, S, ...)                  \
  do {                                                 \
    hipLaunchKernelGGL(K, G, B, SHM, S, __VA_ARGS__);  \
    post_launch(S, #K);                                \
  } while (0)

inline int cblk(long n) { return (int)((n + CTPB - 1) / CTPB); }

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
inline uint16_t f2bf_host(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// batch of nb periodic grids nx x ny x nz, voxel v = b*nv1 + (z*ny + y)*nx + x
struct CDims {
  int nb, nx, ny, nz;
  long nv1, nv;
};
CDims cdims(int nb, int nx, int ny, int nz) {
  CDims d;
  d.nb = nb;
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.nv1 = (long)nx * ny * nz;
  d.nv = d.nv1 * nb;
  return d;
}
CDims half(const CDims &d) { return cdims(d.nb, d.nx / 2, d.ny / 2, d.nz / 2); }

__device__ __forceinline__ void decode_v(const CDims &d, long v, int &b, int &z, int &y, int &x) {
  b = (int)(v / d.nv1);
  long r = v - (long)b * d.nv1;
  x = (int)(r % d.nx);
  r /= d.nx;
  y = (int)(r % d.ny);
  z = (int)(r / d.ny);
}
__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
// the voxel at (z+dz, y+dy, x+dx) of tap
__device__ __forceinline__ long nbr_v(const CDims &d, int b, int z, int y, int x, int tap) {
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  return ((long)(b * d.nz + wrap(z + dz, d.nz)) * d.ny + wrap(y + dy, d.ny)) * d.nx + wrap(x + dx, d.nx);
}

// Y[v][yo + co] = act(sum_{tap,ci} W[co][tap][ci] X[nbr(v,tap)][xo + ci]) for co < co_store.
// W: BF16 [MT*16][27*CI] (rows >= the real output channels are zero).
template <int CI, int MT>
__global__ void __launch_bounds__(256) k_conv_fwd(CDims d, const uint16_t *__restrict__ X, int xs, int xo,
                                                  const uint16_t *__restrict__ W, void *Y, int ys, int yo,
                                                  int co_store, int relu, int out_f32) {
  const int lane = threadIdx.x & 63;
  const long v0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (v0 >= d.nv) return;  // whole wave
  const int j = lane & 15, g = lane >> 4;
  int b, z, y, x;
  decode_v(d, v0 + j, b, z, y, x);
  constexpr int K = 27 * CI, NCH = (K + 15) / 16;
  f32x4_t acc[MT];
#pragma unroll
  for (int m = 0; m < MT; m++) acc[m] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int ch = 0; ch < NCH; ch++) {
    const int k = ch * 16 + g * 4;
    const bool ok = (K % 16 == 0) || (k < K);
    bf16x4_t bv = {0, 0, 0, 0};
    if (ok) {
      const int tap = k / CI, ci = k - tap * CI;
      bv = *reinterpret_cast<const bf16x4_t *>(X + nbr_v(d, b, z, y, x, tap) * xs + xo + ci);
    }
#pragma unroll
    for (int m = 0; m < MT; m++) {
      bf16x4_t av = {0, 0, 0, 0};
      if (ok) av = *reinterpret_cast<const bf16x4_t *>(W + (long)(m * 16 + j) * K + k);
      acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[m], 0, 0, 0);
    }
  }
  const long vo = v0 + j;
#pragma unroll
  for (int m = 0; m < MT; m++) {
    const int co0 = m * 16 + 4 * g;
    if (co0 >= co_store) continue;
    float o[4] = {acc[m][0], acc[m][1], acc[m][2], acc[m][3]};
    if (relu)
      for (int q = 0; q < 4; q++) o[q] = fmaxf(o[q], 0.f);
    const long base = vo * ys + yo + co0;
    if (out_f32) {
      float *Yf = reinterpret_cast<float *>(Y);
      for (int q = 0; q < 4; q++)
        if (co0 + q < co_store) Yf[base + q] = o[q];
    } else {
      uint16_t *Yb = reinterpret_cast<uint16_t *>(Y);
      if (co0 + 4 <= co_store) {
        bf16x4_t h = {(short)f2bf(o[0]), (short)f2bf(o[1]), (short)f2bf(o[2]), (short)f2bf(o[3])};
        *reinterpret_cast<bf16x4_t *>(Yb + base) = h;
      } else {
        for (int q = 0; q < 4; q++)
          if (co0 + q < co_store) Yb[base + q] = f2bf(o[q]);
      }
    }
  }
}

// part[s][tap][co][ci] (co < 16*mt_n, ci < 16*nt_n) = sum over voxels of
// split s of DZ[v][co] X[nbr(v,tap)][xo + ci].  A wave owns (s, tap, co
// tile, ci tile); each lane's 4 voxels of a 16-voxel step share one x-row
// (nx % 4 == 0).
__global__ void __launch_bounds__(256) k_conv_wgrad(CDims d, const uint16_t *__restrict__ DZ, int zs, int co_real,
                                                    const uint16_t *__restrict__ X, int xs, int xo, int ci_real,
                                                    int mt_n, int nt_n, int nsplit, long chunk, float *part) {
  const long wid = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wid >= (long)nsplit * 27 * mt_n * nt_n) return;
  const int nt = (int)(wid % nt_n);
  long t2 = wid / nt_n;
  const int mt = (int)(t2 % mt_n);
  t2 /= mt_n;
  const int tap = (int)(t2 % 27);
  const int s = (int)(t2 / 27);
  const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
  const int co = mt * 16 + i, ci = nt * 16 + i;
  const bool aok = co < co_real, bok = ci < ci_real;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const long vbeg = (long)s * chunk, vend = min(vbeg + chunk, d.nv);
  f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
  int b, z, y, x;
  decode_v(d, vbeg + 4 * g, b, z, y, x);
  for (long vb = vbeg; vb < vend; vb += 16) {
    const long v = vb + 4 * g;
    const long row = ((long)(b * d.nz + wrap(z + dz, d.nz)) * d.ny + wrap(y + dy, d.ny)) * d.nx;
    bf16x4_t av, bv;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      av[q] = aok ? (short)DZ[(v + q) * zs + co] : (short)0;
      bv[q] = bok ? (short)X[(row + wrap(x + q + dx, d.nx)) * xs + xo + ci] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    // advance this lane's voxel by 16
    x += 16;
    while (x >= d.nx) {
      x -= d.nx;
      if (++y == d.ny) {
        y = 0;
        if (++z == d.nz) {
          z = 0;
          ++b;
        }
      }
    }
  }
  const long cop = (long)mt_n * 16, cip = (long)nt_n * 16;
#pragma unroll
  for (int q = 0; q < 4; q++) part[(((long)s * 27 + tap) * cop + mt * 16 + 4 * g + q) * cip + ci] = acc[q];
}

// dW[co][tap][ci] (FP32, layer layout) = sum_s part[s][tap][co][ci]
__global__ void k_wgrad_reduce(const float *__restrict__ part, int nsplit, int co_real, int ci_real, int cop, int cip,
                               float *dw) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)co_real * 27 * ci_real;
  if (t >= n) return;
  const int ci = (int)(t % ci_real);
  const int tap = (int)((t / ci_real) % 27);
  const int co = (int)(t / ((long)ci_real * 27));
  float acc = 0.f;
  for (int s = 0; s < nsplit; s++) acc += part[(((long)s * 27 + tap) * cop + co) * cip + ci];
  dw[t] = acc;
}

// BF16 forward weights [co_pad][27][ci] and the transposed, tap-flipped
// weights of the input gradient [ci_padT][27][cit]: WT[ci][26-tap][co] = W[co][tap][ci]
__global__ void k_prep_weights(const float *__restrict__ p, int ci_n, int co_n, int co_pad, int cit, int ci_padt,
                               uint16_t *wb, uint16_t *wt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nf = (long)co_pad * 27 * ci_n;
  if (t < nf) {
    const int ci = (int)(t % ci_n), tap = (int)((t / ci_n) % 27), co = (int)(t / ((long)ci_n * 27));
    wb[t] = (co < co_n) ? f2bf(p[((long)co * 27 + tap) * ci_n + ci]) : (uint16_t)0;
  }
  const long nt = (long)ci_padt * 27 * cit;
  if (wt && t < nt) {
    const int co = (int)(t % cit), tp = (int)((t / cit) % 27), ci = (int)(t / ((long)cit * 27));
    wt[t] = (ci < ci_n && co < co_n) ? f2bf(p[((long)co * 27 + (26 - tp)) * ci_n + ci]) : (uint16_t)0;
  }
}

// out[vc][c] = mean of the 8 children in[child*is + io + c]  (grid dc coarse)
__global__ void k_pool(CDims dc, const uint16_t *__restrict__ in, int is, int io, int C, uint16_t *out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dc.nv * C) return;
  const int c = (int)(t % C);
  const long vc = t / C;
  int b, Z, Y, X;
  decode_v(dc, vc, b, Z, Y, X);
  const int fx = dc.nx * 2, fy = dc.ny * 2, fz = dc.nz * 2;
  float s = 0.f;
  for (int a = 0; a < 2; a++)
    for (int e = 0; e < 2; e++)
      for (int f = 0; f < 2; f++) {
        const long vf = ((long)(b * fz + 2 * Z + a) * fy + 2 * Y + e) * fx + 2 * X + f;
        s += bf2f(in[vf * is + io + c]);
      }
  out[vc * C + c] = f2bf(s * 0.125f);
}

// out[v*os + oo + c] = in[parent(v)][c]  (grid df fine)
__global__ void k_upsample(CDims df, const uint16_t *__restrict__ in, int C, uint16_t *out, int os, int oo) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= df.nv * C) return;
  const int c = (int)(t % C);
  const long v = t / C;
  int b, z, y, x;
  decode_v(df, v, b, z, y, x);
  const long vc = ((long)(b * (df.nz / 2) + z / 2) * (df.ny / 2) + y / 2) * (df.nx / 2) + x / 2;
  out[v * os + oo + c] = in[vc * C + c];
}

// out[vc][c] = sum of the 8 children in[child*is + io + c]  (grid dc coarse)
__global__ void k_upsample_grad(CDims dc, const uint16_t *__restrict__ in, int is, int io, int C, uint16_t *out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dc.nv * C) return;
  const int c = (int)(t % C);
  const long vc = t / C;
  int b, Z, Y, X;
  decode_v(dc, vc, b, Z, Y, X);
  const int fx = dc.nx * 2, fy = dc.ny * 2, fz = dc.nz * 2;
  float s = 0.f;
  for (int a = 0; a < 2; a++)
    for (int e = 0; e < 2; e++)
      for (int f = 0; f < 2; f++) {
        const long vf = ((long)(b * fz + 2 * Z + a) * fy + 2 * Y + e) * fx + 2 * X + f;
        s += bf2f(in[vf * is + io + c]);
      }
  out[vc * C + c] = f2bf(s);
}

// out[v][c] = skip[v*ss + so + c] + dp[parent(v)][c] / 8  (grid df fine)
__global__ void k_pool_grad_add(CDims df, const uint16_t *__restrict__ skip, int ss, int so,
                                const uint16_t *__restrict__ dp, int C, uint16_t *out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= df.nv * C) return;
  const int c = (int)(t % C);
  const long v = t / C;
  int b, z, y, x;
  decode_v(df, v, b, z, y, x);
  const long vc = ((long)(b * (df.nz / 2) + z / 2) * (df.ny / 2) + y / 2) * (df.nx / 2) + x / 2;
  out[v * C + c] = f2bf(bf2f(skip[v * ss + so + c]) + bf2f(dp[vc * C + c]) * 0.125f);
}

// dz[v][c] = act[v*as + ao + c] > 0 ? dz[v][c] : 0
__global__ void k_relu_mask(long nv, int C, uint16_t *dz, const uint16_t *__restrict__ act, int as, int ao) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv * C) return;
  const long v = t / C;
  const int c = (int)(t - v * C);
  if (!(bf2f(act[v * as + ao + c]) > 0.f)) dz[t] = 0;
}

__device__ __forceinline__ double block_sum_d(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
  __syncthreads();
  return r;
}

// dy[v][c] = (y - t) * inv for c < cout (0 above), partial[blk] = sum (y - t)^2
__global__ void __launch_bounds__(256) k_loss_grad(long nv, int cout, const float *__restrict__ y,
                                                   const float *__restrict__ tg, float inv, uint16_t *dy,
                                                   double *partial) {
  __shared__ double sh[4];
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (v < nv) {
    for (int c = 0; c < CNN_OUT_STRIDE; c++) {
      const float df = (c < cout) ? y[v * CNN_OUT_STRIDE + c] - tg[v * CNN_OUT_STRIDE + c] : 0.f;
      dy[v * CNN_OUT_STRIDE + c] = f2bf(df * inv);
      acc += (double)df * (double)df;
    }
  }
  const double r = block_sum_d(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ void k_sum_partials(const double *__restrict__ p, long n, double scale, double *out) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (long q = threadIdx.x; q < n; q += blockDim.x) acc += p[q];
  const double r = block_sum_d(acc, sh);
  if (threadIdx.x == 0) out[0] = r * scale;
}

__global__ void k_adam(long n, float *p, const float *__restrict__ g, float *m, float *v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] = p[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// batch item b of a training step: pair hash(seed, step, b) mod ntrain
__global__ void k_gather(long nv1, int cin, int nb, const uint16_t *__restrict__ dsx, const float *__restrict__ dst,
                         long ntrain, uint64_t seed, uint64_t step, float tinv, uint16_t *x, float *t) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nv1 * nb) return;
  const int b = (int)(q / nv1);
  const long v = q - (long)b * nv1;
  const long pair = (long)(splitmix64(seed ^ splitmix64(step * 1024 + b)) % (uint64_t)ntrain);
  for (int c = 0; c < cin; c++) x[q * cin + c] = dsx[(pair * nv1 + v) * cin + c];
  for (int c = 0; c < CNN_OUT_STRIDE; c++)
    t[q * CNN_OUT_STRIDE + c] = dst[(pair * nv1 + v) * CNN_OUT_STRIDE + c] * tinv;
}

// sum of squares of n floats into partial[blk]
__global__ void __launch_bounds__(256) k_sumsq_f(long n, const float *__restrict__ a, double *partial) {
  __shared__ double sh[4];
  double acc = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) acc += (double)a[q] * a[q];
  const double r = block_sum_d(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// p[q] *= a, m[q] *= a, v[q] *= a*a for q < n (one layer's weights and Adam moments)
__global__ void k_rescale_layer(long n, float *p, float *m, float *v, float a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  p[t] *= a;
  m[t] *= a;
  v[t] *= a * a;
}

// |t - y|^2 and |t|^2 of one pair into out[0], out[1] (atomics of a few doubles)
__global__ void __launch_bounds__(256) k_err_pair(long nv, const float *__restrict__ y, const float *__restrict__ tg,
                                                  double *partial) {
  __shared__ double sh[4];
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0, n = 0.0;
  if (v < nv)
    for (int c = 0; c < CNN_OUT_STRIDE; c++) {
      const double a = tg[v * CNN_OUT_STRIDE + c], d = a - (double)y[v * CNN_OUT_STRIDE + c];
      e += d * d;
      n += a * a;
    }
  const double re = block_sum_d(e, sh);
  const double rn = block_sum_d(n, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = re;
    partial[2 * blockIdx.x + 1] = rn;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// the network object
// ---------------------------------------------------------------------------

struct CnnLayer {
  int ci, co;      // real channels
  int co_pad;      // 16-multiple rows of the forward weights
  int cit;         // input channels of the input-gradient convolution (co, or 4 for the output layer)
  int ci_padt;     // 16-multiple rows of the transposed weights
  long off;        // offset in the flat parameter vector
  uint16_t *wb = nullptr, *wt = nullptr;
};

struct CnnNet {
  CnnConfig cfg;
  hipStream_t s = nullptr;
  std::vector<CnnLayer> L;
  long np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  // activations / gradients for capacity cap_nv voxels at level 1
  long cap_nv = 0;
  uint16_t *x0 = nullptr, *a1 = nullptr, *cat1 = nullptr, *d1 = nullptr, *u1 = nullptr;
  uint16_t *p1 = nullptr, *a2 = nullptr, *cat2 = nullptr, *d2 = nullptr, *u2 = nullptr;
  uint16_t *p2 = nullptr, *a3 = nullptr, *b3 = nullptr;
  uint16_t *gy = nullptr, *gu1 = nullptr, *gd1 = nullptr, *gcat1 = nullptr, *gu2 = nullptr, *gd2 = nullptr,
           *gcat2 = nullptr, *gb3 = nullptr, *ga3 = nullptr, *gp2 = nullptr, *gs2 = nullptr, *ga2 = nullptr,
           *gp1 = nullptr, *gs1 = nullptr, *ga1 = nullptr;
  float *y = nullptr, *tb = nullptr;
  float *part = nullptr;
  long part_cap = 0;
  double *lpart = nullptr, *dscal = nullptr;
  bool trained = false;  // the output layer predicts targets in their own units
  // dataset
  long ds_cap = 0, ds_count = 0;
  int ds_n[3] = {0, 0, 0};
  uint16_t *dsx = nullptr;
  float *dst = nullptr;
};

namespace {

template <typename T>
T *cmalloc(long n) {
  void *q = nullptr;
  if (hipMalloc(&q, sizeof(T) * (size_t)std::max(1L, n)) != hipSuccess) return nullptr;
  return (T *)q;
}
void cfree(void *q) {
  if (q) hipFree(q);
}

void free_acts(CnnNet *n) {
  for (void *q : {(void *)n->x0, (void *)n->a1, (void *)n->cat1, (void *)n->d1, (void *)n->u1, (void *)n->p1,
                  (void *)n->a2, (void *)n->cat2, (void *)n->d2, (void *)n->u2, (void *)n->p2, (void *)n->a3,
                  (void *)n->b3, (void *)n->gy, (void *)n->gu1, (void *)n->gd1, (void *)n->gcat1, (void *)n->gu2,
                  (void *)n->gd2, (void *)n->gcat2, (void *)n->gb3, (void *)n->ga3, (void *)n->gp2, (void *)n->gs2,
                  (void *)n->ga2, (void *)n->gp1, (void *)n->gs1, (void *)n->ga1, (void *)n->y, (void *)n->tb,
                  (void *)n->lpart})
    cfree(q);
  n->x0 = n->a1 = n->cat1 = n->d1 = n->u1 = n->p1 = n->a2 = n->cat2 = n->d2 = n->u2 = n->p2 = n->a3 = n->b3 = nullptr;
  n->gy = n->gu1 = n->gd1 = n->gcat1 = n->gu2 = n->gd2 = n->gcat2 = n->gb3 = n->ga3 = n->gp2 = n->gs2 = n->ga2 =
      n->gp1 = n->gs1 = n->ga1 = nullptr;
  n->y = n->tb = nullptr;
  n->lpart = nullptr;
  n->cap_nv = 0;
}

// buffers for nv1 voxels at level 1 (nv1 a multiple of 64)
int ensure_acts(CnnNet *n, long nv1) {
  if (nv1 <= n->cap_nv) return 0;
  free_acts(n);
  const CnnConfig &c = n->cfg;
  const long v1 = nv1, v2 = nv1 / 8, v3 = nv1 / 64;
  bool ok = true;
  auto A = [&](uint16_t *&q, long cnt) {
    q = cmalloc<uint16_t>(cnt);
    ok = ok && q;
  };
  A(n->x0, v1 * c.cin);
  A(n->a1, v1 * c.c1);
  A(n->cat1, v1 * (c.c2 + c.c1));
  A(n->d1, v1 * c.c1);
  A(n->u1, v1 * c.c1);
  A(n->p1, v2 * c.c1);
  A(n->a2, v2 * c.c2);
  A(n->cat2, v2 * (c.c3 + c.c2));
  A(n->d2, v2 * c.c2);
  A(n->u2, v2 * c.c2);
  A(n->p2, v3 * c.c2);
  A(n->a3, v3 * c.c3);
  A(n->b3, v3 * c.c3);
  A(n->gy, v1 * CNN_OUT_STRIDE);
  A(n->gu1, v1 * c.c1);
  A(n->gd1, v1 * c.c1);
  A(n->gcat1, v1 * (c.c2 + c.c1));
  A(n->gu2, v2 * c.c2);
  A(n->gd2, v2 * c.c2);
  A(n->gcat2, v2 * (c.c3 + c.c2));
  A(n->gb3, v3 * c.c3);
  A(n->ga3, v3 * c.c3);
  A(n->gp2, v3 * c.c2);
  A(n->gs2, v2 * c.c2);
  A(n->ga2, v2 * c.c2);
  A(n->gp1, v2 * c.c1);
  A(n->gs1, v1 * c.c1);
  A(n->ga1, v1 * c.c1);
  n->y = cmalloc<float>(v1 * CNN_OUT_STRIDE);
  n->tb = cmalloc<float>(v1 * CNN_OUT_STRIDE);
  n->lpart = cmalloc<double>(2 * (long)cblk(v1) + 8);
  ok = ok && n->y && n->tb && n->lpart;
  if (!ok) {
    free_acts(n);
    return LMM_ERR_NOMEM;
  }
  n->cap_nv = nv1;
  return 0;
}

int mt_of(int c) { return (c + 15) / 16; }

void prep_weights(CnnNet *n) {
  for (auto &l : n->L) {
    const long nf = (long)l.co_pad * 27 * l.ci, nt = (long)l.ci_padt * 27 * l.cit;
    CLAUNCH(k_prep_weights, dim3(cblk(std::max(nf, nt))), dim3(CTPB), 0, n->s, n->p + l.off, l.ci, l.co, l.co_pad,
            l.cit, l.ci_padt, l.wb, l.wt);
  }
}

// one convolution: X (stride xs, offset xo, ci channels) -> Y (stride ys, offset yo, co_store channels)
void conv(hipStream_t s, const CDims &d, int ci, int mt, const uint16_t *X, int xs, int xo, const uint16_t *W,
          void *Y, int ys, int yo, int co_store, bool relu, bool out_f32) {
  const dim3 grid((unsigned)((d.nv / 16 + 3) / 4)), blk(256);
  const int r = relu ? 1 : 0, f = out_f32 ? 1 : 0;
#define CONV_CASE(CIV, MTV)                                                                                 \
  if (ci == CIV && mt == MTV) {                                                                             \
    CLAUNCH((k_conv_fwd<CIV, MTV>), grid, blk, 0, s, d, X, xs, xo, W, Y, ys, yo, co_store, r, f);          \
    return;                                                                                                 \
  }
  CONV_CASE(4, 1)
  CONV_CASE(8, 1)
  CONV_CASE(16, 1)
  CONV_CASE(16, 2)
  CONV_CASE(16, 3)
  CONV_CASE(32, 1)
  CONV_CASE(32, 2)
  CONV_CASE(32, 4)
  CONV_CASE(32, 6)
  CONV_CASE(48, 1)
  CONV_CASE(64, 2)
  CONV_CASE(64, 4)
  CONV_CASE(96, 2)
#undef CONV_CASE
  report_error("cnn conv: channel combination without a kernel instance");
}

// dW of layer l from DZ (stride zs, co_real rows) and its input X
void wgrad(CnnNet *n, const CDims &d, const CnnLayer &l, const uint16_t *DZ, int zs, const uint16_t *X, int xs,
           int xo) {
  const int mt_n = mt_of(l.co), nt_n = mt_of(l.ci);
  const long per_split = 27L * mt_n * 16 * nt_n * 16;
  long nsplit = std::max(1L, d.nv / 256);
  nsplit = std::min(nsplit, std::max(1L, 4096L / (27L * mt_n * nt_n)));
  nsplit = std::min(nsplit, std::max(1L, n->part_cap / per_split));
  long chunk = (d.nv + nsplit - 1) / nsplit;
  chunk = (chunk + 15) / 16 * 16;
  nsplit = (d.nv + chunk - 1) / chunk;
  const long waves = nsplit * 27 * mt_n * nt_n;
  CLAUNCH(k_conv_wgrad, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, n->s, d, DZ, zs, l.co, X, xs, xo, l.ci,
          mt_n, nt_n, (int)nsplit, chunk, n->part);
  const long np = (long)l.co * 27 * l.ci;
  CLAUNCH(k_wgrad_reduce, dim3(cblk(np)), dim3(CTPB), 0, n->s, n->part, (int)nsplit, l.co, l.ci, mt_n * 16,
          nt_n * 16, n->g + l.off);
}

void mask(hipStream_t s, const CDims &d, int C, uint16_t *dz, const uint16_t *act, int as, int ao) {
  CLAUNCH(k_relu_mask, dim3(cblk(d.nv * C)), dim3(CTPB), 0, s, d.nv, C, dz, act, as, ao);
}

bool dims_ok(int nb, const int n[3]) {
  return nb >= 1 && n[0] >= 4 && n[1] >= 4 && n[2] >= 4 && n[0] % 4 == 0 && n[1] % 4 == 0 && n[2] % 4 == 0 &&
         ((long)nb * (n[0] / 4) * (n[1] / 4) * (n[2] / 4)) % 16 == 0;
}

void forward_impl(CnnNet *n, const CDims &d1, const uint16_t *x, float *y) {
  const CnnConfig &c = n->cfg;
  const CDims d2 = half(d1), d3 = half(d2);
  const int c21 = c.c2 + c.c1, c32 = c.c3 + c.c2;
  hipStream_t s = n->s;
  auto &L = n->L;
  conv(s, d1, c.cin, mt_of(c.c1), x, c.cin, 0, L[0].wb, n->a1, c.c1, 0, c.c1, true, false);
  conv(s, d1, c.c1, mt_of(c.c1), n->a1, c.c1, 0, L[1].wb, n->cat1, c21, c.c2, c.c1, true, false);
  CLAUNCH(k_pool, dim3(cblk(d2.nv * c.c1)), dim3(CTPB), 0, s, d2, n->cat1, c21, c.c2, c.c1, n->p1);
  conv(s, d2, c.c1, mt_of(c.c2), n->p1, c.c1, 0, L[2].wb, n->a2, c.c2, 0, c.c2, true, false);
  conv(s, d2, c.c2, mt_of(c.c2), n->a2, c.c2, 0, L[3].wb, n->cat2, c32, c.c3, c.c2, true, false);
  CLAUNCH(k_pool, dim3(cblk(d3.nv * c.c2)), dim3(CTPB), 0, s, d3, n->cat2, c32, c.c3, c.c2, n->p2);
  conv(s, d3, c.c2, mt_of(c.c3), n->p2, c.c2, 0, L[4].wb, n->a3, c.c3, 0, c.c3, true, false);
  conv(s, d3, c.c3, mt_of(c.c3), n->a3, c.c3, 0, L[5].wb, n->b3, c.c3, 0, c.c3, true, false);
  CLAUNCH(k_upsample, dim3(cblk(d2.nv * c.c3)), dim3(CTPB), 0, s, d2, n->b3, c.c3, n->cat2, c32, 0);
  conv(s, d2, c32, mt_of(c.c2), n->cat2, c32, 0, L[6].wb, n->d2, c.c2, 0, c.c2, true, false);
  conv(s, d2, c.c2, mt_of(c.c2), n->d2, c.c2, 0, L[7].wb, n->u2, c.c2, 0, c.c2, true, false);
  CLAUNCH(k_upsample, dim3(cblk(d1.nv * c.c2)), dim3(CTPB), 0, s, d1, n->u2, c.c2, n->cat1, c21, 0);
  conv(s, d1, c21, mt_of(c.c1), n->cat1, c21, 0, L[8].wb, n->d1, c.c1, 0, c.c1, true, false);
  conv(s, d1, c.c1, mt_of(c.c1), n->d1, c.c1, 0, L[9].wb, n->u1, c.c1, 0, c.c1, true, false);
  conv(s, d1, c.c1, 1, n->u1, c.c1, 0, L[10].wb, y, CNN_OUT_STRIDE, 0, CNN_OUT_STRIDE, false, true);
}

// backward of the loss whose output gradient is in n->gy; x the network input
void backward_impl(CnnNet *n, const CDims &d1, const uint16_t *x) {
  const CnnConfig &c = n->cfg;
  const CDims d2 = half(d1), d3 = half(d2);
  const int c21 = c.c2 + c.c1, c32 = c.c3 + c.c2;
  hipStream_t s = n->s;
  auto &L = n->L;
  // L10 (linear)
  wgrad(n, d1, L[10], n->gy, CNN_OUT_STRIDE, n->u1, c.c1, 0);
  conv(s, d1, L[10].cit, mt_of(c.c1), n->gy, CNN_OUT_STRIDE, 0, L[10].wt, n->gu1, c.c1, 0, c.c1, false, false);
  mask(s, d1, c.c1, n->gu1, n->u1, c.c1, 0);
  // L9
  wgrad(n, d1, L[9], n->gu1, c.c1, n->d1, c.c1, 0);
  conv(s, d1, c.c1, mt_of(c.c1), n->gu1, c.c1, 0, L[9].wt, n->gd1, c.c1, 0, c.c1, false, false);
  mask(s, d1, c.c1, n->gd1, n->d1, c.c1, 0);
  // L8 (input: cat1 = [up(u2), s1])
  wgrad(n, d1, L[8], n->gd1, c.c1, n->cat1, c21, 0);
  conv(s, d1, c.c1, mt_of(c21), n->gd1, c.c1, 0, L[8].wt, n->gcat1, c21, 0, c21, false, false);
  CLAUNCH(k_upsample_grad, dim3(cblk(d2.nv * c.c2)), dim3(CTPB), 0, s, d2, n->gcat1, c21, 0, c.c2, n->gu2);
  mask(s, d2, c.c2, n->gu2, n->u2, c.c2, 0);
  // L7
  wgrad(n, d2, L[7], n->gu2, c.c2, n->d2, c.c2, 0);
  conv(s, d2, c.c2, mt_of(c.c2), n->gu2, c.c2, 0, L[7].wt, n->gd2, c.c2, 0, c.c2, false, false);
  mask(s, d2, c.c2, n->gd2, n->d2, c.c2, 0);
  // L6 (input: cat2 = [up(b3), s2])
  wgrad(n, d2, L[6], n->gd2, c.c2, n->cat2, c32, 0);
  conv(s, d2, c.c2, mt_of(c32), n->gd2, c.c2, 0, L[6].wt, n->gcat2, c32, 0, c32, false, false);
  CLAUNCH(k_upsample_grad, dim3(cblk(d3.nv * c.c3)), dim3(CTPB), 0, s, d3, n->gcat2, c32, 0, c.c3, n->gb3);
  mask(s, d3, c.c3, n->gb3, n->b3, c.c3, 0);
  // L5
  wgrad(n, d3, L[5], n->gb3, c.c3, n->a3, c.c3, 0);
  conv(s, d3, c.c3, mt_of(c.c3), n->gb3, c.c3, 0, L[5].wt, n->ga3, c.c3, 0, c.c3, false, false);
  mask(s, d3, c.c3, n->ga3, n->a3, c.c3, 0);
  // L4
  wgrad(n, d3, L[4], n->ga3, c.c3, n->p2, c.c2, 0);
  conv(s, d3, c.c3, mt_of(c.c2), n->ga3, c.c3, 0, L[4].wt, n->gp2, c.c2, 0, c.c2, false, false);
  // s2 feeds the pool and the skip: both gradients
  CLAUNCH(k_pool_grad_add, dim3(cblk(d2.nv * c.c2)), dim3(CTPB), 0, s, d2, n->gcat2, c32, c.c3, n->gp2, c.c2,
          n->gs2);
  mask(s, d2, c.c2, n->gs2, n->cat2, c32, c.c3);
  // L3
  wgrad(n, d2, L[3], n->gs2, c.c2, n->a2, c.c2, 0);
  conv(s, d2, c.c2, mt_of(c.c2), n->gs2, c.c2, 0, L[3].wt, n->ga2, c.c2, 0, c.c2, false, false);
  mask(s, d2, c.c2, n->ga2, n->a2, c.c2, 0);
  // L2
  wgrad(n, d2, L[2], n->ga2, c.c2, n->p1, c.c1, 0);
  conv(s, d2, c.c2, mt_of(c.c1), n->ga2, c.c2, 0, L[2].wt, n->gp1, c.c1, 0, c.c1, false, false);
  CLAUNCH(k_pool_grad_add, dim3(cblk(d1.nv * c.c1)), dim3(CTPB), 0, s, d1, n->gcat1, c21, c.c2, n->gp1, c.c1,
          n->gs1);
  mask(s, d1, c.c1, n->gs1, n->cat1, c21, c.c2);
  // L1
  wgrad(n, d1, L[1], n->gs1, c.c1, n->a1, c.c1, 0);
  conv(s, d1, c.c1, mt_of(c.c1), n->gs1, c.c1, 0, L[1].wt, n->ga1, c.c1, 0, c.c1, false, false);
  mask(s, d1, c.c1, n->ga1, n->a1, c.c1, 0);
  // L0 (the input gradient is not needed)
  wgrad(n, d1, L[0], n->ga1, c.c1, x, c.cin, 0);
}

}  // namespace

// ---------------------------------------------------------------------------
// internal API
// ---------------------------------------------------------------------------

CnnNet *cnn_new(const CnnConfig &cfg, uint64_t seed, hipStream_t s, std::string *err) {
  if (cfg.cin % 4 || cfg.c1 % 16 || cfg.c2 % 16 || cfg.c3 % 16 || cfg.cout < 1 || cfg.cout > CNN_OUT_STRIDE ||
      cfg.cin > 96) {
    if (err) *err = "cnn config: cin a multiple of 4, c1/c2/c3 multiples of 16, 1 <= cout <= 4";
    return nullptr;
  }
  CnnNet *n = new CnnNet();
  n->cfg = cfg;
  n->s = s;
  const int io[11][2] = {{cfg.cin, cfg.c1}, {cfg.c1, cfg.c1}, {cfg.c1, cfg.c2},           {cfg.c2, cfg.c2},
                         {cfg.c2, cfg.c3},  {cfg.c3, cfg.c3}, {cfg.c3 + cfg.c2, cfg.c2}, {cfg.c2, cfg.c2},
                         {cfg.c2 + cfg.c1, cfg.c1}, {cfg.c1, cfg.c1}, {cfg.c1, cfg.cout}};
  long off = 0;
  bool ok = true;
  for (int q = 0; q < 11; q++) {
    CnnLayer l;
    l.ci = io[q][0];
    l.co = io[q][1];
    l.co_pad = mt_of(l.co) * 16;
    l.cit = (q == 10) ? CNN_OUT_STRIDE : l.co;
    l.ci_padt = mt_of(l.ci) * 16;
    l.off = off;
    off += 27L * l.ci * l.co;
    l.wb = cmalloc<uint16_t>((long)l.co_pad * 27 * l.ci);
    l.wt = cmalloc<uint16_t>((long)l.ci_padt * 27 * l.cit);
    ok = ok && l.wb && l.wt;
    n->L.push_back(l);
  }
  n->np = off;
  n->p = cmalloc<float>(off);
  n->g = cmalloc<float>(off);
  n->m = cmalloc<float>(off);
  n->v = cmalloc<float>(off);
  n->part_cap = 4L << 20;
  n->part = cmalloc<float>(n->part_cap);
  n->dscal = cmalloc<double>(16);
  ok = ok && n->p && n->g && n->m && n->v && n->part && n->dscal;
  if (!ok) {
    if (err) *err = "cnn: device allocation failed";
    cnn_free(n);
    return nullptr;
  }
  // He initialisation, N(0, 2/fan_in), from a seeded host generator
  std::vector<float> h(off);
  std::mt19937_64 gen(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (const auto &l : n->L) {
    const float sc = std::sqrt(2.f / (27.f * l.ci));
    for (long q = 0; q < 27L * l.ci * l.co; q++) h[l.off + q] = sc * nd(gen);
  }
  hipMemcpyAsync(n->p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice, s);
  hipMemsetAsync(n->m, 0, sizeof(float) * off, s);
  hipMemsetAsync(n->v, 0, sizeof(float) * off, s);
  hipMemsetAsync(n->g, 0, sizeof(float) * off, s);
  prep_weights(n);
  if (hipStreamSynchronize(s) != hipSuccess) {
    if (err) *err = "cnn: initialisation failed";
    cnn_free(n);
    return nullptr;
  }
  return n;
}

void cnn_free(CnnNet *n) {
  if (!n) return;
  if (n->s) hipStreamSynchronize(n->s);
  free_acts(n);
  for (auto &l : n->L) {
    cfree(l.wb);
    cfree(l.wt);
  }
  for (void *q : {(void *)n->p, (void *)n->g, (void *)n->m, (void *)n->v, (void *)n->part, (void *)n->dscal,
                  (void *)n->dsx, (void *)n->dst})
    cfree(q);
  delete n;
}

long cnn_nparams(const CnnNet *n) { return n->np; }
const CnnConfig &cnn_config(const CnnNet *n) { return n->cfg; }
float *cnn_grad_dev(CnnNet *n) { return n->g; }
hipStream_t cnn_stream(CnnNet *n) { return n->s; }
void cnn_set_stream(CnnNet *n, hipStream_t s) { n->s = s; }

int cnn_get_params(CnnNet *n, float *host) {
  if (hipMemcpyAsync(host, n->p, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->s) != hipSuccess) return LMM_ERR_DEVICE;
  return hipStreamSynchronize(n->s) == hipSuccess ? 0 : LMM_ERR_DEVICE;
}

int cnn_set_params(CnnNet *n, const float *host) {
  if (hipMemcpyAsync(n->p, host, sizeof(float) * n->np, hipMemcpyHostToDevice, n->s) != hipSuccess) return LMM_ERR_DEVICE;
  prep_weights(n);
  return hipStreamSynchronize(n->s) == hipSuccess ? 0 : LMM_ERR_DEVICE;
}

int cnn_forward(CnnNet *n, int nb, const int g[3], const uint16_t *x, float *y) {
  if (!dims_ok(nb, g)) return LMM_ERR_ARG;
  const CDims d1 = cdims(nb, g[0], g[1], g[2]);
  if (int rc = ensure_acts(n, d1.nv)) return rc;
  forward_impl(n, d1, x, y);
  return 0;
}

int cnn_loss_grad(CnnNet *n, int nb, const int g[3], const uint16_t *x, const float *t, double *loss_dev) {
  if (!dims_ok(nb, g)) return LMM_ERR_ARG;
  const CDims d1 = cdims(nb, g[0], g[1], g[2]);
  if (int rc = ensure_acts(n, d1.nv)) return rc;
  forward_impl(n, d1, x, n->y);
  const double cnt = (double)d1.nv * n->cfg.cout;
  const int nblk = cblk(d1.nv);
  CLAUNCH(k_loss_grad, dim3(nblk), dim3(CTPB), 0, n->s, d1.nv, n->cfg.cout, n->y, t, (float)(1.0 / cnt), n->gy,
          n->lpart);
  CLAUNCH(k_sum_partials, dim3(1), dim3(256), 0, n->s, n->lpart, (long)nblk, 0.5 / cnt, loss_dev);
  backward_impl(n, d1, x);
  return 0;
}

int cnn_adam_step(CnnNet *n, double lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  CLAUNCH(k_adam, dim3(cblk(n->np)), dim3(CTPB), 0, n->s, n->np, n->p, n->g, n->m, n->v, (float)lr, b1, b2, 1e-8f,
          c1, c2);
  prep_weights(n);
  return 0;
}

int cnn_dataset_reserve(CnnNet *n, long count, const int g[3]) {
  if (count < 1 || !dims_ok(1, g) || (n->ds_count > 0 && (g[0] != n->ds_n[0] || g[1] != n->ds_n[1] || g[2] != n->ds_n[2])))
    return LMM_ERR_ARG;
  if (count <= n->ds_cap) return 0;
  const long nv1 = (long)g[0] * g[1] * g[2];
  uint16_t *x = cmalloc<uint16_t>(count * nv1 * n->cfg.cin);
  float *t = cmalloc<float>(count * nv1 * CNN_OUT_STRIDE);
  if (!x || !t) {
    cfree(x);
    cfree(t);
    return LMM_ERR_NOMEM;
  }
  if (n->ds_count > 0) {
    hipMemcpyAsync(x, n->dsx, sizeof(uint16_t) * n->ds_count * nv1 * n->cfg.cin, hipMemcpyDeviceToDevice, n->s);
    hipMemcpyAsync(t, n->dst, sizeof(float) * n->ds_count * nv1 * CNN_OUT_STRIDE, hipMemcpyDeviceToDevice, n->s);
    hipStreamSynchronize(n->s);
  }
  cfree(n->dsx);
  cfree(n->dst);
  n->dsx = x;
  n->dst = t;
  n->ds_cap = count;
  for (int q = 0; q < 3; q++) n->ds_n[q] = g[q];
  return 0;
}

long cnn_dataset_count(const CnnNet *n) { return n->ds_count; }
long cnn_dataset_capacity(const CnnNet *n) { return n->ds_cap; }
const int *cnn_dataset_grid(const CnnNet *n) { return n->ds_n; }
void cnn_dataset_push(CnnNet *n) { n->ds_count++; }
uint16_t *cnn_dataset_x(CnnNet *n, long i) {
  const long nv1 = (long)n->ds_n[0] * n->ds_n[1] * n->ds_n[2];
  return n->dsx + i * nv1 * n->cfg.cin;
}
float *cnn_dataset_t(CnnNet *n, long i) {
  const long nv1 = (long)n->ds_n[0] * n->ds_n[1] * n->ds_n[2];
  return n->dst + i * nv1 * CNN_OUT_STRIDE;
}

int cnn_train(CnnNet *n, int steps, int batch, double lr, uint64_t seed, long held0, float *loss_host) {
  const long ntrain = std::min(held0, n->ds_count);
  if (steps < 1 || batch < 1 || ntrain < 1) return LMM_ERR_ARG;
  const int *g = n->ds_n;
  if (!dims_ok(batch, g)) return LMM_ERR_ARG;
  const long nv1 = (long)g[0] * g[1] * g[2];
  uint16_t *xb = cmalloc<uint16_t>((long)batch * nv1 * n->cfg.cin);
  float *tb = cmalloc<float>((long)batch * nv1 * CNN_OUT_STRIDE);
  double *lh = cmalloc<double>(std::max(steps, 256));  // loss history, first the partial sums
  if (!xb || !tb || !lh) {
    cfree(xb);
    cfree(tb);
    cfree(lh);
    return LMM_ERR_NOMEM;
  }
  // The network trains on targets divided by their RMS over the training
  // pairs (unit scale, whatever the units of the error); the output layer is
  // linear and bias-free, so multiplying its weights by that RMS afterwards
  // makes it predict the targets themselves.  A network trained before is
  // brought back to unit scale first.
  const long nt_val = ntrain * nv1 * CNN_OUT_STRIDE;
  CLAUNCH(k_sumsq_f, dim3(256), dim3(256), 0, n->s, nt_val, n->dst, lh);
  CLAUNCH(k_sum_partials, dim3(1), dim3(256), 0, n->s, lh, 256L, 1.0, n->dscal);
  double ssq = 0.0;
  hipMemcpyAsync(&ssq, n->dscal, sizeof(double), hipMemcpyDeviceToHost, n->s);
  if (hipStreamSynchronize(n->s) != hipSuccess) {
    cfree(xb);
    cfree(tb);
    cfree(lh);
    return LMM_ERR_DEVICE;
  }
  const double rms = std::sqrt(ssq / ((double)ntrain * nv1 * n->cfg.cout));
  const float tscale = (rms > 0.0 && std::isfinite(rms)) ? (float)rms : 1.f;
  const CnnLayer &lo = n->L[10];
  const long nlo = 27L * lo.ci * lo.co;
  if (n->trained)
    CLAUNCH(k_rescale_layer, dim3(cblk(nlo)), dim3(CTPB), 0, n->s, nlo, n->p + lo.off, n->m + lo.off, n->v + lo.off,
            1.f / tscale);
  prep_weights(n);
  int rc = 0;
  for (int st = 0; st < steps && !rc; st++) {
    CLAUNCH(k_gather, dim3(cblk(nv1 * batch)), dim3(CTPB), 0, n->s, nv1, n->cfg.cin, batch, n->dsx, n->dst, ntrain,
            seed, (uint64_t)st, 1.f / tscale, xb, tb);
    rc = cnn_loss_grad(n, batch, g, xb, tb, lh + st);
    if (!rc) rc = cnn_adam_step(n, lr, st + 1);
  }
  CLAUNCH(k_rescale_layer, dim3(cblk(nlo)), dim3(CTPB), 0, n->s, nlo, n->p + lo.off, n->m + lo.off, n->v + lo.off,
          tscale);
  prep_weights(n);
  n->trained = true;
  std::vector<double> hl(steps);
  hipMemcpyAsync(hl.data(), lh, sizeof(double) * steps, hipMemcpyDeviceToHost, n->s);
  if (hipStreamSynchronize(n->s) != hipSuccess && !rc) rc = LMM_ERR_DEVICE;
  if (loss_host)
    for (int st = 0; st < steps; st++) loss_host[st] = (float)hl[st];
  cfree(xb);
  cfree(tb);
  cfree(lh);
  return rc;
}

int cnn_eval(CnnNet *n, long i0, long i1, double *rel_err) {
  if (i0 < 0 || i1 > n->ds_count || i1 <= i0) return LMM_ERR_ARG;
  const int *g = n->ds_n;
  const long nv1 = (long)g[0] * g[1] * g[2];
  if (int rc = ensure_acts(n, nv1)) return rc;
  const int nblk = cblk(nv1);
  std::vector<double> hp(2 * (size_t)nblk);
  double acc = 0.0;
  for (long i = i0; i < i1; i++) {
    if (int rc = cnn_forward(n, 1, g, cnn_dataset_x(n, i), n->y)) return rc;
    CLAUNCH(k_err_pair, dim3(nblk), dim3(CTPB), 0, n->s, nv1, n->y, cnn_dataset_t(n, i), n->lpart);
    hipMemcpyAsync(hp.data(), n->lpart, sizeof(double) * 2 * nblk, hipMemcpyDeviceToHost, n->s);
    if (hipStreamSynchronize(n->s) != hipSuccess) return LMM_ERR_DEVICE;
    double e = 0.0, t = 0.0;
    for (int b = 0; b < nblk; b++) {
      e += hp[2 * b];
      t += hp[2 * b + 1];
    }
    acc += (t > 0.0) ? std::sqrt(e / t) : 0.0;
  }
  *rel_err = acc / (double)(i1 - i0);
  return 0;
}

int cnn_forward_host(CnnNet *n, int nb, const int g[3], const float *x, float *y) {
  if (!dims_ok(nb, g)) return LMM_ERR_ARG;
  const long nv = (long)nb * g[0] * g[1] * g[2];
  const int cin = n->cfg.cin, cout = n->cfg.cout;
  std::vector<uint16_t> hx((size_t)nv * cin);
  for (size_t q = 0; q < hx.size(); q++) hx[q] = f2bf_host(x[q]);
  uint16_t *dx = cmalloc<uint16_t>(nv * cin);
  float *dy = cmalloc<float>(nv * CNN_OUT_STRIDE);
  int rc = (dx && dy) ? 0 : LMM_ERR_NOMEM;
  std::vector<float> hy((size_t)nv * CNN_OUT_STRIDE);
  if (!rc) {
    hipMemcpyAsync(dx, hx.data(), sizeof(uint16_t) * hx.size(), hipMemcpyHostToDevice, n->s);
    rc = cnn_forward(n, nb, g, dx, dy);
    hipMemcpyAsync(hy.data(), dy, sizeof(float) * hy.size(), hipMemcpyDeviceToHost, n->s);
    if (hipStreamSynchronize(n->s) != hipSuccess && !rc) rc = LMM_ERR_DEVICE;
  }
  if (!rc)
    for (long v = 0; v < nv; v++)
      for (int c = 0; c < cout; c++) y[v * cout + c] = hy[v * CNN_OUT_STRIDE + c];
  cfree(dx);
  cfree(dy);
  return rc;
}

int cnn_loss_grad_host(CnnNet *n, int nb, const int g[3], const float *x, const float *t, double *loss, float *grad) {
  if (!dims_ok(nb, g)) return LMM_ERR_ARG;
  const long nv = (long)nb * g[0] * g[1] * g[2];
  const int cin = n->cfg.cin, cout = n->cfg.cout;
  std::vector<uint16_t> hx((size_t)nv * cin);
  for (size_t q = 0; q < hx.size(); q++) hx[q] = f2bf_host(x[q]);
  std::vector<float> ht((size_t)nv * CNN_OUT_STRIDE, 0.f);
  for (long v = 0; v < nv; v++)
    for (int c = 0; c < cout; c++) ht[v * CNN_OUT_STRIDE + c] = t[v * cout + c];
  uint16_t *dx = cmalloc<uint16_t>(nv * cin);
  float *dt = cmalloc<float>(nv * CNN_OUT_STRIDE);
  int rc = (dx && dt) ? 0 : LMM_ERR_NOMEM;
  if (!rc) {
    hipMemcpyAsync(dx, hx.data(), sizeof(uint16_t) * hx.size(), hipMemcpyHostToDevice, n->s);
    hipMemcpyAsync(dt, ht.data(), sizeof(float) * ht.size(), hipMemcpyHostToDevice, n->s);
    rc = cnn_loss_grad(n, nb, g, dx, dt, n->dscal);
    hipMemcpyAsync(grad, n->g, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->s);
    double hl = 0.0;
    hipMemcpyAsync(&hl, n->dscal, sizeof(double), hipMemcpyDeviceToHost, n->s);
    if (hipStreamSynchronize(n->s) != hipSuccess && !rc) rc = LMM_ERR_DEVICE;
    *loss = hl;
  }
  cfree(dx);
  cfree(dt);
  return rc;
}

// ---------------------------------------------------------------------------
// features of the Stokes coarse-grid correction
// ---------------------------------------------------------------------------

namespace {

// sums of squares of the velocity residual (3 face arrays) and of the
// pressure residual (1 cell array): partial[2*blk], partial[2*blk+1]
__global__ void __launch_bounds__(256) k_stokes_ss(long N, const float *__restrict__ r0, const float *__restrict__ r1,
                                                   const float *__restrict__ r2, const float *__restrict__ rp,
                                                   double *partial) {
  __shared__ double sh[4];
  double au = 0.0, ap = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += stride) {
    const double a = r0[q], b = r1[q], c = r2[q], p = rp[q];
    au += a * a + b * b + c * c;
    ap += p * p;
  }
  const double su = block_sum_d(au, sh);
  const double sp = block_sum_d(ap, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = su;
    partial[2 * blockIdx.x + 1] = sp;
  }
}

__global__ void k_sum_pairs(const double *__restrict__ p, long n, double *out) {
  __shared__ double sh[4];
  double a = 0.0, b = 0.0;
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    a += p[2 * q];
    b += p[2 * q + 1];
  }
  const double ra = block_sum_d(a, sh);
  const double rb = block_sum_d(b, sh);
  if (threadIdx.x == 0) {
    out[0] = ra;
    out[1] = rb;
  }
}

// The residual's scale in velocity units: s = sqrt(mean(r_u^2) dx^4 + mean(r_p^2) dx^2)
__device__ __forceinline__ double stokes_scale(const double *ss, long N, double dx) {
  const double dx2 = dx * dx;
  return sqrt(ss[0] / (3.0 * (double)N) * dx2 * dx2 + ss[1] / (double)N * dx2);
}

__global__ void k_cnn_pack(long N, double dx, const float *__restrict__ r0, const float *__restrict__ r1,
                           const float *__restrict__ r2, const float *__restrict__ rp,
                           const float *__restrict__ beta, const float *__restrict__ a0,
                           const float *__restrict__ a1, const float *__restrict__ a2, const double *__restrict__ ss,
                           uint16_t *x) {
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const double sc = stokes_scale(ss, N, dx);
  const float iu = sc > 0.0 ? (float)(dx * dx / sc) : 0.f, ip = sc > 0.0 ? (float)(dx / sc) : 0.f;
  uint16_t *o = x + v * 8;
  o[0] = f2bf(r0[v] * iu);
  o[1] = f2bf(r1[v] * iu);
  o[2] = f2bf(r2[v] * iu);
  o[3] = f2bf(rp[v] * ip);
  o[4] = f2bf(logf(beta[v]));
  o[5] = f2bf(logf(a0[v]));
  o[6] = f2bf(logf(a1[v]));
  o[7] = f2bf(logf(a2[v]));
}

__global__ void k_cnn_pack_target(long N, double dx, const float *__restrict__ e0, const float *__restrict__ e1,
                                  const float *__restrict__ e2, const float *__restrict__ ep,
                                  const double *__restrict__ ss, float *t) {
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const double sc = stokes_scale(ss, N, dx);
  const float iu = sc > 0.0 ? (float)(1.0 / sc) : 0.f, ip = sc > 0.0 ? (float)(dx / sc) : 0.f;
  t[v * 4 + 0] = e0[v] * iu;
  t[v * 4 + 1] = e1[v] * iu;
  t[v * 4 + 2] = e2[v] * iu;
  t[v * 4 + 3] = ep[v] * ip;
}

__global__ void k_sub_f32(long n, float *out, const float *__restrict__ a, const float *__restrict__ b) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = a[t] - b[t];
}

// sum of the pressure channel of y (for its mean)
__global__ void __launch_bounds__(256) k_sum_p(long N, const float *__restrict__ y, double *partial) {
  __shared__ double sh[4];
  double a = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) a += y[v * 4 + 3];
  const double r = block_sum_d(a, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// phi_u = y_u * s, phi_p = (y_p - mean y_p) * s / dx (the prediction in the
// level's units; a constant pressure is in the operator's null space)
__global__ void k_cnn_unpack(long N, double dx, const float *__restrict__ y, const double *__restrict__ ss,
                             const double *__restrict__ psum, float *p0, float *p1, float *p2, float *pp) {
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const double sc = stokes_scale(ss, N, dx);
  const float su = (float)sc, sp = (float)(sc / dx), mean = (float)(psum[0] / (double)N);
  p0[v] = y[v * 4 + 0] * su;
  p1[v] = y[v * 4 + 1] * su;
  p2[v] = y[v * 4 + 2] * su;
  pp[v] = (y[v * 4 + 3] - mean) * sp;
}

}  // namespace

void launch_stokes_ss(hipStream_t s, const Grid &g, const float *const ru[3], const float *rp, double *partial,
                      double *out) {
  CLAUNCH(k_stokes_ss, dim3(RED_BLOCKS), dim3(256), 0, s, g.N, ru[0], ru[1], ru[2], rp, partial);
  CLAUNCH(k_sum_pairs, dim3(1), dim3(256), 0, s, partial, (long)RED_BLOCKS, out);
}

void launch_cnn_pack(hipStream_t s, const Grid &g, const float *const ru[3], const float *rp, const float *beta,
                     const float *const alpha[3], const double *ss, uint16_t *x) {
  CLAUNCH(k_cnn_pack, dim3(cblk(g.N)), dim3(CTPB), 0, s, g.N, g.dx, ru[0], ru[1], ru[2], rp, beta, alpha[0], alpha[1],
          alpha[2], ss, x);
}

void launch_cnn_pack_target(hipStream_t s, const Grid &g, const float *const eu[3], const float *ep, const double *ss,
                            float *t) {
  CLAUNCH(k_cnn_pack_target, dim3(cblk(g.N)), dim3(CTPB), 0, s, g.N, g.dx, eu[0], eu[1], eu[2], ep, ss, t);
}

void launch_sub_f32(hipStream_t s, float *out, const float *a, const float *b, long n) {
  CLAUNCH(k_sub_f32, dim3(cblk(n)), dim3(CTPB), 0, s, n, out, a, b);
}

void launch_cnn_unpack(hipStream_t s, const Grid &g, const float *y, const double *ss, double *partial, double *psum,
                       float *const phu[3], float *php) {
  CLAUNCH(k_sum_p, dim3(RED_BLOCKS), dim3(256), 0, s, g.N, y, partial);
  CLAUNCH(k_sum_partials, dim3(1), dim3(256), 0, s, partial, (long)RED_BLOCKS, 1.0, psum);
  CLAUNCH(k_cnn_unpack, dim3(cblk(g.N)), dim3(CTPB), 0, s, g.N, g.dx, y, ss, psum, phu[0], phu[1], phu[2], php);
}

}  // namespace lmm
