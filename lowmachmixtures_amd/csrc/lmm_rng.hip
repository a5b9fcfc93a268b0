// Counter-based Gaussian noise on the device.
//
// The reference fills its noise fields from a host MT19937 stream
// (multifab_fill_random.f90:16-63 -> RNGs.c:247-260), one value after another.
// Here every value is an independent function of (seed, step, value index):
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) of the counter (block, step) under the key seed,
// turned into a Box-Muller pair.  Any thread can produce any value, the fill
// is one pass over the output, and a value does not depend on how the fill is
// split.  Value v: block b = v / 2 gives words x0..x3, u1 = ((x1:x0) >> 11 +
// 0.5) 2^-53, u2 = ((x3:x2) >> 11 + 0.5) 2^-53, r = sqrt(-2 log u1);
// v even -> r cos(2 pi u2), v odd -> r sin(2 pi u2).
#include "lmm_kernels.h"

namespace lmm {

namespace {

const char kPhiloxName[] = "k_fill_normal_philox";

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

__device__ __forceinline__ double u53(uint32_t lo, uint32_t hi) {
  const uint64_t a = ((uint64_t)hi << 32) | lo;
  return ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// thread t: block b0 + t, writing the values of its pair in [v0, v0 + count)
__global__ void __launch_bounds__(256) k_fill_normal_philox(double *out, long count, long v0, long b0, long nb,
                                                            uint32_t k0, uint32_t k1, uint32_t s0, uint32_t s1) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb) return;
  const unsigned long long b = (unsigned long long)(b0 + t);
  uint32_t c[4] = {(uint32_t)b, (uint32_t)(b >> 32), s0, s1};
  philox4x32_10(c, k0, k1);
  const double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]);
  const double r = sqrt(-2.0 * log(u1));
  const double a = 6.283185307179586 * u2;
  const long e = 2 * (long)b - v0;  // output index of the even value
  if (e >= 0 && e < count) out[e] = r * cos(a);
  if (e + 1 >= 0 && e + 1 < count) out[e + 1] = r * sin(a);
}

}  // namespace

void launch_fill_normal_philox(hipStream_t s, double *out, long count, long v0, uint64_t seed, uint64_t step) {
  if (count <= 0) return;
  const long b0 = v0 >> 1, b1 = (v0 + count - 1) >> 1, nb = b1 - b0 + 1;
  hipLaunchKernelGGL(k_fill_normal_philox, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, out, count, v0, b0,
                     nb, (uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)step, (uint32_t)(step >> 32));
  post_launch(s, kPhiloxName);
}

}  // namespace lmm
