/*
 * Counter-based Gaussian stream of the CPU oracle (test infrastructure
 * only).  The reference draws its noise from a host MT19937 stream
 * (src_random/RNGs.c, restated in orc_flux.c); the device fill of the noise
 * fields (lmm_fill_m_stochastic_device, no counterpart in the reference) uses
 * this generator instead, and this file is the CPU statement it is checked
 * against.
 *
 *   Philox4x32-10: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 *   "Parallel random numbers: as easy as 1, 2, 3", SC'11.  Ten rounds of
 *     (c0,c1,c2,c3) -> (hi(M1*c2) ^ c1 ^ k0, lo(M1*c2), hi(M0*c0) ^ c3 ^ k1, lo(M0*c0))
 *   with M0 = 0xD2511F53, M1 = 0xCD9E8D57 and the key bumped by the Weyl
 *   constants (0x9E3779B9, 0xBB67AE85) between rounds.
 *
 *   Value v of the stream for (seed, step): block b = v / 2 is Philox of the
 *   counter (b_lo, b_hi, step_lo, step_hi) under the key (seed_lo, seed_hi);
 *   its words give two 53-bit uniforms u1 = ((x1:x0) >> 11 + 0.5) 2^-53 and
 *   u2 = ((x3:x2) >> 11 + 0.5) 2^-53 in (0, 1), and the Box-Muller pair
 *   r = sqrt(-2 log u1): value 2b = r cos(2 pi u2), value 2b+1 = r sin(2 pi u2).
 */
#include <math.h>
#include <stdint.h>

#include "lmm_oracle.h"

static void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0], n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    philox_round(c, k);
  }
  for (int q = 0; q < 4; q++) out[q] = c[q];
}

static double u53(uint32_t lo, uint32_t hi) {
  const uint64_t a = ((uint64_t)hi << 32) | lo;
  return ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_fill_normal_philox(uint64_t seed, uint64_t step, int64_t v0, int64_t count, double *out) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t t = 0; t < count; t++) {
    const uint64_t v = (uint64_t)(v0 + t), b = v >> 1;
    const uint32_t ctr[4] = {(uint32_t)b, (uint32_t)(b >> 32), (uint32_t)step, (uint32_t)(step >> 32)};
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    const double u1 = u53(x[0], x[1]), u2 = u53(x[2], x[3]);
    const double r = sqrt(-2.0 * log(u1));
    const double a = 6.283185307179586 * u2;
    out[t] = (v & 1) ? r * sin(a) : r * cos(a);
  }
}
