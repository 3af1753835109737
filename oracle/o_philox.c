/* Philox4x32-10 counter-based RNG (Salmon, Moraes, Dror & Shaw, "Parallel
 * random numbers: as easy as 1, 2, 3", SC'11).  TEST INFRASTRUCTURE ONLY
 * (see uclbs_oracle.h).  Used to draw the seeded initial-condition noise so
 * that the oracle and the device initialiser see the same numbers. */
#include "uclbs_oracle.h"

static inline void mulhilo(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint32_t h0, l0, h1, l1;
    mulhilo(0xD2511F53u, c0, &h0, &l0);
    mulhilo(0xCD9E8D57u, c2, &h1, &l1);
    uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
    c0 = n0; c1 = l1; c2 = n2; c3 = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* counter = (node lo, node hi, component, 0), key = (seed lo, seed hi);
 * 53-bit uniform = ((out0 >> 5) * 2^26 + (out1 >> 6)) * 2^-53 */
double o_uniform(uint64_t node, uint32_t comp, uint64_t seed) {
  uint32_t ctr[4] = {(uint32_t)node, (uint32_t)(node >> 32), comp, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  o_philox4x32_10(ctr, key, o);
  uint64_t m = ((uint64_t)(o[0] >> 5) << 26) + (uint64_t)(o[1] >> 6);
  return (double)m * (1.0 / 9007199254740992.0);
}
