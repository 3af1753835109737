// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror & Shaw, SC'11),
// used for the seeded initial-condition noise so that every node's draw
// depends only on (seed, node, component) and host and device agree.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace uclbs {

__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// counter (node lo, node hi, component, 0), key (seed lo, seed hi);
// 53-bit uniform in [0,1) from the first two output words.
__host__ __device__ inline double philox_uniform(uint64_t node, uint32_t comp, uint64_t seed) {
  uint32_t c[4] = {(uint32_t)node, (uint32_t)(node >> 32), comp, 0u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t m = ((uint64_t)(c[0] >> 5) << 26) + (uint64_t)(c[1] >> 6);
  return (double)m * (1.0 / 9007199254740992.0);
}

}  // namespace uclbs
