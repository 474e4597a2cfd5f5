// philox.h -- counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).
//
// Every uniform the sampler draws is addressed by a counter instead of a
// running state, so a chain's stream can be replayed from (seed, stream,
// iteration, chain, slot) alone:
//   counter = (iteration t, chain, slot / 4, tag), key = (seed, stream)
//   u = (word[slot % 4] + 0.5) / 2^32
// The slot layout within an iteration is the reference's consumption order
// (src/rust/src/cpu_native.rs:726-787): pairs (u1, u2) per parameter for the
// Box-Muller normals, then the acceptance uniform; DE prepends u_a, u_b,
// u_gamma.  The map to (0, 1) is the reference's (cpu_native.rs:25-28).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpum {

enum PhiloxTag : uint32_t {
  TAG_STEP = 0,   // proposal / acceptance uniforms
  TAG_SWAP = 1,   // parallel-tempering swap uniforms
  TAG_INIT = 2,   // initial values drawn by the host API
  TAG_DA = 3,     // delayed-acceptance stage uniforms
  TAG_DATA = 4,   // synthetic data generation
  TAG_MLP = 5,    // surrogate initialisation and minibatch sampling
};

struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__host__ __device__ inline double u32_to_unit(uint32_t h) {
  return ((double)h + 0.5) / 4294967296.0;
}

__host__ __device__ inline uint32_t u4_word(U4 v, uint32_t i) {
  return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

}  // namespace gpum
