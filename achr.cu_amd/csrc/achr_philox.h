// Counter-based random numbers for the ACHR chains: Philox4x32-10.
//
// Written from the published description (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): a 4x32-bit counter
// and a 2x32-bit key pass through 10 rounds; each round multiplies words 0 and
// 2 by the Philox multipliers, swaps the high/low halves across the word pairs
// and xors in the key; the key is bumped by the Weyl constants between rounds.
//
// The reference draws its step uniforms from cuRAND XORWOW seeded by clock64()
// (ACHRcu/ACHRCuda.cu:291-295, :309, :371), which no run can reproduce. This
// sampler replaces it with a stream that is a pure function of
// (seed, chain, file, model, attempt/step), so a GPU run and the host oracle
// consume bit-identical uniforms.
//
// Stream layout (shared by every kernel in this library and by the oracle):
//   key     = { seed & 0xffffffff, seed >> 32 }
//   counter = { index, chain, file, (model << 4) | tag }
//   tag 0 : warmup-point pick for attempt number `index` (misses included)
//   tag 1 : step-length uniform for accepted step number `index`; a miss
//           reuses it, as randVector[stepCount] does in ACHRCuda.cu:314
//   tag 2 : Gaussian/latent draws of the direction models (index = attempt)
//   tag 3+: reserved
// The double uniform takes the top 53 bits of (w1:w0) and lies in [0, 1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace achr {

struct u32x4 {
  uint32_t v0, v1, v2, v3;
};

__host__ __device__ __forceinline__ void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = static_cast<uint64_t>(a) * static_cast<uint64_t>(b);
  hi = static_cast<uint32_t>(p >> 32);
  lo = static_cast<uint32_t>(p);
}

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kMulA = 0xD2511F53u, kMulB = 0xCD9E8D57u;
  constexpr uint32_t kWeylA = 0x9E3779B9u, kWeylB = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hiA, loA, hiB, loB;
    philox_mulhilo(kMulA, ctr.v0, hiA, loA);
    philox_mulhilo(kMulB, ctr.v2, hiB, loB);
    const u32x4 nxt = {hiB ^ ctr.v1 ^ k0, loB, hiA ^ ctr.v3 ^ k1, loA};
    ctr = nxt;
    k0 += kWeylA;
    k1 += kWeylB;
  }
  return ctr;
}

enum StreamTag : uint32_t { kTagPick = 0u, kTagStep = 1u, kTagDir = 2u };

// Uniform double in [0, 1) from the first two output words.
__host__ __device__ __forceinline__ double u01_from(uint32_t w0, uint32_t w1) {
  const uint64_t bits = ((static_cast<uint64_t>(w1) << 32) | w0) >> 11;
  return static_cast<double>(bits) * 0x1.0p-53;
}

__host__ __device__ __forceinline__ u32x4 stream_block(uint64_t seed, uint32_t index, uint32_t chain,
                                                       uint32_t file, uint32_t model, uint32_t tag) {
  const u32x4 ctr = {index, chain, file, (model << 4) | tag};
  return philox4x32_10(ctr, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}

__host__ __device__ __forceinline__ double stream_u01(uint64_t seed, uint32_t index, uint32_t chain,
                                                      uint32_t file, uint32_t model, uint32_t tag) {
  const u32x4 r = stream_block(seed, index, chain, file, model, tag);
  return u01_from(r.v0, r.v1);
}

// Warmup column for an attempt: (int)(W*u) clamped to W-1 (ACHRCuda.cu:309-310).
__host__ __device__ __forceinline__ int pick_index(double u, int W) {
  int id = static_cast<int>(static_cast<double>(W) * u);
  return id >= W ? W - 1 : id;
}

}  // namespace achr
