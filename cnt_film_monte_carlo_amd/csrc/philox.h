// philox.h -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable on host and device.
//
// It replaces the reference's global std::rand() stream (src/monte_carlo/scatterer.cpp:17,
// src/monte_carlo/scatterer.h:76, src/monte_carlo/monte_carlo.cpp:311,331), whose order
// depends on OpenMP scheduling.  Here every draw is a pure function of
// (seed, exciton, hop, stream), so results do not depend on launch shape or thread order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cntkmc {

struct u32x4 {
  uint32_t v[4];
};

// a ^ b ^ c; on the device one gfx950 three-input bitwise operation (truth table 0x96).
__host__ __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

// On the device the key (the seed) must be the same for every lane of the wave: its ten round
// values are stepped in scalar registers by inline adds, so that the compiler does not keep all
// twenty of them live across a caller's loop.
__host__ __device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                        uint32_t c3, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c0;
    const uint64_t p1 = (uint64_t)M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = xor3(hi1, c1, k0);
    const uint32_t n2 = xor3(hi0, c3, k1);
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
#if defined(__HIP_DEVICE_COMPILE__)
    if (r < 9) {
      asm volatile("s_add_u32 %0, %0, 0x9E3779B9" : "+s"(k0) : : "scc");
      asm volatile("s_add_u32 %0, %0, 0xBB67AE85" : "+s"(k1) : : "scc");
    }
#else
    k0 += W0;
    k1 += W1;
#endif
  }
  u32x4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// Draw streams.  Counter = {c lo, c hi, exciton, stream}; key = seed.
enum : uint32_t {
  STREAM_HOP = 0u,            // counter h >> 1: hops 2P and 2P + 1 share one call, the even hop
                              // taking words 0 (selection) and 1 (dwell), the odd hop 2 and 3
  STREAM_INIT = 1u,           // injection: w0 site, w1 dwell, w2 heading, w3 dwell spare
  STREAM_HOP_REDRAW = 2u,     // + k, counter h: further dwell candidates after a zero word
  STREAM_REINJECT = 0x10000000u,  // counter h: w0 the source site after a drain at hop h
  STREAM_INIT_REDRAW = 0x80000000u,
  STREAM_FLY_DRAIN = 0x40000000u,  // + j: re-injection site after the j-th drain of one flight
  STREAM_CONTACT = 0x20000000u,    // contact mode, a new particle: w0 site, w1 dwell, w3 spare
  STREAM_CONTACT_REDRAW = 0x20000001u,  // + k: its further dwell candidates
  REDRAW_LIMIT = 16u,         // bounded redraw loop; after it the draw is taken as 1
};

__host__ __device__ __forceinline__ u32x4 draw4(uint64_t seed, uint32_t exciton, uint64_t hop,
                                                uint32_t stream) {
  return philox4x32_10((uint32_t)hop, (uint32_t)(hop >> 32), exciton, stream, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}

// 31-bit draw in [0, 2^31-1] = [0, RAND_MAX] of glibc rand().
__host__ __device__ __forceinline__ uint32_t r31(uint32_t w) { return w >> 1; }

// First non-zero 31-bit draw for a dwell time (ff_time rejects rand()==0, scatterer.h:76-77):
// candidates w_a, w_b, then words of the redraw streams.
__host__ __device__ __forceinline__ uint32_t nonzero_draw(uint32_t wa, uint32_t wb, uint64_t seed,
                                                          uint32_t exciton, uint64_t hop,
                                                          uint32_t redraw_stream) {
  uint32_t r = r31(wa);
  if (r != 0u) return r;
  r = r31(wb);
  if (r != 0u) return r;
  for (uint32_t k = 0; k < REDRAW_LIMIT; ++k) {
    const u32x4 w = draw4(seed, exciton, hop, redraw_stream + k);
    if (r31(w.v[0]) != 0u) return r31(w.v[0]);
    if (r31(w.v[1]) != 0u) return r31(w.v[1]);
    if (r31(w.v[2]) != 0u) return r31(w.v[2]);
    if (r31(w.v[3]) != 0u) return r31(w.v[3]);
  }
  return 1u;
}

// First non-zero 31-bit draw from one dwell word: the word, then the redraw streams.
__host__ __device__ __forceinline__ uint32_t nonzero_draw1(uint32_t w, uint64_t seed,
                                                           uint32_t exciton, uint64_t hop,
                                                           uint32_t redraw_stream) {
  const uint32_t r = r31(w);
  if (r != 0u) return r;
  for (uint32_t k = 0; k < REDRAW_LIMIT; ++k) {
    const u32x4 v = draw4(seed, exciton, hop, redraw_stream + k);
    if (r31(v.v[0]) != 0u) return r31(v.v[0]);
    if (r31(v.v[1]) != 0u) return r31(v.v[1]);
    if (r31(v.v[2]) != 0u) return r31(v.v[2]);
    if (r31(v.v[3]) != 0u) return r31(v.v[3]);
  }
  return 1u;
}

// The selection and dwell words of consecutive hops, one Philox call per pair of hops: the
// words of the odd hop of a pair are kept from the call the even hop made.  `have` is cleared
// when a hop index is passed over without drawing.
struct hop_words {
  uint32_t w2 = 0u, w3 = 0u;
  bool have = false;  // w2, w3 are the words of hop h + 1 after the even hop h drew

  __host__ __device__ __forceinline__ void get(uint64_t seed, uint32_t e, uint64_t h,
                                               uint32_t& sel, uint32_t& dw) {
    if ((h & 1u) != 0u && have) {
      sel = w2;
      dw = w3;
      have = false;
      return;
    }
    const u32x4 w = draw4(seed, e, h >> 1, STREAM_HOP);
    const bool odd = (h & 1u) != 0u;
    sel = odd ? w.v[2] : w.v[0];
    dw = odd ? w.v[3] : w.v[1];
    w2 = w.v[2];
    w3 = w.v[3];
    have = !odd;
  }
};

}  // namespace cntkmc
