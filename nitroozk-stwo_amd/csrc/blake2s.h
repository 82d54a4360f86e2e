// Blake2s-256 compression (RFC 7693), host and device.
// The reference hashes with the blake2 crate's Blake2s256 (crates/stwo/src/core/vcs/blake2_hash.rs,
// blake2_merkle.rs:20-58); its SIMD backend compresses 16 nodes per vector
// (prover/backend/simd/blake2s.rs:165-300). Here one lane owns one message: the 16 message
// words and 16 state words stay in VGPRs and every rotate is a single v_alignbit_b32.
#pragma once
#include "field.h"

namespace sh {

#define B2_IV0 0x6A09E667u
#define B2_IV1 0xBB67AE85u
#define B2_IV2 0x3C6EF372u
#define B2_IV3 0xA54FF53Au
#define B2_IV4 0x510E527Fu
#define B2_IV5 0x9B05688Cu
#define B2_IV6 0x1F83D9ABu
#define B2_IV7 0x5BE0CD19u

SHD uint32_t rotr32(uint32_t x, uint32_t n) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_alignbit(x, x, n);
#else
  return (x >> n) | (x << (32 - n));
#endif
}

#define B2_G(a, b, c, d, x, y)        \
  do {                                \
    v[a] = v[a] + v[b] + (x);         \
    v[d] = rotr32(v[d] ^ v[a], 16);   \
    v[c] = v[c] + v[d];               \
    v[b] = rotr32(v[b] ^ v[c], 12);   \
    v[a] = v[a] + v[b] + (y);         \
    v[d] = rotr32(v[d] ^ v[a], 8);    \
    v[c] = v[c] + v[d];               \
    v[b] = rotr32(v[b] ^ v[c], 7);    \
  } while (0)

#define B2_ROUND(m, s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
  do {                                                                                     \
    B2_G(0, 4, 8, 12, m[s0], m[s1]);                                                       \
    B2_G(1, 5, 9, 13, m[s2], m[s3]);                                                       \
    B2_G(2, 6, 10, 14, m[s4], m[s5]);                                                      \
    B2_G(3, 7, 11, 15, m[s6], m[s7]);                                                      \
    B2_G(0, 5, 10, 15, m[s8], m[s9]);                                                      \
    B2_G(1, 6, 11, 12, m[s10], m[s11]);                                                    \
    B2_G(2, 7, 8, 13, m[s12], m[s13]);                                                     \
    B2_G(3, 4, 9, 14, m[s14], m[s15]);                                                     \
  } while (0)

// h: chaining state; m: 16 message words; t: byte count including this block.
SHD void blake2s_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, bool last) {
  uint32_t v[16];
  for (int i = 0; i < 8; ++i) v[i] = h[i];
  v[8] = B2_IV0; v[9] = B2_IV1; v[10] = B2_IV2; v[11] = B2_IV3;
  v[12] = B2_IV4 ^ t; v[13] = B2_IV5; v[14] = last ? ~B2_IV6 : B2_IV6; v[15] = B2_IV7;
  B2_ROUND(m, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  B2_ROUND(m, 14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3);
  B2_ROUND(m, 11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4);
  B2_ROUND(m, 7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8);
  B2_ROUND(m, 9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13);
  B2_ROUND(m, 2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9);
  B2_ROUND(m, 12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11);
  B2_ROUND(m, 13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10);
  B2_ROUND(m, 6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5);
  B2_ROUND(m, 10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0);
  for (int i = 0; i < 8; ++i) h[i] ^= v[i] ^ v[i + 8];
}

// The same compression spread over the 4 lanes of a quad, for latency-bound hashes (the top levels
// of a Merkle tree, the channel): lane j of the quad holds column j of the state matrix
// (v[j], v[4+j], v[8+j], v[12+j]) and runs one G of each step; the diagonal step reads rows
// b, c, d rotated by 1, 2, 3 lanes (DPP quad permutations), then rotates them back. On entry and
// exit lane j holds chaining words hA = h[j] and hB = h[4+j]; every lane holds the full message.
// All 4 lanes of the quad must be active. The permutations set bound_ctrl: they never read a lane
// outside the quad, so it changes no value, and with it the compiler folds most of the moves into
// the add or xor that consumes them (DPP operand), about 38 fewer instructions per compression.
template <int CTRL>
__device__ __forceinline__ uint32_t quad_rot(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
#define B2Q_ROT1 0x39  // lane i reads lane (i + 1) % 4
#define B2Q_ROT2 0x4E  // lane i reads lane (i + 2) % 4
#define B2Q_ROT3 0x93  // lane i reads lane (i + 3) % 4
#define B2Q_SEL(w0, w1, w2, w3) (j2 ? (j1 ? (w3) : (w2)) : (j1 ? (w1) : (w0)))
#define B2Q_G(x, y)               \
  do {                            \
    a = a + b + (x);              \
    d = rotr32(d ^ a, 16);        \
    c = c + d;                    \
    b = rotr32(b ^ c, 12);        \
    a = a + b + (y);              \
    d = rotr32(d ^ a, 8);         \
    c = c + d;                    \
    b = rotr32(b ^ c, 7);         \
  } while (0)
#define B2Q_ROUND(m, s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
  do {                                                                                      \
    B2Q_G(B2Q_SEL(m[s0], m[s2], m[s4], m[s6]), B2Q_SEL(m[s1], m[s3], m[s5], m[s7]));        \
    b = quad_rot<B2Q_ROT1>(b);                                                              \
    c = quad_rot<B2Q_ROT2>(c);                                                              \
    d = quad_rot<B2Q_ROT3>(d);                                                              \
    B2Q_G(B2Q_SEL(m[s8], m[s10], m[s12], m[s14]), B2Q_SEL(m[s9], m[s11], m[s13], m[s15]));  \
    b = quad_rot<B2Q_ROT3>(b);                                                              \
    c = quad_rot<B2Q_ROT2>(c);                                                              \
    d = quad_rot<B2Q_ROT1>(d);                                                              \
  } while (0)
__device__ __forceinline__ void blake2s_compress_quad(uint32_t& hA, uint32_t& hB, const uint32_t msg[16], uint32_t t,
                                                      bool last, uint32_t j) {
  const bool j1 = (j & 1) != 0, j2 = (j & 2) != 0;
  // The message words pass through an empty asm so that each lane's per-round choice among them
  // stays a register select: folded into an indexed load, the array would go to scratch memory.
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    m[k] = msg[k];
    asm volatile("" : "+v"(m[k]));
  }
  uint32_t a = hA, b = hB;
  uint32_t c = B2Q_SEL(B2_IV0, B2_IV1, B2_IV2, B2_IV3);
  uint32_t d = B2Q_SEL(B2_IV4 ^ t, B2_IV5, last ? ~B2_IV6 : B2_IV6, B2_IV7);
  B2Q_ROUND(m, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  B2Q_ROUND(m, 14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3);
  B2Q_ROUND(m, 11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4);
  B2Q_ROUND(m, 7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8);
  B2Q_ROUND(m, 9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13);
  B2Q_ROUND(m, 2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9);
  B2Q_ROUND(m, 12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11);
  B2Q_ROUND(m, 13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10);
  B2Q_ROUND(m, 6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5);
  B2Q_ROUND(m, 10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0);
  hA ^= a ^ c;
  hB ^= b ^ d;
}
// Initial chaining words of lane j (Blake2s-256, no key): h[j] and h[4+j].
__device__ __forceinline__ void blake2s_init_quad(uint32_t j, uint32_t& hA, uint32_t& hB) {
  const bool j1 = (j & 1) != 0, j2 = (j & 2) != 0;
  hA = B2Q_SEL(B2_IV0 ^ 0x01010020u, B2_IV1, B2_IV2, B2_IV3);
  hB = B2Q_SEL(B2_IV4, B2_IV5, B2_IV6, B2_IV7);
}
#undef B2Q_ROUND
#undef B2Q_G
#undef B2Q_SEL

SHD void blake2s_init(uint32_t h[8]) {
  h[0] = B2_IV0 ^ 0x01010020u; h[1] = B2_IV1; h[2] = B2_IV2; h[3] = B2_IV3;
  h[4] = B2_IV4; h[5] = B2_IV5; h[6] = B2_IV6; h[7] = B2_IV7;
}

}  // namespace sh
