// The Blake2s Fiat-Shamir channel with its state in device memory (crates/stwo/src/core/channel/
// blake2s.rs:19-140): mix a root or words, draw secure felts and query positions, and the proof of
// work from the device digest (GrindOps<Blake2sChannel>, prover/backend/simd/grind.rs:20-91). Each
// operation is one single-lane launch on the caller's stream, with every argument in device memory
// or passed by value, so that a sequence of them, and of the trait entries that launch the same
// way, can run without host round trips and be captured in a graph.
//
// State (CHANNEL_WORDS words): digest 0..7, n_draws 8, the grind's result slot 10..11 (u64, ~0 while
// no nonce is known), the proof-of-work prefix 12..19.
#include "blake2s.h"
#include "kernels.h"

namespace sh {

namespace {

// H(digest || n_draws || 0): 37 bytes in one block (the draw of Blake2sChannel::draw_u32s)
__device__ void draw_u32s(uint32_t* st, uint32_t out[8]) {
  uint32_t m[16] = {st[0], st[1], st[2], st[3], st[4], st[5], st[6], st[7], st[8], 0, 0, 0, 0, 0, 0, 0};
  st[8] += 1;
  uint32_t h[8];
  blake2s_init(h);
  blake2s_compress(h, m, 37, true);
  for (int k = 0; k < 8; ++k) out[k] = h[k];
}
// draw_base_felts: redrawn until all 8 words are below 2P (a redraw has probability about 2^-28)
__device__ void draw_base_felts(uint32_t* st, uint32_t out[8]) {
  for (;;) {
    draw_u32s(st, out);
    bool ok = true;
    for (int k = 0; k < 8; ++k) ok &= out[k] < 2 * P;
    if (ok) break;
  }
  for (int k = 0; k < 8; ++k) out[k] = out[k] >= P ? out[k] - P : out[k];
}

__global__ void chan_init_kernel(uint32_t* __restrict__ st, ChanDigest d, uint32_t n_draws) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 8; ++k) st[k] = d.w[k];
  st[8] = n_draws;
  st[9] = 0;
  st[10] = st[11] = ~0u;
  for (int k = 12; k < (int)CHANNEL_WORDS; ++k) st[k] = 0;
}

// digest = H(digest || words), n_draws = 0 (mix_root: the 8 words of a root; mix_felts: 4 words per
// felt; mix_u64: low word, high word)
__global__ void chan_mix_words_kernel(uint32_t* __restrict__ st, const uint32_t* __restrict__ words, uint32_t n) {
  if (threadIdx.x != 0) return;
  uint32_t h[8];
  blake2s_init(h);
  uint32_t m[16];
  const uint32_t total = 8 + n;  // words in the message
  const uint32_t blocks = (total + 15) / 16;
  for (uint32_t b = 0; b < blocks; ++b) {
    for (uint32_t k = 0; k < 16; ++k) {
      const uint32_t w = 16 * b + k;
      m[k] = w < 8 ? st[w] : (w < total ? words[w - 8] : 0u);
    }
    const bool last = b + 1 == blocks;
    blake2s_compress(h, m, last ? 4 * total : 64 * (b + 1), last);
  }
  for (int k = 0; k < 8; ++k) st[k] = h[k];
  st[8] = 0;
}

// draw_secure_felts(n): base felts of successive draws, four at a time (a draw's second four are
// used only when more are wanted)
__global__ void chan_draw_felts_kernel(uint32_t* __restrict__ st, uint32_t n, uint32_t* __restrict__ out) {
  if (threadIdx.x != 0) return;
  uint32_t f[8];
  for (uint32_t i = 0; i < n; i += 2) {
    draw_base_felts(st, f);
    for (int k = 0; k < 4; ++k) out[4 * i + k] = f[k];
    if (i + 1 < n)
      for (int k = 0; k < 4; ++k) out[4 * (i + 1) + k] = f[4 + k];
  }
}

// draw_queries (core/queries.rs:11-29): words of successive draws, each masked to the domain
__global__ void chan_draw_queries_kernel(uint32_t* __restrict__ st, uint32_t log_domain, uint32_t n,
                                         uint32_t* __restrict__ out) {
  if (threadIdx.x != 0) return;
  const uint32_t mask = (uint32_t)(((uint64_t)1 << log_domain) - 1);
  uint32_t u[8];
  for (uint32_t i = 0; i < n;) {
    draw_u32s(st, u);
    for (int k = 0; k < 8 && i < n; ++k, ++i) out[i] = u[k] & mask;
  }
}

// the proof-of-work prefix H(0x12345678 || 0^12 || digest || pow_bits), and the result slot emptied
__global__ void chan_pow_prefix_kernel(uint32_t* __restrict__ st, uint32_t pow_bits) {
  if (threadIdx.x != 0) return;
  uint32_t m[16] = {0x12345678u, 0, 0, 0, st[0], st[1], st[2], st[3], st[4], st[5], st[6], st[7], pow_bits, 0, 0, 0};
  uint32_t h[8];
  blake2s_init(h);
  blake2s_compress(h, m, 52, true);
  for (int k = 0; k < 8; ++k) st[12 + k] = h[k];
  st[10] = st[11] = ~0u;
}

__global__ void chan_copy_nonce_kernel(const uint32_t* __restrict__ st, uint32_t* __restrict__ nonce) {
  if (threadIdx.x != 0) return;
  nonce[0] = st[10];
  nonce[1] = st[11];
}

}  // namespace

void launch_chan_init(uint32_t* st, const ChanDigest& d, uint32_t n_draws, hipStream_t stream) {
  hipLaunchKernelGGL(chan_init_kernel, dim3(1), dim3(64), 0, stream, st, d, n_draws);
}
void launch_chan_mix_words(uint32_t* st, const uint32_t* words, uint32_t n, hipStream_t stream) {
  hipLaunchKernelGGL(chan_mix_words_kernel, dim3(1), dim3(64), 0, stream, st, words, n);
}
void launch_chan_draw_felts(uint32_t* st, uint32_t n, uint32_t* out, hipStream_t stream) {
  hipLaunchKernelGGL(chan_draw_felts_kernel, dim3(1), dim3(64), 0, stream, st, n, out);
}
void launch_chan_draw_queries(uint32_t* st, uint32_t log_domain, uint32_t n, uint32_t* out, hipStream_t stream) {
  hipLaunchKernelGGL(chan_draw_queries_kernel, dim3(1), dim3(64), 0, stream, st, log_domain, n, out);
}
void launch_chan_grind(uint32_t* st, uint32_t pow_bits, uint32_t count, uint32_t* nonce, hipStream_t stream) {
  hipLaunchKernelGGL(chan_pow_prefix_kernel, dim3(1), dim3(64), 0, stream, st, pow_bits);
  launch_grind_dev(st + 12, pow_bits, 0, count, reinterpret_cast<unsigned long long*>(st + 10), stream);
  hipLaunchKernelGGL(chan_copy_nonce_kernel, dim3(1), dim3(64), 0, stream, st, nonce);
}

}  // namespace sh
