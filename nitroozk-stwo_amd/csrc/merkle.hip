// Blake2s Merkle layers (MerkleOps<Blake2sMerkleHasher>::commit_on_layer) and Blake2s proof of
// work (GrindOps::grind).
//
// Node hash (crates/stwo/src/core/vcs/blake2_merkle.rs:37-58):
//   node[i] = Blake2s-256(prev[2i] || prev[2i+1] || col_0[i] || col_1[i] || ...)   (u32 LE)
// One lane hashes one node. The children (64 bytes) are one message block; columns follow in
// blocks of 16 words, read column-major so that 64 consecutive lanes read 256 contiguous bytes
// of each column. A node with no input at all is Blake2s of the empty string.
//
// Grind (prover/backend/simd/grind.rs:20-91): the nonce is the smallest (hi << 32) | lo with
// lo < 2^20 such that the first digest word of H(prefix || lo || hi) has >= pow_bits trailing
// zeros, prefix = H(0x12345678 || 0^12 || digest || pow_bits).
#include "blake2s.h"
#include "kernels.h"

namespace sh {

__global__ void __launch_bounds__(256) merkle_layer_kernel(uint32_t n_nodes, const uint32_t* __restrict__ prev,
                                                           const uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                                           uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  uint32_t h[8];
  blake2s_init(h);
  uint32_t m[16];
  uint32_t t = 0;
  if (prev) {
    const uint4* p4 = reinterpret_cast<const uint4*>(prev + (size_t)i * 16);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint4 v = p4[k];
      m[4 * k] = v.x; m[4 * k + 1] = v.y; m[4 * k + 2] = v.z; m[4 * k + 3] = v.w;
    }
    t = 64;
    blake2s_compress(h, m, t, n_cols == 0);
  } else if (n_cols == 0) {
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = 0;
    blake2s_compress(h, m, 0, true);
  }
  for (uint32_t cs = 0; cs < n_cols; cs += 16) {
    uint32_t nw = n_cols - cs < 16 ? n_cols - cs : 16;
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) m[k] = k < nw ? cols[cs + k][i] : 0u;
    t += 4 * nw;
    blake2s_compress(h, m, t, cs + 16 >= n_cols);
  }
  uint4* o4 = reinterpret_cast<uint4*>(out + (size_t)i * 8);
  o4[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o4[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

void launch_merkle_layer(uint32_t log_size, const uint32_t* prev, const uint32_t* const* cols, uint32_t n_cols,
                         uint32_t* out, hipStream_t stream) {
  uint32_t n = 1u << log_size;
  uint32_t threads = n < 256 ? (n < 64 ? 64 : n) : 256;
  hipLaunchKernelGGL(merkle_layer_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, stream, n, prev, cols,
                     n_cols, out);
}

__global__ void __launch_bounds__(256) grind_kernel(const uint32_t* __restrict__ prefix, uint32_t pow_bits,
                                                    uint32_t hi_base, unsigned long long* __restrict__ result) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t lo = g & ((1u << 20) - 1);
  uint32_t hi = hi_base + (g >> 20);
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) m[k] = prefix[k];
  m[8] = lo;
  m[9] = hi;
#pragma unroll
  for (int k = 10; k < 16; ++k) m[k] = 0;
  uint32_t h[8];
  blake2s_init(h);
  blake2s_compress(h, m, 40, true);
  uint32_t tz = h[0] ? __builtin_ctz(h[0]) : 32u;
  if (tz >= pow_bits) atomicMin(result, ((unsigned long long)hi << 32) | lo);
}

void launch_grind(const uint32_t* prefix, uint32_t pow_bits, uint32_t hi_base, uint32_t n_hi,
                  unsigned long long* result, hipStream_t stream) {
  uint32_t total = n_hi << 20;
  hipLaunchKernelGGL(grind_kernel, dim3(total / 256), dim3(256), 0, stream, prefix, pow_bits, hi_base, result);
}

}  // namespace sh
