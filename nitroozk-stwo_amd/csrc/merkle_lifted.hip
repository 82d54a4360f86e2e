// Lifted Merkle leaves: MerkleOpsLifted<H>::build_leaves (crates/stwo/src/prover/vcs_lifted/ops.rs:8-18;
// CpuBackend crates/stwo/src/prover/backend/cpu/merkle_lifted.rs:12-76) for the two lifted hashers,
// Blake2sMerkleHasher and Poseidon252MerkleHasher (crates/stwo/src/core/vcs_lifted/blake2_merkle.rs,
// poseidon252_merkle.rs).
//
// Columns come in non-decreasing order of log size. A column of log size l, lifted to the lifting
// domain of log size L, takes at leaf i the value of its row ((i >> (L - l + 1)) << 1) | (i & 1):
// the lifted polynomial f(pi^(L-l)(P)) on the bit-reversed domain. The CPU reference carries one
// hasher state per row and clones it each time the row count grows; every lifting step composes
// to that one row map, so leaf i's hash is the hash of the stream of every column's value at its
// mapped row, in column order:
//   Blake2s:     Blake2s-256 of the values as u32 little endian (one message block per 16 values);
//   Poseidon252: poseidon_hash_many of the stream cut into felt252s of 8 values (the last chunk,
//                if shorter, with its length padding), which is what the hasher's 16-value buffer
//                and poseidon_finalize compute.
// One lane computes one leaf, streaming the columns through a table of (address, lift shift); no
// hasher state goes to memory. Lanes i and i + 1 read the same words of a lifted column, which the
// caches serve.
#include "blake2s.h"
#include "felt252.h"
#include "kernels.h"

namespace sh {

namespace {

__device__ __forceinline__ uint32_t lifted_row(uint32_t i, uint32_t shift) {
  return ((i >> (shift + 1)) << 1) | (i & 1u);
}

__global__ void __launch_bounds__(256) lifted_leaves_blake2s_kernel(uint32_t n_leaves,
                                                                    const LiftedCol* __restrict__ cols,
                                                                    uint32_t n_cols, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_leaves) return;
  uint32_t h[8];
  blake2s_init(h);
  uint32_t m[16];
  const uint32_t n_blocks = n_cols == 0 ? 1u : (n_cols + 15) / 16;
  for (uint32_t b = 0; b < n_blocks; ++b) {
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
      const uint32_t c = 16 * b + k;
      uint32_t w = 0;
      if (c < n_cols) {
        const LiftedCol lc = cols[c];
        w = lc.ptr[lifted_row(i, lc.shift)];
      }
      m[k] = w;
    }
    const bool last = b + 1 == n_blocks;
    blake2s_compress(h, m, last ? 4u * n_cols : 64u * (b + 1), last);
  }
  uint4* o4 = reinterpret_cast<uint4*>(out + (size_t)i * 8);
  o4[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o4[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

__global__ void __launch_bounds__(256) lifted_leaves_poseidon252_kernel(uint32_t n_leaves,
                                                                        const LiftedCol* __restrict__ cols,
                                                                        uint32_t n_cols, uint32_t* __restrict__ out) {
  using namespace f252;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_leaves) return;
  Sponge sp;
  for (uint32_t c0 = 0; c0 < n_cols; c0 += 8) {
    const int k = n_cols - c0 < 8 ? (int)(n_cols - c0) : 8;
    uint32_t x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < k; ++j) {
      const LiftedCol lc = cols[c0 + j];
      x[j] = lc.ptr[lifted_row(i, lc.shift)];
    }
    sp.absorb(to_mont(felt_from_m31s(x, k)));
  }
  const F r = from_mont(sp.finish());
#pragma unroll
  for (int j = 0; j < 8; ++j) out[8 * (size_t)i + j] = r.w[j];
}

}  // namespace

void launch_lifted_leaves(uint32_t hasher, uint32_t lifting_log_size, const LiftedCol* cols, uint32_t n_cols,
                          uint32_t* out, hipStream_t stream) {
  const uint32_t n = n_cols ? 1u << lifting_log_size : 1u;  // no columns: the empty hasher's one leaf
  const uint32_t threads = n < 256 ? (n < 64 ? 64 : n) : 256;
  if (hasher == LIFTED_POSEIDON252)
    hipLaunchKernelGGL(lifted_leaves_poseidon252_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, stream,
                       n, cols, n_cols, out);
  else
    hipLaunchKernelGGL(lifted_leaves_blake2s_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, stream, n,
                       cols, n_cols, out);
}

}  // namespace sh
