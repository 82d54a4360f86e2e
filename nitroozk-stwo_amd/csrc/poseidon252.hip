// The reference's second Merkle channel on the device: MerkleOps<Poseidon252MerkleHasher>::
// commit_on_layer (crates/stwo/src/core/vcs/poseidon252_merkle.rs:18-41; the CUDA backend's
// poseidon252_commit_on_first_layer / _with_previous, crates/stwo/src/prover/backend/cuda/
// poseidon252.rs) and GrindOps<Poseidon252Channel>::grind (crates/stwo/src/core/channel/
// poseidon252.rs:146-156).
//
// A node's hash is a felt252, stored canonical as 8 little-endian 32-bit words. One lane computes
// one node: hash(left, right) when the layer injects no columns, else poseidon_hash_many over the
// two children (if any) and the layer's column values packed 8 M31 to a felt252. Each node costs
// one permutation per two absorbed values (214 Montgomery products), so the kernel is VALU-bound.
#include <stdexcept>
#include "felt252.h"
#include "kernels.h"

namespace sh {

using namespace f252;

namespace {

__device__ __forceinline__ F load_felt(const uint32_t* __restrict__ p) {
  F f;
#pragma unroll
  for (int i = 0; i < 8; ++i) f.w[i] = p[i];
  return f;
}

__global__ void __launch_bounds__(256) p252_layer_kernel(uint32_t n, const uint32_t* __restrict__ prev,
                                                         const uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                                         uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  M h;
  if (prev && n_cols == 0) {
    h = hash2(to_mont(load_felt(prev + 16 * (size_t)i)), to_mont(load_felt(prev + 16 * (size_t)i + 8)));
  } else {
    Sponge sp;
    if (prev) {
      sp.absorb(to_mont(load_felt(prev + 16 * (size_t)i)));
      sp.absorb(to_mont(load_felt(prev + 16 * (size_t)i + 8)));
    }
    for (uint32_t c0 = 0; c0 < n_cols; c0 += 8) {
      const int k = n_cols - c0 < 8 ? (int)(n_cols - c0) : 8;
      uint32_t x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int j = 0; j < k; ++j) x[j] = cols[c0 + j][i];
      sp.absorb(to_mont(felt_from_m31s(x, k)));
    }
    h = sp.finish();
  }
  const F r = from_mont(h);
#pragma unroll
  for (int j = 0; j < 8; ++j) out[8 * (size_t)i + j] = r.w[j];
}

// Nonces (hi << 32) | lo with lo < 2^14 (GRIND_LOW_BITS of prover/backend/simd/grind.rs:150-190),
// candidate j of the launch at hi = hi0 + j / 2^14: a hash of (prefixed digest, nonce) with pow_bits
// trailing zero bits (of its low 128 bits, as u128::trailing_zeros of the big-endian bytes 16..32)
// records the nonce with an atomic minimum, which is the search order's first.
constexpr uint32_t P252_GRIND_LOW_BITS = 14;
__global__ void __launch_bounds__(256) p252_grind_kernel(M prefixed, uint32_t pow_bits, uint64_t hi0,
                                                         unsigned long long* __restrict__ best) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nonce = ((hi0 + (j >> P252_GRIND_LOW_BITS)) << 32) | (j & ((1u << P252_GRIND_LOW_BITS) - 1));
  const F h = from_mont(hash2(prefixed, to_mont(from_u64(nonce))));
  bool ok = true;
  for (uint32_t b = 0; b < pow_bits; b += 32) {
    const uint32_t bits = pow_bits - b < 32 ? pow_bits - b : 32;
    const uint32_t mask = bits == 32 ? 0xFFFFFFFFu : ((1u << bits) - 1);
    ok = ok && (h.w[b / 32] & mask) == 0;
  }
  if (ok) atomicMin(best, (unsigned long long)nonce);
}

__global__ void set_u64_kernel(unsigned long long* p, unsigned long long v) { *p = v; }

// The permutation-rate ceiling: every lane permutes a register-resident state `iters` times (no
// memory traffic), with 16 workgroups of 256 lanes per CU, enough waves to fill every SIMD.
__global__ void __launch_bounds__(256) p252_perm_rate_kernel(uint32_t iters, uint32_t seed, uint32_t* out) {
  M s[3] = {ONE, TWO, THREE};
  s[0].w[0] ^= (seed * (threadIdx.x + 1u)) & 0xFFFFu;  // distinct per lane; the limb stays below 2^29
  for (uint32_t it = 0; it < iters; ++it) permute(s);
  const uint32_t x = s[0].w[0] ^ s[1].w[1] ^ s[2].w[2];
  if (x == 0x12345678u) out[blockIdx.x] = x;  // keeps the states live
}

}  // namespace

void launch_p252_layer(uint32_t log_size, const uint32_t* prev, const uint32_t* const* cols, uint32_t n_cols,
                       uint32_t* out, hipStream_t stream) {
  const uint32_t n = 1u << log_size;
  hipLaunchKernelGGL(p252_layer_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, prev, cols, n_cols, out);
}

double measure_permutation_rate(uint32_t* scratch, hipStream_t stream) {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint32_t blocks = (uint32_t)cus * 16, threads = 256, iters = 2;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(p252_perm_rate_kernel, dim3(blocks), dim3(threads), 0, stream, 1u, 3u, scratch);  // warm-up
  (void)hipEventRecord(e0, stream);
  hipLaunchKernelGGL(p252_perm_rate_kernel, dim3(blocks), dim3(threads), 0, stream, iters, 5u, scratch);
  (void)hipEventRecord(e1, stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double perms = (double)blocks * threads * iters;
  return ms > 0 ? perms / (ms * 1e-3) : 0.0;
}

M p252_pow_prefix(const uint32_t digest[8], uint32_t pow_bits) {
  // poseidon_hash_many([POW_PREFIX, digest, n_bits]) (Montgomery form)
  Sponge sp;
  F d;
  for (int i = 0; i < 8; ++i) d.w[i] = digest[i];
  sp.absorb(to_mont(from_u64(0x12345678u)));
  sp.absorb(to_mont(d));
  sp.absorb(to_mont(from_u64(pow_bits)));
  return sp.finish();
}

uint64_t p252_grind(const uint32_t digest[8], uint32_t pow_bits, unsigned long long* d_best, hipStream_t stream) {
  if (pow_bits > 64) throw std::invalid_argument("p252_grind: pow_bits above 64");
  const M pre = p252_pow_prefix(digest, pow_bits);
  uint64_t hi0 = 0, count = 1u << 16;  // candidates per launch: whole chunks of 2^14
  for (;;) {
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, stream, d_best, ~0ull);
    hipLaunchKernelGGL(p252_grind_kernel, dim3((uint32_t)(count / 256)), dim3(256), 0, stream, pre, pow_bits, hi0,
                       d_best);
    unsigned long long best = ~0ull;
    if (hipMemcpyAsync(&best, d_best, 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      throw std::runtime_error("p252_grind: device copy failed");
    if (best != ~0ull) return best;
    hi0 += count >> P252_GRIND_LOW_BITS;
    if (count < (1u << 22)) count <<= 2;
  }
}

void p252_hash_many_host(const uint32_t* values, size_t n, uint32_t out[8]) {
  Sponge sp;
  for (size_t k = 0; k < n; ++k) {
    F v;
    for (int i = 0; i < 8; ++i) v.w[i] = values[8 * k + i];
    sp.absorb(to_mont(v));
  }
  const F r = from_mont(sp.finish());
  for (int i = 0; i < 8; ++i) out[i] = r.w[i];
}

void p252_permute_host(uint32_t state[24]) {
  M s[3];
  for (int j = 0; j < 3; ++j) {
    F v;
    for (int i = 0; i < 8; ++i) v.w[i] = state[8 * j + i];
    s[j] = to_mont(v);
  }
  permute(s);
  for (int j = 0; j < 3; ++j) {
    const F r = from_mont(s[j]);
    for (int i = 0; i < 8; ++i) state[8 * j + i] = r.w[i];
  }
}

bool p252_canonical(const uint32_t* v) {
  F f;
  for (int i = 0; i < 8; ++i) f.w[i] = v[i];
  return !geq_p(f);
}

}  // namespace sh
