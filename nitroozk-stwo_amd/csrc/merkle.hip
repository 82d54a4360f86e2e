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
#include <stdexcept>
#include "blake2s.h"
#include "kernels.h"

namespace sh {

// Column addresses: a table, or (STRIDED) column c at base + c * stride, which needs no address
// loads at all.
struct LeafCols {
  const uint32_t* const* table;
  const uint32_t* base;
  size_t stride;
};

template <bool STRIDED>
__global__ void __launch_bounds__(256) merkle_layer_kernel(uint32_t n_nodes, const uint32_t* __restrict__ prev,
                                                           LeafCols cols, uint32_t n_cols,
                                                           uint32_t* __restrict__ out) {
  CRITICAL_PATH_WAVE();
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
  if (n_cols) {
    // Software pipeline: the column words of the next two blocks are in flight while a block
    // compresses (one wave per SIMD when a tree has few leaves, so nothing else hides the loads).
    // Loads are issued unconditionally and two blocks ahead (a column index past the end is
    // clamped to the last column), and only the last block, which may be partial, is masked when it
    // is consumed: masking every loaded word would make each block wait for all loads in flight.
    auto raw = [&](uint32_t c) {
      c = min(c, n_cols - 1);
      return gmem(STRIDED ? cols.base + (size_t)c * cols.stride : cols.table[c])[i];
    };
    const uint32_t rem = n_cols & 15, n_blocks = (n_cols + 15) >> 4;
    uint32_t nx1[16], nx2[16];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) m[k] = raw(k);
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) nx1[k] = raw(16 + k);
    for (uint32_t b = 0; b < n_blocks; ++b) {
#pragma unroll
      for (uint32_t k = 0; k < 16; ++k) nx2[k] = raw(16 * (b + 2) + k);
      const bool last = b + 1 == n_blocks;
      if (last && rem) {
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) m[k] = k < rem ? m[k] : 0u;
      }
      t += last && rem ? 4 * rem : 64;
      blake2s_compress(h, m, t, last);
#pragma unroll
      for (uint32_t k = 0; k < 16; ++k) { m[k] = nx1[k]; nx1[k] = nx2[k]; }
    }
  }
  uint4* o4 = reinterpret_cast<uint4*>(out + (size_t)i * 8);
  o4[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o4[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

static void launch_leaf_layer(uint32_t log_size, const uint32_t* prev, LeafCols cols, uint32_t n_cols, uint32_t* out,
                              hipStream_t stream) {
  uint32_t n = 1u << log_size;
  uint32_t threads = n < 256 ? (n < 64 ? 64 : n) : 256;
  if (cols.base)
    hipLaunchKernelGGL(merkle_layer_kernel<true>, dim3((n + threads - 1) / threads), dim3(threads), 0, stream, n, prev,
                       cols, n_cols, out);
  else
    hipLaunchKernelGGL(merkle_layer_kernel<false>, dim3((n + threads - 1) / threads), dim3(threads), 0, stream, n,
                       prev, cols, n_cols, out);
}

void launch_merkle_layer(uint32_t log_size, const uint32_t* prev, const uint32_t* const* cols, uint32_t n_cols,
                         uint32_t* out, hipStream_t stream) {
  launch_leaf_layer(log_size, prev, LeafCols{cols, nullptr, 0}, n_cols, out, stream);
}

void launch_merkle_layer_strided(uint32_t log_size, const uint32_t* prev, const uint32_t* base, size_t stride,
                                 uint32_t n_cols, uint32_t* out, hipStream_t stream) {
  launch_leaf_layer(log_size, prev, LeafCols{nullptr, base, stride}, n_cols, out, stream);
}

// The same node hash for at most 16 injected columns whose addresses travel as kernel
// arguments (no pointer table in device memory).
struct ColPtrs16 { const uint32_t* c[16]; };
__global__ void __launch_bounds__(256) merkle_layer_small_kernel(uint32_t n_nodes, const uint32_t* __restrict__ prev,
                                                                 ColPtrs16 cols, uint32_t n_cols,
                                                                 uint32_t* __restrict__ out) {
  CRITICAL_PATH_WAVE();
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
  if (n_cols) {
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) m[k] = k < n_cols ? cols.c[k][i] : 0u;
    blake2s_compress(h, m, t + 4 * n_cols, true);
  }
  uint4* o4 = reinterpret_cast<uint4*>(out + (size_t)i * 8);
  o4[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o4[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

void launch_merkle_layer_small(uint32_t log_size, const uint32_t* prev, const uint32_t* const* host_cols,
                               uint32_t n_cols, uint32_t* out, hipStream_t stream) {
  ColPtrs16 cp{};
  for (uint32_t k = 0; k < n_cols && k < 16; ++k) cp.c[k] = host_cols[k];
  uint32_t n = 1u << log_size;
  uint32_t threads = n < 256 ? (n < 64 ? 64 : n) : 256;
  hipLaunchKernelGGL(merkle_layer_small_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, stream, n, prev,
                     cp, n_cols, out);
}

// Up to 9 levels of a Merkle tree with no injected columns in one launch. Workgroup w reads the
// 2^L children [w 2^L, (w+1) 2^L) of level log_child and writes its 2^(L-j) nodes of level
// log_child - j (j = 1..L); the levels in between stay in LDS. The levels are latency-bound (each
// waits for the one below), so every node is hashed by a quad of lanes (blake2s_compress_quad):
// 4 * 2^(L-1) threads, lane q of a quad holding chaining words q and 4 + q.
constexpr uint32_t SUBTREE_MAX_LEVELS = 9;
struct LayerPtrs { uint32_t* p[SUBTREE_MAX_LEVELS]; };
__global__ void __launch_bounds__(1024) merkle_subtree_kernel(const uint32_t* __restrict__ child, uint32_t L,
                                                              LayerPtrs outs) {
  CRITICAL_PATH_WAVE();
  __shared__ uint32_t lds[(1u << (SUBTREE_MAX_LEVELS - 1)) * 8];
  const uint32_t w = blockIdx.x, nd = threadIdx.x >> 2, q = threadIdx.x & 3;
  uint32_t n = 1u << (L - 1);  // nodes of the first level computed by this workgroup
  for (uint32_t j = 0; j < L; ++j, n >>= 1) {
    uint32_t hA = 0, hB = 0;
    if (nd < n) {
      uint32_t m[16];
      if (j == 0) {
        const uint4* p4 = reinterpret_cast<const uint4*>(child + ((size_t)w * 2 * n + 2 * nd) * 8);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint4 v = p4[k];
          m[4 * k] = v.x; m[4 * k + 1] = v.y; m[4 * k + 2] = v.z; m[4 * k + 3] = v.w;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] = lds[2 * nd * 8 + k];
      }
      blake2s_init_quad(q, hA, hB);
      blake2s_compress_quad(hA, hB, m, 64, true, q);
    }
    __syncthreads();
    if (nd < n) {
      lds[nd * 8 + q] = hA;
      lds[nd * 8 + 4 + q] = hB;
      uint32_t* o = outs.p[j] + ((size_t)w * n + nd) * 8;
      o[q] = hA;
      o[4 + q] = hB;
    }
    __syncthreads();
  }
}

void launch_merkle_subtree(uint32_t log_child, const uint32_t* child, uint32_t levels, uint32_t* const* outs,
                           hipStream_t stream) {
  // levels <= SUBTREE_MAX_LEVELS and <= log_child; outs[j] receives level log_child - 1 - j
  if (levels == 0 || levels > SUBTREE_MAX_LEVELS || levels > log_child) throw std::invalid_argument("merkle subtree: levels");
  LayerPtrs lp{};
  for (uint32_t j = 0; j < levels; ++j) lp.p[j] = outs[j];
  uint32_t wgs = 1u << (log_child - levels);
  uint32_t threads = 4u << (levels - 1);
  if (threads < 64) threads = 64;
  hipLaunchKernelGGL(merkle_subtree_kernel, dim3(wgs), dim3(threads), 0, stream, child, levels, lp);
}

// Candidates are numbered g = hi * 2^20 + lo (lo < 2^20), i.e. in the order of the smallest-nonce
// search; one launch covers [start, start + count).
__device__ __forceinline__ void grind_candidate(const uint32_t* pw, uint32_t pow_bits, uint64_t start, uint32_t count,
                                                unsigned long long* __restrict__ result) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t g = start + i;
  uint32_t lo = (uint32_t)(g & ((1u << 20) - 1));
  uint32_t hi = (uint32_t)(g >> 20);
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) m[k] = pw[k];
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
__global__ void __launch_bounds__(256) grind_kernel(GrindPrefix prefix, uint32_t pow_bits, uint64_t start,
                                                    uint32_t count, unsigned long long* __restrict__ result) {
  grind_candidate(prefix.w, pow_bits, start, count, result);
}
__global__ void __launch_bounds__(256) grind_dev_kernel(const uint32_t* __restrict__ prefix, uint32_t pow_bits,
                                                        uint64_t start, uint32_t count,
                                                        unsigned long long* __restrict__ result) {
  uint32_t pw[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) pw[k] = prefix[k];
  grind_candidate(pw, pow_bits, start, count, result);
}

void launch_grind(const GrindPrefix& prefix, uint32_t pow_bits, uint64_t start, uint32_t count,
                  unsigned long long* result, hipStream_t stream) {
  hipLaunchKernelGGL(grind_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, prefix, pow_bits, start, count,
                     result);
}
void launch_grind_dev(const uint32_t* prefix, uint32_t pow_bits, uint64_t start, uint32_t count,
                      unsigned long long* result, hipStream_t stream) {
  hipLaunchKernelGGL(grind_dev_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, prefix, pow_bits, start,
                     count, result);
}

// ---- device-side Fiat-Shamir channel (crates/stwo/src/core/channel/blake2s.rs:19-140) --------
// state = digest (8 words) + n_draws. One lane runs it so that FRI can mix each layer root and
// draw the next folding coefficient without a host round trip.
__device__ void dev_channel_draw_u32s(uint32_t* st, uint32_t out[8]) {
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) m[k] = st[k];
  m[8] = st[8];  // n_draws, little endian
  m[9] = 0;      // the trailing zero byte, then zero padding
#pragma unroll
  for (int k = 10; k < 16; ++k) m[k] = 0;
  st[8] += 1;
  uint32_t h[8];
  blake2s_init(h);
  blake2s_compress(h, m, 37, true);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[k] = h[k];
}

// Mixes a Merkle root (8 words) into the channel and draws one QM31 (draw_secure_felt). The
// record receives the root (words 0..7) and the drawn value (words 8..11) so that the host can
// replay the transcript.
__global__ void channel_mix_root_draw_kernel(uint32_t* __restrict__ st, const uint32_t* __restrict__ root,
                                             uint32_t* __restrict__ rec) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) { m[k] = st[k]; m[8 + k] = root[k]; rec[k] = root[k]; }
  uint32_t h[8];
  blake2s_init(h);
  blake2s_compress(h, m, 64, true);
#pragma unroll
  for (int k = 0; k < 8; ++k) st[k] = h[k];
  st[8] = 0;
  // draw_base_felts: redraw until all 8 words are below 2P; the value uses the first 4
  for (int attempt = 0; attempt < 64; ++attempt) {
    uint32_t u[8];
    dev_channel_draw_u32s(st, u);
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) ok &= u[k] < 2 * P;
    if (ok) {
#pragma unroll
      for (int k = 0; k < 4; ++k) rec[8 + k] = u[k] >= P ? u[k] - P : u[k];
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) rec[8 + k] = P;  // not canonical: the host replay reports it
}

void launch_channel_mix_root_draw(uint32_t* state, const uint32_t* root, uint32_t* record, hipStream_t stream) {
  hipLaunchKernelGGL(channel_mix_root_draw_kernel, dim3(1), dim3(64), 0, stream, state, root, record);
}

}  // namespace sh
