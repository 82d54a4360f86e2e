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
  if (n_cols) {
    // software pipeline: the next block's 16 column words are loaded while this block compresses
    uint32_t nw = n_cols < 16 ? n_cols : 16;
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) m[k] = k < nw ? cols[k][i] : 0u;
    for (uint32_t cs = 0; cs < n_cols; cs += 16) {
      uint32_t nx[16];
      uint32_t ncs = cs + 16;
      uint32_t nnw = ncs < n_cols ? (n_cols - ncs < 16 ? n_cols - ncs : 16) : 0;
#pragma unroll
      for (uint32_t k = 0; k < 16; ++k) nx[k] = k < nnw ? cols[ncs + k][i] : 0u;
      t += 4 * nw;
      blake2s_compress(h, m, t, ncs >= n_cols);
#pragma unroll
      for (uint32_t k = 0; k < 16; ++k) m[k] = nx[k];
      nw = nnw;
    }
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

// The same node hash for at most 16 injected columns whose addresses travel as kernel
// arguments (no pointer table in device memory).
struct ColPtrs16 { const uint32_t* c[16]; };
__global__ void __launch_bounds__(256) merkle_layer_small_kernel(uint32_t n_nodes, const uint32_t* __restrict__ prev,
                                                                 ColPtrs16 cols, uint32_t n_cols,
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

// Up to 8 levels of a Merkle tree with no injected columns in one launch. Workgroup w reads the
// 2^L children [w 2^L, (w+1) 2^L) of level log_child and writes its 2^(L-j) nodes of level
// log_child - j (j = 1..L); the levels in between stay in LDS. The levels are latency-bound (each
// waits for the one below), so every node is hashed by a quad of lanes (blake2s_compress_quad):
// 4 * 2^(L-1) threads, lane q of a quad holding chaining words q and 4 + q.
struct LayerPtrs { uint32_t* p[8]; };
__global__ void __launch_bounds__(512) merkle_subtree_kernel(const uint32_t* __restrict__ child, uint32_t L,
                                                             LayerPtrs outs) {
  __shared__ uint32_t lds[128 * 8];
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
  // levels <= 8 and <= log_child; outs[j] receives level log_child - 1 - j
  LayerPtrs lp{};
  for (uint32_t j = 0; j < levels; ++j) lp.p[j] = outs[j];
  uint32_t wgs = 1u << (log_child - levels);
  uint32_t threads = 4u << (levels - 1);
  if (threads < 64) threads = 64;
  hipLaunchKernelGGL(merkle_subtree_kernel, dim3(wgs), dim3(threads), 0, stream, child, levels, lp);
}

// Candidates are numbered g = hi * 2^20 + lo (lo < 2^20), i.e. in the order of the smallest-nonce
// search; one launch covers [start, start + count).
__global__ void __launch_bounds__(256) grind_kernel(const uint32_t* __restrict__ prefix, uint32_t pow_bits,
                                                    uint64_t start, uint32_t count,
                                                    unsigned long long* __restrict__ result) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t g = start + i;
  uint32_t lo = (uint32_t)(g & ((1u << 20) - 1));
  uint32_t hi = (uint32_t)(g >> 20);
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

void launch_grind(const uint32_t* prefix, uint32_t pow_bits, uint64_t start, uint32_t count,
                  unsigned long long* result, hipStream_t stream) {
  hipLaunchKernelGGL(grind_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, prefix, pow_bits, start, count,
                     result);
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

// ---- FRI tail: every layer of at most 2^12 values in one workgroup -----------------------------
// Layer l (l = l_start down to last_log + 1) holds 2^l QM31 values as 4 coordinate columns. Per
// layer: the Merkle tree of its 4 columns (leaf = H(c0[i] || c1[i] || c2[i] || c3[i])), the root
// mixed into the channel and a folding coefficient drawn (FriProver::commit_inner_layers,
// prover/fri.rs:100-272), then fold_line into layer l - 1 (core/fri.rs:702-725). Layouts in global
// memory, shared with the host:
//   vals:  layer l at 4 (2^(l_start+1) - 2^(l+1)) words, coordinate k at + k 2^l
//   nodes: tree of layer l at sum_{m = l+1}^{l_start} 8 (2^(m+1) - 1) words; level j of it
//          (2^j nodes of 8 words) at + 8 (2^j - 1)
//   rec:   12 words per layer (root, drawn coefficient), layer l_start first
constexpr uint32_t FRI_TAIL_MAX_LOG = 12;

__device__ __forceinline__ void leaf4(const uint32_t (*V)[1 << FRI_TAIL_MAX_LOG], uint32_t i, uint32_t h[8]) {
  uint32_t m[16] = {V[0][i], V[1][i], V[2][i], V[3][i], 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  blake2s_init(h);
  blake2s_compress(h, m, 16, true);
}
__device__ __forceinline__ void node2(const uint32_t* l, const uint32_t* r, uint32_t h[8]) {
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) { m[k] = l[k]; m[8 + k] = r[k]; }
  blake2s_init(h);
  blake2s_compress(h, m, 64, true);
}
__device__ __forceinline__ void store8(uint32_t* dst, const uint32_t h[8]) {
  uint4* o = reinterpret_cast<uint4*>(dst);
  o[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

__global__ void __launch_bounds__(1024) fri_tail_kernel(uint32_t l_start, uint32_t last_log, uint32_t* __restrict__ vals,
                                                        uint32_t* __restrict__ nodes, uint32_t* __restrict__ state,
                                                        uint32_t* __restrict__ rec, const uint32_t* __restrict__ itree,
                                                        uint32_t tree_len) {
  __shared__ uint32_t V[4][1 << FRI_TAIL_MAX_LOG];
  __shared__ uint32_t Nd[(1 << (FRI_TAIL_MAX_LOG - 1)) * 8];
  __shared__ uint32_t alpha_s[4];
  const uint32_t t = threadIdx.x, T = blockDim.x;
  size_t voff = 0, noff = 0;
  {
    const uint32_t n = 1u << l_start;
    for (uint32_t i = t; i < n; i += T)
#pragma unroll
      for (int k = 0; k < 4; ++k) V[k][i] = vals[(size_t)k * n + i];
  }
  __syncthreads();
  uint32_t layer = 0;
  for (uint32_t l = l_start; l > last_log; --l, ++layer) {
    const uint32_t n = 1u << l;
    uint32_t* tree = nodes + noff;
    uint32_t top;  // level held in Nd after the leaf step
    if (l == FRI_TAIL_MAX_LOG) {
      // 2^12 leaves do not fit Nd: each thread hashes two leaves and their parent
      for (uint32_t i = t; i < n / 2; i += T) {
        uint32_t h0[8], h1[8], p[8];
        leaf4(V, 2 * i, h0);
        leaf4(V, 2 * i + 1, h1);
        store8(tree + 8 * ((n - 1) + 2 * i), h0);
        store8(tree + 8 * ((n - 1) + 2 * i + 1), h1);
        node2(h0, h1, p);
        store8(tree + 8 * ((n / 2 - 1) + i), p);
#pragma unroll
        for (int k = 0; k < 8; ++k) Nd[8 * i + k] = p[k];
      }
      top = l - 1;
    } else if (n > T / 4) {
      for (uint32_t i = t; i < n; i += T) {
        uint32_t h[8];
        leaf4(V, i, h);
        store8(tree + 8 * ((n - 1) + i), h);
#pragma unroll
        for (int k = 0; k < 8; ++k) Nd[8 * i + k] = h[k];
      }
      top = l;
    } else {
      // few leaves: one quad of lanes per leaf (latency-bound)
      const uint32_t i = t >> 2, q = t & 3;
      if (i < n) {
        uint32_t msg[16] = {V[0][i], V[1][i], V[2][i], V[3][i], 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t hA, hB;
        blake2s_init_quad(q, hA, hB);
        blake2s_compress_quad(hA, hB, msg, 16, true, q);
        tree[8 * ((n - 1) + i) + q] = hA;
        tree[8 * ((n - 1) + i) + 4 + q] = hB;
        Nd[8 * i + q] = hA;
        Nd[8 * i + 4 + q] = hB;
      }
      top = l;
    }
    __syncthreads();
    for (uint32_t j = top; j > 0; --j) {
      const uint32_t m = 1u << (j - 1);  // parents at level j - 1 (m <= 1024)
      if (m > T / 4) {
        uint32_t h[8];
        if (t < m) node2(Nd + 16 * t, Nd + 16 * t + 8, h);
        __syncthreads();
        if (t < m) {
#pragma unroll
          for (int k = 0; k < 8; ++k) Nd[8 * t + k] = h[k];
          store8(tree + 8 * ((m - 1) + t), h);
        }
      } else {
        // one quad of lanes per parent: the level is latency-bound
        const uint32_t i = t >> 2, q = t & 3;
        uint32_t hA = 0, hB = 0;
        if (i < m) {
          uint32_t msg[16];
#pragma unroll
          for (int k = 0; k < 16; ++k) msg[k] = Nd[16 * i + k];
          blake2s_init_quad(q, hA, hB);
          blake2s_compress_quad(hA, hB, msg, 64, true, q);
        }
        __syncthreads();
        if (i < m) {
          Nd[8 * i + q] = hA;
          Nd[8 * i + 4 + q] = hB;
          tree[8 * ((m - 1) + i) + q] = hA;
          tree[8 * ((m - 1) + i) + 4 + q] = hB;
        }
      }
      __syncthreads();
    }
    if (t == 0) {
      // mix_root, then draw_secure_felt (redraw until all 8 words are below 2P)
      uint32_t m[16], h[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) { m[k] = state[k]; m[8 + k] = Nd[k]; rec[12 * layer + k] = Nd[k]; }
      blake2s_init(h);
      blake2s_compress(h, m, 64, true);
#pragma unroll
      for (int k = 0; k < 8; ++k) state[k] = h[k];
      state[8] = 0;
      uint32_t a[4] = {P, P, P, P};  // not canonical if no draw is accepted: the host replay reports it
      for (int attempt = 0; attempt < 64; ++attempt) {
        uint32_t u[8];
        dev_channel_draw_u32s(state, u);
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) ok &= u[k] < 2 * P;
        if (ok) {
#pragma unroll
          for (int k = 0; k < 4; ++k) a[k] = u[k] >= P ? u[k] - P : u[k];
          break;
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) { rec[12 * layer + 8 + k] = a[k]; alpha_s[k] = a[k]; }
    }
    __syncthreads();
    const qm31 alpha = {alpha_s[0], alpha_s[1], alpha_s[2], alpha_s[3]};
    // fold_line: f'[i] = f0 + alpha f1, (f0, f1) = ibutterfly(f[2i], f[2i+1], 1/x_i)
    const uint32_t nn = n / 2;
    const uint32_t* lt = itree + tree_len - n;
    uint32_t* nv = vals + voff + 4 * (size_t)n;
    qm31 r[2];
    for (uint32_t i = t, s = 0; i < nn; i += T, ++s) {
      qm31 fx = {V[0][2 * i], V[1][2 * i], V[2][2 * i], V[3][2 * i]};
      qm31 fnx = {V[0][2 * i + 1], V[1][2 * i + 1], V[2][2 * i + 1], V[3][2 * i + 1]};
      qm31 f0 = qm_add(fx, fnx);
      qm31 f1 = qm_mul_m31(qm_sub(fx, fnx), lt[i]);
      r[s] = qm_add(f0, qm_mul(alpha, f1));
    }
    __syncthreads();
    for (uint32_t i = t, s = 0; i < nn; i += T, ++s) {
      V[0][i] = r[s].a; V[1][i] = r[s].b; V[2][i] = r[s].c; V[3][i] = r[s].d;
      nv[i] = r[s].a; nv[nn + i] = r[s].b; nv[2 * nn + i] = r[s].c; nv[3 * nn + i] = r[s].d;
    }
    __syncthreads();
    voff += 4 * (size_t)n;
    noff += 8 * (size_t)(2 * n - 1);
  }
}

void launch_fri_tail(uint32_t l_start, uint32_t last_log, uint32_t* vals, uint32_t* nodes, uint32_t* state,
                     uint32_t* rec, const uint32_t* itree, uint32_t tree_len, hipStream_t stream) {
  hipLaunchKernelGGL(fri_tail_kernel, dim3(1), dim3(1024), 0, stream, l_start, last_log, vals, nodes, state, rec,
                     itree, tree_len);
}

}  // namespace sh
