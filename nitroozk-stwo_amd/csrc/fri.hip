// FRI commit phase on the device (FriProver::commit, prover/fri.rs:100-272), one fused step per
// layer: the fold from the layer above (fold_circle_into_line, core/fri.rs:751-780, for the first
// line layer; fold_line, core/fri.rs:702-725, after it), the Merkle tree of the layer's 4
// coordinate columns (leaf = H(c0[i] || c1[i] || c2[i] || c3[i]), node = H(left || right);
// vcs/blake2_merkle.rs:20-58), and the channel step that mixes the root and draws the next
// folding coefficient (Blake2sChannel::mix_root and draw_secure_felt, channel/blake2s.rs:19-140).
//
// Layers of 2^9 .. 2^18 values run one launch each (fri_layer_kernel) over 2^(l-c) workgroups of
// 256 threads: a workgroup folds and hashes a chunk of 2^c leaves and builds the chunk's subtree;
// the workgroup that finishes last (a ticket counter, no waiting) hashes the levels above the chunk
// roots and runs the channel step. Layers of at most 2^8 values run in one workgroup, all of them
// in one launch (fri_small_kernel), with each layer's values kept in LDS for the next fold.
//
// Inside a workgroup a Merkle level is held in LDS word-major (word k of node i at k * n + i), so a
// lane reads both children's word k with one 8-byte load and a wave's loads hit distinct banks.
// Wide levels hash one node per lane (Blake2s is issue-bound: one wave per SIMD already reaches
// peak throughput); levels of at most 64 parents hash one node per quad of lanes, which has about
// half the latency of one lane.
#include <stdexcept>
#include <type_traits>
#include "blake2s.h"
#include "field.h"
#include "kernels.h"

namespace sh {

namespace {

constexpr uint32_t FL_THREADS = 256;

// Lane k of the quad, broadcast to the 4 lanes (DPP quad_perm [k, k, k, k]).
template <int K>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, K * 0x55, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ uint32_t quad_perm(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}

// An empty asm that keeps a value in a VGPR: a lane-dependent choice among such values stays a
// register select instead of becoming an indexed load from a scratch copy of the array.
__device__ __forceinline__ uint32_t pin(uint32_t x) {
  asm volatile("" : "+v"(x));
  return x;
}

__device__ __forceinline__ void store8(uint32_t* dst, const uint32_t h[8]) {
  GLOBAL_AS u32x4* o = reinterpret_cast<GLOBAL_AS u32x4*>(gmem(dst));
  o[0] = u32x4{h[0], h[1], h[2], h[3]};
  o[1] = u32x4{h[4], h[5], h[6], h[7]};
}

// Inverse circle twiddle of output h: line layer 0 pairs [x, y] -> [y, -y, -x, x].
__device__ __forceinline__ uint32_t fri_circle_itw(const uint32_t* __restrict__ lt0, uint32_t h) {
  uint32_t x = gmem(lt0)[(h >> 2) * 2], y = gmem(lt0)[(h >> 2) * 2 + 1];
  switch (h & 3) {
    case 0: return y;
    case 1: return m31_neg(y);
    case 2: return m31_neg(x);
    default: return x;
  }
}

// One folded value: (f0, f1) = ibutterfly(f[2i], f[2i+1], tw), result f0 + alpha f1.
__device__ __forceinline__ qm31 fri_fold(qm31 fp, qm31 fn, uint32_t tw, qm31 alpha) {
  qm31 f0 = qm_add(fp, fn);
  qm31 f1 = qm_mul_m31(qm_sub(fp, fn), tw);
  return qm_add(f0, qm_mul(alpha, f1));
}
__device__ __forceinline__ uint32_t fri_twiddle(const FriLayerDesc& d, uint32_t i) {
  return d.mode == FRI_FOLD_CIRCLE ? fri_circle_itw(d.tw, i) : gmem(d.tw)[i];
}

// Value i of the layer: read (FRI_LEAVES) or folded from the layer above in global memory, which
// an earlier launch wrote; a folded value is also stored to the layer's columns.
__device__ __forceinline__ qm31 fri_value_global(const FriLayerDesc& d, uint32_t i, qm31 alpha) {
  if (d.mode == FRI_LEAVES) return {gmem(d.src.c[0])[i], gmem(d.src.c[1])[i], gmem(d.src.c[2])[i], gmem(d.src.c[3])[i]};
  u32x2 w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = *reinterpret_cast<const GLOBAL_AS u32x2*>(gmem(d.src.c[k]) + 2 * i);
  qm31 v = fri_fold({w[0].x, w[1].x, w[2].x, w[3].x}, {w[0].y, w[1].y, w[2].y, w[3].y}, fri_twiddle(d, i), alpha);
  gmem(d.dst.c[0])[i] = v.a; gmem(d.dst.c[1])[i] = v.b; gmem(d.dst.c[2])[i] = v.c; gmem(d.dst.c[3])[i] = v.d;
  return v;
}

__device__ __forceinline__ void leaf_hash(qm31 v, uint32_t h[8]) {
  uint32_t m[16] = {v.a, v.b, v.c, v.d, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  blake2s_init(h);
  blake2s_compress(h, m, 16, true);
}

// The m parents of the 2m nodes in `in` (word k of node i at in[k * 2m + i]) into `out` (word k of
// parent p at out[k * m + p]), each parent also stored as 8 consecutive words at g + 8 p. Called by
// every thread of the workgroup; ends with a barrier.
__device__ __forceinline__ void fl_level(const uint32_t* in, uint32_t* out, uint32_t m, uint32_t* __restrict__ g) {
  const uint32_t t = threadIdx.x, si = 2 * m;
  if (m > 64) {
    for (uint32_t p = t; p < m; p += FL_THREADS) {
      uint32_t msg[16], h[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint2 w = *reinterpret_cast<const uint2*>(in + k * si + 2 * p);
        msg[k] = w.x;
        msg[8 + k] = w.y;
      }
      blake2s_init(h);
      blake2s_compress(h, msg, 64, true);
#pragma unroll
      for (int k = 0; k < 8; ++k) out[k * m + p] = h[k];
      store8(g + 8 * p, h);
    }
  } else {
    const uint32_t p = t >> 2, q = t & 3;
    if (p < m) {  // whole quads are active or not
      uint32_t msg[16];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint2 w = *reinterpret_cast<const uint2*>(in + k * si + 2 * p);
        msg[k] = w.x;
        msg[8 + k] = w.y;
      }
      uint32_t hA, hB;
      blake2s_init_quad(q, hA, hB);
      blake2s_compress_quad(hA, hB, msg, 64, true, q);
      out[q * m + p] = hA;
      out[(4 + q) * m + p] = hB;
      gmem(g)[8 * p + q] = hA;
      gmem(g)[8 * p + 4 + q] = hB;
    }
  }
  __syncthreads();
}

// Hashes levels up from the n nodes held in `a` (word-major, row stride n) until one node remains,
// alternating between the buffers a and b; `heap` is the layer's Merkle heap, `lev` the level of
// the nodes in a and `first` the heap index of a's node 0 within that level. Returns the buffer
// that holds the top node (row stride 1 at the end).
__device__ __forceinline__ const uint32_t* fl_levels(uint32_t* a, uint32_t* b, uint32_t n, uint32_t lev, uint32_t first,
                                                     uint32_t* __restrict__ heap) {
  uint32_t* in = a;
  uint32_t* out = b;
  while (n > 1) {
    const uint32_t m = n >> 1;
    --lev;
    fl_level(in, out, m, heap + 8 * (((size_t)1 << lev) - 1 + (size_t)(first >> 1)));
    first >>= 1;
    uint32_t* x = in;
    in = out;
    out = x;
    n = m;
  }
  return in;
}

// The channel step on one quad of lanes (q = lane within the quad; all 4 active): st (digest and
// n_draws, identical in the 4 lanes) absorbs the root and a QM31 is drawn, redrawing until all 8
// drawn words are below 2P (draw_base_felts). On return st and alpha are identical in the 4 lanes;
// alpha is P in every coordinate (not canonical) if 64 draws fail, which the host replay reports.
__device__ __forceinline__ void fl_channel(uint32_t st[9], const uint32_t root[8], uint32_t q, uint32_t alpha[4]) {
  uint32_t m[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) { m[k] = st[k]; m[8 + k] = root[k]; }
  uint32_t hA, hB;
  blake2s_init_quad(q, hA, hB);
  blake2s_compress_quad(hA, hB, m, 64, true, q);
  st[0] = quad_bcast<0>(hA); st[1] = quad_bcast<1>(hA); st[2] = quad_bcast<2>(hA); st[3] = quad_bcast<3>(hA);
  st[4] = quad_bcast<0>(hB); st[5] = quad_bcast<1>(hB); st[6] = quad_bcast<2>(hB); st[7] = quad_bcast<3>(hB);
  st[8] = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) alpha[k] = P;
  for (int attempt = 0; attempt < 64; ++attempt) {
#pragma unroll
    for (int k = 0; k < 8; ++k) m[k] = st[k];
    m[8] = st[8];  // n_draws, little endian, then the zero byte and zero padding
#pragma unroll
    for (int k = 9; k < 16; ++k) m[k] = 0;
    st[8] += 1;
    blake2s_init_quad(q, hA, hB);
    blake2s_compress_quad(hA, hB, m, 37, true, q);
    uint32_t ok = (hA < 2 * P && hB < 2 * P) ? 1u : 0u;
    ok &= quad_perm<0x39>(ok);  // lanes (i, i + 1)
    ok &= quad_perm<0x4E>(ok);  // all four lanes
    if (ok) {
      const uint32_t a = hA >= P ? hA - P : hA;  // lane k holds drawn word k
      alpha[0] = quad_bcast<0>(a); alpha[1] = quad_bcast<1>(a); alpha[2] = quad_bcast<2>(a); alpha[3] = quad_bcast<3>(a);
      break;
    }
  }
}

// Lane q of the quad writes words q and q + 4 of the record (root, then alpha in words 8..11) and of
// the channel state (n_draws from lane 0) to rec and st_out, and alpha word q to alpha_out when
// given. Selects instead of indexing by q keep the arrays in registers.
template <class StPtr, class AlphaPtr>
__device__ __forceinline__ void fl_publish(uint32_t q, const uint32_t st[9], const uint32_t root[8],
                                           const uint32_t alpha[4], uint32_t* __restrict__ rec, StPtr st_out,
                                           AlphaPtr alpha_out) {
  uint32_t r_lo = pin(root[0]), r_hi = pin(root[4]), a = pin(alpha[0]), s_lo = pin(st[0]), s_hi = pin(st[4]);
#pragma unroll
  for (uint32_t k = 1; k < 4; ++k) {
    const uint32_t rl = pin(root[k]), rh = pin(root[4 + k]), ak = pin(alpha[k]), sl = pin(st[k]), sh = pin(st[4 + k]);
    if (q == k) { r_lo = rl; r_hi = rh; a = ak; s_lo = sl; s_hi = sh; }
  }
  gmem(rec)[q] = r_lo;
  gmem(rec)[4 + q] = r_hi;
  gmem(rec)[8 + q] = a;
  st_out[q] = s_lo;
  st_out[4 + q] = s_hi;
  if (q == 0) st_out[8] = st[8];
  if constexpr (!std::is_same_v<AlphaPtr, std::nullptr_t>) alpha_out[q] = a;
}

}  // namespace

// One FRI layer of 2^d.log values (9 <= log <= 18) over 2^(log - c) workgroups of 256 threads,
// 2^c leaves each (7 <= c <= 10, log - c <= 8). `ticket` is zero before the launch.
__global__ void __launch_bounds__(256) fri_layer_kernel(FriLayerDesc d, uint32_t c, uint32_t* __restrict__ state,
                                                        uint32_t* __restrict__ ticket) {
  __shared__ __align__(16) uint32_t bufA[8 << 10];
  __shared__ __align__(16) uint32_t bufB[8 << 9];
  __shared__ uint32_t last;
  const uint32_t t = threadIdx.x, w = blockIdx.x;
  const uint32_t l = d.log, C = 1u << c;
  qm31 alpha = qm_zero();
  if (d.mode != FRI_LEAVES) alpha = {gmem(d.alpha)[0], gmem(d.alpha)[1], gmem(d.alpha)[2], gmem(d.alpha)[3]};
  for (uint32_t r = t; r < C; r += FL_THREADS) {
    const uint32_t i = w * C + r;
    const qm31 v = fri_value_global(d, i, alpha);
    uint32_t h[8];
    leaf_hash(v, h);
#pragma unroll
    for (int k = 0; k < 8; ++k) bufA[k * C + r] = h[k];
    store8(d.nodes + 8 * (((size_t)1 << l) - 1 + i), h);
  }
  __syncthreads();
  const uint32_t* top = fl_levels(bufA, bufB, C, l, w * C, d.nodes);
  uint32_t root[8];
  if (gridDim.x > 1) {
    // The chunk root is published with device-scope atomic exchanges, which are performed where
    // every XCD reads them; wave 0 waits for their results before it takes the ticket. This orders
    // root and ticket without a device-scope fence, which would write back the whole L2 of the
    // XCD once per workgroup.
    if (t < 8) {
      uint32_t* slot = d.nodes + 8 * (((size_t)1 << (l - c)) - 1 + w) + t;
      const uint32_t old = __hip_atomic_exchange(slot, top[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      asm volatile("" ::"v"(old) : "memory");
    }
    if (t == 0)
      last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1 ? 1u : 0u;
    __syncthreads();
    if (!last) return;
    // every other workgroup has taken its ticket: the counter goes back to zero for the next launch
    if (t == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the chunk roots, read with read-modify-write atomics of the same scope (performed where the
    // exchanges were); the operand is opaque so that they are not turned into plain loads
    const uint32_t R = gridDim.x, lev = l - c;
    uint32_t* g = d.nodes + 8 * (((size_t)1 << lev) - 1);
    const uint32_t zero = pin(0u);
    for (uint32_t i = t; i < R; i += FL_THREADS) {
      uint32_t r[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) r[k] = __hip_atomic_fetch_or(g + 8 * i + k, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
      for (int k = 0; k < 8; ++k) bufA[k * R + i] = r[k];
    }
    __syncthreads();
    top = fl_levels(bufA, bufB, R, lev, 0, d.nodes);
  }
  if (t < 4) {
#pragma unroll
    for (int k = 0; k < 8; ++k) root[k] = top[k];
    uint32_t st[9], a[4];
#pragma unroll
    for (int k = 0; k < 9; ++k) st[k] = d.init ? d.st0[k] : state[k];
    fl_channel(st, root, t, a);
    fl_publish(t, st, root, a, d.rec, state, nullptr);
    if (d.init && t < 2) state[14 + t] = ~0u;
  }
}

// Layers of at most 2^8 values in one workgroup of 256 threads, in the order given: each entry
// folds from the entry before it (values in LDS) or, for the first entry, from global memory, then
// (when d.commit) builds the Merkle tree and runs the channel step. The channel state stays in
// LDS between layers and is written back at the end.
__global__ void __launch_bounds__(256) fri_small_kernel(const FriLayerDesc* __restrict__ descs, uint32_t n,
                                                        uint32_t* __restrict__ state) {
  __shared__ uint32_t V[4][256];
  __shared__ __align__(16) uint32_t bufA[8 << 8];
  __shared__ __align__(16) uint32_t bufB[8 << 7];
  __shared__ uint32_t st_s[9], alpha_s[4];
  const uint32_t t = threadIdx.x;
  if (t < 9) st_s[t] = state[t];
  __syncthreads();
  for (uint32_t e = 0; e < n; ++e) {
    const FriLayerDesc d = descs[e];
    const uint32_t N = 1u << d.log;
    qm31 alpha = qm_zero();
    if (d.mode != FRI_LEAVES)
      alpha = e == 0 ? qm31{gmem(d.alpha)[0], gmem(d.alpha)[1], gmem(d.alpha)[2], gmem(d.alpha)[3]}
                     : qm31{alpha_s[0], alpha_s[1], alpha_s[2], alpha_s[3]};
    qm31 v = qm_zero();
    if (t < N) {
      if (e == 0) {
        v = fri_value_global(d, t, alpha);
      } else {
        // d.mode is a fold: the layer above is in V
        const qm31 fp = {V[0][2 * t], V[1][2 * t], V[2][2 * t], V[3][2 * t]};
        const qm31 fn = {V[0][2 * t + 1], V[1][2 * t + 1], V[2][2 * t + 1], V[3][2 * t + 1]};
        v = fri_fold(fp, fn, fri_twiddle(d, t), alpha);
        gmem(d.dst.c[0])[t] = v.a; gmem(d.dst.c[1])[t] = v.b; gmem(d.dst.c[2])[t] = v.c; gmem(d.dst.c[3])[t] = v.d;
      }
    }
    __syncthreads();  // every read of V is done
    if (t < N) { V[0][t] = v.a; V[1][t] = v.b; V[2][t] = v.c; V[3][t] = v.d; }
    __syncthreads();
    if (!d.commit) continue;
    if (N >= 128) {
      if (t < N) {
        uint32_t h[8];
        leaf_hash(v, h);
#pragma unroll
        for (int k = 0; k < 8; ++k) bufA[k * N + t] = h[k];
        store8(d.nodes + 8 * ((size_t)N - 1 + t), h);
      }
    } else {
      const uint32_t p = t >> 2, q = t & 3;
      if (p < N) {
        uint32_t msg[16] = {V[0][p], V[1][p], V[2][p], V[3][p], 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t hA, hB;
        blake2s_init_quad(q, hA, hB);
        blake2s_compress_quad(hA, hB, msg, 16, true, q);
        bufA[q * N + p] = hA;
        bufA[(4 + q) * N + p] = hB;
        gmem(d.nodes)[8 * (N - 1 + p) + q] = hA;
        gmem(d.nodes)[8 * (N - 1 + p) + 4 + q] = hB;
      }
    }
    __syncthreads();
    const uint32_t* top = fl_levels(bufA, bufB, N, d.log, 0, d.nodes);
    if (t < 4) {
      uint32_t root[8], st[9], a[4];
#pragma unroll
      for (int k = 0; k < 8; ++k) root[k] = top[k];
#pragma unroll
      for (int k = 0; k < 9; ++k) st[k] = st_s[k];
      fl_channel(st, root, t, a);
      fl_publish(t, st, root, a, d.rec, st_s, alpha_s);
    }
    __syncthreads();
  }
  if (t < 9) state[t] = st_s[t];
}

// The last layer's interpolation, the channel mix of the last polynomial and the proof-of-work
// prefix (see launch_fri_last_pow). The butterflies follow line_interpolate on the host: the values
// in natural order, then per level (a, b) -> (a + b, (a - b) x^-1) on pairs half a level apart.
__global__ void __launch_bounds__(64) fri_last_pow_kernel(const uint32_t* __restrict__ vals, uint32_t L, uint32_t BL,
                                                          const uint32_t* __restrict__ xi, uint32_t* __restrict__ state,
                                                          uint32_t pow_bits, uint32_t* __restrict__ prefix) {
  __shared__ uint32_t W[4][64];
  __shared__ uint32_t M[8 + 4 * 16];  // the digest, then the last polynomial's coefficients
  const uint32_t t = threadIdx.x, n = 1u << L, bound = 1u << BL;
  if (t < n)
    for (int k = 0; k < 4; ++k) W[k][t] = gmem(vals)[k * n + bitrev(t, L)];
  __syncthreads();
  uint32_t off = 0;
  for (uint32_t ds = n; ds > 1; ds >>= 1) {
    const uint32_t half = ds >> 1;
    if (t < n / 2) {
      const uint32_t ch = (t / half) * ds, i = t % half, x = gmem(xi)[off + i];
      for (int k = 0; k < 4; ++k) {
        const uint32_t a = W[k][ch + i], b = W[k][ch + half + i];
        W[k][ch + i] = m31_add(a, b);
        W[k][ch + half + i] = m31_mul(m31_sub(a, b), x);
      }
    }
    off += half;
    __syncthreads();
  }
  // coefficient j of the last polynomial is the interpolant's coefficient bitrev_L(bitrev_BL(j))
  if (t < bound) {
    const uint32_t src = bitrev(bitrev(t, BL), L), inv_n = gmem(xi)[n - 1];
    for (int k = 0; k < 4; ++k) M[8 + 4 * t + k] = m31_mul(W[k][src], inv_n);
  }
  if (t < 8) M[t] = state[t];
  __syncthreads();
  // the two hashes on one quad of lanes (lane q holds chaining words q and q + 4)
  if (t >= 4) return;
  const uint32_t q = t, nbytes = 32 + 16 * bound;
  uint32_t hA, hB;
  blake2s_init_quad(q, hA, hB);
  for (uint32_t o = 0; o < nbytes; o += 64) {
    uint32_t m[16];
    for (int k = 0; k < 16; ++k) m[k] = o + 4 * k < nbytes ? M[o / 4 + k] : 0u;
    const bool last = o + 64 >= nbytes;
    blake2s_compress_quad(hA, hB, m, last ? nbytes : o + 64, last, q);
  }
  const uint32_t m[16] = {0x12345678u,     0, 0, 0, quad_bcast<0>(hA), quad_bcast<1>(hA), quad_bcast<2>(hA),
                          quad_bcast<3>(hA), quad_bcast<0>(hB), quad_bcast<1>(hB), quad_bcast<2>(hB),
                          quad_bcast<3>(hB), pow_bits, 0, 0, 0};
  state[q] = hA;
  state[4 + q] = hB;
  if (q == 0) state[8] = 0;
  uint32_t pA, pB;
  blake2s_init_quad(q, pA, pB);
  blake2s_compress_quad(pA, pB, m, 52, true, q);
  prefix[q] = pA;
  prefix[4 + q] = pB;
}

void launch_fri_last_pow(const uint32_t* vals, uint32_t L, uint32_t BL, const uint32_t* xi, uint32_t* state,
                         uint32_t pow_bits, uint32_t* prefix, hipStream_t stream) {
  if (L > 6 || BL > L || BL > 4) throw std::invalid_argument("launch_fri_last_pow: last layer too large");
  hipLaunchKernelGGL(fri_last_pow_kernel, dim3(1), dim3(64), 0, stream, vals, L, BL, xi, state, pow_bits, prefix);
}

void launch_fri_layer(const FriLayerDesc& d, uint32_t* state, uint32_t* ticket, hipStream_t stream) {
  // chunk: 2^c leaves, at most 2^8 chunks. Up to 2^14 values, chunks of 2^7 leaves, whose levels
  // above the leaves are all quad levels (at most 64 parents), at the price of one more quad level
  // above the chunk roots.
  const uint32_t c = d.log > 16 ? d.log - 8 : (d.log > 14 ? 8 : 7);
  hipLaunchKernelGGL(fri_layer_kernel, dim3(1u << (d.log - c)), dim3(FL_THREADS), 0, stream, d, c, state, ticket);
}

void launch_fri_small(const FriLayerDesc* descs_dev, uint32_t n, uint32_t* state, hipStream_t stream) {
  hipLaunchKernelGGL(fri_small_kernel, dim3(1), dim3(FL_THREADS), 0, stream, descs_dev, n, state);
}

}  // namespace sh
