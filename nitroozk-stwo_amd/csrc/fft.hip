// Circle FFT over batches of same-size M31 columns (PolyOps::interpolate / evaluate).
//
// Layer structure (prover/backend/cpu/circle.rs:28-80,190-228,295-335):
//   stage 0 (circle layer): pairs (2h, 2h+1), twiddle from line layer 0 pairs [x, y] -> [y,-y,-x,x]
//   stage s >= 1 (line layer s-1): pairs (idx, idx + 2^s), twiddle tree[len - 2^(K-s) + (idx >> (s+1))]
// interpolate runs stages 0..K-1 with inverse butterflies and inverse twiddles, then scales by
// 2^-K; evaluate runs stages K-1..0 with forward butterflies. Evaluating k coefficients onto a
// 2^b-times larger domain skips the top b stages: on zero-padded input they only copy, so the
// first pass reads the coefficient column with the index masked to 2^k.
//
// MI355X mapping: one workgroup owns a tile of 2^S * 2^c words in LDS and runs S consecutive
// stages on it. The first (lowest) stages use contiguous tiles (c = 0); higher stages use
// strided tiles made of 2^c adjacent sub-columns so that global loads stay coalesced.
#include <stdexcept>
#include "butterfly.h"
#include "field.h"
#include "kernels.h"

namespace sh {

// Doubled circle twiddle of pair h (line layer 0 pairs [x, y] -> [y, -y, -x, x]).
__device__ __forceinline__ uint32_t circle_twiddle(const uint32_t* __restrict__ lt0, uint32_t h) {
  const uint32_t x = lt0[(h >> 2) * 2], y = lt0[(h >> 2) * 2 + 1];
  switch (h & 3) {
    case 0: return y;
    case 1: return neg_tw(y);
    case 2: return neg_tw(x);
    default: return x;
  }
}

// Doubled twiddle of the pair whose lower element has global index idx0 at stage s (see top of
// file); tree is a doubled tree.
__device__ __forceinline__ uint32_t stage_twiddle(const uint32_t* __restrict__ tree, uint32_t tree_len,
                                                  uint32_t log_n, uint32_t s, uint32_t idx0) {
  if (s == 0) return circle_twiddle(tree + tree_len - (1u << (log_n - 1)), idx0 >> 1);
  return tree[tree_len - (1u << (log_n - s)) + (idx0 >> (s + 1))];
}

// One round of R consecutive stages (local bits j0 .. j0+R-1 of t) on 2^R elements held in
// registers: 2^(R-1) * R butterflies, one twiddle per distinct butterfly block.
template <bool INV, int R>
__device__ __forceinline__ void radix_round(uint32_t* lds, uint32_t C, uint32_t c, uint32_t e, uint32_t j0,
                                            uint32_t base, uint32_t tile_n, const uint32_t* __restrict__ tree,
                                            uint32_t tree_len, uint32_t log_n) {
  constexpr uint32_t G = 1u << R;
  const uint32_t n_groups = tile_n >> R;
  for (uint32_t g = threadIdx.x; g < n_groups; g += blockDim.x) {
    uint32_t lo = g & (C - 1), r = g >> c;
    // t with R zero bits inserted at j0
    uint32_t t0 = ((r >> j0) << (j0 + R)) | (r & ((1u << j0) - 1));
    uint32_t v[G];
#pragma unroll
    for (uint32_t q = 0; q < G; ++q) v[q] = lds[(t0 | (q << j0)) * C + lo];
    const uint32_t idx_base = base | (t0 << e) | lo;
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
      const int k = INV ? kk : R - 1 - kk;  // bit within the group
      const uint32_t s = e + j0 + k;
      // the twiddle depends only on the group bits above k: 2^(R-1-k) distinct values
#pragma unroll
      for (uint32_t u = 0; u < (G >> (k + 1)); ++u) {
        uint32_t tw = stage_twiddle(tree, tree_len, log_n, s, idx_base | ((u << (k + 1)) << (j0 + e)));
#pragma unroll
        for (uint32_t l = 0; l < (1u << k); ++l) {
          uint32_t q = (u << (k + 1)) | l;
          bfly<INV>(v[q], v[q | (1u << k)], tw);
        }
      }
    }
#pragma unroll
    for (uint32_t q = 0; q < G; ++q) lds[(t0 | (q << j0)) * C + lo] = v[q];
  }
}

template <bool INV>
__global__ void __launch_bounds__(256) fft_pass_kernel(const uint32_t* const* __restrict__ src,
                                                       uint32_t* const* __restrict__ dst, uint32_t log_n,
                                                       uint32_t in_mask, uint32_t e, uint32_t S, uint32_t c,
                                                       const uint32_t* __restrict__ tree, uint32_t tree_len,
                                                       uint32_t scale, uint32_t prio) {
  if (prio) CRITICAL_PATH_WAVE();
  extern __shared__ uint32_t lds[];
  const uint32_t col = blockIdx.y;
  const GLOBAL_AS uint32_t* __restrict__ in = gmem(src[col]);
  GLOBAL_AS uint32_t* __restrict__ out = gmem(dst[col]);
  const uint32_t C = 1u << c;
  const uint32_t tile = blockIdx.x;
  const uint32_t lo_blk = tile & ((1u << (e - c)) - 1);
  const uint32_t hi = tile >> (e - c);
  const uint32_t base = (hi << (e + S)) | (lo_blk << c);
  const uint32_t tile_n = 1u << (S + c);
  for (uint32_t l = threadIdx.x; l < tile_n; l += blockDim.x) {
    uint32_t t = l >> c, lo = l & (C - 1);
    lds[l] = in[(base | (t << e) | lo) & in_mask];
  }
  __syncthreads();
  // stages in rounds of up to 3: ascending for the inverse transform, descending for the forward
  uint32_t done = 0;
  while (done < S) {
    uint32_t R = S - done < 3 ? S - done : 3;
    uint32_t j0 = INV ? done : S - done - R;
    if (R == 3) radix_round<INV, 3>(lds, C, c, e, j0, base, tile_n, tree, tree_len, log_n);
    else if (R == 2) radix_round<INV, 2>(lds, C, c, e, j0, base, tile_n, tree, tree_len, log_n);
    else radix_round<INV, 1>(lds, C, c, e, j0, base, tile_n, tree, tree_len, log_n);
    done += R;
    __syncthreads();
  }
  for (uint32_t l = threadIdx.x; l < tile_n; l += blockDim.x) {
    uint32_t t = l >> c, lo = l & (C - 1);
    uint32_t v = lds[l];
    out[base | (t << e) | lo] = scale ? mul_tw(v, scale) : v;
  }
}

// ---- high stages in registers ------------------------------------------------------------------
// Stages [e, e + S) of a transform whose low e >= 11 stages are done: a lane owns the 2^S values
// at hi 2^(e+S) + t 2^e + lo (t < 2^S) and runs all S stages on them in registers. The 64 lanes of
// a wave take 64 consecutive lo, so each of a lane's loads and stores is one 256-byte segment of the
// wave, and every twiddle (index (hi 2^(S-1-j)) + (t >> (j + 1)) at stage e + j) is the same for the
// whole wave: a scalar load. No LDS and no barrier.
template <bool INV, int S>
__global__ void __launch_bounds__(256) fft_pass_reg_kernel(const uint32_t* const* __restrict__ src,
                                                           uint32_t* const* __restrict__ dst, uint32_t log_n,
                                                           uint32_t in_mask, uint32_t e,
                                                           const uint32_t* __restrict__ tree, uint32_t tree_len,
                                                           uint32_t scale, uint32_t prio) {
  if (prio) CRITICAL_PATH_WAVE();
  const uint32_t col = blockIdx.y;
  const GLOBAL_AS uint32_t* __restrict__ in = gmem(src[col]);
  GLOBAL_AS uint32_t* __restrict__ out = gmem(dst[col]);
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lo = g & ((1u << e) - 1);
  const uint32_t hi = __builtin_amdgcn_readfirstlane(g >> e);  // e >= 6: one value per wave
  const uint32_t base = (hi << (e + S)) | lo;
  uint32_t v[1 << S];
#pragma unroll
  for (uint32_t t = 0; t < (1u << S); ++t) v[t] = in[(base | (t << e)) & in_mask];
#pragma unroll
  for (int jj = 0; jj < S; ++jj) {
    const int j = INV ? jj : S - 1 - jj;  // the inverse ascends the stages, the forward descends
    const uint32_t* __restrict__ row = tree + (tree_len - (1u << (log_n - e - j)));
#pragma unroll
    for (uint32_t u = 0; u < (1u << (S - 1 - j)); ++u) {
      const uint32_t tw = row[(hi << (S - 1 - j)) | u];
#pragma unroll
      for (uint32_t l = 0; l < (1u << j); ++l) {
        const uint32_t q = (u << (j + 1)) | l;
        bfly<INV>(v[q], v[q | (1u << j)], tw);
      }
    }
  }
#pragma unroll
  for (uint32_t t = 0; t < (1u << S); ++t) {
    const uint32_t x = v[t];
    out[base | (t << e)] = scale ? mul_tw(x, scale) : x;
  }
}
constexpr uint32_t PASS_REG_MAX = 6;  // stages per register pass (2^S values per lane)

// ---- low stages of a 2^S-element block in one pass (11 <= S <= 15) ---------------------------
// A workgroup of 2^(S-5) threads owns one block (tile) of 2^S consecutive elements of one column
// and runs stages [0, S) on it as three rounds (bits 0-4, 5-9, 10..S-1); in a round each thread
// holds 32 elements in registers and runs up to 5 stages on them. The first round reads its
// elements straight from global memory and the last round writes its results straight back; only
// the hand-overs between rounds go through LDS, where word e lives at e + e/32 so that every
// round's accesses are free of bank conflicts.
//
// Workgroups go to the 8 XCDs in turn (workgroup n to XCD n mod 8). Renumbering them so that each
// XCD takes a contiguous range of tiles keeps the tiles of one column on one XCD and its L2: when
// evaluation onto a larger domain makes every tile of a column read the same coefficients, only
// the first read reaches HBM.
__device__ __forceinline__ uint32_t lds_pad(uint32_t e) { return e + (e >> 5); }

// Threads per workgroup and element registers per thread for a tile of 2^S elements.
template <int S> struct LowShape {
  static constexpr int X = 0;                       // 2^X groups of 32 elements per thread
  static constexpr uint32_t T = 1u << (S - 5 - X);
  static constexpr uint32_t V = 32u << X;
};

// Element (within the tile) held in register slot g * 2^R + q by thread t in the round at bit A.
template <int S, int A, int R>
__device__ __forceinline__ uint32_t round_elem(uint32_t t, uint32_t g, uint32_t q) {
  constexpr uint32_t T = LowShape<S>::T;
  const uint32_t rho = t + g * T;
  return ((rho >> A) << (A + R)) | (rho & ((1u << A) - 1)) | (q << A);
}

// With one group per lane and A >= 5, slot q's element is the lane's first element plus q 2^A
// (no carries into the lane's bits), and its padded LDS word is the first one's plus
// q (2^A + 2^(A-5)): one base address per lane and immediate offsets, instead of three
// instructions of index arithmetic per slot.
template <int S, int A, int R>
constexpr bool LDS_AFFINE = A >= 5 && ((1u << (5 - R)) << LowShape<S>::X) == 1;
template <int S, int A, int R>
__device__ __forceinline__ void round_from_lds(uint32_t* v, const uint32_t* lds) {
  constexpr uint32_t G = (1u << (5 - R)) << LowShape<S>::X, Q = 1u << R;
  if constexpr (LDS_AFFINE<S, A, R>) {
    const uint32_t* b = lds + lds_pad(round_elem<S, A, R>(threadIdx.x, 0, 0));
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q) v[q] = b[q * ((1u << A) + (1u << (A - 5)))];
    return;
  }
#pragma unroll
  for (uint32_t g = 0; g < G; ++g)
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q) v[g * Q + q] = lds[lds_pad(round_elem<S, A, R>(threadIdx.x, g, q))];
}
template <int S, int A, int R>
__device__ __forceinline__ void round_to_lds(const uint32_t* v, uint32_t* lds) {
  constexpr uint32_t G = (1u << (5 - R)) << LowShape<S>::X, Q = 1u << R;
  if constexpr (LDS_AFFINE<S, A, R>) {
    uint32_t* b = lds + lds_pad(round_elem<S, A, R>(threadIdx.x, 0, 0));
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q) b[q * ((1u << A) + (1u << (A - 5)))] = v[q];
    return;
  }
#pragma unroll
  for (uint32_t g = 0; g < G; ++g)
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q) lds[lds_pad(round_elem<S, A, R>(threadIdx.x, g, q))] = v[g * Q + q];
}
// Global elements gbase | e (index masked with in_mask) of the round's slots. At A = 0 a thread's
// 32 elements are consecutive: 16-byte loads when the column allows them.
// Rounds whose lanes own the 2^A low bits of the tile index (A = 10 with 1024 lanes): element q
// of lane t is the tile's word t + q 2^A, so a buffer access takes the lane offset in a vector
// register and the element offset in a scalar one, with no vector address arithmetic.
template <int S, int A> constexpr bool LANE_LOW_BITS = LowShape<S>::X == 0 && LowShape<S>::T == (1u << A);
__device__ __forceinline__ __amdgpu_buffer_rsrc_t column_rsrc(const void* p, uint32_t words) {
  // raw buffer over the column (stride 0, 32-bit data format); accesses past `words` read 0
  const uint32_t bytes = words >= (1u << 30) ? 0xffffffffu : words * 4u;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

template <int S, int A, int R>
__device__ __forceinline__ void round_from_global(uint32_t* v, const GLOBAL_AS uint32_t* __restrict__ in, uint32_t gbase,
                                                  uint32_t in_mask, bool vec) {
  constexpr uint32_t G = (1u << (5 - R)) << LowShape<S>::X, Q = 1u << R;
  if constexpr (LANE_LOW_BITS<S, A>) {
    // (gbase | t | q 2^A) & in_mask with in_mask >= 2^S - 1: the tile's base is gbase & in_mask
    const __amdgpu_buffer_rsrc_t rs = column_rsrc((const void*)in, in_mask + 1);
    const uint32_t base = __builtin_amdgcn_readfirstlane(gbase & in_mask);
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q)
      v[q] = __builtin_amdgcn_raw_buffer_load_b32(rs, threadIdx.x * 4u, (base + (q << A)) * 4u, 0);
    return;
  }
  if (A == 0 && R == 5 && vec) {
#pragma unroll
    for (uint32_t g = 0; g < G; ++g) {
      const uint32_t e0 = 32 * (threadIdx.x + g * LowShape<S>::T);
#pragma unroll
      for (uint32_t m = 0; m < 8; ++m) {
        const u32x4 w = *reinterpret_cast<const GLOBAL_AS u32x4*>(in + ((gbase | (e0 + 4 * m)) & in_mask));
        v[32 * g + 4 * m] = w.x; v[32 * g + 4 * m + 1] = w.y; v[32 * g + 4 * m + 2] = w.z; v[32 * g + 4 * m + 3] = w.w;
      }
    }
    return;
  }
#pragma unroll
  for (uint32_t g = 0; g < G; ++g)
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q) v[g * Q + q] = in[(gbase | round_elem<S, A, R>(threadIdx.x, g, q)) & in_mask];
}
// ALWAYS_VEC: the caller guarantees 16-byte aligned columns, and the round at bit 0 stores with
// 16-byte writes on one code path.
template <int S, int A, int R, bool ALWAYS_VEC = false>
__device__ __forceinline__ void round_to_global(uint32_t* v, GLOBAL_AS uint32_t* __restrict__ out, uint32_t gbase,
                                                uint32_t scale, bool vec, uint32_t out_words) {
  constexpr uint32_t G = (1u << (5 - R)) << LowShape<S>::X, Q = 1u << R;
  if (scale) {
#pragma unroll
    for (uint32_t i = 0; i < LowShape<S>::V; ++i) v[i] = mul_tw(v[i], scale);
  }
  if constexpr (LANE_LOW_BITS<S, A>) {
    const __amdgpu_buffer_rsrc_t rs = column_rsrc((const void*)out, out_words);
    const uint32_t base = __builtin_amdgcn_readfirstlane(gbase);
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q)
      __builtin_amdgcn_raw_buffer_store_b32(v[q], rs, threadIdx.x * 4u, (base + (q << A)) * 4u, 0);
    return;
  }
  if (A == 0 && R == 5 && (ALWAYS_VEC || vec)) {
#pragma unroll
    for (uint32_t g = 0; g < G; ++g) {
      const uint32_t e0 = 32 * (threadIdx.x + g * LowShape<S>::T);
#pragma unroll
      for (uint32_t m = 0; m < 8; ++m)
        *reinterpret_cast<GLOBAL_AS u32x4*>(out + (gbase | (e0 + 4 * m))) =
            u32x4{v[32 * g + 4 * m], v[32 * g + 4 * m + 1], v[32 * g + 4 * m + 2], v[32 * g + 4 * m + 3]};
    }
    return;
  }
#pragma unroll
  for (uint32_t g = 0; g < G; ++g)
#pragma unroll
    for (uint32_t q = 0; q < Q; ++q) out[gbase | round_elem<S, A, R>(threadIdx.x, g, q)] = v[g * Q + q];
}

// The R stages of the round at bit A on the registers (2^(R-1) * R butterflies per group, one
// twiddle per distinct butterfly block).
template <bool INV, int S, int A, int R>
__device__ __forceinline__ void round_butterflies(uint32_t* v, uint32_t gbase, const uint32_t* __restrict__ tree,
                                                  uint32_t tree_len, uint32_t log_n) {
  constexpr uint32_t T = LowShape<S>::T, G = (1u << (5 - R)) << LowShape<S>::X, Q = 1u << R;
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (uint32_t g = 0; g < G; ++g) {
    const uint32_t rho = t + g * T;
    const uint32_t idx_base = gbase | ((rho >> A) << (A + R)) | (rho & ((1u << A) - 1));
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
      const int k = INV ? kk : R - 1 - kk;
      const uint32_t s = A + k;
      if (A == 0 && k == 0) {
        // stage 0: pair h = (idx_base >> 1) | u, whose quarter u & 3 is known here, so the negated
        // circle twiddles become swapped additions
        // pair h = (idx_base >> 1) + u (idx_base's low 5 bits are zero): its [x, y] twiddle pair is
        // at 2 (h >> 2) = 2 (idx_base >> 3) + 2 (u >> 2)
        const uint32_t* lb = tree + tree_len - (1u << (log_n - 1)) + 2 * (idx_base >> 3);
#pragma unroll
        for (uint32_t u = 0; u < (Q >> 1); u += 4) {
          const uint32_t x = lb[2 * (u >> 2)], y = lb[2 * (u >> 2) + 1];
          bfly<INV, false>(v[g * Q + 2 * u], v[g * Q + 2 * u + 1], y);
          bfly<INV, true>(v[g * Q + 2 * u + 2], v[g * Q + 2 * u + 3], y);
          bfly<INV, true>(v[g * Q + 2 * u + 4], v[g * Q + 2 * u + 5], x);
          bfly<INV, false>(v[g * Q + 2 * u + 6], v[g * Q + 2 * u + 7], x);
        }
        continue;
      }
      // stage s >= 1: the twiddle of block u is tree[len - 2^(log_n - s) + (idx_base >> (s + 1)) + u]
      // (idx_base is zero on the round's bits, so u adds without carries): one base per stage
      // (when the lanes' bits all lie below bit A, the base is the same for the whole workgroup and
      // the twiddles come through the scalar cache)
      constexpr bool UNIFORM_TW = T * G <= (1u << A);
      // The round at bit 5 with one group per lane: the lanes' bits above the round are t >> 5, two
      // values in a wave (its halves), so both halves' twiddles load through the scalar cache (the
      // upper half's 2^(R-1-k) entries further on) and each lane selects its own. No vector load
      // waits here on older loads and stores still in flight.
      constexpr bool HALF_TW = !UNIFORM_TW && A == 5 && G == 1;
      const uint32_t toff = (UNIFORM_TW || HALF_TW) ? __builtin_amdgcn_readfirstlane(idx_base >> (s + 1))
                                                    : idx_base >> (s + 1);
      const uint32_t* tb = tree + tree_len - (1u << (log_n - s)) + toff;
      const bool upper = (threadIdx.x & 32) != 0;
#pragma unroll
      for (uint32_t u = 0; u < (Q >> (k + 1)); ++u) {
        uint32_t tw;
        if constexpr (HALF_TW) {
          // both values made wave-uniform, so that the select is not folded into one per-lane load
          const uint32_t lo = __builtin_amdgcn_readfirstlane(tb[u]);
          const uint32_t hi = __builtin_amdgcn_readfirstlane(tb[u + (1u << (R - 1 - k))]);
          tw = upper ? hi : lo;
        } else {
          tw = tb[u];
        }
#pragma unroll
        for (uint32_t l = 0; l < (1u << k); ++l) {
          const uint32_t q = (u << (k + 1)) | l;
          bfly<INV>(v[g * Q + q], v[g * Q + (q | (1u << k))], tw);
        }
      }
    }
  }
}

template <bool INV, int S>
__global__ void __launch_bounds__(LowShape<S>::T) fft_low_kernel(const uint32_t* const* __restrict__ src,
                                                                 uint32_t* const* __restrict__ dst, uint32_t log_n,
                                                                 uint32_t in_mask, const uint32_t* __restrict__ tree,
                                                                 uint32_t tree_len, uint32_t scale, uint32_t prio) {
  if (prio) CRITICAL_PATH_WAVE();
  constexpr uint32_t N = 1u << S;
  // rounds in execution order: the inverse runs bits 0-4, 5-9, 10..S-1; the forward the reverse
  constexpr int A1 = INV ? 0 : 10, R1 = INV ? 5 : S - 10;
  constexpr int A3 = INV ? 10 : 0, R3 = INV ? S - 10 : 5;
  __shared__ uint32_t lds[N + (N >> 5)];
  const uint32_t n_wg = gridDim.x, hw = blockIdx.x;
  const uint32_t tile = (n_wg & 7) == 0 ? (hw & 7) * (n_wg >> 3) + (hw >> 3) : hw;
  const uint32_t blk_log = log_n - S;
  const uint32_t col = tile >> blk_log, gbase = (tile & ((1u << blk_log) - 1)) << S;
  const GLOBAL_AS uint32_t* __restrict__ in = gmem(src[col]);
  GLOBAL_AS uint32_t* __restrict__ out = gmem(dst[col]);
  uint32_t v[LowShape<S>::V];
  round_from_global<S, A1, R1>(v, in, gbase, in_mask, (reinterpret_cast<uintptr_t>(in) & 15) == 0);
  round_butterflies<INV, S, A1, R1>(v, gbase, tree, tree_len, log_n);
  round_to_lds<S, A1, R1>(v, lds);
  __syncthreads();
  round_from_lds<S, 5, 5>(v, lds);
  round_butterflies<INV, S, 5, 5>(v, gbase, tree, tree_len, log_n);
  round_to_lds<S, 5, 5>(v, lds);
  __syncthreads();
  round_from_lds<S, A3, R3>(v, lds);
  round_butterflies<INV, S, A3, R3>(v, gbase, tree, tree_len, log_n);
  round_to_global<S, A3, R3>(v, out, gbase, scale, (reinterpret_cast<uintptr_t>(out) & 15) == 0, 1u << log_n);
}

// The per-lane twiddles of the round at bit 0 (R = 5, one group per lane), loaded ahead of the
// butterflies: stage 0's circle pairs (x, y) for its four quarters, then the line twiddles of
// stages 1..4 (8 + 4 + 2 + 1, stage k from 16 - 2^(5-k)).
struct Round0Tw {
  uint32_t cx[4], cy[4], t[15];
};
template <int S>
__device__ __forceinline__ void load_round0_tw(Round0Tw& tw, uint32_t gbase, const uint32_t* __restrict__ tree,
                                               uint32_t tree_len, uint32_t log_n) {
  static_assert(LowShape<S>::X == 0, "one group per lane");
  const uint32_t idx_base = gbase | (threadIdx.x << 5);
  const uint32_t* lb = tree + tree_len - (1u << (log_n - 1)) + 2 * (idx_base >> 3);
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) {
    tw.cx[j] = lb[2 * j];
    tw.cy[j] = lb[2 * j + 1];
  }
#pragma unroll
  for (uint32_t k = 1; k < 5; ++k) {
    const uint32_t* tb = tree + tree_len - (1u << (log_n - k)) + (idx_base >> (k + 1));
#pragma unroll
    for (uint32_t u = 0; u < (16u >> k); ++u) tw.t[16 - (32u >> k) + u] = tb[u];
  }
}
// round_butterflies<INV, S, 0, 5> with the twiddles of load_round0_tw.
template <bool INV>
__device__ __forceinline__ void round0_butterflies_pre(uint32_t* v, const Round0Tw& tw) {
#pragma unroll
  for (int kk = 0; kk < 5; ++kk) {
    const int k = INV ? kk : 4 - kk;
    if (k == 0) {
#pragma unroll
      for (uint32_t u = 0; u < 16; u += 4) {
        const uint32_t x = tw.cx[u >> 2], y = tw.cy[u >> 2];
        bfly<INV, false>(v[2 * u], v[2 * u + 1], y);
        bfly<INV, true>(v[2 * u + 2], v[2 * u + 3], y);
        bfly<INV, true>(v[2 * u + 4], v[2 * u + 5], x);
        bfly<INV, false>(v[2 * u + 6], v[2 * u + 7], x);
      }
      continue;
    }
#pragma unroll
    for (uint32_t u = 0; u < (16u >> k); ++u) {
      const uint32_t t = tw.t[16 - (32u >> k) + u];
#pragma unroll
      for (uint32_t l = 0; l < (1u << k); ++l) {
        const uint32_t q = (u << (k + 1)) | l;
        bfly<INV>(v[q], v[q | (1u << k)], t);
      }
    }
  }
}

// The low stages with each workgroup looping over tiles of 2^15 values (output columns 16-byte
// aligned): one workgroup per CU, whose LDS holds one such tile and no second. As soon as a tile's first round has gone to
// LDS, the workgroup issues the first round's loads of its next tile, and computes and stores the
// second and third rounds while they are in flight. Vector memory operations complete in issue
// order, so every vector load a tile still waits for is issued before the prefetch: the rounds at
// bits 5 and 10 take their twiddles through the scalar cache, and the per-lane twiddles of the
// round at bit 0 are loaded ahead (the forward transform's last round: with the tile's own
// prefetch; the inverse transform's first round: together with the next tile's values). The only
// wait on a prefetch is then the next tile's first round.
// The tiles are split into 8 contiguous ranges, one per XCD (workgroup n runs on XCD n mod 8), and
// the workgroups of an XCD take its tiles in turn: tiles of one column run side by side on one XCD
// and share its L2. Every branch on `more` is uniform over the workgroup, so all waves meet the
// same barriers and leave the loop together.
// The looping kernel's first round (peeled off its loop: inside the loop, the round's waits for
// the prefetch then have one path, after the previous tile's stores, and leave those in flight).
template <bool INV, int S>
__device__ __forceinline__ void loop_first_round(uint32_t* v, const Round0Tw& tw, uint32_t* lds, uint32_t gbase,
                                                 const uint32_t* __restrict__ tree, uint32_t tree_len, uint32_t log_n) {
  constexpr int A1 = INV ? 0 : 10, R1 = INV ? 5 : S - 10;
  if constexpr (INV) round0_butterflies_pre<true>(v, tw);
  else round_butterflies<false, S, A1, R1>(v, gbase, tree, tree_len, log_n);
  round_to_lds<S, A1, R1>(v, lds);
}

// 4 waves per SIMD: one 2^15 tile (1024 threads) per CU.
template <bool INV, int S>
__global__ void __launch_bounds__(LowShape<S>::T, 4) fft_low_loop_kernel(const uint32_t* const* __restrict__ src,
                                                                      uint32_t* const* __restrict__ dst, uint32_t log_n,
                                                                      uint32_t in_mask, const uint32_t* __restrict__ tree,
                                                                      uint32_t tree_len, uint32_t scale, uint32_t prio,
                                                                      uint32_t n_tiles) {
  if (prio) CRITICAL_PATH_WAVE();
  constexpr uint32_t N = 1u << S;
  // rounds in execution order: the inverse runs bits 0-4, 5-9, 10..S-1; the forward the reverse
  constexpr int A1 = INV ? 0 : 10, R1 = INV ? 5 : S - 10;
  constexpr int A3 = INV ? 10 : 0, R3 = INV ? S - 10 : 5;
  __shared__ uint32_t lds[N + (N >> 5)];
  const uint32_t per_xcd = gridDim.x >> 3, xcd = blockIdx.x & 7;
  const uint32_t t_end = (uint32_t)(((uint64_t)n_tiles * (xcd + 1)) >> 3);
  uint32_t tile = (uint32_t)(((uint64_t)n_tiles * xcd) >> 3) + (blockIdx.x >> 3);
  if (tile >= t_end) return;
  const uint32_t blk_log = log_n - S, blk_mask = (1u << blk_log) - 1;
  uint32_t v[LowShape<S>::V], w[LowShape<S>::V];
  Round0Tw tw;  // the round at bit 0: the inverse's first (next tile's), the forward's last
  {
    const GLOBAL_AS uint32_t* __restrict__ in = gmem(src[tile >> blk_log]);
    round_from_global<S, A1, R1>(v, in, (tile & blk_mask) << S, in_mask, (reinterpret_cast<uintptr_t>(in) & 15) == 0);
    if constexpr (INV) load_round0_tw<S>(tw, (tile & blk_mask) << S, tree, tree_len, log_n);
    loop_first_round<INV, S>(v, tw, lds, (tile & blk_mask) << S, tree, tree_len, log_n);
  }
  for (;;) {
    const uint32_t gbase = (tile & blk_mask) << S;
    if constexpr (!INV) load_round0_tw<S>(tw, gbase, tree, tree_len, log_n);
    const uint32_t next = tile + per_xcd;
    const bool more = next < t_end;
    {
      // v is free: the next tile's first-round values (the last tile loads its own again, unused:
      // a branch here would make the compiler's waits below cover the prefetch)
      const uint32_t pf = more ? next : tile;
      const GLOBAL_AS uint32_t* __restrict__ nin = gmem(src[pf >> blk_log]);
      round_from_global<S, A1, R1>(v, nin, (pf & blk_mask) << S, in_mask, (reinterpret_cast<uintptr_t>(nin) & 15) == 0);
      if constexpr (INV) load_round0_tw<S>(tw, (pf & blk_mask) << S, tree, tree_len, log_n);
    }
    // the loads above stay here: the compiler may not sink them past this point
    asm volatile("" ::: "memory");
    __syncthreads();
    round_from_lds<S, 5, 5>(w, lds);
    round_butterflies<INV, S, 5, 5>(w, gbase, tree, tree_len, log_n);
    round_to_lds<S, 5, 5>(w, lds);
    __syncthreads();
    round_from_lds<S, A3, R3>(w, lds);
    if constexpr (INV) round_butterflies<true, S, A3, R3>(w, gbase, tree, tree_len, log_n);
    else round0_butterflies_pre<false>(w, tw);
    GLOBAL_AS uint32_t* __restrict__ out = gmem(dst[tile >> blk_log]);
    // one store path (the launch requires 16-byte aligned output columns): with two, the compiler
    // could not tell how many stores are in flight when the next tile's first round waits
    round_to_global<S, A3, R3, true>(w, out, gbase, scale, true, 1u << log_n);
    if (!more) break;
    __syncthreads();  // this tile's last LDS reads end before the next tile's first writes
    tile = next;
    loop_first_round<INV, S>(v, tw, lds, (tile & blk_mask) << S, tree, tree_len, log_n);
  }
}

template <bool INV>
static void launch_low_loop(uint32_t S, uint32_t n_cols, const uint32_t* const* src, uint32_t* const* dst,
                            uint32_t log_n, uint32_t in_mask, const uint32_t* tree, uint32_t tree_len, uint32_t scale,
                            uint32_t prio, hipStream_t stream) {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint32_t n_tiles = n_cols << (log_n - S);
  // one workgroup per CU (its LDS holds one 2^15 tile), a multiple of 8 (the XCD split), no more
  // than the tiles
  uint32_t n_wg = (uint32_t)cus / 8 * 8;
  if (n_wg < 8) n_wg = 8;
  while (n_wg > 8 && n_wg > n_tiles) n_wg -= 8;
  const dim3 grid(n_wg), block(1u << (S - 5));
  hipLaunchKernelGGL((fft_low_loop_kernel<INV, 15>), grid, block, 0, stream, src, dst, log_n, in_mask, tree, tree_len, scale,
                     prio, n_tiles);
}

template <bool INV>
static void launch_low(uint32_t S, uint32_t n_cols, const uint32_t* const* src, uint32_t* const* dst, uint32_t log_n,
                       uint32_t in_mask, const uint32_t* tree, uint32_t tree_len, uint32_t scale, uint32_t prio,
                       hipStream_t stream) {
  const dim3 grid(n_cols << (log_n - S)), block(1u << (S - 5));
  switch (S) {
    case 11: hipLaunchKernelGGL((fft_low_kernel<INV, 11>), grid, block, 0, stream, src, dst, log_n, in_mask, tree, tree_len, scale, prio); break;
    case 12: hipLaunchKernelGGL((fft_low_kernel<INV, 12>), grid, block, 0, stream, src, dst, log_n, in_mask, tree, tree_len, scale, prio); break;
    case 13: hipLaunchKernelGGL((fft_low_kernel<INV, 13>), grid, block, 0, stream, src, dst, log_n, in_mask, tree, tree_len, scale, prio); break;
    case 14: hipLaunchKernelGGL((fft_low_kernel<INV, 14>), grid, block, 0, stream, src, dst, log_n, in_mask, tree, tree_len, scale, prio); break;
    default: hipLaunchKernelGGL((fft_low_kernel<INV, 15>), grid, block, 0, stream, src, dst, log_n, in_mask, tree, tree_len, scale, prio); break;
  }
}

// Host-side pass planning shared by interpolate and evaluate.
struct FftPass { uint32_t e, S, c; };
static std::vector<FftPass> plan_passes(uint32_t k, uint32_t K) {
  // stages [0, k) split into a contiguous first chunk and strided higher chunks
  std::vector<FftPass> p;
  const uint32_t TILE_LOG = 12;
  uint32_t first = k < TILE_LOG ? k : TILE_LOG;
  p.push_back({0, first, 0});
  uint32_t e = first;
  while (e < k) {
    uint32_t S = k - e;
    if (S > 8) S = 8;
    uint32_t c = TILE_LOG - S;
    if (c > e) c = e;
    // the tile must fit in the column: S + c <= K
    p.push_back({e, S, c});
    e += S;
  }
  (void)K;
  return p;
}

void launch_fft(bool inverse, const uint32_t* const* src, uint32_t* const* dst, uint32_t n_cols, uint32_t log_in,
                uint32_t log_n, const uint32_t* tree, uint32_t tree_len, hipStream_t stream, bool background,
                uint32_t low_log) {
  if (n_cols == 0) return;
  const uint32_t prio = background ? 0u : 1u;
  const uint32_t k = inverse ? log_n : log_in;  // number of stages actually run
  const uint32_t inv_n = inverse ? 2 * m31_inv(1u << log_n) : 0;  // doubled, as the twiddles
  auto run_pass = [&](const FftPass& ps, const uint32_t* const* from, uint32_t mask, uint32_t scale) {
    if (ps.c == ~0u) {  // a register pass (after the low stages: e >= 11, 2^(log_n - S) >= 256 lanes)
      const dim3 grid(1u << (log_n - ps.S - 8), n_cols), block(256);
#define PASS_REG(SS)                                                                                            \
  case SS:                                                                                                      \
    if (inverse)                                                                                                \
      hipLaunchKernelGGL((fft_pass_reg_kernel<true, SS>), grid, block, 0, stream, from, dst, log_n, mask, ps.e, \
                         tree, tree_len, scale, prio);                                                          \
    else                                                                                                        \
      hipLaunchKernelGGL((fft_pass_reg_kernel<false, SS>), grid, block, 0, stream, from, dst, log_n, mask,      \
                         ps.e, tree, tree_len, scale, prio);                                                    \
    break;
      switch (ps.S) {
        PASS_REG(1) PASS_REG(2) PASS_REG(3) PASS_REG(4) PASS_REG(5) PASS_REG(6)
        default: throw std::invalid_argument("launch_fft: register pass of more than 6 stages");
      }
#undef PASS_REG
      return;
    }
    const uint32_t tiles = 1u << (log_n - ps.S - ps.c), tile_n = 1u << (ps.S + ps.c);
    const uint32_t threads = tile_n / 2 < 64 ? 64 : (tile_n / 2 > 256 ? 256 : tile_n / 2);
    if (inverse)
      hipLaunchKernelGGL(fft_pass_kernel<true>, dim3(tiles, n_cols), dim3(threads), tile_n * 4, stream, from, dst,
                         log_n, mask, ps.e, ps.S, ps.c, tree, tree_len, scale, prio);
    else
      hipLaunchKernelGGL(fft_pass_kernel<false>, dim3(tiles, n_cols), dim3(threads), tile_n * 4, stream, from, dst,
                         log_n, mask, ps.e, ps.S, ps.c, tree, tree_len, scale, prio);
  };
  const uint32_t full = (1u << log_n) - 1, in_mask = inverse ? full : (1u << log_in) - 1;
  if (k >= 11) {
    // stages [0, S) in one pass of 2^S-element blocks; stages [S, k) in strided passes
    const uint32_t smax = k < 15 ? k : 15;
    const bool loop = (low_log & FFT_LOW_LOOP) != 0;
    low_log &= ~FFT_LOW_LOOP;
    if (low_log != 0 && (low_log < 11 || low_log > smax)) throw std::invalid_argument("launch_fft: low_log out of range");
    const uint32_t S = low_log ? low_log : smax;
    if (loop && S != 15) throw std::invalid_argument("launch_fft: the looping low kernel needs low_log 15");
    // the stages above the low tile: register passes of at most PASS_REG_MAX stages, as even as
    // the count allows (8 stages: 4 + 4, not 6 + 2)
    std::vector<FftPass> high;
    {
      const uint32_t rest = k - S, n_pass = (rest + PASS_REG_MAX - 1) / PASS_REG_MAX;
      for (uint32_t i = 0, e = S; i < n_pass; ++i) {
        const uint32_t st = (rest - (e - S) + (n_pass - i) - 1) / (n_pass - i);
        high.push_back({e, st, ~0u});
        e += st;
      }
    }
    auto low = [&](bool inv, const uint32_t* const* from, uint32_t mask, uint32_t scale) {
      if (inv) (loop ? launch_low_loop<true> : launch_low<true>)(S, n_cols, from, dst, log_n, mask, tree, tree_len, scale, prio, stream);
      else (loop ? launch_low_loop<false> : launch_low<false>)(S, n_cols, from, dst, log_n, mask, tree, tree_len, scale, prio, stream);
    };
    if (inverse) {
      low(true, src, full, high.empty() ? inv_n : 0u);
      for (size_t i = 0; i < high.size(); ++i)
        run_pass(high[i], (const uint32_t* const*)dst, full, i + 1 == high.size() ? inv_n : 0u);
    } else {
      for (size_t ii = high.size(); ii-- > 0;)
        run_pass(high[ii], ii + 1 == high.size() ? src : (const uint32_t* const*)dst,
                 ii + 1 == high.size() ? in_mask : full, 0u);
      low(false, high.empty() ? src : (const uint32_t* const*)dst, high.empty() ? in_mask : full, 0u);
    }
    return;
  }
  auto passes = plan_passes(k, log_n);
  if (inverse) {
    for (size_t i = 0; i < passes.size(); ++i)
      run_pass(passes[i], i == 0 ? src : (const uint32_t* const*)dst, full, i + 1 == passes.size() ? inv_n : 0u);
  } else {
    for (size_t ii = passes.size(); ii-- > 0;) {
      const bool first = ii + 1 == passes.size();
      run_pass(passes[ii], first ? src : (const uint32_t* const*)dst, first ? in_mask : full, 0u);
    }
  }
}

}  // namespace sh
