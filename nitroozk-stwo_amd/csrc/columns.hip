// Column and domain utilities:
//   the twiddle tree of any coset      PolyOps::precompute_twiddles (prover/poly/circle/ops.rs:122;
//                                      layout prover/backend/cpu/circle.rs:274-291)
//   bit reversal of a column           ColumnOps::bit_reverse_column (prover/backend/mod.rs:39-42)
//   circle transforms of 2 and 4 words  interpolate / evaluate on domains of log size 1 and 2
//                                      (prover/backend/cpu/circle.rs:35-58,199-213)
#include "field.h"
#include "kernels.h"

namespace sh {

// Entry e of the tree of the coset (init, L): layer j (0 <= j < L) holds the x-coordinates of the
// first half of the coset doubled j times, bit-reversed, at [2^L - 2^(L-j), 2^L - 2^(L-j-1)); the
// last entry is the pad 1. Point k of the coset doubled j times has index
// (init + k 2^(31-L)) 2^j, and the point of an index is the sum of the generator's powers 2^b
// over its set bits. Inverses by Fermat; the doubled trees hold 2t. A zero x-coordinate (which
// has no inverse) raises *zero_flag.
__global__ void __launch_bounds__(256) twiddle_tree_kernel(uint32_t init, uint32_t L, GenPowers g, uint32_t* tw,
                                                           uint32_t* itw, uint32_t* tw2, uint32_t* itw2,
                                                           uint32_t* zero_flag) {
  const uint32_t len = 1u << L;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < len; e += gridDim.x * blockDim.x) {
    uint32_t x = 1, xi = 1;
    if (e + 1 < len) {
      const uint32_t rem = len - e;                        // in (2^(L-j-1), 2^(L-j)]
      const uint32_t lj = 31u - __builtin_clz(rem - 1);   // log size of layer j = L - 1 - j
      const uint32_t j = L - 1 - lj;
      const uint32_t i = e - (len - (2u << lj));
      uint32_t idx = ((init << j) + (bitrev(i, lj) << (31 - L + j))) & 0x7fffffffu;
      cpm31 p = {1, 0};
      for (uint32_t b = 0; idx; ++b, idx >>= 1)
        if (idx & 1) p = cp_add(p, g.p[b]);
      x = p.x;
      if (x == 0) atomicOr(zero_flag, 1u);
      xi = m31_inv(x);
    }
    tw[e] = x;
    itw[e] = xi;
    tw2[e] = 2 * x;
    itw2[e] = 2 * xi;
  }
}

void launch_twiddle_tree(uint32_t init_index, uint32_t log_size, uint32_t* tw, uint32_t* itw, uint32_t* tw2,
                         uint32_t* itw2, uint32_t* zero_flag, hipStream_t stream) {
  GenPowers g;
  cpm31 c = circle_gen();
  for (int b = 0; b < 31; ++b) { g.p[b] = c; c = cp_double(c); }
  const size_t len = (size_t)1 << log_size;
  const uint32_t blocks = (uint32_t)((len + 255) / 256 < 8192 ? (len + 255) / 256 : 8192);
  hipLaunchKernelGGL(twiddle_tree_kernel, dim3(blocks), dim3(256), 0, stream, init_index & 0x7fffffffu, log_size, g, tw,
                     itw, tw2, itw2, zero_flag);
}

// In-place bit reversal: the lane of the smaller index of each swapped pair moves both words.
__global__ void __launch_bounds__(256) bit_reverse_kernel(uint32_t* col, uint32_t log_n) {
  const uint32_t n = 1u << log_n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t r = bitrev(i, log_n);
    if (i < r) {
      const uint32_t a = col[i], b = col[r];
      col[i] = b;
      col[r] = a;
    }
  }
}

void launch_bit_reverse(uint32_t* col, uint32_t log_n, hipStream_t stream) {
  const size_t n = (size_t)1 << log_n;
  const uint32_t blocks = (uint32_t)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
  hipLaunchKernelGGL(bit_reverse_kernel, dim3(blocks), dim3(256), 0, stream, col, log_n);
}

// Transforms of 2 or 4 words per column, one lane per column. x and y are the half coset's first
// point (their inverses for the inverse transform); scale is 2^-log_n for the inverse. Inputs of
// the forward transform beyond 2^log_in words are zero (extend).
__device__ __forceinline__ void tiny_bfly(uint32_t& a, uint32_t& b, uint32_t t) {
  const uint32_t m = m31_mul(b, t);
  const uint32_t s = m31_add(a, m), d = m31_sub(a, m);
  a = s;
  b = d;
}
__device__ __forceinline__ void tiny_ibfly(uint32_t& a, uint32_t& b, uint32_t t) {
  const uint32_t s = m31_add(a, b), d = m31_mul(m31_sub(a, b), t);
  a = s;
  b = d;
}
__global__ void fft_tiny_kernel(const uint32_t* const* src, uint32_t* const* dst, uint32_t n_cols, uint32_t log_in,
                                uint32_t log_n, uint32_t inverse, uint32_t x, uint32_t y, uint32_t scale) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  const uint32_t n = 1u << log_n, m = 1u << log_in;
  uint32_t v[4] = {0, 0, 0, 0};
  for (uint32_t i = 0; i < m && i < n; ++i) v[i] = src[c][i];
  if (log_n == 1) {
    if (inverse) tiny_ibfly(v[0], v[1], y); else tiny_bfly(v[0], v[1], y);
  } else if (inverse) {
    tiny_ibfly(v[0], v[1], y);
    tiny_ibfly(v[2], v[3], m31_neg(y));
    tiny_ibfly(v[0], v[2], x);
    tiny_ibfly(v[1], v[3], x);
  } else {
    tiny_bfly(v[0], v[2], x);
    tiny_bfly(v[1], v[3], x);
    tiny_bfly(v[0], v[1], y);
    tiny_bfly(v[2], v[3], m31_neg(y));
  }
  for (uint32_t i = 0; i < n; ++i) dst[c][i] = inverse ? m31_mul(v[i], scale) : v[i];
}

void launch_fft_tiny(bool inverse, const uint32_t* const* src, uint32_t* const* dst, uint32_t n_cols, uint32_t log_in,
                     uint32_t log_n, cpm31 half_coset_initial, hipStream_t stream) {
  if (n_cols == 0) return;
  uint32_t x = half_coset_initial.x, y = half_coset_initial.y, scale = 1;
  if (inverse) {
    x = log_n == 2 ? m31_inv(x) : x;
    y = m31_inv(y);
    scale = m31_inv(1u << log_n);
  }
  hipLaunchKernelGGL(fft_tiny_kernel, dim3((n_cols + 63) / 64), dim3(64), 0, stream, src, dst, n_cols, log_in, log_n,
                     inverse ? 1u : 0u, x, y, scale);
}

}  // namespace sh
