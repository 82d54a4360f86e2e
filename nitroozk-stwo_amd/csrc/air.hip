// Constraint evaluation on the evaluation domain (ComponentProver::
// evaluate_constraint_quotients_on_domain) for the two benchmark AIRs, their trace generation,
// and the Poseidon LogUp interaction trace.
//
//   wide Fibonacci   crates/examples/src/wide_fibonacci/mod.rs:34-83
//   Poseidon2        crates/examples/src/poseidon/mod.rs:38-335 (16-wide state, 4+4 full rounds,
//                    14 partial rounds, x^5, all round constants 1234)
//   LogUp            crates/constraint-framework/src/lib.rs:149-258 (finalize_logup_in_pairs),
//                    prover/logup.rs:21-200 (LogupTraceGenerator), simd/prefix_sum.rs
//   domain driver    crates/constraint-framework/src/prover/component_prover.rs:32-169:
//                    row_res = sum_k coeff_k C_k(row), accumulated as row_res * denom_inv[row >> trace_log]
//
// One lane evaluates one evaluation-domain row; trace columns are column-major, so the 64 lanes
// of a wave read 256 contiguous bytes of each column. Constraint coefficients are wave-uniform
// and load through the scalar cache.
#include <stdexcept>
#include "field.h"
#include "kernels.h"

namespace sh {

__device__ __forceinline__ void acc_add_qm(qacc& acc, qm31 v) {
  acc.a += v.a; acc.b += v.b; acc.c += v.c; acc.d += v.d;
}

__global__ void __launch_bounds__(256) wide_fib_comp_kernel(uint32_t n_rows, uint32_t trace_log,
                                                            const uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                                            const qm31* __restrict__ pows,
                                                            const uint32_t* __restrict__ denom_inv, Cols4 acc_out,
                                                            uint32_t accumulate) {
  uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  uint32_t a = cols[0][row], b = cols[1][row];
  qacc acc = qacc_zero();
  for (uint32_t k = 2; k < n_cols; ++k) {
    uint32_t c = cols[k][row];
    uint32_t s = m31_reduce64((uint64_t)a * a + (uint64_t)b * b);
    qacc_mad(acc, pows[k - 2], m31_sub(c, s));
    if (k % 3 == 1) qacc_fold(acc);  // at most 3 unreduced products between folds
    a = b;
    b = c;
  }
  qm31 r = qm_mul_m31(qacc_reduce(acc), denom_inv[row >> trace_log]);
  if (accumulate) r = qm_add(qm31{acc_out.c[0][row], acc_out.c[1][row], acc_out.c[2][row], acc_out.c[3][row]}, r);
  acc_out.c[0][row] = r.a; acc_out.c[1][row] = r.b; acc_out.c[2][row] = r.c; acc_out.c[3][row] = r.d;
}

void launch_wide_fib_composition(uint32_t eval_log, uint32_t trace_log, const uint32_t* const* cols, uint32_t n_cols,
                                 const qm31* pows, const uint32_t* denom_inv, Cols4 acc, bool accumulate,
                                 hipStream_t stream) {
  uint32_t n = 1u << eval_log;
  hipLaunchKernelGGL(wide_fib_comp_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, trace_log, cols, n_cols,
                     pows, denom_inv, acc, (uint32_t)accumulate);
}

// Several small wide-Fibonacci components in one launch (the batched composition of a
// multi-component proof, ComponentProvers::compute_composition_polynomial's batch hook,
// crates/stwo/src/prover/air/component_prover.rs:118-141): components are grouped by evaluation
// size; a block covers 64 rows of one group and its WF_WAVES waves take the group's components in
// turn, their partial sums meeting in LDS. The sum over a group's components times the group's
// denominator equals the per-component accumulation, so the accumulator ends the same.
constexpr uint32_t WF_WAVES = 4;
__global__ void __launch_bounds__(256) wide_fib_comp_grouped_kernel(const WfGroup* __restrict__ groups,
                                                                    uint32_t n_groups,
                                                                    const WfComp* __restrict__ comps) {
  __shared__ uint32_t part[WF_WAVES][4][64];
  const uint32_t lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint32_t g = 0;  // the block's group: the last whose first block is at most this one (uniform)
  for (uint32_t k = 1; k < n_groups; ++k)
    if (groups[k].block_off <= blockIdx.x) g = k;
  const WfGroup& G = groups[g];
  const uint32_t row = (blockIdx.x - G.block_off) * 64 + lane;
  const bool active = row < (1u << G.eval_log);
  qacc acc = qacc_zero();
  if (active) {
    for (uint32_t j = w; j < G.n_comps; j += WF_WAVES) {
      const WfComp& C = comps[G.comp_off + j];
      uint32_t a = C.cols[0][row], b = C.cols[1][row];
      for (uint32_t k = 2; k < C.n_cols; ++k) {
        const uint32_t c = C.cols[k][row];
        const uint32_t s = m31_reduce64((uint64_t)a * a + (uint64_t)b * b);
        qacc_mad(acc, C.pows[k - 2], m31_sub(c, s));
        if (k % 3 == 1) qacc_fold(acc);
        a = b;
        b = c;
      }
      qacc_fold(acc);
    }
  }
  const qm31 r = qacc_reduce(acc);
  part[w][0][lane] = r.a;
  part[w][1][lane] = r.b;
  part[w][2][lane] = r.c;
  part[w][3][lane] = r.d;
  __syncthreads();
  if (w == 0 && active) {
    qm31 tot = qm_zero();
    for (uint32_t v = 0; v < WF_WAVES; ++v)
      tot = qm_add(tot, qm31{part[v][0][lane], part[v][1][lane], part[v][2][lane], part[v][3][lane]});
    tot = qm_mul_m31(tot, G.dinv[row >> G.trace_log]);
    if (G.accumulate) tot = qm_add(qm31{G.acc[0][row], G.acc[1][row], G.acc[2][row], G.acc[3][row]}, tot);
    G.acc[0][row] = tot.a;
    G.acc[1][row] = tot.b;
    G.acc[2][row] = tot.c;
    G.acc[3][row] = tot.d;
  }
}
void launch_wide_fib_composition_grouped(const WfGroup* groups, uint32_t n_groups, const WfComp* comps,
                                         uint32_t n_blocks, hipStream_t stream) {
  hipLaunchKernelGGL(wide_fib_comp_grouped_kernel, dim3(n_blocks), dim3(64 * WF_WAVES), 0, stream, groups, n_groups,
                     comps);
}

// ---- Poseidon2 -------------------------------------------------------------------------------
constexpr uint32_t PSD_COLS_PER_REP = 158;
constexpr uint32_t PSD_ROUND_CONST = 1234;

__device__ __forceinline__ uint32_t pow5(uint32_t x) {
  uint32_t x2 = m31_sqr(x), x4 = m31_sqr(x2);
  return m31_mul(x4, x);
}
__device__ __forceinline__ void apply_m4(uint32_t& x0, uint32_t& x1, uint32_t& x2, uint32_t& x3) {
  uint32_t t0 = m31_add(x0, x1), t02 = m31_add(t0, t0), t1 = m31_add(x2, x3), t12 = m31_add(t1, t1);
  uint32_t t2 = m31_add(m31_add(x1, x1), t1), t3 = m31_add(m31_add(x3, x3), t0);
  uint32_t t4 = m31_add(m31_add(t12, t12), t3), t5 = m31_add(m31_add(t02, t02), t2);
  uint32_t t6 = m31_add(t3, t5), t7 = m31_add(t2, t4);
  x0 = t6; x1 = t5; x2 = t7; x3 = t4;
}
__device__ __forceinline__ void external_matrix(uint32_t* s) {
#pragma unroll
  for (int i = 0; i < 4; ++i) apply_m4(s[4 * i], s[4 * i + 1], s[4 * i + 2], s[4 * i + 3]);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint32_t t = m31_add(m31_add(s[j], s[j + 4]), m31_add(s[j + 8], s[j + 12]));
#pragma unroll
    for (int i = 0; i < 4; ++i) s[4 * i + j] = m31_add(s[4 * i + j], t);
  }
}
__device__ __forceinline__ void internal_matrix(uint32_t* s) {
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) sum += s[i];
  uint32_t sm = m31_reduce64(sum);
#pragma unroll
  for (int i = 0; i < 16; ++i) s[i] = m31_add(m31_rotl(s[i], i + 1), sm);
}
// canonical state words and coefficients: products below 2^62, three between folds
__device__ __forceinline__ qm31 combine16(const PoseidonParams& p, const uint32_t* v) {
  qacc acc = qacc_zero();
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    qacc_mad(acc, p.alpha_pow[i], v[i]);
    if (i % 3 == 2) qacc_fold(acc);
  }
  return qm_sub(qacc_reduce(acc), p.z);
}
__device__ __forceinline__ uint32_t offset_prev_row(uint32_t row, uint32_t trace_log, uint32_t eval_log) {
  // offset_bit_reversed_circle_domain_index(row, trace_log, eval_log, -1) (core/utils.rs)
  uint32_t prev = bitrev(row, eval_log);
  uint32_t half = 1u << (eval_log - 1);
  uint32_t step = 1u << (eval_log - trace_log - 1);
  if (prev < half) prev = (prev + half - step) & (half - 1);
  else prev = ((prev - half + step) & (half - 1)) + half;
  return bitrev(prev, eval_log);
}

// Column access: a table of column addresses (any layout), or columns at a constant stride from
// a base (one allocation; no address loads in the kernel).
struct ColTable {
  const uint32_t* const* p;
  __device__ __forceinline__ const uint32_t* col(uint32_t c) const { return p[c]; }
};
struct ColStrided {
  const uint32_t* base;
  size_t stride;
  __device__ __forceinline__ const uint32_t* col(uint32_t c) const { return base + c * stride; }
};

// One Poseidon2 instance (columns c0 .. c0 + 157): rounds with their 142 constraints; s holds the
// initial state (columns 0..15) on entry and the final state on return. The column values of the
// next round are loaded while the current round is evaluated.
__device__ __forceinline__ void psd_full_round(uint32_t* s, const uint32_t* m, const qm31* __restrict__ pows,
                                               qacc& acc) {
#pragma unroll
  for (int i = 0; i < 16; ++i) s[i] = m31_add(s[i], PSD_ROUND_CONST);
  external_matrix(s);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    qacc_mad(acc, pows[i], m31_sub(pow5(s[i]), m[i]));
    s[i] = m[i];
    if (i % 3 == 2 || i == 15) qacc_fold(acc);
  }
}
template <class Cols>
__device__ __forceinline__ void poseidon_rounds_constraints(uint32_t* s, const Cols& cols, uint32_t c0, uint32_t row,
                                                            const qm31* __restrict__ pows, qacc& acc) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) m[i] = cols.col(c0 + 16 + i)[row];
  // full rounds 0..3 read columns 16 + 16 r ..; the last one prefetches the partial-round columns 80..93
#pragma unroll 1
  for (uint32_t r = 0; r < 4; ++r) {
    uint32_t nx[16];
    const uint32_t base = c0 + 32 + 16 * r;
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) nx[i] = (r < 3 || i < 14) ? cols.col(base + i)[row] : 0u;
    psd_full_round(s, m, pows + 16 * r, acc);
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = nx[i];
  }
  // 14 partial rounds (columns 80..93), prefetching columns 94..109
  {
    uint32_t nx[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) nx[i] = cols.col(c0 + 94 + i)[row];
#pragma unroll
    for (int r = 0; r < 14; ++r) {
      s[0] = m31_add(s[0], PSD_ROUND_CONST);
      internal_matrix(s);
      qacc_mad(acc, pows[64 + r], m31_sub(pow5(s[0]), m[r]));
      s[0] = m[r];
      if (r % 3 == 2 || r == 13) qacc_fold(acc);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = nx[i];
  }
  // full rounds 4..7 read columns 94 + 16 (r - 4) ..
#pragma unroll 1
  for (uint32_t r = 0; r < 4; ++r) {
    uint32_t nx[16];
    const uint32_t base = c0 + 110 + 16 * r;
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) nx[i] = r < 3 ? cols.col(base + i)[row] : 0u;
    psd_full_round(s, m, pows + 78 + 16 * r, acc);
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = nx[i];
  }
}

// A block covers 64 rows; wave w evaluates instances w, w + W, ... (W waves) of those rows, and
// the W partial sums meet in LDS. Each instance contributes its 142 round constraints and its
// LogUp constraint (finalize_logup_in_pairs: cumulative column rep minus column rep - 1, or for the
// last instance minus the previous row's last column, plus the claimed-sum shift).
constexpr uint32_t PSD_MAX_WAVES = 8;
template <class Cols>
__global__ void __launch_bounds__(512, 4) poseidon_comp_kernel(uint32_t eval_log, uint32_t trace_log, Cols trace,
                                                            Cols inter, const PoseidonParams* __restrict__ params,
                                                            const qm31* __restrict__ pows,
                                                            const uint32_t* __restrict__ denom_inv, Cols4 acc_out,
                                                            uint32_t accumulate) {
  __shared__ uint32_t part[PSD_MAX_WAVES][4][64];
  // the wave index is uniform across the wave: readfirstlane lets the compiler keep the instance
  // loop and the constraint coefficients in scalar registers
  const uint32_t lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), W = blockDim.x >> 6;
  const uint32_t row = blockIdx.x * 64 + lane;
  const bool active = row < (1u << eval_log);
  const PoseidonParams& p = *params;
  const uint32_t ipr = p.ipr;
  qacc acc = qacc_zero();
  if (active) {
    for (uint32_t rep = w; rep < ipr; rep += W) {
      const uint32_t c0 = rep * PSD_COLS_PER_REP;
      uint32_t s[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) s[i] = trace.col(c0 + i)[row];
      const qm31 d0 = combine16(p, s);
      poseidon_rounds_constraints(s, trace, c0, row, pows + rep * 142, acc);
      const qm31 d1 = combine16(p, s);
      // pair fraction 1/d0 - 1/d1 = (d1 - d0) / (d0 d1)
      const qm31 num = qm_sub(d1, d0), den = qm_mul(d0, d1);
      const uint32_t i0 = 4 * rep;
      const qm31 cur = {inter.col(i0)[row], inter.col(i0 + 1)[row], inter.col(i0 + 2)[row], inter.col(i0 + 3)[row]};
      qm31 prev_col = qm_zero();
      if (rep > 0)
        prev_col = {inter.col(i0 - 4)[row], inter.col(i0 - 3)[row], inter.col(i0 - 2)[row], inter.col(i0 - 1)[row]};
      qm31 diff;
      if (rep + 1 < ipr) {
        diff = qm_sub(cur, prev_col);
      } else {
        const uint32_t pr = offset_prev_row(row, trace_log, eval_log);
        const qm31 prow = {inter.col(i0)[pr], inter.col(i0 + 1)[pr], inter.col(i0 + 2)[pr], inter.col(i0 + 3)[pr]};
        diff = qm_add(qm_sub(qm_sub(cur, prow), prev_col), p.shift);
      }
      const qm31 cons = qm_sub(qm_mul(diff, den), num);
      acc_add_qm(acc, qm_mul(pows[ipr * 142 + rep], cons));
    }
  }
  const qm31 r = qacc_reduce(acc);
  part[w][0][lane] = r.a;
  part[w][1][lane] = r.b;
  part[w][2][lane] = r.c;
  part[w][3][lane] = r.d;
  __syncthreads();
  if (w == 0 && active) {
    qm31 tot = qm_zero();
    for (uint32_t v = 0; v < W; ++v)
      tot = qm_add(tot, qm31{part[v][0][lane], part[v][1][lane], part[v][2][lane], part[v][3][lane]});
    tot = qm_mul_m31(tot, denom_inv[row >> trace_log]);
    if (accumulate)
      tot = qm_add(qm31{acc_out.c[0][row], acc_out.c[1][row], acc_out.c[2][row], acc_out.c[3][row]}, tot);
    acc_out.c[0][row] = tot.a; acc_out.c[1][row] = tot.b; acc_out.c[2][row] = tot.c; acc_out.c[3][row] = tot.d;
  }
}

void launch_poseidon_composition(uint32_t eval_log, uint32_t trace_log, const uint32_t* const* trace_cols,
                                 const uint32_t* const* inter_cols, const PoseidonParams* params, uint32_t ipr,
                                 const qm31* pows, const uint32_t* denom_inv, Cols4 acc, bool accumulate,
                                 hipStream_t stream) {
  uint32_t n = 1u << eval_log;
  uint32_t waves = ipr < PSD_MAX_WAVES ? ipr : PSD_MAX_WAVES;
  hipLaunchKernelGGL(poseidon_comp_kernel<ColTable>, dim3((n + 63) / 64), dim3(64 * waves), 0, stream, eval_log,
                     trace_log, ColTable{trace_cols}, ColTable{inter_cols}, params, pows, denom_inv, acc,
                     (uint32_t)accumulate);
}
void launch_poseidon_composition_strided(uint32_t eval_log, uint32_t trace_log, const uint32_t* trace_base,
                                         const uint32_t* inter_base, size_t stride, const PoseidonParams* params,
                                         uint32_t ipr, const qm31* pows, const uint32_t* denom_inv, Cols4 acc,
                                         bool accumulate, hipStream_t stream) {
  uint32_t n = 1u << eval_log;
  uint32_t waves = ipr < PSD_MAX_WAVES ? ipr : PSD_MAX_WAVES;
  hipLaunchKernelGGL(poseidon_comp_kernel<ColStrided>, dim3((n + 63) / 64), dim3(64 * waves), 0, stream, eval_log,
                     trace_log, ColStrided{trace_base, stride}, ColStrided{inter_base, stride}, params, pows,
                     denom_inv, acc, (uint32_t)accumulate);
}

// ---- trace generation -----------------------------------------------------------------------
// Seeded inputs: element idx of stream `seed` is the first of splitmix64 outputs
// (seed + (4 idx + a + 1) * golden), a = 0..3, whose top 31 bits are below P.
__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint32_t uniform_m31(uint64_t seed, uint64_t idx) {
  for (uint64_t a = 0; a < 4; ++a) {
    uint32_t v = (uint32_t)(splitmix64_at(seed, idx * 4 + a) >> 33);
    if (v < P) return v;
  }
  return 0;
}

__global__ void gen_wide_fib_kernel(uint32_t n_rows, uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                    uint32_t mode, uint64_t seed) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  uint32_t a = mode == 0 ? 1u : uniform_m31(seed, 2ull * r);
  uint32_t b = mode == 0 ? (r % P) : uniform_m31(seed, 2ull * r + 1);
  cols[0][r] = a;
  cols[1][r] = b;
  for (uint32_t c = 2; c < n_cols; ++c) {
    uint32_t x = m31_reduce64((uint64_t)a * a + (uint64_t)b * b);
    cols[c][r] = x;
    a = b;
    b = x;
  }
}
void launch_gen_wide_fib(uint32_t log_n, uint32_t* const* cols, uint32_t n_cols, uint32_t mode, uint64_t seed,
                         hipStream_t stream) {
  uint32_t n = 1u << log_n;
  hipLaunchKernelGGL(gen_wide_fib_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, cols, n_cols, mode, seed);
}

__global__ void gen_poseidon_kernel(uint32_t n_rows, uint32_t* const* __restrict__ cols, uint32_t ipr, uint32_t mode,
                                    uint64_t seed) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  for (uint32_t rep = 0; rep < ipr; ++rep) {
    uint32_t* const* c = cols + rep * PSD_COLS_PER_REP;
    uint32_t s[16];
    for (uint32_t i = 0; i < 16; ++i)
      s[i] = mode == 0 ? (r + i + rep) % P : uniform_m31(seed, ((uint64_t)r * ipr + rep) * 16 + i);
    uint32_t ci = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) c[ci++][r] = s[i];
    for (int rd = 0; rd < 4; ++rd) {
#pragma unroll
      for (int i = 0; i < 16; ++i) s[i] = m31_add(s[i], PSD_ROUND_CONST);
      external_matrix(s);
#pragma unroll
      for (int i = 0; i < 16; ++i) { s[i] = pow5(s[i]); c[ci++][r] = s[i]; }
    }
    for (int rd = 0; rd < 14; ++rd) {
      s[0] = m31_add(s[0], PSD_ROUND_CONST);
      internal_matrix(s);
      s[0] = pow5(s[0]);
      c[ci++][r] = s[0];
    }
    for (int rd = 0; rd < 4; ++rd) {
#pragma unroll
      for (int i = 0; i < 16; ++i) s[i] = m31_add(s[i], PSD_ROUND_CONST);
      external_matrix(s);
#pragma unroll
      for (int i = 0; i < 16; ++i) { s[i] = pow5(s[i]); c[ci++][r] = s[i]; }
    }
  }
}
void launch_gen_poseidon(uint32_t log_n, uint32_t* const* cols, uint32_t ipr, uint32_t mode, uint64_t seed,
                         hipStream_t stream) {
  uint32_t n = 1u << log_n;
  hipLaunchKernelGGL(gen_poseidon_kernel, dim3((n + 127) / 128), dim3(128), 0, stream, n, cols, ipr, mode, seed);
}

// LogupTraceGenerator for the Poseidon relation: column rep holds sum_{r' <= rep} frac_{r'} per row;
// the last column's coordinate sums (over all rows) are accumulated into sums[0..4].
// One thread per (row, instance): in a block of 256 threads, thread rep * RB + j owns row
// blockIdx.x * RB + j and instance rep (RB = 2^log_rb = 256 / ipr rounded up to a power of two),
// so the field inversions of a row's instances run side by side; the running sums over the
// instances of a row meet in LDS.
__global__ void __launch_bounds__(256) poseidon_logup_kernel(uint32_t n_rows, uint32_t log_rb,
                                                             const uint32_t* const* __restrict__ trace_cols,
                                                             uint32_t* const* __restrict__ inter_cols,
                                                             const PoseidonParams* __restrict__ params,
                                                             unsigned long long* __restrict__ sums) {
  CRITICAL_PATH_WAVE();
  __shared__ uint32_t fr[4][256];
  __shared__ unsigned long long red[4][256];
  const PoseidonParams& p = *params;
  const uint32_t t = threadIdx.x, rb = 1u << log_rb;
  const uint32_t rep = t >> log_rb, j = t & (rb - 1);
  const uint32_t r = blockIdx.x * rb + j;
  const bool live = r < n_rows && rep < p.ipr;
  qm31 frac = qm_zero();
  if (live) {
    const uint32_t* const* c = trace_cols + rep * PSD_COLS_PER_REP;
    uint32_t s0[16], s1[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      s0[i] = gmem(c[i])[r];
      s1[i] = gmem(c[PSD_COLS_PER_REP - 16 + i])[r];
    }
    const qm31 d0 = combine16(p, s0), d1 = combine16(p, s1);
    frac = qm_mul(qm_sub(d1, d0), qm_inv(qm_mul(d0, d1)));
  }
  fr[0][t] = frac.a; fr[1][t] = frac.b; fr[2][t] = frac.c; fr[3][t] = frac.d;
  __syncthreads();
  qm31 cum = qm_zero();
  for (uint32_t q = 0; q <= rep && q < p.ipr; ++q) {
    const uint32_t u = (q << log_rb) + j;
    cum = qm_add(cum, qm31{fr[0][u], fr[1][u], fr[2][u], fr[3][u]});
  }
  if (live) {
    uint32_t* const* ic = inter_cols + 4 * rep;
    gmem(ic[0])[r] = cum.a; gmem(ic[1])[r] = cum.b; gmem(ic[2])[r] = cum.c; gmem(ic[3])[r] = cum.d;
  }
  const bool last = live && rep + 1 == p.ipr;
  red[0][t] = last ? cum.a : 0;
  red[1][t] = last ? cum.b : 0;
  red[2][t] = last ? cum.c : 0;
  red[3][t] = last ? cum.d : 0;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < 4; ++k) red[k][t] += red[k][t + w];
    __syncthreads();
  }
  if (t < 4) atomicAdd(&sums[t], red[t][0]);
}

void launch_poseidon_logup_fracs(uint32_t log_n, uint32_t ipr, const uint32_t* const* trace_cols,
                                 uint32_t* const* inter_cols, const PoseidonParams* params, unsigned long long* sums,
                                 hipStream_t stream) {
  uint32_t log_ipr = 0;  // instance slots per row: the next power of two
  while ((1u << log_ipr) < ipr) ++log_ipr;
  if (log_ipr > 8) throw std::invalid_argument("LogUp: at most 256 instances per row");
  const uint32_t log_rb = 8 - log_ipr, n = 1u << log_n;
  const uint32_t blocks = (n + (1u << log_rb) - 1) >> log_rb;
  hipLaunchKernelGGL(poseidon_logup_kernel, dim3(blocks), dim3(256), 0, stream, n, log_rb, trace_cols, inter_cols,
                     params, sums);
}

// Inclusive prefix sum in coset order of (value - shift), for 4 coordinate columns stored in
// bit-reversed circle-domain order (simd/prefix_sum.rs): coset index c maps to circle index
// c/2 (c even) or n-1-(c-1)/2 (c odd), stored at bitrev(circle index).
__device__ __forceinline__ uint32_t coset_storage(uint32_t c, uint32_t log_n) {
  uint32_t n = 1u << log_n;
  uint32_t circ = (c & 1) ? n - 1 - (c >> 1) : (c >> 1);
  return bitrev(circ, log_n);
}
constexpr uint32_t SCAN_BLOCK = 1024;

// Coordinate k of the shift: given, or (sums non-null) the LogUp column sums scaled by 1/n, as
// Component::cumsum_shift computes it, so that the scan needs no host round trip.
__device__ __forceinline__ uint32_t scan_shift(qm31 shift, const unsigned long long* __restrict__ sums,
                                               uint32_t inv_n, uint32_t k) {
  if (sums) return m31_mul(m31_reduce64(sums[k]), inv_n);
  return k == 0 ? shift.a : k == 1 ? shift.b : k == 2 ? shift.c : shift.d;
}

// Inclusive M31 prefix sum over a workgroup of SCAN_BLOCK lanes: within each wave by shuffles,
// then over the wave totals.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* wave_tot) {
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)v, off);
    if (lane >= off) v = m31_add(v, u);
  }
  if (lane == 63) wave_tot[w] = v;
  __syncthreads();
  if (w == 0) {
    uint32_t x = lane < SCAN_BLOCK / 64 ? wave_tot[lane] : 0;
#pragma unroll
    for (uint32_t off = 1; off < SCAN_BLOCK / 64; off <<= 1) {
      const uint32_t u = (uint32_t)__shfl_up((int)x, off);
      if (lane >= off) x = m31_add(x, u);
    }
    if (lane < SCAN_BLOCK / 64) wave_tot[lane] = x;
  }
  __syncthreads();
  return w ? m31_add(v, wave_tot[w - 1]) : v;
}

__global__ void __launch_bounds__(SCAN_BLOCK) scan_local_kernel(uint32_t log_n, uint32_t* const* __restrict__ cols,
                                                               qm31 shift, const unsigned long long* __restrict__ sums,
                                                               uint32_t inv_n, uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t wave_tot[SCAN_BLOCK / 64];
  const uint32_t c = blockIdx.x * SCAN_BLOCK + threadIdx.x;
  const uint32_t k = blockIdx.y;
  const uint32_t idx = coset_storage(c, log_n);
  GLOBAL_AS uint32_t* col = gmem(cols[k]);
  const uint32_t v = block_scan(m31_sub(col[idx], scan_shift(shift, sums, inv_n, k)), wave_tot);
  col[idx] = v;
  if (threadIdx.x == SCAN_BLOCK - 1) block_sums[k * gridDim.x + blockIdx.x] = v;
}

// Exclusive scan of the per-block totals, one workgroup per coordinate (n_blocks <= SCAN_BLOCK per
// pass; longer lists carry the running total from pass to pass).
__global__ void __launch_bounds__(SCAN_BLOCK) scan_blocks_kernel(uint32_t n_blocks, uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t wave_tot[SCAN_BLOCK / 64];
  __shared__ uint32_t carry_s;
  const uint32_t k = blockIdx.x;
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < n_blocks; b0 += SCAN_BLOCK) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t v = b < n_blocks ? block_sums[k * n_blocks + b] : 0;
    const uint32_t incl = block_scan(v, wave_tot);
    if (b < n_blocks) block_sums[k * n_blocks + b] = m31_add(carry, m31_sub(incl, v));  // exclusive
    if (threadIdx.x == SCAN_BLOCK - 1) carry_s = m31_add(carry, incl);
    __syncthreads();
    carry = carry_s;
    __syncthreads();
  }
}

__global__ void scan_add_kernel(uint32_t log_n, uint32_t* const* __restrict__ cols,
                                const uint32_t* __restrict__ block_sums) {
  uint32_t c = blockIdx.x * SCAN_BLOCK + threadIdx.x;
  uint32_t k = blockIdx.y;
  uint32_t idx = coset_storage(c, log_n);
  GLOBAL_AS uint32_t* col = gmem(cols[k]);
  col[idx] = m31_add(col[idx], block_sums[k * gridDim.x + blockIdx.x]);
}

// Small columns (n < SCAN_BLOCK) are scanned by one workgroup with n lanes.
__global__ void scan_small_kernel(uint32_t log_n, uint32_t* const* __restrict__ cols, qm31 shift,
                                  const unsigned long long* __restrict__ sums, uint32_t inv_n) {
  __shared__ uint32_t sm[SCAN_BLOCK];
  uint32_t n = 1u << log_n;
  uint32_t k = blockIdx.y;
  uint32_t sh = scan_shift(shift, sums, inv_n, k);
  uint32_t c = threadIdx.x;
  uint32_t idx = c < n ? coset_storage(c, log_n) : 0;
  sm[c] = c < n ? m31_sub(cols[k][idx], sh) : 0;
  __syncthreads();
  for (uint32_t off = 1; off < n; off <<= 1) {
    uint32_t add = (c < n && c >= off) ? sm[c - off] : 0;
    __syncthreads();
    if (c < n) sm[c] = m31_add(sm[c], add);
    __syncthreads();
  }
  if (c < n) cols[k][idx] = sm[c];
}

void launch_coset_prefix_sum(uint32_t log_n, uint32_t* const* cols4, qm31 shift, const unsigned long long* shift_sums,
                             uint32_t* scratch, hipStream_t stream) {
  const uint32_t n = 1u << log_n, inv_n = m31_inv(n);
  if (n < SCAN_BLOCK) {
    hipLaunchKernelGGL(scan_small_kernel, dim3(1, 4), dim3(SCAN_BLOCK), 0, stream, log_n, cols4, shift, shift_sums,
                       inv_n);
    return;
  }
  uint32_t nb = n / SCAN_BLOCK;
  hipLaunchKernelGGL(scan_local_kernel, dim3(nb, 4), dim3(SCAN_BLOCK), 0, stream, log_n, cols4, shift, shift_sums,
                     inv_n, scratch);
  hipLaunchKernelGGL(scan_blocks_kernel, dim3(4), dim3(SCAN_BLOCK), 0, stream, nb, scratch);
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb, 4), dim3(SCAN_BLOCK), 0, stream, log_n, cols4, scratch);
}

}  // namespace sh
