/*
 * libstwo_hip: MI355X (gfx950) prover backend for Circle STARKs over M31 with the Blake2s
 * Merkle channel and the Poseidon252 Merkle channel. C ABI with plain pointers and sizes: the boundary a reference-side backend
 * binds, as CudaBackend binds libstwo_cuda through crates/stwo/src/stwo_cuda/bindings.rs.
 *
 * Conventions
 *  - Every function returns an int status: 0 ok, 1 invalid argument, 2 HIP runtime error,
 *    3 unsupported size or configuration, 4 protocol failure (e.g. FRI degree check),
 *    5 unknown component, 6 constraints not satisfied (prove_ex's sanity check; the proof
 *    handle is still set, see the whole-proof entry points). The message of the last failure
 *    is in stwo_hip_last_error().
 *    (The reference exits the process on any device error, CU/utils.cuh:37-40.)
 *  - Column arguments are device pointers to u32 words; "const uint32_t* const* cols" is a
 *    HOST array of device pointers. Every M31 word produced is canonical, in [0, 2^31 - 1).
 *  - QM31 values are passed by value as stwo_qm31 {a, b, c, d} = (a + b i) + (c + d i) u,
 *    as CudaSecureField (bindings.rs:9-43). Secure columns are 4 coordinate columns.
 *  - Evaluations are in bit-reversed circle-domain order of CanonicCoset(log_size);
 *    coefficients are in the FFT basis (crates/stwo/src/prover/poly/circle/ops.rs:125-141).
 *  - One context owns one stream; calls on one context are serialised by the caller.
 *    Different contexts are independent. Calls are stream-ordered unless stated otherwise.
 */
#ifndef STWO_HIP_H
#define STWO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct stwo_hip_ctx stwo_hip_ctx;
typedef struct { uint32_t a, b, c, d; } stwo_qm31;
typedef struct { stwo_qm31 x, y; } stwo_circle_point_qm31;

/* PcsConfig + FriConfig (crates/stwo/src/core/pcs/mod.rs:29-80, core/fri.rs:32-81).
 * lifting_log_size = 0 means None. */
typedef struct {
  uint32_t pow_bits;
  uint32_t log_blowup_factor;
  uint32_t log_last_layer_degree_bound;
  uint32_t n_queries;
  uint32_t line_fold_step;
  uint32_t lifting_log_size;
} stwo_pcs_config;

/* ---------------------------------------------------------------- context / memory ----- */
/* No reference counterpart: the CUDA library uses the legacy default stream and a global
 * pool (CU/cuda_mem_pool.cuh). */
int stwo_hip_create(int device, stwo_hip_ctx** out);
void stwo_hip_destroy(stwo_hip_ctx* ctx);
int stwo_hip_synchronize(stwo_hip_ctx* ctx);
const char* stwo_hip_last_error(stwo_hip_ctx* ctx);

/* cuda_malloc_uint32_t / cuda_alloc_zeroes_uint32_t / cuda_free_memory and the copies
 * (bindings.rs:84-131): Column::zeros / uninitialized / to_cpu / from_iter. stwo_hip_free takes
 * only addresses stwo_hip_alloc returned on the same context (status 1 otherwise); freed blocks
 * are kept by the context and reused, in stream order, by later allocations of the same size. */
int stwo_hip_alloc(stwo_hip_ctx* ctx, size_t n_words, int zeroed, uint32_t** out);
int stwo_hip_free(stwo_hip_ctx* ctx, uint32_t* ptr);
/* The same for n_cols columns of n_words each in one call (a trace's columns: one library call
 * instead of one per column, which is what a host binding pays per call). out receives n_cols
 * addresses; the free stops at the first address the context did not hand out (status 1). */
int stwo_hip_alloc_columns(stwo_hip_ctx* ctx, uint32_t n_cols, size_t n_words, int zeroed, uint32_t** out);
int stwo_hip_free_columns(stwo_hip_ctx* ctx, uint32_t* const* ptrs, uint32_t n_cols);
int stwo_hip_copy_h2d(stwo_hip_ctx* ctx, uint32_t* dst, const uint32_t* src, size_t n_words);
int stwo_hip_copy_d2h(stwo_hip_ctx* ctx, uint32_t* dst, const uint32_t* src, size_t n_words); /* synchronous */
/* Column::batch_at_multi (crates/stwo/src/prover/backend/mod.rs:67-89; batch_gather_multi_uint32,
 * bindings.rs:277-332): out[r * n_cols + c] = cols[c][rows[r]], row-major; synchronous. */
int stwo_hip_batch_at_multi(stwo_hip_ctx* ctx, const uint32_t* const* cols, uint32_t n_cols, const uint32_t* rows,
                            uint32_t n_rows, uint32_t* out_host);

/* ColumnOps::bit_reverse_column (prover/backend/mod.rs:39-42; bit_reverse_base_field, used by
 * prover/backend/cuda/column.rs:14-21): in place, 2^log_size words. */
int stwo_hip_bit_reverse_column(stwo_hip_ctx* ctx, uint32_t* col, uint32_t log_size);
/* Column::at / Column::set (prover/backend/mod.rs:59-62; cuda_get_uint32_t / cuda_set_uint32_t,
 * bindings.rs:84-131). at is synchronous; set is stream-ordered and complete on return. */
int stwo_hip_column_at(stwo_hip_ctx* ctx, const uint32_t* col, size_t index, uint32_t* out);
int stwo_hip_column_set(stwo_hip_ctx* ctx, uint32_t* col, size_t index, uint32_t value);
/* Column::split_at_mid (prover/backend/mod.rs:90-91) and PolyOps::split_at_mid
 * (prover/poly/circle/ops.rs:127-145, whose coefficient halves are p_left and p_right of
 * p = p_left + pi^(L-2)(x) p_right): the halves of a column of n_words (even, at least 2) as
 * borrowed views, *left = col and *right = col + n_words / 2. No device work; the views live as
 * long as col and are not freed on their own. */
int stwo_hip_split_at_mid(stwo_hip_ctx* ctx, uint32_t* col, size_t n_words, uint32_t** left, uint32_t** right);

/* ---------------------------------------------------------------- PolyOps -------------- */
/* PolyOps::extend (prover/poly/circle/ops.rs:86-90; cpu/circle.rs:182-188, cuda/poly.rs:409-420):
 * out (2^new_log_size words) = coeffs (2^log_size words) followed by zeros; out may equal coeffs
 * when that buffer already holds 2^new_log_size words. */
int stwo_hip_extend(stwo_hip_ctx* ctx, const uint32_t* coeffs, uint32_t log_size, uint32_t new_log_size,
                    uint32_t* out);
/* PolyOps::precompute_twiddles(coset) for any Coset::new(coset_initial_index, coset_log_size)
 * (prover/poly/circle/ops.rs:122; precompute_twiddles(initial, step, size), bindings.rs:165-169):
 * the tree and its inverse, built on the device and owned by the handle (freed by
 * stwo_hip_twiddles_free or with the context). Status 1 when a twiddle is zero (no inverse).
 * stwo_hip_twiddles_read copies 2^coset_log_size words of each (host pointers may be NULL). */
typedef struct stwo_hip_twiddles stwo_hip_twiddles;
int stwo_hip_twiddles_create(stwo_hip_ctx* ctx, uint32_t coset_initial_index, uint32_t coset_log_size,
                             stwo_hip_twiddles** out);
int stwo_hip_twiddles_read(stwo_hip_ctx* ctx, const stwo_hip_twiddles* tw, uint32_t* tw_host, uint32_t* itw_host);
int stwo_hip_twiddles_free(stwo_hip_ctx* ctx, stwo_hip_twiddles* tw);
/* PolyOps::interpolate / evaluate on any CircleDomain (ops.rs:23-26,92-96; cpu/circle.rs:28-80,
 * 190-228): the domain is CircleDomain::new(Coset::new(half_coset_initial_index, log_size - 1)),
 * whose half coset must be a doubling of the tree's coset (status 1 otherwise). Evaluations are
 * in bit-reversed domain order. evaluate zero-extends 2^log_coeffs coefficients to the domain. */
int stwo_hip_interpolate_on_domain(stwo_hip_ctx* ctx, const stwo_hip_twiddles* tw, uint32_t half_coset_initial_index,
                                   uint32_t log_size, const uint32_t* const* evals, uint32_t* const* coeffs,
                                   uint32_t n_cols);
int stwo_hip_evaluate_on_domain(stwo_hip_ctx* ctx, const stwo_hip_twiddles* tw, uint32_t half_coset_initial_index,
                                uint32_t log_size, const uint32_t* const* coeffs, uint32_t log_coeffs,
                                uint32_t* const* evals, uint32_t n_cols);
/* PolyOps::eval_at_point_by_folding (ops.rs:77-84; cpu/circle.rs:143-180): the value at point of
 * the polynomial whose bit-reversed evaluations on CanonicCoset(log_size) are evals, by one circle
 * fold and log_size - 1 line folds with the point's coordinates; synchronous. */
int stwo_hip_eval_at_point_by_folding(stwo_hip_ctx* ctx, const uint32_t* evals, uint32_t log_size,
                                      stwo_circle_point_qm31 point, stwo_qm31* out);
/* precompute_twiddles (bindings.rs:165-169; prover/poly/circle/ops.rs:122): caches the twiddle
 * tree of the half coset of CanonicCoset(log_domain) in the context. A tree serves every canonic
 * domain up to its size. tw/itw (host, may be NULL) receive 2^(log_domain-1) words each. */
int stwo_hip_precompute_twiddles(stwo_hip_ctx* ctx, uint32_t log_domain, uint32_t* tw_host, uint32_t* itw_host);
/* interpolate_columns (ops.rs:23-39; ntt_b2n_column, bindings.rs:503-510): evals[i] (bit-reversed
 * evaluations on CanonicCoset(log_size), log_size >= 1) -> coeffs[i]; in place when the pointers
 * are equal. */
int stwo_hip_interpolate_columns(stwo_hip_ctx* ctx, const uint32_t* const* evals, uint32_t* const* coeffs,
                                 uint32_t n_cols, uint32_t log_size);
/* evaluate_polynomials (ops.rs:98-120; ntt_n2b_columns, bindings.rs:512-519): coeffs (2^log_size)
 * -> evals on CanonicCoset(log_size + log_blowup) (2^(log_size+log_blowup) words each). */
int stwo_hip_evaluate_polynomials(stwo_hip_ctx* ctx, const uint32_t* const* coeffs, uint32_t* const* evals,
                                  uint32_t n_cols, uint32_t log_size, uint32_t log_blowup);
/* batch_eval_at_point (ops.rs:41-59; batch_eval_at_points, bindings.rs:180-194): out_host[i] =
 * coeffs[i](point), synchronous. */
int stwo_hip_batch_eval_at_point(stwo_hip_ctx* ctx, const uint32_t* const* coeffs, uint32_t n_cols, uint32_t log_size,
                                 stwo_circle_point_qm31 point, stwo_qm31* out_host);

/* No reference counterpart: batch_eval_at_point for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The same
 * values as stwo_hip_batch_eval_at_point; every device buffer comes from a 32 MiB array that the
 * library holds on each device for the process's lifetime, so the call allocates nothing. Limit:
 * n_cols x 2^max(0, log_size - 11) x 16 bytes of partial sums plus n_cols x 16 bytes of values
 * must fit in 31 MiB (status 3 otherwise), e.g. up to 3960 columns at log_size 20. Calls on one
 * device share the array: concurrent calls from several threads are serialised by a lock, and
 * each call leaves its contents behind. Synchronous. */
int stwo_hip_batch_eval_at_point_noalloc(stwo_hip_ctx* ctx, const uint32_t* const* coeffs, uint32_t n_cols,
                                         uint32_t log_size, stwo_circle_point_qm31 point, stwo_qm31* out_host);

/* barycentric_weights (ops.rs:61-76; barycentric_* in bindings.rs:141-158; cpu/circle.rs:101-132):
 * weights[0..3] (coordinate columns of 2^log_size) receive w_i = V_n(p) / (S_i V_i(p)) over the
 * bit-reversed circle domain of CanonicCoset(log_size). Synchronous. */
int stwo_hip_barycentric_weights(stwo_hip_ctx* ctx, uint32_t log_size, stwo_circle_point_qm31 point,
                                 uint32_t* const weights[4]);
/* barycentric_eval_at_point (ops.rs:61-76; cpu/circle.rs:134-141) for n_cols columns at once:
 * out_host[c] = sum_i evals[c][i] * w_i with the weights above. Synchronous. */
int stwo_hip_batch_barycentric_eval(stwo_hip_ctx* ctx, const uint32_t* const* evals, uint32_t n_cols, uint32_t log_size,
                                    const uint32_t* const weights[4], stwo_qm31* out_host);

/* ---------------------------------------------------------------- MerkleOps ------------ */
/* MerkleOps<Blake2sMerkleHasher>::commit_on_layer (prover/vcs/ops.rs:8-29; commit_on_first_layer /
 * commit_on_layer_with_previous, bindings.rs:262-275): out (2^log_size nodes x 8 words) =
 * Blake2s(prev[2i] || prev[2i+1] || cols[0][i] || ...). prev may be NULL (leaf layer). */
int stwo_hip_commit_on_layer(stwo_hip_ctx* ctx, uint32_t log_size, const uint32_t* prev, const uint32_t* const* cols,
                             uint32_t n_cols, uint32_t* out);

/* MerkleOps<Poseidon252MerkleHasher>::commit_on_layer (the second Merkle channel:
 * core/vcs/poseidon252_merkle.rs:18-41; poseidon252_commit_on_first_layer / _with_previous,
 * stwo_cuda/cuda/poseidon252.cuh, via prover/backend/cuda/poseidon252.rs). A hash is a canonical
 * felt252 as 8 little-endian 32-bit words (starknet FieldElement::to_bytes_be read backwards in
 * 4-byte words). Node i = poseidon_hash(prev[2i], prev[2i+1]) when n_cols is 0, else
 * poseidon_hash_many(prev[2i], prev[2i+1] (when prev is not NULL), the columns' values at i packed
 * 8 per felt252 with the length of a short last group in bits 248..250). out: 2^log_size x 8 words. */
int stwo_hip_commit_on_layer_poseidon252(stwo_hip_ctx* ctx, uint32_t log_size, const uint32_t* prev,
                                         const uint32_t* const* cols, uint32_t n_cols, uint32_t* out);

/* MerkleOpsLifted<H> (prover/vcs_lifted/ops.rs:8-18; CpuBackend prover/backend/cpu/merkle_lifted.rs:
 * 12-87; the CUDA backend's blake2s_build_leaves_fused / blake2s_lift_states / ... and
 * poseidon252_lifted_build_next_layer, stwo_cuda/bindings.rs:223,562). hasher 0: Blake2sMerkleHasher
 * (core/vcs_lifted/blake2_merkle.rs), 1: Poseidon252MerkleHasher (core/vcs_lifted/poseidon252_merkle.rs);
 * hashes are 8 words as for commit_on_layer / commit_on_layer_poseidon252.
 * build_leaves: cols[c] has 2^log_sizes[c] words, log sizes non-decreasing, each in
 * [1, lifting_log_size] (status 1 otherwise). Leaf i hashes, in column order, every column's value
 * at row ((i >> (lifting_log_size - l + 1)) << 1) | (i & 1) (the lifted polynomial on the
 * bit-reversed domain): Blake2s of the values as u32 LE, or poseidon_hash_many of them packed 8 per
 * felt252. out: 2^lifting_log_size x 8 words; with no columns, one hash (the empty hasher's).
 * build_next_layer: out (2^log_size x 8 words) = hash_children(prev[2i], prev[2i+1]). */
int stwo_hip_build_leaves_lifted(stwo_hip_ctx* ctx, uint32_t hasher, const uint32_t* const* cols,
                                 const uint32_t* log_sizes, uint32_t n_cols, uint32_t lifting_log_size,
                                 uint32_t* out);
int stwo_hip_build_next_layer_lifted(stwo_hip_ctx* ctx, uint32_t hasher, const uint32_t* prev, uint32_t log_size,
                                     uint32_t* out);

/* ---------------------------------------------------------------- FriOps --------------- */
/* fold_circle_into_line (prover/fri.rs:46-66; bindings.rs:196-212): dst (4 coordinate columns of
 * 2^(log_src-1)) = dst * alpha^2 + fold(src), src on CanonicCoset(log_src). Uses the cached twiddles. */
int stwo_hip_fold_circle_into_line(stwo_hip_ctx* ctx, uint32_t log_src, const uint32_t* const src[4],
                                   uint32_t* const dst[4], stwo_qm31 alpha);
/* Grouped form of fold_circle_into_line for the batched form (no reference counterpart:
 * CudaBackend folds one secure column per call): n_groups secure columns of one size, src and dst
 * holding 4 coordinate addresses per group, folded with the same alpha in one launch. */
int stwo_hip_fold_circle_into_line_grouped(stwo_hip_ctx* ctx, uint32_t log_src, uint32_t n_groups,
                                           const uint32_t* const* src, uint32_t* const* dst, stwo_qm31 alpha);
/* fold_line (prover/fri.rs:31-44; fold_line, bindings.rs:196-212): src on the line domain of log
 * size log_src (Coset::half_odds(log_src)); fold_step folds, the coefficient squared before each
 * fold after the first (prover/backend/cpu/fri.rs:15-30), so dst has 2^(log_src - fold_step)
 * words per coordinate. */
int stwo_hip_fold_line(stwo_hip_ctx* ctx, uint32_t log_src, const uint32_t* const src[4], uint32_t* const dst[4],
                       stwo_qm31 alpha, uint32_t fold_step);

/* ---------------------------------------------------------------- QuotientOps ---------- */
/* One ColumnSampleBatch (core/pcs/quotients.rs:56-100): a point and, per sampled column, the
 * column index, the sampled value and the random-coefficient power. */
typedef struct {
  stwo_circle_point_qm31 point;
  uint32_t n_columns;
  const uint32_t* column_indices;
  const stwo_qm31* sample_values;
  const stwo_qm31* random_coeffs;
} stwo_sample_batch;
/* accumulate_numerators (prover/pcs/quotient_ops.rs:19-43; accumulate_numerators_batch,
 * bindings.rs:354-381): for batch b, out[4b..4b+4) (2^log_size each) =
 * sum_j (c_j * cols[idx_j] - b_j), and first_linear_host[b] = sum_j a_j. */
int stwo_hip_accumulate_numerators(stwo_hip_ctx* ctx, uint32_t log_size, const uint32_t* const* cols, uint32_t n_cols,
                                   const stwo_sample_batch* batches, uint32_t n_batches, uint32_t* const* out,
                                   stwo_qm31* first_linear_host);
/* AccumulatedNumerators (quotient_ops.rs:45-70). */
typedef struct {
  stwo_circle_point_qm31 sample_point;
  const uint32_t* partial[4];
  uint32_t log_size;
  stwo_qm31 first_linear_term;
} stwo_accumulated_numerators;
/* compute_quotients_and_combine (quotient_ops.rs:34-43; compute_quotients_and_combine,
 * bindings.rs:354-381): out (4 x 2^lifting_log_size) on CanonicCoset(lifting_log_size). */
int stwo_hip_compute_quotients_and_combine(stwo_hip_ctx* ctx, const stwo_accumulated_numerators* accs, uint32_t n_accs,
                                           uint32_t lifting_log_size, uint32_t* const out[4]);

/* ---------------------------------------------------------------- AccumulationOps ------ */
/* accumulate (prover/air/accumulation.rs:127-145; bindings.rs:214-221): dst += src. */
int stwo_hip_accumulate(stwo_hip_ctx* ctx, uint32_t log_size, uint32_t* const dst[4], const uint32_t* const src[4]);
/* lift_and_accumulate (accumulation.rs:127-145; lift_and_accumulate, bindings.rs:214-221): cols4
 * holds 4 coordinate pointers per column, sizes ascending; out (4 x 2^max) = sum of lifts. */
int stwo_hip_lift_and_accumulate(stwo_hip_ctx* ctx, const uint32_t* const* cols4, const uint32_t* log_sizes,
                                 uint32_t n_cols, uint32_t* const out[4]);
/* generate_secure_powers (accumulation.rs:127-145): out_host[i] = felt^i. */
int stwo_hip_generate_secure_powers(stwo_hip_ctx* ctx, stwo_qm31 felt, uint32_t n, stwo_qm31* out_host);

/* ---------------------------------------------------------------- GrindOps ------------- */
/* GrindOps<Blake2sChannel>::grind (core/proof_of_work.rs:3-7; grind_blake2s, bindings.rs:615):
 * smallest nonce (hi << 32) | lo, lo < 2^20, for the channel digest; synchronous. */
int stwo_hip_grind(stwo_hip_ctx* ctx, const uint8_t digest[32], uint32_t pow_bits, uint64_t* nonce);
/* GrindOps<Poseidon252Channel>::grind (prover/backend/simd/grind.rs:150-190, which the CUDA
 * backend calls, cuda/backend.rs:52-56): the first nonce (hi << 32) | lo, lo < 2^14, in (hi, lo)
 * order, with poseidon_hash(poseidon_hash_many(0x12345678, digest, pow_bits), nonce) ending in
 * pow_bits zero bits (channel/poseidon252.rs:146-156). digest: a canonical felt252 (8 words LE);
 * pow_bits <= 64 (status 3 beyond); synchronous. */
int stwo_hip_grind_poseidon252(stwo_hip_ctx* ctx, const uint32_t digest[8], uint32_t pow_bits, uint64_t* nonce);
/* No context, no GPU: starknet-crypto's poseidon_hash_many of n canonical felt252s (8 words LE
 * each), and poseidon_permute_comp of a canonical state of 3, on the host. Status 1 for a
 * non-canonical input. For bindings and channels that hash on the host. */
int stwo_poseidon252_hash_many(const uint32_t* values, size_t n, uint32_t out[8]);
int stwo_poseidon252_permute(uint32_t state[24]);

/* ---------------------------------------------------------------- ComponentProver ------ */
/* evaluate_constraint_quotients_on_domain (bindings.rs:469-490; constraint-framework
 * component_prover.rs:249-499). eval_blob = fnv1a(name) (u32 LE) || component parameters:
 *   "fibonacci_example": nothing more (the sequence length is n_trace);
 *   "poseidon_example":  u32 instances_per_row || stwo_qm31 z || stwo_qm31 alpha.
 * acc (4 x 2^eval_log) += (sum_k coeff_powers[k] C_k(row)) * denom_inv[row >> trace_log].
 * trace / interaction columns are on the evaluation domain CanonicCoset(eval_log).
 * Returns 5 for an unknown component (the reference then falls back to the CPU). */
int stwo_hip_evaluate_constraint_quotients_on_domain(stwo_hip_ctx* ctx, uint32_t* const acc[4],
                                                     const uint32_t* const* trace, uint32_t n_trace,
                                                     const uint32_t* const* interaction, uint32_t n_interaction,
                                                     const stwo_qm31* coeff_powers, uint32_t n_constraints,
                                                     const uint32_t* denom_inv_host, uint32_t trace_log,
                                                     uint32_t eval_log, const uint8_t* eval_blob, size_t blob_len,
                                                     stwo_qm31 cumsum_shift);

/* ---------------------------------------------------------------- whole proofs --------- */
/* prove() over the wide-Fibonacci AIR (crates/examples/src/wide_fibonacci/mod.rs:186-261) and the
 * Poseidon2 AIR (crates/examples/src/poseidon/mod.rs:337-419), with the trace already resident:
 * trace[i] are device columns of 2^log_n_rows words (left untouched). The proof is written in the
 * library's canonical byte encoding. Synchronous. When prove_ex's sanity check fails
 * (prover/mod.rs:127-144: the reference returns Err(ConstraintsNotSatisfied) and no proof) the
 * call returns 6, *out is still set, and stwo_proof_status reports 2: the caller inspects the
 * bytes or frees the handle. */
typedef struct stwo_proof stwo_proof;
int stwo_hip_prove_wide_fibonacci(stwo_hip_ctx* ctx, const uint32_t* const* trace, uint32_t n_cols,
                                  uint32_t log_n_rows, const stwo_pcs_config* config, stwo_proof** out);
int stwo_hip_prove_poseidon(stwo_hip_ctx* ctx, const uint32_t* const* trace, uint32_t instances_per_row,
                            uint32_t log_n_rows, const stwo_pcs_config* config, stwo_proof** out);
/* prove() over several wide-Fibonacci components of mixed sizes (prover/mod.rs:28-150 with a
 * component list; trace columns allocated in component order, as TraceLocationAllocator does):
 * component i has 2^log_n_rows[i] rows and n_cols[i] columns; trace holds all components'
 * columns in order. Synchronous. */
int stwo_hip_prove_wide_fibonacci_multi(stwo_hip_ctx* ctx, const uint32_t* const* trace, uint32_t n_components,
                                        const uint32_t* log_n_rows, const uint32_t* n_cols,
                                        const stwo_pcs_config* config, stwo_proof** out);
/* The same proofs driven through the trait entry points above, one C-ABI call per trait-method
 * invocation, in the order the reference's generic prove() makes them on a backend
 * (prover/mod.rs:28-150, pcs/mod.rs, fri.rs, vcs/prover.rs, quotient_ops.rs): the host side a Rust
 * HipBackend runs above this ABI, in C++. This is the plain per-call form of the prover; the
 * entry points above are the fused, device-resident form. Both produce the same bytes.
 * Synchronous; status 6 as above. */
int stwo_hip_prove_wide_fibonacci_per_call(stwo_hip_ctx* ctx, const uint32_t* const* trace, uint32_t n_components,
                                            const uint32_t* log_n_rows, const uint32_t* n_cols,
                                            const stwo_pcs_config* config, stwo_proof** out);
int stwo_hip_prove_poseidon_per_call(stwo_hip_ctx* ctx, const uint32_t* const* trace, uint32_t instances_per_row,
                                     uint32_t log_n_rows, const stwo_pcs_config* config, stwo_proof** out);
/* verify (core/verifier.rs:19-118 with CommitmentSchemeVerifier::verify_values,
 * core/pcs/verifier.rs:60-158) on the host, for proofs in the library's encoding: the statement is
 * the components' shapes (and, for Poseidon2, the claimed LogUp sum) and the expected PcsConfig,
 * which the proof's own must equal. Returns 0 when the proof is accepted, 4 when it is rejected
 * (err receives the reference's error name, e.g. "OodsNotMatching", "ProofOfWork",
 * "Merkle (tree): RootMismatch"), 1 for invalid arguments. Needs no context and no GPU;
 * thread-safe. */
int stwo_verify_wide_fibonacci(const uint8_t* proof, size_t len, uint32_t n_components, const uint32_t* log_n_rows,
                               const uint32_t* n_cols, const stwo_pcs_config* config, char* err, size_t err_len);
int stwo_verify_poseidon(const uint8_t* proof, size_t len, uint32_t instances_per_row, uint32_t log_n_rows,
                         stwo_qm31 claimed_sum, const stwo_pcs_config* config, char* err, size_t err_len);
/* prove() and verify over Poseidon252MerkleChannel (core/vcs/poseidon252_merkle.rs:78-85 with
 * Poseidon252Channel, core/channel/poseidon252.rs:17-156): the same AIRs, configuration and
 * per-call sequence as stwo_hip_prove_*_per_call, with every tree committed by
 * stwo_hip_commit_on_layer_poseidon252, the proof of work ground by stwo_hip_grind_poseidon252 and
 * the channel on felt252 (the reference's test_wide_fib_prove_with_poseidon and
 * test_poseidon_prove_with_poseidon, crates/examples). Every hash in the proof, and the digest
 * stwo_proof_digest returns, is a felt252 as 32 little-endian bytes; the verifier rejects one at or
 * above the Stark prime as "InvalidStructure". Status codes as above. */
int stwo_hip_prove_wide_fibonacci_poseidon252(stwo_hip_ctx* ctx, const uint32_t* const* trace, uint32_t n_components,
                                               const uint32_t* log_n_rows, const uint32_t* n_cols,
                                               const stwo_pcs_config* config, stwo_proof** out);
int stwo_hip_prove_poseidon_poseidon252(stwo_hip_ctx* ctx, const uint32_t* const* trace, uint32_t instances_per_row,
                                        uint32_t log_n_rows, const stwo_pcs_config* config, stwo_proof** out);
int stwo_verify_wide_fibonacci_poseidon252(const uint8_t* proof, size_t len, uint32_t n_components,
                                           const uint32_t* log_n_rows, const uint32_t* n_cols,
                                           const stwo_pcs_config* config, char* err, size_t err_len);
int stwo_verify_poseidon_poseidon252(const uint8_t* proof, size_t len, uint32_t instances_per_row,
                                     uint32_t log_n_rows, stwo_qm31 claimed_sum, const stwo_pcs_config* config,
                                     char* err, size_t err_len);
int stwo_proof_status(const stwo_proof* p);
size_t stwo_proof_size(const stwo_proof* p);
const uint8_t* stwo_proof_bytes(const stwo_proof* p);
void stwo_proof_digest(const stwo_proof* p, uint8_t out[32]);
void stwo_proof_claimed_sum(const stwo_proof* p, stwo_qm31* out);
/* stage wall times in ms: trace commit, interaction trace, composition, OODS, quotients, FRI, PoW,
 * decommit */
void stwo_proof_stage_ms(const stwo_proof* p, double out[8]);
void stwo_proof_free(stwo_proof* p);

/* Trace generation on the device (crates/examples/src/wide_fibonacci/mod.rs:61-83,
 * poseidon/mod.rs:216-303). mode 0: the reference inputs; mode 1: uniform [0, P) from seed. */
int stwo_hip_gen_trace_wide_fibonacci(stwo_hip_ctx* ctx, uint32_t* const* cols, uint32_t n_cols, uint32_t log_n_rows,
                                      uint32_t mode, uint64_t seed);
int stwo_hip_gen_trace_poseidon(stwo_hip_ctx* ctx, uint32_t* const* cols, uint32_t instances_per_row,
                                uint32_t log_n_rows, uint32_t mode, uint64_t seed);
/* The Poseidon2 AIR's LogUp interaction trace (crates/examples/src/poseidon/mod.rs:305-335 with
 * LogupTraceGenerator, constraint-framework/src/prover/logup.rs:21-200): out receives 4 x
 * instances_per_row columns of 2^log_n_rows (the fractions of each pair of instances as QM31
 * coordinates, the last 4 the cumulative sum shifted by claimed_sum / 2^log_n_rows); z and alpha
 * are the lookup elements drawn from the channel. Synchronous. */
int stwo_hip_gen_interaction_trace_poseidon(stwo_hip_ctx* ctx, const uint32_t* const* trace, uint32_t instances_per_row,
                                            uint32_t log_n_rows, stwo_qm31 z, stwo_qm31 alpha, uint32_t* const* out,
                                            stwo_qm31* claimed_sum);

/* ---------------------------------------------------------------- device channel ----- */
/* Blake2sChannel (core/channel/blake2s.rs:19-140) with its state in device memory: a buffer of
 * STWO_HIP_CHANNEL_WORDS words the caller allocates (digest in words 0..7, n_draws in word 8, the
 * rest the library's). No reference counterpart as an FFI (the CUDA backend keeps the channel on the
 * host); it is the device channel API that the batched form needs: every call below except
 * stwo_hip_channel_read is one launch on the context stream with no host synchronisation, so the
 * FRI commit loop (commit a layer, mix its root, draw the folding coefficient, fold) runs without
 * host round trips and can be captured in a graph. Results equal the host channel's.
 *  - init: digest (32 bytes) and n_draws;
 *  - mix_root: digest = H(digest || root) (Blake2sMerkleChannel::mix_root, 8 device words);
 *  - mix_u32s: digest = H(digest || words) for n device words (mix_felts: 4 words a, b, c, d per
 *    felt; mix_u64: low word, high word), n_draws = 0;
 *  - draw_secure_felts: n felts (4 device words each), as draw_secure_felts(n);
 *  - draw_queries: n_queries raw positions, each draw_u32s word masked to 2^log_domain
 *    (core/queries.rs:11-29; sorting and deduplication are the caller's);
 *  - grind: the smallest nonce of GrindOps<Blake2sChannel>::grind for the device digest into
 *    nonce[0..1] (low, high); under graph capture only the first 2^(pow_bits+5) candidates (at least
 *    2^12) are searched, and nonce reads 0xffffffff 0xffffffff in the e^-32-likely event of a miss;
 *    the caller mixes the nonce (mix_u32s of the two words) as prove() does;
 *  - read: the digest and n_draws (synchronous). */
#define STWO_HIP_CHANNEL_WORDS 32
int stwo_hip_channel_init(stwo_hip_ctx* ctx, uint32_t* state, const uint8_t digest[32], uint32_t n_draws);
int stwo_hip_channel_mix_root(stwo_hip_ctx* ctx, uint32_t* state, const uint32_t* root);
int stwo_hip_channel_mix_u32s(stwo_hip_ctx* ctx, uint32_t* state, const uint32_t* words, uint32_t n);
int stwo_hip_channel_draw_secure_felts(stwo_hip_ctx* ctx, uint32_t* state, uint32_t n, uint32_t* out);
int stwo_hip_channel_draw_queries(stwo_hip_ctx* ctx, uint32_t* state, uint32_t log_domain, uint32_t n_queries,
                                  uint32_t* out);
int stwo_hip_channel_grind(stwo_hip_ctx* ctx, uint32_t* state, uint32_t pow_bits, uint32_t* nonce);
int stwo_hip_channel_read(stwo_hip_ctx* ctx, const uint32_t* state, uint8_t digest[32], uint32_t* n_draws);
/* The folds with the coefficient in device memory (4 words, as drawn by
 * stwo_hip_channel_draw_secure_felts): fold_circle_into_line (dst = dst alpha^2 + fold(src); with
 * dst_is_zero the destination's contents are ignored, as for FriProver's zero first line layer,
 * prover/fri.rs:148-152) and fold_line with fold step 1. */
int stwo_hip_fold_circle_into_line_dev(stwo_hip_ctx* ctx, uint32_t log_src, const uint32_t* const src[4],
                                       uint32_t* const dst[4], const uint32_t* alpha, uint32_t dst_is_zero);
int stwo_hip_fold_line_dev(stwo_hip_ctx* ctx, uint32_t log_src, const uint32_t* const src[4], uint32_t* const dst[4],
                           const uint32_t* alpha);

/* ---------------------------------------------------------------- graphs -------------- */
/* No reference counterpart. Capture and replay of a fixed sequence of calls on one context (HIP
 * stream capture, thread-local mode: begin, the captured calls and end on one host thread):
 * stwo_hip_graph_begin synchronises the context and starts capturing its stream; until
 * stwo_hip_graph_end, only these entry points run, launching without
 * synchronising: the device channel calls above (read excepted), the two _dev folds,
 * stwo_hip_fold_circle_into_line, stwo_hip_fold_line with fold step 1 (both need their twiddle sizes
 * built by an earlier call), and stwo_hip_commit_on_layer with at most 16 columns or columns at a
 * stride of one column. Any other entry point returns 3 without acting, and the capture stays open
 * until stwo_hip_graph_end. stwo_hip_graph_launch replays the graph on the context stream with the
 * same device buffers (their current contents); stwo_hip_synchronize waits for it. A captured
 * sequence that starts with stwo_hip_channel_init starts every replay from that channel state. */
typedef struct stwo_hip_graph stwo_hip_graph;
int stwo_hip_graph_begin(stwo_hip_ctx* ctx);
int stwo_hip_graph_end(stwo_hip_ctx* ctx, stwo_hip_graph** out);
int stwo_hip_graph_launch(stwo_hip_ctx* ctx, const stwo_hip_graph* graph);
int stwo_hip_graph_free(stwo_hip_ctx* ctx, stwo_hip_graph* graph);

/* ---------------------------------------------------------------- FFT plans ---------- */
/* No reference counterpart (the reference's transforms take no launch parameters). A circle FFT
 * of k >= 11 stages runs stages [0, S) in one kernel whose workgroup owns 2^S consecutive values
 * (11 <= S <= min(k, 15)) and the stages above in strided passes. The context picks S per shape:
 * the largest tile when its grid fills the device four times over, else the fastest candidate,
 * timed once per (device, direction, columns, sizes) and kept in the cache directory
 * (STWO_HIP_KERNEL_CACHE, else ~/.cache/stwo_hip: fft_plans_<arch>_<CUs>.txt). Results never depend
 * on S. stwo_hip_fft_low_log returns the plan of a shape (tuning it if needed; 0 when the
 * transform has fewer than 11 stages); stwo_hip_set_fft_low_log fixes S for every later transform
 * of the context (clamped to each shape's range; 0 restores planning). A plan may carry the flag
 * 0x100 (S = 15): the low-stage kernel then runs one workgroup per CU, each looping over tiles and
 * loading its next tile while it computes the current one; a transform whose output columns are not
 * all 16-byte aligned runs without it. */
int stwo_hip_fft_low_log(stwo_hip_ctx* ctx, int inverse, uint32_t n_cols, uint32_t log_in, uint32_t log_n,
                         uint32_t* low_log);
int stwo_hip_set_fft_low_log(stwo_hip_ctx* ctx, uint32_t low_log);

/* ---------------------------------------------------------------- measurement -------- */
/* No reference counterpart (its CUDA timer is a stub, CU/timer.cu:3-15). When enabled, the kernels
 * of each class are bracketed with HIP events on the context stream; stats accumulate the event
 * time, the algorithmic HBM bytes and the launch count. Classes: 0 interpolate, 1 evaluate,
 * 2 merkle, 3 composition, 4 accumulate_numerators, 5 quotients, 6 fri_fold, 7 oods, 8 logup,
 * 9 grind, 10 misc, 11 fri_tail (FRI layers of at most 2^12 values: Merkle, channel and fold). */
int stwo_hip_set_kernel_timing(stwo_hip_ctx* ctx, int enable);
/* No reference counterpart. The device's 32-bit VALU issue rate in lane-instructions per second,
 * measured with independent chains of add / xor / rotate (the operations of a Blake2s round): the
 * peak against which VALU-bound kernels (Blake2s hashing, M31 butterflies) are reported. */
int stwo_hip_measure_int32_issue(stwo_hip_ctx* ctx, double* lane_instr_per_s);
/* No reference counterpart. The device's forward (inverse != 0: inverse) circle-FFT butterflies per
 * second with the transform kernels' own butterfly code on register-resident values and twiddles:
 * the arithmetic ceiling of the interpolate / evaluate kernels, against which their butterflies per
 * second (stwo_hip_kernel_units over the class's event time) are reported. */
int stwo_hip_measure_butterfly_rate(stwo_hip_ctx* ctx, int inverse, double* butterflies_per_s);
/* No reference counterpart. Starknet Poseidon permutations (poseidon_permute_comp) per second on
 * register-resident states with the library's own permutation code: the arithmetic ceiling of
 * stwo_hip_commit_on_layer_poseidon252, whose permutations per second (stwo_hip_kernel_units of the
 * merkle class over its event time) are reported against it. */
int stwo_hip_measure_permutation_rate(stwo_hip_ctx* ctx, double* permutations_per_s);
/* No reference counterpart. Blake2s compressions per second on register-resident states and message
 * blocks with the Merkle kernels' own one-lane compression code: the arithmetic ceiling of
 * stwo_hip_commit_on_layer and of the device prover's Blake2s Merkle layers, whose compressions per
 * second (stwo_hip_kernel_units of the merkle class over its event time) are reported against it. */
int stwo_hip_measure_compression_rate(stwo_hip_ctx* ctx, double* compressions_per_s);
int stwo_hip_kernel_stats(stwo_hip_ctx* ctx, uint32_t kernel_class, double* ms, double* bytes, uint64_t* launches);
/* No reference counterpart. Algorithmic work units of a kernel class since the last reset: circle-FFT
 * butterflies for interpolate (n_cols 2^(log-1) log) and evaluate (n_cols 2^(log-1) k, k the
 * coefficients' log size); for merkle launches, Blake2s compressions (stwo_hip_commit_on_layer, the
 * lifted Blake2s layers and the device prover's trees: one per 64-byte block of a node's message,
 * one for an empty message) or Poseidon permutations (stwo_hip_commit_on_layer_poseidon252); 0 for
 * the other classes. */
int stwo_hip_kernel_units(stwo_hip_ctx* ctx, uint32_t kernel_class, double* units);
int stwo_hip_reset_kernel_stats(stwo_hip_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
