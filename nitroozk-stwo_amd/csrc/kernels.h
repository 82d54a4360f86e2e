// Launch wrappers for the HIP kernels of libstwo_hip. All take device pointers and a stream;
// none allocates or synchronises.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "field.h"

namespace sh {

struct Cols4 { uint32_t* c[4]; };
struct CCols4 { const uint32_t* c[4]; };

// fft.hip
// tree: a doubled twiddle tree (every entry 2t; Ctx::tw2 for evaluation, Ctx::itw2 for
// interpolation). background: the transform runs beside critical-path kernels on another stream
// and leaves them the issue priority. low_log: the tile of the low-stage kernel, 2^low_log values
// per workgroup (11 <= low_log <= min(k, 15) for a transform of k >= 11 stages; 0 = the largest),
// the stages above it run in strided passes. low_log | FFT_LOW_LOOP (low_log 15; every output
// column 16-byte aligned): the low-stage kernel runs one workgroup per CU, each looping over tiles
// and loading its next tile while it computes the current one.
constexpr uint32_t FFT_LOW_LOOP = 0x100;
void launch_fft(bool inverse, const uint32_t* const* src, uint32_t* const* dst, uint32_t n_cols, uint32_t log_in,
                uint32_t log_n, const uint32_t* tree, uint32_t tree_len, hipStream_t stream, bool background = false,
                uint32_t low_log = 0);

// merkle.hip
void launch_merkle_layer(uint32_t log_size, const uint32_t* prev, const uint32_t* const* cols, uint32_t n_cols,
                         uint32_t* out, hipStream_t stream);
// The same with column c at base + c * stride (columns of one allocation): no address table.
void launch_merkle_layer_strided(uint32_t log_size, const uint32_t* prev, const uint32_t* base, size_t stride,
                                 uint32_t n_cols, uint32_t* out, hipStream_t stream);
// The same for at most 16 columns, given as a host array of device addresses.
void launch_merkle_layer_small(uint32_t log_size, const uint32_t* prev, const uint32_t* const* host_cols,
                               uint32_t n_cols, uint32_t* out, hipStream_t stream);
// `levels` (<= 9, <= log_child) column-free levels above a layer of 2^log_child nodes; outs[j]
// receives level log_child - 1 - j.
void launch_merkle_subtree(uint32_t log_child, const uint32_t* child, uint32_t levels, uint32_t* const* outs,
                           hipStream_t stream);
// Scans nonce candidates [start, start + count) in search order (hi * 2^20 + lo, lo < 2^20) and
// atomically keeps the smallest solution in *result.
// prefix: the 8 words of the channel's proof-of-work prefix digest.
struct GrindPrefix { uint32_t w[8]; };
void launch_grind(const GrindPrefix& prefix, uint32_t pow_bits, uint64_t start, uint32_t count,
                  unsigned long long* result, hipStream_t stream);
// The same search with the prefix read from device memory (8 words), for a prefix that an earlier
// kernel on the stream computed.
void launch_grind_dev(const uint32_t* prefix, uint32_t pow_bits, uint64_t start, uint32_t count,
                      unsigned long long* result, hipStream_t stream);

// ops.hip
void launch_fold_circle_into_line(uint32_t log_src, CCols4 src, Cols4 dst, qm31 alpha, const uint32_t* itree,
                                  uint32_t tree_len, hipStream_t stream);
// n_groups secure columns of one size in one launch (srcs, dsts: device tables of 4 addresses per
// group, all 16-byte aligned; log_src >= 3), each dst = dst * alpha^2 + fold(src).
void launch_fold_circle_into_line_grouped(uint32_t log_src, uint32_t n_groups, const uint32_t* const* srcs,
                                          uint32_t* const* dsts, qm31 alpha, const uint32_t* itree, uint32_t tree_len,
                                          hipStream_t stream);
// The same into a destination whose previous contents are ignored (dst = fold(src)).
void launch_fold_circle_into_line_zero(uint32_t log_src, CCols4 src, Cols4 dst, qm31 alpha, const uint32_t* itree,
                                       uint32_t tree_len, hipStream_t stream);
void launch_fold_line(uint32_t log_src, CCols4 src, Cols4 dst, qm31 alpha, const uint32_t* itree, uint32_t tree_len,
                      hipStream_t stream);

// Variants that read the folding coefficient from device memory (drawn by the device channel);
// fold_line squares it n_squarings times first (fold steps > 1). With dst_zero the circle fold
// treats dst as zero (writes without reading it).
void launch_fold_circle_into_line_dev(uint32_t log_src, CCols4 src, Cols4 dst, const qm31* alpha_dev, bool dst_zero,
                                      const uint32_t* itree, uint32_t tree_len, hipStream_t stream);
void launch_fold_line_dev(uint32_t log_src, CCols4 src, Cols4 dst, const qm31* alpha_dev, uint32_t n_squarings,
                          const uint32_t* itree, uint32_t tree_len, hipStream_t stream);
// Device channel (state = 8 digest words + n_draws): mix a root, draw one QM31. record[0..8) gets
// the root, record[8..12) the drawn QM31.
void launch_channel_mix_root_draw(uint32_t* state, const uint32_t* root, uint32_t* record, hipStream_t stream);

// fri.hip: one FRI layer per step, fused: fold from the layer above, Merkle tree of the layer's 4
// coordinate columns, channel step (mix the root, draw the next folding coefficient).
enum FriMode : uint32_t { FRI_LEAVES = 0, FRI_FOLD_CIRCLE = 1, FRI_FOLD_LINE = 2 };
struct FriLayerDesc {
  uint32_t log;     // the layer holds 2^log values
  uint32_t mode;    // FRI_LEAVES: the layer is src itself; folds: src is the layer above (2^(log+1) values)
  uint32_t commit;  // 0 for the last layer: folded only
  uint32_t init;    // nonzero on the launch that opens the commit phase (fri_layer_kernel only): the
                    // channel state comes from st0, and the launch sets the proof-of-work result
                    // slot (state words 14 and 15) to all ones
  CCols4 src;
  Cols4 dst;                // folds: the layer's 4 coordinate columns
  const uint32_t* tw;       // FRI_FOLD_CIRCLE: line layer 0 of the inverse twiddle tree of the layer above
                            // (itree + tree_len - 2^log); FRI_FOLD_LINE: itree + tree_len - 2^(log+1)
  const uint32_t* alpha;    // folds: the 4 words of the folding coefficient drawn for the layer above
  uint32_t* nodes;          // Merkle heap: level j (2^j nodes of 8 words) at 8 (2^j - 1) words, 16-byte aligned
  uint32_t* rec;            // 12 words: root, then the coefficient drawn after it
  uint32_t st0[9];          // init: channel digest (8 words) and n_draws at the start of the phase
};
// One layer, 9 <= log <= 18, over 2^(log - c) workgroups (c = 7 up to log 14, then max(8, log - 8));
// *ticket is zero before the launch, and the launch leaves it at zero. state: channel digest (8
// words) and n_draws.
void launch_fri_layer(const FriLayerDesc& d, uint32_t* state, uint32_t* ticket, hipStream_t stream);
// Layers of at most 2^8 values in one workgroup, in order (descs in device memory); every entry
// after the first is a fold of the one before it. Only the last entry may have commit == 0.
// The FRI commit phase's end on the device (prover/fri.rs:215-238, core/proof_of_work.rs): the
// last layer's 2^L values (4 coordinate columns at vals) are interpolated into line coefficients
// (xi: the inverse x-coordinates of the line domain's points, level by level, 2^L - 1 words, then
// the inverse of 2^L), the 2^BL coefficients of the last polynomial are mixed into the channel
// state, and the proof-of-work prefix for pow_bits is written to prefix (8 words). state gets the
// mixed digest and n_draws = 0. L <= 6, BL <= min(L, 4); one workgroup of 64 threads.
void launch_fri_last_pow(const uint32_t* vals, uint32_t L, uint32_t BL, const uint32_t* xi, uint32_t* state,
                         uint32_t pow_bits, uint32_t* prefix, hipStream_t stream);
void launch_fri_small(const FriLayerDesc* descs_dev, uint32_t n, uint32_t* state, hipStream_t stream);

// Fused accumulate_numerators for one size group. Entries of batch b are
// [batch_start[b], batch_start[b+1]); out holds 4 coordinate pointers per batch. aligned16 says
// that every column address is 16-byte aligned (enables the vectorised form).
struct NumEntry { const uint32_t* col; uint32_t pad[2]; qm31 c; };
void launch_accumulate_numerators(uint32_t log_size, const NumEntry* entries, const uint32_t* batch_start,
                                  const qm31* batch_b, uint32_t n_batches, uint32_t* const* out, bool aligned16,
                                  hipStream_t stream);

// A point's linear term coefficient is first_linear plus the a of each of its parts.
struct QuotPoint { cm31 prx, pix, pry, piy; qm31 first_linear; uint32_t part_begin, part_end, pad[2]; };
struct QuotPart { const uint32_t* c[4]; uint32_t log_size; uint32_t pad[3]; qm31 a; };

// Line coefficients of the quotient batches from the sampled values on the device
// (core/constraints.rs complex_conjugate_line_coeffs with the batch sums of
// core/pcs/quotients.rs:200-298). Sample q of the proof has random coefficient r = qrc^q and value
// vals[vi]; for batch b (entries [first, first + count), sample point y-coordinate py,
// cy = conj(py) - py): entry e gets c = r cy (out_entries[e] = {col, c}), batch_b[b] =
// cy sum r v - py sum r (conj(v) - v), and parts[part].a = sum r (conj(v) - v).
// scratch: 8 words per entry.
struct QCoefEntry { const uint32_t* col; uint32_t q, vi, batch, pad; };
struct QCoefBatch { uint32_t first, count, part, pad; qm31 cy, py; };
void launch_quotient_coeffs(qm31 qrc, const qm31* vals, const QCoefEntry* entries, uint32_t n_entries,
                            const QCoefBatch* batches, uint32_t n_batches, NumEntry* out_entries, qm31* batch_b,
                            QuotPart* parts, qm31* scratch, hipStream_t stream);
void launch_quotients(uint32_t lifting_log, const uint32_t* domain_xy, const QuotPoint* points, uint32_t n_points,
                      const QuotPart* parts, Cols4 out, hipStream_t stream);

void launch_accumulate_lifted(uint32_t log_dst, Cols4 dst, uint32_t log_src, CCols4 src, hipStream_t stream);

// OODS: value of FFT-basis coefficient columns at QM31 points. Tasks are grouped by (point, size):
// group_factors[g*32 + b] is the folding factor of bit b (y, x, pi(x), ...). Scratch: weights holds
// 2^OODS_TABLE_LOG QM31 per group, partials oods_max_chunks(max_log) QM31 per task.
constexpr uint32_t OODS_TABLE_LOG = 16;
constexpr uint32_t OODS_BLOCK_LOG = 11;  // coefficients per block of the chunk kernel
inline uint32_t oods_max_chunks(uint32_t max_log) {
  return max_log > OODS_BLOCK_LOG ? 1u << (max_log - OODS_BLOCK_LOG) : 1u;
}
struct OodsTask { const uint32_t* coeffs; uint32_t log_size; uint32_t out_index; uint32_t pad[2]; };
// Tasks are sorted by group; batch b is tasks [batch_start[b], batch_start[b+1]), at most
// OODS_BATCH of one group, which share each weight read (oods_batch_starts builds the list).
constexpr uint32_t OODS_BATCH = 8;
inline std::vector<uint32_t> oods_batch_starts(const std::vector<uint32_t>& sorted_task_group) {
  std::vector<uint32_t> st;
  for (uint32_t i = 0; i < sorted_task_group.size(); ++i)
    if (st.empty() || i - st.back() == OODS_BATCH || sorted_task_group[i] != sorted_task_group[i - 1]) st.push_back(i);
  st.push_back((uint32_t)sorted_task_group.size());
  return st;
}
void launch_oods(const OodsTask* tasks, const uint32_t* task_group, uint32_t n_tasks, const uint32_t* batch_start,
                 uint32_t n_batches, uint32_t max_log, const qm31* group_factors, const uint32_t* group_log,
                 uint32_t n_groups, qm31* weights, qm31* partials, qm31* out, hipStream_t stream);

// Barycentric weights of CanonicCoset(log_size) for p (4 coordinate columns) and the batched
// sums sum_i evals_c[i] w_i; shifts[j-1] (j = 1 .. log_size-1) as in barycentric_weights_kernel.
// partial holds n_cols * ceil(2^log_size / 2048) QM31 (scratch).
void launch_barycentric_weights(uint32_t log_size, cpqm31 p, qm31 vn, const cpm31* shifts, Cols4 out,
                                hipStream_t stream);
void launch_barycentric_eval(uint32_t log_size, const uint32_t* const* evals, uint32_t n_cols, CCols4 weights,
                             qm31* partial, qm31* out, hipStream_t stream);

void launch_gather_words(const uint32_t* const* bases, const uint32_t* base_idx, const uint32_t* offsets, uint32_t n,
                         uint32_t* out, hipStream_t stream);
// Column::batch_at_multi: out[r * n_cols + c] = cols[c][rows[r]]. A gather of at most
// GATHER_ARG_COLS columns and GATHER_ARG_ROWS rows takes both lists as kernel arguments (no
// transfer before the launch); a larger one reads them from device tables.
constexpr uint32_t GATHER_ARG_COLS = 16, GATHER_ARG_ROWS = 48;
struct GatherArgs {
  const uint32_t* cols[GATHER_ARG_COLS];
  uint32_t rows[GATHER_ARG_ROWS];
};
// out_is_host: out is the device address of coherent pinned host memory (the kernel fences at
// system scope before it ends)
void launch_gather_args(const GatherArgs& g, uint32_t n_cols, uint32_t n_rows, uint32_t* out, bool out_is_host,
                        hipStream_t stream);
void launch_gather_rows(const uint32_t* const* cols, uint32_t n_cols, const uint32_t* rows, uint32_t n_rows,
                        uint32_t* out, hipStream_t stream);

// Decommitment reads for a set of query positions, in one launch. A Merkle tree (or FRI layer)
// decommitment only touches, at each level, the block of 2^w nodes that holds the query's path
// node: k = pos >> (top - level + w) gives block [k 2^w, (k+1) 2^w). An item copies, for every
// query, either the values of a chunk of the columns of one level at the nodes of that block
// (DEC_COLS, node-major rows of row_width words) or the child hashes under the block (DEC_HASH:
// 2^(w+1) nodes of the level below, 8 words each). pos is the query's position in the item's top
// level: mode 0 ((p >> 1) << a) + (p & 1), mode 1 ((p >> a) << 1) + (p & 1), mode 2 p >> a.
// Nodes past the end of their level read as zero.
enum DecKind : uint32_t { DEC_COLS = 0, DEC_HASH = 1 };
struct DecItem {
  const uint32_t* const* cols;  // DEC_COLS: this chunk's column addresses
  const uint32_t* layer;        // DEC_HASH: the child level's nodes
  uint32_t kind;
  uint32_t width;       // words per row read: chunk columns, or 8
  uint32_t row_width;   // words per row in the output block
  uint32_t rows_log;    // rows per query: 2^w (DEC_COLS) or 2^(w+1) (DEC_HASH)
  uint32_t level_log;   // nodes in the level read: 2^level_log
  uint32_t shift;       // k = pos >> shift; first row node = k << rows_log
  uint32_t pos_mode, pos_arg;
  uint32_t out, stride;  // query i's block at out + i * stride
};
constexpr uint32_t DEC_MAX_QUERIES = 64;  // positions per launch (kernel argument)
// Splits into launches of at most DEC_MAX_QUERIES positions; positions[i] is query i.
void launch_decommit_gather(const DecItem* items, uint32_t n_items, const uint32_t* positions, uint32_t n_queries,
                            uint32_t* out, hipStream_t stream);

// Composition (constraint quotients on the evaluation domain), added into acc (accumulate) or
// written to it.
// Grouped wide-Fibonacci composition: groups of components of one evaluation size, 64 rows per block
// (group g's blocks start at block_off), components comp_off .. comp_off + n_comps of `comps`.
struct WfGroup {
  const uint32_t* dinv;
  uint32_t* acc[4];
  uint32_t block_off, eval_log, trace_log, accumulate, comp_off, n_comps;
};
struct WfComp {
  const uint32_t* const* cols;
  const qm31* pows;
  uint32_t n_cols, pad;
};
void launch_wide_fib_composition_grouped(const WfGroup* groups, uint32_t n_groups, const WfComp* comps,
                                         uint32_t n_blocks, hipStream_t stream);
void launch_wide_fib_composition(uint32_t eval_log, uint32_t trace_log, const uint32_t* const* cols, uint32_t n_cols,
                                 const qm31* pows, const uint32_t* denom_inv, Cols4 acc, bool accumulate,
                                 hipStream_t stream);
struct PoseidonParams { qm31 z; qm31 alpha_pow[16]; qm31 shift; uint32_t ipr; uint32_t pad[3]; };
// ipr (instances per row, also in params) sets the launch shape: min(ipr, 8) waves per 64 rows.
void launch_poseidon_composition(uint32_t eval_log, uint32_t trace_log, const uint32_t* const* trace_cols,
                                 const uint32_t* const* inter_cols, const PoseidonParams* params, uint32_t ipr,
                                 const qm31* pows, const uint32_t* denom_inv, Cols4 acc, bool accumulate,
                                 hipStream_t stream);
// The same with trace column c at trace_base + c stride and interaction column c at
// inter_base + c stride (column blocks of one allocation each).
void launch_poseidon_composition_strided(uint32_t eval_log, uint32_t trace_log, const uint32_t* trace_base,
                                         const uint32_t* inter_base, size_t stride, const PoseidonParams* params,
                                         uint32_t ipr, const qm31* pows, const uint32_t* denom_inv, Cols4 acc,
                                         bool accumulate, hipStream_t stream);

// Trace generation and the Poseidon LogUp interaction trace.
void launch_gen_wide_fib(uint32_t log_n, uint32_t* const* cols, uint32_t n_cols, uint32_t mode, uint64_t seed,
                         hipStream_t stream);
void launch_gen_poseidon(uint32_t log_n, uint32_t* const* cols, uint32_t ipr, uint32_t mode, uint64_t seed,
                         hipStream_t stream);
// ipr: instances per row, at most 256 (also in params).
void launch_poseidon_logup_fracs(uint32_t log_n, uint32_t ipr, const uint32_t* const* trace_cols,
                                 uint32_t* const* inter_cols, const PoseidonParams* params, unsigned long long* sums,
                                 hipStream_t stream);
// shift_sums (4 u64 sums of M31 values, may be null) replaces shift by sums / 2^log_n, read on the
// device.
void launch_coset_prefix_sum(uint32_t log_n, uint32_t* const* cols4, qm31 shift, const unsigned long long* shift_sums,
                             uint32_t* scratch, hipStream_t stream);

void launch_fill_zero(uint32_t* p, size_t n, hipStream_t stream);
void launch_copy_words(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t stream);

// columns.hip
// Powers g^(2^b), b < 31, of the circle generator, passed by value to the twiddle-tree kernel.
struct GenPowers { cpm31 p[31]; };
// The twiddle tree of Coset::new(init_index, log_size) and its inverse, plain and doubled
// (2^log_size words each); zero_flag (one word, cleared by the caller) is raised when an
// x-coordinate is zero.
void launch_twiddle_tree(uint32_t init_index, uint32_t log_size, uint32_t* tw, uint32_t* itw, uint32_t* tw2,
                         uint32_t* itw2, uint32_t* zero_flag, hipStream_t stream);
void launch_bit_reverse(uint32_t* col, uint32_t log_n, hipStream_t stream);
// Circle transforms on domains of log size 1 or 2 (the half coset's first point given).
void launch_fft_tiny(bool inverse, const uint32_t* const* src, uint32_t* const* dst, uint32_t n_cols, uint32_t log_in,
                     uint32_t log_n, cpm31 half_coset_initial, hipStream_t stream);

// measure.hip: 32-bit VALU lane-instructions per second on the current device (synchronous; scratch:
// one word per workgroup, 16 x CUs words).
double measure_int32_issue(uint32_t* scratch, hipStream_t stream);
// Forward (or inverse) FFT butterflies per second on register-resident data (butterfly.h).
double measure_butterfly_rate(bool inverse, uint32_t* scratch, hipStream_t stream);
// Blake2s compressions per second with register-resident states and messages, the Merkle kernels'
// one-lane compression code: the ceiling of the Blake2s Merkle layers. scratch: CUs x 16 words.
double measure_compression_rate(uint32_t* scratch, hipStream_t stream);
// The Blake2s channel on the device (channel_dev.hip): state of CHANNEL_WORDS words (digest 0..7,
// n_draws 8, the grind's u64 result slot 10..11, the proof-of-work prefix 12..19). Single-lane
// launches; launch_chan_grind covers the first `count` candidates of the nonce order.
constexpr uint32_t CHANNEL_WORDS = 32;
struct ChanDigest {
  uint32_t w[8];
};
void launch_chan_init(uint32_t* st, const ChanDigest& d, uint32_t n_draws, hipStream_t stream);
void launch_chan_mix_words(uint32_t* st, const uint32_t* words, uint32_t n, hipStream_t stream);
void launch_chan_draw_felts(uint32_t* st, uint32_t n, uint32_t* out, hipStream_t stream);
void launch_chan_draw_queries(uint32_t* st, uint32_t log_domain, uint32_t n, uint32_t* out, hipStream_t stream);
void launch_chan_grind(uint32_t* st, uint32_t pow_bits, uint32_t count, uint32_t* nonce, hipStream_t stream);

// Lifted Merkle leaves (merkle_lifted.hip): column c's value at leaf i is its row
// ((i >> (shift + 1)) << 1) | (i & 1), shift = lifting log size - its log size.
constexpr uint32_t LIFTED_BLAKE2S = 0, LIFTED_POSEIDON252 = 1;
struct LiftedCol {
  const uint32_t* ptr;
  uint32_t shift, pad;
};
// 2^lifting_log_size leaves of 8 words (one leaf when n_cols is 0)
void launch_lifted_leaves(uint32_t hasher, uint32_t lifting_log_size, const LiftedCol* cols, uint32_t n_cols,
                          uint32_t* out, hipStream_t stream);

// Starknet Poseidon permutations per second on register-resident states (poseidon252.hip): the
// ceiling of the Poseidon252 Merkle layers. scratch: CUs x 16 words.
double measure_permutation_rate(uint32_t* scratch, hipStream_t stream);

// poseidon252.hip: the Poseidon252 Merkle channel. Hashes are canonical felt252s, 8 little-endian
// words each. prev: 2^(log_size+1) child hashes or null; cols: n_cols M31 columns of 2^log_size.
void launch_p252_layer(uint32_t log_size, const uint32_t* prev, const uint32_t* const* cols, uint32_t n_cols,
                       uint32_t* out, hipStream_t stream);
// The first nonce (hi << 32) | lo, lo < 2^14, in (hi, lo) order, whose hash with the proof-of-work
// prefix of `digest` (canonical) has pow_bits trailing zero bits; d_best: 8 bytes of device
// scratch. Synchronises the stream.
uint64_t p252_grind(const uint32_t digest[8], uint32_t pow_bits, unsigned long long* d_best, hipStream_t stream);
// Host: poseidon_hash_many of n canonical values (8 words each); poseidon_permute_comp of a
// canonical state of 3; whether 8 words hold a canonical felt252.
void p252_hash_many_host(const uint32_t* values, size_t n, uint32_t out[8]);
void p252_permute_host(uint32_t state[24]);
bool p252_canonical(const uint32_t* v);

}  // namespace sh
