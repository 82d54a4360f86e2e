// Device-resident prover of libstwo_hip: the reference protocol (prove / prove_ex,
// CommitmentSchemeProver, FriProver, MerkleProver) orchestrated over HIP kernels.
#pragma once
#include "runtime.h"

namespace sh {

struct PcsCfg {
  uint32_t pow_bits = 5;
  uint32_t log_blowup = 1;
  uint32_t log_last_layer = 0;
  uint32_t n_queries = 3;
  uint32_t fold_step = 1;
  uint32_t lifting_log_size = 0;  // 0 = None
};

enum AirKind { AIR_WIDE_FIB = 0, AIR_POSEIDON = 1 };

struct Component {
  AirKind kind = AIR_WIDE_FIB;
  uint32_t log_n = 0;
  uint32_t n_cols = 0;  // wide-fib sequence length
  uint32_t ipr = 1;     // Poseidon instances per row
  qm31 z{}, alpha{};
  qm31 alpha_pow[16]{};
  qm31 claimed_sum{};
  size_t trace_start = 0, inter_start = 0;
  uint32_t n_trace_cols() const { return kind == AIR_WIDE_FIB ? n_cols : ipr * 158; }
  uint32_t n_inter_cols() const { return kind == AIR_WIDE_FIB ? 0 : 4 * ipr; }
  uint32_t n_constraints() const { return kind == AIR_WIDE_FIB ? n_cols - 2 : ipr * 143; }
  uint32_t eval_log() const { return kind == AIR_WIDE_FIB ? log_n + 1 : log_n + 2; }
  qm31 cumsum_shift() const { return qm_mul_m31(claimed_sum, m31_inv(1u << log_n)); }
};

// A committed tree: columns (evaluations, optionally coefficients) and its Merkle layers.
struct DevTree {
  std::vector<const uint32_t*> coeffs;  // per column, null when not kept
  std::vector<uint32_t> coeff_logs;
  std::vector<const uint32_t*> evals;
  std::vector<uint32_t> eval_logs;
  std::vector<DevBuf> owned;
  std::vector<DevBuf> layers;  // layers[k]: 2^k nodes of 8 words; layers[0] is the root
  Hash root{};
  // Evaluations of every column on CanonicCoset(ext_log), computed on the side stream ahead of
  // the composition (empty when not precomputed); readers join the side stream first.
  std::vector<DevBuf> ext;
  uint32_t ext_log = 0;
  uint32_t max_log() const { return (uint32_t)layers.size() - 1; }
};

struct ProveResult {
  int status = 0;  // 0 ok, 2 ConstraintsNotSatisfied (reference prove_ex sanity check)
  std::vector<uint8_t> proof;
  Hash digest{};
  qm31 claimed_sum{};
  double stage_ms[8] = {0};  // trace commit, interaction, composition, oods, quotients, fri, pow, decommit
};

// GrindOps::grind on the device: smallest nonce for the prefixed digest (simd/grind.rs order).
// slot: a device u64 that holds ~0 (allocated and set here when null); receives the nonce.
uint64_t device_grind(Ctx& ctx, const Hash& prefix, uint32_t pow_bits, unsigned long long* slot = nullptr);
// Builds the Merkle layers of a tree from its evaluation columns (MerkleProver::commit). With
// read_root false the root stays on the device (layers[0]) and t.root is not set.
void merkle_commit(Ctx& ctx, DevTree& t, bool read_root = true);
// Interpolates columns (any sizes) into new coefficient buffers owned by `into`.
void interpolate_columns(Ctx& ctx, const std::vector<const uint32_t*>& evals, const std::vector<uint32_t>& logs,
                         DevTree& into);
// Evaluates the tree's coefficients with blowup into new evaluation buffers.
void evaluate_polynomials(Ctx& ctx, DevTree& t, uint32_t log_blowup);

ProveResult prove_wide_fib(Ctx& ctx, const std::vector<const uint32_t*>& trace, uint32_t log_n, uint32_t n_cols,
                           const PcsCfg& cfg);
// Several wide-Fibonacci components (2^logs[i] rows, ncols[i] columns each) in one proof; trace
// holds the components' columns in order.
ProveResult prove_wide_fib_multi(Ctx& ctx, const std::vector<const uint32_t*>& trace, const std::vector<uint32_t>& logs,
                                 const std::vector<uint32_t>& ncols, const PcsCfg& cfg);
ProveResult prove_poseidon(Ctx& ctx, const std::vector<const uint32_t*>& trace, uint32_t log_n, uint32_t ipr,
                           const PcsCfg& cfg);

}  // namespace sh

typedef struct stwo_hip_ctx stwo_hip_ctx;
namespace sh {
// The MerkleChannel of a proof: Blake2sMerkleChannel (core/vcs/blake2_merkle.rs:45-59) or
// Poseidon252MerkleChannel (core/vcs/poseidon252_merkle.rs:78-85).
enum class MerkleKind { Blake2s, Poseidon252 };

namespace percall {
// prove() with every device step through the C ABI's trait entry points (percall.cpp): the plain
// per-call form, byte-identical to the device-resident prover (Blake2s), and the prover over the
// Poseidon252 Merkle channel.
ProveResult prove_wide_fib_multi(stwo_hip_ctx* h, const std::vector<const uint32_t*>& trace,
                                 const std::vector<uint32_t>& logs, const std::vector<uint32_t>& ncols,
                                 const PcsCfg& cfg, MerkleKind mc = MerkleKind::Blake2s);
ProveResult prove_poseidon(stwo_hip_ctx* h, const std::vector<const uint32_t*>& trace, uint32_t log_n, uint32_t ipr,
                           const PcsCfg& cfg, MerkleKind mc = MerkleKind::Blake2s);
}  // namespace percall
}  // namespace sh
