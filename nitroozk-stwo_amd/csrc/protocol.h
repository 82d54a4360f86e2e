// Host-side pieces of the reference protocol shared by the two provers of libstwo_hip (the
// device-resident prover and the per-call prover over the trait entry points): the proof
// encoding, sampled-value bookkeeping, the AIRs' point evaluation (prove_ex's sanity check,
// crates/stwo/src/prover/mod.rs:127-144), mask offsets, query folding and the last FRI layer's
// line interpolation.
#pragma once
#include "prover.h"

namespace sh {

struct MerkleDec {
  std::vector<Hash> hash_witness;
  std::vector<uint32_t> column_witness;
};
struct FriLayer {
  std::vector<qm31> witness;
  MerkleDec dec;
  Hash commitment;
};

std::vector<size_t> sorted_unique(std::vector<size_t> v);
// Positions of the next FRI layer: each position shifted down by n, consecutive duplicates dropped.
std::vector<size_t> fold_queries(const std::vector<size_t>& q, uint32_t n);
// core/pcs/utils.rs prepare_preprocessed_query_positions
std::vector<size_t> tree_local_positions(const std::vector<size_t>& q, uint32_t max_log, uint32_t tree_max);
// Columns in the order MerkleProver hashes them: stable by decreasing size.
std::vector<size_t> merkle_order(const std::vector<uint32_t>& logs);

qm31 qm_pow_u(qm31 b, uint64_t e);
// Lexicographic order of the coordinates (the order of compute_fri_quotients' sample points).
bool point_less(const cpqm31& a, const cpqm31& b);
bool point_eq(const cpqm31& a, const cpqm31& b);
cpqm31 cpq_repeated_double(cpqm31 p, uint32_t n);
// SecureField::from_partial_evals: e0 + e1 i + e2 u + e3 iu.
qm31 from_partial_evals(qm31 e0, qm31 e1, qm31 e2, qm31 e3);

// Sampled values of every committed column, flattened. Column k of tree t is global column
// col_base[t] + k; its samples are [start[g], start[g + 1]) of pid (ids into pts) and val.
struct SampleSet {
  std::vector<uint32_t> col_base{0};
  std::vector<uint32_t> start{0};
  std::vector<uint32_t> pid;
  std::vector<qm31> val;
  std::vector<cpqm31> pts;
  uint32_t n_trees() const { return (uint32_t)col_base.size() - 1; }
  uint32_t n_cols(size_t t) const { return col_base[t + 1] - col_base[t]; }
  uint32_t first(size_t t, size_t c) const { return start[col_base[t] + c]; }
  uint32_t count(size_t t, size_t c) const { return start[col_base[t] + c + 1] - start[col_base[t] + c]; }
  const qm31* vals(size_t t, size_t c) const { return val.data() + first(t, c); }
};

// Host point evaluation of a component's constraints from the sampled values.
struct PointEval {
  const SampleSet* sv;
  const Component* c;
  size_t ti = 0, ii = 0;
  qm31* acc;
  qm31 rc, dinv;
  qm31 next_trace() { return sv->vals(1, c->trace_start + ti++)[0]; }
  void add(qm31 x) { *acc = qm_add(qm_mul(*acc, rc), qm_mul(dinv, x)); }
  std::vector<qm31> next_ext(size_t n_off) {
    const qm31* m[4];
    for (int k = 0; k < 4; ++k) m[k] = sv->vals(2, c->inter_start + ii++);
    std::vector<qm31> r;
    for (size_t o = 0; o < n_off; ++o) r.push_back(from_partial_evals(m[0][o], m[1][o], m[2][o], m[3][o]));
    return r;
  }
};
void eval_at_point(PointEval& e);

// Mask offsets of column i of tree t (1 = trace, 2 = interaction) of a component: the rows,
// relative to the current one, at which its constraints read the column. Returns the count.
uint32_t mask_offsets(const Component& c, size_t t, size_t i, long* offs);

struct ProofParts {
  std::vector<Hash> commitments;
  SampleSet sampled;
  std::vector<MerkleDec> decs;
  std::vector<std::vector<uint32_t>> queried;
  uint64_t pow = 0;
  FriLayer first;
  std::vector<FriLayer> inner;
  std::vector<qm31> last_poly;
};
// The library's canonical StarkProof encoding: little endian, counts as u32.
std::vector<uint8_t> encode(const PcsCfg& cfg, const ProofParts& p);

// Host inverse line FFT of the last FRI layer (prover/line.rs:44-64); returns stored-order coeffs.
std::vector<qm31> line_interpolate(std::vector<qm31> v, HCoset dom);

// core/verifier.rs verify over the library's proof encoding, for the given components (trace and
// interaction offsets are assigned here, in order) and the expected configuration: "" when the
// proof is accepted, else the reason (verifier.cpp); mc names the proof's Merkle channel.
std::string verify_proof(const uint8_t* bytes, size_t len, const std::vector<Component>& comps, const PcsCfg& cfg,
                         MerkleKind mc = MerkleKind::Blake2s);

}  // namespace sh
