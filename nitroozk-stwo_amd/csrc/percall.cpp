// prove() driven through the trait entry points of the C ABI, one call per trait-method
// invocation: the sequence of calls that the reference's generic prover makes on a backend
// (paths relative to crates/stwo/src):
//   prove / prove_ex                        prover/mod.rs:28-150
//   CommitmentSchemeProver commit           prover/pcs/mod.rs:58-168 (TreeBuilder, evaluate, MerkleProver)
//   prove_values                            prover/pcs/mod.rs:170-357
//   compute_fri_quotients                   prover/pcs/quotient_ops.rs:73-145
//   FriProver commit / decommit             prover/fri.rs:100-412
//   MerkleProver commit / decommit          prover/vcs/prover.rs:39-217
//   composition                             prover/air/component_prover.rs:99-114, air/accumulation.rs:21-125
// This is the host side a Rust HipBackend would run above the C ABI, restated in C++ because the
// image has no Rust toolchain. Every device step is a stwo_hip_* call from include/stwo_hip.h;
// nothing here launches a kernel or touches the context's internals. It is the "plain per-call"
// form of the prover: calls are stream-ordered on one stream, as a backend's trait calls are on
// the reference's CUDA stream, and the host waits wherever it reads device data (every Merkle root
// for the channel, the sampled values, the decommitment's values, the nonce); columns are allocated
// per call. The device-resident prover (prover.cpp) is the fused form of the same
// protocol; both produce the same bytes. Like the reference's prove(), it is generic over the
// MerkleChannel: Blake2s, or Poseidon252 (felt252 hashes through commit_on_layer_poseidon252 and
// grind_poseidon252, the felt252 channel of p252_channel.cpp).
#include "../../include/stwo_hip.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <tuple>

#include "protocol.h"

namespace sh {
namespace percall {
namespace {

struct Api {
  stwo_hip_ctx* h;
  void ck(int st, const char* what) const {
    if (st != 0) throw HipError(st, std::string(what) + ": " + stwo_hip_last_error(h));
  }
};

// A device column owned by this prover (Column::zeros / uninitialized, freed on drop).
struct Col {
  const Api* api = nullptr;
  uint32_t* p = nullptr;
  size_t n = 0;
  Col() = default;
  Col(const Api& a, size_t words, bool zero) : api(&a), n(words) {
    a.ck(stwo_hip_alloc(a.h, words, zero ? 1 : 0, &p), "alloc");
  }
  Col(const Col&) = delete;
  Col& operator=(const Col&) = delete;
  Col(Col&& o) noexcept { *this = std::move(o); }
  Col& operator=(Col&& o) noexcept {
    std::swap(api, o.api);
    std::swap(p, o.p);
    std::swap(n, o.n);
    return *this;
  }
  ~Col() {
    if (p && api) (void)stwo_hip_free(api->h, p);
  }
};

// The MerkleChannel the generic prover is instantiated with (BackendForChannel<MC>,
// prover/backend/mod.rs:34-37): its channel type and the backend's MerkleOps<MC::H>::commit_on_layer
// and GrindOps<MC::C>::grind entry points.
struct Blake2sMC {
  using Ch = Channel;
  static int commit_on_layer(stwo_hip_ctx* h, uint32_t ls, const uint32_t* prev, const uint32_t* const* cols,
                             uint32_t n, uint32_t* out) {
    return stwo_hip_commit_on_layer(h, ls, prev, cols, n, out);
  }
  static int grind(stwo_hip_ctx* h, const Ch& ch, uint32_t bits, uint64_t* nonce) {
    return stwo_hip_grind(h, ch.digest.data(), bits, nonce);
  }
};
struct Poseidon252MC {
  using Ch = P252Channel;
  static int commit_on_layer(stwo_hip_ctx* h, uint32_t ls, const uint32_t* prev, const uint32_t* const* cols,
                             uint32_t n, uint32_t* out) {
    return stwo_hip_commit_on_layer_poseidon252(h, ls, prev, cols, n, out);
  }
  static int grind(stwo_hip_ctx* h, const Ch& ch, uint32_t bits, uint64_t* nonce) {
    const auto w = ch.digest_words();
    return stwo_hip_grind_poseidon252(h, w.data(), bits, nonce);
  }
};

stwo_qm31 wq(qm31 v) { return {v.a, v.b, v.c, v.d}; }
qm31 rq(stwo_qm31 v) { return {v.a, v.b, v.c, v.d}; }
stwo_circle_point_qm31 wp(const cpqm31& p) { return {wq(p.x), wq(p.y)}; }

// A committed tree: coefficient columns (owned, or the caller's trace is only evaluated), their
// evaluations with blowup and the Merkle layers (layers[k]: 2^k nodes).
struct Tree {
  std::vector<Col> coeff_own;
  std::vector<const uint32_t*> coeffs;
  std::vector<uint32_t> logs;  // coefficient log sizes
  std::vector<Col> evals;
  std::vector<uint32_t> eval_logs;
  std::vector<Col> layers;
  Hash root{};
  uint32_t max_log() const { return (uint32_t)layers.size() - 1; }
};

Hash read_root(const Api& a, const uint32_t* layer0) {
  uint32_t w[8];
  a.ck(stwo_hip_copy_d2h(a.h, w, layer0, 8), "copy_d2h");
  Hash r;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 4; ++j) r[4 * i + j] = (uint8_t)(w[i] >> (8 * j));
  return r;
}

// PolyOps::interpolate_columns, one call per column size (as CudaBackend groups them).
void interpolate_into(const Api& a, const std::vector<const uint32_t*>& evals, const std::vector<uint32_t>& logs,
                      Tree& t) {
  std::map<uint32_t, std::vector<size_t>> by;
  for (size_t i = 0; i < evals.size(); ++i) by[logs[i]].push_back(i);
  t.coeffs.assign(evals.size(), nullptr);
  t.logs = logs;
  t.coeff_own.resize(evals.size());
  for (auto& [lg, idx] : by) {
    std::vector<const uint32_t*> src;
    std::vector<uint32_t*> dst;
    for (size_t i : idx) {
      t.coeff_own[i] = Col(a, (size_t)1 << lg, false);
      t.coeffs[i] = t.coeff_own[i].p;
      src.push_back(evals[i]);
      dst.push_back(t.coeff_own[i].p);
    }
    a.ck(stwo_hip_interpolate_columns(a.h, src.data(), dst.data(), (uint32_t)idx.size(), lg), "interpolate_columns");
  }
}

// MerkleProver::commit: columns by decreasing size, one commit_on_layer per layer.
template <class MC>
void merkle_commit(const Api& a, Tree& t) {
  t.layers.clear();
  if (t.evals.empty()) {
    t.layers.emplace_back(a, 8, false);
    a.ck(MC::commit_on_layer(a.h, 0, nullptr, nullptr, 0, t.layers[0].p), "commit_on_layer");
    t.root = read_root(a, t.layers[0].p);
    return;
  }
  const auto order = merkle_order(t.eval_logs);
  const uint32_t top = t.eval_logs[order[0]];
  std::vector<Col> rev;
  size_t ci = 0;
  for (int ls = (int)top; ls >= 0; --ls) {
    std::vector<const uint32_t*> lc;
    while (ci < order.size() && t.eval_logs[order[ci]] == (uint32_t)ls) lc.push_back(t.evals[order[ci++]].p);
    Col layer(a, (size_t)8 << ls, false);
    a.ck(MC::commit_on_layer(a.h, (uint32_t)ls, rev.empty() ? nullptr : rev.back().p, lc.data(),
                             (uint32_t)lc.size(), layer.p),
         "commit_on_layer");
    rev.push_back(std::move(layer));
  }
  for (size_t i = rev.size(); i-- > 0;) t.layers.push_back(std::move(rev[i]));
  t.root = read_root(a, t.layers[0].p);
}

// CommitmentSchemeProver::commit of a tree given by coefficients: evaluate_polynomials with the
// blowup (one call per size), the Merkle tree, and the root into the channel.
template <class MC>
void commit(const Api& a, Tree& t, const PcsCfg& cfg, typename MC::Ch& ch) {
  std::map<uint32_t, std::vector<size_t>> by;
  for (size_t i = 0; i < t.coeffs.size(); ++i) by[t.logs[i]].push_back(i);
  t.evals.resize(t.coeffs.size());
  t.eval_logs.assign(t.coeffs.size(), 0);
  for (auto& [lg, idx] : by) {
    std::vector<const uint32_t*> src;
    std::vector<uint32_t*> dst;
    for (size_t i : idx) {
      t.evals[i] = Col(a, (size_t)1 << (lg + cfg.log_blowup), false);
      t.eval_logs[i] = lg + cfg.log_blowup;
      src.push_back(t.coeffs[i]);
      dst.push_back(t.evals[i].p);
    }
    a.ck(stwo_hip_evaluate_polynomials(a.h, src.data(), dst.data(), (uint32_t)idx.size(), lg, cfg.log_blowup),
         "evaluate_polynomials");
  }
  merkle_commit<MC>(a, t);
  ch.mix_root(t.root);
}

// MerkleProver::decommit (prover/vcs/prover.rs:84-217): the walk from the largest layer down,
// child hashes and column values fetched with Column::batch_at_multi.
void merkle_decommit(const Api& a, const Tree& t, const std::map<uint32_t, std::vector<size_t>>& qpl,
                     std::vector<uint32_t>* queried, MerkleDec& dec) {
  const auto order = merkle_order(t.eval_logs);
  size_t ci = 0;
  std::vector<size_t> last;
  static const std::vector<size_t> none;
  for (int ls = (int)t.max_log(); ls >= 0; --ls) {
    auto it = qpl.find((uint32_t)ls);
    const std::vector<size_t>& colq = it == qpl.end() ? none : it->second;
    std::vector<size_t> merged;
    for (size_t q : last) merged.push_back(q / 2);
    merged.insert(merged.end(), colq.begin(), colq.end());
    merged = sorted_unique(merged);
    std::vector<const uint32_t*> lc;
    while (ci < order.size() && t.eval_logs[order[ci]] == (uint32_t)ls) lc.push_back(t.evals[order[ci++]].p);
    if ((uint32_t)ls < t.max_log() && !merged.empty()) {
      // children of every merged node, 8 words each: the layer read as 8 interleaved columns
      const uint32_t* lay = t.layers[ls + 1].p;
      const uint32_t* cols8[8];
      for (int k = 0; k < 8; ++k) cols8[k] = lay + k;
      std::vector<uint32_t> rows;
      for (size_t n : merged) { rows.push_back((uint32_t)(16 * n)); rows.push_back((uint32_t)(16 * n + 8)); }
      std::vector<uint32_t> w(rows.size() * 8);
      a.ck(stwo_hip_batch_at_multi(a.h, cols8, 8, rows.data(), (uint32_t)rows.size(), w.data()), "batch_at_multi");
      for (size_t i = 0; i < merged.size(); ++i)
        for (size_t c = 0; c < 2; ++c) {
          const size_t child = 2 * merged[i] + c;
          if (std::binary_search(last.begin(), last.end(), child)) continue;
          Hash hsh;
          for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 4; ++j) hsh[4 * k + j] = (uint8_t)(w[(2 * i + c) * 8 + k] >> (8 * j));
          dec.hash_witness.push_back(hsh);
        }
    }
    if (!lc.empty()) {
      std::vector<uint32_t> qn, wn;
      for (size_t n : merged) (std::binary_search(colq.begin(), colq.end(), n) ? qn : wn).push_back((uint32_t)n);
      if (!qn.empty()) {
        std::vector<uint32_t> v(qn.size() * lc.size());
        a.ck(stwo_hip_batch_at_multi(a.h, lc.data(), (uint32_t)lc.size(), qn.data(), (uint32_t)qn.size(), v.data()),
             "batch_at_multi");
        if (queried) queried->insert(queried->end(), v.begin(), v.end());
      }
      if (!wn.empty()) {
        std::vector<uint32_t> v(wn.size() * lc.size());
        a.ck(stwo_hip_batch_at_multi(a.h, lc.data(), (uint32_t)lc.size(), wn.data(), (uint32_t)wn.size(), v.data()),
             "batch_at_multi");
        dec.column_witness.insert(dec.column_witness.end(), v.begin(), v.end());
      }
    }
    last = merged;
  }
}

// A FRI layer: 4 coordinate columns of 2^log and its tree; fold_step of the fold that follows.
struct FriCol {
  Tree tree;
  uint32_t log = 0, fold_step = 1;
  const uint32_t* c(int k) const { return tree.evals[k].p; }
};
FriCol fri_col(const Api& a, uint32_t log, bool zero) {
  FriCol f;
  f.log = log;
  for (int k = 0; k < 4; ++k) f.tree.evals.emplace_back(a, (size_t)1 << log, zero);
  f.tree.eval_logs.assign(4, log);
  return f;
}

// FriProver::decommit of one layer (prover/fri.rs:291-412): the values of every queried fold coset
// that were not queried, then the Merkle decommitment of the coset positions.
FriLayer fri_decommit_layer(const Api& a, const FriCol& f, const std::vector<size_t>& q, uint32_t w) {
  FriLayer out;
  std::vector<size_t> pos;
  std::vector<uint32_t> wit;
  for (size_t i = 0; i < q.size();) {
    size_t j = i;
    while (j < q.size() && (q[j] >> w) == (q[i] >> w)) ++j;
    const size_t start = (q[i] >> w) << w;
    for (size_t p = start; p < start + ((size_t)1 << w); ++p) {
      pos.push_back(p);
      if (!std::binary_search(q.begin() + i, q.begin() + j, p)) wit.push_back((uint32_t)p);
    }
    i = j;
  }
  if (!wit.empty()) {
    const uint32_t* cols[4] = {f.c(0), f.c(1), f.c(2), f.c(3)};
    std::vector<uint32_t> v(wit.size() * 4);
    a.ck(stwo_hip_batch_at_multi(a.h, cols, 4, wit.data(), (uint32_t)wit.size(), v.data()), "batch_at_multi");
    for (size_t i = 0; i < wit.size(); ++i) out.witness.push_back({v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]});
  }
  merkle_decommit(a, f.tree, {{f.log, pos}}, nullptr, out.dec);
  out.commitment = f.tree.root;
  return out;
}

uint32_t fnv1a(const char* s) {
  uint32_t h = 0x811C9DC5u;
  for (; *s; ++s) { h ^= (uint8_t)*s; h *= 0x01000193u; }
  return h;
}

// prove_ex over committed trees (0 preprocessed, 1 trace, [2 interaction]); appends the
// composition tree. Returns the sanity-check status (0, or 2 for ConstraintsNotSatisfied).
template <class MC>
int prove_ex(const Api& a, typename MC::Ch& ch, std::vector<Tree>& trees, const std::vector<Component>& comps,
             const PcsCfg& cfg, ProofParts& out) {
  // ---- composition (ComponentProvers::compute_composition_polynomial) ----
  const qm31 random_coeff = ch.draw_secure_felt();
  size_t T = 0;
  uint32_t max_log = 0;
  for (auto& c : comps) { T += c.n_constraints(); max_log = std::max(max_log, c.eval_log()); }
  std::vector<stwo_qm31> pw(T);
  a.ck(stwo_hip_generate_secure_powers(a.h, wq(random_coeff), (uint32_t)T, pw.data()), "generate_secure_powers");
  std::vector<qm31> powers(T);
  for (size_t i = 0; i < T; ++i) powers[i] = rq(pw[i]);
  std::map<uint32_t, std::array<Col, 4>> accs;  // DomainEvaluationAccumulator: one per log size
  for (auto& c : comps) {
    const uint32_t E = c.eval_log();
    if (!accs.count(E)) {
      auto& acc = accs[E];
      for (int k = 0; k < 4; ++k) acc[k] = Col(a, (size_t)1 << E, true);
    }
  }
  for (auto& c : comps) {
    const uint32_t E = c.eval_log(), nc = c.n_constraints();
    std::vector<stwo_qm31> cp;
    for (size_t i = 0; i < nc; ++i) cp.push_back(wq(powers[powers.size() - 1 - i]));
    powers.resize(powers.size() - nc);
    // the component's columns on its evaluation domain: the committed evaluations when their
    // size is the domain's, else evaluated there from the coefficients
    std::vector<Col> tmp;
    auto on_domain = [&](const Tree& t, size_t start, size_t n) {
      std::vector<const uint32_t*> r(n);
      std::vector<const uint32_t*> src;
      std::vector<uint32_t*> dst;
      uint32_t lg = 0;
      for (size_t i = 0; i < n; ++i) {
        if (t.eval_logs[start + i] == E) { r[i] = t.evals[start + i].p; continue; }
        tmp.emplace_back(a, (size_t)1 << E, false);
        r[i] = tmp.back().p;
        src.push_back(t.coeffs[start + i]);
        dst.push_back(tmp.back().p);
        lg = t.logs[start + i];
      }
      if (!src.empty())
        a.ck(stwo_hip_evaluate_polynomials(a.h, src.data(), dst.data(), (uint32_t)src.size(), lg, E - lg),
             "evaluate_polynomials");
      return r;
    };
    const auto tc = on_domain(trees[1], c.trace_start, c.n_trace_cols());
    std::vector<const uint32_t*> ic;
    if (c.n_inter_cols()) ic = on_domain(trees[2], c.inter_start, c.n_inter_cols());
    // vanishing-polynomial inverses on the 2^(E - log_n) cosets, bit-reversed
    const uint32_t le = E - c.log_n;
    std::vector<uint32_t> dinv(1u << le);
    for (uint32_t i = 0; i < (1u << le); ++i)
      dinv[bitrev(i, le)] = m31_inv(coset_vanishing_m31(c.log_n, canonic_domain_at(E, i)));
    std::vector<uint8_t> blob(4);
    uint32_t id = fnv1a(c.kind == AIR_WIDE_FIB ? "fibonacci_example" : "poseidon_example");
    std::memcpy(blob.data(), &id, 4);
    if (c.kind == AIR_POSEIDON) {
      blob.resize(4 + 4 + 32);
      std::memcpy(blob.data() + 4, &c.ipr, 4);
      const stwo_qm31 z = wq(c.z), al = wq(c.alpha);
      std::memcpy(blob.data() + 8, &z, 16);
      std::memcpy(blob.data() + 24, &al, 16);
    }
    auto& acc = accs[E];
    uint32_t* a4[4] = {acc[0].p, acc[1].p, acc[2].p, acc[3].p};
    a.ck(stwo_hip_evaluate_constraint_quotients_on_domain(a.h, a4, tc.data(), (uint32_t)tc.size(), ic.data(),
                                                          (uint32_t)ic.size(), cp.data(), nc, dinv.data(), c.log_n, E,
                                                          blob.data(), blob.size(), wq(c.cumsum_shift())),
         "evaluate_constraint_quotients_on_domain");
  }
  // finalize: lift_and_accumulate in increasing size, then 4 interpolations; split_at_mid
  std::array<Col, 4> sum;
  const uint32_t* fin[4];
  if (accs.size() == 1) {
    for (int k = 0; k < 4; ++k) fin[k] = accs.begin()->second[k].p;
  } else {
    std::vector<const uint32_t*> c4;
    std::vector<uint32_t> ls;
    for (auto& [lg, acc] : accs) {
      for (int k = 0; k < 4; ++k) c4.push_back(acc[k].p);
      ls.push_back(lg);
    }
    for (int k = 0; k < 4; ++k) sum[k] = Col(a, (size_t)1 << max_log, false);
    uint32_t* o4[4] = {sum[0].p, sum[1].p, sum[2].p, sum[3].p};
    a.ck(stwo_hip_lift_and_accumulate(a.h, c4.data(), ls.data(), (uint32_t)ls.size(), o4), "lift_and_accumulate");
    for (int k = 0; k < 4; ++k) fin[k] = sum[k].p;
  }
  Tree comp;
  {
    Tree full;
    interpolate_into(a, std::vector<const uint32_t*>(fin, fin + 4), std::vector<uint32_t>(4, max_log), full);
    accs.clear();
    std::vector<const uint32_t*> halves(8);
    for (int k = 0; k < 4; ++k) {
      uint32_t *lft = nullptr, *rgt = nullptr;
      a.ck(stwo_hip_split_at_mid(a.h, full.coeff_own[k].p, (size_t)1 << max_log, &lft, &rgt), "split_at_mid");
      halves[k] = lft;
      halves[4 + k] = rgt;
    }
    comp.coeff_own = std::move(full.coeff_own);  // the views stay valid while the columns live
    comp.coeffs = halves;
    comp.logs.assign(8, max_log - 1);
  }
  commit<MC>(a, comp, cfg, ch);
  trees.push_back(std::move(comp));

  // ---- OODS point and prove_values ----
  cpqm31 oods;
  {
    const qm31 t = ch.draw_secure_felt(), t2 = qm_sqr(t), d = qm_inv(qm_add(t2, qm_one()));
    oods = {qm_mul(qm_sub(qm_one(), t2), d), qm_mul(qm_add(t, t), d)};
  }
  const uint32_t split_log = trees.back().max_log();
  const uint32_t L = cfg.lifting_log_size ? cfg.lifting_log_size : split_log;
  if (L < split_log) throw HipError(3, "lifting log size below the composition size");
  const uint32_t max_deg = L - cfg.log_blowup;
  SampleSet& S = out.sampled;
  for (size_t t = 0; t < trees.size(); ++t) S.col_base.push_back(S.col_base.back() + (uint32_t)trees[t].evals.size());
  const uint32_t NC = S.col_base.back();
  const cpm31 step = index_to_point(subgroup_gen(max_deg));
  std::vector<long> pt_off;
  auto pid_of = [&](long o) -> uint32_t {
    for (size_t i = 0; i < pt_off.size(); ++i)
      if (pt_off[i] == o) return (uint32_t)i;
    pt_off.push_back(o);
    return (uint32_t)pt_off.size() - 1;
  };
  std::vector<std::vector<uint32_t>> cpid(NC);
  for (auto& c : comps)
    for (size_t t = 1; t < 3 && t + 1 < trees.size(); ++t) {
      const size_t start = t == 1 ? c.trace_start : c.inter_start, n = t == 1 ? c.n_trace_cols() : c.n_inter_cols();
      for (size_t i = 0; i < n; ++i) {
        long offs[2];
        const uint32_t k = mask_offsets(c, t, i, offs);
        for (uint32_t j = 0; j < k; ++j) cpid[S.col_base[t] + start + i].push_back(pid_of(offs[j]));
      }
    }
  for (uint32_t g = S.col_base[trees.size() - 1]; g < NC; ++g) cpid[g].assign(1, pid_of(0));
  S.start.assign(NC + 1, 0);
  for (uint32_t g = 0; g < NC; ++g) S.start[g + 1] = S.start[g] + (uint32_t)cpid[g].size();
  for (uint32_t g = 0; g < NC; ++g) S.pid.insert(S.pid.end(), cpid[g].begin(), cpid[g].end());
  for (long o : pt_off) S.pts.push_back(cpq_add(oods, cpq_mul_signed(step, o)));
  const uint32_t Lc = trees.back().max_log();
  S.val.resize(S.pid.size());
  {
    // batch_eval_at_point per (coefficient size, point lifted to the column's size)
    struct Key { uint32_t k, pid, shift; bool operator<(const Key& o) const { return std::tie(k, pid, shift) < std::tie(o.k, o.pid, o.shift); } };
    std::map<Key, std::vector<std::pair<const uint32_t*, uint32_t>>> groups;
    for (size_t t = 0; t < trees.size(); ++t)
      for (size_t c = 0; c < trees[t].evals.size(); ++c)
        for (uint32_t j = 0; j < S.count(t, c); ++j) {
          const uint32_t at = S.first(t, c) + j;
          groups[{trees[t].logs[c], S.pid[at], Lc - trees[t].eval_logs[c]}].push_back({trees[t].coeffs[c], at});
        }
    for (auto& [key, v] : groups) {
      const cpqm31 p = cpq_repeated_double(S.pts[key.pid], key.shift);
      std::vector<const uint32_t*> cols;
      for (auto& e : v) cols.push_back(e.first);
      std::vector<stwo_qm31> r(cols.size());
      a.ck(stwo_hip_batch_eval_at_point(a.h, cols.data(), (uint32_t)cols.size(), key.k, wp(p), r.data()),
           "batch_eval_at_point");
      for (size_t i = 0; i < v.size(); ++i) S.val[v[i].second] = rq(r[i]);
    }
  }
  ch.mix_felts(S.val);

  // ---- FRI quotients (compute_fri_quotients) ----
  const qm31 qrc = ch.draw_secure_felt();
  // samples per column: a periodicity sample first for two-sample columns, consecutive powers
  std::vector<cpqm31> qpts = S.pts;
  std::vector<std::array<uint32_t, 3>> periodic;  // (point id, eval log, new point id)
  std::vector<uint32_t> qstart(NC + 1, 0), qpid, qvi;
  for (size_t t = 0; t < trees.size(); ++t)
    for (size_t c = 0; c < trees[t].evals.size(); ++c) {
      const uint32_t g = S.col_base[t] + (uint32_t)c, n = S.count(t, c), f0 = S.first(t, c);
      qstart[g] = (uint32_t)qpid.size();
      if (n == 2) {
        const uint32_t el = trees[t].eval_logs[c], p1 = S.pid[f0 + 1];
        uint32_t np = UINT32_MAX;
        for (auto& e : periodic)
          if (e[0] == p1 && e[1] == el) np = e[2];
        if (np == UINT32_MAX) {
          np = (uint32_t)qpts.size();
          cpm31 pg = index_to_point(subgroup_gen(Lc));
          for (uint32_t i = 0; i < el; ++i) pg = cp_double(pg);
          qpts.push_back(cpq_add(S.pts[p1], cpq_from(pg)));
          periodic.push_back({p1, el, np});
        }
        qpid.push_back(np);
        qvi.push_back(f0 + 1);
      }
      for (uint32_t j = 0; j < n; ++j) { qpid.push_back(S.pid[f0 + j]); qvi.push_back(f0 + j); }
    }
  qstart[NC] = (uint32_t)qpid.size();
  std::vector<stwo_qm31> rpow(qpid.size());
  a.ck(stwo_hip_generate_secure_powers(a.h, wq(qrc), (uint32_t)qpid.size(), rpow.data()), "generate_secure_powers");
  // per column size group (increasing), a ColumnSampleBatch per point in order of first use
  struct Item { size_t t, c; uint32_t ls; };
  std::vector<Item> items;
  for (size_t t = 0; t < trees.size(); ++t)
    for (size_t c = 0; c < trees[t].evals.size(); ++c) items.push_back({t, c, trees[t].eval_logs[c]});
  std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.ls < y.ls; });
  std::vector<Col> partials;
  std::vector<stwo_accumulated_numerators> accn;
  for (size_t i = 0; i < items.size();) {
    size_t j = i;
    while (j < items.size() && items[j].ls == items[i].ls) ++j;
    std::vector<const uint32_t*> gcols;
    std::vector<uint32_t> bpid;
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> per;  // (column index, sample index)
    for (size_t k = i; k < j; ++k) {
      const uint32_t g = S.col_base[items[k].t] + (uint32_t)items[k].c;
      gcols.push_back(trees[items[k].t].evals[items[k].c].p);
      for (uint32_t q = qstart[g]; q < qstart[g + 1]; ++q) {
        size_t b = std::find(bpid.begin(), bpid.end(), qpid[q]) - bpid.begin();
        if (b == bpid.size()) { bpid.push_back(qpid[q]); per.emplace_back(); }
        per[b].push_back({(uint32_t)(k - i), q});
      }
    }
    const uint32_t ls = items[i].ls, nb = (uint32_t)bpid.size();
    std::vector<stwo_sample_batch> batches(nb);
    std::vector<std::vector<uint32_t>> idx(nb);
    std::vector<std::vector<stwo_qm31>> vals(nb), rnds(nb);
    for (uint32_t b = 0; b < nb; ++b) {
      for (auto& [ci, q] : per[b]) {
        idx[b].push_back(ci);
        vals[b].push_back(wq(S.val[qvi[q]]));
        rnds[b].push_back(rpow[q]);
      }
      batches[b] = {wp(qpts[bpid[b]]), (uint32_t)idx[b].size(), idx[b].data(), vals[b].data(), rnds[b].data()};
    }
    std::vector<uint32_t*> outs;
    const size_t p0 = partials.size();
    for (uint32_t b = 0; b < 4 * nb; ++b) {
      partials.emplace_back(a, (size_t)1 << ls, false);
      outs.push_back(partials.back().p);
    }
    std::vector<stwo_qm31> fl(nb);
    if (nb)
      a.ck(stwo_hip_accumulate_numerators(a.h, ls, gcols.data(), (uint32_t)gcols.size(), batches.data(), nb,
                                          outs.data(), fl.data()),
           "accumulate_numerators");
    for (uint32_t b = 0; b < nb; ++b) {
      stwo_accumulated_numerators an{};
      an.sample_point = wp(qpts[bpid[b]]);
      for (int k = 0; k < 4; ++k) an.partial[k] = partials[p0 + 4 * b + k].p;
      an.log_size = ls;
      an.first_linear_term = fl[b];
      accn.push_back(an);
    }
    i = j;
  }
  FriCol first = fri_col(a, Lc, false);
  {
    uint32_t* o4[4] = {first.tree.evals[0].p, first.tree.evals[1].p, first.tree.evals[2].p, first.tree.evals[3].p};
    a.ck(stwo_hip_compute_quotients_and_combine(a.h, accn.data(), (uint32_t)accn.size(), Lc, o4),
         "compute_quotients_and_combine");
  }
  partials.clear();

  // ---- FRI commit ----
  const uint32_t last_log = cfg.log_last_layer + cfg.log_blowup;
  if (Lc - 1 < last_log) throw HipError(3, "FRI: last layer larger than the first folded layer");
  merkle_commit<MC>(a, first.tree);
  ch.mix_root(first.tree.root);
  qm31 alpha = ch.draw_secure_felt();
  FriCol cur = fri_col(a, Lc - 1, true);
  {
    const uint32_t* s4[4] = {first.c(0), first.c(1), first.c(2), first.c(3)};
    uint32_t* d4[4] = {cur.tree.evals[0].p, cur.tree.evals[1].p, cur.tree.evals[2].p, cur.tree.evals[3].p};
    a.ck(stwo_hip_fold_circle_into_line(a.h, Lc, s4, d4, wq(alpha)), "fold_circle_into_line");
  }
  std::vector<FriCol> inner;
  while (cur.log > last_log) {
    const uint32_t fs = cur.log > last_log + cfg.fold_step ? cfg.fold_step : cur.log - last_log;
    cur.fold_step = fs;
    merkle_commit<MC>(a, cur.tree);
    ch.mix_root(cur.tree.root);
    alpha = ch.draw_secure_felt();
    FriCol nx = fri_col(a, cur.log - fs, false);
    const uint32_t* s4[4] = {cur.c(0), cur.c(1), cur.c(2), cur.c(3)};
    uint32_t* d4[4] = {nx.tree.evals[0].p, nx.tree.evals[1].p, nx.tree.evals[2].p, nx.tree.evals[3].p};
    a.ck(stwo_hip_fold_line(a.h, cur.log, s4, d4, wq(alpha), fs), "fold_line");
    inner.push_back(std::move(cur));
    cur = std::move(nx);
  }
  {
    // the last layer to the host, interpolated on its line domain; the tail must vanish
    const size_t n = (size_t)1 << cur.log;
    std::vector<uint32_t> w(4 * n);
    for (int k = 0; k < 4; ++k) a.ck(stwo_hip_copy_d2h(a.h, w.data() + k * n, cur.c(k), n), "copy_d2h");
    std::vector<qm31> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = {w[i], w[n + i], w[2 * n + i], w[3 * n + i]};
    HCoset dom = HCoset::half_odds(Lc - 1);
    for (uint32_t i = Lc - 1; i > cur.log; --i) dom = dom.doubled();
    std::vector<qm31> ordered = line_interpolate(v, dom);
    for (size_t i = 0; i < n; ++i) { size_t j = bitrev((uint32_t)i, cur.log); if (j > i) std::swap(ordered[i], ordered[j]); }
    const size_t bound = (size_t)1 << cfg.log_last_layer;
    for (size_t i = bound; i < n; ++i)
      if (!qm_eq(ordered[i], qm_zero())) throw HipError(4, "invalid degree");
    ordered.resize(bound);
    for (size_t i = 0; i < bound; ++i) {
      size_t j = bitrev((uint32_t)i, cfg.log_last_layer);
      if (j > i) std::swap(ordered[i], ordered[j]);
    }
    out.last_poly = ordered;
    ch.mix_felts(out.last_poly);
  }

  // ---- proof of work (GrindOps::grind) ----
  a.ck(MC::grind(a.h, ch, cfg.pow_bits, &out.pow), "grind");
  ch.mix_u64(out.pow);

  // ---- queries and decommitments ----
  std::vector<size_t> raw;
  {
    const uint32_t mask = (uint32_t)(((uint64_t)1 << Lc) - 1);
    while (raw.size() < cfg.n_queries)
      for (auto x : ch.draw_u32s())
        if (raw.size() < cfg.n_queries) raw.push_back(x & mask);
  }
  const auto qpos = sorted_unique(raw);
  out.first = fri_decommit_layer(a, first, qpos, 1);
  {
    std::vector<size_t> lq = fold_queries(qpos, 1);
    for (auto& l : inner) {
      out.inner.push_back(fri_decommit_layer(a, l, lq, l.fold_step));
      lq = fold_queries(lq, l.fold_step);
    }
  }
  for (auto& t : trees) {
    // prepare_preprocessed_query_positions, then each column size's queried nodes
    const uint32_t top = t.max_log();
    const auto P = tree_local_positions(qpos, Lc, top);
    std::map<uint32_t, std::vector<size_t>> qpl;
    for (auto ls : std::set<uint32_t>(t.eval_logs.begin(), t.eval_logs.end())) {
      std::vector<size_t> v;
      for (size_t p : P) v.push_back(((p >> (top - ls + 1)) << 1) + (p & 1));
      qpl[ls] = sorted_unique(v);
    }
    MerkleDec d;
    std::vector<uint32_t> queried;
    if (top > 0) merkle_decommit(a, t, qpl, &queried, d);
    out.queried.push_back(std::move(queried));
    out.decs.push_back(std::move(d));
    out.commitments.push_back(t.root);
  }

  // ---- sanity check (prover/mod.rs:127-144) ----
  const size_t lt = S.n_trees() - 1;
  const qm31 lft = from_partial_evals(S.vals(lt, 0)[0], S.vals(lt, 1)[0], S.vals(lt, 2)[0], S.vals(lt, 3)[0]);
  const qm31 rgt = from_partial_evals(S.vals(lt, 4)[0], S.vals(lt, 5)[0], S.vals(lt, 6)[0], S.vals(lt, 7)[0]);
  const qm31 committed = qm_add(lft, qm_mul(cpq_repeated_double(oods, max_deg - 1).x, rgt));
  qm31 acc = qm_zero();
  for (auto& c : comps) {
    PointEval e;
    e.sv = &out.sampled;
    e.c = &c;
    e.acc = &acc;
    e.rc = random_coeff;
    e.dinv = qm_inv(coset_vanishing_qm31(max_deg, oods));
    eval_at_point(e);
  }
  return qm_eq(committed, acc) ? 0 : 2;
}

template <class MC>
ProveResult prove_wide_fib_multi_mc(stwo_hip_ctx* h, const std::vector<const uint32_t*>& trace,
                                    const std::vector<uint32_t>& logs, const std::vector<uint32_t>& ncols,
                                    const PcsCfg& cfg) {
  const Api a{h};
  ProveResult r;
  typename MC::Ch ch;
  std::vector<Tree> trees(2);
  commit<MC>(a, trees[0], cfg, ch);  // the preprocessed tree has no columns
  std::vector<uint32_t> col_logs;
  std::vector<Component> comps;
  size_t start = 0;
  for (size_t i = 0; i < logs.size(); ++i) {
    Component c;
    c.kind = AIR_WIDE_FIB;
    c.log_n = logs[i];
    c.n_cols = ncols[i];
    c.trace_start = start;
    start += ncols[i];
    col_logs.insert(col_logs.end(), ncols[i], logs[i]);
    comps.push_back(c);
  }
  if (start != trace.size()) throw HipError(1, "trace column count does not match the components");
  interpolate_into(a, trace, col_logs, trees[1]);
  commit<MC>(a, trees[1], cfg, ch);
  ProofParts parts;
  r.status = prove_ex<MC>(a, ch, trees, comps, cfg, parts);
  r.proof = encode(cfg, parts);
  r.digest = ch.digest;
  return r;
}

template <class MC>
ProveResult prove_poseidon_mc(stwo_hip_ctx* h, const std::vector<const uint32_t*>& trace, uint32_t log_n,
                              uint32_t ipr, const PcsCfg& cfg) {
  const Api a{h};
  ProveResult r;
  typename MC::Ch ch;
  std::vector<Tree> trees(2);
  commit<MC>(a, trees[0], cfg, ch);
  interpolate_into(a, trace, std::vector<uint32_t>(trace.size(), log_n), trees[1]);
  commit<MC>(a, trees[1], cfg, ch);
  Component c;
  c.kind = AIR_POSEIDON;
  c.log_n = log_n;
  c.ipr = ipr;
  const auto le = ch.draw_secure_felts(2);
  c.z = le[0];
  c.alpha = le[1];
  {
    qm31 p = qm_one();
    for (int i = 0; i < 16; ++i) { c.alpha_pow[i] = p; p = qm_mul(p, c.alpha); }
  }
  // the LogUp interaction trace (trace generation, outside the trait seam), then its commitment
  std::vector<Col> icols;
  std::vector<uint32_t*> ip;
  for (uint32_t i = 0; i < 4 * ipr; ++i) {
    icols.emplace_back(a, (size_t)1 << log_n, false);
    ip.push_back(icols.back().p);
  }
  stwo_qm31 cs{};
  a.ck(stwo_hip_gen_interaction_trace_poseidon(h, trace.data(), ipr, log_n, wq(c.z), wq(c.alpha), ip.data(), &cs),
       "gen_interaction_trace_poseidon");
  c.claimed_sum = rq(cs);
  trees.emplace_back();
  interpolate_into(a, std::vector<const uint32_t*>(ip.begin(), ip.end()), std::vector<uint32_t>(ip.size(), log_n),
                   trees[2]);
  icols.clear();
  commit<MC>(a, trees[2], cfg, ch);
  r.claimed_sum = c.claimed_sum;
  ProofParts parts;
  r.status = prove_ex<MC>(a, ch, trees, {c}, cfg, parts);
  r.proof = encode(cfg, parts);
  r.digest = ch.digest;
  return r;
}

}  // namespace

ProveResult prove_wide_fib_multi(stwo_hip_ctx* h, const std::vector<const uint32_t*>& trace,
                                 const std::vector<uint32_t>& logs, const std::vector<uint32_t>& ncols,
                                 const PcsCfg& cfg, MerkleKind mc) {
  return mc == MerkleKind::Poseidon252 ? prove_wide_fib_multi_mc<Poseidon252MC>(h, trace, logs, ncols, cfg)
                                       : prove_wide_fib_multi_mc<Blake2sMC>(h, trace, logs, ncols, cfg);
}
ProveResult prove_poseidon(stwo_hip_ctx* h, const std::vector<const uint32_t*>& trace, uint32_t log_n, uint32_t ipr,
                           const PcsCfg& cfg, MerkleKind mc) {
  return mc == MerkleKind::Poseidon252 ? prove_poseidon_mc<Poseidon252MC>(h, trace, log_n, ipr, cfg)
                                       : prove_poseidon_mc<Blake2sMC>(h, trace, log_n, ipr, cfg);
}

}  // namespace percall
}  // namespace sh
