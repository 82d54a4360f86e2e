// Oracle (test infrastructure only): FRI prover and verifier, restated from
//   prover/fri.rs:100-412 (commit, decommit, compute_decommitment_positions_and_witness_evals)
//   core/fri.rs:83-780 (FriVerifier, fold_line, fold_coset, fold_circle_into_line)
//   prover/line.rs:44-64 (LineEvaluation::interpolate, line_ifft)
//   core/poly/line.rs (LineDomain, LinePoly::eval_at_point)
// Paths are relative to crates/stwo/src.
#pragma once
#include <functional>
#include "ocore.h"

namespace oracle {

// A column of QM31 values stored as 4 coordinate columns (prover/secure_column.rs:13).
struct SecureCol {
  std::vector<M31> c[4];
  size_t size() const { return c[0].size(); }
  void resize(size_t n) { for (auto& x : c) x.assign(n, M31(0)); }
  QM31 at(size_t i) const { return QM31::from_m31(c[0][i], c[1][i], c[2][i], c[3][i]); }
  void set(size_t i, QM31 v) { for (int k = 0; k < 4; ++k) c[k][i] = v.coord(k); }
  std::vector<const std::vector<M31>*> refs() const { return {&c[0], &c[1], &c[2], &c[3]}; }
};

struct FriConfig {
  uint32_t log_blowup_factor = 1;
  uint32_t log_last_layer_degree_bound = 0;
  uint32_t n_queries = 3;
  uint32_t line_fold_step = 1;
  size_t last_layer_domain_size() const { return (size_t)1 << (log_last_layer_degree_bound + log_blowup_factor); }
};

struct FriLayerProof {
  std::vector<QM31> fri_witness;
  MerkleDecommitment decommitment;
  Hash32 commitment;
};
struct FriProof {
  FriLayerProof first_layer;
  std::vector<FriLayerProof> inner_layers;
  std::vector<QM31> last_layer_poly;  // LinePoly coeffs as stored (bit-reversed order)
};

// Line-domain x coordinate at index i: the x of coset point i.
inline M31 line_at(const Coset& c, size_t i) { return c.at(i).x; }

inline QM31 fold_pair_line(QM31 fx, QM31 fnx, M31 x_inv, QM31 alpha) {
  QM31 f0 = fx, f1 = fnx;
  ibutterfly(f0, f1, x_inv);
  return f0 + alpha * f1;
}

// One fold of a line evaluation on `domain` (core/fri.rs fold_line).
inline SecureCol fold_line(const SecureCol& e, const Coset& domain, QM31 alpha) {
  size_t n = e.size();
  SecureCol r; r.resize(n / 2);
  for (size_t i = 0; i < n / 2; ++i) {
    M31 x = line_at(domain, bit_reverse_index(i << 1, domain.log_size));
    r.set(i, fold_pair_line(e.at(2 * i), e.at(2 * i + 1), inv(x), alpha));
  }
  return r;
}

// dst = dst * alpha^2 + fold(src), src on a circle domain (core/fri.rs fold_circle_into_line).
inline void fold_circle_into_line(SecureCol& dst, const SecureCol& src, const CircleDomain& sd, QM31 alpha) {
  QM31 a2 = alpha * alpha;
  size_t n = src.size();
  std::vector<M31> yinv(n / 2);
  for (size_t i = 0; i < n / 2; ++i) yinv[i] = sd.at(bit_reverse_index(i << 1, sd.log_size())).y;
  yinv = batch_inverse(yinv);
  for (size_t i = 0; i < n / 2; ++i) {
    QM31 f0 = src.at(2 * i), f1 = src.at(2 * i + 1);
    ibutterfly(f0, f1, yinv[i]);
    QM31 fp = alpha * f1 + f0;
    dst.set(i, dst.at(i) * a2 + fp);
  }
}

// LineEvaluation::interpolate: returns LinePoly coeffs (stored, bit-reversed order).
inline std::vector<QM31> line_interpolate(const SecureCol& e, const Coset& domain) {
  std::vector<QM31> v(e.size());
  for (size_t i = 0; i < e.size(); ++i) v[i] = e.at(i);
  bit_reverse(v);
  Coset d = domain;
  while (d.size() > 1) {
    auto xs = d.points();
    for (size_t ch = 0; ch < v.size(); ch += d.size()) {
      size_t half = d.size() / 2;
      for (size_t i = 0; i < half; ++i) ibutterfly(v[ch + i], v[ch + half + i], inv(xs[i].x));
    }
    d = d.doubled();
  }
  M31 li = inv(M31((uint32_t)v.size()));
  for (auto& x : v) x = x * li;
  return v;
}

inline QM31 fold_qm31(const std::vector<QM31>& vals, const std::vector<QM31>& factors, size_t off, size_t n, size_t fi) {
  if (n == 1) return vals[off];
  QM31 l = fold_qm31(vals, factors, off, n / 2, fi + 1);
  QM31 r = fold_qm31(vals, factors, off + n / 2, n / 2, fi + 1);
  return l + r * factors[fi];
}
inline QM31 line_poly_eval(const std::vector<QM31>& coeffs, QM31 x) {
  std::vector<QM31> d;
  size_t L = log2_size(coeffs.size());
  for (size_t i = 0; i < L; ++i) { d.push_back(x); x = double_x(x); }
  return fold_qm31(coeffs, d, 0, coeffs.size(), 0);
}

inline void decommitment_positions_and_witness(const SecureCol& col, const std::vector<size_t>& q, uint32_t fold_step,
                                               std::vector<size_t>& positions, std::vector<QM31>& witness) {
  size_t i = 0;
  while (i < q.size()) {
    size_t j = i;
    while (j < q.size() && (q[j] >> fold_step) == (q[i] >> fold_step)) ++j;
    size_t start = (q[i] >> fold_step) << fold_step;
    size_t k = i;
    for (size_t pos = start; pos < start + ((size_t)1 << fold_step); ++pos) {
      positions.push_back(pos);
      if (k < j && q[k] == pos) ++k;
      else witness.push_back(col.at(pos));
    }
    i = j;
  }
}

struct FriProver {
  FriConfig config;
  const SecureCol* column = nullptr;
  uint32_t column_log = 0;
  MerkleTree first_tree;
  struct Inner { SecureCol eval; MerkleTree tree; uint32_t fold_step; };
  std::vector<Inner> inner;
  std::vector<QM31> last_layer_poly;

  void commit(Channel& ch, const FriConfig& cfg, const SecureCol& col) {
    config = cfg;
    column = &col;
    column_log = log2_size(col.size());
    first_tree = merkle_commit(col.refs());
    ch.mix_root(first_tree.root());
    uint32_t first_inner_log = column_log - 1;
    Coset dom = Coset::half_odds(first_inner_log);
    SecureCol le; le.resize((size_t)1 << first_inner_log);
    QM31 alpha = ch.draw_secure_felt();
    fold_circle_into_line(le, col, canonic_domain(column_log), alpha);
    uint32_t last_log = log2_size(cfg.last_layer_domain_size());
    if (first_inner_log != last_log) {
      while (log2_size(le.size()) > last_log + cfg.line_fold_step) {
        Inner in{le, merkle_commit(le.refs()), cfg.line_fold_step};
        ch.mix_root(in.tree.root());
        QM31 a = ch.draw_secure_felt();
        for (uint32_t s = 0; s < cfg.line_fold_step; ++s) { le = fold_line(le, dom, a); dom = dom.doubled(); a = a * a; }
        inner.push_back(std::move(in));
      }
      uint32_t last_step = log2_size(le.size()) - last_log;
      Inner in{le, merkle_commit(le.refs()), last_step};
      ch.mix_root(in.tree.root());
      QM31 a = ch.draw_secure_felt();
      for (uint32_t s = 0; s < last_step; ++s) { le = fold_line(le, dom, a); dom = dom.doubled(); a = a * a; }
      inner.push_back(std::move(in));
    }
    if (le.size() != cfg.last_layer_domain_size()) throw std::runtime_error("fri last layer size");
    auto coeffs = line_interpolate(le, dom);  // stored order
    std::vector<QM31> ordered = coeffs;
    bit_reverse(ordered);
    size_t bound = (size_t)1 << cfg.log_last_layer_degree_bound;
    for (size_t i = bound; i < ordered.size(); ++i)
      if (!ordered[i].is_zero()) throw std::runtime_error("invalid degree");
    ordered.resize(bound);
    bit_reverse(ordered);
    last_layer_poly = ordered;
    ch.mix_felts(last_layer_poly);
  }

  static FriLayerProof decommit_layer(const SecureCol& col, const MerkleTree& tree, const std::vector<size_t>& q,
                                      uint32_t fold_step) {
    FriLayerProof p;
    std::vector<size_t> pos;
    decommitment_positions_and_witness(col, q, fold_step, pos, p.fri_witness);
    std::map<uint32_t, std::vector<size_t>> qpl{{log2_size(col.size()), pos}};
    merkle_decommit(tree, qpl, col.refs(), p.decommitment);
    p.commitment = tree.root();
    return p;
  }

  FriProof decommit(Channel& ch, std::vector<size_t>& query_positions) {
    auto unsorted = draw_queries(ch, column_log, config.n_queries);
    auto q = sorted_unique(unsorted);
    query_positions = q;
    FriProof pr;
    pr.first_layer = decommit_layer(*column, first_tree, q, 1);
    auto lq = fold_queries(q, 1);
    for (auto& in : inner) {
      pr.inner_layers.push_back(decommit_layer(in.eval, in.tree, lq, in.fold_step));
      lq = fold_queries(lq, in.fold_step);
    }
    pr.last_layer_poly = last_layer_poly;
    return pr;
  }
};

// ---- verifier ----------------------------------------------------------------------------
inline bool rebuild_evals(const std::vector<size_t>& q, const std::vector<QM31>& qe, const std::vector<QM31>& wit,
                          uint32_t fold_step, uint32_t log_domain, std::vector<size_t>& positions,
                          std::vector<std::vector<QM31>>& subsets, std::vector<size_t>& initials) {
  size_t qi = 0, wi = 0, i = 0;
  while (i < q.size()) {
    size_t j = i;
    while (j < q.size() && (q[j] >> fold_step) == (q[i] >> fold_step)) ++j;
    size_t start = (q[i] >> fold_step) << fold_step;
    std::vector<QM31> sub;
    size_t k = i;
    for (size_t pos = start; pos < start + ((size_t)1 << fold_step); ++pos) {
      positions.push_back(pos);
      if (k < j && q[k] == pos) { ++k; sub.push_back(qe[qi++]); }
      else { if (wi >= wit.size()) return false; sub.push_back(wit[wi++]); }
    }
    subsets.push_back(sub);
    initials.push_back(bit_reverse_index(start, log_domain));
    i = j;
  }
  return wi == wit.size();
}

inline QM31 fold_coset(std::vector<QM31> eval, Coset domain, QM31 alpha) {
  uint32_t n = domain.log_size;
  for (uint32_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < ((size_t)1 << (n - i)); j += 2) {
      M31 x = line_at(domain, bit_reverse_index(j, domain.log_size));
      QM31 f0 = eval[j], f1 = eval[j + 1];
      ibutterfly(f0, f1, inv(x));
      eval[j >> 1] = f0 + alpha * f1;
    }
    alpha = alpha * alpha;
    domain = domain.doubled();
  }
  return eval[0];
}

// Returns "" on success, or an error name.
inline std::string fri_verify(Channel& ch, const FriConfig& cfg, const FriProof& proof, uint32_t column_bound,
                              const std::function<std::vector<QM31>(const std::vector<size_t>&)>& first_layer_answers,
                              std::vector<size_t>* out_positions, const std::function<bool(Channel&)>& between) {
  ch.mix_root(proof.first_layer.commitment);
  uint32_t col_log = column_bound + cfg.log_blowup_factor;
  QM31 first_alpha = ch.draw_secure_felt();
  struct IL { Coset domain; QM31 alpha; uint32_t fold_step; const FriLayerProof* p; };
  std::vector<IL> inner;
  uint32_t layer_bound = column_bound - 1;
  Coset dom = Coset::half_odds(layer_bound + cfg.log_blowup_factor);
  for (size_t li = 0; li < proof.inner_layers.size(); ++li) {
    ch.mix_root(proof.inner_layers[li].commitment);
    bool last = li + 1 == proof.inner_layers.size();
    uint32_t fs = cfg.line_fold_step;
    if (last) {
      if (layer_bound < cfg.log_last_layer_degree_bound) return "InvalidNumFriLayers";
      fs = layer_bound - cfg.log_last_layer_degree_bound;
      if (fs < 1 || fs > cfg.line_fold_step) return "InvalidNumFriLayers";
    }
    QM31 a = ch.draw_secure_felt();
    inner.push_back({dom, a, fs, &proof.inner_layers[li]});
    if (layer_bound < fs) return "InvalidNumFriLayers";
    layer_bound -= fs;
    dom = dom.repeated_double(fs);
  }
  if (layer_bound != cfg.log_last_layer_degree_bound) return "InvalidNumFriLayers";
  if (proof.last_layer_poly.size() > ((size_t)1 << cfg.log_last_layer_degree_bound)) return "LastLayerDegreeInvalid";
  ch.mix_felts(proof.last_layer_poly);
  if (!between(ch)) return "ProofOfWork";
  // query positions
  auto unsorted = draw_queries(ch, col_log, cfg.n_queries);
  auto q = sorted_unique(unsorted);
  if (out_positions) *out_positions = q;
  auto evals = first_layer_answers(q);
  if (evals.empty() && !q.empty()) return "Answers";
  // first layer
  std::vector<size_t> pos, initials;
  std::vector<std::vector<QM31>> subsets;
  if (!rebuild_evals(q, evals, proof.first_layer.fri_witness, 1, col_log, pos, subsets, initials))
    return "FirstLayerEvaluationsInvalid";
  {
    std::vector<M31> flat;
    for (auto& s : subsets) for (auto& x : s) for (int k = 0; k < 4; ++k) flat.push_back(x.coord(k));
    std::map<uint32_t, std::vector<size_t>> qpl{{col_log, pos}};
    auto err = merkle_verify(proof.first_layer.commitment, {col_log, col_log, col_log, col_log}, qpl, flat,
                             proof.first_layer.decommitment);
    if (!err.empty()) return "FirstLayerCommitmentInvalid:" + err;
  }
  CircleDomain cdom = canonic_domain(col_log);
  std::vector<QM31> layer_evals;
  for (size_t s = 0; s < subsets.size(); ++s) {
    uint32_t idx = cdom.index_at(initials[s]);
    CirclePoint<M31> p = index_to_point(idx);
    QM31 f0 = subsets[s][0], f1 = subsets[s][1];
    ibutterfly(f0, f1, inv(p.y));
    layer_evals.push_back(f0 + first_alpha * f1);
  }
  auto lq = fold_queries(q, 1);
  for (size_t li = 0; li < inner.size(); ++li) {
    auto& L = inner[li];
    std::vector<size_t> p2, in2;
    std::vector<std::vector<QM31>> sub2;
    if (!rebuild_evals(lq, layer_evals, L.p->fri_witness, L.fold_step, L.domain.log_size, p2, sub2, in2))
      return "InnerLayerEvaluationsInvalid";
    std::vector<M31> flat;
    for (auto& s : sub2) for (auto& x : s) for (int k = 0; k < 4; ++k) flat.push_back(x.coord(k));
    uint32_t cl = L.domain.log_size;
    std::map<uint32_t, std::vector<size_t>> qpl{{cl, p2}};
    auto err = merkle_verify(L.p->commitment, {cl, cl, cl, cl}, qpl, flat, L.p->decommitment);
    if (!err.empty()) return "InnerLayerCommitmentInvalid:" + err;
    std::vector<QM31> folded;
    for (size_t s = 0; s < sub2.size(); ++s) {
      uint32_t init = L.domain.index_at(in2[s]);
      folded.push_back(fold_coset(sub2[s], Coset::make(init, L.fold_step), L.alpha));
    }
    layer_evals = folded;
    lq = fold_queries(lq, L.fold_step);
  }
  for (size_t i = 0; i < lq.size(); ++i) {
    M31 x = line_at(dom, bit_reverse_index(lq[i], dom.log_size));
    if (layer_evals[i] != line_poly_eval(proof.last_layer_poly, QM31(x))) return "LastLayerEvaluationsInvalid";
  }
  return "";
}

}  // namespace oracle
