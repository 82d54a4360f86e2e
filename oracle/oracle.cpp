// Oracle (test infrastructure only): CPU restatement of the reference prover and verifier for
// the Blake2s Merkle channel, with a C API for the test suite and the CPU baseline.
// Only tests/, the smoke check and bench.py's cpu_baseline leg may load the library built
// from this file; the product (HIP) path never links or calls it.
//
// Reference files followed (relative to crates/stwo/src unless noted):
//   prove / prove_ex                 prover/mod.rs:28-150
//   CommitmentSchemeProver           prover/pcs/mod.rs:36-494 (commit, prove_values, decommit)
//   compute_fri_quotients            prover/pcs/quotient_ops.rs:73-145
//   samples / batches / line coeffs  core/pcs/quotients.rs:71-366; core/constraints.rs:119-140
//   CPU quotients                    prover/backend/cpu/quotients.rs:20-80
//   composition accumulator          prover/air/accumulation.rs:21-125; cpu/accumulation.rs
//   component domain evaluation      crates/constraint-framework/src/prover/component_prover.rs:32-169
//   verifier                         core/verifier.rs:19-118; core/pcs/verifier.rs:26-182;
//                                    core/pcs/quotients.rs:120-200 (fri_answers)
//   wide-fib / Poseidon provers      crates/examples/src/wide_fibonacci/mod.rs:186-261;
//                                    crates/examples/src/poseidon/mod.rs:216-419
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <omp.h>
#include "oair.h"

using namespace oracle;

namespace {

struct PcsConfig {
  uint32_t pow_bits = 5;
  FriConfig fri;
  uint32_t lifting_log_size = 0;  // 0 = None
};

struct StarkProof {
  PcsConfig config;
  std::vector<Hash32> commitments;
  std::vector<std::vector<std::vector<QM31>>> sampled_values;
  std::vector<MerkleDecommitment> decommitments;
  std::vector<std::vector<M31>> queried_values;
  uint64_t proof_of_work = 0;
  FriProof fri;
};

// ---- canonical byte encoding (little endian; counts as u32) -------------------------------
struct Writer {
  std::vector<uint8_t> b;
  void u32(uint32_t v) { for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i))); }
  void u64(uint64_t v) { u32((uint32_t)v); u32((uint32_t)(v >> 32)); }
  void hash(const Hash32& h) { b.insert(b.end(), h.begin(), h.end()); }
  void qm31(const QM31& q) { for (int i = 0; i < 4; ++i) u32(q.coord(i).v); }
  void dec(const MerkleDecommitment& d) {
    u32((uint32_t)d.hash_witness.size());
    for (auto& h : d.hash_witness) hash(h);
    u32((uint32_t)d.column_witness.size());
    for (auto& v : d.column_witness) u32(v.v);
  }
  void layer(const FriLayerProof& l) {
    u32((uint32_t)l.fri_witness.size());
    for (auto& q : l.fri_witness) qm31(q);
    dec(l.decommitment);
    hash(l.commitment);
  }
};
std::vector<uint8_t> serialize(const StarkProof& p) {
  Writer w;
  w.u32(p.config.pow_bits);
  w.u32(p.config.fri.log_blowup_factor);
  w.u32(p.config.fri.log_last_layer_degree_bound);
  w.u32(p.config.fri.n_queries);
  w.u32(p.config.fri.line_fold_step);
  w.u32(p.config.lifting_log_size);
  w.u32((uint32_t)p.commitments.size());
  for (auto& h : p.commitments) w.hash(h);
  w.u32((uint32_t)p.sampled_values.size());
  for (auto& t : p.sampled_values) {
    w.u32((uint32_t)t.size());
    for (auto& c : t) { w.u32((uint32_t)c.size()); for (auto& q : c) w.qm31(q); }
  }
  w.u32((uint32_t)p.decommitments.size());
  for (auto& d : p.decommitments) w.dec(d);
  w.u32((uint32_t)p.queried_values.size());
  for (auto& t : p.queried_values) { w.u32((uint32_t)t.size()); for (auto& v : t) w.u32(v.v); }
  w.u64(p.proof_of_work);
  w.layer(p.fri.first_layer);
  w.u32((uint32_t)p.fri.inner_layers.size());
  for (auto& l : p.fri.inner_layers) w.layer(l);
  w.u32((uint32_t)p.fri.last_layer_poly.size());
  for (auto& q : p.fri.last_layer_poly) w.qm31(q);
  return w.b;
}
struct Reader {
  const uint8_t* p;
  size_t n, i = 0;
  void need(size_t k) { if (i + k > n) throw std::runtime_error("proof truncated"); }
  uint32_t u32() { need(4); uint32_t v = 0; for (int k = 0; k < 4; ++k) v |= (uint32_t)p[i + k] << (8 * k); i += 4; return v; }
  uint64_t u64() { uint64_t lo = u32(); return lo | ((uint64_t)u32() << 32); }
  uint32_t count() { uint32_t c = u32(); if (c > (1u << 28)) throw std::runtime_error("bad count"); return c; }
  Hash32 hash() { need(32); Hash32 h; memcpy(h.data(), p + i, 32); i += 32; return h; }
  M31 m31() { uint32_t v = u32(); if (v >= P) throw std::runtime_error("non-canonical M31"); return M31(v); }
  QM31 qm31() { M31 a = m31(), b = m31(), c = m31(), d = m31(); return QM31::from_m31(a, b, c, d); }
  MerkleDecommitment dec() {
    MerkleDecommitment d;
    uint32_t nh = count(); for (uint32_t k = 0; k < nh; ++k) d.hash_witness.push_back(hash());
    uint32_t nc = count(); for (uint32_t k = 0; k < nc; ++k) d.column_witness.push_back(m31());
    return d;
  }
  FriLayerProof layer() {
    FriLayerProof l;
    uint32_t nw = count(); for (uint32_t k = 0; k < nw; ++k) l.fri_witness.push_back(qm31());
    l.decommitment = dec();
    l.commitment = hash();
    return l;
  }
};
StarkProof deserialize(const uint8_t* b, size_t n) {
  Reader r{b, n};
  StarkProof p;
  p.config.pow_bits = r.u32();
  p.config.fri.log_blowup_factor = r.u32();
  p.config.fri.log_last_layer_degree_bound = r.u32();
  p.config.fri.n_queries = r.u32();
  p.config.fri.line_fold_step = r.u32();
  p.config.lifting_log_size = r.u32();
  uint32_t nc = r.count(); for (uint32_t k = 0; k < nc; ++k) p.commitments.push_back(r.hash());
  uint32_t nt = r.count();
  p.sampled_values.resize(nt);
  for (auto& t : p.sampled_values) {
    t.resize(r.count());
    for (auto& c : t) { c.resize(r.count()); for (auto& q : c) q = r.qm31(); }
  }
  uint32_t nd = r.count(); for (uint32_t k = 0; k < nd; ++k) p.decommitments.push_back(r.dec());
  uint32_t nq = r.count();
  p.queried_values.resize(nq);
  for (auto& t : p.queried_values) { t.resize(r.count()); for (auto& v : t) v = r.m31(); }
  p.proof_of_work = r.u64();
  p.fri.first_layer = r.layer();
  uint32_t ni = r.count(); for (uint32_t k = 0; k < ni; ++k) p.fri.inner_layers.push_back(r.layer());
  uint32_t nl = r.count(); for (uint32_t k = 0; k < nl; ++k) p.fri.last_layer_poly.push_back(r.qm31());
  if (r.i != n) throw std::runtime_error("trailing bytes");
  return p;
}

// ---- commitment scheme prover ---------------------------------------------------------------
struct TreeProver {
  std::vector<std::vector<M31>> coeffs;
  std::vector<std::vector<M31>> evals;
  MerkleTree tree;
  std::vector<const std::vector<M31>*> eval_refs() const {
    std::vector<const std::vector<M31>*> r;
    for (auto& e : evals) r.push_back(&e);
    return r;
  }
};

struct SchemeProver {
  PcsConfig config;
  const TwiddleTree* tw;
  std::vector<TreeProver> trees;

  void commit(std::vector<std::vector<M31>> polys, Channel& ch) {
    TreeProver t;
    t.evals.resize(polys.size());
#pragma omp parallel for schedule(dynamic)
    for (long i = 0; i < (long)polys.size(); ++i)
      t.evals[i] = evaluate(polys[i], log2_size(polys[i].size()) + config.fri.log_blowup_factor, *tw);
    t.coeffs = std::move(polys);
    t.tree = merkle_commit(t.eval_refs());
    ch.mix_root(t.tree.root());
    trees.push_back(std::move(t));
  }
  std::vector<std::vector<M31>> interpolate_columns(std::vector<std::vector<M31>> cols) const {
#pragma omp parallel for schedule(dynamic)
    for (long i = 0; i < (long)cols.size(); ++i) interpolate(cols[i], *tw);
    return cols;
  }
};

struct PointSample { CirclePoint<QM31> point; QM31 value; };

bool point_less(const CirclePoint<QM31>& a, const CirclePoint<QM31>& b) {
  for (int k = 0; k < 4; ++k) if (a.x.coord(k).v != b.x.coord(k).v) return a.x.coord(k).v < b.x.coord(k).v;
  for (int k = 0; k < 4; ++k) if (a.y.coord(k).v != b.y.coord(k).v) return a.y.coord(k).v < b.y.coord(k).v;
  return false;
}

// core/pcs/quotients.rs build_samples_with_randomness_and_periodicity
std::vector<std::vector<std::vector<std::pair<PointSample, QM31>>>> samples_with_randomness(
    const std::vector<std::vector<std::vector<PointSample>>>& samples, const std::vector<std::vector<uint32_t>>& log_sizes,
    uint32_t lifting_log_size, QM31 random_coeff) {
  QM31 pw = QM31::one();
  auto next = [&]() { QM31 c = pw; pw = pw * random_coeff; return c; };
  CirclePoint<M31> gen = canonic_coset(lifting_log_size).step;
  std::vector<std::vector<std::vector<std::pair<PointSample, QM31>>>> res(samples.size());
  for (size_t t = 0; t < samples.size(); ++t) {
    for (size_t c = 0; c < samples[t].size(); ++c) {
      std::vector<std::pair<PointSample, QM31>> ns;
      auto& sc = samples[t][c];
      if (!sc.empty()) {
        if (sc.size() == 2) {
          CirclePoint<M31> pg = cp_repeated_double(gen, log_sizes[t][c]);
          ns.push_back({PointSample{cp_add(sc[1].point, to_ef(pg)), sc[1].value}, next()});
        }
        for (auto& s : sc) ns.push_back({s, next()});
      }
      res[t].push_back(ns);
    }
  }
  return res;
}

struct NumeratorData { size_t column_index; QM31 value, rand; };
struct SampleBatch { CirclePoint<QM31> point; std::vector<NumeratorData> cols; };
std::vector<SampleBatch> sample_batches(const std::vector<const std::vector<std::pair<PointSample, QM31>>*>& cols) {
  std::vector<SampleBatch> out;
  for (size_t ci = 0; ci < cols.size(); ++ci)
    for (auto& [s, r] : *cols[ci]) {
      SampleBatch* b = nullptr;
      for (auto& e : out) if (e.point == s.point) { b = &e; break; }
      if (!b) { out.push_back({s.point, {}}); b = &out.back(); }
      b->cols.push_back({ci, s.value, r});
    }
  return out;
}
struct LineCoeffs { QM31 a, b, c; };
LineCoeffs line_coeffs(const CirclePoint<QM31>& p, QM31 value, QM31 alpha) {
  if (p.y == conj(p.y)) throw std::runtime_error("Cannot evaluate a line with a single point");
  QM31 a = conj(value) - value;
  QM31 c = conj(p.y) - p.y;
  QM31 b = value * c - a * p.y;
  return {alpha * a, alpha * b, alpha * c};
}
CM31 quotient_denominator(const CirclePoint<QM31>& sp, CirclePoint<M31> dp) {
  CM31 prx = sp.x.a, pry = sp.y.a, pix = sp.x.b, piy = sp.y.b;
  return (prx - CM31(dp.x)) * piy - (pry - CM31(dp.y)) * pix;
}

struct AccNum { CirclePoint<QM31> point; SecureCol partial; QM31 first_linear; };

SecureCol lift_and_accumulate(std::vector<SecureCol> cols) {
  SecureCol curr; curr.resize(2);
  for (auto& col : cols) {
    uint32_t lr = log2_size(col.size()) - log2_size(curr.size());
    for (size_t i = 0; i < col.size(); ++i) col.set(i, col.at(i) + curr.at(((i >> (lr + 1)) << 1) + (i & 1)));
    curr = std::move(col);
  }
  return curr;
}

// QuotientOps::accumulate_numerators for one batch (prover/pcs/quotient_ops.rs:19-43; CPU form
// prover/backend/cpu/quotients.rs): partial[row] = sum_k (c_k * col_k[row] - b_k),
// first_linear = sum_k a_k, with (a, b, c) the batch's line coefficients times the random power.
AccNum accumulate_batch(const SampleBatch& b, const std::vector<const std::vector<M31>*>& cv) {
  std::vector<LineCoeffs> lc;
  for (auto& nd : b.cols) lc.push_back(line_coeffs(b.point, nd.value, nd.rand));
  size_t n = cv[0]->size();
  AccNum an;
  an.point = b.point;
  an.partial.resize(n);
  an.first_linear = QM31::zero();
  for (auto& l : lc) an.first_linear = an.first_linear + l.a;
#pragma omp parallel for schedule(static) if (n >= 4096)
  for (long row = 0; row < (long)n; ++row) {
    QM31 acc = QM31::zero();
    for (size_t k = 0; k < b.cols.size(); ++k) acc = acc + (lc[k].c * (*cv[b.cols[k].column_index])[row] - lc[k].b);
    an.partial.set(row, acc);
  }
  return an;
}

// QuotientOps::compute_quotients_and_combine (cpu/quotients.rs, simd/quotients.rs:73-121): on every
// row of CanonicCoset(lifting_log_size), bit-reversed, sum over the accumulations of
// (lift(partial)[row] - first_linear * y) / den, den = (Re p.x - x) Im p.y - (Re p.y - y) Im p.x.
SecureCol quotients_and_combine(const std::vector<AccNum>& per_point, uint32_t lifting_log_size) {
  CircleDomain dom = canonic_domain(lifting_log_size);
  size_t N = (size_t)1 << lifting_log_size;
  SecureCol q; q.resize(N);
  auto pts = dom.half_coset.points();
#pragma omp parallel for schedule(static)
  for (long row = 0; row < (long)N; ++row) {
    size_t nat = bit_reverse_index(row, lifting_log_size);
    CirclePoint<M31> dp = nat < N / 2 ? pts[nat] : cp_conj(pts[nat - N / 2]);
    QM31 acc = QM31::zero();
    for (auto& a : per_point) {
      CM31 dinv = inv(quotient_denominator(a.point, dp));
      uint32_t lr = lifting_log_size - log2_size(a.partial.size());
      size_t li = ((row >> (lr + 1)) << 1) + (row & 1);
      QM31 num = a.partial.at(li) - a.first_linear * dp.y;
      acc = acc + mul_cm31(num, dinv);
    }
    q.set(row, acc);
  }
  return q;
}

SecureCol compute_fri_quotients(const std::vector<std::vector<const std::vector<M31>*>>& columns,
                                const std::vector<std::vector<std::vector<PointSample>>>& samples, QM31 random_coeff,
                                uint32_t lifting_log_size) {
  std::vector<std::vector<uint32_t>> log_sizes(columns.size());
  for (size_t t = 0; t < columns.size(); ++t)
    for (auto* c : columns[t]) log_sizes[t].push_back(log2_size(c->size()));
  auto swr = samples_with_randomness(samples, log_sizes, lifting_log_size, random_coeff);
  struct Item { const std::vector<M31>* col; const std::vector<std::pair<PointSample, QM31>>* s; uint32_t ls; };
  std::vector<Item> items;
  for (size_t t = 0; t < columns.size(); ++t)
    for (size_t c = 0; c < columns[t].size(); ++c) items.push_back({columns[t][c], &swr[t][c], log_sizes[t][c]});
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.ls < b.ls; });
  std::vector<AccNum> accs;
  for (size_t i = 0; i < items.size();) {
    size_t j = i;
    while (j < items.size() && items[j].ls == items[i].ls) ++j;
    std::vector<const std::vector<std::pair<PointSample, QM31>>*> sv;
    std::vector<const std::vector<M31>*> cv;
    for (size_t k = i; k < j; ++k) { sv.push_back(items[k].s); cv.push_back(items[k].col); }
    for (auto& b : sample_batches(sv)) accs.push_back(accumulate_batch(b, cv));
    i = j;
  }
  std::stable_sort(accs.begin(), accs.end(), [](const AccNum& a, const AccNum& b) { return point_less(a.point, b.point); });
  std::vector<AccNum> per_point;
  for (size_t i = 0; i < accs.size();) {
    size_t j = i;
    while (j < accs.size() && accs[j].point == accs[i].point) ++j;
    AccNum a;
    a.point = accs[i].point;
    a.first_linear = QM31::zero();
    std::vector<SecureCol> parts;
    for (size_t k = i; k < j; ++k) { a.first_linear = a.first_linear + accs[k].first_linear; parts.push_back(std::move(accs[k].partial)); }
    a.partial = lift_and_accumulate(std::move(parts));
    per_point.push_back(std::move(a));
    i = j;
  }
  return quotients_and_combine(per_point, lifting_log_size);
}

std::vector<size_t> prepare_pp_positions(const std::vector<size_t>& q, uint32_t max_log, uint32_t pp_max_log) {
  std::vector<size_t> r;
  if (pp_max_log == 0) return r;
  if (max_log < pp_max_log) {
    for (size_t p : q) r.push_back(((p >> 1) << (pp_max_log - max_log + 1)) + (p & 1));
    return r;
  }
  for (size_t p : q) r.push_back(((p >> (max_log - pp_max_log + 1)) << 1) + (p & 1));
  return r;
}
std::map<uint32_t, std::vector<size_t>> queries_per_log_size(const std::vector<uint32_t>& col_logs,
                                                              const std::vector<size_t>& q) {
  std::map<uint32_t, std::vector<size_t>> r;
  if (col_logs.empty()) return r;
  uint32_t mx = *std::max_element(col_logs.begin(), col_logs.end());
  for (auto ls : col_logs) {
    if (r.count(ls)) continue;
    uint32_t shift = mx - ls;
    std::vector<size_t> v;
    for (size_t p : q) v.push_back(((p >> (shift + 1)) << 1) + (p & 1));
    r[ls] = sorted_unique(v);
  }
  return r;
}

// ---- composition ------------------------------------------------------------------------------
struct ComponentSet {
  std::vector<Component> comps;
  uint32_t composition_log_degree_bound() const {
    uint32_t m = 0;
    for (auto& c : comps) m = std::max(m, c.max_constraint_log_degree_bound());
    return m;
  }
  size_t total_constraints() const { size_t s = 0; for (auto& c : comps) s += c.n_constraints(); return s; }
};

std::vector<std::vector<M31>> compute_composition(const ComponentSet& cs, const std::vector<TreeProver>& trees,
                                                  QM31 random_coeff, const TwiddleTree& tw) {
  uint32_t max_log = cs.composition_log_degree_bound();
  size_t T = cs.total_constraints();
  std::vector<QM31> powers(T);
  { QM31 c = QM31::one(); for (size_t i = 0; i < T; ++i) { powers[i] = c; c = c * random_coeff; } }
  std::vector<SecureCol> sub(max_log + 1);
  std::vector<bool> has(max_log + 1, false);
  for (auto& comp : cs.comps) {
    size_t nc = comp.n_constraints();
    std::vector<QM31> cp(powers.end() - nc, powers.end());
    powers.resize(powers.size() - nc);
    std::reverse(cp.begin(), cp.end());
    uint32_t eval_log = comp.max_constraint_log_degree_bound();
    uint32_t trace_log = comp.log_n_rows;
    // gather component columns per tree and extend to the evaluation domain where needed
    std::vector<std::vector<std::vector<M31>>> ext(3);
    std::vector<std::vector<const std::vector<M31>*>> refs(3);
    auto gather = [&](int tree, size_t start, size_t n) {
      ext[tree].resize(n);
#pragma omp parallel for schedule(dynamic)
      for (long i = 0; i < (long)n; ++i) {
        const auto& ev = trees[tree].evals[start + i];
        if (log2_size(ev.size()) == eval_log) ext[tree][i] = ev;
        else ext[tree][i] = evaluate(trees[tree].coeffs[start + i], eval_log, tw);
      }
      for (auto& e : ext[tree]) refs[tree].push_back(&e);
    };
    gather(1, comp.trace_col_start, comp.n_trace_cols());
    if (comp.kind == POSEIDON) gather(2, comp.interaction_col_start, comp.n_interaction_cols());
    // denominators: 1 / vanishing(trace coset) on the first 2^log_expand points, bit reversed
    uint32_t le = eval_log - trace_log;
    CircleDomain ed = canonic_domain(eval_log);
    std::vector<M31> dinv;
    for (size_t i = 0; i < ((size_t)1 << le); ++i) dinv.push_back(inv(coset_vanishing(canonic_coset(trace_log), ed.at(i))));
    bit_reverse(dinv);
    if (!has[eval_log]) { sub[eval_log].resize((size_t)1 << eval_log); has[eval_log] = true; }
    SecureCol& acc = sub[eval_log];
    QM31 shift = comp.cumsum_shift();
#pragma omp parallel for schedule(static)
    for (long row = 0; row < (long)(1l << eval_log); ++row) {
      DomainEval e;
      e.cols = &refs;
      e.row = row;
      e.trace_log = trace_log;
      e.eval_log = eval_log;
      e.pows = cp.data();
      e.shift = shift;
      eval_component(e, comp);
      acc.set(row, acc.at(row) + e.row_res * dinv[row >> trace_log]);
    }
  }
  std::vector<SecureCol> present;
  for (uint32_t l = 0; l <= max_log; ++l) if (has[l]) present.push_back(std::move(sub[l]));
  std::vector<std::vector<M31>> out(4);
  if (present.empty()) { for (auto& c : out) c.assign((size_t)1 << max_log, M31(0)); return out; }
  SecureCol lifted = lift_and_accumulate(std::move(present));
  for (int k = 0; k < 4; ++k) { out[k] = lifted.c[k]; interpolate(out[k], tw); }
  return out;
}

std::vector<std::vector<std::vector<CirclePoint<QM31>>>> mask_points(const ComponentSet& cs, size_t n_trees,
                                                                    CirclePoint<QM31> p, uint32_t max_deg) {
  std::vector<std::vector<std::vector<CirclePoint<QM31>>>> r(n_trees);
  CirclePoint<M31> step = canonic_coset(max_deg).step;
  for (auto& c : cs.comps) {
    auto offs = c.mask_offsets();
    for (size_t t = 1; t < offs.size(); ++t)
      for (auto& col : offs[t]) {
        std::vector<CirclePoint<QM31>> pts;
        for (long o : col) {
          CirclePoint<M31> s = o > 0 ? cp_mul(step, (unsigned __int128)o) : cp_mul(cp_conj(step), (unsigned __int128)(-o));
          pts.push_back(cp_add(p, to_ef(s)));
        }
        r[t].push_back(pts);
      }
  }
  return r;
}

QM31 eval_composition_at_point(const ComponentSet& cs, CirclePoint<QM31> p,
                               const std::vector<std::vector<std::vector<QM31>>>& sv, QM31 random_coeff, uint32_t max_deg) {
  QM31 acc = QM31::zero();
  for (auto& c : cs.comps) {
    std::vector<std::vector<const std::vector<QM31>*>> m(3);
    for (size_t i = 0; i < c.n_trace_cols(); ++i) m[1].push_back(&sv[1][c.trace_col_start + i]);
    for (size_t i = 0; i < c.n_interaction_cols(); ++i) m[2].push_back(&sv[2][c.interaction_col_start + i]);
    PointEval e;
    e.mask = &m;
    e.acc = &acc;
    e.random_coeff = random_coeff;
    e.denom_inv = inv(coset_vanishing(canonic_coset(max_deg), p));
    e.shift = c.cumsum_shift();
    eval_component(e, c);
  }
  return acc;
}

QM31 extract_composition_oods(const StarkProof& pr, CirclePoint<QM31> p, uint32_t max_deg, bool& ok) {
  ok = false;
  auto& last = pr.sampled_values.back();
  if (last.size() != 8) return QM31::zero();
  for (auto& c : last) if (c.size() != 1) return QM31::zero();
  ok = true;
  QM31 l = from_partial_evals(last[0][0], last[1][0], last[2][0], last[3][0]);
  QM31 r = from_partial_evals(last[4][0], last[5][0], last[6][0], last[7][0]);
  return l + cp_repeated_double(p, max_deg - 1).x * r;
}

// prover/mod.rs prove_ex. Returns 0 on success, 2 on ConstraintsNotSatisfied (proof still filled).
int prove_ex(const ComponentSet& cs, Channel& ch, SchemeProver& sp, StarkProof& proof) {
  QM31 random_coeff = ch.draw_secure_felt();
  auto comp = compute_composition(cs, sp.trees, random_coeff, *sp.tw);
  std::vector<std::vector<M31>> halves;
  size_t half = comp[0].size() / 2;
  for (int k = 0; k < 4; ++k) halves.emplace_back(comp[k].begin(), comp[k].begin() + half);
  for (int k = 0; k < 4; ++k) halves.emplace_back(comp[k].begin() + half, comp[k].end());
  sp.commit(std::move(halves), ch);
  CirclePoint<QM31> oods = get_random_point(ch);
  uint32_t split_log = sp.trees.back().tree.max_log();
  uint32_t lifting = sp.config.lifting_log_size ? sp.config.lifting_log_size : split_log;
  if (lifting < split_log) throw std::runtime_error("lifting log size");
  uint32_t max_deg = lifting - sp.config.fri.log_blowup_factor;
  auto spts = mask_points(cs, sp.trees.size(), oods, max_deg);
  spts.back() = std::vector<std::vector<CirclePoint<QM31>>>(8, {oods});
  // preprocessed tree: no sampled columns in these AIRs
  spts[0].assign(sp.trees[0].evals.size(), {});
  // ---- prove_values ----
  uint32_t L = sp.trees.back().tree.max_log();
  std::vector<std::vector<std::vector<PointSample>>> samples(sp.trees.size());
  for (size_t t = 0; t < sp.trees.size(); ++t) {
    auto& tr = sp.trees[t];
    samples[t].resize(tr.evals.size());
#pragma omp parallel for schedule(dynamic)
    for (long c = 0; c < (long)tr.evals.size(); ++c) {
      uint32_t el = log2_size(tr.evals[c].size());
      for (auto& pt : spts[t][c]) samples[t][c].push_back({pt, eval_at_point(tr.coeffs[c], cp_repeated_double(pt, L - el))});
    }
  }
  proof.config = sp.config;
  proof.sampled_values.resize(samples.size());
  std::vector<QM31> flat;
  for (size_t t = 0; t < samples.size(); ++t)
    for (auto& c : samples[t]) {
      std::vector<QM31> v;
      for (auto& s : c) { v.push_back(s.value); flat.push_back(s.value); }
      proof.sampled_values[t].push_back(v);
    }
  ch.mix_felts(flat);
  QM31 qrc = ch.draw_secure_felt();
  std::vector<std::vector<const std::vector<M31>*>> cols;
  for (auto& t : sp.trees) cols.push_back(t.eval_refs());
  SecureCol quot = compute_fri_quotients(cols, samples, qrc, L);
  FriProver fri;
  fri.commit(ch, sp.config.fri, quot);
  proof.proof_of_work = ch.grind(sp.config.pow_bits);
  ch.mix_u64(proof.proof_of_work);
  std::vector<size_t> qpos;
  proof.fri = fri.decommit(ch, qpos);
  // Every tree is decommitted at tree-local positions. The reference maps only the preprocessed
  // tree (pcs/mod.rs:310-331); for any other tree smaller than the lifting domain its raw
  // positions run past the columns, while its verifier (core/pcs/quotients.rs
  // reshape_flat_to_per_column) expects tree-local positions. The mapping is the identity for
  // trees of full size.
  for (size_t t = 0; t < sp.trees.size(); ++t) {
    auto& tr = sp.trees[t];
    const auto q = prepare_pp_positions(qpos, L, tr.tree.max_log());
    std::vector<uint32_t> logs;
    for (auto& e : tr.evals) logs.push_back(log2_size(e.size()));
    std::map<uint32_t, std::vector<size_t>> qpl;
    uint32_t mx = tr.tree.max_log();
    for (auto ls : std::set<uint32_t>(logs.begin(), logs.end())) {
      uint32_t shift = mx - ls;
      std::vector<size_t> v;
      for (size_t p : q) v.push_back(((p >> (shift + 1)) << 1) + (p & 1));
      qpl[ls] = sorted_unique(v);
    }
    MerkleDecommitment d;
    proof.queried_values.push_back(merkle_decommit(tr.tree, qpl, tr.eval_refs(), d));
    proof.decommitments.push_back(d);
    proof.commitments.push_back(tr.tree.root());
  }
  bool ok;
  QM31 committed = extract_composition_oods(proof, oods, max_deg, ok);
  QM31 cpu = eval_composition_at_point(cs, oods, proof.sampled_values, random_coeff, max_deg);
  return (ok && committed == cpu) ? 0 : 2;
}

// core/pcs/quotients.rs reshape_flat_to_per_column
std::vector<std::vector<M31>> reshape(const std::vector<M31>& flat, const std::vector<uint32_t>& logs,
                                      const std::vector<size_t>& qp, uint32_t max_log) {
  uint32_t tmax = logs.empty() ? 0 : *std::max_element(logs.begin(), logs.end());
  auto tq = prepare_pp_positions(qp, max_log, tmax);
  std::map<uint32_t, std::vector<size_t>> by;
  for (size_t i = 0; i < logs.size(); ++i) by[logs[i]].push_back(i);
  std::map<uint32_t, std::vector<size_t>> qpl;
  for (auto& [ls, _] : by) {
    std::vector<size_t> v;
    for (size_t p : tq) v.push_back(((p >> (tmax - ls + 1)) << 1) + (p & 1));
    qpl[ls] = sorted_unique(v);
  }
  std::vector<std::vector<M31>> nat(logs.size());
  size_t fi = 0;
  for (auto it = by.rbegin(); it != by.rend(); ++it)
    for (size_t k = 0; k < qpl[it->first].size(); ++k)
      for (size_t ci : it->second) {
        if (fi >= flat.size()) throw std::runtime_error("reshape: too few values");
        nat[ci].push_back(flat[fi++]);
      }
  if (fi != flat.size()) throw std::runtime_error("reshape: too many values");
  std::vector<std::vector<M31>> res(logs.size());
  for (size_t ci = 0; ci < logs.size(); ++ci) {
    auto& pos = qpl[logs[ci]];
    for (size_t p : tq) {
      size_t lifted = ((p >> (tmax - logs[ci] + 1)) << 1) + (p & 1);
      size_t idx = std::lower_bound(pos.begin(), pos.end(), lifted) - pos.begin();
      res[ci].push_back(nat[ci][idx]);
    }
  }
  return res;
}

// core/verifier.rs verify_ex + core/pcs/verifier.rs verify_values. Returns "" on success.
std::string verify(const ComponentSet& cs, Channel& ch, std::vector<std::vector<uint32_t>> tree_logs,
                   const StarkProof& proof) {
  const PcsConfig& cfg = proof.config;
  uint32_t blow = cfg.fri.log_blowup_factor;
  if (proof.commitments.size() != tree_logs.size() + 1) return "InvalidStructure";
  uint32_t split = cs.composition_log_degree_bound() - 1;
  uint32_t lifting = cfg.lifting_log_size ? cfg.lifting_log_size : split + blow;
  uint32_t max_deg = lifting - blow;
  QM31 random_coeff = ch.draw_secure_felt();
  ch.mix_root(proof.commitments.back());
  tree_logs.push_back(std::vector<uint32_t>(8, max_deg + blow));
  CirclePoint<QM31> oods = get_random_point(ch);
  auto spts = mask_points(cs, tree_logs.size(), oods, max_deg);
  spts.back() = std::vector<std::vector<CirclePoint<QM31>>>(8, {oods});
  spts[0].assign(tree_logs[0].size(), {});
  if (proof.sampled_values.size() != spts.size()) return "InvalidStructure";
  for (size_t t = 0; t < spts.size(); ++t) {
    if (proof.sampled_values[t].size() != spts[t].size()) return "InvalidStructure";
    for (size_t c = 0; c < spts[t].size(); ++c)
      if (proof.sampled_values[t][c].size() != spts[t][c].size()) return "InvalidStructure";
  }
  bool ok;
  QM31 comp = extract_composition_oods(proof, oods, max_deg, ok);
  if (!ok) return "InvalidStructure";
  if (comp != eval_composition_at_point(cs, oods, proof.sampled_values, random_coeff, max_deg)) return "OodsNotMatching";
  // verify_values
  std::vector<QM31> flat;
  for (auto& t : proof.sampled_values) for (auto& c : t) for (auto& v : c) flat.push_back(v);
  ch.mix_felts(flat);
  QM31 qrc = ch.draw_secure_felt();
  uint32_t lls = 0;
  for (size_t t = 0; t < tree_logs.size(); ++t)
    for (size_t c = 0; c < tree_logs[t].size(); ++c)
      if (!spts[t][c].empty()) lls = std::max(lls, tree_logs[t][c]);
  uint32_t bound = lls - blow;
  std::string err;
  try {
    err = fri_verify(
        ch, cfg.fri, proof.fri, bound,
        [&](const std::vector<size_t>& qp) -> std::vector<QM31> {
          if (proof.decommitments.size() != tree_logs.size() || proof.queried_values.size() != tree_logs.size())
            throw std::runtime_error("InvalidStructure");
          for (size_t t = 0; t < tree_logs.size(); ++t) {
            uint32_t tmax = tree_logs[t].empty() ? 0 : *std::max_element(tree_logs[t].begin(), tree_logs[t].end());
            auto qpl = queries_per_log_size(tree_logs[t], prepare_pp_positions(qp, lls, tmax));
            auto e = merkle_verify(proof.commitments[t], tree_logs[t], qpl, proof.queried_values[t], proof.decommitments[t]);
            if (!e.empty()) throw std::runtime_error("Merkle:" + e);
          }
          // fri_answers
          std::vector<std::vector<M31>> qv;
          for (size_t t = 0; t < tree_logs.size(); ++t) {
            auto r = reshape(proof.queried_values[t], tree_logs[t], qp, lls);
            for (auto& x : r) qv.push_back(x);
          }
          std::vector<std::vector<std::vector<PointSample>>> samples(spts.size());
          for (size_t t = 0; t < spts.size(); ++t)
            for (size_t c = 0; c < spts[t].size(); ++c) {
              std::vector<PointSample> v;
              for (size_t k = 0; k < spts[t][c].size(); ++k) v.push_back({spts[t][c][k], proof.sampled_values[t][c][k]});
              samples[t].push_back(v);
            }
          auto swr = samples_with_randomness(samples, tree_logs, lls, qrc);
          std::vector<const std::vector<std::pair<PointSample, QM31>>*> all;
          for (auto& t : swr) for (auto& c : t) all.push_back(&c);
          auto batches = sample_batches(all);
          std::vector<std::vector<LineCoeffs>> lcs;
          for (auto& b : batches) {
            std::vector<LineCoeffs> lc;
            for (auto& nd : b.cols) lc.push_back(line_coeffs(b.point, nd.value, nd.rand));
            lcs.push_back(lc);
          }
          CircleDomain dom = canonic_domain(lls);
          std::vector<QM31> ans;
          for (size_t qi = 0; qi < qp.size(); ++qi) {
            CirclePoint<M31> dp = dom.at(bit_reverse_index(qp[qi], lls));
            QM31 acc = QM31::zero();
            for (size_t b = 0; b < batches.size(); ++b) {
              QM31 num = QM31::zero();
              for (size_t k = 0; k < batches[b].cols.size(); ++k) {
                QM31 val = lcs[b][k].c * qv[batches[b].cols[k].column_index][qi];
                num = num + (val - (lcs[b][k].a * dp.y + lcs[b][k].b));
              }
              acc = acc + mul_cm31(num, inv(quotient_denominator(batches[b].point, dp)));
            }
            ans.push_back(acc);
          }
          return ans;
        },
        nullptr,
        [&](Channel& c) {
          if (!c.verify_pow_nonce(cfg.pow_bits, proof.proof_of_work)) return false;
          c.mix_u64(proof.proof_of_work);
          return true;
        });
  } catch (const std::exception& e) {
    return e.what();
  }
  return err;
}

// ---- inputs ------------------------------------------------------------------------------------
uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// Uniform element of [0,P) for element `idx` of the stream `seed` (rejection on 31 bits).
uint32_t uniform_m31(uint64_t seed, uint64_t idx) {
  for (uint64_t a = 0; a < 4; ++a) {
    uint32_t v = (uint32_t)(splitmix64_at(seed, idx * 4 + a) >> 33);
    if (v < P) return v;
  }
  return 0;
}

std::vector<std::vector<M31>> gen_wide_fib(uint32_t log_n, uint32_t n_cols, int mode, uint64_t seed) {
  size_t N = (size_t)1 << log_n;
  std::vector<std::vector<M31>> t(n_cols, std::vector<M31>(N));
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)N; ++r) {
    M31 a = mode == 0 ? M31(1) : M31(uniform_m31(seed, 2 * r));
    M31 b = mode == 0 ? M31::from_u32((uint32_t)r) : M31(uniform_m31(seed, 2 * r + 1));
    t[0][r] = a; t[1][r] = b;
    for (uint32_t c = 2; c < n_cols; ++c) { M31 x = a * a + b * b; t[c][r] = x; a = b; b = x; }
  }
  return t;
}

struct PoseidonTrace {
  std::vector<std::vector<M31>> cols;
  std::vector<std::vector<std::array<M31, 16>>> init, fin;  // [rep][row]
};
PoseidonTrace gen_poseidon(uint32_t log_n, uint32_t ipr, int mode, uint64_t seed) {
  size_t N = (size_t)1 << log_n;
  PoseidonTrace pt;
  pt.cols.assign((size_t)ipr * POSEIDON_COLS_PER_REP, std::vector<M31>(N));
  pt.init.assign(ipr, std::vector<std::array<M31, 16>>(N));
  pt.fin.assign(ipr, std::vector<std::array<M31, 16>>(N));
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)N; ++r) {
    for (uint32_t rep = 0; rep < ipr; ++rep) {
      M31 st[16];
      for (uint32_t i = 0; i < 16; ++i)
        st[i] = mode == 0 ? M31::from_u32((uint32_t)(r + i + rep)) : M31(uniform_m31(seed, ((uint64_t)r * ipr + rep) * 16 + i));
      for (int i = 0; i < 16; ++i) pt.init[rep][r][i] = st[i];
      M31 out[POSEIDON_COLS_PER_REP];
      poseidon_permute_trace(st, out);
      for (uint32_t k = 0; k < POSEIDON_COLS_PER_REP; ++k) pt.cols[rep * POSEIDON_COLS_PER_REP + k][r] = out[k];
      for (int i = 0; i < 16; ++i) pt.fin[rep][r][i] = st[i];
    }
  }
  return pt;
}

// inclusive prefix sum in coset order for values given in bit-reversed circle-domain order
std::vector<M31> inclusive_prefix_sum(std::vector<M31> v) {
  size_t n = v.size();
  bit_reverse(v);
  std::vector<M31> coset(n);
  for (size_t i = 0; i < n / 2; ++i) { coset[2 * i] = v[i]; coset[2 * i + 1] = v[n - 1 - i]; }
  M31 acc(0);
  for (auto& x : coset) { acc = acc + x; x = acc; }
  std::vector<M31> circ(n);
  for (size_t i = 0; i < n / 2; ++i) { circ[i] = coset[2 * i]; circ[n / 2 + i] = coset[n - 1 - 2 * i]; }
  bit_reverse(circ);
  return circ;
}

// LogupTraceGenerator for the Poseidon relation; returns 4*ipr columns and the claimed sum.
std::vector<std::vector<M31>> gen_poseidon_interaction(const PoseidonTrace& pt, uint32_t log_n, uint32_t ipr,
                                                       const LookupElements& le, QM31& claimed_sum) {
  size_t N = (size_t)1 << log_n;
  std::vector<SecureCol> cols(ipr);
  for (uint32_t rep = 0; rep < ipr; ++rep) {
    cols[rep].resize(N);
    std::vector<QM31> den(N), num(N);
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)N; ++r) {
      QM31 d0 = le.combine(pt.init[rep][r].data()), d1 = le.combine(pt.fin[rep][r].data());
      num[r] = d1 - d0;
      den[r] = d0 * d1;
    }
    auto dinv = batch_inverse(den);
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)N; ++r) {
      QM31 v = num[r] * dinv[r];
      if (rep > 0) v = v + cols[rep - 1].at(r);
      cols[rep].set(r, v);
    }
  }
  SecureCol& last = cols.back();
  M31 s[4] = {M31(0), M31(0), M31(0), M31(0)};
  for (size_t r = 0; r < N; ++r) for (int k = 0; k < 4; ++k) s[k] = s[k] + last.c[k][r];
  claimed_sum = QM31::from_m31(s[0], s[1], s[2], s[3]);
  QM31 shift = claimed_sum * inv(M31((uint32_t)N));
  for (int k = 0; k < 4; ++k) {
    for (size_t r = 0; r < N; ++r) last.c[k][r] = last.c[k][r] - shift.coord(k);
    last.c[k] = inclusive_prefix_sum(last.c[k]);
  }
  std::vector<std::vector<M31>> out;
  for (auto& c : cols) for (int k = 0; k < 4; ++k) out.push_back(c.c[k]);
  return out;
}

}  // namespace

// =========================================== C API ===========================================
extern "C" {

struct OracleProveArgs {
  uint32_t air;          // 0 wide-fib, 1 Poseidon
  uint32_t log_n_rows;
  uint32_t n_cols;       // wide-fib sequence length
  uint32_t ipr;          // Poseidon instances per row
  uint32_t input_mode;   // 0 reference inputs, 1 seeded uniform
  uint64_t seed;
  uint32_t pow_bits;
  uint32_t log_blowup;
  uint32_t log_last_layer;
  uint32_t n_queries;
  uint32_t fold_step;
  uint32_t lifting_log_size;  // 0 = None
};

void oracle_free(void* p) { free(p); }
int oracle_num_threads() { return omp_get_max_threads(); }
void oracle_set_threads(int n) { omp_set_num_threads(n); }

// Proves; returns 0 ok, 2 ConstraintsNotSatisfied (the reference's prove_ex sanity check), <0 error.
int oracle_prove(const OracleProveArgs* a, uint8_t** proof_out, size_t* proof_len, uint8_t digest_out[32],
                 uint32_t claimed_sum_out[4], double* prove_ms) {
  try {
    PcsConfig cfg;
    cfg.pow_bits = a->pow_bits;
    cfg.fri.log_blowup_factor = a->log_blowup;
    cfg.fri.log_last_layer_degree_bound = a->log_last_layer;
    cfg.fri.n_queries = a->n_queries;
    cfg.fri.line_fold_step = a->fold_step;
    cfg.lifting_log_size = a->lifting_log_size;
    uint32_t n = a->log_n_rows;
    uint32_t ext = a->air == 0 ? 1 : POSEIDON_LOG_EXPAND;
    TwiddleTree tw = precompute_twiddles(canonic_domain(n + ext + a->log_blowup).half_coset);
    Channel ch;
    SchemeProver sp{cfg, &tw, {}};
    ComponentSet cs;
    Component c;
    c.log_n_rows = n;
    PoseidonTrace pt;
    std::vector<std::vector<M31>> trace;
    if (a->air == 0) {
      c.kind = WIDE_FIB;
      c.n_cols = a->n_cols;
      trace = gen_wide_fib(n, a->n_cols, a->input_mode, a->seed);
    } else {
      c.kind = POSEIDON;
      c.ipr = a->ipr;
      pt = gen_poseidon(n, a->ipr, a->input_mode, a->seed);
      trace = std::move(pt.cols);
    }
    auto t0 = std::chrono::steady_clock::now();
    sp.commit({}, ch);  // empty preprocessed tree
    sp.commit(sp.interpolate_columns(std::move(trace)), ch);
    if (a->air == 1) {
      c.le = LookupElements::draw(ch);
      QM31 cs_sum;
      auto inter = gen_poseidon_interaction(pt, n, a->ipr, c.le, cs_sum);
      c.claimed_sum = cs_sum;
      sp.commit(sp.interpolate_columns(std::move(inter)), ch);
    }
    cs.comps.push_back(c);
    StarkProof proof;
    int st = prove_ex(cs, ch, sp, proof);
    auto t1 = std::chrono::steady_clock::now();
    if (prove_ms) *prove_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    auto bytes = serialize(proof);
    *proof_out = (uint8_t*)malloc(bytes.size());
    memcpy(*proof_out, bytes.data(), bytes.size());
    *proof_len = bytes.size();
    memcpy(digest_out, ch.digest.data(), 32);
    for (int k = 0; k < 4; ++k) claimed_sum_out[k] = c.claimed_sum.coord(k).v;
    return st;
  } catch (const std::exception& e) {
    fprintf(stderr, "oracle_prove: %s\n", e.what());
    return -1;
  }
}

// Verifies a serialized proof; returns 0 when accepted, 1 when rejected (reason in err).
int oracle_verify(const OracleProveArgs* a, const uint8_t* proof, size_t len, const uint32_t claimed_sum[4], char* err,
                  size_t errlen) {
  std::string e;
  try {
    StarkProof p = deserialize(proof, len);
    ComponentSet cs;
    Component c;
    c.log_n_rows = a->log_n_rows;
    Channel ch;
    std::vector<std::vector<uint32_t>> logs;
    uint32_t blow = p.config.fri.log_blowup_factor;
    if (p.commitments.size() < 3) throw std::runtime_error("InvalidStructure");
    ch.mix_root(p.commitments[0]);
    logs.push_back({});
    ch.mix_root(p.commitments[1]);
    if (a->air == 0) {
      c.kind = WIDE_FIB;
      c.n_cols = a->n_cols;
      logs.push_back(std::vector<uint32_t>(a->n_cols, a->log_n_rows + blow));
    } else {
      c.kind = POSEIDON;
      c.ipr = a->ipr;
      logs.push_back(std::vector<uint32_t>(c.n_trace_cols(), a->log_n_rows + blow));
      c.le = LookupElements::draw(ch);
      c.claimed_sum = QM31::from_u32(claimed_sum[0], claimed_sum[1], claimed_sum[2], claimed_sum[3]);
      if (p.commitments.size() < 4) throw std::runtime_error("InvalidStructure");
      ch.mix_root(p.commitments[2]);
      logs.push_back(std::vector<uint32_t>(c.n_interaction_cols(), a->log_n_rows + blow));
    }
    cs.comps.push_back(c);
    e = verify(cs, ch, logs, p);
  } catch (const std::exception& ex) {
    e = ex.what();
  }
  if (err && errlen) { snprintf(err, errlen, "%s", e.c_str()); }
  return e.empty() ? 0 : 1;
}

// Several wide-Fibonacci components in one proof (prove() over a component list, S/prover/mod.rs:28-150):
// component i has 2^logs[i] rows and ncols[i] columns, its trace generated with seed + i; the
// trace tree holds the components' columns in order (TraceLocationAllocator order).
static PcsConfig cfg_of(const OracleProveArgs* a) {
  PcsConfig cfg;
  cfg.pow_bits = a->pow_bits;
  cfg.fri.log_blowup_factor = a->log_blowup;
  cfg.fri.log_last_layer_degree_bound = a->log_last_layer;
  cfg.fri.n_queries = a->n_queries;
  cfg.fri.line_fold_step = a->fold_step;
  cfg.lifting_log_size = a->lifting_log_size;
  return cfg;
}
int oracle_prove_wide_fib_multi(const OracleProveArgs* a, uint32_t n_comp, const uint32_t* logs, const uint32_t* ncols,
                                uint8_t** proof_out, size_t* proof_len, uint8_t digest_out[32], double* prove_ms) {
  try {
    if (n_comp == 0) throw std::runtime_error("no components");
    PcsConfig cfg = cfg_of(a);
    uint32_t max_n = 0;
    for (uint32_t i = 0; i < n_comp; ++i) max_n = std::max(max_n, logs[i]);
    TwiddleTree tw = precompute_twiddles(canonic_domain(max_n + 1 + a->log_blowup).half_coset);
    Channel ch;
    SchemeProver sp{cfg, &tw, {}};
    ComponentSet cs;
    std::vector<std::vector<M31>> trace;
    for (uint32_t i = 0; i < n_comp; ++i) {
      Component c;
      c.kind = WIDE_FIB;
      c.log_n_rows = logs[i];
      c.n_cols = ncols[i];
      c.trace_col_start = trace.size();
      auto t = gen_wide_fib(logs[i], ncols[i], a->input_mode, a->seed + i);
      for (auto& col : t) trace.push_back(std::move(col));
      cs.comps.push_back(c);
    }
    auto t0 = std::chrono::steady_clock::now();
    sp.commit({}, ch);
    sp.commit(sp.interpolate_columns(std::move(trace)), ch);
    StarkProof proof;
    int st = prove_ex(cs, ch, sp, proof);
    auto t1 = std::chrono::steady_clock::now();
    if (prove_ms) *prove_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    auto bytes = serialize(proof);
    *proof_out = (uint8_t*)malloc(bytes.size());
    memcpy(*proof_out, bytes.data(), bytes.size());
    *proof_len = bytes.size();
    memcpy(digest_out, ch.digest.data(), 32);
    return st;
  } catch (const std::exception& e) {
    fprintf(stderr, "oracle_prove_wide_fib_multi: %s\n", e.what());
    return -1;
  }
}
int oracle_verify_wide_fib_multi(const OracleProveArgs* a, uint32_t n_comp, const uint32_t* logs, const uint32_t* ncols,
                                 const uint8_t* proof, size_t len, char* err, size_t errlen) {
  std::string e;
  try {
    (void)a;
    StarkProof p = deserialize(proof, len);
    if (p.commitments.size() < 3) throw std::runtime_error("InvalidStructure");
    uint32_t blow = p.config.fri.log_blowup_factor;
    ComponentSet cs;
    Channel ch;
    std::vector<std::vector<uint32_t>> tlogs(2);
    ch.mix_root(p.commitments[0]);
    ch.mix_root(p.commitments[1]);
    size_t start = 0;
    for (uint32_t i = 0; i < n_comp; ++i) {
      Component c;
      c.kind = WIDE_FIB;
      c.log_n_rows = logs[i];
      c.n_cols = ncols[i];
      c.trace_col_start = start;
      start += ncols[i];
      for (uint32_t k = 0; k < ncols[i]; ++k) tlogs[1].push_back(logs[i] + blow);
      cs.comps.push_back(c);
    }
    e = verify(cs, ch, tlogs, p);
  } catch (const std::exception& ex) {
    e = ex.what();
  }
  if (err && errlen) { snprintf(err, errlen, "%s", e.c_str()); }
  return e.empty() ? 0 : 1;
}

// ---- primitives, for kernel-level parity tests ----
void oracle_blake2s(const uint8_t* data, size_t len, uint8_t out[32]) {
  auto h = Blake2s::hash(data, len);
  memcpy(out, h.data(), 32);
}
// Twiddle tree for the half coset of CanonicCoset(log_domain): 2^(log_domain-1) entries each.
void oracle_twiddles(uint32_t log_domain, uint32_t* tw, uint32_t* itw) {
  TwiddleTree t = precompute_twiddles(canonic_domain(log_domain).half_coset);
  for (size_t i = 0; i < t.tw.size(); ++i) { tw[i] = t.tw[i].v; itw[i] = t.itw[i].v; }
}
void oracle_interpolate(uint32_t log_size, uint32_t* vals) {
  TwiddleTree t = precompute_twiddles(canonic_domain(std::max(log_size, 1u)).half_coset);
  std::vector<M31> v((size_t)1 << log_size);
  for (size_t i = 0; i < v.size(); ++i) v[i] = M31(vals[i]);
  interpolate(v, t);
  for (size_t i = 0; i < v.size(); ++i) vals[i] = v[i].v;
}
void oracle_evaluate(uint32_t log_in, const uint32_t* coeffs, uint32_t log_out, uint32_t* out) {
  TwiddleTree t = precompute_twiddles(canonic_domain(log_out).half_coset);
  std::vector<M31> c((size_t)1 << log_in);
  for (size_t i = 0; i < c.size(); ++i) c[i] = M31(coeffs[i]);
  auto v = evaluate(c, log_out, t);
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i].v;
}
void oracle_eval_at_point(uint32_t log_size, const uint32_t* coeffs, const uint32_t point[8], uint32_t out[4]) {
  std::vector<M31> c((size_t)1 << log_size);
  for (size_t i = 0; i < c.size(); ++i) c[i] = M31(coeffs[i]);
  CirclePoint<QM31> p{QM31::from_u32(point[0], point[1], point[2], point[3]), QM31::from_u32(point[4], point[5], point[6], point[7])};
  QM31 r = eval_at_point(c, p);
  for (int k = 0; k < 4; ++k) out[k] = r.coord(k).v;
}
void oracle_barycentric_eval(uint32_t log_size, const uint32_t* evals, const uint32_t point[8], uint32_t out[4]) {
  CirclePoint<QM31> p{QM31::from_u32(point[0], point[1], point[2], point[3]), QM31::from_u32(point[4], point[5], point[6], point[7])};
  auto w = barycentric_weights(log_size, p);
  QM31 acc = QM31::zero();
  for (size_t i = 0; i < w.size(); ++i) acc = acc + w[i] * M31(evals[i]);
  for (int k = 0; k < 4; ++k) out[k] = acc.coord(k).v;
}
// Barycentric weights, coordinate-major: out[k * 2^log_size + i] = coordinate k of w_i.
void oracle_barycentric_weights(uint32_t log_size, const uint32_t point[8], uint32_t* out) {
  CirclePoint<QM31> p{QM31::from_u32(point[0], point[1], point[2], point[3]), QM31::from_u32(point[4], point[5], point[6], point[7])};
  auto w = barycentric_weights(log_size, p);
  for (size_t i = 0; i < w.size(); ++i)
    for (int k = 0; k < 4; ++k) out[(size_t)k * w.size() + i] = w[i].coord(k).v;
}
// One Merkle layer: 2^log_size hashes, prev may be null; columns are 2^log_size words each.
void oracle_commit_on_layer(uint32_t log_size, const uint8_t* prev, uint32_t n_cols, const uint32_t* const* cols,
                            uint8_t* out) {
  std::vector<Hash32> pv;
  if (prev) { pv.resize((size_t)2 << log_size); memcpy(pv.data(), prev, pv.size() * 32); }
  std::vector<std::vector<M31>> cv(n_cols, std::vector<M31>((size_t)1 << log_size));
  std::vector<const std::vector<M31>*> refs;
  for (uint32_t c = 0; c < n_cols; ++c) {
    for (size_t i = 0; i < cv[c].size(); ++i) cv[c][i] = M31(cols[c][i]);
    refs.push_back(&cv[c]);
  }
  auto r = commit_on_layer(log_size, prev ? &pv : nullptr, refs);
  memcpy(out, r.data(), r.size() * 32);
}
// QuotientOps::accumulate_numerators over columns of one size (prover/pcs/quotient_ops.rs:19-43).
// Batch b: point at points[8b..8b+8), samples [starts[b], starts[b+1]) of (col_idx, value, rand);
// out receives 4 coordinate columns of 2^log_size per batch, first_linear 4 words per batch.
void oracle_accumulate_numerators(uint32_t log_size, uint32_t n_cols, const uint32_t* const* cols, uint32_t n_batches,
                                  const uint32_t* points, const uint32_t* starts, const uint32_t* col_idx,
                                  const uint32_t* values, const uint32_t* rands, uint32_t* const* out,
                                  uint32_t* first_linear) {
  const size_t n = (size_t)1 << log_size;
  std::vector<std::vector<M31>> cv(n_cols, std::vector<M31>(n));
  std::vector<const std::vector<M31>*> refs;
  for (uint32_t c = 0; c < n_cols; ++c) {
    for (size_t i = 0; i < n; ++i) cv[c][i] = M31(cols[c][i]);
    refs.push_back(&cv[c]);
  }
  for (uint32_t b = 0; b < n_batches; ++b) {
    SampleBatch sb;
    const uint32_t* pt = points + 8 * b;
    sb.point = {QM31::from_u32(pt[0], pt[1], pt[2], pt[3]), QM31::from_u32(pt[4], pt[5], pt[6], pt[7])};
    for (uint32_t j = starts[b]; j < starts[b + 1]; ++j)
      sb.cols.push_back({col_idx[j], QM31::from_u32(values[4 * j], values[4 * j + 1], values[4 * j + 2], values[4 * j + 3]),
                         QM31::from_u32(rands[4 * j], rands[4 * j + 1], rands[4 * j + 2], rands[4 * j + 3])});
    AccNum a = accumulate_batch(sb, refs);
    for (int k = 0; k < 4; ++k) {
      for (size_t i = 0; i < n; ++i) out[4 * b + k][i] = a.partial.c[k][i].v;
      first_linear[4 * b + k] = a.first_linear.coord(k).v;
    }
  }
}
// QuotientOps::compute_quotients_and_combine (quotient_ops.rs:34-43): accumulation a has its sample
// point at points[8a..8a+8), 4 partial columns partials[4a..4a+4) of 2^log_sizes[a] words and its
// first linear term at first_linear[4a..4a+4); out: 4 columns of 2^lifting_log_size.
void oracle_compute_quotients_and_combine(uint32_t n_accs, const uint32_t* points, const uint32_t* const* partials,
                                          const uint32_t* log_sizes, const uint32_t* first_linear,
                                          uint32_t lifting_log_size, uint32_t* const* out) {
  std::vector<AccNum> accs(n_accs);
  for (uint32_t a = 0; a < n_accs; ++a) {
    const uint32_t* pt = points + 8 * a;
    accs[a].point = {QM31::from_u32(pt[0], pt[1], pt[2], pt[3]), QM31::from_u32(pt[4], pt[5], pt[6], pt[7])};
    accs[a].first_linear = QM31::from_u32(first_linear[4 * a], first_linear[4 * a + 1], first_linear[4 * a + 2],
                                          first_linear[4 * a + 3]);
    const size_t n = (size_t)1 << log_sizes[a];
    accs[a].partial.resize(n);
    for (int k = 0; k < 4; ++k)
      for (size_t i = 0; i < n; ++i) accs[a].partial.c[k][i] = M31(partials[4 * a + k][i]);
  }
  SecureCol q = quotients_and_combine(accs, lifting_log_size);
  for (int k = 0; k < 4; ++k)
    for (size_t i = 0; i < q.size(); ++i) out[k][i] = q.c[k][i].v;
}
// ComponentProver::evaluate_constraint_quotients_on_domain for one component
// (crates/constraint-framework/src/prover/component_prover.rs:32-169): for every row of the
// evaluation domain, acc[row] += (sum_k pows[k] C_k(row)) * denom_inv[row >> trace_log]. air 0:
// wide Fibonacci with n_trace columns; air 1: Poseidon2 with ipr instances per row, lookup
// elements (z, alpha) and the LogUp cumsum shift. Columns are evaluations on CanonicCoset(eval_log).
void oracle_eval_constraints_on_domain(uint32_t air, uint32_t ipr, uint32_t trace_log, uint32_t eval_log,
                                       const uint32_t* const* trace, uint32_t n_trace, const uint32_t* const* inter,
                                       uint32_t n_inter, const uint32_t* pows, uint32_t n_constraints,
                                       const uint32_t* denom_inv, const uint32_t z[4], const uint32_t alpha[4],
                                       const uint32_t shift[4], uint32_t* const* acc) {
  const size_t n = (size_t)1 << eval_log;
  Component comp;
  comp.kind = air == 0 ? WIDE_FIB : POSEIDON;
  comp.log_n_rows = trace_log;
  comp.n_cols = n_trace;
  comp.ipr = ipr;
  if (air == 1) {
    comp.le.z = QM31::from_u32(z[0], z[1], z[2], z[3]);
    comp.le.alpha = QM31::from_u32(alpha[0], alpha[1], alpha[2], alpha[3]);
    QM31 cur = QM31::one();
    for (uint32_t i = 0; i < POSEIDON_N_STATE; ++i) { comp.le.alpha_powers[i] = cur; cur = cur * comp.le.alpha; }
  }
  if (comp.n_constraints() != n_constraints) throw std::runtime_error("constraint count");
  std::vector<std::vector<std::vector<M31>>> store(3);
  std::vector<std::vector<const std::vector<M31>*>> refs(3);
  auto load = [&](int t, const uint32_t* const* cols, uint32_t nc) {
    store[t].assign(nc, std::vector<M31>(n));
    for (uint32_t c = 0; c < nc; ++c)
      for (size_t i = 0; i < n; ++i) store[t][c][i] = M31(cols[c][i]);
    for (auto& v : store[t]) refs[t].push_back(&v);
  };
  load(1, trace, n_trace);
  load(2, inter, n_inter);
  std::vector<QM31> pw(n_constraints);
  for (uint32_t k = 0; k < n_constraints; ++k)
    pw[k] = QM31::from_u32(pows[4 * k], pows[4 * k + 1], pows[4 * k + 2], pows[4 * k + 3]);
  const QM31 sh = QM31::from_u32(shift[0], shift[1], shift[2], shift[3]);
  for (size_t row = 0; row < n; ++row) {
    DomainEval e;
    e.cols = &refs;
    e.row = row;
    e.trace_log = trace_log;
    e.eval_log = eval_log;
    e.pows = pw.data();
    e.shift = sh;
    eval_component(e, comp);
    QM31 v = QM31::from_u32(acc[0][row], acc[1][row], acc[2][row], acc[3][row]) + e.row_res * M31(denom_inv[row >> trace_log]);
    for (int k = 0; k < 4; ++k) acc[k][row] = v.coord(k).v;
  }
}
// Root of a mixed-degree Merkle tree over columns of lengths 2^log_sizes[i].
void oracle_merkle_root(uint32_t n_cols, const uint32_t* const* cols, const uint32_t* log_sizes, uint8_t root[32]) {
  std::vector<std::vector<M31>> cv(n_cols);
  std::vector<const std::vector<M31>*> refs;
  for (uint32_t c = 0; c < n_cols; ++c) {
    cv[c].resize((size_t)1 << log_sizes[c]);
    for (size_t i = 0; i < cv[c].size(); ++i) cv[c][i] = M31(cols[c][i]);
    refs.push_back(&cv[c]);
  }
  auto t = merkle_commit(refs);
  memcpy(root, t.root().data(), 32);
}
// dst (4 coordinate columns of 2^(log_src-1)) = dst*alpha^2 + fold(src)
void oracle_fold_circle_into_line(uint32_t log_src, const uint32_t* const* src, uint32_t* const* dst, const uint32_t alpha[4]) {
  SecureCol s, d;
  size_t n = (size_t)1 << log_src;
  s.resize(n); d.resize(n / 2);
  for (int k = 0; k < 4; ++k) {
    for (size_t i = 0; i < n; ++i) s.c[k][i] = M31(src[k][i]);
    for (size_t i = 0; i < n / 2; ++i) d.c[k][i] = M31(dst[k][i]);
  }
  fold_circle_into_line(d, s, canonic_domain(log_src), QM31::from_u32(alpha[0], alpha[1], alpha[2], alpha[3]));
  for (int k = 0; k < 4; ++k) for (size_t i = 0; i < n / 2; ++i) dst[k][i] = d.c[k][i].v;
}
// One fold_line step over the line domain half_odds(log_size) doubled `n_doubles` times.
void oracle_fold_line(uint32_t log_size, uint32_t domain_log_first, const uint32_t* const* src, uint32_t* const* dst,
                      const uint32_t alpha[4]) {
  size_t n = (size_t)1 << log_size;
  SecureCol s; s.resize(n);
  for (int k = 0; k < 4; ++k) for (size_t i = 0; i < n; ++i) s.c[k][i] = M31(src[k][i]);
  Coset dom = Coset::half_odds(domain_log_first).repeated_double(domain_log_first - log_size);
  SecureCol r = fold_line(s, dom, QM31::from_u32(alpha[0], alpha[1], alpha[2], alpha[3]));
  for (int k = 0; k < 4; ++k) for (size_t i = 0; i < n / 2; ++i) dst[k][i] = r.c[k][i].v;
}
uint64_t oracle_grind(const uint8_t digest[32], uint32_t bits) {
  Channel ch;
  memcpy(ch.digest.data(), digest, 32);
  return ch.grind(bits);
}
int oracle_verify_pow(const uint8_t digest[32], uint32_t bits, uint64_t nonce) {
  Channel ch;
  memcpy(ch.digest.data(), digest, 32);
  return ch.verify_pow_nonce(bits, nonce) ? 1 : 0;
}
// Channel known-answer helpers: start from the default channel.
void oracle_channel_mix_u32s(const uint32_t* data, size_t n, uint8_t digest_out[32]) {
  Channel ch;
  ch.mix_u32s(std::vector<uint32_t>(data, data + n));
  memcpy(digest_out, ch.digest.data(), 32);
}
void oracle_channel_mix_u64(uint64_t v, uint8_t digest_out[32]) {
  Channel ch;
  ch.mix_u64(v);
  memcpy(digest_out, ch.digest.data(), 32);
}
// Draws 1 u32 batch then n secure felts; returns n_draws after.
uint32_t oracle_channel_draw_count(uint32_t n_felts) {
  Channel ch;
  ch.draw_u32s();
  ch.draw_secure_felts(n_felts);
  return ch.n_draws;
}
void oracle_random_point(const uint8_t digest[32], uint32_t out[8]) {
  Channel ch;
  memcpy(ch.digest.data(), digest, 32);
  auto p = get_random_point(ch);
  for (int k = 0; k < 4; ++k) { out[k] = p.x.coord(k).v; out[4 + k] = p.y.coord(k).v; }
}
void oracle_gen_wide_fib(uint32_t log_n, uint32_t n_cols, uint32_t mode, uint64_t seed, uint32_t* out) {
  auto t = gen_wide_fib(log_n, n_cols, mode, seed);
  size_t N = (size_t)1 << log_n;
  for (uint32_t c = 0; c < n_cols; ++c) for (size_t r = 0; r < N; ++r) out[c * N + r] = t[c][r].v;
}
void oracle_gen_poseidon(uint32_t log_n, uint32_t ipr, uint32_t mode, uint64_t seed, uint32_t* out) {
  auto pt = gen_poseidon(log_n, ipr, mode, seed);
  size_t N = (size_t)1 << log_n;
  for (size_t c = 0; c < pt.cols.size(); ++c) for (size_t r = 0; r < N; ++r) out[c * N + r] = pt.cols[c][r].v;
}
uint32_t oracle_uniform_m31(uint64_t seed, uint64_t idx) { return uniform_m31(seed, idx); }
// Poseidon2 permutation of one state (16 words in/out), for the MDS known-answer checks.
void oracle_poseidon_permute(uint32_t st[16]) {
  M31 s[16], cols[POSEIDON_COLS_PER_REP];
  for (int i = 0; i < 16; ++i) s[i] = M31::from_u32(st[i]);
  poseidon_permute_trace(s, cols);
  for (int i = 0; i < 16; ++i) st[i] = s[i].v;
}
void oracle_apply_m4(uint32_t x[4]) {
  M31 v[4];
  for (int i = 0; i < 4; ++i) v[i] = M31::from_u32(x[i]);
  apply_m4(v);
  for (int i = 0; i < 4; ++i) x[i] = v[i].v;
}
void oracle_internal_matrix(uint32_t x[16]) {
  M31 v[16];
  for (int i = 0; i < 16; ++i) v[i] = M31::from_u32(x[i]);
  internal_matrix(v);
  for (int i = 0; i < 16; ++i) x[i] = v[i].v;
}
void oracle_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) {
  QM31 r = QM31::from_u32(a[0], a[1], a[2], a[3]) * QM31::from_u32(b[0], b[1], b[2], b[3]);
  for (int k = 0; k < 4; ++k) out[k] = r.coord(k).v;
}
void oracle_qm31_inv(const uint32_t a[4], uint32_t out[4]) {
  QM31 r = inv(QM31::from_u32(a[0], a[1], a[2], a[3]));
  for (int k = 0; k < 4; ++k) out[k] = r.coord(k).v;
}
// Circle point of index `idx` (GEN * idx).
void oracle_index_to_point(uint32_t idx, uint32_t out[2]) {
  auto p = index_to_point(idx);
  out[0] = p.x.v; out[1] = p.y.v;
}
size_t oracle_offset_bit_reversed_index(size_t i, uint32_t dlog, uint32_t elog, long off) {
  return offset_bit_reversed_circle_domain_index(i, dlog, elog, off);
}
void oracle_prefix_sum(uint32_t log_n, uint32_t* v) {
  std::vector<M31> x((size_t)1 << log_n);
  for (size_t i = 0; i < x.size(); ++i) x[i] = M31(v[i]);
  x = inclusive_prefix_sum(x);
  for (size_t i = 0; i < x.size(); ++i) v[i] = x[i].v;
}

}  // extern "C"
