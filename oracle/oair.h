// Oracle (test infrastructure only): the two benchmark AIRs and their evaluators.
//   wide Fibonacci      crates/examples/src/wide_fibonacci/mod.rs:34-83
//   Poseidon2           crates/examples/src/poseidon/mod.rs:38-335 (instances per row is a parameter;
//                       the reference fixes N_LOG_INSTANCES_PER_ROW = 0)
//   LogUp batching      crates/constraint-framework/src/lib.rs:149-258 (finalize_logup_batched),
//                       logup.rs (LookupElements::draw/combine, cumsum shift)
//   LogUp trace         crates/constraint-framework/src/prover/logup.rs:21-200;
//                       crates/stwo/src/prover/backend/simd/prefix_sum.rs:123-140
//   domain evaluation   crates/constraint-framework/src/prover/component_prover.rs:32-169,
//                       simd_domain.rs (mask offsets via offset_bit_reversed_circle_domain_index)
//   point evaluation    crates/constraint-framework/src/{component.rs:253-300,point.rs}
#pragma once
#include <functional>
#include "ofri.h"

namespace oracle {

constexpr uint32_t POSEIDON_N_STATE = 16;
constexpr uint32_t POSEIDON_PARTIAL = 14;
constexpr uint32_t POSEIDON_HALF_FULL = 4;
constexpr uint32_t POSEIDON_COLS_PER_REP = POSEIDON_N_STATE * (1 + 2 * POSEIDON_HALF_FULL) + POSEIDON_PARTIAL;  // 158
constexpr uint32_t POSEIDON_LOG_EXPAND = 2;
constexpr uint32_t POSEIDON_CONST = 1234;

struct LookupElements {
  QM31 z, alpha;
  QM31 alpha_powers[POSEIDON_N_STATE];
  static LookupElements draw(Channel& ch) {
    auto v = ch.draw_secure_felts(2);
    LookupElements le;
    le.z = v[0];
    le.alpha = v[1];
    QM31 cur = QM31::one();
    for (uint32_t i = 0; i < POSEIDON_N_STATE; ++i) { le.alpha_powers[i] = cur; cur = cur * le.alpha; }
    return le;
  }
  template <class F>
  QM31 combine(const F* v) const {
    QM31 acc = QM31::zero();
    for (uint32_t i = 0; i < POSEIDON_N_STATE; ++i) acc = acc + alpha_powers[i] * v[i];
    return acc - z;
  }
};

enum ComponentKind { WIDE_FIB = 0, POSEIDON = 1 };

struct Component {
  ComponentKind kind = WIDE_FIB;
  uint32_t log_n_rows = 0;
  uint32_t n_cols = 0;  // wide-fib sequence length
  uint32_t ipr = 1;     // Poseidon instances per row
  LookupElements le{};
  QM31 claimed_sum = QM31::zero();
  size_t trace_col_start = 0;        // offset in tree 1
  size_t interaction_col_start = 0;  // offset in tree 2

  uint32_t n_trace_cols() const { return kind == WIDE_FIB ? n_cols : ipr * POSEIDON_COLS_PER_REP; }
  uint32_t n_interaction_cols() const { return kind == WIDE_FIB ? 0 : 4 * ipr; }
  uint32_t n_constraints() const {
    return kind == WIDE_FIB ? n_cols - 2 : ipr * (8 * POSEIDON_N_STATE + POSEIDON_PARTIAL) + ipr;
  }
  uint32_t max_constraint_log_degree_bound() const {
    return kind == WIDE_FIB ? log_n_rows + 1 : log_n_rows + POSEIDON_LOG_EXPAND;
  }
  // per tree (0 preprocessed, 1 trace, 2 interaction), per column: mask offsets
  std::vector<std::vector<std::vector<long>>> mask_offsets() const {
    std::vector<std::vector<std::vector<long>>> r(kind == WIDE_FIB ? 2 : 3);
    r[1].assign(n_trace_cols(), {0});
    if (kind == POSEIDON) {
      for (uint32_t i = 0; i < 4 * (ipr - 1); ++i) r[2].push_back({0});
      for (uint32_t i = 0; i < 4; ++i) r[2].push_back({-1, 0});
    }
    return r;
  }
  QM31 cumsum_shift() const { return claimed_sum * inv(M31(1u << log_n_rows)); }
};

// ---- evaluators ----------------------------------------------------------------------------
struct DomainEval {
  using F = M31;
  using EF = QM31;
  const std::vector<std::vector<const std::vector<M31>*>>* cols;  // per tree: component's extended columns
  size_t row;
  uint32_t trace_log, eval_log;
  const QM31* pows;  // constraint k gets pows[k]
  size_t ci = 0;
  size_t col_idx[3] = {0, 0, 0};
  QM31 row_res = QM31::zero();
  QM31 shift;
  M31 next_trace() { return (*(*cols)[1][col_idx[1]++])[row]; }
  M31 at_offset(int tree, size_t c, long off) {
    size_t r = off == 0 ? row : offset_bit_reversed_circle_domain_index(row, trace_log, eval_log, off);
    return (*(*cols)[tree][c])[r];
  }
  std::vector<QM31> next_ext(int tree, const std::vector<long>& offs) {
    M31 v[4][2];
    for (int k = 0; k < 4; ++k) {
      size_t c = col_idx[tree]++;
      for (size_t o = 0; o < offs.size(); ++o) v[k][o] = at_offset(tree, c, offs[o]);
    }
    std::vector<QM31> r;
    for (size_t o = 0; o < offs.size(); ++o) r.push_back(QM31::from_m31(v[0][o], v[1][o], v[2][o], v[3][o]));
    return r;
  }
  void add_f(M31 c) { row_res = row_res + pows[ci++] * c; }
  void add_ef(QM31 c) { row_res = row_res + pows[ci++] * c; }
};

inline QM31 from_partial_evals(QM31 e0, QM31 e1, QM31 e2, QM31 e3) {
  return e0 + e1 * QM31::from_u32(0, 1, 0, 0) + e2 * QM31::from_u32(0, 0, 1, 0) + e3 * QM31::from_u32(0, 0, 0, 1);
}

struct PointEval {
  using F = QM31;
  using EF = QM31;
  const std::vector<std::vector<const std::vector<QM31>*>>* mask;  // per tree: component's column masks
  size_t col_idx[3] = {0, 0, 0};
  QM31* acc;
  QM31 random_coeff, denom_inv, shift;
  QM31 next_trace() { return (*(*mask)[1][col_idx[1]++])[0]; }
  std::vector<QM31> next_ext(int tree, const std::vector<long>& offs) {
    const std::vector<QM31>* m[4];
    for (int k = 0; k < 4; ++k) m[k] = (*mask)[tree][col_idx[tree]++];
    std::vector<QM31> r;
    for (size_t o = 0; o < offs.size(); ++o) r.push_back(from_partial_evals((*m[0])[o], (*m[1])[o], (*m[2])[o], (*m[3])[o]));
    return r;
  }
  void add_f(QM31 c) { *acc = *acc * random_coeff + denom_inv * c; }
  void add_ef(QM31 c) { add_f(c); }
};

template <class F> inline F pow5(F x) { F x2 = x * x; F x4 = x2 * x2; return x4 * x; }
template <class F> inline void apply_m4(F* x) {
  F t0 = x[0] + x[1], t02 = t0 + t0, t1 = x[2] + x[3], t12 = t1 + t1;
  F t2 = x[1] + x[1] + t1, t3 = x[3] + x[3] + t0;
  F t4 = t12 + t12 + t3, t5 = t02 + t02 + t2;
  F t6 = t3 + t5, t7 = t2 + t4;
  x[0] = t6; x[1] = t5; x[2] = t7; x[3] = t4;
}
template <class F> inline void external_matrix(F* s) {
  for (int i = 0; i < 4; ++i) apply_m4(s + 4 * i);
  for (int j = 0; j < 4; ++j) {
    F t = s[j] + s[j + 4] + s[j + 8] + s[j + 12];
    for (int i = 0; i < 4; ++i) s[4 * i + j] = s[4 * i + j] + t;
  }
}
template <class F> inline void internal_matrix(F* s) {
  F sum = s[0];
  for (int i = 1; i < 16; ++i) sum = sum + s[i];
  for (int i = 0; i < 16; ++i) s[i] = s[i] * M31(1u << (i + 1)) + sum;
}

template <class F> inline F add_const(F x, M31 c) { return x + c; }

// Full Poseidon2 permutation on one 16-word state (trace generation).
inline void poseidon_permute_trace(M31* st, M31* out_cols /* 158 */) {
  size_t k = 0;
  for (int i = 0; i < 16; ++i) out_cols[k++] = st[i];
  M31 C(POSEIDON_CONST);
  for (uint32_t r = 0; r < POSEIDON_HALF_FULL; ++r) {
    for (int i = 0; i < 16; ++i) st[i] = st[i] + C;
    external_matrix(st);
    for (int i = 0; i < 16; ++i) { st[i] = pow5(st[i]); out_cols[k++] = st[i]; }
  }
  for (uint32_t r = 0; r < POSEIDON_PARTIAL; ++r) {
    st[0] = st[0] + C;
    internal_matrix(st);
    st[0] = pow5(st[0]);
    out_cols[k++] = st[0];
  }
  for (uint32_t r = 0; r < POSEIDON_HALF_FULL; ++r) {
    for (int i = 0; i < 16; ++i) st[i] = st[i] + C;
    external_matrix(st);
    for (int i = 0; i < 16; ++i) { st[i] = pow5(st[i]); out_cols[k++] = st[i]; }
  }
}

template <class E>
inline void eval_component(E& e, const Component& c) {
  using F = typename E::F;
  if (c.kind == WIDE_FIB) {
    F a = e.next_trace(), b = e.next_trace();
    for (uint32_t i = 2; i < c.n_cols; ++i) {
      F x = e.next_trace();
      e.add_f(x - (a * a + b * b));
      a = b;
      b = x;
    }
    return;
  }
  M31 C(POSEIDON_CONST);
  std::vector<std::pair<QM31, QM31>> fracs;  // (numerator, denominator)
  for (uint32_t rep = 0; rep < c.ipr; ++rep) {
    F st[16], init[16];
    for (int i = 0; i < 16; ++i) { st[i] = e.next_trace(); init[i] = st[i]; }
    for (uint32_t r = 0; r < POSEIDON_HALF_FULL; ++r) {
      for (int i = 0; i < 16; ++i) st[i] = add_const(st[i], C);
      external_matrix(st);
      for (int i = 0; i < 16; ++i) st[i] = pow5(st[i]);
      for (int i = 0; i < 16; ++i) { F m = e.next_trace(); e.add_f(st[i] - m); st[i] = m; }
    }
    for (uint32_t r = 0; r < POSEIDON_PARTIAL; ++r) {
      st[0] = add_const(st[0], C);
      internal_matrix(st);
      st[0] = pow5(st[0]);
      F m = e.next_trace();
      e.add_f(st[0] - m);
      st[0] = m;
    }
    for (uint32_t r = 0; r < POSEIDON_HALF_FULL; ++r) {
      for (int i = 0; i < 16; ++i) st[i] = add_const(st[i], C);
      external_matrix(st);
      for (int i = 0; i < 16; ++i) st[i] = pow5(st[i]);
      for (int i = 0; i < 16; ++i) { F m = e.next_trace(); e.add_f(st[i] - m); st[i] = m; }
    }
    fracs.push_back({QM31::one(), c.le.combine(init)});
    fracs.push_back({-QM31::one(), c.le.combine(st)});
  }
  // finalize_logup_in_pairs: batch b = {fracs[2b], fracs[2b+1]}
  QM31 prev_col = QM31::zero();
  for (uint32_t b = 0; b + 1 < c.ipr; ++b) {
    auto f0 = fracs[2 * b], f1 = fracs[2 * b + 1];
    QM31 num = f1.second * f0.first + f0.second * f1.first, den = f0.second * f1.second;
    QM31 cur = e.next_ext(2, {0})[0];
    QM31 diff = cur - prev_col;
    prev_col = cur;
    e.add_ef(diff * den - num);
  }
  auto f0 = fracs[2 * (c.ipr - 1)], f1 = fracs[2 * (c.ipr - 1) + 1];
  QM31 num = f1.second * f0.first + f0.second * f1.first, den = f0.second * f1.second;
  auto m = e.next_ext(2, {-1, 0});
  QM31 diff = m[1] - m[0] - prev_col;
  QM31 shifted = diff + e.shift;
  e.add_ef(shifted * den - num);
}

}  // namespace oracle
