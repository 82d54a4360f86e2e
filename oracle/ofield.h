// Oracle (test infrastructure only): scalar CPU restatement of the reference's field,
// circle-group and domain arithmetic. Only tests/, the smoke check and bench.py's CPU
// baseline leg may load code built from this directory; the product never does.
//
// Reference semantics followed (paths relative to crates/stwo/src):
//   M31  core/fields/m31.rs:14,46-108   (P = 2^31-1, reduce, mul, inverse by pow(P-2))
//   CM31 core/fields/cm31.rs:44-54      (a + bi, i^2 = -1)
//   QM31 core/fields/qm31.rs:14-21,78-88,120-130  (x + yu, u^2 = 2+i; inverse)
//   circle group core/circle.rs (generator (2,1268011823), index arithmetic mod 2^31)
//   cosets/domains core/circle.rs Coset, core/poly/circle/{domain,canonic}.rs
// All values are kept canonical in [0,P).
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>
#include <cassert>

namespace oracle {

constexpr uint32_t P = 0x7fffffffu;

struct M31 {
  uint32_t v = 0;
  M31() = default;
  explicit constexpr M31(uint32_t x) : v(x) {}
  static M31 reduce64(uint64_t x) {  // m31.rs reduce(): valid for x < 2^62
    uint64_t r = (x & P) + (x >> 31);
    r = (r & P) + (r >> 31);
    uint32_t y = (uint32_t)r;
    return M31(y >= P ? y - P : y);
  }
  static M31 from_u32(uint32_t x) { return reduce64(x); }
  bool operator==(M31 o) const { return v == o.v; }
  bool operator!=(M31 o) const { return v != o.v; }
  bool is_zero() const { return v == 0; }
};
inline M31 operator+(M31 a, M31 b) { uint32_t s = a.v + b.v; return M31(s >= P ? s - P : s); }
inline M31 operator-(M31 a, M31 b) { return M31(a.v >= b.v ? a.v - b.v : a.v + P - b.v); }
inline M31 operator-(M31 a) { return M31(a.v == 0 ? 0 : P - a.v); }
inline M31 operator*(M31 a, M31 b) { return M31::reduce64((uint64_t)a.v * b.v); }
inline M31& operator+=(M31& a, M31 b) { a = a + b; return a; }
inline M31& operator-=(M31& a, M31 b) { a = a - b; return a; }
inline M31& operator*=(M31& a, M31 b) { a = a * b; return a; }
inline M31 pow(M31 a, uint64_t e) { M31 r(1); while (e) { if (e & 1) r = r * a; a = a * a; e >>= 1; } return r; }
inline M31 inv(M31 a) { assert(a.v != 0); return pow(a, P - 2); }

struct CM31 {
  M31 a, b;
  CM31() = default;
  constexpr CM31(M31 x, M31 y) : a(x), b(y) {}
  explicit CM31(M31 x) : a(x), b(0) {}
  bool operator==(const CM31& o) const { return a == o.a && b == o.b; }
  bool operator!=(const CM31& o) const { return !(*this == o); }
  bool is_zero() const { return a.is_zero() && b.is_zero(); }
};
inline CM31 operator+(CM31 x, CM31 y) { return CM31(x.a + y.a, x.b + y.b); }
inline CM31 operator-(CM31 x, CM31 y) { return CM31(x.a - y.a, x.b - y.b); }
inline CM31 operator-(CM31 x) { return CM31(-x.a, -x.b); }
inline CM31 operator*(CM31 x, CM31 y) { return CM31(x.a * y.a - x.b * y.b, x.a * y.b + x.b * y.a); }
inline CM31 operator*(CM31 x, M31 s) { return CM31(x.a * s, x.b * s); }
inline CM31 inv(CM31 x) {  // (a - bi) / (a^2 + b^2)
  M31 n = inv(x.a * x.a + x.b * x.b);
  return CM31(x.a * n, -x.b * n);
}

struct QM31 {
  CM31 a, b;  // a + b u
  QM31() = default;
  constexpr QM31(CM31 x, CM31 y) : a(x), b(y) {}
  explicit QM31(M31 x) : a(x, M31(0)), b(M31(0), M31(0)) {}
  static QM31 from_u32(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
    return QM31(CM31(M31(w0), M31(w1)), CM31(M31(w2), M31(w3)));
  }
  static QM31 from_m31(M31 w0, M31 w1, M31 w2, M31 w3) { return QM31(CM31(w0, w1), CM31(w2, w3)); }
  M31 coord(int i) const { return i == 0 ? a.a : i == 1 ? a.b : i == 2 ? b.a : b.b; }
  bool operator==(const QM31& o) const { return a == o.a && b == o.b; }
  bool operator!=(const QM31& o) const { return !(*this == o); }
  bool is_zero() const { return a.is_zero() && b.is_zero(); }
  static QM31 zero() { return QM31(M31(0)); }
  static QM31 one() { return QM31(M31(1)); }
};
inline const CM31 QM31_R(M31(2), M31(1));
inline QM31 operator+(QM31 x, QM31 y) { return QM31(x.a + y.a, x.b + y.b); }
inline QM31 operator-(QM31 x, QM31 y) { return QM31(x.a - y.a, x.b - y.b); }
inline QM31 operator-(QM31 x) { return QM31(-x.a, -x.b); }
inline QM31 operator*(QM31 x, QM31 y) {  // qm31.rs:78-88
  return QM31(x.a * y.a + QM31_R * (x.b * y.b), x.a * y.b + x.b * y.a);
}
inline QM31 operator*(QM31 x, M31 s) { return QM31(x.a * s, x.b * s); }
inline QM31 operator*(M31 s, QM31 x) { return x * s; }
inline QM31 operator+(QM31 x, M31 s) { return QM31(CM31(x.a.a + s, x.a.b), x.b); }
inline QM31 operator-(QM31 x, M31 s) { return QM31(CM31(x.a.a - s, x.a.b), x.b); }
inline QM31 mul_cm31(QM31 x, CM31 c) { return QM31(x.a * c, x.b * c); }
inline QM31& operator+=(QM31& x, QM31 y) { x = x + y; return x; }
inline QM31& operator-=(QM31& x, QM31 y) { x = x - y; return x; }
inline QM31& operator*=(QM31& x, QM31 y) { x = x * y; return x; }
inline QM31 inv(QM31 x) {  // qm31.rs inverse: (a - bu) / (a^2 - (2+i) b^2)
  CM31 b2 = x.b * x.b;
  CM31 ib2(-b2.b, b2.a);
  CM31 denom = x.a * x.a - (b2 + b2 + ib2);
  CM31 di = inv(denom);
  return QM31(x.a * di, -x.b * di);
}
inline QM31 conj(QM31 x) { return QM31(x.a, -x.b); }  // conjugate w.r.t. CM31
inline QM31 square(QM31 x) { return x * x; }

// ---- circle points -------------------------------------------------------------------
template <class F>
struct CirclePoint {
  F x, y;
};
template <class F> inline F field_one();
template <> inline M31 field_one<M31>() { return M31(1); }
template <> inline QM31 field_one<QM31>() { return QM31::one(); }
template <class F> inline F field_zero();
template <> inline M31 field_zero<M31>() { return M31(0); }
template <> inline QM31 field_zero<QM31>() { return QM31::zero(); }

template <class F>
inline CirclePoint<F> cp_add(const CirclePoint<F>& p, const CirclePoint<F>& q) {
  return {p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x};
}
template <class F>
inline CirclePoint<F> cp_double(const CirclePoint<F>& p) { return cp_add(p, p); }
template <class F>
inline F double_x(F x) { F s = x * x; return s + s - field_one<F>(); }
template <class F>
inline CirclePoint<F> cp_conj(const CirclePoint<F>& p) { return {p.x, -p.y}; }
template <class F>
inline CirclePoint<F> cp_repeated_double(CirclePoint<F> p, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) p = cp_double(p);
  return p;
}
template <class F>
inline CirclePoint<F> cp_zero() { return {field_one<F>(), field_zero<F>()}; }
template <class F>
inline CirclePoint<F> cp_mul(CirclePoint<F> p, unsigned __int128 s) {
  CirclePoint<F> r = cp_zero<F>();
  while (s) { if (s & 1) r = cp_add(r, p); p = cp_double(p); s >>= 1; }
  return r;
}
inline CirclePoint<QM31> to_ef(const CirclePoint<M31>& p) { return {QM31(p.x), QM31(p.y)}; }
inline bool operator==(const CirclePoint<QM31>& p, const CirclePoint<QM31>& q) { return p.x == q.x && p.y == q.y; }
inline bool operator==(const CirclePoint<M31>& p, const CirclePoint<M31>& q) { return p.x == q.x && p.y == q.y; }

constexpr uint32_t M31_CIRCLE_LOG_ORDER = 31;
inline const CirclePoint<M31> M31_CIRCLE_GEN{M31(2), M31(1268011823)};

// Circle point index: an element of Z / 2^31, standing for GEN * index.
inline uint32_t idx_reduce(uint64_t i) { return (uint32_t)(i & ((1ull << 31) - 1)); }
inline uint32_t subgroup_gen(uint32_t log_size) { return 1u << (M31_CIRCLE_LOG_ORDER - log_size); }
inline CirclePoint<M31> index_to_point(uint32_t idx) { return cp_mul(M31_CIRCLE_GEN, idx); }

struct Coset {
  uint32_t initial_index;
  uint32_t step_size;
  uint32_t log_size;
  CirclePoint<M31> initial;
  CirclePoint<M31> step;
  static Coset make(uint32_t initial_index, uint32_t log_size) {
    Coset c;
    c.initial_index = idx_reduce(initial_index);
    c.step_size = subgroup_gen(log_size);
    c.log_size = log_size;
    c.initial = index_to_point(c.initial_index);
    c.step = index_to_point(c.step_size);
    return c;
  }
  static Coset odds(uint32_t log_size) { return make(subgroup_gen(log_size + 1), log_size); }
  static Coset half_odds(uint32_t log_size) { return make(subgroup_gen(log_size + 2), log_size); }
  size_t size() const { return (size_t)1 << log_size; }
  uint32_t index_at(size_t i) const { return idx_reduce((uint64_t)initial_index + (uint64_t)step_size * i); }
  CirclePoint<M31> at(size_t i) const { return index_to_point(index_at(i)); }
  Coset doubled() const {
    Coset c;
    c.initial_index = idx_reduce((uint64_t)initial_index * 2);
    c.step_size = idx_reduce((uint64_t)step_size * 2);
    c.log_size = log_size > 0 ? log_size - 1 : 0;
    c.initial = cp_double(initial);
    c.step = cp_double(step);
    return c;
  }
  Coset repeated_double(uint32_t n) const { Coset c = *this; for (uint32_t i = 0; i < n; ++i) c = c.doubled(); return c; }
  // All points, iterated by repeated addition (core/circle.rs CosetIterator).
  std::vector<CirclePoint<M31>> points() const {
    std::vector<CirclePoint<M31>> r(size());
    CirclePoint<M31> cur = initial;
    for (size_t i = 0; i < size(); ++i) { r[i] = cur; cur = cp_add(cur, step); }
    return r;
  }
};

// CircleDomain = half_coset ∪ conjugate(half_coset)  (core/poly/circle/domain.rs:21-95)
struct CircleDomain {
  Coset half_coset;
  uint32_t log_size() const { return half_coset.log_size + 1; }
  size_t size() const { return (size_t)1 << log_size(); }
  uint32_t index_at(size_t i) const {
    size_t h = half_coset.size();
    if (i < h) return half_coset.index_at(i);
    return idx_reduce((1ull << 31) - half_coset.index_at(i - h));
  }
  CirclePoint<M31> at(size_t i) const { return index_to_point(index_at(i)); }
};
inline CircleDomain canonic_domain(uint32_t log_size) { return CircleDomain{Coset::half_odds(log_size - 1)}; }
inline Coset canonic_coset(uint32_t log_size) { return Coset::odds(log_size); }

inline size_t bit_reverse_index(size_t i, uint32_t log_size) {
  if (log_size == 0) return i;
  size_t r = 0;
  for (uint32_t b = 0; b < log_size; ++b) r |= ((i >> b) & 1) << (log_size - 1 - b);
  return r;
}
template <class T>
inline void bit_reverse(std::vector<T>& v) {
  size_t n = v.size();
  uint32_t lg = 0;
  while (((size_t)1 << lg) < n) ++lg;
  for (size_t i = 0; i < n; ++i) {
    size_t j = bit_reverse_index(i, lg);
    if (j > i) std::swap(v[i], v[j]);
  }
}

// core/utils.rs offset_bit_reversed_circle_domain_index
inline size_t offset_bit_reversed_circle_domain_index(size_t i, uint32_t domain_log_size, uint32_t eval_log_size,
                                                      long offset) {
  long prev = (long)bit_reverse_index(i, eval_log_size);
  long half = 1l << (eval_log_size - 1);
  long step = offset * (1l << (eval_log_size - domain_log_size - 1));
  auto rem = [](long a, long m) { long r = a % m; return r < 0 ? r + m : r; };
  if (prev < half) prev = rem(prev + step, half);
  else prev = rem(prev - step, half) + half;
  return bit_reverse_index((size_t)prev, eval_log_size);
}

// core/constraints.rs coset_vanishing
template <class F>
inline F coset_vanishing(const Coset& coset, CirclePoint<F> p) {
  CirclePoint<F> neg_init{F(coset.initial.x), -F(coset.initial.y)};
  CirclePoint<M31> half_step = index_to_point(coset.step_size >> 1);
  p = cp_add(p, neg_init);
  p = cp_add(p, CirclePoint<F>{F(half_step.x), F(half_step.y)});
  F x = p.x;
  for (uint32_t i = 1; i < coset.log_size; ++i) x = double_x(x);
  return x;
}

template <class T>
inline std::vector<T> batch_inverse(const std::vector<T>& v) {
  std::vector<T> r(v.size());
  if (v.empty()) return r;
  r[0] = v[0];
  for (size_t i = 1; i < v.size(); ++i) r[i] = r[i - 1] * v[i];
  T cur = inv(r.back());
  for (size_t i = v.size() - 1; i > 0; --i) { r[i] = r[i - 1] * cur; cur = cur * v[i]; }
  r[0] = cur;
  return r;
}

}  // namespace oracle
