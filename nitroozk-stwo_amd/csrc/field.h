// M31 / CM31 / QM31 arithmetic in 32-bit lanes, usable from host and device code.
//
// Semantics follow crates/stwo/src/core/fields/{m31.rs:14,46-108, cm31.rs:44-54,
// qm31.rs:14-21,78-88}: P = 2^31 - 1, CM31 = M31[i]/(i^2 + 1), QM31 = CM31[u]/(u^2 - 2 - i).
// Every value produced here is canonical, in [0, P): the reference hashes raw stored words
// (prover/backend/simd/blake2s.rs:128,136), so a committed word must never be P.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define SHD __host__ __device__ __forceinline__

// Raises the issue priority of the calling wave above default-priority waves of kernels running
// beside it on another stream (the trace extension overlaps the commitment phase): the kernels on
// the proof's critical path call it first.
#define CRITICAL_PATH_WAVE() __builtin_amdgcn_s_setprio(1)

// Global-memory view of a pointer. Pointers read from memory (column tables, descriptors) are
// generic to the compiler, which then accesses through them with FLAT instructions; those also
// count against the LDS and scalar-load wait counter, so that every LDS wait waits for them too.
// Accesses through gmem() pointers compile to global loads and stores.
#define GLOBAL_AS __attribute__((address_space(1)))
// Native vectors for wide loads and stores through gmem() pointers (HIP's uint2/uint4 classes have
// no constructors from address-space-qualified objects).
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ GLOBAL_AS T* gmem(T* p) {
  return (GLOBAL_AS T*)p;
}

namespace sh {

constexpr uint32_t P = 0x7fffffffu;

SHD uint32_t m31_add(uint32_t a, uint32_t b) { uint32_t s = a + b; return s >= P ? s - P : s; }
// a - b wraps above 2^31 exactly when a < b, and then a - b + P (mod 2^32) is the smaller value.
SHD uint32_t m31_sub(uint32_t a, uint32_t b) {
  uint32_t d = a - b, e = d + P;
  return d < e ? d : e;
}
SHD uint32_t m31_neg(uint32_t a) { return a ? P - a : 0; }
// Folds a value below 2^62 + 2^32 into [0, P).
SHD uint32_t m31_reduce64(uint64_t x) {
  uint64_t r = (x & P) + (x >> 31);          // < 2^32 + 2^31
  uint32_t y = (uint32_t)((r & P) + (r >> 31));  // < 2^31 + 2
  return y >= P ? y - P : y;
}
SHD uint32_t m31_mul(uint32_t a, uint32_t b) {
  uint64_t p = (uint64_t)a * b;                     // < 2^62
  uint32_t lo = (uint32_t)(p & P), hi = (uint32_t)(p >> 31);  // both < 2^31
  uint32_t s = lo + hi;                              // < 2^32
  return s >= P ? s - P : s;                          // lo + hi < 2P, so one subtraction
}
SHD uint32_t m31_sqr(uint32_t a) { return m31_mul(a, a); }
// a * 2^k for 0 < k < 31: multiplying by 2 modulo 2^31 - 1 rotates the 31-bit word.
SHD uint32_t m31_rotl(uint32_t a, uint32_t k) { return ((a << k) & P) | (a >> (31 - k)); }
SHD uint32_t m31_pow(uint32_t a, uint32_t e) {
  uint32_t r = 1;
  while (e) { if (e & 1) r = m31_mul(r, a); a = m31_mul(a, a); e >>= 1; }
  return r;
}
// a^(P-2) with an addition chain: P - 2 = 2^31 - 3 = 0b111...1101.
SHD uint32_t m31_inv(uint32_t a) {
  uint32_t t1 = a;
  uint32_t t2 = m31_mul(m31_sqr(t1), t1);            // a^3  (2 ones)
  uint32_t t4 = m31_mul(m31_sqr(m31_sqr(t2)), t2);   // a^(2^4-1)
  uint32_t t8 = t4;
  for (int i = 0; i < 4; ++i) t8 = m31_sqr(t8);
  t8 = m31_mul(t8, t4);                              // a^(2^8-1)
  uint32_t t16 = t8;
  for (int i = 0; i < 8; ++i) t16 = m31_sqr(t16);
  t16 = m31_mul(t16, t8);                            // a^(2^16-1)
  uint32_t t28 = t16;
  for (int i = 0; i < 8; ++i) t28 = m31_sqr(t28);
  t28 = m31_mul(t28, t8);                            // a^(2^24-1)
  for (int i = 0; i < 4; ++i) t28 = m31_sqr(t28);
  t28 = m31_mul(t28, t4);                            // a^(2^28-1)
  uint32_t r = m31_sqr(t28);                         // a^(2^29-2)
  r = m31_mul(r, t1);                                // a^(2^29-1)
  r = m31_sqr(m31_sqr(r));                           // a^(2^31-4)
  return m31_mul(r, t1);                             // a^(2^31-3)
}

struct cm31 { uint32_t a, b; };
struct qm31 { uint32_t a, b, c, d; };  // (a + b i) + (c + d i) u

SHD cm31 cm_add(cm31 x, cm31 y) { return {m31_add(x.a, y.a), m31_add(x.b, y.b)}; }
SHD cm31 cm_sub(cm31 x, cm31 y) { return {m31_sub(x.a, y.a), m31_sub(x.b, y.b)}; }
SHD cm31 cm_neg(cm31 x) { return {m31_neg(x.a), m31_neg(x.b)}; }
SHD cm31 cm_mul(cm31 x, cm31 y) {
  // (a + bi)(c + di) = (ac - bd) + (ad + bc) i
  uint64_t ac = (uint64_t)x.a * y.a, bd = (uint64_t)x.b * y.b;
  uint64_t ad = (uint64_t)x.a * y.b, bc = (uint64_t)x.b * y.a;
  uint32_t re = m31_sub(m31_reduce64(ac), m31_reduce64(bd));
  return {re, m31_reduce64(ad + bc)};
}
SHD cm31 cm_mul_m31(cm31 x, uint32_t s) { return {m31_mul(x.a, s), m31_mul(x.b, s)}; }
SHD cm31 cm_inv(cm31 x) {
  uint32_t n = m31_inv(m31_reduce64((uint64_t)x.a * x.a + (uint64_t)x.b * x.b));
  return {m31_mul(x.a, n), m31_mul(m31_neg(x.b), n)};
}

SHD qm31 qm_make(cm31 x, cm31 y) { return {x.a, x.b, y.a, y.b}; }
SHD cm31 qm_lo(qm31 q) { return {q.a, q.b}; }
SHD cm31 qm_hi(qm31 q) { return {q.c, q.d}; }
SHD qm31 qm_zero() { return {0, 0, 0, 0}; }
SHD qm31 qm_one() { return {1, 0, 0, 0}; }
SHD qm31 qm_from_m31(uint32_t v) { return {v, 0, 0, 0}; }
SHD bool qm_eq(qm31 x, qm31 y) { return x.a == y.a && x.b == y.b && x.c == y.c && x.d == y.d; }
SHD qm31 qm_add(qm31 x, qm31 y) { return {m31_add(x.a, y.a), m31_add(x.b, y.b), m31_add(x.c, y.c), m31_add(x.d, y.d)}; }
SHD qm31 qm_sub(qm31 x, qm31 y) { return {m31_sub(x.a, y.a), m31_sub(x.b, y.b), m31_sub(x.c, y.c), m31_sub(x.d, y.d)}; }
SHD qm31 qm_neg(qm31 x) { return {m31_neg(x.a), m31_neg(x.b), m31_neg(x.c), m31_neg(x.d)}; }
SHD qm31 qm_add_m31(qm31 x, uint32_t s) { return {m31_add(x.a, s), x.b, x.c, x.d}; }
SHD qm31 qm_mul_m31(qm31 x, uint32_t s) { return {m31_mul(x.a, s), m31_mul(x.b, s), m31_mul(x.c, s), m31_mul(x.d, s)}; }
SHD qm31 qm_mul_cm31(qm31 x, cm31 c) { cm31 l = cm_mul(qm_lo(x), c), h = cm_mul(qm_hi(x), c); return qm_make(l, h); }
SHD qm31 qm_mul(qm31 x, qm31 y) {
  // (x0 + x1 u)(y0 + y1 u) = (x0 y0 + (2 + i) x1 y1) + (x0 y1 + x1 y0) u
  cm31 x0 = qm_lo(x), x1 = qm_hi(x), y0 = qm_lo(y), y1 = qm_hi(y);
  cm31 t = cm_mul(x1, y1);
  cm31 rt = {m31_sub(m31_add(t.a, t.a), t.b), m31_add(m31_add(t.b, t.b), t.a)};  // (2 + i) t
  cm31 lo = cm_add(cm_mul(x0, y0), rt);
  cm31 hi = cm_add(cm_mul(x0, y1), cm_mul(x1, y0));
  return qm_make(lo, hi);
}
SHD qm31 qm_sqr(qm31 x) { return qm_mul(x, x); }
SHD qm31 qm_inv(qm31 x) {
  // (a + bu)^-1 = (a - bu) / (a^2 - (2 + i) b^2)
  cm31 a = qm_lo(x), b = qm_hi(x);
  cm31 b2 = cm_mul(b, b);
  cm31 rb2 = {m31_sub(m31_add(b2.a, b2.a), b2.b), m31_add(m31_add(b2.b, b2.b), b2.a)};
  cm31 den = cm_sub(cm_mul(a, a), rb2);
  cm31 di = cm_inv(den);
  return qm_make(cm_mul(a, di), cm_mul(cm_neg(b), di));
}
SHD qm31 qm_conj(qm31 x) { return {x.a, x.b, m31_neg(x.c), m31_neg(x.d)}; }

// Lazy accumulator for sums of QM31 x M31 products: each product is folded below 2^32 and
// summed in 64 bits, so up to 2^31 terms fit before the final reduction.
struct qacc { uint64_t a, b, c, d; };
SHD qacc qacc_zero() { return {0, 0, 0, 0}; }
SHD uint64_t fold31(uint64_t p) { return (p & P) + (p >> 31); }
SHD void qacc_mac(qacc& acc, qm31 c, uint32_t v) {
  acc.a += fold31((uint64_t)c.a * v);
  acc.b += fold31((uint64_t)c.b * v);
  acc.c += fold31((uint64_t)c.c * v);
  acc.d += fold31((uint64_t)c.d * v);
}
// acc += c * v with no reduction: products are below 2^62, so at most 3 of them may follow a
// qacc_fold (which leaves each coordinate below 2^33) before the next fold.
SHD void qacc_mad(qacc& acc, qm31 c, uint32_t v) {
  acc.a += (uint64_t)c.a * v;
  acc.b += (uint64_t)c.b * v;
  acc.c += (uint64_t)c.c * v;
  acc.d += (uint64_t)c.d * v;
}
SHD void qacc_fold(qacc& acc) {
  acc.a = fold31(acc.a);
  acc.b = fold31(acc.b);
  acc.c = fold31(acc.c);
  acc.d = fold31(acc.d);
}
SHD qm31 qacc_reduce(const qacc& acc) {
  return {m31_reduce64(fold31(acc.a)), m31_reduce64(fold31(acc.b)), m31_reduce64(fold31(acc.c)), m31_reduce64(fold31(acc.d))};
}

// ---- circle points ------------------------------------------------------------------------
struct cpm31 { uint32_t x, y; };
struct cpqm31 { qm31 x, y; };

SHD cpm31 cp_add(cpm31 p, cpm31 q) {
  return {m31_sub(m31_mul(p.x, q.x), m31_mul(p.y, q.y)), m31_add(m31_mul(p.x, q.y), m31_mul(p.y, q.x))};
}
SHD cpm31 cp_double(cpm31 p) { return cp_add(p, p); }
SHD cpm31 cp_conj(cpm31 p) { return {p.x, m31_neg(p.y)}; }
SHD cpqm31 cpq_add(cpqm31 p, cpqm31 q) {
  return {qm_sub(qm_mul(p.x, q.x), qm_mul(p.y, q.y)), qm_add(qm_mul(p.x, q.y), qm_mul(p.y, q.x))};
}
SHD cpqm31 cpq_double(cpqm31 p) { return cpq_add(p, p); }
SHD cpqm31 cpq_from(cpm31 p) { return {qm_from_m31(p.x), qm_from_m31(p.y)}; }
SHD qm31 qm_double_x(qm31 x) { qm31 s = qm_sqr(x); return qm_sub(qm_add(s, s), qm_one()); }
SHD uint32_t m31_double_x(uint32_t x) { uint32_t s = m31_sqr(x); return m31_sub(m31_add(s, s), 1); }

// Circle generator (2, 1268011823) of order 2^31 (core/circle.rs M31_CIRCLE_GEN).
SHD cpm31 circle_gen() { return {2u, 1268011823u}; }
SHD cpm31 index_to_point(uint32_t idx) {
  cpm31 r = {1, 0}, c = circle_gen();
  idx &= 0x7fffffffu;
  while (idx) { if (idx & 1) r = cp_add(r, c); c = cp_double(c); idx >>= 1; }
  return r;
}

SHD uint32_t bitrev(uint32_t i, uint32_t log_n) { return log_n ? (__builtin_bitreverse32(i) >> (32 - log_n)) : i; }

}  // namespace sh
