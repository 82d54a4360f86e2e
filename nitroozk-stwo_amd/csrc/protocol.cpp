// Host-side protocol pieces shared by the device-resident and the per-call provers.
#include "protocol.h"

#include <algorithm>
#include <cstring>

namespace sh {

namespace {

// ---- proof encoding (little endian, counts as u32) ---------------------------------------------
struct Writer {
  std::vector<uint8_t> b;
  size_t n = 0;
  Writer() { b.resize(1 << 16); }
  uint8_t* room(size_t k) {
    if (n + k > b.size()) b.resize(std::max(2 * b.size(), n + k));
    uint8_t* r = b.data() + n;
    n += k;
    return r;
  }
  void u32(uint32_t v) {  // the host is little endian, as is the encoding
    std::memcpy(room(4), &v, 4);
  }
  void u64(uint64_t v) { u32((uint32_t)v); u32((uint32_t)(v >> 32)); }
  void hash(const Hash& h) { std::memcpy(room(32), h.data(), 32); }
  void q(const qm31& x) { uint32_t w[4] = {x.a, x.b, x.c, x.d}; std::memcpy(room(16), w, 16); }
  std::vector<uint8_t> take() { b.resize(n); return std::move(b); }
};
void write_dec(Writer& w, const MerkleDec& d) {
  w.u32((uint32_t)d.hash_witness.size());
  for (auto& h : d.hash_witness) w.hash(h);
  w.u32((uint32_t)d.column_witness.size());
  for (auto v : d.column_witness) w.u32(v);
}
void write_layer(Writer& w, const FriLayer& l) {
  w.u32((uint32_t)l.witness.size());
  for (auto& x : l.witness) w.q(x);
  write_dec(w, l.dec);
  w.hash(l.commitment);
}


qm31 pow5q(qm31 x) { qm31 x2 = qm_sqr(x); return qm_mul(qm_sqr(x2), x); }
void m4q(qm31* x) {
  qm31 t0 = qm_add(x[0], x[1]), t02 = qm_add(t0, t0), t1 = qm_add(x[2], x[3]), t12 = qm_add(t1, t1);
  qm31 t2 = qm_add(qm_add(x[1], x[1]), t1), t3 = qm_add(qm_add(x[3], x[3]), t0);
  qm31 t4 = qm_add(qm_add(t12, t12), t3), t5 = qm_add(qm_add(t02, t02), t2);
  x[0] = qm_add(t3, t5); x[1] = t5; x[2] = qm_add(t2, t4); x[3] = t4;
}
void ext_q(qm31* s) {
  for (int i = 0; i < 4; ++i) m4q(s + 4 * i);
  for (int j = 0; j < 4; ++j) {
    qm31 t = qm_add(qm_add(s[j], s[j + 4]), qm_add(s[j + 8], s[j + 12]));
    for (int i = 0; i < 4; ++i) s[4 * i + j] = qm_add(s[4 * i + j], t);
  }
}
void int_q(qm31* s) {
  qm31 sum = s[0];
  for (int i = 1; i < 16; ++i) sum = qm_add(sum, s[i]);
  for (int i = 0; i < 16; ++i) s[i] = qm_add(qm_mul_m31(s[i], 1u << (i + 1)), sum);
}
qm31 combine_q(const Component& c, const qm31* v) {
  qm31 acc = qm_zero();
  for (int i = 0; i < 16; ++i) acc = qm_add(acc, qm_mul(c.alpha_pow[i], v[i]));
  return qm_sub(acc, c.z);
}

}  // namespace

std::vector<size_t> sorted_unique(std::vector<size_t> v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}
std::vector<size_t> fold_queries(const std::vector<size_t>& q, uint32_t n) {
  std::vector<size_t> r;
  for (size_t x : q) { size_t y = x >> n; if (r.empty() || r.back() != y) r.push_back(y); }
  return r;
}
// core/pcs/utils.rs prepare_preprocessed_query_positions
std::vector<size_t> tree_local_positions(const std::vector<size_t>& q, uint32_t max_log, uint32_t tree_max) {
  std::vector<size_t> r;
  if (tree_max == 0) return r;
  if (max_log < tree_max) {
    for (size_t p : q) r.push_back(((p >> 1) << (tree_max - max_log + 1)) + (p & 1));
    return r;
  }
  for (size_t p : q) r.push_back(((p >> (max_log - tree_max + 1)) << 1) + (p & 1));
  return r;
}

// Columns in the order MerkleProver hashes them: stable by decreasing size.
std::vector<size_t> merkle_order(const std::vector<uint32_t>& logs) {
  std::vector<size_t> o(logs.size());
  for (size_t i = 0; i < o.size(); ++i) o[i] = i;
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) { return logs[a] > logs[b]; });
  return o;
}

qm31 qm_pow_u(qm31 b, uint64_t e) { qm31 r = qm_one(); while (e) { if (e & 1) r = qm_mul(r, b); b = qm_mul(b, b); e >>= 1; } return r; }

bool point_less(const cpqm31& a, const cpqm31& b) {
  uint32_t x[8] = {a.x.a, a.x.b, a.x.c, a.x.d, a.y.a, a.y.b, a.y.c, a.y.d};
  uint32_t y[8] = {b.x.a, b.x.b, b.x.c, b.x.d, b.y.a, b.y.b, b.y.c, b.y.d};
  for (int i = 0; i < 8; ++i) if (x[i] != y[i]) return x[i] < y[i];
  return false;
}
bool point_eq(const cpqm31& a, const cpqm31& b) { return qm_eq(a.x, b.x) && qm_eq(a.y, b.y); }
cpqm31 cpq_repeated_double(cpqm31 p, uint32_t n) { for (uint32_t i = 0; i < n; ++i) p = cpq_double(p); return p; }

qm31 from_partial_evals(qm31 e0, qm31 e1, qm31 e2, qm31 e3) {
  qm31 r = e0;
  r = qm_add(r, qm_mul(e1, qm31{0, 1, 0, 0}));
  r = qm_add(r, qm_mul(e2, qm31{0, 0, 1, 0}));
  r = qm_add(r, qm_mul(e3, qm31{0, 0, 0, 1}));
  return r;
}


void eval_at_point(PointEval& e) {
  const Component& c = *e.c;
  if (c.kind == AIR_WIDE_FIB) {
    qm31 a = e.next_trace(), b = e.next_trace();
    for (uint32_t i = 2; i < c.n_cols; ++i) {
      qm31 x = e.next_trace();
      e.add(qm_sub(x, qm_add(qm_sqr(a), qm_sqr(b))));
      a = b;
      b = x;
    }
    return;
  }
  std::vector<std::pair<qm31, qm31>> fr;
  for (uint32_t rep = 0; rep < c.ipr; ++rep) {
    qm31 s[16], init[16];
    for (int i = 0; i < 16; ++i) { s[i] = e.next_trace(); init[i] = s[i]; }
    auto full = [&]() {
      for (int i = 0; i < 16; ++i) s[i] = qm_add_m31(s[i], 1234);
      ext_q(s);
      for (int i = 0; i < 16; ++i) s[i] = pow5q(s[i]);
      for (int i = 0; i < 16; ++i) { qm31 m = e.next_trace(); e.add(qm_sub(s[i], m)); s[i] = m; }
    };
    for (int r = 0; r < 4; ++r) full();
    for (int r = 0; r < 14; ++r) {
      s[0] = qm_add_m31(s[0], 1234);
      int_q(s);
      s[0] = pow5q(s[0]);
      qm31 m = e.next_trace();
      e.add(qm_sub(s[0], m));
      s[0] = m;
    }
    for (int r = 0; r < 4; ++r) full();
    qm31 d0 = combine_q(c, init), d1 = combine_q(c, s);
    fr.push_back({qm_sub(d1, d0), qm_mul(d0, d1)});
  }
  qm31 prev = qm_zero();
  for (uint32_t b = 0; b + 1 < c.ipr; ++b) {
    qm31 cur = e.next_ext(1)[0];
    qm31 diff = qm_sub(cur, prev);
    prev = cur;
    e.add(qm_sub(qm_mul(diff, fr[b].second), fr[b].first));
  }
  auto m = e.next_ext(2);
  qm31 diff = qm_add(qm_sub(qm_sub(m[1], m[0]), prev), c.cumsum_shift());
  e.add(qm_sub(qm_mul(diff, fr.back().second), fr.back().first));
}

// mask offsets per tree (0 preprocessed, 1 trace, 2 interaction), per column of the component
// Mask offsets of column i of tree t (1 = trace, 2 = interaction) of a component: the rows,
// relative to the current one, at which its constraints read the column. Returns the count.
uint32_t mask_offsets(const Component& c, size_t t, size_t i, long* offs) {
  if (t == 2 && c.kind == AIR_POSEIDON && i >= 4 * (size_t)(c.ipr - 1)) {
    offs[0] = -1;  // the cumulative-sum columns also read the previous row
    offs[1] = 0;
    return 2;
  }
  offs[0] = 0;
  return 1;
}


std::vector<uint8_t> encode(const PcsCfg& cfg, const ProofParts& p) {
  Writer w;
  w.u32(cfg.pow_bits); w.u32(cfg.log_blowup); w.u32(cfg.log_last_layer); w.u32(cfg.n_queries);
  w.u32(cfg.fold_step); w.u32(cfg.lifting_log_size);
  w.u32((uint32_t)p.commitments.size());
  for (auto& h : p.commitments) w.hash(h);
  const SampleSet& S = p.sampled;
  w.u32(S.n_trees());
  for (uint32_t t = 0; t < S.n_trees(); ++t) {
    w.u32(S.n_cols(t));
    for (uint32_t c = 0; c < S.n_cols(t); ++c) {
      w.u32(S.count(t, c));
      for (uint32_t k = 0; k < S.count(t, c); ++k) w.q(S.vals(t, c)[k]);
    }
  }
  w.u32((uint32_t)p.decs.size());
  for (auto& d : p.decs) write_dec(w, d);
  w.u32((uint32_t)p.queried.size());
  for (auto& t : p.queried) { w.u32((uint32_t)t.size()); for (auto v : t) w.u32(v); }
  w.u64(p.pow);
  write_layer(w, p.first);
  w.u32((uint32_t)p.inner.size());
  for (auto& l : p.inner) write_layer(w, l);
  w.u32((uint32_t)p.last_poly.size());
  for (auto& q : p.last_poly) w.q(q);
  return w.take();
}

// Host inverse line FFT of the last FRI layer (prover/line.rs:44-64); returns stored-order coeffs.
std::vector<qm31> line_interpolate(std::vector<qm31> v, HCoset dom) {
  size_t n = v.size();
  uint32_t lg = 0;
  while (((size_t)1 << lg) < n) ++lg;
  for (size_t i = 0; i < n; ++i) { size_t j = bitrev((uint32_t)i, lg); if (j > i) std::swap(v[i], v[j]); }
  HCoset d = dom;
  while (((size_t)1 << d.log_size) > 1) {
    size_t ds = (size_t)1 << d.log_size, half = ds / 2;
    std::vector<uint32_t> xi(half);
    for (size_t i = 0; i < half; ++i) xi[i] = m31_inv(d.at(i).x);
    for (size_t ch = 0; ch < n; ch += ds)
      for (size_t i = 0; i < half; ++i) {
        qm31 a = v[ch + i], b = v[ch + half + i];
        v[ch + i] = qm_add(a, b);
        v[ch + half + i] = qm_mul_m31(qm_sub(a, b), xi[i]);
      }
    d = d.doubled();
  }
  uint32_t li = m31_inv((uint32_t)n);
  for (auto& x : v) x = qm_mul_m31(x, li);
  return v;
}

}  // namespace sh
