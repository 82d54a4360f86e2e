// Oracle (test infrastructure only): scalar CPU restatement of the reference prover's
// polynomial, Merkle, channel and FRI layers. Only tests/, the smoke check and bench.py's CPU
// baseline leg may load code built from this directory.
//
// Reference files followed (relative to crates/stwo/src):
//   twiddle tree            prover/backend/cpu/circle.rs:274-291 (slow_precompute_twiddles)
//   interpolate / evaluate  prover/backend/cpu/circle.rs:28-80,190-228,295-335; core/fft.rs:5-21
//   line twiddles accessor  core/poly/utils.rs:98-114
//   eval_at_point / fold    prover/backend/cpu/circle.rs:82-99; core/poly/utils.rs:27-70
//   barycentric             prover/backend/cpu/circle.rs:101-141; core/constraints.rs
//   Merkle commit/decommit  prover/vcs/prover.rs:39-221; core/vcs/blake2_merkle.rs:37-58
//   Merkle verify           core/vcs/verifier.rs:60-200
//   channel                 core/channel/blake2s.rs:19-140; core/vcs/blake2_merkle.rs:72-77
//   queries                 core/queries.rs:11-58
//   grind                   prover/backend/simd/grind.rs:20-91 (smallest (hi<<32)|lo, lo < 2^20)
#pragma once
#include <map>
#include <set>
#include <algorithm>
#include <stdexcept>
#include <string>
#include "ofield.h"
#include "oblake2s.h"

namespace oracle {

// ---------------------------------------------------------------- twiddles ---------------
struct TwiddleTree {
  Coset root;
  std::vector<M31> tw, itw;
};
inline TwiddleTree precompute_twiddles(const Coset& coset) {
  TwiddleTree t;
  t.root = coset;
  Coset c = coset;
  for (uint32_t j = 0; j < coset.log_size; ++j) {
    size_t i0 = t.tw.size();
    auto pts = c.points();
    for (size_t i = 0; i < c.size() / 2; ++i) t.tw.push_back(pts[i].x);
    std::vector<M31> seg(t.tw.begin() + i0, t.tw.end());
    bit_reverse(seg);
    std::copy(seg.begin(), seg.end(), t.tw.begin() + i0);
    c = c.doubled();
  }
  t.tw.push_back(M31(1));
  t.itw = batch_inverse(t.tw);
  return t;
}
// layers of line twiddles for a canonic circle domain of log size k, layer 0 largest.
inline std::vector<const M31*> line_twiddles(const std::vector<M31>& buf, uint32_t k) {
  uint32_t h = k - 1;  // half coset log size
  if (((size_t)1 << h) > buf.size()) throw std::runtime_error("Not enough twiddles!");
  std::vector<const M31*> r;
  for (int i = (int)h - 1; i >= 0; --i) r.push_back(buf.data() + buf.size() - ((size_t)2 << i));
  return r;
}

// ---------------------------------------------------------------- FFT --------------------
template <class F>
inline void butterfly(F& v0, F& v1, M31 t) { F tmp = v1 * t; v1 = v0 - tmp; v0 = v0 + tmp; }
template <class F>
inline void ibutterfly(F& v0, F& v1, M31 t) { F tmp = v0; v0 = tmp + v1; v1 = (tmp - v1) * t; }

// Inverse circle FFT of bit-reversed evaluations on CanonicCoset(k) into FFT-basis coefficients.
inline void interpolate(std::vector<M31>& v, const TwiddleTree& tt) {
  uint32_t k = 0;
  while (((size_t)1 << k) < v.size()) ++k;
  CircleDomain d = canonic_domain(k);
  if (k == 1) {
    M31 y = d.half_coset.initial.y, n(2);
    M31 yn_inv = inv(y * n), y_inv = yn_inv * n, n_inv = yn_inv * y;
    ibutterfly(v[0], v[1], y_inv);
    v[0] *= n_inv; v[1] *= n_inv;
    return;
  }
  if (k == 2) {
    M31 x = d.half_coset.initial.x, y = d.half_coset.initial.y, n(4);
    M31 xyn_inv = inv(x * y * n);
    M31 x_inv = xyn_inv * y * n, y_inv = xyn_inv * x * n, n_inv = xyn_inv * x * y;
    ibutterfly(v[0], v[1], y_inv);
    ibutterfly(v[2], v[3], -y_inv);
    ibutterfly(v[0], v[2], x_inv);
    ibutterfly(v[1], v[3], x_inv);
    for (auto& e : v) e *= n_inv;
    return;
  }
  auto lt = line_twiddles(tt.itw, k);
  size_t n = v.size();
  // circle layer (layer 0): twiddles [y,-y,-x,x] per [x,y] pair of line layer 0.
  for (size_t h = 0; h < n / 2; ++h) {
    M31 x = lt[0][(h >> 2) * 2], y = lt[0][(h >> 2) * 2 + 1];
    M31 t = (h & 3) == 0 ? y : (h & 3) == 1 ? -y : (h & 3) == 2 ? -x : x;
    ibutterfly(v[2 * h], v[2 * h + 1], t);
  }
  for (size_t li = 0; li < lt.size(); ++li) {
    size_t i = li + 1;
    for (size_t h = 0; h < (n >> (i + 1)); ++h) {
      M31 t = lt[li][h];
      for (size_t l = 0; l < ((size_t)1 << i); ++l) {
        size_t i0 = (h << (i + 1)) + l, i1 = i0 + ((size_t)1 << i);
        ibutterfly(v[i0], v[i1], t);
      }
    }
  }
  M31 ninv = inv(M31((uint32_t)n));
  for (auto& e : v) e *= ninv;
}

// Forward circle FFT: coefficients (zero-padded) onto CanonicCoset(K) in bit-reversed order.
inline std::vector<M31> evaluate(const std::vector<M31>& coeffs, uint32_t K, const TwiddleTree& tt) {
  std::vector<M31> v(coeffs);
  v.resize((size_t)1 << K, M31(0));
  CircleDomain d = canonic_domain(K);
  if (K == 1) { butterfly(v[0], v[1], d.half_coset.initial.y); return v; }
  if (K == 2) {
    M31 x = d.half_coset.initial.x, y = d.half_coset.initial.y;
    butterfly(v[0], v[2], x); butterfly(v[1], v[3], x);
    butterfly(v[0], v[1], y); butterfly(v[2], v[3], -y);
    return v;
  }
  auto lt = line_twiddles(tt.tw, K);
  size_t n = v.size();
  for (int li = (int)lt.size() - 1; li >= 0; --li) {
    size_t i = li + 1;
    for (size_t h = 0; h < (n >> (i + 1)); ++h) {
      M31 t = lt[li][h];
      for (size_t l = 0; l < ((size_t)1 << i); ++l) {
        size_t i0 = (h << (i + 1)) + l, i1 = i0 + ((size_t)1 << i);
        butterfly(v[i0], v[i1], t);
      }
    }
  }
  for (size_t h = 0; h < n / 2; ++h) {
    M31 x = lt[0][(h >> 2) * 2], y = lt[0][(h >> 2) * 2 + 1];
    M31 t = (h & 3) == 0 ? y : (h & 3) == 1 ? -y : (h & 3) == 2 ? -x : x;
    butterfly(v[2 * h], v[2 * h + 1], t);
  }
  return v;
}

// Evaluates FFT-basis coefficients at a QM31 circle point.
inline QM31 eval_at_point(const std::vector<M31>& coeffs, CirclePoint<QM31> p) {
  size_t n = coeffs.size();
  if (n == 1) return QM31(coeffs[0]);
  uint32_t L = 0;
  while (((size_t)1 << L) < n) ++L;
  // factors in order of increasing bit: y, x, pi(x), pi^2(x), ...
  std::vector<QM31> fac;
  fac.push_back(p.y);
  QM31 x = p.x;
  for (uint32_t i = 1; i < L; ++i) { fac.push_back(x); x = double_x(x); }
  std::vector<QM31> cur(n / 2);
  for (size_t i = 0; i < n / 2; ++i) cur[i] = QM31(coeffs[2 * i]) + fac[0] * coeffs[2 * i + 1];
  for (uint32_t lv = 1; lv < L; ++lv) {
    size_t m = cur.size() / 2;
    for (size_t i = 0; i < m; ++i) cur[i] = cur[2 * i] + cur[2 * i + 1] * fac[lv];
    cur.resize(m);
  }
  return cur[0];
}

// Barycentric weights for CanonicCoset(k) at p, bit-reversed order.
inline std::vector<QM31> barycentric_weights(uint32_t k, CirclePoint<QM31> p) {
  CircleDomain d = canonic_domain(k);
  size_t n = d.size();
  Coset gen_coset = Coset::make(1, k);
  std::vector<QM31> denom(n);
  for (size_t i = 0; i < n; ++i) {
    CirclePoint<QM31> cpt = to_ef(d.at(bit_reverse_index(i, k)));
    QM31 m2y = cpt.y * (-M31(2));
    // coset_vanishing_derivative(Coset(generator, k), cpt) = 4^(k-1) * prod_{i=1}^{k-1} V_{CanonicCoset(i)}(cpt)
    QM31 e = QM31::one();
    for (uint32_t j = 1; j < k; ++j) e = e * M31(4);
    QM31 van = QM31::one();
    for (uint32_t j = 1; j < k; ++j) van = van * coset_vanishing(canonic_coset(j), cpt);
    (void)gen_coset;
    QM31 si = m2y * (e * van);
    // point_vanishing(cpt, p) = h.y / (1 + h.x), h = p - cpt
    CirclePoint<QM31> h = cp_add(p, cp_conj(cpt));
    QM31 vi = h.y * inv(QM31::one() + h.x);
    denom[i] = si * vi;
  }
  QM31 vn = coset_vanishing(canonic_coset(k), p);
  auto dinv = batch_inverse(denom);
  for (auto& w : dinv) w = vn * w;
  return dinv;
}

// ---------------------------------------------------------------- Merkle -----------------
struct MerkleDecommitment {
  std::vector<Hash32> hash_witness;
  std::vector<M31> column_witness;
};
struct MerkleTree {
  std::vector<std::vector<Hash32>> layers;  // layers[k] has 2^k nodes; layers[0] = root
  Hash32 root() const { return layers[0][0]; }
  uint32_t max_log() const { return (uint32_t)layers.size() - 1; }
};
inline uint32_t log2_size(size_t n) { uint32_t k = 0; while (((size_t)1 << k) < n) ++k; return k; }

inline std::vector<Hash32> commit_on_layer(uint32_t log_size, const std::vector<Hash32>* prev,
                                           const std::vector<const std::vector<M31>*>& cols) {
  std::vector<Hash32> out((size_t)1 << log_size);
#pragma omp parallel for schedule(static) if (((size_t)1 << log_size) >= 4096)
  for (long i = 0; i < (long)out.size(); ++i) {
    Blake2s h;
    if (prev) { h.update((*prev)[2 * i].data(), 32); h.update((*prev)[2 * i + 1].data(), 32); }
    for (auto* c : cols) h.update_u32((*c)[i].v);
    out[i] = h.finalize();
  }
  return out;
}
// Columns are hashed in stable order of decreasing length (prover/vcs/prover.rs:39-65).
inline std::vector<size_t> merkle_column_order(const std::vector<const std::vector<M31>*>& cols) {
  std::vector<size_t> order(cols.size());
  for (size_t i = 0; i < cols.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cols[a]->size() > cols[b]->size(); });
  return order;
}
inline MerkleTree merkle_commit(const std::vector<const std::vector<M31>*>& cols) {
  MerkleTree t;
  if (cols.empty()) { t.layers.push_back(commit_on_layer(0, nullptr, {})); return t; }
  auto order = merkle_column_order(cols);
  uint32_t max_log = log2_size(cols[order[0]]->size());
  std::vector<std::vector<Hash32>> rev;
  size_t ci = 0;
  for (int ls = (int)max_log; ls >= 0; --ls) {
    std::vector<const std::vector<M31>*> lc;
    while (ci < order.size() && log2_size(cols[order[ci]]->size()) == (uint32_t)ls) lc.push_back(cols[order[ci++]]);
    rev.push_back(commit_on_layer(ls, rev.empty() ? nullptr : &rev.back(), lc));
  }
  std::reverse(rev.begin(), rev.end());
  t.layers = std::move(rev);
  return t;
}
// Returns queried values (node-major per layer, largest layer first) and the decommitment.
inline std::vector<M31> merkle_decommit(const MerkleTree& t, const std::map<uint32_t, std::vector<size_t>>& qpl,
                                        const std::vector<const std::vector<M31>*>& cols, MerkleDecommitment& dec) {
  std::vector<M31> queried;
  auto order = merkle_column_order(cols);
  size_t ci = 0;
  std::vector<size_t> last;
  for (int ls = (int)t.layers.size() - 1; ls >= 0; --ls) {
    bool has_prev = (size_t)ls + 1 < t.layers.size();
    std::set<size_t> prev_set(last.begin(), last.end());
    std::set<size_t> colq;
    auto it = qpl.find((uint32_t)ls);
    if (it != qpl.end()) colq.insert(it->second.begin(), it->second.end());
    std::set<size_t> merged;
    for (size_t q : last) merged.insert(q / 2);
    merged.insert(colq.begin(), colq.end());
    std::vector<const std::vector<M31>*> lc;
    while (ci < order.size() && log2_size(cols[order[ci]]->size()) == (uint32_t)ls) lc.push_back(cols[order[ci++]]);
    std::vector<size_t> qnodes, wnodes;
    for (size_t node : merged) {
      if (has_prev) {
        if (!prev_set.count(2 * node)) dec.hash_witness.push_back(t.layers[ls + 1][2 * node]);
        if (!prev_set.count(2 * node + 1)) dec.hash_witness.push_back(t.layers[ls + 1][2 * node + 1]);
      }
      if (colq.count(node)) qnodes.push_back(node); else wnodes.push_back(node);
    }
    for (size_t node : qnodes) for (auto* c : lc) queried.push_back((*c)[node]);
    for (size_t node : wnodes) for (auto* c : lc) dec.column_witness.push_back((*c)[node]);
    last.assign(merged.begin(), merged.end());
  }
  return queried;
}
inline Hash32 hash_node(const Hash32* l, const Hash32* r, const M31* vals, size_t n) {
  Blake2s h;
  if (l) { h.update(l->data(), 32); h.update(r->data(), 32); }
  for (size_t i = 0; i < n; ++i) h.update_u32(vals[i].v);
  return h.finalize();
}
// core/vcs/verifier.rs verify(); returns "" on success or an error name.
inline std::string merkle_verify(const Hash32& root, const std::vector<uint32_t>& col_log_sizes,
                                 const std::map<uint32_t, std::vector<size_t>>& qpl, const std::vector<M31>& queried,
                                 const MerkleDecommitment& dec) {
  if (col_log_sizes.empty()) return "";
  uint32_t max_log = *std::max_element(col_log_sizes.begin(), col_log_sizes.end());
  std::map<uint32_t, size_t> ncols;
  for (auto l : col_log_sizes) ncols[l]++;
  size_t qi = 0, hw = 0, cw = 0;
  std::vector<std::pair<size_t, Hash32>> last;
  bool have_last = false;
  for (int ls = (int)max_log; ls >= 0; --ls) {
    size_t nc = ncols.count(ls) ? ncols[ls] : 0;
    std::vector<std::pair<size_t, Hash32>> cur;
    std::vector<size_t> colq;
    auto it = qpl.find((uint32_t)ls);
    if (it != qpl.end()) colq = it->second;
    size_t pi = 0, cq = 0;
    while (true) {
      bool hp = pi < last.size(), hc = cq < colq.size();
      if (!hp && !hc) break;
      size_t node = hp ? last[pi].first / 2 : SIZE_MAX;
      if (hc && colq[cq] < node) node = colq[cq];
      Hash32 lh{}, rh{};
      if (have_last) {
        if (pi < last.size() && last[pi].first == 2 * node) lh = last[pi++].second;
        else { if (hw >= dec.hash_witness.size()) return "WitnessTooShort"; lh = dec.hash_witness[hw++]; }
        if (pi < last.size() && last[pi].first == 2 * node + 1) rh = last[pi++].second;
        else { if (hw >= dec.hash_witness.size()) return "WitnessTooShort"; rh = dec.hash_witness[hw++]; }
      }
      // skip any further prev queries under this node (none remain after the two children)
      while (pi < last.size() && last[pi].first / 2 == node) ++pi;
      std::vector<M31> vals;
      if (hc && colq[cq] == node) {
        ++cq;
        if (qi + nc > queried.size()) return "TooFewQueriedValues";
        vals.assign(queried.begin() + qi, queried.begin() + qi + nc);
        qi += nc;
      } else {
        if (cw + nc > dec.column_witness.size()) return "WitnessTooShort";
        vals.assign(dec.column_witness.begin() + cw, dec.column_witness.begin() + cw + nc);
        cw += nc;
      }
      cur.push_back({node, hash_node(have_last ? &lh : nullptr, have_last ? &rh : nullptr, vals.data(), vals.size())});
    }
    last = std::move(cur);
    have_last = true;
  }
  if (hw != dec.hash_witness.size()) return "WitnessTooLong";
  if (qi != queried.size()) return "TooManyQueriedValues";
  if (cw != dec.column_witness.size()) return "WitnessTooLong";
  if (last.size() != 1 || last[0].second != root) return "RootMismatch";
  return "";
}

// ---------------------------------------------------------------- channel ----------------
struct Channel {
  Hash32 digest{};
  uint32_t n_draws = 0;
  void update_digest(const Hash32& d) { digest = d; n_draws = 0; }
  void mix_root(const Hash32& root) {
    Blake2s h; h.update(digest.data(), 32); h.update(root.data(), 32); update_digest(h.finalize());
  }
  void mix_u32s(const std::vector<uint32_t>& data) {
    Blake2s h; h.update(digest.data(), 32);
    for (auto w : data) h.update_u32(w);
    update_digest(h.finalize());
  }
  void mix_u64(uint64_t v) { mix_u32s({(uint32_t)v, (uint32_t)(v >> 32)}); }
  void mix_felts(const std::vector<QM31>& felts) {
    Blake2s h; h.update(digest.data(), 32);
    for (auto& f : felts) for (int i = 0; i < 4; ++i) h.update_u32(f.coord(i).v);
    update_digest(h.finalize());
  }
  std::array<uint32_t, 8> draw_u32s() {
    uint8_t in[37];
    memcpy(in, digest.data(), 32);
    for (int i = 0; i < 4; ++i) in[32 + i] = (uint8_t)(n_draws >> (8 * i));
    in[36] = 0;
    n_draws++;
    auto hh = Blake2s::hash(in, 37);
    std::array<uint32_t, 8> r;
    for (int i = 0; i < 8; ++i) r[i] = (uint32_t)hh[4 * i] | ((uint32_t)hh[4 * i + 1] << 8) | ((uint32_t)hh[4 * i + 2] << 16) | ((uint32_t)hh[4 * i + 3] << 24);
    return r;
  }
  std::array<M31, 8> draw_base_felts() {
    while (true) {
      auto u = draw_u32s();
      bool ok = true;
      for (auto x : u) if (x >= 2 * P) ok = false;
      if (ok) { std::array<M31, 8> r; for (int i = 0; i < 8; ++i) r[i] = M31::reduce64(u[i]); return r; }
    }
  }
  QM31 draw_secure_felt() { auto f = draw_base_felts(); return QM31::from_m31(f[0], f[1], f[2], f[3]); }
  std::vector<QM31> draw_secure_felts(size_t n) {
    std::vector<QM31> r;
    while (r.size() < n) {
      auto f = draw_base_felts();
      r.push_back(QM31::from_m31(f[0], f[1], f[2], f[3]));
      if (r.size() < n) r.push_back(QM31::from_m31(f[4], f[5], f[6], f[7]));
    }
    return r;
  }
  static constexpr uint32_t POW_PREFIX = 0x12345678u;
  Hash32 pow_prefix(uint32_t pow_bits) const {
    Blake2s h; h.update_u32(POW_PREFIX);
    uint8_t z[12] = {0}; h.update(z, 12);
    h.update(digest.data(), 32); h.update_u32(pow_bits);
    return h.finalize();
  }
  bool verify_pow_nonce(uint32_t bits, uint64_t nonce) const {
    Hash32 pre = pow_prefix(bits);
    Blake2s h; h.update(pre.data(), 32); h.update_u32((uint32_t)nonce); h.update_u32((uint32_t)(nonce >> 32));
    auto r = h.finalize();
    unsigned __int128 v = 0;
    for (int i = 15; i >= 0; --i) v = (v << 8) | r[i];
    uint32_t tz = 0;
    if (v == 0) tz = 128; else while (!((v >> tz) & 1)) ++tz;
    return tz >= bits;
  }
  uint64_t grind(uint32_t bits) const {
    if (bits > 32) throw std::runtime_error("pow_bits > 32 is not supported");
    Hash32 pre = pow_prefix(bits);
    for (uint64_t hi = 0;; ++hi) {
      for (uint32_t lo = 0; lo < (1u << 20); ++lo) {
        Blake2s h; h.update(pre.data(), 32); h.update_u32(lo); h.update_u32((uint32_t)hi);
        auto r = h.finalize();
        uint32_t w0 = (uint32_t)r[0] | ((uint32_t)r[1] << 8) | ((uint32_t)r[2] << 16) | ((uint32_t)r[3] << 24);
        uint32_t tz = w0 == 0 ? 32 : __builtin_ctz(w0);
        if (tz >= bits) return (hi << 32) | lo;
      }
    }
  }
};

inline CirclePoint<QM31> get_random_point(Channel& ch) {
  QM31 t = ch.draw_secure_felt();
  QM31 t2 = t * t;
  QM31 d = inv(t2 + QM31::one());
  return {(QM31::one() - t2) * d, (t + t) * d};
}

inline std::vector<size_t> draw_queries(Channel& ch, uint32_t log_size, size_t n) {
  std::vector<size_t> r;
  uint32_t mask = (uint32_t)(((uint64_t)1 << log_size) - 1);
  while (true) {
    auto w = ch.draw_u32s();
    for (auto x : w) { r.push_back(x & mask); if (r.size() == n) return r; }
  }
}
inline std::vector<size_t> sorted_unique(std::vector<size_t> v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}
inline std::vector<size_t> fold_queries(const std::vector<size_t>& q, uint32_t n) {
  std::vector<size_t> r;
  for (size_t x : q) { size_t y = x >> n; if (r.empty() || r.back() != y) r.push_back(y); }
  return r;
}

}  // namespace oracle
