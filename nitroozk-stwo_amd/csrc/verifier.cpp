// Host verifier of libstwo_hip's proofs (paths relative to crates/stwo/src):
//   verify / verify_ex                      core/verifier.rs:19-118
//   CommitmentSchemeVerifier::verify_values core/pcs/verifier.rs:60-158
//   fri_answers, sample batches, line coeffs core/pcs/quotients.rs:60-440; core/constraints.rs:119-134
//   FriVerifier commit / decommit           core/fri.rs:98-290,424-720
//   MerkleVerifier::verify                  core/vcs/verifier.rs:85-195
// It reads the library's canonical proof encoding, replays the channel (Blake2s or Poseidon252,
// the proof's MerkleChannel) and checks every
// Merkle path, the FRI layers, the proof of work, the out-of-domain sample against the AIRs'
// constraints, and the quotients at the queried positions. It runs on the host only, so the
// serving loop can check each proof while the device proves the next.
#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>

#include "protocol.h"

namespace sh {
namespace {

struct Fail : std::runtime_error {
  explicit Fail(const std::string& m) : std::runtime_error(m) {}
};

// ---- decoding (the inverse of encode in protocol.cpp) --------------------------------------------
struct Reader {
  const uint8_t* p;
  size_t n, off = 0;
  void need(size_t k) {
    if (off + k > n) throw Fail("InvalidStructure: proof truncated");
  }
  uint32_t u32() {
    need(4);
    uint32_t v;
    std::memcpy(&v, p + off, 4);
    off += 4;
    return v;
  }
  uint32_t count(size_t max) {
    const uint32_t v = u32();
    if (v > max) throw Fail("InvalidStructure: count out of range");
    return v;
  }
  uint64_t u64() { uint64_t lo = u32(); return lo | (uint64_t)u32() << 32; }
  Hash hash() {
    need(32);
    Hash h;
    std::memcpy(h.data(), p + off, 32);
    off += 32;
    return h;
  }
  uint32_t m31() {
    const uint32_t v = u32();
    if (v >= P) throw Fail("InvalidStructure: non-canonical field element");
    return v;
  }
  qm31 q() { qm31 r; r.a = m31(); r.b = m31(); r.c = m31(); r.d = m31(); return r; }
};
MerkleDec read_dec(Reader& r) {
  MerkleDec d;
  for (uint32_t i = 0, n = r.count(1u << 24); i < n; ++i) d.hash_witness.push_back(r.hash());
  for (uint32_t i = 0, n = r.count(1u << 28); i < n; ++i) d.column_witness.push_back(r.m31());
  return d;
}
FriLayer read_layer(Reader& r) {
  FriLayer l;
  for (uint32_t i = 0, n = r.count(1u << 24); i < n; ++i) l.witness.push_back(r.q());
  l.dec = read_dec(r);
  l.commitment = r.hash();
  return l;
}
struct Decoded {
  PcsCfg cfg;
  std::vector<Hash> commitments;
  std::vector<std::vector<std::vector<qm31>>> sampled;  // [tree][column][sample]
  std::vector<MerkleDec> decs;
  std::vector<std::vector<uint32_t>> queried;
  uint64_t pow = 0;
  FriLayer first;
  std::vector<FriLayer> inner;
  std::vector<qm31> last_poly;
};
Decoded decode(const uint8_t* p, size_t n) {
  Reader r{p, n};
  Decoded d;
  d.cfg.pow_bits = r.u32();
  d.cfg.log_blowup = r.u32();
  d.cfg.log_last_layer = r.u32();
  d.cfg.n_queries = r.u32();
  d.cfg.fold_step = r.u32();
  d.cfg.lifting_log_size = r.u32();
  for (uint32_t i = 0, k = r.count(64); i < k; ++i) d.commitments.push_back(r.hash());
  d.sampled.resize(r.count(64));
  for (auto& t : d.sampled) {
    t.resize(r.count(1u << 24));
    for (auto& c : t)
      for (uint32_t i = 0, k = r.count(64); i < k; ++i) c.push_back(r.q());
  }
  d.decs.resize(r.count(64));
  for (auto& x : d.decs) x = read_dec(r);
  d.queried.resize(r.count(64));
  for (auto& t : d.queried)
    for (uint32_t i = 0, k = r.count(1u << 28); i < k; ++i) t.push_back(r.m31());
  d.pow = r.u64();
  d.first = read_layer(r);
  for (uint32_t i = 0, k = r.count(64); i < k; ++i) d.inner.push_back(read_layer(r));
  for (uint32_t i = 0, k = r.count(1u << 20); i < k; ++i) d.last_poly.push_back(r.q());
  if (r.off != n) throw Fail("InvalidStructure: trailing bytes");
  return d;
}

// ---- Merkle ------------------------------------------------------------------------------------
Hash hash_node(const Hash* left, const Hash* right, const uint32_t* vals, size_t n) {
  Blake2sHost b;
  if (left) { b.update(left->data(), 32); b.update(right->data(), 32); }
  if (n) b.update(vals, 4 * n);  // u32 little endian, the host's order
  return b.finalize();
}
// MerkleVerifier::verify: the root from the queried values and the witnesses, every input consumed.
template <class MC>
void merkle_verify(const Hash& root, const std::vector<uint32_t>& col_logs,
                   const std::map<uint32_t, std::vector<size_t>>& qpl, const std::vector<uint32_t>& queried,
                   const MerkleDec& dec, const char* what) {
  if (col_logs.empty()) return;
  const uint32_t top = *std::max_element(col_logs.begin(), col_logs.end());
  std::map<uint32_t, size_t> ncols;
  for (auto l : col_logs) ++ncols[l];
  size_t qi = 0, hi = 0, ci = 0;
  std::vector<std::pair<size_t, Hash>> last;
  bool have_last = false;
  static const std::vector<size_t> none;
  for (int ls = (int)top; ls >= 0; --ls) {
    const size_t nc = ncols.count((uint32_t)ls) ? ncols[(uint32_t)ls] : 0;
    auto it = qpl.find((uint32_t)ls);
    const std::vector<size_t>& colq = it == qpl.end() ? none : it->second;
    std::vector<size_t> nodes;
    for (auto& e : last) nodes.push_back(e.first / 2);
    nodes.insert(nodes.end(), colq.begin(), colq.end());
    nodes = sorted_unique(nodes);
    std::vector<std::pair<size_t, Hash>> cur;
    size_t li = 0;
    std::vector<uint32_t> vals(nc);
    for (size_t node : nodes) {
      Hash l, r;
      if (have_last) {
        for (int side = 0; side < 2; ++side) {
          Hash& h = side ? r : l;
          if (li < last.size() && last[li].first == 2 * node + side) {
            h = last[li++].second;
          } else {
            if (hi >= dec.hash_witness.size()) throw Fail(std::string("Merkle (") + what + "): WitnessTooShort");
            h = dec.hash_witness[hi++];
          }
        }
      }
      const bool q = std::binary_search(colq.begin(), colq.end(), node);
      const std::vector<uint32_t>& src = q ? queried : dec.column_witness;
      size_t& at = q ? qi : ci;
      if (at + nc > src.size())
        throw Fail(std::string("Merkle (") + what + "): " + (q ? "TooFewQueriedValues" : "WitnessTooShort"));
      std::copy(src.begin() + at, src.begin() + at + nc, vals.begin());
      at += nc;
      cur.push_back({node, MC::node(have_last ? &l : nullptr, &r, vals.data(), nc)});
    }
    last = std::move(cur);
    have_last = true;
  }
  if (hi != dec.hash_witness.size()) throw Fail(std::string("Merkle (") + what + "): WitnessTooLong");
  if (qi != queried.size()) throw Fail(std::string("Merkle (") + what + "): TooManyQueriedValues");
  if (ci != dec.column_witness.size()) throw Fail(std::string("Merkle (") + what + "): WitnessTooLong");
  if (last.size() != 1 || last[0].second != root) throw Fail(std::string("Merkle (") + what + "): RootMismatch");
}
std::map<uint32_t, std::vector<size_t>> queries_per_log_size(const std::vector<uint32_t>& logs,
                                                             const std::vector<size_t>& tree_pos) {
  std::map<uint32_t, std::vector<size_t>> r;
  if (logs.empty()) return r;
  const uint32_t top = *std::max_element(logs.begin(), logs.end());
  for (auto ls : std::set<uint32_t>(logs.begin(), logs.end())) {
    std::vector<size_t> v;
    for (size_t p : tree_pos) v.push_back(((p >> (top - ls + 1)) << 1) + (p & 1));
    r[ls] = sorted_unique(v);
  }
  return r;
}

// ---- circle domains ------------------------------------------------------------------------------
// CircleDomain::index_at for the domain whose half coset is `h`: the half coset, then its conjugate.
uint32_t circle_domain_index_at(const HCoset& h, size_t i) {
  const size_t half = (size_t)1 << h.log_size;
  return i < half ? h.index_at(i) : (0x80000000u - h.index_at(i - half)) & 0x7fffffffu;
}
qm31 eval_line_poly(const std::vector<qm31>& coeffs, qm31 x) {
  // LinePoly::eval_at_point (core/poly/line.rs:138-145): fold (core/poly/utils.rs:27-38) of the
  // stored coefficients with the factors x, pi(x), pi^2(x), ...; the last factor pairs adjacent
  // values, the first the two halves
  std::vector<qm31> fac;
  for (size_t n = coeffs.size(); n > 1; n >>= 1) { fac.push_back(x); x = qm_double_x(x); }
  std::vector<qm31> v = coeffs;
  for (size_t k = fac.size(); k-- > 0;) {
    std::vector<qm31> nv(v.size() / 2);
    for (size_t i = 0; i < nv.size(); ++i) nv[i] = qm_add(v[2 * i], qm_mul(v[2 * i + 1], fac[k]));
    v.swap(nv);
  }
  return v[0];
}

}  // namespace
}  // namespace sh

namespace sh {
namespace {

bool verify_pow_nonce(const Channel& ch, uint32_t bits, uint64_t nonce) {
  const Hash pre = ch.pow_prefix(bits);
  Blake2sHost b;
  b.update(pre.data(), 32);
  b.update(&nonce, 8);
  const Hash h = b.finalize();
  // trailing zeros of the first 16 bytes as a little-endian u128 (core/channel/blake2s.rs:124-140)
  uint32_t tz = 0;
  for (int i = 0; i < 16; ++i) {
    if (h[i] == 0) { tz += 8; continue; }
    tz += (uint32_t)__builtin_ctz(h[i]);
    break;
  }
  return tz >= bits;
}

// compute_decommitment_positions_and_rebuild_evals (core/fri.rs:599-636): per subset of 2^w
// positions holding queries, the positions, the subset's values (queried ones from evals, the rest
// from the witness) and the subset's initial domain index.
struct Sparse {
  std::vector<size_t> positions;
  std::vector<std::vector<qm31>> subsets;
  std::vector<size_t> initial;  // bit-reversed subset start
};
Sparse rebuild(const std::vector<size_t>& q, uint32_t log_domain, const std::vector<qm31>& evals,
               const std::vector<qm31>& witness, uint32_t w, const char* what) {
  Sparse s;
  size_t ei = 0, wi = 0;
  for (size_t i = 0; i < q.size();) {
    size_t j = i;
    while (j < q.size() && (q[j] >> w) == (q[i] >> w)) ++j;
    const size_t start = (q[i] >> w) << w;
    std::vector<qm31> sub;
    size_t k = i;
    for (size_t p = start; p < start + ((size_t)1 << w); ++p) {
      s.positions.push_back(p);
      if (k < j && q[k] == p) {
        sub.push_back(evals[ei++]);
        ++k;
      } else {
        if (wi >= witness.size()) throw Fail(std::string(what) + "EvaluationsInvalid");
        sub.push_back(witness[wi++]);
      }
    }
    s.subsets.push_back(sub);
    s.initial.push_back(bitrev((uint32_t)start, log_domain));
    i = j;
  }
  if (wi != witness.size()) throw Fail(std::string(what) + "EvaluationsInvalid");
  return s;
}
std::vector<uint32_t> flat_coords(const Sparse& s) {
  std::vector<uint32_t> v;
  for (auto& sub : s.subsets)
    for (auto& x : sub) { v.push_back(x.a); v.push_back(x.b); v.push_back(x.c); v.push_back(x.d); }
  return v;
}
// fold_coset (core/fri.rs:728-747) on the line domain Coset::new(initial, n).
qm31 fold_coset(std::vector<qm31> e, uint32_t initial, uint32_t n, qm31 alpha) {
  HCoset d = HCoset::make(initial, n);
  for (uint32_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < ((size_t)1 << (n - i)); j += 2) {
      const uint32_t xi = m31_inv(d.at(bitrev((uint32_t)j, d.log_size)).x);
      const qm31 f0 = qm_add(e[j], e[j + 1]), f1 = qm_mul_m31(qm_sub(e[j], e[j + 1]), xi);
      e[j >> 1] = qm_add(f0, qm_mul(alpha, f1));
    }
    alpha = qm_sqr(alpha);
    d = d.doubled();
  }
  return e[0];
}

// The verifier's side of a MerkleChannel: node hash, channel, proof-of-work check and which 32-byte
// strings are hashes at all (a felt252 must lie below the Stark prime).
struct Blake2sMC {
  using Ch = Channel;
  static Hash node(const Hash* l, const Hash* r, const uint32_t* v, size_t n) { return hash_node(l, r, v, n); }
  static bool pow_ok(const Ch& ch, uint32_t bits, uint64_t nonce) { return verify_pow_nonce(ch, bits, nonce); }
  static bool hash_ok(const Hash&) { return true; }
};
struct Poseidon252MC {
  using Ch = P252Channel;
  static Hash node(const Hash* l, const Hash* r, const uint32_t* v, size_t n) { return p252_hash_node(l, r, v, n); }
  static bool pow_ok(const Ch& ch, uint32_t bits, uint64_t nonce) { return ch.verify_pow_nonce(bits, nonce); }
  static bool hash_ok(const Hash& h) { return p252_hash_canonical(h); }
};

template <class MC>
void check_hashes(const Decoded& d) {
  auto ok = [](const Hash& h) {
    if (!MC::hash_ok(h)) throw Fail("InvalidStructure: a hash is not a felt252");
  };
  for (auto& h : d.commitments) ok(h);
  for (auto& m : d.decs)
    for (auto& h : m.hash_witness) ok(h);
  ok(d.first.commitment);
  for (auto& h : d.first.dec.hash_witness) ok(h);
  for (auto& l : d.inner) {
    ok(l.commitment);
    for (auto& h : l.dec.hash_witness) ok(h);
  }
}

template <class MC>
std::string verify_impl(const uint8_t* bytes, size_t len, std::vector<Component> comps, const PcsCfg& want) {
  const Decoded d = decode(bytes, len);
  check_hashes<MC>(d);
  const PcsCfg& cfg = d.cfg;
  if (cfg.pow_bits != want.pow_bits || cfg.log_blowup != want.log_blowup ||
      cfg.log_last_layer != want.log_last_layer || cfg.n_queries != want.n_queries ||
      cfg.fold_step != want.fold_step || cfg.lifting_log_size != want.lifting_log_size)
    throw Fail("InvalidStructure: the proof's configuration differs from the expected one");
  if (cfg.log_blowup < 1 || cfg.fold_step < 1 || cfg.n_queries < 1 || cfg.pow_bits > 64)
    throw Fail("InvalidStructure: configuration out of range");
  const uint32_t blow = cfg.log_blowup;
  bool inter = false;
  for (auto& c : comps) inter |= c.n_inter_cols() > 0;
  std::vector<std::vector<uint32_t>> tlogs(inter ? 3 : 2);
  for (auto& c : comps) {
    c.trace_start = tlogs[1].size();
    tlogs[1].insert(tlogs[1].end(), c.n_trace_cols(), c.log_n + blow);
    if (inter) {
      c.inter_start = tlogs[2].size();
      tlogs[2].insert(tlogs[2].end(), c.n_inter_cols(), c.log_n + blow);
    }
  }
  if (d.commitments.size() != tlogs.size() + 1) throw Fail("InvalidStructure: commitment count");
  typename MC::Ch ch;
  ch.mix_root(d.commitments[0]);
  ch.mix_root(d.commitments[1]);
  if (inter) {
    const auto le = ch.draw_secure_felts(2);
    for (auto& c : comps) {
      c.z = le[0];
      c.alpha = le[1];
      qm31 p = qm_one();
      for (int i = 0; i < 16; ++i) { c.alpha_pow[i] = p; p = qm_mul(p, c.alpha); }
    }
    ch.mix_root(d.commitments[2]);
  }
  uint32_t bound = 0;
  for (auto& c : comps) bound = std::max(bound, c.eval_log());
  const uint32_t lifting = cfg.lifting_log_size ? cfg.lifting_log_size : bound - 1 + blow;
  if (lifting < blow + 1) throw Fail("InvalidStructure: lifting size");
  const uint32_t max_deg = lifting - blow;
  const qm31 random_coeff = ch.draw_secure_felt();
  ch.mix_root(d.commitments.back());
  tlogs.push_back(std::vector<uint32_t>(8, max_deg + blow));
  cpqm31 oods;
  {
    const qm31 t = ch.draw_secure_felt(), t2 = qm_sqr(t), dn = qm_inv(qm_add(t2, qm_one()));
    oods = {qm_mul(qm_sub(qm_one(), t2), dn), qm_mul(qm_add(t, t), dn)};
  }
  // sample points (mask points of the components, the OODS point for the composition columns)
  SampleSet S;
  const cpm31 step = index_to_point(subgroup_gen(max_deg));
  std::vector<long> pt_off;
  auto pid_of = [&](long o) -> uint32_t {
    for (size_t i = 0; i < pt_off.size(); ++i)
      if (pt_off[i] == o) return (uint32_t)i;
    pt_off.push_back(o);
    return (uint32_t)pt_off.size() - 1;
  };
  for (size_t t = 0; t < tlogs.size(); ++t) S.col_base.push_back(S.col_base.back() + (uint32_t)tlogs[t].size());
  const uint32_t NC = S.col_base.back();
  std::vector<std::vector<uint32_t>> cpid(NC);
  for (auto& c : comps)
    for (size_t t = 1; t < 3 && t + 1 < tlogs.size(); ++t) {
      const size_t start = t == 1 ? c.trace_start : c.inter_start, n = t == 1 ? c.n_trace_cols() : c.n_inter_cols();
      for (size_t i = 0; i < n; ++i) {
        long offs[2];
        const uint32_t k = mask_offsets(c, t, i, offs);
        for (uint32_t j = 0; j < k; ++j) cpid[S.col_base[t] + start + i].push_back(pid_of(offs[j]));
      }
    }
  for (uint32_t g = S.col_base[tlogs.size() - 1]; g < NC; ++g) cpid[g].assign(1, pid_of(0));
  for (long o : pt_off) S.pts.push_back(cpq_add(oods, cpq_mul_signed(step, o)));
  if (d.sampled.size() != tlogs.size()) throw Fail("InvalidStructure: sampled values per tree");
  S.start.assign(NC + 1, 0);
  for (size_t t = 0; t < tlogs.size(); ++t) {
    if (d.sampled[t].size() != tlogs[t].size()) throw Fail("InvalidStructure: sampled columns");
    for (size_t c = 0; c < tlogs[t].size(); ++c) {
      const uint32_t g = S.col_base[t] + (uint32_t)c;
      if (d.sampled[t][c].size() != cpid[g].size()) throw Fail("InvalidStructure: samples per column");
      S.start[g + 1] = S.start[g] + (uint32_t)cpid[g].size();
      S.pid.insert(S.pid.end(), cpid[g].begin(), cpid[g].end());
      S.val.insert(S.val.end(), d.sampled[t][c].begin(), d.sampled[t][c].end());
    }
  }
  // the composition polynomial at the OODS point against the constraints
  {
    const size_t lt = tlogs.size() - 1;
    const qm31 lft = from_partial_evals(S.vals(lt, 0)[0], S.vals(lt, 1)[0], S.vals(lt, 2)[0], S.vals(lt, 3)[0]);
    const qm31 rgt = from_partial_evals(S.vals(lt, 4)[0], S.vals(lt, 5)[0], S.vals(lt, 6)[0], S.vals(lt, 7)[0]);
    const qm31 committed = qm_add(lft, qm_mul(cpq_repeated_double(oods, max_deg - 1).x, rgt));
    qm31 acc = qm_zero();
    for (auto& c : comps) {
      PointEval e;
      e.sv = &S;
      e.c = &c;
      e.acc = &acc;
      e.rc = random_coeff;
      e.dinv = qm_inv(coset_vanishing_qm31(max_deg, oods));
      eval_at_point(e);
    }
    if (!qm_eq(committed, acc)) throw Fail("OodsNotMatching");
  }
  // ---- verify_values ----
  ch.mix_felts(S.val);
  const qm31 qrc = ch.draw_secure_felt();
  uint32_t lls = 0;
  for (size_t t = 0; t < tlogs.size(); ++t)
    for (size_t c = 0; c < tlogs[t].size(); ++c)
      if (S.count(t, c)) lls = std::max(lls, tlogs[t][c]);
  if (lls < blow + 1) throw Fail("InvalidStructure: lifting size");
  // FRI commitment phase
  ch.mix_root(d.first.commitment);
  const qm31 alpha0 = ch.draw_secure_felt();
  struct Inner { uint32_t dom_init, dom_log, fold_step; qm31 alpha; };
  std::vector<Inner> inner;
  uint32_t layer_bound = lls - blow - 1;  // fold_to_line
  HCoset ldom = HCoset::half_odds(layer_bound + blow);
  for (size_t i = 0; i < d.inner.size(); ++i) {
    ch.mix_root(d.inner[i].commitment);
    uint32_t fs = cfg.fold_step;
    if (i + 1 == d.inner.size()) {
      if (layer_bound < cfg.log_last_layer) throw Fail("InvalidNumFriLayers");
      fs = layer_bound - cfg.log_last_layer;
      if (fs < 1 || fs > cfg.fold_step) throw Fail("InvalidNumFriLayers");
    }
    inner.push_back({ldom.initial_index, ldom.log_size, fs, ch.draw_secure_felt()});
    if (fs > layer_bound) throw Fail("InvalidNumFriLayers");
    layer_bound -= fs;
    for (uint32_t k = 0; k < fs; ++k) ldom = ldom.doubled();
  }
  if (layer_bound != cfg.log_last_layer) throw Fail("InvalidNumFriLayers");
  if (d.last_poly.size() > ((size_t)1 << cfg.log_last_layer) || d.last_poly.empty() ||
      (d.last_poly.size() & (d.last_poly.size() - 1)))
    throw Fail("LastLayerDegreeInvalid");
  ch.mix_felts(d.last_poly);
  if (!MC::pow_ok(ch, cfg.pow_bits, d.pow)) throw Fail("ProofOfWork");
  ch.mix_u64(d.pow);
  std::vector<size_t> raw;
  {
    const uint32_t mask = (uint32_t)(((uint64_t)1 << lls) - 1);
    while (raw.size() < cfg.n_queries)
      for (auto x : ch.draw_u32s())
        if (raw.size() < cfg.n_queries) raw.push_back(x & mask);
  }
  const auto qpos = sorted_unique(raw);
  // the trees' decommitments at their tree-local positions
  if (d.decs.size() != tlogs.size() || d.queried.size() != tlogs.size()) throw Fail("InvalidStructure: decommitments");
  std::vector<std::vector<uint32_t>> colq;  // per global column: value at each query, qpos order
  for (size_t t = 0; t < tlogs.size(); ++t) {
    const uint32_t tmax = tlogs[t].empty() ? 0 : *std::max_element(tlogs[t].begin(), tlogs[t].end());
    const auto P = tree_local_positions(qpos, lls, tmax);
    const auto qpl = queries_per_log_size(tlogs[t], P);
    merkle_verify<MC>(d.commitments[t], tlogs[t], qpl, d.queried[t], d.decs[t], "tree");
    // reshape_flat_to_per_column (core/pcs/quotients.rs:360-440): decommit order is by size,
    // largest first, node by node, the size's columns in order
    std::map<uint32_t, std::vector<size_t>> by;
    for (size_t c = 0; c < tlogs[t].size(); ++c) by[tlogs[t][c]].push_back(c);
    std::vector<std::vector<uint32_t>> nat(tlogs[t].size());
    size_t fi = 0;
    for (auto it = by.rbegin(); it != by.rend(); ++it)
      for (size_t k = 0; k < qpl.at(it->first).size(); ++k)
        for (size_t c : it->second) nat[c].push_back(d.queried[t][fi++]);
    for (size_t c = 0; c < tlogs[t].size(); ++c) {
      const auto& pos = qpl.at(tlogs[t][c]);
      std::vector<uint32_t> v;
      for (size_t p : P) {
        const size_t lifted = ((p >> (tmax - tlogs[t][c] + 1)) << 1) + (p & 1);
        v.push_back(nat[c][std::lower_bound(pos.begin(), pos.end(), lifted) - pos.begin()]);
      }
      colq.push_back(std::move(v));
    }
  }
  // fri_answers: samples with the random powers (a periodicity sample first for two-sample
  // columns), batches by point in order of first use, line coefficients, one quotient per query
  struct Num { uint32_t col; qm31 a, b, c; };
  struct Batch { cpqm31 p; std::vector<Num> nums; };
  std::vector<Batch> batches;
  qm31 rp = qm_one();
  const cpm31 lift_gen = index_to_point(subgroup_gen(lls));
  auto add_sample = [&](uint32_t g, const cpqm31& p, const qm31& v) {
    if (qm_eq(p.y, qm_conj(p.y))) throw Fail("InvalidStructure: a sample point with a base-field y");
    size_t b = 0;
    while (b < batches.size() && !point_eq(batches[b].p, p)) ++b;
    if (b == batches.size()) batches.push_back({p, {}});
    const qm31 a = qm_sub(qm_conj(v), v), c = qm_sub(qm_conj(p.y), p.y);
    const qm31 bb = qm_sub(qm_mul(v, c), qm_mul(a, p.y));
    batches[b].nums.push_back({g, qm_mul(rp, a), qm_mul(rp, bb), qm_mul(rp, c)});
    rp = qm_mul(rp, qrc);
  };
  for (size_t t = 0; t < tlogs.size(); ++t)
    for (size_t c = 0; c < tlogs[t].size(); ++c) {
      const uint32_t g = S.col_base[t] + (uint32_t)c, n = S.count(t, c), f0 = S.first(t, c);
      if (n == 2) {
        cpm31 pg = lift_gen;
        for (uint32_t i = 0; i < tlogs[t][c]; ++i) pg = cp_double(pg);
        add_sample(g, cpq_add(S.pts[S.pid[f0 + 1]], cpq_from(pg)), S.val[f0 + 1]);
      }
      for (uint32_t j = 0; j < n; ++j) add_sample(g, S.pts[S.pid[f0 + j]], S.val[f0 + j]);
    }
  std::vector<qm31> answers;
  for (size_t qi = 0; qi < qpos.size(); ++qi) {
    const cpm31 dp = canonic_domain_at(lls, bitrev((uint32_t)qpos[qi], lls));
    qm31 acc = qm_zero();
    for (auto& bt : batches) {
      qm31 num = qm_zero();
      for (auto& nm : bt.nums)
        num = qm_add(num, qm_sub(qm_mul_m31(nm.c, colq[nm.col][qi]), qm_add(qm_mul_m31(nm.a, dp.y), nm.b)));
      const cm31 prx = qm_lo(bt.p.x), pix = qm_hi(bt.p.x), pry = qm_lo(bt.p.y), piy = qm_hi(bt.p.y);
      const cm31 dd = cm_sub(cm_mul(cm_sub(prx, cm31{dp.x, 0}), piy), cm_mul(cm_sub(pry, cm31{dp.y, 0}), pix));
      if (dd.a == 0 && dd.b == 0) throw Fail("InvalidStructure: a query on a sample line");
      acc = qm_add(acc, qm_mul_cm31(num, cm_inv(dd)));
    }
    answers.push_back(acc);
  }
  // FRI decommitment: the first layer
  const Sparse s0 = rebuild(qpos, lls, answers, d.first.witness, 1, "FirstLayer");
  merkle_verify<MC>(d.first.commitment, std::vector<uint32_t>(4, lls), {{lls, s0.positions}}, flat_coords(s0),
                d.first.dec, "first FRI layer");
  const HCoset src_half = HCoset::half_odds(lls - 1);
  std::vector<qm31> ev;
  for (size_t i = 0; i < s0.subsets.size(); ++i) {
    const cpm31 p = index_to_point(circle_domain_index_at(src_half, s0.initial[i]));
    const qm31 e0 = s0.subsets[i][0], e1 = s0.subsets[i][1];
    const qm31 f0 = qm_add(e0, e1), f1 = qm_mul_m31(qm_sub(e0, e1), m31_inv(p.y));
    ev.push_back(qm_add(qm_mul(alpha0, f1), f0));
  }
  std::vector<size_t> q = fold_queries(qpos, 1);
  uint32_t qlog = lls - 1;
  for (size_t i = 0; i < inner.size(); ++i) {
    const Inner& L = inner[i];
    const Sparse s = rebuild(q, qlog, ev, d.inner[i].witness, L.fold_step, "InnerLayer");
    merkle_verify<MC>(d.inner[i].commitment, std::vector<uint32_t>(4, L.dom_log), {{L.dom_log, s.positions}},
                  flat_coords(s), d.inner[i].dec, "inner FRI layer");
    const HCoset dom = HCoset::make(L.dom_init, L.dom_log);
    ev.clear();
    for (size_t k = 0; k < s.subsets.size(); ++k)
      ev.push_back(fold_coset(s.subsets[k], dom.index_at(s.initial[k]), L.fold_step, L.alpha));
    q = fold_queries(q, L.fold_step);
    qlog -= L.fold_step;
  }
  // the last layer's polynomial at the remaining queries
  for (size_t k = 0; k < q.size(); ++k) {
    const uint32_t x = ldom.at(bitrev((uint32_t)q[k], ldom.log_size)).x;
    if (!qm_eq(ev[k], eval_line_poly(d.last_poly, qm_from_m31(x)))) throw Fail("LastLayerEvaluationsInvalid");
  }
  return "";
}

}  // namespace

std::string verify_proof(const uint8_t* bytes, size_t len, const std::vector<Component>& comps, const PcsCfg& cfg,
                         MerkleKind mc) {
  try {
    return mc == MerkleKind::Poseidon252 ? verify_impl<Poseidon252MC>(bytes, len, comps, cfg)
                                         : verify_impl<Blake2sMC>(bytes, len, comps, cfg);
  } catch (const Fail& f) {
    return f.what();
  }
}

}  // namespace sh
