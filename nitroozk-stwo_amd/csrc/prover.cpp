// Device-resident prover of libstwo_hip.
//
// Follows the reference protocol step by step (paths relative to crates/stwo/src):
//   prove / prove_ex                   prover/mod.rs:28-150
//   CommitmentSchemeProver             prover/pcs/mod.rs:36-494
//   compute_fri_quotients              prover/pcs/quotient_ops.rs:73-145
//   samples, batches, line coeffs      core/pcs/quotients.rs:71-366; core/constraints.rs:119-140
//   FriProver commit / decommit        prover/fri.rs:100-412
//   MerkleProver commit / decommit     prover/vcs/prover.rs:39-221
//   composition accumulator            prover/air/accumulation.rs:21-125
//   component domain evaluation        crates/constraint-framework/src/prover/component_prover.rs:32-169
// Column data, Merkle layers, FRI layers and quotients stay in HBM; the host holds the
// Fiat-Shamir channel and reads back only roots, sampled values, the last FRI layer, the
// proof-of-work nonce and the decommitted values.
#include "prover.h"
#include "protocol.h"

#include <algorithm>
#include <array>
#include <functional>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>

namespace sh {

namespace {

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double lap() {
    auto t = std::chrono::steady_clock::now();
    double ms = std::chrono::duration<double, std::milli>(t - t0).count();
    t0 = t;
    return ms;
  }
};

// Host phase marks, printed to stderr after each proof when STWO_HIP_HOST_PHASES is set: the
// host-side time between consecutive marks (including any waits on the device).
struct HostMarks {
  bool on = std::getenv("STWO_HIP_HOST_PHASES") != nullptr;
  std::vector<std::pair<const char*, std::chrono::steady_clock::time_point>> v;
  void mark(const char* name) {
    if (on) v.push_back({name, std::chrono::steady_clock::now()});
  }
  void flush() {
    if (!on || v.empty()) return;
    for (size_t i = 1; i < v.size(); ++i)
      std::fprintf(stderr, "%s=%.0f ", v[i].first,
                   std::chrono::duration<double, std::micro>(v[i].second - v[i - 1].second).count());
    std::fprintf(stderr, "\n");
    v.clear();
  }
};
thread_local HostMarks g_marks;
#define HMARK(name) g_marks.mark(name)

Hash read_hash(Ctx& ctx, const uint32_t* dev) {
  uint32_t w[8];
  ctx.d2h(w, dev, 32);
  Hash h;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 4; ++j) h[4 * i + j] = (uint8_t)(w[i] >> (8 * j));
  return h;
}

Hash words_to_hash(const std::vector<uint32_t>& w, uint32_t at) {
  Hash h;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 4; ++j) h[4 * i + j] = (uint8_t)(w[at + i] >> (8 * j));
  return h;
}

// MerkleProver::decommit (prover/vcs/prover.rs:84-217) and FriProver::decommit
// (prover/fri.rs:215-412) served by one device read per proof. At every level the reference walk
// only touches nodes inside the block of 2^w nodes that holds a query's path node (w = 1 for
// trees, the fold step for FRI layers), so launch_decommit_gather copies, per query, each level's
// column values at that block and the child hashes under it; the walk below then picks the words
// the reference emits, in its order, from that superset.
struct DecObject {
  const DevTree* t = nullptr;
  bool empty = true;  // no query reaches the tree (a tree of one node)
  uint32_t top = 0, w = 1, pos_mode = 2, pos_arg = 0;
  std::vector<uint32_t> cols_out, cols_n, hash_out;  // per level: superset offsets (query 0), column counts
};
struct DecLayout {
  std::vector<DecObject> objs;
  std::vector<DecItem> items;
  size_t words = 0;
  uint32_t n_q = 0;  // query blocks reserved per item
};
// Adds tree t, its query positions given by (mode, arg) of the global query positions.
void dec_add_object(Ctx& ctx, DecLayout& L, const DevTree& t, uint32_t w, uint32_t mode, uint32_t arg) {
  DecObject ob;
  ob.t = &t;
  ob.top = t.max_log();
  ob.w = w;
  ob.pos_mode = mode;
  ob.pos_arg = arg;
  ob.empty = ob.top == 0;  // tree_local_positions: a tree of one node has no queries
  if (ob.empty) { L.objs.push_back(ob); return; }
  ob.cols_out.assign(ob.top + 1, UINT32_MAX);
  ob.cols_n.assign(ob.top + 1, 0);
  ob.hash_out.assign(ob.top + 1, UINT32_MAX);
  const auto order = merkle_order(t.eval_logs);
  size_t ci = 0;
  std::vector<const uint32_t*> cols;
  for (int ls = (int)ob.top; ls >= 0; --ls) {
    const uint32_t shift = ob.top - (uint32_t)ls + w;
    cols.clear();
    while (ci < order.size() && t.eval_logs[order[ci]] == (uint32_t)ls) cols.push_back(t.evals[order[ci++]]);
    if (!cols.empty()) {
      const uint32_t nc = (uint32_t)cols.size(), stride = nc << w, chunk = std::max(1u, 1024u >> w);
      const uint32_t* const* tab = ctx.table_vec(cols);
      ob.cols_out[ls] = (uint32_t)L.words;
      ob.cols_n[ls] = nc;
      for (uint32_t c0 = 0; c0 < nc; c0 += chunk) {
        DecItem it{};
        it.cols = tab + c0;
        it.kind = DEC_COLS;
        it.width = std::min(chunk, nc - c0);
        it.row_width = nc;
        it.rows_log = w;
        it.level_log = (uint32_t)ls;
        it.shift = shift;
        it.pos_mode = mode;
        it.pos_arg = arg;
        it.out = (uint32_t)L.words + c0;
        it.stride = stride;
        L.items.push_back(it);
      }
      L.words += (size_t)stride * L.n_q;
    }
    if ((uint32_t)ls < ob.top) {
      DecItem it{};
      it.layer = t.layers[ls + 1].p;
      it.kind = DEC_HASH;
      it.width = 8;
      it.row_width = 8;
      it.rows_log = w + 1;
      it.level_log = (uint32_t)ls + 1;
      it.shift = shift;
      it.pos_mode = mode;
      it.pos_arg = arg;
      it.out = (uint32_t)L.words;
      it.stride = 8u << (w + 1);
      ob.hash_out[ls] = it.out;
      L.items.push_back(it);
      L.words += (size_t)it.stride * L.n_q;
    }
  }
  if (L.words >= ((size_t)1 << 32)) throw HipError(3, "decommitment read too large");
  L.objs.push_back(std::move(ob));
}
uint32_t dec_pos_host(uint32_t mode, uint32_t a, uint32_t p) {
  return mode == 0 ? ((p >> 1) << a) + (p & 1) : mode == 1 ? ((p >> a) << 1) + (p & 1) : p >> a;
}
// Superset offsets of what the reference decommitment emits, in its order: the first word of each
// witness hash, and runs (offset, length) of queried and witness column values.
struct DecPlan {
  std::vector<uint32_t> hash_at;
  std::vector<std::pair<uint32_t, uint32_t>> qry, wit;
};
// The walk of MerkleProver::decommit for object ob. P[i]: query i's position in the top level;
// qpl: (level, sorted unique queried nodes) for the levels with queried columns.
void dec_plan_object(const DecObject& ob, const std::vector<uint32_t>& P,
                     const std::vector<std::pair<uint32_t, std::vector<size_t>>>& qpl, DecPlan& dp) {
  if (ob.empty) return;
  static const std::vector<size_t> none;
  std::vector<size_t> last, merged;
  const uint32_t w = ob.w, nodes_in_block = 1u << w;
  for (int ls = (int)ob.top; ls >= 0; --ls) {
    const std::vector<size_t>* colq = &none;
    for (auto& e : qpl)
      if (e.first == (uint32_t)ls) colq = &e.second;
    merged.clear();
    for (size_t q : last) merged.push_back(q / 2);
    merged.insert(merged.end(), colq->begin(), colq->end());
    std::sort(merged.begin(), merged.end());
    merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
    const uint32_t shift = ob.top - (uint32_t)ls + w;
    auto query_of = [&](size_t block) -> uint32_t {
      for (uint32_t i = 0; i < P.size(); ++i)
        if ((P[i] >> shift) == block) return i;
      throw HipError(4, "decommitment node outside the queried blocks");
    };
    if ((uint32_t)ls < ob.top) {
      const uint32_t hs = 8u << (w + 1);
      for (size_t node : merged)
        for (size_t child = 2 * node; child < 2 * node + 2; ++child)
          if (!std::binary_search(last.begin(), last.end(), child))
            dp.hash_at.push_back(ob.hash_out[ls] + query_of(child >> (w + 1)) * hs +
                                 (uint32_t)(child & (2 * nodes_in_block - 1)) * 8);
    }
    const uint32_t nc = ob.cols_n[ls];
    if (nc) {
      auto at = [&](size_t node) {
        return ob.cols_out[ls] + query_of(node >> w) * (nc << w) + (uint32_t)(node & (nodes_in_block - 1)) * nc;
      };
      for (size_t node : merged)
        if (std::binary_search(colq->begin(), colq->end(), node)) dp.qry.push_back({at(node), nc});
      for (size_t node : merged)
        if (!std::binary_search(colq->begin(), colq->end(), node)) dp.wit.push_back({at(node), nc});
    }
    last.swap(merged);
  }
}
void dec_finish(const DecPlan& dp, const uint32_t* sup, MerkleDec& dec, std::vector<uint32_t>* queried) {
  dec.hash_witness.resize(dp.hash_at.size());
  for (size_t k = 0; k < dp.hash_at.size(); ++k) std::memcpy(dec.hash_witness[k].data(), sup + dp.hash_at[k], 32);
  for (auto& r : dp.wit) dec.column_witness.insert(dec.column_witness.end(), sup + r.first, sup + r.first + r.second);
  if (queried)
    for (auto& r : dp.qry) queried->insert(queried->end(), sup + r.first, sup + r.first + r.second);
}

// Whether column k is at v[0] + k stride for every k (one allocation, no address table needed).
bool is_strided(const std::vector<const uint32_t*>& v, size_t stride) {
  for (size_t k = 1; k < v.size(); ++k)
    if (v[k] != v[0] + k * stride) return false;
  return !v.empty();
}

void evaluate_into(Ctx& ctx, const std::vector<const uint32_t*>& coeffs, const std::vector<uint32_t>& logs,
                   uint32_t target_log, std::vector<DevBuf>& out, std::vector<DevBuf>& keep, bool on_side = false) {
  // evaluates every column onto CanonicCoset(target_log), grouping equal sizes into one batch; the
  // columns of a batch are consecutive in one allocation (kept in `keep`; out holds views)
  std::map<uint32_t, std::vector<size_t>> by;
  for (size_t i = 0; i < coeffs.size(); ++i) by[logs[i]].push_back(i);
  out.resize(coeffs.size());
  ctx.ensure_twiddles(target_log);
  const size_t n = (size_t)1 << target_log;
  for (auto& [lg, idx] : by) {
    std::vector<const uint32_t*> src;
    std::vector<uint32_t*> dst;
    keep.emplace_back(&ctx, idx.size() * n);
    uint32_t* blk = keep.back().p;
    for (size_t j = 0; j < idx.size(); ++j) {
      const size_t i = idx[j];
      out[i] = DevBuf::view(&ctx, blk + j * n, n);
      src.push_back(coeffs[i]);
      dst.push_back(out[i].p);
    }
    const uint32_t* const* d_src = (const uint32_t* const*)ctx.table_vec(src);
    uint32_t* const* d_dst = (uint32_t* const*)ctx.table_vec(dst);
    const uint32_t low = Ctx::fft_plan_for(
        ctx.fft_low_log(false, (uint32_t)idx.size(), lg, target_log, ctx.tw2.p, ctx.tree_len(), on_side), dst.data(),
        dst.size());
    hipStream_t st = ctx.stream;
    if (on_side) {
      ctx.fork();  // the side stream starts after the upload and everything queued before it
      st = ctx.side;
    }
    KScope ks(ctx, K_EVALUATE, (double)idx.size() * (4.0 * (1u << lg) + 4.0 * (1u << target_log)), st);
    ks.units(fft_butterflies(idx.size(), target_log, lg));
    launch_fft(false, d_src, d_dst, (uint32_t)idx.size(), lg, target_log, ctx.tw2.p, ctx.tree_len(), st, on_side, low);
  }
}

}  // namespace

// Candidates of a grind's first launch: 32 times the expected nonce (about 2^pow_bits), so that a
// miss has probability about e^-32, within [2^12, 2^24].
static uint32_t grind_first_count(uint32_t pow_bits) {
  return 1u << (pow_bits + 5 < 12 ? 12 : (pow_bits + 5 > 24 ? 24 : pow_bits + 5));
}

uint64_t device_grind(Ctx& ctx, const Hash& prefix, uint32_t pow_bits, unsigned long long* slot) {
  if (pow_bits > 32) throw HipError(3, "pow_bits > 32 is not supported");
  GrindPrefix pre;
  for (int i = 0; i < 8; ++i)
    pre.w[i] = (uint32_t)prefix[4 * i] | ((uint32_t)prefix[4 * i + 1] << 8) | ((uint32_t)prefix[4 * i + 2] << 16) |
               ((uint32_t)prefix[4 * i + 3] << 24);
  DevBuf own;
  unsigned long long found = ~0ull;
  if (!slot) {
    own = DevBuf(&ctx, 2);
    slot = (unsigned long long*)own.p;
    HIP_CHECK(hipMemcpyAsync(slot, &found, 8, hipMemcpyHostToDevice, ctx.stream));
  }
  // later launches double the range
  uint64_t count = grind_first_count(pow_bits);
  const uint64_t limit = (uint64_t)0x7fffffff << 20;  // hi < P (simd/grind.rs:47-56)
  for (uint64_t start = 0; found == ~0ull; start += count, count = count < (1u << 24) ? count * 2 : count) {
    if (start >= limit) throw HipError(4, "Grind failed to find a solution.");
    {
      KScope ks(ctx, K_GRIND, 0.0);
      launch_grind(pre, pow_bits, start, (uint32_t)count, slot, ctx.stream);
    }
    ctx.d2h(&found, slot, 8);
  }
  return found;
}

// Wide-Fibonacci components on evaluation domains below 2^WF_GROUP_LOG share one composition launch.
constexpr uint32_t WF_GROUP_LOG = 16;
// Levels of at least 2^MERKLE_WIDE_LOG nodes are hashed one launch per level.
constexpr uint32_t MERKLE_WIDE_LOG = 17;

void merkle_commit(Ctx& ctx, DevTree& t, bool read_root) {
  t.layers.clear();
  if (t.evals.empty()) {
    // a tree without columns is one node, Blake2s of the empty string: known on the host
    t.root = blake2s_hash(nullptr, 0);
    uint32_t w[8];
    for (int i = 0; i < 8; ++i)
      w[i] = (uint32_t)t.root[4 * i] | (uint32_t)t.root[4 * i + 1] << 8 | (uint32_t)t.root[4 * i + 2] << 16 |
             (uint32_t)t.root[4 * i + 3] << 24;
    t.layers.emplace_back(&ctx, 8);
    launch_copy_words(t.layers[0].p, (const uint32_t*)ctx.upload(w, sizeof(w)), 8, ctx.stream);
    return;
  }
  auto order = merkle_order(t.eval_logs);
  uint32_t max_log = t.eval_logs[order[0]];
  std::vector<DevBuf> rev;  // layers from max_log down to 0
  std::vector<uint32_t> has_cols(max_log + 1, 0);
  for (auto l : t.eval_logs) has_cols[l] = 1;
  size_t ci = 0;
  int ls = (int)max_log;
  while (ls >= 0) {
    if (!rev.empty() && !has_cols[ls] && ls >= (int)MERKLE_WIDE_LOG) {
      // a wide level without injected columns: one lane per node, one launch per level (the fused
      // subtree kernel hashes a node with a quad of lanes and waits between levels inside a
      // workgroup, which pays on narrow levels only; on 2^22 nodes it ran at a quarter of this)
      DevBuf layer(&ctx, (size_t)8 << ls);
      KScope ks(ctx, K_MERKLE, 96.0 * (double)(1u << ls));
      ks.units((double)(1u << ls));
      launch_merkle_layer_small((uint32_t)ls, rev.back().p, nullptr, 0, layer.p, ctx.stream);
      rev.push_back(std::move(layer));
      --ls;
      continue;
    }
    if (!rev.empty() && !has_cols[ls]) {
      // run of levels without injected columns: fused, up to 9 levels per launch
      uint32_t levels = 0;
      while ((int)levels < 9 && ls - (int)levels >= 0 && !has_cols[ls - levels]) ++levels;
      uint32_t log_child = (uint32_t)ls + 1;
      std::vector<uint32_t*> outs;
      for (uint32_t j = 0; j < levels; ++j) {
        rev.emplace_back(&ctx, (size_t)8 << (ls - j));
        outs.push_back(rev.back().p);
      }
      const uint32_t* child = rev[rev.size() - levels - 1].p;
      KScope ks(ctx, K_MERKLE, 96.0 * (double)(1u << (log_child - 1)) * 2.0);
      ks.units((double)((1u << log_child) - (1u << (log_child - levels))));  // one per node of the levels
      launch_merkle_subtree(log_child, child, levels, outs.data(), ctx.stream);
      ls -= (int)levels;
      continue;
    }
    std::vector<const uint32_t*> lc;
    while (ci < order.size() && t.eval_logs[order[ci]] == (uint32_t)ls) lc.push_back(t.evals[order[ci++]]);
    DevBuf layer(&ctx, (size_t)8 << ls);
    const uint32_t* prev = rev.empty() ? nullptr : rev.back().p;
    const double comps = blake2s_node_compressions(prev != nullptr, lc.size()) * (double)(1u << ls);
    if (lc.size() <= 16) {
      KScope ks(ctx, K_MERKLE, ((prev ? 64.0 : 0.0) + 4.0 * lc.size() + 32.0) * (double)(1u << ls));
      ks.units(comps);
      launch_merkle_layer_small((uint32_t)ls, prev, lc.data(), (uint32_t)lc.size(), layer.p, ctx.stream);
    } else if (is_strided(lc, (size_t)1 << ls)) {
      KScope ks(ctx, K_MERKLE, ((prev ? 64.0 : 0.0) + 4.0 * lc.size() + 32.0) * (double)(1u << ls));
      ks.units(comps);
      launch_merkle_layer_strided((uint32_t)ls, prev, lc[0], (size_t)1 << ls, (uint32_t)lc.size(), layer.p, ctx.stream);
    } else {
      const uint32_t* const* cp = (const uint32_t* const*)ctx.table_vec(lc);
      KScope ks(ctx, K_MERKLE, ((prev ? 64.0 : 0.0) + 4.0 * lc.size() + 32.0) * (double)(1u << ls));
      ks.units(comps);
      launch_merkle_layer((uint32_t)ls, prev, cp, (uint32_t)lc.size(), layer.p, ctx.stream);
    }
    rev.push_back(std::move(layer));
    --ls;
  }
  for (size_t i = rev.size(); i-- > 0;) t.layers.push_back(std::move(rev[i]));
  if (read_root) t.root = read_hash(ctx, t.layers[0].p);
}

void interpolate_columns(Ctx& ctx, const std::vector<const uint32_t*>& evals, const std::vector<uint32_t>& logs,
                         DevTree& into) {
  std::map<uint32_t, std::vector<size_t>> by;
  for (size_t i = 0; i < evals.size(); ++i) by[logs[i]].push_back(i);
  size_t base = into.coeffs.size();
  into.coeffs.resize(base + evals.size());
  into.coeff_logs.resize(base + evals.size());
  for (auto& [lg, idx] : by) {
    if (lg < 3) throw HipError(3, "interpolate: log size < 3 is not supported");
    ctx.ensure_twiddles(lg);
    std::vector<const uint32_t*> src;
    std::vector<uint32_t*> dst;
    // the group's columns are consecutive in one allocation
    into.owned.emplace_back(&ctx, idx.size() << lg);
    uint32_t* blk = into.owned.back().p;
    for (size_t j = 0; j < idx.size(); ++j) {
      const size_t i = idx[j];
      src.push_back(evals[i]);
      dst.push_back(blk + (j << lg));
      into.coeffs[base + i] = dst.back();
      into.coeff_logs[base + i] = lg;
    }
    const uint32_t* const* d_src = (const uint32_t* const*)ctx.table_vec(src);
    uint32_t* const* d_dst = (uint32_t* const*)ctx.table_vec(dst);
    const uint32_t low = Ctx::fft_plan_for(ctx.fft_low_log(true, (uint32_t)idx.size(), lg, lg, ctx.itw2.p, ctx.tree_len()),
                                           dst.data(), dst.size());
    KScope ks(ctx, K_INTERPOLATE, (double)idx.size() * 8.0 * (1u << lg));
    ks.units(fft_butterflies(idx.size(), lg, lg));
    launch_fft(true, d_src, d_dst, (uint32_t)idx.size(), lg, lg, ctx.itw2.p, ctx.tree_len(), ctx.stream, false, low);
  }
}

void evaluate_polynomials(Ctx& ctx, DevTree& t, uint32_t log_blowup) {
  std::map<uint32_t, std::vector<size_t>> by;
  for (size_t i = 0; i < t.coeffs.size(); ++i) by[t.coeff_logs[i]].push_back(i);
  t.evals.assign(t.coeffs.size(), nullptr);
  t.eval_logs.assign(t.coeffs.size(), 0);
  for (auto& [lg, idx] : by) {
    uint32_t K = lg + log_blowup;
    ctx.ensure_twiddles(K);
    std::vector<const uint32_t*> src;
    std::vector<uint32_t*> dst;
    // the group's columns are consecutive in one allocation
    t.owned.emplace_back(&ctx, idx.size() << K);
    uint32_t* blk = t.owned.back().p;
    for (size_t j = 0; j < idx.size(); ++j) {
      const size_t i = idx[j];
      src.push_back(t.coeffs[i]);
      dst.push_back(blk + (j << K));
      t.evals[i] = dst.back();
      t.eval_logs[i] = K;
    }
    const uint32_t* const* d_src = (const uint32_t* const*)ctx.table_vec(src);
    uint32_t* const* d_dst = (uint32_t* const*)ctx.table_vec(dst);
    const uint32_t low = Ctx::fft_plan_for(ctx.fft_low_log(false, (uint32_t)idx.size(), lg, K, ctx.tw2.p, ctx.tree_len()),
                                           dst.data(), dst.size());
    KScope ks(ctx, K_EVALUATE, (double)idx.size() * (4.0 * (1u << lg) + 4.0 * (1u << K)));
    ks.units(fft_butterflies(idx.size(), K, lg));
    launch_fft(false, d_src, d_dst, (uint32_t)idx.size(), lg, K, ctx.tw2.p, ctx.tree_len(), ctx.stream, false, low);
  }
}

namespace {

// CommitmentSchemeProver::commit + TreeBuilder::commit for trees given as coefficients.
// `before_root` runs on the host after the tree is queued and before its root is read.
void commit_tree(Ctx& ctx, DevTree& t, const PcsCfg& cfg, Channel& ch, const std::function<void()>& before_root = {}) {
  evaluate_polynomials(ctx, t, cfg.log_blowup);
  merkle_commit(ctx, t, false);
  if (before_root) before_root();
  if (!t.evals.empty()) t.root = read_hash(ctx, t.layers[0].p);
  ch.mix_root(t.root);
  ctx.reset_arena();  // the root read synchronised the stream
}

// The composition's launches up to its random coefficient: per component the accumulator, the
// columns on the evaluation domain (extensions queued here when needed), the vanishing
// denominators and the column tables. Laid out before the last commitment's root is read, so the
// host does it while the device still works on that commitment.
struct CompLaunch {
  uint32_t E = 0;
  bool accumulate = false, join_side = false;
  Cols4 a4{};
  std::vector<const uint32_t*> tcols, icols;
  const uint32_t* const* d_t = nullptr;
  const uint32_t* const* d_i = nullptr;
  const uint32_t* d_dinv = nullptr;
  double cbytes = 0;
  std::vector<DevBuf> ext_t, ext_i, ext_keep;  // alive until the composition is queued
};
struct CompPrep {
  bool ready = false;
  uint32_t max_log = 0;
  size_t T = 0;
  std::map<uint32_t, std::vector<DevBuf>> accs;
  std::vector<CompLaunch> items;
};

void prepare_composition(Ctx& ctx, std::vector<DevTree>& trees, const std::vector<Component>& comps, CompPrep& P) {
  for (auto& c : comps) { P.max_log = std::max(P.max_log, c.eval_log()); P.T += c.n_constraints(); }
  P.items.resize(comps.size());
  for (size_t ci = 0; ci < comps.size(); ++ci) {
    const Component& c = comps[ci];
    CompLaunch& L = P.items[ci];
    const uint32_t E = c.eval_log();
    L.E = E;
    ctx.ensure_twiddles(E);
    auto& acc = P.accs[E];
    // the first component of an evaluation size writes the accumulator, later ones add to it
    L.accumulate = !acc.empty();
    if (acc.empty()) for (int k = 0; k < 4; ++k) acc.emplace_back(&ctx, (size_t)1 << E);
    L.a4 = {{acc[0].p, acc[1].p, acc[2].p, acc[3].p}};
    // columns on the evaluation domain: committed evaluations when sizes match, else extended
    auto cols_on_domain = [&](const DevTree& t, size_t start, size_t n, std::vector<DevBuf>& ext) {
      std::vector<const uint32_t*> r(n);
      bool need = false;
      for (size_t i = 0; i < n; ++i) need |= t.eval_logs[start + i] != E;
      if (!need) { for (size_t i = 0; i < n; ++i) r[i] = t.evals[start + i]; return r; }
      if (t.ext_log == E && t.ext.size() >= start + n) {
        for (size_t i = 0; i < n; ++i) r[i] = t.ext[start + i].p;
        L.join_side = true;
        return r;
      }
      std::vector<const uint32_t*> cf(t.coeffs.begin() + start, t.coeffs.begin() + start + n);
      std::vector<uint32_t> lg(t.coeff_logs.begin() + start, t.coeff_logs.begin() + start + n);
      evaluate_into(ctx, cf, lg, E, ext, L.ext_keep);
      for (size_t i = 0; i < n; ++i) r[i] = ext[i].p;
      return r;
    };
    L.tcols = cols_on_domain(trees[1], c.trace_start, c.n_trace_cols(), L.ext_t);
    const uint32_t le = E - c.log_n;
    std::vector<uint32_t> dinv(1u << le);
    for (uint32_t i = 0; i < (1u << le); ++i) dinv[i] = m31_inv(coset_vanishing_m31(c.log_n, canonic_domain_at(E, i)));
    std::vector<uint32_t> dinv_br(dinv.size());
    for (uint32_t i = 0; i < dinv.size(); ++i) dinv_br[bitrev(i, le)] = dinv[i];
    L.d_dinv = ctx.table_vec(dinv_br);
    L.d_t = (const uint32_t* const*)ctx.table_vec(L.tcols);
    L.cbytes = (4.0 * (c.n_trace_cols() + c.n_inter_cols()) + 32.0) * (double)(1u << E);
    if (c.kind != AIR_WIDE_FIB) {
      L.icols = cols_on_domain(trees[2], c.inter_start, c.n_inter_cols(), L.ext_i);
      L.d_i = (const uint32_t* const*)ctx.table_vec(L.icols);
    }
  }
  P.ready = true;
}

// A FRI layer column (4 coordinate columns of 2^log) with its Merkle tree.
struct FriCol {
  std::vector<DevBuf> c;  // 4 coordinate columns
  uint32_t log = 0;
  DevTree tree;
  uint32_t fold_step = 1;
  CCols4 cc() const { return {{c[0].p, c[1].p, c[2].p, c[3].p}}; }
  Cols4 mc() { return {{c[0].p, c[1].p, c[2].p, c[3].p}}; }
};
// A FRI layer's decommit (prover/fri.rs:291-378) with
// compute_decommitment_positions_and_witness_evals (prover/fri.rs:383-412): the values of every
// query's fold coset that are not queried, then the Merkle decommitment of the coset positions. q: the layer's sorted
// unique query positions; P[i]: query i's position in the layer.
struct FriDecPlan {
  std::vector<uint32_t> witness_at;  // first word of each witness value (4 coordinates)
  DecPlan mp;
};
FriDecPlan plan_fri_layer(const DecObject& ob, const std::vector<size_t>& q, const std::vector<uint32_t>& P) {
  FriDecPlan fp;
  const uint32_t w = ob.w;
  std::vector<size_t> pos;
  size_t i = 0;
  while (i < q.size()) {
    size_t j = i;
    while (j < q.size() && (q[j] >> w) == (q[i] >> w)) ++j;
    const size_t start = (q[i] >> w) << w;
    size_t k = i;
    for (size_t p = start; p < start + ((size_t)1 << w); ++p) {
      pos.push_back(p);
      if (k < j && q[k] == p) {
        ++k;
        continue;
      }
      uint32_t qi = UINT32_MAX;
      for (uint32_t t = 0; t < P.size() && qi == UINT32_MAX; ++t)
        if ((P[t] >> w) == (p >> w)) qi = t;
      if (qi == UINT32_MAX) throw HipError(4, "FRI witness outside the queried blocks");
      fp.witness_at.push_back(ob.cols_out[ob.top] + qi * (4u << w) + (uint32_t)(p & ((1u << w) - 1)) * 4);
    }
    i = j;
  }
  std::vector<std::pair<uint32_t, std::vector<size_t>>> qpl{{ob.top, pos}};
  dec_plan_object(ob, P, qpl, fp.mp);
  return fp;
}
FriLayer finish_fri_layer(const FriDecPlan& fp, const uint32_t* sup, const Hash& commitment) {
  FriLayer L;
  for (uint32_t at : fp.witness_at) L.witness.push_back({sup[at], sup[at + 1], sup[at + 2], sup[at + 3]});
  dec_finish(fp.mp, sup, L.dec, nullptr);
  L.commitment = commitment;
  return L;
}

// prove_ex + prove_values. `trees` holds the committed preprocessed, trace (and interaction) trees.
int prove_ex(Ctx& ctx, Channel& ch, std::vector<DevTree>& trees, const std::vector<Component>& comps,
             const PcsCfg& cfg, ProofParts& out, double* stage_ms, CompPrep* prep = nullptr) {
  Timer tm;
  HMARK("prove_ex");
  // ---------------- composition ----------------
  CompPrep own_prep;
  if (!prep || !prep->ready) {
    prepare_composition(ctx, trees, comps, own_prep);
    prep = &own_prep;
  }
  qm31 random_coeff = ch.draw_secure_felt();
  const uint32_t max_log = prep->max_log;
  const size_t T = prep->T;
  // random_coeff^i, in four interleaved chains (exact, so the same values as one chain)
  std::vector<qm31> powers(T);
  {
    const qm31 a2 = qm_sqr(random_coeff), a4 = qm_sqr(a2);
    qm31 p[4] = {qm_one(), random_coeff, a2, qm_mul(a2, random_coeff)};
    for (size_t i = 0; i < T; i += 4)
      for (size_t j = 0; j < 4 && i + j < T; ++j) {
        powers[i + j] = p[j];
        p[j] = qm_mul(p[j], a4);
      }
  }
  HMARK("x_pows");
  auto& accs = prep->accs;
  // each component's constraint coefficients: its slice of the powers, reversed, in component order
  std::vector<const qm31*> comp_pows(comps.size());
  for (size_t ci = 0; ci < comps.size(); ++ci) {
    const size_t nc = comps[ci].n_constraints();
    std::vector<qm31> cp(powers.end() - nc, powers.end());
    powers.resize(powers.size() - nc);
    std::reverse(cp.begin(), cp.end());
    comp_pows[ci] = ctx.stage_vec(cp);
  }
  HMARK("x_tables");
  // Small wide-Fibonacci components (evaluation domains below 2^WF_GROUP_LOG) go in one grouped
  // launch, by evaluation size, when every component of that size is one: a separate launch of a
  // 2^11-row component occupies a few CUs for tens of microseconds.
  std::vector<char> grouped(comps.size(), 0);
  {
    std::map<uint32_t, std::vector<size_t>> small;
    std::set<uint32_t> other;
    for (size_t ci = 0; ci < comps.size(); ++ci) {
      const uint32_t E = prep->items[ci].E;
      if (comps[ci].kind == AIR_WIDE_FIB && E < WF_GROUP_LOG) small[E].push_back(ci);
      else other.insert(E);
    }
    for (uint32_t E : other) small.erase(E);
    size_t n_small = 0;
    for (auto& kv : small) n_small += kv.second.size();
    if (n_small >= 2) {
      std::vector<WfGroup> gs;
      std::vector<WfComp> cs;
      uint32_t blocks = 0;
      double bytes = 0;
      bool join = false;
      for (auto& [E, idx] : small) {
        const CompLaunch& L0 = prep->items[idx[0]];
        WfGroup G{};
        G.dinv = L0.d_dinv;
        for (int k = 0; k < 4; ++k) G.acc[k] = L0.a4.c[k];
        G.block_off = blocks;
        G.eval_log = E;
        G.trace_log = comps[idx[0]].log_n;
        G.accumulate = L0.accumulate ? 1u : 0u;
        G.comp_off = (uint32_t)cs.size();
        G.n_comps = (uint32_t)idx.size();
        for (size_t ci : idx) {
          const CompLaunch& L = prep->items[ci];
          // one evaluation size: the same trace size, denominators (equal tables) and accumulator
          if (comps[ci].log_n != G.trace_log || L.a4.c[0] != G.acc[0])
            throw HipError(4, "grouped composition: components of one size disagree");
          cs.push_back(WfComp{L.d_t, comp_pows[ci], comps[ci].n_cols, 0u});
          grouped[ci] = 1;
          bytes += L.cbytes;
          join |= L.join_side;
        }
        blocks += ((1u << E) + 63) / 64;
        gs.push_back(G);
      }
      if (join) ctx.join();
      const WfGroup* dg = ctx.stage_vec(gs);
      const WfComp* dc = ctx.stage_vec(cs);
      ctx.commit_staged();
      KScope ks(ctx, K_COMPOSITION, bytes);
      launch_wide_fib_composition_grouped(dg, (uint32_t)gs.size(), dc, blocks, ctx.stream);
    }
  }
  for (size_t ci = 0; ci < comps.size(); ++ci) {
    if (grouped[ci]) continue;
    const Component& c = comps[ci];
    CompLaunch& L = prep->items[ci];
    const uint32_t E = L.E;
    const qm31* d_pows = comp_pows[ci];
    if (L.join_side) ctx.join();
    if (c.kind == AIR_WIDE_FIB) {
      ctx.commit_staged();
      KScope ks(ctx, K_COMPOSITION, L.cbytes);
      launch_wide_fib_composition(E, c.log_n, L.d_t, c.n_cols, d_pows, L.d_dinv, L.a4, L.accumulate, ctx.stream);
    } else {
      PoseidonParams pp{};
      pp.z = c.z;
      for (int i = 0; i < 16; ++i) pp.alpha_pow[i] = c.alpha_pow[i];
      pp.shift = c.cumsum_shift();
      pp.ipr = c.ipr;
      const PoseidonParams* d_pp = (const PoseidonParams*)ctx.upload(&pp, sizeof(pp));
      KScope ks(ctx, K_COMPOSITION, L.cbytes);
      const size_t stride = (size_t)1 << E;
      if (is_strided(L.tcols, stride) && is_strided(L.icols, stride))
        launch_poseidon_composition_strided(E, c.log_n, L.tcols[0], L.icols[0], stride, d_pp, c.ipr, d_pows, L.d_dinv,
                                            L.a4, L.accumulate, ctx.stream);
      else
        launch_poseidon_composition(E, c.log_n, L.d_t, L.d_i, d_pp, c.ipr, d_pows, L.d_dinv, L.a4, L.accumulate,
                                    ctx.stream);
    }
  }
  // finalize: lift smaller accumulations into the largest, then interpolate 4 coordinate columns
  auto& big = accs[max_log];
  Cols4 b4 = {{big[0].p, big[1].p, big[2].p, big[3].p}};
  for (auto& [lg, a] : accs)
    if (lg != max_log) launch_accumulate_lifted(max_log, b4, lg, {{a[0].p, a[1].p, a[2].p, a[3].p}}, ctx.stream);
  HMARK("x_compkernel");
  DevTree comp;
  {
    std::vector<const uint32_t*> ev = {big[0].p, big[1].p, big[2].p, big[3].p};
    std::vector<uint32_t> lg(4, max_log);
    DevTree full;
    interpolate_columns(ctx, ev, lg, full);
    // split_at_mid: left halves then right halves, as views into the coefficient buffers
    size_t half = (size_t)1 << (max_log - 1);
    for (int k = 0; k < 4; ++k) { comp.coeffs.push_back(full.coeffs[k]); comp.coeff_logs.push_back(max_log - 1); }
    for (int k = 0; k < 4; ++k) { comp.coeffs.push_back(full.coeffs[k] + half); comp.coeff_logs.push_back(max_log - 1); }
    for (auto& o : full.owned) comp.owned.push_back(std::move(o));
  }
  accs.clear();
  HMARK("comp_launch");
  // the composition commitment is queued; its root is read after the host has laid out the OODS
  // and quotient work below, which depends on the proof's shape but not on the OODS point
  evaluate_polynomials(ctx, comp, cfg.log_blowup);
  merkle_commit(ctx, comp, false);
  trees.push_back(std::move(comp));
  HMARK("comp_queue");

  // ---------------- OODS ----------------
  cpqm31 oods;
  uint32_t split_log = trees.back().max_log();
  uint32_t L = cfg.lifting_log_size ? cfg.lifting_log_size : split_log;
  if (L < split_log) throw HipError(3, "lifting log size below the composition size");
  uint32_t max_deg = L - cfg.log_blowup;
  // sample points: the mask offsets of every trace and interaction column, and the OODS point
  // itself for the composition columns (prover/mod.rs:95-118, constraint-framework mask points)
  SampleSet& S = out.sampled;
  for (size_t t = 0; t < trees.size(); ++t) S.col_base.push_back(S.col_base.back() + (uint32_t)trees[t].evals.size());
  const uint32_t NC = S.col_base.back();
  const cpm31 step = index_to_point(subgroup_gen(max_deg));
  std::vector<long> pt_off;  // sample point p is the OODS point plus pt_off[p] steps
  auto pid_of = [&](long o) -> uint32_t {
    for (size_t i = 0; i < pt_off.size(); ++i)
      if (pt_off[i] == o) return (uint32_t)i;
    pt_off.push_back(o);
    return (uint32_t)pt_off.size() - 1;
  };
  std::vector<uint8_t> ccount(NC, 0);
  std::vector<std::array<uint32_t, 2>> cpid(NC);
  for (auto& c : comps)
    for (size_t t = 1; t < 3 && t + 1 < trees.size(); ++t) {
      size_t start = t == 1 ? c.trace_start : c.inter_start;
      size_t n = t == 1 ? c.n_trace_cols() : c.n_inter_cols();
      for (size_t i = 0; i < n; ++i) {
        long offs[2];
        uint32_t k = mask_offsets(c, t, i, offs);
        uint32_t g = S.col_base[t] + (uint32_t)(start + i);
        if (ccount[g] + k > 2) throw HipError(3, "more than two mask points on a column");
        for (uint32_t j = 0; j < k; ++j) cpid[g][ccount[g]++] = pid_of(offs[j]);
      }
    }
  for (uint32_t g = S.col_base[trees.size() - 1]; g < NC; ++g) { cpid[g][0] = pid_of(0); ccount[g] = 1; }
  S.start.assign(NC + 1, 0);
  for (uint32_t g = 0; g < NC; ++g) S.start[g + 1] = S.start[g] + ccount[g];
  S.pid.resize(S.start[NC]);
  for (uint32_t g = 0; g < NC; ++g)
    for (uint32_t j = 0; j < ccount[g]; ++j) S.pid[S.start[g] + j] = cpid[g][j];
  uint32_t Lc = trees.back().max_log();  // prove_values lifting size: the last tree
  // one OODS task per sample, grouped by (coefficient log size, point lifted to the column size)
  std::vector<OodsTask> tasks;
  tasks.reserve(S.pid.size());
  std::vector<uint32_t> task_group, group_log;
  task_group.reserve(S.pid.size());
  struct GKey { uint32_t k, pid, shift; };
  std::vector<GKey> group_key;
  uint32_t max_coeff_log = 0;
  for (size_t t = 0; t < trees.size(); ++t)
    for (size_t c = 0; c < trees[t].evals.size(); ++c)
      for (uint32_t j = 0; j < S.count(t, c); ++j) {
        const uint32_t pid = S.pid[S.first(t, c) + j];
        const uint32_t shift = Lc - trees[t].eval_logs[c];
        const uint32_t k = trees[t].coeff_logs[c];
        uint32_t g = 0;
        while (g < group_key.size() && !(group_key[g].k == k && group_key[g].pid == pid && group_key[g].shift == shift)) ++g;
        if (g == group_key.size()) {
          group_key.push_back({k, pid, shift});
          group_log.push_back(k);
        }
        tasks.push_back(OodsTask{trees[t].coeffs[c], k, (uint32_t)tasks.size(), {0, 0}});
        task_group.push_back(g);
        max_coeff_log = std::max(max_coeff_log, k);
      }
  const size_t n_vals = tasks.size();
  {
    // tasks of one group side by side, in batches that share each weight read (out_index keeps
    // every value's place)
    std::vector<uint32_t> perm(n_vals);
    for (uint32_t i = 0; i < n_vals; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return task_group[a] < task_group[b]; });
    std::vector<OodsTask> st(n_vals);
    std::vector<uint32_t> sg(n_vals);
    for (uint32_t i = 0; i < n_vals; ++i) { st[i] = tasks[perm[i]]; sg[i] = task_group[perm[i]]; }
    tasks.swap(st);
    task_group.swap(sg);
  }
  S.val.resize(n_vals);
  DevBuf oods_partials, oods_vals, oods_weights;
  const uint32_t* dvals_p = nullptr;
  const OodsTask* d_otasks = nullptr;
  const uint32_t *d_otg = nullptr, *d_ogl = nullptr, *d_obs = nullptr;
  uint32_t n_obatches = 0;
  double obytes = 0;
  if (n_vals) {
    uint32_t max_chunks = oods_max_chunks(max_coeff_log);
    oods_partials = DevBuf(&ctx, (size_t)n_vals * max_chunks * 4);
    oods_vals = DevBuf(&ctx, n_vals * 4);
    oods_weights = DevBuf(&ctx, group_key.size() * ((size_t)4 << OODS_TABLE_LOG));
    dvals_p = oods_vals.p;
    for (auto& tk : tasks) obytes += 4.0 * (double)(1u << tk.log_size);
    d_otasks = ctx.table_vec(tasks);
    d_otg = ctx.table_vec(task_group);
    d_ogl = ctx.table_vec(group_log);
    const std::vector<uint32_t> bstart = oods_batch_starts(task_group);
    d_obs = ctx.table_vec(bstart);
    n_obatches = (uint32_t)bstart.size() - 1;
  }
  HMARK("oods_prep");
  // ---- quotient plan: everything that does not depend on the sampled values, built on the host
  // while the OODS kernels run ----
  // Quotient samples per column (core/pcs/quotients.rs:80-105,306-366): a column sampled at two
  // points first gets a periodicity sample at its second point shifted by the generator of its
  // evaluation domain, then its samples; consecutive powers of the random coefficient in order.
  // Point ids past the sample points' are the periodicity points, listed in `periodic`.
  std::vector<std::array<uint32_t, 3>> periodic;  // (point id, eval log, new point id)
  std::vector<uint32_t> qstart(NC + 1, 0), qpid, qvi;  // per sample: point, index of its value in S.val
  uint32_t n_qpoints = (uint32_t)pt_off.size();
  {
    for (size_t t = 0; t < trees.size(); ++t)
      for (size_t c = 0; c < trees[t].evals.size(); ++c) {
        const uint32_t g = S.col_base[t] + (uint32_t)c, n = S.count(t, c), f0 = S.first(t, c);
        const uint32_t* pids = S.pid.data() + f0;
        qstart[g] = (uint32_t)qpid.size();
        if (n == 2) {
          const uint32_t el = trees[t].eval_logs[c];
          uint32_t np = UINT32_MAX;
          for (auto& e : periodic)
            if (e[0] == pids[1] && e[1] == el) np = e[2];
          if (np == UINT32_MAX) {
            np = n_qpoints++;
            periodic.push_back({pids[1], el, np});
          }
          qpid.push_back(np);
          qvi.push_back(f0 + 1);
        }
        for (uint32_t j = 0; j < n; ++j) {
          qpid.push_back(pids[j]);
          qvi.push_back(f0 + j);
        }
      }
    qstart[NC] = (uint32_t)qpid.size();
  }
  // ColumnSampleBatch per (size group, point); line coefficients (core/constraints.rs
  // complex_conjugate_line_coeffs)
  //   a = r (conj(v) - v), c = r (conj(p.y) - p.y), b = r (v (conj(p.y) - p.y) - (conj(v) - v) p.y)
  // so that over a batch A = sum r (conj(v) - v) and B = cy sum r v - p.y A, with cy = conj(p.y) - p.y.
  // The coefficients themselves need the sampled values and the random coefficient, so they are
  // computed on the device (launch_quotient_coeffs); the host lays out entries and batches here.
  struct QGroup {
    uint32_t ls;
    size_t n_cols;
    bool aligned;
    uint32_t ent_off, batch_off, n_batches;  // into the proof's entry and batch arrays
    const uint32_t* d_starts;
    uint32_t* const* d_outs;
  };
  struct Acc { cpqm31 point; uint32_t pid, log; const uint32_t* c[4]; uint32_t batch; };
  std::vector<Acc> acc_list;
  std::vector<DevBuf> partial_bufs;
  std::vector<QGroup> qgroups;
  std::vector<QCoefEntry> qents;
  std::vector<QCoefBatch> qbatches;
  {
    struct Item { size_t t, c; uint32_t ls; };
    std::vector<Item> items;
    items.reserve(NC);
    for (size_t t = 0; t < trees.size(); ++t)
      for (size_t c = 0; c < trees[t].evals.size(); ++c) items.push_back({t, c, trees[t].eval_logs[c]});
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.ls < b.ls; });
    std::vector<int> batch_of(n_qpoints, -1);
    std::vector<uint32_t> bpid;
    std::vector<std::vector<std::pair<uint32_t, const uint32_t*>>> per;
    for (size_t i = 0; i < items.size();) {
      size_t j = i;
      while (j < items.size() && items[j].ls == items[i].ls) ++j;
      QGroup G;
      G.ls = items[i].ls;
      G.n_cols = j - i;
      G.aligned = true;
      G.ent_off = (uint32_t)qents.size();
      G.batch_off = (uint32_t)qbatches.size();
      // a batch per point, in order of first use; its entries in column order
      bpid.clear();
      per.clear();
      for (size_t k = i; k < j; ++k) {
        const uint32_t* col = trees[items[k].t].evals[items[k].c];
        G.aligned &= ((uintptr_t)col & 15) == 0;
        const uint32_t g = S.col_base[items[k].t] + (uint32_t)items[k].c;
        for (uint32_t q = qstart[g]; q < qstart[g + 1]; ++q) {
          const uint32_t pid = qpid[q];
          if (batch_of[pid] < 0) {
            batch_of[pid] = (int)bpid.size();
            bpid.push_back(pid);
            per.emplace_back();
          }
          per[batch_of[pid]].push_back({q, col});
        }
      }
      std::vector<uint32_t> starts;
      std::vector<uint32_t*> outs;
      for (size_t bi = 0; bi < bpid.size(); ++bi) {
        const uint32_t pid = bpid[bi];
        batch_of[pid] = -1;
        QCoefBatch qb{};
        qb.first = (uint32_t)qents.size();
        qb.count = (uint32_t)per[bi].size();
        starts.push_back(qb.first - G.ent_off);
        for (auto& [q, col] : per[bi]) qents.push_back({col, q, qvi[q], (uint32_t)qbatches.size(), 0});
        Acc a;
        a.pid = pid;
        a.log = G.ls;
        a.batch = (uint32_t)qbatches.size();
        for (int k = 0; k < 4; ++k) {
          partial_bufs.emplace_back(&ctx, (size_t)1 << G.ls);
          outs.push_back(partial_bufs.back().p);
          a.c[k] = partial_bufs.back().p;
        }
        acc_list.push_back(a);
        qbatches.push_back(qb);
      }
      G.n_batches = (uint32_t)bpid.size();
      starts.push_back((uint32_t)qents.size() - G.ent_off);
      G.d_starts = ctx.table_vec(starts);
      G.d_outs = (uint32_t* const*)ctx.table_vec(outs);
      qgroups.push_back(G);
      i = j;
    }
  }
  const QCoefEntry* d_qents = ctx.table_vec(qents);
  DevBuf num_entries(&ctx, std::max<size_t>(qents.size(), 1) * (sizeof(NumEntry) / 4));
  DevBuf batch_b(&ctx, std::max<size_t>(qbatches.size(), 1) * 4);
  DevBuf qterms(&ctx, std::max<size_t>(qents.size(), 1) * 8);
  HMARK("x_plan");

  // ---- the composition root, the OODS point (prover/mod.rs:85-95) and the OODS launch ----
  trees.back().root = read_hash(ctx, trees.back().layers[0].p);
  ch.mix_root(trees.back().root);
  ctx.reset_arena();
  HMARK("comp_commit");
  stage_ms[2] = tm.lap();
  {
    qm31 t = ch.draw_secure_felt();
    qm31 t2 = qm_sqr(t);
    qm31 d = qm_inv(qm_add(t2, qm_one()));
    oods = {qm_mul(qm_sub(qm_one(), t2), d), qm_mul(qm_add(t, t), d)};
  }
  for (long o : pt_off) S.pts.push_back(cpq_add(oods, cpq_mul_signed(step, o)));
  if (n_vals) {
    std::vector<qm31> group_factors(32 * group_key.size(), qm_zero());
    for (size_t g = 0; g < group_key.size(); ++g) {
      const cpqm31 q = cpq_repeated_double(S.pts[group_key[g].pid], group_key[g].shift);
      qm31* f = group_factors.data() + 32 * g;
      f[0] = q.y;
      qm31 x = q.x;
      for (uint32_t b = 1; b < group_key[g].k; ++b) { f[b] = x; x = qm_double_x(x); }
    }
    const qm31* df = ctx.upload_vec(group_factors);
    KScope ks(ctx, K_OODS, obytes);
    launch_oods(d_otasks, d_otg, (uint32_t)n_vals, d_obs, n_obatches, max_coeff_log, df, d_ogl,
                (uint32_t)group_key.size(), (qm31*)oods_weights.p, (qm31*)oods_partials.p, (qm31*)oods_vals.p,
                ctx.stream);
  }
  HMARK("oods_launch");
  // sample point values: the mask points, then the periodicity points
  std::vector<cpqm31> qsample_pts = S.pts;
  qsample_pts.resize(n_qpoints);
  {
    const cpm31 gen = index_to_point(subgroup_gen(Lc));
    for (auto& e : periodic) {
      cpm31 pg = gen;
      for (uint32_t i = 0; i < e[1]; ++i) pg = cp_double(pg);
      qsample_pts[e[2]] = cpq_add(qsample_pts[e[0]], cpq_from(pg));
    }
  }
  for (auto& a : acc_list) {
    a.point = qsample_pts[a.pid];
    QCoefBatch& qb = qbatches[a.batch];
    qb.py = a.point.y;
    qb.cy = qm_sub(qm_conj(qb.py), qb.py);
  }
  // quotient points in the order of their coordinates, each with its batches as parts
  std::stable_sort(acc_list.begin(), acc_list.end(), [](const Acc& a, const Acc& b) { return point_less(a.point, b.point); });
  std::vector<QuotPoint> qpts;
  std::vector<QuotPart> qparts;
  for (size_t i = 0; i < acc_list.size();) {
    size_t j = i;
    while (j < acc_list.size() && point_eq(acc_list[j].point, acc_list[i].point)) ++j;
    QuotPoint qp{};
    const cpqm31& p = acc_list[i].point;
    qp.prx = qm_lo(p.x); qp.pix = qm_hi(p.x); qp.pry = qm_lo(p.y); qp.piy = qm_hi(p.y);
    qp.first_linear = qm_zero();  // the parts' a, written on the device, make up the coefficient
    qp.part_begin = (uint32_t)qparts.size();
    for (size_t k = i; k < j; ++k) {
      QuotPart pa{};
      for (int c = 0; c < 4; ++c) pa.c[c] = acc_list[k].c[c];
      pa.log_size = acc_list[k].log;
      qbatches[acc_list[k].batch].part = (uint32_t)qparts.size();
      qparts.push_back(pa);
    }
    qp.part_end = (uint32_t)qparts.size();
    qpts.push_back(qp);
    i = j;
  }
  // one transfer for the three point-dependent tables
  const QCoefBatch* d_qbatches = ctx.stage_vec(qbatches);
  QuotPart* d_qparts = (QuotPart*)ctx.stage_vec(qparts);
  const QuotPoint* d_qpts = ctx.stage_vec(qpts);
  ctx.commit_staged();
  HMARK("x_qplan");

  // ---- FRI layout, also while the OODS kernels run: the layers' buffers, Merkle heaps and kernel
  // descriptors. The launches are recorded and queued after the quotients (the channel state they
  // start from is known only then). ----
  // The channel runs on the device during the commit phase (one record of root + folding
  // coefficient per layer), so that no layer waits for the host; the host replays the
  // transcript from the recorded roots afterwards and checks every coefficient.
  FriCol first;
  first.log = Lc;
  for (int k = 0; k < 4; ++k) first.c.emplace_back(&ctx, (size_t)1 << Lc);
  const uint32_t* dxy = ctx.domain_xy(Lc);
  ctx.ensure_twiddles(Lc);
  uint32_t last_log = cfg.log_last_layer + cfg.log_blowup;
  if (Lc - 1 < last_log) throw HipError(3, "FRI: last layer larger than the first folded layer");
  uint32_t n_layers = 1;
  for (uint32_t l = Lc - 1; l > last_log;) {
    uint32_t fs = l > last_log + cfg.fold_step ? cfg.fold_step : l - last_log;
    ++n_layers;
    l -= fs;
  }
  const size_t last_n = (size_t)1 << last_log;
  // device block: channel state (9 words; words 14 and 15 hold the proof-of-work result), the
  // transcript records (12 words per layer) and the last layer's values, which the last fold
  // writes in place, so that one read-back takes both
  DevBuf chan(&ctx, 16 + 12 * (size_t)n_layers + 4 * last_n + 8);
  uint32_t* dstate = chan.p;
  uint32_t* dtickets = ctx.fri_ticket_words(n_layers);
  uint32_t* drec = chan.p + 16;
  uint32_t* dlast = drec + 12 * (size_t)n_layers;
  // The end of the phase on the device as well (launch_fri_last_pow, then the grind from its
  // prefix), so that the proof of work needs no round trip of its own: for a last layer folded in
  // place of at most 2^6 values. The prefix goes after the last layer; the host checks the mixed
  // digest and the prefix against its own.
  uint32_t* dprefix = dlast + 4 * last_n;
  const bool dev_pow = cfg.fold_step == 1 && last_log <= 6 && cfg.log_last_layer <= 4 && cfg.pow_bits <= 32;
  const uint32_t* d_last_xi = nullptr;
  if (dev_pow) {
    // the inverse x-coordinates of the last layer's line domain, level by level, then 1 / 2^L
    HCoset dom = HCoset::half_odds(Lc - 1);
    for (uint32_t i = Lc - 1; i > last_log; --i) dom = dom.doubled();
    std::vector<uint32_t> xi;
    for (HCoset d = dom; d.log_size > 0; d = d.doubled())
      for (size_t i = 0; i < ((size_t)1 << d.log_size) / 2; ++i) xi.push_back(m31_inv(d.at(i).x));
    xi.push_back(m31_inv((uint32_t)last_n));
    d_last_xi = ctx.table_vec(xi);
  }
  // the channel state at the start of the phase goes to the first layer launch as an argument when
  // that launch is a fused layer (fri_st0_by_value), else it is uploaded to dstate
  std::array<uint32_t, 9> fri_st0{};
  bool fri_st0_by_value = false;
  auto mix_and_draw = [&ctx, dstate, drec](FriCol& f, uint32_t layer) {
    f.tree.evals = {f.c[0].p, f.c[1].p, f.c[2].p, f.c[3].p};
    f.tree.eval_logs.assign(4, f.log);
    merkle_commit(ctx, f.tree, false);
    KScope ks(ctx, K_MISC, 0.0);
    launch_channel_mix_root_draw(dstate, f.tree.layers[0].p, drec + 12 * layer, ctx.stream);
    return (const qm31*)(drec + 12 * layer + 8);
  };
  std::vector<FriCol> inner;
  inner.reserve(n_layers);  // FriCol addresses stay valid while layers are appended
  FriCol cur;
  uint32_t layer = 1;
  std::vector<std::function<void()>> fri_launches;  // in stream order
  std::vector<FriLayerDesc> small;
  DevBuf fri_buf;
  if (cfg.fold_step == 1) {
    // Fused layers (fri.hip): a layer of 2^9 .. 2^18 values is one launch of launch_fri_layer;
    // layers of at most 2^8 values go to one launch_fri_small at the end; larger layers (and the
    // fold into a last layer above 2^8) use the separate fold, Merkle and channel kernels.
    constexpr uint32_t FAST_MAX = 18, SMALL_MAX = 8;
    // Line layers of at most 2^FAST_MAX values share one block: the values of layer l (l_start
    // down to last_log) at 4 (2^(l_start+1) - 2^(l+1)) words, coordinate k at + k 2^l, then the
    // Merkle heaps of the committed ones, then the first layer's heap when it is fused too.
    const uint32_t l_start = std::min(FAST_MAX, Lc - 1);
    const size_t tv_words = 4 * (((size_t)2 << l_start) - ((size_t)1 << last_log));
    std::vector<size_t> heap_off(l_start + 1, 0);
    size_t tn_words = 0;
    for (uint32_t l = l_start; l > last_log; --l) {
      heap_off[l] = tv_words + tn_words;
      tn_words += 8 * (((size_t)2 << l) - 1);
    }
    const bool first_fused = Lc <= FAST_MAX;
    const size_t fh_words = first_fused ? 8 * (((size_t)2 << Lc) - 1) : 0;
    fri_buf = DevBuf(&ctx, tv_words + tn_words + fh_words);
    auto heap_views = [&](FriCol& f, uint32_t* heap) {
      f.tree.evals = {f.c[0].p, f.c[1].p, f.c[2].p, f.c[3].p};
      f.tree.eval_logs.assign(4, f.log);
      f.tree.layers.clear();
      for (uint32_t j = 0; j <= f.log; ++j)
        f.tree.layers.push_back(DevBuf::view(&ctx, heap + 8 * (((size_t)1 << j) - 1), (size_t)8 << j));
    };
    auto line_col = [&](uint32_t l) {
      FriCol f;
      f.log = l;
      f.fold_step = 1;
      if (l <= l_start) {
        uint32_t* v = fri_buf.p + 4 * (((size_t)2 << l_start) - ((size_t)2 << l));
        for (int k = 0; k < 4; ++k) f.c.push_back(DevBuf::view(&ctx, v + ((size_t)k << l), (size_t)1 << l));
      } else {
        for (int k = 0; k < 4; ++k) f.c.emplace_back(&ctx, (size_t)1 << l);
      }
      return f;
    };
    // one step per layer: li = 0 is the circle layer (the quotients), li = n_layers the last layer
    FriCol* prev = &first;
    for (uint32_t li = 0; li <= n_layers; ++li) {
      const uint32_t l = Lc - li;
      const bool commit = li < n_layers;
      // the layer's FriCol at its final place (inner or cur), so that recorded launches can refer to it
      FriCol* cf = &first;
      if (li > 0) {
        if (commit) {
          inner.push_back(line_col(l));
          cf = &inner.back();
        } else {
          cur.log = l;
          cur.fold_step = 1;
          for (int k = 0; k < 4; ++k) cur.c.push_back(DevBuf::view(&ctx, dlast + ((size_t)k << l), (size_t)1 << l));
          cf = &cur;
        }
      }
      FriLayerDesc d{};
      d.log = l;
      d.mode = li == 0 ? FRI_LEAVES : li == 1 ? FRI_FOLD_CIRCLE : FRI_FOLD_LINE;
      d.commit = commit ? 1u : 0u;
      d.src = li == 0 ? first.cc() : prev->cc();
      d.dst = cf->mc();
      d.tw = li == 1 ? ctx.itw.p + ctx.tree_len() - ((size_t)1 << l) : ctx.itw.p + ctx.tree_len() - ((size_t)2 << l);
      d.alpha = li > 0 ? drec + 12 * (li - 1) + 8 : nullptr;
      d.rec = commit ? drec + 12 * li : nullptr;
      if (commit && l <= FAST_MAX) {
        d.nodes = li == 0 ? fri_buf.p + tv_words + tn_words : fri_buf.p + heap_off[l];
        heap_views(*cf, d.nodes);
      }
      if (l <= SMALL_MAX) {
        small.push_back(d);
      } else if (commit && l <= FAST_MAX) {
        if (li == 0) fri_st0_by_value = true;
        fri_launches.push_back([&ctx, d, l, li, dstate, dtickets, &fri_st0] {
          KScope ks(ctx, K_FRI_FOLD, (li == 0 ? 16.0 : 48.0) * (double)(1u << l) + 64.0 * (double)(1u << l));
          FriLayerDesc dd = d;
          if (li == 0) {
            dd.init = 1;
            std::copy(fri_st0.begin(), fri_st0.end(), dd.st0);
          }
          launch_fri_layer(dd, dstate, dtickets + li, ctx.stream);
        });
      } else {
        if (li > 0) {
          const CCols4 src = li == 1 ? first.cc() : prev->cc();
          const Cols4 dst = cf->mc();
          fri_launches.push_back([&ctx, d, l, li, src, dst, Lc] {
            KScope ks(ctx, K_FRI_FOLD, 48.0 * (double)(1u << l));
            const qm31* a = (const qm31*)d.alpha;
            if (li == 1) launch_fold_circle_into_line_dev(Lc, src, dst, a, true, ctx.itw.p, ctx.tree_len(), ctx.stream);
            else launch_fold_line_dev(l + 1, src, dst, a, 0, ctx.itw.p, ctx.tree_len(), ctx.stream);
          });
        }
        if (commit) fri_launches.push_back([cf, li, mix_and_draw] { mix_and_draw(*cf, li); });
      }
      if (li > 0 && commit) prev = cf;
    }
    if (!small.empty()) {
      double tb = 0;
      for (auto& d : small) tb += 64.0 * (double)(1u << d.log);
      const FriLayerDesc* d_small = ctx.table_vec(small);
      const uint32_t n_small = (uint32_t)small.size();
      fri_launches.push_back([&ctx, tb, d_small, n_small, dstate] {
        KScope ks(ctx, K_FRI_TAIL, tb);
        launch_fri_small(d_small, n_small, dstate, ctx.stream);
      });
    }
    layer = n_layers;
  }
  HMARK("x_friprep");

  // the uploads above stay in the staging arena until the quotient kernels have run
  if (n_vals) ctx.d2h(S.val.data(), dvals_p, n_vals * 16);
  HMARK("oods_gpu");
  ch.mix_felts(S.val);
  HMARK("oods_post");
  stage_ms[3] = tm.lap();

  // ---------------- FRI quotients ----------------
  const qm31 qrc = ch.draw_secure_felt();
  NumEntry* d_num = (NumEntry*)num_entries.p;
  qm31* d_bb = (qm31*)batch_b.p;
  {
    KScope ks(ctx, K_NUMERATORS, 32.0 * (double)qents.size());
    launch_quotient_coeffs(qrc, (const qm31*)dvals_p, d_qents, (uint32_t)qents.size(), d_qbatches,
                           (uint32_t)qbatches.size(), d_num, d_bb, d_qparts, (qm31*)qterms.p, ctx.stream);
  }
  HMARK("q_samples");
  for (auto& G : qgroups) {
    if (!G.n_batches) continue;
    KScope ks(ctx, K_NUMERATORS, (4.0 * G.n_cols + 16.0 * G.n_batches) * (double)(1u << G.ls));
    launch_accumulate_numerators(G.ls, d_num + G.ent_off, G.d_starts, d_bb + G.batch_off, G.n_batches, G.d_outs,
                                 G.aligned, ctx.stream);
  }
  HMARK("q_numer");
  double qbytes = 24.0 * (double)(1u << Lc);
  for (auto& pa : qparts) qbytes += 16.0 * (double)(1u << pa.log_size);
  {
    KScope ksq(ctx, K_QUOTIENTS, qbytes);
    launch_quotients(Lc, dxy, d_qpts, (uint32_t)qpts.size(), d_qparts, first.mc(), ctx.stream);
  }
  HMARK("q_launch");
  partial_bufs.clear();
  stage_ms[4] = tm.lap();

  // ---------------- FRI commit ----------------
  {
    for (int i = 0; i < 8; ++i)
      fri_st0[i] = (uint32_t)ch.digest[4 * i] | (uint32_t)ch.digest[4 * i + 1] << 8 |
                   (uint32_t)ch.digest[4 * i + 2] << 16 | (uint32_t)ch.digest[4 * i + 3] << 24;
    fri_st0[8] = ch.n_draws;
    if (!fri_st0_by_value) {
      std::vector<uint32_t> st(16, 0);
      std::copy(fri_st0.begin(), fri_st0.end(), st.begin());
      st[14] = st[15] = ~0u;  // the proof-of-work result slot (u64), empty
      launch_copy_words(dstate, ctx.upload_vec(st), st.size(), ctx.stream);
    }
  }
  for (auto& f : fri_launches) f();
  if (dev_pow) {
    KScope ks(ctx, K_GRIND, 0.0);
    launch_fri_last_pow(dlast, last_log, cfg.log_last_layer, d_last_xi, dstate, cfg.pow_bits, dprefix, ctx.stream);
    launch_grind_dev(dprefix, cfg.pow_bits, 0, grind_first_count(cfg.pow_bits), (unsigned long long*)(dstate + 14),
                     ctx.stream);
  }
  if (cfg.fold_step != 1) {
    const qm31* alpha0 = mix_and_draw(first, 0);
    cur.log = Lc - 1;
    for (int k = 0; k < 4; ++k) cur.c.emplace_back(&ctx, (size_t)1 << cur.log);
    {
      // the first line layer starts at zero (FriProver::commit): the fold writes it without reading
      KScope ks(ctx, K_FRI_FOLD, 16.0 * (double)(1u << Lc) + 16.0 * (double)(1u << (Lc - 1)));
      launch_fold_circle_into_line_dev(Lc, first.cc(), cur.mc(), alpha0, true, ctx.itw.p, ctx.tree_len(), ctx.stream);
    }
    auto fold_layer = [&](FriCol& src, uint32_t fs, const qm31* a) {
      FriCol* s = &src;
      FriCol tmp;
      for (uint32_t st = 0; st < fs; ++st) {
        FriCol nx;
        nx.log = s->log - 1;
        for (int k = 0; k < 4; ++k) nx.c.emplace_back(&ctx, (size_t)1 << nx.log);
        KScope ks(ctx, K_FRI_FOLD, 24.0 * (double)(1u << s->log));
        launch_fold_line_dev(s->log, s->cc(), nx.mc(), a, st, ctx.itw.p, ctx.tree_len(), ctx.stream);
        tmp = std::move(nx);
        s = &tmp;
      }
      return tmp;
    };
    while (cur.log != last_log) {
      uint32_t fs = cur.log > last_log + cfg.fold_step ? cfg.fold_step : cur.log - last_log;
      cur.fold_step = fs;
      const qm31* a = mix_and_draw(cur, layer++);
      FriCol nx = fold_layer(cur, fs, a);
      inner.push_back(std::move(cur));
      cur = std::move(nx);
    }
  }
  if (layer != n_layers) throw HipError(4, "FRI layer count mismatch");
  HMARK("fri_launch");
  // Decommitment read layout: it depends only on the proof's shape, so the item table is a cached
  // table from the second proof of a shape on.
  DecLayout dl;
  dl.n_q = cfg.n_queries;
  for (auto& t : trees) {
    const uint32_t mx = t.max_log();
    if (Lc < mx) dec_add_object(ctx, dl, t, 1, 0, mx - Lc + 1);
    else dec_add_object(ctx, dl, t, 1, 1, Lc - mx + 1);
  }
  dec_add_object(ctx, dl, first.tree, 1, 2, 0);
  {
    uint32_t a = 1;
    for (auto& l : inner) {
      dec_add_object(ctx, dl, l.tree, l.fold_step, 2, a);
      a += l.fold_step;
    }
  }
  const DecItem* ditems = ctx.table_vec(dl.items);
  DevBuf dsup(&ctx, std::max<size_t>(dl.words, 1));
  HMARK("x_declayout");
  // Sanity check (prover/mod.rs:127-144): the composition value at the OODS point against the
  // constraints evaluated from the sampled values. It only sets the status, so it runs on the
  // host while the FRI kernels are in flight.
  int status;
  {
    const size_t lt = S.n_trees() - 1;
    qm31 lft = from_partial_evals(S.vals(lt, 0)[0], S.vals(lt, 1)[0], S.vals(lt, 2)[0], S.vals(lt, 3)[0]);
    qm31 rgt = from_partial_evals(S.vals(lt, 4)[0], S.vals(lt, 5)[0], S.vals(lt, 6)[0], S.vals(lt, 7)[0]);
    qm31 committed = qm_add(lft, qm_mul(cpq_repeated_double(oods, max_deg - 1).x, rgt));
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
    status = qm_eq(committed, acc) ? 0 : 2;
  }
  HMARK("sanity");
  std::vector<uint32_t> fri_hb;  // the FRI block as read back
  {
    // one read-back: the transcript records and the last layer's evaluations
    for (int k = 0; k < 4; ++k)
      if (cur.c[k].p != dlast + k * last_n) launch_copy_words(dlast + k * last_n, cur.c[k].p, last_n, ctx.stream);
    // the whole block: channel state and proof-of-work slot, records, last layer, prefix
    fri_hb.assign(16 + 12 * (size_t)n_layers + 4 * last_n + 8, 0);
    ctx.d2h(fri_hb.data(), dstate, fri_hb.size() * 4);
    const uint32_t* hb = fri_hb.data() + 16;
    ctx.reset_arena();
    HMARK("fri_wait");
    auto replay = [&](DevTree& t, uint32_t l) {
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 4; ++j) t.root[4 * i + j] = (uint8_t)(hb[12 * l + i] >> (8 * j));
      ch.mix_root(t.root);
      qm31 a = ch.draw_secure_felt();
      const uint32_t* r = hb + 12 * l + 8;
      if (!qm_eq(a, qm31{r[0], r[1], r[2], r[3]})) throw HipError(4, "device channel diverged from the transcript");
    };
    replay(first.tree, 0);
    for (uint32_t l = 1; l < n_layers; ++l) replay(inner[l - 1].tree, l);
    const size_t n = last_n;
    const uint32_t* w = hb + 12 * (size_t)n_layers;
    std::vector<qm31> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = {w[i], w[n + i], w[2 * n + i], w[3 * n + i]};
    HCoset dom = HCoset::half_odds(Lc - 1);
    for (uint32_t i = Lc - 1; i > cur.log; --i) dom = dom.doubled();
    auto coeffs = line_interpolate(v, dom);
    std::vector<qm31> ordered(coeffs);
    uint32_t lg = cur.log;
    for (size_t i = 0; i < n; ++i) { size_t j = bitrev((uint32_t)i, lg); if (j > i) std::swap(ordered[i], ordered[j]); }
    size_t bound = (size_t)1 << cfg.log_last_layer;
    for (size_t i = bound; i < n; ++i)
      if (!qm_eq(ordered[i], qm_zero())) throw HipError(4, "invalid degree");
    ordered.resize(bound);
    uint32_t bl = cfg.log_last_layer;
    for (size_t i = 0; i < bound; ++i) { size_t j = bitrev((uint32_t)i, bl); if (j > i) std::swap(ordered[i], ordered[j]); }
    out.last_poly = ordered;
    ch.mix_felts(out.last_poly);
  }
  HMARK("fri_host");
  stage_ms[5] = tm.lap();

  // ---------------- proof of work ----------------
  if (dev_pow) {
    const uint32_t* dw = fri_hb.data();
    for (int i = 0; i < 8; ++i) {
      const uint32_t hw = (uint32_t)ch.digest[4 * i] | (uint32_t)ch.digest[4 * i + 1] << 8 |
                          (uint32_t)ch.digest[4 * i + 2] << 16 | (uint32_t)ch.digest[4 * i + 3] << 24;
      if (dw[i] != hw) throw HipError(4, "device channel diverged from the transcript");
    }
    const Hash prefix = ch.pow_prefix(cfg.pow_bits);
    const uint32_t* dp = dw + 16 + 12 * (size_t)n_layers + 4 * last_n;
    for (int i = 0; i < 32; ++i)
      if (prefix[i] != (uint8_t)(dp[i / 4] >> (8 * (i % 4)))) throw HipError(4, "device proof-of-work prefix differs");
    out.pow = (uint64_t)dw[14] | (uint64_t)dw[15] << 32;
    // a nonce beyond the first launch's range (probability about e^-32): the search continues
    if (out.pow == ~0ull) out.pow = device_grind(ctx, prefix, cfg.pow_bits);
  } else {
    out.pow = device_grind(ctx, ch.pow_prefix(cfg.pow_bits), cfg.pow_bits, (unsigned long long*)(dstate + 14));
  }
  ch.mix_u64(out.pow);
  HMARK("pow");
  stage_ms[6] = tm.lap();

  // ---------------- decommit ----------------
  std::vector<size_t> unsorted;
  {
    uint32_t mask = (uint32_t)(((uint64_t)1 << Lc) - 1);
    while (unsorted.size() < cfg.n_queries) {
      auto w = ch.draw_u32s();
      for (auto x : w) { if (unsorted.size() < cfg.n_queries) unsorted.push_back(x & mask); }
    }
  }
  const auto qpos = sorted_unique(unsorted);
  const std::vector<uint32_t> qp32(qpos.begin(), qpos.end());
  const uint32_t nq = (uint32_t)qp32.size();
  launch_decommit_gather(ditems, (uint32_t)dl.items.size(), qp32.data(), nq, dsup.p, ctx.stream);
  const uint32_t* sup = (const uint32_t*)ctx.readback_buf(std::max<size_t>(dl.words, 1) * 4);
  if (dl.words) HIP_CHECK(hipMemcpyAsync((void*)sup, dsup.p, dl.words * 4, hipMemcpyDeviceToHost, ctx.stream));
  HMARK("x_queries");
  // the reference walks, on the host while the read is in flight
  auto positions = [&](const DecObject& ob) {
    std::vector<uint32_t> P(nq);
    for (uint32_t i = 0; i < nq; ++i) P[i] = dec_pos_host(ob.pos_mode, ob.pos_arg, qp32[i]);
    return P;
  };
  std::vector<DecPlan> tplans(trees.size());
  for (size_t t = 0; t < trees.size(); ++t) {
    const DecObject& ob = dl.objs[t];
    if (ob.empty) continue;
    const auto P = positions(ob);
    // prepare_preprocessed_query_positions, then each column size's queried nodes
    std::vector<std::pair<uint32_t, std::vector<size_t>>> qpl;
    for (auto ls : std::set<uint32_t>(trees[t].eval_logs.begin(), trees[t].eval_logs.end())) {
      const uint32_t shift = ob.top - ls;
      std::vector<size_t> v;
      for (uint32_t p : P) v.push_back(((size_t)(p >> (shift + 1)) << 1) + (p & 1));
      qpl.push_back({ls, sorted_unique(v)});
    }
    dec_plan_object(ob, P, qpl, tplans[t]);
  }
  std::vector<FriDecPlan> fplans;
  {
    std::vector<size_t> lq = qpos;
    for (size_t j = 0; j <= inner.size(); ++j) {
      const DecObject& ob = dl.objs[trees.size() + j];
      fplans.push_back(plan_fri_layer(ob, lq, positions(ob)));
      lq = fold_queries(lq, ob.w);
    }
  }
  HMARK("dec_plan");
  ctx.sync_main();
  HMARK("dec_gather");
  out.first = finish_fri_layer(fplans[0], sup, first.tree.root);
  for (size_t j = 0; j < inner.size(); ++j) out.inner.push_back(finish_fri_layer(fplans[j + 1], sup, inner[j].tree.root));
  for (size_t t = 0; t < trees.size(); ++t) {
    MerkleDec d;
    std::vector<uint32_t> queried;
    dec_finish(tplans[t], sup, d, &queried);
    out.queried.push_back(std::move(queried));
    out.decs.push_back(std::move(d));
    out.commitments.push_back(trees[t].root);
  }
  ctx.reset_arena();
  HMARK("dec_finish");
  stage_ms[7] = tm.lap();

  return status;
}

}  // namespace

ProveResult prove_wide_fib(Ctx& ctx, const std::vector<const uint32_t*>& trace, uint32_t log_n, uint32_t n_cols,
                           const PcsCfg& cfg) {
  ProofArenaScope arena(ctx);
  ProveResult r;
  Timer tm;
  Channel ch;
  std::vector<DevTree> trees(2);
  ctx.ensure_twiddles(log_n + 1 + cfg.log_blowup);
  commit_tree(ctx, trees[0], cfg, ch);  // empty preprocessed tree
  interpolate_columns(ctx, trace, std::vector<uint32_t>(trace.size(), log_n), trees[1]);
  std::vector<Component> comps(1);
  comps[0].kind = AIR_WIDE_FIB;
  comps[0].log_n = log_n;
  comps[0].n_cols = n_cols;
  CompPrep prep;
  commit_tree(ctx, trees[1], cfg, ch, [&] { prepare_composition(ctx, trees, comps, prep); });
  r.stage_ms[0] = tm.lap();
  ProofParts parts;
  r.status = prove_ex(ctx, ch, trees, comps, cfg, parts, r.stage_ms, &prep);
  r.proof = encode(cfg, parts);
  r.digest = ch.digest;
  return r;
}

ProveResult prove_wide_fib_multi(Ctx& ctx, const std::vector<const uint32_t*>& trace, const std::vector<uint32_t>& logs,
                                 const std::vector<uint32_t>& ncols, const PcsCfg& cfg) {
  ProofArenaScope arena(ctx);
  ProveResult r;
  Timer tm;
  Channel ch;
  std::vector<DevTree> trees(2);
  uint32_t max_n = 0;
  std::vector<uint32_t> col_logs;
  std::vector<Component> comps;
  size_t start = 0;
  for (size_t i = 0; i < logs.size(); ++i) {
    max_n = std::max(max_n, logs[i]);
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
  ctx.ensure_twiddles(max_n + 1 + cfg.log_blowup);
  commit_tree(ctx, trees[0], cfg, ch);  // empty preprocessed tree
  interpolate_columns(ctx, trace, col_logs, trees[1]);
  CompPrep prep;
  commit_tree(ctx, trees[1], cfg, ch, [&] { prepare_composition(ctx, trees, comps, prep); });
  r.stage_ms[0] = tm.lap();
  ProofParts parts;
  r.status = prove_ex(ctx, ch, trees, comps, cfg, parts, r.stage_ms, &prep);
  r.proof = encode(cfg, parts);
  r.digest = ch.digest;
  return r;
}

ProveResult prove_poseidon(Ctx& ctx, const std::vector<const uint32_t*>& trace, uint32_t log_n, uint32_t ipr,
                           const PcsCfg& cfg) {
  ProofArenaScope arena(ctx);
  ProveResult r;
  Timer tm;
  Channel ch;
  HMARK("begin");
  std::vector<DevTree> trees(2);
  ctx.ensure_twiddles(log_n + 2 + cfg.log_blowup);
  commit_tree(ctx, trees[0], cfg, ch);
  HMARK("pre_tree");
  interpolate_columns(ctx, trace, std::vector<uint32_t>(trace.size(), log_n), trees[1]);
  HMARK("tr_interp");
  // The extension of the trace to the composition domain needs only its coefficients. It runs on
  // the side stream once the commitment's transforms are queued (two transforms side by side
  // would split the CUs, whose LDS holds one transform block each), beside the trace's Merkle
  // tree and the LogUp interaction phase, whose kernels take issue priority.
  evaluate_polynomials(ctx, trees[1], cfg.log_blowup);
  {
    const uint32_t E = log_n + 2;
    evaluate_into(ctx, trees[1].coeffs, trees[1].coeff_logs, E, trees[1].ext, trees[1].owned, true);
    trees[1].ext_log = E;
  }
  merkle_commit(ctx, trees[1]);
  ch.mix_root(trees[1].root);
  ctx.reset_arena();
  HMARK("tr_commit");
  r.stage_ms[0] = tm.lap();
  Component c;
  c.kind = AIR_POSEIDON;
  c.log_n = log_n;
  c.ipr = ipr;
  std::vector<Component> comps;  // {c}, once the LogUp draws are in it; the composition refers to it
  CompPrep prep;
  auto le = ch.draw_secure_felts(2);
  c.z = le[0];
  c.alpha = le[1];
  { qm31 p = qm_one(); for (int i = 0; i < 16; ++i) { c.alpha_pow[i] = p; p = qm_mul(p, c.alpha); } }
  // LogUp interaction trace on the device
  DevTree inter_src;
  std::vector<DevBuf> icols;
  for (uint32_t i = 0; i < 4 * ipr; ++i) icols.emplace_back(&ctx, (size_t)1 << log_n);
  {
    PoseidonParams pp{};
    pp.z = c.z;
    for (int i = 0; i < 16; ++i) pp.alpha_pow[i] = c.alpha_pow[i];
    pp.ipr = ipr;
    std::vector<uint32_t*> ip;
    for (auto& b : icols) ip.push_back(b.p);
    // read back once, after the interaction commit: the 4 column sums (words 0..8, u64) and the
    // interaction root (words 8..16); the prefix sum takes its shift from the sums on the device
    DevBuf sums(&ctx, 16, true);
    const uint32_t* const* d_tr = (const uint32_t* const*)ctx.table_vec(trace);
    uint32_t* const* d_ip = (uint32_t* const*)ctx.table_vec(ip);
    const PoseidonParams* d_pp = (const PoseidonParams*)ctx.upload(&pp, sizeof(pp));
    {
      KScope ks(ctx, K_LOGUP, (double)(1u << log_n) * (128.0 * ipr + 16.0 * ipr));
      launch_poseidon_logup_fracs(log_n, ipr, d_tr, d_ip, d_pp, (unsigned long long*)sums.p, ctx.stream);
    }
    HMARK("logup");
    std::vector<uint32_t*> last4(ip.end() - 4, ip.end());
    uint32_t nb = (1u << log_n) >= 1024 ? (1u << log_n) / 1024 : 1;
    DevBuf scratch(&ctx, 4 * nb);
    {
      KScope ks2(ctx, K_LOGUP, 32.0 * (double)(1u << log_n));
      launch_coset_prefix_sum(log_n, (uint32_t* const*)ctx.table_vec(last4), qm_zero(),
                              (const unsigned long long*)sums.p, scratch.p, ctx.stream);
    }
    std::vector<const uint32_t*> ev;
    for (auto& b : icols) ev.push_back(b.p);
    trees.emplace_back();
    interpolate_columns(ctx, ev, std::vector<uint32_t>(ev.size(), log_n), trees[2]);
    evaluate_polynomials(ctx, trees[2], cfg.log_blowup);
    // the interaction columns' extension follows the trace's on the side stream
    const uint32_t E = log_n + 2;
    evaluate_into(ctx, trees[2].coeffs, trees[2].coeff_logs, E, trees[2].ext, trees[2].owned, true);
    trees[2].ext_log = E;
    merkle_commit(ctx, trees[2], false);
    launch_copy_words(sums.p + 8, trees[2].layers[0].p, 8, ctx.stream);
    comps.assign(1, c);
    prepare_composition(ctx, trees, comps, prep);
    uint32_t w[16];
    ctx.d2h(w, sums.p, sizeof(w));
    unsigned long long s[4];
    std::memcpy(s, w, sizeof(s));
    c.claimed_sum = {m31_reduce64(s[0] % P), m31_reduce64(s[1] % P), m31_reduce64(s[2] % P), m31_reduce64(s[3] % P)};
    comps[0].claimed_sum = c.claimed_sum;
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 4; ++j) trees[2].root[4 * i + j] = (uint8_t)(w[8 + i] >> (8 * j));
    ch.mix_root(trees[2].root);
    ctx.reset_arena();
  }
  icols.clear();
  HMARK("inter_commit");
  r.stage_ms[1] = tm.lap();
  r.claimed_sum = c.claimed_sum;
  ProofParts parts;
  r.status = prove_ex(ctx, ch, trees, comps, cfg, parts, r.stage_ms, &prep);
  r.proof = encode(cfg, parts);
  r.digest = ch.digest;
  HMARK("encode");
  g_marks.flush();
  return r;
}

}  // namespace sh
