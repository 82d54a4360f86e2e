// C entry points of libstwo_hip (declared in include/stwo_hip.h).
#include "../../include/stwo_hip.h"

#include <cstdio>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>

#include "compiler.h"
#include "prover.h"
#include "protocol.h"

using namespace sh;

struct stwo_hip_twiddles {
  uint32_t init_index = 0, log_size = 0;  // the root coset, Coset::new(init_index, log_size)
  DevBuf tw, itw, tw2, itw2;
};
struct stwo_hip_ctx {
  Ctx* ctx = nullptr;
  std::string err;
  std::set<stwo_hip_twiddles*> trees;  // freed with the context if the caller has not
};
struct stwo_proof {
  ProveResult r;
};

// File-scope device array behind the no-allocation entry point (see resolve_scratch below).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {

template <class F>
int guarded(stwo_hip_ctx* h, F&& f, bool may_recycle = true, bool capture_ok = false) {
  if (!h || !h->ctx) return 1;
  try {
    // the calling thread may differ from the one that created the context: select its device
    int cur = -1;
    HIP_CHECK(hipGetDevice(&cur));
    if (cur != h->ctx->device) HIP_CHECK(hipSetDevice(h->ctx->device));
    if (h->ctx->capturing && !capture_ok)
      throw HipError(3, "not capturable: this entry point synchronises, allocates or stages host data");
    // between entry points no host code holds a table address, so a full table store is emptied
    // here (the no-allocation entry point skips this: it may free a spilled table)
    if (may_recycle && !h->ctx->capturing) h->ctx->recycle_tables();
    f(*h->ctx);
    return 0;
  } catch (const HipError& e) {
    h->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    h->err = e.what();
    return 1;
  }
}
inline qm31 q(stwo_qm31 v) { return {v.a, v.b, v.c, v.d}; }
inline stwo_qm31 q(qm31 v) { return {v.a, v.b, v.c, v.d}; }
inline cpqm31 cp(stwo_circle_point_qm31 p) { return {q(p.x), q(p.y)}; }
void check_canonical(const stwo_qm31& v) {
  if (v.a >= P || v.b >= P || v.c >= P || v.d >= P) throw HipError(1, "non-canonical QM31 argument");
}
PcsCfg cfg_of(const stwo_pcs_config* c) {
  PcsCfg r;
  if (!c) return r;
  r.pow_bits = c->pow_bits;
  r.log_blowup = c->log_blowup_factor;
  r.log_last_layer = c->log_last_layer_degree_bound;
  r.n_queries = c->n_queries;
  r.fold_step = c->line_fold_step;
  r.lifting_log_size = c->lifting_log_size;
  if (r.log_blowup < 1 || r.log_blowup > 16 || r.log_last_layer > 10 || r.fold_step < 1 || r.n_queries < 1)
    throw HipError(1, "invalid PcsConfig");
  return r;
}
// prove_ex's sanity check (crates/stwo/src/prover/mod.rs:127-144): the reference returns
// Err(ConstraintsNotSatisfied) and no proof. Here the call returns status 6 and *out still holds
// the proof, which the caller inspects or frees.
void reject_unsatisfied(const stwo_proof& p) {
  if (p.r.status != 0)
    throw HipError(6, "ConstraintsNotSatisfied: the composition's OODS value differs from the constraints' "
                      "(prove_ex's sanity check); the proof handle is set and must be freed");
}
// A circle transform of same-size columns on the domain whose half coset starts at half_init:
// domains of log size 1 and 2 by the explicit formulas, larger ones by the FFT kernels over a
// doubled tree whose last 2^(log_n - 1) entries are the domain's line twiddles.
// `low` is the plan's low-stage tile (Ctx::fft_low_log), taken before the caller's timing scope.
void run_fft(Ctx& c, bool inverse, const uint32_t* const* ds, uint32_t* const* dd, uint32_t n_cols, uint32_t log_in,
             uint32_t log_n, const uint32_t* tree2, uint32_t tree_len, uint32_t half_init, uint32_t low) {
  if (log_n <= 2) launch_fft_tiny(inverse, ds, dd, n_cols, log_in, log_n, index_to_point(half_init), c.stream);
  else launch_fft(inverse, ds, dd, n_cols, log_in, log_n, tree2, tree_len, c.stream, false, low);
}
uint32_t plan_fft(Ctx& c, bool inverse, uint32_t n_cols, uint32_t log_in, uint32_t log_n, const uint32_t* tree2,
                  uint32_t tree_len, uint32_t* const* dst) {
  return log_n <= 2 ? 0u : Ctx::fft_plan_for(c.fft_low_log(inverse, n_cols, log_in, log_n, tree2, tree_len), dst, n_cols);
}
uint32_t canonic_half_init(uint32_t log_n) { return HCoset::half_odds(log_n - 1).initial_index; }
void check_tree(const stwo_hip_twiddles* t, uint32_t half_init, uint32_t log_n) {
  // domain.half_coset.is_doubling_of(twiddles.root_coset) (prover/backend/cpu/circle.rs:195)
  if (!t) throw HipError(1, "null twiddle tree");
  if (log_n < 1 || log_n - 1 > t->log_size) throw HipError(1, "domain larger than the twiddle tree's coset");
  const uint32_t d = t->log_size - (log_n - 1);
  const uint32_t want = d >= 31 ? 0u : (uint32_t)(((uint64_t)t->init_index << d) & 0x7fffffffu);
  if ((half_init & 0x7fffffffu) != want) throw HipError(1, "the domain's half coset is not a doubling of the tree's coset");
}

// Scratch of the no-allocation entry point (stwo_hip_batch_eval_at_point_noalloc): one copy per
// device for the process's lifetime. Host code never uses the symbol's name as an address: the
// device address of each device's copy comes from hipGetSymbolAddress, taken once per device
// when a context is created on it, and calls on one device are serialised by its lock.
struct DeviceScratch {
  uint8_t* ptr = nullptr;
  std::mutex mu;
};
std::mutex g_scratch_mu;
std::map<int, std::unique_ptr<DeviceScratch>> g_scratch;  // by device
// With the device current: the device address of its scratch copy (status 2 if the runtime cannot
// resolve it).
int resolve_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto& e = g_scratch[device];
  if (e && e->ptr) return 0;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return 2;
  if (!e) e = std::make_unique<DeviceScratch>();
  e->ptr = (uint8_t*)p;
  return 0;
}
DeviceScratch& scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(device);
  if (it == g_scratch.end() || !it->second || !it->second->ptr) throw HipError(2, "no-allocation scratch not resolved");
  return *it->second;
}

uint32_t fnv1a(const char* s) {
  uint32_t h = 0x811C9DC5u;
  for (; *s; ++s) { h ^= (uint8_t)*s; h *= 0x01000193u; }
  return h;
}

}  // namespace

extern "C" {

int stwo_hip_create(int device, stwo_hip_ctx** out) {
  if (!out) return 1;
  auto* h = new stwo_hip_ctx();
  try {
    h->ctx = new Ctx(device);  // makes the device current
  } catch (const std::exception& e) {
    delete h;
    *out = nullptr;
    return 2;
  }
  if (resolve_scratch(device) != 0) {
    delete h->ctx;
    delete h;
    *out = nullptr;
    return 2;
  }
  *out = h;
  return 0;
}
void stwo_hip_destroy(stwo_hip_ctx* h) {
  if (!h) return;
  for (auto* t : h->trees) delete t;
  h->trees.clear();
  delete h->ctx;
  delete h;
}
int stwo_hip_synchronize(stwo_hip_ctx* h) { return guarded(h, [&](Ctx& c) { c.sync(); c.reset_arena(); }); }
const char* stwo_hip_last_error(stwo_hip_ctx* h) { return h ? h->err.c_str() : "no context"; }

int stwo_hip_alloc(stwo_hip_ctx* h, size_t n, int zeroed, uint32_t** out) {
  return guarded(h, [&](Ctx& c) {
    if (!out) throw HipError(1, "null out");
    *out = (uint32_t*)c.col_alloc(n ? n * 4 : 4);
    if (zeroed) launch_fill_zero(*out, n, c.stream);
  });
}
int stwo_hip_free(stwo_hip_ctx* h, uint32_t* p) {
  return guarded(h, [&](Ctx& c) { if (p) c.col_free(p); });
}
int stwo_hip_alloc_columns(stwo_hip_ctx* h, uint32_t n_cols, size_t n, int zeroed, uint32_t** out) {
  return guarded(h, [&](Ctx& c) {
    if (!out && n_cols) throw HipError(1, "null out");
    for (uint32_t k = 0; k < n_cols; ++k) {
      out[k] = (uint32_t*)c.col_alloc(n ? n * 4 : 4);
      if (zeroed) launch_fill_zero(out[k], n, c.stream);
    }
  });
}
int stwo_hip_free_columns(stwo_hip_ctx* h, uint32_t* const* ptrs, uint32_t n_cols) {
  return guarded(h, [&](Ctx& c) {
    for (uint32_t k = 0; k < n_cols; ++k)
      if (ptrs[k]) c.col_free(ptrs[k]);
  });
}
int stwo_hip_copy_h2d(stwo_hip_ctx* h, uint32_t* dst, const uint32_t* src, size_t n) {
  return guarded(h, [&](Ctx& c) {
    HIP_CHECK(hipMemcpyAsync(dst, src, n * 4, hipMemcpyHostToDevice, c.stream));
    c.sync();
  });
}
int stwo_hip_copy_d2h(stwo_hip_ctx* h, uint32_t* dst, const uint32_t* src, size_t n) {
  return guarded(h, [&](Ctx& c) { c.d2h(dst, src, n * 4); });
}
static_assert(GATHER_ARG_COLS * GATHER_ARG_ROWS <= Ctx::gather_host_words,
              "an argument-block gather must fit the coherent host buffer");
int stwo_hip_batch_at_multi(stwo_hip_ctx* h, const uint32_t* const* cols, uint32_t n_cols, const uint32_t* rows,
                            uint32_t n_rows, uint32_t* out_host) {
  return guarded(h, [&](Ctx& c) {
    const size_t n = (size_t)n_cols * n_rows;
    if (!n) return;
    if (!cols || !rows || !out_host) throw HipError(1, "null argument");
    if (n_cols <= GATHER_ARG_COLS && n_rows <= GATHER_ARG_ROWS) {
      // a decommitment's reads: a few rows of a layer's hashes or of a few columns, written by the
      // kernel straight into the context's coherent host buffer (GATHER_ARG_COLS x GATHER_ARG_ROWS
      // words fit in it)
      GatherArgs g{};
      for (uint32_t k = 0; k < n_cols; ++k) g.cols[k] = cols[k];
      for (uint32_t r = 0; r < n_rows; ++r) g.rows[r] = rows[r];
      launch_gather_args(g, n_cols, n_rows, c.gather_host_dev, true, c.stream);
      c.sync_main();
      std::memcpy(out_host, c.gather_host, n * 4);
      c.reset_arena();
      return;
    }
    uint32_t* o = c.gather_out(n);
    {
      // the column table repeats between the queried and witness reads of a layer: kept in the
      // table store; the rows go through the staging arena
      const std::vector<const uint32_t*> bases(cols, cols + n_cols);
      const std::vector<uint32_t> rv(rows, rows + n_rows);
      launch_gather_rows(c.table_vec(bases), n_cols, c.upload_vec(rv), n_rows, o, c.stream);
    }
    c.d2h(out_host, o, n * 4);
    c.reset_arena();
  });
}

int stwo_hip_precompute_twiddles(stwo_hip_ctx* h, uint32_t log_domain, uint32_t* tw, uint32_t* itw) {
  return guarded(h, [&](Ctx& c) {
    if (log_domain < 2 || log_domain > 30) throw HipError(3, "log_domain out of range");
    c.ensure_twiddles(log_domain);
    size_t n = (size_t)1 << (log_domain - 1);
    // the cached tree may be larger: a smaller domain's tree is the tail of the larger one
    size_t off = c.tree_len() - n;
    if (tw) {
      std::vector<uint32_t> t(n);
      c.d2h(t.data(), c.tw.p + off, n * 4);
      // layers of the smaller tree are the last layers of the larger tree, followed by the pad
      memcpy(tw, t.data(), n * 4);
    }
    if (itw) c.d2h(itw, c.itw.p + off, n * 4);
  });
}

int stwo_hip_interpolate_columns(stwo_hip_ctx* h, const uint32_t* const* evals, uint32_t* const* coeffs,
                                 uint32_t n_cols, uint32_t log_size) {
  return guarded(h, [&](Ctx& c) {
    if (log_size < 1 || log_size > 30) throw HipError(3, "log_size must be in [1, 30]");
    c.ensure_twiddles(log_size < 2 ? 2 : log_size);
    std::vector<const uint32_t*> s(evals, evals + n_cols);
    std::vector<uint32_t*> d(coeffs, coeffs + n_cols);
    const uint32_t* const* ds = (const uint32_t* const*)c.upload_vec(s);
    uint32_t* const* dd = (uint32_t* const*)c.upload_vec(d);
    {
      const uint32_t low = plan_fft(c, true, n_cols, log_size, log_size, c.itw2.p, c.tree_len(), d.data());
      KScope ks(c, K_INTERPOLATE, (double)n_cols * 8.0 * (double)(1u << log_size));
      ks.units(fft_butterflies(n_cols, log_size, log_size));
      run_fft(c, true, ds, dd, n_cols, log_size, log_size, c.itw2.p, c.tree_len(), canonic_half_init(log_size), low);
    }
    c.end_stream_ordered();
  });
}
int stwo_hip_evaluate_polynomials(stwo_hip_ctx* h, const uint32_t* const* coeffs, uint32_t* const* evals,
                                  uint32_t n_cols, uint32_t log_size, uint32_t log_blowup) {
  return guarded(h, [&](Ctx& c) {
    if (log_size < 1 || log_size + log_blowup > 30) throw HipError(3, "sizes out of range");
    const uint32_t log_n = log_size + log_blowup;
    c.ensure_twiddles(log_n < 2 ? 2 : log_n);
    std::vector<const uint32_t*> s(coeffs, coeffs + n_cols);
    std::vector<uint32_t*> d(evals, evals + n_cols);
    const uint32_t* const* ds = (const uint32_t* const*)c.upload_vec(s);
    uint32_t* const* dd = (uint32_t* const*)c.upload_vec(d);
    {
      const uint32_t low = plan_fft(c, false, n_cols, log_size, log_n, c.tw2.p, c.tree_len(), d.data());
      KScope ks(c, K_EVALUATE, (double)n_cols * (4.0 * (double)(1u << log_size) + 4.0 * (double)(1u << log_n)));
      ks.units(fft_butterflies(n_cols, log_n, log_size));
      run_fft(c, false, ds, dd, n_cols, log_size, log_n, c.tw2.p, c.tree_len(), canonic_half_init(log_n), low);
    }
    c.end_stream_ordered();
  });
}
int stwo_hip_batch_eval_at_point(stwo_hip_ctx* h, const uint32_t* const* coeffs, uint32_t n_cols, uint32_t log_size,
                                 stwo_circle_point_qm31 point, stwo_qm31* out_host) {
  return guarded(h, [&](Ctx& c) {
    if (log_size > 30) throw HipError(3, "log_size out of range");
    check_canonical(point.x);
    check_canonical(point.y);
    std::vector<OodsTask> tasks;
    std::vector<qm31> f(32, qm_zero());
    cpqm31 p = cp(point);
    f[0] = p.y;
    qm31 x = p.x;
    for (uint32_t b = 1; b < log_size; ++b) { f[b] = x; x = qm_double_x(x); }
    std::vector<uint32_t> tg(n_cols, 0), gl(1, log_size);
    for (uint32_t i = 0; i < n_cols; ++i) tasks.push_back({coeffs[i], log_size, i, {0, 0}});
    uint32_t max_chunks = oods_max_chunks(log_size);
    DevBuf partials(&c, (size_t)n_cols * max_chunks * 4), vals(&c, (size_t)n_cols * 4), w(&c, (size_t)4 << OODS_TABLE_LOG);
    const std::vector<uint32_t> bs = oods_batch_starts(tg);
    launch_oods(c.upload_vec(tasks), c.upload_vec(tg), n_cols, c.upload_vec(bs), (uint32_t)bs.size() - 1, log_size,
                c.upload_vec(f), c.upload_vec(gl), 1, (qm31*)w.p, (qm31*)partials.p, (qm31*)vals.p, c.stream);
    c.d2h(out_host, vals.p, (size_t)n_cols * 16);
    c.reset_arena();
  });
}

int stwo_hip_batch_eval_at_point_noalloc(stwo_hip_ctx* h, const uint32_t* const* coeffs, uint32_t n_cols,
                                         uint32_t log_size, stwo_circle_point_qm31 point, stwo_qm31* out_host) {
  return guarded(h, [&](Ctx& c) {
    if (log_size > 30) throw HipError(3, "log_size out of range");
    if (!coeffs || !out_host) throw HipError(1, "null argument");
    check_canonical(point.x);
    check_canonical(point.y);
    if (n_cols == 0) return;
    // the same work as stwo_hip_batch_eval_at_point, every device buffer carved from the scratch:
    // weights table, values, per-chunk partial sums (none needs clearing: each is written before
    // it is read)
    const uint32_t max_chunks = oods_max_chunks(log_size);
    const size_t w_bytes = ((size_t)4 << OODS_TABLE_LOG) * 4, v_bytes = ((size_t)n_cols * 16 + 255) & ~(size_t)255;
    const size_t p_bytes = (size_t)n_cols * max_chunks * 16;
    if (w_bytes + v_bytes + p_bytes > NOALLOC_SCRATCH_BYTES)
      throw HipError(3, "batch_eval_at_point_noalloc: n_cols x 2^max(0, log_size - 11) x 16 bytes of partial sums "
                        "exceed the 32 MiB scratch (31 MiB for partials and values)");
    std::vector<OodsTask> tasks;
    std::vector<qm31> f(32, qm_zero());
    const cpqm31 p = cp(point);
    f[0] = p.y;
    qm31 x = p.x;
    for (uint32_t b = 1; b < log_size; ++b) { f[b] = x; x = qm_double_x(x); }
    std::vector<uint32_t> tg(n_cols, 0), gl(1, log_size);
    for (uint32_t i = 0; i < n_cols; ++i) tasks.push_back({coeffs[i], log_size, i, {0, 0}});
    const std::vector<uint32_t> bs = oods_batch_starts(tg);
    // the small tables go through the context's staging buffers, allocated when the context was
    // created; the call is refused rather than let them overflow into an allocation
    const size_t tab = ((tasks.size() * sizeof(OodsTask) + 255) & ~(size_t)255) + ((tg.size() * 4 + 255) & ~(size_t)255) +
                       ((bs.size() * 4 + 255) & ~(size_t)255) + 512 + 256;
    if (c.pinned_off + tab > c.pinned_cap || c.arena_off + tab > c.arena_cap || (size_t)n_cols * 16 > c.d2h_cap)
      throw HipError(3, "batch_eval_at_point_noalloc: tables exceed the context's staging buffers");
    DeviceScratch& sc = scratch_of(c.device);
    std::lock_guard<std::mutex> lk(sc.mu);
    uint8_t* base = sc.ptr;
    qm31* w = (qm31*)base;
    qm31* vals = (qm31*)(base + w_bytes);
    qm31* partials = (qm31*)(base + w_bytes + v_bytes);
    const OodsTask* d_tasks = c.stage_vec(tasks);
    const uint32_t* d_tg = c.stage_vec(tg);
    const uint32_t* d_bs = c.stage_vec(bs);
    const qm31* d_f = c.stage_vec(f);
    const uint32_t* d_gl = c.stage_vec(gl);
    c.commit_staged();
    launch_oods(d_tasks, d_tg, n_cols, d_bs, (uint32_t)bs.size() - 1, log_size, d_f, d_gl, 1, w, partials, vals,
                c.stream);
    c.d2h(out_host, vals, (size_t)n_cols * 16);
    c.reset_arena();
  }, /*may_recycle=*/false);
}

int stwo_hip_barycentric_weights(stwo_hip_ctx* h, uint32_t log_size, stwo_circle_point_qm31 point,
                                 uint32_t* const weights[4]) {
  return guarded(h, [&](Ctx& c) {
    if (log_size < 1 || log_size > 30) throw HipError(3, "log_size out of range");
    check_canonical(point.x);
    check_canonical(point.y);
    for (int k = 0; k < 4; ++k)
      if (!weights[k]) throw HipError(1, "null weight column");
    cpqm31 p = cp(point);
    std::vector<cpm31> shifts;
    for (uint32_t j = 1; j < log_size; ++j) {
      HCoset hc = HCoset::odds(j);
      shifts.push_back(cp_add(cp_conj(index_to_point(hc.initial_index)), index_to_point(hc.step_size >> 1)));
    }
    const qm31 vn = coset_vanishing_qm31(log_size, p);
    launch_barycentric_weights(log_size, p, vn, shifts.data(), {{weights[0], weights[1], weights[2], weights[3]}},
                               c.stream);
    c.sync();
  });
}
int stwo_hip_batch_barycentric_eval(stwo_hip_ctx* h, const uint32_t* const* evals, uint32_t n_cols, uint32_t log_size,
                                    const uint32_t* const weights[4], stwo_qm31* out_host) {
  return guarded(h, [&](Ctx& c) {
    if (log_size > 30) throw HipError(3, "log_size out of range");
    if (n_cols == 0) return;
    for (uint32_t i = 0; i < n_cols; ++i)
      if (!evals[i]) throw HipError(1, "null evaluation column");
    std::vector<const uint32_t*> ev(evals, evals + n_cols);
    const uint32_t blocks = ((1u << log_size) + 2047) / 2048;
    DevBuf partial(&c, (size_t)n_cols * blocks * 4), vals(&c, (size_t)n_cols * 4);
    launch_barycentric_eval(log_size, (const uint32_t* const*)c.upload_vec(ev), n_cols,
                            {{weights[0], weights[1], weights[2], weights[3]}}, (qm31*)partial.p, (qm31*)vals.p,
                            c.stream);
    c.d2h(out_host, vals.p, (size_t)n_cols * 16);
    c.reset_arena();
  });
}

int stwo_hip_commit_on_layer(stwo_hip_ctx* h, uint32_t log_size, const uint32_t* prev, const uint32_t* const* cols,
                             uint32_t n_cols, uint32_t* out) {
  return guarded(h, [&](Ctx& c) {
    if (log_size > 30) throw HipError(3, "log_size out of range");
    std::vector<const uint32_t*> v(cols, cols + n_cols);
    // columns laid out at a fixed stride of one column need no address table
    bool strided = n_cols > 1;
    for (uint32_t k = 1; strided && k < n_cols; ++k) strided = v[k] == v[0] + ((size_t)k << log_size);
    // under graph capture, at most 16 columns travel as kernel arguments (no staged table)
    const bool small = c.capturing && !strided;
    if (small && n_cols > 16) throw HipError(3, "not capturable: more than 16 columns not at one stride");
    const uint32_t* const* d = !strided && !small && n_cols ? (const uint32_t* const*)c.upload_vec(v) : nullptr;
    {
      KScope ks(c, K_MERKLE, ((prev ? 64.0 : 0.0) + 4.0 * n_cols + 32.0) * (double)(1u << log_size));
      ks.units(blake2s_node_compressions(prev != nullptr, n_cols) * (double)(1u << log_size));
      if (strided) launch_merkle_layer_strided(log_size, prev, v[0], (size_t)1 << log_size, n_cols, out, c.stream);
      else if (small) launch_merkle_layer_small(log_size, prev, v.data(), n_cols, out, c.stream);
      else launch_merkle_layer(log_size, prev, d, n_cols, out, c.stream);
    }
    c.end_stream_ordered();
  }, true, true);
}

// The twiddle trees cover log_domain; under graph capture they must exist already (building them
// allocates and uploads).
static void need_twiddles(Ctx& c, uint32_t log_domain) {
  if (!c.capturing) return c.ensure_twiddles(log_domain);
  if (c.tw_log < log_domain)
    throw HipError(3, "not capturable: the twiddles of this size are not built yet (one call of the size "
                      "outside the capture builds them)");
}

int stwo_hip_fold_circle_into_line(stwo_hip_ctx* h, uint32_t log_src, const uint32_t* const src[4],
                                   uint32_t* const dst[4], stwo_qm31 alpha) {
  return guarded(h, [&](Ctx& c) {
    if (log_src < 3 || log_src > 30) throw HipError(3, "log_src must be in [3, 30]");
    check_canonical(alpha);
    need_twiddles(c, log_src);
    {
      KScope ks(c, K_FRI_FOLD, 16.0 * (double)(1u << log_src) + 32.0 * (double)(1u << (log_src - 1)));
      launch_fold_circle_into_line(log_src, {{src[0], src[1], src[2], src[3]}}, {{dst[0], dst[1], dst[2], dst[3]}},
                                   q(alpha), c.itw.p, c.tree_len(), c.stream);
    }
    c.end_stream_ordered();
  }, true, true);
}
int stwo_hip_fold_circle_into_line_dev(stwo_hip_ctx* h, uint32_t log_src, const uint32_t* const src[4],
                                       uint32_t* const dst[4], const uint32_t* alpha, uint32_t dst_is_zero) {
  return guarded(h, [&](Ctx& c) {
    if (log_src < 3 || log_src > 30) throw HipError(3, "log_src must be in [3, 30]");
    if (!alpha) throw HipError(1, "null alpha");
    need_twiddles(c, log_src);
    {
      KScope ks(c, K_FRI_FOLD, 16.0 * (double)(1u << log_src) + 32.0 * (double)(1u << (log_src - 1)));
      launch_fold_circle_into_line_dev(log_src, {{src[0], src[1], src[2], src[3]}},
                                       {{dst[0], dst[1], dst[2], dst[3]}}, (const qm31*)alpha, dst_is_zero != 0, c.itw.p,
                                       c.tree_len(), c.stream);
    }
    c.end_stream_ordered();
  }, true, true);
}
int stwo_hip_fold_line_dev(stwo_hip_ctx* h, uint32_t log_src, const uint32_t* const src[4], uint32_t* const dst[4],
                           const uint32_t* alpha) {
  return guarded(h, [&](Ctx& c) {
    if (log_src < 1 || log_src > 29) throw HipError(3, "log_src must be in [1, 29]");
    if (!alpha) throw HipError(1, "null alpha");
    need_twiddles(c, log_src + 1);
    {
      KScope ks(c, K_FRI_FOLD, 24.0 * (double)(1u << log_src));
      launch_fold_line_dev(log_src, {{src[0], src[1], src[2], src[3]}}, {{dst[0], dst[1], dst[2], dst[3]}},
                           (const qm31*)alpha, 0, c.itw.p, c.tree_len(), c.stream);
    }
    c.end_stream_ordered();
  }, true, true);
}
int stwo_hip_fold_circle_into_line_grouped(stwo_hip_ctx* h, uint32_t log_src, uint32_t n_groups,
                                           const uint32_t* const* src, uint32_t* const* dst, stwo_qm31 alpha) {
  return guarded(h, [&](Ctx& c) {
    if (log_src < 3 || log_src > 30) throw HipError(3, "log_src must be in [3, 30]");
    if (n_groups == 0) return;
    if (n_groups > 65535) throw HipError(3, "at most 65535 groups per call");
    check_canonical(alpha);
    c.ensure_twiddles(log_src);
    bool aligned = true;
    for (uint32_t k = 0; k < 4 * n_groups; ++k) aligned &= (((uintptr_t)src[k] | (uintptr_t)dst[k]) & 15) == 0;
    const double bytes = (double)n_groups * (16.0 * (double)(1u << log_src) + 32.0 * (double)(1u << (log_src - 1)));
    KScope ks(c, K_FRI_FOLD, bytes);
    if (aligned) {
      std::vector<const uint32_t*> sv(src, src + 4 * (size_t)n_groups);
      std::vector<uint32_t*> dv(dst, dst + 4 * (size_t)n_groups);
      launch_fold_circle_into_line_grouped(log_src, n_groups, (const uint32_t* const*)c.upload_vec(sv),
                                           (uint32_t* const*)c.upload_vec(dv), q(alpha), c.itw.p, c.tree_len(),
                                           c.stream);
    } else {
      for (uint32_t g = 0; g < n_groups; ++g)
        launch_fold_circle_into_line(log_src, {{src[4 * g], src[4 * g + 1], src[4 * g + 2], src[4 * g + 3]}},
                                     {{dst[4 * g], dst[4 * g + 1], dst[4 * g + 2], dst[4 * g + 3]}}, q(alpha), c.itw.p,
                                     c.tree_len(), c.stream);
    }
    c.end_stream_ordered();
  });
}
int stwo_hip_fold_line(stwo_hip_ctx* h, uint32_t log_src, const uint32_t* const src[4], uint32_t* const dst[4],
                       stwo_qm31 alpha, uint32_t fold_step) {
  return guarded(h, [&](Ctx& c) {
    if (log_src < 1 || log_src > 29) throw HipError(3, "log_src must be in [1, 29]");
    if (fold_step < 1 || fold_step > log_src) throw HipError(1, "fold_step must be in [1, log_src]");
    check_canonical(alpha);
    if (c.capturing && fold_step != 1) throw HipError(3, "not capturable: fold_step above 1 allocates");
    need_twiddles(c, log_src + 1);
    // fold_step folds, the coefficient squared before each after the first (prover/backend/cpu/fri.rs:15-30);
    // the intermediate layers live in one temporary buffer: layers after even steps at its start
    // (at most 4 x 2^(log_src-1) words), after odd steps from word 2^(log_src+1) on (at most
    // 4 x 2^(log_src-2) words)
    const size_t tmp_words = fold_step > 1 ? (size_t)3 << log_src : 0;
    DevBuf tmp;
    if (tmp_words) tmp = DevBuf(&c, tmp_words);
    CCols4 from = {{src[0], src[1], src[2], src[3]}};
    qm31 a = q(alpha);
    for (uint32_t st = 0; st < fold_step; ++st) {
      const uint32_t ls = log_src - st;
      const size_t half = (size_t)1 << (ls - 1);
      Cols4 to = {{dst[0], dst[1], dst[2], dst[3]}};
      if (st + 1 < fold_step) {
        // alternate between the two halves of the temporary buffer
        uint32_t* base = tmp.p + ((st & 1) ? ((size_t)2 << log_src) : 0);
        for (int k = 0; k < 4; ++k) to.c[k] = base + (size_t)k * half;
      }
      {
        KScope ks(c, K_FRI_FOLD, 24.0 * (double)(1u << ls));
        launch_fold_line(ls, from, to, a, c.itw.p, c.tree_len(), c.stream);
      }
      from = {{to.c[0], to.c[1], to.c[2], to.c[3]}};
      a = qm_mul(a, a);
    }
    c.end_stream_ordered();
  }, true, true);
}

int stwo_hip_accumulate_numerators(stwo_hip_ctx* h, uint32_t log_size, const uint32_t* const* cols, uint32_t n_cols,
                                   const stwo_sample_batch* batches, uint32_t n_batches, uint32_t* const* out,
                                   stwo_qm31* first_linear_host) {
  return guarded(h, [&](Ctx& c) {
    std::vector<NumEntry> ents;
    std::vector<uint32_t> starts;
    std::vector<qm31> Bs;
    for (uint32_t b = 0; b < n_batches; ++b) {
      const auto& bt = batches[b];
      cpqm31 p = cp(bt.point);
      if (qm_eq(p.y, qm_conj(p.y))) throw HipError(1, "Cannot evaluate a line with a single point");
      starts.push_back((uint32_t)ents.size());
      qm31 A = qm_zero(), B = qm_zero();
      for (uint32_t j = 0; j < bt.n_columns; ++j) {
        if (bt.column_indices[j] >= n_cols) throw HipError(1, "column index out of range");
        qm31 v = q(bt.sample_values[j]), al = q(bt.random_coeffs[j]);
        qm31 aa = qm_sub(qm_conj(v), v), cc = qm_sub(qm_conj(p.y), p.y);
        qm31 bb = qm_sub(qm_mul(v, cc), qm_mul(aa, p.y));
        NumEntry e{};
        e.col = cols[bt.column_indices[j]];
        e.c = qm_mul(al, cc);
        ents.push_back(e);
        A = qm_add(A, qm_mul(al, aa));
        B = qm_add(B, qm_mul(al, bb));
      }
      Bs.push_back(B);
      if (first_linear_host) first_linear_host[b] = q(A);
    }
    starts.push_back((uint32_t)ents.size());
    bool aligned = true;
    for (uint32_t j = 0; j < n_cols; ++j) aligned &= ((uintptr_t)cols[j] & 15) == 0;
    std::vector<uint32_t*> ov(out, out + 4 * (size_t)n_batches);
    const NumEntry* d_ents = c.stage_vec(ents);
    const uint32_t* d_starts = c.stage_vec(starts);
    const qm31* d_bs = c.stage_vec(Bs);
    uint32_t* const* d_out = (uint32_t* const*)c.upload_vec(ov);
    launch_accumulate_numerators(log_size, d_ents, d_starts, d_bs, n_batches, d_out, aligned, c.stream);
    c.end_stream_ordered();
  });
}
int stwo_hip_compute_quotients_and_combine(stwo_hip_ctx* h, const stwo_accumulated_numerators* accs, uint32_t n_accs,
                                           uint32_t lifting_log_size, uint32_t* const out[4]) {
  return guarded(h, [&](Ctx& c) {
    if (lifting_log_size < 1 || lifting_log_size > 30) throw HipError(3, "lifting_log_size out of range");
    std::vector<QuotPoint> pts;
    std::vector<QuotPart> parts;
    for (uint32_t a = 0; a < n_accs; ++a) {
      if (accs[a].log_size > lifting_log_size) throw HipError(1, "accumulation larger than the lifting domain");
      cpqm31 p = cp(accs[a].sample_point);
      QuotPoint qp{};
      qp.prx = qm_lo(p.x); qp.pix = qm_hi(p.x); qp.pry = qm_lo(p.y); qp.piy = qm_hi(p.y);
      qp.first_linear = q(accs[a].first_linear_term);
      qp.part_begin = (uint32_t)parts.size();
      QuotPart pa{};
      for (int k = 0; k < 4; ++k) pa.c[k] = accs[a].partial[k];
      pa.log_size = accs[a].log_size;
      parts.push_back(pa);
      qp.part_end = (uint32_t)parts.size();
      pts.push_back(qp);
    }
    launch_quotients(lifting_log_size, c.domain_xy(lifting_log_size), c.upload_vec(pts), n_accs, c.upload_vec(parts),
                     {{out[0], out[1], out[2], out[3]}}, c.stream);
    c.end_stream_ordered();
  });
}

int stwo_hip_accumulate(stwo_hip_ctx* h, uint32_t log_size, uint32_t* const dst[4], const uint32_t* const src[4]) {
  return guarded(h, [&](Ctx& c) {
    launch_accumulate_lifted(log_size, {{dst[0], dst[1], dst[2], dst[3]}}, log_size, {{src[0], src[1], src[2], src[3]}},
                             c.stream);
    c.end_stream_ordered();
  });
}
int stwo_hip_lift_and_accumulate(stwo_hip_ctx* h, const uint32_t* const* cols4, const uint32_t* log_sizes,
                                 uint32_t n_cols, uint32_t* const out[4]) {
  return guarded(h, [&](Ctx& c) {
    if (n_cols == 0) throw HipError(1, "no columns");
    for (uint32_t i = 1; i < n_cols; ++i)
      if (log_sizes[i] < log_sizes[i - 1]) throw HipError(1, "columns must be sorted by increasing size");
    if (log_sizes[0] < 1) throw HipError(1, "A column must be of length at least 2.");
    uint32_t L = log_sizes[n_cols - 1];
    Cols4 o = {{out[0], out[1], out[2], out[3]}};
    for (int k = 0; k < 4; ++k) launch_copy_words(out[k], cols4[4 * (n_cols - 1) + k], (size_t)1 << L, c.stream);
    for (uint32_t i = 0; i + 1 < n_cols; ++i)
      launch_accumulate_lifted(L, o, log_sizes[i], {{cols4[4 * i], cols4[4 * i + 1], cols4[4 * i + 2], cols4[4 * i + 3]}},
                               c.stream);
    c.end_stream_ordered();
  });
}
int stwo_hip_generate_secure_powers(stwo_hip_ctx* h, stwo_qm31 felt, uint32_t n, stwo_qm31* out) {
  return guarded(h, [&](Ctx&) {
    qm31 p = qm_one(), f = q(felt);
    for (uint32_t i = 0; i < n; ++i) { out[i] = q(p); p = qm_mul(p, f); }
  });
}

int stwo_hip_grind(stwo_hip_ctx* h, const uint8_t digest[32], uint32_t pow_bits, uint64_t* nonce) {
  return guarded(h, [&](Ctx& c) {
    Channel ch;
    memcpy(ch.digest.data(), digest, 32);
    *nonce = device_grind(c, ch.pow_prefix(pow_bits), pow_bits);
  });
}

int stwo_hip_commit_on_layer_poseidon252(stwo_hip_ctx* h, uint32_t log_size, const uint32_t* prev,
                                         const uint32_t* const* cols, uint32_t n_cols, uint32_t* out) {
  return guarded(h, [&](Ctx& c) {
    if (log_size > 30) throw HipError(3, "log_size out of range");
    if (!out || (n_cols && !cols)) throw HipError(1, "null argument");
    std::vector<const uint32_t*> v(cols, cols + n_cols);
    const uint32_t* const* d = n_cols ? (const uint32_t* const*)c.upload_vec(v) : nullptr;
    {
      // permutations per node: one for two children alone, else one per absorbed pair plus the last
      const double absorbed = (prev ? 2.0 : 0.0) + (double)((n_cols + 7) / 8);
      const double perms = prev && n_cols == 0 ? 1.0 : std::floor(absorbed / 2) + 1;
      KScope ks(c, K_MERKLE, ((prev ? 64.0 : 0.0) + 4.0 * n_cols + 32.0) * (double)(1u << log_size));
      ks.units(perms * (double)(1u << log_size));
      launch_p252_layer(log_size, prev, d, n_cols, out, c.stream);
    }
    c.end_stream_ordered();
  });
}
int stwo_hip_build_leaves_lifted(stwo_hip_ctx* h, uint32_t hasher, const uint32_t* const* cols,
                                 const uint32_t* log_sizes, uint32_t n_cols, uint32_t lifting_log_size,
                                 uint32_t* out) {
  return guarded(h, [&](Ctx& c) {
    if (hasher > LIFTED_POSEIDON252) throw HipError(1, "unknown lifted hasher");
    if (!out || (n_cols && (!cols || !log_sizes))) throw HipError(1, "null argument");
    if (lifting_log_size > 30) throw HipError(3, "lifting_log_size out of range");
    std::vector<LiftedCol> t(n_cols);
    double bytes = 32.0 * (double)(n_cols ? 1u << lifting_log_size : 1u);
    for (uint32_t i = 0; i < n_cols; ++i) {
      if (log_sizes[i] < 1) throw HipError(1, "a column must be of length >= 2");
      if (i && log_sizes[i] < log_sizes[i - 1])
        throw HipError(1, "columns must come in non-decreasing order of length");
      if (log_sizes[i] > lifting_log_size) throw HipError(1, "a column is longer than the lifting domain");
      t[i] = LiftedCol{cols[i], lifting_log_size - log_sizes[i], 0u};
      bytes += 4.0 * (double)(1u << log_sizes[i]);
    }
    const LiftedCol* d = n_cols ? (const LiftedCol*)c.upload_vec(t) : nullptr;
    {
      KScope ks(c, K_MERKLE, bytes);
      if (hasher == LIFTED_POSEIDON252)
        ks.units((std::floor((double)((n_cols + 7) / 8) / 2) + 1) * (double)(n_cols ? 1u << lifting_log_size : 1u));
      else
        ks.units(blake2s_node_compressions(false, n_cols) * (double)(n_cols ? 1u << lifting_log_size : 1u));
      launch_lifted_leaves(hasher, lifting_log_size, d, n_cols, out, c.stream);
    }
    c.end_stream_ordered();
  });
}
int stwo_hip_build_next_layer_lifted(stwo_hip_ctx* h, uint32_t hasher, const uint32_t* prev, uint32_t log_size,
                                     uint32_t* out) {
  return guarded(h, [&](Ctx& c) {
    if (hasher > LIFTED_POSEIDON252) throw HipError(1, "unknown lifted hasher");
    if (!out || !prev) throw HipError(1, "null argument");
    if (log_size > 29) throw HipError(3, "log_size out of range");
    {
      KScope ks(c, K_MERKLE, 96.0 * (double)(1u << log_size));
      if (hasher == LIFTED_POSEIDON252) {
        ks.units((double)(1u << log_size));
        launch_p252_layer(log_size, prev, nullptr, 0, out, c.stream);
      } else {
        ks.units((double)(1u << log_size));
        launch_merkle_layer_small(log_size, prev, nullptr, 0, out, c.stream);
      }
    }
    c.end_stream_ordered();
  });
}
int stwo_hip_grind_poseidon252(stwo_hip_ctx* h, const uint32_t digest[8], uint32_t pow_bits, uint64_t* nonce) {
  return guarded(h, [&](Ctx& c) {
    if (!digest || !nonce) throw HipError(1, "null argument");
    if (!p252_canonical(digest)) throw HipError(1, "digest is not a canonical felt252");
    if (pow_bits > 64) throw HipError(3, "pow_bits above 64");
    DevBuf best(&c, 2);
    KScope ks(c, K_GRIND, 0);
    *nonce = p252_grind(digest, pow_bits, reinterpret_cast<unsigned long long*>(best.p), c.stream);
  });
}
int stwo_poseidon252_hash_many(const uint32_t* values, size_t n, uint32_t out[8]) {
  if ((n && !values) || !out) return 1;
  for (size_t k = 0; k < n; ++k)
    if (!p252_canonical(values + 8 * k)) return 1;
  p252_hash_many_host(values, n, out);
  return 0;
}
int stwo_poseidon252_permute(uint32_t state[24]) {
  if (!state) return 1;
  for (int k = 0; k < 3; ++k)
    if (!p252_canonical(state + 8 * k)) return 1;
  p252_permute_host(state);
  return 0;
}

int stwo_hip_evaluate_constraint_quotients_on_domain(stwo_hip_ctx* h, uint32_t* const acc[4],
                                                     const uint32_t* const* trace, uint32_t n_trace,
                                                     const uint32_t* const* interaction, uint32_t n_interaction,
                                                     const stwo_qm31* coeff_powers, uint32_t n_constraints,
                                                     const uint32_t* denom_inv_host, uint32_t trace_log,
                                                     uint32_t eval_log, const uint8_t* blob, size_t blob_len,
                                                     stwo_qm31 cumsum_shift) {
  return guarded(h, [&](Ctx& c) {
    if (!blob || blob_len < 4) throw HipError(1, "eval_blob too short");
    if (eval_log <= trace_log || eval_log > 30) throw HipError(3, "eval_log must exceed trace_log");
    uint32_t id;
    memcpy(&id, blob, 4);
    std::vector<qm31> pows(n_constraints);
    for (uint32_t i = 0; i < n_constraints; ++i) pows[i] = q(coeff_powers[i]);
    std::vector<uint32_t> dinv(denom_inv_host, denom_inv_host + (1u << (eval_log - trace_log)));
    Cols4 a4 = {{acc[0], acc[1], acc[2], acc[3]}};
    std::vector<const uint32_t*> tv(trace, trace + n_trace);
    if (id == fnv1a("fibonacci_example")) {
      if (n_trace < 2 || n_constraints != n_trace - 2) throw HipError(1, "wide-fib: n_constraints must be n_trace - 2");
      launch_wide_fib_composition(eval_log, trace_log, (const uint32_t* const*)c.upload_vec(tv), n_trace,
                                  c.upload_vec(pows), c.upload_vec(dinv), a4, true, c.stream);
    } else if (id == fnv1a("poseidon_example")) {
      if (blob_len < 4 + 4 + 32) throw HipError(1, "poseidon eval_blob too short");
      PoseidonParams pp{};
      memcpy(&pp.ipr, blob + 4, 4);
      stwo_qm31 z, al;
      memcpy(&z, blob + 8, 16);
      memcpy(&al, blob + 24, 16);
      if (pp.ipr == 0 || n_trace != 158 * pp.ipr || n_interaction != 4 * pp.ipr || n_constraints != 143 * pp.ipr)
        throw HipError(1, "poseidon: column or constraint counts do not match instances_per_row");
      pp.z = q(z);
      qm31 a = q(al), p = qm_one();
      for (int i = 0; i < 16; ++i) { pp.alpha_pow[i] = p; p = qm_mul(p, a); }
      pp.shift = q(cumsum_shift);
      std::vector<const uint32_t*> iv(interaction, interaction + n_interaction);
      launch_poseidon_composition(eval_log, trace_log, (const uint32_t* const*)c.upload_vec(tv),
                                  (const uint32_t* const*)c.upload_vec(iv), (const PoseidonParams*)c.upload(&pp, sizeof(pp)),
                                  pp.ipr, c.upload_vec(pows), c.upload_vec(dinv), a4, true, c.stream);
    } else if (id == fnv1a("expr_component")) {
      // a constraint program, compiled into a kernel on first use: u32 program words, the
      // program, u32 parameter count, the parameters (QM31 each); parameter 0 is cumsum_shift
      if (blob_len < 8) throw HipError(1, "expr_component eval_blob too short");
      uint32_t nw;
      memcpy(&nw, blob + 4, 4);
      if (blob_len < 8 + 4 * (size_t)nw + 4) throw HipError(1, "expr_component eval_blob too short");
      std::vector<uint32_t> words(nw);
      memcpy(words.data(), blob + 8, 4 * (size_t)nw);
      uint32_t np;
      memcpy(&np, blob + 8 + 4 * (size_t)nw, 4);
      if (blob_len != 12 + 4 * (size_t)nw + 16 * (size_t)np) throw HipError(1, "expr_component eval_blob length");
      std::vector<qm31> params(1 + (size_t)np);
      params[0] = q(cumsum_shift);
      memcpy(params.data() + 1, blob + 12 + 4 * (size_t)nw, 16 * (size_t)np);
      for (auto& x : params)
        if (x.a >= P || x.b >= P || x.c >= P || x.d >= P) throw HipError(1, "non-canonical parameter");
      if (compiled_component_constraints(words.data(), words.size()) != n_constraints)
        throw HipError(1, "expr_component: n_constraints differs from the program's");
      std::vector<const uint32_t*> iv(interaction, interaction + n_interaction);
      const auto* d_t = (const uint32_t* const*)c.upload_vec(tv);
      const auto* d_i = n_interaction ? (const uint32_t* const*)c.upload_vec(iv) : nullptr;
      const qm31* d_p = c.upload_vec(params);
      const qm31* d_w = c.upload_vec(pows);
      const uint32_t* d_d = c.upload_vec(dinv);
      KScope ks(c, K_COMPOSITION, (4.0 * (n_trace + n_interaction) + 32.0) * (double)(1u << eval_log));
      evaluate_compiled_component(c.device, words.data(), words.size(), eval_log, trace_log, d_t, n_trace, d_i,
                                  n_interaction, d_p, (uint32_t)params.size(), d_w, n_constraints, d_d, a4, c.stream);
    } else {
      throw HipError(5, "no device kernel for this component");
    }
    c.end_stream_ordered();
  });
}

int stwo_hip_prove_wide_fibonacci(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t n_cols, uint32_t log_n_rows,
                                  const stwo_pcs_config* config, stwo_proof** out) {
  return guarded(h, [&](Ctx& c) {
    if (!out) throw HipError(1, "null out");
    if (n_cols < 3) throw HipError(1, "wide-fib needs at least 3 columns");
    if (log_n_rows < 3 || log_n_rows > 26) throw HipError(3, "log_n_rows must be in [3, 26]");
    PcsCfg cfg = cfg_of(config);
    std::vector<const uint32_t*> t(trace, trace + n_cols);
    auto* p = new stwo_proof();
    try {
      p->r = prove_wide_fib(c, t, log_n_rows, n_cols, cfg);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    reject_unsatisfied(*p);
  });
}
int stwo_hip_prove_wide_fibonacci_multi(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t n_components,
                                        const uint32_t* log_n_rows, const uint32_t* n_cols,
                                        const stwo_pcs_config* config, stwo_proof** out) {
  return guarded(h, [&](Ctx& c) {
    if (!out || !log_n_rows || !n_cols) throw HipError(1, "null argument");
    if (n_components == 0) throw HipError(1, "no components");
    std::vector<uint32_t> logs(log_n_rows, log_n_rows + n_components), nc(n_cols, n_cols + n_components);
    size_t total = 0;
    for (uint32_t i = 0; i < n_components; ++i) {
      if (nc[i] < 3) throw HipError(1, "wide-fib needs at least 3 columns per component");
      if (logs[i] < 3 || logs[i] > 26) throw HipError(3, "log_n_rows must be in [3, 26]");
      total += nc[i];
    }
    PcsCfg cfg = cfg_of(config);
    std::vector<const uint32_t*> t(trace, trace + total);
    auto* p = new stwo_proof();
    try {
      p->r = prove_wide_fib_multi(c, t, logs, nc, cfg);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    reject_unsatisfied(*p);
  });
}
int stwo_hip_prove_poseidon(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t ipr, uint32_t log_n_rows,
                            const stwo_pcs_config* config, stwo_proof** out) {
  return guarded(h, [&](Ctx& c) {
    if (!out) throw HipError(1, "null out");
    if (ipr < 1 || ipr > 64) throw HipError(1, "instances_per_row must be in [1, 64]");
    if (log_n_rows < 3 || log_n_rows > 24) throw HipError(3, "log_n_rows must be in [3, 24]");
    PcsCfg cfg = cfg_of(config);
    std::vector<const uint32_t*> t(trace, trace + 158 * ipr);
    auto* p = new stwo_proof();
    try {
      p->r = prove_poseidon(c, t, log_n_rows, ipr, cfg);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    reject_unsatisfied(*p);
  });
}
static int prove_wide_fib_per_call(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t n_components,
                                   const uint32_t* log_n_rows, const uint32_t* n_cols, const stwo_pcs_config* config,
                                   stwo_proof** out, MerkleKind mc) {
  return guarded(h, [&](Ctx&) {
    if (!out || !log_n_rows || !n_cols) throw HipError(1, "null argument");
    if (n_components == 0) throw HipError(1, "no components");
    std::vector<uint32_t> logs(log_n_rows, log_n_rows + n_components), nc(n_cols, n_cols + n_components);
    size_t total = 0;
    for (uint32_t i = 0; i < n_components; ++i) {
      if (nc[i] < 3) throw HipError(1, "wide-fib needs at least 3 columns per component");
      if (logs[i] < 3 || logs[i] > 26) throw HipError(3, "log_n_rows must be in [3, 26]");
      total += nc[i];
    }
    PcsCfg cfg = cfg_of(config);
    std::vector<const uint32_t*> t(trace, trace + total);
    auto* p = new stwo_proof();
    try {
      p->r = percall::prove_wide_fib_multi(h, t, logs, nc, cfg, mc);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    reject_unsatisfied(*p);
  });
}
static int prove_poseidon_per_call(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t ipr, uint32_t log_n_rows,
                                   const stwo_pcs_config* config, stwo_proof** out, MerkleKind mc) {
  return guarded(h, [&](Ctx&) {
    if (!out) throw HipError(1, "null out");
    if (ipr < 1 || ipr > 64) throw HipError(1, "instances_per_row must be in [1, 64]");
    if (log_n_rows < 3 || log_n_rows > 24) throw HipError(3, "log_n_rows must be in [3, 24]");
    PcsCfg cfg = cfg_of(config);
    std::vector<const uint32_t*> t(trace, trace + 158 * ipr);
    auto* p = new stwo_proof();
    try {
      p->r = percall::prove_poseidon(h, t, log_n_rows, ipr, cfg, mc);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    reject_unsatisfied(*p);
  });
}
int stwo_hip_prove_wide_fibonacci_per_call(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t n_components,
                                            const uint32_t* log_n_rows, const uint32_t* n_cols,
                                            const stwo_pcs_config* config, stwo_proof** out) {
  return prove_wide_fib_per_call(h, trace, n_components, log_n_rows, n_cols, config, out, MerkleKind::Blake2s);
}
int stwo_hip_prove_poseidon_per_call(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t ipr, uint32_t log_n_rows,
                                     const stwo_pcs_config* config, stwo_proof** out) {
  return prove_poseidon_per_call(h, trace, ipr, log_n_rows, config, out, MerkleKind::Blake2s);
}
int stwo_hip_prove_wide_fibonacci_poseidon252(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t n_components,
                                               const uint32_t* log_n_rows, const uint32_t* n_cols,
                                               const stwo_pcs_config* config, stwo_proof** out) {
  return prove_wide_fib_per_call(h, trace, n_components, log_n_rows, n_cols, config, out, MerkleKind::Poseidon252);
}
int stwo_hip_prove_poseidon_poseidon252(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t ipr,
                                        uint32_t log_n_rows, const stwo_pcs_config* config, stwo_proof** out) {
  return prove_poseidon_per_call(h, trace, ipr, log_n_rows, config, out, MerkleKind::Poseidon252);
}
int stwo_hip_gen_interaction_trace_poseidon(stwo_hip_ctx* h, const uint32_t* const* trace, uint32_t ipr,
                                            uint32_t log_n_rows, stwo_qm31 z, stwo_qm31 alpha, uint32_t* const* out,
                                            stwo_qm31* claimed_sum) {
  return guarded(h, [&](Ctx& c) {
    if (!trace || !out || !claimed_sum) throw HipError(1, "null argument");
    if (ipr < 1 || ipr > 64 || log_n_rows < 3 || log_n_rows > 26) throw HipError(1, "bad trace shape");
    check_canonical(z);
    check_canonical(alpha);
    PoseidonParams pp{};
    pp.z = q(z);
    qm31 a = q(alpha), pw = qm_one();
    for (int i = 0; i < 16; ++i) { pp.alpha_pow[i] = pw; pw = qm_mul(pw, a); }
    pp.ipr = ipr;
    std::vector<const uint32_t*> tv(trace, trace + 158 * ipr);
    std::vector<uint32_t*> ov(out, out + 4 * ipr);
    // fractions per row, the column sums (u64 per coordinate) for the cumulative-sum shift, then
    // the coset-order prefix sum of the last 4 columns shifted by sum / 2^log_n
    DevBuf sums(&c, 8, true);
    {
      KScope ks(c, K_LOGUP, (double)(1u << log_n_rows) * (128.0 * ipr + 16.0 * ipr));
      launch_poseidon_logup_fracs(log_n_rows, ipr, (const uint32_t* const*)c.upload_vec(tv),
                                  (uint32_t* const*)c.upload_vec(ov), (const PoseidonParams*)c.upload(&pp, sizeof(pp)),
                                  (unsigned long long*)sums.p, c.stream);
    }
    std::vector<uint32_t*> last4(ov.end() - 4, ov.end());
    const uint32_t nb = (1u << log_n_rows) >= 1024 ? (1u << log_n_rows) / 1024 : 1;
    DevBuf scratch(&c, 4 * (size_t)nb);
    {
      KScope ks(c, K_LOGUP, 32.0 * (double)(1u << log_n_rows));
      launch_coset_prefix_sum(log_n_rows, (uint32_t* const*)c.upload_vec(last4), qm_zero(),
                              (const unsigned long long*)sums.p, scratch.p, c.stream);
    }
    unsigned long long sv[4];
    c.d2h(sv, sums.p, sizeof(sv));
    *claimed_sum = {m31_reduce64(sv[0] % P), m31_reduce64(sv[1] % P), m31_reduce64(sv[2] % P), m31_reduce64(sv[3] % P)};
    c.reset_arena();
  });
}

static int verify_wide_fib(const uint8_t* proof, size_t len, uint32_t n_components, const uint32_t* log_n_rows,
                           const uint32_t* n_cols, const stwo_pcs_config* config, char* err, size_t err_len,
                           MerkleKind mc) {
  std::string e;
  int st = 0;
  try {
    if (!proof || !log_n_rows || !n_cols || !config || n_components == 0) throw HipError(1, "null argument");
    std::vector<Component> comps;
    for (uint32_t i = 0; i < n_components; ++i) {
      if (n_cols[i] < 3 || log_n_rows[i] < 1 || log_n_rows[i] > 28) throw HipError(1, "bad component shape");
      Component c;
      c.kind = AIR_WIDE_FIB;
      c.log_n = log_n_rows[i];
      c.n_cols = n_cols[i];
      comps.push_back(c);
    }
    e = verify_proof(proof, len, comps, cfg_of(config), mc);
    st = e.empty() ? 0 : 4;
  } catch (const HipError& x) {
    e = x.what();
    st = x.code;
  } catch (const std::exception& x) {
    e = x.what();
    st = 4;
  }
  if (err && err_len) snprintf(err, err_len, "%s", e.c_str());
  return st;
}
static int verify_poseidon(const uint8_t* proof, size_t len, uint32_t ipr, uint32_t log_n_rows, stwo_qm31 claimed_sum,
                           const stwo_pcs_config* config, char* err, size_t err_len, MerkleKind mc) {
  std::string e;
  int st = 0;
  try {
    if (!proof || !config) throw HipError(1, "null argument");
    if (ipr < 1 || ipr > 64 || log_n_rows < 1 || log_n_rows > 28) throw HipError(1, "bad component shape");
    check_canonical(claimed_sum);
    Component c;
    c.kind = AIR_POSEIDON;
    c.log_n = log_n_rows;
    c.ipr = ipr;
    c.claimed_sum = q(claimed_sum);
    e = verify_proof(proof, len, {c}, cfg_of(config), mc);
    st = e.empty() ? 0 : 4;
  } catch (const HipError& x) {
    e = x.what();
    st = x.code;
  } catch (const std::exception& x) {
    e = x.what();
    st = 4;
  }
  if (err && err_len) snprintf(err, err_len, "%s", e.c_str());
  return st;
}
int stwo_verify_wide_fibonacci(const uint8_t* proof, size_t len, uint32_t n_components, const uint32_t* log_n_rows,
                               const uint32_t* n_cols, const stwo_pcs_config* config, char* err, size_t err_len) {
  return verify_wide_fib(proof, len, n_components, log_n_rows, n_cols, config, err, err_len, MerkleKind::Blake2s);
}
int stwo_verify_poseidon(const uint8_t* proof, size_t len, uint32_t ipr, uint32_t log_n_rows, stwo_qm31 claimed_sum,
                         const stwo_pcs_config* config, char* err, size_t err_len) {
  return verify_poseidon(proof, len, ipr, log_n_rows, claimed_sum, config, err, err_len, MerkleKind::Blake2s);
}
int stwo_verify_wide_fibonacci_poseidon252(const uint8_t* proof, size_t len, uint32_t n_components,
                                           const uint32_t* log_n_rows, const uint32_t* n_cols,
                                           const stwo_pcs_config* config, char* err, size_t err_len) {
  return verify_wide_fib(proof, len, n_components, log_n_rows, n_cols, config, err, err_len, MerkleKind::Poseidon252);
}
int stwo_verify_poseidon_poseidon252(const uint8_t* proof, size_t len, uint32_t ipr, uint32_t log_n_rows,
                                     stwo_qm31 claimed_sum, const stwo_pcs_config* config, char* err,
                                     size_t err_len) {
  return verify_poseidon(proof, len, ipr, log_n_rows, claimed_sum, config, err, err_len, MerkleKind::Poseidon252);
}

int stwo_proof_status(const stwo_proof* p) { return p ? p->r.status : -1; }
size_t stwo_proof_size(const stwo_proof* p) { return p ? p->r.proof.size() : 0; }
const uint8_t* stwo_proof_bytes(const stwo_proof* p) { return p ? p->r.proof.data() : nullptr; }
void stwo_proof_digest(const stwo_proof* p, uint8_t out[32]) { if (p) memcpy(out, p->r.digest.data(), 32); }
void stwo_proof_claimed_sum(const stwo_proof* p, stwo_qm31* out) { if (p && out) *out = q(p->r.claimed_sum); }
void stwo_proof_stage_ms(const stwo_proof* p, double out[8]) { if (p) memcpy(out, p->r.stage_ms, sizeof(double) * 8); }
void stwo_proof_free(stwo_proof* p) { delete p; }

int stwo_hip_gen_trace_wide_fibonacci(stwo_hip_ctx* h, uint32_t* const* cols, uint32_t n_cols, uint32_t log_n_rows,
                                      uint32_t mode, uint64_t seed) {
  return guarded(h, [&](Ctx& c) {
    if (n_cols < 2 || log_n_rows > 28) throw HipError(1, "bad trace shape");
    std::vector<uint32_t*> v(cols, cols + n_cols);
    launch_gen_wide_fib(log_n_rows, (uint32_t* const*)c.upload_vec(v), n_cols, mode, seed, c.stream);
    c.sync();
    c.reset_arena();
  });
}
int stwo_hip_gen_trace_poseidon(stwo_hip_ctx* h, uint32_t* const* cols, uint32_t ipr, uint32_t log_n_rows, uint32_t mode,
                                uint64_t seed) {
  return guarded(h, [&](Ctx& c) {
    if (ipr < 1 || ipr > 64 || log_n_rows > 26) throw HipError(1, "bad trace shape");
    std::vector<uint32_t*> v(cols, cols + 158 * ipr);
    launch_gen_poseidon(log_n_rows, (uint32_t* const*)c.upload_vec(v), ipr, mode, seed, c.stream);
    c.sync();
    c.reset_arena();
  });
}

// ---- ColumnOps / Column ------------------------------------------------------------------------
int stwo_hip_bit_reverse_column(stwo_hip_ctx* h, uint32_t* col, uint32_t log_size) {
  return guarded(h, [&](Ctx& c) {
    if (!col) throw HipError(1, "null column");
    if (log_size > 31) throw HipError(3, "log_size out of range");
    launch_bit_reverse(col, log_size, c.stream);
    c.end_stream_ordered();
  });
}
int stwo_hip_column_at(stwo_hip_ctx* h, const uint32_t* col, size_t index, uint32_t* out) {
  return guarded(h, [&](Ctx& c) {
    if (!col || !out) throw HipError(1, "null argument");
    c.d2h(out, col + index, 4);
  });
}
int stwo_hip_column_set(stwo_hip_ctx* h, uint32_t* col, size_t index, uint32_t value) {
  return guarded(h, [&](Ctx& c) {
    if (!col) throw HipError(1, "null column");
    const uint32_t* v = (const uint32_t*)c.upload(&value, 4);
    launch_copy_words(col + index, v, 1, c.stream);
    c.sync();
    c.reset_arena();
  });
}
int stwo_hip_split_at_mid(stwo_hip_ctx* h, uint32_t* col, size_t n_words, uint32_t** left, uint32_t** right) {
  return guarded(h, [&](Ctx&) {
    if (!col || !left || !right) throw HipError(1, "null argument");
    if (n_words < 2 || (n_words & 1)) throw HipError(1, "split_at_mid needs an even length of at least 2");
    *left = col;
    *right = col + n_words / 2;
  });
}

// ---- PolyOps: extend, twiddles of any coset, transforms on any domain, folding evaluation ------
int stwo_hip_extend(stwo_hip_ctx* h, const uint32_t* coeffs, uint32_t log_size, uint32_t new_log_size, uint32_t* out) {
  return guarded(h, [&](Ctx& c) {
    if (!coeffs || !out) throw HipError(1, "null argument");
    if (new_log_size < log_size || new_log_size > 31) throw HipError(1, "new log size must be at least the old one");
    const size_t n = (size_t)1 << log_size, m = (size_t)1 << new_log_size;
    if (out != coeffs) launch_copy_words(out, coeffs, n, c.stream);
    if (m > n) launch_fill_zero(out + n, m - n, c.stream);
    c.end_stream_ordered();
  });
}
int stwo_hip_twiddles_create(stwo_hip_ctx* h, uint32_t coset_initial_index, uint32_t coset_log_size,
                             stwo_hip_twiddles** out) {
  return guarded(h, [&](Ctx& c) {
    if (!out) throw HipError(1, "null out");
    if (coset_log_size > 29) throw HipError(3, "coset log size must be at most 29");
    auto t = std::make_unique<stwo_hip_twiddles>();
    t->init_index = coset_initial_index & 0x7fffffffu;
    t->log_size = coset_log_size;
    build_twiddle_tree(c, t->init_index, coset_log_size, t->tw, t->itw, t->tw2, t->itw2);
    h->trees.insert(t.get());
    *out = t.release();
  });
}
int stwo_hip_twiddles_read(stwo_hip_ctx* h, const stwo_hip_twiddles* t, uint32_t* tw_host, uint32_t* itw_host) {
  return guarded(h, [&](Ctx& c) {
    if (!t || !h->trees.count(const_cast<stwo_hip_twiddles*>(t))) throw HipError(1, "unknown twiddle tree");
    const size_t n = (size_t)1 << t->log_size;
    if (tw_host) c.d2h(tw_host, t->tw.p, n * 4);
    if (itw_host) c.d2h(itw_host, t->itw.p, n * 4);
  });
}
int stwo_hip_twiddles_free(stwo_hip_ctx* h, stwo_hip_twiddles* t) {
  return guarded(h, [&](Ctx& c) {
    if (!t || !h->trees.erase(t)) throw HipError(1, "unknown twiddle tree");
    c.sync();
    delete t;
  });
}
int stwo_hip_interpolate_on_domain(stwo_hip_ctx* h, const stwo_hip_twiddles* t, uint32_t half_coset_initial_index,
                                   uint32_t log_size, const uint32_t* const* evals, uint32_t* const* coeffs,
                                   uint32_t n_cols) {
  return guarded(h, [&](Ctx& c) {
    if (!t || !h->trees.count(const_cast<stwo_hip_twiddles*>(t))) throw HipError(1, "unknown twiddle tree");
    check_tree(t, half_coset_initial_index, log_size);
    if (n_cols == 0) return;
    std::vector<const uint32_t*> s(evals, evals + n_cols);
    std::vector<uint32_t*> d(coeffs, coeffs + n_cols);
    {
      const uint32_t low = plan_fft(c, true, n_cols, log_size, log_size, t->itw2.p, 1u << t->log_size, d.data());
      KScope ks(c, K_INTERPOLATE, (double)n_cols * 8.0 * (double)(1u << log_size));
      ks.units(fft_butterflies(n_cols, log_size, log_size));
      run_fft(c, true, (const uint32_t* const*)c.upload_vec(s), (uint32_t* const*)c.upload_vec(d), n_cols, log_size,
              log_size, t->itw2.p, 1u << t->log_size, half_coset_initial_index, low);
    }
    c.end_stream_ordered();
  });
}
int stwo_hip_evaluate_on_domain(stwo_hip_ctx* h, const stwo_hip_twiddles* t, uint32_t half_coset_initial_index,
                                uint32_t log_size, const uint32_t* const* coeffs, uint32_t log_coeffs,
                                uint32_t* const* evals, uint32_t n_cols) {
  return guarded(h, [&](Ctx& c) {
    if (!t || !h->trees.count(const_cast<stwo_hip_twiddles*>(t))) throw HipError(1, "unknown twiddle tree");
    check_tree(t, half_coset_initial_index, log_size);
    if (log_coeffs < 1 || log_coeffs > log_size) throw HipError(1, "coefficients larger than the domain");
    if (n_cols == 0) return;
    std::vector<const uint32_t*> s(coeffs, coeffs + n_cols);
    std::vector<uint32_t*> d(evals, evals + n_cols);
    {
      const uint32_t low = plan_fft(c, false, n_cols, log_coeffs, log_size, t->tw2.p, 1u << t->log_size, d.data());
      KScope ks(c, K_EVALUATE, (double)n_cols * (4.0 * (double)(1u << log_coeffs) + 4.0 * (double)(1u << log_size)));
      ks.units(fft_butterflies(n_cols, log_size, log_coeffs));
      run_fft(c, false, (const uint32_t* const*)c.upload_vec(s), (uint32_t* const*)c.upload_vec(d), n_cols, log_coeffs,
              log_size, t->tw2.p, 1u << t->log_size, half_coset_initial_index, low);
    }
    c.end_stream_ordered();
  });
}
int stwo_hip_eval_at_point_by_folding(stwo_hip_ctx* h, const uint32_t* evals, uint32_t log_size,
                                      stwo_circle_point_qm31 point, stwo_qm31* out) {
  return guarded(h, [&](Ctx& c) {
    if (!evals || !out) throw HipError(1, "null argument");
    if (log_size < 3 || log_size > 30) throw HipError(3, "log_size must be in [3, 30]");
    check_canonical(point.x);
    check_canonical(point.y);
    c.ensure_twiddles(log_size);
    // folding coefficients (core/poly/utils.rs:45-70): p.y for the circle fold, then p.x and its
    // successive doublings for the line folds
    const cpqm31 p = cp(point);
    const size_t half = (size_t)1 << (log_size - 1);
    // zero coordinates 1..3 of the M31 column, then two ping-pong line layers of 4 coordinates
    DevBuf zero(&c, half * 2, true), la(&c, 4 * half), lb(&c, 4 * (half / 2 ? half / 2 : 1));
    CCols4 src = {{evals, zero.p, zero.p, zero.p}};
    Cols4 a4 = {{la.p, la.p + half, la.p + 2 * half, la.p + 3 * half}};
    launch_fold_circle_into_line_zero(log_size, src, a4, p.y, c.itw.p, c.tree_len(), c.stream);
    qm31 alpha = p.x;
    Cols4 cur = a4;
    DevBuf* bufs[2] = {&la, &lb};
    for (uint32_t ls = log_size - 1, k = 1; ls >= 1; --ls, ++k) {
      const size_t nh = (size_t)1 << (ls - 1);
      DevBuf& nb = *bufs[k & 1];
      Cols4 nxt = {{nb.p, nb.p + nh, nb.p + 2 * nh, nb.p + 3 * nh}};
      launch_fold_line(ls, {{cur.c[0], cur.c[1], cur.c[2], cur.c[3]}}, nxt, alpha, c.itw.p, c.tree_len(), c.stream);
      cur = nxt;
      alpha = qm_double_x(alpha);
    }
    uint32_t v[4];
    for (int k = 0; k < 4; ++k) c.d2h(&v[k], cur.c[k], 4);
    const qm31 r = qm_mul_m31({v[0], v[1], v[2], v[3]}, m31_inv(1u << log_size));
    *out = q(r);
  });
}

// ---- the channel on the device and graph capture ----
static uint32_t* need_state(uint32_t* st) {
  if (!st) throw HipError(1, "null channel state");
  return st;
}
int stwo_hip_channel_init(stwo_hip_ctx* h, uint32_t* state, const uint8_t digest[32], uint32_t n_draws) {
  return guarded(h, [&](Ctx& c) {
    if (!digest) throw HipError(1, "null digest");
    ChanDigest d;
    for (int i = 0; i < 8; ++i)
      d.w[i] = (uint32_t)digest[4 * i] | (uint32_t)digest[4 * i + 1] << 8 | (uint32_t)digest[4 * i + 2] << 16 |
               (uint32_t)digest[4 * i + 3] << 24;
    launch_chan_init(need_state(state), d, n_draws, c.stream);
  }, true, true);
}
int stwo_hip_channel_mix_root(stwo_hip_ctx* h, uint32_t* state, const uint32_t* root) {
  return guarded(h, [&](Ctx& c) {
    if (!root) throw HipError(1, "null root");
    launch_chan_mix_words(need_state(state), root, 8, c.stream);
  }, true, true);
}
int stwo_hip_channel_mix_u32s(stwo_hip_ctx* h, uint32_t* state, const uint32_t* words, uint32_t n) {
  return guarded(h, [&](Ctx& c) {
    if (n && !words) throw HipError(1, "null words");
    launch_chan_mix_words(need_state(state), words, n, c.stream);
  }, true, true);
}
int stwo_hip_channel_draw_secure_felts(stwo_hip_ctx* h, uint32_t* state, uint32_t n, uint32_t* out) {
  return guarded(h, [&](Ctx& c) {
    if (n && !out) throw HipError(1, "null out");
    launch_chan_draw_felts(need_state(state), n, out, c.stream);
  }, true, true);
}
int stwo_hip_channel_draw_queries(stwo_hip_ctx* h, uint32_t* state, uint32_t log_domain, uint32_t n_queries,
                                  uint32_t* out) {
  return guarded(h, [&](Ctx& c) {
    if (n_queries && !out) throw HipError(1, "null out");
    if (log_domain > 31) throw HipError(1, "log_domain above 31");
    launch_chan_draw_queries(need_state(state), log_domain, n_queries, out, c.stream);
  }, true, true);
}
int stwo_hip_channel_grind(stwo_hip_ctx* h, uint32_t* state, uint32_t pow_bits, uint32_t* nonce) {
  return guarded(h, [&](Ctx& c) {
    if (!nonce) throw HipError(1, "null nonce");
    if (pow_bits > 32) throw HipError(3, "pow_bits > 32 is not supported");
    uint32_t* st = need_state(state);
    // the first launch covers 32 times the expected nonce (a miss has probability about e^-32)
    const uint32_t bits = pow_bits + 5 < 12 ? 12 : (pow_bits + 5 > 24 ? 24 : pow_bits + 5);
    uint64_t count = (uint64_t)1 << bits;
    {
      KScope ks(c, K_GRIND, 0.0);
      launch_chan_grind(st, pow_bits, (uint32_t)count, nonce, c.stream);
    }
    if (c.capturing) return;
    unsigned long long found = ~0ull;
    c.d2h(&found, st + 10, 8);
    // beyond the first range: further launches on the device prefix, doubling, as device_grind
    const uint64_t limit = (uint64_t)0x7fffffff << 20;
    for (uint64_t start = count; found == ~0ull; start += count) {
      if (start >= limit) throw HipError(4, "Grind failed to find a solution.");
      count = count < (1u << 24) ? count * 2 : count;
      launch_grind_dev(st + 12, pow_bits, start, (uint32_t)count, reinterpret_cast<unsigned long long*>(st + 10),
                       c.stream);
      c.d2h(&found, st + 10, 8);
    }
    const uint32_t w[2] = {(uint32_t)found, (uint32_t)(found >> 32)};
    HIP_CHECK(hipMemcpyAsync(nonce, w, 8, hipMemcpyHostToDevice, c.stream));
    c.sync();
  }, true, true);
}
int stwo_hip_channel_read(stwo_hip_ctx* h, const uint32_t* state, uint8_t digest[32], uint32_t* n_draws) {
  return guarded(h, [&](Ctx& c) {
    if (!state || !digest || !n_draws) throw HipError(1, "null argument");
    uint32_t w[9];
    c.d2h(w, state, sizeof(w));
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 4; ++j) digest[4 * i + j] = (uint8_t)(w[i] >> (8 * j));
    *n_draws = w[8];
  });
}

struct stwo_hip_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
};
int stwo_hip_graph_begin(stwo_hip_ctx* h) {
  return guarded(h, [&](Ctx& c) {
    if (c.capturing) throw HipError(1, "a capture is already open on this context");
    c.sync();  // nothing of earlier calls is left in flight, on either stream
    c.reset_arena();
    HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    c.capturing = true;
  });
}
int stwo_hip_graph_end(stwo_hip_ctx* h, stwo_hip_graph** out) {
  if (!h || !h->ctx) return 1;
  Ctx& c = *h->ctx;
  if (!c.capturing) {
    h->err = "no capture is open on this context";
    return 1;
  }
  // the capture ends whatever happens next
  c.capturing = false;
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur != c.device) (void)hipSetDevice(c.device);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c.stream, &g);
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    h->err = std::string("hipStreamEndCapture failed: ") + hipGetErrorString(e);
    return 2;
  }
  if (!out) {
    (void)hipGraphDestroy(g);
    h->err = "null out";
    return 1;
  }
  auto* gr = new stwo_hip_graph();
  gr->graph = g;
  gr->device = c.device;
  const hipError_t ie = hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    delete gr;
    h->err = std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ie);
    return 2;
  }
  *out = gr;
  return 0;
}
int stwo_hip_graph_launch(stwo_hip_ctx* h, const stwo_hip_graph* g) {
  return guarded(h, [&](Ctx& c) {
    if (!g || !g->exec) throw HipError(1, "null graph");
    if (g->device != c.device) throw HipError(1, "the graph was captured on another device");
    HIP_CHECK(hipGraphLaunch(g->exec, c.stream));
  });
}
int stwo_hip_graph_free(stwo_hip_ctx* h, stwo_hip_graph* g) {
  return guarded(h, [&](Ctx& c) {
    if (!g) return;
    c.sync();  // a launch of it may still run
    if (g->exec) HIP_CHECK(hipGraphExecDestroy(g->exec));
    if (g->graph) HIP_CHECK(hipGraphDestroy(g->graph));
    delete g;
  });
}

int stwo_hip_measure_int32_issue(stwo_hip_ctx* h, double* lane_instr_per_s) {
  return guarded(h, [&](Ctx& c) {
    if (!lane_instr_per_s) throw HipError(1, "null out");
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
    DevBuf scratch(&c, (size_t)cus * 16);
    c.sync();
    *lane_instr_per_s = measure_int32_issue(scratch.p, c.stream);
  });
}

int stwo_hip_measure_butterfly_rate(stwo_hip_ctx* h, int inverse, double* butterflies_per_s) {
  return guarded(h, [&](Ctx& c) {
    if (!butterflies_per_s) throw HipError(1, "null out");
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
    DevBuf scratch(&c, (size_t)cus * 16);
    c.sync();
    *butterflies_per_s = measure_butterfly_rate(inverse != 0, scratch.p, c.stream);
  });
}

int stwo_hip_measure_compression_rate(stwo_hip_ctx* h, double* compressions_per_s) {
  return guarded(h, [&](Ctx& c) {
    if (!compressions_per_s) throw HipError(1, "null out");
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
    DevBuf scratch(&c, (size_t)cus * 16);
    c.sync();
    *compressions_per_s = measure_compression_rate(scratch.p, c.stream);
  });
}

int stwo_hip_measure_permutation_rate(stwo_hip_ctx* h, double* permutations_per_s) {
  return guarded(h, [&](Ctx& c) {
    if (!permutations_per_s) throw HipError(1, "null out");
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
    DevBuf scratch(&c, (size_t)cus * 16);
    c.sync();
    *permutations_per_s = measure_permutation_rate(scratch.p, c.stream);
  });
}

int stwo_hip_set_fft_low_log(stwo_hip_ctx* h, uint32_t low_log) {
  return guarded(h, [&](Ctx& c) {
    const uint32_t lg = low_log & ~FFT_LOW_LOOP;
    if (low_log != 0 && (lg < 11 || lg > 15 || ((low_log & FFT_LOW_LOOP) && lg != 15) || (low_log & ~(FFT_LOW_LOOP | 0xFFu))))
      throw HipError(1, "low_log must be 0, 11..15, or 15 with the loop flag 0x100");
    c.fft_low_override = low_log;
  });
}
int stwo_hip_fft_low_log(stwo_hip_ctx* h, int inverse, uint32_t n_cols, uint32_t log_in, uint32_t log_n,
                         uint32_t* low_log) {
  return guarded(h, [&](Ctx& c) {
    if (!low_log) throw HipError(1, "null out");
    if (log_in > log_n || log_n > 30 || (inverse && log_in != log_n)) throw HipError(3, "sizes out of range");
    *low_log = 0;
    if (log_n <= 2 || n_cols == 0) return;
    c.ensure_twiddles(log_n);
    *low_log = inverse ? c.fft_low_log(true, n_cols, log_in, log_n, c.itw2.p, c.tree_len())
                       : c.fft_low_log(false, n_cols, log_in, log_n, c.tw2.p, c.tree_len());
    c.sync();
  });
}

int stwo_hip_set_kernel_timing(stwo_hip_ctx* h, int enable) {
  return guarded(h, [&](Ctx& c) { c.sync(); c.timing = enable != 0; });
}
int stwo_hip_kernel_stats(stwo_hip_ctx* h, uint32_t cls, double* ms, double* bytes, uint64_t* launches) {
  return guarded(h, [&](Ctx& c) {
    if (cls >= K_NUM_CLASSES) throw HipError(1, "kernel class out of range");
    c.sync();
    c.collect_timing();
    if (ms) *ms = c.kstats[cls].ms;
    if (bytes) *bytes = c.kstats[cls].bytes;
    if (launches) *launches = c.kstats[cls].launches;
  });
}
int stwo_hip_kernel_units(stwo_hip_ctx* h, uint32_t cls, double* units) {
  return guarded(h, [&](Ctx& c) {
    if (cls >= K_NUM_CLASSES) throw HipError(1, "kernel class out of range");
    if (!units) throw HipError(1, "null out");
    c.sync();
    c.collect_timing();
    *units = c.kstats[cls].units;
  });
}
int stwo_hip_reset_kernel_stats(stwo_hip_ctx* h) {
  return guarded(h, [&](Ctx& c) {
    c.sync();
    c.collect_timing();
    for (auto& k : c.kstats) k = KernelStat();
  });
}

}  // extern "C"
