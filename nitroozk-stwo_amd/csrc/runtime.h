// Host runtime of libstwo_hip: device context, stream-ordered memory, pinned staging,
// twiddle / domain caches, and the host-side Fiat-Shamir channel.
#pragma once
#include <hip/hip_runtime.h>
#include <array>
#include <unordered_map>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include "blake2s.h"
#include "field.h"
#include "kernels.h"

namespace sh {

struct HipError : std::runtime_error {
  int code;
  HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
#define HIP_CHECK(x)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) throw HipError(2, std::string(#x " failed: ") + hipGetErrorString(e_)); \
  } while (0)

struct Ctx;

// The directory of files kept across processes (compiled constraint kernels, FFT plans):
// STWO_HIP_KERNEL_CACHE, else ~/.cache/stwo_hip; created on first use. Empty if neither is set.
std::string cache_directory();

// Owning device buffer of u32 words, freed stream-ordered.
struct DevBuf {
  uint32_t* p = nullptr;
  size_t n = 0;
  Ctx* ctx = nullptr;
  bool owned = true;  // false for proof-arena memory, released all at once at the end of a proof
  DevBuf() = default;
  DevBuf(Ctx* c, size_t words, bool zero = false);
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
  DevBuf& operator=(DevBuf&& o) noexcept;
  ~DevBuf();
  void reset();
  // A non-owning view of device memory owned elsewhere (a proof arena block).
  static DevBuf view(Ctx* c, uint32_t* p, size_t words) {
    DevBuf b;
    b.p = p;
    b.n = words;
    b.ctx = c;
    b.owned = false;
    return b;
  }
};

using Hash = std::array<uint8_t, 32>;

struct StageTimes {
  double ms[16] = {0};
};

// Kernel classes timed with HIP events when timing is enabled (stwo_hip_set_kernel_timing).
enum KernelClass {
  K_INTERPOLATE = 0, K_EVALUATE, K_MERKLE, K_COMPOSITION, K_NUMERATORS, K_QUOTIENTS, K_FRI_FOLD, K_OODS,
  K_LOGUP, K_GRIND, K_MISC, K_FRI_TAIL, K_NUM_CLASSES
};
struct KernelStat {
  double ms = 0;
  double bytes = 0;  // algorithmic HBM bytes
  double units = 0;  // algorithmic work units (butterflies for the transform classes)
  uint64_t launches = 0;
};

struct Ctx {
  int device = 0;
  bool timing = false;
  // The stream is being captured into a graph (stwo_hip_graph_begin .. _end): only entry points
  // that launch on the stream with no synchronisation, allocation or staged upload may run, and
  // kernel timing records no events.
  bool capturing = false;
  struct Pending { int cls; double bytes, units; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  KernelStat kstats[K_NUM_CLASSES];
  hipEvent_t get_event();
  void collect_timing();  // after a sync: folds finished event pairs into kstats
  hipStream_t stream = nullptr;
  // Second stream for independent work that overlaps the main stream: fork() orders it after
  // everything queued on the main stream so far, join() orders the main stream after it.
  hipStream_t side = nullptr;  // created on the first fork()
  hipMemPool_t pool = nullptr;  // this context's device memory pool
  // Column allocations handed to callers (stwo_hip_alloc): released blocks are kept per size and
  // handed out again in stream order, so a serving loop that allocates and frees the same trace
  // shapes makes no allocator calls after its first request. Cached bytes are capped.
  void* col_alloc(size_t bytes);
  void col_free(void* p);
  std::unordered_map<void*, size_t> col_live;
  std::unordered_map<size_t, std::vector<void*>> col_cache;
  size_t col_cached_bytes = 0;
  static constexpr size_t COL_CACHE_CAP = (size_t)32 << 30;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void fork();
  void join();
  bool side_busy = false;  // work forked to the side stream since the last full sync()
  std::string last_error;
  // twiddle tree for the half coset of CanonicCoset(tw_log): 2^(tw_log-1) words each
  uint32_t tw_log = 0;
  DevBuf tw, itw;
  DevBuf tw2, itw2;  // the same trees with every entry doubled (2t < 2^32), read by the FFT kernels
  std::map<uint32_t, DevBuf> domain_points;  // bit-reversed (x, y) of CanonicCoset(log)
  // Ticket counters of the fused FRI layer launches, one per layer: zeroed once here, and every
  // launch leaves its counter at zero again.
  DevBuf fri_tickets;
  uint32_t* fri_ticket_words(size_t n);
  // Device output of Column::batch_at_multi gathers, reused call after call (grown on demand).
  DevBuf gather_buf;
  uint32_t* gather_out(size_t n);
  // small-upload arena (device) and pinned staging (host)
  uint8_t* pinned = nullptr;
  size_t pinned_cap = 0, pinned_off = 0;
  uint8_t* dev_arena = nullptr;
  size_t arena_cap = 0, arena_off = 0;
  bool uploads_in_flight = false;
  size_t staged_from = 0;  // start of staged bytes whose transfer is not queued yet
  std::vector<uint8_t*> overflow;  // arena overflow allocations, freed at reset
  // proof arena: HBM chunks kept across proofs; a proof bump-allocates its buffers from them
  std::vector<std::pair<uint8_t*, size_t>> bump_chunks;
  size_t bump_chunk = 0, bump_off = 0;
  int bump_depth = 0;
  void* bump_alloc(size_t bytes);
  void bump_reset() { bump_chunk = 0; bump_off = 0; }

  explicit Ctx(int dev);
  ~Ctx();
  // Copies host bytes to device through pinned staging; valid on the stream until reset_arena().
  void* upload(const void* src, size_t bytes);
  template <class T> const T* upload_vec(const std::vector<T>& v) { return (const T*)upload(v.data(), v.size() * sizeof(T)); }
  // Staging without the copy: the bytes land in pinned memory and the returned device address is
  // valid once commit_staged() (or the next upload) has queued the transfer.
  void* stage(const void* src, size_t bytes);
  void commit_staged();
  template <class T> const T* stage_vec(const std::vector<T>& v) { return (const T*)stage(v.data(), v.size() * sizeof(T)); }
  // Tables that repeat from call to call (column address tables, per-shape constants) are kept in
  // device memory keyed by their bytes: a table seen before costs no transfer. A table's address
  // stays valid until the next recycle_tables(), across reset_arena() (the prover lays tables out
  // before a commitment's root is read and launches on them after). Once the store is full, a
  // table goes to a buffer of its own, with the same lifetime.
  const void* upload_table(const void* src, size_t bytes);
  template <class T> const T* table_vec(const std::vector<T>& v) { return (const T*)upload_table(v.data(), v.size() * sizeof(T)); }
  struct TableEntry { size_t off, bytes; };
  std::unordered_multimap<uint64_t, TableEntry> table_cache;  // content hash -> table in table_dev
  uint8_t* table_dev = nullptr;
  uint8_t* table_pinned = nullptr;
  size_t table_cap = 0, table_off = 0;
  std::vector<void*> table_spill;  // tables that did not fit the store
  // Empties the store (after a full sync) once it is half full or has spilled. Called between
  // top-level calls only, where no host code holds a table address.
  void recycle_tables();
  void sync();       // both streams
  void sync_main();  // the main stream only
  void reset_arena();  // rewinds the staging arena; a no-op while uploads are queued
  // The end of a stream-ordered entry point: nothing waits for the device, except that staging
  // areas more than half full are drained (a sync) before the arena is rewound. A no-op while
  // capturing.
  void end_stream_ordered();
  void ensure_twiddles(uint32_t log_domain);
  // Circle-FFT launch plans (fft_plan.cpp): the low-stage tile (2^S values per workgroup) for a
  // transform shape, as launch_fft's low_log; 0 for transforms of fewer than 11 stages. background:
  // the transform runs beside critical-path kernels on another stream, and its plan leaves out the
  // looping kernel (whose workgroups would hold every CU until the transform ends).
  uint32_t fft_low_log(bool inverse, uint32_t n_cols, uint32_t log_in, uint32_t log_n, const uint32_t* tree2,
                       uint32_t tree_len, bool background = false);
  uint32_t fft_low_override = 0;  // a fixed tile for every shape (stwo_hip_set_fft_low_log); 0 = planned
  // The plan for output columns `dst` (host copies of their addresses): without the looping kernel
  // unless every column is 16-byte aligned.
  static uint32_t fft_plan_for(uint32_t plan, uint32_t* const* dst, size_t n_cols);
  int cu_count = 0;
  const uint32_t* domain_xy(uint32_t log);
  uint32_t tree_len() const { return 1u << (tw_log - 1); }
  void d2h(void* dst, const void* src, size_t bytes);  // synchronous
  // Coherent pinned host memory that small gathers write directly from the device (zero copy; the
  // kernel ends with a system-scope fence, the host reads after the stream synchronises).
  uint32_t* gather_host = nullptr;
  uint32_t* gather_host_dev = nullptr;  // the device's address of the same memory
  static constexpr size_t gather_host_words = 16384;
  uint8_t* d2h_pinned = nullptr;  // read-backs of up to d2h_cap bytes land here first
  static constexpr size_t d2h_cap = 4u << 20;
  // Pinned host memory for asynchronous read-backs on the main stream (kept and grown as needed;
  // the contents are valid after the next sync).
  uint8_t* rb_pinned = nullptr;
  size_t rb_cap = 0;
  void* readback_buf(size_t bytes);
};

// While alive, DevBufs of the context come from the proof arena; the arena is rewound (after a
// stream sync) when the outermost scope ends.
struct ProofArenaScope {
  Ctx& c;
  explicit ProofArenaScope(Ctx& ctx) : c(ctx) { c.bump_depth++; }
  ~ProofArenaScope() {
    if (--c.bump_depth == 0) {
      (void)hipStreamSynchronize(c.stream);
      if (c.side) (void)hipStreamSynchronize(c.side);
      c.bump_reset();
    }
  }
};
// Suspends the proof arena (for cached buffers that outlive a proof).
struct PersistentAllocScope {
  Ctx& c;
  int saved;
  explicit PersistentAllocScope(Ctx& ctx) : c(ctx), saved(ctx.bump_depth) { c.bump_depth = 0; }
  ~PersistentAllocScope() { c.bump_depth = saved; }
};

// Times the kernels launched during its lifetime on a stream of the context (class, algorithmic
// bytes); the main stream unless another is given.
struct KScope {
  Ctx& c;
  int cls;
  double bytes, n_units = 0;
  hipStream_t st;
  hipEvent_t a = nullptr;
  KScope(Ctx& ctx, int k, double b, hipStream_t s = nullptr) : c(ctx), cls(k), bytes(b), st(s ? s : ctx.stream) {
    if (c.timing && !c.capturing) { a = c.get_event(); (void)hipEventRecord(a, st); }
  }
  // The launch group's algorithmic work units (KernelStat::units).
  KScope& units(double u) { n_units = u; return *this; }
  ~KScope() {
    if (c.timing && a) {
      hipEvent_t b = c.get_event();
      (void)hipEventRecord(b, st);
      c.pending.push_back({cls, bytes, n_units, a, b});
    }
  }
};
// Blake2s compressions of one Merkle node over its two children (when prev) and n_cols column words:
// one per 64-byte block of the message, one for an empty message.
inline double blake2s_node_compressions(bool prev, size_t n_cols) {
  const size_t bytes = (prev ? 64 : 0) + 4 * n_cols;
  return bytes == 0 ? 1.0 : (double)((bytes + 63) / 64);
}
// Butterflies of a circle FFT over n_cols columns of 2^log_n values that runs `stages` stages
// (interpolation: all log_n; evaluation from 2^k coefficients: k, the top stages only copy).
inline double fft_butterflies(size_t n_cols, uint32_t log_n, uint32_t stages) {
  if (log_n == 0 || stages == 0) return 0.0;
  return (double)n_cols * (double)((size_t)1 << (log_n - 1)) * (double)stages;
}

// Several host arrays sent to the device in one transfer: add() stages each array and returns its
// device address; commit() queues the copy and must precede the first launch that reads them.
struct UploadBatch {
  Ctx& c;
  bool committed = false;
  explicit UploadBatch(Ctx& ctx) : c(ctx) { c.commit_staged(); }
  template <class T> const T* add(const std::vector<T>& v) { return (const T*)c.stage(v.data(), v.size() * sizeof(T)); }
  const void* add_bytes(const void* p, size_t n) { return c.stage(p, n); }
  void commit() { c.commit_staged(); committed = true; }
  ~UploadBatch() { if (!committed) c.commit_staged(); }
};

// The twiddle tree of Coset::new(init_index, log_size), plain and doubled (2^log_size words each).
void build_twiddle_tree(Ctx& c, uint32_t init_index, uint32_t log_size, DevBuf& tw, DevBuf& itw, DevBuf& tw2,
                        DevBuf& itw2);

// ---- host Blake2s and channel ------------------------------------------------------------
struct Blake2sHost {
  uint32_t h[8];
  uint8_t buf[64];
  size_t buflen = 0;
  uint64_t t = 0;
  Blake2sHost() { blake2s_init(h); }
  void update(const void* data, size_t len);
  void update_u32(uint32_t w) { uint8_t b[4] = {(uint8_t)w, (uint8_t)(w >> 8), (uint8_t)(w >> 16), (uint8_t)(w >> 24)}; update(b, 4); }
  Hash finalize();
};
Hash blake2s_hash(const void* data, size_t len);

// Blake2sChannel (crates/stwo/src/core/channel/blake2s.rs:19-140).
struct Channel {
  Hash digest{};
  uint32_t n_draws = 0;
  void update_digest(const Hash& d) { digest = d; n_draws = 0; }
  void mix_root(const Hash& root);
  void mix_u32s(const uint32_t* data, size_t n);
  void mix_u64(uint64_t v) { uint32_t d[2] = {(uint32_t)v, (uint32_t)(v >> 32)}; mix_u32s(d, 2); }
  void mix_felts(const std::vector<qm31>& f);
  std::array<uint32_t, 8> draw_u32s();
  std::array<uint32_t, 8> draw_base_felts();
  qm31 draw_secure_felt();
  std::vector<qm31> draw_secure_felts(size_t n);
  Hash pow_prefix(uint32_t pow_bits) const;
};

// Poseidon252Channel (core/channel/poseidon252.rs:17-156): the digest is a felt252 below the Stark
// prime, kept as its 32 little-endian bytes; a draw permutes (digest, n_draws, 3). The part of
// Channel's interface that prove() uses (mix_u32s is not), except that a draw yields 7 words and the
// proof of work is checked here.
struct P252Channel {
  Hash digest{};
  uint32_t n_draws = 0;
  void update_digest(const Hash& d) { digest = d; n_draws = 0; }
  void mix_root(const Hash& root);
  void mix_u64(uint64_t v);
  void mix_felts(const std::vector<qm31>& f);
  Hash draw_felt252();
  std::array<uint32_t, 7> draw_u32s();
  std::array<uint32_t, 8> draw_base_felts();
  qm31 draw_secure_felt();
  std::vector<qm31> draw_secure_felts(size_t n);
  std::array<uint32_t, 8> digest_words() const;
  bool verify_pow_nonce(uint32_t pow_bits, uint64_t nonce) const;
};
// Poseidon252MerkleHasher::hash_node (core/vcs/poseidon252_merkle.rs:18-41) on the host: children
// (left may be null: a leaf) and n column values, 8 per felt252.
Hash p252_hash_node(const Hash* left, const Hash* right, const uint32_t* vals, size_t n);
bool p252_hash_canonical(const Hash& h);

// ---- host circle-domain helpers -------------------------------------------------------------
inline uint32_t subgroup_gen(uint32_t log) { return 1u << (31 - log); }
struct HCoset {
  uint32_t initial_index, step_size, log_size;
  static HCoset make(uint32_t init, uint32_t log) { return {init & 0x7fffffffu, subgroup_gen(log), log}; }
  static HCoset odds(uint32_t log) { return make(subgroup_gen(log + 1), log); }
  static HCoset half_odds(uint32_t log) { return make(subgroup_gen(log + 2), log); }
  uint32_t index_at(size_t i) const { return (uint32_t)(((uint64_t)initial_index + (uint64_t)step_size * i) & 0x7fffffffu); }
  cpm31 at(size_t i) const { return index_to_point(index_at(i)); }
  HCoset doubled() const { return {(initial_index * 2) & 0x7fffffffu, (step_size * 2) & 0x7fffffffu, log_size ? log_size - 1 : 0}; }
};
// Point i of the circle domain of CanonicCoset(log) (natural order).
cpm31 canonic_domain_at(uint32_t log, size_t i);
// Vanishing polynomial of CanonicCoset(log) (core/constraints.rs coset_vanishing).
uint32_t coset_vanishing_m31(uint32_t coset_log, cpm31 p);
qm31 coset_vanishing_qm31(uint32_t coset_log, cpqm31 p);
cpqm31 cpq_mul_signed(cpm31 step, long off);

}  // namespace sh
