// Host runtime of libstwo_hip (see runtime.h).
#include "runtime.h"
#include <cstring>
#include <cstdlib>

#include <algorithm>

namespace sh {

DevBuf::DevBuf(Ctx* c, size_t words, bool zero) : n(words), ctx(c) {
  if (words) {
    if (c->bump_depth > 0) {
      p = (uint32_t*)c->bump_alloc(words * 4);
      owned = false;
    } else {
      HIP_CHECK(hipMallocFromPoolAsync((void**)&p, words * 4, c->pool, c->stream));
    }
    if (zero) launch_fill_zero(p, words, c->stream);
  }
}
void* Ctx::bump_alloc(size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  while (bump_chunk < bump_chunks.size()) {
    auto& ch = bump_chunks[bump_chunk];
    if (bump_off + bytes <= ch.second) {
      void* r = ch.first + bump_off;
      bump_off += bytes;
      return r;
    }
    ++bump_chunk;
    bump_off = 0;
  }
  size_t cap = bytes > ((size_t)1 << 30) ? bytes : ((size_t)1 << 30);
  uint8_t* d;
  HIP_CHECK(hipMalloc((void**)&d, cap));
  bump_chunks.push_back({d, cap});
  bump_chunk = bump_chunks.size() - 1;
  bump_off = bytes;
  return d;
}
DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    reset();
    p = o.p; n = o.n; ctx = o.ctx; owned = o.owned;
    o.p = nullptr; o.n = 0;
  }
  return *this;
}
void DevBuf::reset() {
  if (p && owned) (void)hipFreeAsync(p, ctx->stream);
  p = nullptr;
  n = 0;
}
DevBuf::~DevBuf() { reset(); }

Ctx::Ctx(int dev) : device(dev) {
  HIP_CHECK(hipSetDevice(dev));
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  // A pool of its own: blocks freed by this context are never handed to another context's stream,
  // so contexts used side by side do not wait on each other's frees.
  hipMemPoolProps props{};
  props.allocType = hipMemAllocationTypePinned;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = dev;
  HIP_CHECK(hipMemPoolCreate(&pool, &props));
  HIP_CHECK(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, dev));
  uint64_t thr = UINT64_MAX;  // keep freed blocks cached in the pool
  HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
  pinned_cap = 64u << 20;
  HIP_CHECK(hipHostMalloc((void**)&pinned, pinned_cap, hipHostMallocDefault));
  HIP_CHECK(hipHostMalloc((void**)&d2h_pinned, d2h_cap, hipHostMallocDefault));
  HIP_CHECK(hipHostMalloc((void**)&gather_host, gather_host_words * 4, hipHostMallocCoherent | hipHostMallocMapped));
  HIP_CHECK(hipHostGetDevicePointer((void**)&gather_host_dev, gather_host, 0));
  arena_cap = 64u << 20;
  HIP_CHECK(hipMalloc((void**)&dev_arena, arena_cap));
}
Ctx::~Ctx() {
  if (side) (void)hipStreamSynchronize(side);
  (void)hipStreamSynchronize(stream);
  for (auto& p : pending) { event_pool.push_back(p.a); event_pool.push_back(p.b); }
  for (auto e : event_pool) (void)hipEventDestroy(e);
  tw.reset();
  itw.reset();
  tw2.reset();
  itw2.reset();
  domain_points.clear();
  fri_tickets.reset();
  gather_buf.reset();
  for (auto& [bytes, v] : col_cache)
    for (void* p : v) (void)hipFreeAsync(p, stream);
  for (auto& [p, bytes] : col_live) (void)hipFreeAsync(p, stream);
  for (auto* p : overflow) (void)hipFree(p);
  (void)hipStreamSynchronize(stream);
  for (auto& ch : bump_chunks) (void)hipFree(ch.first);
  if (dev_arena) (void)hipFree(dev_arena);
  if (pinned) (void)hipHostFree(pinned);
  if (rb_pinned) (void)hipHostFree(rb_pinned);
  if (d2h_pinned) (void)hipHostFree(d2h_pinned);
  if (gather_host) (void)hipHostFree(gather_host);
  if (table_dev) (void)hipFree(table_dev);
  if (table_pinned) (void)hipHostFree(table_pinned);
  (void)hipEventDestroy(ev_fork);
  (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
  (void)hipStreamSynchronize(stream);
  if (pool) (void)hipMemPoolDestroy(pool);
  (void)hipStreamDestroy(stream);
}

void* Ctx::stage(const void* src, size_t bytes) {
  size_t a = (bytes + 255) & ~(size_t)255;
  if (a == 0) a = 256;
  if (pinned_off + a > pinned_cap || arena_off + a > arena_cap) {
    // rare large upload: queue what is staged, then a dedicated allocation and synchronous copy
    commit_staged();
    uint8_t* d;
    HIP_CHECK(hipMalloc((void**)&d, a));
    HIP_CHECK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    uploads_in_flight = false;
    overflow.push_back(d);
    return d;
  }
  uint8_t* h = pinned + pinned_off;
  uint8_t* d = dev_arena + arena_off;
  std::memcpy(h, src, bytes);
  pinned_off += a;
  arena_off += a;
  return d;
}
void Ctx::commit_staged() {
  if (pinned_off > staged_from) {
    HIP_CHECK(hipMemcpyAsync(dev_arena + staged_from, pinned + staged_from, pinned_off - staged_from,
                             hipMemcpyHostToDevice, stream));
    uploads_in_flight = true;
  }
  staged_from = pinned_off;
}
void* Ctx::col_alloc(size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  void* p = nullptr;
  auto it = col_cache.find(bytes);
  if (it != col_cache.end() && !it->second.empty()) {
    p = it->second.back();
    it->second.pop_back();
    col_cached_bytes -= bytes;
  } else {
    HIP_CHECK(hipMallocFromPoolAsync(&p, bytes, pool, stream));
  }
  col_live[p] = bytes;
  return p;
}
void Ctx::col_free(void* p) {
  auto it = col_live.find(p);
  if (it == col_live.end()) throw HipError(1, "free of an address not allocated by this context");
  const size_t bytes = it->second;
  col_live.erase(it);
  if (col_cached_bytes + bytes > COL_CACHE_CAP) {
    HIP_CHECK(hipFreeAsync(p, stream));
    return;
  }
  col_cache[bytes].push_back(p);  // reused by later work on this context's stream, in order
  col_cached_bytes += bytes;
}
// 64-bit content hash of a table (8 bytes at a time, multiply-xorshift mixing).
static uint64_t table_hash(const void* src, size_t bytes) {
  const uint8_t* p = (const uint8_t*)src;
  uint64_t h = 0x9E3779B97F4A7C15ull ^ bytes;
  size_t i = 0;
  // four independent lanes over 32-byte steps (a column table of 1264 addresses is 10 KB), folded
  // into h, then the remaining words one at a time
  if (bytes >= 64) {
    uint64_t l[4] = {h, h ^ 0x6A09E667F3BCC908ull, h ^ 0xBB67AE8584CAA73Bull, h ^ 0x3C6EF372FE94F82Bull};
    for (; i + 32 <= bytes; i += 32)
      for (int k = 0; k < 4; ++k) {
        uint64_t w;
        std::memcpy(&w, p + i + 8 * k, 8);
        l[k] = (l[k] ^ w) * 0xBF58476D1CE4E5B9ull;
        l[k] ^= l[k] >> 31;
      }
    for (int k = 0; k < 4; ++k) {
      h = (h ^ l[k]) * 0xBF58476D1CE4E5B9ull;
      h ^= h >> 31;
    }
  }
  for (; i + 8 <= bytes; i += 8) {
    uint64_t w;
    std::memcpy(&w, p + i, 8);
    h = (h ^ w) * 0xBF58476D1CE4E5B9ull;
    h ^= h >> 31;
  }
  uint64_t w = 0;
  std::memcpy(&w, p + i, bytes - i);
  h = (h ^ w) * 0x94D049BB133111EBull;
  return h ^ (h >> 29);
}
const void* Ctx::upload_table(const void* src, size_t bytes) {
  const uint64_t key = table_hash(src, bytes);
  // a hit is confirmed against the pinned copy of the table's bytes
  auto range = table_cache.equal_range(key);
  for (auto it = range.first; it != range.second; ++it)
    if (it->second.bytes == bytes && std::memcmp(table_pinned + it->second.off, src, bytes) == 0)
      return table_dev + it->second.off;
  const size_t a = (bytes + 255) & ~(size_t)255;
  if (!table_dev) {
    // STWO_HIP_TABLE_STORE_BYTES (at least 4096) sizes the store; 16 MiB by default
    table_cap = 16u << 20;
    if (const char* e = std::getenv("STWO_HIP_TABLE_STORE_BYTES")) {
      const unsigned long long v = std::strtoull(e, nullptr, 10);
      if (v >= 4096 && v <= (1ull << 30)) table_cap = (size_t)v & ~(size_t)255;
    }
    HIP_CHECK(hipMalloc((void**)&table_dev, table_cap));
    HIP_CHECK(hipHostMalloc((void**)&table_pinned, table_cap, hipHostMallocDefault));
  }
  if (a == 0) return table_dev;  // an empty table is never read
  if (table_off + a > table_cap) {
    // not into the staging arena: reset_arena() rewinds it while the caller may still hold the
    // address. The source is host memory of the caller's, so the copy completes before return.
    void* d = col_alloc(a);
    table_spill.push_back(d);
    HIP_CHECK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, stream));
    sync_main();
    return d;
  }
  // the pinned copy is not overwritten before recycle_tables(), which synchronises first, so the
  // queued transfer may read it at any time
  std::memcpy(table_pinned + table_off, src, bytes);
  HIP_CHECK(hipMemcpyAsync(table_dev + table_off, table_pinned + table_off, bytes, hipMemcpyHostToDevice, stream));
  const void* d = table_dev + table_off;
  table_cache.emplace(key, TableEntry{table_off, bytes});
  table_off += a;
  return d;
}
void Ctx::recycle_tables() {
  if (table_off <= table_cap / 2 && table_spill.empty()) return;
  sync();  // no queued transfer or kernel still reads a table
  for (void* p : table_spill) col_free(p);
  table_spill.clear();
  table_cache.clear();
  table_off = 0;
}
void* Ctx::upload(const void* src, size_t bytes) {
  void* d = stage(src, bytes);
  commit_staged();
  return d;
}
void Ctx::fork() {
  if (!side) HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));  // default priority
  commit_staged();
  side_busy = true;
  HIP_CHECK(hipEventRecord(ev_fork, stream));
  HIP_CHECK(hipStreamWaitEvent(side, ev_fork, 0));
}
void Ctx::join() {
  if (!side) return;
  HIP_CHECK(hipEventRecord(ev_join, side));
  HIP_CHECK(hipStreamWaitEvent(stream, ev_join, 0));
}
void Ctx::sync_main() {
  commit_staged();
  HIP_CHECK(hipStreamSynchronize(stream));
  uploads_in_flight = false;
}
void Ctx::sync() {
  commit_staged();
  if (side) HIP_CHECK(hipStreamSynchronize(side));
  side_busy = false;
  HIP_CHECK(hipStreamSynchronize(stream));
  uploads_in_flight = false;
  // event pairs are folded into kstats in bulk (each elapsed-time query costs host time)
  if (pending.size() >= 16384) collect_timing();
}
hipEvent_t Ctx::get_event() {
  if (!event_pool.empty()) { hipEvent_t e = event_pool.back(); event_pool.pop_back(); return e; }
  hipEvent_t e;
  HIP_CHECK(hipEventCreate(&e));
  return e;
}
void Ctx::collect_timing() {
  for (auto& p : pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      kstats[p.cls].ms += ms;
      kstats[p.cls].bytes += p.bytes;
      kstats[p.cls].units += p.units;
      kstats[p.cls].launches += 1;
    }
    event_pool.push_back(p.a);
    event_pool.push_back(p.b);
  }
  pending.clear();
}
void Ctx::end_stream_ordered() {
  if (capturing) return;
  if (arena_off > arena_cap / 2 || pinned_off > pinned_cap / 2) sync();
  reset_arena();
}
void Ctx::reset_arena() {
  // staging still read by queued copies (no sync since the last upload) is kept until a later call
  // (and, while forked work may still read the arena, until a full sync)
  if (uploads_in_flight || side_busy) return;
  commit_staged();
  if (uploads_in_flight) return;
  for (auto* p : overflow) (void)hipFree(p);
  overflow.clear();
  pinned_off = 0;
  arena_off = 0;
  staged_from = 0;
}
void* Ctx::readback_buf(size_t bytes) {
  if (bytes > rb_cap) {
    sync_main();  // no read-back into the old buffer is still queued
    if (rb_pinned) HIP_CHECK(hipHostFree(rb_pinned));
    rb_pinned = nullptr;
    rb_cap = std::max(bytes, (size_t)1 << 20);
    HIP_CHECK(hipHostMalloc((void**)&rb_pinned, rb_cap, hipHostMallocDefault));
  }
  return rb_pinned;
}
void Ctx::d2h(void* dst, const void* src, size_t bytes) {
  // through pinned memory: a read-back into pageable memory goes through the runtime's own
  // staging, which costs tens of microseconds more per call
  if (bytes <= d2h_cap) {
    HIP_CHECK(hipMemcpyAsync(d2h_pinned, src, bytes, hipMemcpyDeviceToHost, stream));
    sync_main();
    std::memcpy(dst, d2h_pinned, bytes);
    return;
  }
  HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
  sync_main();
}

// Twiddle tree of the coset (init_index, log_size) (prover/backend/cpu/circle.rs:274-291): layer j =
// x-coordinates of the first half of the coset doubled j times, bit-reversed; pad 1. Built on the
// device (one lane per entry: the point by the generator's powers, the inverse by Fermat), so
// that a tree of 2^29 entries costs milliseconds instead of a serial host loop.
void build_twiddle_tree(Ctx& c, uint32_t init_index, uint32_t log_size, DevBuf& tw, DevBuf& itw, DevBuf& tw2,
                        DevBuf& itw2) {
  if (log_size > 30) throw HipError(3, "coset log size must be at most 30");
  PersistentAllocScope keep(c);
  const size_t len = (size_t)1 << log_size;
  tw = DevBuf(&c, len);
  itw = DevBuf(&c, len);
  tw2 = DevBuf(&c, len);
  itw2 = DevBuf(&c, len);
  DevBuf flag(&c, 1, true);
  launch_twiddle_tree(init_index, log_size, tw.p, itw.p, tw2.p, itw2.p, flag.p, c.stream);
  uint32_t zero = 0;
  c.d2h(&zero, flag.p, 4);
  if (zero) throw HipError(1, "the coset has a point with x = 0, whose twiddle has no inverse");
}

// The canonic tree (root coset: the half coset of CanonicCoset(log_domain)) serves every smaller
// canonic domain: line twiddles are read from the end of the buffer (core/poly/utils.rs:98-114).
void Ctx::ensure_twiddles(uint32_t log_domain) {
  if (log_domain <= tw_log) return;
  const uint32_t R = log_domain - 1;  // root coset log size
  build_twiddle_tree(*this, HCoset::half_odds(R).initial_index, R, tw, itw, tw2, itw2);
  tw_log = log_domain;
}

cpm31 canonic_domain_at(uint32_t log, size_t i) {
  HCoset hc = HCoset::half_odds(log - 1);
  size_t h = (size_t)1 << (log - 1);
  if (i < h) return hc.at(i);
  return index_to_point((0x80000000u - hc.index_at(i - h)) & 0x7fffffffu);
}

uint32_t* Ctx::fri_ticket_words(size_t n) {
  if (fri_tickets.n < n) {
    PersistentAllocScope keep(*this);
    fri_tickets = DevBuf(this, n < 64 ? 64 : n, true);
  }
  return fri_tickets.p;
}

uint32_t* Ctx::gather_out(size_t n) {
  if (gather_buf.n < n) {
    PersistentAllocScope keep(*this);
    gather_buf = DevBuf(this, n < ((size_t)1 << 14) ? ((size_t)1 << 14) : n);
  }
  return gather_buf.p;
}

const uint32_t* Ctx::domain_xy(uint32_t log) {
  auto it = domain_points.find(log);
  if (it != domain_points.end()) return it->second.p;
  PersistentAllocScope keep(*this);
  size_t n = (size_t)1 << log, h = n / 2;
  std::vector<uint32_t> xy(2 * n);
  HCoset hc = HCoset::half_odds(log - 1);
  cpm31 p = index_to_point(hc.initial_index), st = index_to_point(hc.step_size);
  for (size_t j = 0; j < h; ++j) {
    size_t r0 = bitrev((uint32_t)j, log), r1 = bitrev((uint32_t)(j + h), log);
    xy[2 * r0] = p.x; xy[2 * r0 + 1] = p.y;
    xy[2 * r1] = p.x; xy[2 * r1 + 1] = m31_neg(p.y);  // conjugate half
    p = cp_add(p, st);
  }
  DevBuf b(this, 2 * n);
  HIP_CHECK(hipMemcpyAsync(b.p, xy.data(), 2 * n * 4, hipMemcpyHostToDevice, stream));
  sync();
  const uint32_t* r = b.p;
  domain_points.emplace(log, std::move(b));
  return r;
}

uint32_t coset_vanishing_m31(uint32_t coset_log, cpm31 p) {
  HCoset c = HCoset::odds(coset_log);
  cpm31 init = index_to_point(c.initial_index);
  cpm31 half = index_to_point(c.step_size >> 1);
  p = cp_add(cp_add(p, cp_conj(init)), half);
  uint32_t x = p.x;
  for (uint32_t i = 1; i < coset_log; ++i) x = m31_double_x(x);
  return x;
}
qm31 coset_vanishing_qm31(uint32_t coset_log, cpqm31 p) {
  HCoset c = HCoset::odds(coset_log);
  cpm31 init = index_to_point(c.initial_index);
  cpm31 half = index_to_point(c.step_size >> 1);
  p = cpq_add(cpq_add(p, cpq_from(cp_conj(init))), cpq_from(half));
  qm31 x = p.x;
  for (uint32_t i = 1; i < coset_log; ++i) x = qm_double_x(x);
  return x;
}
cpqm31 cpq_mul_signed(cpm31 step, long off) {
  cpm31 base = off > 0 ? step : cp_conj(step);
  unsigned long k = off > 0 ? off : -off;
  cpm31 r = {1, 0};
  while (k) { if (k & 1) r = cp_add(r, base); base = cp_double(base); k >>= 1; }
  return cpq_from(r);
}

// ---- Blake2s (host) -----------------------------------------------------------------------
static void words_from_bytes(const uint8_t* b, uint32_t m[16]) {
  for (int i = 0; i < 16; ++i)
    m[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) | ((uint32_t)b[4 * i + 3] << 24);
}
void Blake2sHost::update(const void* data, size_t len) {
  const uint8_t* p = (const uint8_t*)data;
  // whole blocks straight from the input while more than one block remains (the last block is
  // held back: it may be the final one)
  if (buflen == 64 && len) {
    t += 64;
    uint32_t m[16];
    words_from_bytes(buf, m);
    blake2s_compress(h, m, (uint32_t)t, false);
    buflen = 0;
  }
  if (buflen == 0)
    while (len > 64) {
      t += 64;
      uint32_t m[16];
      words_from_bytes(p, m);
      blake2s_compress(h, m, (uint32_t)t, false);
      p += 64;
      len -= 64;
    }
  while (len) {
    if (buflen == 64) {
      t += 64;
      uint32_t m[16];
      words_from_bytes(buf, m);
      blake2s_compress(h, m, (uint32_t)t, false);
      buflen = 0;
    }
    size_t k = std::min(len, 64 - buflen);
    memcpy(buf + buflen, p, k);
    buflen += k;
    p += k;
    len -= k;
  }
}
Hash Blake2sHost::finalize() {
  t += buflen;
  memset(buf + buflen, 0, 64 - buflen);
  uint32_t m[16];
  words_from_bytes(buf, m);
  blake2s_compress(h, m, (uint32_t)t, true);
  Hash out;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(h[i] >> (8 * j));
  return out;
}
Hash blake2s_hash(const void* data, size_t len) { Blake2sHost b; b.update(data, len); return b.finalize(); }

void Channel::mix_root(const Hash& root) {
  Blake2sHost b; b.update(digest.data(), 32); b.update(root.data(), 32); update_digest(b.finalize());
}
void Channel::mix_u32s(const uint32_t* data, size_t n) {
  Blake2sHost b; b.update(digest.data(), 32);
  for (size_t i = 0; i < n; ++i) b.update_u32(data[i]);
  update_digest(b.finalize());
}
void Channel::mix_felts(const std::vector<qm31>& f) {
  // the felts' words in order, little endian (the host's order): the vector's bytes as they are
  static_assert(sizeof(qm31) == 16, "qm31 is 4 packed words");
  Blake2sHost b;
  b.update(digest.data(), 32);
  b.update(f.data(), f.size() * sizeof(qm31));
  update_digest(b.finalize());
}
std::array<uint32_t, 8> Channel::draw_u32s() {
  uint8_t in[37];
  memcpy(in, digest.data(), 32);
  for (int i = 0; i < 4; ++i) in[32 + i] = (uint8_t)(n_draws >> (8 * i));
  in[36] = 0;
  n_draws++;
  Hash h = blake2s_hash(in, 37);
  std::array<uint32_t, 8> r;
  for (int i = 0; i < 8; ++i) r[i] = (uint32_t)h[4 * i] | ((uint32_t)h[4 * i + 1] << 8) | ((uint32_t)h[4 * i + 2] << 16) | ((uint32_t)h[4 * i + 3] << 24);
  return r;
}
std::array<uint32_t, 8> Channel::draw_base_felts() {
  while (true) {
    auto u = draw_u32s();
    bool ok = true;
    for (auto x : u) ok &= x < 2 * P;
    if (ok) { for (auto& x : u) x = x >= P ? x - P : x; return u; }
  }
}
qm31 Channel::draw_secure_felt() { auto f = draw_base_felts(); return {f[0], f[1], f[2], f[3]}; }
std::vector<qm31> Channel::draw_secure_felts(size_t n) {
  std::vector<qm31> r;
  while (r.size() < n) {
    auto f = draw_base_felts();
    r.push_back({f[0], f[1], f[2], f[3]});
    if (r.size() < n) r.push_back({f[4], f[5], f[6], f[7]});
  }
  return r;
}
Hash Channel::pow_prefix(uint32_t pow_bits) const {
  Blake2sHost b;
  b.update_u32(0x12345678u);
  uint8_t z[12] = {0};
  b.update(z, 12);
  b.update(digest.data(), 32);
  b.update_u32(pow_bits);
  return b.finalize();
}

}  // namespace sh
