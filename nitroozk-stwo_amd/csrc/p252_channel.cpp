// The Poseidon252 Merkle channel on the host (paths relative to crates/stwo/src):
//   Poseidon252Channel        core/channel/poseidon252.rs:17-156
//   Poseidon252MerkleHasher   core/vcs/poseidon252_merkle.rs:18-71 (hash_node, 8 M31 per felt252)
//   Poseidon252MerkleChannel  core/vcs/poseidon252_merkle.rs:78-85 (mix_root = poseidon_hash)
// A felt252 travels as its 32 little-endian bytes (Hash), the same 8 u32 words the device layers
// hold; arithmetic runs in Montgomery form (felt252.h) and every value leaves in normal form.
#include "felt252.h"
#include "runtime.h"

namespace sh {
namespace {

using f252::F;
using f252::M;

F felt_of(const Hash& h) {
  F f;
  for (int i = 0; i < 8; ++i)
    f.w[i] = (uint32_t)h[4 * i] | (uint32_t)h[4 * i + 1] << 8 | (uint32_t)h[4 * i + 2] << 16 |
             (uint32_t)h[4 * i + 3] << 24;
  return f;
}
Hash hash_of(const F& f) {
  Hash h;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 4; ++j) h[4 * i + j] = (uint8_t)(f.w[i] >> (8 * j));
  return h;
}
M mont(const Hash& h) { return f252::to_mont(felt_of(h)); }
Hash normal(const M& m) { return hash_of(f252::from_mont(m)); }
M mont_u64(uint64_t v) { return f252::to_mont(f252::from_u64(v)); }

// f << 31 | x for f below 2^221 and x below 2^31 (no reduction: the result stays below 2^252)
void shl31_or(F& f, uint32_t x) {
  for (int i = 7; i > 0; --i) f.w[i] = (f.w[i] << 31) | (f.w[i - 1] >> 1);
  f.w[0] = (f.w[0] << 31) | x;
}
// bits [at, at + n) of a normal-form felt, n <= 32
uint32_t bits_at(const F& f, uint32_t at, uint32_t n) {
  const uint32_t w = at / 32, s = at % 32;
  uint64_t v = f.w[w];
  if (w + 1 < 8) v |= (uint64_t)f.w[w + 1] << 32;
  return (uint32_t)((v >> s) & ((n == 32 ? 0 : (1ull << n)) - 1));
}

}  // namespace

bool p252_hash_canonical(const Hash& h) { return !f252::geq_p(felt_of(h)); }

Hash p252_hash_node(const Hash* left, const Hash* right, const uint32_t* vals, size_t n) {
  if (left && n == 0) return normal(f252::hash2(mont(*left), mont(*right)));
  f252::Sponge sp;
  if (left) {
    sp.absorb(mont(*left));
    sp.absorb(mont(*right));
  }
  for (size_t c = 0; c < n; c += 8) {
    const int k = n - c < 8 ? (int)(n - c) : 8;
    uint32_t x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < k; ++j) x[j] = vals[c + j];
    sp.absorb(f252::to_mont(f252::felt_from_m31s(x, k)));
  }
  return normal(sp.finish());
}

void P252Channel::mix_root(const Hash& root) { update_digest(normal(f252::hash2(mont(digest), mont(root)))); }

void P252Channel::mix_u64(uint64_t v) { update_digest(normal(f252::hash2(mont(digest), mont_u64(v)))); }

void P252Channel::mix_felts(const std::vector<qm31>& f) {
  // pairs of secure felts as one felt252: 1, then their 8 (or 4) coordinates, 31 bits each
  f252::Sponge sp;
  sp.absorb(mont(digest));
  for (size_t i = 0; i < f.size(); i += 2) {
    F x = f252::from_u64(1);
    for (size_t j = i; j < i + 2 && j < f.size(); ++j) {
      shl31_or(x, f[j].a);
      shl31_or(x, f[j].b);
      shl31_or(x, f[j].c);
      shl31_or(x, f[j].d);
    }
    sp.absorb(f252::to_mont(x));
  }
  update_digest(normal(sp.finish()));
}

Hash P252Channel::draw_felt252() {
  M s[3] = {mont(digest), mont_u64(n_draws), f252::THREE};
  f252::permute(s);
  ++n_draws;
  return normal(s[0]);
}

std::array<uint32_t, 7> P252Channel::draw_u32s() {
  const F f = felt_of(draw_felt252());
  std::array<uint32_t, 7> r;
  for (int i = 0; i < 7; ++i) r[i] = f.w[i];
  return r;
}

std::array<uint32_t, 8> P252Channel::draw_base_felts() {
  const F f = felt_of(draw_felt252());
  std::array<uint32_t, 8> r;
  for (int i = 0; i < 8; ++i) {
    const uint32_t x = bits_at(f, 31 * (uint32_t)i, 31);
    r[i] = x == 0x7fffffffu ? 0u : x;  // BaseField::reduce of a 31-bit value
  }
  return r;
}

qm31 P252Channel::draw_secure_felt() {
  const auto f = draw_base_felts();
  return {f[0], f[1], f[2], f[3]};
}

std::vector<qm31> P252Channel::draw_secure_felts(size_t n) {
  std::vector<qm31> out;
  out.reserve(n);
  while (out.size() < n) {
    const auto f = draw_base_felts();
    out.push_back({f[0], f[1], f[2], f[3]});
    if (out.size() < n) out.push_back({f[4], f[5], f[6], f[7]});
  }
  return out;
}

std::array<uint32_t, 8> P252Channel::digest_words() const {
  const F f = felt_of(digest);
  std::array<uint32_t, 8> r;
  for (int i = 0; i < 8; ++i) r[i] = f.w[i];
  return r;
}

bool P252Channel::verify_pow_nonce(uint32_t pow_bits, uint64_t nonce) const {
  f252::Sponge sp;
  sp.absorb(mont_u64(0x12345678u));  // POW_PREFIX
  sp.absorb(mont(digest));
  sp.absorb(mont_u64(pow_bits));
  const F h = f252::from_mont(f252::hash2(sp.finish(), mont_u64(nonce)));
  // trailing zeros of the low 128 bits (u128 of the big-endian bytes 16..32)
  uint32_t tz = 0;
  for (int i = 0; i < 4; ++i) {
    if (h.w[i] == 0) {
      tz += 32;
      continue;
    }
    tz += (uint32_t)__builtin_ctz(h.w[i]);
    break;
  }
  return tz >= pow_bits;
}

}  // namespace sh
