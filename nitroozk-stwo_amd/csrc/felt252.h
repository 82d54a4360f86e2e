// The Stark field (p = 2^251 + 17 2^192 + 1) and Starknet's Poseidon permutation, on the host and
// the device: the hash of the reference's second Merkle channel (Poseidon252MerkleChannel,
// crates/stwo/src/core/vcs/poseidon252_merkle.rs; crates/stwo/src/core/channel/poseidon252.rs),
// which calls starknet-crypto's poseidon_hash, poseidon_hash_many and poseidon_permute_comp.
//
// An element enters and leaves as eight little-endian 32-bit limbs (F), canonical. Inside the
// permutation it is in Montgomery form (M): nine limbs of 29 bits in 32-bit registers, R = 2^261.
// Radix 2^29 leaves room in each limb: a product column (at most 9 products of limbs below 2^30)
// accumulates in one 64-bit register with no carry handling, and sums are limb-wise additions whose
// carries are propagated once. p = 1 (mod 2^29), so the Montgomery factor of each column is minus
// its low limb, and p's limbs are zero except limb 0 (1), limb 6 (17 2^18) and limb 8 (2^19).
//
// The permutation (state of 3): 4 full rounds (add 3 constants, cube all three), 83 partial rounds
// (add one constant to the last element, cube it), 4 full rounds; each round ends with the MDS mix
// t = s0 + s1 + s2 -> (t + 2 s0, t - 2 s1, t - 3 s2). The 107 round constants, in that order, are
// the published table (crates/stwo/src/stwo_cuda/cuda/poseidon252_constants.cuh), converted to
// Montgomery form at compile time.
#pragma once
#include "field.h"

namespace sh {
namespace f252 {

struct F {
  uint32_t w[8];
};

constexpr uint32_t PL[8] = {1u, 0u, 0u, 0u, 0u, 0u, 17u, 0x08000000u};

SHD constexpr F zero() { return F{{0, 0, 0, 0, 0, 0, 0, 0}}; }
SHD constexpr F from_u64(uint64_t v) { return F{{(uint32_t)v, (uint32_t)(v >> 32), 0, 0, 0, 0, 0, 0}}; }
SHD constexpr bool is_zero(const F& a) {
  uint32_t o = 0;
  for (int i = 0; i < 8; ++i) o |= a.w[i];
  return o == 0;
}
SHD constexpr bool equal(const F& a, const F& b) {
  uint32_t o = 0;
  for (int i = 0; i < 8; ++i) o |= a.w[i] ^ b.w[i];
  return o == 0;
}
SHD constexpr bool geq_p(const F& a) {
  for (int i = 7; i >= 0; --i)
    if (a.w[i] != PL[i]) return a.w[i] > PL[i];
  return true;
}
SHD constexpr F sub_p(const F& a) {
  F r = zero();
  uint32_t borrow = 0;
  for (int i = 0; i < 8; ++i) {
    const uint64_t d = (uint64_t)a.w[i] - PL[i] - borrow;
    r.w[i] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
  }
  return r;
}
// a + b for a, b < p (the sum stays below 2^253)
SHD constexpr F add(const F& a, const F& b) {
  F r = zero();
  uint64_t c = 0;
  for (int i = 0; i < 8; ++i) {
    c += (uint64_t)a.w[i] + b.w[i];
    r.w[i] = (uint32_t)c;
    c >>= 32;
  }
  return geq_p(r) ? sub_p(r) : r;
}
// a - b for a, b < p
SHD constexpr F sub(const F& a, const F& b) {
  F r = zero();
  uint32_t borrow = 0;
  for (int i = 0; i < 8; ++i) {
    const uint64_t d = (uint64_t)a.w[i] - b.w[i] - borrow;
    r.w[i] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
  }
  if (borrow) {
    uint64_t c = 0;
    for (int i = 0; i < 8; ++i) {
      c += (uint64_t)r.w[i] + PL[i];
      r.w[i] = (uint32_t)c;
      c >>= 32;
    }
  }
  return r;
}
// 2^n mod p (canonical doublings; compile time)
constexpr F pow2_mod_p(int n) {
  F x = from_u64(1);
  for (int i = 0; i < n; ++i) x = add(x, x);
  return x;
}

constexpr F RC_NORMAL[107] = {
    {{0xac77fe6fu, 0x80b7fd8eu, 0x8109ae47u, 0xe58e2ad9u, 0x562a30b9u, 0x39dd92f9u, 0xea556a23u, 0x06861759u}},
    {{0xf27508c4u, 0xe4ab1a22u, 0xabf28a64u, 0x3da43f76u, 0x3d00425au, 0xffc8397au, 0x995d5af9u, 0x03827681u}},
    {{0x704a8309u, 0x7a0dbe17u, 0x279b2d68u, 0x2cac75dcu, 0x277dc7cbu, 0xe7f760a2u, 0x2fad44d0u, 0x03a3956du}},
    {{0xe67ef4cdu, 0x882031afu, 0x26a2d1c6u, 0x291c78f9u, 0x214aa759u, 0xf13c4282u, 0x7d421fe1u, 0x0626c47au}},
    {{0xca05d6a6u, 0xadfe17bau, 0x948d772cu, 0x2d327fccu, 0x18cfd41fu, 0x5bd6df55u, 0xe1650503u, 0x078985f8u}},
    {{0x341696ffu, 0x142dcf34u, 0xb456dc1du, 0xc6e26a68u, 0x75341243u, 0x204c6598u, 0x867514a3u, 0x05427f10u}},
    {{0xdce699d7u, 0x0465e60eu, 0xd9d4cb8bu, 0x847cd2c5u, 0x3f0883c5u, 0x45436173u, 0x36e4c729u, 0x05af083fu}},
    {{0xfd388882u, 0xc77e709bu, 0x2ff84b1au, 0x2d3975f9u, 0xf7b352a5u, 0x54fa3f74u, 0xbde3d06du, 0x07d71701u}},
    {{0x7a6b1d3eu, 0xe3958446u, 0x99ffea44u, 0xac1b64f6u, 0x320a1c5eu, 0x9c26f8a6u, 0x88201900u, 0x0603da06u}},
    {{0xf6453a5du, 0x613b9721u, 0xfd13a36bu, 0xeebd8870u, 0xf47ad1cdu, 0x8e79ce13u, 0x6bde2f28u, 0x04332a6fu}},
    {{0x5cba78e7u, 0xc9e8caebu, 0xaeff60d6u, 0x4b9a813au, 0x4dec2e7eu, 0x5310a04cu, 0x61664c68u, 0x053d0ebfu}},
    {{0xeaabfebcu, 0x1bfc2d15u, 0x928494eeu, 0x6c82f99fu, 0x88028d2au, 0x5ae5ebcbu, 0x89484583u, 0x05346a68u}},
    {{0xf54a3f9au, 0x92864c00u, 0x9dd5a995u, 0x3433b6abu, 0x445954bfu, 0x3453cc97u, 0x1df4258cu, 0x04b085ebu}},
    {{0x961f6d96u, 0x5ca5d7b0u, 0x610e6f2du, 0xfdfe0a8eu, 0x79b2a169u, 0x965f12a0u, 0x9d508285u, 0x0731cfd1u}},
    {{0x3587eeaau, 0xf8079c1eu, 0x9a73d743u, 0x2ecd9e1fu, 0x4936b3e7u, 0xcc6f7a77u, 0x5339852bu, 0x0217d08bu}},
    {{0x08518c6eu, 0x959e2670u, 0x10c81b20u, 0xcb966ba5u, 0xc850dab3u, 0x63599b13u, 0xd633b0fdu, 0x000c935du}},
    {{0xe7b8b2beu, 0x467fe940u, 0x87103971u, 0x8cf5a0a3u, 0x3f79a7d9u, 0xee187ed2u, 0x78dd6772u, 0x052af8d3u}},
    {{0x95f8ab75u, 0x6362cda5u, 0xb79b2e50u, 0xaa6e28a7u, 0x9f12fcceu, 0x1ec9918bu, 0x98b2682fu, 0x0294851cu}},
    {{0xab086064u, 0xc204b169u, 0x63471494u, 0x9da8c300u, 0x418d4f58u, 0x824d1021u, 0x0bacc280u, 0x011b5999u}},
    {{0xeb2acd6au, 0xb599d49cu, 0xe6bbdcd0u, 0xeb2ce690u, 0xae099b9bu, 0x91960d59u, 0xe3d7753fu, 0x04b4df56u}},
    {{0xb5161ac4u, 0x5f1cf09eu, 0xf30629fcu, 0x66243534u, 0xd8ff06e4u, 0x3ecae9fbu, 0x15a757dcu, 0x005eecfau}},
    {{0xa4279059u, 0x7f8448bdu, 0xf81b1567u, 0x82e5a7ceu, 0x34a1ec52u, 0x46592276u, 0x8b9680e0u, 0x0680bfddu}},
    {{0xacd85053u, 0xb8877d97u, 0x1cedd791u, 0x5c17e10bu, 0x94f7a306u, 0x14e29307u, 0xab0a1a7au, 0x01d0bf8fu}},
    {{0x7e32f464u, 0x876701a3u, 0x10f14c48u, 0xc0fc11a6u, 0x53c0d412u, 0x54ba2070u, 0x9f808aceu, 0x02c2c8c7u}},
    {{0x54afa562u, 0x630ad56cu, 0x11ad7bceu, 0x74d8234cu, 0x9b858d34u, 0x52aae19au, 0xd01d20ecu, 0x0354ec9eu}},
    {{0xdeec1e55u, 0x94e6ea28u, 0x892340e9u, 0xe4136ac5u, 0xa42287f4u, 0xc38bb5d1u, 0xcf6427bau, 0x030df20fu}},
    {{0xfc9c6778u, 0x72eda296u, 0x4091e314u, 0x41d99b42u, 0xdeff61e2u, 0x3040bafbu, 0x9c64e7eeu, 0x0528cf32u}},
    {{0x30bd1960u, 0xff6e7579u, 0x69e6c5b5u, 0xbf0c3e0cu, 0x0df5435eu, 0x63478966u, 0x4f623534u, 0x040416f2u}},
    {{0xccc9ea38u, 0xd5444435u, 0x1e88339cu, 0x15ba21b1u, 0xbac7dce3u, 0xfd488ae3u, 0x936e2ed9u, 0x00380c8fu}},
    {{0x1f255b75u, 0x45e0e229u, 0x832829d2u, 0x03ad0541u, 0x392efd2du, 0xf1a8e344u, 0x5603d176u, 0x01cc4f5du}},
    {{0x99c09752u, 0x23c502aau, 0xe015ea54u, 0xe011d6c8u, 0x99f5d142u, 0x9539310du, 0xaf5da91fu, 0x05728917u}},
    {{0x17219aa1u, 0xc788feccu, 0xd17c56d7u, 0x90925107u, 0x48f0f235u, 0x46e295a3u, 0xa9e86e1au, 0x00efb450u}},
    {{0x774dca99u, 0xa6c26693u, 0x77de6c53u, 0x84b8fcd9u, 0x6b342d07u, 0xe1a02561u, 0xd36c421au, 0x02020d74u}},
    {{0x83989287u, 0xa2b7b4f5u, 0x9f4bf84fu, 0x33505096u, 0xe511dc82u, 0x2705558au, 0xb75fd3bfu, 0x07cfb309u}},
    {{0x812d75e9u, 0xc785a784u, 0xa617eb98u, 0x9d7b7d02u, 0xece62680u, 0x5365e009u, 0xb2e9349au, 0x04651e48u}},
    {{0x71c0dcc7u, 0x7a3b5c8fu, 0xaad0f094u, 0xcae822d9u, 0x19da923cu, 0x22d02697u, 0xb270f651u, 0x00d77627u}},
    {{0x32908ea8u, 0xd3b7758cu, 0x7f0dd8b7u, 0x82b00337u, 0x37773a8bu, 0xd571b3feu, 0x641b54c4u, 0x0199ad3du}},
    {{0x37cdfe6eu, 0x319e72d1u, 0xbe2da289u, 0x82b2d297u, 0x72a73334u, 0x973e2e91u, 0x0a8ecfd3u, 0x044f3364u}},
    {{0x6701abf7u, 0x6b6f366bu, 0xf5216eabu, 0x7f801234u, 0xdcf1e6beu, 0x89d00a02u, 0x894d9641u, 0x07e4adf9u}},
    {{0xe711ee90u, 0xca559fb4u, 0xb5471dfdu, 0xa7c6aecfu, 0x8a9817edu, 0xf5ff808fu, 0xb3c90452u, 0x03641fa5u}},
    {{0xcc3f3533u, 0xd0eee2c2u, 0x6e2f0e02u, 0x306df32eu, 0x8fa923fcu, 0x897a49a7u, 0xefd2fcbdu, 0x03de5729u}},
    {{0x82d6c994u, 0xdb77bd28u, 0xc8f06c9bu, 0x06563500u, 0xca0be20cu, 0x7f622966u, 0x1a3fc1e2u, 0x06269189u}},
    {{0x04baa0beu, 0x86e96e06u, 0x9885ae1du, 0x1590cc4fu, 0x2205763du, 0x688739f7u, 0xf11c18e4u, 0x06608d3bu}},
    {{0xf593a316u, 0x98c25cbbu, 0xcbdda4fdu, 0x07249f51u, 0xe779116du, 0x3419726cu, 0x39cac71eu, 0x011c9c9bu}},
    {{0x7c462ad7u, 0x4dd68c55u, 0xea40c684u, 0x748e3312u, 0x74da27b9u, 0xaef0850fu, 0x8203269cu, 0x061e23b5u}},
    {{0xcf83bdc8u, 0xce010855u, 0xb25951e4u, 0xd9878440u, 0x010bc2a3u, 0xf870a572u, 0xab1d9488u, 0x04182cd9u}},
    {{0x4f3ecb71u, 0x65f6a59fu, 0x9d306f99u, 0xf1df2fe8u, 0x23116d15u, 0xf9055e68u, 0x4a096793u, 0x0520fe6cu}},
    {{0x2fd4afa1u, 0x3fbe4d4eu, 0xf47723eau, 0x99ed6d71u, 0x3dcd3dfau, 0x29e09309u, 0xd6e58101u, 0x0346b2b2u}},
    {{0xe9f7e6cfu, 0xea72877cu, 0xe8e973acu, 0xe8b682c8u, 0x3684bee4u, 0x8ec1dd2au, 0x23e7f144u, 0x01359ca9u}},
    {{0xb95d9675u, 0x92420f37u, 0xb145a1b3u, 0xde9deadau, 0xd24de86fu, 0x00dfefdau, 0x55cf3078u, 0x047c655fu}},
    {{0xca8441a8u, 0xa38ad87du, 0xf2d58b7au, 0x98efd925u, 0x9c285a95u, 0xa968cd7cu, 0x16555fa8u, 0x04ab291fu}},
    {{0x17c629c5u, 0x2c9c7c5du, 0x59510e39u, 0x22c44ef9u, 0x7d86c21au, 0x223d1f6fu, 0x7c782101u, 0x039f409cu}},
    {{0xcafb6052u, 0xbfacead0u, 0x5587a3bau, 0xe1a2f9fbu, 0xcd6beb02u, 0xd86eecb0u, 0x782f882au, 0x044be36bu}},
    {{0x3a0d538bu, 0x107679a5u, 0x94c52ce7u, 0x3cd1ceb3u, 0xb5b50720u, 0x2906db6eu, 0xe9b504adu, 0x050a1dfdu}},
    {{0xeab183abu, 0x83e7aebbu, 0x1017dae0u, 0x680bdd7fu, 0xd11ac6c3u, 0x7b181c0du, 0x4da89e28u, 0x05c753c1u}},
    {{0x5fc0520cu, 0xa5774652u, 0x5f7b4f4fu, 0x6eee93e1u, 0x3b180f38u, 0x06c8015au, 0xed322321u, 0x02cf6306u}},
    {{0x2c1c3a09u, 0x4d7d8e50u, 0x689bd4f5u, 0x1dfc6380u, 0x34742909u, 0x420873cfu, 0x34d52e27u, 0x02c20146u}},
    {{0xf86ed8a1u, 0x50fc09cdu, 0x45f010d8u, 0x975c4463u, 0x058e33e5u, 0x2fdacde2u, 0xbd93e02bu, 0x03cfb9c5u}},
    {{0x1f6b0299u, 0x9ea4e128u, 0xf72cb50au, 0x6e311e84u, 0x11fc5f80u, 0x97933f14u, 0xa383cf38u, 0x0363fa71u}},
    {{0x91b05df8u, 0xbcb960dau, 0x3686cffeu, 0xb4901b2au, 0x76271676u, 0x6947b3fcu, 0x57067ee1u, 0x07281996u}},
    {{0x7e4ee021u, 0x652a078du, 0x9872df5au, 0xc9bde34au, 0x28e8921du, 0x4f0723b7u, 0x7d27f3d3u, 0x03fdfbd4u}},
    {{0x98eff449u, 0x945c4533u, 0xa7d24197u, 0xe53cc02cu, 0x7858eb7du, 0x7dc0efbeu, 0x9440cacdu, 0x07f24137u}},
    {{0xa5c25f0bu, 0x04354e52u, 0x7574d5f1u, 0xcbb457a2u, 0xf3727797u, 0x4e3bf056u, 0x1ea9a000u, 0x05b2e877u}},
    {{0xbda2e23eu, 0xa057a081u, 0xe6bf822eu, 0xe910245bu, 0x333146e1u, 0xa7e0b3b0u, 0xe708de44u, 0x00a8ddbcu}},
    {{0xaa9b0f61u, 0xedce2e19u, 0x33835613u, 0xc12270e5u, 0x90a0f551u, 0x31aa47cdu, 0xdaca24e4u, 0x02d521e0u}},
    {{0x30e08ae6u, 0x41a6266cu, 0x7aaee003u, 0x03eef7b0u, 0x93f855afu, 0xf7d5ac3bu, 0x2aa54d2cu, 0x06cdbc0fu}},
    {{0x36ed9bf3u, 0x0941112bu, 0xcbe92c7au, 0x1ceee452u, 0x6794c096u, 0x4c5da3adu, 0x7111ec8fu, 0x03dd96a1u}},
    {{0xc36687dcu, 0xc0c5e732u, 0x96a65d9eu, 0xa2cfb921u, 0x06159344u, 0x7ac07fddu, 0xdeedc5c0u, 0x05eafb1eu}},
    {{0x65c207e6u, 0x312cfcd0u, 0x7c7f449bu, 0xbcec5a7bu, 0x60feaebdu, 0x324577b2u, 0x7b09eda9u, 0x04ab038du}},
    {{0x7fdef220u, 0x495fc80cu, 0x2570f68eu, 0x463c5805u, 0x4e235608u, 0x05d2b0d9u, 0x1e4df6b5u, 0x04ca7198u}},
    {{0x7e413d2fu, 0x573c15b7u, 0xe2759ffeu, 0x7e9b59d7u, 0x9c8ea813u, 0x2aa41989u, 0xda4617e3u, 0x06dee9c6u}},
    {{0x60c6f186u, 0x1e625d04u, 0x4770e454u, 0x5b4a795au, 0x6065a430u, 0x4dcbe239u, 0xb34ddab8u, 0x058f9e60u}},
    {{0xfc9af632u, 0xfcecef17u, 0xf274bea8u, 0x2d290906u, 0xdb6c3404u, 0xe6c9c735u, 0x802a10c1u, 0x047b7b4au}},
    {{0xa2bd430cu, 0xb9463a06u, 0xe0bfc0fbu, 0x084a074bu, 0x6a186774u, 0x096ecc93u, 0xb9ad7171u, 0x01849bcdu}},
    {{0x466736ccu, 0xc4d699b1u, 0x217547deu, 0xea3b49b4u, 0xddaecd8au, 0x8af5767bu, 0xe0443834u, 0x041870fbu}},
    {{0x84738b07u, 0x965029d2u, 0xd4da68d1u, 0xc0ec42e3u, 0x557daf28u, 0xfa02aa64u, 0x598076a9u, 0x0226c04eu}},
    {{0xd6dac0fbu, 0x5335b2a7u, 0x71f269a9u, 0x475cec43u, 0x655c3ddbu, 0x42eb92d6u, 0xf0485a5bu, 0x01f0e971u}},
    {{0x26431563u, 0x7b40dd39u, 0xf795047au, 0xebdf3703u, 0xa47ee3f4u, 0xf994d35au, 0x2907dccbu, 0x009f31ccu}},
    {{0xc64bd7d2u, 0xffe8cca4u, 0x98c84451u, 0xc22cfbc1u, 0x58ce5a42u, 0xe31ce4dfu, 0x78f3b641u, 0x04b40cceu}},
    {{0xe50f8b9fu, 0xf72cb93cu, 0x62fdfb70u, 0xf3ac7369u, 0x3de4a226u, 0x3e456317u, 0x89e4bd8au, 0x01916604u}},
    {{0x7b467410u, 0x7ed8dd4cu, 0x103373dcu, 0xa814b475u, 0x293f2daeu, 0xf74eb01fu, 0x618db971u, 0x018c0919u}},
    {{0x7eadc803u, 0xfb240bd5u, 0x941c3dc1u, 0xef37002eu, 0x121577d0u, 0x845c8753u, 0x3848530eu, 0x035b6025u}},
    {{0x957451d5u, 0xfd029134u, 0x13d97fa6u, 0x581ad9b4u, 0x71a5f362u, 0xc8b43a9fu, 0x1575ae91u, 0x01ae99dbu}},
    {{0xe172647bu, 0xd1dfb357u, 0xb58a37d2u, 0x7d7cb8feu, 0xbd5d7d33u, 0x28148ebcu, 0xf3f8a0f7u, 0x03e6e1d0u}},
    {{0xa209b8f3u, 0xc38dd1f4u, 0x619e16f5u, 0x6ee3e9adu, 0x1b55d270u, 0x659e1a17u, 0xffa8f96au, 0x018bc36du}},
    {{0x893de4edu, 0x83fa5b64u, 0x9a1cc3abu, 0x99a16efeu, 0xa107ff91u, 0x2b54afc3u, 0x1afb6a30u, 0x02c7a3efu}},
    {{0x69de4bd7u, 0x9856e944u, 0x4b0c6ddeu, 0x5e4fe4e8u, 0xe92f6ae8u, 0xf5e27dd8u, 0x89bed07bu, 0x053a7bd8u}},
    {{0xf03ccea9u, 0xd3cc0cc2u, 0xce9a0bf9u, 0x6bec5a02u, 0xa35995f8u, 0xfda704acu, 0x7ffc6318u, 0x04d383ffu}},
    {{0xc326f5cdu, 0xd1292519u, 0x03037710u, 0xa96e0b35u, 0x8a65b21cu, 0x3d07ef7eu, 0x62fb8ad5u, 0x04667b67u}},
    {{0x22e5921fu, 0xb7fb5364u, 0x9589f02eu, 0x6055dcccu, 0xea98bcd4u, 0x42a93c39u, 0x3e75cf0bu, 0x002cc8b4u}},
    {{0x82bf9c60u, 0x4b2ce255u, 0xbe59c55fu, 0xba4df0e7u, 0xbfd76086u, 0x38751447u, 0x8680871du, 0x06b32ee9u}},
    {{0x387102a3u, 0x2d635cb1u, 0x0609f086u, 0x34efac1fu, 0x358b82e7u, 0xaa3b3c81u, 0x7c7182fau, 0x03e90792u}},
    {{0x24fbc6eau, 0x93ea5e84u, 0x73f6a937u, 0x8d2f4379u, 0x8e512af7u, 0x5f025372u, 0x7b3a0897u, 0x03f3a505u}},
    {{0xca68d8f5u, 0x7a946bb6u, 0x7224726au, 0x1c21e101u, 0xfd8aa582u, 0xf8aa74b3u, 0x73724779u, 0x014b491du}},
    {{0xabb814efu, 0x61695ea4u, 0x4ad13738u, 0x7aca84c5u, 0x4fe3b936u, 0xe7f60e51u, 0x7bbfc30au, 0x05c8278cu}},
    {{0x755c8074u, 0xba0b4502u, 0xba3bf99bu, 0x67018a82u, 0xe65e44b4u, 0x5a08a7ddu, 0x7f9836eeu, 0x06485193u}},
    {{0xf1d971dcu, 0xf9ea5e73u, 0xb85f422fu, 0x882be77cu, 0xfb52b441u, 0x9eca450fu, 0x25129476u, 0x06a9ac84u}},
    {{0x545064deu, 0x1884f5cau, 0x1d27f0fcu, 0xc6a81d99u, 0x2f7bd098u, 0xc9a2b71fu, 0x710b0d04u, 0x037ec35bu}},
    {{0x2bfeb7a1u, 0xf46d6a24u, 0x1538bdffu, 0x72c6d0a6u, 0x883040dau, 0x35119816u, 0x5b052c02u, 0x005334f7u}},
    {{0xa50aa431u, 0xcd49d6dcu, 0xf2ec6547u, 0x8f80ee4au, 0x9d871ae6u, 0xc1020ccau, 0xcbd9e056u, 0x05d0af4fu}},
    {{0x901b1e65u, 0x5ab99537u, 0x0f1d53bau, 0xb4699dc0u, 0x46d24e00u, 0x4114a19cu, 0xe2fba569u, 0x030131bcu}},
    {{0x2e788965u, 0x4957660fu, 0x3161be3cu, 0x538f93f1u, 0xed2e641cu, 0xb34c0750u, 0xa7c1ae86u, 0x05646a95u}},
    {{0x7d4ec01au, 0xc90df47bu, 0x59692b52u, 0x78581f52u, 0x0a11f43eu, 0x9fac3623u, 0xd7b430c7u, 0x04b9f291u}},
    {{0xbae01a13u, 0xd890e74au, 0x6ceb4d73u, 0x3e00becfu, 0x27072dc8u, 0x1a98f191u, 0x3d3480f4u, 0x05006d39u}},
    {{0x6a679d80u, 0x202dc3b2u, 0xe6b346fdu, 0xacf4f702u, 0x15143106u, 0x0e7cb8a1u, 0x199f3b26u, 0x062c9d42u}},
    {{0xdc5717a5u, 0xbf1d9ac2u, 0xabd8af54u, 0x7606836eu, 0xa13b7f2cu, 0xf180b1a8u, 0x92db5099u, 0x051274d0u}},
    {{0xbfff0c5bu, 0x861a8db0u, 0x0d9cdbedu, 0xf415e14fu, 0xaa4ca528u, 0x7ad0fb7au, 0xb8eb75e1u, 0x061fc552u}},
};
// ---- Montgomery form in radix 2^29 -----------------------------------------------------------
constexpr uint32_t MASK29 = (1u << 29) - 1;
struct M {
  uint32_t w[9];
};

// F (any value below 2^256) -> nine 29-bit limbs
SHD constexpr M to_radix29(const F& a) {
  M r{};
  for (int i = 0; i < 9; ++i) {
    const int bit = 29 * i, w = bit / 32, sh = bit % 32;
    uint64_t v = a.w[w];
    if (w + 1 < 8) v |= (uint64_t)a.w[w + 1] << 32;
    r.w[i] = (uint32_t)(v >> sh) & MASK29;
  }
  return r;
}
// nine 29-bit limbs of a value below 2^256 -> F
SHD constexpr F from_radix29(const M& a) {
  F r = zero();
  for (int i = 0; i < 9; ++i) {
    const int bit = 29 * i, w = bit / 32, sh = bit % 32;
    const uint64_t v = (uint64_t)a.w[i] << sh;
    r.w[w] |= (uint32_t)v;
    if (w + 1 < 8) r.w[w + 1] |= (uint32_t)(v >> 32);
  }
  return r;
}

// a b 2^-261 (mod p) by product scanning: column k gathers a_i b_j (i + j = k) and the reduction
// terms m_i p_j in 64-bit accumulators; for k < 9 the factor m_k = -acc (mod 2^29) clears the
// column (p's limb 0 is 1: acc + m_k is a multiple of 2^29, and its quotient is the carry out),
// above that the columns are the result's limbs. Inputs: limbs below 2^30 (a column then stays
// below 2^64); output: limbs 0-7 below 2^29, value below a b / 2^261 + p (below 2 p for inputs
// below 32 p).
SHD constexpr M mul29(const M& a, const M& b) {
  uint32_t m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  M r{};
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 17; ++k) {
    // two partial sums per column (products of even and odd i, the even one starting from the
    // carry in): chains of multiply-adds of which only one waits for the previous column
    uint64_t e = acc, o = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int j = k - i;
      if (j >= 0 && j < 9) {
        if (i & 1)
          o += (uint64_t)a.w[i] * b.w[j];
        else
          e += (uint64_t)a.w[i] * b.w[j];
      }
    }
    if (k >= 6 && k - 6 < 9) o += (uint64_t)m[k - 6] * (17u << 18);
    if (k >= 8 && k - 8 < 9) e += (uint64_t)m[k - 8] << 19;
    acc = e + o;
    if (k < 9) {
      m[k] = (0u - (uint32_t)acc) & MASK29;
      acc = (acc + m[k]) >> 29;
    } else {
      r.w[k - 9] = (uint32_t)acc & MASK29;
      acc >>= 29;
    }
  }
  r.w[8] = (uint32_t)acc;
  return r;
}

constexpr M R2M = to_radix29(pow2_mod_p(522));  // R^2 mod p
// canonical F -> Montgomery form (below 1.04 p)
SHD constexpr M to_mont(const F& a) { return mul29(to_radix29(a), R2M); }
// Montgomery form (limbs below 2^29, value below 2^261) -> canonical F
SHD constexpr F from_mont(const M& a) {
  const F r = from_radix29(mul29(a, M{{1, 0, 0, 0, 0, 0, 0, 0, 0}}));  // at most p
  return geq_p(r) ? sub_p(r) : r;
}

constexpr M ONE = to_mont(from_u64(1));
constexpr M TWO = to_mont(from_u64(2));
constexpr M THREE = to_mont(from_u64(3));

struct RoundConstants {
  M c[107];
};
constexpr RoundConstants make_round_constants() {
  RoundConstants t{};
  for (int i = 0; i < 107; ++i) t.c[i] = to_mont(RC_NORMAL[i]);
  return t;
}
constexpr RoundConstants RC = make_round_constants();

// ---- the permutation's arithmetic: reduced lazily, no branches ----------------------------------
// limb-wise sum
SHD M addl(const M& a, const M& b) {
  M r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.w[i] = a.w[i] + b.w[i];
  return r;
}
// carries of limbs below 2^32 - 8: limbs 0-7 below 2^29 after it
SHD M carry(const M& x) {
  M r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint32_t v = x.w[i] + c;
    r.w[i] = v & MASK29;
    c = v >> 29;
  }
  r.w[8] = x.w[8] + c;
  return r;
}
// x - (q - 1) p for q = floor(x / 2^251) >= 1 (x itself for q = 0): congruent to x and below
// 2^251 + p < 2 p, for x with normalized limbs and q at most 120 (a value below 2^257.9: the limb-6
// term 17 2^18 (q - 1) stays below 2^29). (q - 1) p touches limbs 0, 6 and 8 only; the borrows are
// -1 or 0.
SHD M reduce_top(const M& x) {
  const uint32_t q = x.w[8] >> 19;
  const uint32_t k = q - (q != 0u ? 1u : 0u);
  const uint32_t sub[9] = {k, 0u, 0u, 0u, 0u, 0u, (17u << 18) * k, 0u, k << 19};
  M r;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int32_t v = (int32_t)x.w[i] - (int32_t)sub[i] + c;
    r.w[i] = (uint32_t)v & MASK29;
    c = v >> 29;  // arithmetic: -1 or 0
  }
  r.w[8] = (uint32_t)((int32_t)x.w[8] - (int32_t)sub[8] + c);
  return r;
}
// The value v plus a multiple of 2^29 moved down from each limb to the one below: the same value
// with limbs 0-7 at least D (2^29 - 1), so that subtracting D times a normalized element limb by
// limb leaves every limb non-negative.
constexpr M borrowed(const M& v, uint32_t D) {
  M r = v;
  for (int i = 0; i < 8; ++i) {
    r.w[i] += D << 29;
    r.w[i + 1] -= D;
  }
  return r;
}
constexpr M from_small_multiple_of_p(uint32_t k) {  // k p in radix 2^29, k < 32
  return M{{k, 0u, 0u, 0u, 0u, 0u, (17u << 18) * k, 0u, k << 19}};
}
// 5 p and 7 p, borrowed for subtracting 2 s1 and 3 s2 (their top limbs 5 2^19 and 7 2^19 exceed
// 2 and 3 times the top limb of elements below 2.01 p, plus the borrow); 21 p for 2 s1 with s1
// below 10.3 p (the partial rounds that follow a round without reduction)
constexpr M K5P = borrowed(from_small_multiple_of_p(5), 2);
constexpr M K7P = borrowed(from_small_multiple_of_p(7), 3);
constexpr M K21P = borrowed(from_small_multiple_of_p(21), 2);

// t = s0 + s1 + s2 -> (t + 2 s0, t - 2 s1, t - 3 s2) for normalized inputs below 2.01 p. The
// unnormalized outputs stay below 9.1 p, 10.1 p and 12.1 p (limbs below 7 2^29); the first two are
// brought below 2.01 p, the third (which is cubed next) only normalized.
SHD void mix(M* s) {
  M a, b, c;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const uint32_t t = s[0].w[i] + s[1].w[i] + s[2].w[i];
    a.w[i] = t + 2u * s[0].w[i];
    b.w[i] = t + K5P.w[i] - 2u * s[1].w[i];
    c.w[i] = t + K7P.w[i] - 3u * s[2].w[i];
  }
  s[0] = reduce_top(carry(a));
  s[1] = reduce_top(carry(b));
  s[2] = carry(c);
}
// The partial rounds' mix, reducing s0 and s1 every other round (s2, cubed next, is only
// normalized). REDUCE = false takes s0, s1 below 2.01 p and leaves them normalized below 9.3 p and
// 10.3 p (c below 12.3 p); REDUCE = true takes those, subtracts 2 s1 from 21 p, and brings a (below
// 39.2 p) and b (below 41.7 p) below 2 p (c below 27.7 p, its cube below 1.06 p). Limbs stay below
// 2^32 - 8 either way: every input is normalized.
template <bool REDUCE>
SHD void mix_partial(M* s) {
  M a, b, c;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const uint32_t t = s[0].w[i] + s[1].w[i] + s[2].w[i];
    a.w[i] = t + 2u * s[0].w[i];
    b.w[i] = t + (REDUCE ? K21P.w[i] : K5P.w[i]) - 2u * s[1].w[i];
    c.w[i] = t + K7P.w[i] - 3u * s[2].w[i];
  }
  if (REDUCE) {
    s[0] = reduce_top(carry(a));
    s[1] = reduce_top(carry(b));
  } else {
    s[0] = carry(a);
    s[1] = carry(b);
  }
  s[2] = carry(c);
}
SHD M cube(const M& x) { return mul29(mul29(x, x), x); }
// each element plus its constant (limbs below 2^30, value below 14 p) cubed: below 1.2 p
SHD void full_round(M* s, int k) {
  for (int i = 0; i < 3; ++i) s[i] = cube(addl(s[i], RC.c[k + i]));
  mix(s);
}
// poseidon_permute_comp on a Montgomery-form state: elements with limbs below 2^30 and values
// below 3.1 p (Montgomery forms of canonical values, or a permutation's outputs plus an absorbed
// value). Outputs: s0 and s1 below 2.01 p, s2 below 12.1 p, all normalized (from_mont reduces
// fully). The 83 partial rounds run as 41 pairs (without, then with the reduction of s0 and s1)
// and one round without it, whose outputs (below 12.3 p plus a constant: under the full rounds'
// 14 p) go to the last full rounds. One copy of the code per program: the device calls it, it is
// not inlined.
__host__ __device__ __attribute__((noinline)) inline void permute(M* s) {
  for (int i = 0; i < 3; ++i) s[i] = carry(s[i]);
  int k = 0;
  for (int r = 0; r < 4; ++r, k += 3) full_round(s, k);
#ifdef __HIP_DEVICE_COMPILE__
#pragma unroll 1
#endif
  for (int r = 0; r < 41; ++r, k += 2) {
    s[2] = cube(addl(s[2], RC.c[k]));
    mix_partial<false>(s);
    s[2] = cube(addl(s[2], RC.c[k + 1]));
    mix_partial<true>(s);
  }
  s[2] = cube(addl(s[2], RC.c[k]));
  mix_partial<false>(s);
  ++k;
  for (int r = 0; r < 4; ++r, k += 3) full_round(s, k);
}
// poseidon_hash(x, y) (Montgomery form in and out)
SHD M hash2(const M& x, const M& y) {
  M s[3] = {x, y, TWO};
  permute(s);
  return s[0];
}
// poseidon_hash_many, absorbing one value at a time (Montgomery form; absorbed values below 1.1 p)
struct Sponge {
  M s[3] = {M{}, M{}, M{}};
  M pending = M{};
  bool odd = false;
  SHD void absorb(const M& v) {
    if (!odd) {
      pending = v;
      odd = true;
      return;
    }
    s[0] = addl(s[0], pending);
    s[1] = addl(s[1], v);
    permute(s);
    odd = false;
  }
  SHD M finish() {
    if (odd) {
      s[0] = addl(s[0], pending);
      s[1] = addl(s[1], ONE);
    } else {
      s[0] = addl(s[0], ONE);
    }
    permute(s);
    return s[0];
  }
};
// construct_felt252_from_m31s (normal form): felt = felt << 31 | x for each of n <= 8 values; a
// chunk shorter than 8 adds n 2^248
SHD F felt_from_m31s(const uint32_t* x, int n) {
  F f = zero();
  for (int k = 0; k < n; ++k) {
    for (int i = 7; i > 0; --i) f.w[i] = (f.w[i] << 31) | (f.w[i - 1] >> 1);
    f.w[0] = (f.w[0] << 31) | x[k];
  }
  if (n < 8) f.w[7] += (uint32_t)n << 24;
  return f;
}

}  // namespace f252
}  // namespace sh
