"""The Poseidon252 Merkle channel on the GPU against the restatement (tests/p252.py, pinned by the
reference's known answers): commit_on_layer for leaf layers of 1..100 columns (short last groups
of the 8-per-felt252 packing included), layers with children and columns, children only, a whole
mixed-degree tree, and the proof-of-work grind's nonce in the reference's search order."""
import numpy as np
import pytest

import p252 as Q
from hip_lib import SH

pytestmark = pytest.mark.gpu
M31 = (1 << 31) - 1


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hb():
    b = SH.HipBackend(0)
    yield b
    b.close()


def words_to_felts(col, n):
    w = col.to_cpu()
    return [SH.felt252_from_words(w[8 * i:8 * i + 8]) for i in range(n)]


@pytest.mark.parametrize("log_size,n_cols", [(0, 1), (3, 3), (3, 8), (5, 13), (4, 16), (7, 100), (2, 0)])
def test_leaf_layer(hb, log_size, n_cols):
    say("poseidon252 leaves", log_size, n_cols)
    rng = np.random.default_rng(500 + 7 * log_size + n_cols)
    cols = [rng.integers(0, M31, 1 << log_size, dtype=np.uint32) for _ in range(n_cols)]
    dev = [hb.from_cpu(c) for c in cols]
    out = hb.commit_on_layer_poseidon252(log_size, None, dev)
    got = words_to_felts(out, 1 << log_size)
    exp = [Q.hash_node(None, [int(c[i]) for c in cols]) for i in range(1 << log_size)]
    assert got == exp


@pytest.mark.parametrize("log_size,n_cols", [(0, 0), (3, 0), (2, 5), (4, 9)])
def test_layer_with_children(hb, log_size, n_cols):
    say("poseidon252 inner", log_size, n_cols)
    rng = np.random.default_rng(600 + log_size + n_cols)
    kids = [int.from_bytes(rng.bytes(32), "little") % Q.P for _ in range(2 << log_size)]
    prev = hb.from_cpu(np.concatenate([SH.felt252_to_words(k) for k in kids]))
    cols = [rng.integers(0, M31, 1 << log_size, dtype=np.uint32) for _ in range(n_cols)]
    out = hb.commit_on_layer_poseidon252(log_size, prev, [hb.from_cpu(c) for c in cols])
    got = words_to_felts(out, 1 << log_size)
    exp = [Q.hash_node((kids[2 * i], kids[2 * i + 1]), [int(c[i]) for c in cols]) for i in range(1 << log_size)]
    assert got == exp


def test_mixed_degree_tree(hb):
    """Columns of 2^5 injected at the leaves, columns of 2^3 at layer 3 (MerkleProver::commit's
    order), then hash-only layers up to the root."""
    rng = np.random.default_rng(61)
    big = [rng.integers(0, M31, 32, dtype=np.uint32) for _ in range(11)]
    small = [rng.integers(0, M31, 8, dtype=np.uint32) for _ in range(3)]
    layer = hb.commit_on_layer_poseidon252(5, None, [hb.from_cpu(c) for c in big])
    exp = [Q.hash_node(None, [int(c[i]) for c in big]) for i in range(32)]
    for ls in range(4, -1, -1):
        inj = small if ls == 3 else []
        layer = hb.commit_on_layer_poseidon252(ls, layer, [hb.from_cpu(c) for c in inj])
        exp = [Q.hash_node((exp[2 * i], exp[2 * i + 1]), [int(c[i]) for c in inj]) for i in range(1 << ls)]
        assert words_to_felts(layer, 1 << ls) == exp, ls
    say("root", hex(exp[0]))


@pytest.mark.parametrize("pow_bits", [0, 1, 4, 8])
def test_grind_matches_restatement(hb, pow_bits):
    digest = (0x1234567 + 999 * pow_bits) % Q.P
    n = hb.grind_poseidon252(digest, pow_bits)
    say("grind", pow_bits, n)
    assert n == Q.grind(digest, pow_bits)
    c = Q.Channel()
    c.digest = digest
    assert c.verify_pow_nonce(pow_bits, n)


def test_grind_beyond_one_chunk(hb):
    """16 bits need more than one chunk of 2^14 low values on average: the first valid nonce may
    have hi > 0, and every nonce before it in (hi, lo) order fails."""
    digest = 0x0777
    n = hb.grind_poseidon252(digest, 16)
    c = Q.Channel()
    c.digest = digest
    assert c.verify_pow_nonce(16, n)
    say("grind 16 bits", hex(n))
    hi, lo = n >> 32, n & 0xFFFFFFFF
    assert lo < (1 << 14)
    # spot-check the order: the restatement agrees on the chunk, and no nonce of it below lo passes
    if hi == 0:
        assert all(not c.verify_pow_nonce(16, x) for x in range(max(0, lo - 64), lo))


def test_permutation_rate_and_merkle_units(hb):
    """The permutation ceiling is plausible (one permutation is about 60 thousand VALU instructions,
    the device issues tens of trillions per second), and the merkle class counts one permutation per
    two absorbed felt252s plus one, or one for children alone."""
    rate = hb.measure_permutation_rate()
    say(f"permutations per second: {rate:.3e}")
    assert 5e7 < rate < 5e9
    hb.reset_kernel_stats()
    hb.set_kernel_timing(True)
    rng = np.random.default_rng(3)
    cols = [hb.from_cpu(rng.integers(0, M31, 1 << 6, dtype=np.uint32)) for _ in range(19)]
    leaves = hb.commit_on_layer_poseidon252(6, None, cols)  # 3 felts: 2 permutations per node
    mid = hb.commit_on_layer_poseidon252(5, leaves, cols[:5])  # 2 children + 1 felt: 2 per node
    hb.commit_on_layer_poseidon252(4, mid, [])  # children only: 1 per node
    hb.set_kernel_timing(False)
    assert hb.kernel_units()["merkle"] == 2 * 64 + 2 * 32 + 16
