"""MerkleOpsLifted on the device (stwo_hip_build_leaves_lifted, stwo_hip_build_next_layer_lifted)
against the restatement in tests/lifted.py (pinned in tests/test_lifted.py), for both lifted
hashers: single and mixed size groups, all columns at the lifting size, a lifting size above the
largest column, no columns, odd 8-value remainders for Poseidon252, a whole lifted tree to its
root, and the argument checks (unsorted sizes, a column of one value, a column above the lifting
size)."""
import random

import numpy as np
import pytest

import lifted as LF
import p252
from hip_lib import SH

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hb():
    b = SH.HipBackend(0)
    yield b
    b.close()


def to_hashes(words, n, hasher):
    w = np.asarray(words, np.uint32)
    if hasher == "blake2s":
        return [w[8 * i:8 * i + 8].tobytes() for i in range(n)]
    return [SH.felt252_from_words(w[8 * i:8 * i + 8]) for i in range(n)]


def from_hashes(hs, hasher):
    if hasher == "blake2s":
        return np.frombuffer(b"".join(hs), np.uint32).copy()
    return np.concatenate([SH.felt252_to_words(h) for h in hs])


CASES = [([], 3), ([1], 1), ([3, 3, 3], 3), ([1, 2, 2, 4], 5), ([2] * 17 + [3] * 2, 4),
         ([1, 3, 3, 5, 5, 5, 6], 8), ([3] * 9, 4), ([2, 4] + [4] * 15, 6), ([5] * 33, 5), ([1] * 40, 7)]


@pytest.mark.parametrize("hasher", ["blake2s", "poseidon252"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_build_leaves(hb, hasher, case):
    logs, L = CASES[case]
    say("lifted leaves", hasher, logs, L)
    rng = random.Random(500 + case)
    cols = [[rng.randrange(LF.P31) for _ in range(1 << l)] for l in logs]
    dev = [hb.from_cpu(np.array(c, np.uint32)) for c in cols]
    out = hb.build_leaves_lifted(dev, logs, L, hasher)
    n = (1 << L) if cols else 1
    assert to_hashes(out.to_cpu(), n, hasher) == LF.build_leaves_stream(cols, L, hasher)


@pytest.mark.parametrize("hasher", ["blake2s", "poseidon252"])
def test_tree_to_root(hb, hasher):
    rng = random.Random(9)
    logs, L = [4, 6, 6, 8, 9, 9, 9], 10
    say("lifted tree", hasher, logs, L)
    cols = [[rng.randrange(LF.P31) for _ in range(1 << l)] for l in logs]
    dev = [hb.from_cpu(np.array(c, np.uint32)) for c in cols]
    layer = hb.build_leaves_lifted(dev, logs, L, hasher)
    # the leaves on a sample of rows, then every level up to the root against the restatement's
    exp = LF.build_leaves_stream(cols, L, hasher) if hasher == "blake2s" else None
    got = to_hashes(layer.to_cpu(), 1 << L, hasher)
    if exp is not None:
        assert got == exp
    else:
        for i in (0, 1, 2, 3, 511, 512, 1022, 1023):
            words = [int(c[LF.lifted_row(i, L - l)]) for c, l in zip(cols, logs)]
            assert got[i] == LF.poseidon252_stream(words)
    level = got
    for ls in range(L - 1, -1, -1):
        layer = hb.build_next_layer_lifted(layer, ls, hasher)
        level = LF.build_next_layer(level, hasher)
        assert to_hashes(layer.to_cpu(), 1 << ls, hasher) == level
    assert len(level) == 1


def test_next_layer_from_host_hashes(hb):
    rng = random.Random(4)
    hs = [rng.randrange(p252.P) for _ in range(16)]
    prev = hb.from_cpu(from_hashes(hs, "poseidon252"))
    out = hb.build_next_layer_lifted(prev, 3, "poseidon252")
    assert to_hashes(out.to_cpu(), 8, "poseidon252") == LF.build_next_layer(hs, "poseidon252")


def test_argument_checks(hb):
    c2 = hb.from_cpu(np.arange(4, dtype=np.uint32))
    c1 = hb.from_cpu(np.arange(2, dtype=np.uint32))
    c0 = hb.from_cpu(np.arange(1, dtype=np.uint32))
    for cols, logs, L in [([c2, c1], [2, 1], 3), ([c0], [0], 2), ([c2], [2], 1)]:
        with pytest.raises(SH.StwoHipError):
            hb.build_leaves_lifted(cols, logs, L, "blake2s")
