"""Lifted Merkle leaves (MerkleOpsLifted, crates/stwo/src/prover/vcs_lifted/ops.rs) in the
restatement: the row map ((i >> (L - l + 1)) << 1) | (i & 1) is the lifted polynomial's, checked
against the circle domain (the point at bit-reversed index i of CanonicCoset(L), doubled L - l
times, is the point at the mapped index of CanonicCoset(l)) and against the example in
prover/backend/cpu/merkle_lifted.rs:17-45; the stream formulation of the leaves equals the
CpuBackend's stateful algorithm (per-row hashers cloned as the row count grows, 16-column chunks,
Poseidon252's 16-value buffer) for both hashers, over column sets with several size groups, long
groups and odd buffer remainders; a planted row-map mistake breaks the agreement. Host only."""
import random

import pytest

import lifted as LF
import pystark as S


@pytest.mark.parametrize("l,L", [(1, 1), (1, 3), (2, 5), (3, 3), (3, 6), (4, 7)])
def test_row_map_is_the_lifted_polynomial(l, L):
    small, big = S.domain_bitrev(l), S.domain_bitrev(L)
    for i, pt in enumerate(big):
        for _ in range(L - l):
            pt = S.mp_add(pt, pt)
        assert pt == small[LF.lifted_row(i, L - l)]


def test_row_map_matches_the_reference_example():
    # a..h on a domain of 8 (bit-reversed a e c g b f d h) lifted to 16 (merkle_lifted.rs:17-45)
    col = list("aecgbfdh")
    assert [col[LF.lifted_row(i, 1)] for i in range(16)] == list("aeaecgcgbfbfdhdh")


def random_cols(rng, logs):
    return [[rng.randrange(LF.P31) for _ in range(1 << l)] for l in logs]


CASES = [([], 3), ([1], 1), ([2], 4), ([3, 3, 3], 3), ([1, 2, 2, 4], 5), ([2] * 17 + [3] * 2, 4),
         ([1, 3, 3, 5, 5, 5, 6], 6), ([3] * 9, 4), ([2, 4] + [4] * 15, 6)]


@pytest.mark.parametrize("hasher", ["blake2s", "poseidon252"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_stream_equals_the_stateful_algorithm(hasher, case):
    logs, L = CASES[case]
    rng = random.Random(100 + case)
    cols = random_cols(rng, logs)
    a = LF.build_leaves_stream(cols, L, hasher)
    b = LF.build_leaves_stateful(cols, L, hasher)
    assert a == b and len(a) == (1 << L if cols else 1)


def test_planted_row_map_mistake_breaks_agreement(monkeypatch):
    rng = random.Random(7)
    cols = random_cols(rng, [1, 2, 4])
    good = LF.build_leaves_stateful(cols, 5, "blake2s")
    monkeypatch.setattr(LF, "lifted_row", lambda i, s: (i >> s) if s else i)
    assert LF.build_leaves_stream(cols, 5, "blake2s") != good


def test_next_layer_is_hash_children():
    rng = random.Random(3)
    leaves = LF.build_leaves_stream(random_cols(rng, [2, 3]), 3, "blake2s")
    nxt = LF.build_next_layer(leaves, "blake2s")
    import hashlib
    assert nxt[1] == hashlib.blake2s(leaves[2] + leaves[3]).digest()
    pl = LF.build_leaves_stream(random_cols(rng, [2, 3]), 3, "poseidon252")
    import p252
    assert LF.build_next_layer(pl, "poseidon252")[2] == p252.permute([pl[4], pl[5], 2])[0]
