"""Lifted Merkle leaves, restated twice in pure Python (test infrastructure only; no product code
imports it). Paths relative to crates/stwo/src.

    build_leaves_stream      each leaf hashes, in column order, every column's value at its lifted
                             row ((i >> (L - l + 1)) << 1) | (i & 1): the formulation the device
                             kernels compute.
    build_leaves_stateful    the CpuBackend algorithm step by step (prover/backend/cpu/merkle_lifted.rs:
                             12-76): one hasher per row, groups of equal log size in 16-column
                             chunks, every hasher cloned from row (idx >> (ratio + 1) << 1) + (idx & 1)
                             when the row count grows, then finalised; with the incremental hashers
                             MerkleHasherLifted defines (core/vcs_lifted/blake2_merkle.rs:16-35: a
                             Blake2s stream of u32 LE values; core/vcs_lifted/poseidon252_merkle.rs:
                             18-99: a 16-value buffer, two felt252s of 8 values per permutation,
                             poseidon_finalize for the rest).
    build_next_layer         hash_children of adjacent pairs (prover/backend/cpu/merkle_lifted.rs:78-86).

The two leaf restatements are checked against each other, and the row map against the circle
domain's points (tests/test_lifted.py)."""
import hashlib

import p252

P31 = (1 << 31) - 1


def lifted_row(i, shift):
    return ((i >> (shift + 1)) << 1) | (i & 1)


def le32(v):
    return int(v).to_bytes(4, "little")


def blake2s_words(words):
    return hashlib.blake2s(b"".join(le32(w) for w in words), digest_size=32).digest()


def poseidon252_stream(words):
    felts = [p252.felt_from_m31s(list(words[k:k + 8])) for k in range(0, len(words), 8)]
    return p252.hash_many(felts)


def build_leaves_stream(cols, lifting_log_size, hasher):
    """cols: lists of ints, lengths powers of two >= 2 in non-decreasing order."""
    h = blake2s_words if hasher == "blake2s" else poseidon252_stream
    if not cols:
        return [h([])]
    logs = [len(c).bit_length() - 1 for c in cols]
    return [h([int(c[lifted_row(i, lifting_log_size - l)]) for c, l in zip(cols, logs)])
            for i in range(1 << lifting_log_size)]


class _Blake2sLifted:
    def __init__(self):
        self.h = hashlib.blake2s(digest_size=32)

    def clone(self):
        c = _Blake2sLifted()
        c.h = self.h.copy()
        return c

    def update_leaf(self, values):
        for v in values:
            self.h.update(le32(v))

    def finalize(self):
        return self.h.digest()


class _Poseidon252Lifted:
    def __init__(self):
        self.state, self.buffer = [0, 0, 0], []

    def clone(self):
        c = _Poseidon252Lifted()
        c.state, c.buffer = list(self.state), list(self.buffer)
        return c

    def update_leaf(self, values):
        stream = self.buffer + [int(v) for v in values]
        k = 0
        while len(stream) - k >= 16:
            a = p252.felt_from_m31s(stream[k:k + 8])
            b = p252.felt_from_m31s(stream[k + 8:k + 16])
            self.state[0] = (self.state[0] + a) % p252.P
            self.state[1] = (self.state[1] + b) % p252.P
            self.state = p252.permute(self.state)
            k += 16
        self.buffer = stream[k:]

    def finalize(self):
        rem = [p252.felt_from_m31s(self.buffer[k:k + 8]) for k in range(0, len(self.buffer), 8)]
        s = list(self.state)
        while len(rem) >= 2:
            s[0] = (s[0] + rem[0]) % p252.P
            s[1] = (s[1] + rem[1]) % p252.P
            s = p252.permute(s)
            rem = rem[2:]
        if len(rem) == 1:
            s[0] = (s[0] + rem[0]) % p252.P
        s[len(rem)] = (s[len(rem)] + 1) % p252.P
        return p252.permute(s)[0]


def build_leaves_stateful(cols, lifting_log_size, hasher):
    mk = _Blake2sLifted if hasher == "blake2s" else _Poseidon252Lifted
    if not cols:
        return [mk().finalize()]
    assert len(cols[0]) >= 2
    prev, prev_log = [mk(), mk()], 1
    groups = []
    for c in cols:
        lg = len(c).bit_length() - 1
        if groups and groups[-1][0] == lg:
            groups[-1][1].append(c)
        else:
            groups.append((lg, [c]))
    for lg, group in groups:
        ratio = lg - prev_log
        prev = [prev[(idx >> (ratio + 1) << 1) + (idx & 1)].clone() for idx in range(1 << lg)]
        for k in range(0, len(group), 16):
            chunk = group[k:k + 16]
            for i, hs in enumerate(prev):
                hs.update_leaf([int(col[i]) for col in chunk])
        prev_log = lg
    ratio = lifting_log_size - prev_log
    if ratio > 0:
        prev = [prev[(idx >> (ratio + 1) << 1) + (idx & 1)].clone() for idx in range(1 << lifting_log_size)]
    return [hs.finalize() for hs in prev]


def build_next_layer(prev, hasher):
    if hasher == "blake2s":
        return [hashlib.blake2s(prev[2 * i] + prev[2 * i + 1], digest_size=32).digest() for i in range(len(prev) // 2)]
    return [p252.hash2(prev[2 * i], prev[2 * i + 1]) for i in range(len(prev) // 2)]
