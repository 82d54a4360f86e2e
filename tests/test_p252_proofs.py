"""Whole proofs over the Poseidon252 Merkle channel, checked on the host: proofs made by the
pure-Python restatement (tests/pystark.py over tests/p252.py) are accepted by the library's host
verifier (stwo_verify_*_poseidon252: its C++ felt252 channel, Merkle hasher and proof-of-work check
are written independently of the Python), single-bit tamperings and wrong statements are rejected,
a proof is bound to its channel (a Blake2s proof fails the Poseidon252 verifier and the reverse),
a hash at or above the Stark prime is rejected as InvalidStructure, and the Poseidon2 AIR's proofs
are rejected with OodsNotMatching as over Blake2s. Host only: no GPU."""
import numpy as np
import pytest

import p252
import pystark as S
from hip_lib import SH

PC = "poseidon252"
WF = dict(pow_bits=5, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)
CONFIGS = [WF, dict(pow_bits=8, log_blowup=2, log_last_layer=2, n_queries=7, fold_step=1),
           dict(pow_bits=3, log_blowup=1, log_last_layer=1, n_queries=5, fold_step=2),
           dict(pow_bits=4, log_blowup=1, log_last_layer=3, n_queries=4, fold_step=3)]


def cfg_of(d):
    return SH.PcsConfig.make(d["pow_bits"], d["log_blowup"], d["log_last_layer"], d["n_queries"], d["fold_step"], 0)


def py_proof(log_n, n_cols, cfg, seed=0):
    pb, digest, ok = S.prove(S.WideFib(log_n, n_cols, 1, seed), cfg, S.Poseidon252MerkleChannel)
    assert ok
    return pb, digest


@pytest.mark.parametrize("ci", range(len(CONFIGS)))
@pytest.mark.parametrize("log_n,n_cols", [(4, 8), (6, 19)])
def test_host_verifier_accepts_python_proofs(ci, log_n, n_cols):
    pb, digest = py_proof(log_n, n_cols, CONFIGS[ci], seed=ci)
    assert 0 < digest < p252.P
    assert SH.verify_wide_fibonacci(pb, [log_n], [n_cols], cfg_of(CONFIGS[ci]), merkle_channel=PC) == ""


def test_rejects_tampering_and_wrong_statements():
    log_n, n_cols = 5, 12
    pb, _ = py_proof(log_n, n_cols, WF, seed=3)
    assert SH.verify_wide_fibonacci(pb, [log_n], [n_cols + 1], cfg_of(WF), merkle_channel=PC) != ""
    assert SH.verify_wide_fibonacci(pb, [log_n + 1], [n_cols], cfg_of(WF), merkle_channel=PC) != ""
    assert SH.verify_wide_fibonacci(pb[:-1], [log_n], [n_cols], cfg_of(WF),
                                    merkle_channel=PC).startswith("InvalidStructure")
    rng = np.random.default_rng(5)
    for k in range(30):
        b = bytearray(pb)
        pos = int(rng.integers(24, len(b)))  # past the configuration words
        b[pos] ^= 1 << int(rng.integers(0, 8))
        assert SH.verify_wide_fibonacci(bytes(b), [log_n], [n_cols], cfg_of(WF), merkle_channel=PC) != "", (k, pos)


def test_a_proof_is_bound_to_its_channel():
    log_n, n_cols = 4, 8
    pb, _ = py_proof(log_n, n_cols, WF)
    assert SH.verify_wide_fibonacci(pb, [log_n], [n_cols], cfg_of(WF)) != ""
    pb2, _, ok = S.prove(S.WideFib(log_n, n_cols, 1, 0), WF)
    assert ok and SH.verify_wide_fibonacci(pb2, [log_n], [n_cols], cfg_of(WF)) == ""
    assert SH.verify_wide_fibonacci(pb2, [log_n], [n_cols], cfg_of(WF), merkle_channel=PC) != ""


def test_hash_above_the_prime_is_invalid_structure():
    log_n, n_cols = 4, 8
    pb, _ = py_proof(log_n, n_cols, WF)
    # the first commitment follows 6 configuration words and the tree count
    at = 7 * 4
    for bad in (p252.P, (1 << 256) - 1):
        b = pb[:at] + bad.to_bytes(32, "little") + pb[at + 32:]
        assert SH.verify_wide_fibonacci(b, [log_n], [n_cols], cfg_of(WF),
                                        merkle_channel=PC).startswith("InvalidStructure")
    # a canonical but wrong root is a root mismatch or a transcript failure, not a structure error
    b = pb[:at] + (p252.P - 1).to_bytes(32, "little") + pb[at + 32:]
    e = SH.verify_wide_fibonacci(b, [log_n], [n_cols], cfg_of(WF), merkle_channel=PC)
    assert e != "" and not e.startswith("InvalidStructure")


def test_poseidon_air_is_rejected_as_over_blake2s():
    log_n, ipr = 4, 1
    air = S.Poseidon(log_n, ipr, 1, 33)
    pb, _, ok = S.prove(air, WF, S.Poseidon252MerkleChannel)
    assert not ok
    cs = tuple(int(x) for x in air.claimed_sum)
    assert SH.verify_poseidon(pb, ipr, log_n, cs, cfg_of(WF), merkle_channel=PC) == "OodsNotMatching"
