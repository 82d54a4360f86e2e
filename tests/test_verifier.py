"""The library's host verifier (stwo_verify_*, core/verifier.rs restated in the product) against
the oracle: it accepts every honest wide-Fibonacci proof the oracle makes (one and several
components, blowup 1-2, fold steps 1-3, last-layer bounds up to 2^3, PoW 3-10), rejects the same
single-bit tamperings and wrong statements the oracle's verifier rejects, and rejects the
Poseidon2 proofs with OodsNotMatching as the reference protocol does. Host only: no GPU."""
import hashlib

import numpy as np
import pytest

import oracle_lib as O
from hip_lib import SH


def cfg_of(d):
    return SH.PcsConfig.make(d["pow_bits"], d["log_blowup"], d["log_last_layer"], d["n_queries"], d["fold_step"], 0)


WF = dict(pow_bits=5, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)
CONFIGS = [WF, dict(pow_bits=8, log_blowup=2, log_last_layer=2, n_queries=7, fold_step=1),
           dict(pow_bits=3, log_blowup=1, log_last_layer=1, n_queries=5, fold_step=2),
           dict(pow_bits=4, log_blowup=1, log_last_layer=3, n_queries=4, fold_step=3),
           dict(pow_bits=10, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)]


@pytest.mark.parametrize("ci", range(len(CONFIGS)))
@pytest.mark.parametrize("log_n,n_cols", [(4, 8), (9, 20)])
def test_accepts_honest_wide_fibonacci(ci, log_n, n_cols):
    cfg = CONFIGS[ci]
    st, proof, _, _, _ = O.prove(air="wide_fib", log_n_rows=log_n, n_cols=n_cols, input_mode=1, seed=ci, **cfg)
    assert st == 0
    assert SH.verify_wide_fibonacci(proof, [log_n], [n_cols], cfg_of(cfg)) == ""


def test_accepts_multi_component():
    logs, ncols = [8, 5, 6, 7], [10, 40, 8, 17]
    st, proof, _, _ = O.prove_multi(logs, ncols, input_mode=1, seed=7, **WF)
    assert st == 0
    assert SH.verify_wide_fibonacci(proof, logs, ncols, cfg_of(WF)) == ""
    # a different statement (component sizes swapped) is rejected
    assert SH.verify_wide_fibonacci(proof, [5, 8, 6, 7], [40, 10, 8, 17], cfg_of(WF)) != ""


def test_rejects_tampering_like_the_oracle():
    log_n, n_cols = 7, 12
    _, proof, _, _, _ = O.prove(air="wide_fib", log_n_rows=log_n, n_cols=n_cols, input_mode=1, seed=3, **WF)
    rng = np.random.default_rng(1)
    for k in range(40):
        b = bytearray(proof)
        pos = int(rng.integers(24, len(b)))  # past the configuration words
        b[pos] ^= 1 << int(rng.integers(0, 8))
        b = bytes(b)
        mine = SH.verify_wide_fibonacci(b, [log_n], [n_cols], cfg_of(WF))
        theirs = O.verify(b, air="wide_fib", log_n_rows=log_n, n_cols=n_cols, **WF)
        assert mine != "", (k, pos)
        assert theirs != "", (k, pos)


def test_rejects_wrong_statement_and_config():
    _, proof, _, _, _ = O.prove(air="wide_fib", log_n_rows=6, n_cols=10, input_mode=1, seed=3, **WF)
    assert SH.verify_wide_fibonacci(proof, [6], [10], cfg_of(WF)) == ""
    assert SH.verify_wide_fibonacci(proof, [6], [11], cfg_of(WF)) != ""
    assert SH.verify_wide_fibonacci(proof, [7], [10], cfg_of(WF)) != ""
    assert "configuration" in SH.verify_wide_fibonacci(proof, [6], [10], cfg_of(dict(WF, pow_bits=4)))
    assert SH.verify_wide_fibonacci(proof[:-1], [6], [10], cfg_of(WF)).startswith("InvalidStructure")


@pytest.mark.parametrize("log_n,ipr", [(5, 1), (6, 2)])
def test_poseidon_is_rejected_as_the_reference_rejects_it(log_n, ipr):
    st, proof, _, cs, _ = O.prove(air="poseidon", log_n_rows=log_n, ipr=ipr, input_mode=1, seed=33, **WF)
    assert st == 2
    cs = tuple(int(x) for x in cs)
    assert SH.verify_poseidon(proof, ipr, log_n, cs, cfg_of(WF)) == "OodsNotMatching"
    assert O.verify(proof, claimed_sum=cs, air="poseidon", log_n_rows=log_n, ipr=ipr, **WF) == "OodsNotMatching"
