"""The drop-in seam end to end: whole proofs driven through the trait entry points of the C ABI,
one call per trait-method invocation (interpolate_columns, evaluate_polynomials, commit_on_layer,
evaluate_constraint_quotients_on_domain, lift_and_accumulate, split_at_mid, batch_eval_at_point,
accumulate_numerators, compute_quotients_and_combine, fold_circle_into_line, fold_line, grind,
batch_at_multi), as the reference's generic prove() drives a backend. Their bytes must equal the
oracle's and the fused prover's, and the oracle's verifier must accept the wide-Fibonacci ones."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from hip_lib import SH

pytestmark = pytest.mark.gpu
WF = dict(pow_bits=5, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hb():
    b = SH.HipBackend(0)
    yield b
    b.close()


def cfg_of(d):
    return SH.PcsConfig.make(d["pow_bits"], d["log_blowup"], d["log_last_layer"], d["n_queries"], d["fold_step"], 0)


@pytest.mark.parametrize("log_n,n_cols,cfg", [
    (3, 16, WF), (8, 100, WF), (11, 37, WF),
    (9, 20, dict(pow_bits=8, log_blowup=2, log_last_layer=2, n_queries=7, fold_step=1)),
    (9, 20, dict(pow_bits=3, log_blowup=1, log_last_layer=1, n_queries=5, fold_step=2)),
    (10, 24, dict(pow_bits=4, log_blowup=1, log_last_layer=3, n_queries=4, fold_step=3)),
    (16, 100, WF)])
def test_per_call_wide_fibonacci_byte_identical(hb, log_n, n_cols, cfg):
    say("per-call wide-fib", log_n, n_cols, cfg)
    trace = hb.gen_trace_wide_fibonacci(n_cols, log_n, 1, 31)
    pr = hb.prove_wide_fibonacci_per_call([trace], [log_n], cfg_of(cfg))
    st, proof, dig, _, _ = O.prove(air="wide_fib", log_n_rows=log_n, n_cols=n_cols, input_mode=1, seed=31, **cfg)
    assert pr.status == st == 0
    assert pr.bytes == proof and pr.digest == dig
    assert O.verify(pr.bytes, air="wide_fib", log_n_rows=log_n, n_cols=n_cols, **cfg) == ""
    fused = hb.prove_wide_fibonacci(trace, log_n, cfg_of(cfg))
    assert fused.bytes == pr.bytes
    hb.free_columns(trace)


@pytest.mark.parametrize("logs,ncols", [([5, 7, 4], [16, 9, 30]), ([12, 9, 10, 11, 9, 10], [24, 100, 8, 60, 33, 12])])
def test_per_call_multi_component_byte_identical(hb, logs, ncols):
    say("per-call components", logs, ncols)
    traces = [hb.gen_trace_wide_fibonacci(nc, lg, 1, 7 + i) for i, (lg, nc) in enumerate(zip(logs, ncols))]
    pr = hb.prove_wide_fibonacci_per_call(traces, logs, cfg_of(WF))
    st, proof, dig, _ = O.prove_multi(logs, ncols, input_mode=1, seed=7, **WF)
    assert pr.status == st == 0
    assert pr.bytes == proof and pr.digest == dig
    assert O.verify_multi(pr.bytes, logs, ncols, **WF) == ""


@pytest.mark.parametrize("log_n,ipr", [(5, 1), (6, 2), (7, 8), (15, 8)])
def test_per_call_poseidon_byte_identical(hb, log_n, ipr):
    say("per-call poseidon", log_n, ipr)
    cfg = dict(WF, pow_bits=10) if log_n == 15 else WF
    trace = hb.gen_trace_poseidon(ipr, log_n, 1, 33)
    pr = hb.prove_poseidon_per_call(trace, ipr, log_n, cfg_of(cfg), allow_unsatisfied=True)
    st, proof, dig, cs, _ = O.prove(air="poseidon", log_n_rows=log_n, ipr=ipr, input_mode=1, seed=33, **cfg)
    assert pr.status == st
    assert pr.bytes == proof and pr.digest == dig
    fused = hb.prove_poseidon(trace, ipr, log_n, cfg_of(cfg), allow_unsatisfied=True)
    assert fused.bytes == pr.bytes and fused.claimed_sum == tuple(map(int, cs))
    hb.free_columns(trace)


def test_interaction_trace_is_deterministic(hb):
    """The interaction trace entry (which the per-call Poseidon proofs above pin through their
    commitments) gives canonical words and the same columns and claimed sum on every call."""
    log_n, ipr = 6, 2
    trace = hb.gen_trace_poseidon(ipr, log_n, 1, 5)
    z, alpha = (11, 22, 33, 44), (5, 6, 7, 8)
    cols, cs = hb.gen_interaction_trace_poseidon(trace, ipr, log_n, z, alpha)
    assert len(cols) == 4 * ipr and all(0 <= v < O.P for v in cs)
    assert all((c.to_cpu() < O.P).all() for c in cols)
    cols2, cs2 = hb.gen_interaction_trace_poseidon(trace, ipr, log_n, z, alpha)
    assert cs2 == cs and all(np.array_equal(a.to_cpu(), b.to_cpu()) for a, b in zip(cols, cols2))
