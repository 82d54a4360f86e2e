"""Whole proofs over Poseidon252MerkleChannel on the device (stwo_hip_prove_*_poseidon252: the
per-call prover with commit_on_layer_poseidon252, grind_poseidon252 and the felt252 channel): their
bytes and final digest equal the pure-Python restatement's (tests/pystark.py over tests/p252.py)
on small traces, the host verifier accepts them at every size up to 2^16 x 100, a proof is bound
to its channel, and the Poseidon2 AIR's proofs carry the sanity-check status and the
OodsNotMatching rejection they carry over Blake2s."""
import sys
import time

import pytest

import pystark as S
from hip_lib import SH

pytestmark = pytest.mark.gpu
PC = "poseidon252"
WF = dict(pow_bits=5, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)
CONFIGS = [WF, dict(pow_bits=8, log_blowup=2, log_last_layer=2, n_queries=7, fold_step=1),
           dict(pow_bits=3, log_blowup=1, log_last_layer=1, n_queries=5, fold_step=2),
           dict(pow_bits=4, log_blowup=1, log_last_layer=3, n_queries=4, fold_step=3)]


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


@pytest.mark.parametrize("log_n,n_cols,ci", [(3, 16, 0), (4, 8, 0), (6, 19, 1), (6, 19, 2), (7, 24, 3), (8, 100, 0)])
def test_wide_fibonacci_matches_the_restatement(hb, log_n, n_cols, ci):
    cfg = CONFIGS[ci]
    say("poseidon252 wide-fib", log_n, n_cols, cfg)
    trace = hb.gen_trace_wide_fibonacci(n_cols, log_n, 1, 31)
    pr = hb.prove_wide_fibonacci_per_call([trace], [log_n], cfg_of(cfg), merkle_channel=PC)
    pb, dig, ok = S.prove(S.WideFib(log_n, n_cols, 1, 31), cfg, S.Poseidon252MerkleChannel)
    assert ok and pr.status == 0
    assert pr.bytes == pb and pr.digest == dig.to_bytes(32, "little")
    assert SH.verify_wide_fibonacci(pr.bytes, [log_n], [n_cols], cfg_of(cfg), merkle_channel=PC) == ""
    hb.free_columns(trace)


@pytest.mark.parametrize("log_n,n_cols", [(12, 37), (16, 100)])
def test_large_wide_fibonacci_is_accepted(hb, log_n, n_cols):
    say("poseidon252 wide-fib large", log_n, n_cols)
    trace = hb.gen_trace_wide_fibonacci(n_cols, log_n, 1, 5)
    hb.prove_wide_fibonacci_per_call([trace], [log_n], cfg_of(WF), merkle_channel=PC)  # warm
    t0 = time.perf_counter()
    pr = hb.prove_wide_fibonacci_per_call([trace], [log_n], cfg_of(WF), merkle_channel=PC)
    say(f"  prove {1e3 * (time.perf_counter() - t0):.1f} ms, {len(pr.bytes)} bytes")
    assert pr.status == 0
    assert SH.verify_wide_fibonacci(pr.bytes, [log_n], [n_cols], cfg_of(WF), merkle_channel=PC) == ""
    assert SH.verify_wide_fibonacci(pr.bytes, [log_n], [n_cols], cfg_of(WF)) != ""
    blake = hb.prove_wide_fibonacci_per_call([trace], [log_n], cfg_of(WF))
    assert blake.bytes != pr.bytes
    assert SH.verify_wide_fibonacci(blake.bytes, [log_n], [n_cols], cfg_of(WF), merkle_channel=PC) != ""
    hb.free_columns(trace)


def test_multi_component_is_accepted(hb):
    logs, ncols = [9, 6, 8, 7], [24, 100, 8, 33]
    say("poseidon252 components", logs, ncols)
    traces = [hb.gen_trace_wide_fibonacci(nc, lg, 1, 7 + i) for i, (lg, nc) in enumerate(zip(logs, ncols))]
    pr = hb.prove_wide_fibonacci_per_call(traces, logs, cfg_of(WF), merkle_channel=PC)
    assert pr.status == 0
    assert SH.verify_wide_fibonacci(pr.bytes, logs, ncols, cfg_of(WF), merkle_channel=PC) == ""
    swapped = [ncols[1], ncols[0]] + ncols[2:]
    assert SH.verify_wide_fibonacci(pr.bytes, logs, swapped, cfg_of(WF), merkle_channel=PC) != ""
    for t in traces:
        hb.free_columns(t)


@pytest.mark.parametrize("log_n,ipr", [(4, 1), (5, 2)])
def test_poseidon_matches_the_restatement(hb, log_n, ipr):
    say("poseidon252 poseidon2", log_n, ipr)
    trace = hb.gen_trace_poseidon(ipr, log_n, 1, 33)
    pr = hb.prove_poseidon_per_call(trace, ipr, log_n, cfg_of(WF), allow_unsatisfied=True, merkle_channel=PC)
    air = S.Poseidon(log_n, ipr, 1, 33)
    pb, dig, ok = S.prove(air, WF, S.Poseidon252MerkleChannel)
    assert not ok and pr.status == 2
    assert pr.bytes == pb and pr.digest == dig.to_bytes(32, "little")
    cs = tuple(int(x) for x in air.claimed_sum)
    assert SH.verify_poseidon(pr.bytes, ipr, log_n, cs, cfg_of(WF), merkle_channel=PC) == "OodsNotMatching"
    hb.free_columns(trace)


def test_poseidon_at_2e18_instances(hb):
    log_n, ipr = 15, 8
    cfg = dict(WF, pow_bits=10)
    say("poseidon252 poseidon2 2^18 instances")
    trace = hb.gen_trace_poseidon(ipr, log_n, 1, 2026)
    t0 = time.perf_counter()
    pr = hb.prove_poseidon_per_call(trace, ipr, log_n, cfg_of(cfg), allow_unsatisfied=True, merkle_channel=PC)
    say(f"  prove {1e3 * (time.perf_counter() - t0):.1f} ms, {len(pr.bytes)} bytes")
    assert pr.status == 2
    cs = pr.claimed_sum
    assert SH.verify_poseidon(pr.bytes, ipr, log_n, cs, cfg_of(cfg), merkle_channel=PC) == "OodsNotMatching"
    hb.free_columns(trace)
