"""GPU parity at the benchmark configurations' full sizes.

- Wide-Fibonacci 2^16 rows x 100 columns, 5 PoW bits, blowup 1, 3 queries: proof bytes equal the
  oracle's and the oracle's verifier accepts them.
- Poseidon2 2^18 instances (2^15 rows x 8 per row, 1264 + 32 columns): proof bytes, digest and
  claimed sum equal the oracle's.
- Kernel batch 2^20-word columns: 16 columns through interpolate, evaluate with blowup 1 (to 2^21),
  the Merkle root and fold_circle_into_line compared word for word with the oracle; all 256
  columns through size-independent properties (interpolate(evaluate(c)) = c, Merkle nodes that
  hash their children, leaves that hash their rows, and a fold that is linear in its source).
"""
import hashlib
import sys

import numpy as np
import pytest

import oracle_lib as O
from hip_lib import SH

P = O.P
pytestmark = pytest.mark.gpu


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


def test_wide_fibonacci_2e16_x100_byte_identical(hb):
    cfg = dict(pow_bits=5, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)
    say("wide-fib 2^16 x 100, pow 5")
    trace = hb.gen_trace_wide_fibonacci(100, 16, 1, 2026)
    pr = hb.prove_wide_fibonacci(trace, 16, cfg_of(cfg))
    st, proof, dig, _, _ = O.prove(air="wide_fib", log_n_rows=16, n_cols=100, input_mode=1, seed=2026, **cfg)
    assert pr.status == st == 0
    assert pr.bytes == proof and pr.digest == dig
    assert O.verify(pr.bytes, air="wide_fib", log_n_rows=16, n_cols=100, **cfg) == ""
    hb.free_columns(trace)


@pytest.mark.parametrize("pow_bits", [10])
def test_poseidon2_2e18_byte_identical(hb, pow_bits):
    cfg = dict(pow_bits=pow_bits, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)
    say("poseidon2 2^18 instances (2^15 rows x 8), pow", pow_bits)
    trace = hb.gen_trace_poseidon(8, 15, 1, 2026)
    pr = hb.prove_poseidon(trace, 8, 15, cfg_of(cfg), allow_unsatisfied=True)
    st, proof, dig, cs, _ = O.prove(air="poseidon", log_n_rows=15, ipr=8, input_mode=1, seed=2026, **cfg)
    assert pr.status == st
    assert pr.claimed_sum == tuple(map(int, cs))
    assert pr.bytes == proof and pr.digest == dig
    hb.free_columns(trace)


LOG = 20


def test_kernel_batch_16_columns_match_oracle(hb):
    say("kernel batch: 16 columns of 2^20 against the oracle")
    rng = np.random.default_rng(20)
    host = [rng.integers(0, P, 1 << LOG, dtype=np.uint32) for _ in range(16)]
    dev = [hb.from_cpu(h) for h in host]
    coeffs = hb.interpolate_columns(dev, LOG)
    ext = hb.evaluate_polynomials(coeffs, LOG, 1)
    oc = [O.interpolate(h) for h in host]
    oe = [O.evaluate(c, LOG + 1) for c in oc]
    for i in range(16):
        assert np.array_equal(coeffs[i].to_cpu(), oc[i]), i
        assert np.array_equal(ext[i].to_cpu(), oe[i]), i
    # Merkle root of the 16 extended columns, layer by layer through commit_on_layer
    layer = hb.commit_on_layer(LOG + 1, None, ext)
    for j in range(LOG, -1, -1):
        layer = hb.commit_on_layer(j, layer, [])
    assert layer.to_cpu().view(np.uint8).tobytes() == O.merkle_root(oe)
    # fold_circle_into_line of the extended columns as 4 QM31 columns into a nonzero destination
    alpha = tuple(int(x) for x in rng.integers(0, P, 4))
    for j in range(4):
        d0 = [rng.integers(0, P, 1 << LOG, dtype=np.uint32) for _ in range(4)]
        dd = [hb.from_cpu(x) for x in d0]
        hb.fold_circle_into_line(dd, ext[4 * j:4 * j + 4], alpha, LOG + 1)
        exp = O.fold_circle_into_line(oe[4 * j:4 * j + 4], d0, alpha)
        for k in range(4):
            assert np.array_equal(dd[k].to_cpu(), exp[k]), (j, k)


def test_kernel_batch_256_columns_properties(hb):
    n_cols = 256
    say("kernel batch: 256 columns of 2^20, properties")
    rng = np.random.default_rng(256)
    evals = []
    for _ in range(n_cols):
        evals.append(hb.from_cpu(rng.integers(0, P, 1 << LOG, dtype=np.uint32)))
    coeffs = hb.interpolate_columns(evals, LOG)
    ext = hb.evaluate_polynomials(coeffs, LOG, 1)
    # interpolate(evaluate(c)) = c padded with zeros, for every column
    back = hb.interpolate_columns(ext, LOG + 1)
    for i in range(n_cols):
        b = back[i].to_cpu()
        assert np.array_equal(b[: 1 << LOG], coeffs[i].to_cpu()) and not b[1 << LOG:].any(), i
        back[i].free()
    # the first and last columns' evaluations against the oracle
    for i in (0, n_cols - 1):
        assert np.array_equal(ext[i].to_cpu(), O.evaluate(coeffs[i].to_cpu(), LOG + 1)), i
    # Merkle: leaves hash their rows (sampled), inner nodes hash their children (every node, oracle)
    leaves = hb.commit_on_layer(LOG + 1, None, ext)
    lv = leaves.to_cpu().view(np.uint8).reshape(-1, 32)
    rows = sorted(set(int(r) for r in rng.integers(0, 2 << LOG, 48)) | {0, (2 << LOG) - 1})
    vals = np.asarray(hb.batch_at_multi(ext, rows), np.uint32).reshape(len(rows), n_cols)
    for r, v in zip(rows, vals):
        assert lv[r].tobytes() == hashlib.blake2s(v.astype("<u4").tobytes(), digest_size=32).digest(), r
    prev, prev_host = leaves, lv.reshape(-1)
    for j in range(LOG, -1, -1):
        node = hb.commit_on_layer(j, prev, [])
        nh = node.to_cpu().view(np.uint8)
        assert nh.tobytes() == O.commit_on_layer(j, prev_host, []).tobytes(), j
        prev.free()
        prev, prev_host = node, nh
    # fold: linear in the source, fold(a + b) = fold(a) + fold(b) (destinations start at zero)
    alpha = tuple(int(x) for x in rng.integers(0, P, 4))
    outs = []
    for j in range(n_cols // 4):
        dd = [hb.column(1 << LOG, zeroed=True) for _ in range(4)]
        hb.fold_circle_into_line(dd, ext[4 * j:4 * j + 4], alpha, LOG + 1)
        outs.append(dd)
    s = [hb.column(2 << LOG, zeroed=True) for _ in range(4)]
    for j in (0, 1):
        hb.accumulate(s, ext[4 * j:4 * j + 4], LOG + 1)
    fs = [hb.column(1 << LOG, zeroed=True) for _ in range(4)]
    hb.fold_circle_into_line(fs, s, alpha, LOG + 1)
    for k in range(4):
        lhs = fs[k].to_cpu().astype(np.uint64)
        rhs = (outs[0][k].to_cpu().astype(np.uint64) + outs[1][k].to_cpu()) % P
        assert np.array_equal(lhs, rhs), k
    # and the first and last QM31 columns against the oracle
    for j in (0, n_cols // 4 - 1):
        src = [ext[4 * j + k].to_cpu() for k in range(4)]
        exp = O.fold_circle_into_line(src, [np.zeros(1 << LOG, np.uint32)] * 4, alpha)
        for k in range(4):
            assert np.array_equal(outs[j][k].to_cpu(), exp[k]), (j, k)
