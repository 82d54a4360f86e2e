"""GPU parity: every HIP entry point, called through the C ABI, against the oracle on the same
seeded inputs; whole proofs byte-identical to the oracle's and accepted by its verifier."""
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


def rand_cols(rng, n_cols, log_n):
    return [rng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(n_cols)]


def qm(rng):
    return [int(x) for x in rng.integers(0, P, 4)]


def random_point(seed):
    dig = np.frombuffer(hashlib.blake2s(bytes([seed]), digest_size=32).digest(), np.uint8).copy()
    pt = np.zeros(8, np.uint32)
    O.lib().oracle_random_point(O.u8ptr(dig), O.u32ptr(pt))
    return [int(x) for x in pt]


@pytest.mark.parametrize("log_n", [3, 4, 5, 8, 11, 12, 13, 14, 16, 18])
def test_interpolate_columns(hb, log_n):
    say("interpolate log", log_n)
    rng = np.random.default_rng(log_n)
    cols = rand_cols(rng, 3, log_n)
    dev = [hb.from_cpu(c) for c in cols]
    out = hb.interpolate_columns(dev, log_n)
    for c, o in zip(cols, out):
        assert np.array_equal(o.to_cpu(), O.interpolate(c))


@pytest.mark.parametrize("log_n,blow", [(3, 1), (4, 2), (7, 1), (10, 3), (11, 1), (12, 1), (13, 1), (15, 1), (15, 2),
                                        (16, 1), (17, 1)])
def test_evaluate_polynomials(hb, log_n, blow):
    say("evaluate log", log_n, "blowup", blow)
    rng = np.random.default_rng(100 + log_n)
    cols = rand_cols(rng, 2, log_n)
    dev = [hb.from_cpu(c) for c in cols]
    out = hb.evaluate_polynomials(dev, log_n, blow)
    for c, o in zip(cols, out):
        assert np.array_equal(o.to_cpu(), O.evaluate(c, log_n + blow))


def test_fft_round_trip_large(hb):
    # size-independent property at a full-size column: interpolate(evaluate(c)) == c
    log_n = 20
    say("fft round trip log", log_n)
    rng = np.random.default_rng(7)
    c = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    d = hb.from_cpu(c)
    ev = hb.evaluate_polynomials([d], log_n, 1)[0]
    back = hb.interpolate_columns([ev], log_n + 1)[0].to_cpu()
    assert np.array_equal(back[: 1 << log_n], c) and not back[1 << log_n:].any()


def test_alloc_and_free_columns(hb):
    say("alloc columns")
    cols = hb.columns(5, 1 << 12, zeroed=True)
    assert len({c.ptr for c in cols}) == 5
    assert all(not c.to_cpu().any() for c in cols)
    data = np.arange(1 << 12, dtype=np.uint32)
    hb.lib.stwo_hip_copy_h2d(hb.ctx, cols[2].ptr, data.ctypes.data, data.size)
    assert np.array_equal(cols[2].to_cpu(), data)
    hb.free_columns(cols)
    assert all(c.ptr is None for c in cols)
    again = hb.columns(5, 1 << 12)  # served from the context's cache of freed blocks
    hb.free_columns(again)


def test_twiddles_match_oracle(hb):
    tw, itw = hb.precompute_twiddles(10)
    otw, oitw = O.twiddles(10)
    assert np.array_equal(tw, otw) and np.array_equal(itw, oitw)


@pytest.mark.parametrize("log_n,n_cols,with_prev", [(0, 0, False), (3, 1, False), (5, 16, False), (6, 17, True),
                                                     (4, 0, True), (10, 33, True), (12, 100, False)])
def test_commit_on_layer(hb, log_n, n_cols, with_prev):
    say("merkle layer", log_n, n_cols, with_prev)
    rng = np.random.default_rng(log_n * 100 + n_cols)
    cols = rand_cols(rng, n_cols, log_n)
    prev = rng.integers(0, 256, (2 << log_n) * 32, dtype=np.uint8) if with_prev else None
    dprev = hb.from_cpu(prev.view(np.uint32)) if with_prev else None
    out = hb.commit_on_layer(log_n, dprev, [hb.from_cpu(c) for c in cols]).to_cpu()
    exp = O.commit_on_layer(log_n, prev, cols)
    assert out.view(np.uint8).tobytes() == exp.tobytes()


@pytest.mark.parametrize("log_n,n_cols,with_prev", [(6, 17, False), (8, 31, True), (11, 100, False), (12, 16, False)])
def test_commit_on_layer_strided_columns(hb, log_n, n_cols, with_prev):
    """Columns at a fixed stride of one column in one allocation take the address-free kernel."""
    say("merkle layer strided", log_n, n_cols, with_prev)
    rng = np.random.default_rng(log_n * 7 + n_cols)
    cols = rand_cols(rng, n_cols, log_n)
    big = hb.from_cpu(np.concatenate(cols))
    views = [big.ptr + 4 * (c << log_n) for c in range(n_cols)]
    prev = rng.integers(0, 256, (2 << log_n) * 32, dtype=np.uint8) if with_prev else None
    dprev = hb.from_cpu(prev.view(np.uint32)) if with_prev else None
    out = hb.commit_on_layer(log_n, dprev, views).to_cpu()
    exp = O.commit_on_layer(log_n, prev, cols)
    assert out.view(np.uint8).tobytes() == exp.tobytes()


@pytest.mark.parametrize("log_src", [3, 5, 10, 16])
def test_fold_circle_into_line(hb, log_src):
    say("fold circle", log_src)
    rng = np.random.default_rng(log_src)
    src = rand_cols(rng, 4, log_src)
    dst = rand_cols(rng, 4, log_src - 1)
    alpha = qm(rng)
    ddst = [hb.from_cpu(c) for c in dst]
    hb.fold_circle_into_line(ddst, [hb.from_cpu(c) for c in src], alpha, log_src)
    exp = O.fold_circle_into_line(src, dst, alpha)
    for d, e in zip(ddst, exp):
        assert np.array_equal(d.to_cpu(), e)


@pytest.mark.parametrize("log_src", [1, 2, 4, 9, 15])
def test_fold_line(hb, log_src):
    say("fold line", log_src)
    rng = np.random.default_rng(50 + log_src)
    src = rand_cols(rng, 4, log_src)
    alpha = qm(rng)
    out = hb.fold_line([hb.from_cpu(c) for c in src], alpha, log_src)
    exp = O.fold_line(src, log_src, alpha)
    for d, e in zip(out, exp):
        assert np.array_equal(d.to_cpu(), e)


@pytest.mark.parametrize("log_n", [0, 1, 3, 6, 11, 12, 15])
def test_batch_eval_at_point(hb, log_n):
    say("eval at point", log_n)
    rng = np.random.default_rng(200 + log_n)
    cols = rand_cols(rng, 3, log_n)
    pt = random_point(log_n)
    got = hb.batch_eval_at_point([hb.from_cpu(c) for c in cols], log_n, pt)
    for c, g in zip(cols, got):
        assert tuple(g) == tuple(map(int, O.eval_at_point(c, pt)))


@pytest.mark.parametrize("log_n", [4, 9, 13])
def test_barycentric_weights_and_eval(hb, log_n):
    say("barycentric", log_n)
    rng = np.random.default_rng(300 + log_n)
    pt = random_point(log_n + 1)
    w = hb.barycentric_weights(log_n, pt)
    got_w = np.stack([x.to_cpu() for x in w])
    assert np.array_equal(got_w, O.barycentric_weights(log_n, pt))
    cols = rand_cols(rng, 3, log_n)
    got = hb.batch_barycentric_eval([hb.from_cpu(c) for c in cols], log_n, w)
    for c, g in zip(cols, got):
        assert tuple(g) == tuple(map(int, O.barycentric_eval(c, pt)))


def test_grind_matches_oracle(hb):
    for bits in (1, 5, 10, 14):
        digest = hashlib.blake2s(b"pow" + bytes([bits]), digest_size=32).digest()
        say("grind bits", bits)
        assert hb.grind(digest, bits) == O.grind(digest, bits)


def test_batch_at_multi(hb):
    rng = np.random.default_rng(3)
    cols = rand_cols(rng, 5, 8)
    rows = [0, 7, 255, 3]
    got = hb.batch_at_multi([hb.from_cpu(c) for c in cols], rows)
    exp = [cols[c][r] for r in rows for c in range(5)]
    assert list(got) == exp


def test_batch_at_multi_both_paths(hb):
    """Gathers of at most 16 columns and 48 rows pass their lists as kernel arguments, larger ones
    as device tables; both write into one reused output buffer, grown on demand."""
    rng = np.random.default_rng(4)
    host = rand_cols(rng, 40, 10)
    dev = [hb.from_cpu(c) for c in host]
    shapes = [(16, 48), (17, 48), (16, 49), (1, 1), (40, 700), (3, 2), (40, 5), (40, 5), (16, 48)]
    for n_cols, n_rows in shapes:
        pick = sorted(rng.choice(40, n_cols, replace=False).tolist())
        rows = [int(r) for r in rng.integers(0, 1 << 10, n_rows)]
        got = hb.batch_at_multi([dev[c] for c in pick], rows)
        exp = [int(host[c][r]) for r in rows for c in pick]
        assert [int(v) for v in got] == exp, (n_cols, n_rows)


def test_trace_generation_matches_oracle(hb):
    say("trace gen")
    for mode in (0, 1):
        cols = hb.gen_trace_wide_fibonacci(7, 6, mode, 11)
        exp = O.gen_wide_fib(6, 7, mode, 11)
        assert all(np.array_equal(c.to_cpu(), e) for c, e in zip(cols, exp))
        pcols = hb.gen_trace_poseidon(2, 5, mode, 12)
        pexp = O.gen_poseidon(5, 2, mode, 12)
        assert all(np.array_equal(c.to_cpu(), e) for c, e in zip(pcols, pexp))


def test_accumulation_ops(hb):
    rng = np.random.default_rng(9)
    a = rand_cols(rng, 4, 6)
    b = rand_cols(rng, 4, 6)
    da = [hb.from_cpu(x) for x in a]
    hb.accumulate(da, [hb.from_cpu(x) for x in b], 6)
    for x, y, d in zip(a, b, da):
        assert np.array_equal(d.to_cpu(), (x.astype(np.uint64) + y) % P)
    # lift_and_accumulate: out[i] = sum_g col_g[(i >> (r+1) << 1) + (i & 1)]
    small = rand_cols(rng, 4, 3)
    big = rand_cols(rng, 4, 6)
    out = hb.lift_and_accumulate([[hb.from_cpu(x) for x in small], [hb.from_cpu(x) for x in big]], [3, 6])
    i = np.arange(1 << 6)
    li = ((i >> 4) << 1) + (i & 1)
    for k in range(4):
        assert np.array_equal(out[k].to_cpu(), (big[k].astype(np.uint64) + small[k][li]) % P)
    pw = hb.generate_secure_powers([3, 0, 0, 0], 5)
    assert [p[0] for p in pw] == [1, 3, 9, 27, 81]


# ---------------------------------------------------------------- whole proofs --------------
WF = dict(pow_bits=5, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)


def cfg_of(d):
    return SH.PcsConfig.make(d["pow_bits"], d["log_blowup"], d["log_last_layer"], d["n_queries"], d["fold_step"], 0)


@pytest.mark.parametrize("log_n,n_cols,mode", [(3, 16, 0), (5, 16, 1), (8, 100, 1), (10, 37, 0), (12, 100, 1),
                                               (18, 4, 1)])
def test_wide_fib_proof_byte_identical(hb, log_n, n_cols, mode):
    say("wide-fib proof", log_n, n_cols, mode)
    trace = hb.gen_trace_wide_fibonacci(n_cols, log_n, mode, 21)
    pr = hb.prove_wide_fibonacci(trace, log_n, cfg_of(WF))
    st, proof, dig, _, _ = O.prove(air="wide_fib", log_n_rows=log_n, n_cols=n_cols, input_mode=mode, seed=21, **WF)
    assert pr.status == st == 0
    assert pr.bytes == proof
    assert pr.digest == dig
    assert O.verify(pr.bytes, air="wide_fib", log_n_rows=log_n, n_cols=n_cols, **WF) == ""


@pytest.mark.parametrize("cfg", [dict(pow_bits=8, log_blowup=2, log_last_layer=2, n_queries=7, fold_step=1),
                                 dict(pow_bits=3, log_blowup=1, log_last_layer=1, n_queries=5, fold_step=2),
                                 dict(pow_bits=4, log_blowup=1, log_last_layer=8, n_queries=4, fold_step=1)])
def test_wide_fib_other_configs(hb, cfg):
    say("wide-fib config", cfg)
    trace = hb.gen_trace_wide_fibonacci(20, 9, 1, 5)
    pr = hb.prove_wide_fibonacci(trace, 9, cfg_of(cfg))
    st, proof, dig, _, _ = O.prove(air="wide_fib", log_n_rows=9, n_cols=20, input_mode=1, seed=5, **cfg)
    assert pr.bytes == proof and pr.digest == dig
    assert O.verify(pr.bytes, air="wide_fib", log_n_rows=9, n_cols=20, **cfg) == ""


@pytest.mark.parametrize("logs,ncols,cfg", [
    ([5, 7, 4], [16, 9, 30], WF),
    ([6, 6], [12, 20], dict(WF, log_blowup=2)),
    ([8, 5, 6, 7], [10, 40, 8, 17], dict(WF, pow_bits=3, n_queries=5)),
    ([12, 9, 10, 11, 9, 10], [24, 100, 8, 60, 33, 12], WF),
])
def test_wide_fib_multi_component_byte_identical(hb, logs, ncols, cfg):
    say("wide-fib components", logs, ncols)
    traces = [hb.gen_trace_wide_fibonacci(nc, lg, 1, 7 + i) for i, (lg, nc) in enumerate(zip(logs, ncols))]
    pr = hb.prove_wide_fibonacci_multi(traces, logs, cfg_of(cfg))
    st, proof, dig, _ = O.prove_multi(logs, ncols, input_mode=1, seed=7, **cfg)
    assert pr.status == st == 0
    assert pr.bytes == proof and pr.digest == dig
    assert O.verify_multi(pr.bytes, logs, ncols, **cfg) == ""


@pytest.mark.parametrize("log_n,ipr,mode", [(5, 1, 0), (6, 2, 1), (7, 8, 1), (10, 1, 1)])
def test_poseidon_proof_byte_identical(hb, log_n, ipr, mode):
    say("poseidon proof", log_n, ipr, mode)
    trace = hb.gen_trace_poseidon(ipr, log_n, mode, 33)
    pr = hb.prove_poseidon(trace, ipr, log_n, cfg_of(WF), allow_unsatisfied=True)
    st, proof, dig, cs, _ = O.prove(air="poseidon", log_n_rows=log_n, ipr=ipr, input_mode=mode, seed=33, **WF)
    assert pr.claimed_sum == tuple(map(int, cs))
    assert pr.status == st
    assert pr.bytes == proof
    assert pr.digest == dig


def test_poseidon_unsatisfied_proof_is_an_error_by_default(hb):
    """prove_ex returns Err(ConstraintsNotSatisfied) for this AIR (prover/mod.rs:127-144): the
    library reports status 6 unless the caller asks for the bytes."""
    trace = hb.gen_trace_poseidon(1, 5, 1, 33)
    with pytest.raises(SH.StwoHipError) as e:
        hb.prove_poseidon(trace, 1, 5, cfg_of(WF))
    assert e.value.status == 6
    pr = hb.prove_poseidon(trace, 1, 5, cfg_of(WF), allow_unsatisfied=True)
    assert pr.status == 2 and len(pr.bytes) > 0


def test_compression_rate_and_merkle_units(hb):
    """The Blake2s compression ceiling is plausible (a compression is about a thousand VALU
    instructions, the device issues tens of trillions per second), and the merkle class counts one
    compression per 64-byte block of a node's message, one for an empty message."""
    rate = hb.measure_compression_rate()
    say(f"compressions per second: {rate:.3e}")
    assert 1e9 < rate < 1e12
    hb.reset_kernel_stats()
    hb.set_kernel_timing(True)
    rng = np.random.default_rng(5)
    cols = [hb.from_cpu(rng.integers(0, P, 1 << 6, dtype=np.uint32)) for _ in range(19)]
    leaves = hb.commit_on_layer(6, None, cols)  # 76 bytes: 2 compressions per node
    mid = hb.commit_on_layer(5, leaves, cols[:5])  # 64 + 20 bytes: 2 per node
    hb.commit_on_layer(4, mid, [])  # children only: 1 per node
    hb.commit_on_layer(0, None, [])  # the empty message: 1
    hb.set_kernel_timing(False)
    assert hb.kernel_units()["merkle"] == 2 * 64 + 2 * 32 + 16 + 1


def test_stream_ordered_chain(hb):
    """Trait entries return without waiting for the device: a chain of them (interpolate, evaluate
    with blowup, the Merkle tree of the evaluations level by level, a fold) with no read in between
    gives the oracle's results when read at the end."""
    say("stream-ordered chain")
    rng = np.random.default_rng(21)
    log_n, n_cols = 12, 8
    cols = rand_cols(rng, n_cols, log_n)
    coeffs = hb.interpolate_columns([hb.from_cpu(c) for c in cols], log_n)
    ev = hb.evaluate_polynomials(coeffs, log_n, 1)
    layers = [hb.commit_on_layer(log_n + 1, None, ev)]
    for ls in range(log_n, -1, -1):
        layers.append(hb.commit_on_layer(ls, layers[-1], []))
    alpha = qm(rng)
    dst = [hb.column(1 << log_n, zeroed=True) for _ in range(4)]
    hb.fold_circle_into_line(dst, ev[:4], alpha, log_n + 1)
    # the reads, in reverse order of the launches
    exp_ev = [O.evaluate(O.interpolate(c), log_n + 1) for c in cols]
    fold = O.fold_circle_into_line(exp_ev[:4], [np.zeros(1 << log_n, np.uint32)] * 4, alpha)
    for d, e in zip(dst, fold):
        assert np.array_equal(d.to_cpu(), e)
    prev = O.commit_on_layer(log_n + 1, None, exp_ev)
    leaf = prev
    for ls in range(log_n, -1, -1):
        prev = O.commit_on_layer(ls, prev, [])
    assert layers[-1].to_cpu().view(np.uint8).tobytes() == prev.tobytes()
    assert layers[0].to_cpu().view(np.uint8).tobytes() == leaf.tobytes()
    for e, x in zip(ev, exp_ev):
        assert np.array_equal(e.to_cpu(), x)
