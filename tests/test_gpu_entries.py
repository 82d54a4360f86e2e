"""GPU parity of the trait entry points that the library's own prover does not call in the same
form: accumulate_numerators and compute_quotients_and_combine (line coefficients from the host
arguments), and evaluate_constraint_quotients_on_domain for both component blobs. Each is called
through the C ABI and compared word for word with the oracle on the same seeded inputs."""
import hashlib
import struct
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
    return tuple(int(x) for x in rng.integers(0, P, 4))


def random_point(tag):
    dig = np.frombuffer(hashlib.blake2s(tag, digest_size=32).digest(), np.uint8).copy()
    pt = np.zeros(8, np.uint32)
    O.lib().oracle_random_point(O.u8ptr(dig), O.u32ptr(pt))
    return [int(x) for x in pt]


def make_batches(rng, n_cols, n_batches, tag):
    out = []
    for b in range(n_batches):
        k = int(rng.integers(1, n_cols + 1))
        idx = sorted(rng.choice(n_cols, size=k, replace=False).tolist())
        out.append((random_point(tag + bytes([b])), [(int(i), qm(rng), qm(rng)) for i in idx]))
    return out


@pytest.mark.parametrize("log_n,n_cols,n_batches,misalign", [(3, 5, 1, False), (6, 20, 3, True), (10, 37, 2, False),
                                                             (12, 64, 4, True), (16, 100, 2, False)])
def test_accumulate_numerators(hb, log_n, n_cols, n_batches, misalign):
    say("accumulate_numerators", log_n, n_cols, n_batches, misalign)
    rng = np.random.default_rng(1000 + log_n)
    cols = rand_cols(rng, n_cols, log_n)
    if misalign:
        # columns that are not 16-byte aligned take the kernel's word-load path
        big = hb.from_cpu(np.concatenate([np.zeros(1, np.uint32)] + cols))
        dev = [big.ptr + 4 * (1 + (c << log_n)) for c in range(n_cols)]
    else:
        dev = [hb.from_cpu(c) for c in cols]
    batches = make_batches(rng, n_cols, n_batches, b"acc%d" % log_n)
    parts, fl = hb.accumulate_numerators(dev, log_n, batches)
    eparts, efl = O.accumulate_numerators(cols, batches)
    assert fl == efl
    for b in range(n_batches):
        for k in range(4):
            assert np.array_equal(parts[b][k].to_cpu(), eparts[b][k]), (b, k)


@pytest.mark.parametrize("logs,lifting", [([4], 4), ([3, 5, 5], 5), ([7, 9, 10, 10], 10), ([12, 14, 16], 16)])
def test_compute_quotients_and_combine(hb, logs, lifting):
    say("compute_quotients_and_combine", logs, lifting)
    rng = np.random.default_rng(2000 + lifting)
    accs, dev = [], []
    for i, lg in enumerate(logs):
        part = rand_cols(rng, 4, lg)
        pt = random_point(b"q%d-%d" % (lifting, i))
        fl = qm(rng)
        accs.append((pt, part, fl))
        dev.append((pt, [hb.from_cpu(c) for c in part], lg, fl))
    got = hb.compute_quotients_and_combine(dev, lifting)
    exp = O.compute_quotients_and_combine(accs, lifting)
    for k in range(4):
        assert np.array_equal(got[k].to_cpu(), exp[k]), k


def test_numerators_then_quotients_match_oracle_chain(hb):
    """The two quotient entry points chained as compute_fri_quotients chains them (two size groups,
    the smaller lifted), against the oracle's chain."""
    rng = np.random.default_rng(77)
    groups = [(9, 12, 2), (11, 30, 3)]
    oaccs, daccs = [], []
    for lg, nc, nb in groups:
        cols = rand_cols(rng, nc, lg)
        batches = make_batches(rng, nc, nb, b"chain%d" % lg)
        parts, fl = hb.accumulate_numerators([hb.from_cpu(c) for c in cols], lg, batches)
        eparts, efl = O.accumulate_numerators(cols, batches)
        for b, (pt, _) in enumerate(batches):
            daccs.append((pt, parts[b], lg, fl[b]))
            oaccs.append((pt, eparts[b], efl[b]))
    got = hb.compute_quotients_and_combine(daccs, 11)
    exp = O.compute_quotients_and_combine(oaccs, 11)
    for k in range(4):
        assert np.array_equal(got[k].to_cpu(), exp[k])


def denom_inv(log_expand):
    # any nonzero words: the entry point multiplies by them as given
    rng = np.random.default_rng(5 + log_expand)
    return rng.integers(1, P, 1 << log_expand, dtype=np.uint32)


@pytest.mark.parametrize("trace_log,eval_log,n_cols", [(3, 4, 5), (6, 7, 16), (8, 10, 40), (12, 13, 100)])
def test_constraint_quotients_wide_fibonacci(hb, trace_log, eval_log, n_cols):
    say("constraints wide-fib", trace_log, eval_log, n_cols)
    rng = np.random.default_rng(3000 + trace_log)
    trace = rand_cols(rng, n_cols, eval_log)
    acc = rand_cols(rng, 4, eval_log)
    pows = [qm(rng) for _ in range(n_cols - 2)]
    dinv = denom_inv(eval_log - trace_log)
    dacc = [hb.from_cpu(c) for c in acc]
    blob = struct.pack("<I", SH.fnv1a("fibonacci_example"))
    hb.evaluate_constraint_quotients_on_domain(dacc, [hb.from_cpu(c) for c in trace], [], pows, dinv, trace_log,
                                               eval_log, blob)
    exp = O.eval_constraints_on_domain("wide_fib", 1, trace_log, eval_log, trace, [], pows, dinv, acc)
    for k in range(4):
        assert np.array_equal(dacc[k].to_cpu(), exp[k]), k


@pytest.mark.parametrize("trace_log,eval_log,ipr", [(4, 6, 1), (5, 7, 2), (7, 9, 8)])
def test_constraint_quotients_poseidon(hb, trace_log, eval_log, ipr):
    say("constraints poseidon", trace_log, eval_log, ipr)
    rng = np.random.default_rng(4000 + trace_log)
    trace = rand_cols(rng, 158 * ipr, eval_log)
    inter = rand_cols(rng, 4 * ipr, eval_log)
    acc = rand_cols(rng, 4, eval_log)
    pows = [qm(rng) for _ in range(143 * ipr)]
    z, alpha, shift = qm(rng), qm(rng), qm(rng)
    dinv = denom_inv(eval_log - trace_log)
    dacc = [hb.from_cpu(c) for c in acc]
    blob = struct.pack("<II", SH.fnv1a("poseidon_example"), ipr) + struct.pack("<8I", *z, *alpha)
    hb.evaluate_constraint_quotients_on_domain(dacc, [hb.from_cpu(c) for c in trace], [hb.from_cpu(c) for c in inter],
                                               pows, dinv, trace_log, eval_log, blob, shift)
    exp = O.eval_constraints_on_domain("poseidon", ipr, trace_log, eval_log, trace, inter, pows, dinv, acc, z=z,
                                       alpha=alpha, shift=shift)
    for k in range(4):
        assert np.array_equal(dacc[k].to_cpu(), exp[k]), k


def test_constraint_quotients_unknown_component(hb):
    acc = [hb.column(1 << 4, zeroed=True) for _ in range(4)]
    trace = [hb.column(1 << 4, zeroed=True) for _ in range(3)]
    blob = struct.pack("<I", SH.fnv1a("no_such_component"))
    with pytest.raises(SH.StwoHipError) as e:
        hb.evaluate_constraint_quotients_on_domain(acc, trace, [], [(0, 0, 0, 0)], denom_inv(1), 3, 4, blob)
    assert e.value.status == 5
