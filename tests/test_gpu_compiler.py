"""The composition compiler on the GPU: constraint programs recorded by framework.py, compiled
into kernels at run time through evaluate_constraint_quotients_on_domain (blob
"expr_component"), give the same accumulator words as the hand-written kernels and the oracle,
for the wide-Fibonacci and Poseidon2 AIRs."""
import importlib.util
import os
import struct
import sys
import time

import numpy as np
import pytest

import oracle_lib as O
from hip_lib import SH, ROOT

P = O.P
pytestmark = pytest.mark.gpu
_spec = importlib.util.spec_from_file_location("framework", os.path.join(ROOT, "nitroozk-stwo_amd", "framework.py"))
F = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(F)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hb():
    b = SH.HipBackend(0)
    yield b
    b.close()


def rand_cols(rng, n, log_n):
    return [rng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(n)]


def qm(rng):
    return tuple(int(x) for x in rng.integers(0, P, 4))


def run(hb, blob, trace, inter, pows, dinv, acc, trace_log, eval_log, shift=(0, 0, 0, 0)):
    dacc = [hb.from_cpu(c) for c in acc]
    hb.evaluate_constraint_quotients_on_domain(dacc, [hb.from_cpu(c) for c in trace], [hb.from_cpu(c) for c in inter],
                                               pows, dinv, trace_log, eval_log, blob, shift)
    return [d.to_cpu() for d in dacc]


@pytest.mark.parametrize("trace_log,eval_log,n_cols", [(3, 4, 3), (6, 7, 16), (10, 11, 100)])
def test_compiled_wide_fibonacci(hb, trace_log, eval_log, n_cols):
    rng = np.random.default_rng(trace_log + n_cols)
    trace = rand_cols(rng, n_cols, eval_log)
    acc = rand_cols(rng, 4, eval_log)
    pows = [qm(rng) for _ in range(n_cols - 2)]
    dinv = rng.integers(1, P, 1 << (eval_log - trace_log), dtype=np.uint32)
    t0 = time.perf_counter()
    got = run(hb, F.expr_blob(F.wide_fibonacci_program(n_cols), []), trace, [], pows, dinv, acc, trace_log, eval_log)
    say("compiled wide-fib %d columns: first call %.2f s" % (n_cols, time.perf_counter() - t0))
    hand = run(hb, struct.pack("<I", SH.fnv1a("fibonacci_example")), trace, [], pows, dinv, acc, trace_log, eval_log)
    exp = O.eval_constraints_on_domain("wide_fib", 1, trace_log, eval_log, trace, [], pows, dinv, acc)
    for k in range(4):
        assert np.array_equal(got[k], hand[k]) and np.array_equal(got[k], exp[k]), k


@pytest.mark.parametrize("trace_log,eval_log,ipr", [(4, 6, 1), (5, 7, 2), (7, 9, 8)])
def test_compiled_poseidon(hb, trace_log, eval_log, ipr):
    rng = np.random.default_rng(100 + ipr)
    trace = rand_cols(rng, 158 * ipr, eval_log)
    inter = rand_cols(rng, 4 * ipr, eval_log)
    acc = rand_cols(rng, 4, eval_log)
    pows = [qm(rng) for _ in range(143 * ipr)]
    z, alpha, shift = qm(rng), qm(rng), qm(rng)
    dinv = rng.integers(1, P, 1 << (eval_log - trace_log), dtype=np.uint32)
    t0 = time.perf_counter()
    blob = F.expr_blob(F.poseidon_program(ipr), F.poseidon_params(z, alpha))
    got = run(hb, blob, trace, inter, pows, dinv, acc, trace_log, eval_log, shift)
    say("compiled poseidon ipr %d: first call %.2f s" % (ipr, time.perf_counter() - t0))
    hand_blob = struct.pack("<II", SH.fnv1a("poseidon_example"), ipr) + struct.pack("<8I", *z, *alpha)
    hand = run(hb, hand_blob, trace, inter, pows, dinv, acc, trace_log, eval_log, shift)
    for k in range(4):
        assert np.array_equal(got[k], hand[k]), k
    if ipr <= 2:
        exp = O.eval_constraints_on_domain("poseidon", ipr, trace_log, eval_log, trace, inter, pows, dinv, acc, z=z,
                                           alpha=alpha, shift=shift)
        for k in range(4):
            assert np.array_equal(got[k], exp[k]), k


def test_malformed_program_is_rejected(hb):
    prog = F.wide_fibonacci_program(4)
    body = bytearray(prog.to_bytes())
    # node 0 made to read node 5: an operand that does not precede its node
    struct.pack_into("<5I", body, 20, F.OP_ADD, 5, 0, 0, 0)
    n = len(body) // 4
    h = SH.fnv1a("expr_component")
    blob = struct.pack("<II", h, n) + bytes(body) + struct.pack("<I", 0)
    acc = [hb.column(16, zeroed=True) for _ in range(4)]
    trace = [hb.column(16, zeroed=True) for _ in range(4)]
    with pytest.raises(SH.StwoHipError) as e:
        hb.evaluate_constraint_quotients_on_domain(acc, trace, [], [(0, 0, 0, 0)] * 2, np.ones(2, np.uint32), 3, 4,
                                                   blob)
    assert e.value.status == 1
