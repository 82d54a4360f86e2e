"""The no-allocation entry point stwo_hip_batch_eval_at_point_noalloc, called through the C ABI,
against stwo_hip_batch_eval_at_point (the entry point whose work it repeats without allocating):
two calls in a row, the second smaller, then a size beyond the scratch's limit."""
import ctypes
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


def random_point(tag):
    dig = np.frombuffer(hashlib.blake2s(tag, digest_size=32).digest(), np.uint8).copy()
    pt = np.zeros(8, np.uint32)
    O.lib().oracle_random_point(O.u8ptr(dig), O.u32ptr(pt))
    return [int(x) for x in pt]


def call_noalloc(hb, cols, log_n, pt):
    out = (SH.QM31 * len(cols))()
    p = SH.CirclePointQM31(SH.QM31.of(pt[:4]), SH.QM31.of(pt[4:]))
    st = hb.lib.stwo_hip_batch_eval_at_point_noalloc(hb.ctx, SH._ptrs(cols), len(cols), log_n, p, out)
    return st, [o.tuple() for o in out]


def test_noalloc_matches_batch_eval_at_point_twice():
    hb = SH.HipBackend(0)
    try:
        for log_n, n_cols, tag in ((18, 40, b"na1"), (9, 5, b"na2")):
            say("noalloc", log_n, n_cols)
            rng = np.random.default_rng(log_n)
            cols = [hb.from_cpu(rng.integers(0, P, 1 << log_n, dtype=np.uint32)) for _ in range(n_cols)]
            pt = random_point(tag)
            st, got = call_noalloc(hb, cols, log_n, pt)
            assert st == 0, hb.lib.stwo_hip_last_error(hb.ctx)
            exp = hb.batch_eval_at_point(cols, log_n, pt)
            assert [tuple(g) for g in got] == [tuple(e) for e in exp]
        # beyond the limit: 4096 columns of 2^20 need 32 MiB of partial sums
        big = [hb.column(1 << 4, zeroed=True)] * 4096
        st, _ = call_noalloc(hb, big, 20, random_point(b"na3"))
        assert st == 3
    finally:
        hb.close()
