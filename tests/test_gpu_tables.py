"""The prover's table store under pressure. With a 4 KiB store (STWO_HIP_TABLE_STORE_BYTES) every
proof overflows it: tables laid out before a commitment's root is read, and launched on after the
staging arena has been rewound, go to buffers of their own, and the store is emptied between
calls. Proofs of mixed shapes on such contexts, in sequence and from two threads at once, equal
the proofs of a context with the default 16 MiB store byte for byte, and the wide-Fibonacci ones
verify."""
import os
import threading

import pytest

from hip_lib import SH

pytestmark = pytest.mark.gpu
CFG = {"wide_fib": (5, 1, 0, 3, 1, 0), "poseidon": (10, 1, 0, 3, 1, 0)}
SHAPES = [("wide_fib", 6, 8), ("poseidon", 5, 1), ("wide_fib", 9, 30), ("wide_fib", 11, 57), ("poseidon", 6, 2),
          ("wide_fib", 9, 31)]


def prove(hb, kind, log_n, width, seed):
    cfg = SH.PcsConfig.make(*CFG[kind])
    if kind == "wide_fib":
        tr = hb.gen_trace_wide_fibonacci(width, log_n, 1, seed)
        pr = hb.prove_wide_fibonacci(tr, log_n, cfg)
    else:
        tr = hb.gen_trace_poseidon(width, log_n, 1, seed)
        pr = hb.prove_poseidon(tr, width, log_n, cfg, allow_unsatisfied=True)
    hb.free_columns(tr)
    return pr.bytes


def test_small_table_store_gives_the_same_proofs():
    ref_hb = SH.HipBackend(0)
    try:
        ref = {(k, n, w): prove(ref_hb, k, n, w, 300 + i) for i, (k, n, w) in enumerate(SHAPES)}
    finally:
        ref_hb.close()
    old = os.environ.get("STWO_HIP_TABLE_STORE_BYTES")
    os.environ["STWO_HIP_TABLE_STORE_BYTES"] = "4096"
    try:
        # one context in sequence, three passes over the shapes
        hb = SH.HipBackend(0)
        try:
            for rep in range(3):
                for i, (k, n, w) in enumerate(SHAPES):
                    print("sequential pass", rep, k, n, w, flush=True)
                    assert prove(hb, k, n, w, 300 + i) == ref[(k, n, w)], (rep, k, n, w)
        finally:
            hb.close()
        # two contexts from two threads
        errors, lock = [], threading.Lock()

        def worker(order):
            b = SH.HipBackend(0)
            try:
                for rep in range(2):
                    for i in order:
                        k, n, w = SHAPES[i]
                        if prove(b, k, n, w, 300 + i) != ref[(k, n, w)]:
                            with lock:
                                errors.append((rep, k, n, w))
            except Exception as e:
                with lock:
                    errors.append(repr(e))
            finally:
                b.close()

        idx = list(range(len(SHAPES)))
        threads = [threading.Thread(target=worker, args=(o,)) for o in (idx, idx[::-1])]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors, errors
    finally:
        if old is None:
            del os.environ["STWO_HIP_TABLE_STORE_BYTES"]
        else:
            os.environ["STWO_HIP_TABLE_STORE_BYTES"] = old
    for (k, n, w), b in ref.items():
        if k == "wide_fib":
            assert SH.verify_wide_fibonacci(b, [n], [w], SH.PcsConfig.make(*CFG[k])) == ""
