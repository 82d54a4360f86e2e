"""A small serving run on the GPU: mixed wide-Fibonacci and Poseidon2 requests proved concurrently
on two contexts from two host threads (the batched form), each proof compared byte for byte with
the oracle's and checked by the library's host verifier (wide-Fibonacci accepted; Poseidon2
rejected with OodsNotMatching, as the reference's own verifier rejects it)."""
import queue
import random
import sys
import threading

import pytest

import oracle_lib as O
from hip_lib import SH

pytestmark = pytest.mark.gpu
CFG = {"wide_fib": dict(pow_bits=5, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1),
       "poseidon": dict(pow_bits=10, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)}


def cfg_of(d):
    return SH.PcsConfig.make(d["pow_bits"], d["log_blowup"], d["log_last_layer"], d["n_queries"], d["fold_step"], 0)


def test_concurrent_mixed_requests_match_the_oracle_and_verify():
    rnd = random.Random(5)
    reqs = []
    for i in range(12):
        if rnd.random() < 0.7:
            reqs.append((i, "wide_fib", rnd.randint(6, 12), rnd.randint(8, 40)))
        else:
            reqs.append((i, "poseidon", rnd.randint(4, 8), rnd.choice([1, 2, 8])))
    work = queue.Queue()
    for r in reqs:
        work.put(r)
    results, errors, lock = {}, [], threading.Lock()

    def worker():
        hb = SH.HipBackend(0)
        try:
            while True:
                try:
                    i, kind, log_n, width = work.get_nowait()
                except queue.Empty:
                    return
                if kind == "wide_fib":
                    tr = hb.gen_trace_wide_fibonacci(width, log_n, 1, 100 + i)
                    pr = hb.prove_wide_fibonacci(tr, log_n, cfg_of(CFG[kind]))
                else:
                    tr = hb.gen_trace_poseidon(width, log_n, 1, 100 + i)
                    pr = hb.prove_poseidon(tr, width, log_n, cfg_of(CFG[kind]), allow_unsatisfied=True)
                hb.free_columns(tr)
                with lock:
                    results[i] = pr
        except Exception as e:
            with lock:
                errors.append(repr(e))
        finally:
            hb.close()

    threads = [threading.Thread(target=worker) for _ in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert len(results) == len(reqs)
    for i, kind, log_n, width in reqs:
        pr = results[i]
        print("request", i, kind, log_n, width, flush=True)
        sys.stdout.flush()
        if kind == "wide_fib":
            st, proof, dig, _, _ = O.prove(air="wide_fib", log_n_rows=log_n, n_cols=width, input_mode=1, seed=100 + i,
                                           **CFG[kind])
            assert pr.bytes == proof and pr.digest == dig
            assert SH.verify_wide_fibonacci(pr.bytes, [log_n], [width], cfg_of(CFG[kind])) == ""
        else:
            st, proof, dig, cs, _ = O.prove(air="poseidon", log_n_rows=log_n, ipr=width, input_mode=1, seed=100 + i,
                                            **CFG[kind])
            assert pr.bytes == proof and pr.digest == dig
            assert SH.verify_poseidon(pr.bytes, width, log_n, pr.claimed_sum, cfg_of(CFG[kind])) == "OodsNotMatching"
