"""Writes tests/golden/transcripts.json: whole proofs of tiny wide-Fibonacci and Poseidon2 instances
computed by the pure-Python restatement of the protocol (tests/pystark.py, written from the
reference's definitions and the mathematics). Each case records its parameters, the proof's length
and SHA-256, the final channel digest and whether prove_ex's sanity check passes; the smallest
proofs are stored whole (hex).

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import pystark as S  # noqa: E402

WF = dict(pow_bits=5, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)
CASES = [
    # air, log_n, width (columns or instances per row), mode, seed, config
    ("wide_fib", 3, 8, 0, 0, WF),
    ("wide_fib", 3, 16, 1, 2024, WF),
    ("wide_fib", 4, 5, 0, 0, dict(pow_bits=3, log_blowup=2, log_last_layer=0, n_queries=5, fold_step=1)),
    ("wide_fib", 5, 6, 1, 3, dict(pow_bits=4, log_blowup=1, log_last_layer=2, n_queries=4, fold_step=1)),
    ("wide_fib", 5, 4, 1, 11, dict(pow_bits=2, log_blowup=1, log_last_layer=1, n_queries=6, fold_step=2)),
    ("wide_fib", 4, 3, 0, 0, dict(pow_bits=6, log_blowup=3, log_last_layer=1, n_queries=7, fold_step=3)),
    ("wide_fib", 6, 10, 1, 5, dict(pow_bits=8, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)),
    ("poseidon", 3, 1, 0, 0, WF),
    ("poseidon", 3, 2, 1, 9, WF),
    ("poseidon", 4, 1, 1, 33, dict(pow_bits=4, log_blowup=1, log_last_layer=1, n_queries=4, fold_step=1)),
]


def make():
    out = []
    for air, log_n, width, mode, seed, cfg in CASES:
        a = S.WideFib(log_n, width, mode, seed) if air == "wide_fib" else S.Poseidon(log_n, width, mode, seed)
        proof, digest, ok = S.prove(a, cfg)
        rec = dict(air=air, log_n_rows=log_n, width=width, input_mode=mode, seed=seed, config=cfg,
                   proof_len=len(proof), proof_sha256=hashlib.sha256(proof).hexdigest(), digest=digest.hex(),
                   constraints_satisfied=ok)
        if len(proof) <= 1600:
            rec["proof_hex"] = proof.hex()
        out.append(rec)
        print(air, log_n, width, len(proof), ok, flush=True)
    return out


if __name__ == "__main__":
    recs = make()
    with open(os.path.join(HERE, "transcripts.json"), "w") as f:
        json.dump(recs, f, indent=1)
        f.write("\n")
