"""Benchmark: prove() on one MI355X through the C ABI of libstwo_hip.

Default workload: Poseidon2 at 2^18 instances, 8 instances per row (2^15 rows x 1264 trace
columns + 32 LogUp interaction columns), log_blowup 1, 3 queries, last-layer degree bound 2^0,
the reference's default 10 PoW bits; seeded uniform [0, P) initial states. One step = one whole
proof from a trace resident in HBM: trace commit, LogUp interaction trace and commit,
composition, OODS, quotients, FRI, proof of work and decommitment, proof bytes on the host.

Prints one JSON line (see the keys below). The metric comes from a first timed region of
`--steps` proofs; a second timed region of the same length records a HIP event pair around
every kernel launch group on the library's stream, and the roofline entry is the kernel class
with the largest device time there.
The CPU baseline is the oracle (a multi-threaded C++ port of the reference prover) proving
the same workload with the same seed once, on the host's cores; its proof bytes are compared
with the GPU proof.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def load_pkg():
    import importlib.util
    spec = importlib.util.spec_from_file_location("stwo_hip", os.path.join(ROOT, "nitroozk-stwo_amd", "stwo_hip.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def device_sync(hb):
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    hb.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["poseidon", "wide_fib"], default="poseidon")
    ap.add_argument("--log-instances", type=int, default=18)
    ap.add_argument("--ipr", type=int, default=8, help="Poseidon instances per row")
    ap.add_argument("--n-cols", type=int, default=100, help="wide-fib columns")
    ap.add_argument("--pow-bits", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box of this pool has one MI355X; --gpus %d is not run "
              "(see DESIGN.md for how the prover shards)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    m = load_pkg()
    hb = m.HipBackend(0)
    cfg_d = dict(pow_bits=args.pow_bits, log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1)
    cfg = m.PcsConfig.make(args.pow_bits, 1, 0, 3, 1, 0)
    if args.workload == "poseidon":
        log_rows = args.log_instances - (args.ipr.bit_length() - 1)
        trace = hb.gen_trace_poseidon(args.ipr, log_rows, 1, args.seed)
        prove = lambda: hb.prove_poseidon(trace, args.ipr, log_rows, cfg)
        metric = "prove ms (Poseidon2 2^%d)" % args.log_instances
        workload = {"workload": "poseidon2", "log_instances": args.log_instances, "instances_per_row": args.ipr,
                    "log_n_rows": log_rows, "trace_columns": 158 * args.ipr, "interaction_columns": 4 * args.ipr}
        oracle_kw = dict(air="poseidon", log_n_rows=log_rows, ipr=args.ipr)
    else:
        log_rows = args.log_instances
        trace = hb.gen_trace_wide_fibonacci(args.n_cols, log_rows, 1, args.seed)
        prove = lambda: hb.prove_wide_fibonacci(trace, log_rows, cfg)
        metric = "prove ms (wide-Fibonacci 2^%d x %d)" % (log_rows, args.n_cols)
        workload = {"workload": "wide_fibonacci", "log_n_rows": log_rows, "trace_columns": args.n_cols}
        oracle_kw = dict(air="wide_fib", log_n_rows=log_rows, n_cols=args.n_cols)
    workload.update({"log_blowup": 1, "n_queries": 3, "pow_bits": args.pow_bits, "log_last_layer_degree_bound": 0,
                     "channel": "blake2s", "seed": args.seed})

    device_sync(hb)  # brings up torch's device context before anything is timed
    for _ in range(args.warmup):
        prove()

    def timed(n):
        device_sync(hb)
        times, proofs = [], []
        t0 = time.perf_counter()
        for _ in range(n):
            s = time.perf_counter()
            proofs.append(prove())
            times.append((time.perf_counter() - s) * 1e3)
        device_sync(hb)
        return (time.perf_counter() - t0) * 1e3, times, proofs

    # timed region 1: the metric (no per-kernel events on the stream)
    total_ms, times, proofs = timed(args.steps)
    # timed region 2: the same steps with a HIP event pair around every kernel launch group on the
    # library's stream, for the per-kernel table and the roofline entry
    hb.reset_kernel_stats()
    hb.set_kernel_timing(True)
    ev_total_ms, _, ev_proofs = timed(args.steps)
    hb.set_kernel_timing(False)
    stats = hb.kernel_stats()
    proofs += ev_proofs
    ms_per_step = total_ms / args.steps
    last = proofs[-1]
    deterministic = all(p.bytes == last.bytes for p in proofs)

    # roofline: kernel class with the largest device time
    dom = max((k for k in stats if k != "grind"), key=lambda k: stats[k][0])
    ms, by, n = stats[dom]
    achieved = by / (ms / 1e3) / 1e9
    kernels = {k: {"ms_per_step": round(v[0] / args.steps, 3), "launches_per_step": v[2] // args.steps,
                   "GB_per_s": round(v[1] / (v[0] / 1e3) / 1e9, 1) if v[0] > 0 and v[1] > 0 else None}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])}
    stage = {k: round(v, 2) for k, v in last.stage_ms.items()}

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        import oracle_lib as O
        threads = O.lib().oracle_num_threads()
        st, obytes, odig, _, oms = O.prove(input_mode=1, seed=args.seed, **cfg_d, **oracle_kw)
        parity = obytes == last.bytes and odig == last.digest
        cpu = {"value": round(oms, 1), "unit": "ms", "cores": threads, "kind": "port",
               "sample": "one whole %s proof of the same workload and seed (trace generation excluded), "
                         "C++ port of the reference prover on the host's cores" % workload["workload"]}

    out = {
        "metric": metric,
        "value": round(ms_per_step, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (M31 field, exact)",
        "data": "synthetic: seeded uniform [0,P) inputs generated on the device (SplitMix64, seed %d)" % args.seed,
        "config": workload,
        "proofs_per_s": round(1e3 / ms_per_step, 2),
        "ms_per_step_with_kernel_events": round(ev_total_ms / args.steps, 3),
        "p50_ms": round(statistics.median(times), 3),
        "p99_ms": round(sorted(times)[max(0, int(round(0.99 * len(times))) - 1)], 3),
        "proof_status": last.status,
        "proof_bytes": len(last.bytes),
        "deterministic": deterministic,
        "byte_identical_to_oracle": parity,
        "stage_ms": stage,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    hb.close()


if __name__ == "__main__":
    main()
