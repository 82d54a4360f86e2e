"""ctypes loader for the oracle library (oracle/_build/liboracle.so).

The oracle is test infrastructure: a CPU restatement of the reference prover. Only the test
suite, the smoke check and bench.py's CPU-baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
P = (1 << 31) - 1

_lib = None


class ProveArgs(ctypes.Structure):
    _fields_ = [
        ("air", ctypes.c_uint32),
        ("log_n_rows", ctypes.c_uint32),
        ("n_cols", ctypes.c_uint32),
        ("ipr", ctypes.c_uint32),
        ("input_mode", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("pow_bits", ctypes.c_uint32),
        ("log_blowup", ctypes.c_uint32),
        ("log_last_layer", ctypes.c_uint32),
        ("n_queries", ctypes.c_uint32),
        ("fold_step", ctypes.c_uint32),
        ("lifting_log_size", ctypes.c_uint32),
    ]


def make_args(air="wide_fib", log_n_rows=5, n_cols=16, ipr=1, input_mode=0, seed=0, pow_bits=5,
              log_blowup=1, log_last_layer=0, n_queries=3, fold_step=1, lifting_log_size=0):
    return ProveArgs(0 if air == "wide_fib" else 1, log_n_rows, n_cols, ipr, input_mode, seed, pow_bits,
                     log_blowup, log_last_layer, n_queries, fold_step, lifting_log_size)


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_prove.argtypes = [ctypes.POINTER(ProveArgs), ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t),
                                   u8p, u32p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_prove.restype = ctypes.c_int
        L.oracle_verify.argtypes = [ctypes.POINTER(ProveArgs), u8p, ctypes.c_size_t, u32p, ctypes.c_char_p,
                                    ctypes.c_size_t]
        L.oracle_verify.restype = ctypes.c_int
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_grind.restype = ctypes.c_uint64
        L.oracle_grind.argtypes = [u8p, ctypes.c_uint32]
        L.oracle_verify_pow.argtypes = [u8p, ctypes.c_uint32, ctypes.c_uint64]
        L.oracle_channel_mix_u64.argtypes = [ctypes.c_uint64, u8p]
        L.oracle_channel_mix_u32s.argtypes = [u32p, ctypes.c_size_t, u8p]
        L.oracle_channel_draw_count.restype = ctypes.c_uint32
        L.oracle_uniform_m31.restype = ctypes.c_uint32
        L.oracle_uniform_m31.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_gen_wide_fib.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, u32p]
        L.oracle_gen_poseidon.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, u32p]
        L.oracle_offset_bit_reversed_index.restype = ctypes.c_size_t
        L.oracle_offset_bit_reversed_index.argtypes = [ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                                       ctypes.c_long]
        L.oracle_blake2s.argtypes = [ctypes.c_char_p, ctypes.c_size_t, u8p]
        L.oracle_prove_wide_fib_multi.argtypes = [ctypes.POINTER(ProveArgs), ctypes.c_uint32, u32p, u32p,
                                                  ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t), u8p,
                                                  ctypes.POINTER(ctypes.c_double)]
        L.oracle_verify_wide_fib_multi.argtypes = [ctypes.POINTER(ProveArgs), ctypes.c_uint32, u32p, u32p, u8p,
                                                   ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
        _lib = L
    return _lib


def u32ptr(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def u8ptr(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def ptr_array(cols):
    arr = (ctypes.POINTER(ctypes.c_uint32) * len(cols))()
    for i, c in enumerate(cols):
        arr[i] = u32ptr(c)
    return arr


def prove(**kw):
    """Returns (status, proof_bytes, final_digest, claimed_sum, prove_ms)."""
    L = lib()
    a = make_args(**kw)
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    dig = np.zeros(32, np.uint8)
    cs = np.zeros(4, np.uint32)
    ms = ctypes.c_double()
    st = L.oracle_prove(ctypes.byref(a), ctypes.byref(out), ctypes.byref(n), u8ptr(dig), u32ptr(cs), ctypes.byref(ms))
    if st < 0:
        raise RuntimeError("oracle_prove failed")
    data = ctypes.string_at(out, n.value)
    L.oracle_free(out)
    return st, data, dig.tobytes(), cs, ms.value


def verify(proof, claimed_sum=None, **kw):
    """Returns '' when the proof is accepted, else the rejection reason."""
    L = lib()
    a = make_args(**kw)
    buf = np.frombuffer(proof, np.uint8).copy()
    cs = np.zeros(4, np.uint32) if claimed_sum is None else np.asarray(claimed_sum, np.uint32).copy()
    err = ctypes.create_string_buffer(256)
    L.oracle_verify(ctypes.byref(a), u8ptr(buf), len(proof), u32ptr(cs), err, 256)
    return err.value.decode()


def prove_multi(logs, ncols, **kw):
    """Several wide-Fibonacci components (component i seeded with seed + i) in one proof.
    Returns (status, proof_bytes, final_digest, prove_ms)."""
    L = lib()
    a = make_args(**kw)
    lg = np.asarray(logs, np.uint32).copy()
    nc = np.asarray(ncols, np.uint32).copy()
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    dig = np.zeros(32, np.uint8)
    ms = ctypes.c_double()
    st = L.oracle_prove_wide_fib_multi(ctypes.byref(a), len(lg), u32ptr(lg), u32ptr(nc), ctypes.byref(out),
                                       ctypes.byref(n), u8ptr(dig), ctypes.byref(ms))
    if st < 0:
        raise RuntimeError("oracle_prove_wide_fib_multi failed")
    data = ctypes.string_at(out, n.value)
    L.oracle_free(out)
    return st, data, dig.tobytes(), ms.value


def verify_multi(proof, logs, ncols, **kw):
    """Returns '' when the multi-component proof is accepted, else the rejection reason."""
    L = lib()
    a = make_args(**kw)
    lg = np.asarray(logs, np.uint32).copy()
    nc = np.asarray(ncols, np.uint32).copy()
    buf = np.frombuffer(proof, np.uint8).copy()
    err = ctypes.create_string_buffer(256)
    L.oracle_verify_wide_fib_multi(ctypes.byref(a), len(lg), u32ptr(lg), u32ptr(nc), u8ptr(buf), len(proof), err, 256)
    return err.value.decode()


def interpolate(vals):
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().oracle_interpolate(ctypes.c_uint32(int(np.log2(len(v)))), u32ptr(v))
    return v


def evaluate(coeffs, log_out):
    c = np.ascontiguousarray(coeffs, np.uint32)
    out = np.zeros(1 << log_out, np.uint32)
    lib().oracle_evaluate(ctypes.c_uint32(int(np.log2(len(c)))), u32ptr(c), ctypes.c_uint32(log_out), u32ptr(out))
    return out


def twiddles(log_domain):
    n = 1 << (log_domain - 1)
    tw = np.zeros(n, np.uint32)
    itw = np.zeros(n, np.uint32)
    lib().oracle_twiddles(ctypes.c_uint32(log_domain), u32ptr(tw), u32ptr(itw))
    return tw, itw


def eval_at_point(coeffs, point):
    c = np.ascontiguousarray(coeffs, np.uint32)
    p = np.asarray(point, np.uint32).copy()
    out = np.zeros(4, np.uint32)
    lib().oracle_eval_at_point(ctypes.c_uint32(int(np.log2(len(c)))), u32ptr(c), u32ptr(p), u32ptr(out))
    return out


def barycentric_weights(log_n, point):
    p = np.asarray(point, dtype=np.uint32)
    out = np.zeros(4 << log_n, dtype=np.uint32)
    lib().oracle_barycentric_weights(ctypes.c_uint32(log_n), u32ptr(p), u32ptr(out))
    return out.reshape(4, 1 << log_n)


def barycentric_eval(evals, point):
    e = np.ascontiguousarray(evals, np.uint32)
    p = np.asarray(point, np.uint32).copy()
    out = np.zeros(4, np.uint32)
    lib().oracle_barycentric_eval(ctypes.c_uint32(int(np.log2(len(e)))), u32ptr(e), u32ptr(p), u32ptr(out))
    return out


def commit_on_layer(log_size, prev, cols):
    out = np.zeros((1 << log_size) * 32, np.uint8)
    cols = [np.ascontiguousarray(c, np.uint32) for c in cols]
    arr = ptr_array(cols)
    prev_p = None if prev is None else u8ptr(np.ascontiguousarray(prev, np.uint8))
    lib().oracle_commit_on_layer(ctypes.c_uint32(log_size), prev_p, ctypes.c_uint32(len(cols)), arr, u8ptr(out))
    return out


def merkle_root(cols):
    cols = [np.ascontiguousarray(c, np.uint32) for c in cols]
    logs = np.array([int(np.log2(len(c))) for c in cols], np.uint32)
    root = np.zeros(32, np.uint8)
    lib().oracle_merkle_root(ctypes.c_uint32(len(cols)), ptr_array(cols), u32ptr(logs), u8ptr(root))
    return root.tobytes()


def fold_circle_into_line(src4, dst4, alpha):
    src4 = [np.ascontiguousarray(c, np.uint32) for c in src4]
    dst4 = [np.ascontiguousarray(c, np.uint32).copy() for c in dst4]
    a = np.asarray(alpha, np.uint32).copy()
    lib().oracle_fold_circle_into_line(ctypes.c_uint32(int(np.log2(len(src4[0])))), ptr_array(src4),
                                       ptr_array(dst4), u32ptr(a))
    return dst4


def fold_line(src4, domain_log_first, alpha):
    src4 = [np.ascontiguousarray(c, np.uint32) for c in src4]
    n = len(src4[0])
    dst4 = [np.zeros(n // 2, np.uint32) for _ in range(4)]
    a = np.asarray(alpha, np.uint32).copy()
    lib().oracle_fold_line(ctypes.c_uint32(int(np.log2(n))), ctypes.c_uint32(domain_log_first), ptr_array(src4),
                           ptr_array(dst4), u32ptr(a))
    return dst4


def grind(digest, bits):
    d = np.frombuffer(digest, np.uint8).copy()
    return lib().oracle_grind(u8ptr(d), bits)


def verify_pow(digest, bits, nonce):
    d = np.frombuffer(digest, np.uint8).copy()
    return bool(lib().oracle_verify_pow(u8ptr(d), bits, nonce))


def blake2s(data: bytes):
    out = np.zeros(32, np.uint8)
    lib().oracle_blake2s(data, len(data), u8ptr(out))
    return out.tobytes()


def gen_wide_fib(log_n, n_cols, mode, seed):
    out = np.zeros(n_cols << log_n, np.uint32)
    lib().oracle_gen_wide_fib(log_n, n_cols, mode, seed, u32ptr(out))
    return out.reshape(n_cols, 1 << log_n)


def gen_poseidon(log_n, ipr, mode, seed):
    out = np.zeros((158 * ipr) << log_n, np.uint32)
    lib().oracle_gen_poseidon(log_n, ipr, mode, seed, u32ptr(out))
    return out.reshape(158 * ipr, 1 << log_n)


def _q4(v):
    return np.asarray(v, np.uint32).reshape(4)


def accumulate_numerators(cols, batches):
    """QuotientOps::accumulate_numerators over columns of one size. batches: list of
    (point8, [(col_idx, value4, rand4), ...]). Returns ([4 partial columns per batch], [first_linear4])."""
    cols = [np.ascontiguousarray(c, np.uint32) for c in cols]
    n = len(cols[0])
    log_size = n.bit_length() - 1
    pts = np.array([list(pt) for pt, _ in batches], np.uint32).reshape(-1)
    starts, idx, vals, rnds = [0], [], [], []
    for _, ents in batches:
        for ci, v, r in ents:
            idx.append(ci)
            vals += list(v)
            rnds += list(r)
        starts.append(len(idx))
    starts = np.array(starts, np.uint32)
    idx = np.array(idx, np.uint32)
    vals = np.array(vals, np.uint32)
    rnds = np.array(rnds, np.uint32)
    outs = [np.zeros(n, np.uint32) for _ in range(4 * len(batches))]
    fl = np.zeros(4 * len(batches), np.uint32)
    lib().oracle_accumulate_numerators(ctypes.c_uint32(log_size), ctypes.c_uint32(len(cols)), ptr_array(cols),
                                       ctypes.c_uint32(len(batches)), u32ptr(pts), u32ptr(starts), u32ptr(idx),
                                       u32ptr(vals), u32ptr(rnds), ptr_array(outs), u32ptr(fl))
    return [outs[4 * b:4 * b + 4] for b in range(len(batches))], [tuple(int(x) for x in fl[4 * b:4 * b + 4])
                                                                   for b in range(len(batches))]


def compute_quotients_and_combine(accs, lifting_log_size):
    """accs: list of (point8, [4 partial columns], first_linear4). Returns 4 columns of 2^lifting_log_size."""
    pts = np.array([list(a[0]) for a in accs], np.uint32).reshape(-1)
    parts = [np.ascontiguousarray(c, np.uint32) for a in accs for c in a[1]]
    logs = np.array([len(a[1][0]).bit_length() - 1 for a in accs], np.uint32)
    fl = np.array([list(a[2]) for a in accs], np.uint32).reshape(-1)
    out = [np.zeros(1 << lifting_log_size, np.uint32) for _ in range(4)]
    lib().oracle_compute_quotients_and_combine(ctypes.c_uint32(len(accs)), u32ptr(pts), ptr_array(parts), u32ptr(logs),
                                               u32ptr(fl), ctypes.c_uint32(lifting_log_size), ptr_array(out))
    return out


def eval_constraints_on_domain(air, ipr, trace_log, eval_log, trace, inter, pows, denom_inv, acc4, z=(0, 0, 0, 0),
                               alpha=(0, 0, 0, 0), shift=(0, 0, 0, 0)):
    """ComponentProver::evaluate_constraint_quotients_on_domain for one component; returns the new acc4."""
    trace = [np.ascontiguousarray(c, np.uint32) for c in trace]
    inter = [np.ascontiguousarray(c, np.uint32) for c in inter]
    acc = [np.ascontiguousarray(c, np.uint32).copy() for c in acc4]
    pw = np.array([list(p) for p in pows], np.uint32).reshape(-1)
    dinv = np.ascontiguousarray(denom_inv, np.uint32)
    lib().oracle_eval_constraints_on_domain(
        ctypes.c_uint32(0 if air == "wide_fib" else 1), ctypes.c_uint32(ipr), ctypes.c_uint32(trace_log),
        ctypes.c_uint32(eval_log), ptr_array(trace), ctypes.c_uint32(len(trace)),
        ptr_array(inter) if inter else None, ctypes.c_uint32(len(inter)), u32ptr(pw), ctypes.c_uint32(len(pows)),
        u32ptr(dinv), u32ptr(_q4(z).copy()), u32ptr(_q4(alpha).copy()), u32ptr(_q4(shift).copy()), ptr_array(acc))
    return acc
