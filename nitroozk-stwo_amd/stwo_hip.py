"""Python host side of libstwo_hip: a HipBackend with the reference's prover-backend trait
methods (PolyOps, MerkleOps<Blake2sMerkleHasher>, FriOps, QuotientOps, AccumulationOps,
GrindOps, ComponentProver) and the two example provers, over the C ABI in include/stwo_hip.h.

Device memory is owned by DeviceColumn objects (the analogue of BaseFieldVec,
crates/stwo/src/stwo_cuda/base_field_vec.rs:7-140). Every call goes to the HIP library; there is
no CPU fallback, and constructing a backend fails loudly when the library or the GPU is missing.
"""
import collections.abc
import ctypes
import operator
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libstwo_hip.so")
P = (1 << 31) - 1

STATUS = {0: "ok", 1: "invalid argument", 2: "HIP error", 3: "unsupported size", 4: "protocol failure",
          5: "unknown component", 6: "constraints not satisfied"}
ST_UNSATISFIED = 6


class StwoHipError(RuntimeError):
    """A non-zero status from the library; .status holds the C-ABI code."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


class QM31(ctypes.Structure):
    _fields_ = [("a", ctypes.c_uint32), ("b", ctypes.c_uint32), ("c", ctypes.c_uint32), ("d", ctypes.c_uint32)]

    @staticmethod
    def of(v):
        return QM31(*[int(x) for x in v])

    def tuple(self):
        return (self.a, self.b, self.c, self.d)


class CirclePointQM31(ctypes.Structure):
    _fields_ = [("x", QM31), ("y", QM31)]


class PcsConfig(ctypes.Structure):
    """PcsConfig + FriConfig (crates/stwo/src/core/pcs/mod.rs:29-80, core/fri.rs:32-81)."""
    _fields_ = [("pow_bits", ctypes.c_uint32), ("log_blowup_factor", ctypes.c_uint32),
                ("log_last_layer_degree_bound", ctypes.c_uint32), ("n_queries", ctypes.c_uint32),
                ("line_fold_step", ctypes.c_uint32), ("lifting_log_size", ctypes.c_uint32)]

    @staticmethod
    def make(pow_bits=5, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3, line_fold_step=1,
             lifting_log_size=0):
        return PcsConfig(pow_bits, log_blowup_factor, log_last_layer_degree_bound, n_queries, line_fold_step,
                         lifting_log_size)


class SampleBatch(ctypes.Structure):
    _fields_ = [("point", CirclePointQM31), ("n_columns", ctypes.c_uint32),
                ("column_indices", ctypes.POINTER(ctypes.c_uint32)), ("sample_values", ctypes.POINTER(QM31)),
                ("random_coeffs", ctypes.POINTER(QM31))]


class AccumulatedNumerators(ctypes.Structure):
    _fields_ = [("sample_point", CirclePointQM31), ("partial", ctypes.c_void_p * 4), ("log_size", ctypes.c_uint32),
                ("first_linear_term", QM31)]


_U32P = ctypes.POINTER(ctypes.c_uint32)
_VPP = ctypes.POINTER(ctypes.c_void_p)

EXPORTS = [
    "stwo_hip_create", "stwo_hip_destroy", "stwo_hip_synchronize", "stwo_hip_last_error", "stwo_hip_alloc",
    "stwo_hip_free", "stwo_hip_alloc_columns", "stwo_hip_free_columns", "stwo_hip_copy_h2d", "stwo_hip_copy_d2h", "stwo_hip_batch_at_multi",
    "stwo_hip_precompute_twiddles", "stwo_hip_interpolate_columns", "stwo_hip_evaluate_polynomials",
    "stwo_hip_batch_eval_at_point", "stwo_hip_barycentric_weights", "stwo_hip_batch_barycentric_eval",
    "stwo_hip_commit_on_layer", "stwo_hip_fold_circle_into_line",
    "stwo_hip_fold_line", "stwo_hip_accumulate_numerators", "stwo_hip_compute_quotients_and_combine",
    "stwo_hip_accumulate", "stwo_hip_lift_and_accumulate", "stwo_hip_generate_secure_powers", "stwo_hip_grind",
    "stwo_hip_evaluate_constraint_quotients_on_domain", "stwo_hip_prove_wide_fibonacci", "stwo_hip_prove_poseidon",
    "stwo_hip_prove_wide_fibonacci_multi",
    "stwo_proof_status", "stwo_proof_size", "stwo_proof_bytes", "stwo_proof_digest", "stwo_proof_claimed_sum",
    "stwo_proof_stage_ms", "stwo_proof_free", "stwo_hip_gen_trace_wide_fibonacci", "stwo_hip_gen_trace_poseidon",
    "stwo_hip_set_kernel_timing", "stwo_hip_kernel_stats", "stwo_hip_reset_kernel_stats",
    "stwo_hip_bit_reverse_column", "stwo_hip_column_at", "stwo_hip_column_set", "stwo_hip_split_at_mid",
    "stwo_hip_extend", "stwo_hip_twiddles_create", "stwo_hip_twiddles_read", "stwo_hip_twiddles_free",
    "stwo_hip_interpolate_on_domain", "stwo_hip_evaluate_on_domain", "stwo_hip_eval_at_point_by_folding",
    "stwo_hip_measure_int32_issue", "stwo_hip_prove_wide_fibonacci_per_call", "stwo_hip_prove_poseidon_per_call",
    "stwo_hip_gen_interaction_trace_poseidon", "stwo_hip_batch_eval_at_point_noalloc",
    "stwo_verify_wide_fibonacci", "stwo_verify_poseidon", "stwo_hip_fold_circle_into_line_grouped",
    "stwo_hip_measure_butterfly_rate", "stwo_hip_kernel_units", "stwo_hip_fft_low_log", "stwo_hip_set_fft_low_log",
    "stwo_hip_commit_on_layer_poseidon252", "stwo_hip_grind_poseidon252", "stwo_poseidon252_hash_many",
    "stwo_poseidon252_permute", "stwo_hip_prove_wide_fibonacci_poseidon252", "stwo_hip_prove_poseidon_poseidon252",
    "stwo_verify_wide_fibonacci_poseidon252", "stwo_verify_poseidon_poseidon252", "stwo_hip_measure_permutation_rate",
    "stwo_hip_measure_compression_rate",
    "stwo_hip_build_leaves_lifted", "stwo_hip_build_next_layer_lifted",
    "stwo_hip_channel_init", "stwo_hip_channel_mix_root", "stwo_hip_channel_mix_u32s",
    "stwo_hip_channel_draw_secure_felts", "stwo_hip_channel_draw_queries", "stwo_hip_channel_grind",
    "stwo_hip_channel_read", "stwo_hip_fold_circle_into_line_dev", "stwo_hip_fold_line_dev",
    "stwo_hip_graph_begin", "stwo_hip_graph_end", "stwo_hip_graph_launch", "stwo_hip_graph_free",
]
CHANNEL_WORDS = 32
LIFTED_HASHERS = {"blake2s": 0, "poseidon252": 1}

# The MerkleChannel of a proof: the suffix of its prove / verify entry points.
MERKLE_CHANNELS = {"blake2s": ("_per_call", ""), "poseidon252": ("_poseidon252", "_poseidon252")}


def _merkle_channel(name):
    if name not in MERKLE_CHANNELS:
        raise ValueError(f"unknown Merkle channel {name!r} (one of {sorted(MERKLE_CHANNELS)})")
    return MERKLE_CHANNELS[name]

KERNEL_CLASSES = ["interpolate", "evaluate", "merkle", "composition", "accumulate_numerators", "quotients",
                  "fri_fold", "oods", "logup", "grind", "misc", "fri_tail"]

_lib = None


def load_library(path=LIB_PATH):
    """Loads libstwo_hip.so; raises when it is missing (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise StwoHipError("libstwo_hip.so not built: run __graft_entry__.build() (expected at %s)" % path)
    L = ctypes.CDLL(path)
    c = ctypes
    L.stwo_hip_create.argtypes = [c.c_int, c.POINTER(c.c_void_p)]
    L.stwo_hip_destroy.argtypes = [c.c_void_p]
    L.stwo_hip_destroy.restype = None
    L.stwo_hip_synchronize.argtypes = [c.c_void_p]
    L.stwo_hip_last_error.argtypes = [c.c_void_p]
    L.stwo_hip_last_error.restype = c.c_char_p
    L.stwo_hip_alloc.argtypes = [c.c_void_p, c.c_size_t, c.c_int, c.POINTER(c.c_void_p)]
    L.stwo_hip_free.argtypes = [c.c_void_p, c.c_void_p]
    L.stwo_hip_alloc_columns.argtypes = [c.c_void_p, c.c_uint32, c.c_size_t, c.c_int, _VPP]
    L.stwo_hip_free_columns.argtypes = [c.c_void_p, _VPP, c.c_uint32]
    L.stwo_hip_copy_h2d.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_size_t]
    L.stwo_hip_copy_d2h.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_size_t]
    L.stwo_hip_batch_at_multi.argtypes = [c.c_void_p, _VPP, c.c_uint32, _U32P, c.c_uint32, _U32P]
    L.stwo_hip_precompute_twiddles.argtypes = [c.c_void_p, c.c_uint32, _U32P, _U32P]
    L.stwo_hip_interpolate_columns.argtypes = [c.c_void_p, _VPP, _VPP, c.c_uint32, c.c_uint32]
    L.stwo_hip_evaluate_polynomials.argtypes = [c.c_void_p, _VPP, _VPP, c.c_uint32, c.c_uint32, c.c_uint32]
    L.stwo_hip_batch_eval_at_point.argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32, CirclePointQM31,
                                               c.POINTER(QM31)]
    L.stwo_hip_batch_eval_at_point_noalloc.argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32, CirclePointQM31,
                                                       c.POINTER(QM31)]
    L.stwo_hip_barycentric_weights.argtypes = [c.c_void_p, c.c_uint32, CirclePointQM31, _VPP]
    L.stwo_hip_batch_barycentric_eval.argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32, _VPP, c.POINTER(QM31)]
    L.stwo_hip_commit_on_layer.argtypes = [c.c_void_p, c.c_uint32, c.c_void_p, _VPP, c.c_uint32, c.c_void_p]
    L.stwo_hip_fold_circle_into_line.argtypes = [c.c_void_p, c.c_uint32, _VPP, _VPP, QM31]
    L.stwo_hip_fold_line.argtypes = [c.c_void_p, c.c_uint32, _VPP, _VPP, QM31, c.c_uint32]
    L.stwo_hip_fold_circle_into_line_grouped.argtypes = [c.c_void_p, c.c_uint32, c.c_uint32, _VPP, _VPP, QM31]
    L.stwo_hip_accumulate_numerators.argtypes = [c.c_void_p, c.c_uint32, _VPP, c.c_uint32, c.POINTER(SampleBatch),
                                                 c.c_uint32, _VPP, c.POINTER(QM31)]
    L.stwo_hip_compute_quotients_and_combine.argtypes = [c.c_void_p, c.POINTER(AccumulatedNumerators), c.c_uint32,
                                                         c.c_uint32, _VPP]
    L.stwo_hip_accumulate.argtypes = [c.c_void_p, c.c_uint32, _VPP, _VPP]
    L.stwo_hip_lift_and_accumulate.argtypes = [c.c_void_p, _VPP, _U32P, c.c_uint32, _VPP]
    L.stwo_hip_generate_secure_powers.argtypes = [c.c_void_p, QM31, c.c_uint32, c.POINTER(QM31)]
    L.stwo_hip_grind.argtypes = [c.c_void_p, c.c_char_p, c.c_uint32, c.POINTER(c.c_uint64)]
    L.stwo_hip_evaluate_constraint_quotients_on_domain.argtypes = [
        c.c_void_p, _VPP, _VPP, c.c_uint32, _VPP, c.c_uint32, c.POINTER(QM31), c.c_uint32, _U32P, c.c_uint32,
        c.c_uint32, c.c_char_p, c.c_size_t, QM31]
    L.stwo_hip_prove_wide_fibonacci.argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32, c.POINTER(PcsConfig),
                                                c.POINTER(c.c_void_p)]
    L.stwo_hip_prove_poseidon.argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32, c.POINTER(PcsConfig),
                                          c.POINTER(c.c_void_p)]
    L.stwo_hip_prove_wide_fibonacci_multi.argtypes = [c.c_void_p, _VPP, c.c_uint32, _U32P, _U32P,
                                                      c.POINTER(PcsConfig), c.POINTER(c.c_void_p)]
    L.stwo_proof_status.argtypes = [c.c_void_p]
    L.stwo_proof_size.argtypes = [c.c_void_p]
    L.stwo_proof_size.restype = c.c_size_t
    L.stwo_proof_bytes.argtypes = [c.c_void_p]
    L.stwo_proof_bytes.restype = c.c_void_p
    L.stwo_proof_digest.argtypes = [c.c_void_p, c.c_char_p]
    L.stwo_proof_digest.restype = None
    L.stwo_proof_claimed_sum.argtypes = [c.c_void_p, c.POINTER(QM31)]
    L.stwo_proof_claimed_sum.restype = None
    L.stwo_proof_stage_ms.argtypes = [c.c_void_p, c.POINTER(c.c_double)]
    L.stwo_proof_stage_ms.restype = None
    L.stwo_proof_free.argtypes = [c.c_void_p]
    L.stwo_proof_free.restype = None
    L.stwo_hip_gen_trace_wide_fibonacci.argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32, c.c_uint32,
                                                    c.c_uint64]
    L.stwo_hip_gen_trace_poseidon.argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint64]
    L.stwo_hip_set_kernel_timing.argtypes = [c.c_void_p, c.c_int]
    L.stwo_hip_kernel_stats.argtypes = [c.c_void_p, c.c_uint32, c.POINTER(c.c_double), c.POINTER(c.c_double),
                                        c.POINTER(c.c_uint64)]
    L.stwo_hip_reset_kernel_stats.argtypes = [c.c_void_p]
    L.stwo_hip_bit_reverse_column.argtypes = [c.c_void_p, c.c_void_p, c.c_uint32]
    L.stwo_hip_column_at.argtypes = [c.c_void_p, c.c_void_p, c.c_size_t, _U32P]
    L.stwo_hip_column_set.argtypes = [c.c_void_p, c.c_void_p, c.c_size_t, c.c_uint32]
    L.stwo_hip_split_at_mid.argtypes = [c.c_void_p, c.c_void_p, c.c_size_t, c.POINTER(c.c_void_p),
                                        c.POINTER(c.c_void_p)]
    L.stwo_hip_extend.argtypes = [c.c_void_p, c.c_void_p, c.c_uint32, c.c_uint32, c.c_void_p]
    L.stwo_hip_twiddles_create.argtypes = [c.c_void_p, c.c_uint32, c.c_uint32, c.POINTER(c.c_void_p)]
    L.stwo_hip_twiddles_read.argtypes = [c.c_void_p, c.c_void_p, _U32P, _U32P]
    L.stwo_hip_twiddles_free.argtypes = [c.c_void_p, c.c_void_p]
    L.stwo_hip_interpolate_on_domain.argtypes = [c.c_void_p, c.c_void_p, c.c_uint32, c.c_uint32, _VPP, _VPP,
                                                 c.c_uint32]
    L.stwo_hip_evaluate_on_domain.argtypes = [c.c_void_p, c.c_void_p, c.c_uint32, c.c_uint32, _VPP, c.c_uint32,
                                              _VPP, c.c_uint32]
    L.stwo_hip_eval_at_point_by_folding.argtypes = [c.c_void_p, c.c_void_p, c.c_uint32, CirclePointQM31,
                                                    c.POINTER(QM31)]
    L.stwo_hip_measure_int32_issue.argtypes = [c.c_void_p, c.POINTER(c.c_double)]
    L.stwo_hip_measure_butterfly_rate.argtypes = [c.c_void_p, c.c_int, c.POINTER(c.c_double)]
    L.stwo_hip_measure_permutation_rate.argtypes = [c.c_void_p, c.POINTER(c.c_double)]
    L.stwo_hip_measure_compression_rate.argtypes = [c.c_void_p, c.POINTER(c.c_double)]
    L.stwo_hip_build_leaves_lifted.argtypes = [c.c_void_p, c.c_uint32, _VPP, _U32P, c.c_uint32, c.c_uint32,
                                               c.c_void_p]
    L.stwo_hip_build_next_layer_lifted.argtypes = [c.c_void_p, c.c_uint32, c.c_void_p, c.c_uint32, c.c_void_p]
    L.stwo_hip_channel_init.argtypes = [c.c_void_p, c.c_void_p, c.c_char_p, c.c_uint32]
    L.stwo_hip_channel_mix_root.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p]
    L.stwo_hip_channel_mix_u32s.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_uint32]
    L.stwo_hip_channel_draw_secure_felts.argtypes = [c.c_void_p, c.c_void_p, c.c_uint32, c.c_void_p]
    L.stwo_hip_channel_draw_queries.argtypes = [c.c_void_p, c.c_void_p, c.c_uint32, c.c_uint32, c.c_void_p]
    L.stwo_hip_channel_grind.argtypes = [c.c_void_p, c.c_void_p, c.c_uint32, c.c_void_p]
    L.stwo_hip_channel_read.argtypes = [c.c_void_p, c.c_void_p, c.c_char_p, c.POINTER(c.c_uint32)]
    L.stwo_hip_fold_circle_into_line_dev.argtypes = [c.c_void_p, c.c_uint32, _VPP, _VPP, c.c_void_p, c.c_uint32]
    L.stwo_hip_fold_line_dev.argtypes = [c.c_void_p, c.c_uint32, _VPP, _VPP, c.c_void_p]
    L.stwo_hip_graph_begin.argtypes = [c.c_void_p]
    L.stwo_hip_graph_end.argtypes = [c.c_void_p, c.POINTER(c.c_void_p)]
    L.stwo_hip_graph_launch.argtypes = [c.c_void_p, c.c_void_p]
    L.stwo_hip_graph_free.argtypes = [c.c_void_p, c.c_void_p]
    L.stwo_hip_kernel_units.argtypes = [c.c_void_p, c.c_uint32, c.POINTER(c.c_double)]
    L.stwo_hip_fft_low_log.argtypes = [c.c_void_p, c.c_int, c.c_uint32, c.c_uint32, c.c_uint32, _U32P]
    L.stwo_hip_commit_on_layer_poseidon252.argtypes = [c.c_void_p, c.c_uint32, c.c_void_p, _VPP, c.c_uint32,
                                                       c.c_void_p]
    L.stwo_hip_grind_poseidon252.argtypes = [c.c_void_p, _U32P, c.c_uint32, c.POINTER(c.c_uint64)]
    L.stwo_poseidon252_hash_many.argtypes = [_U32P, c.c_size_t, _U32P]
    L.stwo_poseidon252_permute.argtypes = [_U32P]
    L.stwo_hip_set_fft_low_log.argtypes = [c.c_void_p, c.c_uint32]
    for sfx in ("", "_poseidon252"):
        getattr(L, "stwo_verify_wide_fibonacci" + sfx).argtypes = [c.c_char_p, c.c_size_t, c.c_uint32, _U32P, _U32P,
                                                                  c.POINTER(PcsConfig), c.c_char_p, c.c_size_t]
        getattr(L, "stwo_verify_poseidon" + sfx).argtypes = [c.c_char_p, c.c_size_t, c.c_uint32, c.c_uint32, QM31,
                                                            c.POINTER(PcsConfig), c.c_char_p, c.c_size_t]
    for sfx in ("_per_call", "_poseidon252"):
        getattr(L, "stwo_hip_prove_wide_fibonacci" + sfx).argtypes = [c.c_void_p, _VPP, c.c_uint32, _U32P, _U32P,
                                                                     c.POINTER(PcsConfig), c.POINTER(c.c_void_p)]
        getattr(L, "stwo_hip_prove_poseidon" + sfx).argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32,
                                                               c.POINTER(PcsConfig), c.POINTER(c.c_void_p)]
    L.stwo_hip_gen_interaction_trace_poseidon.argtypes = [c.c_void_p, _VPP, c.c_uint32, c.c_uint32, QM31, QM31, _VPP,
                                                          c.POINTER(QM31)]
    _lib = L
    return L


class ColumnList(collections.abc.Sequence):
    """Columns allocated together (HipBackend.columns), as an immutable sequence that keeps the
    array of their addresses: a library call taking them all (a whole trace, every proof) passes the
    array as it is, and a column's DeviceColumn object is made only when the column is indexed, so
    a trace of a thousand columns costs no Python object per column unless one is used. A column
    freed through its object leaves a null address in the array."""

    def __init__(self, backend, arr, n_cols, n_words):
        self._backend, self._arr, self._n, self._words = backend, arr, n_cols, n_words
        self._cols = [None] * n_cols
        self._made = []  # indices that have a DeviceColumn

    def _addrs(self):
        return np.frombuffer(self._arr, dtype=np.uint64, count=self._n)

    def __len__(self):
        return self._n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(self._n))]
        i = operator.index(i)
        if i < 0:
            i += self._n
        if not 0 <= i < self._n:
            raise IndexError("column index out of range")
        col = self._cols[i]
        if col is None:
            addr = self._arr[i]
            col = DeviceColumn(self._backend, self._words, ptr=addr if addr else 0)
            if not addr:
                col.ptr = None  # freed already
            col._slot = (self, i)
            self._cols[i] = col
            self._made.append(i)
        return col

    def _live(self):
        return self._n == 0 or bool(np.all(self._addrs()))

    def _freed(self):
        ctypes.memset(ctypes.addressof(self._arr), 0, self._n * ctypes.sizeof(ctypes.c_void_p))
        for i in self._made:
            self._cols[i].ptr = None

    def __del__(self):
        # columns never indexed have no DeviceColumn to free them (those that do free their own)
        try:
            b = self._backend
            if not b.ctx or self._n == 0:
                return
            a = self._addrs().copy()
            a[self._made] = 0
            live = np.ascontiguousarray(a[a != 0])
            if live.size:
                b.lib.stwo_hip_free_columns(b.ctx, live.ctypes.data_as(ctypes.POINTER(ctypes.c_void_p)), live.size)
        except Exception:
            pass


def _min_words(cols):
    """The shortest column's length, without making a DeviceColumn per column of a ColumnList."""
    if isinstance(cols, ColumnList):
        return cols._words
    return min((c.n for c in cols), default=0)


def _ptrs(cols):
    if isinstance(cols, ColumnList):
        return cols._arr
    arr = (ctypes.c_void_p * max(1, len(cols)))()
    for i, col in enumerate(cols):
        arr[i] = col.ptr if isinstance(col, DeviceColumn) else col
    return arr


def _concat_ptrs(lists):
    """The address array of several column sequences, copied array by array when they are
    ColumnLists."""
    n = sum(len(t) for t in lists)
    arr = (ctypes.c_void_p * max(1, n))()
    off = 0
    for t in lists:
        src = _ptrs(t)
        ctypes.memmove(ctypes.addressof(arr) + off * ctypes.sizeof(ctypes.c_void_p), src,
                       len(t) * ctypes.sizeof(ctypes.c_void_p))
        off += len(t)
    return arr


class DeviceColumn:
    """A device-resident column of u32 words (M31 values), freed with its backend."""

    def __init__(self, backend, n_words, zeroed=False, ptr=None, parent=None):
        self.backend = backend
        self.n = int(n_words)
        # a view (split_at_mid) borrows its parent's memory and keeps the parent alive
        self.parent = parent
        if ptr is not None:  # an address the backend already allocated (columns()), or a view
            self.ptr = ptr
            return
        p = ctypes.c_void_p()
        backend._check(backend.lib.stwo_hip_alloc(backend.ctx, self.n, int(zeroed), ctypes.byref(p)))
        self.ptr = p.value

    @property
    def log_size(self):
        return self.n.bit_length() - 1

    def free(self):
        if self.ptr and self.parent is None:
            self.backend.lib.stwo_hip_free(self.backend.ctx, self.ptr)
            slot = getattr(self, "_slot", None)
            if slot is not None:  # a column of a ColumnList: its address leaves the list's array
                slot[0]._arr[slot[1]] = None
        self.ptr = None

    # Column::at / set / split_at_mid (crates/stwo/src/prover/backend/mod.rs:59-62,90-91)
    def at(self, index):
        v = ctypes.c_uint32()
        self.backend._check(self.backend.lib.stwo_hip_column_at(self.backend.ctx, self.ptr, index, ctypes.byref(v)))
        return v.value

    def set(self, index, value):
        self.backend._check(self.backend.lib.stwo_hip_column_set(self.backend.ctx, self.ptr, index, value))

    def split_at_mid(self):
        lft, rgt = ctypes.c_void_p(), ctypes.c_void_p()
        self.backend._check(self.backend.lib.stwo_hip_split_at_mid(self.backend.ctx, self.ptr, self.n,
                                                                   ctypes.byref(lft), ctypes.byref(rgt)))
        root = self.parent if self.parent is not None else self
        return (DeviceColumn(self.backend, self.n // 2, ptr=lft.value, parent=root),
                DeviceColumn(self.backend, self.n // 2, ptr=rgt.value, parent=root))

    def to_cpu(self):
        out = np.empty(self.n, np.uint32)
        self.backend._check(self.backend.lib.stwo_hip_copy_d2h(self.backend.ctx, out.ctypes.data, self.ptr, self.n))
        return out

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class TwiddleTree:
    """A twiddle tree of any coset, owned by the device (TwiddleTree<B>, prover/poly/twiddles.rs:9-14)."""

    def __init__(self, backend, coset_initial_index, coset_log_size):
        self.backend = backend
        self.log_size = coset_log_size
        h = ctypes.c_void_p()
        backend._check(backend.lib.stwo_hip_twiddles_create(backend.ctx, coset_initial_index, coset_log_size,
                                                            ctypes.byref(h)))
        self.handle = h.value

    def read(self):
        n = 1 << self.log_size
        tw, itw = np.empty(n, np.uint32), np.empty(n, np.uint32)
        self.backend._check(self.backend.lib.stwo_hip_twiddles_read(self.backend.ctx, self.handle,
                                                                    tw.ctypes.data_as(_U32P), itw.ctypes.data_as(_U32P)))
        return tw, itw

    def free(self):
        if self.handle and self.backend.ctx:
            self.backend._check(self.backend.lib.stwo_hip_twiddles_free(self.backend.ctx, self.handle))
        self.handle = None


class Proof:
    def __init__(self, backend, handle):
        L = backend.lib
        self.status = L.stwo_proof_status(handle)
        n = L.stwo_proof_size(handle)
        self.bytes = ctypes.string_at(L.stwo_proof_bytes(handle), n)
        d = ctypes.create_string_buffer(32)
        L.stwo_proof_digest(handle, d)
        self.digest = d.raw
        cs = QM31()
        L.stwo_proof_claimed_sum(handle, ctypes.byref(cs))
        self.claimed_sum = cs.tuple()
        ms = (ctypes.c_double * 8)()
        L.stwo_proof_stage_ms(handle, ms)
        self.stage_ms = dict(zip(["trace_commit", "interaction", "composition", "oods", "quotients", "fri", "pow",
                                  "decommit"], list(ms)))
        L.stwo_proof_free(handle)


class HipBackend:
    """Drop-in counterpart of SimdBackend's trait methods (crates/stwo/src/prover/backend/mod.rs:20-37)."""

    def __init__(self, device=0):
        self.lib = load_library()
        p = ctypes.c_void_p()
        st = self.lib.stwo_hip_create(device, ctypes.byref(p))
        if st != 0:
            raise StwoHipError("stwo_hip_create failed (status %d): is a GPU visible?" % st, st)
        self.ctx = p.value

    def close(self):
        if self.ctx:
            self.lib.stwo_hip_destroy(self.ctx)
            self.ctx = None

    def _check(self, st):
        if st != 0:
            msg = self.lib.stwo_hip_last_error(self.ctx)
            raise StwoHipError("%s: %s" % (STATUS.get(st, st), msg.decode() if msg else ""), st)

    # ---- columns ----
    def column(self, n_words, zeroed=False):
        return DeviceColumn(self, n_words, zeroed)

    def columns(self, n_cols, n_words, zeroed=False):
        """n_cols columns of n_words each, allocated in one library call."""
        arr = (ctypes.c_void_p * max(1, n_cols))()
        self._check(self.lib.stwo_hip_alloc_columns(self.ctx, n_cols, n_words, int(zeroed), arr))
        return ColumnList(self, arr, n_cols, n_words)

    def free_columns(self, cols):
        """Frees columns in one library call."""
        if isinstance(cols, ColumnList):
            if cols._live():
                st = self.lib.stwo_hip_free_columns(self.ctx, cols._arr, len(cols))  # the kept array
            else:  # some columns were freed one by one
                a = cols._addrs()
                live = [int(x) for x in a[a != 0]]
                st = self.lib.stwo_hip_free_columns(self.ctx, _ptrs(live), len(live)) if live else 0
            cols._freed()
            self._check(st)
            return
        live = [col for col in cols if col.ptr]
        if not live:
            return
        st = self.lib.stwo_hip_free_columns(self.ctx, _ptrs(live), len(live))
        for col in live:
            col.ptr = None
        self._check(st)

    def from_cpu(self, arr):
        a = np.ascontiguousarray(arr, np.uint32)
        col = DeviceColumn(self, a.size)
        self._check(self.lib.stwo_hip_copy_h2d(self.ctx, col.ptr, a.ctypes.data, a.size))
        return col

    def batch_at_multi(self, cols, rows):
        rows = np.ascontiguousarray(rows, np.uint32)
        out = np.empty(len(rows) * len(cols), np.uint32)
        self._check(self.lib.stwo_hip_batch_at_multi(self.ctx, _ptrs(cols), len(cols), rows.ctypes.data_as(_U32P),
                                                     len(rows), out.ctypes.data_as(_U32P)))
        return out

    def synchronize(self):
        self._check(self.lib.stwo_hip_synchronize(self.ctx))

    # ---- PolyOps ----
    def precompute_twiddles(self, log_domain):
        n = 1 << (log_domain - 1)
        tw = np.empty(n, np.uint32)
        itw = np.empty(n, np.uint32)
        self._check(self.lib.stwo_hip_precompute_twiddles(self.ctx, log_domain, tw.ctypes.data_as(_U32P),
                                                          itw.ctypes.data_as(_U32P)))
        return tw, itw

    def interpolate_columns(self, evals, log_size, out=None):
        out = out if out is not None else [self.column(1 << log_size) for _ in evals]
        self._check(self.lib.stwo_hip_interpolate_columns(self.ctx, _ptrs(evals), _ptrs(out), len(evals), log_size))
        return out

    def evaluate_polynomials(self, coeffs, log_size, log_blowup, out=None):
        out = out if out is not None else [self.column(1 << (log_size + log_blowup)) for _ in coeffs]
        self._check(self.lib.stwo_hip_evaluate_polynomials(self.ctx, _ptrs(coeffs), _ptrs(out), len(coeffs), log_size,
                                                           log_blowup))
        return out

    def bit_reverse_column(self, col, log_size):
        """ColumnOps::bit_reverse_column, in place."""
        self._check(self.lib.stwo_hip_bit_reverse_column(self.ctx, col.ptr, log_size))

    def extend(self, coeffs, log_size, new_log_size):
        """PolyOps::extend: the coefficients followed by zeros, in a new column."""
        out = self.column(1 << new_log_size)
        self._check(self.lib.stwo_hip_extend(self.ctx, coeffs.ptr, log_size, new_log_size, out.ptr))
        return out

    def split_at_mid(self, coeffs):
        """PolyOps::split_at_mid: (p_left, p_right) as views of the coefficient column."""
        return coeffs.split_at_mid()

    def twiddles(self, coset_initial_index, coset_log_size):
        """PolyOps::precompute_twiddles for Coset::new(coset_initial_index, coset_log_size)."""
        return TwiddleTree(self, coset_initial_index, coset_log_size)

    def interpolate_on_domain(self, tree, half_coset_initial_index, log_size, evals):
        out = [self.column(1 << log_size) for _ in evals]
        self._check(self.lib.stwo_hip_interpolate_on_domain(self.ctx, tree.handle, half_coset_initial_index, log_size,
                                                            _ptrs(evals), _ptrs(out), len(evals)))
        return out

    def evaluate_on_domain(self, tree, half_coset_initial_index, log_size, coeffs, log_coeffs):
        out = [self.column(1 << log_size) for _ in coeffs]
        self._check(self.lib.stwo_hip_evaluate_on_domain(self.ctx, tree.handle, half_coset_initial_index, log_size,
                                                         _ptrs(coeffs), log_coeffs, _ptrs(out), len(coeffs)))
        return out

    def eval_at_point_by_folding(self, evals, log_size, point):
        out = QM31()
        pt = CirclePointQM31(QM31.of(point[:4]), QM31.of(point[4:]))
        self._check(self.lib.stwo_hip_eval_at_point_by_folding(self.ctx, evals.ptr, log_size, pt, ctypes.byref(out)))
        return out.tuple()

    def batch_eval_at_point(self, coeffs, log_size, point):
        out = (QM31 * len(coeffs))()
        pt = CirclePointQM31(QM31.of(point[:4]), QM31.of(point[4:]))
        self._check(self.lib.stwo_hip_batch_eval_at_point(self.ctx, _ptrs(coeffs), len(coeffs), log_size, pt, out))
        return [o.tuple() for o in out]

    def barycentric_weights(self, log_size, point):
        """PolyOps::barycentric_weights: 4 device coordinate columns of 2^log_size."""
        w = [self.column(1 << log_size) for _ in range(4)]
        pt = CirclePointQM31(QM31.of(point[:4]), QM31.of(point[4:]))
        self._check(self.lib.stwo_hip_barycentric_weights(self.ctx, log_size, pt, _ptrs(w)))
        return w

    def batch_barycentric_eval(self, evals, log_size, weights):
        """PolyOps::barycentric_eval_at_point for several columns: list of QM31 tuples."""
        out = (QM31 * len(evals))()
        self._check(self.lib.stwo_hip_batch_barycentric_eval(self.ctx, _ptrs(evals), len(evals), log_size,
                                                             _ptrs(weights), out))
        return [o.tuple() for o in out]

    # ---- MerkleOps ----
    def commit_on_layer(self, log_size, prev, cols, out=None):
        out = out if out is not None else self.column(8 << log_size)
        self._check(self.lib.stwo_hip_commit_on_layer(self.ctx, log_size, prev.ptr if prev is not None else None,
                                                      _ptrs(cols), len(cols), out.ptr))
        return out

    def commit_on_layer_poseidon252(self, log_size, prev, cols, out=None):
        """MerkleOps<Poseidon252MerkleHasher>::commit_on_layer: 2^log_size canonical felt252 hashes,
        8 little-endian words each (felt252_to_words / felt252_from_words convert)."""
        out = out if out is not None else self.column(8 << log_size)
        self._check(self.lib.stwo_hip_commit_on_layer_poseidon252(
            self.ctx, log_size, prev.ptr if prev is not None else None, _ptrs(cols), len(cols), out.ptr))
        return out

    # ---- MerkleOpsLifted ----
    def build_leaves_lifted(self, cols, log_sizes, lifting_log_size, hasher="blake2s"):
        """MerkleOpsLifted::build_leaves: columns in non-decreasing order of log size, lifted to
        2^lifting_log_size leaves (one leaf for no columns); hasher "blake2s" or "poseidon252"."""
        logs = np.array(log_sizes, np.uint32).reshape(-1)
        out = self.column(8 << lifting_log_size if len(cols) else 8)
        self._check(self.lib.stwo_hip_build_leaves_lifted(self.ctx, LIFTED_HASHERS[hasher], _ptrs(cols),
                                                          logs.ctypes.data_as(_U32P), len(cols), lifting_log_size,
                                                          out.ptr))
        return out

    def build_next_layer_lifted(self, prev, log_size, hasher="blake2s"):
        """MerkleOpsLifted::build_next_layer: 2^log_size parents of the 2^(log_size + 1) hashes in prev."""
        out = self.column(8 << log_size)
        self._check(self.lib.stwo_hip_build_next_layer_lifted(self.ctx, LIFTED_HASHERS[hasher], prev.ptr, log_size,
                                                              out.ptr))
        return out

    # ---- the channel on the device, and graphs ----
    def channel_new(self, digest=bytes(32), n_draws=0):
        """A device channel state (CHANNEL_WORDS words), initialised (stream-ordered)."""
        st = self.column(CHANNEL_WORDS, zeroed=True)
        self.channel_init(st, digest, n_draws)
        return st

    def channel_init(self, st, digest=bytes(32), n_draws=0):
        self._check(self.lib.stwo_hip_channel_init(self.ctx, st.ptr, bytes(digest), n_draws))

    def channel_mix_root(self, st, root, offset=0):
        """root: a device column; its words offset..offset + 8 are the root."""
        self._check(self.lib.stwo_hip_channel_mix_root(self.ctx, st.ptr, root.ptr + 4 * offset))

    def channel_mix_u32s(self, st, words, n=None):
        n = words.n if n is None else n
        self._check(self.lib.stwo_hip_channel_mix_u32s(self.ctx, st.ptr, words.ptr, n))

    def channel_draw_secure_felts(self, st, n, out=None):
        out = out if out is not None else self.column(4 * n)
        self._check(self.lib.stwo_hip_channel_draw_secure_felts(self.ctx, st.ptr, n, out.ptr))
        return out

    def channel_draw_queries(self, st, log_domain, n_queries, out=None):
        out = out if out is not None else self.column(n_queries)
        self._check(self.lib.stwo_hip_channel_draw_queries(self.ctx, st.ptr, log_domain, n_queries, out.ptr))
        return out

    def channel_grind(self, st, pow_bits, nonce=None):
        nonce = nonce if nonce is not None else self.column(2)
        self._check(self.lib.stwo_hip_channel_grind(self.ctx, st.ptr, pow_bits, nonce.ptr))
        return nonce

    def channel_read(self, st):
        d = ctypes.create_string_buffer(32)
        n = ctypes.c_uint32()
        self._check(self.lib.stwo_hip_channel_read(self.ctx, st.ptr, d, ctypes.byref(n)))
        return d.raw, n.value

    def fold_circle_into_line_dev(self, dst4, src4, alpha_col, log_src, dst_is_zero=False):
        self._check(self.lib.stwo_hip_fold_circle_into_line_dev(self.ctx, log_src, _ptrs(src4), _ptrs(dst4),
                                                                alpha_col.ptr, int(bool(dst_is_zero))))

    def fold_line_dev(self, dst4, src4, alpha_col, log_src):
        self._check(self.lib.stwo_hip_fold_line_dev(self.ctx, log_src, _ptrs(src4), _ptrs(dst4), alpha_col.ptr))

    def graph_begin(self):
        self._check(self.lib.stwo_hip_graph_begin(self.ctx))

    def graph_end(self):
        g = ctypes.c_void_p()
        self._check(self.lib.stwo_hip_graph_end(self.ctx, ctypes.byref(g)))
        return g

    def graph_launch(self, g):
        self._check(self.lib.stwo_hip_graph_launch(self.ctx, g))

    def graph_free(self, g):
        self._check(self.lib.stwo_hip_graph_free(self.ctx, g))

    # ---- FriOps ----
    def fold_circle_into_line(self, dst4, src4, alpha, log_src):
        self._check(self.lib.stwo_hip_fold_circle_into_line(self.ctx, log_src, _ptrs(src4), _ptrs(dst4),
                                                            QM31.of(alpha)))

    def fold_circle_into_line_grouped(self, dst4s, src4s, alpha, log_src):
        """Grouped fold: dst4s[g] = dst4s[g] * alpha^2 + fold(src4s[g]) for every group, one launch."""
        src = [c for g in src4s for c in g]
        dst = [c for g in dst4s for c in g]
        self._check(self.lib.stwo_hip_fold_circle_into_line_grouped(self.ctx, log_src, len(src4s), _ptrs(src),
                                                                    _ptrs(dst), QM31.of(alpha)))

    def fold_line(self, src4, alpha, log_src, fold_step=1):
        dst = [self.column(1 << (log_src - fold_step)) for _ in range(4)]
        self._check(self.lib.stwo_hip_fold_line(self.ctx, log_src, _ptrs(src4), _ptrs(dst), QM31.of(alpha),
                                                fold_step))
        return dst

    # ---- QuotientOps ----
    def accumulate_numerators(self, cols, log_size, batches):
        """batches: list of (point8, [(col_idx, value4, rand4), ...]); returns (partials, first_linear)."""
        keep = []
        arr = (SampleBatch * len(batches))()
        for i, (pt, entries) in enumerate(batches):
            idx = np.array([e[0] for e in entries], np.uint32)
            vals = (QM31 * len(entries))(*[QM31.of(e[1]) for e in entries])
            rnds = (QM31 * len(entries))(*[QM31.of(e[2]) for e in entries])
            keep += [idx, vals, rnds]
            arr[i].point = CirclePointQM31(QM31.of(pt[:4]), QM31.of(pt[4:]))
            arr[i].n_columns = len(entries)
            arr[i].column_indices = idx.ctypes.data_as(_U32P)
            arr[i].sample_values = vals
            arr[i].random_coeffs = rnds
        outs = [self.column(1 << log_size) for _ in range(4 * len(batches))]
        fl = (QM31 * len(batches))()
        self._check(self.lib.stwo_hip_accumulate_numerators(self.ctx, log_size, _ptrs(cols), len(cols), arr,
                                                            len(batches), _ptrs(outs), fl))
        return [outs[4 * b:4 * b + 4] for b in range(len(batches))], [f.tuple() for f in fl]

    def compute_quotients_and_combine(self, accs, lifting_log_size):
        """accs: list of (point8, partial4 columns, log_size, first_linear4)."""
        arr = (AccumulatedNumerators * len(accs))()
        for i, (pt, part, ls, fl) in enumerate(accs):
            arr[i].sample_point = CirclePointQM31(QM31.of(pt[:4]), QM31.of(pt[4:]))
            for k in range(4):
                arr[i].partial[k] = part[k].ptr
            arr[i].log_size = ls
            arr[i].first_linear_term = QM31.of(fl)
        out = [self.column(1 << lifting_log_size) for _ in range(4)]
        self._check(self.lib.stwo_hip_compute_quotients_and_combine(self.ctx, arr, len(accs), lifting_log_size,
                                                                    _ptrs(out)))
        return out

    # ---- AccumulationOps ----
    def accumulate(self, dst4, src4, log_size):
        self._check(self.lib.stwo_hip_accumulate(self.ctx, log_size, _ptrs(dst4), _ptrs(src4)))

    def lift_and_accumulate(self, cols4_list, log_sizes):
        flat = [c for cols4 in cols4_list for c in cols4]
        ls = np.array(log_sizes, np.uint32)
        out = [self.column(1 << int(max(log_sizes))) for _ in range(4)]
        self._check(self.lib.stwo_hip_lift_and_accumulate(self.ctx, _ptrs(flat), ls.ctypes.data_as(_U32P), len(ls),
                                                          _ptrs(out)))
        return out

    def generate_secure_powers(self, felt, n):
        out = (QM31 * n)()
        self._check(self.lib.stwo_hip_generate_secure_powers(self.ctx, QM31.of(felt), n, out))
        return [o.tuple() for o in out]

    # ---- GrindOps ----
    def grind(self, digest, pow_bits):
        nonce = ctypes.c_uint64()
        self._check(self.lib.stwo_hip_grind(self.ctx, digest, pow_bits, ctypes.byref(nonce)))
        return nonce.value

    def grind_poseidon252(self, digest, pow_bits):
        """GrindOps<Poseidon252Channel>::grind for a channel digest (a felt252 as an int)."""
        nonce = ctypes.c_uint64()
        d = felt252_to_words(digest)
        self._check(self.lib.stwo_hip_grind_poseidon252(self.ctx, d.ctypes.data_as(_U32P), pow_bits,
                                                        ctypes.byref(nonce)))
        return nonce.value

    # ---- ComponentProver ----
    def evaluate_constraint_quotients_on_domain(self, acc4, trace, interaction, coeff_powers, denom_inv, trace_log,
                                                eval_log, eval_blob, cumsum_shift=(0, 0, 0, 0)):
        pows = (QM31 * len(coeff_powers))(*[QM31.of(p) for p in coeff_powers])
        dinv = np.ascontiguousarray(denom_inv, np.uint32)
        self._check(self.lib.stwo_hip_evaluate_constraint_quotients_on_domain(
            self.ctx, _ptrs(acc4), _ptrs(trace), len(trace), _ptrs(interaction), len(interaction), pows,
            len(coeff_powers), dinv.ctypes.data_as(_U32P), trace_log, eval_log, eval_blob, len(eval_blob),
            QM31.of(cumsum_shift)))

    # ---- measurement ----
    def set_kernel_timing(self, enable):
        self._check(self.lib.stwo_hip_set_kernel_timing(self.ctx, int(enable)))

    def measure_int32_issue(self):
        """32-bit VALU lane-instructions per second of this device (the VALU roofline's peak)."""
        v = ctypes.c_double()
        self._check(self.lib.stwo_hip_measure_int32_issue(self.ctx, ctypes.byref(v)))
        return v.value

    def fft_low_log(self, inverse, n_cols, log_in, log_n):
        """The low-stage tile log (2^S values per workgroup) this context uses for a transform
        shape; 0 for transforms of fewer than 11 stages."""
        v = ctypes.c_uint32()
        self._check(self.lib.stwo_hip_fft_low_log(self.ctx, int(bool(inverse)), n_cols, log_in, log_n,
                                                  ctypes.byref(v)))
        return v.value

    def set_fft_low_log(self, low_log):
        """Fix the low-stage tile log for every later transform (0 = planned per shape)."""
        self._check(self.lib.stwo_hip_set_fft_low_log(self.ctx, low_log))

    def measure_butterfly_rate(self, inverse=False):
        """Forward (or inverse) circle-FFT butterflies per second with register-resident data: the
        ceiling of the transform kernels' arithmetic on this device."""
        v = ctypes.c_double()
        self._check(self.lib.stwo_hip_measure_butterfly_rate(self.ctx, int(bool(inverse)), ctypes.byref(v)))
        return v.value

    def measure_compression_rate(self):
        """Blake2s compressions per second with register-resident states and blocks: the ceiling of
        the Blake2s Merkle layers' hashing on this device."""
        v = ctypes.c_double()
        self._check(self.lib.stwo_hip_measure_compression_rate(self.ctx, ctypes.byref(v)))
        return v.value

    def measure_permutation_rate(self):
        """Starknet Poseidon permutations per second with register-resident states: the ceiling of
        the Poseidon252 Merkle layers on this device."""
        v = ctypes.c_double()
        self._check(self.lib.stwo_hip_measure_permutation_rate(self.ctx, ctypes.byref(v)))
        return v.value

    def reset_kernel_stats(self):
        self._check(self.lib.stwo_hip_reset_kernel_stats(self.ctx))

    def kernel_units(self):
        """{class: algorithmic work units} for the classes that count them (butterflies for
        interpolate and evaluate; Blake2s compressions or Poseidon permutations for merkle)."""
        out = {}
        for i, name in enumerate(KERNEL_CLASSES):
            u = ctypes.c_double()
            self._check(self.lib.stwo_hip_kernel_units(self.ctx, i, ctypes.byref(u)))
            if u.value:
                out[name] = u.value
        return out

    def kernel_stats(self):
        """{class: (total_ms, algorithmic_bytes, launches)} for the classes that ran."""
        out = {}
        for i, name in enumerate(KERNEL_CLASSES):
            ms, by, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
            self._check(self.lib.stwo_hip_kernel_stats(self.ctx, i, ctypes.byref(ms), ctypes.byref(by),
                                                       ctypes.byref(n)))
            if n.value:
                out[name] = (ms.value, by.value, n.value)
        return out

    # ---- trace generation and whole proofs ----
    def gen_trace_wide_fibonacci(self, n_cols, log_n_rows, mode=1, seed=0, out=None):
        """The trace on the device, into `out` (n_cols columns of 2^log_n_rows words) when given."""
        cols = out if out is not None else self.columns(n_cols, 1 << log_n_rows)
        if len(cols) != n_cols or _min_words(cols) < (1 << log_n_rows):
            raise ValueError("out must hold n_cols columns of 2^log_n_rows words")
        self._check(self.lib.stwo_hip_gen_trace_wide_fibonacci(self.ctx, _ptrs(cols), n_cols, log_n_rows, mode,
                                                               seed))
        return cols

    def gen_trace_poseidon(self, instances_per_row, log_n_rows, mode=1, seed=0, out=None):
        """The trace on the device, into `out` (158 columns per instance of a row) when given."""
        cols = out if out is not None else self.columns(158 * instances_per_row, 1 << log_n_rows)
        if len(cols) != 158 * instances_per_row or _min_words(cols) < (1 << log_n_rows):
            raise ValueError("out must hold 158 columns of 2^log_n_rows words per instance of a row")
        self._check(self.lib.stwo_hip_gen_trace_poseidon(self.ctx, _ptrs(cols), instances_per_row, log_n_rows, mode,
                                                         seed))
        return cols

    def _proof(self, st, h, allow_unsatisfied):
        """A proof from a whole-proof call. Status 6 is prove_ex's ConstraintsNotSatisfied: the
        reference returns no proof, so this raises unless the caller asked for the bytes anyway."""
        if st == ST_UNSATISFIED and h.value:
            pr = Proof(self, h.value)
            if allow_unsatisfied:
                return pr
        self._check(st)
        return Proof(self, h.value)

    def prove_wide_fibonacci(self, trace, log_n_rows, config, allow_unsatisfied=False):
        h = ctypes.c_void_p()
        st = self.lib.stwo_hip_prove_wide_fibonacci(self.ctx, _ptrs(trace), len(trace), log_n_rows,
                                                    ctypes.byref(config), ctypes.byref(h))
        return self._proof(st, h, allow_unsatisfied)

    def prove_wide_fibonacci_multi(self, traces, log_n_rows, config, allow_unsatisfied=False):
        """traces[i]: the columns of component i (2^log_n_rows[i] rows each)."""
        logs = np.array(log_n_rows, np.uint32)
        ncols = np.array([len(t) for t in traces], np.uint32)
        h = ctypes.c_void_p()
        st = self.lib.stwo_hip_prove_wide_fibonacci_multi(self.ctx, _concat_ptrs(traces), len(traces),
                                                          logs.ctypes.data_as(_U32P), ncols.ctypes.data_as(_U32P),
                                                          ctypes.byref(config), ctypes.byref(h))
        return self._proof(st, h, allow_unsatisfied)

    def prove_wide_fibonacci_per_call(self, traces, log_n_rows, config, allow_unsatisfied=False,
                                      merkle_channel="blake2s"):
        """The same proof through the trait entry points, one C-ABI call per trait-method call (the
        plain per-call form). traces[i]: the columns of component i. merkle_channel: "blake2s" or
        "poseidon252" (Poseidon252MerkleChannel: felt252 hashes and channel)."""
        fn = getattr(self.lib, "stwo_hip_prove_wide_fibonacci" + _merkle_channel(merkle_channel)[0])
        logs = np.array(log_n_rows, np.uint32)
        ncols = np.array([len(t) for t in traces], np.uint32)
        h = ctypes.c_void_p()
        st = fn(self.ctx, _concat_ptrs(traces), len(traces), logs.ctypes.data_as(_U32P), ncols.ctypes.data_as(_U32P),
                ctypes.byref(config), ctypes.byref(h))
        return self._proof(st, h, allow_unsatisfied)

    def prove_poseidon_per_call(self, trace, instances_per_row, log_n_rows, config, allow_unsatisfied=False,
                                merkle_channel="blake2s"):
        fn = getattr(self.lib, "stwo_hip_prove_poseidon" + _merkle_channel(merkle_channel)[0])
        h = ctypes.c_void_p()
        st = fn(self.ctx, _ptrs(trace), instances_per_row, log_n_rows, ctypes.byref(config), ctypes.byref(h))
        return self._proof(st, h, allow_unsatisfied)

    def gen_interaction_trace_poseidon(self, trace, instances_per_row, log_n_rows, z, alpha):
        """The LogUp interaction columns (4 per instance of a row) and the claimed sum."""
        cols = self.columns(4 * instances_per_row, 1 << log_n_rows)
        cs = QM31()
        self._check(self.lib.stwo_hip_gen_interaction_trace_poseidon(self.ctx, _ptrs(trace), instances_per_row,
                                                                     log_n_rows, QM31.of(z), QM31.of(alpha),
                                                                     _ptrs(cols), ctypes.byref(cs)))
        return cols, cs.tuple()

    def prove_poseidon(self, trace, instances_per_row, log_n_rows, config, allow_unsatisfied=False):
        """The Poseidon2 AIR's proofs fail prove_ex's sanity check in this protocol version (its
        constraint degree exceeds what the lifted protocol supports), so a caller that wants the
        bytes passes allow_unsatisfied=True and finds status 2 on the proof."""
        h = ctypes.c_void_p()
        st = self.lib.stwo_hip_prove_poseidon(self.ctx, _ptrs(trace), instances_per_row, log_n_rows,
                                              ctypes.byref(config), ctypes.byref(h))
        return self._proof(st, h, allow_unsatisfied)


def felt252_to_words(v):
    """A felt252 (int) as the library's 8 little-endian 32-bit words."""
    return np.array([(int(v) >> (32 * i)) & 0xFFFFFFFF for i in range(8)], np.uint32)


def felt252_from_words(w):
    return sum(int(x) << (32 * i) for i, x in enumerate(np.asarray(w, np.uint32)[:8]))


def poseidon252_hash_many(values):
    """starknet-crypto's poseidon_hash_many on the host (no GPU)."""
    L = load_library()
    v = np.concatenate([felt252_to_words(x) for x in values]) if values else np.zeros(8, np.uint32)
    out = np.zeros(8, np.uint32)
    if L.stwo_poseidon252_hash_many(v.ctypes.data_as(_U32P), len(values), out.ctypes.data_as(_U32P)) != 0:
        raise StwoHipError("non-canonical felt252", 1)
    return felt252_from_words(out)


def poseidon252_permute(state):
    """poseidon_permute_comp of 3 felt252s on the host (no GPU)."""
    L = load_library()
    v = np.concatenate([felt252_to_words(x) for x in state])
    if L.stwo_poseidon252_permute(v.ctypes.data_as(_U32P)) != 0:
        raise StwoHipError("non-canonical felt252", 1)
    return [felt252_from_words(v[8 * i:8 * i + 8]) for i in range(3)]


def verify_wide_fibonacci(proof: bytes, log_n_rows, n_cols, config, merkle_channel="blake2s"):
    """verify() of a wide-Fibonacci proof (one or several components) on the host: "" when it is
    accepted, else the verifier's error. log_n_rows / n_cols: per component."""
    L = load_library()
    fn = getattr(L, "stwo_verify_wide_fibonacci" + _merkle_channel(merkle_channel)[1])
    logs = np.array(log_n_rows, np.uint32).reshape(-1)
    nc = np.array(n_cols, np.uint32).reshape(-1)
    err = ctypes.create_string_buffer(256)
    st = fn(proof, len(proof), len(logs), logs.ctypes.data_as(_U32P), nc.ctypes.data_as(_U32P), ctypes.byref(config),
            err, 256)
    if st == 1:
        raise StwoHipError("invalid argument: " + err.value.decode(), st)
    return "" if st == 0 else err.value.decode()


def verify_poseidon(proof: bytes, instances_per_row, log_n_rows, claimed_sum, config, merkle_channel="blake2s"):
    L = load_library()
    fn = getattr(L, "stwo_verify_poseidon" + _merkle_channel(merkle_channel)[1])
    err = ctypes.create_string_buffer(256)
    st = fn(proof, len(proof), instances_per_row, log_n_rows, QM31.of(claimed_sum), ctypes.byref(config), err, 256)
    if st == 1:
        raise StwoHipError("invalid argument: " + err.value.decode(), st)
    return "" if st == 0 else err.value.decode()


def fnv1a(name: str) -> int:
    """Component id of the reference's dispatcher (crates/constraint-framework/src/lib.rs fnv1a_eval_id_gen)."""
    h = 0x811C9DC5
    for b in name.encode():
        h ^= b
        h = (h * 0x01000193) & 0xFFFFFFFF
    return h
