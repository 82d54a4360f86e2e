"""The constraint recorder of the composition compiler (nitroozk-stwo_amd/framework.py), pinned on
the host: its programs for the wide-Fibonacci and Poseidon2 AIRs, interpreted on random values,
give the constraint values of the independent Python restatement (tests/pystark.py) in the same
order, and a planted change to a recorded AIR changes them."""
import importlib.util
import os
import struct

import numpy as np
import pytest

import pystark as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("framework", os.path.join(ROOT, "nitroozk-stwo_amd", "framework.py"))
F = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(F)
P = S.P


def interpret(prog_bytes, mask, params):
    """Evaluates every node of a program in QM31 (base values embedded), returns the constraints."""
    w = struct.unpack("<%dI" % (len(prog_bytes) // 4), prog_bytes)
    assert w[0] == F.MAGIC and w[1] == 1
    n_nodes, n_cons = w[2], w[3]
    vals = []
    for i in range(n_nodes):
        op, a, b, c, d = w[5 + 5 * i: 10 + 5 * i]
        assert a < i or op in (F.OP_COL, F.OP_CONST, F.OP_PARAM)
        if op == F.OP_COL:
            off = c - (1 << 32) if c >= (1 << 31) else c
            v = mask(a, b, off)
        elif op == F.OP_CONST:
            v = S.q_from(a)
        elif op == F.OP_PARAM:
            v = params[a]
        elif op == F.OP_ADD:
            v = S.q_add(vals[a], vals[b])
        elif op == F.OP_SUB:
            v = S.q_sub(vals[a], vals[b])
        elif op == F.OP_MUL:
            v = S.q_mul(vals[a], vals[b])
        elif op == F.OP_NEG:
            v = S.q_neg(vals[a])
        elif op == F.OP_EXT:
            v = S.from_coords([vals[a], vals[b], vals[c], vals[d]])
        else:
            raise AssertionError(op)
        vals.append(v)
    roots = w[5 + 5 * n_nodes: 5 + 5 * n_nodes + n_cons]
    assert len(roots) == n_cons and 5 + 5 * n_nodes + n_cons == len(w)
    return [vals[r] for r in roots]


def rand_q(rng):
    return tuple(int(x) for x in rng.integers(0, P, 4))


@pytest.mark.parametrize("n_cols", [3, 10, 57])
def test_wide_fibonacci_program_matches_restatement(n_cols):
    rng = np.random.default_rng(n_cols)
    tv = [int(x) for x in rng.integers(0, P, n_cols)]
    mask = lambda tree, col, off: S.q_from(tv[col])
    got = interpret(F.wide_fibonacci_program(n_cols).to_bytes(), mask, [])
    exp = S.WideFib(3, n_cols).constraints(lambda t, c, o: S.q_from(tv[c]))
    assert got == exp and len(got) == n_cols - 2


@pytest.mark.parametrize("ipr", [1, 2, 3])
def test_poseidon_program_matches_restatement(ipr):
    rng = np.random.default_rng(40 + ipr)
    tv = [int(x) for x in rng.integers(0, P, 158 * ipr)]
    iv = {(c, o): int(rng.integers(0, P)) for c in range(4 * ipr) for o in (-1, 0)}
    z, alpha, shift = rand_q(rng), rand_q(rng), rand_q(rng)
    params = [shift] + F.poseidon_params(z, alpha)
    mask = lambda tree, col, off: S.q_from(tv[col] if tree == 1 else iv[(col, off)])
    prog = F.poseidon_program(ipr)
    got = interpret(prog.to_bytes(), mask, params)
    alpha_pows = [S.q_pow(alpha, i) for i in range(16)]
    exp = S.poseidon_constraints(ipr, lambda c: S.q_from(tv[c]), lambda c, o: S.q_from(iv[(c, o)]), z, alpha_pows,
                                 shift)
    assert len(got) == 143 * ipr
    assert got == exp


def test_planted_round_constant_changes_the_program():
    rng = np.random.default_rng(9)
    tv = [int(x) for x in rng.integers(0, P, 158)]
    iv = {(c, o): int(rng.integers(0, P)) for c in range(4) for o in (-1, 0)}
    params = [rand_q(rng)] + F.poseidon_params(rand_q(rng), rand_q(rng))
    mask = lambda tree, col, off: S.q_from(tv[col] if tree == 1 else iv[(col, off)])
    good = interpret(F.poseidon_program(1).to_bytes(), mask, params)
    old = F.PSD_CONST
    try:
        F.PSD_CONST = old + 1
        bad = interpret(F.poseidon_program(1).to_bytes(), mask, params)
    finally:
        F.PSD_CONST = old
    assert bad != good


def test_program_sharing_and_blob_layout():
    prog = F.wide_fibonacci_program(5)
    # a*a for the shared column is recorded once per distinct product
    assert len(prog.nodes) == len(set(prog.nodes))
    blob = F.expr_blob(prog, [(1, 2, 3, 4)])
    h, n = struct.unpack("<II", blob[:8])
    assert n * 4 == len(prog.to_bytes())
    assert struct.unpack("<I", blob[8 + 4 * n: 12 + 4 * n])[0] == 1
