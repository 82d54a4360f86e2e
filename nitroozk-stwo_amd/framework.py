"""Constraint expressions for the composition compiler of libstwo_hip.

The host-side counterpart of the reference's constraint framework for this path
(crates/constraint-framework/src/{lib.rs,logup.rs,expr/}): an AIR describes its constraints once
through an EvalAtRow-like recorder (next_trace_mask, next_interaction_mask, add_constraint,
add_to_relation, finalize_logup_in_pairs), which records an expression DAG instead of computing
values, as the reference's ExprEvaluator does. The DAG serialises to a small program that
stwo_hip_evaluate_constraint_quotients_on_domain compiles into a HIP kernel at run time (blob id
fnv1a("expr_component")), so any component gets a device kernel without hand-written code.

Program words (u32, little endian):
    magic 0x58454853 ("SHEX"), version 1, n_nodes, n_constraints, n_params
    n_nodes nodes of 5 words (op, a, b, c, d):
        1 COL    a = tree (1 trace, 2 interaction), b = column, c = row offset (signed, in trace rows)
        2 CONST  a = M31 value
        3 PARAM  a = index into the parameter table (0 = the cumulative-sum shift, then the blob's)
        4 ADD a b   5 SUB a b   6 MUL a b   7 NEG a
        8 EXT  a b c d: the secure value a + b i + c u + d iu of four values (from_partial_evals)
    n_constraints node indices: constraint k (coefficient power k) is that node's value.
Operands refer to earlier nodes. COL and CONST are base-field (M31) values; PARAM and EXT are
secure (QM31); arithmetic is secure when an operand is.
"""
import struct

MAGIC = 0x58454853
OP_COL, OP_CONST, OP_PARAM, OP_ADD, OP_SUB, OP_MUL, OP_NEG, OP_EXT = range(1, 9)
P = (1 << 31) - 1


class Expr:
    __slots__ = ("prog", "idx", "secure")

    def __init__(self, prog, idx, secure):
        self.prog, self.idx, self.secure = prog, idx, secure

    def _wrap(self, o):
        if isinstance(o, Expr):
            return o
        return self.prog.const(o)

    def __add__(self, o):
        return self.prog.node(OP_ADD, self, self._wrap(o))

    __radd__ = __add__

    def __sub__(self, o):
        return self.prog.node(OP_SUB, self, self._wrap(o))

    def __rsub__(self, o):
        return self.prog.node(OP_SUB, self._wrap(o), self)

    def __mul__(self, o):
        return self.prog.node(OP_MUL, self, self._wrap(o))

    __rmul__ = __mul__

    def __neg__(self):
        return self.prog.node(OP_NEG, self)


class Program:
    """An expression DAG with structural sharing (identical nodes are recorded once)."""

    def __init__(self):
        self.nodes, self.secure, self.memo = [], [], {}
        self.constraints = []
        self.n_params = 0

    def _add(self, key, secure):
        if key in self.memo:
            return Expr(self, self.memo[key], self.secure[self.memo[key]])
        self.nodes.append(key)
        self.secure.append(secure)
        self.memo[key] = len(self.nodes) - 1
        return Expr(self, len(self.nodes) - 1, secure)

    def col(self, tree, col, offset=0):
        return self._add((OP_COL, tree, col, offset & 0xFFFFFFFF, 0), False)

    def const(self, v):
        return self._add((OP_CONST, int(v) % P, 0, 0, 0), False)

    def param(self, k):
        self.n_params = max(self.n_params, k)
        return self._add((OP_PARAM, k, 0, 0, 0), True)

    def node(self, op, a, b=None):
        sec = a.secure or (b is not None and b.secure)
        return self._add((op, a.idx, 0 if b is None else b.idx, 0, 0), sec)

    def ext(self, a, b, c, d):
        return self._add((OP_EXT, a.idx, b.idx, c.idx, d.idx), True)

    def to_bytes(self):
        words = [MAGIC, 1, len(self.nodes), len(self.constraints), self.n_params]
        for n in self.nodes:
            words += list(n)
        words += [c.idx for c in self.constraints]
        return struct.pack("<%dI" % len(words), *words)


class EvalAtRow:
    """The recorder an AIR's evaluate() runs against (EvalAtRow, crates/constraint-framework/src/lib.rs)."""

    def __init__(self):
        self.prog = Program()
        self.trace_col = 0
        self.inter_col = 0
        self.fracs = []  # (numerator, denominator) pairs added to the relation, in order

    def next_trace_mask(self):
        e = self.prog.col(1, self.trace_col, 0)
        self.trace_col += 1
        return e

    def next_interaction_mask(self, offsets):
        r = [self.prog.col(2, self.inter_col, o) for o in offsets]
        self.inter_col += 1
        return r

    def add_constraint(self, e):
        self.prog.constraints.append(e)

    def combine_ef(self, a, b, c, d):
        return self.prog.ext(a, b, c, d)

    def param(self, k):
        return self.prog.param(k)

    # ---- LogUp (crates/constraint-framework/src/lib.rs:149-258, logup.rs:85-100) ----
    def add_to_relation(self, numerator, denominator):
        self.fracs.append((numerator, denominator))

    def finalize_logup_in_pairs(self, shift):
        """One constraint per pair of fractions: (cum_b - cum_{b-1}) d0 d1 - (n0 d1 + n1 d0), the
        last pair against the previous row's last cumulative column plus the claimed-sum shift."""
        n = len(self.fracs) // 2
        prev = None
        for b in range(n):
            (n0, d0), (n1, d1) = self.fracs[2 * b], self.fracs[2 * b + 1]
            num = n0 * d1 + n1 * d0
            den = d0 * d1
            if b + 1 < n:
                cur = self.combine_ef(*[m[0] for m in (self.next_interaction_mask([0]) for _ in range(4))])
                diff = cur if prev is None else cur - prev
                prev = cur
            else:
                masks = [self.next_interaction_mask([-1, 0]) for _ in range(4)]
                prev_row = self.combine_ef(*[m[0] for m in masks])
                cur = self.combine_ef(*[m[1] for m in masks])
                diff = cur - prev_row + shift
                if prev is not None:
                    diff = diff - prev
            self.add_constraint(diff * den - num)
        self.fracs = []


def lookup_combine(e, z, alpha_pows, values):
    """LookupElements::combine: sum alpha^i v_i - z."""
    acc = None
    for v, a in zip(values, alpha_pows):
        t = a * v
        acc = t if acc is None else acc + t
    return acc - z


# ---- the two benchmark AIRs, written once against the recorder ------------------------------------
def wide_fibonacci_program(n_cols):
    """crates/examples/src/wide_fibonacci/mod.rs:34-59: col(i) = col(i-2)^2 + col(i-1)^2."""
    e = EvalAtRow()
    a, b = e.next_trace_mask(), e.next_trace_mask()
    for _ in range(2, n_cols):
        c = e.next_trace_mask()
        e.add_constraint(c - (a * a + b * b))
        a, b = b, c
    return e.prog


PSD_STATE, PSD_PARTIAL, PSD_HALF_FULL, PSD_CONST = 16, 14, 4, 1234


def _m4(x):
    t0 = x[0] + x[1]
    t02 = t0 + t0
    t1 = x[2] + x[3]
    t12 = t1 + t1
    t2 = x[1] + x[1] + t1
    t3 = x[3] + x[3] + t0
    t4 = t12 + t12 + t3
    t5 = t02 + t02 + t2
    return [t3 + t5, t5, t2 + t4, t4]


def _external(s):
    s = sum((_m4(s[4 * i:4 * i + 4]) for i in range(4)), [])
    for j in range(4):
        t = s[j] + s[j + 4] + (s[j + 8] + s[j + 12])
        for i in range(4):
            s[4 * i + j] = s[4 * i + j] + t
    return s


def _internal(s):
    tot = s[0]
    for x in s[1:]:
        tot = tot + x
    return [x * (1 << (i + 1)) + tot for i, x in enumerate(s)]


def _pow5(x):
    x2 = x * x
    return x2 * x2 * x


def poseidon_program(ipr):
    """crates/examples/src/poseidon/mod.rs:150-210 (eval_poseidon_constraints) for ipr instances
    per row, with the LogUp relation of each instance's initial and final state. Parameters: 1 = z,
    2..17 = alpha^0 .. alpha^15 (the lookup elements), 0 = the cumulative-sum shift."""
    e = EvalAtRow()
    z = e.param(1)
    alpha = [e.param(2 + i) for i in range(16)]
    for _ in range(ipr):
        s = [e.next_trace_mask() for _ in range(PSD_STATE)]
        init = list(s)
        for half in range(2):
            if half == 1:
                for _ in range(PSD_PARTIAL):
                    s[0] = s[0] + PSD_CONST
                    s = _internal(s)
                    s[0] = _pow5(s[0])
                    m = e.next_trace_mask()
                    e.add_constraint(s[0] - m)
                    s[0] = m
            for _ in range(PSD_HALF_FULL):
                s = _external([x + PSD_CONST for x in s])
                s = [_pow5(x) for x in s]
                for i in range(PSD_STATE):
                    m = e.next_trace_mask()
                    e.add_constraint(s[i] - m)
                    s[i] = m
        e.add_to_relation(e.prog.const(1), lookup_combine(e, z, alpha, init))
        e.add_to_relation(e.prog.const(P - 1), lookup_combine(e, z, alpha, s))
    e.finalize_logup_in_pairs(e.param(0))
    return e.prog


def poseidon_params(z, alpha):
    """The parameter table of poseidon_program after the shift: z, then alpha^0 .. alpha^15."""
    def qmul(a, b):
        a0, a1, b0, b1 = (a[0], a[1]), (a[2], a[3]), (b[0], b[1]), (b[2], b[3])
        cm = lambda x, y: ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)
        t = cm(a1, b1)
        r = cm((2, 1), t)
        c0 = cm(a0, b0)
        x, y = cm(a0, b1), cm(a1, b0)
        return ((c0[0] + r[0]) % P, (c0[1] + r[1]) % P, (x[0] + y[0]) % P, (x[1] + y[1]) % P)
    out, p = [tuple(z)], (1, 0, 0, 0)
    for _ in range(16):
        out.append(p)
        p = qmul(p, tuple(alpha))
    return out


def expr_blob(program, params):
    """eval_blob of evaluate_constraint_quotients_on_domain for a compiled component:
    fnv1a("expr_component") || u32 program words || program || u32 n_params || n_params QM31."""
    body = program.to_bytes()
    h = 0x811C9DC5
    for ch in b"expr_component":
        h ^= ch
        h = (h * 0x01000193) & 0xFFFFFFFF
    flat = [int(x) for p in params for x in p]
    return (struct.pack("<II", h, len(body) // 4) + body + struct.pack("<I", len(params)) +
            struct.pack("<%dI" % len(flat), *flat))
