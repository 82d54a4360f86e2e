"""Pure-Python restatement of the circle-STARK prover over M31 with the Blake2s channel, for the
wide-Fibonacci AIR (crates/examples/src/wide_fibonacci/mod.rs) and the Poseidon2 AIR with its LogUp
relation (crates/examples/src/poseidon/mod.rs), at tiny sizes.

Test infrastructure: it pins the C++ oracle (and, through it, the HIP prover). It is written from
the reference's definitions and from the mathematics, independently of the oracle's code, and it
avoids the index formulas the fast implementations use wherever a geometric statement exists:

- field and circle arithmetic on Python integers (crates/stwo/src/core/fields, core/circle.rs);
- polynomials only through the FFT basis y^b0 x^b1 pi(x)^b2 ... (prover/poly/circle/ops.rs:125-141):
  evaluation by brute force, interpolation by the even/odd recursion on explicit point sets;
- lifting as evaluation at the repeatedly doubled point, mask offsets as point shifts, the
  quotient as one sum over every (column, sample) with its random power (no batching or
  sorting), folds on explicit (point, conjugate) and (x, -x) pairs;
- Merkle trees, the channel and proof of work on hashlib's Blake2s (RFC 7693).

Reference files followed (relative to crates/stwo/src unless noted):
  channel, PoW            core/channel/blake2s.rs:19-140; core/vcs/blake2_merkle.rs:17-58
  queries                 core/queries.rs:11-60
  prove / prove_ex        prover/mod.rs:28-150
  commitment scheme       prover/pcs/mod.rs:146-494; core/pcs/utils.rs:200-225
  FRI quotients           core/pcs/quotients.rs:200-345; core/constraints.rs:1-40,119-134
  FRI prover              prover/fri.rs:100-412; prover/backend/cpu/fri.rs:15-120; prover/line.rs:54-99
  Merkle prover           prover/vcs/prover.rs:39-217
  composition             prover/air/accumulation.rs:21-125; prover/air/component_prover.rs:99-114;
                          crates/constraint-framework/src/prover/component_prover.rs:32-169
  LogUp                   crates/constraint-framework/src/{lib.rs:140-258, logup.rs:62-100,
                          prover/logup.rs:21-215}; prover/backend/simd/prefix_sum.rs:15-22
  OODS sample points      crates/constraint-framework/src/component.rs:264-276; core/proof.rs
  random point            core/circle.rs:169-181

The proof byte encoding is this library's (the reference derives serde but fixes no format):
little-endian u32 words, counts as u32, fields in StarkProof order.
"""
import hashlib

import p252

P = (1 << 31) - 1
GEN_INDEX_LOG = 31
M31_GEN = (2, 1268011823)


# ---------------------------------------------------------------- fields ----------------------
def m_inv(a):
    assert a % P
    return pow(a, P - 2, P)


def c_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def c_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def c_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def c_inv(a):
    n = (a[0] * a[0] + a[1] * a[1]) % P
    ni = m_inv(n)
    return (a[0] * ni % P, (-a[1]) * ni % P)


QZERO = (0, 0, 0, 0)
QONE = (1, 0, 0, 0)


def q(*v):
    return tuple(x % P for x in v)


def q_from(m):
    return (m % P, 0, 0, 0)


def q_add(a, b):
    return tuple((x + y) % P for x, y in zip(a, b))


def q_sub(a, b):
    return tuple((x - y) % P for x, y in zip(a, b))


def q_neg(a):
    return tuple((-x) % P for x in a)


def q_mul(a, b):
    """(a0 + a1 u)(b0 + b1 u) with u^2 = 2 + i (core/fields/qm31.rs)."""
    a0, a1, b0, b1 = (a[0], a[1]), (a[2], a[3]), (b[0], b[1]), (b[2], b[3])
    r = c_mul((2, 1), c_mul(a1, b1))
    lo = c_add(c_mul(a0, b0), r)
    hi = c_add(c_mul(a0, b1), c_mul(a1, b0))
    return (lo[0], lo[1], hi[0], hi[1])


def q_mul_m(a, m):
    return tuple(x * m % P for x in a)


def q_conj(a):
    """QM31 complex conjugate: (a0 + a1 u) -> (a0 - a1 u) (core/fields/mod.rs:413)."""
    return (a[0], a[1], (-a[2]) % P, (-a[3]) % P)


def q_inv(a):
    """Inverse by the norm to CM31: a^-1 = conj(a) / (a * conj(a))."""
    n = q_mul(a, q_conj(a))
    assert n[2] == 0 and n[3] == 0
    ni = c_inv((n[0], n[1]))
    return q_mul(q_conj(a), (ni[0], ni[1], 0, 0))


def q_pow(a, e):
    r = QONE
    while e:
        if e & 1:
            r = q_mul(r, a)
        a = q_mul(a, a)
        e >>= 1
    return r


# ---------------------------------------------------------------- circle -----------------------
def mp_add(p, r):
    return ((p[0] * r[0] - p[1] * r[1]) % P, (p[0] * r[1] + p[1] * r[0]) % P)


def mp_conj(p):
    return (p[0], (-p[1]) % P)


def mp_mul(p, k):
    acc, cur = (1, 0), p
    while k:
        if k & 1:
            acc = mp_add(acc, cur)
        cur = mp_add(cur, cur)
        k >>= 1
    return acc


def index_point(i):
    """The circle point of index i (multiple of the generator of order 2^31)."""
    return mp_mul(M31_GEN, i % (1 << GEN_INDEX_LOG))


def qp_add(p, r):
    return (q_sub(q_mul(p[0], r[0]), q_mul(p[1], r[1])), q_add(q_mul(p[0], r[1]), q_mul(p[1], r[0])))


def qp_double(p):
    return qp_add(p, p)


def qp_repeated_double(p, n):
    for _ in range(n):
        p = qp_double(p)
    return p


def qp_from(mp):
    return (q_from(mp[0]), q_from(mp[1]))


def double_x(x):
    return (2 * x * x - 1) % P


def q_double_x(x):
    s = q_mul(x, x)
    return q_sub(q_add(s, s), QONE)


class Coset:
    """initial_index + <step> of 2^log points (core/circle.rs:284-420)."""

    def __init__(self, initial_index, log_size):
        self.initial_index = initial_index % (1 << 31)
        self.log_size = log_size
        self.step_index = 1 << (31 - log_size)

    def index_at(self, i):
        return (self.initial_index + self.step_index * i) % (1 << 31)

    def at(self, i):
        return index_point(self.index_at(i))

    def points(self):
        return [self.at(i) for i in range(1 << self.log_size)]


def canonic_coset(n):
    """CanonicCoset(n).coset() = Coset::odds(n) = G_{2^(n+1)} + <G_{2^n}>."""
    return Coset(1 << (31 - (n + 1)), n)


def half_odds(n):
    return Coset(1 << (31 - (n + 2)), n)


def circle_domain(n):
    """CircleDomain of CanonicCoset(n) in natural order: the half coset Coset::half_odds(n-1), then
    the conjugates of its points (core/poly/circle/domain.rs:57-68)."""
    half = half_odds(n - 1).points()
    return half + [mp_conj(p) for p in half]


def bitrev(i, n):
    r = 0
    for _ in range(n):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def domain_bitrev(n):
    """Points of CanonicCoset(n)'s circle domain in bit-reversed order (storage order of evaluations)."""
    nat = circle_domain(n)
    return [nat[bitrev(r, n)] for r in range(1 << n)]


def coset_vanishing(coset, p):
    """core/constraints.rs coset_vanishing: shift p by -initial + step/2 and double the x
    coordinate log_size - 1 times. p is a QM31 point."""
    shift = mp_add(mp_conj(index_point(coset.initial_index)), index_point(coset.step_index // 2))
    pp = qp_add(p, qp_from(shift))
    x = pp[0]
    for _ in range(1, coset.log_size):
        x = q_double_x(x)
    return x


# ---------------------------------------------------------------- polynomials ------------------
def basis_value(k, L, x, y, mul, one, dx):
    """FFT basis element k of a 2^L coefficient polynomial at (x, y): y^b0 x^b1 pi(x)^b2 ...
    (prover/poly/circle/ops.rs:125-141), over a field given by its mul/one/double_x."""
    v = one
    if k & 1:
        v = mul(v, y)
    xx = x
    for b in range(1, L):
        if (k >> b) & 1:
            v = mul(v, xx)
        xx = dx(xx)
    return v


def eval_m(coeffs, pt):
    """Value of FFT-basis coefficients at an M31 point (brute force)."""
    L = len(coeffs).bit_length() - 1
    mul = lambda a, b: a * b % P
    return sum(c * basis_value(k, L, pt[0], pt[1], mul, 1, double_x) for k, c in enumerate(coeffs)) % P


def eval_q(coeffs, pt):
    """Value of FFT-basis (M31) coefficients at a QM31 point (brute force)."""
    L = len(coeffs).bit_length() - 1
    acc = QZERO
    for k, c in enumerate(coeffs):
        if c:
            acc = q_add(acc, q_mul_m(basis_value(k, L, pt[0], pt[1], q_mul, QONE, q_double_x), c))
    return acc


def evaluate(coeffs, n):
    """Evaluations on CanonicCoset(n) in bit-reversed order."""
    return [eval_m(coeffs, p) for p in domain_bitrev(n)]


def _interp_line(xs, vals):
    """Coefficients of the line basis 1, x, pi(x), x pi(x), ... (index bit j <-> pi^j(x)) of the
    function given by its values at the points xs, a set closed under x -> -x: g(x) = g0(pi(x)) +
    x g1(pi(x)), with g0, g1 found from each pair (x, -x) and interpolated on the doubled set."""
    if len(xs) == 1:
        return [vals[0] % P]
    where = {x: i for i, x in enumerate(xs)}
    seen, nx, e, o = set(), [], [], []
    for i, x in enumerate(xs):
        if x in seen:
            continue
        j = where[(-x) % P]
        seen.add(x)
        seen.add((-x) % P)
        a, b = vals[i], vals[j]
        inv2 = m_inv(2)
        nx.append(double_x(x))
        e.append((a + b) * inv2 % P)
        o.append((a - b) * inv2 * m_inv(x) % P)
    ce, co = _interp_line(nx, e), _interp_line(nx, o)
    out = []
    for a, b in zip(ce, co):
        out += [a, b]
    return out


def interpolate(vals, n):
    """FFT-basis coefficients of the evaluations vals on CanonicCoset(n) (bit-reversed order):
    f(x, y) = f0(x) + y f1(x), with f0, f1 from each (point, conjugate) pair, then the line
    recursion on their x-coordinates."""
    pts = domain_bitrev(n)
    where = {p: i for i, p in enumerate(pts)}
    seen, xs, e, o = set(), [], [], []
    inv2 = m_inv(2)
    for i, p in enumerate(pts):
        if p in seen:
            continue
        c = mp_conj(p)
        j = where[c]
        seen.add(p)
        seen.add(c)
        a, b = vals[i], vals[j]
        xs.append(p[0])
        e.append((a + b) * inv2 % P)
        o.append((a - b) * inv2 * m_inv(p[1]) % P)
    ce, co = _interp_line(xs, e), _interp_line(xs, o)
    out = []
    for a, b in zip(ce, co):
        out += [a, b]
    return out


# ---------------------------------------------------------------- hashing / channel ------------
def blake2s(data):
    return hashlib.blake2s(data, digest_size=32).digest()


def le32(v):
    return (v % (1 << 32)).to_bytes(4, "little")


def hash_node(children, values):
    """Blake2sMerkleHasher::hash_node: left || right (if any) || column values as u32 LE."""
    data = b""
    if children is not None:
        data += children[0] + children[1]
    return blake2s(data + b"".join(le32(v) for v in values))


class Channel:
    """Blake2sChannel (core/channel/blake2s.rs:19-140)."""

    POW_PREFIX = 0x12345678

    def __init__(self):
        self.digest = bytes(32)
        self.n_draws = 0

    def update_digest(self, d):
        self.digest = d
        self.n_draws = 0

    def mix_root(self, root):
        self.update_digest(blake2s(self.digest + root))

    def mix_u32s(self, words):
        self.update_digest(blake2s(self.digest + b"".join(le32(w) for w in words)))

    def mix_u64(self, v):
        self.mix_u32s([v & 0xFFFFFFFF, v >> 32])

    def mix_felts(self, felts):
        self.update_digest(blake2s(self.digest + b"".join(le32(c) for f in felts for c in f)))

    def draw_u32s(self):
        h = blake2s(self.digest + le32(self.n_draws) + b"\x00")
        self.n_draws += 1
        return [int.from_bytes(h[4 * i:4 * i + 4], "little") for i in range(8)]

    def draw_base_felts(self):
        while True:
            w = self.draw_u32s()
            if all(x < 2 * P for x in w):
                return [x % P for x in w]

    def draw_secure_felt(self):
        return tuple(self.draw_base_felts()[:4])

    def draw_secure_felts(self, n):
        out, buf = [], []
        while len(out) < n:
            if len(buf) < 4:
                buf += self.draw_base_felts()
            out.append(tuple(buf[:4]))
            buf = buf[4:]
        return out

    def pow_prefix(self, bits):
        return blake2s(le32(self.POW_PREFIX) + bytes(12) + self.digest + le32(bits))

    def verify_pow_nonce(self, bits, nonce):
        h = blake2s(self.pow_prefix(bits) + nonce.to_bytes(8, "little"))
        v = int.from_bytes(h[:16], "little")
        tz = 128 if v == 0 else (v & -v).bit_length() - 1
        return tz >= bits


def grind(channel, bits):
    """Smallest nonce in the search order of GrindOps (simd/grind.rs:20-91): n = hi * 2^32 + lo,
    lo < 2^20, tried as hi * 2^20 + lo; for the small bit counts used here the first solution has
    hi = 0, so the order is plain counting up from 0."""
    n = 0
    while True:
        if channel.verify_pow_nonce(bits, n):
            return n
        n += 1
        assert n < (1 << 20), "solution beyond hi = 0 not handled in this restatement"


class Blake2sMerkleChannel:
    """Blake2sMerkleChannel (core/vcs/blake2_merkle.rs:45-59): Blake2s nodes and channel."""
    Channel = Channel
    hash_node = staticmethod(hash_node)
    grind = staticmethod(grind)


class Poseidon252MerkleChannel:
    """Poseidon252MerkleChannel (core/vcs/poseidon252_merkle.rs:78-85): felt252 Poseidon nodes and
    channel (tests/p252.py); grinding in the (hi, lo < 2^14) order of GrindOps<Poseidon252Channel>.
    A hash is an int below the Stark prime; the proof carries it as 32 little-endian bytes."""
    Channel = p252.Channel
    hash_node = staticmethod(p252.hash_node)

    @staticmethod
    def grind(channel, bits):
        return p252.grind(channel.digest, bits)


def draw_queries(channel, log_domain, n_queries):
    """core/queries.rs:11-29: the words of each draw (8 for Blake2s, 7 for Poseidon252), each masked
    to the domain."""
    out = []
    mask = (1 << log_domain) - 1
    while True:
        for w in channel.draw_u32s():
            out.append(w & mask)
            if len(out) == n_queries:
                return out


def get_random_point(channel):
    """core/circle.rs:169-181: t drawn, x = (1 - t^2)/(1 + t^2), y = 2t/(1 + t^2)."""
    t = channel.draw_secure_felt()
    t2 = q_mul(t, t)
    inv = q_inv(q_add(t2, QONE))
    return (q_mul(q_sub(QONE, t2), inv), q_mul(q_add(t, t), inv))


# ---------------------------------------------------------------- Merkle -----------------------
class MerkleTree:
    """MerkleProver::commit (prover/vcs/prover.rs:39-65): columns by length descending; for each
    log size from the largest down to 0, node i = H(children 2i, 2i+1 of the layer below, then the
    values at i of the columns of that size, in that order)."""

    def __init__(self, columns, hash_node=hash_node):
        cols = sorted(columns, key=lambda c: -len(c))  # stable
        if not cols:
            self.layers = [[hash_node(None, [])]]
            return
        top = len(cols[0]).bit_length() - 1
        layers, prev, ci = [], None, 0
        for ls in range(top, -1, -1):
            here = []
            while ci < len(cols) and len(cols[ci]).bit_length() - 1 == ls:
                here.append(cols[ci])
                ci += 1
            layer = [hash_node(None if prev is None else (prev[2 * i], prev[2 * i + 1]), [c[i] for c in here])
                     for i in range(1 << ls)]
            layers.append(layer)
            prev = layer
        self.layers = layers[::-1]  # layers[k] has 2^k nodes
        self.cols = cols

    def root(self):
        return self.layers[0][0]

    def decommit(self, queries_per_log_size, columns):
        """MerkleProver::decommit (prover/vcs/prover.rs:84-217). Returns (queried values,
        hash witness, column witness)."""
        cols = sorted(columns, key=lambda c: -len(c))
        queried, hash_w, col_w = [], [], []
        last = []
        ci = 0
        for ls in range(len(self.layers) - 1, -1, -1):
            has_prev = ls + 1 < len(self.layers)
            prev_set = set(last)
            colq = set(queries_per_log_size.get(ls, []))
            merged = sorted(set(q // 2 for q in last) | colq)
            here = []
            while ci < len(cols) and len(cols[ci]).bit_length() - 1 == ls:
                here.append(cols[ci])
                ci += 1
            qnodes, wnodes = [], []
            for node in merged:
                if has_prev:
                    for ch in (2 * node, 2 * node + 1):
                        if ch not in prev_set:
                            hash_w.append(self.layers[ls + 1][ch])
                (qnodes if node in colq else wnodes).append(node)
            if here:
                for node in qnodes:
                    queried += [c[node] for c in here]
                for node in wnodes:
                    col_w += [c[node] for c in here]
            last = merged
        return queried, hash_w, col_w


# ---------------------------------------------------------------- FRI --------------------------
def fold_circle_into_line(values, points, alpha):
    """One circle-to-line fold over explicit points (prover/backend/cpu/fri.rs, core/fri.rs:751-780):
    each (p, conj p) pair at positions (2i, 2i+1) gives f0 = f(p) + f(conj p),
    f1 = (f(p) - f(conj p)) / p.y and the value f0 + alpha f1 at p.x."""
    out, xs = [], []
    for i in range(0, len(values), 2):
        p, pc = points[i], points[i + 1]
        assert pc == mp_conj(p)
        f0 = q_add(values[i], values[i + 1])
        f1 = q_mul_m(q_sub(values[i], values[i + 1]), m_inv(p[1]))
        out.append(q_add(f0, q_mul(alpha, f1)))
        xs.append(p[0])
    return out, xs


def fold_line(values, xs, alpha):
    """One line fold (prover/backend/cpu/fri.rs fold_line_cpu): each (x, -x) pair at positions
    (2i, 2i+1) gives f0 = f(x) + f(-x), f1 = (f(x) - f(-x)) / x and f0 + alpha f1 at pi(x)."""
    out, nx = [], []
    for i in range(0, len(values), 2):
        x, xn = xs[i], xs[i + 1]
        assert xn == (-x) % P
        f0 = q_add(values[i], values[i + 1])
        f1 = q_mul_m(q_sub(values[i], values[i + 1]), m_inv(x))
        out.append(q_add(f0, q_mul(alpha, f1)))
        nx.append(double_x(x))
    return out, nx


def line_coefficients(values, xs):
    """Coefficients of the line basis in natural order (1, x, pi(x), x pi(x), ...) of QM31 values
    at the points xs, coordinate by coordinate."""
    per = [_interp_line(xs, [v[k] for v in values]) for k in range(4)]
    return [tuple(per[k][i] for k in range(4)) for i in range(len(values))]


def decommit_positions(n_values, positions, fold_step):
    """compute_decommitment_positions_and_witness_evals (prover/fri.rs:383-412): for each fold
    coset holding queries, all its 2^fold_step positions; the witness is the values at positions
    that were not queried."""
    out_pos, wit = [], []
    groups = {}
    for p in positions:
        groups.setdefault(p >> fold_step, []).append(p)
    for g in sorted(groups):
        qs = set(groups[g])
        for pos in range(g << fold_step, (g + 1) << fold_step):
            out_pos.append(pos)
            if pos not in qs:
                wit.append(pos)
    return out_pos, wit


class FriProver:
    """FriProver::commit / decommit (prover/fri.rs:100-360)."""

    def __init__(self, channel, cfg, quotient, L, hash_node=hash_node):
        self.cfg = cfg
        self.first = quotient  # QM31 values on CanonicCoset(L), bit-reversed
        self.L = L
        cols = [[v[k] for v in quotient] for k in range(4)]
        self.first_tree = MerkleTree(cols, hash_node)
        channel.mix_root(self.first_tree.root())
        last_log = cfg["log_last_layer"] + cfg["log_blowup"]
        alpha = channel.draw_secure_felt()
        vals, xs = fold_circle_into_line(quotient, domain_bitrev(L), alpha)
        self.inner = []  # (values, tree, fold_step)
        step = cfg["fold_step"]

        def fold_n(vals, xs, alpha, k):
            for _ in range(k):
                vals, xs = fold_line(vals, xs, alpha)
                alpha = q_mul(alpha, alpha)
            return vals, xs

        if L - 1 != last_log:
            while len(vals).bit_length() - 1 > last_log + step:
                tree = MerkleTree([[v[k] for v in vals] for k in range(4)], hash_node)
                self.inner.append((vals, tree, step))
                channel.mix_root(tree.root())
                alpha = channel.draw_secure_felt()
                vals, xs = fold_n(vals, xs, alpha, step)
            k = len(vals).bit_length() - 1 - last_log
            tree = MerkleTree([[v[kk] for v in vals] for kk in range(4)], hash_node)
            self.inner.append((vals, tree, k))
            channel.mix_root(tree.root())
            alpha = channel.draw_secure_felt()
            vals, xs = fold_n(vals, xs, alpha, k)
        # last layer: natural-order coefficients, the tail must vanish; the polynomial is stored
        # (and mixed) with its coefficients in bit-reversed order (core/poly/line.rs:115-170)
        coeffs = line_coefficients(vals, xs)
        bound = 1 << cfg["log_last_layer"]
        assert all(c == QZERO for c in coeffs[bound:]), "invalid degree"
        kept = coeffs[:bound]
        lb = bound.bit_length() - 1
        self.last_poly = [kept[bitrev(i, lb)] for i in range(bound)]
        channel.mix_felts(self.last_poly)

    def decommit(self, channel):
        raw = draw_queries(channel, self.L, self.cfg["n_queries"])
        queries = sorted(set(raw))
        layers = []
        pos, wit = decommit_positions(len(self.first), queries, 1)
        cols = [[v[k] for v in self.first] for k in range(4)]
        _, hw, cw = self.first_tree.decommit({self.L: pos}, cols)
        layers.append(([self.first[p] for p in wit], hw, cw, self.first_tree.root()))
        lq = sorted(set(p >> 1 for p in queries))
        for vals, tree, step in self.inner:
            pos, wit = decommit_positions(len(vals), lq, step)
            lg = len(vals).bit_length() - 1
            _, hw, cw = tree.decommit({lg: pos}, [[v[k] for v in vals] for k in range(4)])
            layers.append(([vals[p] for p in wit], hw, cw, tree.root()))
            lq = sorted(set(p >> step for p in lq))
        return queries, layers


# ---------------------------------------------------------------- AIRs ---------------------------
def uniform_m31(seed, idx):
    """This library's seeded input stream: SplitMix64 outputs (seed + (4 idx + a + 1) * golden) for
    a = 0..3, the first whose top 31 bits are below P."""
    M = (1 << 64) - 1
    for a in range(4):
        z = (seed + (idx * 4 + a + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        v = z >> 33
        if v < P:
            return v
    return 0


def from_coords(v4):
    """SecureField from 4 coordinate values: v0 + v1 i + v2 u + v3 iu (QM31 values allowed)."""
    return q_add(q_add(v4[0], q_mul(v4[1], (0, 1, 0, 0))), q_add(q_mul(v4[2], (0, 0, 1, 0)), q_mul(v4[3], (0, 0, 0, 1))))


class WideFib:
    """wide_fibonacci/mod.rs:34-83: columns col(i) = col(i-2)^2 + col(i-1)^2, one constraint per
    column from the third on. mode 0: a = 1, b = row index (the reference example's inputs);
    mode 1: seeded uniform a, b."""

    def __init__(self, log_n, n_cols, mode=0, seed=0):
        self.log_n, self.n_cols, self.mode, self.seed = log_n, n_cols, mode, seed
        self.eval_log = log_n + 1
        self.n_constraints = n_cols - 2

    def trace(self):
        cols = [[0] * (1 << self.log_n) for _ in range(self.n_cols)]
        for r in range(1 << self.log_n):
            if self.mode == 0:
                a, b = 1, r % P
            else:
                a, b = uniform_m31(self.seed, 2 * r), uniform_m31(self.seed, 2 * r + 1)
            cols[0][r], cols[1][r] = a, b
            for c in range(2, self.n_cols):
                x = (a * a + b * b) % P
                cols[c][r] = x
                a, b = b, x
        return cols

    def interaction(self, channel):
        return None

    def mask_offsets(self, tree):
        return [[0]] * self.n_cols if tree == 1 else []

    def constraints(self, mask):
        """mask(tree, col, offset) -> QM31 value; the constraint values in evaluation order."""
        v = [mask(1, c, 0) for c in range(self.n_cols)]
        out = []
        for i in range(2, self.n_cols):
            a, b = v[i - 2], v[i - 1]
            out.append(q_sub(v[i], q_add(q_mul(a, a), q_mul(b, b))))
        return out


# Poseidon2 over 16 M31 words: 4 + 4 full rounds and 14 partial rounds, x^5, every round constant
# 1234 (crates/examples/src/poseidon/mod.rs:38-210).
PSD_STATE, PSD_PARTIAL, PSD_HALF_FULL, PSD_CONST = 16, 14, 4, 1234
PSD_COLS = PSD_STATE * (1 + 2 * PSD_HALF_FULL) + PSD_PARTIAL  # 158


def psd_m4(x, add):
    t0 = add(x[0], x[1])
    t02 = add(t0, t0)
    t1 = add(x[2], x[3])
    t12 = add(t1, t1)
    t2 = add(add(x[1], x[1]), t1)
    t3 = add(add(x[3], x[3]), t0)
    t4 = add(add(t12, t12), t3)
    t5 = add(add(t02, t02), t2)
    return [add(t3, t5), t5, add(t2, t4), t4]


def psd_external(s, add):
    """circ(2 M4, M4, M4, M4): M4 on each block of 4, then every word plus its column's sum."""
    s = sum((psd_m4(s[4 * i:4 * i + 4], add) for i in range(4)), [])
    for j in range(4):
        t = add(add(s[j], s[j + 4]), add(s[j + 8], s[j + 12]))
        for i in range(4):
            s[4 * i + j] = add(s[4 * i + j], t)
    return s


def psd_internal(s, add, mul_m):
    """s_i <- s_i * 2^(i+1) + sum(s)."""
    tot = s[0]
    for x in s[1:]:
        tot = add(tot, x)
    return [add(mul_m(x, 1 << (i + 1)), tot) for i, x in enumerate(s)]


def psd_ops(field):
    if field == "m31":
        return (lambda a, b: (a + b) % P), (lambda a, m: a * m % P), (lambda a: pow(a, 5, P)), (lambda a: (a + PSD_CONST) % P)
    return q_add, q_mul_m, (lambda a: q_mul(q_mul(q_mul(a, a), q_mul(a, a)), a)), (lambda a: q_add(a, q_from(PSD_CONST)))


def psd_permute_trace(state):
    """Trace generation of one instance: the 158 column values (poseidon/mod.rs:216-303)."""
    add, mul_m, pow5, addc = psd_ops("m31")
    cols = list(state)
    s = list(state)
    for _ in range(PSD_HALF_FULL):
        s = psd_external([addc(x) for x in s], add)
        s = [pow5(x) for x in s]
        cols += s
    for _ in range(PSD_PARTIAL):
        s[0] = addc(s[0])
        s = psd_internal(s, add, mul_m)
        s[0] = pow5(s[0])
        cols.append(s[0])
    for _ in range(PSD_HALF_FULL):
        s = psd_external([addc(x) for x in s], add)
        s = [pow5(x) for x in s]
        cols += s
    return cols, s


def logup_combine(z, alpha_pows, values):
    """LookupElements::combine (crates/constraint-framework/src/logup.rs:85-100): sum alpha^i v_i - z."""
    acc = QZERO
    for v, a in zip(values, alpha_pows):
        acc = q_add(acc, q_mul(a, v if isinstance(v, tuple) else q_from(v)))
    return q_sub(acc, z)


def poseidon_constraints(ipr, trace_at, inter_at, z, alpha_pows, shift):
    """eval_poseidon_constraints (poseidon/mod.rs:150-210) with finalize_logup_in_pairs
    (crates/constraint-framework/src/lib.rs:188-258), in QM31. trace_at(c): value of trace column c;
    inter_at(c, offset): value of interaction column c at the point shifted by offset trace steps.
    Returns the 143 * ipr constraint values in evaluation order: each instance's 142 round
    constraints, then one LogUp constraint per instance."""
    add, mul_m, pow5, addc = psd_ops("qm31")
    out, fracs = [], []
    for rep in range(ipr):
        base = rep * PSD_COLS
        k = base
        s = [trace_at(k + i) for i in range(16)]
        k += 16
        init = list(s)
        for half in range(2):
            if half == 1:
                for _ in range(PSD_PARTIAL):
                    s[0] = addc(s[0])
                    s = psd_internal(s, add, mul_m)
                    s[0] = pow5(s[0])
                    m = trace_at(k)
                    k += 1
                    out.append(q_sub(s[0], m))
                    s[0] = m
            for _ in range(PSD_HALF_FULL):
                s = psd_external([addc(x) for x in s], add)
                s = [pow5(x) for x in s]
                for i in range(16):
                    m = trace_at(k)
                    k += 1
                    out.append(q_sub(s[i], m))
                    s[i] = m
        assert k == base + PSD_COLS
        fracs.append((QONE, logup_combine(z, alpha_pows, init)))
        fracs.append((q_neg(QONE), logup_combine(z, alpha_pows, s)))
    prev_col = QZERO
    for b in range(ipr):
        (n0, d0), (n1, d1) = fracs[2 * b], fracs[2 * b + 1]
        num = q_add(q_mul(n0, d1), q_mul(n1, d0))
        den = q_mul(d0, d1)
        cur = from_coords([inter_at(4 * b + i, 0) for i in range(4)])
        if b + 1 < ipr:
            diff = q_sub(cur, prev_col)
            prev_col = cur
        else:
            prev_row = from_coords([inter_at(4 * b + i, -1) for i in range(4)])
            diff = q_add(q_sub(q_sub(cur, prev_row), prev_col), shift)
        out.append(q_sub(q_mul(diff, den), num))
    return out


class Poseidon:
    """The Poseidon2 AIR with its LogUp relation (poseidon/mod.rs:38-419), ipr instances per row.
    mode 0: initial state word i of instance rep at row r is r + i + rep (the reference's gen_trace);
    mode 1: seeded uniform."""

    def __init__(self, log_n, ipr, mode=0, seed=0):
        self.log_n, self.ipr, self.mode, self.seed = log_n, ipr, mode, seed
        self.eval_log = log_n + 2
        self.n_constraints = 143 * ipr
        self.init, self.final = None, None

    def trace(self):
        N = 1 << self.log_n
        cols = [[0] * N for _ in range(PSD_COLS * self.ipr)]
        self.init = [[None] * N for _ in range(self.ipr)]
        self.final = [[None] * N for _ in range(self.ipr)]
        for r in range(N):
            for rep in range(self.ipr):
                if self.mode == 0:
                    st = [(r + i + rep) % P for i in range(16)]
                else:
                    st = [uniform_m31(self.seed, (r * self.ipr + rep) * 16 + i) for i in range(16)]
                vals, fin = psd_permute_trace(st)
                for k, v in enumerate(vals):
                    cols[rep * PSD_COLS + k][r] = v
                self.init[rep][r], self.final[rep][r] = st, fin
        return cols

    def interaction(self, channel):
        """PoseidonElements::draw, then gen_interaction_trace (poseidon/mod.rs:305-335) with
        LogupTraceGenerator (crates/constraint-framework/src/prover/logup.rs): column rep holds the
        running sum over instances 0..rep of 1/combine(init) - 1/combine(final); the last one is
        shifted by claimed_sum / 2^n and summed over the rows in the order of the trace coset."""
        self.z, alpha = channel.draw_secure_felts(2)
        self.alpha_pows = [q_pow(alpha, i) for i in range(16)]
        N = 1 << self.log_n
        run = [QZERO] * N
        cols = []
        for rep in range(self.ipr):
            for r in range(N):
                d0 = logup_combine(self.z, self.alpha_pows, self.init[rep][r])
                d1 = logup_combine(self.z, self.alpha_pows, self.final[rep][r])
                run[r] = q_add(run[r], q_mul(q_sub(d1, d0), q_inv(q_mul(d0, d1))))
            cols.append(list(run))
        last = cols[-1]
        claimed = QZERO
        for v in last:
            claimed = q_add(claimed, v)
        self.claimed_sum = claimed
        self.shift = q_mul_m(claimed, m_inv(N))
        shifted = [q_sub(v, self.shift) for v in last]
        where = {p: r for r, p in enumerate(domain_bitrev(self.log_n))}
        coset = canonic_coset(self.log_n)
        acc, summed = QZERO, [None] * N
        for k in range(N):
            r = where[coset.at(k)]
            acc = q_add(acc, shifted[r])
            summed[r] = acc
        cols[-1] = summed
        out = []
        for c in cols:
            out += [[v[k] for v in c] for k in range(4)]
        return out

    def mask_offsets(self, tree):
        if tree == 1:
            return [[0]] * (PSD_COLS * self.ipr)
        if tree == 2:
            return [[0]] * (4 * (self.ipr - 1)) + [[-1, 0]] * 4
        return []

    def constraints(self, mask):
        return poseidon_constraints(self.ipr, lambda c: mask(1, c, 0), lambda c, o: mask(2, c, o), self.z,
                                    self.alpha_pows, self.shift)


# ---------------------------------------------------------------- proof encoding ----------------
class Writer:
    def __init__(self):
        self.b = bytearray()

    def u32(self, v):
        self.b += le32(v)

    def u64(self, v):
        self.b += v.to_bytes(8, "little")

    def hash(self, h):
        # Blake2s: 32 bytes; felt252: its 32 little-endian bytes
        self.b += h.to_bytes(32, "little") if isinstance(h, int) else h

    def qm(self, v):
        for c in v:
            self.u32(c)

    def dec(self, hw, cw):
        self.u32(len(hw))
        for h in hw:
            self.hash(h)
        self.u32(len(cw))
        for v in cw:
            self.u32(v)

    def layer(self, lay):
        wit, hw, cw, root = lay
        self.u32(len(wit))
        for v in wit:
            self.qm(v)
        self.dec(hw, cw)
        self.hash(root)


# ---------------------------------------------------------------- prove ---------------------------
def prove(air, cfg, mc=Blake2sMerkleChannel):
    """prove() over one component with an empty preprocessed tree, as the reference examples build
    it (crates/examples/src/wide_fibonacci/mod.rs:186-261, poseidon/mod.rs:337-419), with
    lifting_log_size None. cfg: pow_bits, log_blowup, log_last_layer, n_queries, fold_step.
    mc: the MerkleChannel (Blake2sMerkleChannel or Poseidon252MerkleChannel).
    Returns (proof bytes, final channel digest, composition matches at the OODS point)."""
    blow = cfg["log_blowup"]
    n = air.log_n
    ch = mc.Channel()
    t0 = MerkleTree([], mc.hash_node)
    ch.mix_root(t0.root())
    trees = [(t0, [], [])]  # (tree, evaluations, coefficients)

    def commit(cols, log):
        coeffs = [interpolate(c, log) for c in cols]
        evals = [evaluate(c, log + blow) for c in coeffs]
        t = MerkleTree(evals, mc.hash_node)
        ch.mix_root(t.root())
        trees.append((t, evals, coeffs))

    commit(air.trace(), n)
    inter = air.interaction(ch)
    if inter is not None:
        commit(inter, n)

    # composition on CanonicCoset(E): sum_k alpha^(N-1-k) C_k(p) / V(p), V the vanishing polynomial
    # of the trace coset; mask offsets shift the point by the trace step
    alpha = ch.draw_secure_felt()
    E, nc = air.eval_log, air.n_constraints
    pows = [q_pow(alpha, nc - 1 - k) for k in range(nc)]
    tcoset = canonic_coset(n)
    tstep = index_point(1 << (31 - n))
    comp_vals = []
    for pt in domain_bitrev(E):
        shifted = {0: pt, -1: mp_add(pt, mp_conj(tstep))}
        mask = lambda t, c, o: q_from(eval_m(trees[t][2][c], shifted[o]))
        acc = QZERO
        for w, cv in zip(pows, air.constraints(mask)):
            acc = q_add(acc, q_mul(w, cv))
        comp_vals.append(q_mul(acc, q_inv(coset_vanishing(tcoset, qp_from(pt)))))
    comp_coeffs = [interpolate([v[k] for v in comp_vals], E) for k in range(4)]
    half = 1 << (E - 1)
    comp_polys = [c[:half] for c in comp_coeffs] + [c[half:] for c in comp_coeffs]
    commit_cols = [evaluate(c, E - 1 + blow) for c in comp_polys]
    t_comp = MerkleTree(commit_cols, mc.hash_node)
    ch.mix_root(t_comp.root())
    trees.append((t_comp, commit_cols, comp_polys))

    # OODS: mask points oods + offset * step of CanonicCoset(lifting - blowup); a column committed at
    # log size l is sampled at its point doubled (lifting - l) times (prover/pcs/mod.rs:146-300)
    oods = get_random_point(ch)
    lifting = E - 1 + blow
    max_deg = lifting - blow
    dstep = index_point(1 << (31 - max_deg))
    sample_pts = [[]]
    for t in range(1, len(trees) - 1):
        sample_pts.append([[qp_add(oods, qp_from(mp_mul(dstep, o) if o >= 0 else mp_mul(mp_conj(dstep), -o)))
                            for o in offs] for offs in air.mask_offsets(t)])
    sample_pts.append([[oods]] * 8)
    sampled = []
    for t, (tree, evals, coeffs) in enumerate(trees):
        tv = []
        for c, pts in enumerate(sample_pts[t]):
            r = lifting - (len(evals[c]).bit_length() - 1)
            tv.append([eval_q(coeffs[c], qp_repeated_double(p, r)) for p in pts])
        sampled.append(tv)
    # the split composition recombines to the composition: f = f_L + pi^(E-2)(x) f_R
    full = [eval_q(c, oods) for c in comp_coeffs]
    piE = qp_repeated_double(oods, E - 2)[0]
    left = from_coords([v[0] for v in sampled[-1][:4]])
    right = from_coords([v[0] for v in sampled[-1][4:]])
    assert from_coords(full) == q_add(left, q_mul(piE, right))
    # prove_ex's sanity check: the composition at the OODS point from the committed halves against
    # the constraints evaluated on the sampled mask values (point evaluation accumulator:
    # acc = acc * alpha + C_k / V(oods), with V of CanonicCoset(max_deg))
    mvals = {}
    for t in range(1, len(trees) - 1):
        for c, offs in enumerate(air.mask_offsets(t)):
            for o, v in zip(offs, sampled[t][c]):
                mvals[(t, c, o)] = v
    vinv = q_inv(coset_vanishing(canonic_coset(max_deg), oods))
    cacc = QZERO
    for cv in air.constraints(lambda t, c, o: mvals[(t, c, o)]):
        cacc = q_add(q_mul(cacc, alpha), q_mul(vinv, cv))
    satisfied = cacc == q_add(left, q_mul(qp_repeated_double(oods, max_deg - 1)[0], right))
    ch.mix_felts([v for tv in sampled for col in tv for v in col])

    # FRI quotients on the lifting domain: one sum over every (column, sample) with its random
    # power, in tree / column / sample order, a column with two samples getting first a
    # periodicity sample (core/pcs/quotients.rs:306-345, 200-230; core/constraints.rs:119-134)
    qrc = ch.draw_secure_felt()
    lgen = index_point(1 << (31 - lifting))
    terms = []
    rp = QONE
    for t, (tree, evals, coeffs) in enumerate(trees):
        for c, pts in enumerate(sample_pts[t]):
            if not pts:
                continue
            col_log = len(evals[c]).bit_length() - 1
            samples = list(zip(pts, sampled[t][c]))
            if len(samples) == 2:
                period = mp_mul(lgen, 1 << col_log)
                samples = [(qp_add(samples[1][0], qp_from(period)), samples[1][1])] + samples
            for spt, val in samples:
                a = q_sub(q_conj(val), val)
                cc = q_sub(q_conj(spt[1]), spt[1])
                b = q_sub(q_mul(val, cc), q_mul(a, spt[1]))
                terms.append((coeffs[c], lifting - col_log, spt, q_mul(rp, a), q_mul(rp, b), q_mul(rp, cc)))
                rp = q_mul(rp, qrc)
    quot = []
    for dp in domain_bitrev(lifting):
        acc = QZERO
        for coeffs, r, spt, a, b, c in terms:
            # the column lifted to the lifting domain: its polynomial at the point doubled r times
            dr = dp
            for _ in range(r):
                dr = mp_add(dr, dr)
            f = eval_m(coeffs, dr)
            num = q_sub(q_mul_m(c, f), q_add(q_mul_m(a, dp[1]), b))
            prx, pix = (spt[0][0], spt[0][1]), (spt[0][2], spt[0][3])
            pry, piy = (spt[1][0], spt[1][1]), (spt[1][2], spt[1][3])
            den = c_sub(c_mul(c_sub(prx, (dp[0], 0)), piy), c_mul(c_sub(pry, (dp[1], 0)), pix))
            di = c_inv(den)
            acc = q_add(acc, q_mul(num, (di[0], di[1], 0, 0)))
        quot.append(acc)

    fri = FriProver(ch, cfg, quot, lifting, mc.hash_node)
    nonce = mc.grind(ch, cfg["pow_bits"])
    ch.mix_u64(nonce)
    queries, fri_layers = fri.decommit(ch)

    # every tree is opened at tree-local positions (prepare_preprocessed_query_positions,
    # core/pcs/utils.rs:200-218), the identity for trees of the lifting size
    queried, decs = [], []
    for tree, evals, _ in trees:
        tmax = len(tree.layers) - 1
        if tmax == 0:
            qpos = []
        elif lifting < tmax:
            qpos = [((p >> 1) << (tmax - lifting + 1)) + (p & 1) for p in queries]
        else:
            qpos = [((p >> (lifting - tmax + 1)) << 1) + (p & 1) for p in queries]
        qpl = {}
        for ls in sorted(set(len(c).bit_length() - 1 for c in evals)):
            shift = tmax - ls
            qpl[ls] = sorted(set(((p >> (shift + 1)) << 1) + (p & 1) for p in qpos))
        qv, hw, cw = tree.decommit(qpl, evals)
        queried.append(qv)
        decs.append((hw, cw))

    w = Writer()
    for key in ("pow_bits", "log_blowup", "log_last_layer", "n_queries", "fold_step"):
        w.u32(cfg[key])
    w.u32(0)  # lifting_log_size: None
    w.u32(len(trees))
    for tree, _, _ in trees:
        w.hash(tree.root())
    w.u32(len(sampled))
    for tv in sampled:
        w.u32(len(tv))
        for col in tv:
            w.u32(len(col))
            for v in col:
                w.qm(v)
    w.u32(len(decs))
    for hw, cw in decs:
        w.dec(hw, cw)
    w.u32(len(queried))
    for qv in queried:
        w.u32(len(qv))
        for v in qv:
            w.u32(v)
    w.u64(nonce)
    w.layer(fri_layers[0])
    w.u32(len(fri_layers) - 1)
    for lay in fri_layers[1:]:
        w.layer(lay)
    w.u32(len(fri.last_poly))
    for v in fri.last_poly:
        w.qm(v)
    return bytes(w.b), ch.digest, satisfied


def prove_wide_fib(log_n, n_cols, cfg, mode=0, seed=0):
    pb, dig, ok = prove(WideFib(log_n, n_cols, mode, seed), cfg)
    assert ok, "wide-Fibonacci constraints must hold"
    return pb, dig
