"""Independent pins of the oracle: every oracle function the HIP library is checked against is
itself checked here against the pure-Python restatement in tests/pystark.py (written from the
reference's definitions and from the mathematics) or against hashlib, on random non-constant
inputs, and whole proofs are checked against committed transcripts (tests/golden/transcripts.json,
made by tests/golden/make_golden.py from the same restatement)."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import pyref
import pystark as S

P = O.P
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "transcripts.json")))


def gid(rec):
    return "%s-%d-%d-%d" % (rec["air"], rec["log_n_rows"], rec["width"], rec["config"]["log_blowup"])


def oracle_proof(rec):
    kw = dict(rec["config"])
    if rec["air"] == "wide_fib":
        return O.prove(air="wide_fib", log_n_rows=rec["log_n_rows"], n_cols=rec["width"],
                       input_mode=rec["input_mode"], seed=rec["seed"], **kw)
    return O.prove(air="poseidon", log_n_rows=rec["log_n_rows"], ipr=rec["width"], input_mode=rec["input_mode"],
                   seed=rec["seed"], **kw)


# ---------------------------------------------------------------- whole transcripts -----------------
@pytest.mark.parametrize("rec", GOLDEN, ids=gid)
def test_oracle_proof_matches_golden_transcript(rec):
    st, proof, digest, _, _ = oracle_proof(rec)
    assert len(proof) == rec["proof_len"]
    assert hashlib.sha256(proof).hexdigest() == rec["proof_sha256"]
    assert digest.hex() == rec["digest"]
    assert (st == 0) == rec["constraints_satisfied"]
    if "proof_hex" in rec:
        assert proof.hex() == rec["proof_hex"]


@pytest.mark.parametrize("rec", GOLDEN[:3] + GOLDEN[7:8], ids=gid)
def test_python_restatement_reproduces_golden_transcript(rec):
    air = (S.WideFib(rec["log_n_rows"], rec["width"], rec["input_mode"], rec["seed"]) if rec["air"] == "wide_fib"
           else S.Poseidon(rec["log_n_rows"], rec["width"], rec["input_mode"], rec["seed"]))
    proof, digest, ok = S.prove(air, rec["config"])
    assert hashlib.sha256(proof).hexdigest() == rec["proof_sha256"] and digest.hex() == rec["digest"]
    assert ok == rec["constraints_satisfied"]


def test_wide_fib_golden_proof_is_accepted():
    for rec in GOLDEN:
        if rec["air"] == "wide_fib" and "proof_hex" in rec:
            err = O.verify(bytes.fromhex(rec["proof_hex"]), air="wide_fib", log_n_rows=rec["log_n_rows"],
                           n_cols=rec["width"], **rec["config"])
            assert err == ""


def test_planted_mistakes_change_the_transcript(monkeypatch):
    """The transcript is sensitive to the conventions it pins: shifted composition weights, a fold
    with a wrong coefficient, or a different OODS point each change the proof."""
    rec = GOLDEN[0]
    cfg = rec["config"]

    def run():
        return S.prove(S.WideFib(rec["log_n_rows"], rec["width"], rec["input_mode"], rec["seed"]), cfg)[0]

    good = hashlib.sha256(run()).hexdigest()
    assert good == rec["proof_sha256"]
    orig_pow = S.q_pow
    monkeypatch.setattr(S, "q_pow", lambda a, e: orig_pow(a, e + 1))
    assert hashlib.sha256(run()).hexdigest() != good
    monkeypatch.setattr(S, "q_pow", orig_pow)
    orig_fold = S.fold_line
    monkeypatch.setattr(S, "fold_line", lambda v, xs, a: orig_fold(v, xs, S.q_add(a, S.QONE)))
    assert hashlib.sha256(run()).hexdigest() != good
    monkeypatch.setattr(S, "fold_line", orig_fold)
    orig_pt = S.get_random_point
    monkeypatch.setattr(S, "get_random_point", lambda ch: S.qp_double(orig_pt(ch)))
    assert hashlib.sha256(run()).hexdigest() != good


# ---------------------------------------------------------------- polynomials ------------------
@pytest.mark.parametrize("log_n", [1, 2, 3, 5])
def test_python_interpolation_inverts_bruteforce_evaluation(log_n):
    rng = np.random.default_rng(log_n)
    coeffs = [int(x) for x in rng.integers(0, P, 1 << log_n)]
    vals = pyref.evaluate_bruteforce(coeffs, log_n)
    assert S.evaluate(coeffs, log_n) == vals
    assert S.interpolate(vals, log_n) == coeffs


@pytest.mark.parametrize("log_n", [3, 6])
def test_oracle_fft_matches_python(log_n):
    rng = np.random.default_rng(10 + log_n)
    vals = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    coeffs = O.interpolate(vals)
    assert [int(c) for c in coeffs] == S.interpolate([int(v) for v in vals], log_n)
    ext = O.evaluate(coeffs, log_n + 2)
    assert [int(v) for v in ext] == S.evaluate([int(c) for c in coeffs], log_n + 2)


def test_coset_vanishing_vanishes_exactly_on_the_coset():
    n = 3
    coset = S.canonic_coset(n)
    on = set(coset.points())
    for p in S.circle_domain(n + 2):
        v = S.coset_vanishing(coset, S.qp_from(p))
        assert (v == S.QZERO) == (p in on)


# ---------------------------------------------------------------- FRI folds ------------------------
def rq(rng):
    return tuple(int(x) for x in rng.integers(0, P, 4))


@pytest.mark.parametrize("log_src", [3, 4, 6, 8])
def test_fold_circle_into_line_on_random_input(log_src):
    """dst' = dst alpha^2 + (f(p) + f(conj p)) + alpha (f(p) - f(conj p)) / p.y at every pair."""
    rng = np.random.default_rng(log_src)
    src = [rng.integers(0, P, 1 << log_src, dtype=np.uint32) for _ in range(4)]
    dst = [rng.integers(0, P, 1 << (log_src - 1), dtype=np.uint32) for _ in range(4)]
    alpha = rq(rng)
    got = O.fold_circle_into_line(src, dst, alpha)
    vals = [tuple(int(src[k][i]) for k in range(4)) for i in range(1 << log_src)]
    folded, _ = S.fold_circle_into_line(vals, S.domain_bitrev(log_src), alpha)
    a2 = S.q_mul(alpha, alpha)
    for i, f in enumerate(folded):
        d = tuple(int(dst[k][i]) for k in range(4))
        assert tuple(int(got[k][i]) for k in range(4)) == S.q_add(S.q_mul(d, a2), f)


@pytest.mark.parametrize("log_src", [1, 2, 5, 7])
def test_fold_line_on_random_input(log_src):
    """f' = (f(x) + f(-x)) + alpha (f(x) - f(-x)) / x over the line domain half_odds(log_src)."""
    rng = np.random.default_rng(20 + log_src)
    src = [rng.integers(0, P, 1 << log_src, dtype=np.uint32) for _ in range(4)]
    alpha = rq(rng)
    got = O.fold_line(src, log_src, alpha)
    dom = S.half_odds(log_src)
    xs = [dom.at(S.bitrev(i, log_src))[0] for i in range(1 << log_src)]
    vals = [tuple(int(src[k][i]) for k in range(4)) for i in range(1 << log_src)]
    folded, _ = S.fold_line(vals, xs, alpha)
    for i, f in enumerate(folded):
        assert tuple(int(got[k][i]) for k in range(4)) == f


# ---------------------------------------------------------------- channel / proof of work ---------
@pytest.mark.parametrize("seed", range(4))
def test_random_point_matches_python(seed):
    dig = hashlib.blake2s(b"point" + bytes([seed]), digest_size=32).digest()
    pt = np.zeros(8, np.uint32)
    O.lib().oracle_random_point(O.u8ptr(np.frombuffer(dig, np.uint8).copy()), O.u32ptr(pt))
    ch = S.Channel()
    ch.digest = dig
    x, y = S.get_random_point(ch)
    assert tuple(int(v) for v in pt) == x + y
    # a point of the circle: x^2 + y^2 = 1
    assert S.q_add(S.q_mul(x, x), S.q_mul(y, y)) == S.QONE


@pytest.mark.parametrize("bits", [1, 4, 8, 11])
def test_grind_prefix_and_smallest_nonce_with_hashlib(bits):
    """H(H(0x12345678 LE || 12 zero bytes || digest || bits LE) || nonce LE): the smallest nonce whose
    hash has at least `bits` trailing zero bits (core/channel/blake2s.rs:124-140), by hashlib."""
    digest = hashlib.blake2s(b"grind-pin" + bytes([bits]), digest_size=32).digest()
    prefix = hashlib.blake2s((0x12345678).to_bytes(4, "little") + bytes(12) + digest + bits.to_bytes(4, "little"),
                             digest_size=32).digest()
    n = 0
    while True:
        h = int.from_bytes(hashlib.blake2s(prefix + n.to_bytes(8, "little"), digest_size=32).digest()[:16], "little")
        if h & ((1 << bits) - 1) == 0:
            break
        n += 1
    assert O.grind(digest, bits) == n


# ---------------------------------------------------------------- quotients ------------------------
def random_circle_point(rng):
    ch = S.Channel()
    ch.digest = bytes(int(x) for x in rng.integers(0, 256, 32))
    return S.get_random_point(ch)


def quotient_term(spt, val, rc, f, dp):
    """rc (c f - a y - b) / den for one sample (complex_conjugate_line_coeffs, denominator_inverses)."""
    a = S.q_sub(S.q_conj(val), val)
    c = S.q_sub(S.q_conj(spt[1]), spt[1])
    b = S.q_sub(S.q_mul(val, c), S.q_mul(a, spt[1]))
    num = S.q_mul(rc, S.q_sub(S.q_mul_m(c, f), S.q_add(S.q_mul_m(a, dp[1]), b)))
    prx, pix, pry, piy = (spt[0][0], spt[0][1]), (spt[0][2], spt[0][3]), (spt[1][0], spt[1][1]), (spt[1][2], spt[1][3])
    den = S.c_sub(S.c_mul(S.c_sub(prx, (dp[0], 0)), piy), S.c_mul(S.c_sub(pry, (dp[1], 0)), pix))
    di = S.c_inv(den)
    return S.q_mul(num, (di[0], di[1], 0, 0))


def test_quotient_pipeline_matches_direct_sum():
    """accumulate_numerators per batch, then compute_quotients_and_combine with partials of several
    sizes lifted to the lifting domain, equals the direct sum over every (column, sample) of
    rc (c f~(q) - a q.y - b) / den(q), where f~(q) is the column's value at q doubled
    (lifting - log) times, found by its point in the smaller domain."""
    rng = np.random.default_rng(42)
    L = 5
    groups = [(5, 3), (3, 2)]  # (log size, columns)
    pts = [random_circle_point(rng) for _ in range(2)]
    accs, direct_terms = [], []
    for ls, nc in groups:
        cols = [rng.integers(0, P, 1 << ls, dtype=np.uint32) for _ in range(nc)]
        batches = []
        for pt in pts:
            ents = [(c, rq(rng), rq(rng)) for c in range(nc)]
            batches.append((list(pt[0] + pt[1]), ents))
            for c, v, r in ents:
                direct_terms.append((ls, cols[c], pt, v, r))
        parts, fls = O.accumulate_numerators(cols, batches)
        for (pt8, _), part, fl in zip(batches, parts, fls):
            accs.append((pt8, part, fl))
    got = O.compute_quotients_and_combine(accs, L)
    for row, dp in enumerate(S.domain_bitrev(L)):
        acc = S.QZERO
        for ls, col, pt, v, r in direct_terms:
            dr = dp
            for _ in range(L - ls):
                dr = S.mp_add(dr, dr)
            f = int(col[S.domain_bitrev(ls).index(dr)])
            acc = S.q_add(acc, quotient_term(pt, v, r, f, dp))
        assert tuple(int(got[k][row]) for k in range(4)) == acc


# ---------------------------------------------------------------- Poseidon AIR ---------------------
@pytest.mark.parametrize("ipr", [1, 2])
def test_poseidon_constraints_match_python_on_random_columns(ipr):
    """Every round constraint and the LogUp constraint of each instance, on random (non-trace) column
    values, with the previous row of the last interaction columns found as the point shifted by
    minus one trace step."""
    rng = np.random.default_rng(7 + ipr)
    tl, el = 2, 4
    trace = [rng.integers(0, P, 1 << el, dtype=np.uint32) for _ in range(158 * ipr)]
    inter = [rng.integers(0, P, 1 << el, dtype=np.uint32) for _ in range(4 * ipr)]
    pows = [rq(rng) for _ in range(143 * ipr)]
    dinv = rng.integers(1, P, 1 << (el - tl), dtype=np.uint32)
    acc = [np.zeros(1 << el, np.uint32) for _ in range(4)]
    z, alpha, shift = rq(rng), rq(rng), rq(rng)
    got = O.eval_constraints_on_domain("poseidon", ipr, tl, el, trace, inter, pows, dinv, acc, z=z, alpha=alpha,
                                       shift=shift)
    apow = [S.q_pow(alpha, i) for i in range(16)]
    pts = S.domain_bitrev(el)
    where = {p: r for r, p in enumerate(pts)}
    step = S.index_point(1 << (31 - tl))
    for row, p in enumerate(pts):
        prev = where[S.mp_add(p, S.mp_conj(step))]
        cons = S.poseidon_constraints(ipr, lambda c: S.q_from(int(trace[c][row])),
                                      lambda c, o: S.q_from(int(inter[c][row if o == 0 else prev])), z, apow, shift)
        tot = S.QZERO
        for w, cv in zip(pows, cons):
            tot = S.q_add(tot, S.q_mul(w, cv))
        tot = S.q_mul_m(tot, int(dinv[row >> tl]))
        assert tuple(int(got[k][row]) for k in range(4)) == tot


def test_poseidon_trace_matches_python():
    """Trace generation (both input modes) against the Python permutation."""
    for mode in (0, 1):
        cols = O.gen_poseidon(3, 2, mode, 77)
        air = S.Poseidon(3, 2, mode, 77)
        exp = air.trace()
        assert [[int(v) for v in c] for c in cols] == exp


def test_poseidon_trace_satisfies_the_round_constraints():
    """On a genuine trace row, the 142 round constraints of each instance vanish."""
    air = S.Poseidon(2, 1, 1, 5)
    cols = air.trace()
    for r in range(4):
        cons = S.poseidon_constraints(1, lambda c: S.q_from(cols[c][r]), lambda c, o: S.QZERO, S.QONE,
                                      [S.QONE] * 16, S.QZERO)
        assert all(c == S.QZERO for c in cons[:142])
