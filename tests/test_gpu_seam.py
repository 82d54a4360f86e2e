"""GPU parity of the rest of the backend-trait seam: ColumnOps::bit_reverse_column, Column::at /
set / split_at_mid, PolyOps::extend / split_at_mid / precompute_twiddles for any coset /
interpolate and evaluate on any circle domain (including domains of 2 and 4 points) /
eval_at_point_by_folding, and fold_line with fold steps above 1.

The expected values come from the mathematics (brute-force FFT-basis evaluation at the domain's
points, tests/pyref.py), from the oracle, or from identities the reference states (p = p_left +
pi^(L-2)(x) p_right for split_at_mid; folding evaluates the interpolant at the point)."""
import hashlib
import sys

import numpy as np
import pytest

import oracle_lib as O
import pyref
import pystark as S
from hip_lib import SH

P = O.P
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hb():
    b = SH.HipBackend(0)
    yield b
    b.close()


def random_point(tag):
    dig = np.frombuffer(hashlib.blake2s(tag, digest_size=32).digest(), np.uint8).copy()
    pt = np.zeros(8, np.uint32)
    O.lib().oracle_random_point(O.u8ptr(dig), O.u32ptr(pt))
    return [int(x) for x in pt]


def domain_point(half_init, log_n, i):
    """Point i (natural order) of CircleDomain::new(Coset::new(half_init, log_n - 1)): the half
    coset, then its conjugate (core/poly/circle/domain.rs:57-68)."""
    h = log_n - 1
    step = 1 << (31 - h)
    if i < (1 << h):
        return pyref.index_point(half_init + step * i)
    return pyref.index_point(-(half_init + step * (i - (1 << h))))


def evaluate_on(coeffs, half_init, log_n):
    L = len(coeffs).bit_length() - 1
    out = []
    for r in range(1 << log_n):
        x, y = domain_point(half_init, log_n, pyref.bitrev(r, log_n))
        out.append(sum(int(c) * pyref.basis_eval(k, L, x, y) for k, c in enumerate(coeffs)) % P)
    return np.array(out, np.uint32)


def coset_tree(init, log_size):
    """slow_precompute_twiddles (prover/backend/cpu/circle.rs:274-291) by brute force."""
    tw = []
    for j in range(log_size):
        lj = log_size - j
        ini, step = (init << j) % (1 << 31), (1 << (31 - log_size + j)) % (1 << 31)
        xs = [pyref.index_point(ini + step * k)[0] for k in range(1 << (lj - 1))]
        tw += [xs[pyref.bitrev(i, lj - 1)] for i in range(len(xs))]
    tw.append(1)
    return np.array(tw, np.uint32)


# ---------------------------------------------------------------- ColumnOps / Column -----------
@pytest.mark.parametrize("log_n", [0, 1, 5, 12, 20])
def test_bit_reverse_column(hb, log_n):
    say("bit_reverse_column", log_n)
    rng = np.random.default_rng(log_n)
    v = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    d = hb.from_cpu(v)
    hb.bit_reverse_column(d, log_n)
    idx = np.array([pyref.bitrev(i, log_n) for i in range(1 << log_n)]) if log_n <= 12 else None
    got = d.to_cpu()
    if idx is not None:
        assert np.array_equal(got, v[idx])
    hb.bit_reverse_column(d, log_n)  # an involution
    assert np.array_equal(d.to_cpu(), v)
    if idx is None:
        for i in (0, 1, 12345, (1 << log_n) - 2):
            assert got[i] == v[pyref.bitrev(i, log_n)]


def test_column_at_set_and_split(hb):
    rng = np.random.default_rng(4)
    v = rng.integers(0, P, 64, dtype=np.uint32)
    d = hb.from_cpu(v)
    assert [d.at(i) for i in (0, 7, 63)] == [int(v[0]), int(v[7]), int(v[63])]
    d.set(7, 12345)
    v[7] = 12345
    assert np.array_equal(d.to_cpu(), v)
    lft, rgt = d.split_at_mid()
    assert np.array_equal(lft.to_cpu(), v[:32]) and np.array_equal(rgt.to_cpu(), v[32:])
    rgt.set(0, 9)
    assert d.at(32) == 9
    a, b = rgt.split_at_mid()
    assert b.at(0) == v[48]
    del lft, rgt, a, b
    assert d.at(0) == v[0]


# ---------------------------------------------------------------- PolyOps ------------------------
@pytest.mark.parametrize("log_n,new_log", [(3, 3), (3, 5), (10, 12)])
def test_extend(hb, log_n, new_log):
    rng = np.random.default_rng(log_n + new_log)
    c = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    e = hb.extend(hb.from_cpu(c), log_n, new_log).to_cpu()
    assert np.array_equal(e[: 1 << log_n], c) and not e[1 << log_n:].any()
    # evaluating the extension on CanonicCoset(new_log + 1) equals evaluating the original there
    ev1 = hb.evaluate_polynomials([hb.from_cpu(e)], new_log, 1)[0].to_cpu()
    ev2 = hb.evaluate_polynomials([hb.from_cpu(c)], log_n, new_log + 1 - log_n)[0].to_cpu()
    assert np.array_equal(ev1, ev2)


@pytest.mark.parametrize("log_n", [2, 4, 9])
def test_split_at_mid_identity(hb, log_n):
    """p(z) = p_left(z) + pi^(L-2)(z.x) p_right(z) (prover/poly/circle/ops.rs:127-141)."""
    rng = np.random.default_rng(70 + log_n)
    c = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    d = hb.from_cpu(c)
    lft, rgt = hb.split_at_mid(d)
    pt = random_point(b"split%d" % log_n)
    full = hb.batch_eval_at_point([d], log_n, pt)[0]
    pl, pr = hb.batch_eval_at_point([lft, rgt], log_n - 1, pt)
    x = tuple(pt[:4])
    for _ in range(log_n - 2):
        x = S.q_double_x(x)
    assert tuple(full) == S.q_add(tuple(pl), S.q_mul(x, tuple(pr)))


# odd initial indices: no point of such a coset or its doublings (log size below 30) has x = 0
@pytest.mark.parametrize("init,log_size", [(1 << 27 | 1, 4), (12345, 6), (1 << 20 | 3, 9), ((1 << 31) - 7, 3)])
def test_twiddles_of_any_coset(hb, init, log_size):
    say("twiddles coset", init, log_size)
    t = hb.twiddles(init, log_size)
    tw, itw = t.read()
    exp = coset_tree(init, log_size)
    assert np.array_equal(tw, exp)
    assert all(int(a) * int(b) % P == 1 for a, b in zip(tw, itw))
    t.free()


def test_canonic_twiddles_built_on_device_match_oracle_large(hb):
    tw, itw = hb.precompute_twiddles(20)
    otw, oitw = O.twiddles(20)
    assert np.array_equal(tw, otw) and np.array_equal(itw, oitw)


def test_twiddles_reject_a_zero_coordinate(hb):
    # the coset through (0, 1) (index 2^29) has a twiddle x = 0, which has no inverse
    with pytest.raises(SH.StwoHipError) as e:
        hb.twiddles(1 << 29, 3)
    assert e.value.status == 1


@pytest.mark.parametrize("tree_init,tree_log,doublings,log_coeffs", [
    (12345, 6, 0, 7), (12345, 6, 2, 3), (1 << 20 | 3, 9, 0, 8), (1 << 20 | 3, 9, 1, 9), (777, 4, 3, 2),
    (777, 4, 4, 1)])
def test_evaluate_and_interpolate_on_any_domain(hb, tree_init, tree_log, doublings, log_coeffs):
    """Evaluation on CircleDomain(half coset = the tree's coset doubled `doublings` times) against
    brute-force FFT-basis evaluation at the domain's points; interpolation inverts it."""
    half_log = tree_log - doublings
    log_n = half_log + 1
    half_init = (tree_init << doublings) % (1 << 31)
    say("on domain", tree_init, tree_log, doublings, log_coeffs, "->", log_n)
    t = hb.twiddles(tree_init, tree_log)
    rng = np.random.default_rng(tree_init % 1000 + doublings)
    c = rng.integers(0, P, 1 << log_coeffs, dtype=np.uint32)
    ev = hb.evaluate_on_domain(t, half_init, log_n, [hb.from_cpu(c)], log_coeffs)[0].to_cpu()
    assert np.array_equal(ev, evaluate_on(c, half_init, log_n))
    back = hb.interpolate_on_domain(t, half_init, log_n, [hb.from_cpu(ev)])[0].to_cpu()
    assert np.array_equal(back[: 1 << log_coeffs], c) and not back[1 << log_coeffs:].any()
    t.free()


def test_domain_must_be_a_doubling_of_the_tree(hb):
    t = hb.twiddles(12345, 6)
    with pytest.raises(SH.StwoHipError):
        hb.evaluate_on_domain(t, 12346, 7, [hb.column(8, zeroed=True)], 3)
    with pytest.raises(SH.StwoHipError):
        hb.evaluate_on_domain(t, 12345, 8, [hb.column(8, zeroed=True)], 3)
    t.free()


@pytest.mark.parametrize("log_n", [1, 2])
def test_canonic_transforms_of_two_and_four_points(hb, log_n):
    rng = np.random.default_rng(log_n)
    c = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    ev = hb.evaluate_polynomials([hb.from_cpu(c)], log_n, 0)[0].to_cpu()
    assert list(ev) == pyref.evaluate_bruteforce([int(x) for x in c], log_n)
    back = hb.interpolate_columns([hb.from_cpu(ev)], log_n)[0].to_cpu()
    assert np.array_equal(back, c)
    if log_n == 1:  # blowup from 2 to 8 points
        ev3 = hb.evaluate_polynomials([hb.from_cpu(c)], 1, 2)[0].to_cpu()
        assert list(ev3) == pyref.evaluate_bruteforce([int(x) for x in c], 3)


@pytest.mark.parametrize("log_n", [3, 6, 11])
def test_eval_at_point_by_folding(hb, log_n):
    say("eval_at_point_by_folding", log_n)
    rng = np.random.default_rng(90 + log_n)
    ev = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    pt = random_point(b"fold%d" % log_n)
    got = hb.eval_at_point_by_folding(hb.from_cpu(ev), log_n, pt)
    assert tuple(got) == tuple(map(int, O.eval_at_point(O.interpolate(ev), pt)))


# ---------------------------------------------------------------- FriOps ---------------------------
@pytest.mark.parametrize("log_src,fold_step", [(5, 2), (9, 3), (12, 4), (3, 3)])
def test_fold_line_fold_step(hb, log_src, fold_step):
    say("fold_line step", log_src, fold_step)
    rng = np.random.default_rng(log_src * 10 + fold_step)
    src = [rng.integers(0, P, 1 << log_src, dtype=np.uint32) for _ in range(4)]
    alpha = tuple(int(x) for x in rng.integers(0, P, 4))
    got = hb.fold_line([hb.from_cpu(x) for x in src], alpha, log_src, fold_step)
    exp, a = src, alpha
    for k in range(fold_step):
        exp = O.fold_line(exp, log_src - k, a)
        a = S.q_mul(a, a)
    for g, e in zip(got, exp):
        assert np.array_equal(g.to_cpu(), e)


@pytest.mark.parametrize("log_src", [4, 13])
def test_folds_on_unaligned_columns(hb, log_src):
    """Columns that are not 16-byte aligned take the one-word path of both folds."""
    rng = np.random.default_rng(500 + log_src)
    src = [rng.integers(0, P, 1 << log_src, dtype=np.uint32) for _ in range(4)]
    dst = [rng.integers(0, P, 1 << (log_src - 1), dtype=np.uint32) for _ in range(4)]
    alpha = tuple(int(x) for x in rng.integers(0, P, 4))
    big = hb.from_cpu(np.concatenate([np.zeros(1, np.uint32)] + src))
    views = [big.ptr + 4 * (1 + (k << log_src)) for k in range(4)]
    dbig = hb.from_cpu(np.concatenate([np.zeros(1, np.uint32)] + dst))
    dviews = [dbig.ptr + 4 * (1 + (k << (log_src - 1))) for k in range(4)]
    hb.fold_circle_into_line(dviews, views, alpha, log_src)
    exp = O.fold_circle_into_line(src, dst, alpha)
    out = dbig.to_cpu()[1:]
    for k in range(4):
        assert np.array_equal(out[k << (log_src - 1):(k + 1) << (log_src - 1)], exp[k])
    got = hb.fold_line(views, alpha, log_src)
    for g, e in zip(got, O.fold_line(src, log_src, alpha)):
        assert np.array_equal(g.to_cpu(), e)


def test_int32_issue_rate_is_plausible(hb):
    """The add / xor / rotate mix issues some tens of T lane-instructions per second on 256 CUs
    (about 42 T measured on MI355X; plain adds alone issue faster)."""
    v = hb.measure_int32_issue()
    say("int32 issue rate %.3g lane-instr/s" % v)
    assert 1e13 < v < 1e14


def test_butterfly_rate_is_plausible(hb):
    """About 11 VALU instructions per M31 butterfly at tens of T lane-instructions per second: a few
    T butterflies per second, forward and inverse."""
    f, i = hb.measure_butterfly_rate(False), hb.measure_butterfly_rate(True)
    say("butterflies per second: forward %.3g, inverse %.3g" % (f, i))
    assert 5e11 < f < 2e13 and 5e11 < i < 2e13


def test_kernel_units_count_butterflies(hb):
    """The transform classes count n_cols 2^(log-1) stages butterflies per launch group."""
    rng = np.random.default_rng(77)
    k, n = 10, 3
    evals = [hb.from_cpu(rng.integers(0, P, 1 << k, dtype=np.uint32)) for _ in range(n)]
    coeffs = [hb.column(1 << k) for _ in range(n)]
    ext = [hb.column(2 << k) for _ in range(n)]
    hb.precompute_twiddles(k + 1)
    hb.reset_kernel_stats()
    hb.set_kernel_timing(True)
    hb.interpolate_columns(evals, k, out=coeffs)
    hb.evaluate_polynomials(coeffs, k, 1, out=ext)
    hb.set_kernel_timing(False)
    u = hb.kernel_units()
    assert u["interpolate"] == n * (1 << (k - 1)) * k
    assert u["evaluate"] == n * (1 << k) * k
    hb.reset_kernel_stats()
    assert hb.kernel_units() == {}


@pytest.mark.parametrize("log_src,n_groups,misalign", [(3, 2, False), (10, 5, False), (12, 3, True)])
def test_grouped_fold_matches_per_group(hb, log_src, n_groups, misalign):
    say("grouped fold", log_src, n_groups, misalign)
    rng = np.random.default_rng(700 + log_src)
    srcs = [[rng.integers(0, P, 1 << log_src, dtype=np.uint32) for _ in range(4)] for _ in range(n_groups)]
    dsts = [[rng.integers(0, P, 1 << (log_src - 1), dtype=np.uint32) for _ in range(4)] for _ in range(n_groups)]
    alpha = tuple(int(x) for x in rng.integers(0, P, 4))
    if misalign:
        big = hb.from_cpu(np.concatenate([np.zeros(1, np.uint32)] + [c for g in srcs for c in g]))
        dsrc = [[big.ptr + 4 * (1 + ((4 * g + k) << log_src)) for k in range(4)] for g in range(n_groups)]
    else:
        dsrc = [[hb.from_cpu(c) for c in g] for g in srcs]
    ddst = [[hb.from_cpu(c) for c in g] for g in dsts]
    hb.fold_circle_into_line_grouped(ddst, dsrc, alpha, log_src)
    for g in range(n_groups):
        exp = O.fold_circle_into_line(srcs[g], dsts[g], alpha)
        for k in range(4):
            assert np.array_equal(ddst[g][k].to_cpu(), exp[k]), (g, k)
