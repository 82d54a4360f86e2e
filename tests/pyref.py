"""Pure-Python big-int restatement of the circle-STARK math, for pinning the oracle at tiny sizes.

Written from the mathematics (the reference's definitions are cited), independently of the
C++ oracle: modular arithmetic over P = 2^31 - 1, the circle group x^2 + y^2 = 1 with
generator (2, 1268011823) (crates/stwo/src/core/circle.rs), canonic domains
(core/poly/circle/canonic.rs, domain.rs) and the FFT basis y^b0 * x^b1 * pi(x)^b2 * ...
(prover/poly/circle/ops.rs:125-141), evaluated by brute force.
"""
P = (1 << 31) - 1
GEN = (2, 1268011823)


def add(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def mul(p, k):
    r, c = (1, 0), p
    while k:
        if k & 1:
            r = add(r, c)
        c = add(c, c)
        k >>= 1
    return r


def index_point(i):
    return mul(GEN, i % (1 << 31))


def canonic_domain_point(log_size, i):
    """CircleDomain of CanonicCoset(log_size): half coset half_odds(log_size-1) then its conjugate."""
    h = log_size - 1
    init = 1 << (31 - (h + 2))
    step = 1 << (31 - h)
    if i < (1 << h):
        return index_point(init + step * i)
    j = i - (1 << h)
    return index_point(-(init + step * j))


def bitrev(i, n):
    return int(format(i, "0%db" % n)[::-1], 2) if n else i


def pi(x):
    return (2 * x * x - 1) % P


def basis_eval(n_idx, L, x, y):
    v = 1
    if n_idx & 1:
        v = v * y % P
    xx = x
    for b in range(1, L):
        if (n_idx >> b) & 1:
            v = v * xx % P
        xx = pi(xx)
    return v


def evaluate_bruteforce(coeffs, log_domain):
    """Evaluations of FFT-basis coefficients on CanonicCoset(log_domain), bit-reversed order."""
    L = max(1, (len(coeffs) - 1).bit_length()) if len(coeffs) > 1 else 0
    L = len(coeffs).bit_length() - 1
    out = []
    for r in range(1 << log_domain):
        x, y = canonic_domain_point(log_domain, bitrev(r, log_domain))
        out.append(sum(c * basis_eval(k, L, x, y) for k, c in enumerate(coeffs)) % P)
    return out


# ---- QM31 big-int arithmetic (CM31 = M31[i]/(i^2+1), QM31 = CM31[u]/(u^2-2-i)) ----
def cm_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def qm_mul(a, b):
    a0, a1 = (a[0], a[1]), (a[2], a[3])
    b0, b1 = (b[0], b[1]), (b[2], b[3])
    t = cm_mul(a1, b1)
    r = cm_mul((2, 1), t)
    c0 = ((cm_mul(a0, b0)[0] + r[0]) % P, (cm_mul(a0, b0)[1] + r[1]) % P)
    x, y = cm_mul(a0, b1), cm_mul(a1, b0)
    c1 = ((x[0] + y[0]) % P, (x[1] + y[1]) % P)
    return (c0[0], c0[1], c1[0], c1[1])


def qm_add(a, b):
    return tuple((x + y) % P for x, y in zip(a, b))


def qm_from(x):
    return (x % P, 0, 0, 0)


def qm_eval_at_point(coeffs, px, py):
    """Brute-force FFT-basis evaluation at a QM31 point (px, py)."""
    L = len(coeffs).bit_length() - 1
    acc = (0, 0, 0, 0)
    for k, c in enumerate(coeffs):
        term = qm_from(c)
        if k & 1:
            term = qm_mul(term, py)
        xx = px
        for b in range(1, L):
            if (k >> b) & 1:
                term = qm_mul(term, xx)
            xx2 = qm_mul(xx, xx)
            xx = qm_add(xx2, qm_add(xx2, (P - 1, 0, 0, 0)))
        acc = qm_add(acc, term)
    return acc
