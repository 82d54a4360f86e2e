"""The Poseidon252 restatement (tests/p252.py) pinned on the host: the reference's known answers for
the Merkle hasher (crates/stwo/src/core/vcs/poseidon252_merkle.rs:104-124) and the channel
(crates/stwo/src/core/channel/poseidon252.rs:223-253), the MDS mix against its matrix, the packing
against its bit layout, and a planted change to a round constant failing them."""
import p252 as Q


def test_reference_merkle_vectors():
    assert Q.hash_node(None, [0, 1]) == \
        2978883932528585652864046122079599882777358126302490183268546077323303473078
    assert Q.hash_node((1, 2), [3]) == \
        3286095315900630438551061262740794783852190427874264245042874292062185873630


def test_reference_channel_vectors():
    c = Q.Channel()
    c.mix_u64(0x1111222233334444)
    assert c.digest == 0x07cecc0ee3d858c843fe63165f038353f9f80f52dd8d32eead9f635e2f7d8b8e
    c = Q.Channel()
    c.mix_u32s([1, 2, 3, 4, 5, 6, 7, 8, 9])
    assert c.digest == 0x06c7fc11690eb272bcc81115e801ad52de4e6271ddff3f97a2b75315e3572ced


def test_mix_is_the_mds_matrix():
    s = [5, 7, 11]
    m = [[3, 1, 1], [1, -1, 1], [1, 1, -2]]
    assert Q.mix(s) == [sum(m[i][j] * s[j] for j in range(3)) % Q.P for i in range(3)]


def test_packing_layout():
    word = [1, 2, 3, 4, 5, 6, 7, 8]
    f = Q.felt_from_m31s(word)
    assert f == sum(w << (31 * (7 - i)) for i, w in enumerate(word))
    assert Q.felt_from_m31s([9, 10, 11]) == (9 << 62) + (10 << 31) + 11 + (3 << 248)


def test_channel_draw_counts_and_uniqueness():
    c = Q.Channel()
    a = c.draw_u32s()
    assert c.n_draws == 1 and a != c.draw_u32s()
    d = [c.draw_secure_felt() for _ in range(5)]
    assert len(set(d)) == 5 and all(all(x < (1 << 31) - 1 for x in q) for q in d)
    # channel/poseidon252.rs:169-180: draw_u32s, then 9 secure felts from 5 draws of 8 base felts
    c = Q.Channel()
    c.draw_u32s()
    nine = c.draw_secure_felts(9)
    assert c.n_draws == 6
    # the flattened stream: the first draw's 8 base felts are the first two secure felts
    c2 = Q.Channel()
    c2.draw_u32s()
    b = c2.draw_base_felts()
    assert nine[0] == tuple(b[:4]) and nine[1] == tuple(b[4:])
    # channel/poseidon252.rs:203-214: 5 then 4 felts, all distinct
    c = Q.Channel()
    f = c.draw_secure_felts(5) + c.draw_secure_felts(4)
    assert len(set(f)) == 9


def test_planted_constant_fails_the_vectors():
    old = Q.RC[50]
    try:
        Q.RC[50] = (old + 1) % Q.P
        assert Q.hash_node(None, [0, 1]) != \
            2978883932528585652864046122079599882777358126302490183268546077323303473078
    finally:
        Q.RC[50] = old


# ---- the library's host code (felt252.h, the same code the device runs) against the restatement ----
def test_library_host_permutation_and_hash_many():
    import random
    from hip_lib import SH
    rnd = random.Random(3)
    for _ in range(4):
        st = [rnd.randrange(Q.P) for _ in range(3)]
        assert SH.poseidon252_permute(st) == Q.permute(st)
    for n in range(0, 7):
        vals = [rnd.randrange(Q.P) for _ in range(n)]
        assert SH.poseidon252_hash_many(vals) == Q.hash_many(vals)
    # the reference's vector through the library: hash_node(None, [0, 1]) is hash_many of one packed felt
    assert SH.poseidon252_hash_many([Q.felt_from_m31s([0, 1])]) == \
        2978883932528585652864046122079599882777358126302490183268546077323303473078
    # edge values: 0, 1, p - 1
    assert SH.poseidon252_permute([0, 1, Q.P - 1]) == Q.permute([0, 1, Q.P - 1])


def test_library_lazy_reduction_on_extreme_states():
    """The permutation keeps values unreduced between rounds (below 2^256, limbs of 29 bits); states
    at the top of the range, with every 29-bit limb full, and with long runs of equal elements drive
    those bounds hardest. Chains of permutations feed outputs back in as the sponge does."""
    import random
    from hip_lib import SH
    rnd = random.Random(11)
    P = Q.P
    edge = [0, 1, 2, P - 1, P - 2, (1 << 251) - 1, 1 << 251, P - (1 << 192), (1 << 232) - 1,
            sum(((1 << 29) - 1) << (29 * i) for i in range(8)) % P, (P - 1) // 2, (P + 1) // 2]
    states = [[a, b, c] for a in edge[:6] for b in edge[3:9] for c in edge[6:]][:120]
    states += [[rnd.randrange(P) for _ in range(3)] for _ in range(40)]
    for st in states:
        assert SH.poseidon252_permute(st) == Q.permute(st), st
    # chained: the library's outputs fed back agree with the restatement's chain
    a = b = [P - 1, P - 2, (1 << 251) - 1]
    for _ in range(25):
        a, b = SH.poseidon252_permute(a), Q.permute(b)
        assert a == b
    for n in (1, 2, 9, 16, 33):
        vals = [rnd.choice(edge) if rnd.random() < 0.5 else rnd.randrange(P) for _ in range(n)]
        assert SH.poseidon252_hash_many(vals) == Q.hash_many(vals)


def test_library_rejects_non_canonical():
    import pytest
    from hip_lib import SH
    with pytest.raises(Exception):
        SH.poseidon252_hash_many([Q.P])
