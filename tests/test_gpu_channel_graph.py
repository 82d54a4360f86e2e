"""The Blake2s channel on the device (stwo_hip_channel_*) and graph capture (stwo_hip_graph_*).

The device channel's transcript (mix_root, mix_u32s, draw_secure_felts, draw_queries, grind) equals
the pure-Python restatement's (tests/pystark.py, hashlib). An FRI commit loop driven only by
device-side calls (each layer's Merkle tree by commit_on_layer, its root mixed on the device, the
folding coefficient drawn on the device and used from device memory by the _dev folds, then
queries and the proof of work) equals the same loop driven from the host with the host-side
restatement of the channel and the trait folds. Captured in a graph, that loop replays to the
eager results on several replays in a row, on new input data written into the same buffers, with
unrelated eager work and synchronisations between the replays; an entry point that cannot be
captured is refused during the capture with status 3 and leaves the capture usable."""
import random

import numpy as np
import pytest

import pystark as S
from hip_lib import SH

pytestmark = pytest.mark.gpu
M31 = (1 << 31) - 1


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hb():
    b = SH.HipBackend(0)
    yield b
    b.close()


def put(hb, col, arr):
    a = np.ascontiguousarray(arr, np.uint32)
    assert a.size <= col.n
    hb._check(hb.lib.stwo_hip_copy_h2d(hb.ctx, col.ptr, a.ctypes.data, a.size))


def words_bytes(w):
    return np.asarray(w, np.uint32).tobytes()


def felts(words):
    w = [int(x) for x in words]
    return [tuple(w[4 * i:4 * i + 4]) for i in range(len(w) // 4)]


def test_device_channel_matches_restatement(hb):
    rng = random.Random(12)
    digest = bytes(rng.randrange(256) for _ in range(32))
    st = hb.channel_new(digest, 3)
    ch = S.Channel()
    ch.digest, ch.n_draws = digest, 3
    root = [rng.randrange(1 << 32) for _ in range(8)]
    rc = hb.from_cpu(np.array(root, np.uint32))
    hb.channel_mix_root(st, rc)
    ch.mix_root(words_bytes(root))
    assert felts(hb.channel_draw_secure_felts(st, 5).to_cpu()) == ch.draw_secure_felts(5)
    words = [rng.randrange(1 << 32) for _ in range(11)]
    hb.channel_mix_u32s(st, hb.from_cpu(np.array(words, np.uint32)))
    ch.mix_u32s(words)
    assert [int(x) for x in hb.channel_draw_queries(st, 14, 13).to_cpu()] == S.draw_queries(ch, 14, 13)
    assert hb.channel_read(st) == (ch.digest, ch.n_draws)
    for bits in (1, 6, 10):
        nonce = hb.channel_grind(st, bits)
        nv = [int(x) for x in nonce.to_cpu()]
        exp = S.grind(ch, bits)
        assert nv[0] | nv[1] << 32 == exp
        hb.channel_mix_u32s(st, nonce, 2)
        ch.mix_u64(exp)
        assert hb.channel_read(st) == (ch.digest, ch.n_draws)


L0, LAST = 12, 3


class Loop:
    """The buffers of one FRI commit loop: a secure column of 2^L0 values, then line layers down to
    2^LAST, each layer's Merkle levels, the drawn coefficients, queries and nonce."""

    def __init__(self, hb):
        self.first = [hb.column(1 << L0) for _ in range(4)]
        self.layers = {k: [hb.column(1 << k) for _ in range(4)] for k in range(LAST, L0)}
        self.trees = {k: {ls: hb.column(8 << ls) for ls in range(k + 1)} for k in range(LAST + 1, L0 + 1)}
        self.alphas = {k: hb.column(4) for k in range(LAST + 1, L0 + 1)}
        self.st = hb.column(SH.CHANNEL_WORDS, zeroed=True)
        self.queries = hb.column(6)
        self.nonce = hb.column(2)

    def cols(self, k):
        return self.first if k == L0 else self.layers[k]

    def run(self, hb, digest):
        """The loop through device-side calls only (capturable)."""
        hb.channel_init(self.st, digest, 0)
        for k in range(L0, LAST, -1):
            t = self.trees[k]
            hb.commit_on_layer(k, None, self.cols(k), out=t[k])
            for ls in range(k - 1, -1, -1):
                hb.commit_on_layer(ls, t[ls + 1], [], out=t[ls])
            hb.channel_mix_root(self.st, t[0])
            hb.channel_draw_secure_felts(self.st, 1, out=self.alphas[k])
            if k == L0:
                hb.fold_circle_into_line_dev(self.layers[k - 1], self.first, self.alphas[k], k, dst_is_zero=True)
            else:
                hb.fold_line_dev(self.layers[k - 1], self.layers[k], self.alphas[k], k)
        hb.channel_mix_u32s(self.st, self.layers[LAST][0], 1 << LAST)  # the last layer's first coordinate
        hb.channel_draw_queries(self.st, L0, 6, out=self.queries)
        hb.channel_grind(self.st, 5, nonce=self.nonce)

    def results(self, hb):
        out = {"st": hb.channel_read(self.st), "q": self.queries.to_cpu().tolist(),
               "nonce": self.nonce.to_cpu().tolist()}
        for k in range(LAST, L0):
            out["l%d" % k] = [c.to_cpu().tolist() for c in self.layers[k]]
        for k in range(LAST + 1, L0 + 1):
            out["r%d" % k] = self.trees[k][0].to_cpu().tolist()
            out["a%d" % k] = self.alphas[k].to_cpu().tolist()
        return out


def random_first(rng):
    return [np.array([rng.randrange(M31) for _ in range(1 << L0)], np.uint32) for _ in range(4)]


def test_device_loop_equals_host_driven_loop(hb):
    rng = random.Random(5)
    data = random_first(rng)
    lp = Loop(hb)
    for c, d in zip(lp.first, data):
        put(hb, c, d)
    digest = bytes(range(32))
    lp.run(hb, digest)
    hb.synchronize()
    got = lp.results(hb)
    # the host-driven loop: roots read back, the restatement's channel, the trait folds
    ch = S.Channel()
    ch.digest = digest
    cur = [hb.from_cpu(d) for d in data]
    for k in range(L0, LAST, -1):
        leaves = hb.commit_on_layer(k, None, cur)
        lvl = leaves
        for ls in range(k - 1, -1, -1):
            lvl = hb.commit_on_layer(ls, lvl, [])
        root = lvl.to_cpu()
        assert root.tolist() == got["r%d" % k]
        ch.mix_root(root.tobytes())
        a = ch.draw_secure_felt()
        assert list(a) == got["a%d" % k]
        if k == L0:
            nxt = [hb.column(1 << (k - 1), zeroed=True) for _ in range(4)]
            hb.fold_circle_into_line(nxt, cur, a, k)
        else:
            nxt = hb.fold_line(cur, a, k)
        assert [c.to_cpu().tolist() for c in nxt] == got["l%d" % (k - 1)]
        cur = nxt
    ch.mix_u32s([int(x) for x in cur[0].to_cpu()])
    assert got["q"] == S.draw_queries(ch, L0, 6)
    n = S.grind(ch, 5)
    assert got["nonce"] == [n & 0xFFFFFFFF, n >> 32]
    assert got["st"] == (ch.digest, ch.n_draws)


def test_captured_loop_replays(hb):
    rng = random.Random(8)
    data_a, data_b = random_first(rng), random_first(rng)
    lp = Loop(hb)
    digest = bytes(rng.randrange(256) for _ in range(32))

    def load(data):
        for c, d in zip(lp.first, data):
            put(hb, c, d)

    load(data_a)
    lp.run(hb, digest)  # eager; also builds the twiddles the capture needs
    hb.synchronize()
    res_a = lp.results(hb)
    load(data_b)
    lp.run(hb, digest)
    hb.synchronize()
    res_b = lp.results(hb)
    assert res_a != res_b

    hb.graph_begin()
    try:
        # an entry point that synchronises is refused, and the capture goes on
        with pytest.raises(SH.StwoHipError) as e:
            hb.channel_read(lp.st)
        assert e.value.status == 3
        lp.run(hb, digest)
    finally:
        g = hb.graph_end()
    say("captured the FRI commit loop")
    other = hb.from_cpu(np.arange(1 << 10, dtype=np.uint32))
    try:
        for rep, (data, exp) in enumerate([(data_a, res_a), (data_a, res_a), (data_b, res_b), (data_a, res_a)]):
            load(data)
            hb.commit_on_layer(10, None, [other, other, other])  # unrelated eager work between replays
            hb.synchronize()
            hb.graph_launch(g)
            hb.synchronize()
            assert lp.results(hb) == exp, rep
    finally:
        hb.graph_free(g)
