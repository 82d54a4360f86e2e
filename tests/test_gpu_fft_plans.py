"""Circle-FFT launch plans on the GPU: every low-stage tile (2^11 .. 2^15 values per workgroup, the
stages above it in strided passes) gives the oracle's interpolation and evaluation, including
evaluation from fewer coefficients than the domain; the planner keeps the largest tile for
transforms that fill the device, returns a tile in range otherwise, and persists what it timed.
The looping kernel (plan flag 0x100) gives the same values."""
import os

import numpy as np
import pytest

import oracle_lib as O
from hip_lib import SH

pytestmark = pytest.mark.gpu
P = (1 << 31) - 1


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hb():
    b = SH.HipBackend(0)
    yield b
    b.close()


@pytest.mark.parametrize("low", [11, 12, 13, 14, 15, 15 | 0x100])
def test_every_tile_matches_the_oracle(hb, low):
    """0x100: the low stages in the looping kernel (workgroups loop over tiles, next tile
    prefetched), forward and inverse."""
    say("fft tile", low)
    rng = np.random.default_rng(900 + low)
    hb.set_fft_low_log(low)
    try:
        for log_n, blow, n in [(15, 1, 3), (16, 1, 2), (13, 2, 2)]:
            cols = [rng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(n)]
            dev = [hb.from_cpu(c) for c in cols]
            coeffs = hb.interpolate_columns(dev, log_n)
            ext = hb.evaluate_polynomials(coeffs, log_n, blow)
            for c, cf, e in zip(cols, coeffs, ext):
                exp_c = O.interpolate(c)
                assert np.array_equal(cf.to_cpu(), exp_c), (low, log_n)
                assert np.array_equal(e.to_cpu(), O.evaluate(exp_c, log_n + blow)), (low, log_n, blow)
    finally:
        hb.set_fft_low_log(0)


def test_plans_are_in_range_and_persist(hb, tmp_path):
    old = os.environ.get("STWO_HIP_KERNEL_CACHE")
    os.environ["STWO_HIP_KERNEL_CACHE"] = str(tmp_path)
    try:
        # 15 stages filling the device many times over: the largest tile without timing, in the
        # looping kernel when forward
        assert hb.fft_low_log(False, 4096, 15, 16) == 15 | 0x100
        assert hb.fft_low_log(True, 4096, 15, 15) == 15
        # more stages: timed (on a subset of the columns), a tile in range
        s0 = hb.fft_low_log(True, 4096, 16, 16)
        assert 11 <= (s0 & 0xff) <= 15 and (not s0 & 0x100 or s0 & 0xff == 15)
        assert hb.fft_low_log(False, 7, 10, 12) == 0  # 10 stages: strided passes only
        # few columns: a timed choice in range, the same on the next query (the looping kernel is a
        # candidate with 2^15-value tiles)
        s1 = hb.fft_low_log(True, 5, 17, 17)
        s2 = hb.fft_low_log(False, 3, 14, 16)
        say("planned tiles %#x %#x" % (s1, s2))
        for plan, top in ((s1, 15), (s2, 14)):
            assert 11 <= (plan & 0xff) <= top and (plan & ~0x1ff) == 0 and (not plan & 0x100 or plan & 0xff == 15)
        assert hb.fft_low_log(True, 5, 17, 17) == s1
        lines = []
        for f in os.listdir(tmp_path):
            if f.startswith("fft_plans_"):
                lines += open(os.path.join(tmp_path, f)).read().split("\n")
        assert "1 5 17 17 %d" % s1 in lines and "0 3 14 16 %d" % s2 in lines
    finally:
        if old is None:
            del os.environ["STWO_HIP_KERNEL_CACHE"]
        else:
            os.environ["STWO_HIP_KERNEL_CACHE"] = old


def test_looping_plan_falls_back_for_unaligned_outputs(hb):
    """The looping kernel writes 16 bytes at a time; output columns 4 bytes off alignment take the
    plain kernel under the same plan and still get the oracle's values."""
    rng = np.random.default_rng(31)
    log_n, n = 15, 2
    cols = [rng.integers(0, P, 1 << log_n, dtype=np.uint32) for _ in range(n)]
    coeffs = hb.interpolate_columns([hb.from_cpu(c) for c in cols], log_n)
    big = hb.column(1 + n * (2 << log_n), zeroed=True)
    out = [big.ptr + 4 * (1 + i * (2 << log_n)) for i in range(n)]
    hb.set_fft_low_log(15 | 0x100)
    try:
        hb.evaluate_polynomials(coeffs, log_n, 1, out=out)
    finally:
        hb.set_fft_low_log(0)
    got = big.to_cpu()[1:]
    for i, c in enumerate(cols):
        assert np.array_equal(got[i * (2 << log_n):(i + 1) * (2 << log_n)], O.evaluate(O.interpolate(c), log_n + 1))
