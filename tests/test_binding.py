"""The Python binding's column list (HipBackend.columns), on the host with a stand-in for the
library's entry points: the address array passed to whole-trace calls, column objects made only
when indexed, and every column freed exactly once whichever way it is freed (one by one, all
together, or when the list is collected)."""
import ctypes
import gc

import pytest

from hip_lib import SH


class FakeLib:
    def __init__(self):
        self.freed = []

    def stwo_hip_free(self, ctx, p):
        self.freed.append(p)
        return 0

    def stwo_hip_free_columns(self, ctx, arr, n):
        self.freed.extend(arr[i] for i in range(n))
        return 0


class FakeBackend:
    def __init__(self):
        self.lib = FakeLib()
        self.ctx = 1

    def _check(self, st):
        assert st == 0

    free_columns = SH.HipBackend.free_columns


def make(b, addrs, words=8):
    arr = (ctypes.c_void_p * len(addrs))(*addrs)
    return SH.ColumnList(b, arr, len(addrs), words)


def test_indexing_and_address_array():
    b = FakeBackend()
    cols = make(b, [16, 32, 48, 64])
    assert len(cols) == 4 and cols._made == []
    assert SH._ptrs(cols) is cols._arr  # whole-trace calls take the kept array
    assert cols[1].ptr == 32 and cols[-1].ptr == 64 and cols[1] is cols[1]
    assert [c.ptr for c in cols[0:2]] == [16, 32]
    assert sorted(cols._made) == [0, 1, 3]
    with pytest.raises(IndexError):
        cols[4]
    assert SH._min_words(cols) == 8


def test_free_all_at_once():
    b = FakeBackend()
    cols = make(b, [16, 32, 48])
    first = cols[0]
    b.free_columns(cols)
    assert sorted(b.lib.freed) == [16, 32, 48]
    assert first.ptr is None and all(c.ptr is None for c in cols)
    b.free_columns(cols)  # nothing left to free
    assert sorted(b.lib.freed) == [16, 32, 48]


def test_free_one_then_the_rest():
    b = FakeBackend()
    cols = make(b, [16, 32, 48])
    cols[1].free()
    assert b.lib.freed == [32] and not cols._live()
    b.free_columns(cols)
    assert sorted(b.lib.freed) == [16, 32, 48]


def test_collected_list_frees_what_is_left():
    b = FakeBackend()
    cols = make(b, [16, 32, 48, 64])
    cols[2].free()
    kept = cols[3]  # an indexed column frees itself when its object goes
    del cols, kept
    gc.collect()
    assert sorted(b.lib.freed) == [16, 32, 48, 64]
