"""Host logic of the Engine's allocation pool (biomedisa_amd/_hip.py), with a
stand-in for the library's bmx_malloc/bmx_free: no device needed."""
import ctypes as C

import pytest

from biomedisa_amd import _hip


class FakeLib:
    def __init__(self, limit):
        self.limit, self.live, self.next, self.mallocs, self.frees = limit, {}, 0x1000, 0, 0

    def bmx_malloc(self, ctx, nbytes, out):
        if sum(self.live.values()) + nbytes > self.limit:
            return _hip.BMX_E_OOM
        self.next += 0x1000
        self.live[self.next] = nbytes
        out._obj.value = self.next
        self.mallocs += 1
        return _hip.BMX_OK

    def bmx_free(self, ctx, ptr):
        del self.live[ptr]
        self.frees += 1
        return _hip.BMX_OK

    def bmx_close(self, ctx):
        return _hip.BMX_OK

    def bmx_last_error(self):
        return b"out of memory"


@pytest.fixture
def fake(monkeypatch):
    lib = FakeLib(limit=1 << 40)
    monkeypatch.setattr(_hip, "load", lambda: lib)
    eng = object.__new__(_hip.Engine)
    eng.ctx, eng.device, eng._pool, eng._pooled = 1, 0, {}, 0
    return lib, eng


def test_large_buffers_are_reused(fake):
    lib, eng = fake
    big = _hip.POOL_MIN_BYTES
    a = eng.alloc(big)
    pa = a.ptr
    a.free()
    assert lib.frees == 0 and eng.pooled_bytes() == big
    b = eng.alloc(big)
    assert b.ptr == pa and lib.mallocs == 1 and eng.pooled_bytes() == 0
    c = eng.alloc(big + 1)          # another size: a new allocation
    assert c.ptr != pa and lib.mallocs == 2


def test_small_buffers_are_not_pooled(fake):
    lib, eng = fake
    a = eng.alloc(_hip.POOL_MIN_BYTES - 1)
    a.free()
    assert lib.frees == 1 and eng.pooled_bytes() == 0


def test_pool_is_capped(fake, monkeypatch):
    lib, eng = fake
    monkeypatch.setattr(_hip, "POOL_MAX_BYTES", 3 * _hip.POOL_MIN_BYTES)
    bufs = [eng.alloc(_hip.POOL_MIN_BYTES) for _ in range(5)]
    for b in bufs:
        b.free()
    assert eng.pooled_bytes() == 3 * _hip.POOL_MIN_BYTES and lib.frees == 2


def test_out_of_memory_empties_the_pool_and_retries(fake):
    lib, eng = fake
    lib.limit = 3 * _hip.POOL_MIN_BYTES
    a = eng.alloc(2 * _hip.POOL_MIN_BYTES)
    a.free()                         # pooled: the device still holds it
    b = eng.alloc(3 * _hip.POOL_MIN_BYTES)
    assert b.ptr is not None and eng.pooled_bytes() == 0 and lib.frees == 1
    with pytest.raises(_hip.BmxError) as e:
        eng.alloc(_hip.POOL_MIN_BYTES)
    assert e.value.code == _hip.BMX_E_OOM


def test_close_returns_the_pool(fake):
    lib, eng = fake
    eng.alloc(_hip.POOL_MIN_BYTES).free()
    eng.close()
    assert lib.live == {} and eng.ctx is None
