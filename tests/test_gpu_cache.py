"""The stream-ordered caching allocator (clue_cache_*): bin rounding, reuse on the same stream and
after the recorded event, release under the HBM budget, error paths, and a clustering run whose
buffers all come from the cache."""
import ctypes as C

import numpy as np
import pytest

import clue_hip
import helpers
import synth

pytestmark = pytest.mark.gpu


def stats(dev):
    v = [C.c_uint64(), C.c_uint64(), C.c_int64(), C.c_int64()]
    assert dev.lib.clue_cache_stats(dev.ctx, *[C.byref(x) for x in v]) == 0
    return tuple(x.value for x in v)


@pytest.fixture
def cdev():
    d = clue_hip.Device(0, budget_bytes=256 << 20)
    yield d
    d.close()


def test_bins_and_reuse_on_one_stream(cdev):
    lib, ctx = cdev.lib, cdev.ctx
    print("case cache bins", flush=True)
    a = lib.clue_cache_alloc(ctx, 1000, None)
    assert a and stats(cdev) == (1024, 0, 0, 1)
    assert lib.clue_cache_free(ctx, a) == 0 and stats(cdev) == (0, 1024, 0, 1)
    b = lib.clue_cache_alloc(ctx, 600, None)  # same 1 KiB bin, same stream: reused at once
    assert b == a and stats(cdev) == (1024, 0, 1, 1)
    c = lib.clue_cache_alloc(ctx, 1025, None)  # next bin: a new block
    assert c and c != a and stats(cdev) == (1024 + 2048, 0, 1, 2)
    tiny = lib.clue_cache_alloc(ctx, 1, None)  # the smallest bin is 256 B
    assert stats(cdev)[0] == 1024 + 2048 + 256
    for p in (b, c, tiny):
        assert lib.clue_cache_free(ctx, p) == 0
    assert stats(cdev)[:2] == (0, 1024 + 2048 + 256)
    assert lib.clue_cache_release(ctx) == 0 and stats(cdev)[:2] == (0, 0)
    assert lib.clue_cache_free(ctx, a) == -1  # already released
    assert lib.clue_cache_free(ctx, None) == 0


def test_reuse_across_streams_after_the_event(cdev):
    lib, ctx = cdev.lib, cdev.ctx
    s1, s2 = lib.clue_stream_create(ctx), lib.clue_stream_create(ctx)
    print("case cache streams", flush=True)
    a = lib.clue_cache_alloc(ctx, 32 << 20, s1)
    src = lib.clue_cache_alloc(ctx, 32 << 20, s1)
    for _ in range(8):
        assert lib.clue_memcpy_dtod(ctx, a, src, 32 << 20, s1) == 0
    assert lib.clue_cache_free(ctx, a) == 0
    assert lib.clue_stream_sync(ctx, s1) == 0  # the block's event has completed
    b = lib.clue_cache_alloc(ctx, 20 << 20, s2)
    assert b == a
    assert lib.clue_cache_free(ctx, b) == 0 and lib.clue_cache_free(ctx, src) == 0
    assert lib.clue_stream_sync(ctx, s2) == 0
    for s in (s1, s2):
        assert lib.clue_stream_destroy(ctx, s) == 0


def test_budget_releases_idle_blocks(cdev):
    lib, ctx = cdev.lib, cdev.ctx
    print("case cache budget", flush=True)
    blocks = [lib.clue_cache_alloc(ctx, 64 << 20, None) for _ in range(3)]  # 192 of 256 MiB
    assert all(blocks)
    for p in blocks:
        assert lib.clue_cache_free(ctx, p) == 0
    assert stats(cdev)[1] == 192 << 20
    big = lib.clue_cache_alloc(ctx, 100 << 20, None)  # 128 MiB bin: fits only once the idle blocks go
    assert big and stats(cdev)[:2] == (128 << 20, 0)
    assert not lib.clue_cache_alloc(ctx, 200 << 20, None)  # 256 MiB bin over the budget
    assert "budget" in cdev.error()
    assert not lib.clue_cache_alloc(ctx, (1 << 34) + 1, None)  # exact size above the largest bin
    assert lib.clue_cache_free(ctx, big) == 0
    assert lib.clue_cache_free(ctx, 12345) == -1


def test_clustering_in_cached_buffers(cdev, oracle_lib):
    lib, ctx = cdev.lib, cdev.ctx
    ev = synth.clue2d_event(120000, 17)
    n = ev["x"].size
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], 3.0, 10.0, 2.0)
    print("case cache clustering", flush=True)
    for rep in range(3):  # the second and third runs take every buffer from the cache
        ptr = {k: lib.clue_cache_alloc(ctx, n * 4, None) for k in clue_hip.FIELDS_2D_IN + clue_hip.FIELDS_2D_OUT}
        assert all(ptr.values())
        for k in clue_hip.FIELDS_2D_IN:
            a = np.ascontiguousarray(ev[k], dtype=np.int32 if k == "layer" else np.float32)
            assert lib.clue_memcpy_htod(ctx, ptr[k], a.ctypes.data, n * 4, None) == 0
        p = clue_hip.params2d(3.0, 10.0, 2.0)
        assert lib.clue2d_run_device(ctx, C.byref(p), n, *[ptr[k] for k in clue_hip.FIELDS_2D_IN],
                                     *[ptr[k] for k in clue_hip.FIELDS_2D_OUT], None) == 0
        got = {}
        for k in clue_hip.FIELDS_2D_OUT:
            got[k] = np.empty(n, np.float32 if k in ("rho", "delta") else np.int32)
            assert lib.clue_memcpy_dtoh(ctx, got[k].ctypes.data, ptr[k], n * 4, None) == 0
        cdev.check()
        helpers.assert_bit_exact(got, want, helpers.OUT2D, f"cached buffers run {rep}")
        for v in ptr.values():
            assert lib.clue_cache_free(ctx, v) == 0
    hits, misses = stats(cdev)[2:]
    assert misses == 9 and hits == 18
