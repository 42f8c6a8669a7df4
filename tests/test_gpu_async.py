"""clue_stream_callback: the completion callback the external-work modules hand their holder to.

The callback copies a pinned host buffer that a device-to-host copy enqueued before it fills; the
copy it sees must be the finished clustering, equal to the oracle's."""
import ctypes as C
import time

import numpy as np
import pytest

import clue_hip
import helpers
import synth

pytestmark = pytest.mark.gpu

CALLBACK = C.CFUNCTYPE(None, C.c_void_p)


def test_stream_callback_follows_the_enqueued_work(dev, oracle_lib):
    ev = synth.clue2d_event(300000, 41)
    n = ev["x"].size
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], 3.0, 10.0, 2.0)
    din = dev.upload_soa(ev, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(n, clue_hip.FIELDS_2D_OUT)
    host = dev.pinned_copy(np.full(n, -7, np.int32))
    snapshots = []
    cb = CALLBACK(lambda user: snapshots.append((user, host.copy())))
    fn = C.cast(cb, C.c_void_p)
    print("case stream callback 2d 300000", flush=True)
    for rep in range(3):
        host[:] = -7
        dev.clue2d_device(din, dout, n, clue_hip.params2d(3.0, 10.0, 2.0))
        assert dev.lib.clue_memcpy_dtoh(dev.ctx, host.ctypes.data, dout["clusterIndex"].ptr, n * 4, None) == 0
        assert dev.lib.clue_stream_callback(dev.ctx, None, fn, rep + 1) == 0
        dev.check()
        deadline = time.monotonic() + 5.0
        while len(snapshots) < rep + 1 and time.monotonic() < deadline:
            time.sleep(0.001)
    assert [s[0] for s in snapshots] == [1, 2, 3]
    for _, got in snapshots:
        helpers.assert_bit_exact({"clusterIndex": got}, want, ("clusterIndex",), "callback snapshot")
    assert dev.lib.clue_stream_callback(dev.ctx, None, None, None) == -1
    for a in list(din.values()) + list(dout.values()):
        a.free()


def _batch(dev, packed, offs, fields_in, fields_out):
    N = int(offs[-1])
    return dev.upload_soa(packed, fields_in), dev.alloc_soa(N, fields_out)


def test_batches_enqueue_back_to_back_without_waiting(dev, oracle_lib):
    """Two different CLUE2D batches and two CLUE3D batches are enqueued on one stream with no host
    synchronisation between them (the second of each kind while the first still runs, smaller so
    that the shared workspace does not grow); every event of all four matches the oracle. The block
    tables travel through the context's pinned ring and stream-ordered copies, so the second batch's
    tables cannot overwrite the first's before its kernels ran."""
    s = dev.lib.clue_stream_create(dev.ctx)
    assert s
    b2 = [synth.pack([synth.clue2d_event(n, 70 + i, shuffle=(i % 2 == 1)) for i, n in enumerate(ns)])
          for ns in ([900_000, 4_000, 250_000, 1_000], [20_000, 333_333, 7_777])]
    b3 = [synth.clue3d_batch(k, first_event=40 * k) for k in (24, 9)]
    io2 = [_batch(dev, p, o, clue_hip.FIELDS_2D_IN, clue_hip.FIELDS_2D_OUT) for p, o in b2]
    io3 = [_batch(dev, p, o, clue_hip.FIELDS_3D_IN, clue_hip.FIELDS_3D_OUT) for p, o in b3]
    p3 = clue_hip.params3d(S=2)
    print("case back-to-back batches: 2D 1155000 + 361110 points, 3D 24 + 9 events", flush=True)
    dev.sync()
    for (din, dout), (_, offs) in zip(io2, b2):
        dev.clue2d_batch(din, dout, offs, stream=s)
    for (din, dout), (_, offs) in zip(io3, b3):
        dev.clue3d_batch(din, dout, offs, p3, stream=s)
    dev.check(s)
    import oracle
    for k, ((din, dout), (packed, offs)) in enumerate(zip(io2, b2)):
        helpers.assert_bit_exact(dev.download_soa(dout), oracle_lib.clue2d_batch(packed, offs), helpers.OUT2D,
                                 f"2d batch {k}")
    for k, ((din, dout), (packed, offs)) in enumerate(zip(io3, b3)):
        helpers.assert_bit_exact(dev.download_soa(dout), oracle_lib.clue3d_batch(packed, offs, oracle.clue3d_params(S=2)),
                                 helpers.OUT3D, f"3d batch {k}")
    for din, dout in io2 + io3:
        for a in list(din.values()) + list(dout.values()):
            a.free()
    assert dev.lib.clue_stream_destroy(dev.ctx, s) == 0
