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
