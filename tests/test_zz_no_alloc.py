"""clue2d_run_device_noalloc: the allocation-free variant of clue2d_run_device.

It runs the same kernels on the library's fixed per-GPU scratch array instead of the context's
workspace, so its results must equal clue2d_run_device's bit for bit. The array keeps the previous
call's contents, so two calls in a row are checked, the second on a smaller event (stale words of the
first call must not leak into it), and the largest accepted event and the rejection above it."""
import numpy as np
import pytest

import clue_hip
import helpers
import synth

pytestmark = pytest.mark.gpu


def _run(dev, ev, noalloc, p=None):
    n = ev["x"].size
    din = dev.upload_soa(ev, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(n, clue_hip.FIELDS_2D_OUT)
    for a in dout.values():
        a.upload(np.full(n, -7, a.dtype))
    if noalloc:
        dev.clue2d_device_noalloc(din, dout, n, p)
    else:
        dev.clue2d_device(din, dout, n, p)
    dev.check()
    out = dev.download_soa(dout)
    for a in list(din.values()) + list(dout.values()):
        a.free()
    return out


def test_noalloc_matches_run_device_twice():
    dev = clue_hip.Device(0)
    try:
        for k, (n, seed, shuffle) in enumerate([(400_000, 71, False), (37_000, 72, True)]):
            ev = synth.clue2d_event(n, seed, shuffle=shuffle)
            print(f"case noalloc call {k} n={n} seed={seed}", flush=True)
            want = _run(dev, ev, noalloc=False)
            got = _run(dev, ev, noalloc=True)
            helpers.assert_bit_exact(got, want, helpers.OUT2D, f"noalloc call {k}")
    finally:
        dev.close()


def test_noalloc_other_parameters_and_limits():
    dev = clue_hip.Device(0)
    try:
        ins, want, extra = helpers.load_case("clue2d", "hgcal_config")
        dc, rhoc, odf = (float(v) for v in extra["params"])
        print(f"case noalloc golden hgcal_config n={ins['x'].size}", flush=True)
        got = _run(dev, ins, noalloc=True, p=clue_hip.params2d(dc, rhoc, odf))
        helpers.assert_bit_exact(got, want, helpers.OUT2D, "noalloc hgcal_config")

        n = clue_hip.NOALLOC_MAX_POINTS
        ev = synth.clue2d_event(n, 73)
        print(f"case noalloc largest event n={n}", flush=True)
        helpers.assert_bit_exact(_run(dev, ev, noalloc=True), _run(dev, ev, noalloc=False), helpers.OUT2D,
                                 "noalloc largest")

        big = {k: np.zeros(n + 1, v.dtype) for k, v in ev.items()}
        din = dev.upload_soa(big, clue_hip.FIELDS_2D_IN)
        dout = dev.alloc_soa(n + 1, clue_hip.FIELDS_2D_OUT)
        with pytest.raises(clue_hip.ClueError) as e:
            dev.clue2d_device_noalloc(din, dout, n + 1)
        assert e.value.code == -7
        for a in list(din.values()) + list(dout.values()):
            a.free()
    finally:
        dev.close()
