"""The continuous-batching serving loop: results bit-exact, identical alone and in the mix."""
import time

import numpy as np
import pytest

import clue_hip
import helpers
import oracle
import synth

pytestmark = pytest.mark.gpu


def mixed_events():
    rng = np.random.Generator(np.random.PCG64(21))
    evs = []
    for k in range(24):
        if rng.random() < 0.7:
            n = int(round(10 ** rng.uniform(3, 5.3)))
            evs.append((2, synth.clue2d_event(n, 900 + k, shuffle=(k % 4 == 0))))
        else:
            evs.append((3, synth.clue3d_event(900 + k)))
    return evs


def run_all(server, evs):
    ids = [server.submit2d(e) if d == 2 else server.submit3d(e) for d, e in evs]
    return [server.wait(i) for i in ids]


def test_server_mix_equals_alone_equals_oracle(dev, oracle_lib):
    evs = mixed_events()
    p3 = clue_hip.params3d(S=2)
    print("case server mix", flush=True)
    mix = clue_hip.Server(dev, p3=p3, max_points_2d=2_000_000, max_points_3d=400_000, max_events=32, slots=2)
    got_mix = run_all(mix, evs)
    stats = mix.stats()
    mix.close()
    assert stats["events"] == len(evs) and stats["batches"] < len(evs)  # events were batched together
    print("case server alone", flush=True)
    alone = clue_hip.Server(dev, p3=p3, max_points_2d=2_000_000, max_points_3d=400_000, max_events=1, slots=1)
    got_alone = run_all(alone, evs)
    assert alone.stats()["batches"] == len(evs)
    alone.close()
    for k, ((d, e), (om, lm), (oa, la)) in enumerate(zip(evs, got_mix, got_alone)):
        keys = helpers.OUT2D if d == 2 else helpers.OUT3D
        helpers.assert_bit_exact(om, oa, keys, f"event {k} mix vs alone")
        want = (oracle_lib.clue2d(e["x"], e["y"], e["layer"], e["weight"]) if d == 2
                else oracle_lib.clue3d(e, oracle.clue3d_params(S=2)))
        helpers.assert_bit_exact(om, want, keys, f"event {k} vs oracle")
        assert lm > 0 and la > 0


def test_server_rejects_bad_events(dev):
    s = clue_hip.Server(dev, max_points_2d=100_000, max_points_3d=10_000, max_events=8, slots=1)
    ev = synth.clue2d_event(2000, 1)
    ev["layer"][3] = -1
    with pytest.raises(clue_hip.ClueError) as e:
        s.submit2d(ev)
    assert e.value.code == -4
    with pytest.raises(clue_hip.ClueError) as e:
        s.submit2d(synth.clue2d_event(200_000, 2))
    assert e.value.code == -3
    out, lat = s.wait(s.submit2d(synth.clue2d_event(3000, 3)))  # the server keeps working
    assert out["rho"].size == 3000
    s.close()


def test_server_poisson_stream(dev):
    rng = np.random.Generator(np.random.PCG64(5))
    pool = [synth.clue2d_event(int(round(10 ** rng.uniform(3, 5))), 950 + k) for k in range(16)]
    s = clue_hip.Server(dev, max_points_2d=2_000_000, max_points_3d=100_000, max_events=32, slots=3)
    print("case server poisson", flush=True)
    t0 = time.perf_counter()
    t = 0.0
    ids = []
    for k in range(300):
        t += rng.exponential(1.0 / 3000.0)
        while time.perf_counter() - t0 < t:
            pass
        ids.append(s.submit2d(pool[k % len(pool)]))
    lat = np.array([s.wait(i)[1] for i in ids])
    st = s.stats()
    s.close()
    assert st["events"] == 300 and np.all(lat > 0)
    print(f"poisson: p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms, "
          f"{st['batches']} batches", flush=True)


def test_server_flags_a_bad_event_alone(dev, oracle_lib):
    """A CLUE3D event whose phi has no tile bin (NaN, infinite) fails alone with CLUE_ERR_BIN; the
    events batched with it are exact, and the context's own error flag stays clear."""
    p3 = clue_hip.params3d(S=2)
    evs = [synth.clue3d_event(960 + k) for k in range(5)]
    bad = {k: v.copy() for k, v in evs[2].items()}
    bad["phi"][[3, 40]] = np.array([np.nan, np.inf], np.float32)
    s = clue_hip.Server(dev, p3=p3, max_points_2d=100_000, max_points_3d=200_000, max_events=16, slots=2)
    print("case server bad-phi event in a batch", flush=True)
    ids = [s.submit3d(bad if k == 2 else e) for k, e in enumerate(evs)]
    for k, i in enumerate(ids):
        if k == 2:
            with pytest.raises(clue_hip.ClueError) as e:
                s.wait(i)
            assert e.value.code == -5
        else:
            out, _ = s.wait(i)
            helpers.assert_bit_exact(out, oracle_lib.clue3d(evs[k], oracle.clue3d_params(S=2)), helpers.OUT3D,
                                     f"event {k} beside the bad one")
    s.close()
    dev.check()


def test_server_destroy_finishes_running_batches(dev, oracle_lib):
    """Destroying a server with batches still queued and on the device finishes them first: every
    caller buffer, pinned or pageable, holds the event's result afterwards."""
    evs = [synth.clue2d_event(n, 980 + k) for k, n in enumerate([300_000, 2_000, 150_000, 40_000, 600_000, 9_000,
                                                                  75_000, 1_000])]
    s = clue_hip.Server(dev, max_points_2d=1_000_000, max_points_3d=1_000, max_events=2, slots=3)
    print("case server destroy with batches in flight", flush=True)
    outs = []
    for k, e in enumerate(evs):
        n = e["x"].size
        if k % 2 == 0:  # pinned outputs: copied by DMA straight from the device
            o = {f: dev.pinned_copy(np.full(n, -7, np.int32 if f in ("nearestHigher", "isSeed", "clusterIndex")
                                            else np.float32)) for f in clue_hip.FIELDS_2D_OUT}
        else:
            o = None
        rid = s.submit2d(e, o)
        outs.append(s.outputs(rid))
    s.close()
    for k, (e, o) in enumerate(zip(evs, outs)):
        helpers.assert_bit_exact(o, oracle_lib.clue2d(e["x"], e["y"], e["layer"], e["weight"]), helpers.OUT2D,
                                 f"event {k} after destroy")
