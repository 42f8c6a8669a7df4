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
