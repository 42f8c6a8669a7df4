"""CLUE2D on the GPU, bit-exact against the reference fixtures and the CPU oracle."""
import numpy as np
import pytest

import clue_hip
import helpers
import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", helpers.golden_names("clue2d"))
def test_golden_host_path(dev, name):
    ins, want, extra = helpers.load_case("clue2d", name)
    dc, rhoc, odf = (float(v) for v in extra["params"])
    print(f"case clue2d host {name} n={ins['x'].size}", flush=True)
    got = dev.clue2d_host(ins, clue_hip.params2d(dc, rhoc, odf))
    helpers.assert_bit_exact(got, want, helpers.OUT2D, name)


@pytest.mark.parametrize("name", helpers.golden_names("clue2d"))
def test_golden_device_path(dev, name):
    ins, want, extra = helpers.load_case("clue2d", name)
    dc, rhoc, odf = (float(v) for v in extra["params"])
    n = ins["x"].size
    print(f"case clue2d device {name} n={n}", flush=True)
    din = dev.upload_soa(ins, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(n, clue_hip.FIELDS_2D_OUT)
    dev.clue2d_device(din, dout, n, clue_hip.params2d(dc, rhoc, odf))
    dev.check()
    helpers.assert_bit_exact(dev.download_soa(dout), want, helpers.OUT2D, name)
    for a in list(din.values()) + list(dout.values()):
        a.free()


def test_one_million_hit_event_vs_oracle(dev, oracle_lib):
    ev = synth.clue2d_event(1_000_000, 0)
    print("case clue2d 1e6-hit event", flush=True)
    got = dev.clue2d_host(ev)
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"])
    helpers.assert_bit_exact(got, want, helpers.OUT2D, "1e6")
    helpers.check_invariants_2d(ev, got, 3.0, 10.0, 2.0)


def test_dense_column_of_70000_hits(dev, oracle_lib):
    """70,000 hits in one 5 cm column of one layer (about 1400 candidates per hit, bin-table runs far
    beyond 16 bits), plus a C1-model event on other layers."""
    rng = np.random.Generator(np.random.PCG64(77))
    n = 70_000
    base = synth.clue2d_event(2_000, 77)
    base["layer"] = (base["layer"] % 50 + 50).astype(np.int32)
    ev = {"x": np.concatenate([rng.uniform(0.5, 4.5, n).astype(np.float32), base["x"]]),
          "y": np.concatenate([rng.uniform(-249.0, 249.0, n).astype(np.float32), base["y"]]),
          "layer": np.concatenate([np.full(n, 7, np.int32), base["layer"]]),
          "weight": np.concatenate([rng.exponential(1.0, n).astype(np.float32), base["weight"]])}
    print(f"case clue2d dense column n={ev['x'].size}", flush=True)
    got = dev.clue2d_host(ev)
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"])
    helpers.assert_bit_exact(got, want, helpers.OUT2D, "dense column")


def mixed_batch():
    sizes = [1000, 0, 1, 23_456, 150_000, 3_333, 70_000, 2_048, 2_049, 120]
    evs = [synth.clue2d_event(n, 500 + i, shuffle=(i % 3 == 0)) for i, n in enumerate(sizes)]
    return synth.pack(evs)


def test_batch_equals_per_event_equals_oracle(dev, oracle_lib):
    packed, offs = mixed_batch()
    N = int(offs[-1])
    print(f"case clue2d mixed batch events={offs.size - 1} n={N}", flush=True)
    din = dev.upload_soa(packed, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_2D_OUT)
    dev.clue2d_batch(din, dout, offs)
    dev.check()
    batch = dev.download_soa(dout)
    want = oracle_lib.clue2d_batch(packed, offs)
    helpers.assert_bit_exact(batch, want, helpers.OUT2D, "batch vs oracle")
    # each event alone on the device gives the same bits as in the mix
    for e in range(offs.size - 1):
        a, b = int(offs[e]), int(offs[e + 1])
        if b == a:
            continue
        one = dev.clue2d_host({k: packed[k][a:b] for k in clue_hip.FIELDS_2D_IN})
        helpers.assert_bit_exact(one, {k: v[a:b] for k, v in batch.items()}, helpers.OUT2D, f"event {e} alone")


def test_plan_graph_replays_are_identical(dev, oracle_lib):
    packed, offs = mixed_batch()
    N = int(offs[-1])
    print(f"case clue2d graph replay n={N}", flush=True)
    din = dev.upload_soa(packed, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_2D_OUT)
    want = oracle_lib.clue2d_batch(packed, offs)
    for use_graph in (False, True):
        plan = dev.clue2d_plan(din, dout, offs, use_graph=use_graph)
        for rep in range(3):
            for a in dout.values():
                a.upload(np.full(N, -7, a.dtype))
            plan.launch()
            dev.check()
            helpers.assert_bit_exact(dev.download_soa(dout), want, helpers.OUT2D, f"graph={use_graph} replay {rep}")
        plan.destroy()


def test_invalid_layer_is_reported(dev):
    ev = synth.clue2d_event(5000, 9)
    ev["layer"][17] = 100
    with pytest.raises(clue_hip.ClueError) as e:
        dev.clue2d_host(ev)
    assert e.value.code == -4
    ev["layer"][17] = 3
    dev.clue2d_host(ev)  # the context stays usable


def test_c3_batch_properties_full_size(dev, oracle_lib):
    counts = synth.c3_hit_counts(256)
    packed, offs = synth.clue2d_batch(counts)
    N = int(offs[-1])
    print(f"case clue2d 256-event batch n={N}", flush=True)
    din = dev.upload_soa(packed, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_2D_OUT)
    plan = dev.clue2d_plan(din, dout, offs, use_graph=True)
    plan.launch()
    dev.check()
    out = dev.download_soa(dout)
    for e in range(offs.size - 1):
        a, b = int(offs[e]), int(offs[e + 1])
        helpers.check_invariants_2d({k: v[a:b] for k, v in packed.items()}, {k: v[a:b] for k, v in out.items()},
                                    3.0, 10.0, 2.0)
    for e in (0, int(np.argmin(counts)), int(np.argmax(counts)), 128):
        a, b = int(offs[e]), int(offs[e + 1])
        want = oracle_lib.clue2d(packed["x"][a:b], packed["y"][a:b], packed["layer"][a:b], packed["weight"][a:b])
        helpers.assert_bit_exact({k: v[a:b] for k, v in out.items()}, want, helpers.OUT2D, f"c3 event {e}")
    plan.destroy()


@pytest.mark.parametrize("params", helpers.EXTREME_PARAMS)
def test_extreme_coordinates_and_distances(dev, oracle_lib, params):
    """NaN, infinite and out-of-grid coordinates, zero / negative / NaN / very large distances, alone and
    batched beside an ordinary 50000-hit event (staged and unstaged neighbourhoods in one launch)."""
    dc, rhoc, odf = params
    p = clue_hip.params2d(dc, rhoc, odf)
    ev = helpers.extreme_event()
    print(f"case clue2d extremes {params} n={ev['x'].size}", flush=True)
    got = dev.clue2d_host(ev, p)
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], dc, rhoc, odf)
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"extremes {params}")
    packed, offs = synth.pack([ev, synth.clue2d_event(50000, 3), helpers.extreme_event(seed=6)])
    N = int(offs[-1])
    din = dev.upload_soa(packed, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_2D_OUT)
    dev.clue2d_batch(din, dout, offs, p)
    dev.check()
    got = dev.download_soa(dout)
    for a in list(din.values()) + list(dout.values()):
        a.free()
    helpers.assert_bit_exact(got, oracle_lib.clue2d_batch(packed, offs, dc, rhoc, odf), helpers.OUT2D,
                             f"extremes batched {params}")


def test_batch_of_68_million_hits(dev, oracle_lib):
    """68 copies of one 1e6-hit event: 4.26e6 radix histogram counters, more partial sums than the
    two-launch scan takes, so the radix passes run the three-launch scan; every copy must equal the
    oracle's result for the event alone."""
    ev = synth.clue2d_event(1_000_000, 11)
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"])
    E, n = 68, ev["x"].size
    packed = {k: np.tile(v, E) for k, v in ev.items()}
    offs = np.arange(E + 1, dtype=np.int64) * n
    print(f"case clue2d 68-copy batch n={E * n}", flush=True)
    din = dev.upload_soa(packed, clue_hip.FIELDS_2D_IN)
    del packed
    dout = dev.alloc_soa(E * n, clue_hip.FIELDS_2D_OUT)
    dev.clue2d_batch(din, dout, offs)
    dev.check()
    for k in helpers.OUT2D:
        got = dout[k].download().reshape(E, n)
        w = np.asarray(want[k]).reshape(1, n)
        if got.dtype == np.float32:
            got, w = got.view(np.int32), w.view(np.int32)
        bad = np.nonzero((got != w).any(axis=1))[0]
        assert bad.size == 0, f"{k}: copies {bad[:8]} differ from the oracle"
    for a in list(din.values()) + list(dout.values()):
        a.free()
