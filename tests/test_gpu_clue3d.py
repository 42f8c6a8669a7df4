"""CLUE3D on the GPU, bit-exact against the reference fixtures and the CPU oracle."""
import numpy as np
import pytest

import clue_hip
import helpers
import oracle
import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", helpers.golden_names("clue3d"))
def test_golden_host_path(dev, name):
    ins, want, extra = helpers.load_case("clue3d", name)
    print(f"case clue3d host {name} n={ins['layer'].size}", flush=True)
    got = dev.clue3d_host(ins, clue_hip.params3d(S=int(extra["S"])))
    helpers.assert_bit_exact(got, want, helpers.OUT3D, name)


@pytest.mark.parametrize("name", helpers.golden_names("clue3d"))
def test_golden_device_path(dev, name):
    ins, want, extra = helpers.load_case("clue3d", name)
    n = ins["layer"].size
    print(f"case clue3d device {name} n={n}", flush=True)
    din = dev.upload_soa(ins, clue_hip.FIELDS_3D_IN)
    dout = dev.alloc_soa(n, clue_hip.FIELDS_3D_OUT)
    dev.clue3d_device(din, dout, n, clue_hip.params3d(S=int(extra["S"])))
    dev.check()
    helpers.assert_bit_exact(dev.download_soa(dout), want, helpers.OUT3D, name)
    for a in list(din.values()) + list(dout.values()):
        a.free()


def test_batch_equals_per_event_equals_oracle(dev, oracle_lib):
    evs = [synth.clue3d_event(i) for i in range(6)] + [synth.clue3d_event(7, n_layers=94, mean_per_layer=30)]
    packed, offs = synth.pack(evs)
    N = int(offs[-1])
    print(f"case clue3d batch n={N}", flush=True)
    din = dev.upload_soa(packed, clue_hip.FIELDS_3D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_3D_OUT)
    for S in (2, 3):
        dev.clue3d_batch(din, dout, offs, clue_hip.params3d(S=S))
        dev.check()
        batch = dev.download_soa(dout)
        want = oracle_lib.clue3d_batch(packed, offs, oracle.clue3d_params(S=S))
        helpers.assert_bit_exact(batch, want, helpers.OUT3D, f"batch S={S}")
        for e in range(offs.size - 1):
            a, b = int(offs[e]), int(offs[e + 1])
            one = dev.clue3d_host({k: v[a:b] for k, v in packed.items()}, clue_hip.params3d(S=S))
            helpers.assert_bit_exact(one, {k: v[a:b] for k, v in batch.items()}, helpers.OUT3D, f"event {e}")


def test_plan_graph_c4_properties(dev, oracle_lib):
    packed, offs = synth.clue3d_batch(512)
    N = int(offs[-1])
    print(f"case clue3d 512-event batch n={N}", flush=True)
    din = dev.upload_soa(packed, clue_hip.FIELDS_3D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_3D_OUT)
    plan = dev.clue3d_plan(din, dout, offs, clue_hip.params3d(S=2), use_graph=True)
    for rep in range(2):
        plan.launch()
        dev.check()
        out = dev.download_soa(dout)
        for e in (0, 255, 511):
            a, b = int(offs[e]), int(offs[e + 1])
            want = oracle_lib.clue3d({k: v[a:b] for k, v in packed.items()}, oracle.clue3d_params(S=2))
            helpers.assert_bit_exact({k: v[a:b] for k, v in out.items()}, want, helpers.OUT3D, f"c4 event {e}")
    for e in range(0, 512, 37):
        a, b = int(offs[e]), int(offs[e + 1])
        helpers.check_invariants_3d({k: v[a:b] for k, v in packed.items()}, {k: v[a:b] for k, v in out.items()})
    plan.destroy()


def test_autotuned_plan(dev, oracle_lib):
    """An autotuned CLUE3D plan: the tuning launch and the graph replays with the chosen variants are
    bit-exact."""
    packed, offs = synth.clue3d_batch(24)
    N = int(offs[-1])
    print(f"case clue3d autotuned plan n={N}", flush=True)
    p = oracle.clue3d_params(S=2)
    want = oracle_lib.clue3d_batch(packed, offs, p)
    din = dev.upload_soa(packed, clue_hip.FIELDS_3D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_3D_OUT)
    plan = dev.clue3d_plan(din, dout, offs, clue_hip.params3d(S=2), use_graph=True, autotune=True)
    assert plan.tuning == {"density_points": 512, "density_stage": 2048, "nh_points": 512, "nh_stage": 2048,
                           "nh_index_pass": False}
    for rep in range(3):
        for a in dout.values():
            a.upload(np.full(N, -7, a.dtype))
        plan.launch()
        dev.check()
        helpers.assert_bit_exact(dev.download_soa(dout), want, helpers.OUT3D, f"autotuned launch {rep}")
    t = plan.tuning
    assert t["density_points"] in (256, 512, 1024) and t["density_stage"] in (1024, 2048), t
    assert t["nh_points"] in (256, 512, 1024) and t["nh_stage"] in (1024, 2048), t
    print(f"tuning {t}", flush=True)
    plan.destroy()
    for a in list(din.values()) + list(dout.values()):
        a.free()


@pytest.mark.parametrize("kw", [{"S": 3, "densityOnSameLayer": 1, "nearestHigherOnSameLayer": 1},
                                {"S": 0}, {"S": 40}, {"S": 1, "nearestHigherOnSameLayer": 1}])
def test_parameters(dev, oracle_lib, kw):
    """Same-layer density / nearest higher, no sibling layers, and a sibling reach beyond the staged
    run lists (every block scans its windows from global memory)."""
    ev = synth.clue3d_event(7, n_layers=60, mean_per_layer=60.0)
    print(f"case clue3d parameters {kw} n={ev['layer'].size}", flush=True)
    got = dev.clue3d_host(ev, clue_hip.params3d(**kw))
    helpers.assert_bit_exact(got, oracle_lib.clue3d(ev, oracle.clue3d_params(**kw)), helpers.OUT3D, f"{kw}")


def test_dense_row_of_70000_clusters(dev, oracle_lib):
    """70,000 layer clusters in one (layer, eta) row (about 2800 candidates per cluster, blocks too wide
    to stage), plus a C4-model event."""
    rng = np.random.Generator(np.random.PCG64(78))
    n = 70_000
    eta = rng.uniform(1.56, 1.59, n).astype(np.float32)
    phi = rng.uniform(-np.pi, np.pi, n).astype(np.float32)
    layer = np.full(n, 12, np.int32)
    z = (320.0 + 1.2 * layer.astype(np.float64)).astype(np.float32)
    roz = (1.0 / np.sinh(eta.astype(np.float64))).astype(np.float32)
    rr = z.astype(np.float64) * roz.astype(np.float64)
    dense = {"x": (rr * np.cos(phi)).astype(np.float32), "y": (rr * np.sin(phi)).astype(np.float32), "z": z,
             "eta": eta, "phi": phi, "r_over_absz": roz, "radius": np.full(n, 1.8, np.float32), "layer": layer,
             "energy": rng.exponential(2.0, n).astype(np.float32), "isSilicon": np.ones(n, np.int32)}
    other = synth.clue3d_event(9)
    ev = {k: np.concatenate([dense[k], other[k]]) for k in other}
    print(f"case clue3d dense row n={ev['layer'].size}", flush=True)
    got = dev.clue3d_host(ev, clue_hip.params3d(S=2))
    helpers.assert_bit_exact(got, oracle_lib.clue3d(ev, oracle.clue3d_params(S=2)), helpers.OUT3D, "dense row")


def test_invalid_layer_is_reported(dev):
    ev = synth.clue3d_event(2)
    ev["layer"][5] = 94
    with pytest.raises(clue_hip.ClueError) as e:
        dev.clue3d_host(ev)
    assert e.value.code == -4
