"""The library's CLUE2D and CLUE3D kernels run on the CPU through an emulation of the HIP execution model (one
thread per lane, real barriers, wave ballots and shuffles through a per-wave mailbox,
AddressSanitizer + UndefinedBehaviorSanitizer), with the library's own launch sequences and grids,
bit-exact against the golden fixtures and the oracle:

* the global-sort path (radix passes, bin table, density, nearest higher, seed ranks, labels), the
  path the device runs; device scratch starts poisoned, so reads of unwritten memory show up;
* the CLUE3D path (radix passes, bin table, density and nearest higher over the sibling layers
  with the staged run-list walk and the unstaged row scan, seed ranks, labels);
* the fused segment path (layer bucket pass, segment lists, fused segment kernels, seed numbering),
  which the library keeps disabled on the device (clue_set_path refuses it) until it has also
  passed on the GPU."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import helpers
import oracle
import synth

EMU = Path(__file__).resolve().parent / "cpu_emu"
EXE = EMU / "_build" / "fused2d_emu"
EXE_GLOBAL = EMU / "_build" / "global2d_emu"
EXE_3D = EMU / "_build" / "global3d_emu"
# the ThreadSanitizer builds report every race except the one k_labels intends (tsan.supp)
ENV = {**os.environ, "TSAN_OPTIONS": f"suppressions={EMU / 'tsan.supp'} history_size=4"}


@pytest.fixture(scope="module")
def emulator():
    r = subprocess.run(["make", "-s", "-C", str(EMU)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return EXE


def run_emulated(exe, tmp_path, ev, offs, dc=3.0, rhoc=10.0, odf=2.0, wgs=2, extra=()):
    N = int(offs[-1])
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(np.int32(offs.size - 1).tobytes())
        f.write(np.array([dc, rhoc, odf], np.float32).tobytes())
        f.write(np.asarray(offs, np.int64).tobytes())
        for k, dt in (("x", np.float32), ("y", np.float32), ("layer", np.int32), ("weight", np.float32)):
            f.write(np.ascontiguousarray(ev[k], dt).tobytes())
    fused = Path(exe).name.startswith("fused2d")
    args = [str(exe), str(src), str(dst)] + ([str(wgs)] if fused else [str(a) for a in extra])
    r = subprocess.run(args, capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    d = np.fromfile(dst, np.int32)
    o = d[2:].reshape(5, N)
    out = {"rho": o[0].view(np.float32), "delta": o[1].view(np.float32), "nearestHigher": o[2], "isSeed": o[3],
           "clusterIndex": o[4]}
    return int(d[0]), int(d[1]), out


@pytest.mark.parametrize("name", ["one_hit", "two_layers", "duplicates", "grid_ties", "dc_boundary", "bin_edges",
                                  "dense_bin"])
def test_emulated_fused_golden(emulator, tmp_path, name):
    ins, want, extra = helpers.load_case("clue2d", name)
    dc, rhoc, odf = (float(v) for v in extra["params"])
    err, _, got = run_emulated(emulator, tmp_path, ins, np.array([0, ins["x"].size]), dc, rhoc, odf)
    assert err == 0
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"emulated {name}")


def class_batch():
    """Several events (one empty) whose layers fall in all three fused size classes and on their edges."""
    rng = np.random.Generator(np.random.PCG64(41))
    evs = []
    for n_layers, per in ((3, [1, 2, 1023]), (3, [1024, 1025, 7]), (2, [4096, 4097]), (0, [])):
        xs, ys, ls = [], [], []
        for l, n in enumerate(per):
            c = rng.uniform(-60, 60, (max(1, n // 20), 2))
            p = c[rng.integers(0, c.shape[0], n)] + rng.normal(0, 1.0, (n, 2))
            xs.append(p[:, 0]); ys.append(p[:, 1]); ls.append(np.full(n, 3 * l + 1))
        if not per:
            evs.append({"x": np.zeros(0, np.float32), "y": np.zeros(0, np.float32),
                        "layer": np.zeros(0, np.int32), "weight": np.zeros(0, np.float32)})
            continue
        x = np.concatenate(xs).astype(np.float32)
        perm = rng.permutation(x.size)
        evs.append({"x": x[perm], "y": np.concatenate(ys).astype(np.float32)[perm],
                    "layer": np.concatenate(ls).astype(np.int32)[perm],
                    "weight": rng.exponential(1.0, x.size).astype(np.float32)})
    return synth.pack(evs)


def test_emulated_fused_batch_and_classes(emulator, tmp_path, oracle_lib):
    """Several events (one empty) whose layers fall in all three size classes and on their edges."""
    packed, offs = class_batch()
    err, big, got = run_emulated(emulator, tmp_path, packed, offs)
    assert err == 0 and big == 4097
    want = oracle_lib.clue2d_batch(packed, offs)
    helpers.assert_bit_exact(got, want, helpers.OUT2D, "emulated batch")


def test_emulated_capacity_is_flagged(emulator, tmp_path):
    """A layer above the fused capacity raises the capacity error bit and is left to the global path."""
    n = 10113
    rng = np.random.Generator(np.random.PCG64(42))
    ev = {"x": rng.uniform(-250, 250, n).astype(np.float32), "y": rng.uniform(-250, 250, n).astype(np.float32),
          "layer": np.zeros(n, np.int32), "weight": np.ones(n, np.float32)}
    err, big, _ = run_emulated(emulator, tmp_path, ev, np.array([0, n]))
    assert big == n and err & 4


@pytest.mark.parametrize("name", ["grid_ties", "bin_edges"])
def test_emulated_global_golden(emulator, tmp_path, name):
    ins, want, extra = helpers.load_case("clue2d", name)
    dc, rhoc, odf = (float(v) for v in extra["params"])
    err, _, got = run_emulated(EXE_GLOBAL, tmp_path, ins, np.array([0, ins["x"].size]), dc, rhoc, odf)
    assert err == 0
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"emulated global {name}")


def test_emulated_global_batch(emulator, tmp_path, oracle_lib):
    """Three events (one empty, one spanning two sort blocks) in one batch."""
    evs = [synth.clue2d_event(n, 70 + i) for i, n in enumerate([2500, 0, 700])]
    packed, offs = synth.pack(evs)
    err, _, got = run_emulated(EXE_GLOBAL, tmp_path, packed, offs)
    assert err == 0
    helpers.assert_bit_exact(got, oracle_lib.clue2d_batch(packed, offs), helpers.OUT2D, "emulated global batch")


def large_layer_event(seed=43):
    """One event with a layer in each size class of k_segsort2d: 12000 points on layer 3 (above the
    10240 it keeps in LDS), 8000 on layer 50 (the middle class, more than the 5120 whose records it
    keeps in registers), both with one densely filled bin, and 300 on layers 0 and 99."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xs, ys, ls = [], [], []
    for lay, n, dense in ((3, 12000, 2000), (50, 8000, 700), (0, 300, 0), (99, 300, 0)):
        xs += [rng.uniform(-250, 250, n - dense), rng.uniform(10.0, 14.9, dense)]
        ys += [rng.uniform(-250, 250, n - dense), rng.uniform(-4.9, -0.1, dense)]
        ls.append(np.full(n, lay))
    x = np.concatenate(xs).astype(np.float32)
    y = np.concatenate(ys).astype(np.float32)
    layer = np.concatenate(ls).astype(np.int32)
    perm = rng.permutation(x.size)
    return {"x": x[perm], "y": y[perm], "layer": layer[perm], "weight": rng.exponential(1.0, x.size).astype(np.float32)}


@pytest.mark.parametrize("exe", ["global2d_emu", "global2d_emu_ldsz", "global2d_emu_tsan"])
def test_emulated_global_large_layer(emulator, tmp_path, oracle_lib, exe):
    """Layers of every size class of the per-layer bin sort, the largest on its global-slot path."""
    ev = large_layer_event()
    err, _, got = run_emulated(EMU / "_build" / exe, tmp_path, ev, np.array([0, ev["x"].size]))
    assert err == 0
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"])
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"emulated large layer {exe}")


def test_emulated_global_wide_layer(emulator, tmp_path, oracle_lib):
    """A layer of 70,000 points: its bin-table entries exceed 16-bit offsets from the layer's start and
    take the 32-bit table, next to a small layer that keeps the 16-bit one."""
    rng = np.random.Generator(np.random.PCG64(44))
    n = 70000
    ev = {"x": np.concatenate([rng.uniform(-250, 250, n), rng.uniform(-250, 250, 500)]).astype(np.float32),
          "y": np.concatenate([rng.uniform(-250, 250, n), rng.uniform(-250, 250, 500)]).astype(np.float32),
          "layer": np.concatenate([np.full(n, 7), np.full(500, 8)]).astype(np.int32)}
    ev["weight"] = rng.exponential(1.0, ev["x"].size).astype(np.float32)
    err, _, got = run_emulated(EXE_GLOBAL, tmp_path, ev, np.array([0, ev["x"].size]))
    assert err == 0
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"])
    helpers.assert_bit_exact(got, want, helpers.OUT2D, "emulated wide layer")


@pytest.mark.parametrize("variants", [(1, 1), (2, 3), (3, 2), (4, 5), (5, 4)])
def test_emulated_global_variants(emulator, tmp_path, variants):
    """Every (points per workgroup, LDS staging) variant of density and nearest higher gives the same bits."""
    ins, want, extra = helpers.load_case("clue2d", "bin_edges")
    dc, rhoc, odf = (float(v) for v in extra["params"])
    err, _, got = run_emulated(EXE_GLOBAL, tmp_path, ins, np.array([0, ins["x"].size]), dc, rhoc, odf,
                               extra=variants)
    assert err == 0
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"emulated variants {variants}")


@pytest.mark.parametrize("exe", ["global2d_emu", "global2d_emu_ldsz", "global2d_emu_tsan"])
@pytest.mark.parametrize("variants", [(0, 0), (1, 1), (4, 5)])
def test_emulated_index_pass(emulator, tmp_path, oracle_lib, exe, variants):
    """The nearest-higher step's index-pass output form (one index-order record per point, tile-order
    follower links, outputs copied to the points' indices by the labels step) gives the oracle's bits:
    a three-event batch with an empty event and hits in random order, and the layer-ordered golden
    'bin_edges' case, under AddressSanitizer, with LDS redzones and under ThreadSanitizer."""
    evs = [synth.clue2d_event(n, 90 + i, shuffle=True) for i, n in enumerate([3000, 0, 900])]
    packed, offs = synth.pack(evs)
    err, _, got = run_emulated(EMU / "_build" / exe, tmp_path, packed, offs, extra=variants + (1,))
    assert err == 0
    helpers.assert_bit_exact(got, oracle_lib.clue2d_batch(packed, offs), helpers.OUT2D, f"index pass {exe}")
    ins, want, extra = helpers.load_case("clue2d", "bin_edges")
    dc, rhoc, odf = (float(v) for v in extra["params"])
    err, _, got = run_emulated(EMU / "_build" / exe, tmp_path, ins, np.array([0, ins["x"].size]), dc, rhoc, odf,
                               extra=variants + (1,))
    assert err == 0
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"index pass bin_edges {exe}")
    ev = helpers.extreme_event(n_normal=200)
    for dc, rhoc, odf in (helpers.EXTREME_PARAMS[4], helpers.EXTREME_PARAMS[5]):
        err, _, got = run_emulated(EMU / "_build" / exe, tmp_path, ev, np.array([0, ev["x"].size]), dc, rhoc, odf,
                                   extra=variants + (1,))
        assert err == 0
        want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], dc, rhoc, odf)
        helpers.assert_bit_exact(got, want, helpers.OUT2D, f"index pass extremes {exe} {dc}")


@pytest.mark.parametrize("name", helpers.golden_names("clue2d"))
def test_emulated_index_pass_golden(emulator, tmp_path, name):
    """Every CLUE2D golden case through the index-pass output form."""
    ins, want, extra = helpers.load_case("clue2d", name)
    dc, rhoc, odf = (float(v) for v in extra["params"])
    err, _, got = run_emulated(EXE_GLOBAL, tmp_path, ins, np.array([0, ins["x"].size]), dc, rhoc, odf,
                               extra=(0, 0, 1))
    assert err == 0
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"index pass {name}")


@pytest.mark.parametrize("params", [helpers.EXTREME_PARAMS[k] for k in (4, 5)])
def test_emulated_global_extremes(emulator, tmp_path, oracle_lib, params):
    """NaN, infinite and out-of-grid coordinates with a NaN and a very large distance:
    the staging reach and the table columns staged per block must still hold every lookup (the GPU
    test runs every parameter set)."""
    ev = helpers.extreme_event(n_normal=200)
    dc, rhoc, odf = params
    err, _, got = run_emulated(EXE_GLOBAL, tmp_path, ev, np.array([0, ev["x"].size]), dc, rhoc, odf)
    assert err == 0
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], dc, rhoc, odf)
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"emulated extremes {params}")


# ------------------------------------------------------------------ CLUE3D

FIELDS_3D_F = ("z", "eta", "phi", "r_over_absz", "radius", "energy")


def run_emulated3d(tmp_path, ev, offs, params, variants=(), exe=EXE_3D):
    N = int(offs[-1])
    src, dst = tmp_path / "in3.bin", tmp_path / "out3.bin"
    with open(src, "wb") as f:
        f.write(np.int32(len(offs) - 1).tobytes())
        f.write(bytes(params))
        f.write(np.asarray(offs, np.int64).tobytes())
        for k in FIELDS_3D_F:
            f.write(np.ascontiguousarray(ev[k], np.float32).tobytes())
        for k in ("layer", "isSilicon"):
            f.write(np.ascontiguousarray(ev[k], np.int32).tobytes())
    r = subprocess.run([str(exe), str(src), str(dst)] + [str(v) for v in variants], capture_output=True,
                       text=True, timeout=900, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    d = np.fromfile(dst, np.int32)
    o = d[1:].reshape(7, N)
    out = {"rho": o[0].view(np.float32), "delta_f": o[1].view(np.float32), "delta_l": o[2], "nh_layer": o[3],
           "nh_idx": o[4], "isSeed": o[5], "tracksterIndex": o[6]}
    return int(d[0]), out


@pytest.mark.parametrize("name", ["single", "phi_wrap", "eta_row34", "side_split", "scintillator", "ties3d",
                                  "grid3d_ties", "c4_event1_S3"])
def test_emulated_clue3d_golden(emulator, tmp_path, name):
    ins, want, extra = helpers.load_case("clue3d", name)
    err, got = run_emulated3d(tmp_path, ins, np.array([0, ins["layer"].size]),
                              oracle.clue3d_params(S=int(extra["S"])))
    assert err == 0
    helpers.assert_bit_exact(got, want, helpers.OUT3D, f"emulated clue3d {name}")


def test_emulated_clue3d_batch(emulator, tmp_path, oracle_lib):
    """Three events of the C4 model at lower occupancy (one empty) in one batch, density over +/-2
    sibling layers."""
    evs = [synth.clue3d_event(e, n_layers=30, mean_per_layer=80.0) for e in range(2)]
    evs.insert(1, {k: v[:0] for k, v in evs[0].items()})
    packed, offs = synth.pack(evs)
    p = oracle.clue3d_params(S=2)
    err, got = run_emulated3d(tmp_path, packed, offs, p)
    assert err == 0
    helpers.assert_bit_exact(got, oracle_lib.clue3d_batch(packed, offs, p), helpers.OUT3D, "emulated clue3d batch")


@pytest.mark.parametrize("variants", [(1, 2), (2, 3), (3, 1)])
def test_emulated_clue3d_variants(emulator, tmp_path, oracle_lib, variants):
    """Every (points per workgroup, LDS staging) variant gives the same bits; the 1024-point stage
    leaves most blocks of this event unstaged, so the row scan from global memory runs too."""
    ev = synth.clue3d_event(5, n_layers=25)
    p = oracle.clue3d_params(S=2)
    err, got = run_emulated3d(tmp_path, ev, np.array([0, ev["layer"].size]), p, variants)
    assert err == 0
    helpers.assert_bit_exact(got, oracle_lib.clue3d(ev, p), helpers.OUT3D, f"emulated clue3d variants {variants}")


@pytest.mark.parametrize("kw", [{"S": 3, "densityOnSameLayer": 1, "nearestHigherOnSameLayer": 1},
                                {"S": 0}, {"S": 40}, {"S": 1, "nearestHigherOnSameLayer": 1}])
def test_emulated_clue3d_params(emulator, tmp_path, oracle_lib, kw):
    """Same-layer density / nearest higher, no sibling layers, and a reach beyond the run list's layer
    field (every block scans from global memory)."""
    ev = synth.clue3d_event(7, n_layers=60, mean_per_layer=60.0)
    p = oracle.clue3d_params(**kw)
    err, got = run_emulated3d(tmp_path, ev, np.array([0, ev["layer"].size]), p)
    assert err == 0
    helpers.assert_bit_exact(got, oracle_lib.clue3d(ev, p), helpers.OUT3D, f"emulated clue3d {kw}")


def test_emulated_clue3d_bad_phi_is_flagged(emulator, tmp_path, oracle_lib):
    """NaN and infinite phi have no tile bin: the run flags kErrBinRange, stays inside every buffer
    (AddressSanitizer, LDS index checks) and leaves the other events of the batch exact."""
    evs = [synth.clue3d_event(e, n_layers=20, mean_per_layer=60.0) for e in range(3)]
    bad = {k: v.copy() for k, v in evs[1].items()}
    bad["phi"][[0, 5, 17]] = np.array([np.nan, np.inf, -np.inf], np.float32)
    packed, offs = synth.pack([evs[0], bad, evs[2]])
    p = oracle.clue3d_params(S=2)
    err, got = run_emulated3d(tmp_path, packed, offs, p)
    assert err & 2
    for e in (0, 2):
        a, b = int(offs[e]), int(offs[e + 1])
        helpers.assert_bit_exact({k: v[a:b] for k, v in got.items()}, oracle_lib.clue3d(evs[e], p), helpers.OUT3D,
                                 f"event {e} beside a bad-phi event")


# ------------------------------------------------------------------ sanitizer builds

SANITIZED = ("ldsz", "tsan")  # LDS arrays with redzones (AddressSanitizer); ThreadSanitizer


def sanitized(name, variant):
    return EMU / "_build" / f"{name}_{variant}"


@pytest.mark.parametrize("variant", SANITIZED)
def test_sanitized_fused(emulator, tmp_path, oracle_lib, variant):
    """The fused segment kernels with every LDS array bounds-checked, and under ThreadSanitizer (a
    barrier missing between two lanes' accesses to one LDS word is a reported race), on the size-class
    batch and on a single event of four 1500-point layers, each run with two workgroups per class."""
    packed, offs = class_batch()
    err, big, got = run_emulated(sanitized("fused2d_emu", variant), tmp_path, packed, offs)
    assert err == 0 and big == 4097
    helpers.assert_bit_exact(got, oracle_lib.clue2d_batch(packed, offs), helpers.OUT2D, f"fused {variant}")
    rng = np.random.Generator(np.random.PCG64(5))
    n = 6000
    ev = {"x": rng.uniform(-250, 250, n).astype(np.float32), "y": rng.uniform(-250, 250, n).astype(np.float32),
          "layer": np.repeat(np.arange(4, dtype=np.int32), n // 4)[rng.permutation(n)],
          "weight": rng.exponential(1.0, n).astype(np.float32)}
    err, big, got = run_emulated(sanitized("fused2d_emu", variant), tmp_path, ev, np.array([0, n]))
    assert err == 0 and big == 1500
    helpers.assert_bit_exact(got, oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"]), helpers.OUT2D,
                             f"fused 4 x 1500 {variant}")


@pytest.mark.parametrize("variant", SANITIZED)
def test_sanitized_global2d(emulator, tmp_path, oracle_lib, variant):
    """The global-sort path's kernels, LDS bounds-checked and under ThreadSanitizer: a three-event batch and
    the extreme-coordinate event with a very large distance."""
    evs = [synth.clue2d_event(n, 80 + i) for i, n in enumerate([2500, 0, 700])]
    packed, offs = synth.pack(evs)
    err, _, got = run_emulated(sanitized("global2d_emu", variant), tmp_path, packed, offs)
    assert err == 0
    helpers.assert_bit_exact(got, oracle_lib.clue2d_batch(packed, offs), helpers.OUT2D, f"global {variant}")
    ev = helpers.extreme_event(n_normal=200)
    dc, rhoc, odf = helpers.EXTREME_PARAMS[5]
    err, _, got = run_emulated(sanitized("global2d_emu", variant), tmp_path, ev, np.array([0, ev["x"].size]),
                               dc, rhoc, odf)
    assert err == 0
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], dc, rhoc, odf)
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"global extremes {variant}")


@pytest.mark.parametrize("variant", SANITIZED)
def test_sanitized_clue3d(emulator, tmp_path, oracle_lib, variant):
    """The CLUE3D kernels, LDS bounds-checked and under ThreadSanitizer, on a two-event batch (density over
    +/-2 sibling layers) and with a reach beyond the run list's layer field."""
    evs = [synth.clue3d_event(e, n_layers=30, mean_per_layer=80.0) for e in range(2)]
    packed, offs = synth.pack(evs)
    for p in (oracle.clue3d_params(S=2), oracle.clue3d_params(S=40)):
        err, got = run_emulated3d(tmp_path, packed, offs, p, exe=sanitized("global3d_emu", variant))
        assert err == 0
        helpers.assert_bit_exact(got, oracle_lib.clue3d_batch(packed, offs, p), helpers.OUT3D, f"clue3d {variant}")


EXE_SCAN = EMU / "_build" / "scan_emu"


@pytest.mark.parametrize("n,direct_max", [(0, 2048), (1, 2048), (2047, 2048), (2049, 2048), (2048 * 12 + 3, 2048),
                                          (2048 * 12 + 3, 0), (2048 * 6, 6), (2048 * 6 + 1, 6)])
def test_emulated_scan(emulator, tmp_path, n, direct_max):
    """The device-wide exclusive scan, in place and out of place, in its two-launch form (each block
    sums the partials before it) and its three-launch form (a single-block scan of the partials), with
    the partials buffer poisoned, against numpy's cumulative sum."""
    v = np.random.Generator(np.random.PCG64(n)).integers(-3, 50, n).astype(np.int32)
    src, dst = tmp_path / "scan_in.bin", tmp_path / "scan_out.bin"
    with open(src, "wb") as f:
        f.write(np.int64(n).tobytes())
        f.write(v.tobytes())
    r = subprocess.run([str(EXE_SCAN), str(src), str(dst), str(direct_max)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    d = np.fromfile(dst, np.int32)
    want = (np.cumsum(v, dtype=np.int64) - v).astype(np.int32)
    assert np.array_equal(d[:n], want)
    assert np.array_equal(d[n:], want)
