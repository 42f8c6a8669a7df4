"""The library's CLUE2D kernels run on the CPU through an emulation of the HIP execution model (one
thread per lane, real barriers, wave ballots and shuffles through a per-wave mailbox,
AddressSanitizer + UndefinedBehaviorSanitizer), with the library's own launch sequences and grids,
bit-exact against the golden fixtures and the oracle:

* the global-sort path (radix passes, bin table, density, nearest higher, seed ranks, labels), the
  path the device runs; device scratch starts poisoned, so reads of unwritten memory show up;
* the fused segment path (layer bucket pass, segment lists, fused segment kernels, seed numbering),
  which the library keeps disabled on the device (clue_set_path refuses it) until it has also
  passed on the GPU."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import helpers
import synth

EMU = Path(__file__).resolve().parent / "cpu_emu"
EXE = EMU / "_build" / "fused2d_emu"
EXE_GLOBAL = EMU / "_build" / "global2d_emu"


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
    args = [str(exe), str(src), str(dst)] + ([str(wgs)] if exe == EXE else [str(a) for a in extra])
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
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


def test_emulated_fused_batch_and_classes(emulator, tmp_path, oracle_lib):
    """Several events (one empty) whose layers fall in all three size classes and on their edges."""
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
    packed, offs = synth.pack(evs)
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


@pytest.mark.parametrize("variants", [(1, 1), (2, 3), (3, 2)])
def test_emulated_global_variants(emulator, tmp_path, variants):
    """Every (points per workgroup, LDS staging) variant of density and nearest higher gives the same bits."""
    ins, want, extra = helpers.load_case("clue2d", "bin_edges")
    dc, rhoc, odf = (float(v) for v in extra["params"])
    err, _, got = run_emulated(EXE_GLOBAL, tmp_path, ins, np.array([0, ins["x"].size]), dc, rhoc, odf,
                               extra=variants)
    assert err == 0
    helpers.assert_bit_exact(got, want, helpers.OUT2D, f"emulated variants {variants}")
