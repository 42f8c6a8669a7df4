"""The CLUE2D validator and CSV writer modules on the CPU, run from the plugin library by a small
harness on a product computed by the oracle.

The reference's result files (data/output/reference/ref_<N>k_20_25_2.csv) are not available here, so
each test writes one in the reference writer's format (CLUEOutputProducer.cc:36-42) from the
oracle's output for a synthetic toyDetector layer, at the parameters those files were made with
(dc = 20, rhoc = 25, outlierDeltaFactor = 2). The checks follow the reference's acceptance rule
(CLUEValidator.cc:42-199): 1e-3 absolute on floats, delta clamped at 999, seed-mapped cluster ids,
results only checked at those parameters.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import helpers
import synth

REPO = Path(__file__).resolve().parents[1]
PKG = REPO / "heterogeneous-clue_amd"
HARNESS_SRC = REPO / "tests" / "modules" / "run_module.cc"
HARNESS = REPO / "tests" / "modules" / "_build" / "run_module"
PLUGINS = PKG / "libclue_plugins.so"


@pytest.fixture(scope="module")
def harness():
    if not PLUGINS.exists():
        pytest.skip("plugin library not built")
    HARNESS.parent.mkdir(exist_ok=True)
    # built under a name of this process and renamed into place, so that test processes running side
    # by side never execute a half-written binary
    tmp = HARNESS.with_name(f"{HARNESS.name}.{os.getpid()}")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", str(PKG / "framework"), "-I", str(REPO / "include"),
                    "-o", str(tmp), str(HARNESS_SRC), "-L", str(PKG), "-lclue_framework", "-ldl",
                    f"-Wl,-rpath,{PKG}"], check=True, timeout=120)
    os.replace(tmp, HARNESS)
    return HARNESS


@pytest.fixture(scope="module")
def toy(oracle_lib):
    layer = synth.toy_detector_layer(1000)
    ev = synth.expand_toy_detector(layer)
    out = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], 20.0, 25.0, 2.0)
    assert out["isSeed"].sum() > 50 and (out["clusterIndex"] == -1).any() and (out["delta"] > 999).any()
    return ev, out, helpers.clue2d_csv_lines(ev, out)


def layout(root: Path, lines, name="toyDetector_1k"):
    """data/input/<name>.bin and data/output/reference/ref_1k_20_25_2.csv under root."""
    (root / "data" / "input").mkdir(parents=True, exist_ok=True)
    ref = root / "data" / "output" / "reference"
    ref.mkdir(parents=True, exist_ok=True)
    (ref / "ref_1k_20_25_2.csv").write_text("\n".join(lines) + "\n")
    inp = root / "data" / "input" / f"{name}.bin"
    inp.write_bytes(b"")
    return inp


def write_product(path: Path, ev, out):
    n = ev["x"].size
    with open(path, "wb") as f:
        f.write(np.int32(n).tobytes())
        for a, dt in ((ev["x"], np.float32), (ev["y"], np.float32), (ev["layer"], np.int32),
                      (ev["weight"], np.float32), (out["rho"], np.float32), (out["delta"], np.float32),
                      (out["nearestHigher"], np.int32), (out["isSeed"], np.int32), (out["clusterIndex"], np.int32)):
            f.write(np.ascontiguousarray(a, dtype=dt).tobytes())


def run(harness, tmp_path, ev, out, inp, params=(20, 25, 2), strict=False, outdir=None):
    prod = tmp_path / "product.bin"
    write_product(prod, ev, out)
    args = [str(harness), str(PLUGINS), str(prod), str(inp)] + [str(p) for p in params] + [str(int(strict))]
    if outdir is not None:
        args.append(str(outdir))
    return subprocess.run(args, capture_output=True, text=True, timeout=120)


def copy(d):
    return {k: v.copy() for k, v in d.items()}


def test_accepts_its_own_reference(harness, toy, tmp_path):
    ev, out, lines = toy
    inp = layout(tmp_path, lines)
    for strict in (False, True):
        r = run(harness, tmp_path, ev, out, inp, strict=strict)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "ref_1k_20_25_2.csv" in r.stdout and "Input variables are correct!" in r.stdout
        for field in ("rho", "delta", "nearestHigher", "isSeed", "cluster IDs"):
            assert f"Output {field} -> Ok" in r.stdout
        assert "CLUE output is correct!" in r.stdout
        assert ("identical to the reference file" in r.stdout) == strict


def test_tolerance_is_one_thousandth(harness, toy, tmp_path):
    ev, out, lines = toy
    inp = layout(tmp_path, lines)
    k = int(np.flatnonzero(out["rho"] > 1)[7])
    near = copy(out)
    near["rho"][k] += np.float32(0.0005)
    assert run(harness, tmp_path, ev, near, inp).returncode == 0
    r = run(harness, tmp_path, ev, near, inp, strict=True)
    assert r.returncode == 1 and f"row {k} is" in r.stdout
    far = copy(out)
    far["rho"][k] += np.float32(0.002)
    r = run(harness, tmp_path, ev, far, inp)
    assert r.returncode == 1 and f"rho differs from the reference at i={k}," in r.stdout
    assert "Output x -> Ok" in r.stdout and "Output rho -> Ok" not in r.stdout


def test_delta_is_compared_after_clamping(harness, toy, tmp_path):
    ev, out, lines = toy
    inp = layout(tmp_path, lines)
    big = copy(out)
    k = int(np.flatnonzero(out["delta"] > 999)[0])
    big["delta"][k] = np.float32(5000.0)
    for strict in (False, True):
        assert run(harness, tmp_path, ev, big, inp, strict=strict).returncode == 0
    small = copy(out)
    small["delta"][k] = np.float32(998.0)
    r = run(harness, tmp_path, ev, small, inp)
    assert r.returncode == 1 and f"delta differs from the reference at i={k}," in r.stdout


def test_cluster_ids_are_mapped_through_seeds(harness, toy, tmp_path):
    ev, out, lines = toy
    inp = layout(tmp_path, lines)
    ids = out["clusterIndex"]
    n_clusters = int(ids.max()) + 1
    perm = np.random.default_rng(3).permutation(n_clusters).astype(np.int32)
    relabelled = copy(out)
    relabelled["clusterIndex"] = np.where(ids >= 0, perm[np.maximum(ids, 0)], -1).astype(np.int32)
    r = run(harness, tmp_path, ev, relabelled, inp)
    assert r.returncode == 0 and "Output cluster IDs -> Ok" in r.stdout, r.stdout
    r = run(harness, tmp_path, ev, relabelled, inp, strict=True)
    assert r.returncode == 1 and "(strict)" in r.stdout
    # a follower moved to another cluster is caught
    moved = copy(out)
    k = int(np.flatnonzero((ids >= 0) & (out["isSeed"] == 0))[0])
    moved["clusterIndex"][k] = (ids[k] + 1) % n_clusters
    r = run(harness, tmp_path, ev, moved, inp)
    assert r.returncode == 1 and f"cluster ids differ from the reference at i={k}," in r.stdout
    # an outlier given a cluster id maps to 0 unless that id is seeded (reference: map default)
    outl = copy(out)
    k = int(np.flatnonzero(ids == -1)[0])
    outl["clusterIndex"][k] = 0
    r = run(harness, tmp_path, ev, outl, inp)
    assert r.returncode == 1 and f"i={k}, original=0" in r.stdout


@pytest.mark.parametrize("field,kind", [("x", "in"), ("y", "in"), ("layer", "in"), ("weight", "in"),
                                        ("nearestHigher", "out"), ("isSeed", "out")])
def test_each_field_is_checked(harness, toy, tmp_path, field, kind):
    ev, out, lines = toy
    inp = layout(tmp_path, lines)
    e2, o2 = copy(ev), copy(out)
    tgt = (e2 if kind == "in" else o2)
    key = {"x": "x", "y": "y", "layer": "layer", "weight": "weight"}.get(field, field)
    k = 4321
    if tgt[key].dtype == np.float32:
        tgt[key][k] += np.float32(0.01)
    else:
        tgt[key][k] += 1 if field != "isSeed" else (1 - 2 * tgt[key][k])
    r = run(harness, tmp_path, e2, o2, inp)
    assert r.returncode == 1 and f"{field} differs from the reference at i={k}," in r.stdout


def test_other_parameters_check_inputs_only(harness, toy, tmp_path):
    ev, out, lines = toy
    inp = layout(tmp_path, lines)
    junk = copy(out)
    junk["rho"][:] = 0
    junk["clusterIndex"][:] = 7
    r = run(harness, tmp_path, ev, junk, inp, params=(3, 10, 2))
    assert r.returncode == 0 and "Input variables are correct!" in r.stdout
    assert "cannot check results" in r.stdout and "CLUE output is correct!" not in r.stdout


def test_inputs_without_reference_file(harness, toy, tmp_path):
    ev, out, lines = toy
    inp = layout(tmp_path, lines, name="raw2D")
    r = run(harness, tmp_path, ev, out, inp)
    assert r.returncode == 0 and "There is no reference output data" in r.stdout
    inp = layout(tmp_path, lines, name="toyDetector_10k")  # ref_10k_20_25_2.csv is absent
    r = run(harness, tmp_path, ev, out, inp)
    assert r.returncode == 1 and "ref_10k_20_25_2.csv" in r.stdout and "cannot open reference file" in r.stdout


def test_short_or_malformed_reference(harness, toy, tmp_path):
    ev, out, lines = toy
    inp = layout(tmp_path, lines[:-1])
    r = run(harness, tmp_path, ev, out, inp)
    assert r.returncode == 1 and f"{ev['x'].size} points, the reference file holds {ev['x'].size - 1}" in r.stdout
    bad = list(lines)
    f = bad[6].split(",")
    f[5] = "abc"
    bad[6] = ",".join(f)
    inp = layout(tmp_path, bad)
    r = run(harness, tmp_path, ev, out, inp)
    assert r.returncode == 1 and "Bad Input:" in r.stdout and "in line 6" in r.stdout
    assert "the reference file holds 5" in r.stdout


def test_csv_writer_prints_the_reference_format(harness, toy, tmp_path):
    ev, out, lines = toy
    inp = layout(tmp_path, lines)
    r = run(harness, tmp_path, ev, out, inp, outdir=tmp_path / "csv")
    assert r.returncode == 0, r.stdout + r.stderr
    written = (tmp_path / "csv" / "Event_1_20.00_25.00_2.00.csv").read_text()
    assert written == "\n".join(lines) + "\n"
