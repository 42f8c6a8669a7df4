"""The driver running the HIP modules through the plugin interface, outputs compared bit for bit."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import helpers
import oracle
import synth

pytestmark = pytest.mark.gpu
PKG = Path(__file__).resolve().parents[1] / "heterogeneous-clue_amd"


def read_dump(path, dim):
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(raw[:4].view(np.uint32)[0])
    cols = helpers.OUT2D if dim == 2 else helpers.OUT3D
    out, pos = {}, 4
    for k in cols:
        dt = np.float32 if k in ("rho", "delta", "delta_f") else np.int32
        out[k] = raw[pos:pos + 4 * n].view(dt)
        pos += 4 * n
    assert pos == raw.size
    return out


def test_driver_clue2d_two_streams(tmp_path, oracle_lib):
    evs = [synth.clue2d_event(n, 70 + i, shuffle=(i == 1)) for i, n in enumerate((3000, 25000, 1200, 60000))]
    f = tmp_path / "raw2D.bin"
    synth.write_raw(f, evs, dim=2)
    cfg = tmp_path / "cfg.csv"
    cfg.write_text("3,10,2,1\n")
    print("case driver clue2d", flush=True)
    r = subprocess.run([str(PKG / "clue_hip_main"), "--dim", "2", "--inputFile", str(f), "--configFile", str(cfg),
                        "--numberOfThreads", "2", "--maxEvents", "6", "--outputDir", str(tmp_path / "out"),
                        "--dumpDir", str(tmp_path / "dump")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Processed 6 events in" in r.stdout and "events/s" in r.stdout
    for eid in range(1, 7):
        ev = evs[(eid - 1) % len(evs)]
        got = read_dump(tmp_path / "dump" / f"Event_{eid}.bin", 2)
        want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], 3.0, 10.0, 2.0)
        helpers.assert_bit_exact(got, want, helpers.OUT2D, f"driver event {eid}")
        csv = tmp_path / "out" / f"Event_{eid}_3.00_10.00_2.00.csv"
        lines = csv.read_text().splitlines()
        assert lines[0] == "index,x,y,layer,weight,rho,delta,nh,isSeed,clusterId"
        assert len(lines) == ev["x"].size + 1
        last = lines[-1].split(",")
        assert int(last[-1]) == int(want["clusterIndex"][-1]) and int(last[-2]) == int(want["isSeed"][-1])


def test_driver_clue3d_with_parameters(tmp_path, oracle_lib):
    evs = [synth.clue3d_event(i) for i in range(3)]
    f = tmp_path / "raw3D.bin"
    synth.write_raw(f, evs, dim=3)
    cfg = tmp_path / "clue3d.txt"
    cfg.write_text("# C4 parameters\ndensitySiblingLayers,2\ncriticalDensity,0.6\n")
    print("case driver clue3d", flush=True)
    r = subprocess.run([str(PKG / "clue_hip_main"), "--dim", "3", "--inputFile", str(f), "--configFile", str(cfg),
                        "--numberOfThreads", "2", "--dumpDir", str(tmp_path / "dump")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Processed 3 events" in r.stdout
    for eid in range(1, 4):
        got = read_dump(tmp_path / "dump" / f"Event_{eid}.bin", 3)
        want = oracle_lib.clue3d(evs[eid - 1], oracle.clue3d_params(S=2))
        helpers.assert_bit_exact(got, want, helpers.OUT3D, f"driver 3d event {eid}")


def test_driver_reports_bad_layer(tmp_path):
    ev = synth.clue2d_event(2000, 3)
    ev["layer"][10] = 150
    f = tmp_path / "raw2D.bin"
    synth.write_raw(f, [ev], dim=2)
    cfg = tmp_path / "cfg.csv"
    cfg.write_text("3,10,2,0\n")
    print("case driver bad layer", flush=True)
    r = subprocess.run([str(PKG / "clue_hip_main"), "--dim", "2", "--inputFile", str(f), "--configFile", str(cfg)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "Caught std::runtime_error" in r.stdout and "layer" in r.stdout


def toy_layout(root, lines):
    """data/input/toyDetector_1k.bin (written by the caller) and the reference result file the
    validator reads for it, data/output/reference/ref_1k_20_25_2.csv."""
    (root / "data" / "input").mkdir(parents=True)
    ref = root / "data" / "output" / "reference"
    ref.mkdir(parents=True)
    (ref / "ref_1k_20_25_2.csv").write_text("\n".join(lines) + "\n")
    return root / "data" / "input" / "toyDetector_1k.bin"


def test_driver_toy_detector_validation(tmp_path, oracle_lib):
    layer = synth.toy_detector_layer(1000)
    ev = synth.expand_toy_detector(layer)
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], 20.0, 25.0, 2.0)
    inp = toy_layout(tmp_path, helpers.clue2d_csv_lines(ev, want))
    synth.write_toy_detector(inp, layer)
    main = str(PKG / "clue_hip_main")
    print("case driver toyDetector --validation", flush=True)
    # --validation after --inputFile selects config/toyDetector_config.csv (20, 25, 2); 10 events by default
    r = subprocess.run([main, "--inputFile", str(inp), "--validation", "--numberOfThreads", "2",
                        "--dumpDir", str(tmp_path / "dump")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Processing 10 events" in r.stdout and "dc = 20" in r.stderr
    assert r.stdout.count("CLUE output is correct!") == 10
    for eid in (1, 10):
        got = read_dump(tmp_path / "dump" / f"Event_{eid}.bin", 2)
        helpers.assert_bit_exact(got, want, helpers.OUT2D, f"toyDetector event {eid}")
    print("case driver toyDetector --strictValidation", flush=True)
    r = subprocess.run([main, "--inputFile", str(inp), "--strictValidation", "--maxEvents", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("identical to the reference file") == 2
    print("case driver toyDetector --runForMinutes 0", flush=True)
    r = subprocess.run([main, "--inputFile", str(inp), "--validation", "--runForMinutes", "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Processing for about 0 minutes" in r.stdout and "Processed 1 events" in r.stdout
    print("case driver toyDetector other parameters", flush=True)
    cfg = tmp_path / "cfg.csv"
    cfg.write_text("3,10,2,0\n")
    r = subprocess.run([main, "--inputFile", str(inp), "--validation", "--configFile", str(cfg), "--maxEvents", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "cannot check results" in r.stdout, r.stdout + r.stderr


def test_driver_validation_failure_and_3d(tmp_path, oracle_lib):
    layer = synth.toy_detector_layer(600, seed=11)
    ev = synth.expand_toy_detector(layer)
    want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], 20.0, 25.0, 2.0)
    lines = helpers.clue2d_csv_lines(ev, want)
    k = int(np.flatnonzero(want["rho"] > 1)[3])
    f = lines[k + 1].split(",")
    f[5] = "%g" % (float(f[5]) + 0.01)
    lines[k + 1] = ",".join(f)
    inp = toy_layout(tmp_path, lines)
    synth.write_toy_detector(inp, layer)
    print("case driver toyDetector corrupted reference", flush=True)
    r = subprocess.run([str(PKG / "clue_hip_main"), "--inputFile", str(inp), "--validation", "--maxEvents", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "Caught std::runtime_error" in r.stdout and f"rho differs from the reference at i={k}," in r.stdout
    evs = [synth.clue3d_event(5)]
    f3 = tmp_path / "raw3D.bin"
    synth.write_raw(f3, evs, dim=3)
    print("case driver clue3d --validation", flush=True)
    r = subprocess.run([str(PKG / "clue_hip_main"), "--dim", "3", "--inputFile", str(f3), "--validation"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Validation not available for CLUE 3D" in r.stderr


def test_driver_more_streams_than_threads(tmp_path, oracle_lib):
    """4 EDM streams on 1 thread: the external-work modules hand the thread to the other streams while
    an event is on the device; every event's outputs stay bit-exact."""
    evs = [synth.clue2d_event(n, 90 + i) for i, n in enumerate((40000, 2000, 150000, 7000, 0, 60000))]
    f = tmp_path / "raw2D.bin"
    synth.write_raw(f, evs, dim=2)
    cfg = tmp_path / "cfg.csv"
    cfg.write_text("3,10,2,0\n")
    print("case driver 4 streams 1 thread 2d", flush=True)
    r = subprocess.run([str(PKG / "clue_hip_main"), "--inputFile", str(f), "--configFile", str(cfg),
                        "--numberOfThreads", "1", "--numberOfStreams", "4", "--maxEvents", "12",
                        "--dumpDir", str(tmp_path / "dump")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Processed 12 events" in r.stdout
    for eid in range(1, 13):
        ev = evs[(eid - 1) % len(evs)]
        got = read_dump(tmp_path / "dump" / f"Event_{eid}.bin", 2)
        want = oracle_lib.clue2d(ev["x"], ev["y"], ev["layer"], ev["weight"], 3.0, 10.0, 2.0)
        helpers.assert_bit_exact(got, want, helpers.OUT2D, f"4-stream event {eid}")
    evs3 = [synth.clue3d_event(20 + i) for i in range(3)]
    f3 = tmp_path / "raw3D.bin"
    synth.write_raw(f3, evs3, dim=3)
    print("case driver 3 streams 1 thread 3d", flush=True)
    r = subprocess.run([str(PKG / "clue_hip_main"), "--dim", "3", "--inputFile", str(f3), "--numberOfThreads", "1",
                        "--numberOfStreams", "3", "--maxEvents", "6", "--dumpDir", str(tmp_path / "dump3")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for eid in range(1, 7):
        got = read_dump(tmp_path / "dump3" / f"Event_{eid}.bin", 3)
        want = oracle_lib.clue3d(evs3[(eid - 1) % 3], oracle.clue3d_params())
        helpers.assert_bit_exact(got, want, helpers.OUT3D, f"3-stream 3d event {eid}")
