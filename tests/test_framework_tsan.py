"""The framework's scheduler (task pool, stream schedules, counted holders, external work completed
on foreign threads) under ThreadSanitizer: the self-test and the driver with more streams than
threads, built for the host with g++ -fsanitize=thread (no device code involved)."""
import os
import subprocess
from pathlib import Path

import pytest

import synth

REPO = Path(__file__).resolve().parents[1]
PKG = REPO / "heterogeneous-clue_amd"
OUT = REPO / "tests" / "modules" / "_build" / "tsan"
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=thread", "-I", str(PKG / "framework"), "-I", str(REPO / "include")]
FW = [str(PKG / "framework" / "Framework" / f) for f in ("PluginFactory.cc", "TaskPool.cc")]
ENV = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")


def build(args, timeout=300):
    r = subprocess.run(["g++"] + FLAGS + args + ["-pthread"], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0 and "tsan" in r.stderr.lower():
        pytest.skip("ThreadSanitizer runtime not available: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.fixture(scope="module")
def tsan_dir():
    OUT.mkdir(parents=True, exist_ok=True)
    return OUT


def test_selftest_under_tsan(tsan_dir):
    exe = tsan_dir / "fw_selftest"
    build(["-o", str(exe), str(PKG / "bin" / "fw_selftest.cc")] + FW)
    for _ in range(3):
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=ENV)
        assert r.returncode == 0 and "ThreadSanitizer" not in r.stdout + r.stderr, (r.stdout + r.stderr)[-4000:]
        assert "framework self-test passed" in r.stdout


def test_driver_scheduler_under_tsan(tsan_dir, tmp_path):
    if not (PKG / "libclue_plugins.so").exists():
        pytest.skip("plugin library not built")
    lib = tsan_dir / "libclue_framework.so"
    build(["-fPIC", "-shared", "-o", str(lib)] + FW)
    for so in ("libclue_plugins.so", "libclue_hip.so"):  # found next to the instrumented framework
        (tsan_dir / so).write_bytes((PKG / so).read_bytes())
    exe = tsan_dir / "clue_hip_main"
    build(["-o", str(exe), str(PKG / "bin" / "main.cc"), "-L", str(tsan_dir), "-lclue_framework", "-ldl",
           f"-Wl,-rpath,{tsan_dir}"])
    f = tmp_path / "raw2D.bin"
    synth.write_raw(f, [synth.clue2d_event(300, i) for i in range(3)], dim=2)
    cfg = tmp_path / "cfg.csv"
    cfg.write_text("3,10,2,0\n")
    for extra in (["--numberOfThreads", "3", "--numberOfStreams", "7", "--maxEvents", "2000"],
                  ["--numberOfThreads", "4", "--numberOfStreams", "4", "--runForMinutes", "0"]):
        r = subprocess.run([str(exe), "--inputFile", str(f), "--configFile", str(cfg), "--empty",
                            "--plugin", str(tsan_dir / "libclue_plugins.so")] + extra,
                           capture_output=True, text=True, timeout=300, env=ENV)
        assert r.returncode == 0 and "ThreadSanitizer" not in r.stdout + r.stderr, (r.stdout + r.stderr)[-4000:]
        assert "Processed" in r.stdout
