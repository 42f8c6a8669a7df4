"""Host-side pieces that need no GPU: generator, reference file format, C ABI exports."""
import re
from pathlib import Path

import numpy as np
import pytest

import synth

REPO = Path(__file__).resolve().parents[1]


def test_generator_is_deterministic_and_shaped():
    a = synth.clue2d_event(12345, 3)
    b = synth.clue2d_event(12345, 3)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert a["x"].size == 12345 and a["x"].dtype == np.float32 and a["layer"].dtype == np.int32
    counts = np.bincount(a["layer"], minlength=100)
    assert counts.max() - counts.min() <= 1
    assert np.all(a["weight"] >= 0)
    c = synth.clue2d_event(12345, 4)
    assert not np.array_equal(a["x"], c["x"])


def test_c3_counts_log_uniform():
    n = synth.c3_hit_counts(256)
    assert n.size == 256 and n.min() >= 1000 and n.max() <= 1_000_000
    lg = np.log10(n)
    assert 4.2 < lg.mean() < 4.8


def test_clue3d_event_model():
    e = synth.clue3d_event(0)
    assert e["layer"].max() < 50
    assert np.all(np.abs(e["eta"]) >= 1.5) and np.all(np.abs(e["eta"]) <= 3.0)
    assert np.all(e["phi"] >= -np.float32(np.pi)) and np.all(e["phi"] < np.float32(np.pi))
    assert 50 * 150 < e["layer"].size < 50 * 250


def test_raw_format_round_trip(tmp_path):
    evs = [synth.clue2d_event(500, i) for i in range(3)]
    p = tmp_path / "raw2D.bin"
    synth.write_raw(p, evs, dim=2)
    assert p.stat().st_size == sum(4 + 16 * e["x"].size for e in evs)
    back = synth.read_raw(p, dim=2)
    for a, b in zip(evs, back):
        for k in a:
            assert np.array_equal(a[k], b[k])
    e3 = [synth.clue3d_event(0)]
    p3 = tmp_path / "raw3D.bin"
    synth.write_raw(p3, e3, dim=3)
    back3 = synth.read_raw(p3, dim=3)
    for k in e3[0]:
        assert np.array_equal(e3[0][k], back3[0][k])
    with open(p, "ab") as f:
        f.write(b"\x05\x00\x00\x00\x00")
    with pytest.raises(ValueError):
        synth.read_raw(p, dim=2)


def declared_functions():
    hdr = (REPO / "include" / "clue_hip.h").read_text()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(clue\w*)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    import clue_hip
    lib = clue_hip.load_library()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(clue_hip.EXPORTS) == names


def test_library_reports_no_device_without_gpu():
    import clue_hip
    lib = clue_hip.load_library()
    n = lib.clue_device_count()
    assert n >= 0
    if n == 0:
        assert lib.clue_create(0, 0) is None
        p = lib.clue3d_default_params()
        assert p.densitySiblingLayers == 3 and abs(p.densityXYDistanceSqr - 3.24) < 1e-6


def test_header_constants_match_the_binding():
    """Every #define of include/clue_hip.h has the value the Python binding uses."""
    import clue_hip
    hdr = (REPO / "include" / "clue_hip.h").read_text()
    defs = {k: int(v) for k, v in re.findall(r"#define (CLUE_\w+) \(?(-?\d+)\)?", hdr)}
    py = {"CLUE_OK": clue_hip.CLUE_OK, "CLUE_FUSED_CAPACITY": clue_hip.FUSED_CAPACITY,
          "CLUE_NOALLOC_MAX_POINTS": clue_hip.NOALLOC_MAX_POINTS}
    py.update({name: code for code, name in clue_hip.ERRORS.items()})
    py.update({f"CLUE_{k}": getattr(clue_hip, k) for k in dir(clue_hip) if k.startswith(("PATH_", "PLAN_"))})
    assert sorted(defs) == sorted(py), (sorted(defs), sorted(py))
    for k, v in defs.items():
        assert py[k] == v, k
