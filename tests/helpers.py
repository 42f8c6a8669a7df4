"""Fixture loading and bit-exact comparison helpers for the parity tests."""
from __future__ import annotations

from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
OUT2D = ("rho", "delta", "nearestHigher", "isSeed", "clusterIndex")
OUT3D = ("rho", "delta_f", "delta_l", "nh_layer", "nh_idx", "isSeed", "tracksterIndex")


def golden_names(prefix: str) -> list[str]:
    return sorted(p.stem[len(prefix) + 1:] for p in GOLDEN.glob(f"{prefix}_*.npz"))


def load_case(prefix: str, name: str):
    z = np.load(GOLDEN / f"{prefix}_{name}.npz", allow_pickle=False)
    ins = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    outs = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    extra = {k: z[k] for k in z.files if not (k.startswith("in_") or k.startswith("out_"))}
    return ins, outs, extra


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    return a.view(np.int32) if a.dtype == np.float32 else a.astype(np.int32)


def assert_bit_exact(got: dict, want: dict, keys, what: str = ""):
    for k in keys:
        g, w = bits(got[k]), bits(want[k])
        assert g.shape == w.shape, f"{what} {k}: shape {g.shape} != {w.shape}"
        bad = np.nonzero(g != w)[0]
        if bad.size:
            i = int(bad[0])
            raise AssertionError(f"{what} {k}: {bad.size} of {g.size} differ; first at {i}: "
                                 f"got {got[k][i]!r} want {want[k][i]!r}")


def check_invariants_2d(ev: dict, out: dict, dc: float, rhoc: float, odf: float):
    """Size-independent properties of a CLUE2D result (CLUEAlgoKernels.h:25-153)."""
    x, y, w, lay = ev["x"], ev["y"], ev["weight"], ev["layer"]
    rho, delta, nh = out["rho"], out["delta"], out["nearestHigher"]
    seed, ci = out["isSeed"], out["clusterIndex"]
    n = x.size
    dm = np.float32(odf) * np.float32(dc)
    # every point counts itself with weight 1 (weights are non-negative here)
    assert np.all(rho >= w)
    has = nh >= 0
    j = nh[has]
    i = np.nonzero(has)[0]
    # nearest higher is on the same layer and is "higher" with the index tie-break
    assert np.all(lay[j] == lay[i])
    assert np.all((rho[j] > rho[i]) | ((rho[j] == rho[i]) & (j > i)))
    # delta is the fp32 distance to the nearest higher, and within dm
    dx = (x[i] - x[j]).astype(np.float32)
    dy = (y[i] - y[j]).astype(np.float32)
    d = np.sqrt((dx * dx).astype(np.float32) + (dy * dy).astype(np.float32)).astype(np.float32)
    assert np.array_equal(d.view(np.int32), delta[i].view(np.int32))
    assert np.all(delta[i] <= dm)
    assert np.all(delta[~has] == np.float32(np.finfo(np.float32).max))
    # seeds are numbered by index; followers carry their root's id
    s = np.nonzero(seed)[0]
    assert np.array_equal(ci[s], np.arange(s.size, dtype=np.int32))
    follow = has & (seed == 0) & ~((delta > dm) & (rho < rhoc))
    assert np.all(ci[follow] == ci[nh[follow]])
    assert np.all((ci >= -1) & (ci < max(s.size, 1)))
    assert n == ci.size


def check_invariants_3d(ev: dict, out: dict):
    """Size-independent properties of a CLUE3D result (CLUE3DAlgoKernels.h:286-580)."""
    rho, nhl, nhi = out["rho"], out["nh_layer"], out["nh_idx"]
    seed, ti, dl = out["isSeed"], out["tracksterIndex"], out["delta_l"]
    lay = ev["layer"]
    has = nhl >= 0
    i = np.nonzero(has)[0]
    j = nhi[has]
    assert np.all(rho[j] > rho[i])
    # nearestHigher.first is the searched layer. For |eta| >= 3.15 the eta window's row 34 is the
    # next layer's first row in the reference's contiguous tile array, so the hit found there sits
    # on layer first + 1.
    aliased = lay[j] != nhl[has]
    assert np.all(lay[j][aliased] == nhl[has][aliased] + 1)
    assert np.all(np.abs(ev["eta"][i][aliased]) >= np.float32(3.15) - np.float32(1e-6))
    assert np.all(dl[has] == np.abs(nhl[has] - lay[i]))
    assert np.all(nhl[has] != lay[i])
    s = np.nonzero(seed)[0]
    assert np.array_equal(ti[s], np.arange(s.size, dtype=np.int32))


CSV_HEADER_2D = "index,x,y,layer,weight,rho,delta,nh,isSeed,clusterId"


def _g(v) -> str:
    """A float32 as C++ iostreams print it by default (printf %g, 6 significant digits)."""
    return "%g" % float(np.float32(v))


def clue2d_csv_lines(ev: dict, out: dict) -> list[str]:
    """The result CSV of one CLUE2D event, in the reference writer's format: header, then one row per
    point with delta clamped at 999."""
    rows = [CSV_HEADER_2D]
    for i in range(ev["x"].size):
        d = out["delta"][i]
        rows.append(",".join([str(i), _g(ev["x"][i]), _g(ev["y"][i]), str(int(ev["layer"][i])), _g(ev["weight"][i]),
                              _g(out["rho"][i]), _g(999 if d > 999 else d), str(int(out["nearestHigher"][i])),
                              str(int(out["isSeed"][i])), str(int(out["clusterIndex"][i]))]))
    return rows


EXTREME_PARAMS = [(3.0, 10.0, 2.0), (0.0, 1.0, 2.0), (-1.0, 1.0, 2.0), (300.0, 10.0, 2.0), (3.0, 10.0, 1e6),
                  (float("nan"), 10.0, 2.0)]


def extreme_event(seed: int = 5, n_normal: int = 500) -> dict:
    """A CLUE2D event whose points include NaN, infinite and huge coordinates, points just inside and
    outside the [-250, 250) grid and on bin edges, mixed with ordinary blob points."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(-240, 240, (12, 2))
    pts = c[rng.integers(0, 12, n_normal)] + rng.normal(0, 1.5, (n_normal, 2))
    special = [np.nan, np.inf, -np.inf, 1e30, -1e30, 1e9, -1e9, 250.0, -250.0, 249.99998, -250.00002, 1.07374e10,
               0.0, -0.0, 5.0, -5.0, 245.0, 1e-40]
    sx, sy = [], []
    for a in special:
        for b in (0.0, 3.0, np.nan, 250.0, -1e30):
            sx.append(a)
            sy.append(b)
            sx.append(b)
            sy.append(a)
    x = np.concatenate([pts[:, 0], sx]).astype(np.float32)
    y = np.concatenate([pts[:, 1], sy]).astype(np.float32)
    n = x.size
    layer = rng.integers(0, 4, n).astype(np.int32) * 25
    weight = rng.exponential(1.0, n).astype(np.float32)
    p = rng.permutation(n)
    return {"x": x[p], "y": y[p], "layer": layer[p], "weight": weight[p]}
