"""Seeded synthetic events for the CLUE2D / CLUE3D clusterizers.

Every event is a function of (seed, event id) only, so any event of any workload can be
regenerated identically on any machine (numpy PCG64 streams).

CLUE2D model (per layer): 90 % of the hits sit in Gaussian blobs (sigma 1.0 cm, Poisson(20)
hits per blob, centres uniform in [-250, 250]^2 cm, the last blob truncated to the quota),
10 % are uniform noise over the same square; weights are Exp(mean 1). Hits of an event are
spread evenly over 100 layers (remainder to the low layers).

CLUE3D model (per event): 50 layers x Poisson(200) layer-clusters; 80 % belong to showers
(axis eta0 ~ U[1.5, 3.0], phi0 ~ U[-pi, pi), per-cluster eta/phi ~ N(axis, 0.01), each shower
spanning 5-15 consecutive layers), 20 % are noise; energy ~ Exp(mean 2 GeV);
z = 320 + 1.2 * layer cm, r_over_absz = 1 / sinh(eta), x, y = z * r_over_absz * (cos, sin)(phi),
isSilicon = 1, radius = 1.8 cm. |eta| is clamped to [1.5, 3.0] and phi wrapped to [-pi, pi).

The binary writers/readers follow the reference input format (src/serial/bin/Source.cc:8-58,
88-144): a sequence of events, each `uint32 n` followed by n records of 4 (2D: x, y, layer,
weight) or 10 (3D: x, y, z, eta, phi, r_over_absz, radius, layer, energy, isSilicon) float32.
"""
from __future__ import annotations

import math

import numpy as np

SEED_BASE = 0xC1ADE
NLAYERS_2D = 100


def _rng(seed: int, event: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64([int(seed) & 0xFFFFFFFF, int(event)]))


def layer_counts(n_hits: int, n_layers: int = NLAYERS_2D) -> np.ndarray:
    c = np.full(n_layers, n_hits // n_layers, dtype=np.int64)
    c[: n_hits % n_layers] += 1
    return c


def clue2d_event(n_hits: int, event: int = 0, seed: int = SEED_BASE, n_layers: int = NLAYERS_2D,
                 blob_frac: float = 0.9, blob_mean: float = 20.0, sigma: float = 1.0,
                 half: float = 250.0, shuffle: bool = False) -> dict:
    """One CLUE2D event: dict of float32 x, y, weight and int32 layer arrays."""
    rng = _rng(seed, event)
    xs, ys, ls = [], [], []
    for lay, n_l in enumerate(layer_counts(n_hits, n_layers)):
        n_l = int(n_l)
        if n_l == 0:
            continue
        n_blob = int(round(blob_frac * n_l))
        n_noise = n_l - n_blob
        if n_blob > 0:
            sizes = rng.poisson(blob_mean, size=max(4, int(1.5 * n_blob / blob_mean) + 4))
            sizes = sizes[sizes > 0]
            while sizes.sum() < n_blob:
                extra = rng.poisson(blob_mean, size=8)
                sizes = np.concatenate([sizes, extra[extra > 0]])
            cs = np.cumsum(sizes)
            k = int(np.searchsorted(cs, n_blob)) + 1
            sizes = sizes[:k].copy()
            sizes[-1] -= int(cs[k - 1] - n_blob)
            centres = rng.uniform(-half, half, size=(k, 2))
            pts = np.repeat(centres, sizes, axis=0) + rng.normal(0.0, sigma, size=(n_blob, 2))
            xs.append(pts[:, 0])
            ys.append(pts[:, 1])
            ls.append(np.full(n_blob, lay, dtype=np.int32))
        if n_noise > 0:
            noise = rng.uniform(-half, half, size=(n_noise, 2))
            xs.append(noise[:, 0])
            ys.append(noise[:, 1])
            ls.append(np.full(n_noise, lay, dtype=np.int32))
    if xs:
        x = np.concatenate(xs).astype(np.float32)
        y = np.concatenate(ys).astype(np.float32)
        layer = np.concatenate(ls).astype(np.int32)
    else:
        x = np.zeros(0, np.float32)
        y = np.zeros(0, np.float32)
        layer = np.zeros(0, np.int32)
    weight = rng.exponential(1.0, size=x.size).astype(np.float32)
    if shuffle:
        p = rng.permutation(x.size)
        x, y, layer, weight = x[p], y[p], layer[p], weight[p]
    return {"x": x, "y": y, "layer": layer, "weight": weight}


def c3_hit_counts(n_events: int = 256, seed: int = SEED_BASE, lo: float = 3.0, hi: float = 6.0) -> np.ndarray:
    """Hits per event log-uniform in [10^lo, 10^hi] (the 256-event mixed batch)."""
    rng = _rng(seed ^ 0x5EED, 1 << 20)
    return np.rint(10.0 ** rng.uniform(lo, hi, size=n_events)).astype(np.int64)


def pack(events: list[dict]) -> tuple[dict, np.ndarray]:
    """Concatenate events into one SoA plus int64 offsets[E+1]."""
    offs = np.zeros(len(events) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([e["layer"].size for e in events])
    keys = events[0].keys() if events else []
    packed = {k: np.concatenate([e[k] for e in events]) for k in keys}
    return packed, offs


def clue2d_batch(counts, seed: int = SEED_BASE, first_event: int = 0, **kw) -> tuple[dict, np.ndarray]:
    return pack([clue2d_event(int(n), first_event + i, seed, **kw) for i, n in enumerate(counts)])


# ------------------------------------------------------------------ CLUE3D

def clue3d_event(event: int = 0, seed: int = SEED_BASE + 3, n_layers: int = 50, mean_per_layer: float = 200.0,
                 shower_frac: float = 0.8, n_showers: int = 40, sigma: float = 0.01,
                 energy_mean: float = 2.0) -> dict:
    """One CLUE3D event: float32 x, y, z, eta, phi, r_over_absz, radius, energy; int32 layer, isSilicon."""
    rng = _rng(seed, event)
    eta0 = rng.uniform(1.5, 3.0, size=n_showers)
    phi0 = rng.uniform(-math.pi, math.pi, size=n_showers)
    length = rng.integers(5, 16, size=n_showers)
    first = rng.integers(0, n_layers, size=n_showers)
    last = np.minimum(first + length - 1, n_layers - 1)
    etas, phis, lays = [], [], []
    for lay in range(n_layers):
        n_l = int(rng.poisson(mean_per_layer))
        active = np.nonzero((first <= lay) & (last >= lay))[0]
        n_sh = int(round(shower_frac * n_l)) if active.size else 0
        n_no = n_l - n_sh
        if n_sh:
            which = active[rng.integers(0, active.size, size=n_sh)]
            etas.append(eta0[which] + rng.normal(0.0, sigma, size=n_sh))
            phis.append(phi0[which] + rng.normal(0.0, sigma, size=n_sh))
            lays.append(np.full(n_sh, lay, np.int32))
        if n_no:
            etas.append(rng.uniform(1.5, 3.0, size=n_no))
            phis.append(rng.uniform(-math.pi, math.pi, size=n_no))
            lays.append(np.full(n_no, lay, np.int32))
    eta = np.clip(np.concatenate(etas), 1.5, 3.0).astype(np.float32) if etas else np.zeros(0, np.float32)
    phi64 = np.concatenate(phis) if phis else np.zeros(0)
    phi64 = (phi64 + math.pi) % (2.0 * math.pi) - math.pi
    phi = phi64.astype(np.float32)
    phi = np.where(phi >= np.float32(math.pi), phi - np.float32(2.0 * math.pi), phi).astype(np.float32)
    layer = np.concatenate(lays).astype(np.int32) if lays else np.zeros(0, np.int32)
    n = eta.size
    z = (320.0 + 1.2 * layer.astype(np.float64)).astype(np.float32)
    roz = (1.0 / np.sinh(eta.astype(np.float64))).astype(np.float32)
    rr = z.astype(np.float64) * roz.astype(np.float64)
    x = (rr * np.cos(phi.astype(np.float64))).astype(np.float32)
    y = (rr * np.sin(phi.astype(np.float64))).astype(np.float32)
    energy = rng.exponential(energy_mean, size=n).astype(np.float32)
    return {"x": x, "y": y, "z": z, "eta": eta, "phi": phi, "r_over_absz": roz,
            "radius": np.full(n, 1.8, np.float32), "layer": layer, "energy": energy,
            "isSilicon": np.ones(n, np.int32)}


def clue3d_batch(n_events: int, seed: int = SEED_BASE + 3, first_event: int = 0, **kw) -> tuple[dict, np.ndarray]:
    return pack([clue3d_event(first_event + i, seed, **kw) for i in range(n_events)])


# ------------------------------------------------------------------ reference binary format

_FIELDS_2D = ("x", "y", "layer", "weight")
_FIELDS_3D = ("x", "y", "z", "eta", "phi", "r_over_absz", "radius", "layer", "energy", "isSilicon")


def write_raw(path, events: list[dict], dim: int = 2) -> None:
    fields = _FIELDS_2D if dim == 2 else _FIELDS_3D
    with open(path, "wb") as f:
        for ev in events:
            n = ev["layer"].size
            rec = np.empty((n, len(fields)), dtype=np.float32)
            for k, name in enumerate(fields):
                rec[:, k] = ev[name].astype(np.float32)
            f.write(np.uint32(n).tobytes())
            f.write(rec.tobytes())


def toy_detector_layer(n_points: int, seed: int = SEED_BASE + 7) -> dict:
    """One layer of a toy-detector event sized for dc = 20, rhoc = 25: the CLUE2D model on layer 0
    with blobs of Poisson(40) hits, sigma 5 cm, 80 % of the hits in blobs."""
    return clue2d_event(n_points, 0, seed, n_layers=1, blob_frac=0.8, blob_mean=40.0, sigma=5.0)


def write_toy_detector(path, layer_event: dict) -> None:
    """A toyDetector file: the records {x, y, layer, weight} (float32) of one layer, no header
    (the reference reads it in src/serial/bin/Source.cc:60-79)."""
    rec = np.stack([layer_event[k].astype(np.float32) for k in _FIELDS_2D], axis=1)
    with open(path, "wb") as f:
        f.write(np.ascontiguousarray(rec).tobytes())


def expand_toy_detector(layer_event: dict, n_layers: int = NLAYERS_2D) -> dict:
    """The event a toyDetector file stands for: its records repeated on layers 0..99, layer
    int(float32(layer) + l) for copy l."""
    lay = layer_event["layer"].astype(np.float32)
    return {
        "x": np.tile(layer_event["x"].astype(np.float32), n_layers),
        "y": np.tile(layer_event["y"].astype(np.float32), n_layers),
        "layer": np.concatenate([(lay + np.float32(l)).astype(np.int32) for l in range(n_layers)]),
        "weight": np.tile(layer_event["weight"].astype(np.float32), n_layers),
    }


def read_raw(path, dim: int = 2) -> list[dict]:
    fields = _FIELDS_2D if dim == 2 else _FIELDS_3D
    data = np.fromfile(path, dtype=np.uint8)
    out, pos = [], 0
    while pos < data.size:
        if pos + 4 > data.size:
            raise ValueError("truncated event header")
        n = int(data[pos:pos + 4].view(np.uint32)[0])
        pos += 4
        nbytes = n * 4 * len(fields)
        if pos + nbytes > data.size:
            raise ValueError("truncated event records")
        rec = data[pos:pos + nbytes].view(np.float32).reshape(n, len(fields))
        pos += nbytes
        ev = {}
        for k, name in enumerate(fields):
            col = rec[:, k]
            ev[name] = col.astype(np.int32) if name in ("layer", "isSilicon") else col.copy()
        out.append(ev)
    return out
