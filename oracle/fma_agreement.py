"""Agreement between the oracle and the reference built with its own flags. TEST INFRASTRUCTURE ONLY.

The oracle (and the HIP kernels) follow the reference's serial kernels compiled without FMA
contraction. The reference's own build (its Makefile:27-28: -O2 -march=native, GCC's default
-ffp-contract=fast) contracts a*b+c into fused multiply-adds on FMA hosts, e.g. the distance
dx*dx+dy*dy of CLUEAlgoKernels.h:10-12 and CLUE3D's reduceRange and distance expressions, so its
results may differ in the last bits where a distance sits at a threshold. This script runs both on
the golden events (tests/golden) and on seeded synthetic events and reports, per output field, the
fraction of points whose outputs are bit-identical.

usage: python3 oracle/fma_agreement.py   (after `make -C oracle ref-native`)
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(REPO / "heterogeneous-clue_amd"))
sys.path.insert(0, str(REPO / "tests"))

import oracle  # noqa: E402

NATIVE_LIB = HERE / "_ref" / "libclue_ref_native.so"
OUT2D = ("rho", "delta", "nearestHigher", "isSeed", "clusterIndex")
OUT3D = ("rho", "delta_f", "delta_l", "nh_layer", "nh_idx", "isSeed", "tracksterIndex")


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int32) if a.dtype == np.float32 else a.astype(np.int32)


def _events():
    """(dim, name, inputs, params) of the golden events and the synthetic ones."""
    import helpers
    import synth
    for name in helpers.golden_names("clue2d"):
        ins, _, extra = helpers.load_case("clue2d", name)
        yield 2, f"golden {name}", ins, tuple(float(v) for v in extra["params"])
    for name in helpers.golden_names("clue3d"):
        if name == "eta_row34":
            # its last-layer clusters make the reference read past the end of its tile array (the eta
            # window's row 34 of layer 93): undefined behaviour that the natively built library
            # turns into a fault, while the oracle defines it (an empty row)
            continue
        ins, _, extra = helpers.load_case("clue3d", name)
        yield 3, f"golden {name}", ins, oracle.clue3d_params(S=int(extra["S"]))
    for e in range(4):
        yield 2, f"synthetic 1e5 hits #{e}", synth.clue2d_event(100_000, 500 + e), (3.0, 10.0, 2.0)
    for e in range(8):
        yield 3, f"synthetic clue3d #{e}", synth.clue3d_event(500 + e), oracle.clue3d_params(S=2)


def agreement() -> dict:
    """Per dimension and field: [points bit-identical, points], plus the events that differ."""
    O = oracle.Oracle()
    R = oracle.Reference(NATIVE_LIB)
    tot = {2: {k: [0, 0] for k in OUT2D}, 3: {k: [0, 0] for k in OUT3D}}
    differing = []
    for dim, name, ins, prm in _events():
        if dim == 2:
            a = O.clue2d(ins["x"], ins["y"], ins["layer"], ins["weight"], *prm)
            b = R.clue2d(ins["x"], ins["y"], ins["layer"], ins["weight"], *prm)
        else:
            a, b = O.clue3d(ins, prm), R.clue3d(ins, prm)
        diff = False
        for k in tot[dim]:
            same = int(np.count_nonzero(_bits(a[k]) == _bits(b[k])))
            tot[dim][k][0] += same
            tot[dim][k][1] += a[k].size
            diff = diff or same != a[k].size
        if diff:
            differing.append(name)
    return {"fields": tot, "events_differing": differing}


def main():
    if not NATIVE_LIB.exists():
        sys.exit(f"{NATIVE_LIB} is not built: make -C {HERE} ref-native")
    r = agreement()
    print("| CLUE | field | bit-identical points | of | fraction |")
    print("|---|---|---:|---:|---:|")
    for dim, fields in r["fields"].items():
        for k, (same, n) in fields.items():
            print(f"| {dim}D | {k} | {same} | {n} | {same / n:.6f} |")
    print(f"events with any difference: {len(r['events_differing'])}: {', '.join(r['events_differing'])}")


if __name__ == "__main__":
    main()
