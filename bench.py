#!/usr/bin/env python3
"""Benchmark of the CLUE2D batched path on one MI355X.

Workload (the configuration the headline metric is quoted on): 256 CLUE2D events with hit counts
log-uniform in [1e3, 1e6] (about 3.4e7 hits), 100 layers, blob + noise model, dc=3, rhoc=10,
outlierDeltaFactor=2, seeded. A step is one pass of the batched path over the whole batch: one
hipGraph launch of the recorded kernel sequence, inputs already resident in HBM, timed with HIP
events on the launch stream. The same events are also run as per-event launches (one launch
sequence + stream synchronisation per event, as the reference's GPU backend does) and the labels
of both paths are compared bit for bit.

Prints one JSON line (rank 0, N=1). `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "heterogeneous-clue_amd"))

import clue_hip  # noqa: E402
import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_HIT_2D = 36  # in: x, y, weight f32 + layer i32; out: rho, delta f32 + nh, isSeed, clusterIndex i32
BYTES_PER_CLUSTER_3D = 56


TRAFFIC_PROFILE = REPO / "profiles" / "clue2d_batched_256ev_hbm_traffic.md"


def measured_traffic(key="traffic_bytes_per_launch"):
    """HBM bytes per launch of the 256-event batch (the whole graph replay, or one kernel with
    key = "traffic_bytes_<kernel>"), from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the
    copy-probe calibration recorded there), or None when absent."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith(key + ":"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ptr(a: clue_hip.DeviceArray, off: int) -> int:
    return a.ptr + off * a.dtype.itemsize


def per_event_run(dev, din, dout, offs, p, reps=1):
    """Every event as its own launch sequence followed by a stream synchronisation."""
    lib, ctx = dev.lib, dev.ctx
    E = offs.size - 1
    best = None
    for _ in range(reps):
        dev.sync()
        t0 = time.perf_counter()
        for e in range(E):
            a, b = int(offs[e]), int(offs[e + 1])
            if b == a:
                continue
            rc = lib.clue2d_run_device(ctx, C.byref(p), b - a, *[ptr(din[k], a) for k in clue_hip.FIELDS_2D_IN],
                                       *[ptr(dout[k], a) for k in clue_hip.FIELDS_2D_OUT], None)
            if rc:
                raise clue_hip.ClueError(rc, dev.error())
            lib.clue_stream_sync(ctx, None)
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    dev.check()
    return best


def cpu_baseline(packed, offs, p, gpu_out):
    """The reference's own serial kernels (oracle/_ref) on the host cores, event-parallel."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    if oracle.available_reference():
        R, kind = oracle.Reference(), "reference"
        t0 = time.perf_counter()
        out = R.clue2d_batch(packed, offs, p.dc, p.rhoc, p.outlierDeltaFactor, threads=threads)
        t = time.perf_counter() - t0
    else:
        R, kind, threads = oracle.Oracle(), "port", 1
        t0 = time.perf_counter()
        out = R.clue2d_batch(packed, offs, p.dc, p.rhoc, p.outlierDeltaFactor)
        t = time.perf_counter() - t0
    same = all(np.array_equal(np.ascontiguousarray(out[k]).view(np.int32),
                              np.ascontiguousarray(gpu_out[k]).view(np.int32)) for k in out)
    n = int(offs[-1])
    # one thread on one event: the 1e5-hit (100 layers x 1e3) and the 1e6-hit (100 x 1e4) events
    single = {}
    for name, hits in (("1e5_hits", 100_000), ("1e6_hits", 1_000_000)):
        ev, eo = synth.pack([synth.clue2d_event(hits, 0)])
        t1 = time.perf_counter()
        R.clue2d_batch(ev, eo, p.dc, p.rhoc, p.outlierDeltaFactor, **({"threads": 1} if kind == "reference" else {}))
        single[name] = round(hits / (time.perf_counter() - t1) / 1e6, 3)
    return {"value": round(n / t / 1e6, 3), "unit": "Mhits/s", "cores": threads, "kind": kind,
            "sample": f"all {offs.size - 1} events of the workload ({n} hits), one event per thread at a time",
            "seconds": round(t, 3), "labels_bit_exact_vs_gpu": bool(same),
            "single_thread_one_event_mhits_per_s": single, "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def per_event_run3d(dev, din, dout, offs, p):
    """CLUE3D: every event as its own launch sequence followed by a stream synchronisation."""
    lib, ctx = dev.lib, dev.ctx
    t0 = time.perf_counter()
    for e in range(offs.size - 1):
        a, b = int(offs[e]), int(offs[e + 1])
        if b == a:
            continue
        rc = lib.clue3d_run_device(ctx, C.byref(p), b - a, *[ptr(din[k], a) for k in clue_hip.FIELDS_3D_IN],
                                   *[ptr(dout[k], a) for k in clue_hip.FIELDS_3D_OUT], None)
        if rc:
            raise clue_hip.ClueError(rc, dev.error())
        lib.clue_stream_sync(ctx, None)
    t = time.perf_counter() - t0
    dev.check()
    return t


def clue3d_leg(dev, n_events, steps, warmup):
    """C4: 512 CLUE3D events of 50 layers x Poisson(200) layer clusters (S=2), batched graph vs per-event."""
    packed, offs = synth.clue3d_batch(n_events)
    N = int(offs[-1])
    p = clue_hip.params3d(S=2)
    din = dev.upload_soa(packed, clue_hip.FIELDS_3D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_3D_OUT)
    plan = dev.clue3d_plan(din, dout, offs, p, use_graph=True, autotune=True)
    for _ in range(max(2, warmup)):  # the first launch tunes, the second records the graph
        plan.launch()
    dev.check()
    lib, ctx = dev.lib, dev.ctx
    dev.sync()
    lib.clue_mark(ctx, 2, None)
    for _ in range(steps):
        plan.launch()
    lib.clue_mark(ctx, 3, None)
    dev.sync()
    dev.check()
    ms = lib.clue_elapsed_ms(ctx, 2, 3) / steps
    tuning = plan.tuning
    batch_out = dev.download_soa(dout)
    pout = dev.alloc_soa(N, clue_hip.FIELDS_3D_OUT)
    t = per_event_run3d(dev, din, pout, offs, p)
    per_out = dev.download_soa(pout)
    same = all(np.array_equal(per_out[k].view(np.int32), batch_out[k].view(np.int32)) for k in per_out)
    plan.destroy()
    for a in list(din.values()) + list(dout.values()) + list(pout.values()):
        a.free()
    value = N / (ms * 1e-3) / 1e6
    algo = BYTES_PER_CLUSTER_3D * N
    out = {"workload": f"clue3d_batched_{n_events}_events_50_layers_poisson200", "events": n_events, "clusters": N,
           "densitySiblingLayers": 2, "autotuned": tuning, "ms_per_step": round(ms, 4), "value": round(value, 2),
           "unit": "Mclusters/s", "per_event": {"value": round(N / t / 1e6, 2), "unit": "Mclusters/s",
                                                "ms_total": round(t * 1e3, 3)},
           "batched_over_per_event": round(value / (N / t / 1e6), 2), "labels_bit_exact_vs_per_event": bool(same),
           "roofline": {"bound": "hbm", "achieved": round(algo / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes_per_launch": algo}}
    log(f"clue3d: {ms:.3f} ms/step ({value:.1f} Mclusters/s), per-event {t * 1e3:.1f} ms, bit-exact={same}")
    return out


def shuffled_leg(dev, counts, p, steps, warmup):
    """The same 256 events with their hits in random order (as raw input files need not be ordered by
    layer): an autotuned graph plan timed like the headline step; four events checked against the
    CPU oracle."""
    packed, offs = synth.clue2d_batch(counts, shuffle=True)
    N = int(offs[-1])
    din = dev.upload_soa(packed, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_2D_OUT)
    plan = dev.clue2d_plan(din, dout, offs, p, use_graph=True, autotune=True)
    for _ in range(max(2, warmup)):
        plan.launch()
    dev.check()
    lib, ctx = dev.lib, dev.ctx
    dev.sync()
    lib.clue_mark(ctx, 6, None)
    for _ in range(steps):
        plan.launch()
    lib.clue_mark(ctx, 7, None)
    dev.sync()
    dev.check()
    ms = lib.clue_elapsed_ms(ctx, 6, 7) / steps
    out = dev.download_soa(dout)
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle
    O = oracle.Oracle()
    checked = [0, 1, int(np.argmax(np.diff(offs))), len(counts) - 1]
    same = True
    for e in checked:
        a, b = int(offs[e]), int(offs[e + 1])
        want = O.clue2d(packed["x"][a:b], packed["y"][a:b], packed["layer"][a:b], packed["weight"][a:b],
                        p.dc, p.rhoc, p.outlierDeltaFactor)
        same = same and all(np.array_equal(np.ascontiguousarray(out[k][a:b]).view(np.int32),
                                           np.ascontiguousarray(want[k]).view(np.int32)) for k in want)
    tuning = plan.tuning
    plan.destroy()
    for arr in list(din.values()) + list(dout.values()):
        arr.free()
    value = N / (ms * 1e-3) / 1e6
    log(f"shuffled hit order: {ms:.3f} ms/step ({value:.1f} Mhits/s), oracle events {checked} bit-exact={same}")
    return {"workload": "clue2d_batched_256_events_loguniform_1e3_1e6_shuffled_hit_order", "hits": N,
            "ms_per_step": round(ms, 4), "value": round(value, 2), "unit": "Mhits/s", "autotuned": tuning,
            "events_checked_against_oracle": checked, "labels_bit_exact_vs_oracle": bool(same)}


def pcie_leg(dev, packed, din, dout, plan, N, steps, batch_out):
    """The batched step with its host transfers: pinned host -> device copies of the inputs, the graph,
    device -> host copies of the outputs, all on the launch stream (PCIe-inclusive rate)."""
    lib, ctx = dev.lib, dev.ctx
    hin = {k: dev.pinned_copy(np.ascontiguousarray(packed[k], dtype=din[k].dtype)) for k in clue_hip.FIELDS_2D_IN}
    hout = {k: dev.pinned_copy(np.zeros(N, dtype=dout[k].dtype)) for k in clue_hip.FIELDS_2D_OUT}
    dev.sync()
    lib.clue_mark(ctx, 4, None)
    for _ in range(steps):
        for k in clue_hip.FIELDS_2D_IN:
            dev._chk(lib.clue_memcpy_htod(ctx, din[k].ptr, hin[k].ctypes.data, hin[k].nbytes, None))
        plan.launch()
        for k in clue_hip.FIELDS_2D_OUT:
            dev._chk(lib.clue_memcpy_dtoh(ctx, hout[k].ctypes.data, dout[k].ptr, hout[k].nbytes, None))
    lib.clue_mark(ctx, 5, None)
    dev.sync()
    dev.check()
    ms = lib.clue_elapsed_ms(ctx, 4, 5) / steps
    same = all(np.array_equal(hout[k].view(np.int32), batch_out[k].view(np.int32)) for k in hout)
    moved = sum(a.nbytes for a in hin.values()) + sum(a.nbytes for a in hout.values())
    log(f"pcie-inclusive: {ms:.3f} ms/step ({N / ms / 1e3:.1f} Mhits/s, {moved / ms / 1e6:.1f} GB/s over PCIe)")
    return {"value": round(N / (ms * 1e-3) / 1e6, 2), "unit": "Mhits/s", "ms_per_step": round(ms, 4),
            "bytes_over_pcie_per_step": moved, "steps": steps, "labels_bit_exact_vs_batched": bool(same)}


def c2_run(args):
    """One event of 1e6 hits (100 layers x 1e4) as a single-event plan: the configuration the per-kernel
    counter profiles of a single large event are taken on."""
    ev, offs = synth.pack([synth.clue2d_event(1_000_000, 0)])
    N = int(offs[-1])
    dev = clue_hip.Device(0)
    p = clue_hip.params2d(3.0, 10.0, 2.0)
    din = dev.upload_soa(ev, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_2D_OUT)
    plan = dev.clue2d_plan(din, dout, offs, p, use_graph=True, autotune=True)
    for _ in range(max(2, args.warmup)):
        plan.launch()
    dev.check()
    lib, ctx = dev.lib, dev.ctx
    dev.sync()
    lib.clue_mark(ctx, 0, None)
    for _ in range(args.steps):
        plan.launch()
    lib.clue_mark(ctx, 1, None)
    dev.sync()
    dev.check()
    ms = lib.clue_elapsed_ms(ctx, 0, 1) / args.steps
    algo = BYTES_PER_HIT_2D * N
    out = {"metric": "Mhits/s, one 1e6-hit event", "value": round(N / (ms * 1e-3) / 1e6, 2), "unit": "Mhits/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "clue2d_single_event_1e6_hits_100_layers", "hits": N, "launch": "hipGraph",
                      "autotuned": plan.tuning},
           "roofline": {"bound": "hbm", "achieved": round(algo / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes_per_launch": algo}}
    plan.destroy()
    dev.close()
    print(json.dumps(out), flush=True)


SERVE_BUDGET = 64 << 30  # HBM budget the serving loop admits events against


def result_ring(dev, pool, sets=48):
    """Pinned result buffers a front end reuses: `sets` sets, each large enough for any pool event."""
    n2 = max(ev["layer"].size for d, ev in pool if d == 2)
    n3 = max(ev["layer"].size for d, ev in pool if d == 3)
    ring = []
    for _ in range(sets):
        o2 = {k: dev.pinned_copy(np.zeros(n2, clue_hip._dt(k))) for k in clue_hip.FIELDS_2D_OUT}
        o3 = {k: dev.pinned_copy(np.zeros(n3, clue_hip._dt(k))) for k in clue_hip.FIELDS_3D_OUT}
        ring.append((o2, o3))
    return ring


def serving_run(dev, pool, order, arrivals, slots, max_events, p3, ring):
    """Open-loop Poisson stream: submit each event at its arrival time, wait for all. Results go to a
    ring of pinned buffers; a set is reused only after the request that last used it has completed
    (with more sets than requests in flight, that never waits)."""
    srv = clue_hip.Server(dev, p3=p3, max_points_2d=4_000_000, max_points_3d=1_000_000, max_events=max_events,
                          slots=slots)
    ids = []
    lat = {}
    owner = [None] * len(ring)
    t0 = time.perf_counter()
    for j, (k, t) in enumerate(zip(order, arrivals)):
        while time.perf_counter() - t0 < t:
            pass
        d, ev = pool[k]
        r = j % len(ring)
        if owner[r] is not None:
            lat[owner[r]] = srv.wait(owner[r], 120_000)[1]
        n = ev["layer"].size
        outs = {f: a[:n] for f, a in ring[r][0 if d == 2 else 1].items()}
        rid = srv.submit2d(ev, outs) if d == 2 else srv.submit3d(ev, outs)
        owner[r] = rid
        ids.append(rid)
    for i in ids:
        if i not in lat:
            lat[i] = srv.wait(i, 120_000)[1]
    lat = np.array([lat[i] for i in ids])
    span = time.perf_counter() - t0
    st = srv.stats()
    st["device_bytes_in_use"] = int(dev.lib.clue_bytes_in_use(dev.ctx))
    srv.close()
    return lat, span, st


def serving_leg(rate, seconds, seed=11):  # noqa: C901
    """C5: events arrive as a Poisson stream, 70 % CLUE2D (1e3-1e6 hits, log-uniform) and 30 % CLUE3D
    (50 layers x Poisson(200) layer clusters); the continuous-batching server (3 slots, up to 64 events
    per batch) and per-event processing (1 slot, 1 event per batch) see the same stream. Each runs
    on a context of its own with a 64 GB HBM budget, which its slots share and admit batches
    against."""
    dev = clue_hip.Device(0, budget_bytes=SERVE_BUDGET)
    rng = np.random.Generator(np.random.PCG64(seed))
    n2 = synth.c3_hit_counts(96, seed=seed)
    pool = [(2, synth.clue2d_event(int(n), 5000 + k)) for k, n in enumerate(n2)]
    pool += [(3, synth.clue3d_event(6000 + k)) for k in range(48)]
    # events wait in pinned host memory, as a front end receiving them by DMA would hold them
    pool = [(d, {k: dev.pinned_copy(v) for k, v in ev.items()}) for d, ev in pool]
    idx2 = [k for k, (d, _) in enumerate(pool) if d == 2]
    idx3 = [k for k, (d, _) in enumerate(pool) if d == 3]
    n_events = int(rate * seconds)
    order = [int(rng.choice(idx2)) if rng.random() < 0.7 else int(rng.choice(idx3)) for _ in range(n_events)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_events))
    hits = sum(pool[k][1]["layer"].size for k in order)
    p3 = clue_hip.params3d(S=2)
    out = {"rate_events_per_s": rate, "events": n_events, "hits": int(hits),
           "mix": "70% CLUE2D 1e3-1e6 hits log-uniform, 30% CLUE3D 50 layers x Poisson(200)",
           "latency": "submission to results in the caller's host buffers (H2D, kernels, D2H included)",
           "host_buffers": "inputs and results in pinned host memory (results: a ring of 48 reused sets)"}
    ring = result_ring(dev, pool)
    for name, slots, maxev in (("batched", 3, 64), ("per_event", 1, 1)):
        serving_run(dev, pool, order[:50], arrivals[:50] * 0.0, slots, maxev, p3, ring)  # warm the slots
        lat, span, st = serving_run(dev, pool, order, arrivals, slots, maxev, p3, ring)
        out[name] = {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                     "mean_ms": round(float(lat.mean()), 3), "mhits_per_s": round(hits / span / 1e6, 1),
                     "batches": st["batches"], "slots": slots, "max_events_per_batch": maxev,
                     "hbm_budget_bytes": SERVE_BUDGET, "device_bytes_in_use": st["device_bytes_in_use"]}
        log(f"serving {name} at {rate}/s: p50 {out[name]['p50_ms']} ms p99 {out[name]['p99_ms']} ms, "
            f"{st['batches']} batches")
    dev.close()
    return out


def copy_probe(nbytes):
    """Three device-to-device copies of nbytes each: the reference kernel whose FETCH_SIZE / WRITE_SIZE
    counters scale the counters of the clustering kernels to bytes."""
    dev = clue_hip.Device(0)
    a, b = dev.array(nbytes // 4, np.int32), dev.array(nbytes // 4, np.int32)
    for _ in range(3):
        dev._chk(dev.lib.clue_memcpy_dtod(dev.ctx, b.ptr, a.ptr, nbytes, None))
    dev.sync()
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}), flush=True)
    a.free()
    b.free()
    dev.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--events", type=int, default=256)
    ap.add_argument("--per-event-reps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-per-event", action="store_true", help="skip the per-event comparison leg")
    ap.add_argument("--serve-rates", default="500,1000,1500", help="serving leg: event rates (events/s), comma-separated")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="serving leg: length of the stream")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-3d", action="store_true", help="skip the CLUE3D (512-event) leg")
    ap.add_argument("--workload", choices=("c3", "c2", "c3s", "c4"), default="c3",
                    help="c3: the 256-event batch (default, the headline line); c2: only one 1e6-hit event; "
                         "c3s: the 256-event batch with its hits in random order as the timed step (profiling); "
                         "c4: only the 512-event CLUE3D batch (profiling)")
    ap.add_argument("--no-pcie", action="store_true", help="skip the PCIe-inclusive figure")
    ap.add_argument("--no-shuffled", action="store_true", help="skip the shuffled-hit-order leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run 3 device-to-device copies of BYTES (reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); events are independent, so N GPUs "
              f"would each run their own batch. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.workload == "c2":
        c2_run(args)
        return
    if args.workload == "c4":
        dev = clue_hip.Device(0)
        print(json.dumps(clue3d_leg(dev, 512, args.steps, args.warmup)), flush=True)
        dev.close()
        return

    counts = synth.c3_hit_counts(args.events)
    packed, offs = synth.clue2d_batch(counts, shuffle=args.workload == "c3s")
    N = int(offs[-1])
    log(f"workload: {args.events} events, {N} hits")
    dev = clue_hip.Device(0)
    p = clue_hip.params2d(3.0, 10.0, 2.0)
    din = dev.upload_soa(packed, clue_hip.FIELDS_2D_IN)
    dout = dev.alloc_soa(N, clue_hip.FIELDS_2D_OUT)
    # the first launch autotunes the density / nearest-higher variants on this batch, then the graph is recorded
    plan = dev.clue2d_plan(din, dout, offs, p, use_graph=True, autotune=True)
    for _ in range(max(2, args.warmup)):
        plan.launch()
    dev.check()
    tuning = plan.tuning
    lib, ctx = dev.lib, dev.ctx
    dev.sync()
    t0 = time.perf_counter()
    lib.clue_mark(ctx, 0, None)
    for _ in range(args.steps):
        plan.launch()
    lib.clue_mark(ctx, 1, None)
    dev.sync()
    wall = time.perf_counter() - t0
    dev.check()
    ms = lib.clue_elapsed_ms(ctx, 0, 1) / args.steps
    log(f"batched graph: {ms:.3f} ms/step (wall {wall * 1e3 / args.steps:.3f})")
    value = N / (ms * 1e-3) / 1e6
    batch_out = dev.download_soa(dout)

    result = {
        "metric": "Mhits/s batched vs per-event launches (bit-exact labels); p50/p99 latency",
        "value": round(value, 2), "unit": "Mhits/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "clue2d_batched_256_events_loguniform_1e3_1e6", "events": args.events, "hits": N,
                   "layers": 100, "dc": 3.0, "rhoc": 10.0, "outlierDeltaFactor": 2.0, "launch": "hipGraph",
                   "autotuned": tuning},
    }
    algo_bytes = BYTES_PER_HIT_2D * N
    # the dominant kernel, re-run on the plan's stream between HIP events right after the timed steps
    kms = {"k_density2d": plan.kernel_ms(0, reps=10), "k_nh2d": plan.kernel_ms(1, reps=10)}
    kname = max(kms, key=kms.get)
    kachieved = algo_bytes / (kms[kname] * 1e-3) / 1e9
    ktraffic = measured_traffic(f"traffic_bytes_{kname}") if args.events == 256 else None
    achieved = algo_bytes / (ms * 1e-3) / 1e9
    traffic = measured_traffic() if args.events == 256 else None
    result["roofline"] = {"bound": "hbm", "achieved": round(kachieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(kachieved / HBM_PEAK_GBS, 4), "traffic": ktraffic,
                          "kernel": kname, "kernel_ms": round(kms[kname], 4),
                          "kernel_ms_all": {k: round(v, 4) for k, v in kms.items()},
                          "algorithmic_bytes_per_launch": algo_bytes,
                          "replay": {"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                                     "traffic": traffic,
                                     "kernels": "all kernels of one graph replay of the batched path"}}
    if traffic:
        result["roofline"]["traffic_source"] = str(TRAFFIC_PROFILE.relative_to(REPO))
        result["roofline"]["replay"]["traffic_over_algorithmic"] = round(traffic / algo_bytes, 2)
        result["roofline"]["replay"]["hbm_gbs_measured_traffic"] = round(traffic / (ms * 1e-3) / 1e9, 1)
    log(f"dominant kernel {kname}: {kms[kname]:.3f} ms ({kachieved:.0f} GB/s algorithmic), all {kms}")

    if not args.no_pcie:
        result["pcie_inclusive"] = pcie_leg(dev, packed, din, dout, plan, N, min(args.steps, 5), batch_out)

    if not args.no_shuffled:
        result["c3_shuffled"] = shuffled_leg(dev, counts, p, args.steps, args.warmup)

    if not args.no_per_event:
        pout = dev.alloc_soa(N, clue_hip.FIELDS_2D_OUT)
        t = per_event_run(dev, din, pout, offs, p, reps=args.per_event_reps)
        per_out = dev.download_soa(pout)
        same = all(np.array_equal(per_out[k].view(np.int32), batch_out[k].view(np.int32)) for k in per_out)
        pe = N / t / 1e6
        result["per_event"] = {"value": round(pe, 2), "unit": "Mhits/s", "ms_total": round(t * 1e3, 3),
                               "launch": "per event: launch sequence + stream sync",
                               "labels_bit_exact_vs_batched": bool(same)}
        result["batched_over_per_event"] = round(value / pe, 2)
        log(f"per-event: {t * 1e3:.2f} ms total ({pe:.1f} Mhits/s), bit-exact={same}")

    if not args.no_3d:
        result["clue3d"] = clue3d_leg(dev, 512, args.steps, args.warmup)

    if not args.no_serve:
        result["serving"] = [serving_leg(float(r), args.serve_seconds) for r in args.serve_rates.split(",")]

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(packed, offs, p, batch_out)
        log(f"cpu baseline: {result['cpu_baseline']}")

    plan.destroy()
    dev.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
