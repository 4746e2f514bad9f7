"""Benchmark of the photon transport and detection path on one MI355X.

Default workload (the configuration the metric is quoted on): cylinder phantom of
200^3 voxels over a 1 cm cube (water cylinder r = 0.4 cm with a cortical
bone rod r = 0.1 cm, in air), uniform F-18 activity in the cylinder with
positron range and annihilation, 1e8 decays with emission times over 10 s,
8 flat LSO panels (20 x 20 crystals of 1.5 x 1.5 x 10 mm) in an octagon at
8 cm, 12% FWHM energy blur, 350-650 keV window, 300 ps time blur, 4 ns
coincidence window.  FP32 transport, FP64 time.

--workload body runs the large case instead: a 256 x 256 x 512 body
phantom (ellipsoidal water body, lungs, bone spine, 20 hot spheres at 8:1)
in 4 rings x 12 LSO panels, 1e9 F-18 decays.  --workload batch runs 64
independent blob phantoms of 32^3..256^3 voxels (seed 3) with 1e5..1e7
decays each (log-uniform, 1.47e8 in total) as one batched run.

Launch shapes are picked by the library's tuner during the warm-up run.

One step = one full run of the path over the workload's decays (decay sampling,
phantom and panel transport, digitizer, coincidence sorter), inputs resident
in HBM, outputs kept in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "histories/sec + coincidences/sec on 1 MI355X, surrogate vs plain MC"
HBM_PEAK_GBS = 8000.0
QUEUE_BYTES = 6 * 4 + 8 + 16   # one queued photon: 6 floats, a double, an int4


WORKLOADS = {
    "cylinder": dict(config=2, decays=1e8, seed=2,
                     text="cylinder phantom 200^3 over 1 cm (water r=0.4, bone rod r=0.1, air), uniform F-18 "
                          "with positron range, %.0e decays over 10 s, 8-panel LSO octagon at 8 cm, "
                          "plain MC (no surrogate)"),
    "batch": dict(config=3, decays=0, seed=3,
                  text="64 blob phantoms 32^3..256^3 over 10 cm (water/bone/lung/air, seed 3), F-18 cylinder "
                       "sources, 1e5..1e7 decays each (%.3g in total) over 10 s, one batched run on the 8-panel "
                       "octagon, plain MC (no surrogate)"),
    "body": dict(config=5, decays=1e9, seed=5,
                 text="body phantom 256x256x512 at 1 mm (water ellipsoid 23x17x48 cm, lungs 0.3 g/cc, bone spine, "
                      "20 hot spheres 8:1), F-18 with positron range, %.0e decays over 60 s, 4 rings x 12 LSO "
                      "panels (20x20 crystals, 40 cm ring), plain MC (no surrogate)"),
}


def setup(eng, wl, tabs, batch_data=None):
    """Load the workload into an engine; returns run(n_decays) -> RunResult."""
    from gpet_amd import workloads as W

    if wl["config"] == 3:
        phantoms, jobs, src, iso = batch_data
        W.setup_batch(eng, tabs, phantoms)

        def run(scale=1.0):
            js = [dict(j, n_decays=max(1, int(round(j["n_decays"] * scale)))) for j in jobs]
            return eng.run_batch(js, src, iso)
        return run
    W.setup_engine(eng, tabs, config=wl["config"])
    src, iso, t0, t1 = W.workload_sources(wl["config"])

    def run(n):
        return eng.run_source(src, iso, t0, t1, n_decays=int(n))
    return run


def cpu_baseline(n_sample: int, n_full: int, threads: int, wl: dict, batch_data=None):
    """The oracle (host scalar build, OpenMP) timed on a bounded sample of the
    same workload on this machine's cores (batched workload: every job's
    decay count scaled by n_sample / n_full)."""
    from gpet_amd import engine
    import ctypes as C

    tabs = engine.load_tables()
    opt = engine.make_options(seed=wl["seed"], positron_range=1)
    eng = engine.Engine(opt, lib_path=os.path.join(ROOT, "oracle", "liboracle.so"), prefix="gpet_cpu_")
    f = eng.lib.gpet_cpu_set_threads
    f.argtypes = [C.c_void_p, C.c_int32]
    f(eng.ctx, threads)
    run = setup(eng, wl, tabs, batch_data)
    arg = (lambda n: n / n_full) if wl["config"] == 3 else (lambda n: n)
    run(arg(max(n_sample // 20, 1000)))   # warm-up
    t = time.perf_counter()
    r = run(arg(n_sample))
    dt = time.perf_counter() - t
    eng.close()
    return r.tallies.n_histories / dt, r.tallies.n_coinc / dt, dt, r.tallies.n_histories


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="cylinder")
    ap.add_argument("--decays", type=float, default=0.0, help="decays per step (default: the workload's)")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not part of this build "
              "(histories shard by index; see the design notes).", file=sys.stderr)
        sys.exit(2)

    from gpet_amd import capi, engine
    from gpet_amd import workloads as W

    wl = WORKLOADS[args.workload]
    tabs = engine.load_tables()
    opt = engine.make_options(seed=wl["seed"], positron_range=1, batch=args.batch, flags=capi.FLAG_AUTOTUNE)
    eng = engine.Engine(opt)
    batch_data = W.batch_workload() if wl["config"] == 3 else None
    if batch_data is not None:
        n_full = sum(j["n_decays"] for j in batch_data[1])
        n_dec = int(args.decays or n_full)
        step_arg = n_dec / n_full
    else:
        n_dec = int(args.decays or wl["decays"])
        n_full = n_dec
        step_arg = n_dec
    run = setup(eng, wl, tabs, batch_data)

    for _ in range(args.warmup):
        run(step_arg)
    t_k = [0.0, 0.0]
    l_k = [0, 0]
    h_timed = 0
    hist = coinc = 0
    b_alg = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r = run(step_arg)   # returns after device sync
        T = r.tallies
        hist += T.n_histories
        coinc += T.n_coinc
        # HIP-event durations of the launches the library timed (the eager
        # first batch of a graph run, every batch otherwise)
        for k in range(2):
            t_k[k] += T.kernel_seconds[k]
            l_k[k] += T.kernel_launches[k]
        h_timed += T.kernel_histories
        b_alg += 4.0 * T.range_steps + 8.0 * T.phantom_steps + QUEUE_BYTES * T.n_panel_entries
    wall = time.perf_counter() - t_start
    last = T
    launch = eng.get_launch()
    eng.close()

    value = hist / wall
    # dominant kernel: the decay + positron range + phantom kernel.
    # Algorithmic bytes per history (4 B per positron-walk voxel, 8 B per
    # Woodcock step in the phantom, one queue record per panel entry) times
    # the histories of an average timed launch, over its average duration.
    avg_launch = t_k[0] / max(l_k[0], 1)
    bytes_per_launch = b_alg / max(hist, 1) * h_timed / max(l_k[0], 1)
    achieved = bytes_per_launch / avg_launch / 1e9 if avg_launch > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["text"] % n_dec,
                   "decays": n_dec, "launch": launch, "graph": bool(last.graph), "jobs": last.n_jobs,
                   "precision": "fp32 transport, fp64 time"},
        "coincidences_per_s": coinc / wall,
        "singles_per_step": last.n_singles,
        "coincidences_per_step": last.n_coinc,
        "kernel_avg_ms": {"phantom": 1e3 * avg_launch, "panel": 1e3 * t_k[1] / max(l_k[1], 1),
                          "timed_launches": l_k[0], "histories_per_timed_launch": h_timed / max(l_k[0], 1),
                          "launches_per_step": last.launches},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_phantom", "bytes_per_launch": bytes_per_launch, "avg_launch_ms": 1e3 * avg_launch},
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        probe = 200000
        rate, _, dt, _ = cpu_baseline(probe, n_full, threads, wl, batch_data)
        n_sample = int(min(max(rate * args.cpu_seconds, probe), 4e8, n_full))
        rate, crate, dt, n_done = cpu_baseline(n_sample, n_full, threads, wl, batch_data)
        out["cpu_baseline"] = {"value": rate, "unit": "histories/s", "cores": threads, "kind": "port",
                               "sample": "%d decays of the same workload (%.1f s), oracle build -O3 OpenMP" % (n_done, dt),
                               "coincidences_per_s": crate}
        out["gpu_over_cpu"] = value / rate
    print(json.dumps(out))


if __name__ == "__main__":
    main()
