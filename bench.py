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
in 4 rings x 12 LSO panels, 1e9 F-18 decays.

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
    "body": dict(config=5, decays=1e9, seed=5,
                 text="body phantom 256x256x512 at 1 mm (water ellipsoid 23x17x48 cm, lungs 0.3 g/cc, bone spine, "
                      "20 hot spheres 8:1), F-18 with positron range, %.0e decays over 60 s, 4 rings x 12 LSO "
                      "panels (20x20 crystals, 40 cm ring), plain MC (no surrogate)"),
}


def cpu_baseline(n_sample: int, threads: int, wl: dict):
    """The oracle (host scalar build, OpenMP) timed on a bounded sample of the
    same workload on this machine's cores."""
    from gpet_amd import engine
    from gpet_amd import workloads as W
    import ctypes as C

    tabs = engine.load_tables()
    opt = engine.make_options(seed=wl["seed"], positron_range=1)
    eng = engine.Engine(opt, lib_path=os.path.join(ROOT, "oracle", "liboracle.so"), prefix="gpet_cpu_")
    f = eng.lib.gpet_cpu_set_threads
    f.argtypes = [C.c_void_p, C.c_int32]
    f(eng.ctx, threads)
    W.setup_engine(eng, tabs, config=wl["config"])
    src, iso, t0, t1 = W.workload_sources(wl["config"])
    eng.run_source(src, iso, t0, t1, n_decays=max(n_sample // 20, 1000))   # warm-up
    t = time.perf_counter()
    r = eng.run_source(src, iso, t0, t1, n_decays=n_sample)
    dt = time.perf_counter() - t
    eng.close()
    return n_sample / dt, r.tallies.n_coinc / dt, dt


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

    from gpet_amd import engine
    from gpet_amd import workloads as W

    wl = WORKLOADS[args.workload]
    n_dec = int(args.decays or wl["decays"])
    tabs = engine.load_tables()
    opt = engine.make_options(seed=wl["seed"], positron_range=1, batch=args.batch)
    eng = engine.Engine(opt)
    W.setup_engine(eng, tabs, config=wl["config"])
    src, iso, t0, t1 = W.workload_sources(wl["config"])

    for _ in range(args.warmup):
        eng.run_source(src, iso, t0, t1, n_decays=n_dec)
    t_k = [0.0, 0.0]
    l_k = [0, 0]
    hist = coinc = 0
    b_alg = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r = eng.run_source(src, iso, t0, t1, n_decays=n_dec)   # returns after device sync
        T = r.tallies
        hist += T.n_histories
        coinc += T.n_coinc
        for k in range(2):
            t_k[k] += T.kernel_seconds[k]
            l_k[k] += T.kernel_launches[k]
        b_alg += 4.0 * T.range_steps + 8.0 * T.phantom_steps + QUEUE_BYTES * T.n_panel_entries
    wall = time.perf_counter() - t_start
    last = T
    eng.close()

    value = hist / wall
    # dominant kernel: the decay + positron range + phantom kernel
    avg_launch = t_k[0] / max(l_k[0], 1)
    bytes_per_launch = b_alg / max(l_k[0], 1)
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
                   "decays": n_dec, "batch": args.batch or (1 << 22), "precision": "fp32 transport, fp64 time"},
        "coincidences_per_s": coinc / wall,
        "singles_per_step": last.n_singles,
        "coincidences_per_step": last.n_coinc,
        "kernel_ms": {"phantom": 1e3 * t_k[0] / args.steps, "panel": 1e3 * t_k[1] / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_phantom", "bytes_per_launch": bytes_per_launch, "avg_launch_ms": 1e3 * avg_launch},
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        probe = 200000
        rate, _, dt = cpu_baseline(probe, threads, wl)
        n_sample = int(min(max(rate * args.cpu_seconds, probe), 4e8))
        rate, crate, dt = cpu_baseline(n_sample, threads, wl)
        out["cpu_baseline"] = {"value": rate, "unit": "histories/s", "cores": threads, "kind": "port",
                               "sample": "%d decays of the same workload (%.1f s), oracle build -O3 OpenMP" % (n_sample, dt),
                               "coincidences_per_s": crate}
        out["gpu_over_cpu"] = value / rate
    print(json.dumps(out))


if __name__ == "__main__":
    main()
