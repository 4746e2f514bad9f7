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
    # batch: histories per phantom-kernel launch (0: the library's 2^23).  The
    # cylinder workload's 1e8 decays run in three launches of 2^25: per-launch
    # fixed costs (queue-region prefixes, counter flushes, the last waves of each
    # grid) then weigh less (2^25 measured 3.6% above 2^24).  The body workload
    # keeps the default: its event buffer (1.2 events per decay) plus a 2^24
    # batch's headroom does not fit with the digitizer's work buffers.
    "cylinder": dict(config=2, decays=1e8, seed=2, batch=1 << 25, counters="k_phantom_cylinder200_1e8_counters.md",
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


def measured_traffic(wl):
    """HBM bytes per history of the phantom stage (k_phantom + k_longwalk)
    from the committed rocprofv3 counter summary of this workload
    (FETCH_SIZE x 2.0 + WRITE_SIZE, scaled by a plain copy's counters), or
    None."""
    import re

    name = wl.get("counters")
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None, None
    m = re.search(r"^traffic_bytes_per_history:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return (float(m.group(1)), "profiles/" + name) if m else (None, None)


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


SURR_FLOPS_PER_ROW = 2.0 * (256 * 8 + 3 * 256 * 264 + 192 * 264)   # forward, stored shapes


def timed(eng, run, step_arg, warmup, steps):
    """Warm-up runs, then `steps` timed runs; sums of the tallies."""
    for _ in range(warmup):
        run(step_arg)
    M = dict(t_k=[0.0, 0.0], l_k=[0, 0], h_timed=0, hist=0, coinc=0, b_alg=0.0, t_panel_step=0.0)
    t_start = time.perf_counter()
    for _ in range(steps):
        r = run(step_arg)   # returns after device sync
        T = r.tallies
        M["hist"] += T.n_histories
        M["coinc"] += T.n_coinc
        # HIP-event durations of the launches the library timed (the eager
        # first batch of a graph run, every batch otherwise)
        for k in range(2):
            M["t_k"][k] += T.kernel_seconds[k]
            M["l_k"][k] += T.kernel_launches[k]
        M["h_timed"] += T.kernel_histories
        M["b_alg"] += 4.0 * T.range_steps + 8.0 * T.phantom_steps + QUEUE_BYTES * T.n_panel_entries
    M["wall"] = time.perf_counter() - t_start
    M["value"] = M["hist"] / M["wall"]
    M["last"] = T
    M["t_panel_step"] = M["t_k"][1] / max(M["l_k"][1], 1) * T.launches
    return M


def train_surrogate(eng, wl, args):
    """Harvest panel crossings from plain runs of the workload and train the
    surrogate on the GPU (outside the timed region)."""
    from gpet_amd import workloads as W

    src, iso, t0, t1 = W.workload_sources(wl["config"])
    t = time.perf_counter()
    n_runs = int(args.harvest_runs)
    for _ in range(n_runs):
        eng.surrogate_harvest(src, iso, t0, t1, n_decays=int(wl["decays"]), capacity=int(1.2e8))
    t_h = time.perf_counter() - t
    rep = eng.surrogate_train(steps=args.train_steps, batch=65536, lr=1e-3, seed=1, reinit=True)
    return {"records": rep.records, "harvest_runs": n_runs, "harvest_s": t_h, "train_steps": rep.steps,
            "train_s": rep.seconds, "train_tflops": rep.flops / max(rep.seconds, 1e-9) / 1e12,
            "loss_first": rep.loss_first, "loss_last": rep.loss_last}


def surrogate_accuracy(eng, run, step_arg, n_panels):
    """Per-crystal singles and the singles energy spectrum of one plain and
    one surrogate run of the workload (seed-matched decays)."""
    import numpy as np
    from scipy import stats

    out = {}
    res = {}
    for sur in (False, True):
        eng.surrogate_enable(sur)
        r = run(step_arg)
        t, cs, eh = eng.tallies(n_crystals=n_panels * 400, n_ebins=60, e_lo=350e3, e_hi=650e3)
        res[sur] = (r.tallies.n_singles, r.tallies.n_coinc, cs.astype(float), eh.astype(float))
    eng.surrogate_enable(False)
    (sp, cp, csp, ehp), (ss, cs_, css, ehs) = res[False], res[True]
    out["singles_ratio"] = ss / max(sp, 1)
    out["coincidences_ratio"] = cs_ / max(cp, 1)
    panel_p = csp.reshape(n_panels, -1).sum(1)
    panel_s = css.reshape(n_panels, -1).sum(1)
    out["per_panel_max_rel_diff"] = float(np.max(np.abs(panel_s / panel_p - 1)))
    big = csp > 0.5 * csp.max()
    out["per_crystal_rel_diff_median"] = float(np.median(np.abs(css[big] / csp[big] - 1)))
    # two-sample KS on the binned spectra (bin centres weighted by counts)
    centres = np.linspace(350e3, 650e3, 61)[:-1] + 2.5e3
    cdf_p = np.cumsum(ehp) / ehp.sum()
    cdf_s = np.cumsum(ehs) / ehs.sum()
    d = float(np.max(np.abs(cdf_p - cdf_s)))
    ne = ehp.sum() * ehs.sum() / (ehp.sum() + ehs.sum())
    out["energy_ks_d"] = d
    out["energy_ks_p"] = float(stats.kstwo.sf(d, int(round(ne)))) if ne > 0 else None
    out["energy_mean_rel_diff"] = float((centres @ ehs / ehs.sum()) / (centres @ ehp / ehp.sum()) - 1)
    return out


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
    ap.add_argument("--no-tune", action="store_true", help="keep the library's default launch shape")
    ap.add_argument("--no-graph", action="store_true", help="launch batches one by one (profiling counters)")
    ap.add_argument("--surrogate", action="store_true",
                    help="harvest + train the detector-response surrogate, then time surrogate mode against plain MC")
    ap.add_argument("--harvest-runs", type=int, default=7, help="plain runs of the workload harvested for training")
    ap.add_argument("--train-steps", type=int, default=8000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not part of this build "
              "(histories shard by index; see the design notes).", file=sys.stderr)
        sys.exit(2)

    from gpet_amd import capi, engine
    from gpet_amd import workloads as W

    wl = WORKLOADS[args.workload]
    tabs = engine.load_tables()
    flags = (0 if args.no_tune else capi.FLAG_AUTOTUNE) | (capi.FLAG_NO_GRAPH if args.no_graph else 0)
    opt = engine.make_options(seed=wl["seed"], positron_range=1, batch=args.batch or wl.get("batch", 0), flags=flags)
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

    surr = None
    if args.surrogate and wl["config"] == 3:
        print("--surrogate runs on the cylinder or body workload", file=sys.stderr)
        sys.exit(2)
    if args.surrogate:
        surr = train_surrogate(eng, wl, args)
    # plain MC first (the surrogate comparison needs it), then surrogate mode
    plain = timed(eng, run, step_arg, args.warmup, args.steps)
    S = None
    if surr is not None:
        eng.surrogate_enable(True)
        S = timed(eng, run, step_arg, args.warmup, args.steps)
        eng.surrogate_enable(False)
    launch = eng.get_launch()
    acc = surrogate_accuracy(eng, run, step_arg, 48 if wl["config"] == 5 else 8) if surr is not None else None
    eng.close()

    M = S if S is not None else plain
    last = M["last"]
    # dominant stage: the decay + positron range + phantom kernel and the
    # kernel that finishes its long positron walks (launched back to back,
    # timed together by the library's HIP events).  Algorithmic bytes per
    # history (4 B per positron-walk voxel, 8 B per Woodcock step in the
    # phantom, one queue record per panel entry) times the histories of an
    # average timed launch, over its average duration.
    avg_launch = M["t_k"][0] / max(M["l_k"][0], 1)
    bytes_per_launch = M["b_alg"] / max(M["hist"], 1) * M["h_timed"] / max(M["l_k"][0], 1)
    achieved = bytes_per_launch / avg_launch / 1e9 if avg_launch > 0 else 0.0
    tpb, tsrc = measured_traffic(wl) if S is None else (None, None)
    traffic = tpb * M["h_timed"] / max(M["l_k"][0], 1) if tpb else None
    mode = "surrogate (BF16 MLP detector response, plain fallback)" if S is not None else "plain MC (no surrogate)"
    out = {
        "metric": METRIC,
        "value": M["value"],
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * M["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (wl["text"] % n_dec).replace("plain MC (no surrogate)", mode),
                   "decays": n_dec, "launch": launch, "graph": bool(last.graph), "jobs": last.n_jobs,
                   "precision": "fp32 transport, fp64 time"},
        "coincidences_per_s": M["coinc"] / M["wall"],
        "singles_per_step": last.n_singles,
        "coincidences_per_step": last.n_coinc,
        "kernel_avg_ms": {"phantom": 1e3 * avg_launch, "panel": 1e3 * M["t_k"][1] / max(M["l_k"][1], 1),
                          "timed_launches": M["l_k"][0],
                          "histories_per_timed_launch": M["h_timed"] / max(M["l_k"][0], 1),
                          "launches_per_step": last.launches},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_from": tsrc,
                     "kernel": "k_phantom + k_longwalk", "bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": 1e3 * avg_launch},
    }
    if S is not None:
        flops = SURR_FLOPS_PER_ROW * last.n_surrogate
        out["surrogate"] = dict(surr, plain_histories_per_s=plain["value"],
                                plain_coincidences_per_s=plain["coinc"] / plain["wall"],
                                histories_ratio=M["value"] / plain["value"],
                                coincidences_ratio=(M["coinc"] / M["wall"]) / (plain["coinc"] / plain["wall"]),
                                entries_per_step=last.n_surrogate, inference_ms_per_step=1e3 * last.seconds_surrogate,
                                plain_panel_kernel_ms_per_step=1e3 * plain["t_panel_step"],
                                mfma_roofline={"bound": "mfma", "achieved": flops / max(last.seconds_surrogate, 1e-9) / 1e12,
                                               "peak": 2500.0, "unit": "TFLOP/s",
                                               "frac": flops / max(last.seconds_surrogate, 1e-9) / 2.5e15},
                                accuracy=acc)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        probe = 200000
        rate, _, dt, _ = cpu_baseline(probe, n_full, threads, wl, batch_data)
        n_sample = int(min(max(rate * args.cpu_seconds, probe), 4e8, n_full))
        rate, crate, dt, n_done = cpu_baseline(n_sample, n_full, threads, wl, batch_data)
        out["cpu_baseline"] = {"value": rate, "unit": "histories/s", "cores": threads, "kind": "port",
                               "sample": "%d decays of the same workload (%.1f s), oracle build -O3 OpenMP" % (n_done, dt),
                               "coincidences_per_s": crate}
        out["gpu_over_cpu"] = M["value"] / rate
    print(json.dumps(out))


if __name__ == "__main__":
    main()
