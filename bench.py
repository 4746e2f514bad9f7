"""Benchmark of the photon transport and detection path on one MI355X.

Headline workload (the configuration the metric is quoted on): cylinder
phantom of 200^3 voxels over a 1 cm cube (water cylinder r = 0.4 cm with a
cortical bone rod r = 0.1 cm, in air), uniform F-18 activity in the cylinder
with positron range and annihilation, 1e8 decays with emission times over
10 s, 8 flat LSO panels (20 x 20 crystals of 1.5 x 1.5 x 10 mm) in an octagon
at 8 cm, 12% FWHM energy blur, 350-650 keV window, 300 ps time blur, 4 ns
coincidence window.  FP32 transport, FP64 time.

One step = one full run of the path over the workload's decays (decay
sampling, positron walk, phantom and panel transport, adder/readout,
digitizer, coincidence sorter), inputs resident in HBM, outputs kept in HBM.
Launch shapes are picked by the library's tuner during the warm-up run.

The same JSON line also carries, measured in the same process:
  host_sink   the headline step with singles and coincidences returned to
              pinned host buffers (the cost of handing the outputs over),
              with the copies overlapping the next run and without;
  surrogate   the detector-response surrogate on the headline workload:
              harvest of plain-MC panel crossings and GPU training (outside
              the timed region), then timed surrogate-mode steps against
              plain MC, and the accuracy block (per-crystal singles, energy
              spectrum);
  configs     the batched workload (64 blob phantoms of 32^3..256^3, seed 3,
              1.47e8 decays in one batched run) and the body workload
              (256 x 256 x 512 phantom, 4 rings x 12 LSO panels, 1e9 decays),
              each with its own roofline and CPU baseline;
  cpu_baseline the CPU oracle on this machine's cores and on one thread.

--workload batch|body makes that workload the headline instead.
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
SIMDS, CLOCK_HZ = 1024, 2.4e9      # MI355X: 256 CUs x 4 SIMDs, peak engine clock
QUEUE_BYTES = 6 * 4 + 8 + 16       # one queued photon: 6 floats, a double, an int4
MAP_BYTES = 16                     # one walk-map record (uniform box) per range-walk group


WORKLOADS = {
    # batch: histories per phantom-kernel launch (0: the library's 2^23).  The
    # cylinder workload's 1e8 decays run in three launches of 2^25 and the
    # batched workload's 1.47e8 in five: per-launch fixed costs (queue-region
    # prefixes, counter flushes, the last waves of each grid) then weigh less
    # (batched: 3.64e9 / 3.70e9 / 3.75e9 histories/s at 2^23 / 2^24 / 2^25).
    # The body workload runs in 2^24 (1.15e9 -> 1.17e9 histories/s against
    # 2^23): its event buffer (1.2 events per decay) and the digitizer's work
    # buffers leave no room for a 2^25 batch's headroom.
    "cylinder": dict(config=2, decays=1e8, seed=2, batch=1 << 25, counters="k_phantom_cylinder200_1e8_counters.md",
                     text="cylinder phantom 200^3 over 1 cm (water r=0.4, bone rod r=0.1, air), uniform F-18 "
                          "with positron range, %.0e decays over 10 s, 8-panel LSO octagon at 8 cm, "
                          "plain MC (no surrogate)"),
    "batch": dict(config=3, decays=0, seed=3, batch=1 << 25,
                  text="64 blob phantoms 32^3..256^3 over 10 cm (water/bone/lung/air, seed 3), F-18 cylinder "
                       "sources, 1e5..1e7 decays each (%.3g in total) over 10 s, one batched run on the 8-panel "
                       "octagon, plain MC (no surrogate)"),
    "body": dict(config=5, decays=1e9, seed=5, batch=1 << 24, harvest=(1, 1e8), surrogate_decays=2e8,
                 text="body phantom 256x256x512 at 1 mm (water ellipsoid 23x17x48 cm, lungs 0.3 g/cc, bone spine, "
                      "20 hot spheres 8:1), F-18 with positron range, %.0e decays over 60 s, 4 rings x 12 LSO "
                      "panels (20x20 crystals, 40 cm ring), plain MC (no surrogate)"),
}


def profile_figures(wl):
    """HBM bytes and VALU instructions per history of the phantom stage
    (k_phantom + k_longwalk) from the committed rocprofv3 counter summary of
    the workload, or Nones."""
    import re

    name = wl.get("counters")
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None, None, None
    txt = open(path).read()
    t = re.search(r"^traffic_bytes_per_history:\s*([0-9.eE+]+)", txt, re.M)
    v = re.search(r"^valu_insts_per_history:\s*([0-9.eE+]+)", txt, re.M)
    return (float(t.group(1)) if t else None, float(v.group(1)) if v else None, "profiles/" + name)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def setup(eng, wl, tabs, batch_data=None):
    """Load the workload into an engine; returns run(arg, **sink) -> RunResult
    (arg: decays, or the scale of every job's decays for the batched one)."""
    from gpet_amd import workloads as W

    if wl["config"] == 3:
        phantoms, jobs, src, iso = batch_data
        W.setup_batch(eng, tabs, phantoms)

        def run(scale=1.0, **kw):
            js = [dict(j, n_decays=max(1, int(round(j["n_decays"] * scale)))) for j in jobs]
            return eng.run_batch(js, src, iso, **kw)
        return run
    W.setup_engine(eng, tabs, config=wl["config"])
    src, iso, t0, t1 = W.workload_sources(wl["config"])

    def run(n, **kw):
        return eng.run_source(src, iso, t0, t1, n_decays=int(n), **kw)
    return run


class CpuBaseline:
    """The oracle (host scalar build, OpenMP over histories) on this
    machine's cores, timed on bounded samples of the same workload (batched
    workload: every job's decay count scaled by sample / whole)."""

    def __init__(self, wl, n_full, batch_data=None):
        import ctypes as C

        from gpet_amd import engine

        tabs = engine.load_tables()
        opt = engine.make_options(seed=wl["seed"], positron_range=1)
        self.eng = engine.Engine(opt, lib_path=os.path.join(ROOT, "oracle", "liboracle.so"), prefix="gpet_cpu_")
        self.set_threads = self.eng.lib.gpet_cpu_set_threads
        self.set_threads.argtypes = [C.c_void_p, C.c_int32]
        self.run = setup(self.eng, wl, tabs, batch_data)
        self.arg = (lambda n: n / n_full) if wl["config"] == 3 else (lambda n: n)
        self.n_full = n_full

    def time(self, n):
        t = time.perf_counter()
        r = self.run(self.arg(n))
        dt = time.perf_counter() - t
        return r.tallies.n_histories / dt, r.tallies.n_coinc / dt, dt, r.tallies.n_histories

    def rate(self, threads, seconds):
        """(rate, coincidence rate, seconds, decays) on a sample sized to
        about `seconds` of work (at most the whole workload)."""
        self.set_threads(self.eng.ctx, threads)
        probe = 100000 if threads > 1 else 20000
        self.time(probe // 10)                        # warm-up
        r0, _, _, _ = self.time(probe)
        n = int(min(max(r0 * seconds, probe), 4e8, self.n_full))
        return self.time(n)

    def close(self):
        self.eng.close()


def timed(run, step_arg, warmup, steps, **kw):
    """Warm-up runs, then `steps` timed runs; sums of the tallies."""
    for _ in range(warmup):
        run(step_arg, **kw)
    M = dict(t_k=[0.0, 0.0], l_k=[0, 0], h_timed=0, hist=0, coinc=0, b_alg=0.0, t_panel_step=0.0, singles=0)
    t_start = time.perf_counter()
    for _ in range(steps):
        r = run(step_arg, **kw)   # returns after device sync
        T = r.tallies
        M["hist"] += T.n_histories
        M["coinc"] += T.n_coinc
        M["singles"] += T.n_singles
        # HIP-event durations of the launches the library timed (the eager
        # first batch of a graph run, every batch otherwise)
        for k in range(2):
            M["t_k"][k] += T.kernel_seconds[k]
            M["l_k"][k] += T.kernel_launches[k]
        M["h_timed"] += T.kernel_histories
        M["b_alg"] += (4.0 * T.range_steps + MAP_BYTES * T.range_jumps + 8.0 * T.phantom_steps
                       + QUEUE_BYTES * T.n_panel_entries)
    M["wall"] = time.perf_counter() - t_start
    M["value"] = M["hist"] / M["wall"]
    M["last"] = T
    M["t_panel_step"] = M["t_k"][1] / max(M["l_k"][1], 1) * T.launches
    return M


def roofline(M, wl):
    """The dominant stage (decay + positron walk + phantom transport kernel
    and the kernel that finishes its long positron walks, launched back to
    back and timed together by the library's HIP events): algorithmic bytes
    per history (4 B per positron-walk voxel, 16 B per walk-map record, 8 B
    per Woodcock step in the phantom, one queue record per panel entry)
    times the histories of an average timed launch, over its average
    duration; and the VALU issue fraction from the committed counters."""
    avg_launch = M["t_k"][0] / max(M["l_k"][0], 1)
    h_launch = M["h_timed"] / max(M["l_k"][0], 1)
    bytes_per_launch = M["b_alg"] / max(M["hist"], 1) * h_launch
    achieved = bytes_per_launch / avg_launch / 1e9 if avg_launch > 0 else 0.0
    tpb, vpb, src = profile_figures(wl)
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": tpb * h_launch if tpb else None, "traffic_from": src,
           "kernel": "k_phantom + k_longwalk", "bytes_per_launch": bytes_per_launch,
           "avg_launch_ms": 1e3 * avg_launch, "histories_per_launch": h_launch}
    if vpb and avg_launch > 0:
        # the stage's VALU instruction stream at 4 cycles per wave64
        # instruction on every SIMD, over the measured launch: near or above
        # 1 means instruction issue, not memory, bounds the stage (above 1:
        # some instructions issue in fewer than 4 cycles)
        out["issue"] = {"valu_insts_per_history": vpb,
                        "valu_cycles_over_launch": vpb * h_launch * 4.0 / (SIMDS * CLOCK_HZ * avg_launch),
                        "from": src}
    return out


def summary(M, launch):
    last = M["last"]
    return {"value": M["value"], "unit": "histories/s", "coincidences_per_s": M["coinc"] / M["wall"],
            "ms_per_step": 1e3 * M["wall"] / max(M["steps"], 1), "singles_per_step": last.n_singles,
            "coincidences_per_step": last.n_coinc,
            "kernel_avg_ms": {"phantom": 1e3 * M["t_k"][0] / max(M["l_k"][0], 1),
                              "panel": 1e3 * M["t_k"][1] / max(M["l_k"][1], 1),
                              "timed_launches": M["l_k"][0], "launches_per_step": last.launches},
            "launch": launch, "graph": bool(last.graph), "jobs": last.n_jobs}


def host_sink_step(eng, wl, step_arg, warmup, steps, last):
    """The step with singles and coincidences copied into pinned host
    buffers (allocated once, reused): synchronously at the end of every run,
    and asynchronously (gpet_sink.async_copy), where each run's copies run on
    the copy stream beside the next run's transport and the clock stops
    after gpet_sink_wait for the last run's."""
    import ctypes as C

    from gpet_amd import capi, hipmem
    from gpet_amd import workloads as W

    ns = int(last.n_singles * 1.05) + 1024
    nc = int(last.n_coinc * 1.05) + 1024
    sb = hipmem.PinnedBuffer(ns * 48)
    cb = hipmem.PinnedBuffer(2 * nc * 48)
    sink = capi.Sink()
    sink.singles, sink.singles_cap = C.c_void_p(sb.ptr), ns
    sink.coinc, sink.coinc_cap = C.c_void_p(cb.ptr), nc
    src, iso, t0, t1 = W.workload_sources(wl["config"])
    S = (capi.Source * len(src))(*src)
    I = (capi.Isotope * len(iso))(*iso)

    def step():
        eng._ok(eng.f["run_source"](eng.ctx, S, len(src), I, len(iso), float(t0), float(t1), int(step_arg),
                                    C.byref(sink)), "run_source")

    res = {}
    for mode in ("sync", "async"):
        sink.async_copy = 1 if mode == "async" else 0
        for _ in range(warmup):
            step()
        eng.sink_wait()
        t = time.perf_counter()
        for _ in range(steps):
            step()
        eng.sink_wait()
        res[mode] = time.perf_counter() - t
    nbytes = 48 * (min(sink.n_singles, ns) + 2 * min(sink.n_coinc, nc))
    sb.free()
    cb.free()
    dt = res["async"]
    return {"value": steps * int(step_arg) / dt, "unit": "histories/s", "ms_per_step": 1e3 * dt / steps,
            "sync_value": steps * int(step_arg) / res["sync"], "sync_ms_per_step": 1e3 * res["sync"] / steps,
            "bytes_per_step": nbytes, "singles": int(sink.n_singles), "coincidences": int(sink.n_coinc),
            "note": "singles and coincidence pairs copied to pinned host buffers: value with each run's copies "
                    "overlapping the next run's transport (async_copy, clock stopped after gpet_sink_wait), "
                    "sync_value copying at the end of every run"}


def train_surrogate(eng, wl, args):
    """Harvest panel crossings from plain runs of the workload and train the
    surrogate on the GPU (outside the timed region).  The workload may name
    its own harvest (runs, decays per run): the body scanner's 1e9-decay run
    has ~1.3 in-range entries per decay, so one run of 1e8 fills the set."""
    from gpet_amd import workloads as W

    src, iso, t0, t1 = W.workload_sources(wl["config"])
    t = time.perf_counter()
    n_runs, n_dec = wl.get("harvest", (int(args.harvest_runs), int(wl["decays"])))
    for _ in range(n_runs):
        eng.surrogate_harvest(src, iso, t0, t1, n_decays=int(n_dec), capacity=int(1.5e8))
    t_h = time.perf_counter() - t
    rep = eng.surrogate_train(steps=args.train_steps, batch=65536, lr=1e-3, seed=1, reinit=True)
    return {"records": rep.records, "harvest_runs": n_runs, "harvest_s": t_h, "train_steps": rep.steps,
            "train_s": rep.seconds, "train_tflops": rep.flops / max(rep.seconds, 1e-9) / 1e12,
            "loss_first": rep.loss_first, "loss_last": rep.loss_last}


def independent_context(eng, wl, tabs):
    """A second context of the workload with another Philox key (seed + 1000),
    the first one's launch shape and its trained network (saved and loaded,
    with the digitizer settings it was trained for): its surrogate-mode runs
    share no random numbers with the plain runs of the first."""
    import tempfile

    from gpet_amd import engine

    other = engine.Engine(engine.make_options(seed=wl["seed"] + 1000, positron_range=1, batch=eng.options.batch))
    run = setup(other, wl, tabs)
    L = eng.get_launch()
    other.set_launch(block=tuple(L["block"]), wpe=tuple(L["wpe"]), blocks_per_cu=tuple(L["blocks_per_cu"]),
                     batch=L["batch"], transit=tuple(L["transit"]))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "surrogate.bin")
        eng.surrogate_save(path)
        other.surrogate_load(path)
    return other, run


def surrogate_accuracy(eng, run, step_arg, n_panels, wl, tabs):
    """Per-crystal singles and the singles energy spectrum of one plain run
    of the workload and one surrogate-mode run with independent random
    numbers (a second context, another seed): chi-square homogeneity of the
    per-crystal counts and of the spectrum (1 keV bins), two-sample KS of the
    spectrum from the binned CDFs."""
    import numpy as np
    from scipy import stats

    res = {}
    eng.surrogate_enable(False)
    r = run(step_arg)
    t, cs, eh = eng.tallies(n_crystals=n_panels * 400, n_ebins=300, e_lo=350e3, e_hi=650e3)
    res[False] = (r.tallies.n_singles, r.tallies.n_coinc, cs.astype(float), eh.astype(float))
    other, orun = independent_context(eng, wl, tabs)
    try:
        other.surrogate_enable(True)
        r = orun(step_arg)
        t, cs, eh = other.tallies(n_crystals=n_panels * 400, n_ebins=300, e_lo=350e3, e_hi=650e3)
        res[True] = (r.tallies.n_singles, r.tallies.n_coinc, cs.astype(float), eh.astype(float))
    finally:
        other.close()
    (sp, cp, csp, ehp), (ss, cs_, css, ehs) = res[False], res[True]
    out = {"seeds": "plain MC with the workload's seed, surrogate mode with seed + 1000 (independent samples)",
           "singles_ratio": ss / max(sp, 1), "coincidences_ratio": cs_ / max(cp, 1)}
    panel_p = csp.reshape(n_panels, -1).sum(1)
    panel_s = css.reshape(n_panels, -1).sum(1)
    out["per_panel_max_rel_diff"] = float(np.max(np.abs(panel_s / panel_p - 1)))

    def homogeneity(a, b):
        keep = (a + b) > 0
        a, b = a[keep], b[keep]
        table = np.vstack([a, b])
        chi2, p, dof, _ = stats.chi2_contingency(table, correction=False)
        return float(chi2), int(dof), float(p)
    c2, dof, p = homogeneity(csp, css)
    out["per_crystal_chi2"] = {"chi2": c2, "dof": dof, "p": p}
    big = csp > 0.5 * csp.max()
    out["per_crystal_rel_diff_median"] = float(np.median(np.abs(css[big] / csp[big] - 1)))
    out["counts_per_crystal_median"] = float(np.median(csp[big]))
    c2, dof, p = homogeneity(ehp, ehs)
    out["energy_chi2"] = {"chi2": c2, "dof": dof, "p": p}
    cdf_p = np.cumsum(ehp) / ehp.sum()
    cdf_s = np.cumsum(ehs) / ehs.sum()
    d = float(np.max(np.abs(cdf_p - cdf_s)))
    ne = ehp.sum() * ehs.sum() / (ehp.sum() + ehs.sum())
    out["energy_ks_d"] = d
    out["energy_ks_p"] = float(stats.kstwo.sf(d, int(round(ne)))) if ne > 0 else None
    centres = np.linspace(350e3, 650e3, 301)[:-1] + 0.5e3
    out["energy_mean_rel_diff"] = float((centres @ ehs / ehs.sum()) / (centres @ ehp / ehp.sum()) - 1)
    return out


def surrogate_block(eng, run, step_arg, wl, args, plain, tabs):
    """The surrogate measured in the same run as plain MC: training outside
    the timed region, then the timed surrogate-mode steps and the accuracy
    runs.  plain: the plain-MC measurement at step_arg (None: measured here,
    after the headline's own timing, at the surrogate's decays)."""
    if plain is None:
        plain = timed(run, step_arg, 1, args.steps)
    surr = train_surrogate(eng, wl, args)
    eng.surrogate_enable(True)
    S = timed(run, step_arg, args.warmup, args.steps)
    S["steps"] = args.steps
    eng.surrogate_enable(False)
    last = S["last"]
    acc = surrogate_accuracy(eng, run, step_arg, 48 if wl["config"] == 5 else 8, wl, tabs)
    flops = SURR_FLOPS_PER_ROW * last.n_surrogate
    return dict(surr, value=S["value"], unit="histories/s", coincidences_per_s=S["coinc"] / S["wall"],
                ms_per_step=1e3 * S["wall"] / args.steps,
                plain_histories_per_s=plain["value"], plain_coincidences_per_s=plain["coinc"] / plain["wall"],
                histories_ratio=S["value"] / plain["value"],
                coincidences_ratio=(S["coinc"] / S["wall"]) / (plain["coinc"] / plain["wall"]),
                entries_per_step=last.n_surrogate, inference_ms_per_step=1e3 * last.seconds_surrogate,
                plain_panel_kernel_ms_per_step=1e3 * plain["t_panel_step"],
                plain_ms_per_step=1e3 * plain["wall"] / args.steps,
                amdahl_bound=1.0 / max(1e-9, 1.0 - plain["t_panel_step"] / (plain["wall"] / args.steps)),
                mfma_roofline={"bound": "mfma", "achieved": flops / max(last.seconds_surrogate, 1e-9) / 1e12,
                               "peak": 2500.0, "unit": "TFLOP/s",
                               "frac": flops / max(last.seconds_surrogate, 1e-9) / 2.5e15,
                               "kernel": "surrogate inference (network + sampling) per step"},
                accuracy=acc)


SURR_FLOPS_PER_ROW = 2.0 * (256 * 8 + 3 * 256 * 264 + 192 * 264)   # forward, stored shapes


def surrogate_fresh(name, args, tabs):
    """The surrogate measurement of a workload whose plain steps fill most
    of the GPU's memory (the body workload's 1e9-decay events and digitizer
    buffers): a new context at the workload's surrogate decays, plain MC and
    surrogate mode both timed there."""
    from gpet_amd import capi, engine

    wl = WORKLOADS[name]
    flags = (0 if args.no_tune else capi.FLAG_AUTOTUNE) | (capi.FLAG_NO_GRAPH if args.no_graph else 0)
    eng = engine.Engine(engine.make_options(seed=wl["seed"], positron_range=1, batch=wl.get("batch", 0), flags=flags))
    try:
        run = setup(eng, wl, tabs)
        n = int(wl["surrogate_decays"])
        out = surrogate_block(eng, run, n, wl, args, None, tabs)
        out["decays"] = n
        return out
    finally:
        eng.close()


def run_workload(name, args, tabs, tune=True, want_engine=False):
    """Plain-MC measurement of one workload; returns (summary, M, engine or
    None, run, step_arg, n_full, batch_data)."""
    from gpet_amd import capi, engine
    from gpet_amd import workloads as W

    wl = WORKLOADS[name]
    flags = (capi.FLAG_AUTOTUNE if tune else 0) | (capi.FLAG_NO_GRAPH if args.no_graph else 0)
    opt = engine.make_options(seed=wl["seed"], positron_range=1, batch=args.batch or wl.get("batch", 0), flags=flags)
    eng = engine.Engine(opt)
    batch_data = W.batch_workload() if wl["config"] == 3 else None
    if batch_data is not None:
        n_full = sum(j["n_decays"] for j in batch_data[1])
        n_dec = int(args.decays or n_full) if name == args.workload else n_full
        step_arg = n_dec / n_full
    else:
        n_dec = int(args.decays or wl["decays"]) if name == args.workload else int(wl["decays"])
        n_full = n_dec
        step_arg = n_dec
    run = setup(eng, wl, tabs, batch_data)
    if args.shape and name == args.workload:
        wp, wq, bp, bq = (int(v) for v in args.shape.split(","))
        cur = eng.get_launch()
        eng.set_launch(block=tuple(cur["block"]), wpe=(wp, wq), blocks_per_cu=(bp, bq), batch=cur["batch"],
                       transit=tuple(cur["transit"]))
    M = timed(run, step_arg, args.warmup, args.steps)
    M["steps"] = args.steps
    launch = eng.get_launch()
    S = summary(M, launch)
    S["workload"] = wl["text"] % n_dec
    S["decays"] = n_dec
    S["roofline"] = roofline(M, wl)
    if not want_engine:
        eng.close()
        eng = None
    return S, M, eng, run, step_arg, n_full, batch_data


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="cylinder")
    ap.add_argument("--decays", type=float, default=0.0, help="decays per step of the headline workload")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="length of the CPU baseline sample (all cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the library's default launch shape")
    ap.add_argument("--shape", default="", help="launch shape of the headline workload instead of the tuner's: "
                    "wpe_phantom,wpe_panel,blocks_per_cu_phantom,blocks_per_cu_panel (implies --no-tune)")
    ap.add_argument("--no-graph", action="store_true", help="launch batches one by one (profiling counters)")
    ap.add_argument("--no-configs", action="store_true", help="measure the headline workload only")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-host-sink", action="store_true")
    ap.add_argument("--harvest-runs", type=int, default=7, help="plain runs of the workload harvested for training")
    ap.add_argument("--train-steps", type=int, default=48000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not implemented (a run would shard "
              "its histories by index range; each range gives the same per-history results on any GPU).",
              file=sys.stderr)
        sys.exit(2)

    from gpet_amd import engine

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    tabs = engine.load_tables()
    name = args.workload
    wl = WORKLOADS[name]
    log("headline workload %s" % name)
    S, M, eng, run, step_arg, n_full, batch_data = run_workload(name, args, tabs,
                                                                tune=not (args.no_tune or args.shape),
                                                                want_engine=True)
    out = {
        "metric": METRIC,
        "value": S["value"],
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": S["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": S["workload"], "decays": S["decays"], "launch": S["launch"], "graph": S["graph"],
                   "jobs": S["jobs"], "precision": "fp32 transport, fp64 time"},
        "coincidences_per_s": S["coincidences_per_s"],
        "singles_per_step": S["singles_per_step"],
        "coincidences_per_step": S["coincidences_per_step"],
        "kernel_avg_ms": S["kernel_avg_ms"],
        "roofline": S["roofline"],
    }
    if not args.no_host_sink and wl["config"] != 3:
        log("host sink")
        out["host_sink"] = host_sink_step(eng, wl, step_arg, args.warmup, args.steps, M["last"])
    if not args.no_surrogate and wl["config"] != 3:
        log("surrogate: harvest, train, timed steps, accuracy")
        if "surrogate_decays" in wl:
            eng.close()
            eng = None
            out["surrogate"] = surrogate_fresh(name, args, tabs)
        else:
            out["surrogate"] = surrogate_block(eng, run, step_arg, wl, args, M, tabs)
    if eng is not None:
        eng.close()
    if not args.no_cpu:
        log("cpu baseline (all cores, one thread)")
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cb = CpuBaseline(wl, n_full, batch_data)
        rate, crate, dt, n_done = cb.rate(threads, args.cpu_seconds)
        rate1, crate1, dt1, n1 = cb.rate(1, args.cpu_seconds / 2)
        cb.close()
        out["cpu_baseline"] = {"value": rate, "unit": "histories/s", "cores": threads, "kind": "port",
                               "sample": "%d decays of the same workload (%.1f s), oracle build -O3 OpenMP"
                                         % (n_done, dt),
                               "coincidences_per_s": crate, "cpu_model": cpu_model(),
                               "single_thread": {"value": rate1, "coincidences_per_s": crate1, "cores": 1,
                                                 "sample": "%d decays (%.1f s)" % (n1, dt1)}}
        out["gpu_over_cpu"] = out["value"] / rate
        out["gpu_over_cpu_1thread"] = out["value"] / rate1
    if not args.no_configs:
        out["configs"] = {}
        for other in ("batch", "body"):
            if other == name:
                continue
            log("config %s" % other)
            So, _, _, _, _, nf, bd = run_workload(other, args, tabs, tune=not args.no_tune)
            if "surrogate_decays" in WORKLOADS[other] and not args.no_surrogate:
                log("config %s: surrogate" % other)
                So["surrogate"] = surrogate_fresh(other, args, tabs)
            if not args.no_cpu:
                threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
                cb = CpuBaseline(WORKLOADS[other], nf, bd)
                rate, crate, dt, n_done = cb.rate(threads, args.cpu_seconds / 2)
                cb.close()
                So["cpu_baseline"] = {"value": rate, "unit": "histories/s", "cores": threads, "kind": "port",
                                      "sample": "%d decays of the same workload (%.1f s)" % (n_done, dt),
                                      "coincidences_per_s": crate}
                So["gpu_over_cpu"] = So["value"] / rate
            out["configs"][other] = So
    print(json.dumps(out))


if __name__ == "__main__":
    main()
