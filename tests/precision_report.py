"""FP32-vs-FP64 error of the transport path.

Runs one workload through two CPU builds of the oracle (test
infrastructure): the FP32 build, which is bit-identical to the GPU path,
and the FP64 build of the same physics on the same Philox draws.  Histories
whose discrete path is the same in both (interaction counts, panel, number
of events) are compared value by value: deposited energy, interaction
position and timestamps of the events after adder/readout, and the phantom
exit state of every photon.

usage: python tests/precision_report.py [config] [decays]
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def run(lib, config, n, seed=17):
    sys.path.insert(0, ROOT)
    from gpet_amd import engine, workloads as W
    eng = engine.Engine(engine.make_options(seed=seed, positron_range=1 if config == 2 else 0),
                        lib_path=lib, prefix="gpet_cpu_")
    W.setup_engine(eng, engine.load_tables(), config=config)
    src, iso, t0, t1 = W.workload_sources(config)
    r = eng.run_source(src, iso, t0, t1, n_decays=n, adder_cap=4 * n, trace_cap=2 * n, singles_cap=n)
    eng.close()
    return r


def pct(x):
    x = np.abs(np.asarray(x, float))
    if len(x) == 0:
        return {"n": 0}
    return {"n": int(len(x)), "median": float(np.median(x)), "p99": float(np.percentile(x, 99)),
            "max": float(x.max())}


def report(config=2, n=200000):
    lo = os.path.join(ROOT, "oracle", "liboracle.so")
    hi = os.path.join(ROOT, "oracle", "liboracle64.so")
    a, b = run(lo, config, n), run(hi, config, n)
    ta, tb = a.trace, b.trace
    alive = (ta["status"] & 8) != 0
    keys = ("n_steps", "n_compton", "n_rayleigh", "panel", "n_int_panel", "n_events", "status")
    same = alive.copy()
    for k in keys:
        same &= ta[k] == tb[k]
    out = {"config": config, "decays": n, "photons": int(alive.sum()),
           "same_path_fraction": float(same.sum() / max(alive.sum(), 1))}
    # phantom exit state of photons with the same path
    dx = np.sqrt((ta["x"][same] - tb["x"][same]).astype(float) ** 2 + (ta["y"][same] - tb["y"][same]).astype(float) ** 2
                 + (ta["z"][same] - tb["z"][same]).astype(float) ** 2)
    out["exit_position_cm"] = pct(dx)
    out["exit_energy_rel"] = pct((ta["E"][same].astype(float) - tb["E"][same]) / tb["E"][same])
    out["exit_time_us"] = pct(ta["t"][same] - tb["t"][same])
    # events of photons with the same path, matched by (parn, siten)
    ok_parn = set()
    for h, k in zip(ta["history"][same], ta["photon"][same]):
        ok_parn.add(int(2 * h + k))
    ea = {(int(e["parn"]), int(e["siten"])): e for e in a.adder if int(e["parn"]) in ok_parn}
    eb = {(int(e["parn"]), int(e["siten"])): e for e in b.adder if int(e["parn"]) in ok_parn}
    common = sorted(set(ea) & set(eb))
    dE = [(float(ea[k]["E"]) - float(eb[k]["E"])) / float(eb[k]["E"]) for k in common if eb[k]["E"] > 0]
    dX = [np.sqrt(sum((float(ea[k][c]) - float(eb[k][c])) ** 2 for c in "xyz")) for k in common]
    dT = [float(ea[k]["t"]) - float(eb[k]["t"]) for k in common]
    out["events_matched"] = len(common)
    out["events_only_one_side"] = len(set(ea) ^ set(eb))
    out["deposited_energy_rel"] = pct(dE)
    out["interaction_position_cm"] = pct(dX)
    out["timestamp_us"] = pct(dT)
    return out


if __name__ == "__main__":
    cfg = int(sys.argv[1]) if len(sys.argv) > 1 else 2
    n = int(float(sys.argv[2])) if len(sys.argv) > 2 else 200000
    print(json.dumps(report(cfg, n), indent=1))
