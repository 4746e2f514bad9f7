"""Digitizer chain of the oracle vs an O(n^2) numpy restatement on small
synthetic event lists (gPET.cu:380-424, gPET_kernals.cu:607-698,814-837,
plus the time blur and coincidence sorter the reference lacks)."""
import ctypes as C

import numpy as np
import pytest

from conftest import ORACLE_LIB, oracle_engine

SEED = 77


def normals(lib, parn, siten, k):
    out = np.zeros(k, np.float32)
    lib.gpet_cpu_stream(SEED, parn, 4 | (siten << 4), 1, k, out.ctypes.data_as(C.c_void_p))
    return out


def brute(events, dig, lib, module_n=1, crystal_n=400):
    """Straightforward restatement of the chain."""
    ev = []
    for e in events:
        e = e.copy()
        k = 2 + (6 if dig.space_blur > 0 else 0) + (2 if dig.time_blur_fwhm_us > 0 else 0)
        z = normals(lib, int(e["parn"]), int(e["siten"]), k // 2)
        E = np.float32(e["E"])
        if dig.blur_policy == 0:
            R = np.float32(np.sqrt(np.float32(dig.e_ref) / E)) * np.float32(dig.r_ref)
        else:
            R = np.float32(dig.r_ref) + np.float32(dig.slope) * (E - np.float32(dig.e_ref)) / np.float32(1e6)
        R = max(R, np.float32(0))
        e["E"] = E + ((z[0] * R) * E) / np.float32(2.35482)
        i = 1
        if dig.space_blur > 0:
            for f in ("x", "y", "z"):
                e[f] = np.float32(e[f]) + np.float32(dig.space_blur) * z[i]
                i += 1
        if dig.time_blur_fwhm_us > 0:
            e["t"] = e["t"] + float(z[i]) * (dig.time_blur_fwhm_us / 2.35482)
        if not (e["E"] < np.float32(dig.e_threshold) or e["E"] > np.float32(dig.e_upholder)):
            ev.append(e)
    n_thr = len(ev)

    def site(e):
        lv = dig.dead_level
        return {0: 0, 1: e["pann"], 2: e["pann"] * module_n + e["modn"]}.get(lv, e["siten"])

    for e in ev:   # setSitenum rewrites siten whenever the level is not 3
        e["siten"] = site(e)
    if dig.dead_time_us > 0:
        alive = []
        for s in sorted(set(site(e) for e in ev)):
            grp = sorted([e for e in ev if site(e) == s], key=lambda e: (e["t"], e["parn"], e["siten"]))
            tdead = None
            for e in grp:
                if tdead is not None and e["t"] < tdead + dig.dead_time_us:
                    if dig.dead_type == 0:
                        tdead = e["t"]
                    continue
                tdead = e["t"]
                alive.append(e)
        ev = alive
    n_dead = len(ev)
    sg = [e for e in ev if not (e["E"] < np.float32(dig.win_min) or e["E"] > np.float32(dig.win_max))]
    sg.sort(key=lambda e: (e["t"], e["parn"], e["siten"]))
    pairs = []
    W = dig.coinc_window_us
    i = 0
    while W > 0 and i < len(sg):
        j = i + 1
        while j < len(sg) and sg[j]["t"] - sg[i]["t"] <= W:
            j += 1
        m = j - i
        if m == 2 or (m > 2 and dig.coinc_policy == 1):
            for a in range(i, j):
                for b in range(a + 1, j):
                    pairs.append((sg[a], sg[b]))
        i = j
    return n_thr, n_dead, sg, pairs


def synthetic(n, rate_us, rng):
    from gpet_amd import capi
    ev = np.zeros(n, capi.EVENT_DTYPE)
    ev["parn"] = np.arange(n) * 2 + rng.integers(0, 2, n)
    ev["eventid"] = ev["parn"] // 2
    ev["pann"] = rng.integers(0, 8, n)
    ev["modn"] = 0
    ev["cryn"] = rng.integers(0, 400, n)
    ev["siten"] = ev["pann"]            # module-level readout
    ev["t"] = np.cumsum(rng.exponential(rate_us, n))
    # a few near-simultaneous pairs so coincidences exist
    k = rng.choice(n - 1, n // 4, replace=False)
    ev["t"][k + 1] = ev["t"][k] + rng.uniform(0, 0.006, len(k))
    ev["E"] = rng.uniform(20e3, 700e3, n).astype(np.float32)
    ev["x"] = rng.uniform(-1, 0, n)
    ev["y"] = rng.uniform(-1.5, 1.5, n)
    ev["z"] = rng.uniform(-1.5, 1.5, n)
    return ev


@pytest.mark.parametrize("case", ["plain", "paralyzable", "nonparalyzable", "takeall", "policy0_space"])
def test_digitizer_matches_brute_force(tables, case):
    from gpet_amd import capi, engine, workloads as W
    lib = C.CDLL(ORACLE_LIB)
    lib.gpet_cpu_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int32, C.c_int64, C.c_void_p]
    kw = {"plain": {}, "paralyzable": dict(dead_time_us=0.8, dead_type=0, dead_level=1),
          "nonparalyzable": dict(dead_time_us=0.8, dead_type=1, dead_level=3),
          "takeall": dict(coinc_policy=1, coinc_window_us=0.02),
          "policy0_space": dict(blur_policy=0, space_blur=0.1, r_ref=0.15, dead_time_us=0.3, dead_level=2)}[case]
    dig = W.digitizer(**kw)
    eng = oracle_engine(engine.make_options(seed=SEED))
    W.setup_engine(eng, tables, config=1, phantom_n=8, dig=dig)
    rng = np.random.default_rng(hash(case) % 1000)
    ev = synthetic(1500, 0.5, rng)
    s = capi.Sink()
    out_s = np.zeros(2000, capi.EVENT_DTYPE)
    out_c = np.zeros(8000, capi.EVENT_DTYPE)
    s.singles, s.singles_cap = out_s.ctypes.data_as(C.c_void_p), len(out_s)
    s.coinc, s.coinc_cap = out_c.ctypes.data_as(C.c_void_p), len(out_c) // 2
    f = eng.lib.gpet_cpu_digitize
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(capi.Sink)]
    assert f(eng.ctx, ev.ctypes.data_as(C.c_void_p), len(ev), C.byref(s)) == 0
    t, _, _ = eng.tallies()
    n_thr, n_dead, sg, pairs = brute(ev, dig, lib)
    assert t.n_threshold == n_thr and t.n_deadtime == n_dead
    got = out_s[: s.n_singles]
    want = np.array(sg, dtype=capi.EVENT_DTYPE)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    gp = out_c[: 2 * s.n_coinc].reshape(-1, 2)
    assert len(gp) == len(pairs) and len(pairs) > 0
    for (a, b), (pa, pb) in zip(gp, pairs):
        assert a.tobytes() == pa.tobytes() and b.tobytes() == pb.tobytes()
    if case in ("paralyzable", "nonparalyzable", "policy0_space"):
        assert n_dead < n_thr
    eng.close()


def test_energy_blur_width(tables):
    """Policy 1 with Rref = 12% at E: blurred energies have sigma = 0.12 E / 2.35482."""
    from gpet_amd import capi, engine, workloads as W
    dig = W.digitizer(e_threshold=0, win_min=0, win_max=2e6, coinc_window_us=0, time_blur_fwhm_us=300e-6)
    eng = oracle_engine(engine.make_options(seed=3))
    W.setup_engine(eng, tables, config=1, phantom_n=8, dig=dig)
    n = 100000
    ev = np.zeros(n, capi.EVENT_DTYPE)
    ev["parn"] = np.arange(n)
    ev["siten"] = 3
    ev["t"] = np.arange(n) * 10.0
    ev["E"] = 511e3
    s = capi.Sink()
    out = np.zeros(n, capi.EVENT_DTYPE)
    s.singles, s.singles_cap = out.ctypes.data_as(C.c_void_p), n
    f = eng.lib.gpet_cpu_digitize
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(capi.Sink)]
    assert f(eng.ctx, ev.ctypes.data_as(C.c_void_p), n, C.byref(s)) == 0
    E = out["E"].astype(float)
    assert abs(E.mean() - 511e3) < 4 * 0.12 * 511e3 / 2.35482 / np.sqrt(n)
    assert abs(E.std() / (0.12 * 511e3 / 2.35482) - 1) < 0.01
    dt = out["t"] - np.sort(ev["t"])
    assert abs(np.std(dt) / (300e-6 / 2.35482) - 1) < 0.01
    eng.close()
