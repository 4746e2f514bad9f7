"""Batched runs, graph replays, launch shapes and the tuner on the GPU, each
checked bit for bit against the oracle."""
import numpy as np
import pytest

from conftest import first_diff, hip_engine, oracle_engine, same_bytes
from test_batch import small_batch

pytestmark = pytest.mark.gpu


def check_same(a, b):
    for f in ("trace", "adder", "singles", "coinc"):
        x, y = getattr(a, f), getattr(b, f)
        assert same_bytes(x, y), f + ": " + first_diff(x.ravel(), y.ravel())
    for f in ("n_photons", "n_panel_entries", "n_hits", "n_adder", "n_singles", "n_coinc", "n_coinc_true",
              "phantom_steps", "range_steps", "panel_steps", "dropped_events", "majorant_violations"):
        assert getattr(a.tallies, f) == getattr(b.tallies, f), f


@pytest.fixture(scope="module")
def batch_case(tables):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    phantoms, jobs, src, iso = small_batch()
    dig = W.digitizer(dead_time_us=1.5, dead_level=1, coinc_policy=1, coinc_window_us=0.05)
    n = sum(j["n_decays"] for j in jobs)
    caps = dict(singles_cap=n, coinc_cap=4 * n, adder_cap=4 * n, trace_cap=2 * n)
    o = oracle_engine(engine.make_options(seed=5, positron_range=1))
    W.setup_batch(o, tables, phantoms, dig=dig)
    ref = o.run_batch(jobs, src, iso, **caps)
    o.close()
    return phantoms, jobs, src, iso, dig, caps, ref


def gpu_run(tables, case, flags=0, batch=6000, launch=None, repeat=1):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    phantoms, jobs, src, iso, dig, caps, _ = case
    e = hip_engine(engine.make_options(seed=5, positron_range=1, batch=batch, flags=flags))
    try:
        W.setup_batch(e, tables, phantoms, dig=dig)
        if launch:
            e.set_launch(**launch)
        out = []
        for _ in range(repeat):
            out.append(e.run_batch(jobs, src, iso, **caps))
            # eager work on the same stream between replays
            e.tallies(n_crystals=8 * 400, n_ebins=50, e_lo=0.0, e_hi=1e6)
        return out
    finally:
        e.close()


def test_batch_parity(tables, batch_case):
    """Six jobs on five phantom slots of 12^3..40^3 voxels, launches of 6000
    histories that straddle jobs and slots."""
    ref = batch_case[-1]
    (r,) = gpu_run(tables, batch_case)
    assert r.tallies.n_jobs == 6 and r.tallies.launches > 20
    check_same(r, ref)
    assert np.array_equal(r.job_counts, ref.job_counts)


def test_graph_replays(tables, batch_case):
    """The captured graph is replayed by later runs with the same arguments;
    every replay and the eager launch path give the oracle's results."""
    from gpet_amd import capi

    ref = batch_case[-1]
    runs = gpu_run(tables, batch_case, repeat=3)
    assert all(r.tallies.graph == 1 for r in runs)
    for r in runs:
        check_same(r, ref)
    (eager,) = gpu_run(tables, batch_case, flags=capi.FLAG_NO_GRAPH)
    assert eager.tallies.graph == 0
    check_same(eager, ref)
    # with transit rounds in the captured batches
    runs = gpu_run(tables, batch_case, repeat=2,
                   launch=dict(block=(256, 256), wpe=(0, 0), blocks_per_cu=(4, 4), batch=6000, transit=(1, 2)))
    assert all(r.tallies.graph == 1 for r in runs)
    for r in runs:
        check_same(r, ref)


@pytest.mark.parametrize("launch", [
    dict(block=(128, 128), wpe=(0, 0), blocks_per_cu=(1, 8)),
    dict(block=(256, 256), wpe=(6, 6), blocks_per_cu=(4, 4)),
    dict(block=(64, 192), wpe=(5, 5), blocks_per_cu=(2, 16)),
    dict(block=(256, 256), wpe=(0, 0), blocks_per_cu=(4, 4), transit=(1,)),
    dict(block=(128, 256), wpe=(5, 0), blocks_per_cu=(2, 4), transit=(1, 1, 2, 3)),
])
def test_launch_shapes(tables, batch_case, launch):
    ref = batch_case[-1]
    (r,) = gpu_run(tables, batch_case, launch=dict(launch, batch=6000))
    assert r.tallies.launch.as_dict()["block"] == list(launch["block"])
    check_same(r, ref)


def test_autotune(tables, batch_case):
    from gpet_amd import capi

    ref = batch_case[-1]
    first, second = gpu_run(tables, batch_case, flags=capi.FLAG_AUTOTUNE, repeat=2)
    assert first.tallies.seconds_tune > 0 and second.tallies.seconds_tune == 0
    check_same(first, ref)
    check_same(second, ref)


def test_batch_noise_and_delayed_parity(tables):
    """Background singles per job window and delayed-window pairs in a
    batched run: the oracle's singles, pairs, delayed pairs and job counts."""
    from gpet_amd import engine
    from gpet_amd import workloads as W

    phantoms, jobs, src, iso = small_batch()
    dig = W.digitizer(dead_time_us=1.5, dead_level=1, coinc_policy=1, coinc_window_us=0.05, noise_rate=3.0e6,
                      noise_e_mean=450e3, noise_e_sigma=40e3, delay_us=0.4)
    n = sum(j["n_decays"] for j in jobs)
    caps = dict(singles_cap=4 * n + 10**5, coinc_cap=4 * n + 10**5, adder_cap=4 * n, delayed_cap=4 * n + 10**5)
    out = {}
    for name, mk in (("gpu", hip_engine), ("oracle", oracle_engine)):
        e = mk(engine.make_options(seed=5, positron_range=1, batch=6000))
        try:
            W.setup_batch(e, tables, phantoms, dig=dig)
            out[name] = e.run_batch(jobs, src, iso, **caps)
        finally:
            e.close()
    a, b = out["gpu"], out["oracle"]
    assert a.tallies.n_noise == b.tallies.n_noise > 0
    assert a.tallies.n_delayed == b.tallies.n_delayed > 0
    for f in ("adder", "singles", "coinc", "delayed"):
        x, y = getattr(a, f), getattr(b, f)
        assert same_bytes(x, y), f + ": " + first_diff(x.ravel(), y.ravel())
    assert np.array_equal(a.job_counts, b.job_counts)
