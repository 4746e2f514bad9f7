"""Parity of the HIP path against the oracle on the benchmark workloads at
their stated grid sizes (decay counts scaled down so the oracle finishes in
seconds):

* the cylinder workload on its 200^3 grid over 1 cm (100k decays, hits on);
* the body workload on its full 256 x 256 x 512 grid in the 48-panel
  scanner (20k decays);
* the batched workload: 64 blob phantoms of 32^3..256^3 voxels (seed 3,
  3.1e8 voxels in 64 phantom slots), every job's decay count scaled by
  2e5 / 1.47e8 (at least one decay per job), as one batched run.

Every per-photon trace record, event, single, coincidence, tally and per-job
count must be bit-identical.
"""
import numpy as np
import pytest

from conftest import first_diff, hip_engine, oracle_engine, same_bytes
from test_parity_gpu import assert_parity, run_both

pytestmark = pytest.mark.gpu


def test_cylinder_200_parity(tables):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    opt = engine.make_options(seed=23, positron_range=1, record_hits=1)
    src, iso, t0, t1 = W.workload_sources(2)
    n = 100000
    a, b = run_both(tables, opt, lambda e: W.setup_engine(e, tables, config=2),
                    lambda e: e.run_source(src, iso, t0, t1, n_decays=n, singles_cap=n, coinc_cap=n,
                                           adder_cap=4 * n, hits_cap=20 * n, trace_cap=2 * n))
    assert a.tallies.n_singles > 1000 and a.tallies.n_coinc > 100
    # positron walks on the 50 um grid: voxel crossings and uniform-box jumps
    assert a.tallies.range_steps > 2 * n and a.tallies.range_jumps > n
    assert_parity(a, b, hits=True)


def test_body_full_grid_parity(tables):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    opt = engine.make_options(seed=29, positron_range=1, record_hits=1)
    src, iso, t0, t1 = W.workload_sources(5)
    n = 20000
    a, b = run_both(tables, opt, lambda e: W.setup_engine(e, tables, config=5),
                    lambda e: e.run_source(src, iso, t0, t1, n_decays=n, singles_cap=4 * n, coinc_cap=4 * n,
                                           adder_cap=8 * n, hits_cap=40 * n, trace_cap=2 * n))
    assert a.tallies.n_singles > 1000 and a.tallies.n_coinc > 100
    assert a.tallies.phantom_steps > 4 * n         # photons cross the body in many steps
    assert a.arrays[0][8 * 400:].sum() > 0
    assert_parity(a, b, hits=True)


def test_batch_full_size_parity(tables):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    phantoms, jobs, src, iso = W.batch_workload()
    assert len(phantoms) == 64
    sizes = [p[2][0] for p in phantoms]
    assert min(sizes) >= 32 and max(sizes) <= 256 and max(sizes) > 200
    total = sum(j["n_decays"] for j in jobs)
    scale = 2.0e5 / total
    jobs = [dict(j, n_decays=max(1, int(round(j["n_decays"] * scale)))) for j in jobs]
    n = sum(j["n_decays"] for j in jobs)
    caps = dict(singles_cap=n, coinc_cap=n, adder_cap=4 * n, trace_cap=2 * n)
    opt = engine.make_options(seed=31, positron_range=1)
    out = {}
    for name, mk in (("gpu", hip_engine), ("oracle", oracle_engine)):
        e = mk(opt)
        try:
            W.setup_batch(e, tables, phantoms)
            out[name] = e.run_batch(jobs, src, iso, **caps)
        finally:
            e.close()
    a, b = out["gpu"], out["oracle"]
    assert a.tallies.n_jobs == 64 and a.tallies.n_histories == n
    assert a.tallies.n_coinc > 100
    for f in ("trace", "adder", "singles", "coinc"):
        x, y = getattr(a, f), getattr(b, f)
        assert same_bytes(x, y), f + ": " + first_diff(x.ravel(), y.ravel())
    for f in ("n_photons", "n_panel_entries", "n_adder", "n_singles", "n_coinc", "n_coinc_true",
              "phantom_steps", "range_steps", "panel_steps", "dropped_events", "majorant_violations"):
        assert getattr(a.tallies, f) == getattr(b.tallies, f), f
    assert np.array_equal(a.job_counts, b.job_counts)
    assert (a.job_counts[:, 0] > 0).all()
