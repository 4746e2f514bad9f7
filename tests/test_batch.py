"""Batched runs on the oracle: one job equals a plain run, jobs never mix
in deadtime or coincidences, outputs come out in (job, time) order."""
import numpy as np
import pytest

from conftest import oracle_engine, same_bytes


def small_batch(n_phantoms=5, seed=7, decays=(20000, 40000)):
    from gpet_amd import workloads as W
    rng = np.random.default_rng(seed)
    phantoms = [W.blob_phantom(int(n), rng) for n in rng.integers(12, 41, n_phantoms)]
    counts = rng.integers(decays[0], decays[1], n_phantoms + 1)
    sources, jobs = [], []
    for k in range(n_phantoms + 1):
        slot = k if k < n_phantoms else 1          # the last job reuses slot 1
        sources.append(W.source(1, c3=4.0, c4=8.0, natom=1.0e12))
        jobs.append(dict(phantom=slot, first_source=k, n_sources=1, t0=0.0, t1=0.02 * (k + 1),
                         n_decays=int(counts[k])))
    return phantoms, jobs, sources, [W.isotope_f18()]


def job_of(job_first, eventid):
    return np.searchsorted(job_first, eventid, side="right") - 1


def test_one_job_equals_run_source(tables):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    opt = engine.make_options(seed=4, positron_range=1)
    src, iso, t0, t1 = W.workload_sources(2)
    n = 20000
    a = oracle_engine(opt)
    W.setup_engine(a, tables, config=2, phantom_n=50)
    ra = a.run_source(src, iso, t0, t1, n_decays=n, singles_cap=n, coinc_cap=n, adder_cap=4 * n, trace_cap=2 * n)
    b = oracle_engine(opt)
    W.setup_engine(b, tables, config=2, phantom_n=50)
    rb = b.run_batch([dict(phantom=0, first_source=0, n_sources=1, t0=t0, t1=t1, n_decays=n)], src, iso,
                     singles_cap=n, coinc_cap=n, adder_cap=4 * n, trace_cap=2 * n)
    for f in ("trace", "adder", "singles", "coinc"):
        assert same_bytes(getattr(ra, f), getattr(rb, f)), f
    assert rb.job_counts.tolist() == [[n, ra.n_singles, ra.n_coinc]]
    a.close()
    b.close()


@pytest.mark.parametrize("dead_time_us", [0.0, 1.5])
def test_jobs_are_separate(tables, dead_time_us):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    phantoms, jobs, src, iso = small_batch()
    dig = W.digitizer(dead_time_us=dead_time_us, dead_level=1, coinc_policy=1, coinc_window_us=0.05)
    e = oracle_engine(engine.make_options(seed=5, positron_range=1))
    W.setup_batch(e, tables, phantoms, dig=dig)
    n = sum(j["n_decays"] for j in jobs)
    r = e.run_batch(jobs, src, iso, singles_cap=n, coinc_cap=4 * n, adder_cap=4 * n)
    jf = np.concatenate([[0], np.cumsum([j["n_decays"] for j in jobs])])
    assert r.job_counts[:, 0].tolist() == [j["n_decays"] for j in jobs]
    assert r.job_counts[:, 1].sum() == r.n_singles and r.job_counts[:, 2].sum() == r.n_coinc
    js = job_of(jf, r.singles["eventid"])
    # (job, t) order
    assert np.all(np.diff(js) >= 0)
    for j in range(len(jobs)):
        t = r.singles["t"][js == j]
        assert np.all(np.diff(t) >= 0)
        assert len(t) == r.job_counts[j, 1]
        # a job's times lie in its own window (plus time of flight and blur)
        assert t.min() > -1.0 and t.max() < jobs[j]["t1"] * 1e6 + 1.0
    # both singles of a pair belong to one job
    jc = job_of(jf, r.coinc["eventid"])
    assert np.array_equal(jc[:, 0], jc[:, 1])
    assert r.n_coinc > 50
    e.close()
