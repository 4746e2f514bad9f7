"""Background singles (the reference's addnoise, gPET_kernals.cu:699-735) and
the delayed-window randoms estimate, on the oracle.

What pins them: the statistics of a Poisson stream (count, exponential
gaps, uniform crystals, the energy's mean and spread), the crystal-centre
geometry, an exact identity (a delay shorter than any time gap gives back
the prompt pairs) and the randoms estimate itself (on singles that are all
uncorrelated, the delayed pairs estimate the prompt pairs)."""
import numpy as np
import pytest
from scipy import stats

from conftest import oracle_engine


def noise_run(tables, rate, t1, n_decays=10, delay_us=0.0, policy=0, seed=4, window_us=4e-3, e_sigma=30e3):
    from gpet_amd import engine
    from gpet_amd import workloads as W
    dig = W.digitizer(e_threshold=0.0, r_ref=0.0, time_blur_fwhm_us=0.0, win_min=0.0, win_max=1.0e7,
                      coinc_window_us=window_us, coinc_policy=policy, noise_rate=rate, noise_e_mean=400e3,
                      noise_e_sigma=e_sigma, delay_us=delay_us)
    e = oracle_engine(engine.make_options(seed=seed))
    panels, _ = W.setup_engine(e, tables, config=2, phantom_n=40, dig=dig)
    src, iso, _, _ = W.workload_sources(2)
    cap = int(rate * t1 * 1.5) + 10 * n_decays + 1000
    r = e.run_source(src, iso, 0.0, t1, n_decays=n_decays, singles_cap=cap, coinc_cap=4 * cap,
                     adder_cap=4 * n_decays + 100, delayed_cap=4 * cap)
    e.close()
    return r, panels


def test_noise_stream_statistics(tables):
    from gpet_amd import engine
    rate, T = 2.0e6, 0.01                      # 2e4 expected singles over 10 ms
    r, panels = noise_run(tables, rate, T)
    s = r.singles
    nz = s[s["parn"] < 0]
    n = r.tallies.n_noise
    assert len(nz) == n                        # every background single passes the open chain
    mu = rate * T
    assert abs(n - mu) < 5 * np.sqrt(mu)
    # parn -1 .. -n, all in job 0 (eventid 0), times inside the window
    assert np.array_equal(np.sort(-nz["parn"]), np.arange(1, n + 1))
    assert np.all(nz["eventid"] == 0)
    t = np.sort(nz["t"])
    assert t[0] >= 0.0 and t[-1] < T * 1e6
    # exponential gaps of mean 1/rate (us)
    gaps = np.diff(t)
    ks = stats.kstest(gaps, "expon", args=(0, 1e6 / rate))
    assert ks.pvalue > 1e-3, ks
    # the arrival times are uniform over the window
    assert stats.kstest(t / (T * 1e6), "uniform").pvalue > 1e-3
    # uniform panels and crystals
    npan = len(panels)
    mod_n, cry_n, mod_ny, cry_ny = engine.panel_counts(panels[0])
    for key, k in (("pann", npan), ("cryn", cry_n)):
        h = np.bincount(nz[key], minlength=k)
        assert len(h) == k
        assert stats.chisquare(h).pvalue > 1e-3, key
    assert np.all(nz["modn"] == 0) and mod_n == 1
    # energy ~ N(400 keV, 30 keV) (no blur in this digitizer)
    E = nz["E"].astype(np.float64)
    assert abs(E.mean() - 400e3) < 5 * 30e3 / np.sqrt(n)
    assert abs(E.std() / 30e3 - 1) < 5 / np.sqrt(2 * n)
    # site: readout depth 2 (module level) -> pann * moduleN + modn
    assert np.array_equal(nz["siten"], nz["pann"] * mod_n + nz["modn"])


def test_noise_positions_are_crystal_centres(tables):
    """x is the middle of the crystal depth, y and z the centre of the
    crystal's cell on the crystal lattice (panel-local coordinates)."""
    from gpet_amd import engine
    r, panels = noise_run(tables, 1.0e6, 0.004)
    nz = r.singles[r.singles["parn"] < 0]
    P = panels[0]
    mod_n, cry_n, mod_ny, cry_ny = engine.panel_counts(P)
    assert np.allclose(nz["x"], P.dirx * P.lenx / 2)
    ly, lz = nz["cryn"] % cry_ny, nz["cryn"] // cry_ny
    y0 = -P.leny / 2 + ly * (P.cry + P.cspy)
    z0 = -P.lenz / 2 + lz * (P.crz + P.cspz)
    assert np.allclose(nz["y"], y0 + P.cry / 2, atol=1e-5)
    assert np.allclose(nz["z"], z0 + P.crz / 2, atol=1e-5)


def test_noise_is_deterministic_and_seeded(tables):
    a, _ = noise_run(tables, 5.0e5, 0.004, seed=4)
    b, _ = noise_run(tables, 5.0e5, 0.004, seed=4)
    c, _ = noise_run(tables, 5.0e5, 0.004, seed=5)
    assert a.singles.tobytes() == b.singles.tobytes()
    assert a.singles.tobytes() != c.singles.tobytes()


def test_noise_never_counts_as_true(tables):
    """Background singles carry the job's first history as eventid; a pair
    of two of them (or one with history 0's photon) is not a true pair."""
    r, _ = noise_run(tables, 5.0e8, 2e-5, window_us=4e-3)
    c = r.coinc
    both_real = (c[:, 0]["parn"] >= 0) & (c[:, 1]["parn"] >= 0)
    same = c[:, 0]["eventid"] == c[:, 1]["eventid"]
    assert r.tallies.n_coinc > 100
    assert r.tallies.n_coinc_true == int(np.sum(same & both_real))
    assert np.sum(same & ~both_real) > 0      # the case the rule exists for


@pytest.mark.parametrize("policy", [0, 1])
def test_tiny_delay_gives_the_prompt_pairs(tables, policy):
    """A delay shorter than every time gap between singles turns the delayed
    window [delay, delay + W] into the prompt window (0, W]: same pairs."""
    r, _ = noise_run(tables, 2.0e8, 1e-4, n_decays=2000, delay_us=1e-12, policy=policy)
    t = np.sort(r.singles["t"])
    assert np.min(np.diff(t)) > 1e-9            # the identity's premise
    assert r.tallies.n_coinc > 500
    assert r.tallies.n_delayed == r.tallies.n_coinc
    assert r.delayed.tobytes() == r.coinc.tobytes()


@pytest.mark.parametrize("policy", [0, 1])
def test_delayed_window_estimates_randoms(tables, policy):
    """Singles that are all uncorrelated (background only): every prompt pair
    is random, and the delayed window estimates their number."""
    r, _ = noise_run(tables, 1.0e8, 2e-4, n_decays=1, delay_us=0.5, policy=policy)
    p, d = r.tallies.n_coinc, r.tallies.n_delayed
    assert p > 1000
    assert abs(d - p) < 5 * np.sqrt(p + d), (p, d)
    # delayed partners lie delay..delay+W after their opener
    dt = r.delayed[:, 1]["t"] - r.delayed[:, 0]["t"]
    if policy == 0:
        assert np.all((dt >= 0.5) & (dt <= 0.5 + 4e-3))


def test_randoms_with_decays(tables):
    """Decays plus background: the delayed count tracks the random prompts
    (prompts minus trues) within the kill-multiples bias."""
    r, _ = noise_run(tables, 5.0e7, 2e-4, n_decays=100000, delay_us=0.5)
    T = r.tallies
    randoms = T.n_coinc - T.n_coinc_true
    assert T.n_coinc_true > 200 and randoms > 200
    assert abs(T.n_delayed - randoms) < 0.2 * randoms + 5 * np.sqrt(randoms)


def test_noise_in_batched_runs(tables):
    """Each job gets background singles in its own window, carrying its first
    history as eventid; the per-job counts include them."""
    from gpet_amd import engine
    from gpet_amd import workloads as W
    from test_batch import small_batch
    phantoms, jobs, src, iso = small_batch()
    dig = W.digitizer(noise_rate=2.0e6, noise_e_mean=450e3, noise_e_sigma=20e3, coinc_window_us=0.05,
                      delay_us=0.3)
    e = oracle_engine(engine.make_options(seed=5, positron_range=1))
    W.setup_batch(e, tables, phantoms, dig=dig)
    n = sum(j["n_decays"] for j in jobs)
    r = e.run_batch(jobs, src, iso, singles_cap=4 * n + 10**5, coinc_cap=4 * n + 10**5, delayed_cap=4 * n + 10**5)
    e.close()
    s = r.singles
    nz = s[s["parn"] < 0]
    first = np.cumsum([0] + [j["n_decays"] for j in jobs])[:-1]
    assert r.tallies.n_noise > 0 and len(nz) > 0
    assert set(np.unique(nz["eventid"])) <= set(first.tolist())
    for k, j in enumerate(jobs):
        sel = nz[nz["eventid"] == first[k]]
        assert len(sel) > 0
        assert np.all((sel["t"] >= j["t0"] * 1e6 - 1e-3) & (sel["t"] <= j["t1"] * 1e6 + 1e-3))
    assert r.job_counts[:, 1].sum() == r.n_singles
