"""The surrogate on the GPU: its BF16 MFMA forward and backward against the
oracle's FP64 restatement of the network, training on harvested crossings,
and surrogate-mode runs against plain MC."""
import os
import tempfile

import numpy as np
import pytest

from conftest import hip_engine, oracle_engine, same_bytes
from test_surrogate import layer_slices, masked_params, random_params, sample_rows

pytestmark = pytest.mark.gpu


def pair(tables):
    from gpet_amd import engine
    from gpet_amd import workloads as W
    g, o = hip_engine(engine.make_options()), oracle_engine(engine.make_options())
    for e in (g, o):
        W.setup_engine(e, tables, config=2, phantom_n=20)
    return g, o


def test_forward_matches_oracle(tables):
    """BF16 weights/activations with FP32 accumulation (GPU) against the same
    BF16 roundings with FP64 accumulation (oracle)."""
    g, o = pair(tables)
    try:
        w = masked_params(random_params(layer_slices()[1], scale=0.06, seed=3))
        g.surrogate_set_params(w)
        o.surrogate_set_params(w)
        x, _ = sample_rows(1000, seed=9)
        hg, ho = g.surrogate_forward(x), o.surrogate_forward(x)
        assert np.abs(hg - ho).max() <= 5e-3 * np.abs(ho).max()
        # rows are independent: a permutation permutes the outputs exactly
        perm = np.random.default_rng(0).permutation(len(x))
        assert np.array_equal(g.surrogate_forward(x[perm]), hg[perm])
    finally:
        g.close()
        o.close()


def test_gradient_matches_oracle(tables):
    """Loss and every layer's parameter gradient (backward GEMMs, SiLU',
    split-K weight gradients) against the oracle."""
    g, o = pair(tables)
    try:
        w = masked_params(random_params(layer_slices()[1], scale=0.06, seed=4))
        g.surrogate_set_params(w)
        o.surrogate_set_params(w)
        x, y = sample_rows(5000, seed=5)     # > one 4096-row split of the weight-gradient GEMMs
        lg, gg = g.surrogate_grad(x, y)
        lo, go = o.surrogate_grad(x, y)
        assert lg == pytest.approx(lo, rel=1e-4)
        for off, n, kp, k in layer_slices()[0]:
            a, b = gg[off:off + n * kp], go[off:off + n * kp]
            assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b)
    finally:
        g.close()
        o.close()


# training as the benchmark does it: 7 plain runs of the cylinder workload
# (1e8 decays each, ~1.1e8 in-range panel entries), 48000 Adam steps
HARVEST_RUNS, HARVEST_DECAYS, TRAIN_STEPS = 7, int(1e8), 48000


@pytest.fixture(scope="module")
def trained(tables):
    """Harvest C2 crossings, train on the GPU; yields the engine."""
    from gpet_amd import engine
    from gpet_amd import workloads as W
    e = hip_engine(engine.make_options(seed=2, positron_range=1, batch=1 << 25))
    W.setup_engine(e, tables, config=2)
    src, iso, t0, t1 = W.workload_sources(2)
    for _ in range(HARVEST_RUNS):
        e.surrogate_harvest(src, iso, t0, t1, n_decays=HARVEST_DECAYS, capacity=int(1.2e8))
    rep = e.surrogate_train(steps=TRAIN_STEPS, batch=65536, lr=1e-3, seed=1, reinit=True)
    yield e, rep, (src, iso, t0, t1)
    e.close()


def spectra(e, run, n_ebins=30):
    r = run()
    t, cs, eh = e.tallies(n_crystals=8 * 400, n_ebins=n_ebins, e_lo=350e3, e_hi=650e3)
    return r, cs.astype(float), eh.astype(float)


def test_harvest_records(trained):
    e, rep, _ = trained
    R = e.surrogate_records()
    assert len(R) == rep.records > 1e8
    ev = R[:, 6] > 0.5
    assert 0.45 < ev.mean() < 0.6                    # 1 - exp(-mu L) for 1 cm LSO, minus escapes
    assert 0.6 < (R[ev, 7] > 0.9999).mean() < 0.8    # photopeak fraction of detected entries
    assert np.all((R[:, 8] >= 0) & (R[:, 8] <= 1))
    assert rep.loss_last < rep.loss_first - 2.0


def test_surrogate_mode_against_plain(trained):
    """Same decays and phantom transport; in-range panel entries answered by
    the network (the fused inference kernel).  Regression bounds on what the
    trained network reproduces (singles, per panel, spectrum CDF), what an
    untrained one does not, and determinism.  The acceptance bar itself is
    the next test."""
    e, rep, (src, iso, t0, t1) = trained
    n = int(1e7)

    def run():
        return e.run_source(src, iso, t0, t1, n_decays=n, singles_cap=2 * n // 10)

    e.surrogate_enable(False)
    rp, cs_p, eh_p = spectra(e, run)
    e.surrogate_enable(True)
    rs, cs_s, eh_s = spectra(e, run)
    rs2 = run()
    assert same_bytes(rs.singles, rs2.singles)
    assert rs.tallies.n_surrogate > 0.9 * rp.tallies.n_panel_entries
    assert abs(rs.n_singles / rp.n_singles - 1) < 0.02
    panels_p, panels_s = cs_p.reshape(8, 400).sum(1), cs_s.reshape(8, 400).sum(1)
    assert np.max(np.abs(panels_s / panels_p - 1)) < 0.03
    cdf_p, cdf_s = np.cumsum(eh_p) / eh_p.sum(), np.cumsum(eh_s) / eh_s.sum()
    d_trained = np.max(np.abs(cdf_p - cdf_s))
    assert d_trained < 0.005
    # the network before training (same architecture and initialisation)
    w = e.surrogate_params()
    e.surrogate_train(steps=1, batch=1024, lr=1e-12, seed=1, reinit=True)
    ru, _, eh_u = spectra(e, run)
    d_untrained = np.max(np.abs(cdf_p - np.cumsum(eh_u) / eh_u.sum()))
    assert abs(ru.n_singles / rp.n_singles - 1) > 3 * abs(rs.n_singles / rp.n_singles - 1)
    assert d_untrained > 3 * d_trained
    e.surrogate_set_params(w)
    e.surrogate_enable(False)


def acceptance(cs_p, eh_p, cs_s, eh_s):
    """The acceptance figures: per-crystal singles (chi-square homogeneity of
    the two count vectors, and the median relative difference over crystals
    with at least 1e4 counts), the singles energy spectrum (two-sample KS on
    1 keV bins) and the total singles ratio."""
    from scipy import stats

    keep = (cs_p + cs_s) > 0
    _, p_cry, _, _ = stats.chi2_contingency(np.vstack([cs_p[keep], cs_s[keep]]), correction=False)
    big = cs_p >= 1e4
    med = float(np.median(np.abs(cs_s[big] / cs_p[big] - 1))) if big.any() else None
    d = float(np.max(np.abs(np.cumsum(eh_p) / eh_p.sum() - np.cumsum(eh_s) / eh_s.sum())))
    ne = eh_p.sum() * eh_s.sum() / (eh_p.sum() + eh_s.sum())
    p_ks = float(stats.kstwo.sf(d, int(round(ne))))
    return {"crystal_chi2_p": float(p_cry), "crystal_median_rel_diff": med, "energy_ks_d": d, "energy_ks_p": p_ks,
            "singles_ratio": float(cs_s.sum() / cs_p.sum())}


@pytest.fixture(scope="module")
def acceptance_figures(trained, tables):
    """Plain MC with one seed and surrogate mode with another, 1e9 decays
    each (~5.4e7 singles, ~1.7e4 per crystal, where the counting noise of a
    crystal's ratio is ~1.1%)."""
    from gpet_amd import engine
    from gpet_amd import workloads as W

    e, _, (src, iso, t0, t1) = trained
    n = int(1e9)
    path = os.path.join(tempfile.mkdtemp(), "surrogate.bin")
    e.surrogate_save(path)
    e.surrogate_enable(False)
    r, cs_p, eh_p = spectra(e, lambda: e.run_source(src, iso, t0, t1, n_decays=n), n_ebins=300)
    f = hip_engine(engine.make_options(seed=77, positron_range=1, batch=1 << 25))
    try:
        W.setup_engine(f, tables, config=2)
        f.surrogate_load(path)
        f.surrogate_enable(True)
        r2, cs_s, eh_s = spectra(f, lambda: f.run_source(src, iso, t0, t1, n_decays=n), n_ebins=300)
    finally:
        f.close()
    a = acceptance(cs_p, eh_p, cs_s, eh_s)
    print("surrogate acceptance:", a)
    return a


def test_surrogate_acceptance_bar(acceptance_figures):
    """The stated bar for surrogate mode on independent seeds: the singles
    energy spectrum with two-sample KS p > 0.01, total singles within 2%,
    and per-crystal singles within 2% of plain MC: chi-square homogeneity
    p > 0.01, or a median relative difference <= 2% over the crystals with
    at least 1e4 counts."""
    a = acceptance_figures
    assert a["energy_ks_p"] > 0.01, a
    assert abs(a["singles_ratio"] - 1) <= 0.02, a
    assert a["crystal_chi2_p"] > 0.01 or a["crystal_median_rel_diff"] <= 0.02, a


def test_surrogate_crystal_homogeneity(acceptance_figures):
    """Per-crystal chi-square homogeneity alone (p > 0.01) at ~1.7e4 counts
    per crystal: it needs the crystal ratios to agree to well under the
    1.1% counting noise.  Reported as an expected failure, with the figures,
    while the network's per-crystal differences exceed that."""
    a = acceptance_figures
    if not a["crystal_chi2_p"] > 0.01:
        pytest.xfail("per-crystal chi-square homogeneity: %s" % a)


def test_save_load_and_gate(trained, tables):
    from gpet_amd import engine
    from gpet_amd import workloads as W
    e, _, _ = trained
    path = os.path.join(tempfile.mkdtemp(), "surrogate.bin")
    e.surrogate_save(path)
    f = hip_engine(engine.make_options(seed=2, positron_range=1))
    try:
        W.setup_engine(f, tables, config=2)
        f.surrogate_load(path)
        assert np.array_equal(f.surrogate_params(), e.surrogate_params())
        x, _ = sample_rows(64, seed=2)
        assert np.array_equal(f.surrogate_forward(x), e.surrogate_forward(x))
        # a readout the network was not trained for is refused
        f.set_digitizer(W.digitizer(readout_policy=0, readout_depth=3))
        f.surrogate_enable(True)
        src, iso, t0, t1 = W.workload_sources(2)
        with pytest.raises(engine.GpetError):
            f.run_source(src, iso, t0, t1, n_decays=1000)
    finally:
        f.close()
