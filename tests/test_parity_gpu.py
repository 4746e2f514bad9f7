"""Per-history parity of the HIP path against the CPU oracle.

Both run the same histories on the same Philox streams with the same
operation order, so every per-photon record, every event after
adder/readout, every single and every coincidence must be bit-identical.
"""
import numpy as np
import pytest

from conftest import first_diff, hip_engine, oracle_engine, same_bytes

pytestmark = pytest.mark.gpu


def run_both(tables, opt, setup, call):
    res = {}
    for name, mk in (("gpu", hip_engine), ("oracle", oracle_engine)):
        eng = mk(opt)
        try:
            setup(eng)
            res[name] = call(eng)
            res[name].arrays = tallies_of(eng)
        finally:
            eng.close()
    return res["gpu"], res["oracle"]


def tallies_of(eng):
    t, cs, eh, sg = eng.tallies(n_crystals=48 * 400, n_ebins=100, e_lo=0.0, e_hi=1e6, n_s=64, n_phi=48, s_max=8.0,
                                with_sinogram=True)
    return cs, eh, sg


def assert_parity(a, b, hits=False):
    assert a.tallies.n_histories == b.tallies.n_histories
    assert same_bytes(a.trace, b.trace), "trace: " + first_diff(a.trace, b.trace)
    assert same_bytes(a.adder, b.adder), "adder events: " + first_diff(a.adder, b.adder)
    assert same_bytes(a.singles, b.singles), "singles: " + first_diff(a.singles, b.singles)
    assert same_bytes(a.coinc, b.coinc), "coincidences: " + first_diff(a.coinc.ravel(), b.coinc.ravel())
    for x, y in zip(a.arrays, b.arrays):
        assert np.array_equal(x, y), "tally arrays differ"
    assert a.arrays[0].sum() == a.tallies.n_singles
    for f in ("n_photons", "n_panel_entries", "n_hits", "n_adder", "n_threshold", "n_deadtime", "n_singles",
              "n_coinc", "n_coinc_true", "dropped_events", "majorant_violations", "phantom_steps", "range_steps",
              "panel_steps"):
        assert getattr(a.tallies, f) == getattr(b.tallies, f), f
    if hits:
        # one lane writes a photon's hits in the order they happen: a stable
        # sort by photon id puts both lists in the same order
        ka = np.argsort(a.hit_id[:, 0], kind="stable")
        kb = np.argsort(b.hit_id[:, 0], kind="stable")
        assert np.array_equal(a.hit_id[ka], b.hit_id[kb])
        assert np.array_equal(a.hit_f[ka].view(np.uint32), b.hit_f[kb].view(np.uint32))


def with_transit(setup, transit):
    """setup, then (if given) the photon-regrouping schedule of the launch."""
    def f(e):
        setup(e)
        if transit is not None:
            L = e.get_launch()
            e.set_launch(block=tuple(L["block"]), wpe=tuple(L["wpe"]), blocks_per_cu=tuple(L["blocks_per_cu"]),
                         batch=L["batch"], transit=transit)
    return f


# transit: photons still in the phantom after the phantom kernel's first
# steps are continued by later rounds, compacted per block (results must not
# change)
@pytest.mark.parametrize("config,n,prange,transit", [(2, 30000, 1, None), (1, 30000, 0, None),
                                                     (2, 30000, 1, (1,)), (1, 30000, 0, (1, 1, 2))])
def test_source_parity(tables, config, n, prange, transit):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    opt = engine.make_options(seed=11 + config, positron_range=prange, record_hits=1, batch=7000)
    src, iso, t0, t1 = W.workload_sources(config)
    setup = with_transit(lambda e: W.setup_engine(e, tables, config=config, phantom_n=100 if config == 2 else 64),
                         transit)
    a, b = run_both(tables, opt, setup,
                    lambda e: e.run_source(src, iso, t0, t1, n_decays=n, singles_cap=n, coinc_cap=n, adder_cap=4 * n,
                                           hits_cap=20 * n, trace_cap=2 * n))
    assert a.tallies.n_singles > 100 and a.tallies.n_coinc > 10
    assert_parity(a, b, hits=True)


@pytest.mark.parametrize("transit", [None, (2, 3, 5, 8)])
def test_body_parity(tables, transit):
    """Body workload (ellipsoid background + hot spheres, lungs, spine) in
    the 4-ring 48-panel scanner, on a 32 x 32 x 64 grid of the same extent
    (photons take many steps in it: with transit rounds most of them are
    handed on several times)."""
    from gpet_amd import engine
    from gpet_amd import workloads as W

    opt = engine.make_options(seed=5, positron_range=1, record_hits=1, batch=5000)
    src, iso, t0, t1 = W.workload_sources(5)
    n = 12000
    a, b = run_both(tables, opt, with_transit(lambda e: W.setup_engine(e, tables, config=5, phantom_n=32), transit),
                    lambda e: e.run_source(src, iso, t0, t1, n_decays=n, singles_cap=n, coinc_cap=n, adder_cap=4 * n,
                                           hits_cap=20 * n, trace_cap=2 * n))
    if transit is not None:
        assert a.tallies.launch.as_dict()["transit"] == list(transit)
    assert a.tallies.n_singles > 100 and a.tallies.n_coinc > 10
    assert a.arrays[0][8 * 400:].sum() > 0      # panels beyond the first ring see singles
    assert_parity(a, b, hits=True)


DIGI_CASES = {
    # paralyzable deadtime per crystal, winner-takes-all at crystal depth,
    # take-all-goods multiples, spatial blur
    "crystal_paralyzable": dict(readout_depth=3, readout_policy=0, dead_level=3, dead_type=0, dead_time_us=2.0,
                                space_blur=0.05, coinc_policy=1, coinc_window_us=0.05, win_min=100e3, win_max=700e3),
    # non-paralyzable per panel (site rewritten by setSitenum), blur policy 0
    "panel_nonparalyzable": dict(dead_level=1, dead_type=1, dead_time_us=1.0, blur_policy=0, r_ref=0.15),
    # site rewrite to module level without deadtime, panel-level readout
    "module_no_deadtime": dict(readout_depth=1, readout_policy=0, dead_level=2, dead_time_us=0.0),
}


@pytest.mark.parametrize("case", sorted(DIGI_CASES))
def test_deadtime_and_policies(tables, case):
    """Digitizer variants at a rate high enough for deadtime and multiples to
    matter."""
    from gpet_amd import engine
    from gpet_amd import workloads as W

    dig = W.digitizer(**DIGI_CASES[case])
    opt = engine.make_options(seed=5, positron_range=0)
    src, iso, _, _ = W.workload_sources(1)
    n = 40000
    a, b = run_both(tables, opt, lambda e: W.setup_engine(e, tables, config=1, dig=dig),
                    lambda e: e.run_source(src, iso, 0.0, 1e-3, n_decays=n, singles_cap=n, coinc_cap=4 * n,
                                           adder_cap=4 * n, trace_cap=2 * n))
    if DIGI_CASES[case].get("dead_time_us", 0) > 0:
        assert a.tallies.n_deadtime < a.tallies.n_threshold
    assert_parity(a, b)


# phase-space runs finish their photons in the phantom kernel: a launch
# with regrouping on must give the same results
@pytest.mark.parametrize("ptype,transit", [(0, None), (1, None), (0, (1, 2)), (1, (1,))])
def test_psf_parity(tables, ptype, transit):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    rng = np.random.default_rng(3)
    n = 20000
    rec = np.zeros((n, 8))
    rec[:, 0:3] = rng.uniform(-0.3, 0.3, (n, 3))
    rec[:, 3] = np.sort(rng.uniform(0, 1e5, n))
    d = rng.normal(size=(n, 3))
    rec[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 7] = 511e3 if ptype == 1 else rng.uniform(1e4, 6e5, n)
    opt = engine.make_options(seed=9, positron_range=1)
    a, b = run_both(tables, opt, with_transit(lambda e: W.setup_engine(e, tables, config=2, phantom_n=80), transit),
                    lambda e: e.run_psf(rec, ptype, singles_cap=n, coinc_cap=n, adder_cap=4 * n, trace_cap=2 * n))
    assert a.tallies.n_adder > 10
    assert_parity(a, b)


@pytest.mark.parametrize("sphere", [(0.0, 0.0, 0.0, 0.0), (0.05, -0.02, 0.0, 40.0)])
def test_phase_space_parity(tables, sphere):
    """Source and phantom-exit phase-space records (OUTPUTPSF 1/2) and the
    recording sphere (RECORDPSF -1)."""
    from gpet_amd import capi, engine
    from gpet_amd import workloads as W

    opt = engine.make_options(seed=21, positron_range=1, flags=capi.FLAG_PSF_SOURCE | capi.FLAG_PSF_PHANTOM,
                              record_sphere=sphere)
    src, iso, t0, t1 = W.workload_sources(2)
    n = 20000
    a, b = run_both(tables, opt, lambda e: W.setup_engine(e, tables, config=2, phantom_n=100),
                    lambda e: e.run_source(src, iso, t0, t1, n_decays=n, singles_cap=n, coinc_cap=n,
                                           adder_cap=4 * n, trace_cap=2 * n, psf_cap=2 * n))
    assert len(a.psf_source) > 30000 and len(a.psf_phantom) > 20000
    assert same_bytes(a.psf_source, b.psf_source)
    assert same_bytes(a.psf_phantom, b.psf_phantom)
    assert_parity(a, b)


@pytest.mark.parametrize("policy,dead", [(0, 0.0), (1, 0.005)])
def test_noise_and_delayed_parity(tables, policy, dead):
    """Background singles (addnoise) and the delayed-window pairs, on top of
    decays: the same singles, prompt pairs, delayed pairs and delayed
    sinogram as the oracle."""
    from gpet_amd import engine
    from gpet_amd import workloads as W

    dig = W.digitizer(coinc_policy=policy, dead_time_us=dead, dead_level=3, noise_rate=8.0e7, noise_e_mean=430e3,
                      noise_e_sigma=60e3, delay_us=0.25)
    opt = engine.make_options(seed=21, positron_range=1, batch=9000)
    src, iso, _, _ = W.workload_sources(2)
    n, t1 = 40000, 4e-4
    cap = 4 * n + int(8.0e7 * t1 * 2)

    def tallies_delayed(eng):
        t, cs, eh, sg, sgd = eng.tallies(n_crystals=48 * 400, n_ebins=100, e_lo=0.0, e_hi=1e6, n_s=64, n_phi=48,
                                         s_max=8.0, with_delayed=True)
        return cs, eh, sg, sgd

    res = {}
    for name, mk in (("gpu", hip_engine), ("oracle", oracle_engine)):
        eng = mk(opt)
        try:
            W.setup_engine(eng, tables, config=2, phantom_n=100, dig=dig)
            res[name] = eng.run_source(src, iso, 0.0, t1, n_decays=n, singles_cap=cap, coinc_cap=cap, adder_cap=cap,
                                       trace_cap=2 * n, delayed_cap=cap)
            res[name].arrays = tallies_delayed(eng)
        finally:
            eng.close()
    a, b = res["gpu"], res["oracle"]
    assert a.tallies.n_noise == b.tallies.n_noise > 20000
    assert a.tallies.n_delayed == b.tallies.n_delayed > 100
    assert same_bytes(a.delayed, b.delayed), "delayed: " + first_diff(a.delayed.ravel(), b.delayed.ravel())
    assert a.arrays[3].sum() > 0
    assert_parity(a, b)


@pytest.mark.parametrize("dims", [(7, 5, 9), (1, 1, 1), (13, 4, 6)])
def test_positron_walk_small_grids(tables, dims):
    """Positron walks on grids whose sizes are not multiples of the 4-voxel
    brick, with empty (zero-density) voxels, from a box source filling the
    phantom: walks that cross empty voxels, end inside or leave through
    every face (on the 1-voxel grid every walk leaves at its first crossing
    and goes on in air) agree with the oracle history by history."""
    from gpet_amd import engine
    from gpet_amd import workloads as W

    rng = np.random.default_rng(sum(dims))
    nx, ny, nz = dims
    mat = rng.choice([W.AIR, W.WATER, W.BONE], size=(nz, ny, nx)).astype(np.int32)
    den = rng.choice(np.array([0.0, 0.3, 1.0, 1.85], np.float32), size=(nz, ny, nx))
    vox = 0.01
    size = (nx * vox, ny * vox, nz * vox)
    off = tuple(-s / 2 for s in size)
    src = [W.source(0, c3=size[0] * 0.98, c4=size[1] * 0.98, c5=size[2] * 0.98, natom=1.0e11)]
    iso = [W.isotope_f18()]
    n = 20000
    opt = engine.make_options(seed=31, positron_range=1, batch=6000)

    def setup(e):
        W.setup_engine(e, tables, config=2, phantom_n=8)
        e.set_phantom(mat, den, dims, off, size)

    a, b = run_both(tables, opt, setup,
                    lambda e: e.run_source(src, iso, 0.0, 10.0, n_decays=n, singles_cap=n, coinc_cap=n,
                                           adder_cap=4 * n, trace_cap=n))
    assert a.tallies.n_singles > 100
    if min(dims) > 1:
        assert a.tallies.range_steps > n
    assert_parity(a, b)


def test_negative_density_rejected(tables):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    e = hip_engine(engine.make_options(seed=1))
    try:
        e.set_tables(tables)
        mat, den, dims, off, size = W.cylinder_phantom(8)
        den = den.copy()
        den[3, 4, 5] = -1.0
        with pytest.raises(engine.GpetError, match="negative phantom density"):
            e.set_phantom(mat, den, dims, off, size)
    finally:
        e.close()
