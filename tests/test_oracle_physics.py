"""The oracle's physics against closed forms (chi-square / KS at p > 0.01,
or exact identities).  Each check is chosen so that a plausible mistake (a
dropped term, a wrong sign or index) fails it."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from conftest import ORACLE_LIB, oracle_engine

MC2 = 510999.1


@pytest.fixture(scope="module")
def lib():
    l = C.CDLL(ORACLE_LIB)
    l.gpet_cpu_sample_kn.argtypes = [C.c_float, C.c_uint64, C.c_int64, C.c_void_p, C.c_void_p]
    l.gpet_cpu_rotate.argtypes = [C.c_void_p, C.c_float, C.c_float]
    l.gpet_cpu_sample_ek.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_void_p]
    l.gpet_cpu_sample_shape.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_void_p]
    l.gpet_cpu_emit_pairs.argtypes = [C.c_float, C.c_uint64, C.c_int64, C.c_void_p, C.c_void_p]
    l.gpet_cpu_positron_range.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_uint64, C.c_int64, C.c_void_p]
    l.gpet_cpu_crystal_search.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_float,
                                          C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    return l


def chi2_p(counts, expect):
    e = expect * counts.sum() / expect.sum()
    m = e > 5
    return stats.chisquare(counts[m], e[m] * counts[m].sum() / e[m].sum()).pvalue


@pytest.mark.parametrize("E", [511e3, 200e3, 30e3])
def test_klein_nishina_sampler(lib, E):
    """Analytic KN sampling (gPET_kernals.cu:90-126) vs dsigma/dOmega."""
    n = 400000
    ef = np.zeros(n, np.float32)
    ct = np.zeros(n, np.float32)
    lib.gpet_cpu_sample_kn(E, 5, n, ef.ctypes.data_as(C.c_void_p), ct.ctypes.data_as(C.c_void_p))
    k = E / MC2
    assert np.all(ef >= 1 / (1 + 2 * k) - 1e-6) and np.all(ef <= 1 + 1e-6)
    # energy fraction and angle satisfy the Compton relation
    assert np.allclose(ct, 1 - (1 - ef) / (ef * k), atol=2e-5 / min(k, 1))
    edges = np.linspace(-1, 1, 81)
    cnt, _ = np.histogram(ct, edges)
    c = np.linspace(-1, 1, 80001)
    eps = 1 / (1 + k * (1 - c))
    f = eps ** 2 * (eps + 1 / eps - (1 - c ** 2))
    cdf = np.concatenate([[0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(c))])
    expect = np.diff(np.interp(edges, c, cdf))
    assert chi2_p(cnt.astype(float), expect) > 0.01


def test_rotate_identities(lib):
    rng = np.random.default_rng(4)
    for _ in range(2000):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        cth = rng.uniform(-1, 1)
        v = np.array(d, np.float32)
        lib.gpet_cpu_rotate(v.ctypes.data_as(C.c_void_p), cth, rng.uniform(0, 1))
        assert abs(np.linalg.norm(v) - 1) < 2e-6
        assert abs(float(np.dot(v, d)) - cth) < 2e-6
    # special case w = +-1: the new direction has w' = +-cos(theta)
    for w in (1.0, -1.0):
        v = np.array([0, 0, w], np.float32)
        lib.gpet_cpu_rotate(v.ctypes.data_as(C.c_void_p), 0.3, 0.25)
        assert abs(v[2] - 0.3 * w) < 1e-7 and abs(np.linalg.norm(v) - 1) < 1e-6
    # the azimuth is uniform: u = 0.25 turns by 90 degrees about the axis
    a = np.array([0, 0, 1], np.float32)
    b = np.array([0, 0, 1], np.float32)
    lib.gpet_cpu_rotate(a.ctypes.data_as(C.c_void_p), 0.0, 0.0)
    lib.gpet_cpu_rotate(b.ctypes.data_as(C.c_void_p), 0.0, 0.25)
    assert abs(float(np.dot(a, b))) < 1e-6


def test_positron_energy_spectrum(lib):
    """sampleEkPositron (gPET_kernals.cu:420-443): rejection from the F-18
    polynomial; accepted density is clamp(p(E), 0, envelope)."""
    from gpet_amd import workloads as W
    iso = W.isotope_f18()
    n = 300000
    ek = np.zeros(n, np.float32)
    lib.gpet_cpu_sample_ek(C.byref(iso), 8, n, ek.ctypes.data_as(C.c_void_p))
    E = ek / 1e6 + 0.511
    c = np.array(iso.coef[:], float)
    assert E.min() >= 0.511 - 1e-6 and E.max() <= c[0] + 1e-6
    edges = np.linspace(0.511, c[0], 61)
    x = np.linspace(0.511, c[0], 60001)
    p = np.clip(np.polyval(c[2:], x), 0, c[1])
    cdf = np.concatenate([[0], np.cumsum(0.5 * (p[1:] + p[:-1]) * np.diff(x))])
    expect = np.diff(np.interp(edges, x, cdf))
    cnt, _ = np.histogram(E, edges)
    assert chi2_p(cnt.astype(float), expect) > 0.01
    # the fit of data/isotopes.txt has a mean kinetic energy of ~0.275 MeV
    # (measured F-18: ~0.25 MeV)
    assert 0.24 < (E - 0.511).mean() < 0.30


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_source_shapes_uniform(lib, shape):
    from gpet_amd import workloads as W
    s = W.source(shape, centre=(0.3, -0.2, 0.1), c3=0.8 if shape else 0.6, c4=0.5 if shape != 2 else 0.0, c5=0.4)
    n = 200000
    p = np.zeros((n, 3), np.float32)
    lib.gpet_cpu_sample_shape(C.byref(s), 3, n, p.ctypes.data_as(C.c_void_p))
    q = p.astype(float) - [0.3, -0.2, 0.1]
    if shape == 0:
        for k, L in enumerate((0.6, 0.5, 0.4)):
            assert stats.kstest(q[:, k] / L + 0.5, "uniform").pvalue > 0.01
    elif shape == 1:
        r2 = (q[:, 0] ** 2 + q[:, 1] ** 2) / 0.64
        assert stats.kstest(r2, "uniform").pvalue > 0.01
        assert stats.kstest(q[:, 2] / 0.5 + 0.5, "uniform").pvalue > 0.01
        phi = (np.arctan2(q[:, 1], q[:, 0]) / (2 * np.pi)) % 1.0
        assert stats.kstest(phi, "uniform").pvalue > 0.01
    elif shape == 3:
        # uniform in the ellipsoid: the scaled point is uniform in the unit
        # ball, and each coordinate's second moment is a^2/5
        u = q / [0.8, 0.5, 0.4]
        rn = np.linalg.norm(u, axis=1)
        assert rn.max() <= 1.0 + 1e-6
        assert stats.kstest(rn ** 3, "uniform").pvalue > 0.01
        assert stats.kstest((u[:, 2] / rn + 1) / 2, "uniform").pvalue > 0.01
        for k, a in enumerate((0.8, 0.5, 0.4)):
            assert abs((q[:, k] ** 2).mean() / (a * a / 5) - 1) < 0.01
    else:
        r3 = (np.linalg.norm(q, axis=1) / 0.8) ** 3
        assert stats.kstest(np.minimum(r3, 1), "uniform").pvalue > 0.01
        assert stats.kstest((q[:, 2] / np.linalg.norm(q, axis=1) + 1) / 2, "uniform").pvalue > 0.01


def test_pair_noncollinearity(lib):
    """Photon 2 makes angle pi - |delta| with photon 1, delta ~ N(0, sigma),
    and E1 + E2 = 2 mc^2 (gPET_kernals.cu:536-555)."""
    sigma = 0.0037056
    n = 200000
    v = np.zeros((n, 6), np.float32)
    e = np.zeros((n, 2), np.float32)
    lib.gpet_cpu_emit_pairs(sigma, 2, n, v.ctypes.data_as(C.c_void_p), e.ctypes.data_as(C.c_void_p))
    d1, d2 = v[:, :3].astype(float), v[:, 3:].astype(float)
    assert np.abs(np.linalg.norm(d1, axis=1) - 1).max() < 2e-6
    ang = np.pi - np.arccos(np.clip(np.sum(d1 * d2, axis=1), -1, 1))
    # arccos near -1 is ill-conditioned in float: compare with the sine form
    ang = np.arcsin(np.clip(np.linalg.norm(np.cross(d1, d2), axis=1), 0, 1))
    assert stats.kstest(ang / sigma, stats.halfnorm.cdf).pvalue > 0.01
    assert np.abs(e.sum(axis=1) - 2 * np.float32(MC2)).max() < 0.2
    # isotropy of the first photon
    assert stats.kstest((d1[:, 2] + 1) / 2, "uniform").pvalue > 0.01


def _homogeneous(tables, dens):
    from gpet_amd import engine
    eng = oracle_engine(engine.make_options())
    eng.set_tables(tables)
    n = 40
    mat = np.ones((n, n, n), np.int32)
    eng.set_phantom(mat, np.full((n, n, n), dens, np.float32), (n, n, n), (-1, -1, -1), (2, 2, 2))
    return eng


@pytest.mark.parametrize("dens", [1.0, 2.0])
def test_positron_range_in_homogeneous_medium(lib, tables, dens):
    """In a homogeneous phantom of density rho the displacement is r/rho with
    r = |sigma * N3|: Maxwell with scale sigma/rho, sigma = Rex/2 and
    Rex = 0.1*5.44040782*E^2/(0.369516529+E) cm (gPET_kernals.cu:356-363)."""
    eng = _homogeneous(tables, dens)
    ek = 250e3
    n = 50000
    out = np.zeros((n, 3), np.float32)
    x0 = np.array([0.013, -0.021, 0.007], np.float32)
    assert lib.gpet_cpu_positron_range(eng.ctx, x0.ctypes.data_as(C.c_void_p), ek, 4, n,
                                       out.ctypes.data_as(C.c_void_p)) == 0
    E = ek / 1e6
    sigma = 0.1 * 5.44040782 * E * E / (0.369516529 + E) / 2
    r = np.linalg.norm(out.astype(float) - x0, axis=1)
    assert stats.kstest(r, stats.maxwell(scale=sigma / dens).cdf).pvalue > 0.01
    eng.close()


def test_crystal_search_brute_force(lib, tables):
    """crystalSearch (gPET_kernals.cu:1236-1279) vs containment in every
    module and crystal box, on a panel with 2 x 3 modules of 4 x 5 crystals."""
    from gpet_amd import engine, workloads as W
    txt = W.geometry_text(n_panels=1)
    panels, mats = W.octagon_scanner(n_panels=1)
    p = panels[0]
    p.leny, p.lenz = 2 * 0.62 + 0.03, 3 * 0.95 + 2 * 0.04
    p.mody, p.modz, p.mspy, p.mspz = 0.62, 0.95, 0.03, 0.04
    p.cry, p.crz, p.cspy, p.cspz = 0.14, 0.17, 0.02, 0.025
    eng = oracle_engine(engine.make_options())
    eng.set_tables(tables)
    eng.set_detector([p], mats, None)
    mn, cn, mny, cny = engine.panel_counts(p)
    assert (mny, cny) == (2, 4) and mn == 6 and cn == 20
    rng = np.random.default_rng(6)
    M = C.c_int32()
    L = C.c_int32()
    checked = 0
    for _ in range(20000):
        y = np.float32(rng.uniform(-p.leny / 2, p.leny / 2))
        z = np.float32(rng.uniform(-p.lenz / 2, p.lenz / 2))
        m = lib.gpet_cpu_crystal_search(eng.ctx, 0, np.float32(-0.5), y, z, C.byref(M), C.byref(L))
        yy, zz = float(y) + p.leny / 2, float(z) + p.lenz / 2
        want = (1, None, None)
        near = False
        for my in range(2):
            for mz in range(3):
                y0, z0 = my * (p.mody + p.mspy), mz * (p.modz + p.mspz)
                for ly in range(4):
                    for lz in range(5):
                        a, b = y0 + ly * (p.cry + p.cspy), z0 + lz * (p.crz + p.cspz)
                        if min(abs(yy - a), abs(yy - a - p.cry), abs(zz - b), abs(zz - b - p.crz)) < 1e-5:
                            near = True
                        if a <= yy <= a + p.cry and b <= zz <= b + p.crz:
                            want = (0, mz * 2 + my, lz * 4 + ly)
        if near:
            continue
        checked += 1
        assert m == want[0], (y, z)
        if m == 0:
            assert (M.value, L.value) == want[1:], (y, z)
    assert checked > 19000
    eng.close()


def _pencil(tables, config, rec, dig=None, prange=0):
    from gpet_amd import engine, workloads as W
    eng = oracle_engine(engine.make_options(seed=21, positron_range=prange, record_hits=1))
    W.setup_engine(eng, tables, config=config, dig=dig)
    r = eng.run_psf(rec, 1, trace_cap=2 * len(rec), hits_cap=20 * len(rec), adder_cap=4 * len(rec))
    eng.close()
    r.trace = r.trace[0::2]   # slot 2h holds the photon of record h
    return r


def test_narrow_beam_attenuation_in_water(tables):
    """Uncollided fraction of a 511 keV pencil beam through 1.6 cm of water is
    exp(-mu L) with mu the lerped total cross section (Woodcock is exact)."""
    n = 100000
    rec = np.zeros((n, 8))
    rec[:, 3] = 1.0
    rec[:, 4] = 1.0
    rec[:, 7] = 511e3
    r = _pencil(tables, 1, rec)
    tr = r.trace
    unc = (tr["n_compton"] == 0) & (tr["n_rayleigh"] == 0) & ((tr["status"] & 2) != 0)
    e0, idl = float(tables.e0[0]), float(tables.idl[0])
    x = (511e3 - e0) * idl
    j = int(x)
    mu = tables.mu[0][1, j] + (x - j) * (tables.mu[0][1, j + 1] - tables.mu[0][1, j])
    p = np.exp(-mu * 1.6)
    assert abs(unc.mean() - p) < 4 * np.sqrt(p * (1 - p) / n)
    # the photoelectric share follows mu_pe = mu_tot - mu_compt - mu_rayl
    assert 0.80 < p < 0.90


def test_narrow_beam_through_one_crystal_column(tables):
    """A pencil beam along the depth of one LSO crystal: the fraction with no
    real interaction is exp(-mu rho 1 cm) (rho = 7.4); the first hit is in
    the aimed crystal."""
    n = 100000
    rec = np.zeros((n, 8))
    # panel 0: face at global y = -8, local y = global x, local z = -global z;
    # crystal (ly, lz) = (10, 10) centre at local (0.075, 0.075)
    rec[:, 0] = 0.075
    rec[:, 1] = -7.9
    rec[:, 2] = -0.075
    rec[:, 3] = 1.0
    rec[:, 5] = -1.0
    rec[:, 7] = 511e3
    r = _pencil(tables, 2, rec)
    tr = r.trace
    assert np.all(tr["panel"] == 0)
    none = (tr["n_int_panel"] == 0).mean()
    e0, idl = float(tables.e0[0]), float(tables.idl[0])
    x = (511e3 - e0) * idl
    j = int(x)
    mu = tables.mu[0][7, j] + (x - j) * (tables.mu[0][7, j + 1] - tables.mu[0][7, j])
    p = np.exp(-mu * 7.4 * 1.0)
    assert abs(none - p) < 4 * np.sqrt(p * (1 - p) / n)
    first = {}
    for hid in r.hit_id:
        first.setdefault(hid[0], hid)
    # (a first Rayleigh scatter is not recorded and can move the photon on)
    assert np.mean([h[3] == 210 and h[2] == 0 for h in first.values()]) > 0.99
    # energy bookkeeping: deposits of a photon never exceed its energy
    dep = np.zeros(n)
    for hid, hf in zip(r.hit_id, r.hit_f):
        if hid[4] in (1, 2, 4):
            dep[hid[0]] += hf[0]
    assert dep.max() <= 511e3 * (1 + 1e-6)
    # a photoelectric absorption deposits everything: full-energy events exist
    assert (np.abs(dep - 511e3) < 1.0).mean() > 0.2


def test_panel_entry_against_ray_plane(tables):
    """Panel entry (gPET_kernals.cu:958-1009) vs a numpy ray-plane
    intersection for photons leaving the origin isotropically."""
    from gpet_amd import workloads as W
    n = 20000
    rng = np.random.default_rng(8)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rec = np.zeros((n, 8))
    rec[:, 0:3] = [[0.0, 0.0, 0.0]]
    rec[:, 3] = 1.0
    rec[:, 4:7] = d
    rec[:, 7] = 511e3
    panels, _ = W.octagon_scanner()
    r = _pencil(tables, 1, rec)
    tr = r.trace
    # only photons that left the water cube without interacting keep the
    # original direction
    ok = (tr["n_compton"] == 0) & (tr["n_rayleigh"] == 0) & ((tr["status"] & 2) != 0)
    exp_panel = np.full(n, -1)
    for i in np.nonzero(ok)[0]:
        x = np.array([tr["x"][i], tr["y"][i], tr["z"][i]], float)
        v = d[i]
        for k, p in enumerate(panels):
            ux, uy, uz = np.array(p.ux[:]), np.array(p.uy[:]), np.array(p.uz[:])
            rel = x - [p.ox, p.oy, p.oz]
            tx, tvx = rel @ ux, v @ ux
            if tvx * p.dirx < 0:
                continue
            s = -tx / tvx
            ty, tz = rel @ uy + s * (v @ uy), rel @ uz + s * (v @ uz)
            if abs(ty) < p.leny / 2 - 1e-4 and abs(tz) < p.lenz / 2 - 1e-4:
                exp_panel[i] = k
                break
            if abs(abs(ty) - p.leny / 2) < 1e-4 or abs(abs(tz) - p.lenz / 2) < 1e-4:
                exp_panel[i] = -2
                break
    sel = ok & (exp_panel != -2)
    assert np.array_equal(tr["panel"][sel], exp_panel[sel])
    # geometric acceptance of eight 3 x 3 cm panels at 8 cm
    assert 0.05 < (tr["panel"][ok] >= 0).mean() < 0.12
