"""Independent pins for four oracle functions that the GPU path matches bit
for bit, so that a misreading of the reference shared by both would still
fail a test:

  * the inverse-CDF surface read (gPET_kernals.cu:66-88, 128-146),
  * decay counts and emission times (gPET_kernals.cu:497-559, gPET.cu:260-283),
  * the positron walk through heterogeneous media (gPET_kernals.cu:380-415),
  * adder and readout (gPET_kernals.cu:737-813).

Each check is computed without the oracle's code (scipy distributions,
closed forms, brute force from recorded hits).  Each test also applies a
named plausible mistake to the same data in numpy and asserts that the
check rejects it, which shows the check has the power to catch it."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy import stats

from conftest import ROOT, oracle_engine
from gpet_amd import capi, engine, tablegen
from gpet_amd import workloads as W


@pytest.fixture(scope="module")
def eng_tab(tables):
    eng = oracle_engine(engine.make_options())
    eng.set_tables(tables)
    lib = eng.lib
    lib.gpet_cpu_surface.restype = C.c_float
    lib.gpet_cpu_surface.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_float]
    lib.gpet_cpu_surface_n.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p, C.c_int64,
                                       C.c_void_p]
    lib.gpet_cpu_binomial.restype = C.c_int64
    lib.gpet_cpu_binomial.argtypes = [C.c_double, C.c_double, C.c_uint64, C.c_int32]
    yield eng
    eng.close()


# --------------------------------------------------------------- surfaces --

BLOCKS = tablegen.load_blocks(os.path.join(ROOT, "gpet_amd", "data", "gctd_qblocks.json"))
SURF_KIND = {0: "compton", 1: "rayleigh"}
SURF_MAT = {1: "Water", 7: "LSO"}


def _uniforms(n, seed):
    """Uniforms in (0, 1] with 24 bits, as the transport draws them."""
    r = np.random.default_rng(seed)
    return ((r.integers(0, 1 << 24, n) + 1) * 2.0 ** -24).astype(np.float32)


def _surface_cdf(kind, mat, E, n_cos=201):
    """CDF of cos(theta) that the table generator integrates: the density
    (Klein-Nishina x S(q), or (1 + cos^2)/2 x F(q)^2) on its cos grid, by the
    trapezoid rule (the grid of the gCTD surfaces the generator reproduces)."""
    b = BLOCKS[kind][SURF_MAT[mat]]
    c = np.linspace(-1.0, 1.0, n_cos)
    d = tablegen.density(kind, E, c, np.asarray(b["q"]), np.asarray(b["v"]))
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (d[1:] + d[:-1]) * np.diff(c))])
    return c, cdf / cdf[-1]


def _chi2_vs_cdf(samples, c, cdf, nbins=50):
    edges = np.interp(np.linspace(0, 1, nbins + 1), cdf, c)
    edges[0], edges[-1] = -1.0, 1.0
    cnt, _ = np.histogram(samples, edges)
    return stats.chisquare(cnt).pvalue


def _numpy_read(S, ide, idcp, E, u, mutation=None):
    """Bilinear read of a [ncp, ne] surface with clamping, in float32, with an
    optional planted mistake:  'swap'  E and CP axes exchanged in the flat
    index;  'cp_scale'  u scaled by ncp instead of ncp - 1;  'e_next'  the
    energy column one node up."""
    f32 = np.float32
    ncp, ne = S.shape
    x = np.clip(f32(E) * f32(ide), f32(0), f32(ne - 1))
    y = u * (f32(ncp) if mutation == "cp_scale" else f32(idcp))
    y = np.clip(y, f32(0), f32(ncp - 1))
    ix = min(int(x), ne - 2)
    if mutation == "e_next":
        ix = min(ix + 1, ne - 2)
    iy = np.minimum(y.astype(np.int64), ncp - 2)
    fx = f32(x - f32(ix))
    fy = (y - iy.astype(np.float32)).astype(np.float32)
    flat = S.reshape(-1)
    if mutation == "swap":
        idx = lambda r, cc: flat[(cc * ncp + r) % flat.size]
    else:
        idx = lambda r, cc: flat[r * ne + cc]
    v00, v01 = idx(iy, ix), idx(iy, ix + 1)
    v10, v11 = idx(iy + 1, ix), idx(iy + 1, ix + 1)
    a = v00 + fx * (v01 - v00)
    b = v10 + fx * (v11 - v10)
    return np.clip(a + fy * (b - a), f32(-1), f32(1)).astype(np.float32)


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("mat", [1, 7])
def test_surface_read_known_answers(eng_tab, tables, which, mat):
    """At a table node the read returns the stored value; at the centre of a
    cell, the mean of the four corners (bilinear)."""
    f = eng_tab.lib.gpet_cpu_surface
    S = tables.surf[which][mat]
    ide, idcp = float(tables.ide[which]), float(tables.idcp[which])
    for ie in (0, 37, 127, 128, 149):
        for icp in (0, 1, 150, 298, 299):
            E = np.float32(ie / ide)
            u = np.float32(icp / idcp)
            got = f(eng_tab.ctx, which, mat, E, u)
            assert abs(got - S[icp, ie]) <= 2e-6, (ie, icp, got, S[icp, ie])
            Em = np.float32((ie + 0.5) / ide)
            um = np.float32((icp + 0.5) / idcp)
            want = 0.25 * (float(S[icp, ie]) + float(S[icp, ie + 1]) + float(S[icp + 1, ie]) + float(S[icp + 1, ie + 1]))
            assert abs(f(eng_tab.ctx, which, mat, Em, um) - want) <= 3e-6, (ie, icp)
    # the top node row: u = 1 reads the last CP row
    assert abs(f(eng_tab.ctx, which, mat, np.float32(128 / ide), np.float32(1.0)) - S[300, 128]) <= 2e-6
    # a read one energy node off misses the stored values (the distribution
    # check below cannot see a 4 keV shift)
    # (the check above fails as soon as one node misses by more than 2e-6)
    miss = 0
    for ie in (0, 37, 127, 128, 149):
        for icp in (0, 1, 150, 298, 299):
            u = np.array([icp / idcp], np.float32)
            miss += abs(_numpy_read(S, ide, idcp, np.float32(ie / ide), u, "e_next")[0] - S[icp, ie]) > 2e-6
    assert miss >= 3


@pytest.mark.parametrize("which,mat,E", [(0, 1, 511e3), (0, 7, 511e3), (0, 1, 202e3), (0, 7, 350e3),
                                         (1, 1, 200e3), (1, 7, 511e3)])
def test_surface_read_samples_its_distribution(eng_tab, tables, which, mat, E):
    """cos(theta) = surface(E, u) for uniform u follows the angular
    distribution the surface was built from (chi-square over 50 equiprobable
    bins, p > 0.01), including energies between nodes (202, 350 keV).  The
    same check rejects an E/CP axis swap and u scaled by ncp instead of
    ncp - 1."""
    n = 400000
    u = _uniforms(n, 11 + which * 7 + mat)
    out = np.zeros(n, np.float32)
    eng_tab.lib.gpet_cpu_surface_n(eng_tab.ctx, which, mat, E, u.ctypes.data, n, out.ctypes.data)
    c, cdf = _surface_cdf(SURF_KIND[which], mat, E)
    assert _chi2_vs_cdf(out, c, cdf) > 0.01
    S = tables.surf[which][mat]
    ide, idcp = float(tables.ide[which]), float(tables.idcp[which])
    # the numpy read without a mistake is the oracle's read ...
    assert np.array_equal(_numpy_read(S, ide, idcp, E, u), out)
    # ... and each planted mistake fails the distribution check
    for mut in ("swap", "cp_scale"):
        bad = _numpy_read(S, ide, idcp, E, u, mut)
        assert _chi2_vs_cdf(bad, c, cdf) < 1e-4, mut


# ---------------------------------------------------------- decay sampling --

def _binom_draws(lib, n, p, k, seed=3):
    return np.array([lib.gpet_cpu_binomial(n, p, seed, i) for i in range(k)], np.int64)


def _chi2_binom(x, n, p):
    lo, hi = int(x.min()), int(x.max())
    ks = np.arange(lo, hi + 1)
    cnt = np.bincount(x - lo, minlength=len(ks)).astype(float)
    exp = stats.binom.pmf(ks, n, p) * len(x)
    # merge sparse tails into neighbouring bins (expected >= 5)
    c, e, acc_c, acc_e = [], [], 0.0, 0.0
    for ci, ei in zip(cnt, exp):
        acc_c += ci
        acc_e += ei
        if acc_e >= 5:
            c.append(acc_c)
            e.append(acc_e)
            acc_c = acc_e = 0.0
    c[-1] += acc_c
    e[-1] += acc_e + (1.0 - stats.binom.cdf(hi, n, p) + stats.binom.cdf(lo - 1, n, p)) * len(x)
    c, e = np.array(c), np.array(e)
    return stats.chisquare(c, e * c.sum() / e.sum()).pvalue


@pytest.mark.parametrize("n,p,k", [(40, 0.5, 40000), (1000, 0.05, 40000), (2000, 0.3, 20000),
                                   (10 ** 6, 0.3, 4000)])
def test_binomial_counts(eng_tab, n, p, k):
    """The per-source decay count is Binomial(n, p) (the reference decides
    atom by atom, gPET_kernals.cu:497-530): chi-square against the binomial
    pmf.  A Poisson(np) count (ignoring 1 - p) fails the same check at p = 0.5
    and 0.3."""
    x = _binom_draws(eng_tab.lib, n, p, k)
    assert x.min() >= 0 and x.max() <= n
    assert _chi2_binom(x, n, p) > 0.01
    if p >= 0.3:
        pois = np.random.default_rng(1).poisson(n * p, k)
        assert _chi2_binom(np.clip(pois, 0, n), n, p) < 1e-4


def test_binomial_counts_large_mean(eng_tab):
    """A C2-sized source: n = 9.5e10 surviving atoms, p = 1.05e-3 (mean 1e8,
    sd 1e4).  The standardized counts are N(0, 1) (KS) and their variance is
    n p (1 - p) to within its sampling error."""
    n, p, k = 9.5e10, 1.05e-3, 600
    x = _binom_draws(eng_tab.lib, n, p, k).astype(float)
    mu, sd = n * p, np.sqrt(n * p * (1 - p))
    z = (x - mu) / sd
    assert stats.kstest(z, "norm").pvalue > 0.01
    assert abs(z.var(ddof=1) - 1.0) < 5 * np.sqrt(2.0 / k)
    assert abs(z.mean()) < 4 / np.sqrt(k)


def _point_sources(k, natom, spacing=0.5):
    return [W.source(2, centre=(spacing * (i % 10) - 2.25, spacing * (i // 10) - 2.25, 0.0), natom=natom)
            for i in range(k)]


def _decay_engine(tables, seed=5):
    eng = oracle_engine(engine.make_options(seed=seed))
    eng.set_tables(tables)
    eng.set_digitizer(W.digitizer(coinc_window_us=0.0))
    return eng


def test_run_decay_counts_per_source(tables):
    """gpet_run_source with n_decays = 0: source i decays Binomial(navail,
    pdec) times, navail = floor(natom 2^(-t0/T)) atoms surviving to t0 and
    pdec = 1 - 2^(-(t1-t0)/T) (the reference's per-atom loop over
    [t0, t1), gPET_kernals.cu:497-530, gPET.cu:260-283).  Per-source counts
    come from the photons' start points (point sources, no positron range).
    A plan that forgets the decay before t0 (navail = natom) fails."""
    iso = W.isotope_pure_pairs()
    iso.half_life_s = 3.0
    k, natom, t0, t1 = 100, 200.0, 2.0, 5.0
    eng = _decay_engine(tables)
    r = eng.run_source(_point_sources(k, natom), [iso], t0, t1, 0, trace_cap=10 ** 6)
    eng.close()
    tr = r.trace[r.trace["photon"] == 0]
    assert len(tr) == r.tallies.n_histories
    src = np.round((tr["y0"] + 2.25) / 0.5).astype(int) * 10 + np.round((tr["x0"] + 2.25) / 0.5).astype(int)
    cnt = np.bincount(src, minlength=k)
    navail = np.floor(natom * 2.0 ** (-t0 / iso.half_life_s))
    pdec = 1 - 2.0 ** (-(t1 - t0) / iso.half_life_s)
    z = (cnt.sum() - k * navail * pdec) / np.sqrt(k * navail * pdec * (1 - pdec))
    assert abs(z) < 4
    assert stats.kstest((cnt - navail * pdec) / np.sqrt(navail * pdec * (1 - pdec)), "norm").pvalue > 1e-3
    z_bad = (cnt.sum() - k * natom * pdec) / np.sqrt(k * natom * pdec * (1 - pdec))
    assert abs(z_bad) > 10


def test_emission_times_truncated_exponential(tables):
    """Decay times within [t0, t1) follow the exponential law truncated to
    the window: F(t) = (1 - e^-(t-t0)/tau) / (1 - e^-(t1-t0)/tau), tau = T/ln 2
    (KS on the photons' start times; T = 3 s over a 10 s window, so the
    truncation matters).  Planted mistakes fail the same KS: tau = T (no
    ln 2) and times uniform in the window."""
    iso = W.isotope_pure_pairs()
    iso.half_life_s = 3.0
    t0, t1 = 2.0, 12.0
    eng = _decay_engine(tables)
    r = eng.run_source([W.source(2, natom=1e12)], [iso], t0, t1, 100000, trace_cap=10 ** 6)
    eng.close()
    t = r.trace[r.trace["photon"] == 0]["t0"] * 1e-6
    assert len(t) == 100000 and t.min() >= t0 and t.max() < t1
    tau = iso.half_life_s / np.log(2.0)

    def cdf(x, tau=tau):
        return -np.expm1(-(x - t0) / tau) / -np.expm1(-(t1 - t0) / tau)
    assert stats.kstest(t, cdf).pvalue > 0.01
    # the uniforms behind the times, then the times under each mistake
    u = cdf(t)
    pd_ = -np.expm1(-(t1 - t0) / iso.half_life_s)
    no_ln2 = t0 - iso.half_life_s * np.log1p(-u * pd_)
    assert stats.kstest(no_ln2, cdf).pvalue < 1e-4
    assert stats.kstest(t0 + u * (t1 - t0), cdf).pvalue < 1e-4


# --------------------------------------------------------- positron walk --

RHO_OUT = 0.0012905   # density outside the phantom, gPET_kernals.cu:407


def _walk_engine(tables, dens, off, size):
    eng = oracle_engine(engine.make_options())
    eng.set_tables(tables)
    nz, ny, nx = dens.shape
    eng.set_phantom(np.ones(dens.shape, np.int32), dens.astype(np.float32), (nx, ny, nz), off, size)
    return eng


def _oracle_walk(eng, p0, d, rr):
    eng.lib.gpet_cpu_range_walk.argtypes = [C.c_void_p] * 4 + [C.c_int64, C.c_void_p, C.c_void_p]
    n = len(rr)
    out = np.zeros((n, 3), np.float32)
    steps = np.zeros(n, np.int64)
    p0 = np.ascontiguousarray(p0, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    rr = np.ascontiguousarray(rr, np.float32)
    st = eng.lib.gpet_cpu_range_walk(eng.ctx, p0.ctypes.data, d.ctypes.data, rr.ctypes.data, n, out.ctypes.data,
                                     steps.ctypes.data)
    assert st == 0
    return out.astype(np.float64), steps


def _brute_walk(dens, off, vox, p0, d, rr, mutation=None):
    """End point of a walk by brute force in FP64: every plane crossing of the
    ray inside the grid, sorted; the segment ending at a crossing is charged
    at the density of the voxel entered there (the reference's rule, Q8),
    found from the midpoint of the following segment; outside the grid the
    density is RHO_OUT.  Planted mistakes: 'current' charges the voxel being
    left; 'stop_at_edge' ends the walk at the grid boundary; 'air_0012'
    uses 0.0012 outside."""
    nz, ny, nx = dens.shape
    n = np.array([nx, ny, nz])
    off = np.asarray(off, float)
    vox = np.asarray(vox, float)
    p0 = np.asarray(p0, float)
    d = np.asarray(d, float)
    rho_out = 0.0012 if mutation == "air_0012" else RHO_OUT

    def voxel_at(p):
        i = np.floor((p - off) / vox).astype(int)
        if np.all(i >= 0) and np.all(i < n):
            return dens[i[2], i[1], i[0]]
        return None

    ts = []
    for a in range(3):
        if d[a] == 0:
            continue
        planes = off[a] + vox[a] * np.arange(n[a] + 1)
        t = (planes - p0[a]) / d[a]
        ts.append(t[t > 0])
    ts = np.sort(np.concatenate(ts))
    # keep crossings up to the first exit from the grid
    s, tcur = 0.0, 0.0
    cur = voxel_at(p0)
    for k, tk in enumerate(ts):
        nxt_t = ts[k + 1] if k + 1 < len(ts) else tk + 1.0
        ent = voxel_at(p0 + d * 0.5 * (tk + nxt_t))
        rho = ent if ent is not None else rho_out
        if mutation == "current":
            rho = cur
        sn = s + (tk - tcur) * rho
        if sn >= rr:
            return p0 + d * (tk + (rr - sn) / rho)
        if ent is None:
            if mutation == "stop_at_edge":
                return p0 + d * tk
            return p0 + d * (tk + (rr - sn) / rho_out)
        s, tcur, cur = sn, tk, ent
    raise AssertionError("walk left no exit")


def test_walk_two_slabs_closed_form(tables):
    """Along +x through two slabs (rho1 = 1 for x < 0.5 cm, rho2 = 3 beyond,
    voxels of 0.05 cm, grid to x = 1 cm) from the middle of voxel 0: charging
    each segment at the entered voxel's density moves the slab boundary one
    voxel earlier, to b = 0.45 cm, so the end point is
      x0 + r/rho1                                   for r <= rho1 (b - x0),
      b + (r - rho1 (b - x0))/rho2                  up to the last voxel,
    and past the grid's last crossing (at 0.95 cm) the rest at RHO_OUT.  The
    same formula with the boundary at 0.5 cm (the current voxel's density)
    misses the oracle on the walks ending near the boundary."""
    nx = 20
    dens = np.ones((2, 2, nx), np.float32)
    dens[:, :, 10:] = 3.0
    off, vox = (0.0, 0.0, 0.0), (0.05, 0.05, 0.05)
    eng = _walk_engine(tables, dens, off, (1.0, 0.1, 0.1))
    rr = np.linspace(0.001, 2.0, 400).astype(np.float32)
    x0 = 0.025
    p0 = np.tile([x0, 0.03, 0.04], (len(rr), 1))
    d = np.tile([1.0, 0.0, 0.0], (len(rr), 1))
    out, steps = _oracle_walk(eng, p0, d, rr)
    eng.close()

    def closed(r, b):
        x_last = 0.95           # the crossing into the last voxel: beyond it, air
        r1 = 1.0 * (b - x0)
        r2 = r1 + 3.0 * (x_last - b)
        return np.where(r <= r1, x0 + r, np.where(r <= r2, b + (r - r1) / 3.0, x_last + (r - r2) / RHO_OUT))
    want = closed(rr.astype(float), 0.45)
    # FP32: in air the remaining range (rr - s, a difference of numbers near
    # rr) is divided by RHO_OUT, which scales its rounding by 775
    tol = 2e-6 + 2e-5 * want + np.where(want > 0.95, 4 * 1.2e-7 * rr / RHO_OUT, 0.0)
    assert np.all(np.abs(out[:, 0] - want) <= tol)
    assert np.allclose(out[:, 1:], p0[:, 1:], atol=1e-7)
    bad = closed(rr.astype(float), 0.50)
    assert np.sum(np.abs(bad - out[:, 0]) > 1e-3) > 10
    # voxels read inside the phantom: 19 for walks that reach the air, 1 for
    # one that ends in its first voxel (the first segment is charged at the
    # density of the voxel it enters)
    assert steps[-1] == 19 and steps[0] == 1


@pytest.mark.parametrize("seed", [1, 2])
def test_walk_heterogeneous_brute_force(tables, seed):
    """Random walks in a random 12 x 10 x 8 phantom of four densities
    (0.0012, 0.3, 1.0, 1.85; voxels 0.1 x 0.08 x 0.12 cm) against the
    brute-force crossing list, including walks that leave the phantom and
    continue in air.  Each planted mistake ('current' voxel density, stop at
    the grid edge, 0.0012 outside) disagrees with the oracle on many walks."""
    rng = np.random.default_rng(seed)
    dens = rng.choice(np.array([0.0012, 0.3, 1.0, 1.85], np.float32), size=(8, 10, 12))
    off = (-0.6, -0.4, -0.48)
    vox = (0.1, 0.08, 0.12)
    size = (1.2, 0.8, 0.96)
    eng = _walk_engine(tables, dens, off, size)
    n = 1500
    p0 = np.column_stack([rng.uniform(off[a] + 0.01, off[a] + size[a] - 0.01, n) for a in range(3)])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p0 = p0.astype(np.float32)
    d = d.astype(np.float32)
    rr = rng.uniform(0.0, 0.6, n).astype(np.float32)
    out, _ = _oracle_walk(eng, p0, d, rr)
    eng.close()
    want = np.array([_brute_walk(dens, off, vox, p0[i], d[i], float(rr[i])) for i in range(n)])
    dist = np.linalg.norm(out - want, axis=1)
    scale = 1.0 + np.linalg.norm(want - p0, axis=1)
    assert np.max(dist / scale) < 1e-4, np.argmax(dist / scale)
    left = np.any((want < np.array(off)) | (want > np.array(off) + np.array(size)), axis=1)
    assert left.sum() > 50    # some walks exit into air
    for mut in ("current", "stop_at_edge", "air_0012"):
        bad = np.array([_brute_walk(dens, off, vox, p0[i], d[i], float(rr[i]), mut) for i in range(n)])
        wrong = np.linalg.norm(out - bad, axis=1) / scale > 1e-3
        assert wrong.sum() > 20, mut


@pytest.mark.parametrize("seed", [3, 4])
def test_walk_jumps_over_uniform_boxes(tables, seed):
    """A 44 x 40 x 36 phantom of uniform 8^3 blocks (densities 0.0012, 0.3,
    1, 1.85) has many uniform bricks, so the walk crosses whole boxes of
    equal density in one step (range_jump); its end points must still be
    the brute-force crossing list's (every segment charged at the density
    of the voxel it enters), for walks that end inside boxes, at box edges,
    in other blocks and in the air outside the grid.  Planted mistakes
    fail as in the voxel-by-voxel test."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = 44, 40, 36
    blocks = rng.choice(np.array([0.0012, 0.3, 1.0, 1.85], np.float32), size=(nz // 8 + 1, ny // 8 + 1, nx // 8 + 1))
    # blocks aligned with the walk map's bricks (padded voxel i + 1 in brick
    # (i + 1) // 4), so that box exits meet density changes
    iz, iy, ix = np.meshgrid((np.arange(nz) + 1) // 8, (np.arange(ny) + 1) // 8, (np.arange(nx) + 1) // 8,
                             indexing="ij")
    dens = blocks[iz, iy, ix].astype(np.float32)
    vox = (0.02, 0.025, 0.03)
    size = (nx * vox[0], ny * vox[1], nz * vox[2])
    off = tuple(-v / 2 for v in size)
    eng = _walk_engine(tables, dens, off, size)
    n = 3000
    p0 = np.column_stack([rng.uniform(off[a] + 0.01, off[a] + size[a] - 0.01, n) for a in range(3)])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # some walks along the axes and the diagonals (ties between axes)
    d[:200] = np.eye(3)[rng.integers(0, 3, 200)] * rng.choice([-1.0, 1.0], (200, 1))
    d[200:300] = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0) * rng.choice([-1.0, 1.0], (100, 3))
    p0 = p0.astype(np.float32)
    d = d.astype(np.float32)
    rr = rng.uniform(0.0, 0.8, n).astype(np.float32)
    out, steps = _oracle_walk(eng, p0, d, rr)
    eng.close()
    want = np.array([_brute_walk(dens, off, vox, p0[i], d[i], float(rr[i])) for i in range(n)])
    dist = np.linalg.norm(out - want, axis=1)
    scale = 1.0 + np.linalg.norm(want - p0, axis=1)
    assert np.max(dist / scale) < 1e-4, np.argmax(dist / scale)
    # the walk crossed far fewer voxels one by one than the brute-force lists hold
    crossings = np.array([_crossings_inside(off, vox, (nx, ny, nz), p0[i], d[i], want[i]) for i in range(n)])
    assert steps.sum() < 0.5 * crossings.sum()
    for mut in ("current", "stop_at_edge", "air_0012"):
        bad = np.array([_brute_walk(dens, off, vox, p0[i], d[i], float(rr[i]), mut) for i in range(n)])
        wrong = np.linalg.norm(out - bad, axis=1) / scale > 1e-3
        assert wrong.sum() > 20, mut


def _crossings_inside(off, vox, n, p0, d, end):
    """Voxel boundaries inside the grid between p0 and end (brute force)."""
    L = float(np.dot(np.asarray(end, float) - p0, d))
    c = 0
    for a in range(3):
        if d[a] == 0:
            continue
        planes = off[a] + vox[a] * np.arange(n[a] + 1)
        t = (planes - p0[a]) / d[a]
        c += int(np.sum((t > 0) & (t <= L)))
    return c


# --------------------------------------------------------- adder / readout --

def _module_scanner():
    """The 8-panel octagon with 2 x 2 modules of 10 x 10 crystals per panel
    (so panel, module and crystal readout differ)."""
    panels, mats = W.octagon_scanner()
    for p in panels:
        p.mody = p.modz = 10 * 0.15 - 0.0005
        p.mspy = p.mspz = 0.01
        p.leny = p.lenz = 2 * p.mody + 0.01
    return panels, mats


def _events_from_hits(hit_id, hit_f, depth, policy, module_n, crystal_n, mutation=None):
    """The adder and readout of photonde (gPET_kernals.cu:737-813) rebuilt
    from the hit records, photon by photon, in float32:
    adder: a deposit in a crystal already in the photon's list adds its
    energy and moves the position to the energy-weighted centroid, else it
    is appended (at most 4 entries; the rest are dropped).
    readout: depth 3 keeps the list; otherwise sites become the world (0),
    the panel (1) or the module (2; policy 1 always uses 2), and entries of
    one site merge into the first one: energy-weighted centroid (policy 1)
    or the entry with the larger energy (policy 0, ties to the later one).
    Planted mistakes: 'unweighted' centroid; 'winner_min' keeps the
    smaller energy; 'no_force' ignores policy 1's forced module depth."""
    f32 = np.float32
    out = []
    order = np.argsort(hit_id[:, 0], kind="stable")
    hit_id, hit_f = hit_id[order], hit_f[order]
    bounds = np.flatnonzero(np.diff(hit_id[:, 0])) + 1
    for idx in np.split(np.arange(len(hit_id)), bounds):
        ev = []
        for i in idx:
            parn, pann, modn, cryn, typ = (int(v) for v in hit_id[i])
            if typ not in (1, 2, 4):
                continue
            E, t, x, y, z = (f32(v) for v in hit_f[i])
            site = pann * module_n * crystal_n + modn * crystal_n + cryn
            for e in ev:
                if e["site"] == site:
                    sE = e["E"] + E
                    if mutation == "unweighted":
                        e["x"], e["y"], e["z"] = (e["x"] + x) / f32(2), (e["y"] + y) / f32(2), (e["z"] + z) / f32(2)
                    else:
                        e["x"] = (e["x"] * e["E"] + x * E) / sE
                        e["y"] = (e["y"] * e["E"] + y * E) / sE
                        e["z"] = (e["z"] * e["E"] + z * E) / sE
                    e["E"] = sE
                    break
            else:
                if len(ev) < 4:
                    ev.append(dict(parn=parn, pann=pann, modn=modn, cryn=cryn, site=site, E=E, t=t, x=x, y=y, z=z))
        if depth != 3:
            dd = 2 if (policy == 1 and mutation != "no_force") else depth
            for e in ev:
                e["site"] = {0: 0, 1: e["pann"], 2: e["pann"] * module_n + e["modn"]}[dd]
            merged = []
            used = [False] * len(ev)
            for i, e0 in enumerate(ev):
                if used[i]:
                    continue
                e0 = dict(e0)
                for j in range(i + 1, len(ev)):
                    e = ev[j]
                    if used[j] or e["site"] != e0["site"]:
                        continue
                    if policy == 1:
                        sE = e0["E"] + e["E"]
                        for a in ("x", "y", "z"):
                            e0[a] = (e0[a] * e0["E"] + e[a] * e["E"]) / sE
                        e0["E"] = sE
                    elif mutation == "winner_min":
                        e0 = e0 if e0["E"] < e["E"] else dict(e)
                    else:
                        e0 = e0 if e0["E"] > e["E"] else dict(e)
                    used[j] = True
                merged.append(e0)
            ev = merged
        out.extend(sorted(ev, key=lambda e: e["site"]))
    return out


def _compare(ev, adder):
    if len(ev) != len(adder):
        return False
    for e, a in zip(ev, adder):
        for k in ("parn", "pann", "modn", "cryn"):
            if e[k] != a[k]:
                return False
        if e["site"] != a["siten"] or a["eventid"] != e["parn"] // 2:
            return False
        for k in ("E", "x", "y", "z"):
            if np.float32(e[k]) != a[k]:
                return False
        if abs(float(e["t"]) - a["t"]) > 1e-6 * abs(a["t"]) + 1e-9:
            return False
    return True


@pytest.mark.parametrize("depth,policy", [(0, 0), (1, 0), (2, 0), (3, 0), (2, 1), (0, 1)])
def test_adder_readout_from_hits(tables, depth, policy):
    """The oracle's adder output equals the adder and readout rebuilt from its
    own hit records, at every readout depth and both policies, on a scanner
    of 2 x 2 modules per panel (2e4 pairs from a point source in a water
    cube; events compared field by field, times to the hits' float32).  The
    rebuilt chain with a planted mistake (unweighted centroid, winner
    keeping the smaller energy, policy 1 without the forced module depth)
    disagrees whenever the case exercises that rule."""
    eng = oracle_engine(engine.make_options(seed=9, record_hits=1))
    eng.set_tables(tables)
    eng.set_phantom(*W.water_cube(16, 0.1))
    panels, mats = _module_scanner()
    eng.set_detector(panels, mats, np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 1], np.float32))
    eng.set_digitizer(W.digitizer(readout_depth=depth, readout_policy=policy, coinc_window_us=0.0))
    src, iso, t0, t1 = W.workload_sources(1)
    r = eng.run_source(src, iso, t0, t1, 20000, adder_cap=200000, hits_cap=400000)
    eng.close()
    module_n, crystal_n, _, _ = engine.panel_counts(panels[0])
    assert (module_n, crystal_n) == (4, 100)
    assert r.n_hits < 400000 and r.n_adder < 200000
    ev = _events_from_hits(r.hit_id, r.hit_f, depth, policy, module_n, crystal_n)
    assert _compare(ev, r.adder), "adder output differs from the rebuilt chain"
    # the rules this case exercises, each broken once
    muts = ["unweighted"]
    if depth != 3 and policy == 0:
        muts.append("winner_min")
    if policy == 1 and depth != 2:
        muts.append("no_force")
    for mut in muts:
        bad = _events_from_hits(r.hit_id, r.hit_f, depth, policy, module_n, crystal_n, mut)
        assert not _compare(bad, r.adder), mut
