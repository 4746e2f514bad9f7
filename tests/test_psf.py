"""Phase-space outputs (the reference's OUTPUTPSF 1/2 files) and the
recording sphere (RECORDPSF -1, getDistance) on the oracle."""
import os
import tempfile

import numpy as np
import pytest

from conftest import oracle_engine


def run(tables, flags, sphere=(0, 0, 0, 0), n=4000, trace=True):
    from gpet_amd import capi, engine
    from gpet_amd import workloads as W
    e = oracle_engine(engine.make_options(seed=6, positron_range=1, flags=flags, record_sphere=sphere))
    W.setup_engine(e, tables, config=2, phantom_n=50)
    src, iso, t0, t1 = W.workload_sources(2)
    r = e.run_source(src, iso, t0, t1, n_decays=n, trace_cap=2 * n if trace else 0, psf_cap=2 * n)
    e.close()
    return r


def test_source_and_phantom_records(tables):
    from gpet_amd import capi
    r = run(tables, capi.FLAG_PSF_SOURCE | capi.FLAG_PSF_PHANTOM)
    tr = r.trace
    alive = (tr["status"] & 8) != 0
    src = r.psf_source
    assert len(src) == alive.sum() == r.tallies.n_photons
    assert np.array_equal(src["parn"], np.flatnonzero(alive))          # ordered by photon id
    sel = tr[alive]
    for a, b in (("x", "x0"), ("y", "y0"), ("z", "z0"), ("E", "E0")):
        assert np.array_equal(src[a], sel[b])
    assert np.array_equal(src["t"], sel["t0"])
    assert np.array_equal(src["eventid"], src["parn"] // 2)
    assert np.allclose(src["vx"] ** 2 + src["vy"] ** 2 + src["vz"] ** 2, 1, atol=1e-5)
    # the two photons of a decay start at the same point, nearly back to back
    pairs = src.reshape(-1, 2)
    assert np.array_equal(pairs["x"][:, 0], pairs["x"][:, 1])
    dots = (pairs["vx"][:, 0] * pairs["vx"][:, 1] + pairs["vy"][:, 0] * pairs["vy"][:, 1]
            + pairs["vz"][:, 0] * pairs["vz"][:, 1])
    assert np.all(dots < -0.999)
    esc = (tr["status"] & 2) != 0
    ph = r.psf_phantom
    assert np.array_equal(ph["parn"], np.flatnonzero(esc))
    for a in ("x", "y", "z", "vx", "vy", "vz", "E", "t"):
        assert np.array_equal(ph[a], tr[esc][a])


def test_recording_sphere(tables):
    """Escaped photons are moved onto the sphere along their direction, time
    advanced by the path over c."""
    from gpet_amd import capi
    c, R = np.array([0.05, -0.02, 0.0]), 40.0
    a = run(tables, capi.FLAG_PSF_PHANTOM)
    b = run(tables, capi.FLAG_PSF_PHANTOM, sphere=(c[0], c[1], c[2], R))
    pa, pb = a.psf_phantom, b.psf_phantom
    assert np.array_equal(pa["parn"], pb["parn"]) and len(pa) > 1000
    p = np.stack([pb["x"], pb["y"], pb["z"]], 1).astype(float)
    q = np.stack([pa["x"], pa["y"], pa["z"]], 1).astype(float)
    v = np.stack([pa["vx"], pa["vy"], pa["vz"]], 1).astype(float)
    # photons leaving the phantom inside the sphere (most; positrons that left
    # the phantom can annihilate far outside it)
    inside = np.linalg.norm(q - c, axis=1) < R
    assert inside.mean() > 0.8
    assert np.allclose(np.linalg.norm(p[inside] - c, axis=1), R, rtol=1e-5)
    d = p - q
    s = np.einsum("ij,ij->i", d, v)
    assert np.all(s[inside] > 0)                          # inside: the forward (far) root
    assert np.allclose(np.cross(d, v), 0, atol=2e-5)      # every photon moves along its direction
    assert np.allclose(pb["t"] - pa["t"], s / 29979.2458, rtol=1e-4, atol=1e-9)
    # outside the sphere: on it (a root exists) or unchanged (the line misses)
    out = ~inside
    on = np.isclose(np.linalg.norm(p[out] - c, axis=1), R, rtol=1e-5)
    assert np.all(on | np.all(d[out] == 0, axis=1))


def test_sphere_root_rules():
    """getDistance's root choice (gPET_kernals.cu:148-171) restated in numpy:
    inside -> far root; outside approaching -> near root; outside receding ->
    (-b + sqrt)/2a (negative); miss -> 0."""
    def dist(x, v, c, r):
        x = np.asarray(x, float) - c
        a, b, cc = v @ v, 2 * v @ x, x @ x - r * r
        dsc = b * b - 4 * a * cc
        if dsc < 0:
            return 0.0
        if cc < 0 or b >= 0:
            return (-b + np.sqrt(dsc)) / (2 * a)
        return (-b - np.sqrt(dsc)) / (2 * a)
    c = np.zeros(3)
    assert dist([0, 0, 0], np.array([1.0, 0, 0]), c, 2.0) == pytest.approx(2.0)
    assert dist([-5, 0, 0], np.array([1.0, 0, 0]), c, 2.0) == pytest.approx(3.0)
    assert dist([5, 0, 0], np.array([1.0, 0, 0]), c, 2.0) == pytest.approx(-3.0)
    assert dist([5, 5, 0], np.array([1.0, 0, 0]), c, 2.0) == 0.0


def test_psf_writer(tables):
    from gpet_amd import capi, engine
    r = run(tables, capi.FLAG_PSF_SOURCE, n=500, trace=False)
    d = tempfile.mkdtemp()
    lib = engine.host_lib()
    rec = np.ascontiguousarray(r.psf_source)
    assert lib.gpet_write_psf(d.encode(), b"source", rec.ctypes.data, len(rec), 0) == 0
    out = np.fromfile(os.path.join(d, "outsource.dat"), np.float32).reshape(-1, 7)
    ids = np.fromfile(os.path.join(d, "idsource.dat"), np.int32)
    t = np.fromfile(os.path.join(d, "timesource.dat"), np.float64)
    assert len(out) == len(ids) == len(t) == len(rec)
    assert np.array_equal(out[:, 6], rec["E"]) and np.array_equal(out[:, 3], rec["vx"])
    assert np.array_equal(ids, rec["eventid"]) and np.array_equal(t, rec["t"])
