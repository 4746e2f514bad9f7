"""Drop-in file formats: the reference's own input files parse as main.cu,
read_file_ro, readSource and loadIsotopes read them; error paths return
status codes instead of exiting."""
import os

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")


def test_input_file_fields():
    from gpet_amd import engine
    inp = engine.read_input(os.path.join(GOLD, "input_PET.in"))
    assert inp.device == 0 and abs(inp.noncol_sigma - 0.0037056) < 1e-9
    assert list(inp.dims) == [200, 200, 200]
    assert list(inp.offset) == [-0.5, -0.5, -0.5] and list(inp.size) == [1, 1, 1]
    assert inp.mat_file == b"input/cylinder_phantom_mat.dat"
    assert inp.n_histories == 1000000 and inp.use_psf == 0 and inp.ptype == 0 and inp.prange == 0
    assert (inp.t0, inp.t1) == (0.0, 120.0)
    assert list(inp.record_sphere) == [0, 0, 0, 5] and inp.eabs == 1000.0
    assert inp.geo_file == b"input/config8.geo" and inp.n_surfaces == 1
    assert list(inp.surfaces[:10]) == [0] * 9 + [1]
    assert (inp.readout_depth, inp.readout_policy) == (2, 1) and inp.e_threshold == 50000
    assert (inp.blur_policy, inp.e_ref, inp.r_ref, inp.slope, inp.space_blur) == (1, 662000, np.float32(0.05), 0, 0)
    assert (inp.dead_level, inp.dead_type, inp.dead_time) == (3, 0, np.float32(2.2))
    assert (inp.win_min, inp.win_max) == (30000, 700000)
    assert inp.time_blur_fwhm_us == 0 and inp.coinc_window_us == 0 and inp.seed == 1
    d = engine.input_digitizer(inp)
    assert d.e_upholder == 2e6 and d.readout_depth == 2


def test_input_file_extension_fields(tmp_path):
    from gpet_amd import engine
    txt = open(os.path.join(GOLD, "input_PET.in")).read()
    txt += "\ntime blur FWHM (us)\n0.0003\ncoincidence window (us), multiples policy\n0.004 1\nseed\n12345\n"
    p = tmp_path / "in.in"
    p.write_text(txt)
    inp = engine.read_input(str(p))
    assert inp.time_blur_fwhm_us == 0.0003 and inp.coinc_window_us == 0.004 and inp.coinc_policy == 1
    assert inp.seed == 12345
    assert inp.noise_rate == 0 and inp.delay_us == 0
    txt += "background singles: rate (1/s), energy mean and sigma (eV)\n2e5 450000 25000\ndelay (us)\n0.1\n"
    p.write_text(txt)
    inp = engine.read_input(str(p))
    assert (inp.noise_rate, inp.noise_e_mean, inp.noise_e_sigma, inp.delay_us) == (2e5, 450000, 25000, 0.1)
    d = engine.input_digitizer(inp)
    assert (d.noise_rate, d.noise_e_mean, d.noise_e_sigma, d.delay_us) == (2e5, 450000, 25000, 0.1)


def test_config8_geometry():
    """Panel 0 from the file, 7 more rotated by 45 degrees about z."""
    from gpet_amd import engine
    panels, mats = engine.read_geometry(os.path.join(GOLD, "config8.geo"))
    assert len(panels) == 8
    assert (mats[0].material, mats[0].density) == (7, np.float32(7.4))
    assert (mats[1].material, mats[1].density) == (0, np.float32(0.001025))
    p0 = panels[0]
    assert (p0.lenx, p0.leny, p0.lenz) == (2, np.float32(15.91), np.float32(22.99))
    assert (p0.ox, p0.oy, p0.oz) == (0, -22.5, 0) and list(p0.ux) == [0, 1, 0]
    for i, p in enumerate(panels):
        a = np.deg2rad(45 * i)
        assert p.panel == i
        assert np.allclose([p.ox, p.oy], [22.5 * np.sin(a), -22.5 * np.cos(a)], atol=1e-5)
        assert np.allclose(p.ux[:], [-np.sin(a), np.cos(a), 0], atol=1e-6)
    # 9 x 13 modules of 8 x 8 crystals (initialize.cu:1074-1086)
    assert engine.panel_counts(p0) == (9 * 13, 64, 9, 8)


def test_multi_ring_extension(tmp_path):
    from gpet_amd import engine, workloads as W
    p = tmp_path / "r.geo"
    p.write_text(W.geometry_text(n_panels=12, radius=20.0, rings=4, ring_spacing=3.1))
    panels, _ = engine.read_geometry(str(p))
    assert len(panels) == 48 and [q.panel for q in panels] == list(range(48))
    zs = sorted(set(round(q.oz, 4) for q in panels))
    assert np.allclose(zs, [-4.65, -1.55, 1.55, 4.65])


def test_sources_and_isotopes():
    from gpet_amd import engine
    src = engine.read_sources(os.path.join(GOLD, "source.txt"))
    assert len(src) == 7 and src[0].natom == 9000000 and src[6].natom == 78750000
    assert (src[1].isotope, src[1].shape) == (0, 1) and abs(src[1].coef[0] + 0.7) < 1e-7
    pt = engine.read_sources(os.path.join(GOLD, "pointsource.txt"))
    assert len(pt) == 1 and pt[0].natom == 1762974000 and pt[0].isotope == 3 and pt[0].shape == 2
    iso = engine.read_isotopes(os.path.join(ROOT, "gpet_amd", "data", "isotopes.txt"))
    assert len(iso) == 4 and abs(iso[0].half_life_s - 6586.26) < 1e-3 and abs(iso[0].beta_ratio - 0.97) < 1e-7
    assert abs(iso[3].coef[7] + 29.796) < 1e-4


def test_psf_roundtrip(tmp_path):
    from gpet_amd import engine
    rec = np.arange(80, dtype=np.float64).reshape(10, 8)
    p = tmp_path / "psf.dat"
    rec.tofile(p)
    with open(p, "ab") as fh:
        fh.write(b"\0" * 17)   # a partial trailing record is ignored
    assert np.array_equal(engine.read_psf(str(p)), rec)


def test_event_writer_roundtrip(tmp_path):
    from gpet_amd import capi, engine
    ev = np.zeros(5, capi.EVENT_DTYPE)
    ev["parn"] = np.arange(5)
    ev["t"] = np.arange(5) * 1.5
    p = tmp_path / "singles.dat"
    engine.write_events(str(p), ev)
    engine.write_events(str(p), ev, append=True)
    back = np.fromfile(p, dtype=capi.EVENT_DTYPE)
    assert len(back) == 10 and np.array_equal(back[5:], ev)
    engine.write_events(str(p), ev[:2])
    assert len(np.fromfile(p, dtype=capi.EVENT_DTYPE)) == 2


def test_error_paths(tmp_path):
    from gpet_amd import engine
    with pytest.raises(engine.GpetError, match="missing"):
        engine.load_tables(str(tmp_path))
    with pytest.raises(engine.GpetError, match="missing"):
        engine.read_geometry(str(tmp_path / "nope.geo"))
    # wrong ndata in a 1-D table
    d = tmp_path / "bad"
    d.mkdir()
    src = engine.TABLES_DIR
    for f in os.listdir(src):
        txt = open(os.path.join(src, f)).read()
        if f.endswith(".lamph"):
            txt = txt.replace("  4096   1", "  4095   1", 1)
        (d / f).write_text(txt)
    with pytest.raises(engine.GpetError, match="dimension"):
        engine.load_tables(str(d))
    # too many quadric surfaces (main.cu:186-190)
    txt = open(os.path.join(GOLD, "input_PET.in")).read().replace("(in cm):\n1\n0 0 0", "(in cm):\n9\n0 0 0")
    p = tmp_path / "in.in"
    p.write_text(txt)
    with pytest.raises(engine.GpetError, match="capacity"):
        engine.read_input(str(p))
    # a rotation axis that is not a coordinate axis is rejected (Q14)
    g = open(os.path.join(GOLD, "config8.geo")).read().replace("\n0 0 1\n", "\n0.6 0 0.8\n")
    q = tmp_path / "g.geo"
    q.write_text(g)
    with pytest.raises(engine.GpetError, match="unsupported"):
        engine.read_geometry(str(q))
