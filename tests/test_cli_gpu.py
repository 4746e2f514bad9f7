"""The drop-in executable end to end: gpet_amd/bin/gpet reads an
input_PET.in (the reference's field order, main.cu:50-184, plus the
optional extension lines), the phantom files, config8.geo and a source
file, runs on the GPU and writes its outputs; the numpy reader reads them
back and they equal the oracle's run of the same inputs."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import first_diff, oracle_engine, same_bytes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")


def make_run_dir(d):
    from gpet_amd import workloads as W
    os.makedirs(os.path.join(d, "input"), exist_ok=True)
    mat, den, dims, off, size = W.cylinder_phantom(40)
    np.ascontiguousarray(mat, np.int32).tofile(os.path.join(d, "input", "mat.dat"))
    np.ascontiguousarray(den, np.float32).tofile(os.path.join(d, "input", "den.dat"))
    with open(os.path.join(d, "input", "source.txt"), "w") as f:
        f.write("1\nnatom isotope shape cx cy cz c3 c4 c5#\n1.0e11 0 1 0 0 0 0.4 1.0 0\n")
    with open(os.path.join(GOLD, "config8.geo")) as f:
        geo = f.read()
    with open(os.path.join(d, "input", "config8.geo"), "w") as f:
        f.write(geo)
    txt = open(os.path.join(GOLD, "input_PET.in")).read()
    rep = [("200 200 200", "%d %d %d" % tuple(dims)), ("input/cylinder_phantom_mat.dat", "input/mat.dat"),
           ("input/cylinder_phantom_den.dat", "input/den.dat"), ("input/pointsource.txt", "input/source.txt"),
           ("0 120", "0 0.002"), ("3 0 2.2", "3 0 0.01")]
    for a, b in rep:
        assert txt.count(a) == 1, a
        txt = txt.replace(a, b)
    lines = txt.split("\n")
    k = lines.index("positron range consideration (0 for no, 1 for yes):")
    lines[k + 1] = "1"
    txt = "\n".join(lines).rstrip("\n") + "\n"
    txt += ("time blur FWHM (us)\n0.0003\ncoincidence window (us), multiples policy\n0.004 1\nseed\n77\n"
            "background singles: rate (1/s), energy mean and sigma (eV)\n2e6 450000 30000\ndelay (us)\n0.1\n")
    with open(os.path.join(d, "input_PET.in"), "w") as f:
        f.write(txt)
    return (mat, den, dims, off, size)


def test_cli_matches_oracle(tables, tmp_path):
    from gpet_amd import capi, engine, reader
    d = str(tmp_path)
    mat, den, dims, off, size = make_run_dir(d)
    exe = os.path.join(ROOT, "gpet_amd", "bin", "gpet")
    cmd = [exe, "input_PET.in", "--data", engine.TABLES_DIR, "--out", "output", "--hits", "--listmode",
           "--psf-source", "--psf-phantom", "--sinogram", "32", "24", "30"]
    print("running", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=240)
    assert p.returncode == 0, p.stderr
    summary = json.loads(p.stdout.strip().splitlines()[-1])
    out = reader.read_output_dir(os.path.join(d, "output"))

    inp = engine.read_input(os.path.join(d, "input_PET.in"))
    assert inp.seed == 77 and inp.delay_us == 0.1
    e = oracle_engine(engine.make_options(seed=inp.seed, eabs=inp.eabs, noncol_sigma=inp.noncol_sigma,
                                          positron_range=inp.prange, record_hits=1,
                                          flags=capi.FLAG_PSF_SOURCE | capi.FLAG_PSF_PHANTOM))
    e.set_tables(tables)
    e.set_phantom(mat, den, dims, off, size)
    panels, mats = engine.read_geometry(os.path.join(d, "input", "config8.geo"))
    e.set_detector(panels, mats, np.array(inp.surfaces[:10], np.float32))
    e.set_digitizer(engine.input_digitizer(inp))
    src = engine.read_sources(os.path.join(d, "input", "source.txt"))
    iso = engine.read_isotopes(os.path.join(engine.TABLES_DIR, "isotopes.txt"))
    n = summary["histories"]
    cap = 20 * n + 100000
    ref = e.run_source(src, iso, inp.t0, inp.t1, singles_cap=cap, coinc_cap=cap, adder_cap=cap, hits_cap=20 * cap,
                       delayed_cap=cap, psf_cap=2 * n + 2)
    t, cs, eh, sg, sgd = e.tallies(n_s=32, n_phi=24, s_max=30.0, with_delayed=True)
    e.close()

    assert ref.tallies.n_histories == n > 10000
    assert summary["singles"] == ref.n_singles and summary["coincidences"] == ref.n_coinc
    assert summary["background_singles"] == ref.tallies.n_noise > 0
    assert summary["delayed_coincidences"] == ref.tallies.n_delayed > 0
    for name, a, b in (("adder", out["adder"], ref.adder), ("singles", out["singles"], ref.singles),
                       ("coincidences", out["coincidences"], ref.coinc), ("delayed", out["delayed"], ref.delayed)):
        assert same_bytes(a, b), name + ": " + first_diff(a.ravel(), b.ravel())
    for name, a, b in (("psf_source", out["psf_source"], ref.psf_source),
                       ("psf_phantom", out["psf_phantom"], ref.psf_phantom)):
        assert len(a) == len(b) > 0, name
        for f in ("x", "y", "z", "vx", "vy", "vz", "E", "eventid", "t"):
            assert np.array_equal(a[f], b[f]), name + "." + f
    h = out["hits"]
    ka = np.argsort(h["parn"], kind="stable")
    kb = np.argsort(ref.hit_id[:, 0], kind="stable")
    assert np.array_equal(reader.hits_matrix(h[ka])[:, :5], ref.hit_id[kb])
    lm = out["listmode"]
    assert len(lm) == ref.n_coinc + len(ref.delayed)
    assert int(np.sum((lm["flags"] & capi.LM_TRUE) != 0)) == ref.tallies.n_coinc_true
    assert np.array_equal(out["sinogram"], sg.astype(np.float32))
    assert np.array_equal(out["sinogram_delayed"], sgd.astype(np.float32))
