"""The exporters (list-mode, sinogram) and the numpy reader that replaces
output/readOutput.m:1-54, on the oracle's outputs written by the library's
own host-side writers: every file reads back to what was written, and the
list-mode records agree with the pairs they came from (crystal numbering,
times, energies, flags)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import oracle_engine


@pytest.fixture(scope="module")
def run(tables):
    from gpet_amd import capi, engine
    from gpet_amd import workloads as W
    dig = W.digitizer(coinc_policy=1, noise_rate=5.0e7, noise_e_mean=450e3, noise_e_sigma=40e3, delay_us=0.2)
    e = oracle_engine(engine.make_options(seed=8, positron_range=1, record_hits=1,
                                          flags=capi.FLAG_PSF_SOURCE | capi.FLAG_PSF_PHANTOM))
    panels, _ = W.setup_engine(e, tables, config=2, phantom_n=40, dig=dig)
    src, iso, _, _ = W.workload_sources(2)
    n = 30000
    cap = 10 * n
    r = e.run_source(src, iso, 0.0, 3e-4, n_decays=n, singles_cap=cap, coinc_cap=cap, adder_cap=cap,
                     hits_cap=20 * n, delayed_cap=cap, psf_cap=2 * n)
    t, cs, eh, sg, sgd = e.tallies(n_s=32, n_phi=24, s_max=8.0, with_delayed=True)
    e.close()
    return r, panels, sg, sgd


def test_events_and_pairs_round_trip(run, tmp_path):
    from gpet_amd import engine, reader
    r = run[0]
    engine.write_events(str(tmp_path / "singles.dat"), r.singles)
    engine.write_events(str(tmp_path / "coincidences.dat"), r.coinc.ravel())
    s = reader.read_events(str(tmp_path / "singles.dat"))
    assert s.tobytes() == r.singles.tobytes() and len(s) > 100
    m = reader.events_matrix(s)
    assert m.shape == (len(s), 11)
    assert np.array_equal(m[:, 0], s["parn"]) and np.array_equal(m[:, 6], s["t"])
    assert np.array_equal(m[:, 7], s["E"].astype(np.float64))
    p = reader.read_pairs(str(tmp_path / "coincidences.dat"))
    assert p.shape == r.coinc.shape and p.tobytes() == r.coinc.tobytes()
    # a trailing partial record is ignored, as readOutput.m does
    with open(tmp_path / "singles.dat", "ab") as f:
        f.write(b"\0" * 20)
    assert len(reader.read_events(str(tmp_path / "singles.dat"))) == len(s)


def test_hits_round_trip(run, tmp_path):
    from gpet_amd import engine, reader
    r = run[0]
    lib = engine.host_lib()
    ids = np.ascontiguousarray(r.hit_id)
    fs = np.ascontiguousarray(r.hit_f)
    assert lib.gpet_write_hits(str(tmp_path / "HitsID.dat").encode(), str(tmp_path / "Hits.dat").encode(),
                               ids.ctypes.data_as(C.c_void_p), fs.ctypes.data_as(C.c_void_p), len(ids), 0) == 0
    h = reader.read_hits(str(tmp_path / "HitsID.dat"), str(tmp_path / "Hits.dat"))
    assert len(h) == len(ids) > 100
    assert np.array_equal(reader.hits_matrix(h)[:, :5], ids)
    assert np.array_equal(h["Edep"], fs[:, 0]) and np.array_equal(h["z"], fs[:, 4])
    assert set(np.unique(h["type"])) <= {1, 2, 3, 4}


def test_phase_space_round_trip(run, tmp_path):
    from gpet_amd import engine, reader
    r = run[0]
    lib = engine.host_lib()
    for name, rec in (("source", r.psf_source), ("phantom", r.psf_phantom)):
        rec = np.ascontiguousarray(rec)
        assert lib.gpet_write_psf(str(tmp_path).encode(), name.encode(), rec.ctypes.data_as(C.c_void_p), len(rec),
                                  0) == 0
        p = reader.read_psf(str(tmp_path), name)
        assert len(p) == len(rec) > 100
        for f in ("x", "y", "z", "vx", "vy", "vz", "E", "eventid", "t"):
            assert np.array_equal(p[f], rec[f]), f
        assert reader.psf_matrix(p).shape == (len(p), 9)


def test_listmode_records(run, tmp_path):
    from gpet_amd import capi, engine, reader
    r, panels = run[0], run[1]
    mod_n, cry_n, _, _ = engine.panel_counts(panels[0])
    path = str(tmp_path / "coincidences.lm")
    engine.write_listmode(path, r.coinc, r.delayed, mod_n, cry_n, len(panels))
    hdr, lm = reader.read_listmode(path)
    n_p, n_d = len(r.coinc), len(r.delayed)
    assert n_p > 100 and n_d > 100
    assert int(hdr["number of prompts"]) == n_p and int(hdr["number of delayed"]) == n_d
    assert len(lm) == n_p + n_d
    pairs = np.concatenate([r.coinc, r.delayed])
    a, b = pairs[:, 0], pairs[:, 1]
    for ev, key in ((a, "crystal_a"), (b, "crystal_b")):
        assert np.array_equal(lm[key], (ev["pann"].astype(np.int64) * mod_n + ev["modn"]) * cry_n + ev["cryn"])
    (pa, ma, ca), (pb, mb, cb) = reader.listmode_panels(lm, hdr)
    assert np.array_equal(pa, a["pann"]) and np.array_equal(mb, b["modn"]) and np.array_equal(cb, b["cryn"])
    assert np.array_equal(lm["t_us"], a["t"])
    assert np.array_equal(lm["dt_ps"], ((b["t"] - a["t"]) * 1e6).astype(np.float32))
    assert np.all(np.abs(lm["dt_ps"][:n_p]) <= 4e3 * 1.0001)          # inside the 4 ns window
    assert np.array_equal(lm["e_a_kev"], (a["E"] * np.float32(1e-3)).astype(np.float32))
    fl = lm["flags"]
    assert np.all(fl[n_p:] & capi.LM_DELAYED) and not np.any(fl[:n_p] & capi.LM_DELAYED)
    bg = (a["parn"] < 0) | (b["parn"] < 0)
    assert np.array_equal((fl & capi.LM_BACKGROUND) != 0, bg)
    assert int(np.sum((fl[:n_p] & capi.LM_TRUE) != 0)) == r.tallies.n_coinc_true
    assert not np.any(fl[n_p:] & capi.LM_TRUE)
    # the library's own conversion gives the same records
    assert engine.listmode_records(r.coinc, mod_n, cry_n).tobytes() == lm[:n_p].tobytes()
    # append mode: the header counts both runs
    engine.write_listmode(path, r.coinc, r.delayed, mod_n, cry_n, len(panels), append=True)
    hdr2, lm2 = reader.read_listmode(path)
    assert int(hdr2["number of prompts"]) == 2 * n_p and len(lm2) == 2 * len(lm)
    assert lm2[len(lm):].tobytes() == lm.tobytes()


def test_sinogram_round_trip(run, tmp_path):
    from gpet_amd import engine, reader
    r, sg, sgd = run[0], run[2], run[3]
    assert sg.sum() > 0 and sgd.sum() > 0
    assert sgd.sum() <= r.tallies.n_delayed and sg.sum() <= r.tallies.n_coinc
    engine.write_sinogram(str(tmp_path / "sinogram"), sg, 8.0)
    hdr, a = reader.read_sinogram(str(tmp_path / "sinogram"))
    assert a.shape == (24, 32) and np.array_equal(a, sg.astype(np.float32))
    assert int(hdr["matrix size [1]"]) == 32 and int(hdr["matrix size [2]"]) == 24
    assert abs(float(hdr["scale factor (mm/pixel) [1]"]) - 160.0 / 32) < 1e-6
    assert hdr["INTERFILE"] == "" and hdr["name of data file"] == "sinogram.s"


def test_read_output_dir(run, tmp_path):
    from gpet_amd import engine, reader
    r, panels, sg, sgd = run
    d = str(tmp_path)
    engine.write_events(os.path.join(d, "singles.dat"), r.singles)
    engine.write_events(os.path.join(d, "adder.dat"), r.adder)
    engine.write_events(os.path.join(d, "coincidences.dat"), r.coinc.ravel())
    engine.write_events(os.path.join(d, "delayed.dat"), r.delayed.ravel())
    mod_n, cry_n, _, _ = engine.panel_counts(panels[0])
    engine.write_listmode(os.path.join(d, "coincidences.lm"), r.coinc, r.delayed, mod_n, cry_n, len(panels))
    engine.write_sinogram(os.path.join(d, "sinogram"), sg, 8.0)
    engine.write_sinogram(os.path.join(d, "sinogram_delayed"), sgd, 8.0)
    out = reader.read_output_dir(d)
    assert {"singles", "adder", "coincidences", "delayed", "listmode", "sinogram", "sinogram_delayed"} <= set(out)
    assert out["adder"].tobytes() == r.adder.tobytes()
    assert out["delayed"].tobytes() == r.delayed.tobytes()
    assert len(out["listmode"]) == len(r.coinc) + len(r.delayed)
    assert np.array_equal(out["sinogram_delayed"], sgd.astype(np.float32))
