"""Cross-section tables: parsing, interpolation, generated surfaces and the
Woodcock majorant."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import ORACLE_LIB, ROOT, oracle_engine

GOLD = os.path.join(ROOT, "tests", "golden")


def test_parsed_nodes_match_reference_text(tables):
    """Node values of input4gPET.{lamph,compt,rayle} for Water (1) and LSO (7)."""
    nodes = json.load(open(os.path.join(GOLD, "table_nodes.json")))
    for k, name in ((0, "lamph"), (1, "compt"), (3, "rayle")):
        for m, rows in nodes[name].items():
            for i, e, v in rows:
                assert tables.mu[k][int(m), i] == np.float32(v), (name, m, i)
                assert np.float32(tables.e0[k] + i / np.float64(tables.idl[k])) == pytest.approx(e, rel=1e-6)
    assert tables.nmat == 10 and tables.eminph == 1000.0 and tables.emax == 1.0e6


def test_lerp_hits_nodes_and_is_linear(tables):
    from gpet_amd import engine
    eng = oracle_engine(engine.make_options())
    eng.set_tables(tables)
    f = eng.lib.gpet_cpu_lerp_mu
    f.restype = C.c_float
    f.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float]
    e0, idl = float(tables.e0[0]), float(tables.idl[0])
    for i in (0, 5, 2091, 4000):
        E = np.float32(e0 + i / idl)
        x = (E - np.float32(e0)) * np.float32(idl)
        if x == np.float32(i):
            assert f(eng.ctx, 0, 7, E) == tables.mu[0][7, i]
    # midpoint of an interval is the mean of the nodes to float rounding
    i = 2091
    Em = np.float32(e0 + (i + 0.5) / idl)
    ref = 0.5 * (float(tables.mu[0][7, i]) + float(tables.mu[0][7, i + 1]))
    assert abs(f(eng.ctx, 0, 7, Em) - ref) <= 1e-5 * ref
    eng.close()


def test_generated_surface_file_matches_generator(tables):
    """The loaded .cmpsf surface equals gpet_amd.tablegen's output (as written)."""
    from gpet_amd import tablegen
    blocks = tablegen.load_blocks(os.path.join(ROOT, "gpet_amd", "data", "gctd_qblocks.json"))
    cps = np.linspace(0, 1, tablegen.PET_NCP)
    es = np.linspace(0, tablegen.PET_EMAX, tablegen.PET_NE)
    for k, kind in ((0, "compton"), (1, "rayleigh")):
        for m, (pet, src) in enumerate(tablegen.PET_MATERIALS):
            if m not in (1, 7):
                continue
            b = blocks[kind][src]
            s = np.clip(tablegen.surface(kind, b["q"], b["v"], cps, es), -1, 1)
            want = np.array([[np.float32(float("%.7e" % v)) for v in row] for row in s], np.float32)
            assert np.array_equal(tables.surf[k][m], want), (kind, pet)
    # 1/dCP and 1/dE as the reference computes them from the header text
    assert tables.idcp[0] == np.float32(1.0) / np.float32(float("%e" % (1.0 / 300)))
    assert tables.ide[0] == np.float32(1.0) / np.float32(float("%e" % 4000.0))


def test_generator_reproduces_gctd_surfaces():
    """Known answer: regenerating the gCTD Water and LSO surfaces (101 CP x 51
    E) from their own S(q)/F(q) blocks reproduces the gCTD files.  The E = 0
    column is excluded (its convention is not recoverable)."""
    from gpet_amd import tablegen
    g = np.load(os.path.join(GOLD, "gctd_surfaces.npz"))
    for mat in ("Water", "LSO"):
        for kind, tol_all, tol_511 in (("compton", 5e-4, 1e-4), ("rayleigh", 1.5e-2, 5e-3)):
            pre = "%s_%s_" % (kind, mat)
            s = tablegen.surface(kind, g[pre + "q"], g[pre + "v"], g[pre + "cp"], g[pre + "e"], n_cos=201)
            d = np.abs(s - g[pre + "surf"])[:, 1:]
            assert d.max() < tol_all, (kind, mat, d.max())
            i511 = int(np.argmin(np.abs(g[pre + "e"] - 511e3)))
            d511 = np.abs(s[:, i511] - g[pre + "surf"][:, i511])
            assert d511.max() < tol_511
            assert np.median(d511) < 5e-5
            assert np.median(d) < 1e-4
    # a wrong momentum-transfer scale fails the test
    pre = "compton_Water_"
    bad = tablegen.surface("compton", g[pre + "q"] * 2.0, g[pre + "v"], g[pre + "cp"], g[pre + "e"])
    assert np.abs(bad - g[pre + "surf"])[:, 1:].max() > 5e-3


def test_generated_compton_surface_tends_to_klein_nishina():
    """With S(q) = 1 the Compton surface is the inverse CDF of pure
    Klein-Nishina, which has a closed form at 511 keV for CP = 0.5 checked by
    direct integration."""
    from gpet_amd import tablegen
    q = np.logspace(-6, 2, 64)
    s = tablegen.surface("compton", q, np.ones_like(q), np.array([0.5]), np.array([511e3]), n_cos=20001)[0, 0]
    c = np.linspace(-1, 1, 2000001)
    k = 511e3 / tablegen.MC2_EV
    eps = 1 / (1 + k * (1 - c))
    f = eps ** 2 * (eps + 1 / eps - (1 - c ** 2))
    cdf = np.cumsum(f)
    cdf /= cdf[-1]
    assert abs(s - c[np.searchsorted(cdf, 0.5)]) < 2e-4


def majorant_setup(tables):
    from gpet_amd import engine, workloads as W
    eng = oracle_engine(engine.make_options())
    W.setup_engine(eng, tables, config=2, phantom_n=50)
    return eng


def test_majorant_invariant_and_tightness(tables):
    """lambda(E) * mu_j(E) * rho_max_j <= 1 on a dense grid for all present
    materials (the reference's spline majorant breaks this, Q18) and is tight
    to within 0.2%."""
    eng = majorant_setup(tables)
    lam, e0, idl = eng.majorant(0)
    mats = {0: 0.0012, 1: 1.0, 3: 1.85}
    E = np.linspace(1000.0, 999000.0, 400001).astype(np.float32)

    def lerp(T, e0_, idl_, x):
        x = np.clip((x - np.float32(e0_)) * np.float32(idl_), 0, len(T) - 1).astype(np.float32)
        j = np.minimum(x.astype(np.int64), len(T) - 2)
        f = x - j.astype(np.float32)
        return T[j] + f * (T[j + 1] - T[j])

    L = lerp(lam, e0, idl, E).astype(np.float64)
    mu = np.max([lerp(tables.mu[0][m], tables.e0[0], tables.idl[0], E).astype(np.float64) * r
                 for m, r in mats.items()], axis=0)
    prod = L * mu
    assert prod.max() <= 1.0
    assert np.median(prod) > 0.998
    eng.close()


def test_product_and_oracle_build_the_same_majorant(tables):
    from gpet_amd import engine
    eng = majorant_setup(tables)
    lam_o, e0_o, idl_o = eng.majorant(0)
    eng.close()
    lib = engine.host_lib()
    rho = np.zeros(10, np.float32)
    rho[0], rho[1], rho[3] = 0.0012, 1.0, 1.85
    out = np.zeros(4096, np.float32)
    e0, idl = C.c_float(), C.c_float()
    t = tables.as_c()
    assert lib.gpet_build_majorant(C.byref(t), rho.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                   C.byref(e0), C.byref(idl)) == 0
    assert np.array_equal(out, lam_o) and e0.value == e0_o and idl.value == idl_o
