"""Philox4x32-10 and the shared elementary functions."""
import ctypes as C
import json
import os

import numpy as np
import pytest
from scipy import stats

from conftest import ORACLE_LIB, ROOT

GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def lib():
    l = C.CDLL(ORACLE_LIB)
    l.gpet_cpu_mathf.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int64]
    l.gpet_cpu_mathd.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int64]
    l.gpet_cpu_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int32, C.c_int64, C.c_void_p]
    return l


def philox(lib, ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.gpet_cpu_philox(c, k, o)
    return list(o)


def test_philox_random123_vectors(lib):
    kat = json.load(open(os.path.join(GOLD, "philox_kat.json")))
    for v in kat["random123"]:
        assert philox(lib, v["ctr"], v["key"]) == v["out"]


def test_philox_matches_rocrand_engine(lib):
    """Counter {block, 0, id_lo, id_hi} with key = seed equals rocRAND's
    engine(seed, subsequence=id, offset=4*block)."""
    kat = json.load(open(os.path.join(GOLD, "philox_kat.json")))
    assert len(kat["rocrand"]) == 27
    for v in kat["rocrand"]:
        blk, sub, s = v["offset"] // 4, v["subsequence"], v["seed"]
        ctr = [blk & 0xFFFFFFFF, blk >> 32, sub & 0xFFFFFFFF, sub >> 32]
        assert philox(lib, ctr, [s & 0xFFFFFFFF, s >> 32]) == v["out"]


def stream(lib, seed, id_, aux, kind, n):
    out = np.zeros(n, np.float64 if kind == 2 else (np.uint32 if kind == 3 else np.float32))
    lib.gpet_cpu_stream(seed, id_, aux, kind, n, out.ctypes.data_as(C.c_void_p))
    return out


def test_stream_is_consecutive_philox_blocks(lib):
    raw = stream(lib, 99, 12345, 2, 3, 12)
    for b in range(3):
        assert list(raw[4 * b:4 * b + 4]) == philox(lib, [b, 2, 12345, 0], [99, 0])


def test_uniform_range_and_distribution(lib):
    u = stream(lib, 1, 7, 0, 0, 200000)
    assert u.min() > 0.0 and u.max() <= 1.0
    assert stats.kstest(u, "uniform").pvalue > 0.01
    d = stream(lib, 1, 7, 0, 2, 100000)
    assert d.min() > 0.0 and d.max() <= 1.0
    assert stats.kstest(d, "uniform").pvalue > 0.01


def test_normal_distribution(lib):
    z = stream(lib, 3, 1, 0, 1, 200000).astype(np.float64)
    assert stats.kstest(z, "norm").pvalue > 0.01


def mathf(lib, kind, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x)
    lib.gpet_cpu_mathf(kind, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), len(x))
    return y


def ulp_err(y, ref):
    ref32 = ref.astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(y.astype(np.float64) - ref) / spacing


def test_logf_error_bound(lib):
    rng = np.random.default_rng(0)
    x = np.concatenate([np.exp(rng.uniform(-80, 80, 200000)), rng.uniform(0.5, 2.0, 200000),
                        (np.arange(1, 1 << 16) * 2.0 ** -24)]).astype(np.float32)
    y = mathf(lib, 0, x)
    e = ulp_err(y, np.log(x.astype(np.float64)))
    far = np.abs(np.log(x.astype(np.float64))) > 1e-3
    assert e[far].max() <= 2.0
    # near 1 the absolute error is bounded by one ulp of 1
    assert np.abs(y[~far] - np.log(x[~far].astype(np.float64))).max() < 2e-7


def test_expf_error_bound(lib):
    x = np.linspace(-87, 88, 400001).astype(np.float32)
    e = ulp_err(mathf(lib, 1, x), np.exp(x.astype(np.float64)))
    assert e.max() <= 2.0


def test_sincos2pi_error_bound(lib):
    u = np.concatenate([np.linspace(0, 1, 200001), (np.arange(1, 1 << 16) * 2.0 ** -24)]).astype(np.float32)
    s = mathf(lib, 2, u)
    c = mathf(lib, 3, u)
    a = 2 * np.pi * u.astype(np.float64)
    assert np.abs(s - np.sin(a)).max() < 3e-7
    assert np.abs(c - np.cos(a)).max() < 3e-7
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 6e-7


def test_sincosf_error_bound(lib):
    x = np.linspace(-50, 50, 200001).astype(np.float32)
    assert np.abs(mathf(lib, 4, x) - np.sin(x.astype(np.float64))).max() < 3e-7
    assert np.abs(mathf(lib, 5, x) - np.cos(x.astype(np.float64))).max() < 3e-7
    small = np.linspace(-0.02, 0.02, 10001).astype(np.float32)
    assert ulp_err(mathf(lib, 5, small), np.cos(small.astype(np.float64))).max() <= 1.0


def test_cbrtf_error_bound(lib):
    x = np.concatenate([np.linspace(1e-6, 1, 100000), np.exp(np.linspace(-60, 60, 10000))]).astype(np.float32)
    assert ulp_err(mathf(lib, 6, x), np.cbrt(x.astype(np.float64))).max() <= 2.0


def test_log_double_error_bound(lib):
    rng = np.random.default_rng(1)
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 100000)), rng.uniform(0.3, 3, 100000)])
    y = np.zeros_like(x)
    lib.gpet_cpu_mathd(0, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), len(x))
    ref = np.log(x)
    far = np.abs(ref) > 1e-3
    assert (np.abs(y - ref)[far] / np.abs(ref[far])).max() < 8e-16
    t = -rng.uniform(1e-12, 0.2, 100000)
    y = np.zeros_like(t)
    lib.gpet_cpu_mathd(1, t.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), len(t))
    assert (np.abs(y - np.log1p(t)) / np.abs(np.log1p(t))).max() < 8e-16


def test_atan2_error_bound(lib):
    rng = np.random.default_rng(9)
    y = rng.normal(size=200000).astype(np.float32)
    x = rng.normal(size=200000).astype(np.float32)
    x[:100] = 0.0
    y[100:200] = 0.0
    inp = np.concatenate([y, x])
    out = np.zeros(len(y), np.float32)
    lib.gpet_cpu_mathf(7, inp.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), len(y))
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.abs(out - ref).max() < 4e-7
