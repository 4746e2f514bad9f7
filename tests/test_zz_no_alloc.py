"""The allocation-free entry point gpet_run_psf_noalloc against gpet_run_psf.

Both transport the same phase-space records (positrons, then photons) on the
same Philox streams; the events the allocation-free call writes into device
memory must be, as a set, byte-identical to gpet_run_psf's adder events.
Two calls in a row per particle type, the second with fewer records, check
that the shared scratch memory carries nothing from one call to the next.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import hip_engine

pytestmark = pytest.mark.gpu


def records(n, ptype, seed):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 8))
    rec[:, 0:3] = rng.uniform(-0.3, 0.3, (n, 3))
    rec[:, 3] = np.sort(rng.uniform(0, 1e5, n))
    d = rng.normal(size=(n, 3))
    rec[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 7] = 511e3 if ptype == 1 else rng.uniform(1e4, 6e5, n)
    return rec


def noalloc(eng, rec, ptype, cap):
    """Runs the allocation-free call on device copies of rec (device memory
    from the library's own HIP runtime); returns (status, count, events)."""
    from gpet_amd import capi, hipmem

    f = eng.lib.gpet_run_psf_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    f.restype = C.c_int32
    rec = np.ascontiguousarray(rec, np.float64)
    d_rec = hipmem.DeviceBuffer(rec.nbytes)
    d_rec.upload(rec)
    d_ev = hipmem.DeviceBuffer(max(cap, 1) * 48)
    d_ev.fill(0xAB)
    d_n = hipmem.DeviceBuffer(4)
    d_n.fill(0x7F)
    hipmem.synchronize()
    st = f(eng.ctx, d_rec.ptr, len(rec), ptype, d_ev.ptr, cap, d_n.ptr, None)
    hipmem.synchronize()
    n = int(d_n.download(4).view(np.uint32)[0])
    ev = d_ev.download(min(n, cap) * 48).view(capi.EVENT_DTYPE)
    for b in (d_rec, d_ev, d_n):
        b.free()
    return st, n, ev


def as_set(ev):
    return np.sort(np.ascontiguousarray(ev).view("V48"))


@pytest.mark.parametrize("ptype", [0, 1])
def test_noalloc_matches_run_psf(tables, ptype):
    from gpet_amd import engine
    from gpet_amd import workloads as W

    opt = engine.make_options(seed=9, positron_range=1)
    eng = hip_engine(opt)
    try:
        W.setup_engine(eng, tables, config=2, phantom_n=80)
        for n, seed in ((20000, 3), (7000, 4)):
            rec = records(n, ptype, seed)
            ref = eng.run_psf(rec, ptype, adder_cap=4 * n)
            assert ref.n_adder > 10
            st, cnt, ev = noalloc(eng, rec, ptype, 4 * n)
            assert st == 0, eng.f["last_error"](eng.ctx)
            assert cnt == ref.n_adder
            assert np.array_equal(as_set(ev), as_set(ref.adder)), "events differ (n=%d)" % n
        # a buffer too small: the count still reports every event, the
        # first `cap` are written
        st, cnt, ev = noalloc(eng, rec, ptype, 10)
        assert st == 0 and cnt == ref.n_adder and len(ev) == 10
        # more records than the scratch memory holds: rejected up front
        big = records(65537, ptype, 5)
        st, _, _ = noalloc(eng, big, ptype, 16)
        assert st == -5
    finally:
        eng.close()
