"""Asynchronous sink copies (gpet_sink.async_copy, gpet_sink_wait).

Two runs in a row on one context copy their singles and coincidence pairs
into pinned host buffers on the copy stream; the second run's digitizer has
to wait for the first run's copies, and gpet_sink_wait for the second's.
Both buffers must then hold, byte for byte, what the oracle writes for the
same decays, and nothing past the reported counts may be touched.
"""
import numpy as np
import pytest

from conftest import hip_engine, oracle_engine, same_bytes

pytestmark = pytest.mark.gpu

FILL = 0xAB


def test_async_sink_two_runs_match_oracle(tables):
    from gpet_amd import capi, engine, hipmem
    from gpet_amd import workloads as W

    opt = engine.make_options(seed=17, positron_range=1, batch=7000)
    src, iso, t0, t1 = W.workload_sources(2)
    sizes = (30000, 12000)   # the second run is smaller: the first run's counts must not leak into it

    ref = {}
    ora = oracle_engine(opt)
    try:
        W.setup_engine(ora, tables, config=2, phantom_n=100)
        for n in sizes:
            ref[n] = ora.run_source(src, iso, t0, t1, n_decays=n, singles_cap=n, coinc_cap=n)
    finally:
        ora.close()

    eng = hip_engine(opt)
    bufs = []
    try:
        W.setup_engine(eng, tables, config=2, phantom_n=100)
        for n in sizes:
            sb = hipmem.PinnedBuffer(48 * n)
            cb = hipmem.PinnedBuffer(2 * 48 * n)
            sb.array[:] = FILL
            cb.array[:] = FILL
            print("async sink run: config 2, n_decays", n, flush=True)
            sink = eng.run_source_async(src, iso, t0, t1, n, (sb.ptr, n), (cb.ptr, n))
            bufs.append((n, sb, cb, int(sink.n_singles), int(sink.n_coinc)))
        eng.sink_wait()
        eng.sink_wait()   # nothing pending: returns at once
        for n, sb, cb, ns, nc in bufs:
            r = ref[n]
            assert ns == r.tallies.n_singles and nc == r.tallies.n_coinc
            assert ns > 100 and nc > 10
            singles = sb.array[:48 * ns].view(capi.EVENT_DTYPE)
            pairs = cb.array[:2 * 48 * nc].view(capi.EVENT_DTYPE)
            assert same_bytes(singles, np.ascontiguousarray(r.singles).ravel()), "singles of run %d" % n
            assert same_bytes(pairs, np.ascontiguousarray(r.coinc).ravel()), "pairs of run %d" % n
            assert np.all(sb.array[48 * ns:] == FILL) and np.all(cb.array[2 * 48 * nc:] == FILL)
    finally:
        eng.close()
        for _, sb, cb, _, _ in bufs:
            sb.free()
            cb.free()
