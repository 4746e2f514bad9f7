"""Python host mirror of the reference's init -> run sequence over the C ABI.

The reference exposes no operator or plugin API (main.cu:26-229 is the only
caller of its host functions); this module is the same sequence for Python
callers and tests:

    eng = Engine(options)                    # gpet_create
    eng.set_tables(load_tables(dir))         # init(): rlamph, rcompt, ...
    eng.set_phantom(mat, dens, dims, off, size)   # loadPhantom + iniwck
    eng.set_detector(panels, mats, quadrics)      # read_file_ro + iniPanel + iniwck
    eng.set_digitizer(dig)
    res = eng.run_source(sources, isotopes, t0, t1)   # sampleParticle
    res = eng.run_psf(records, ptype)                 # simulateParticle

``Engine`` binds the HIP library (gpet_amd/libgpet.so) by default and raises
if it is missing.  The same class can bind any library exporting the C ABI
under another prefix.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpet.so")
TABLES_DIR = os.path.join(HERE, "data", "tables")


class GpetError(RuntimeError):
    pass


_host_lib = None


def host_lib() -> C.CDLL:
    """The HIP library, for its host-only functions (parsers, writers)."""
    global _host_lib
    if _host_lib is None:
        if not os.path.exists(LIB_PATH):
            raise GpetError("gpet_amd/libgpet.so is not built; run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        for name, (rt, at) in capi.HOST_FUNCS.items():
            capi.bind(lib, name, rt, at)
        _host_lib = lib
    return _host_lib


def _check(st: int, what: str, detail: str = "") -> None:
    if st != 0:
        raise GpetError("%s failed: %s%s" % (what, capi.STATUS.get(st, str(st)), (" - " + detail) if detail else ""))


# ---------------------------------------------------------------- tables --

@dataclass
class TableSet:
    """Numpy copy of gpet_tables."""
    nmat: int
    eminph: float
    emax: float
    e0: np.ndarray          # [4]
    idl: np.ndarray         # [4]
    mu: list                # 4 x [nmat, 4096] float32
    idcp: np.ndarray        # [2]
    ide: np.ndarray         # [2]
    surf: list              # 2 x [nmat, 301, 151] float32
    _keep: list = field(default_factory=list, repr=False)

    def as_c(self) -> capi.Tables:
        t = capi.Tables()
        t.nmat = self.nmat
        t.eminph = self.eminph
        t.emax = self.emax
        t.n1d = capi.GPET_NTAB
        self._keep = []
        for k in range(4):
            t.e0[k] = float(self.e0[k])
            t.idl[k] = float(self.idl[k])
            a = np.ascontiguousarray(self.mu[k], dtype=np.float32)
            self._keep.append(a)
            t.mu[k] = a.ctypes.data_as(C.POINTER(C.c_float))
        t.ncp = capi.GPET_NCP
        t.ne = capi.GPET_NE
        for k in range(2):
            t.idcp[k] = float(self.idcp[k])
            t.ide[k] = float(self.ide[k])
            a = np.ascontiguousarray(self.surf[k], dtype=np.float32)
            self._keep.append(a)
            t.surf[k] = a.ctypes.data_as(C.POINTER(C.c_float))
        return t


def load_tables(data_dir: str = TABLES_DIR) -> TableSet:
    lib = host_lib()
    t = capi.Tables()
    _check(lib.gpet_tables_load(data_dir.encode(), C.byref(t)), "loading tables from " + data_dir)
    try:
        n = t.nmat
        mu = [np.ctypeslib.as_array(t.mu[k], shape=(n * capi.GPET_NTAB,)).reshape(n, capi.GPET_NTAB).copy()
              for k in range(4)]
        surf = [np.ctypeslib.as_array(t.surf[k], shape=(n * capi.GPET_NCP * capi.GPET_NE,))
                .reshape(n, capi.GPET_NCP, capi.GPET_NE).copy() for k in range(2)]
        return TableSet(n, t.eminph, t.emax, np.array(t.e0[:], np.float32), np.array(t.idl[:], np.float32), mu,
                        np.array(t.idcp[:], np.float32), np.array(t.ide[:], np.float32), surf)
    finally:
        lib.gpet_tables_free(C.byref(t))


# ------------------------------------------------------------ file formats --

def read_geometry(path: str):
    lib = host_lib()
    n = C.c_int32()
    arr = (capi.Panel * capi.GPET_MAXPANELS)()
    mats = (capi.PanelMaterial * 2)()
    _check(lib.gpet_read_geometry(path.encode(), arr, capi.GPET_MAXPANELS, C.byref(n), mats), "reading " + path)
    return [arr[i] for i in range(n.value)], [mats[0], mats[1]]


def read_sources(path: str):
    lib = host_lib()
    n = C.c_int32()
    arr = (capi.Source * 4096)()
    _check(lib.gpet_read_sources(path.encode(), arr, 4096, C.byref(n)), "reading " + path)
    return [arr[i] for i in range(n.value)]


def read_isotopes(path: str):
    lib = host_lib()
    n = C.c_int32()
    arr = (capi.Isotope * 256)()
    _check(lib.gpet_read_isotopes(path.encode(), arr, 256, C.byref(n)), "reading " + path)
    return [arr[i] for i in range(n.value)]


def read_psf(path: str) -> np.ndarray:
    lib = host_lib()
    n = C.c_int64()
    _check(lib.gpet_read_psf(path.encode(), None, 0, C.byref(n)), "reading " + path)
    out = np.zeros((n.value, 8), np.float64)
    _check(lib.gpet_read_psf(path.encode(), out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)), "reading " + path)
    return out[: n.value]


def read_input(path: str) -> capi.Input:
    lib = host_lib()
    inp = capi.Input()
    _check(lib.gpet_read_input(path.encode(), C.byref(inp)), "reading " + path)
    return inp


def input_digitizer(inp: capi.Input) -> capi.Digitizer:
    d = capi.Digitizer()
    host_lib().gpet_input_digitizer(C.byref(inp), C.byref(d))
    return d


def panel_counts(p0: capi.Panel):
    lib = host_lib()
    a, b, c, d = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    _check(lib.gpet_panel_counts(C.byref(p0), C.byref(a), C.byref(b), C.byref(c), C.byref(d)), "panel counts")
    return a.value, b.value, c.value, d.value


def write_events(path: str, ev: np.ndarray, append: bool = False) -> None:
    ev = np.ascontiguousarray(ev, dtype=capi.EVENT_DTYPE)
    _check(host_lib().gpet_write_events(path.encode(), ev.ctypes.data_as(C.c_void_p), len(ev), int(append)),
           "writing " + path)


def listmode_records(pairs: np.ndarray, module_n: int, crystal_n: int, flags: int = 0) -> np.ndarray:
    """List-mode records (capi.LM_DTYPE) of coincidence pairs ([n, 2] events)."""
    pairs = np.ascontiguousarray(pairs, dtype=capi.EVENT_DTYPE).reshape(-1, 2)
    out = np.zeros(len(pairs), capi.LM_DTYPE)
    _check(host_lib().gpet_listmode_records(pairs.ctypes.data_as(C.c_void_p), len(pairs), module_n, crystal_n,
                                            flags, out.ctypes.data_as(C.c_void_p)), "list-mode records")
    return out


def write_listmode(path: str, prompts: np.ndarray, delayed=None, module_n: int = 1, crystal_n: int = 1,
                   n_panels: int = 1, append: bool = False) -> None:
    """<path> (prompt pairs, then delayed pairs) and its header <path>h."""
    p = np.ascontiguousarray(prompts, dtype=capi.EVENT_DTYPE).reshape(-1, 2)
    d = np.ascontiguousarray(delayed if delayed is not None else np.zeros((0, 2), capi.EVENT_DTYPE),
                             dtype=capi.EVENT_DTYPE).reshape(-1, 2)
    _check(host_lib().gpet_write_listmode(path.encode(), p.ctypes.data_as(C.c_void_p), len(p),
                                          d.ctypes.data_as(C.c_void_p), len(d), module_n, crystal_n, n_panels,
                                          int(append)), "writing " + path)


def write_sinogram(prefix: str, sino: np.ndarray, s_max: float) -> None:
    """<prefix>.s (float32 [n_phi, n_s]) and the Interfile header <prefix>.hs."""
    s = np.ascontiguousarray(sino, dtype=np.int64)
    n_phi, n_s = s.shape
    _check(host_lib().gpet_write_sinogram(prefix.encode(), s.ctypes.data_as(C.c_void_p), n_s, n_phi, float(s_max)),
           "writing " + prefix)


# ------------------------------------------------------------------ runs --

@dataclass
class RunResult:
    singles: np.ndarray
    coinc: np.ndarray        # [n, 2] events
    adder: np.ndarray
    hit_id: np.ndarray       # [n, 5]
    hit_f: np.ndarray        # [n, 5]
    trace: np.ndarray
    tallies: capi.Tallies
    n_singles: int = 0
    n_coinc: int = 0
    n_adder: int = 0
    n_hits: int = 0
    job_counts: np.ndarray = None   # batched runs: [n_jobs, 3] histories, singles, coincidences
    psf_source: np.ndarray = None   # capi.PSF_DTYPE records (GPET_FLAG_PSF_SOURCE)
    psf_phantom: np.ndarray = None  # capi.PSF_DTYPE records (GPET_FLAG_PSF_PHANTOM)
    delayed: np.ndarray = None      # [n, 2] delayed-window pairs (digitizer delay_us > 0)
    n_delayed: int = 0


def make_options(seed=1, eabs=1.0e3, noncol_sigma=0.0037056, positron_range=0, record_hits=0, batch=0,
                 device=0, flags=0, record_sphere=(0.0, 0.0, 0.0, 0.0)) -> capi.Options:
    o = capi.Options()
    o.device = device
    o.seed = seed
    o.eabs = eabs
    o.noncol_sigma = noncol_sigma
    o.positron_range = positron_range
    o.record_hits = record_hits
    o.batch = batch
    o.flags = flags
    for k in range(4):
        o.record_sphere[k] = record_sphere[k]
    return o


class Engine:
    def __init__(self, options: capi.Options, lib_path: str = LIB_PATH, prefix: str = "gpet_"):
        if not os.path.exists(lib_path):
            raise GpetError("%s is not built; run __graft_entry__.build()" % lib_path)
        self.lib = C.CDLL(lib_path)
        self.prefix = prefix
        self.f = {}
        for name, (rt, at) in capi.CTX_FUNCS.items():
            self.f[name] = capi.bind(self.lib, prefix + name, rt, at)
        self.options = options
        ctx = C.c_void_p()
        st = self.f["create"](C.byref(ctx), C.byref(options))
        _check(st, prefix + "create")
        self.ctx = ctx
        self._keep = []

    def close(self):
        if getattr(self, "ctx", None):
            self.f["destroy"](self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ok(self, st, what):
        if st != 0:
            msg = self.f["last_error"](self.ctx) if self.ctx else b""
            _check(st, self.prefix + what, (msg or b"").decode(errors="replace"))

    def set_tables(self, tabs: TableSet):
        t = tabs.as_c()
        self._ok(self.f["set_tables"](self.ctx, C.byref(t)), "set_tables")

    def set_phantom(self, mat: np.ndarray, dens: np.ndarray, dims, offset, size, slot=None):
        """Phantom of single runs, or of batch slot `slot`."""
        mat = np.ascontiguousarray(mat, dtype=np.int32).ravel()
        dens = np.ascontiguousarray(dens, dtype=np.float32).ravel()
        d = (C.c_int32 * 3)(*[int(x) for x in dims])
        o = (C.c_float * 3)(*[float(x) for x in offset])
        s = (C.c_float * 3)(*[float(x) for x in size])
        if slot is None:
            self._ok(self.f["set_phantom"](self.ctx, mat.ctypes.data_as(C.c_void_p),
                                           dens.ctypes.data_as(C.c_void_p), d, o, s), "set_phantom")
        else:
            self._ok(self.f["set_phantom_slot"](self.ctx, int(slot), mat.ctypes.data_as(C.c_void_p),
                                                dens.ctypes.data_as(C.c_void_p), d, o, s), "set_phantom_slot")

    def set_launch(self, block=(256, 256), wpe=(0, 0), blocks_per_cu=(16, 8), batch=1 << 23, transit=None):
        """transit: (steps in the phantom kernel, steps of transit rounds
        1..3); None keeps the library's current setting."""
        L = capi.Launch()
        if transit is None:
            self._ok(self.f["get_launch"](self.ctx, C.byref(L)), "get_launch")
        else:
            t = list(transit) + [0] * (4 - len(transit))
            for k in range(4):
                L.transit[k] = t[k]
        L.block[0], L.block[1] = block
        L.wpe[0], L.wpe[1] = wpe
        L.blocks_per_cu[0], L.blocks_per_cu[1] = blocks_per_cu
        L.batch = batch
        self._ok(self.f["set_launch"](self.ctx, C.byref(L)), "set_launch")

    def get_launch(self) -> dict:
        L = capi.Launch()
        self._ok(self.f["get_launch"](self.ctx, C.byref(L)), "get_launch")
        return L.as_dict()

    def set_detector(self, panels, mats, quadrics=None):
        arr = (capi.Panel * len(panels))(*panels)
        m = (capi.PanelMaterial * 2)(*mats)
        q = None
        nq = 0
        if quadrics is not None and len(quadrics):
            qa = np.ascontiguousarray(quadrics, dtype=np.float32).ravel()
            self._keep.append(qa)
            q = qa.ctypes.data_as(C.c_void_p)
            nq = len(qa) // 10
        self._ok(self.f["set_detector"](self.ctx, arr, len(panels), m, q, nq), "set_detector")

    def set_digitizer(self, dig: capi.Digitizer):
        self._ok(self.f["set_digitizer"](self.ctx, C.byref(dig)), "set_digitizer")

    def majorant(self, which: int):
        out = np.zeros(capi.GPET_NWCK, np.float32)
        e0, idl = C.c_float(), C.c_float()
        self._ok(self.f["get_majorant"](self.ctx, which, out.ctypes.data_as(C.c_void_p), C.byref(e0), C.byref(idl)),
                 "get_majorant")
        return out, e0.value, idl.value

    def _sink(self, singles_cap, coinc_cap, adder_cap, hits_cap, trace_cap, out_dir, psf_cap=0, delayed_cap=0,
              append=False):
        bufs = {
            "delayed": np.zeros(2 * delayed_cap, capi.EVENT_DTYPE),
            "psf_source": np.zeros(psf_cap, capi.PSF_DTYPE),
            "psf_phantom": np.zeros(psf_cap, capi.PSF_DTYPE),
            "singles": np.zeros(singles_cap, capi.EVENT_DTYPE),
            "coinc": np.zeros(2 * coinc_cap, capi.EVENT_DTYPE),
            "adder": np.zeros(adder_cap, capi.EVENT_DTYPE),
            "hit_id": np.zeros((hits_cap, 5), np.int32),
            "hit_f": np.zeros((hits_cap, 5), np.float32),
            "trace": np.zeros(trace_cap, capi.TRACE_DTYPE),
        }
        s = capi.Sink()

        def ptr(a):
            return a.ctypes.data_as(C.c_void_p) if a.size else None

        s.singles, s.singles_cap = ptr(bufs["singles"]), singles_cap
        s.coinc, s.coinc_cap = ptr(bufs["coinc"]), coinc_cap
        s.adder, s.adder_cap = ptr(bufs["adder"]), adder_cap
        s.hit_id, s.hit_f, s.hits_cap = ptr(bufs["hit_id"]), ptr(bufs["hit_f"]), hits_cap
        s.trace, s.trace_cap = ptr(bufs["trace"]), trace_cap
        s.psf_source, s.psf_source_cap = ptr(bufs["psf_source"]), psf_cap
        s.psf_phantom, s.psf_phantom_cap = ptr(bufs["psf_phantom"]), psf_cap
        s.delayed, s.delayed_cap = ptr(bufs["delayed"]), delayed_cap
        s.out_dir = out_dir.encode() if out_dir else None
        s.out_append = 1 if append else 0
        return s, bufs

    def _result(self, s, bufs) -> RunResult:
        t = capi.Tallies()
        self._ok(self.f["get_tallies"](self.ctx, C.byref(t)), "get_tallies")
        ns = min(s.n_singles, s.singles_cap)
        nc = min(s.n_coinc, s.coinc_cap)
        na = min(s.n_adder, s.adder_cap)
        nh = min(s.n_hits, s.hits_cap)
        nt = min(s.n_trace, s.trace_cap)
        r = RunResult(bufs["singles"][:ns], bufs["coinc"][: 2 * nc].reshape(-1, 2), bufs["adder"][:na],
                      bufs["hit_id"][:nh], bufs["hit_f"][:nh], bufs["trace"][:nt], t,
                      s.n_singles, s.n_coinc, s.n_adder, s.n_hits)
        r.psf_source = bufs["psf_source"][:min(s.n_psf_source, s.psf_source_cap)]
        r.psf_phantom = bufs["psf_phantom"][:min(s.n_psf_phantom, s.psf_phantom_cap)]
        r.delayed = bufs["delayed"][: 2 * min(s.n_delayed, s.delayed_cap)].reshape(-1, 2)
        r.n_delayed = s.n_delayed
        return r

    def run_source(self, sources, isotopes, t0: float, t1: float, n_decays: int = 0, *, singles_cap=0,
                   coinc_cap=0, adder_cap=0, hits_cap=0, trace_cap=0, out_dir=None, psf_cap=0, delayed_cap=0,
                   append=False) -> RunResult:
        src = (capi.Source * len(sources))(*sources)
        iso = (capi.Isotope * len(isotopes))(*isotopes)
        s, bufs = self._sink(singles_cap, coinc_cap, adder_cap, hits_cap, trace_cap, out_dir, psf_cap, delayed_cap,
                             append)
        self._ok(self.f["run_source"](self.ctx, src, len(sources), iso, len(isotopes), float(t0), float(t1),
                                      int(n_decays), C.byref(s)), "run_source")
        return self._result(s, bufs)

    def run_batch(self, jobs, sources, isotopes, *, singles_cap=0, coinc_cap=0, adder_cap=0, hits_cap=0,
                  trace_cap=0, out_dir=None, delayed_cap=0, append=False) -> RunResult:
        """jobs: sequence of dicts {phantom, first_source, n_sources, t0, t1, n_decays}."""
        J = (capi.Job * len(jobs))()
        for k, j in enumerate(jobs):
            J[k].phantom = j["phantom"]
            J[k].first_source = j["first_source"]
            J[k].n_sources = j["n_sources"]
            J[k].t0 = j["t0"]
            J[k].t1 = j["t1"]
            J[k].n_decays = j.get("n_decays", 0)
        src = (capi.Source * len(sources))(*sources)
        iso = (capi.Isotope * len(isotopes))(*isotopes)
        s, bufs = self._sink(singles_cap, coinc_cap, adder_cap, hits_cap, trace_cap, out_dir, 0, delayed_cap, append)
        jc = np.zeros((len(jobs), 3), np.int64)
        s.job_counts = jc.ctypes.data_as(C.c_void_p)
        self._ok(self.f["run_batch"](self.ctx, J, len(jobs), src, len(sources), iso, len(isotopes), C.byref(s)),
                 "run_batch")
        r = self._result(s, bufs)
        r.job_counts = jc
        return r

    # ---- detector-response surrogate ---------------------------------------
    def surrogate_harvest(self, sources, isotopes, t0, t1, n_decays=0, capacity=10 ** 7):
        src = (capi.Source * len(sources))(*sources)
        iso = (capi.Isotope * len(isotopes))(*isotopes)
        self._ok(self.f["surrogate_harvest"](self.ctx, src, len(sources), iso, len(isotopes), float(t0), float(t1),
                                             int(n_decays), int(capacity)), "surrogate_harvest")

    def surrogate_train(self, steps, batch=65536, lr=2e-3, seed=1, reinit=False) -> capi.SurrogateReport:
        cfg = capi.SurrogateConfig()
        cfg.steps, cfg.batch, cfg.lr, cfg.seed, cfg.reinit = steps, batch, lr, seed, int(reinit)
        rep = capi.SurrogateReport()
        self._ok(self.f["surrogate_train"](self.ctx, C.byref(cfg), C.byref(rep)), "surrogate_train")
        return rep

    def surrogate_enable(self, on=True):
        self._ok(self.f["surrogate_enable"](self.ctx, int(bool(on))), "surrogate_enable")

    def surrogate_save(self, path):
        self._ok(self.f["surrogate_save"](self.ctx, path.encode()), "surrogate_save")

    def surrogate_load(self, path):
        self._ok(self.f["surrogate_load"](self.ctx, path.encode()), "surrogate_load")

    def surrogate_params(self) -> np.ndarray:
        n = C.c_int64()
        self._ok(self.f["surrogate_get_params"](self.ctx, None, 0, C.byref(n)), "surrogate_get_params")
        w = np.zeros(n.value, np.float32)
        self._ok(self.f["surrogate_get_params"](self.ctx, w.ctypes.data_as(C.c_void_p), n.value, C.byref(n)),
                 "surrogate_get_params")
        return w

    def surrogate_set_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float32)
        self._ok(self.f["surrogate_set_params"](self.ctx, w.ctypes.data_as(C.c_void_p), len(w)),
                 "surrogate_set_params")

    def surrogate_n_params(self) -> int:
        n = C.c_int64()
        self._ok(self.f["surrogate_get_params"](self.ctx, None, 0, C.byref(n)), "surrogate_get_params")
        return n.value

    def surrogate_records(self, limit=None) -> np.ndarray:
        """Harvested training records [n, 12]: 6 features, event, fE, depth, dy, dz, 0."""
        n = C.c_int64()
        self._ok(self.f["surrogate_records"](self.ctx, None, 0, C.byref(n)), "surrogate_records")
        m = n.value if limit is None else min(n.value, int(limit))
        out = np.zeros((m, 12), np.float32)
        if m:
            self._ok(self.f["surrogate_records"](self.ctx, out.ctypes.data_as(C.c_void_p), m, C.byref(n)),
                     "surrogate_records")
        return out

    def surrogate_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, capi.SURR_NX)
        h = np.zeros((len(x), capi.SURR_NH), np.float32)
        self._ok(self.f["surrogate_forward"](self.ctx, x.ctypes.data_as(C.c_void_p), len(x),
                                             h.ctypes.data_as(C.c_void_p)), "surrogate_forward")
        return h

    def surrogate_grad(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, capi.SURR_NX)
        y = np.ascontiguousarray(y, np.float32).reshape(-1, capi.SURR_NY)
        g = np.zeros(self.surrogate_n_params(), np.float32)
        loss = C.c_float()
        self._ok(self.f["surrogate_grad"](self.ctx, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), len(x),
                                          C.byref(loss), g.ctypes.data_as(C.c_void_p)), "surrogate_grad")
        return loss.value, g

    def run_psf(self, rec: np.ndarray, ptype: int, *, singles_cap=0, coinc_cap=0, adder_cap=0, hits_cap=0,
                trace_cap=0, out_dir=None, psf_cap=0, delayed_cap=0, append=False) -> RunResult:
        rec = np.ascontiguousarray(rec, dtype=np.float64).reshape(-1, 8)
        s, bufs = self._sink(singles_cap, coinc_cap, adder_cap, hits_cap, trace_cap, out_dir, psf_cap, delayed_cap,
                             append)
        self._ok(self.f["run_psf"](self.ctx, rec.ctypes.data_as(C.c_void_p), len(rec), int(ptype), C.byref(s)),
                 "run_psf")
        return self._result(s, bufs)

    def run_source_async(self, sources, isotopes, t0: float, t1: float, n_decays: int, singles, coinc):
        """run_source with the singles and coincidence pairs copied
        asynchronously into the caller's pinned buffers (hipmem.PinnedBuffer
        or any host memory given as (address, capacity in events)): the call
        returns once the digitizer has ended, the copies overlap the next run,
        and the buffers hold the results after sink_wait() or when the next
        run reaches its digitizer.  Returns the sink (n_singles, n_coinc)."""
        s = capi.Sink()
        s.singles, s.singles_cap = C.c_void_p(singles[0]), int(singles[1])
        s.coinc, s.coinc_cap = C.c_void_p(coinc[0]), int(coinc[1])
        s.async_copy = 1
        S = (capi.Source * len(sources))(*sources)
        I = (capi.Isotope * len(isotopes))(*isotopes)
        self._ok(self.f["run_source"](self.ctx, S, len(sources), I, len(isotopes), float(t0), float(t1),
                                      int(n_decays), C.byref(s)), "run_source")
        return s

    def sink_wait(self):
        """Blocks until the last run's asynchronous sink copies have ended
        (gpet_sink_wait; the HIP library only)."""
        f = self.lib.gpet_sink_wait
        f.argtypes = [C.c_void_p]
        f.restype = C.c_int32
        self._ok(f(self.ctx), "sink_wait")

    def tallies(self, n_crystals=0, n_ebins=0, e_lo=0.0, e_hi=1.0, n_s=0, n_phi=0, s_max=0.0, with_sinogram=False,
                with_delayed=False):
        """Run statistics, and on request per-crystal singles, the singles
        energy histogram, the coincidence sinogram [n_phi, n_s] and (with
        with_delayed) the delayed-window sinogram after it."""
        t = capi.Tallies()
        cs = np.zeros(n_crystals, np.int64)
        eh = np.zeros(n_ebins, np.int64)
        sg = np.zeros((n_phi, n_s), np.int64)
        sgd = np.zeros((n_phi, n_s), np.int64)
        if n_s and n_phi and s_max > 0:
            t.sinogram = sg.ctypes.data_as(C.c_void_p)
            if with_delayed:
                t.sinogram_delayed = sgd.ctypes.data_as(C.c_void_p)
            t.n_s, t.n_phi, t.s_max = n_s, n_phi, s_max
        t.crystal_singles = cs.ctypes.data_as(C.c_void_p) if n_crystals else None
        t.n_crystals = n_crystals
        t.energy_hist = eh.ctypes.data_as(C.c_void_p) if n_ebins else None
        t.n_ebins = n_ebins
        t.e_lo = e_lo
        t.e_hi = e_hi
        self._ok(self.f["get_tallies"](self.ctx, C.byref(t)), "get_tallies")
        if with_delayed:
            return t, cs, eh, sg, sgd
        if with_sinogram:
            return t, cs, eh, sg
        return t, cs, eh
