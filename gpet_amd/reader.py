"""Readers for every file the path writes, in numpy only (no library, no
GPU): what output/readOutput.m:1-54 does for the reference's files, plus the
coincidence, delayed-window, list-mode and sinogram outputs.

    from gpet_amd import reader
    out = reader.read_output_dir("output")     # every file present
    s = reader.read_events("output/singles.dat")
    m = reader.events_matrix(s)                # readOutput.m's 11 columns

Record layouts:
  singles.dat, adder.dat      48-byte events: int32 parn, pann, modn, cryn,
                              siten, eventid; float64 t (us); float32 E (eV),
                              x, y, z (panel-local cm)
  coincidences.dat,
  delayed.dat                 the same events, two consecutive per pair
  HitsID.dat / Hits.dat       int32 {parn, panel, module, crystal, type}
                              and float32 {Edep eV, t us, x, y, z} per hit
  out<name>.dat, id<name>.dat,
  time<name>.dat              float32 x y z vx vy vz E, int32 eventid,
                              float64 t (us): phase-space records
  coincidences.lm (+ .lmh)    32-byte list-mode records and their header
  <prefix>.s (+ .hs)          float32 sinogram [views, tangential bins] and
                              its Interfile-style header
"""
import os

import numpy as np

EVENT_DTYPE = np.dtype([("parn", "<i4"), ("pann", "<i4"), ("modn", "<i4"), ("cryn", "<i4"), ("siten", "<i4"),
                        ("eventid", "<i4"), ("t", "<f8"), ("E", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
HIT_DTYPE = np.dtype([("parn", "<i4"), ("panel", "<i4"), ("module", "<i4"), ("crystal", "<i4"), ("type", "<i4"),
                      ("Edep", "<f4"), ("t", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
PSF_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("vx", "<f4"), ("vy", "<f4"), ("vz", "<f4"),
                      ("E", "<f4"), ("eventid", "<i4"), ("t", "<f8")])
LM_DTYPE = np.dtype([("crystal_a", "<u4"), ("crystal_b", "<u4"), ("t_us", "<f8"), ("dt_ps", "<f4"),
                     ("e_a_kev", "<f4"), ("e_b_kev", "<f4"), ("flags", "<u4")])
LM_TRUE, LM_DELAYED, LM_BACKGROUND = 1, 2, 4
HIT_COMPTON, HIT_COMPTON_ABSORBED, HIT_RAYLEIGH, HIT_PHOTO = 1, 2, 3, 4

assert EVENT_DTYPE.itemsize == 48 and LM_DTYPE.itemsize == 32


def _records(path, dtype):
    """Whole records of a binary file (a trailing partial record is ignored,
    as readOutput.m's floor((stop-start)/48))."""
    n = os.path.getsize(path) // dtype.itemsize
    return np.fromfile(path, dtype=dtype, count=n)


def read_events(path: str) -> np.ndarray:
    """singles.dat / adder.dat: structured events."""
    return _records(path, EVENT_DTYPE)


def events_matrix(ev: np.ndarray) -> np.ndarray:
    """readOutput.m's singles matrix: columns parn, panel, module, crystal,
    site, eventid, t, E, x, y, z (float64)."""
    return np.stack([ev[f].astype(np.float64) for f in EVENT_DTYPE.names], axis=1)


def read_pairs(path: str) -> np.ndarray:
    """coincidences.dat / delayed.dat: [n, 2] events."""
    ev = read_events(path)
    return ev[: 2 * (len(ev) // 2)].reshape(-1, 2)


def read_hits(ids_path: str, f_path: str) -> np.ndarray:
    """HitsID.dat + Hits.dat (readOutput.m:1-16): structured hits."""
    ids = np.fromfile(ids_path, dtype="<i4")
    fs = np.fromfile(f_path, dtype="<f4")
    n = min(len(ids), len(fs)) // 5
    out = np.zeros(n, HIT_DTYPE)
    ids = ids[: 5 * n].reshape(n, 5)
    fs = fs[: 5 * n].reshape(n, 5)
    for k, name in enumerate(HIT_DTYPE.names[:5]):
        out[name] = ids[:, k]
    for k, name in enumerate(HIT_DTYPE.names[5:]):
        out[name] = fs[:, k]
    return out


def hits_matrix(h: np.ndarray) -> np.ndarray:
    """readOutput.m's hit matrix: parn, panel, module, crystal, type, Edep,
    t, x, y, z (float64)."""
    return np.stack([h[f].astype(np.float64) for f in HIT_DTYPE.names], axis=1)


def read_psf(directory: str, name: str = "source") -> np.ndarray:
    """out<name>.dat, id<name>.dat, time<name>.dat (readOutput.m:37-54)."""
    xyz = np.fromfile(os.path.join(directory, "out%s.dat" % name), dtype="<f4")
    ids = np.fromfile(os.path.join(directory, "id%s.dat" % name), dtype="<i4")
    t = np.fromfile(os.path.join(directory, "time%s.dat" % name), dtype="<f8")
    n = min(len(xyz) // 7, len(ids), len(t))
    out = np.zeros(n, PSF_DTYPE)
    xyz = xyz[: 7 * n].reshape(n, 7)
    for k, f in enumerate(("x", "y", "z", "vx", "vy", "vz", "E")):
        out[f] = xyz[:, k]
    out["eventid"] = ids[:n]
    out["t"] = t[:n]
    return out


def psf_matrix(p: np.ndarray) -> np.ndarray:
    """readOutput.m's PSF matrix: x y z vx vy vz E, id, t (float64)."""
    return np.stack([p[f].astype(np.float64) for f in PSF_DTYPE.names], axis=1)


def _header(path: str) -> dict:
    """key := value lines (Interfile style); '!' prefixes dropped, values
    kept as text."""
    out = {}
    for line in open(path):
        if ":=" not in line:
            continue
        k, v = line.split(":=", 1)
        out[k.strip().lstrip("!").strip()] = v.strip()
    return out


def read_listmode(path: str):
    """coincidences.lm and its header coincidences.lmh -> (header, records)."""
    hdr = _header(path + "h")
    size = int(hdr.get("record size (bytes)", LM_DTYPE.itemsize))
    if size != LM_DTYPE.itemsize:
        raise ValueError("unexpected list-mode record size %d" % size)
    data = os.path.join(os.path.dirname(path), hdr.get("name of data file", os.path.basename(path)))
    return hdr, _records(data, LM_DTYPE)


def listmode_panels(rec: np.ndarray, hdr: dict):
    """(panel, module, crystal) of both singles of list-mode records."""
    mn = int(hdr["modules per panel"])
    cn = int(hdr["crystals per module"])
    out = []
    for c in (rec["crystal_a"], rec["crystal_b"]):
        c = c.astype(np.int64)
        out.append((c // (mn * cn), (c // cn) % mn, c % cn))
    return out


def read_sinogram(prefix: str):
    """<prefix>.s and <prefix>.hs -> (header, float32 array [views, bins])."""
    hdr = _header(prefix + ".hs")
    n_s = int(hdr["matrix size [1]"])
    n_phi = int(hdr["matrix size [2]"])
    if hdr.get("number format", "float") != "float" or int(hdr.get("number of bytes per pixel", 4)) != 4:
        raise ValueError("unexpected sinogram number format")
    data = os.path.join(os.path.dirname(prefix), hdr["name of data file"])
    a = np.fromfile(data, dtype="<f4", count=n_s * n_phi)
    return hdr, a.reshape(n_phi, n_s)


def read_output_dir(directory: str) -> dict:
    """Every output file present in a directory, by name."""
    out = {}
    j = lambda n: os.path.join(directory, n)  # noqa: E731
    for n in ("singles", "adder"):
        if os.path.exists(j(n + ".dat")):
            out[n] = read_events(j(n + ".dat"))
    for n in ("coincidences", "delayed"):
        if os.path.exists(j(n + ".dat")):
            out[n] = read_pairs(j(n + ".dat"))
    if os.path.exists(j("HitsID.dat")) and os.path.exists(j("Hits.dat")):
        out["hits"] = read_hits(j("HitsID.dat"), j("Hits.dat"))
    for n in ("source", "phantom"):
        if all(os.path.exists(j(p + n + ".dat")) for p in ("out", "id", "time")):
            out["psf_" + n] = read_psf(directory, n)
    if os.path.exists(j("coincidences.lmh")):
        out["listmode_header"], out["listmode"] = read_listmode(j("coincidences.lm"))
    for n in ("sinogram", "sinogram_delayed"):
        if os.path.exists(j(n + ".hs")):
            out[n + "_header"], out[n] = read_sinogram(j(n))
    return out
