"""Generator for the Compton and Rayleigh inverse-CDF surfaces.

The reference reads two tables that its repository does not ship,
``data/input4gPET.cmpsf`` and ``data/input4gPET.rayff`` (initialize.cu:483-566
and :664-748).  Each holds, per material, a scattering-function block
(``ndata logqmin logqmax dlogq`` then rows ``q logq S``) followed by a surface
of cos(theta) values on a grid of cumulative probability CP (rows) and photon
energy E (columns): ``nCP CPmin CPmax dCP nE Emin Emax dE``, the CP row, the E
row, then nCP rows of nE values.  The reference samples these surfaces with
``cos = surf[mat](CP=u, E)`` (gPET_kernals.cu:66-88, :128-146).

This module writes those files from scattering functions:

* Compton: p(cos) proportional to Klein-Nishina(E, cos) * S(q),
* Rayleigh: p(cos) proportional to (1 + cos^2)/2 * F(q)^2,

with momentum transfer q = (E / m_e c^2) * sqrt(2 (1 - cos)) in units of
m_e c.  S and F are interpolated log-log in q.  The density is evaluated on a
uniform cos grid of ``n_cos`` nodes, integrated with the trapezoid rule, and
the CDF is inverted by linear interpolation.  With ``n_cos = 201`` this
reproduces the surfaces of the sibling gCTD data set (data/input4gCTD.*) from
their own scattering-function blocks: Compton to within 5e-4 in cos(theta)
everywhere above E = 0, Rayleigh to within 1e-6 at 511 keV (the known-answer
test lives in tests/test_tablegen.py).  That is the convention used here for
the PET tables.
"""
from __future__ import annotations

import json
import os

import numpy as np

MC2_EV = 510.9991e3  # constants.h:34

# PET material order of data/input4gPET.matter:17-176 and the gCTD
# scattering-function block used for each.  Materials absent from the gCTD
# set borrow the block of the closest composition; only the shape of S or F
# enters an inverse CDF, not its normalisation.
PET_MATERIALS = [
    ("DryAir", "DryAir"),
    ("Water", "Water"),
    ("TissueICRU", "TissueICRP"),
    ("CorticalBone", "TissueICRP"),     # approximation: no bone block available
    ("MuscleStriated", "TissueICRP"),   # approximation
    ("Brain", "TissueICRP"),            # approximation
    ("PMMA", "PMMA"),
    ("LSO", "LSO"),
    ("LYSO", "LYSO"),
    ("Pb", "LYSO"),                     # approximation: no lead block available
]

# Surface grid used for the PET tables: 301 CP nodes x 151 energy nodes
# (constants.h:22-27), E from 0 to 600 keV so that 511 keV lies well inside.
PET_NCP = 301
PET_NE = 151
PET_EMAX = 600.0e3


def _interp_loglog(x, xs, ys):
    lx = np.log(np.maximum(x, 1e-300))
    return np.exp(np.interp(lx, np.log(xs), np.log(ys)))


def density(kind: str, energy: float, cos: np.ndarray, q: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Unnormalised angular density on the cos grid."""
    e = max(float(energy), 1.0)
    k = e / MC2_EV
    qq = k * np.sqrt(np.maximum(2.0 * (1.0 - cos), 0.0))
    if kind == "compton":
        s = _interp_loglog(qq, q, v)
        s = np.where(qq < q[0], v[0] * (qq / q[0]) ** 2, s)
        eps = 1.0 / (1.0 + k * (1.0 - cos))
        kn = eps * eps * (eps + 1.0 / eps - (1.0 - cos * cos))
        return kn * s
    if kind == "rayleigh":
        f = _interp_loglog(qq, q, v)
        f = np.where(qq < q[0], v[0], f)
        return 0.5 * (1.0 + cos * cos) * f * f
    raise ValueError(kind)


def surface(kind: str, q, v, cps, energies, n_cos: int = 201) -> np.ndarray:
    """cos(theta) at [icp, ie] such that CDF_E(cos) = CP."""
    q = np.asarray(q, float)
    v = np.asarray(v, float)
    c = np.linspace(-1.0, 1.0, n_cos)
    out = np.empty((len(cps), len(energies)))
    for ie, e in enumerate(energies):
        f = density(kind, e, c, q, v)
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(c))])
        cdf /= cdf[-1]
        out[:, ie] = np.interp(cps, cdf, c)
    return out


def read_surface_file(path: str):
    """Parse a .cmpsf/.rayff file into a list of per-material dicts."""
    lines = open(path).read().split("\n")
    mats = []
    i = 2
    while i < len(lines):
        if "ndata" in lines[i] and "logq" in lines[i]:
            nd = int(lines[i + 1].split()[0])
            rows = [lines[i + 3 + j].split() for j in range(nd)]
            q = np.array([float(r[0]) for r in rows])
            v = np.array([float(r[2]) for r in rows])
            k = i + 3 + nd
            while "nCP" not in lines[k]:
                k += 1
            h = lines[k + 1].split()
            ncp, ne = int(h[0]), int(h[4])
            cp = np.array(lines[k + 2].split(), float)
            e = np.array(lines[k + 3].split(), float)
            s = np.array([lines[k + 4 + r].split() for r in range(ncp)], float)
            mats.append({"q": q, "v": v, "cp": cp, "e": e, "surf": s, "header": h})
            i = k + 4 + ncp
        else:
            i += 1
    return mats


def write_surface_file(path: str, kind: str, blocks: dict, materials=PET_MATERIALS,
                       ncp: int = PET_NCP, ne: int = PET_NE, emax: float = PET_EMAX,
                       n_cos: int = 201) -> None:
    """Write a table in the drop-in text format read by initialize.cu:483-566."""
    cps = np.linspace(0.0, 1.0, ncp)
    energies = np.linspace(0.0, emax, ne)
    dcp = 1.0 / (ncp - 1)
    de = emax / (ne - 1)
    label = "sf" if kind == "compton" else "ff"
    with open(path, "w") as fh:
        fh.write(" This file is part of set:\n input4gPET (generated by gpet_amd.tablegen)\n")
        for pet_name, src in materials:
            b = blocks[src]
            q = np.asarray(b["q"], float)
            v = np.asarray(b["v"], float)
            lq = np.log(q)
            fh.write(" ndata, logqmin, logqmax, dlogq:\n")
            fh.write("  %d  %e  %e  %e\n" % (len(q), lq[0], lq[-1], (lq[-1] - lq[0]) / (len(q) - 1)))
            fh.write(" q(m_e*c) -- logq -- %s   [%s, from %s]\n" % (label, pet_name, src))
            for qi, lqi, vi in zip(q, lq, v):
                fh.write(" %e  %e  %e\n" % (qi, lqi, vi))
            fh.write("\n nCP, CPmin, CPmax, dCP, nE, Emin, Emax, dE:\n")
            fh.write("  %d %e %e %e %d %e %e %e\n" % (ncp, 0.0, 1.0, dcp, ne, 0.0, emax, de))
            fh.write(" " + " ".join("%e" % x for x in cps) + " \n")
            fh.write(" " + " ".join("%e" % x for x in energies) + " \n")
            s = surface(kind, q, v, cps, energies, n_cos=n_cos)
            s = np.clip(s, -1.0, 1.0)
            for r in range(ncp):
                fh.write(" " + " ".join("%.7e" % x for x in s[r]) + " \n")
            fh.write("\n")


def load_blocks(path: str) -> dict:
    with open(path) as fh:
        return json.load(fh)


def generate(out_dir: str, blocks_json: str) -> None:
    blocks = load_blocks(blocks_json)
    os.makedirs(out_dir, exist_ok=True)
    write_surface_file(os.path.join(out_dir, "input4gPET.cmpsf"), "compton", blocks["compton"])
    write_surface_file(os.path.join(out_dir, "input4gPET.rayff"), "rayleigh", blocks["rayleigh"])
