"""Writes the golden fixtures of tests/golden/ (run in a checkout that has the
reference next to it; the outputs are committed, this script is not needed
at test time).

Fixtures (data only):
  philox_kat.json       Philox4x32-10 known answers: the three published
                        Random123 vectors, and 27 outputs of rocRAND's
                        host-callable philox4x32_10_engine (rocrand_kat.cpp).
  gctd_surfaces.npz     Compton and Rayleigh inverse-CDF surfaces of the gCTD
                        data set (Water, LSO) with their S(q)/F(q) blocks: the
                        known answer of the table generator.
  table_nodes.json      (E, mu/rho) rows of input4gPET.{lamph,compt,rayle} at
                        chosen nodes for Water and LSO: pins parsing and lerp.
  input_PET.in, config8.geo, source.txt, pointsource.txt
                        the reference's own input files (small text fixtures).
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def philox(ref_root):
    exe = os.path.join("/tmp", "gpet_rocrand_kat")
    subprocess.check_call(["/opt/rocm/bin/hipcc", "-O1", "-o", exe, os.path.join(HERE, "rocrand_kat.cpp")])
    rows = []
    for line in subprocess.check_output([exe], text=True).splitlines():
        s, sub, off, *v = map(int, line.split())
        rows.append({"seed": s, "subsequence": sub, "offset": off, "out": v})
    published = [
        {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
        {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2, "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
        {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], "key": [0xa4093822, 0x299f31d0],
         "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
    ]
    with open(os.path.join(HERE, "philox_kat.json"), "w") as fh:
        json.dump({"random123": published, "rocrand": rows}, fh, indent=0)


def gctd(ref_root):
    from gpet_amd.tablegen import read_surface_file
    names = []
    for line in open(os.path.join(ref_root, "data", "input4gCTD.matter")):
        if "MATERIAL:" in line:
            names.append(line.split("MATERIAL:")[1].strip())
    out = {}
    for kind, ext in (("compton", "cmpsf"), ("rayleigh", "rayff")):
        mats = read_surface_file(os.path.join(ref_root, "data", "input4gCTD." + ext))
        for n, m in zip(names, mats):
            if n in ("Water", "LSO"):
                out["%s_%s_q" % (kind, n)] = m["q"]
                out["%s_%s_v" % (kind, n)] = m["v"]
                out["%s_%s_surf" % (kind, n)] = m["surf"].astype(np.float32)
                out["%s_%s_cp" % (kind, n)] = m["cp"]
                out["%s_%s_e" % (kind, n)] = m["e"]
    np.savez_compressed(os.path.join(HERE, "gctd_surfaces.npz"), **out)


def table_nodes(ref_root):
    nodes = [0, 1, 100, 1000, 2090, 2091, 2092, 4094, 4095]
    res = {}
    for t in ("lamph", "compt", "rayle"):
        lines = open(os.path.join(ref_root, "data", "input4gPET." + t)).read().split("\n")
        blocks = []
        i = 2
        while i < len(lines):
            if "ndata" in lines[i]:
                rows = [lines[i + 3 + k].split() for k in range(4096)]
                blocks.append(rows)
                i += 3 + 4096
            else:
                i += 1
        res[t] = {str(m): [[k, float(blocks[m][k][0]), float(blocks[m][k][1])] for k in nodes] for m in (1, 7)}
    with open(os.path.join(HERE, "table_nodes.json"), "w") as fh:
        json.dump(res, fh, indent=0)


def inputs(ref_root):
    for src, dst in (("input_PET.in", "input_PET.in"), ("input/config8.geo", "config8.geo"),
                     ("input/source.txt", "source.txt"), ("input/pointsource.txt", "pointsource.txt")):
        shutil.copyfile(os.path.join(ref_root, src), os.path.join(HERE, dst))


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    philox(ref)
    gctd(ref)
    table_nodes(ref)
    inputs(ref)
    print("fixtures written to", HERE)
