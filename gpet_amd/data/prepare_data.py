"""Data files of the photon transport path.

Two jobs:

``import`` (run once, in a checkout that has the reference data next to it)
    Compresses the cross-section tables data/input4gPET.{matter,lamph,compt,
    phote,rayle} and data/isotopes.txt of the reference into this directory
    (xz), and extracts the scattering-function blocks (S(q) and F(q), 64 nodes
    per material) of data/input4gCTD.{cmpsf,rayff} into gctd_qblocks.json.
    Provenance: utaresearch/gPET, data/ directory; input data in the drop-in
    format, not code.

``materialize`` (run by the build)
    Expands the compressed tables into ``tables/`` and generates the two
    surfaces the reference does not ship (input4gPET.cmpsf, input4gPET.rayff)
    with gpet_amd.tablegen, so that ``tables/`` is a complete data directory
    in the reference's format.
"""
from __future__ import annotations

import json
import lzma
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = ("matter", "lamph", "compt", "phote", "rayle")


def import_reference(ref_root: str) -> None:
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from gpet_amd.tablegen import read_surface_file

    for t in TABLES:
        src = os.path.join(ref_root, "data", "input4gPET." + t)
        with open(src, "rb") as fh:
            raw = fh.read()
        with open(os.path.join(HERE, "input4gPET.%s.xz" % t), "wb") as fh:
            fh.write(lzma.compress(raw, preset=9))
    shutil.copyfile(os.path.join(ref_root, "data", "isotopes.txt"), os.path.join(HERE, "isotopes.txt"))

    names = []
    with open(os.path.join(ref_root, "data", "input4gCTD.matter")) as fh:
        for line in fh:
            if "MATERIAL:" in line:
                names.append(line.split("MATERIAL:")[1].strip())
    out = {}
    for kind, ext in (("compton", "cmpsf"), ("rayleigh", "rayff")):
        mats = read_surface_file(os.path.join(ref_root, "data", "input4gCTD." + ext))
        assert len(mats) == len(names), (len(mats), names)
        out[kind] = {n: {"q": m["q"].tolist(), "v": m["v"].tolist()} for n, m in zip(names, mats)}
    with open(os.path.join(HERE, "gctd_qblocks.json"), "w") as fh:
        json.dump(out, fh)


def tables_dir() -> str:
    return os.path.join(HERE, "tables")


def materialize(force: bool = False) -> str:
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from gpet_amd.tablegen import generate

    out = tables_dir()
    os.makedirs(out, exist_ok=True)
    for t in TABLES:
        dst = os.path.join(out, "input4gPET." + t)
        if force or not os.path.exists(dst):
            with open(os.path.join(HERE, "input4gPET.%s.xz" % t), "rb") as fh:
                raw = lzma.decompress(fh.read())
            with open(dst, "wb") as fh:
                fh.write(raw)
    shutil.copyfile(os.path.join(HERE, "isotopes.txt"), os.path.join(out, "isotopes.txt"))
    if force or not os.path.exists(os.path.join(out, "input4gPET.rayff")):
        generate(out, os.path.join(HERE, "gctd_qblocks.json"))
    return out


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "import":
        import_reference(sys.argv[2])
    elif len(sys.argv) >= 2 and sys.argv[1] == "materialize":
        print(materialize(force="--force" in sys.argv))
    else:
        print("usage: prepare_data.py import <reference_root> | materialize [--force]")
        sys.exit(2)
