"""Per-kernel averages of rocprofv3 counter CSVs (--pmc ... --output-format csv).

usage: python -m gpet_amd.pmc_summary <counter_collection.csv> [...]
"""
import collections
import csv
import sys


def kernel_name(s):
    return s.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]


def summarize(paths):
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    for p in paths:
        for r in csv.DictReader(open(p)):
            k = kernel_name(r["Kernel_Name"])
            per[k][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[(k, r["Counter_Name"])].add((p, r["Dispatch_Id"]))
    out = {}
    for k, cs in per.items():
        out[k] = {c: v / max(len(disp[(k, c)]), 1) for c, v in cs.items()}
        out[k]["dispatches"] = max(len(disp[(k, c)]) for c in cs)
    return out


if __name__ == "__main__":
    for k, v in sorted(summarize(sys.argv[1:]).items()):
        print(k, " ".join("%s=%.6g" % kv for kv in sorted(v.items())))
