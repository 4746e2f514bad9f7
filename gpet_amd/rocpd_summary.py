"""Kernel summary of a rocprofv3 database (rocpd SQLite), as text.

usage: python -m gpet_amd.rocpd_summary <results.db> [top_n] [--by-grid]
Prints, per kernel (per kernel and grid size with --by-grid, which keeps the
launch-shape tuner's trial dispatches apart from the chosen shape's): the
dispatches, total and average duration, share of GPU time, and
VGPR/SGPR/scratch/LDS of the dispatch.
"""
import sqlite3
import sys


def summary(path, top=15, by_grid=False):
    db = sqlite3.connect(path)
    group = "name, grid_x" if by_grid else "name"
    rows = db.execute(
        "select name, count(*), sum(duration), avg(duration), max(vgpr_count), max(sgpr_count), "
        "max(scratch_size), max(lds_size), max(workgroup_x), max(grid_x) from kernels group by %s "
        "order by sum(duration) desc" % group).fetchall()
    total = sum(r[2] for r in rows) or 1
    out = ["| kernel | calls | total us | avg us | % GPU | VGPR | SGPR | scratch B | LDS B | block | grid |",
           "|---|---|---|---|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        name = r[0].replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
        out.append("| %s | %d | %.1f | %.2f | %.2f | %s | %s | %s | %s | %s | %s |"
                   % (name, r[1], r[2] / 1e3, r[3] / 1e3, 100.0 * r[2] / total, r[4], r[5], r[6], r[7], r[8], r[9]))
    return "\n".join(out)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    print(summary(args[0], int(args[1]) if len(args) > 1 else 15, "--by-grid" in sys.argv))
