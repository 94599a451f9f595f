"""Text summary of a rocprofv3 run (its rocpd SQLite output).

    python lgpsf-hessian_amd/rocprof_summary.py RESULTS.db [--top N]

prints per-kernel statistics grouped by kernel name AND launch grid (the
same kernel runs on every AMG level and at several panel counts, and only
a (kernel, grid) row has one algorithmic byte count), and, when the run
collected counters (`rocprofv3 --pmc ...`), the average counter value per
dispatch for each (kernel, grid).
"""
import argparse
import re
import sqlite3
import sys


def short(name):
    name = re.sub(r"\(.*", "", name)
    return name.replace("void ", "")


def kernel_rows(c, top):
    q = ("select name, grid_x, grid_y, count(*), sum(duration), avg(duration), "
         "min(duration), max(duration), max(vgpr_count), max(lds_size) "
         "from kernels group by name, grid_x, grid_y "
         "order by sum(duration) desc")
    rows = list(c.execute(q))
    total = sum(r[4] for r in rows)
    out = ["total kernel time %.1f ms over %d dispatches"
           % (total / 1e6, sum(r[3] for r in rows)), "",
           "| kernel | grid x | grid y | calls | total ms | avg us | min us | max us | vgpr | lds |",
           "|---|---|---|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append("| %s | %d | %d | %d | %.2f | %.1f | %.1f | %.1f | %d | %d |"
                   % (short(r[0]), r[1], r[2], r[3], r[4] / 1e6, r[5] / 1e3,
                      r[6] / 1e3, r[7] / 1e3, r[8], r[9]))
    return out


def counter_rows(c, top):
    try:
        q = ("select kernel_name, grid_size_x, grid_size_y, counter_name, "
             "count(distinct dispatch_id), sum(value) / count(distinct dispatch_id) "
             "from counters_collection group by kernel_name, grid_size_x, "
             "grid_size_y, counter_name order by 6 desc")
        rows = list(c.execute(q))
    except sqlite3.Error:
        return []
    if not rows:
        return []
    out = ["", "| kernel | grid x | grid y | counter | dispatches | avg per dispatch |",
           "|---|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append("| %s | %d | %d | %s | %d | %.1f |"
                   % (short(r[0]), r[1], r[2], r[3], r[4], r[5]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=40)
    a = ap.parse_args()
    c = sqlite3.connect(a.db)
    lines = []
    try:
        lines += kernel_rows(c, a.top)
    except sqlite3.Error as e:
        print("no kernel table: %s" % e, file=sys.stderr)
    lines += counter_rows(c, a.top)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
