"""Summaries of rocprofv3 output databases (rocpd SQLite) for the engine's kernels.

    rocprofv3 --kernel-trace --stats -d OUT -o run -- python3 bench.py ...
    python -m vardepth_amd.profiling kernels OUT/run_results.db
    rocprofv3 --pmc FETCH_SIZE -d OUT -o run -- python3 bench.py --steps 1 --warmup 0 --no-cpu --copy-calib
    python -m vardepth_amd.profiling pmc OUT/run_results.db

`pmc` also reports the 256 MiB device copy that ``bench.py --copy-calib``
runs, so FETCH_SIZE / WRITE_SIZE can be scaled to bytes per dispatch.
"""
from __future__ import annotations

import sqlite3
import sys

COPY_BYTES = 256 << 20


def kernels(db: str):
    """[(name, dispatches, avg_ms, min_ms, max_ms, total_ms)] sorted by total time."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), avg(duration), min(duration), max(duration), sum(duration) "
                       "from kernels group by name order by sum(duration) desc").fetchall()
    return [(r[0], r[1], r[2] / 1e6, r[3] / 1e6, r[4] / 1e6, r[5] / 1e6) for r in rows]


def pmc(db: str):
    """[(dispatch_id, kernel name, counter, value)] in dispatch order."""
    con = sqlite3.connect(db)
    return con.execute("select k.dispatch_id, k.name, p.counter_name, sum(p.counter_value) "
                       "from pmc_events p join kernels k on k.dispatch_id = p.dispatch_id "
                       "group by k.dispatch_id, p.counter_name order by k.start").fetchall()


def copy_calibration(rows, counter: str):
    """counter value (KiB) of the largest copy kernel, and bytes / (value*1024)."""
    copies = [r for r in rows if "copy" in r[1].lower() and r[2] == counter]
    if not copies:
        return None, None
    best = max(copies, key=lambda r: r[3])
    return best[3], COPY_BYTES / (best[3] * 1024.0)


def main(argv):
    if len(argv) < 2 or argv[0] not in ("kernels", "pmc"):
        print(__doc__)
        return 2
    what, dbs = argv[0], argv[1:]
    for db in dbs:
        print("#", db)
        if what == "kernels":
            print("%-70s %6s %12s %12s %12s %12s" % ("kernel", "calls", "avg_ms", "min_ms", "max_ms", "total_ms"))
            for r in kernels(db):
                print("%-70s %6d %12.3f %12.3f %12.3f %12.3f" % ((r[0][:70],) + tuple(r[1:])))
        else:
            rows = pmc(db)
            for counter in sorted({r[2] for r in rows}):
                val, factor = copy_calibration(rows, counter)
                if val is not None:
                    print("calibration %s: 256 MiB copy reads %.1f KiB -> factor %.3f" % (counter, val, factor))
            for r in rows:
                print("%-6d %-60s %-16s %16.1f" % (r[0], r[1][:60], r[2], r[3]))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
