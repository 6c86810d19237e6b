"""Summaries of rocprofv3 output databases (rocpd SQLite) for the engine's kernels.

    rocprofv3 --kernel-trace --stats -d OUT -o run -- python3 bench.py ...
    python -m vardepth_amd.profiling kernels OUT/run_results.db
    rocprofv3 --pmc FETCH_SIZE -d OUT -o run -- python3 bench.py --steps 1 --warmup 0 --no-cpu --copy-calib
    python -m vardepth_amd.profiling pmc OUT/run_results.db

`pmc` also reports the 256 MiB device copy that ``bench.py --copy-calib``
runs, so FETCH_SIZE / WRITE_SIZE can be scaled to bytes per dispatch.

    python -m vardepth_amd.profiling sq OUT/run_results.db
        per kernel (summed over its dispatches): duration, clock from
        GRBM_GUI_ACTIVE (summed over the 8 XCDs), wave64 VALU instructions per
        SIMD cycle over the 1024 SIMDs and its fraction of the 0.5 per cycle
        that FP32 wave64 instructions issue at (2 cycles each)
    python -m vardepth_amd.profiling traffic FETCH_DB WRITE_DB [CALIB_DB]
        HBM bytes per dispatch per kernel, each counter scaled by what it
        reads for the 256 MiB copy (taken from CALIB_DB when given)
    python -m vardepth_amd.profiling counters PMC_DB
        per kernel, every counter of the pass summed over its dispatches
    python -m vardepth_amd.profiling passes KT_DB N
        the last N passes of the batch (each starts at a k_prep_reads
        dispatch): per kernel, dispatches, average and total duration, i.e.
        the timed passes after an autotuning sweep, without its candidates
    python -m vardepth_amd.profiling spans KT_DB N
        the last N passes: span of the FP32 class kernels (first start to
        last end) and the sum of their durations; the two differ when
        consecutive classes overlap on two streams
    python -m vardepth_amd.profiling families KT_DB [FETCH_DB WRITE_DB]
        kernels grouped by family (the template instances of one kernel, e.g.
        the per-class k_fwd32<A, R, U>): variants, dispatches per variant, the
        sum over variants of the average duration (one pass of the batch runs
        each variant once) and, with the PMC databases, the summed calibrated
        bytes per pass
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


XCDS = 8
SIMDS = 1024
VALU_PER_CYCLE = 0.5


def by_kernel(db: str):
    """{kernel name: (dispatches, total duration ns, {counter: total value})}."""
    con = sqlite3.connect(db)
    dur = dict(con.execute("select name, sum(end - start) from kernels group by name").fetchall())
    calls = dict(con.execute("select name, count(*) from kernels group by name").fetchall())
    out = {}
    for name, counter, value in con.execute(
            "select k.name, p.counter_name, sum(p.counter_value) from pmc_events p "
            "join kernels k on k.dispatch_id = p.dispatch_id group by k.name, p.counter_name").fetchall():
        out.setdefault(name, (calls[name], dur[name], {}))[2][counter] = value
    return out


def sq_table(db: str):
    rows = []
    for name, (n, dur, c) in sorted(by_kernel(db).items(), key=lambda kv: -kv[1][1]):
        if "SQ_INSTS_VALU" not in c or "GRBM_GUI_ACTIVE" not in c:
            continue
        cycles = c["GRBM_GUI_ACTIVE"] / XCDS
        ipc = c["SQ_INSTS_VALU"] / (SIMDS * cycles) if cycles else 0.0
        rows.append((name, n, dur / 1e6, cycles / dur if dur else 0.0, c["SQ_INSTS_VALU"], ipc, ipc / VALU_PER_CYCLE,
                     c.get("SQ_WAVES", 0), c.get("SQ_INSTS_SALU", 0), c.get("SQ_INSTS_LDS", 0)))
    return rows


def traffic_table(fetch_db: str, write_db: str, calib_db: str | None = None):
    """[(kernel, dispatches, fetch bytes/dispatch, write bytes/dispatch)] and the two factors."""
    fr, wr = pmc(fetch_db), pmc(write_db)
    cf = copy_calibration(pmc(calib_db) if calib_db else fr, "FETCH_SIZE")[1]
    cw = copy_calibration(pmc(calib_db) if calib_db else wr, "WRITE_SIZE")[1]
    if cf is None or cw is None:
        raise SystemExit("no 256 MiB copy in the calibration database (run bench.py with --copy-calib)")
    agg = {}
    for rows, key, f in ((fr, "FETCH_SIZE", cf), (wr, "WRITE_SIZE", cw)):
        for _, name, counter, value in rows:
            if counter == key and "copy" not in name.lower():
                e = agg.setdefault(name, {"FETCH_SIZE": [], "WRITE_SIZE": []})
                e[key].append(value * 1024.0 * f)
    out = []
    for name, e in agg.items():
        nf, nw = len(e["FETCH_SIZE"]), len(e["WRITE_SIZE"])
        out.append((name, max(nf, nw), sum(e["FETCH_SIZE"]) / max(nf, 1), sum(e["WRITE_SIZE"]) / max(nw, 1)))
    return sorted(out, key=lambda r: -(r[2] + r[3])), cf, cw


def family_of(name: str) -> str:
    base = name.split("(")[0].split("<")[0]
    return base.replace("void ", "").split("::")[-1].strip()


def family_table(kt_db: str, fetch_db: str | None = None, write_db: str | None = None):
    fam = {}
    for name, n, avg, _, _, total in kernels(kt_db):
        e = fam.setdefault(family_of(name), {"variants": 0, "calls": 0, "avg_sum": 0.0, "total": 0.0,
                                             "fetch": 0.0, "write": 0.0})
        e["variants"] += 1
        e["calls"] += n
        e["avg_sum"] += avg
        e["total"] += total
    if fetch_db and write_db:
        rows, _, _ = traffic_table(fetch_db, write_db)
        for name, _, fb, wb in rows:
            e = fam.get(family_of(name))
            if e is not None:
                e["fetch"] += fb
                e["write"] += wb
    return sorted(fam.items(), key=lambda kv: -kv[1]["total"])


def last_passes(db: str, n: int):
    """[(name, dispatches, avg_ms, total_ms)] of the kernels dispatched from the
    start of the n-th last k_prep_reads on, sorted by total time."""
    con = sqlite3.connect(db)
    starts = [r[0] for r in con.execute("select start from kernels where name like '%k_prep_reads%' "
                                        "order by start desc limit ?", (int(n),)).fetchall()]
    if not starts:
        return []
    rows = con.execute("select name, count(*), avg(duration), sum(duration) from kernels where start >= ? "
                       "group by name order by sum(duration) desc", (min(starts),)).fetchall()
    return [(r[0], r[1], r[2] / 1e6, r[3] / 1e6) for r in rows]


def class_spans(db: str, n: int):
    """[(span_ms, sum_ms, dispatches)] of the FP32 class kernels in each of the
    last n forward passes (a pass runs from one k_prep_reads to the next).
    The span runs from the first class kernel's start to the last one's end;
    with consecutive classes on two streams it is the class kernels' time,
    and the sum of their durations counts the overlapped time twice."""
    con = sqlite3.connect(db)
    starts = [r[0] for r in con.execute("select start from kernels where name like '%k_prep_reads%' "
                                        "order by start").fetchall()]
    out = []
    for i in range(max(0, len(starts) - int(n)), len(starts)):
        cond = "start >= ?" + (" and start < ?" if i + 1 < len(starts) else "")
        args = (starts[i], starts[i + 1]) if i + 1 < len(starts) else (starts[i],)
        r = con.execute("select min(start), max(start + duration), sum(duration), count(*) from kernels where "
                        + cond + " and (name like '%k_fwd32<%' or name like '%k_fwd32t<%')", args).fetchone()
        if r[3]:
            out.append(((r[1] - r[0]) / 1e6, r[2] / 1e6, r[3]))
    return out


def counter_table(db: str):
    """(counter names, [(kernel, dispatches, {counter: sum})]) sorted by kernel time."""
    con = sqlite3.connect(db)
    rows = con.execute("select k.name, p.counter_name, sum(p.counter_value), count(distinct k.dispatch_id), "
                       "sum(k.duration) from pmc_events p join kernels k on k.dispatch_id = p.dispatch_id "
                       "group by k.name, p.counter_name").fetchall()
    names = sorted({r[1] for r in rows})
    table = {}
    for name, counter, value, n, dur in rows:
        e = table.setdefault(name, {"n": n, "dur": 0.0, "c": {}})
        e["c"][counter] = value
        e["dur"] = max(e["dur"], dur)
    order = sorted(table.items(), key=lambda kv: -kv[1]["dur"])
    return names, [(k, v["n"], v["c"]) for k, v in order]


def main(argv):
    if len(argv) >= 2 and argv[0] == "counters":
        names, rows = counter_table(argv[1])
        print("%-60s %5s " % ("kernel", "calls") + " ".join("%14s" % n[-14:] for n in names))
        for k, n, c in rows:
            print("%-60s %5d " % (k[:60], n) + " ".join("%14.4g" % c.get(x, 0) for x in names))
        return 0
    if len(argv) >= 3 and argv[0] == "spans":
        print("%10s %10s %6s" % ("span ms", "sum ms", "calls"))
        for span, total, calls in class_spans(argv[1], int(argv[2])):
            print("%10.3f %10.3f %6d" % (span, total, calls))
        return 0
    if len(argv) >= 3 and argv[0] == "passes":
        print("%-72s %6s %10s %10s" % ("kernel", "calls", "avg ms", "total ms"))
        for r in last_passes(argv[1], int(argv[2])):
            print("%-72s %6d %10.3f %10.3f" % (r[0][:72], r[1], r[2], r[3]))
        return 0
    if len(argv) >= 2 and argv[0] == "families":
        rows = family_table(*argv[1:4])
        print("%-22s %8s %7s %14s %12s %14s %14s" % ("family", "variants", "calls", "sum avg ms", "total ms",
                                                     "fetch B/pass", "write B/pass"))
        for name, e in rows:
            print("%-22s %8d %7d %14.3f %12.3f %14.4g %14.4g" % (name, e["variants"], e["calls"], e["avg_sum"],
                                                                e["total"], e["fetch"], e["write"]))
        return 0
    if len(argv) >= 2 and argv[0] == "sq":
        print("%-60s %5s %10s %6s %12s %8s %6s %6s" % ("kernel", "calls", "ms", "GHz", "VALU", "per-cyc", "util",
                                                       "waves"))
        for r in sq_table(argv[1]):
            print("%-60s %5d %10.3f %6.2f %12.4g %8.3f %6.3f %6d" % (r[0][:60], r[1], r[2], r[3], r[4], r[5], r[6],
                                                                   r[7]))
        return 0
    if len(argv) >= 3 and argv[0] == "traffic":
        rows, cf, cw = traffic_table(argv[1], argv[2], argv[3] if len(argv) > 3 else None)
        print("calibration: FETCH_SIZE x%.3f, WRITE_SIZE x%.3f (256 MiB copy)" % (cf, cw))
        print("%-60s %5s %14s %14s" % ("kernel", "calls", "fetch B/disp", "write B/disp"))
        for r in rows:
            print("%-60s %5d %14.4g %14.4g" % (r[0][:60], r[1], r[2], r[3]))
        return 0
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
