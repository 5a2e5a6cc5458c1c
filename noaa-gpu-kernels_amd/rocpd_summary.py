"""Per-kernel duration summary from a rocprofv3 results database.

rocprofv3 --kernel-trace writes <dir>/<name>_results.db (SQLite, rocpd
schema); its --stats CSV is not always written (for instance when the tool
itself exits abnormally after the program finished).  This prints the same
per-kernel statistics from the database:

    python noaa-gpu-kernels_amd/rocpd_summary.py prof/run_results.db [--dispatches NAME]

--dispatches NAME also lists every dispatch of the kernels whose name contains
NAME in launch order (one template instance can serve several modes of a run).
"""
import sqlite3
import sys


def summarize(path: str) -> str:
    db = sqlite3.connect(path)
    rows = db.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
        "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = [f"{'calls':>6} {'total_ms':>10} {'avg_us':>12} {'min_us':>12} {'max_us':>12} {'pct':>6}  kernel"]
    for name, n, tot, avg, mn, mx in rows:
        out.append(f"{n:>6} {tot / 1e6:>10.3f} {avg / 1e3:>12.1f} {mn / 1e3:>12.1f} {mx / 1e3:>12.1f} "
                   f"{100.0 * tot / total:>5.1f}%  {name[:110]}")
    return "\n".join(out)


def dispatches(path: str, name: str) -> str:
    db = sqlite3.connect(path)
    rows = db.execute("select name, start, duration from kernels where name like ? order by start",
                      (f"%{name}%",)).fetchall()
    out = [f"{'#':>4} {'duration_us':>14}  kernel"]
    for i, (n, _, d) in enumerate(rows):
        out.append(f"{i:>4} {d / 1e3:>14.1f}  {n[:110]}")
    return "\n".join(out)


if __name__ == "__main__":
    print(summarize(sys.argv[1]))
    if len(sys.argv) > 3 and sys.argv[2] == "--dispatches":
        print()
        print(dispatches(sys.argv[1], sys.argv[3]))
