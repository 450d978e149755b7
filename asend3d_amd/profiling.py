"""Per-kernel summaries of a rocprofv3 kernel trace (`--kernel-trace
--output-format csv`), optionally restricted to a time window delimited by
launches of named kernels.

    python -m asend3d_amd.profiling TRACE.csv [--after NAME:K] [--until NAME:K] [--top 25] [--gaps 15]
                                     [--by-grid]

--after NAME:K starts the window at the end of the K-th (1-based) launch of
the kernel whose name contains NAME; --until NAME:K ends it at the end of the
K-th launch.  Prints a markdown table: kernel, launches, total, average and
share of the window's kernel time, plus the window's span; --gaps N adds the N
longest idle intervals of the device (no kernel running) with the kernels on
either side, which is where the host held the stream.  --by-grid keys the
table by kernel and grid size, which separates the launches of one kernel on
the different levels of a multigrid hierarchy.
"""
import argparse
import csv
import sys
from collections import defaultdict


def load(path, by_grid=False):
    rows = []
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            name = r["Kernel_Name"]
            if by_grid:
                g = r.get("Grid_Size_X") or r.get("Grid_Size") or "?"
                name = f"{name} @grid {g}"
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), name))
    rows.sort()
    return rows


def nth_end(rows, name, k):
    seen = 0
    for s, e, n in rows:
        if name in n:
            seen += 1
            if seen == k:
                return e
    raise SystemExit(f"kernel containing {name!r} launched {seen} times, fewer than {k}")


def short(name, width=90):
    """Kernel name without return type, namespaces' anonymous part and
    argument list."""
    n = name.replace("(anonymous namespace)::", "")
    grid = ""
    if " @grid " in n:
        n, grid = n.split(" @grid ")
        grid = f" [grid {grid}]"
    n = n.split("(")[0] + grid
    if n.startswith("void "):
        n = n[5:]
    return n if len(n) <= width else n[:width - 3] + "..."


def summarize(rows, t0=None, t1=None, top=25):
    sel = [(s, e, n) for s, e, n in rows if (t0 is None or s >= t0) and (t1 is None or e <= t1)]
    if not sel:
        raise SystemExit("no kernels in the window")
    agg = defaultdict(lambda: [0, 0])
    for s, e, n in sel:
        agg[n][0] += 1
        agg[n][1] += e - s
    busy = sum(v[1] for v in agg.values())
    span = max(e for _, e, _ in sel) - min(s for s, _, _ in sel)
    out = [f"window: {len(sel)} launches, kernel time {busy / 1e6:.3f} ms, span {span / 1e6:.3f} ms "
           f"(first launch to last end)", "",
           "| kernel | launches | total ms | avg us | share |", "|---|---:|---:|---:|---:|"]
    for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"| `{short(n)}` | {c} | {t / 1e6:.3f} | {t / c / 1e3:.2f} | {100.0 * t / busy:.1f}% |")
    return "\n".join(out)


def gaps(rows, t0=None, t1=None, top=15):
    sel = [(s, e, n) for s, e, n in rows if (t0 is None or s >= t0) and (t1 is None or e <= t1)]
    out, busy_end, prev = [], None, None
    for s, e, n in sel:
        if busy_end is not None and s > busy_end:
            out.append((s - busy_end, prev, n))
        if busy_end is None or e > busy_end:
            busy_end, prev = e, n
    out.sort(key=lambda g: -g[0])
    idle = sum(g[0] for g in out)
    lines = [f"idle: {idle / 1e6:.3f} ms in {len(out)} gaps", "", "| gap us | after | before |", "|---:|---|---|"]
    for g, a, b in out[:top]:
        lines.append(f"| {g / 1e3:.1f} | `{short(a, 60)}` | `{short(b, 60)}` |")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--after", default=None)
    ap.add_argument("--until", default=None)
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--gaps", type=int, default=0)
    ap.add_argument("--by-grid", action="store_true")
    a = ap.parse_args()
    rows = load(a.trace, a.by_grid)
    t0 = t1 = None
    if a.after:
        name, k = a.after.rsplit(":", 1)
        t0 = nth_end(rows, name, int(k))
    if a.until:
        name, k = a.until.rsplit(":", 1)
        t1 = nth_end(rows, name, int(k))
    print(summarize(rows, t0, t1, a.top))
    if a.gaps:
        print()
        print(gaps(rows, t0, t1, a.gaps))


if __name__ == "__main__":
    sys.exit(main())
