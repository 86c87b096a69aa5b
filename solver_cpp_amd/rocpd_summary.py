"""Summarise rocprofv3 output databases (rocpd SQLite) as short text.

  python -m solver_cpp_amd.rocpd_summary trace <results.db> [--top N]
  python -m solver_cpp_amd.rocpd_summary pmc <results.db> [--kernel SUBSTR]
  python -m solver_cpp_amd.rocpd_summary traffic <fetch.db> --write-db <write.db>
         --kernel NAME --copy-bytes B
  python -m solver_cpp_amd.rocpd_summary l2 <results.db> [--kernel SUBSTR]

`traffic` turns one FETCH_SIZE run and one WRITE_SIZE run into HBM bytes per
launch of one kernel: each counter is scaled by what it reads for the plain
device copy of B bytes in the same run (the largest copyBuffer dispatch;
B/2 read and B/2 written), and the kernel's median (or, with --stat mean,
mean: for a kernel whose work varies from launch to launch) over its working
dispatches (FETCH_SIZE above 1 MiB, WRITE_SIZE above 64 KiB: gated
launches that returned at once are left out) is reported, with a final machine-readable line
`traffic_bytes_per_launch <kernel> <bytes>`.

`trace` lists kernels by total time (calls, total and mean duration in us;
dispatches faster than --active-us are also reported separately: those are
launches whose device-side stop flag made them return at once).  `pmc`
lists the median and largest counter value per dispatch for each kernel
(FETCH_SIZE / WRITE_SIZE in KiB, other counters as counts).  `l2` takes a
run with TCC_HIT_sum, TCC_MISS_sum and TCC_REQ_sum and lists, per kernel,
the L2 hits, misses and requests per working dispatch and the hit rate
hits / (hits + misses).
"""
import argparse
import collections
import sqlite3
import statistics


def short(name):
    return name.split("(")[0].replace("void ", "")


def trace(db, top=20, active_us=10.0, match=None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, duration from kernels").fetchall()
    agg = collections.defaultdict(list)
    for name, dur in rows:
        agg[short(name)].append(dur / 1000.0)
    tot = sum(sum(v) for v in agg.values())
    out = [f"{'kernel':70s} {'calls':>6s} {'total_ms':>9s} {'mean_us':>8s} {'%':>5s} "
           f"{'active':>6s} {'active_mean_us':>14s}"]
    items = sorted(agg.items(), key=lambda kv: -sum(kv[1]))
    if match:
        items = [kv for kv in items if match in kv[0]]
    for k, v in items[:top]:
        act = [d for d in v if d >= active_us]
        am = statistics.mean(act) if act else 0.0
        out.append(f"{k[:70]:70s} {len(v):6d} {sum(v) / 1e3:9.2f} {statistics.mean(v):8.2f} "
                   f"{100 * sum(v) / tot:5.1f} {len(act):6d} {am:14.2f}")
    out.append(f"total kernel time {tot / 1e3:.2f} ms over {len(rows)} dispatches")
    return "\n".join(out)


def pmc(db, match=None):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, value from counters_collection").fetchall()
    agg = collections.defaultdict(list)
    for k, c, v in rows:
        agg[(short(k), c)].append(v)
    out = [f"{'kernel':70s} {'counter':12s} {'dispatches':>10s} {'median':>12s} {'max':>12s}"]
    for (k, c), vs in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        if match and match not in k:
            continue
        out.append(f"{k[:70]:70s} {c:12s} {len(vs):10d} {statistics.median(vs):12.1f} {max(vs):12.1f}")
    return "\n".join(out)


def l2(db, match=None, min_req=1000.0):
    """Per kernel: TCC hits, misses and requests summed over its dispatches,
    divided by its working dispatches (those with at least min_req requests:
    a gated launch that returned at once adds next to nothing to the sums),
    and the hit rate hits / (hits + misses)."""
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, value from counters_collection").fetchall()
    tot = collections.defaultdict(float)
    work = collections.Counter()
    for k, c, v in rows:
        k = short(k)
        tot[(k, c)] += v
        if c == "TCC_REQ_sum" and v >= min_req:
            work[k] += 1
    out = [f"{'kernel':70s} {'dispatches':>10s} {'hit_M':>9s} {'miss_M':>9s} {'req_M':>9s} {'hit_rate':>8s}"]
    for k, nd in sorted(work.items(), key=lambda kv: -tot[(kv[0], "TCC_REQ_sum")]):
        if match and match not in k:
            continue
        h, m, q = (tot[(k, c)] / nd for c in ("TCC_HIT_sum", "TCC_MISS_sum", "TCC_REQ_sum"))
        out.append(f"{k[:70]:70s} {nd:10d} {h / 1e6:9.2f} {m / 1e6:9.2f} {q / 1e6:9.2f} "
                   f"{h / max(h + m, 1.0):8.3f}")
    return "\n".join(out)


def _counter(db, counter):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, value from counters_collection").fetchall()
    return [(short(k), v) for k, c, v in rows if c == counter]


def traffic(fetch_db, write_db, kernel, copy_bytes, stat="median"):
    fr, wr = _counter(fetch_db, "FETCH_SIZE"), _counter(write_db, "WRITE_SIZE")
    cf = max(v for k, v in fr if "copyBuffer" in k)
    cw = max(v for k, v in wr if "copyBuffer" in k)
    f_factor = (copy_bytes / 2) / (cf * 1024.0)
    w_factor = (copy_bytes / 2) / (cw * 1024.0)
    kf = [v for k, v in fr if k == kernel and v > 1024]
    kw = [v for k, v in wr if k == kernel and v > 64]
    fn = statistics.mean if stat == "mean" else statistics.median
    mf, mw = fn(kf), fn(kw)
    t = mf * 1024.0 * f_factor + mw * 1024.0 * w_factor
    out = [f"# calibration: copy of {copy_bytes} bytes: FETCH_SIZE {cf:.0f} KiB (factor {f_factor:.3f}), "
           f"WRITE_SIZE {cw:.0f} KiB (factor {w_factor:.3f})",
           f"# {kernel}: {len(kf)} working dispatches, {stat} FETCH_SIZE {mf:.1f} KiB, WRITE_SIZE {mw:.1f} KiB",
           f"# HBM bytes per launch = {mf:.1f} KiB x {f_factor:.3f} + {mw:.1f} KiB x {w_factor:.3f} "
           f"= {t / 1e6:.1f} MB",
           f"traffic_bytes_per_launch {kernel.replace(' ', '')} {t:.0f}"]
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "pmc", "traffic", "l2"])
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=20)
    ap.add_argument("--active-us", type=float, default=10.0)
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--write-db", default=None)
    ap.add_argument("--copy-bytes", type=float, default=0)
    ap.add_argument("--stat", choices=["median", "mean"], default="median")
    a = ap.parse_args()
    if a.mode == "trace":
        print(trace(a.db, a.top, a.active_us, a.kernel))
    elif a.mode == "pmc":
        print(pmc(a.db, a.kernel))
    elif a.mode == "l2":
        print(l2(a.db, a.kernel))
    else:
        print(traffic(a.db, a.write_db, a.kernel, a.copy_bytes, a.stat))


if __name__ == "__main__":
    main()
