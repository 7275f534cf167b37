"""Summaries of rocprofv3 result databases (rocpd SQLite) for profiles/.

    python -m twisty_amd.rocprof_summary kernels DB [--skip-first N] [--match SUBSTR]
    python -m twisty_amd.rocprof_summary traffic --fetch DB --write DB \
        --calib-fetch DB --calib-write DB --calib-bytes BYTES --match SUBSTR [--json OUT]

`kernels` prints per-kernel dispatch count and mean/min/max duration.
`counters` prints the mean per dispatch of every counter collected for the
matching kernel (optionally also as JSON), plus VALU and LDS utilisation derived from them
(SQ_INSTS_VALU wave instructions x 4 cycles / (GRBM_GUI_ACTIVE per XCD x
SIMDs), SQ_LDS_BANK_CONFLICT / SQ_INSTS_LDS).
`traffic` prints FETCH_SIZE and WRITE_SIZE per dispatch of the matching kernel,
scaled into bytes by the counters a plain copy of known size produced
(copy kernel from twisty_copy_calib), and optionally writes the per-launch HBM
bytes to a JSON file read by bench.py.
"""
import argparse
import json
import sqlite3
import statistics
import sys


def _kernel_rows(db):
    con = sqlite3.connect(db)
    return con.execute("select name, duration from kernels order by start").fetchall()


def kernels(db, skip_first=0, match=None, first=0):
    """Per-kernel dispatch statistics, in dispatch order: the first skip_first
    dispatches of each kernel are left out and, with first > 0, only the next
    first dispatches are kept."""
    rows = _kernel_rows(db)
    by = {}
    for name, dur in rows:
        by.setdefault(name, []).append(dur)
    lines = ["| kernel | dispatches | mean ms | min ms | max ms |", "|---|---|---|---|---|"]
    out = {}
    for name, durs in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        if match and match not in name:
            continue
        d = durs[skip_first:] if len(durs) > skip_first else durs
        if first > 0:
            d = d[:first]
        ms = [x / 1e6 for x in d]
        out[name] = {"dispatches": len(ms), "mean_ms": statistics.fmean(ms), "min_ms": min(ms), "max_ms": max(ms)}
        lines.append(f"| `{name}` | {len(ms)} | {out[name]['mean_ms']:.4f} | {min(ms):.4f} | {max(ms):.4f} |")
    return "\n".join(lines), out


def _counter(db, counter, match):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, value from counters_collection where counter_name = ? order by id",
                       (counter,)).fetchall()
    return [v for k, v in rows if match in k]


def traffic(fetch_db, write_db, calib_fetch_db, calib_write_db, calib_bytes, match, calib_match="copy_kernel"):
    f = _counter(fetch_db, "FETCH_SIZE", match)
    w = _counter(write_db, "WRITE_SIZE", match)
    cf = _counter(calib_fetch_db, "FETCH_SIZE", calib_match)
    cw = _counter(calib_write_db, "WRITE_SIZE", calib_match)
    kib = calib_bytes / 1024.0
    sf = kib / statistics.fmean(cf)  # real bytes per counted KiB, / 1024
    sw = kib / statistics.fmean(cw)
    fetch_b = statistics.fmean(f) * 1024.0 * sf
    write_b = statistics.fmean(w) * 1024.0 * sw
    res = {"kernel_match": match, "dispatches_fetch": len(f), "dispatches_write": len(w),
           "fetch_kib_per_dispatch": statistics.fmean(f), "write_kib_per_dispatch": statistics.fmean(w),
           "calib_bytes": calib_bytes, "calib_fetch_kib": statistics.fmean(cf), "calib_write_kib": statistics.fmean(cw),
           "fetch_scale": sf, "write_scale": sw, "fetch_bytes_per_launch": fetch_b,
           "write_bytes_per_launch": write_b, "bytes_per_launch": fetch_b + write_b}
    text = "\n".join([
        f"| quantity | value |", "|---|---|",
        f"| FETCH_SIZE per dispatch (KiB, counted) | {res['fetch_kib_per_dispatch']:.3f} |",
        f"| WRITE_SIZE per dispatch (KiB, counted) | {res['write_kib_per_dispatch']:.3f} |",
        f"| copy of {calib_bytes} B: FETCH_SIZE / WRITE_SIZE (KiB) | {res['calib_fetch_kib']:.1f} / {res['calib_write_kib']:.1f} |",
        f"| scale: real bytes per counted byte, fetch / write | {sf:.4f} / {sw:.4f} |",
        f"| HBM bytes read per launch | {fetch_b:.0f} |",
        f"| HBM bytes written per launch | {write_b:.0f} |",
        f"| HBM bytes per launch | {fetch_b + write_b:.0f} |",
    ])
    return text, res


def counters(db, match, simds=1024, xcds=8):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, value from counters_collection").fetchall()
    by = {}
    for k, c, v in rows:
        if match in k:
            by.setdefault(c, []).append(v)
    mean = {c: statistics.fmean(v) for c, v in by.items()}
    lines = ["| counter | mean per dispatch |", "|---|---|"]
    lines += [f"| {c} | {v:.4g} |" for c, v in sorted(mean.items())]
    derived = {}
    if "SQ_INSTS_VALU" in mean and "GRBM_GUI_ACTIVE" in mean:
        cycles = mean["GRBM_GUI_ACTIVE"] / xcds  # GRBM_GUI_ACTIVE sums over the XCDs
        derived["kernel_cycles"] = cycles
        derived["valu_issue_utilisation"] = mean["SQ_INSTS_VALU"] * 4.0 / (cycles * simds)
    if "SQ_LDS_BANK_CONFLICT" in mean and "SQ_INSTS_LDS" in mean:
        derived["lds_bank_conflicts_per_lds_instruction"] = mean["SQ_LDS_BANK_CONFLICT"] / mean["SQ_INSTS_LDS"]
    lines += [f"| derived: {k} | {v:.4g} |" for k, v in derived.items()]
    return "\n".join(lines), {**mean, **derived}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--skip-first", type=int, default=0)
    k.add_argument("--match")
    k.add_argument("--first", type=int, default=0)
    t = sub.add_parser("traffic")
    for n in ("fetch", "write", "calib-fetch", "calib-write"):
        t.add_argument(f"--{n}", required=True)
    t.add_argument("--calib-bytes", type=int, required=True)
    t.add_argument("--match", required=True)
    t.add_argument("--json")
    c = sub.add_parser("counters")
    c.add_argument("db")
    c.add_argument("--match", required=True)
    c.add_argument("--json")
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        text, _ = kernels(a.db, a.skip_first, a.match, a.first)
        print(text)
    elif a.cmd == "counters":
        text, res = counters(a.db, a.match)
        print(text)
        if a.json:
            with open(a.json, "w") as f:
                json.dump({"kernel_match": a.match, **res}, f, indent=1)
    else:
        text, res = traffic(a.fetch, a.write, a.calib_fetch, a.calib_write, a.calib_bytes, a.match)
        print(text)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
