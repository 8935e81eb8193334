"""Measure the DP planner's cost table on this GPU and write it as the
library's built-in table.

    python -m halign.tune [--header PATH] [--summary PATH]

Runs halign_tune (every ring shape: equal-length synthetic jobs filling the
chip, two bands and two lengths per shape; every query-fixed shape: unbanded
jobs whose query fills 1..S slots, two target lengths) on device 0, prints the table and
writes it as csrc/dp_tune_table.h (rebuild the library afterwards). Run it
again after any change to the DP kernels: the table is what the automatic
kernel policy plans with.
"""
import argparse
import os
import sys
import time

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_HEADER = os.path.join(HERE, "..", "csrc", "dp_tune_table.h")


def header_text(rows, device_name):
    lines = [
        "// dp_tune_table.h -- planner cost table measured by halign.tune (halign_tune",
        f"// over every ring and query-fixed shape) on {device_name}. Fields: waves",
        "// (-1: query-fixed), lanes, ring (query-fixed: blocks), form (1: uniform",
        "// tickets, 2: two workgroups per SIMD), jobs in flight, bands measured, the",
        "// bands (16-cell columns; query-fixed: query blocks), ns per anti-diagonal",
        "// of one job at each band, ns per job.",
        "static const halign_tune_entry kTuneDefault[] = {",
    ]
    for r in rows:
        n = len(r["ncol"])
        cols = list(r["ncol"]) + [0] * (4 - n)
        ns = list(r["ns_diag"]) + [0.0] * (4 - n)
        lines.append("    {%d, %d, %d, %d, %d, %d, {%s}, {%s}, %.1f}," % (
            r["waves"], r["lanes"], r["ring"], r["uniform"], r["jobs_in_flight"], n,
            ", ".join(str(c) for c in cols), ", ".join("%.2f" % x for x in ns), r["ns_job"]))
    lines.append("};")
    return "\n".join(lines) + "\n"


def shape_name(r):
    s = r["ring"] // r["lanes"]
    if r["waves"] < 0:
        return f"q{r['lanes']}x{s}"
    if r["waves"]:
        return f"{r['waves']}x{s}{'d' if r['uniform'] == 2 else ''}"
    return f"sub{r['lanes']}x{s}{'u' if r['uniform'] else ''}"


def table_text(rows):
    out = [f"{'shape':>10s} {'ring':>5s} {'in flight':>9s}  ns per anti-diagonal at n_col (band)"
           f"{'':>24s} {'ns/job':>9s}"]
    for r in rows:
        pts = "  ".join(f"{c:4d}: {t:7.1f}" for c, t in zip(r["ncol"], r["ns_diag"]))
        out.append(f"{shape_name(r):>10s} {r['ring']:5d} {r['jobs_in_flight']:9d}  {pts:<60s} {r['ns_job']:9.0f}")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--header", default=DEFAULT_HEADER, help="where to write the C table ('' = do not write)")
    ap.add_argument("--summary", default="", help="also write the printed table to this file")
    args = ap.parse_args(argv)
    ctx = _lib.Context(0)
    t0 = time.perf_counter()
    ctx.tune()
    dt = time.perf_counter() - t0
    rows = ctx.tune_table()
    ctx.close()
    text = table_text(rows)
    print(text)
    print(f"measured in {dt:.1f} s", file=sys.stderr)
    if args.header:
        with open(args.header, "w") as f:
            f.write(header_text(rows, "an MI355X (gfx950)"))
    if args.summary:
        with open(args.summary, "w") as f:
            f.write(text + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
