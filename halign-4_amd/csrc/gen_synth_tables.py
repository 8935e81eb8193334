#!/usr/bin/env python3
"""Writes synth_tables.h: the 4,096-entry length quantile table of the uneven
batch workload (lognormal with median 1,500 bp and sigma 0.8, clamped to
[200, 20,000] bp). Entry i is the (i + 0.5) / 4096 quantile, rounded to the
nearest integer; a query's length is the entry picked by the top 12 bits of a
generator draw. The table is committed, so every machine draws the same
lengths whatever its libm. Uses only the Python standard library.

Run from this directory:  python gen_synth_tables.py > synth_tables.h
"""
import math
from statistics import NormalDist

N, MEDIAN, SIGMA, LO, HI = 4096, 1500.0, 0.8, 200.0, 20000.0


def table():
    nd = NormalDist()
    out = []
    for i in range(N):
        x = MEDIAN * math.exp(SIGMA * nd.inv_cdf((i + 0.5) / N))
        out.append(int(min(max(x, LO), HI) + 0.5))
    return out


def main():
    t = table()
    print("/* synth_tables.h -- written by gen_synth_tables.py; do not edit. */")
    print("#ifndef SYNTH_TABLES_H")
    print("#define SYNTH_TABLES_H")
    print("/* lognormal(median 1500, sigma 0.8) quantiles (i + 0.5) / 4096, clamped to [200, 20000] */")
    print(f"static const int kLognormalLen[{N}] = {{")
    for k in range(0, N, 12):
        print("    " + ", ".join(str(v) for v in t[k:k + 12]) + ",")
    print("};")
    print("#endif")


if __name__ == "__main__":
    main()
