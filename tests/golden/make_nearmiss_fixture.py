#!/usr/bin/env python3
"""Write tests/golden/near_misses.json from data the reference ships.

Sources (read as text; nothing is executed):
  * exhaustive/evidence/*-compact.tar.gz, member */near-misses.json: the
    near-miss catalog of each archived run, one entry per primitive near miss
    with its six values, the non-square pair, whether the six values sum to a
    square ("square_total") and the number of Euler-square splits
    ("euler_squares"), as the reference's auditor wrote them;
  * exhaustive/analysis_1e12.txt: "primitive_near_misses" of every M: line
    with S < 1e12 (values, non-square pair, square_total).
Only the near-misses.json members are read from the archives.

The fixture keeps the distinct entries: "catalog" (with euler_squares, from the
archives) and "below_1e12" (from the analysis file).

Usage:  python3 tests/golden/make_nearmiss_fixture.py
"""
import glob
import json
import os
import tarfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/exhaustive"
OUT = os.path.join(HERE, "near_misses.json")


def key(entry):
    return tuple(entry["values"]), tuple(entry["nonsquare_pair"])


def main():
    catalog = {}
    sources = []
    for path in sorted(glob.glob(os.path.join(REF, "evidence", "*-compact.tar.gz"))):
        with tarfile.open(path, "r:gz") as tf:
            for m in tf.getmembers():
                if not (m.isfile() and m.name.endswith("/near-misses.json")):
                    continue
                entries = json.load(tf.extractfile(m))
                sources.append(f"exhaustive/evidence/{os.path.basename(path)}::{m.name} ({len(entries)} entries)")
                for e in entries:
                    rec = {"values": e["values"], "nonsquare_pair": e["nonsquare_pair"],
                           "square_total": e["square_total"], "euler_squares": e["euler_squares"]}
                    old = catalog.setdefault(key(rec), rec)
                    if old != rec:
                        raise SystemExit(f"conflicting catalog entries for {key(rec)}")
    with open(os.path.join(REF, "analysis_1e12.txt")) as f:
        ana = json.load(f)
    below = {}
    for e in ana["primitive_near_misses"]:
        rec = {"values": e["values"], "nonsquare_pair": e["nonsquare_pair"], "square_total": e["square_total"]}
        below.setdefault(key(rec), rec)
    out = {"generator": "tests/golden/make_nearmiss_fixture.py", "sources": sources + ["exhaustive/analysis_1e12.txt"],
           "analysis_1e12": {k: ana[k] for k in ("S_exclusive", "primitive_six_member_near_misses",
                                                  "near_misses_with_square_total", "positive_sextuples")},
           "catalog": sorted(catalog.values(), key=lambda r: r["values"]),
           "below_1e12": sorted(below.values(), key=lambda r: r["values"])}
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"{len(out['catalog'])} catalog entries, {len(out['below_1e12'])} below 1e12 -> {OUT}")


if __name__ == "__main__":
    main()
