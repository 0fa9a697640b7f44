"""Write tests/golden/matpower_cases.npz: the numeric tables of the MATPOWER
case files the reference ships (src/datasets/*.m), as data.

Run in the container that has the reference sources:

    python tests/golden/make_matpower_fixtures.py

Per case: ``baseMVA``, ``bus`` (13 columns: id type Pd Qd Gs Bs area Vm Va
baseKV zone Vmax Vmin), ``gen`` (10 columns: bus Pg Qg Qmax Qmin Vg mBase
status Pmax Pmin) and ``branch`` (11 columns: f t r x b rateA rateB rateC
ratio angle status), in file order, with +-Inf kept.  The tests write
MATPOWER text back from these tables to exercise the readers.  The tables
are read with a tokenizer of this script's own (not the readers under test).
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/datasets"
CASES = ["case14", "case30", "case57", "case118", "case300", "case1354pegase",
         "case2869pegase", "case9241pegase"]
WIDTH = {"bus": 13, "gen": 10, "branch": 11}


def table(text, name):
    m = re.search(r"mpc\." + name + r"\s*=\s*\[(.*?)\]\s*;", text, re.S)
    body = re.sub(r"%[^\n]*", "", m.group(1))
    rows = []
    for piece in re.split(r"[;\n]", body):
        toks = piece.split()
        if toks:
            rows.append([float(t.replace("Inf", "inf")) for t in toks[:WIDTH[name]]])
    return np.array(rows)


def main():
    out = {}
    for name in CASES:
        text = open(os.path.join(SRC, name + ".m")).read()
        base = float(re.search(r"mpc\.baseMVA\s*=\s*([-+0-9.eE]+)", text).group(1))
        out[f"{name}/baseMVA"] = np.array(base)
        for blk in ("bus", "gen", "branch"):
            out[f"{name}/{blk}"] = table(text, blk)
    path = os.path.join(HERE, "matpower_cases.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    sys.exit(main())
