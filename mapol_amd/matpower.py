"""MATPOWER v2 case reader with the reference loader's unit conventions.

Follows ``loadMatpowerFormat`` (src/files/loadMatpowerFormat.h:14-209):

* bus rows: ``id type Pd Qd Gs Bs area Vm Va baseKV zone Vmax Vmin``;
  loads, shunts and generation are divided by ``baseMVA``
  (loadMatpowerFormat.h:86-90); ``Va`` is converted to radians (:88);
* bus ids are remapped to 0-based positions in file order (:91,119-124,
  186-198);
* gen rows: ``bus Pg Qg Qmax Qmin Vg mBase status Pmax Pmin``; the reference
  adds every generator's ``Pg``/``Qg`` to its bus whatever the status column
  says (:125-128) and overwrites the bus ``VG`` (:139-140).  ``Inf`` limits
  become +-inf;
* branch rows: ``f t r x b rateA rateB rateC ratio angle status``
  (:173-183); rateA (MVA) is kept, divided by ``baseMVA``, for flow checks.

Loads and generation stay separate in :class:`mapol_amd.synth.Case` (the
net injection the reference stores is ``sum(Pg) - Pd``), and each
generator's status is kept so the reference's "all generators count"
behaviour is the default (``honour_gen_status=False``).
"""
from __future__ import annotations

import re

import numpy as np

from .synth import Case


def _block(text: str, name: str):
    m = re.search(r"mpc\." + name + r"\s*=\s*\[(.*?)\];", text, re.S)
    if not m:
        return np.zeros((0, 0))
    rows = []
    for line in m.group(1).splitlines():
        line = line.split("%", 1)[0].strip().rstrip(";").strip()
        if not line:
            continue
        vals = []
        for tok in line.replace(";", " ").split():
            t = tok.strip()
            if t in ("Inf", "inf", "+Inf"):
                vals.append(np.inf)
            elif t in ("-Inf", "-inf"):
                vals.append(-np.inf)
            else:
                vals.append(float(t))
        rows.append(vals)
    width = max(len(r) for r in rows)
    return np.array([r + [0.0] * (width - len(r)) for r in rows])


def parse(text: str, name: str = "", honour_gen_status: bool = False) -> Case:
    mv = re.search(r"mpc\.version\s*=\s*'([^']*)'", text)
    if mv and mv.group(1) != "2":
        raise ValueError("MATPOWER case input: old version")
    base = float(re.search(r"mpc\.baseMVA\s*=\s*([-+0-9.eE]+)", text).group(1))
    bus = _block(text, "bus")
    gen = _block(text, "gen")
    br = _block(text, "branch")
    ids = bus[:, 0].astype(np.int64)
    pos = {int(b): i for i, b in enumerate(ids)}
    n = len(ids)
    gen_bus = np.array([pos[int(b)] for b in gen[:, 0]], dtype=np.int32)
    gstat = (gen[:, 7] > 0).astype(np.uint8)
    if not honour_gen_status:
        gstat = np.ones(len(gen_bus), dtype=np.uint8)
    VGbus = np.ones(n)
    for k, b in enumerate(gen_bus):
        VGbus[b] = gen[k, 5]
    f = np.array([pos[int(b)] for b in br[:, 0]], dtype=np.int32)
    t = np.array([pos[int(b)] for b in br[:, 1]], dtype=np.int32)
    btype = bus[:, 1].astype(np.int32)
    return Case(
        bus_type=btype,
        Pd=bus[:, 2] / base, Qd=bus[:, 3] / base,
        Gs=bus[:, 4] / base, Bs=bus[:, 5] / base,
        Vm=bus[:, 7].copy(), Va=bus[:, 8] * np.pi / 180.0,
        br_from=f, br_to=t, R=br[:, 2].copy(), X=br[:, 3].copy(),
        B=br[:, 4].copy(), tap=br[:, 8].copy(), shift=br[:, 9].copy(),
        br_status=(br[:, 10] > 0).astype(np.uint8),
        gen_bus=gen_bus, Pg=gen[:, 1] / base, Qg=gen[:, 2] / base,
        Qmax=gen[:, 3] / base, Qmin=gen[:, 4] / base, VG=gen[:, 5].copy(),
        gen_status=gstat, baseMVA=base, name=name,
        rate=br[:, 5] / base)


def write(path: str, baseMVA: float, bus, gen, branch) -> None:
    """MATPOWER version-2 text of the three tables (the inverse of the
    reader: every value written in full precision, Inf as ``Inf``)."""
    def fmt(v):
        v = float(v)
        if np.isinf(v):
            return "Inf" if v > 0 else "-Inf"
        return repr(v) if v != int(v) or abs(v) >= 1e15 else str(int(v))

    lines = ["function mpc = case", "mpc.version = '2';",
             f"mpc.baseMVA = {float(baseMVA)!r};"]
    for name, tab in (("bus", bus), ("gen", gen), ("branch", branch)):
        lines.append(f"mpc.{name} = [")
        lines += ["\t" + "\t".join(fmt(v) for v in row) + ";" for row in tab]
        lines.append("];")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def load(path: str, honour_gen_status: bool = False) -> Case:
    with open(path) as fh:
        text = fh.read()
    import os
    return parse(text, os.path.splitext(os.path.basename(path))[0],
                 honour_gen_status)
