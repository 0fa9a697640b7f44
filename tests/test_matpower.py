"""MATPOWER case ingestion through the C-ABI (``bpf_load_matpower``, host only).

The case tables of the eight MATPOWER files the reference ships
(src/datasets/*.m, stored as data in tests/golden/matpower_cases.npz) are
written back as MATPOWER text, with the formatting variations such files
have (comments, tabs, rows ending in ';' or not, Inf limits), and read by:

* the C-ABI reader, checked field by field against the reference loader's
  conventions (src/files/loadMatpowerFormat.h:14-209) applied to the tables
  directly in numpy;
* the Python reader (mapol_amd/matpower.py), which must agree exactly.
"""
import os

import numpy as np
import pytest

from mapol_amd import engine, matpower

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["case14", "case30", "case57", "case118", "case300", "case1354pegase",
         "case2869pegase", "case9241pegase"]

pytestmark = pytest.mark.skipif(not os.path.exists(engine.LIB_PATH),
                                reason="libbatchpf.so not built")


@pytest.fixture(scope="module")
def tables():
    return np.load(os.path.join(HERE, "golden", "matpower_cases.npz"))


def _fmt(v):
    if np.isinf(v):
        return "Inf" if v > 0 else "-Inf"
    return repr(float(v)) if v != int(v) else str(int(v))


def write_case(path, base, bus, gen, branch, style=0):
    """MATPOWER v2 text; ``style`` varies separators and comments"""
    sep = "\t" if style % 2 else " "
    end = ";" if style % 3 != 2 else ""
    lines = ["function mpc = synthcase", "%% written by the tests",
             "mpc.version = '2';", "", "%% system MVA base",
             f"mpc.baseMVA = {base:g};", ""]
    for name, tab in (("bus", bus), ("gen", gen), ("branch", branch)):
        lines.append(f"%% {name} data")
        lines.append(f"mpc.{name} = [")
        for k, row in enumerate(tab):
            txt = sep + sep.join(_fmt(v) for v in row) + end
            if style and k % 7 == 3:
                txt += "\t% trailing comment"
            lines.append(txt)
            if style and k % 11 == 5:
                lines.append("\t% a comment line inside the block")
        lines.append("];")
        lines.append("")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def expected(base, bus, gen, branch, honour):
    """the reference loader's conventions, from the tables"""
    pos = {int(b): i for i, b in enumerate(bus[:, 0])}
    vg = gen[:, 5]
    return dict(
        bus_type=bus[:, 1].astype(np.int32), Pd=bus[:, 2] / base,
        Qd=bus[:, 3] / base, Gs=bus[:, 4] / base, Bs=bus[:, 5] / base,
        Vm=bus[:, 7], Va=bus[:, 8] * np.pi / 180,
        br_from=np.array([pos[int(b)] for b in branch[:, 0]], np.int32),
        br_to=np.array([pos[int(b)] for b in branch[:, 1]], np.int32),
        R=branch[:, 2], X=branch[:, 3], B=branch[:, 4], tap=branch[:, 8],
        shift=branch[:, 9], br_status=(branch[:, 10] != 0).astype(np.uint8),
        rate=branch[:, 5] / base,
        gen_bus=np.array([pos[int(b)] for b in gen[:, 0]], np.int32),
        Pg=gen[:, 1] / base, Qg=gen[:, 2] / base, Qmax=gen[:, 3] / base,
        Qmin=gen[:, 4] / base, VG=vg,
        gen_status=((gen[:, 7] > 0) if honour else np.ones(len(gen), bool))
        .astype(np.uint8))


FIELDS = ["bus_type", "Pd", "Qd", "Gs", "Bs", "Vm", "Va", "br_from", "br_to",
          "R", "X", "B", "tap", "shift", "br_status", "rate", "gen_bus", "Pg",
          "Qg", "Qmax", "Qmin", "VG", "gen_status"]


@pytest.mark.parametrize("name", CASES)
def test_reader_follows_the_reference_conventions(tables, tmp_path, name):
    base = float(tables[f"{name}/baseMVA"])
    bus, gen, branch = (tables[f"{name}/{b}"] for b in ("bus", "gen", "branch"))
    for style in (0, 1, 2):
        path = str(tmp_path / f"{name}_{style}.m")
        write_case(path, base, bus, gen, branch, style)
        for honour in (False, True):
            case, rate = engine.load_matpower(path, honour_gen_status=honour)
            want = expected(base, bus, gen, branch, honour)
            assert (case.nbus, case.nbranch, case.ngen) == (len(bus), len(branch), len(gen))
            assert case.baseMVA == base
            for f in FIELDS:
                got = rate if f == "rate" else getattr(case, f)
                np.testing.assert_array_equal(got, want[f], err_msg=f"{name} {f}")
            # the Python reader agrees exactly
            py = matpower.load(path, honour_gen_status=honour)
            for f in FIELDS:
                np.testing.assert_array_equal(getattr(py, f), getattr(case, f),
                                              err_msg=f"python {name} {f}")


def test_reader_errors(tables, tmp_path):
    base = float(tables["case14/baseMVA"])
    bus, gen, branch = (tables[f"case14/{b}"].copy() for b in ("bus", "gen", "branch"))
    bad = branch.copy()
    bad[3, 1] = 999
    p = str(tmp_path / "unknown_bus.m")
    write_case(p, base, bus, gen, bad)
    with pytest.raises(engine.EngineError, match="unknown bus"):
        engine.load_matpower(p)
    p = str(tmp_path / "old.m")
    write_case(p, base, bus, gen, branch)
    text = open(p).read().replace("mpc.version = '2';", "mpc.version = '1';")
    open(p, "w").write(text)
    with pytest.raises(engine.EngineError, match="old version"):
        engine.load_matpower(p)
    with pytest.raises(engine.EngineError, match="cannot open"):
        engine.load_matpower(str(tmp_path / "missing.m"))
    p = str(tmp_path / "nobase.m")
    write_case(p, base, bus, gen, branch)
    open(p, "w").write(open(p).read().replace("mpc.baseMVA", "mpc.baseXXX"))
    with pytest.raises(engine.EngineError, match="baseMVA"):
        engine.load_matpower(p)
