"""The loss-minimisation driver (include/mapol_pso.hpp, the mirror of the
reference's pso::run, src/pso/run.h:58-168 and Particula.h:88-187), compiled
as a C++ caller of the C-ABI.

* Swarm mechanics without a GPU, on a known fitness (a sum of squares around
  a target on the controls' grids): every evaluated position within its
  bounds, OLTC and SHC controls on their steps, personal bests the best
  positions seen, recorded global bests strictly decreasing and the target
  approached.
* Loss minimisation of MATPOWER case14 (the reference's default case) on the
  GPU with the reference's control vector (AVR per PV bus, SHC per shunt
  bus, OLTC per tapped branch;
  main.cu:232-264): the oracle, solving the case with the controls of the
  defaults and of the final global best, reproduces the losses the driver
  reports for its final global best and for a probe away from the
  defaults (the defaults are MATPOWER's solved point; whether a short run
  beats them depends on the draws).
The swarm draws come from rand() seeded by the caller (the reference seeds
with time(0), so no trajectory of the reference can be reproduced; the
driver's steps follow run.h/Particula.h and are checked here by their
invariants)."""
import os
import subprocess

import numpy as np
import pytest

from mapol_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
AVR, OLTC, SHC = 1, 2, 3

pytestmark = pytest.mark.skipif(not os.path.exists(engine.LIB_PATH),
                                reason="libbatchpf.so not built")


def _build(tmp_path):
    exe = str(tmp_path / "pso_demo")
    lib = os.path.dirname(engine.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "pso_demo.cpp"), "-L", lib,
           "-lbatchpf", f"-Wl,-rpath,{lib}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def _bests(lines):
    return [(int(l.split()[2]), float(l.split()[-1])) for l in lines
            if l.startswith("best it")]


def test_swarm_mechanics_on_a_known_fitness(tmp_path):
    out = subprocess.run([_build(tmp_path)], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    bests = _bests(out)
    assert bests[0][0] == 0
    its = [b[0] for b in bests]
    fits = [b[1] for b in bests]
    assert its == sorted(its) and len(set(its)) == len(its)
    assert all(b < a for a, b in zip(fits, fits[1:]))
    stats = dict(zip(*[iter(next(l for l in out if l.startswith("evaluations")).split())] * 2))
    assert int(stats["evaluations"]) == 24 * 61
    assert int(stats["out_of_bounds"]) == 0
    assert int(stats["off_grid"]) == 0
    assert int(stats["stale_best"]) == 0
    final = np.array([float(v) for v in next(l for l in out if l.startswith("final")).split()[1:]])
    target = np.array([1.03, 0.97, 1.05, 0.93, 0.12])
    # the discrete controls reach their targets' steps, the rest comes close
    np.testing.assert_allclose(final[2:], target[2:], atol=1e-12)
    assert fits[-1] < 1e-4 and fits[-1] == pytest.approx(((final - target) ** 2).sum(), rel=1e-12)
    sol = next(l for l in out if l.startswith("solution")).split(" ", 1)[1]
    assert sol.startswith("Solucao:") and sol.count(";") == 3 + 5


@pytest.mark.gpu
def test_loss_minimisation_matches_oracle(tmp_path):
    import oracle
    from mapol_amd import matpower
    from tests.test_matpower import write_case
    print("[pso] case14, 32 particles, 12 iterations", flush=True)
    z = np.load(os.path.join(ROOT, "tests", "golden", "matpower_cases.npz"))
    path = str(tmp_path / "case14.m")
    write_case(path, float(z["case14/baseMVA"]), z["case14/bus"],
               z["case14/gen"], z["case14/branch"])
    res = subprocess.run([_build(tmp_path), path, "7", "32", "12"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    out = res.stdout.splitlines()
    default = float(next(l for l in out if l.startswith("default loss")).split()[-1])
    bests = _bests(out)
    fits = [b[1] for b in bests]
    assert all(b < a for a, b in zip(fits, fits[1:]))
    assert fits[0] <= default          # particle 0 starts at the defaults
    ctl = [(int(l.split()[1]), int(l.split()[2]), float(l.split()[3]), float(l.split()[4]))
           for l in out if l.startswith("control")]
    assert {k for k, _, _, _ in ctl} == {AVR, SHC, OLTC}
    for k, i, v, dflt in ctl:
        if v == dflt:
            continue                   # a default may lie outside the range
        lo, hi, q = {AVR: (0.9, 1.1, 0), OLTC: (0.9, 1.1, 20), SHC: (0.0, 0.18, 3)}[k]
        assert lo - 1e-12 <= v <= hi + 1e-12   # a step can round past max
        if q:
            steps = (v - lo) * q / (hi - lo) if k == OLTC else v * q / hi
            assert abs(steps - round(steps)) < 1e-9
    case = matpower.load(path)
    orc = oracle.Oracle(case)

    def oracle_loss(controls):
        vset = np.full(case.nbus, np.nan)
        bsh = np.full(case.nbus, np.nan)
        tap = np.full(case.nbranch, np.nan)
        for k, i, v in controls:
            {AVR: vset, SHC: bsh, OLTC: tap}[k][i] = v
        r = orc.solve(vset=vset, tap=tap, bsh=bsh)
        assert r["converged"]
        return r["loss"]

    r0 = orc.solve()
    assert r0["converged"] and default == pytest.approx(r0["loss"], rel=1e-9)
    # the reported best, and a probe away from the defaults
    assert fits[-1] == pytest.approx(oracle_loss([c[:3] for c in ctl]), rel=1e-9)
    probe = float(next(l for l in out if l.startswith("probe loss")).split()[-1])
    pc = [(int(l.split()[1]), int(l.split()[2]), float(l.split()[3]))
          for l in out if l.startswith("probe ") and not l.startswith("probe loss")]
    assert probe != pytest.approx(default, rel=1e-6)
    assert probe == pytest.approx(oracle_loss(pc), rel=1e-9)
