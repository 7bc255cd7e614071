"""Pin the CPU FP64 oracle adjoint (oracle.c: oracle_far_adjoint) against
central finite differences of the (separately pinned) oracle forward.

The loss is L = sum_g W_g (S_g - T_g)^2, so dL/dS = 2 W (S - T).  Finite
differences are taken on one particle's own spectrum (the other particles'
contribution is held fixed), so the large loss value does not cancel, and
Richardson-extrapolated over a ladder of step sizes; the estimate whose
neighbours agree best is used.  The grid keeps |dphase| < pi everywhere, so
the filter mask cannot change under the perturbation.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from adjoint_cases import FD, PROBES, adjoint_case
from oracle import oracle

@pytest.mark.parametrize("comp", ["total", "cartesian"])
def test_oracle_adjoint_matches_finite_differences(comp):
    tr, axes, dt, S0, T, W, G = adjoint_case(comp)
    grads, gw = oracle.far_adjoint(comp, tr, axes, dt, G)
    fd = FD(comp, tr, axes, dt, S0, T, W)
    worst = 0.0
    for (p, c, s) in PROBES:
        est = fd.best(p, c, s)
        err = abs(est - grads[p][c][s]) / abs(grads[p][c][s])
        worst = max(worst, err)
        assert err < 1e-6, (p, c, s, est, grads[p][c][s])
    for p in range(len(tr)):
        est = fd.best(p, 6, 0)
        assert abs(est - gw[p]) / abs(gw[p]) < 1e-6
    print(f"oracle adjoint vs FD ({comp}): worst relative error {worst:.2e}")


def test_oracle_adjoint_last_sample_has_no_position_gradient():
    tr, axes, dt, S0, T, W, G = adjoint_case("total")
    grads, _ = oracle.far_adjoint("total", tr, axes, dt, G)
    for g in grads:
        assert g[0][-1] == 0.0 and g[1][-1] == 0.0 and g[2][-1] == 0.0   # x of the last sample is unused
        assert g[3][-1] != 0.0                                            # its u closes the last interval


ADJ_MUTANTS = {
    "adj_no_phase_term": ("gph -= Hc[c] * Av[c] * qim;", "gph -= 0.0 * qim;"),
    "adj_gm_sign": ("const v3 gm = sub3(sub3(scl3(2.0 * c1 * c2 * nmg, nv), scl3(c1, gAv)), scl3(c2 * c2 * ag, nv));",
                    "const v3 gm = add3(sub3(scl3(2.0 * c1 * c2 * nmg, nv), scl3(c1, gAv)), scl3(c2 * c2 * ag, nv));"),
    "adj_dt_sign": ("ab_[3 * s + 0] += 0.5 * gm.x - dtInv * ga.x;", "ab_[3 * s + 0] += 0.5 * gm.x + dtInv * ga.x;"),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name", sorted(ADJ_MUTANTS))
def test_planted_adjoint_mistakes_are_caught(name, tmp_path):
    src = open(os.path.join(os.path.dirname(oracle.__file__), "oracle.c")).read()
    old, new = ADJ_MUTANTS[name]
    assert src.count(old) == 1
    mut = tmp_path / "oracle_mut.c"
    mut.write_text(src.replace(old, new))
    so = tmp_path / "liboracle_mut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-o", str(so),
                    str(mut), "-lm"], check=True)
    L = oracle.lib(str(so))
    tr, axes, dt, S0, T, W, G = adjoint_case("total")
    grads, gw = oracle.far_adjoint("total", tr, axes, dt, G, library=L)
    fd = FD("total", tr, axes, dt, S0, T, W)   # the pinned forward
    bad = 0
    for (p, c, s) in PROBES:
        est = fd.best(p, c, s)
        if not abs(est - grads[p][c][s]) <= 1e-6 * abs(est):
            bad += 1
    assert bad > 0, f"mutant {name} passed every finite-difference probe"
