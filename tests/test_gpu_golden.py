"""The HIP path through the C ABI against the committed golden fixtures
(tests/golden/), independent of running the oracle: per-subdomain Schur
complements, coarse bases and matrices, the S and M operators, the reduced
right-hand side and the solutions, with the FP64 tolerances of the parity
tests."""
import os
import sys

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import CubeProblem

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name,kind", [("cube2x2.npz", "poisson"), ("elast2x2.npz", "elasticity")])
def test_small_fixtures(name, kind):
    _say(f"case golden {name}")
    g = np.load(os.path.join(GOLD, name), allow_pickle=False)
    p = CubeProblem(2, 2, kind=kind, rhs="random", seed=7)
    b, (it, reason, cond) = bddcml_amd.solve_cube(p, tol=1e-12, maxit=300)
    assert reason == 0 and abs(it - int(g["iterations"])) <= 1
    assert np.array_equal(b.interface_dofs(), g["interface_dofs"])
    for s in range(p.nsub):
        isub = s + 1
        assert _rel(b.sub_matrix(isub, "S"), g[f"S_{isub}"]) < 1e-12
        gd, ind = b.sub_maps(isub)
        assert np.array_equal(gd, g[f"gdofs_{isub}"])
        pos = {int(c): k for k, c in enumerate(g[f"indrowc_{isub}"])}
        perm = np.array([pos[int(c)] for c in ind])
        phi, ac = b.sub_coarse(isub)
        assert _rel(phi, g[f"phi_{isub}"][:, perm]) < 1e-10
        assert _rel(ac, g[f"ac_{isub}"][np.ix_(perm, perm)]) < 1e-10
    assert _rel(b.apply_schur(g["x"]), g["Sx"]) < 1e-11
    assert _rel(b.apply_pc(g["r"]), g["Mr"]) < 1e-10
    assert _rel(b.reduced_rhs(), g["g"]) < 1e-12
    assert _rel(b.interface_solution(), g["u"]) < 1e-10
    b.finalize()


def test_poisson32_fixture():
    """The reference's poisson_on_cube case: 32^3 Q1, 4^3 subdomains, rhs h^3,
    PCG to 1e-8 (35,937 dofs)."""
    _say("case golden poisson32")
    g = np.load(os.path.join(GOLD, "poisson32.npz"), allow_pickle=False)
    p = CubeProblem(8, 4, rhs="const")
    b, (it, reason, cond) = bddcml_amd.solve_cube(p, tol=1e-8, maxit=500)
    assert reason == 0
    assert abs(it - int(g["iterations"])) <= 1
    assert abs(cond - float(g["cond"])) <= 1e-6 * float(g["cond"])
    c = b.counts()
    assert [c["ncorner"], c["nedge"], c["nface"], c["ndofc"], c["nig"]] == list(g["counts"])
    assert _rel(b.interface_solution(), g["u"]) < 1e-9
    b.finalize()
