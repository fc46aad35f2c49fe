"""Committed fixtures (tests/golden/, written by make_golden.py from the
oracle) against the oracle and the synthetic-input generator as they are
now, and against the product's host plan: a change that moves any of them
fails here before it reaches the GPU tests."""
import os

import numpy as np
import pytest

from bddcml_amd.cube import CubeProblem, uniform01

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_rng_streams():
    g = _load("rng.npz")
    for s in (2, 5, 6):
        for k, off in (("lo", 0), ("hi", 1 << 40)):
            v = uniform01(s, np.arange(64, dtype=np.uint64) + np.uint64(off))
            assert np.array_equal(v, g[f"seed{s}_{k}"])


@pytest.mark.parametrize("name,kind", [("cube2x2.npz", "poisson"), ("elast2x2.npz", "elasticity")])
def test_oracle_matches_small_fixtures(oracle_lib, name, kind):
    g = _load(name)
    p = CubeProblem(2, 2, kind=kind, rhs="random", seed=7)
    o, (it, reason, cond) = oracle_lib.solve_cube(p, tol=1e-12, maxit=300)
    assert reason == 0 and it == int(g["iterations"])
    assert np.array_equal(o.interface_dofs(), g["interface_dofs"])
    for s in range(p.nsub):
        isub = s + 1
        assert np.array_equal(o.sub_interface_gdofs(isub), g[f"gdofs_{isub}"])
        assert np.abs(o.sub_schur(isub) - g[f"S_{isub}"]).max() <= 1e-13 * np.abs(g[f"S_{isub}"]).max()
        assert np.abs(o.sub_phi(isub) - g[f"phi_{isub}"]).max() <= 1e-12
        assert np.abs(o.sub_coarse(isub) - g[f"ac_{isub}"]).max() <= 1e-12 * np.abs(g[f"ac_{isub}"]).max()
    assert np.abs(o.apply_schur(g["x"]) - g["Sx"]).max() <= 1e-12 * np.abs(g["Sx"]).max()
    assert np.abs(o.apply_pc(g["r"]) - g["Mr"]).max() <= 1e-12 * np.abs(g["Mr"]).max()
    assert np.abs(o.interface_solution() - g["u"]).max() <= 1e-12 * np.abs(g["u"]).max()
    o.close()


def test_oracle_matches_poisson32(oracle_lib):
    """The reference's poisson_on_cube case (32^3, 4^3 subdomains)."""
    g = _load("poisson32.npz")
    p = CubeProblem(8, 4, rhs="const")
    o, (it, reason, cond) = oracle_lib.solve_cube(p, tol=1e-8, maxit=500)
    assert reason == 0
    assert it == int(g["iterations"])
    assert abs(cond - float(g["cond"])) <= 1e-10 * float(g["cond"])
    assert np.array_equal(np.array(list(o.counts().values())), g["counts"])
    assert np.abs(o.interface_solution() - g["u"]).max() <= 1e-12 * np.abs(g["u"]).max()
    o.close()
