"""PCG iterations replayed from one captured graph (option pcg_graph, the
default) against the same iterations as separate launches: the graph runs
the same kernels in the same order, so iteration counts, residual histories
and solutions must be bit-identical; the graph is kept across solves on one
setup and must still give the oracle's solutions there, and on three
levels."""
import sys

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import CubeProblem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nel,ns,kind,nsub2", [(4, 3, "poisson", 0), (3, 2, "elasticity", 0),
                                              (3, 4, "poisson", 8)])
def test_graph_iterations_equal_launches(oracle_lib, nel, ns, kind, nsub2):
    _say(f"case pcg graph vs launches nel={nel} ns={ns} kind={kind} nsub2={nsub2}")
    p = CubeProblem(nel, ns, kind=kind, rhs="random", seed=6)
    out = {}
    for graph in (0, 1):
        b, (it, r, c) = bddcml_amd.solve_cube(p, tol=1e-11, maxit=300, nsub2=nsub2,
                                              options={"pcg_graph": graph})
        try:
            assert r == 0 and it >= 3
            out[graph] = (it, c, b.residual_history(), b.interface_solution())
        finally:
            b.finalize()
    assert out[0][0] == out[1][0]
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])
    assert np.array_equal(out[0][3], out[1][3])
    if nsub2 == 0:
        o, (it_o, r_o, c_o) = oracle_lib.solve_cube(p, tol=1e-11, maxit=300)
        assert abs(out[1][0] - it_o) <= 1
        assert _relmax(out[1][3], o.interface_solution()) < 1e-10
        o.close()


def test_graph_kept_across_solves(oracle_lib):
    """New right-hand sides on the same setup reuse the captured graph; each
    solve matches the oracle's solve of that problem."""
    _say("case pcg graph kept across solves")
    p1 = CubeProblem(4, 3, rhs="random", seed=3)
    b, (it1, r1, _) = bddcml_amd.solve_cube(p1, tol=1e-11, maxit=300)
    try:
        o1, (it1_o, _, _) = oracle_lib.solve_cube(p1, tol=1e-11, maxit=300)
        assert r1 == 0 and abs(it1 - it1_o) <= 1
        assert _relmax(b.interface_solution(), o1.interface_solution()) < 1e-10
        o1.close()
        for seed in (4, 5):
            p2 = CubeProblem(4, 3, rhs="random", seed=seed, dirichlet="random")
            for d in p2.subdomains():
                b.change_subdomain_data(d)
            b.setup_new_data()
            it2, r2, _ = b.solve(0, 1e-11, 300, 50)
            o2, (it2_o, _, _) = oracle_lib.solve_cube(p2, tol=1e-11, maxit=300)
            assert r2 == 0 and abs(it2 - it2_o) <= 1
            assert _relmax(b.interface_solution(), o2.interface_solution()) < 1e-10
            o2.close()
    finally:
        b.finalize()
