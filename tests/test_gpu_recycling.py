"""Krylov recycling on the device PCG against the oracle (which solves every
right-hand side from scratch).

The reference (src/module_krylov.f90:186-216, 308-455, 547-613, 663-726,
1859-2036) keeps the PCG directions p / sqrt(p.Sp) and S p / sqrt(p.Sp) of
every solve, corrects the initial guess of a later solve by V V^T r0 and
A-orthogonalizes each new direction against the kept ones by two classical
Gram-Schmidt passes.  Checked here:
- first solve (empty basis): iteration count within +-1 and interface
  solution within 1e-10 of the oracle;
- the kept basis stays A-orthonormal, ||V^T W - I||_F < 1e-6 (the reference
  warns above 1e-4), with one direction per iteration;
- a related right-hand side on the same factorization (the first plus 1 %
  of another): solution within 1e-8 of the oracle's (by linearity from two
  oracle solves), in fewer iterations than the same problem solved without
  recycling;
- the same right-hand side once more: the initial-guess correction alone
  meets the tolerance or nearly so (at most 2 iterations).
"""
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


@pytest.mark.parametrize("nel,ns,kind", [(4, 3, "poisson"), (3, 3, "elasticity")])
def test_recycling_matches_oracle_and_saves_iterations(oracle_lib, nel, ns, kind):
    _say(f"case recycling nel={nel} ns={ns} kind={kind}")
    p1 = CubeProblem(nel, ns, kind=kind, rhs="random", seed=3)
    p2 = CubeProblem(nel, ns, kind=kind, rhs="random", seed=4)
    tol = 1e-10

    def mixed():
        for d1, d2 in zip(p1.subdomains(), p2.subdomains()):
            d = dict(d1)
            d["rhs"] = d1["rhs"] + 0.01 * d2["rhs"]
            yield d

    # iterations for the related right-hand side without recycling (the
    # library holds one solver at a time, so this runs first)
    c = bddcml_amd.Bddcml()
    c.init(2, [p1.nsub, 1], numbase=p1.numbase)
    for d in mixed():
        c.upload_subdomain_data(d)
    c.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
    it2_ref, r2_ref, _ = c.solve(0, tol, 300, 50)
    assert r2_ref == 0
    c.finalize()

    b = bddcml_amd.Bddcml()
    b.init(2, [p1.nsub, 1], numbase=p1.numbase)
    for d in p1.subdomains():
        b.upload_subdomain_data(d)
    b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
    it1, r1, _ = b.solve(0, tol, 300, 50, recycling=1, max_number_of_stored_vectors=100)
    o1, (it1_o, r1_o, _) = oracle_lib.solve_cube(p1, tol=tol, maxit=300)
    assert r1 == r1_o == 0
    assert abs(it1 - it1_o) <= 1
    o1_u = o1.interface_solution()
    assert _relmax(b.interface_solution(), o1_u) < 1e-10
    o1_loc = {d["isub"]: o1.local_solution(d["isub"]) for d in p1.subdomains()}
    o1.close()
    nvec, orth = b.recycling_basis()
    _say(f"  first solve {it1} iterations, {nvec} directions kept, ||V^T W - I|| = {orth:.2e}")
    assert nvec == min(it1, 100)
    assert orth < 1e-6

    # a related right-hand side (the first plus a 1 % change), the case
    # recycling is for: by linearity its solution is u1 + 0.01 u2
    o2, (_, r2_o, _) = oracle_lib.solve_cube(p2, tol=tol, maxit=300)
    assert r2_o == 0
    u12 = o1_u + 0.01 * o2.interface_solution()
    loc12 = {d["isub"]: o1_loc[d["isub"]] + 0.01 * o2.local_solution(d["isub"])
             for d in p2.subdomains()}
    o2.close()

    for d in mixed():
        b.change_subdomain_data(d)
    b.setup_new_data()
    it2, r2, _ = b.solve(0, tol, 300, 50, recycling=1, max_number_of_stored_vectors=100)
    _say(f"  related right-hand side: {it2} iterations with recycling, {it2_ref} without")
    assert r2 == 0
    assert it2 < it2_ref
    assert _relmax(b.interface_solution(), u12) < 1e-8
    for d in p2.subdomains():
        ls = b.download_local_solution(d["isub"], d["ndofs"])
        assert np.abs(ls - loc12[d["isub"]]).max() <= 1e-8 * np.abs(u12).max()
    nvec, orth = b.recycling_basis()
    assert nvec == min(it1 + it2, 100)
    assert orth < 1e-6

    for d in mixed():
        b.change_subdomain_data(d)
    b.setup_new_data()
    it3, r3, _ = b.solve(0, tol, 300, 50, recycling=1, max_number_of_stored_vectors=100)
    _say(f"  same right-hand side again: {it3} iterations")
    assert r3 == 0 and it3 <= 2
    assert _relmax(b.interface_solution(), u12) < 1e-8
    b.finalize()
