"""Three-level BDDC on the device against the three-level oracle.

Level 2's elements are level 1's subdomains, its nodes level 1's coarse
nodes, its element matrices the coarse matrices Ac_i; one application of
its BDDC replaces the coarse solve (src/module_levels.f90:1297-1530,
3849-4249).  The oracle composes two CPU instances the same way
(oracle/multilevel.py).  Checked per case:
- the second level's mesh and globs: subdomain count, corners, edges, faces,
  coarse dofs and interface dofs equal to the oracle's (integers, exact);
- FP64: the preconditioner M r within 1e-10 relative of the oracle's, the
  interface solution within 1e-10 relative with PCG iteration counts within
  +-1 and the Lanczos lambda_max within 1e-6 relative;
- mixed precision (FP32 factors of level 1): interface solution within 1e-8
  of the FP64 oracle, iterations within +-1 (the layered boxes, whose
  stiff slabs make S' ill-conditioned and PCG's count sensitive to the
  factor's precision: 15 % + 2, FP64 5 % + 1, as for two levels);
- one second-level subdomain: identical to two levels (the coarse problem is
  then solved directly);
- the levels below the first with their triangular sweeps instead of the
  explicit inverses (option explicit_inverse = 0): the same bars as FP64.
"""
import sys

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import BoxProblem, CubeProblem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = [
    ("cube-poisson", lambda: CubeProblem(4, 4, kind="poisson", rhs="random", seed=3), 8, 0),
    ("cube-poisson-16", lambda: CubeProblem(3, 4, kind="poisson", rhs="random", seed=4), 16, 1),
    ("cube-elasticity", lambda: CubeProblem(3, 3, kind="elasticity", rhs="random", seed=5), 4, 0),
    ("boxes-elasticity", lambda: BoxProblem(12, 24, seed=3, layers=3, thick=1), 5, 1),
]


@pytest.mark.parametrize("name,make,nsub2,wt", CASES, ids=[c[0] for c in CASES])
def test_three_levels_match_oracle(oracle_lib, name, make, nsub2, wt):
    from oracle.multilevel import solve_three_levels
    _say(f"case three_levels {name} nsub2={nsub2} wt={wt}")
    p = make()
    tol = 1e-10
    t, (it_o, r_o, cond_o) = solve_three_levels(p, nsub2, weights_type=wt, tol=tol, maxit=400)
    assert r_o == 0
    u_o = t.o1.interface_solution()
    c2 = t.o2.counts()
    b, (it, r, cond) = bddcml_amd.solve_cube(p, weights_type=wt, tol=tol, maxit=400, nsub2=nsub2)
    _say(f"  oracle {it_o} iterations cond {cond_o:.6g}; device {it} iterations cond {cond:.6g}")
    lc = b.level2_counts()
    assert lc is not None
    assert lc["nsub"] == nsub2 and lc["ndof"] == t.ndofc
    for k in ("ncorner", "nedge", "nface", "ndofc", "nig"):
        assert lc[k] == c2[k], k
    layered = isinstance(p, BoxProblem)
    assert r == 0
    assert abs(it - it_o) <= (max(1, 0.05 * it_o) if layered else 1)
    assert abs(cond - cond_o) <= 1e-6 * cond_o
    assert _relmax(b.interface_solution(), u_o) < 1e-10
    rng = np.random.default_rng(7)
    x = rng.standard_normal(len(u_o))
    assert _relmax(b.apply_pc(x), t.o1.apply_pc(x)) < 1e-10
    b.finalize()

    _say(f"case three_levels {name} mixed precision")
    b, (it, r, _) = bddcml_amd.solve_cube(p, weights_type=wt, tol=tol, maxit=400, nsub2=nsub2,
                                          options={"precision": 32})
    _say(f"  mixed: {it} iterations")
    assert r == 0
    assert (it <= 1.15 * it_o + 2) if layered else abs(it - it_o) <= 1
    assert _relmax(b.interface_solution(), u_o) < 1e-8
    b.finalize()

    # the levels below the first by default apply explicit inverses (S'^-1
    # and the interior factor's) by tile GEMVs; explicit_inverse = 0 keeps
    # the triangular sweeps there: the same operator to rounding
    _say(f"case three_levels {name} sweeps (explicit_inverse 0)")
    b, (it, r, _) = bddcml_amd.solve_cube(p, weights_type=wt, tol=tol, maxit=400, nsub2=nsub2,
                                          options={"explicit_inverse": 0})
    assert r == 0
    assert abs(it - it_o) <= (max(1, 0.05 * it_o) if layered else 1)
    assert _relmax(b.interface_solution(), u_o) < 1e-10
    assert _relmax(b.apply_pc(x), t.o1.apply_pc(x)) < 1e-10
    b.finalize()
    t.close()


def test_one_second_level_subdomain_is_two_levels(oracle_lib):
    _say("case three_levels one second-level subdomain")
    p = CubeProblem(3, 3, kind="poisson", rhs="random", seed=6)
    b2, res2 = bddcml_amd.solve_cube(p, tol=1e-10, maxit=300)
    u2 = b2.interface_solution()
    b2.finalize()
    b3, res3 = bddcml_amd.solve_cube(p, tol=1e-10, maxit=300, nsub2=1)
    assert b3.level2_counts() is None
    assert res3[0] == res2[0] and res3[1] == 0
    assert _relmax(b3.interface_solution(), u2) < 1e-13
    b3.finalize()


def test_three_levels_bicgstab_and_new_data(oracle_lib):
    """BiCGSTAB with three levels, then a new right-hand side on the kept
    factorization (change_subdomain_data + setup_new_data, level 2 untouched),
    both against the three-level oracle: iterations +-1, interface and local
    solutions within 1e-10."""
    from oracle.multilevel import solve_three_levels
    _say("case three_levels bicgstab and new data")
    p = CubeProblem(3, 4, kind="poisson", rhs="random", seed=11)
    q = CubeProblem(3, 4, kind="poisson", rhs="random", seed=12)
    tol = 1e-10
    t, (it_o, r_o, _) = solve_three_levels(p, 8, method=1, tol=tol, maxit=300)
    assert r_o == 0
    u_o = t.o1.interface_solution()
    t.close()
    b = bddcml_amd.Bddcml()
    b.init(3, [p.nsub, 8, 1], numbase=p.numbase)
    for d in p.subdomains():
        b.upload_subdomain_data(d)
    b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
    it, r, _ = b.solve(1, tol, 300, 50)
    _say(f"  BiCGSTAB: device {it}, oracle {it_o} iterations")
    assert r == 0 and abs(it - it_o) <= 1
    assert _relmax(b.interface_solution(), u_o) < 1e-10

    t, (it_o, r_o, _) = solve_three_levels(q, 8, method=0, tol=tol, maxit=300)
    assert r_o == 0
    for d in q.subdomains():
        b.change_subdomain_data(d)
    b.setup_new_data()
    it, r, _ = b.solve(0, tol, 300, 50)
    _say(f"  new right-hand side, PCG: device {it}, oracle {it_o} iterations")
    assert r == 0 and abs(it - it_o) <= 1
    u_o = t.o1.interface_solution()
    assert _relmax(b.interface_solution(), u_o) < 1e-10
    for d in q.subdomains():
        ls = b.download_local_solution(d["isub"], d["ndofs"])
        assert np.abs(ls - t.o1.local_solution(d["isub"])).max() <= 1e-10 * np.abs(u_o).max()
    t.close()
    b.finalize()


DEEP_CASES = [
    ("poisson-4-levels", lambda: CubeProblem(3, 4, kind="poisson", rhs="random", seed=13),
     [64, 8, 2, 1]),
    ("elasticity-4-levels", lambda: CubeProblem(2, 3, kind="elasticity", rhs="random", seed=14),
     [27, 8, 2, 1]),
    ("poisson-5-levels", lambda: CubeProblem(2, 4, kind="poisson", rhs="random", seed=15),
     [64, 16, 4, 2, 1]),
]


@pytest.mark.parametrize("name,make,nsublev", DEEP_CASES, ids=[c[0] for c in DEEP_CASES])
def test_deeper_levels_match_oracle(oracle_lib, name, make, nsublev):
    """Four and five levels: every level l+1 is built from level l's coarse
    nodes and coarse matrices and replaces its coarse solve by one BDDC
    application, recursively (src/module_levels.f90:1297-1530, 3849-4249);
    the oracle chains one CPU instance per level the same way.  Counts of
    every intermediate level exact; FP64 M r and interface solution within
    1e-10 relative, iterations +-1, lambda_max within 1e-6; FP32 local
    factors: interface solution within 1e-8, iterations +-1."""
    from oracle.multilevel import solve_levels
    _say(f"case levels {name} nsublev={nsublev}")
    p = make()
    tol = 1e-10
    t, (it_o, r_o, cond_o) = solve_levels(p, nsublev, tol=tol, maxit=400)
    assert r_o == 0
    u_o = t.o1.interface_solution()
    b, (it, r, cond) = bddcml_amd.solve_cube(p, tol=tol, maxit=400, nsublev=nsublev)
    _say(f"  oracle {it_o} iterations cond {cond_o:.6g}; device {it} iterations cond {cond:.6g}")
    for lev in range(2, len(nsublev)):
        lc, oc = b.level_counts(lev), t.levels[lev - 1].counts()
        assert lc is not None and lc["nsub"] == nsublev[lev - 1]
        assert lc["ndof"] == t.ndofcs[lev - 2]
        for k in ("ncorner", "nedge", "nface", "ndofc", "nig"):
            assert lc[k] == oc[k], (lev, k)
    assert b.level_counts(len(nsublev)) is None
    assert r == 0 and abs(it - it_o) <= 1
    assert abs(cond - cond_o) <= 1e-6 * cond_o
    assert _relmax(b.interface_solution(), u_o) < 1e-10
    rng = np.random.default_rng(8)
    x = rng.standard_normal(len(u_o))
    assert _relmax(b.apply_pc(x), t.o1.apply_pc(x)) < 1e-10
    b.finalize()

    _say(f"case levels {name} mixed precision")
    b, (it, r, _) = bddcml_amd.solve_cube(p, tol=tol, maxit=400, nsublev=nsublev,
                                          options={"precision": 32})
    _say(f"  mixed: {it} iterations")
    assert r == 0 and abs(it - it_o) <= 1
    assert _relmax(b.interface_solution(), u_o) < 1e-8
    b.finalize()
    t.close()
