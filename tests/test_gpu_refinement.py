"""GPU parity of the refined FP32 form (option "refinement"): interior
factorization, Schur complements and S' factor in FP32 on the first level,
every solve refined in FP64 from the residual of the uploaded subdomain
triplets.

Each case prints its name before touching the GPU.  Bar: the interface and
local solutions within 1e-8 relative of the FP64 oracle's (solved to 1e-10
or 1e-12), the refinement having run (at least one correction) and stopped
on the FP64 interface residual below the tolerance.  The FP32 Schur
complements are checked against the oracle's at 1e-5 relative, which shows
that they are FP32 (they differ from the FP64 ones by more than 1e-12) and
still the same operator.  The iteration count sums the first solve and the
corrections; it is bounded by twice the oracle's plus 4.
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


def _check_solution(b, o, p, bound):
    u_ref = o.interface_solution()
    assert _relmax(b.interface_solution(), u_ref) < bound
    for d in p.subdomains():
        ls = b.download_local_solution(d["isub"], d["ndofs"])
        assert np.abs(ls - o.local_solution(d["isub"])).max() <= bound * np.abs(u_ref).max()


def _check_refined(b, it, it_o, tol):
    t = b.timings()
    _say(f"  iterations {it} (oracle {it_o}), corrections {int(t['refinements'])}, "
         f"interface residual {t['refine_relres']:.3e}")
    assert t["refinements"] >= 1
    assert t["refine_relres"] < tol
    assert it <= 2 * it_o + 4


@pytest.mark.parametrize("nel,ns,kind,wt,method", [(4, 3, "poisson", 0, 0),
                                                   (3, 3, "poisson", 1, 0),
                                                   (3, 2, "elasticity", 0, 0),
                                                   (4, 3, "poisson", 0, 1)],
                         ids=["poisson", "poisson_weights1", "elasticity", "bicgstab"])
def test_refined_cube(oracle_lib, nel, ns, kind, wt, method):
    _say(f"case refined cube nel={nel} ns={ns} kind={kind} wt={wt} method={method}")
    tol = 1e-10
    p = CubeProblem(nel, ns, kind=kind, rhs="random", seed=31)
    o, (it_o, r_o, _) = oracle_lib.solve_cube(p, weights_type=wt, method=method, tol=tol,
                                              maxit=300)
    b, (it, r, _) = bddcml_amd.solve_cube(p, weights_type=wt, method=method, tol=tol, maxit=300,
                                          options={"refinement": 1})
    try:
        assert r == r_o == 0
        _check_refined(b, it, it_o, tol)
        for isub in (1, p.nsub):
            S = b.sub_matrix(isub, "S")
            S_ref = o.sub_schur(isub)
            assert _relmax(S, S_ref) < 1e-5, f"FP32 S of subdomain {isub}"
        _check_solution(b, o, p, 1e-8)
    finally:
        o.close()
        b.finalize()


@pytest.mark.parametrize("forms", [1, 2], ids=["many_subdomain_forms", "few_subdomain_forms"])
def test_refined_headline_shape(oracle_lib, forms):
    """27 subdomains of 16^3 elements (the subdomain shape of the 128^3
    workload), both launch forms."""
    _say(f"case refined headline shape launch_forms={forms}")
    tol = 1e-10
    p = CubeProblem(16, 3, rhs="random", seed=2)
    o, (it_o, r_o, _) = oracle_lib.solve_cube(p, tol=tol, maxit=300)
    b, (it, r, _) = bddcml_amd.solve_cube(p, tol=tol, maxit=300,
                                          options={"refinement": 1, "launch_forms": forms})
    try:
        assert r == r_o == 0
        _check_refined(b, it, it_o, tol)
        _check_solution(b, o, p, 1e-8)
    finally:
        o.close()
        b.finalize()


def test_refined_with_explicit_inverse(oracle_lib):
    """The refined form with the explicit S'^-1 formed at setup (FP32 tiles,
    applied by the tile GEMV) as its local preconditioner."""
    _say("case refined cube with explicit_inverse 1")
    tol = 1e-10
    p = CubeProblem(4, 3, kind="poisson", rhs="random", seed=34)
    o, (it_o, r_o, _) = oracle_lib.solve_cube(p, tol=tol, maxit=300)
    b, (it, r, _) = bddcml_amd.solve_cube(p, tol=tol, maxit=300,
                                          options={"refinement": 1, "explicit_inverse": 1})
    try:
        assert r == r_o == 0
        _check_refined(b, it, it_o, tol)
        _check_solution(b, o, p, 1e-8)
    finally:
        o.close()
        b.finalize()


def test_refined_uneven_boxes(oracle_lib):
    """Uneven elasticity boxes with stiff layers (coefficient jump 1e3): the
    FP32 interior factors of the layered subdomains are the least accurate,
    so the refinement takes more corrections; the solution still meets
    1e-8."""
    _say("case refined boxes n=10 ns=12 elasticity")
    tol = 1e-12
    p = BoxProblem(10, 12, seed=3, kind="elasticity", layers=6, jump=1e3, thick=1)
    o, (it_o, r_o, _) = oracle_lib.solve_cube(p, tol=tol, maxit=800)
    b, (it, r, _) = bddcml_amd.solve_cube(p, tol=tol, maxit=800, options={"refinement": 1})
    try:
        assert r == r_o == 0
        _check_refined(b, it, it_o, tol)
        _check_solution(b, o, p, 1e-8)
    finally:
        o.close()
        b.finalize()


def test_refined_three_levels():
    """Three levels: the first level in FP32 with refinement, the second
    level FP64; against the oracle's three-level solve."""
    from oracle.multilevel import solve_three_levels
    _say("case refined three levels nel=3 ns=4 nsub2=8")
    tol = 1e-12
    p = CubeProblem(3, 4, kind="poisson", rhs="random", seed=23)
    t, (it_o, r_o, _) = solve_three_levels(p, 8, tol=tol, maxit=400)
    b, (it, r, _) = bddcml_amd.solve_cube(p, tol=tol, maxit=400, nsub2=8,
                                          options={"refinement": 1})
    try:
        assert r == r_o == 0
        _check_refined(b, it, it_o, tol)
        _check_solution(b, t.o1, p, 1e-8)
    finally:
        b.finalize()


def test_refined_global_memory_forms(oracle_lib):
    """The residual, the sweeps and the preconditioner in their
    global-memory forms (those of subdomains too large for LDS)."""
    _say("case refined cube sweep_in_global=1")
    tol = 1e-10
    p = CubeProblem(4, 3, kind="poisson", rhs="random", seed=32)
    o, (it_o, r_o, _) = oracle_lib.solve_cube(p, tol=tol, maxit=300)
    b, (it, r, _) = bddcml_amd.solve_cube(p, tol=tol, maxit=300,
                                          options={"refinement": 1, "sweep_in_global": 1})
    try:
        assert r == r_o == 0
        _check_refined(b, it, it_o, tol)
        _check_solution(b, o, p, 1e-8)
    finally:
        b.set_option("sweep_in_global", 0)
        o.close()
        b.finalize()


@pytest.mark.parametrize("options", [{}, {"refinement": 1}], ids=["fp64", "refined"])
def test_repeated_solve(oracle_lib, options):
    """A second bddcml_solve on the same setup and right-hand side (its
    initial guess the first solution) keeps the solution: the forward sweep
    of the right-hand side, which the first back substitution overwrote, is
    redone.  The second solve starts from a residual at rounding level
    relative to which it cannot reach the tolerance, so it ends by
    stagnation or the iteration limit; only the solution is checked."""
    _say(f"case repeated solve {options}")
    tol = 1e-10
    p = CubeProblem(4, 3, kind="poisson", rhs="random", seed=33)
    o, _ = oracle_lib.solve_cube(p, tol=tol, maxit=300)
    b, (it, r, _) = bddcml_amd.solve_cube(p, tol=tol, maxit=300, options=options)
    try:
        assert r == 0
        bound = 1e-8 if options else 1e-10
        _check_solution(b, o, p, bound)
        b.solve(0, tol, 40, 5)
        _check_solution(b, o, p, 1e-8 if options else 1e-9)
    finally:
        o.close()
        b.finalize()


@pytest.mark.parametrize("options", [{}, {"precision": 32}, {"refinement": 1}],
                         ids=["fp64", "mixed", "refined"])
def test_results_are_bitwise_reproducible(options):
    """Every reduction runs in a fixed order (CSR gathers, two-stage dots,
    the fixed butterflies of the residual and GEMV kernels), so two setups
    and solves of the same problem give the same bits."""
    _say(f"case reproducible {options}")
    p = CubeProblem(4, 3, kind="elasticity", rhs="random", seed=44)
    out = []
    for _ in range(2):
        b, (it, r, c) = bddcml_amd.solve_cube(p, tol=1e-10, maxit=300, options=options)
        try:
            assert r == 0
            out.append((it, c, b.interface_solution(), b.residual_history()))
        finally:
            b.finalize()
    (it0, c0, u0, h0), (it1, c1, u1, h1) = out
    assert it0 == it1 and c0 == c1
    assert np.array_equal(u0, u1)
    assert np.array_equal(h0, h1)
