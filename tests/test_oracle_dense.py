"""Pins the CPU oracle against dense numpy algebra on small cubes.

Each check is chosen so that a plausible mistake in the oracle fails it:
a wrong sign or dropped term in the Schur complement, a transposed or
unweighted coarse basis, a wrong weight or a missing exchange all change at
least one of these quantities by far more than the tolerances below.
"""
import numpy as np
import pytest

from bddcml_amd.cube import CubeProblem
from tests.dense_ref import bddc_dense, kkt_blocks, schur, sub_dense

CASES = [(2, 2), (3, 2), (2, 3), (3, 3)]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"nel{c[0]}_ns{c[1]}")
def solved(request, oracle_lib):
    nel, ns = request.param
    p = CubeProblem(nel, ns, rhs="random", seed=7)
    o, res = oracle_lib.solve_cube(p, tol=1e-12, maxit=200)
    yield p, o, res
    o.close()


def test_solution_matches_direct_solve(solved):
    p, o, (it, reason, cond) = solved
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    assert reason == 0
    scale = np.abs(u).max()
    ui = o.interface_solution()
    assert np.abs(ui - u[o.interface_dofs()]).max() <= 1e-10 * scale
    for d in p.subdomains():
        ls = o.local_solution(d["isub"])
        assert np.abs(ls - u[d["isvgvn"] - 1]).max() <= 1e-10 * scale


def test_schur_complements(solved):
    p, o, _ = solved
    for d in p.subdomains():
        isub = d["isub"]
        A = sub_dense(d)
        gd = o.sub_interface_gdofs(isub)
        l_of = {int(g): k for k, g in enumerate(d["isvgvn"] - 1)}
        idx_i = np.array([l_of[int(g)] for g in gd])
        idx_o = np.setdiff1d(np.arange(d["ndofs"]), idx_i)
        S_ref = schur(A, idx_i, idx_o)
        S = o.sub_schur(isub)
        assert np.abs(S - S_ref).max() <= 1e-12 * np.abs(S_ref).max()
        np.testing.assert_allclose(S, S.T, atol=1e-14 * np.abs(S).max())


def test_coarse_basis_and_matrix(solved):
    p, o, _ = solved
    for d in p.subdomains():
        isub = d["isub"]
        S = o.sub_schur(isub)
        Cm, _ = o.sub_constraints(isub)
        K, Phi_ref, Ac_ref = kkt_blocks(S, Cm)
        Phi = o.sub_phi(isub)
        Ac = o.sub_coarse(isub)
        assert np.abs(Phi - Phi_ref).max() <= 1e-10 * max(1.0, np.abs(Phi_ref).max())
        assert np.abs(Ac - Ac_ref).max() <= 1e-10 * np.abs(Ac_ref).max()
        # energy identity Ac = Phi^T S Phi and the constraint identity C Phi = I
        assert np.abs(Phi.T @ S @ Phi - Ac).max() <= 1e-10 * np.abs(Ac).max()
        np.testing.assert_allclose(Cm @ Phi, np.eye(Cm.shape[0]), atol=1e-10)


def test_operators_match_dense_bddc(solved):
    p, o, _ = solved
    Sg, Mg, _, Acg = bddc_dense(o, p)
    rng = np.random.default_rng(3)
    n = Sg.shape[0]
    x = rng.standard_normal(n)
    y = o.apply_schur(x)
    assert np.abs(y - Sg @ x).max() <= 1e-12 * np.abs(Sg @ x).max()
    z = o.apply_pc(x)
    assert np.abs(z - Mg @ x).max() <= 1e-10 * np.abs(Mg @ x).max()
    # coarse matrix is SPD
    assert np.linalg.eigvalsh(Acg).min() > 0


def test_bddc_spectrum_and_condition_estimate(solved):
    """lambda(M S) >= 1 (BDDC with a partition of unity), and the Lanczos
    estimate reported by PCG does not exceed lambda_max(M S)."""
    p, o, (it, reason, cond) = solved
    Sg, Mg, _, _ = bddc_dense(o, p)
    ig = o.interface_dofs()
    A, f, fixed = p.global_dense()
    free = ~fixed[ig]
    Sf = Sg[np.ix_(free, free)]
    Mf = Mg[np.ix_(free, free)]
    ev = np.sort(np.linalg.eigvals(Mf @ Sf).real)
    assert ev[0] >= 1.0 - 1e-8
    assert cond <= ev[-1] * (1 + 1e-8)
    if it > 3:
        assert cond >= 0.5 * ev[-1]


def test_pcg_iteration_count_matches_numpy_pcg(solved):
    p, o, (it, reason, cond) = solved
    Sg, Mg, _, _ = bddc_dense(o, p)
    g = o.reduced_rhs()
    ig = o.interface_dofs()
    A, f, fixed = p.global_dense()
    g = np.where(fixed[ig], 0.0, g)
    u = np.zeros_like(g)
    r = g - Sg @ u
    r0 = np.linalg.norm(r)
    z = Mg @ r
    pvec = z.copy()
    rz = r @ z
    k_ref = None
    for k in range(1, 200):
        ap = Sg @ pvec
        alpha = rz / (pvec @ ap)
        u += alpha * pvec
        r -= alpha * ap
        if np.linalg.norm(r) / r0 < 1e-12:
            k_ref = k
            break
        z = Mg @ r
        rz_new = r @ z
        pvec = z + rz_new / rz * pvec
        rz = rz_new
    assert k_ref is not None
    assert abs(it - k_ref) <= 1


def test_weights_partition_of_unity(solved):
    p, o, _ = solved
    acc = np.zeros(p.ndof)
    for d in p.subdomains():
        isub = d["isub"]
        np.add.at(acc, o.sub_interface_gdofs(isub), o.sub_weights(isub))
    ig = o.interface_dofs()
    np.testing.assert_allclose(acc[ig], 1.0, rtol=0, atol=1e-15)


def test_floating_subdomain_kernel(oracle_lib):
    """The centre subdomain of a 3x3x3 split does not touch the Dirichlet
    boundary, so its Schur complement annihilates constants."""
    p = CubeProblem(2, 3)
    o, _ = oracle_lib.solve_cube(p)
    centre = 13 + p.numbase
    S = o.sub_schur(centre)
    assert np.abs(S @ np.ones(S.shape[0])).max() <= 1e-13 * np.abs(S).max()
    o.close()


def test_bicgstab_reaches_same_solution(oracle_lib):
    p = CubeProblem(3, 2, rhs="random", seed=11)
    o1, (it1, r1, c1) = oracle_lib.solve_cube(p, method=0, tol=1e-11)
    o2, (it2, r2, c2) = oracle_lib.solve_cube(p, method=1, tol=1e-11)
    assert r1 == 0 and r2 == 0 and c2 == -1.0
    u1, u2 = o1.interface_solution(), o2.interface_solution()
    assert np.abs(u1 - u2).max() <= 1e-9 * np.abs(u1).max()
    o1.close()
    o2.close()


def test_dotprod_subdomain_is_global_norm(oracle_lib):
    """Sum over subdomains of v_s' W v_s equals the global squared norm for a
    vector that is the restriction of a global vector (reference
    examples/poisson_on_cube.f90:603-613 uses it for the R^n norm)."""
    p = CubeProblem(2, 3, rhs="random", seed=5)
    o, _ = oracle_lib.solve_cube(p, tol=1e-10)
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    tot = 0.0
    for d in p.subdomains():
        ls = o.local_solution(d["isub"])
        tot += o.dotprod_subdomain(d["isub"], ls, ls)
    assert abs(tot - u @ u) <= 1e-9 * (u @ u)
    o.close()


def test_elasticity_small(oracle_lib):
    p = CubeProblem(2, 2, kind="elasticity", rhs="random", seed=4)
    o, (it, reason, cond) = oracle_lib.solve_cube(p, tol=1e-11)
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    assert reason == 0
    assert np.abs(o.interface_solution() - u[o.interface_dofs()]).max() <= 1e-9 * np.abs(u).max()
    o.close()


@pytest.mark.parametrize("nel,ns", [(2, 2), (3, 2)])
def test_nonzero_dirichlet_matches_direct_solve(oracle_lib, nel, ns):
    """Dirichlet values moved to the right-hand side (module_sm.f90:431-600,
    module_dd.f90:1081-1256): the whole solution, fixed dofs included,
    against the dense solve with u_fixed = g."""
    p = CubeProblem(nel, ns, rhs="random", seed=11, dirichlet="random")
    o, (it, reason, cond) = oracle_lib.solve_cube(p, tol=1e-12, maxit=200)
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    assert reason == 0
    scale = np.abs(u).max()
    assert np.abs(o.interface_solution() - u[o.interface_dofs()]).max() <= 1e-10 * scale
    for d in p.subdomains():
        ls = o.local_solution(d["isub"])
        assert np.abs(ls - u[d["isvgvn"] - 1]).max() <= 1e-10 * scale
        fx = d["ifix"] > 0
        assert np.abs(ls[fx] - d["fixv"][fx]).max() <= 1e-12
    o.close()


@pytest.mark.parametrize("n,ns,kind", [(6, 5, "elasticity"), (6, 7, "poisson")])
def test_uneven_boxes_with_stiff_layers(oracle_lib, n, ns, kind):
    """Subdomains of uneven size from random box splits, with coefficient
    jumps (stiff slabs), against the dense solve."""
    from bddcml_amd.cube import BoxProblem
    p = BoxProblem(n, ns, seed=3, kind=kind, layers=3, jump=100.0, thick=1)
    o, (it, reason, cond) = oracle_lib.solve_cube(p, tol=1e-12, maxit=500)
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    assert reason == 0
    scale = np.abs(u).max()
    assert np.abs(o.interface_solution() - u[o.interface_dofs()]).max() <= 1e-10 * scale
    for d in p.subdomains():
        assert np.abs(o.local_solution(d["isub"]) - u[d["isvgvn"] - 1]).max() <= 1e-10 * scale
    o.close()


def _dof_data(problem, seed=21):
    """Positive per-subdomain dof data for weights of type 3: 0.5 + u with u
    uniform[0,1) from the seed's stream at (subdomain, local dof)."""
    from bddcml_amd.cube import uniform01
    out = {}
    for d in problem.subdomains():
        idx = np.arange(d["ndofs"], dtype=np.uint64) + np.uint64(100000 * d["isub"])
        out[d["isub"]] = 0.5 + uniform01(seed, idx)
    return out


def _oracle_with(oracle_lib, p, weights_type, dof_data=None, tol=1e-10, method=0):
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d, dof_data=None if dof_data is None else dof_data[d["isub"]])
    o.setup(1, 1, weights_type)
    res = o.solve(method, tol, 500, 50)
    return o, res


@pytest.mark.parametrize("weights_type", [1, 3])
def test_weights_closed_form(oracle_lib, weights_type):
    """Weights of type 1 (subdomain diagonal stiffness) and 3 (dof data)
    against the closed form rho_i / sum_j rho_j from the uploaded triplets.
    The stiff slabs make the subdomain diagonals differ across an interface,
    so type 1 is not type 0: using the assembled diagonal instead of the
    subdomain's own, or dropping the denominator's neighbour sum, fails."""
    from bddcml_amd.cube import BoxProblem
    from tests.dense_ref import weights_closed_form
    p = BoxProblem(6, 5, seed=3, kind="poisson", layers=3, jump=100.0, thick=1)
    dd = _dof_data(p) if weights_type == 3 else None
    o, (it, reason, cond) = _oracle_with(oracle_lib, p, weights_type, dd)
    assert reason == 0
    ref = weights_closed_form(p, o, weights_type, dd)
    far_from_cardinality = 0.0
    for d in p.subdomains():
        isub = d["isub"]
        w = o.sub_weights(isub)
        np.testing.assert_allclose(w, ref[isub], rtol=1e-14, atol=0)
        w0 = weights_closed_form(p, o, 0)[isub]
        far_from_cardinality = max(far_from_cardinality, np.abs(w - w0).max())
    assert far_from_cardinality > 0.2
    # the weighted problem still solves the global system
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    assert np.abs(o.interface_solution() - u[o.interface_dofs()]).max() <= 1e-8 * np.abs(u).max()
    o.close()


@pytest.mark.parametrize("case", ["cube3x2", "cube2x3", "boxes_w1"])
def test_bicgstab_matches_numpy_bicgstab(oracle_lib, case):
    """The oracle's BiCGSTAB against a numpy BiCGSTAB on the dense S and M
    (both built independently of the oracle's solves): iteration count,
    converged reason and every half- and full-step relative residual.  A
    missing or misplaced half-step test changes the recorded history, and a
    wrong omega or beta changes every value after the first step."""
    from bddcml_amd.cube import BoxProblem
    from tests.dense_ref import bicgstab_dense
    if case == "cube3x2":
        p, wt, tol = CubeProblem(3, 2, rhs="random", seed=11), 0, 1e-10
    elif case == "cube2x3":
        p, wt, tol = CubeProblem(2, 3, rhs="random", seed=12), 0, 1e-11
    else:
        p, wt, tol = BoxProblem(6, 5, seed=3, kind="poisson", layers=3, jump=100.0, thick=1), 1, 1e-10
    o, (it, reason, cond) = _oracle_with(oracle_lib, p, wt, tol=tol, method=1)
    assert reason == 0 and cond == -1.0
    Sg, Mg, _, _ = bddc_dense(o, p)
    g = o.reduced_rhs()
    A, f, fixed = p.global_dense()
    g = np.where(fixed[o.interface_dofs()], 0.0, g)
    it_ref, reason_ref, hist_ref, u_ref = bicgstab_dense(Sg, Mg, g, tol)
    hist = o.residual_history()
    assert (it, reason) == (it_ref, reason_ref)
    assert len(hist) == len(hist_ref)
    # residuals agree to rounding amplified by the Krylov recurrences
    np.testing.assert_allclose(hist, hist_ref, rtol=1e-6, atol=1e-3 * tol)
    assert np.abs(o.interface_solution() - u_ref).max() <= 1e-9 * np.abs(u_ref).max()
    o.close()


def test_bicgstab_nondecrease_exit(oracle_lib):
    """ndecrmax = 1: the first half or full step whose residual does not
    fall below the previous one stops with converged_reason -2 at the same
    iteration as the numpy restatement."""
    from bddcml_amd.cube import BoxProblem
    from tests.dense_ref import bicgstab_dense
    p = BoxProblem(6, 5, seed=3, kind="poisson", layers=3, jump=100.0, thick=1)
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, 0)
    it, reason, cond = o.solve(1, 1e-14, 500, 1)
    Sg, Mg, _, _ = bddc_dense(o, p)
    A, f, fixed = p.global_dense()
    g = np.where(fixed[o.interface_dofs()], 0.0, o.reduced_rhs())
    it_ref, reason_ref, hist_ref, _ = bicgstab_dense(Sg, Mg, g, 1e-14, ndecrmax=1)
    assert (it, reason) == (it_ref, reason_ref)
    assert reason == -2
    o.close()
