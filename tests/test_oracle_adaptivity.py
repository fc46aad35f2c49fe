"""Pins the oracle's adaptive constraints (oracle/adaptivity.py).

- The pair operators are checked against an independent elementwise form:
  x^T A x = |J x|_S^2 with J x the weighted jump (w_j (x_a - x_b) on the
  i copy of a common dof, w_i (x_b - x_a) on the j copy, 0 elsewhere) of
  the projected x, and the projection P against the constraints it must
  satisfy (D^T P x = 0, x - P x in range(D)).
- Every returned eigenpair satisfies A v = lambda B v, v^T B v = 1, v in
  range(P); eigenvalues are >= 0 and the Rayleigh quotient of random vectors
  of range(P) never exceeds the largest one.
- The preconditioner depends only on the span of each face's rows: loading
  the adaptive rows with or without the face orthonormalization gives the
  same M r (a wrong orthonormalization changes the span and fails).
- The effect: on elasticity with stiff channels crossing the faces, the
  adaptive constraints bring the condition number near the largest pair
  estimate (far below the arithmetic-constraint value), and the solution
  still equals the dense direct solve.
"""
import numpy as np
import pytest

from bddcml_amd.cube import CubeProblem
from oracle.adaptivity import adaptive_constraints, orthonormal_rows, pair_eigen, pair_problem


def _channels(n, jump):
    def coef(e):
        e = np.asarray(e)
        y, z = (e // n) % n, e // (n * n)
        c = np.ones(e.shape)
        c[(y % 3 == 1) & (z % 3 == 1)] = jump
        return c
    return coef


def _problem(nel=3, ns=3, kind="elasticity"):
    return CubeProblem(nel, ns, kind=kind, rhs="random", seed=2, coef=_channels(nel * ns, 1e4))


def _setup(oracle_lib, p, wt=1):
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, wt)
    return o


def _subs(o, p, wt=1):
    from oracle.adaptivity import _rho
    out = {}
    for d in p.subdomains():
        isub = d["isub"]
        C, ind = o.sub_constraints(isub)
        out[isub] = dict(gd=o.sub_interface_gdofs(isub), C=C, ind=ind, S=o.sub_schur(isub),
                         rho=_rho(o, d, isub, wt, None))
    return out


def test_pair_operators_and_eigenpairs(oracle_lib):
    p = _problem()
    o = _setup(oracle_lib, p)
    subs = _subs(o, p)
    rng = np.random.default_rng(0)
    # pairs: centre subdomain 14 with its six face neighbours
    for j in (5, 11, 13, 15, 17, 23):
        i_, j_ = sorted((14, j))
        si, sj = subs[i_], subs[j_]
        pp = pair_problem(si, sj)
        n, ni = pp["A"].shape[0], pp["ni"]
        # projection: D^T P x = 0 and x - P x in range(D)
        common = np.intersect1d(si["ind"], sj["ind"])
        D = np.zeros((n, len(common)))
        for k, c in enumerate(common):
            D[:ni, k] = si["C"][np.nonzero(si["ind"] == c)[0][0]]
            D[ni:, k] = -sj["C"][np.nonzero(sj["ind"] == c)[0][0]]
        x = rng.standard_normal(n)
        Px = pp["P"] @ x
        assert np.abs(D.T @ Px).max() <= 1e-12 * np.abs(x).max() * np.abs(D).max()
        coef = np.linalg.lstsq(D, x - Px, rcond=None)[0]
        assert np.abs(D @ coef - (x - Px)).max() <= 1e-10 * np.abs(x).max()
        # quadratic forms against the elementwise jump
        wi = si["rho"][pp["ci"]] / (si["rho"][pp["ci"]] + sj["rho"][pp["cj"]])
        jx = np.zeros(n)
        a, b = pp["ci"], ni + pp["cj"]
        jx[a] = (1.0 - wi) * (Px[a] - Px[b])
        jx[b] = wi * (Px[b] - Px[a])
        S = pp["S"]
        assert abs(x @ pp["A"] @ x - jx @ S @ jx) <= 1e-10 * abs(jx @ S @ jx) + 1e-14
        assert abs(x @ pp["B"] @ x - Px @ S @ Px) <= 1e-10 * abs(Px @ S @ Px)
        # eigenpairs
        lam, V = pair_eigen(pp)
        assert lam.min() >= -1e-8 * lam.max()
        for k in range(min(10, len(lam))):
            v = V[:, k]
            r = pp["A"] @ v - lam[k] * (pp["B"] @ v)
            assert np.abs(r).max() <= 1e-9 * max(lam[0], 1.0) * np.abs(pp["B"] @ v).max() + 1e-12
            assert abs(v @ pp["B"] @ v - 1.0) <= 1e-10
            assert np.abs(pp["P"] @ v - v).max() <= 1e-10 * np.abs(v).max()
        for _ in range(20):
            y = pp["P"] @ rng.standard_normal(n)
            assert (y @ pp["A"] @ y) / (y @ pp["B"] @ y) <= lam[0] * (1 + 1e-10)
    o.close()


def test_orthonormal_rows_keep_the_span():
    rng = np.random.default_rng(1)
    M = rng.standard_normal((4, 30))
    M = np.vstack([M, M[0] + 2 * M[1]])  # a dependent row is dropped
    Q = orthonormal_rows(M)
    assert Q.shape[0] == 4
    np.testing.assert_allclose(Q @ Q.T, np.eye(4), atol=1e-13)
    # same row space
    P = Q.T @ Q
    np.testing.assert_allclose(P @ M.T, M.T, atol=1e-12)


def test_preconditioner_depends_only_on_face_spans(oracle_lib, monkeypatch):
    """M r with orthonormalized face rows equals M r with the raw rows."""
    import oracle.adaptivity as ad
    p = _problem(4, 2)
    o1 = _setup(oracle_lib, p)
    rep1 = adaptive_constraints(o1, p, threshold=10.0, weights_type=1)
    assert sum(q["nadaptive"] for q in rep1["pairs"]) > 0

    def raw_rows(M):
        # the same rows without orthonormalization (full rank here)
        return M.copy()

    monkeypatch.setattr(ad, "orthonormal_rows", raw_rows)
    o2 = _setup(oracle_lib, p)
    rep2 = adaptive_constraints(o2, p, threshold=10.0, weights_type=1)
    assert rep2["ndofc"] == rep1["ndofc"]
    x = np.random.default_rng(2).standard_normal(o1.counts()["nig"])
    z1, z2 = o1.apply_pc(x), o2.apply_pc(x)
    assert np.abs(z1 - z2).max() <= 1e-9 * np.abs(z1).max()
    o1.close()
    o2.close()


@pytest.mark.parametrize("nel,ns", [(3, 3), (4, 2)])
def test_adaptive_constraints_condition_and_solution(oracle_lib, nel, ns):
    p = _problem(nel, ns)
    o = _setup(oracle_lib, p)
    it0, r0, c0 = o.solve(0, 1e-10, 500, 50)
    rep = adaptive_constraints(o, p, threshold=10.0, weights_type=1)
    it1, r1, c1 = o.solve(0, 1e-10, 500, 50)
    assert r0 == 0 and r1 == 0
    nad = [q["nadaptive"] for q in rep["pairs"]]
    assert sum(nad) > 0 and max(nad) < 10  # the eigenvector count does not bind
    assert c1 <= 1.1 * max(rep["est"], 1.0)
    assert c1 < c0 / 5 and it1 < it0
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    assert np.abs(o.interface_solution() - u[o.interface_dofs()]).max() <= 1e-9 * np.abs(u).max()
    o.close()
