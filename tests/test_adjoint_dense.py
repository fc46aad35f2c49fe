"""The adjoint gradient used by bddcml_amd.adjoint, pinned on dense algebra:
for A(k) u = f with A = sum_e k_e A_e (Dirichlet dofs eliminated),
dJ/dk_e = -lam_e^T A_e u_e with A lam = u - u_t (lam = 0 at fixed dofs),
against central finite differences of J = 0.5 |u - u_t|^2."""
import numpy as np

from bddcml_amd.cube import CubeProblem, lognormal, uniform01


def _solve(k, p0):
    p = CubeProblem(p0.nel_sub, p0.nsub_edge, rhs="random", seed=5, coef=lambda e: k[e])
    A, f, fixed = p.global_dense()
    return p, np.linalg.solve(A, f), fixed


def _elements(p):
    """global element id -> global dofs of its 8 nodes (node order of inet)"""
    out = {}
    b = p.numbase
    for d in p.subdomains():
        inet = (d["inet"] - b).reshape(-1, 8)
        g = d["isvgvn"] - b
        for e, nodes in zip(d["isegn"] - b, inet):
            out[int(e)] = g[nodes]
    return out


def test_adjoint_gradient_matches_finite_differences():
    p0 = CubeProblem(2, 2, rhs="random", seed=5)
    k = lognormal(5, np.arange(p0.nelem))
    ut = 2.0 * uniform01(6, np.arange(p0.ndof, dtype=np.uint64)) - 1.0
    p, u, fixed = _solve(k, p0)
    A, _, _ = p.global_dense()
    rhs = np.where(fixed, 0.0, u - ut)
    lam = np.linalg.solve(A, rhs)
    assert np.abs(lam[fixed]).max() == 0.0
    ae = p.kref
    el = _elements(p)
    grad = np.array([-lam[el[e]] @ ae @ u[el[e]] for e in range(p.nelem)])
    J = lambda uu: 0.5 * np.sum((uu - ut) ** 2)
    for e in (0, 5, 17, 33, p.nelem - 1):
        h = 1e-6 * k[e]
        kp = k.copy(); kp[e] += h
        km = k.copy(); km[e] -= h
        fd = (J(_solve(kp, p0)[1]) - J(_solve(km, p0)[1])) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-6 * max(abs(grad).max(), 1e-30), (e, fd, grad[e])
