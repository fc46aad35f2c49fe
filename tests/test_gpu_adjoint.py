"""Adjoint BDDC solve on the kept factorization and the element gradient
kernel through the C ABI (bddcml_amd.adjoint), against the dense gradient,
central finite differences of GPU forward solves, and through the
torch.autograd.Function shim."""
import sys

import numpy as np
import pytest

from bddcml_amd.cube import CubeProblem, lognormal, uniform01

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _make(nel, ns):
    return lambda k: CubeProblem(nel, ns, rhs="random", seed=5, coef=lambda e: k[e])


def _dense_gradient(p, k, ut):
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    lam = np.linalg.solve(A, np.where(fixed, 0.0, u - ut))
    b = p.numbase
    g = np.zeros(p.nelem)
    for d in p.subdomains():
        inet = (d["inet"] - b).reshape(-1, 8)
        gd = d["isvgvn"] - b
        for e, nodes in zip(d["isegn"] - b, inet):
            g[e] = -lam[gd[nodes]] @ p.kref @ u[gd[nodes]]
    return u, g


def test_gradient_against_dense_and_finite_differences():
    from bddcml_amd.adjoint import AdjointSolver
    _say("case adjoint gradient 8^3 elements, 2^3 subdomains")
    nel, ns = 4, 2
    p0 = CubeProblem(nel, ns)
    k = lognormal(5, np.arange(p0.nelem))
    ut = 2.0 * uniform01(6, np.arange(p0.ndof, dtype=np.uint64)) - 1.0
    sol = AdjointSolver(_make(nel, ns), p0.kref, tol=1e-12)
    u = sol.forward(k)
    lam = sol.adjoint(u - ut)
    g = sol.gradient(u, lam)
    u_ref, g_ref = _dense_gradient(sol.problem, k, ut)
    assert np.abs(u - u_ref).max() <= 1e-10 * np.abs(u_ref).max()
    assert np.abs(g - g_ref).max() <= 1e-8 * np.abs(g_ref).max()
    J = lambda uu: 0.5 * np.sum((uu - ut) ** 2)
    idx = (np.arange(8) * 61 + 7) % p0.nelem
    for e in idx:
        h = 1e-6 * k[e]
        kp = k.copy(); kp[e] += h
        km = k.copy(); km[e] -= h
        fd = (J(sol.forward(kp)) - J(sol.forward(km))) / (2 * h)
        assert abs(fd - g[e]) <= 1e-5 * np.abs(g).max(), (int(e), fd, g[e])
    sol.close()


def test_torch_autograd_shim():
    import torch
    from bddcml_amd.adjoint import AdjointSolver, BddcSolve
    _say("case adjoint torch autograd")
    nel, ns = 4, 2
    p0 = CubeProblem(nel, ns)
    k_np = lognormal(5, np.arange(p0.nelem))
    ut = torch.from_numpy(2.0 * uniform01(6, np.arange(p0.ndof, dtype=np.uint64)) - 1.0)
    sol = AdjointSolver(_make(nel, ns), p0.kref, tol=1e-12)
    k = torch.tensor(k_np, dtype=torch.float64, requires_grad=True)
    u = BddcSolve.apply(k, sol)
    J = 0.5 * ((u - ut) ** 2).sum()
    J.backward()
    _, g_ref = _dense_gradient(sol.problem, k_np, ut.numpy())
    assert np.abs(k.grad.numpy() - g_ref).max() <= 1e-8 * np.abs(g_ref).max()
    sol.close()


def test_torch_autograd_two_solves_one_backward():
    """Two BddcSolve applies before one backward: each backward sets its own
    coefficients up again when the solver holds the other factorization, so
    both gradients match the dense ones."""
    import torch
    from bddcml_amd.adjoint import AdjointSolver, BddcSolve
    _say("case adjoint torch autograd, two solves then one backward")
    nel, ns = 4, 2
    p0 = CubeProblem(nel, ns)
    k1_np = lognormal(5, np.arange(p0.nelem))
    k2_np = lognormal(8, np.arange(p0.nelem))
    ut = torch.from_numpy(2.0 * uniform01(6, np.arange(p0.ndof, dtype=np.uint64)) - 1.0)
    sol = AdjointSolver(_make(nel, ns), p0.kref, tol=1e-12)
    k1 = torch.tensor(k1_np, dtype=torch.float64, requires_grad=True)
    k2 = torch.tensor(k2_np, dtype=torch.float64, requires_grad=True)
    u1 = BddcSolve.apply(k1, sol)
    u2 = BddcSolve.apply(k2, sol)
    J = 0.5 * ((u1 - ut) ** 2).sum() + 0.5 * ((u2 - ut) ** 2).sum()
    J.backward()
    _, g1 = _dense_gradient(_make(nel, ns)(k1_np), k1_np, ut.numpy())
    _, g2 = _dense_gradient(_make(nel, ns)(k2_np), k2_np, ut.numpy())
    assert np.abs(g1 - g2).max() > 1e-3 * np.abs(g1).max()
    assert np.abs(k1.grad.numpy() - g1).max() <= 1e-8 * np.abs(g1).max()
    assert np.abs(k2.grad.numpy() - g2).max() <= 1e-8 * np.abs(g2).max()
    sol.close()
