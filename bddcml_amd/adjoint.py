"""Gradients of functionals of the solution with respect to element
coefficients, through the drop-in interface.

For A(k) u = f with A(k) = sum_e k_e A_e (A_e the element matrix per unit
coefficient, Dirichlet dofs eliminated) and J(u), the adjoint lam solves
A lam = dJ/du with homogeneous Dirichlet values, and

    dJ/dk_e = -lam_e^T A_e u_e .

The adjoint solve reuses the factorization and BDDC preconditioner of the
forward solve (A is symmetric): new right-hand side through
bddcml_change_subdomain_data_c + bddcml_setup_new_data_c, then
bddcml_solve_c.  The element products run in the HIP kernel behind
bddcml_amd_element_gradient.  ``BddcSolve`` wraps the forward solve as a
torch.autograd.Function.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Bddcml, _Checked, lib


class AdjointSolver:
    """Forward solves, adjoint solves and element gradients for a family of
    problems that differ only in their element coefficients.

    make_problem(k) must return a problem object (``bddcml_amd.cube``
    generators) whose element matrices are A_e scaled by k[isegn]; ``ae`` is
    A_e (nke x nke, row major in element dof order)."""

    def __init__(self, make_problem, ae, tol=1e-10, maxit=1000, precision=64):
        self.make_problem = make_problem
        self.ae = np.ascontiguousarray(ae, dtype=np.float64)
        self.tol = tol
        self.maxit = maxit
        self.precision = precision
        self.b = None
        self.subs = None
        self.problem = None
        self.k = None  # coefficients of the factorization held by self.b
        self.iterations = []

    def close(self):
        if self.b is not None:
            self.b.finalize()
            self.b = None
        self.k = None

    def forward(self, k):
        """Set up and solve for coefficients k (one per global element);
        returns the global solution."""
        self.close()
        k = np.array(k, dtype=np.float64)
        p = self.make_problem(k)
        self.problem = p
        self.subs = list(p.subdomains())
        b = Bddcml()
        b.init(2, [p.nsub, 1], numbase=p.numbase)
        b.set_option("precision", self.precision)
        for d in self.subs:
            b.upload_subdomain_data(d)
        b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
        it, reason, _ = b.solve(0, self.tol, self.maxit, 50)
        if reason != 0:
            raise RuntimeError(f"forward solve did not converge ({reason})")
        self.iterations.append(it)
        self.b = b
        self.k = k
        return b.download_global_solution(p.ndof)

    def holds(self, k):
        """True when the kept factorization is the one of coefficients k."""
        return self.k is not None and np.array_equal(self.k, np.asarray(k, dtype=np.float64))

    def adjoint(self, dJdu):
        """Solve A lam = dJ/du (homogeneous Dirichlet) on the factorization
        of the last forward solve; returns the global lam."""
        p, b = self.problem, self.b
        g = np.asarray(dJdu, dtype=np.float64)
        for d in self.subs:
            idx = d["isvgvn"] - p.numbase
            rhs = np.where(d["ifix"] > 0, 0.0, g[idx])
            b.change_subdomain_data(dict(isub=d["isub"], ifix=d["ifix"], fixv=np.zeros(len(idx)),
                                         rhs=rhs, is_rhs_complete=1, sol=np.zeros(len(idx))))
        b.setup_new_data()
        it, reason, _ = b.solve(0, self.tol, self.maxit, 50)
        if reason != 0:
            raise RuntimeError(f"adjoint solve did not converge ({reason})")
        self.iterations.append(it)
        return b.download_global_solution(p.ndof)

    def gradient(self, u, lam):
        """dJ/dk_e = -lam_e^T A_e u_e on the device."""
        p = self.problem
        L = _Checked(lib())
        L.bddcml_amd_set_option(b"throw_on_error", 1.0)
        dp = C.POINTER(C.c_double)
        u = np.ascontiguousarray(u, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros(p.nelem)
        L.bddcml_amd_element_gradient(self.ae.ctypes.data_as(dp), self.ae.shape[0], 0,
                                      u.ctypes.data_as(dp), lam.ctypes.data_as(dp),
                                      out.ctypes.data_as(dp))
        return out


def _torch_function():
    import torch

    class BddcSolve(torch.autograd.Function):
        """u = A(k)^{-1} f as a differentiable function of the element
        coefficients k (float64 tensor of nelem entries)."""

        @staticmethod
        def forward(ctx, k, solver):
            kk = k.detach().cpu().numpy().copy()
            u = solver.forward(kk)
            ctx.solver = solver
            ctx.k = kk
            ctx.u = u
            return torch.from_numpy(u).to(k.device)

        @staticmethod
        def backward(ctx, grad_u):
            # the solver holds one factorization: when another forward ran
            # since this one (several solves before one backward), set this
            # one's coefficients up again before the adjoint solve
            if not ctx.solver.holds(ctx.k):
                ctx.solver.forward(ctx.k)
            lam = ctx.solver.adjoint(grad_u.detach().cpu().numpy())
            g = ctx.solver.gradient(ctx.u, lam)
            return torch.from_numpy(g).to(grad_u.device), None

    return BddcSolve


def __getattr__(name):
    if name == "BddcSolve":
        return _torch_function()
    raise AttributeError(name)
