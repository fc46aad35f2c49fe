"""Dense numpy algebra used to pin the CPU oracle on small cubes.

Everything here is computed from the assembled subdomain matrices with
textbook dense linear algebra (numpy.linalg), independent of the oracle's
skyline / LU code paths.
"""
import numpy as np


def sub_dense(d):
    """Dense subdomain matrix after Dirichlet elimination that keeps the
    original diagonal (reference src/module_sm.f90:431-600), 0-based."""
    b = 1 if d["isub"] >= 1 and d["i_sparse"].min() >= 1 else 0
    n = d["ndofs"]
    A = np.zeros((n, n))
    i = d["i_sparse"] - b
    j = d["j_sparse"] - b
    a = d["a_sparse"]
    np.add.at(A, (i, j), a)
    off = i != j
    np.add.at(A, (j[off], i[off]), a[off])
    fixed = d["ifix"] > 0
    dg = np.diag(A).copy()
    A[fixed, :] = 0.0
    A[:, fixed] = 0.0
    A[fixed, fixed] = dg[fixed]
    return A


def schur(A, idx_i, idx_o):
    A22 = A[np.ix_(idx_i, idx_i)]
    if len(idx_o) == 0:
        return A22
    A11 = A[np.ix_(idx_o, idx_o)]
    A12 = A[np.ix_(idx_o, idx_i)]
    return A22 - A12.T @ np.linalg.solve(A11, A12)


def kkt_blocks(S, C):
    """Inverse blocks of [S C^T; C 0]: K (interface block), Phi, Ac = -Lambda."""
    n, m = S.shape[0], C.shape[0]
    M = np.zeros((n + m, n + m))
    M[:n, :n] = S
    M[:n, n:] = C.T
    M[n:, :n] = C
    Minv = np.linalg.inv(M)
    K = Minv[:n, :n]
    Phi = Minv[:n, n:]
    Ac = -Minv[n:, n:]
    return K, Phi, Ac


def bddc_dense(orc, problem):
    """Assemble the BDDC preconditioner M and the interface Schur complement
    S as dense matrices over the oracle's unique interface numbering, using
    numpy for every solve; the oracle supplies only the combinatorial data
    (interface dofs, weights, constraint rows, coarse numbering)."""
    ig = orc.interface_dofs()
    pos = {int(g): k for k, g in enumerate(ig)}
    nig = len(ig)
    ndofc = orc.counts()["ndofc"]
    Sg = np.zeros((nig, nig))
    parts = []
    for d in problem.subdomains():
        isub = d["isub"]
        A = sub_dense(d)
        gd = orc.sub_interface_gdofs(isub)
        b = problem.numbase
        gloc = d["isvgvn"] - b
        l_of = {int(g): k for k, g in enumerate(gloc)}
        idx_i = np.array([l_of[int(g)] for g in gd])
        idx_o = np.setdiff1d(np.arange(d["ndofs"]), idx_i)
        S = schur(A, idx_i, idx_o)
        Cm, indrowc = orc.sub_constraints(isub)
        w = orc.sub_weights(isub)
        R = np.zeros((len(gd), nig))
        R[np.arange(len(gd)), [pos[int(g)] for g in gd]] = 1.0
        Rc = np.zeros((len(indrowc), ndofc))
        Rc[np.arange(len(indrowc)), indrowc] = 1.0
        Sg += R.T @ S @ R
        K, Phi, Ac = kkt_blocks(S, Cm)
        parts.append(dict(isub=isub, R=R, Rc=Rc, W=np.diag(w), K=K, Phi=Phi, Ac=Ac, S=S, C=Cm))
    Acg = sum(p["Rc"].T @ p["Ac"] @ p["Rc"] for p in parts)
    Acinv = np.linalg.inv(Acg)
    Mg = np.zeros((nig, nig))
    for p in parts:
        RW = p["W"] @ p["R"]
        Mg += RW.T @ p["K"] @ RW
    coarse = sum(p["Rc"].T @ p["Phi"].T @ p["W"] @ p["R"] for p in parts)
    Mg += coarse.T @ Acinv @ coarse
    return Sg, Mg, parts, Acg


def bicgstab_dense(S, M, g, tol, maxit=500, ndecrmax=50):
    """BiCGSTAB (van der Vorst 1992) on dense S and M, in the step order of
    the reference src/module_krylov.f90:1101-1455: zero initial guess,
    residuals relative to the initial one, the stopping test and the
    non-decrease counter after the half step (1271) and after the full step
    (1412).  Returns (iterations, reason, relative residual after every half
    and full step, solution)."""
    u = np.zeros_like(g)
    r = g - S @ u
    normrhs = np.linalg.norm(r)
    hist = []
    rhoold = alpha = omega = 1.0
    v = np.zeros_like(g)
    p = np.zeros_like(g)
    rs = r.copy()
    ndecr, lastres = 0, 1.0
    for it in range(1, maxit + 1):
        rho = r @ rs
        beta = rho * alpha / (rhoold * omega)
        p = r + beta * (p - omega * v)
        y = M @ p
        v = S @ y
        alpha = rho / (v @ rs)
        u = u + alpha * y
        s = r - alpha * v
        relres = np.linalg.norm(s) / normrhs
        hist.append(relres)
        if relres < tol:
            return it, 0, np.array(hist), u
        if relres < lastres:
            ndecr = 0
        else:
            ndecr += 1
            if ndecr >= ndecrmax:
                return it, -2, np.array(hist), u
        lastres = relres
        z = M @ s
        t = S @ z
        omega = (s @ t) / (t @ t)
        u = u + omega * z
        r = s - omega * t
        relres = np.linalg.norm(r) / normrhs
        hist.append(relres)
        if relres < tol:
            return it, 0, np.array(hist), u
        if it == maxit:
            return it, -1, np.array(hist), u
        if relres < lastres:
            ndecr = 0
        else:
            ndecr += 1
            if ndecr >= ndecrmax:
                return it, -2, np.array(hist), u
        lastres = relres
        rhoold = rho
    return maxit, -1, np.array(hist), u


def weights_closed_form(problem, orc, weights_type, dof_data=None):
    """Interface weights w_i = rho_i / sum_j rho_j over the subdomains j
    sharing a dof (src/module_dd.f90:11632-11641), with rho the diagonal of
    the subdomain's own (unassembled) matrix for type 1 (zeros replaced by 1;
    src/module_dd.f90:11881-11934, 11690-11697) or the subdomain's dof_data
    for type 3 (src/module_dd.f90:12020-12060).  Computed from the uploaded
    triplets with numpy; the oracle supplies only the interface dof lists.
    Returns {isub: w in the oracle's interface order}."""
    rho, tot = {}, np.zeros(problem.ndof)
    b = problem.numbase
    for d in problem.subdomains():
        isub = d["isub"]
        gd = orc.sub_interface_gdofs(isub)
        l_of = {int(gg): k for k, gg in enumerate(d["isvgvn"] - b)}
        loc = np.array([l_of[int(gg)] for gg in gd])
        if weights_type == 1:
            dg = np.diag(sub_dense(d))
            r = dg[loc].copy()
            r[r == 0.0] = 1.0
        elif weights_type == 3:
            r = np.asarray(dof_data[isub], dtype=float)[loc]
        else:
            r = np.ones(len(loc))
        rho[isub] = r
        np.add.at(tot, gd, r)
    return {isub: rho[isub] / tot[orc.sub_interface_gdofs(isub)] for isub in rho}
