"""Adaptive BDDC constraints on the CPU: numpy restatement of the reference's
face eigenproblems and of the loading of their eigenvectors as constraints.

TEST INFRASTRUCTURE ONLY (the checker of the library's adaptive setup); the
product never imports it.

For every face (a glob shared by exactly two subdomains i < j; pairs as in
dd_create_pairs, src/module_dd.f90:10785-11009) the reference solves
(src/module_adaptivity.f90:296-2515, operator 2560-3024)

    A v = lambda B v,   A = P (I - E)^T S (I - E) P,   B = P S P,

on the concatenated interfaces of i and j (problem size ndofi_i + ndofi_j):
S = diag(S_i, S_j); P = I - Q Q^T with Q an orthonormal basis of the columns
of D_ij = [C_i; -C_j] restricted to the constraint rows the two subdomains
share (its QR, :970-1010; the null projection :3068-3094); E = R R^T W, W the
pair weights rho_i / (rho_i + rho_j) on the dofs common to i and j and 1
elsewhere, R R^T sums the two copies of a common dof into both
(adaptivity_apply_R/RT :3115-3155).  The reference runs LOBPCG (BLOPEX) on
(-A, B); this restatement solves the same pencil densely (as the reference's
explicit-Schur branch does with DSYGV, :1500-1560), on the range of P with
the null space of B (floating subdomains: constants on both) taken out.

The neigvec = min(10, ndofcomm - ncommon_crows) largest eigenvalues are kept
(:1185-1195); those >= threshold give constraints c = -P (I-E^T) S (I-E) P v
(:1900-1935), whose i part and (negated) j part are loaded on the face of
each subdomain, restricted to the face's dofs (dd_load_adaptive_constraints,
src/module_dd.f90:4386-4457), after the face's arithmetic rows.  Every face
block is then orthonormalized by QR with column pivoting, keeping the
columns whose |R_kk| >= 0.01 |R_11| (dd_orthogonalize_constraints,
src/module_dd.f90:4637-4802); coarse dofs are renumbered node by node
(corners, edges, faces) with the new counts per face.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

NEIGVECX = 10  # src/module_adaptivity.f90:46


def _rho(o, d, isub, weights_type, dof_data):
    """Weight coefficients of subdomain isub at its interface dofs
    (dd_get_my_coefficients_for_weights, src/module_dd.f90:11660-11725)."""
    gd = o.sub_interface_gdofs(isub)
    if weights_type == 0:
        return np.ones(len(gd))
    b = 1 if d["i_sparse"].min() >= 1 else 0
    l_of = {int(g): k for k, g in enumerate(d["isvgvn"] - b)}
    loc = np.array([l_of[int(g)] for g in gd])
    if weights_type == 1:
        i, j, a = d["i_sparse"] - b, d["j_sparse"] - b, d["a_sparse"]
        dg = np.bincount(i[i == j], a[i == j], minlength=d["ndofs"])
        r = dg[loc].copy()
        r[r == 0.0] = 1.0
        return r
    if weights_type == 3:
        return np.asarray(dof_data[isub], dtype=float)[loc]
    raise ValueError(f"weights type {weights_type}")


def pair_problem(sub_i, sub_j):
    """Dense A, B (reference semantics), the projection basis Q, the pair
    weights and the common-dof map of one pair.  sub_*: dict with gd
    (interface global dofs), C (constraint rows), ind (their global coarse
    dofs), S (Schur complement), rho (weight coefficients)."""
    ni, nj = len(sub_i["gd"]), len(sub_j["gd"])
    n = ni + nj
    common = np.intersect1d(sub_i["ind"], sub_j["ind"])
    D = np.zeros((n, len(common)))
    for k, c in enumerate(common):
        D[:ni, k] = sub_i["C"][np.nonzero(sub_i["ind"] == c)[0][0]]
        D[ni:, k] = -sub_j["C"][np.nonzero(sub_j["ind"] == c)[0][0]]
    Q = np.linalg.qr(D)[0] if len(common) else np.zeros((n, 0))
    # common interface dofs: positions in i and in j
    pos_j = {int(g): k for k, g in enumerate(sub_j["gd"])}
    ci = np.array([k for k, g in enumerate(sub_i["gd"]) if int(g) in pos_j], dtype=int)
    cj = np.array([pos_j[int(sub_i["gd"][k])] for k in ci], dtype=int)
    w = np.ones(n)
    den = sub_i["rho"][ci] + sub_j["rho"][cj]
    w[ci] = sub_i["rho"][ci] / den
    w[ni + cj] = sub_j["rho"][cj] / den
    # E = R R^T W: on a common dof both copies get w_i x_i + w_j x_j
    E = np.eye(n)
    for a, b in zip(ci, ni + cj):
        E[a, :] = 0.0
        E[b, :] = 0.0
        E[a, a] = E[b, a] = w[a]
        E[a, b] = E[b, b] = w[b]
    S = sla.block_diag(sub_i["S"], sub_j["S"])
    P = np.eye(n) - Q @ Q.T
    IE = np.eye(n) - E
    A = P @ IE.T @ S @ IE @ P
    B = P @ S @ P
    return dict(A=A, B=B, P=P, Q=Q, S=S, IE=IE, w=w, ci=ci, cj=cj, ni=ni, nj=nj,
                ndofcomm=len(ci), ncommon=len(common))


def pair_eigen(pp):
    """All eigenpairs of the pencil (A, B) on the range of P where B is
    positive definite, largest first: (lambda, V with V^T B V = I)."""
    n = pp["A"].shape[0]
    # orthonormal basis of range(P)
    Qf = np.linalg.qr(pp["Q"], mode="complete")[0] if pp["Q"].shape[1] else np.eye(n)
    Z = Qf[:, pp["Q"].shape[1]:]
    Az = Z.T @ pp["A"] @ Z
    Bz = Z.T @ pp["B"] @ Z
    Az = 0.5 * (Az + Az.T)
    Bz = 0.5 * (Bz + Bz.T)
    bl, bv = np.linalg.eigh(Bz)
    keep = bl > 1e-12 * bl.max()
    T = bv[:, keep] / np.sqrt(bl[keep])
    lam, y = np.linalg.eigh(T.T @ Az @ T)
    order = np.argsort(lam)[::-1]
    return lam[order], Z @ (T @ y[:, order])


def orthonormal_rows(M):
    """Rows spanning the rows of M, orthonormal, by QR with column pivoting
    of M^T, keeping columns while |R_kk| >= 0.01 |R_11|
    (dd_orthogonalize_constraints, src/module_dd.f90:4637-4802)."""
    if M.shape[0] == 0:
        return M
    Qm, R, piv = sla.qr(M.T, mode="economic", pivoting=True)
    d = np.abs(np.diag(R))
    nvalid = 0
    if d.size and d[0] > 1e-12:
        for k in range(len(d)):
            if d[k] < 0.01 * d[0]:
                break
            nvalid = k + 1
    return Qm[:, :nvalid].T.copy()


def adaptive_constraints(o, problem, threshold=1.5, weights_type=0, dof_data=None,
                         neigvecx=NEIGVECX):
    """Compute the face eigenproblems of a set-up Oracle `o` (arithmetic
    constraints), load the adaptive constraints, orthonormalize every face
    block, renumber the coarse dofs and rebuild the oracle's coarse problem.
    Returns a report: per pair (face coarse node, i, j), the kept
    eigenvalues, the number of adaptive constraints and the pair estimate
    (src/module_adaptivity.f90:2150-2165), and the new coarse dof count."""
    counts = o.counts()
    nndfc, _ = o.coarse_nodes()
    kdofc = np.concatenate([[0], np.cumsum(nndfc)])
    first_face = counts["ncorner"] + counts["nedge"]
    subs = {}
    for d in problem.subdomains():
        isub = d["isub"]
        C, ind = o.sub_constraints(isub)
        subs[isub] = dict(gd=o.sub_interface_gdofs(isub), C=C, ind=ind, S=o.sub_schur(isub),
                          rho=_rho(o, d, isub, weights_type, dof_data),
                          node=np.searchsorted(kdofc, ind, side="right") - 1)
    face_subs = {}
    for isub, s in subs.items():
        for nd in np.unique(s["node"]):
            if nd >= first_face:
                face_subs.setdefault(int(nd), []).append(isub)
    # adaptive rows per (subdomain, face): over the subdomain's interface
    extra = {}
    report = []
    for f in sorted(face_subs):
        if len(face_subs[f]) != 2:
            continue
        i, j = sorted(face_subs[f])
        pp = pair_problem(subs[i], subs[j])
        neigvec = max(0, min(neigvecx, pp["ndofcomm"] - pp["ncommon"]))
        lam, V = pair_eigen(pp)
        lam, V = lam[:neigvec], V[:, :neigvec]
        nad = int(np.count_nonzero(lam >= threshold))
        est = (lam[nad] if nad < neigvec else lam[nad - 1]) if neigvec else 1.0
        cons = -(pp["P"] @ pp["IE"].T @ pp["S"] @ pp["IE"] @ pp["P"] @ V[:, :nad])
        ni = pp["ni"]
        extra[(i, f)] = cons[:ni].T
        extra[(j, f)] = -cons[ni:].T
        report.append(dict(face=f, isub=i, jsub=j, eigenvalues=lam, nadaptive=nad, est=float(est),
                           ndofi=(ni, pp["nj"]), ncommon=pp["ncommon"], ndofcomm=pp["ndofcomm"]))
    # new rows per coarse node of every subdomain; faces orthonormalized
    rows_of = {}
    new_nndf = np.array(nndfc, dtype=int)
    for isub, s in subs.items():
        for nd in np.unique(s["node"]):
            r = np.nonzero(s["node"] == nd)[0]
            M = s["C"][r]
            if nd >= first_face:
                sup = np.nonzero(np.any(M != 0.0, axis=0))[0]
                blk = M[:, sup]
                if (isub, int(nd)) in extra:
                    blk = np.vstack([blk, extra[(isub, int(nd))][:, sup]])
                q = orthonormal_rows(blk)
                M = np.zeros((q.shape[0], M.shape[1]))
                M[:, sup] = q
                new_nndf[nd] = q.shape[0]
            rows_of[(isub, int(nd))] = M
    # both subdomains of a face must agree on its count
    for f, ss in face_subs.items():
        ks = {rows_of[(s_, f)].shape[0] for s_ in ss}
        if len(ks) != 1:
            raise RuntimeError(f"face {f}: subdomains disagree on the number of constraints {ks}")
    new_kdofc = np.concatenate([[0], np.cumsum(new_nndf)])
    ndofc = int(new_kdofc[-1])
    for isub, s in subs.items():
        Cs, inds = [], []
        for nd in np.unique(s["node"]):
            M = rows_of[(isub, int(nd))]
            Cs.append(M)
            inds.append(new_kdofc[nd] + np.arange(M.shape[0]))
        o.set_constraints(isub, np.vstack(Cs), np.concatenate(inds).astype(np.int32), ndofc)
    o.refresh_coarse()
    return dict(pairs=report, ndofc=ndofc, nndfc=new_nndf,
                est=max([p["est"] for p in report], default=-1.0))
