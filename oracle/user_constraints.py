"""User constraints on edges on the CPU: numpy restatement of the reference's
loading of user constraints and of the orthogonalization of the edge blocks.

TEST INFRASTRUCTURE ONLY (the checker of the library's user constraints);
the product never imports it.

On the first level and on edges only (src/module_levels.f90:2709-2716), each
edge coarse node gets, after its arithmetic rows, one row per user
constraint and nodal component: the value user_constraints(iconstr, node) of
each node of the edge at that component's dof (dd_load_user_constraints,
src/module_dd.f90:4459-4565, appended by dd_append_cnode_constraits
:4567-4635; skipped when all values are zero).  The edge block is then
orthonormalized by QR with column pivoting keeping |R_kk| >= 0.01 |R_11|
(dd_orthogonalize_constraints, :4637-4802; oracle.adaptivity.orthonormal_rows)
and the coarse dofs are renumbered with the new counts.  The block is built
over the edge's dofs in global order from the lowest-numbered subdomain
holding the edge (the user's values agree on shared nodes) and given to
every subdomain of the edge.  Needs the arithmetic rows (use_arith = 1):
their supports give the edge's dofs and components.
"""
from __future__ import annotations

import numpy as np

from oracle.adaptivity import orthonormal_rows


def _dof_node_comp(d):
    """(local node, component) of every local dof of subdomain data d."""
    nndf = np.asarray(d["nndf"])
    node = np.repeat(np.arange(len(nndf)), nndf)
    comp = np.concatenate([np.arange(k) for k in nndf]) if len(nndf) else np.zeros(0, dtype=int)
    return node, comp


def user_constraints(o, problem, values):
    """Load user constraints into a set-up Oracle `o` (arithmetic
    constraints): values[isub] is the (nuser x nnods) array of subdomain
    isub (local nodes).  Renumbers the coarse dofs and rebuilds the coarse
    problem; returns dict(ndofc, nndfc, edge_rows={edge: rows over its sorted
    global dofs}, edge_gdofs)."""
    counts = o.counts()
    nndfc, _ = o.coarse_nodes()
    kdofc = np.concatenate([[0], np.cumsum(nndfc)])
    e0, e1 = counts["ncorner"], counts["ncorner"] + counts["nedge"]
    subs = {}
    for d in problem.subdomains():
        isub = d["isub"]
        C, ind = o.sub_constraints(isub)
        gd = o.sub_interface_gdofs(isub)
        b = 1 if np.asarray(d["isvgvn"]).min() >= 1 else 0
        l_of = {int(g): k for k, g in enumerate(np.asarray(d["isvgvn"]) - b)}
        node, comp = _dof_node_comp(d)
        loc = np.array([l_of[int(g)] for g in gd])
        subs[isub] = dict(C=C, ind=ind, gd=gd, node=node[loc], comp=comp[loc],
                          cnode=np.searchsorted(kdofc, ind, side="right") - 1,
                          vals=np.asarray(values[isub], dtype=float))
    # every edge: lowest-numbered subdomain holding it builds the block
    edge_rows, edge_gdofs = {}, {}
    for isub in sorted(subs):
        s = subs[isub]
        for e in np.unique(s["cnode"]):
            e = int(e)
            if not (e0 <= e < e1) or e in edge_rows:
                continue
            rows = np.nonzero(s["cnode"] == e)[0]
            sup = np.nonzero(np.any(s["C"][rows] != 0.0, axis=0))[0]
            sup = sup[np.argsort(s["gd"][sup], kind="stable")]
            M = [s["C"][r, sup] for r in rows]
            U = []
            for ic in range(s["vals"].shape[0]):
                for dcomp in range(int(s["comp"][sup].max()) + 1):
                    v = np.where(s["comp"][sup] == dcomp, s["vals"][ic, s["node"][sup]], 0.0)
                    U.append(v)
            if any(np.any(u != 0.0) for u in U):
                M += U
            edge_rows[e] = orthonormal_rows(np.array(M))
            edge_gdofs[e] = s["gd"][sup]
    new_nndf = np.array(nndfc, dtype=int)
    for e, q in edge_rows.items():
        new_nndf[e] = q.shape[0]
    new_kdofc = np.concatenate([[0], np.cumsum(new_nndf)])
    ndofc = int(new_kdofc[-1])
    for isub, s in subs.items():
        pos = {int(g): k for k, g in enumerate(s["gd"])}
        Cs, inds = [], []
        for nd in np.unique(s["cnode"]):
            nd = int(nd)
            if nd in edge_rows:
                q = edge_rows[nd]
                M = np.zeros((q.shape[0], len(s["gd"])))
                cols = [pos[int(g)] for g in edge_gdofs[nd]]
                M[:, cols] = q
            else:
                r = np.nonzero(s["cnode"] == nd)[0]
                M = s["C"][r]
            Cs.append(M)
            inds.append(new_kdofc[nd] + np.arange(M.shape[0]))
        o.set_constraints(isub, np.vstack(Cs), np.concatenate(inds).astype(np.int32), ndofc)
    o.refresh_coarse()
    return dict(ndofc=ndofc, nndfc=new_nndf, edge_rows=edge_rows, edge_gdofs=edge_gdofs)
