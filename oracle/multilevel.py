"""Multilevel BDDC composed from CPU oracle instances (one per level).

TEST INFRASTRUCTURE ONLY: the checker for the library's multilevel path.
The product (bddcml_amd) never imports this module.

Level 2 restated from the reference (paths relative to its root):
- its elements are level 1's subdomains, its nodes level 1's coarse nodes
  with their coarse dofs and coordinates (src/module_levels.f90:1484-1530,
  src/module_dd.f90:4941-4992), its element matrices the subdomain coarse
  matrices Ac_i, rows in the order of the subdomain's constraints
  (dd_gather_matrix, src/module_dd.f90:1382-1842);
- the division into subdomains: the reference calls METIS
  (pp_divide_mesh); here the level-1 subdomain centroids are bisected
  recursively (:func:`bisect`), the same rule as the library's;
- each subdomain's nodes sorted by global id, its elements ascending
  (pp_create_submesh, src/module_pp.f90:1443-1503); no Dirichlet data,
  complete right-hand sides, neighbouring 1 (src/module_levels.f90:1502-1506);
- in the preconditioner, one application of level 2's BDDC replaces the
  coarse solve (src/module_levels.f90:3849-4249; orc_apply_level).
"""
from __future__ import annotations

import numpy as np

from .oracle import COARSE_FN, Oracle


def bisect(cent: np.ndarray, nparts: int) -> np.ndarray:
    """Recursive coordinate bisection of the rows of ``cent`` into
    ``nparts`` parts: the axis of largest extent (the first on ties), points
    ordered by (coordinate, index), the first ``n * (p // 2) // p`` to the
    lower half, numbered depth first."""
    n = cent.shape[0]
    part = np.full(n, -1, dtype=np.int64)

    def rec(ids, p, base):
        if p == 1:
            part[ids] = base
            return
        c = cent[ids]
        ext = c.max(axis=0) - c.min(axis=0)
        ax = 0
        for d in range(1, c.shape[1]):
            if ext[d] > ext[ax]:
                ax = d
        ids = ids[np.lexsort((ids, c[:, ax]))]
        pl = p // 2
        nl = len(ids) * pl // p
        rec(ids[:nl], pl, base)
        rec(ids[nl:], p - pl, base + pl)

    rec(np.arange(n), nparts, 0)
    return part


def _seq_mean(rows: np.ndarray) -> np.ndarray:
    """Mean of the rows summed one after another (the library's rounding)."""
    acc = np.zeros(rows.shape[1])
    for r in rows:
        acc = acc + r
    return acc / rows.shape[0]


def level2_subdomains(o1: Oracle, nsub1: int, nsub2: int, numbase: int = 1,
                      find_components: int = 1, use_dual_mesh_graph: int = 0):
    """The second level's subdomain data (upload dictionaries, 1-based like
    the first level's when numbase is 1) built from a set-up first level.
    Returns (subdomains, part, nnodc, ndofc)."""
    cnt = o1.counts()
    nnodc = cnt["ncorner"] + cnt["nedge"] + cnt["nface"]
    ndofc = cnt["ndofc"]
    nndfc, xyzc = o1.coarse_nodes()
    kdofc = np.concatenate([[0], np.cumsum(nndfc)]).astype(np.int64)
    cn = [o1.sub_cnodes(i + numbase) for i in range(nsub1)]
    cent = np.array([_seq_mean(xyzc[np.sort(c)]) for c in cn])
    part = bisect(cent, nsub2)
    subs = []
    b = numbase
    for j in range(nsub2):
        elems = np.flatnonzero(part == j)
        if len(elems) == 0:
            raise ValueError(f"second-level subdomain {j} has no elements")
        nodes = np.unique(np.concatenate([cn[e] for e in elems]))
        loc = {int(g): k for k, g in enumerate(nodes)}
        inet = [loc[int(c)] for e in elems for c in cn[e]]
        nnet = [len(cn[e]) for e in elems]
        nndf = nndfc[nodes]
        isvgvn = np.concatenate([kdofc[g] + np.arange(nndfc[g]) for g in nodes]).astype(np.int64)
        ii, jj, aa = [], [], []
        for e in elems:
            ac = o1.sub_coarse(e + b)
            _, ind = o1.sub_constraints(e + b)
            dl = np.searchsorted(isvgvn, ind)
            assert np.all(isvgvn[dl] == ind)
            for a in range(len(ind)):
                for c in range(len(ind)):
                    if dl[a] <= dl[c]:
                        ii.append(dl[a])
                        jj.append(dl[c])
                        aa.append(ac[a, c])
        ndofs = int(nndf.sum())
        subs.append(dict(
            isub=j + b, nelems=len(elems), nnods=len(nodes), ndofs=ndofs,
            inet=np.asarray(inet) + b, nnet=np.asarray(nnet), nndf=nndf,
            isngn=nodes + b, isvgvn=isvgvn + b, xyz=xyzc[nodes],
            ifix=np.zeros(ndofs, dtype=np.int32), fixv=np.zeros(ndofs), rhs=np.zeros(ndofs),
            is_rhs_complete=1, sol=np.zeros(ndofs), matrixtype=1,
            i_sparse=np.asarray(ii) + b, j_sparse=np.asarray(jj) + b, a_sparse=np.asarray(aa),
            is_assembled=0, find_components=find_components,
            use_dual_mesh_graph=use_dual_mesh_graph, neighbouring=1))
    return subs, part, nnodc, ndofc


class MultiLevels:
    """A chain of oracle instances, one per level below the last:
    ``levels[0]`` holds the uploaded problem, each next one the mesh built
    from its predecessor's coarse nodes and coarse matrices, and every
    instance's coarse solve is one application of the next one's BDDC
    (the last one's a direct solve).  ``nsublev`` as bddcml_init: subdomains
    per level, the last 1; a level of one subdomain solves its coarse problem
    exactly (all its dofs are interior), as the direct solve does.  Solve and
    probes through ``o1``."""

    def __init__(self, subdomains, nsub, nnod, ndof, nsublev, numbase=1, use_corner=1,
                 use_arith=1, weights_type=0):
        subdomains = list(subdomains)
        if nsublev[0] != nsub or nsublev[-1] != 1:
            raise ValueError("nsublev must start with nsub and end with 1")
        self.levels = [Oracle(nsub, nnod, ndof, 3, 3, numbase)]
        for d in subdomains:
            self.levels[0].upload(d)
        self.levels[0].setup(use_corner, use_arith, weights_type)
        d0 = subdomains[0]
        self.subs = [subdomains]
        self.parts, self.ndofcs, self._fns = [], [], []
        for l in range(1, len(nsublev) - 1):
            prev = self.levels[-1]
            subs, part, nnodc, ndofc = level2_subdomains(
                prev, nsublev[l - 1], nsublev[l], numbase, d0["find_components"],
                d0["use_dual_mesh_graph"])
            nxt = Oracle(nsublev[l], nnodc, ndofc, 3, 3, numbase)
            for d in subs:
                nxt.upload(d)
            nxt.setup(use_corner, use_arith, weights_type)
            self._link(prev, nxt, ndofc)
            self.levels.append(nxt)
            self.subs.append(subs)
            self.parts.append(part)
            self.ndofcs.append(ndofc)
        self.o1 = self.levels[0]

    def _link(self, prev, nxt, ndofc):
        def coarse(ctx, rc, uc):
            r = np.ctypeslib.as_array(rc, shape=(ndofc,)).copy()
            np.ctypeslib.as_array(uc, shape=(ndofc,))[:] = nxt.apply_level(r)

        fn = COARSE_FN(coarse)  # kept alive as long as prev may call it
        self._fns.append(fn)
        prev.L.orc_set_coarse_solver(prev.h, fn, None)

    def solve(self, method=0, tol=1e-8, maxit=500, ndecrmax=50):
        return self.o1.solve(method, tol, maxit, ndecrmax)

    def close(self):
        for o in self.levels:
            if o.h:
                o.L.orc_set_coarse_solver(o.h, COARSE_FN(), None)
        for o in self.levels:
            o.close()


class ThreeLevels(MultiLevels):
    """Three levels, nsublev = [nsub, nsub2, 1]: ``o1`` and ``o2`` are the two
    instances, ``subs2``/``part``/``ndofc`` the second level's data."""

    def __init__(self, subdomains, nsub, nnod, ndof, nsub2, numbase=1, use_corner=1,
                 use_arith=1, weights_type=0):
        super().__init__(subdomains, nsub, nnod, ndof, [nsub, nsub2, 1], numbase, use_corner,
                         use_arith, weights_type)
        self.o2 = self.levels[1]
        self.subs2, self.part, self.ndofc = self.subs[1], self.parts[0], self.ndofcs[0]


def solve_levels(problem, nsublev, use_corner=1, use_arith=1, weights_type=0, method=0,
                 tol=1e-8, maxit=500, ndecrmax=50):
    """Multilevel solve of a CubeProblem / BoxProblem with nsublev as
    bddcml_init; returns the MultiLevels object (caller closes) and
    (num_iter, reason, cond)."""
    t = MultiLevels(problem.subdomains(), problem.nsub, problem.nnod, problem.ndof, nsublev,
                    problem.numbase, use_corner, use_arith, weights_type)
    return t, t.solve(method, tol, maxit, ndecrmax)


def solve_three_levels(problem, nsub2, use_corner=1, use_arith=1, weights_type=0, method=0,
                       tol=1e-8, maxit=500, ndecrmax=50):
    """Three-level solve of a CubeProblem / BoxProblem; returns the
    ThreeLevels object (caller closes) and (num_iter, reason, cond)."""
    t = ThreeLevels(problem.subdomains(), problem.nsub, problem.nnod, problem.ndof, nsub2,
                    problem.numbase, use_corner, use_arith, weights_type)
    return t, t.solve(method, tol, maxit, ndecrmax)
