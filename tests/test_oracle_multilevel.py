"""Pins the three-level oracle (oracle/multilevel.py with orc_apply_level and
orc_set_coarse_solver) without the reference, which cannot be built here:

- the bisection: a known answer on a grid of points (blocks), ties broken
  by index;
- the second level's mesh: its assembled matrix equals the assembled
  first-level coarse matrix sum_i R_ci^T Ac_i R_ci, every coarse dof is on
  the second level, nodes sorted, coordinates of corners and globs;
- one second-level subdomain: the level apply is the exact coarse solve
  Ac^{-1}, so three levels reproduce two (iterations and solution);
- several second-level subdomains: the level operator X is symmetric and
  lambda(X Ac) >= 1 (BDDC's lower bound on the coarse problem), hence
  lambda_min(M3 S) >= 1 on the first level and cond(M3 S) >= cond(M2 S);
  the three-level PCG reaches the direct solution of the global system.
"""
import numpy as np
import pytest

from bddcml_amd.cube import CubeProblem
from oracle.multilevel import ThreeLevels, bisect, level2_subdomains


def test_bisect_known_answer():
    # 4 x 4 x 4 grid of points, x fastest: 8 parts are 2 x 2 x 2 blocks
    g = np.array([(x, y, z) for z in range(4) for y in range(4) for x in range(4)], float)
    part = bisect(g, 8)
    for p in range(8):
        pts = g[part == p]
        assert len(pts) == 8
        assert np.ptp(pts[:, 0]) == 1 and np.ptp(pts[:, 1]) == 1 and np.ptp(pts[:, 2]) == 1
    # the first cut halves the longest axis (all tie: x, the first), the
    # next ones y, then z
    assert set(part[g[:, 0] < 2]) == {0, 1, 2, 3}
    assert set(part[(g[:, 0] < 2) & (g[:, 1] < 2)]) == {0, 1}
    assert set(part[(g[:, 0] < 2) & (g[:, 1] < 2) & (g[:, 2] < 2)]) == {0}
    # a flat 4 x 4 x 2 grid: the short z axis is cut last
    flat = g[g[:, 2] < 2]
    pf = bisect(flat, 8)
    for p in range(8):
        pts = flat[pf == p]
        assert len(pts) == 4 and np.ptp(pts[:, 0]) == 0 and np.ptp(pts[:, 2]) == 1
    # equal coordinates: the lower indices go to the lower part
    same = np.zeros((6, 3))
    assert list(bisect(same, 2)) == [0, 0, 0, 1, 1, 1]
    # uneven counts: n * (p // 2) // p to the lower half
    assert np.bincount(bisect(np.arange(7.0)[:, None], 3)).tolist() == [2, 2, 3]


def _assembled_coarse(o1, nsub, ndofc):
    Ac = np.zeros((ndofc, ndofc))
    for i in range(nsub):
        a = o1.sub_coarse(i + 1)
        _, ind = o1.sub_constraints(i + 1)
        Ac[np.ix_(ind, ind)] += a
    return Ac


@pytest.mark.parametrize("kind,nsub2", [("poisson", 8), ("elasticity", 4)])
def test_second_level_mesh(kind, nsub2):
    p = CubeProblem(3, 4 if kind == "poisson" else 3, kind=kind, rhs="random", seed=3)
    t = ThreeLevels(p.subdomains(), p.nsub, p.nnod, p.ndof, nsub2, p.numbase)
    nc = t.ndofc
    Ac = _assembled_coarse(t.o1, p.nsub, nc)
    A2 = np.zeros((nc, nc))
    seen = np.zeros(nc, bool)
    for d in t.subs2:
        g = d["isvgvn"] - 1
        assert np.all(np.diff(d["isngn"]) > 0)          # nodes sorted by global id
        assert np.all(np.diff(g) > 0)
        seen[g] = True
        i, j = d["i_sparse"] - 1, d["j_sparse"] - 1
        assert np.all(i <= j)                            # upper triangle
        np.add.at(A2, (g[i], g[j]), d["a_sparse"])
        off = i != j
        np.add.at(A2, (g[j[off]], g[i[off]]), d["a_sparse"][off])
    assert seen.all()
    assert np.abs(A2 - Ac).max() <= 1e-13 * np.abs(Ac).max()
    # 4^3 (3^3) subdomains cut into blocks: equal element counts
    counts = np.bincount(t.part)
    assert counts.max() - counts.min() <= 1 if kind == "poisson" else counts.max() <= 9
    # coordinates: corners at subdomain vertices (multiples of 1/ns), face
    # globs at face centres (interior, one coordinate on a subdomain plane)
    nndf, xyz = t.o1.coarse_nodes()
    ns = p.nsub_edge
    c = t.o1.counts()
    corners = xyz[:c["ncorner"]] * ns
    assert np.abs(corners - np.round(corners)).max() < 1e-12 or kind == "poisson"
    faces = xyz[c["ncorner"] + c["nedge"]:] * ns
    on_plane = np.abs(faces - np.round(faces)) < 1e-12
    assert np.all(on_plane.sum(axis=1) >= 1)
    t.close()


def test_one_second_level_subdomain_is_exact():
    p = CubeProblem(3, 3, kind="poisson", rhs="random", seed=6)
    t = ThreeLevels(p.subdomains(), p.nsub, p.nnod, p.ndof, 1, p.numbase)
    nc = t.ndofc
    Ac = _assembled_coarse(t.o1, p.nsub, nc)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(nc)
    assert np.abs(t.o2.apply_level(r) - np.linalg.solve(Ac, r)).max() < 1e-12 * np.abs(
        np.linalg.solve(Ac, r)).max()
    it3, r3, cond3 = t.solve(0, 1e-10, 300, 50)
    u3 = t.o1.interface_solution()
    t.close()
    from oracle.oracle import solve_cube
    o, (it2, r2, cond2) = solve_cube(p, tol=1e-10, maxit=300)
    assert (it3, r3) == (it2, r2) == (it2, 0)
    assert abs(cond3 - cond2) < 1e-12 * cond2
    assert np.abs(u3 - o.interface_solution()).max() < 1e-13 * np.abs(u3).max()
    o.close()


@pytest.mark.parametrize("kind,ns,nsub2", [("poisson", 4, 8), ("elasticity", 3, 4)])
def test_several_second_level_subdomains(kind, ns, nsub2):
    p = CubeProblem(2, ns, kind=kind, rhs="random", seed=8)
    t = ThreeLevels(p.subdomains(), p.nsub, p.nnod, p.ndof, nsub2, p.numbase)
    nc = t.ndofc
    Ac = _assembled_coarse(t.o1, p.nsub, nc)
    X = np.array([t.o2.apply_level(e) for e in np.eye(nc)]).T
    assert np.abs(X - X.T).max() < 1e-12 * np.abs(X).max()
    L = np.linalg.cholesky(Ac)
    lam = np.linalg.eigvalsh(L.T @ X @ L)
    assert lam[0] > 1 - 1e-10
    assert lam[-1] > 1 + 1e-3                        # the level is inexact
    # first level: lambda_min(M3 S) >= 1, cond above the two-level one
    n = t.o1.counts()["nig"]
    E = np.eye(n)
    S = np.array([t.o1.apply_schur(e) for e in E]).T
    M3 = np.array([t.o1.apply_pc(e) for e in E]).T
    free = np.abs(np.diag(S)) > 0
    lam3 = np.sort(np.linalg.eigvals(M3[np.ix_(free, free)] @ S[np.ix_(free, free)]).real)
    assert lam3[0] > 1 - 1e-9
    it, r, cond = t.solve(0, 1e-12, 400, 50)
    assert r == 0
    assert abs(cond - lam3[-1]) < 1e-6 * lam3[-1]
    # direct solution of the global system
    K, f, _ = p.global_dense()
    u = np.linalg.solve(K, f)
    sol = np.zeros(p.ndof)
    for d in p.subdomains():
        sol[d["isvgvn"] - 1] = t.o1.local_solution(d["isub"])
    assert np.abs(sol - u).max() < 1e-9 * np.abs(u).max()
    t.close()


def test_four_levels():
    """Four levels: a third level of one subdomain solves the second level's
    coarse problem exactly, so [64, 8, 1, 1] reproduces three levels
    [64, 8, 1]; [64, 8, 2, 1] keeps lambda_min(M S) >= 1 and reaches the
    direct solution."""
    from oracle.multilevel import MultiLevels
    p = CubeProblem(2, 4, kind="poisson", rhs="random", seed=9)
    t3 = MultiLevels(p.subdomains(), p.nsub, p.nnod, p.ndof, [64, 8, 1], p.numbase)
    it3, r3, c3 = t3.solve(0, 1e-11, 300, 50)
    u3 = t3.o1.interface_solution()
    t3.close()
    t4 = MultiLevels(p.subdomains(), p.nsub, p.nnod, p.ndof, [64, 8, 1, 1], p.numbase)
    it4, r4, c4 = t4.solve(0, 1e-11, 300, 50)
    assert (it4, r4) == (it3, r3 == 0 and 0)
    assert abs(c4 - c3) < 1e-10 * c3
    assert np.abs(t4.o1.interface_solution() - u3).max() < 1e-12 * np.abs(u3).max()
    t4.close()
    t = MultiLevels(p.subdomains(), p.nsub, p.nnod, p.ndof, [64, 8, 2, 1], p.numbase)
    assert len(t.levels) == 3
    n = t.o1.counts()["nig"]
    E = np.eye(n)
    S = np.array([t.o1.apply_schur(e) for e in E]).T
    M = np.array([t.o1.apply_pc(e) for e in E]).T
    lam = np.sort(np.linalg.eigvals(M @ S).real)
    assert lam[0] > 1 - 1e-9
    it, r, cond = t.solve(0, 1e-12, 400, 50)
    assert r == 0 and it >= it3
    assert abs(cond - lam[-1]) < 1e-6 * lam[-1]
    K, f, _ = p.global_dense()
    u = np.linalg.solve(K, f)
    sol = np.zeros(p.ndof)
    for d in p.subdomains():
        sol[d["isvgvn"] - 1] = t.o1.local_solution(d["isub"])
    assert np.abs(sol - u).max() < 1e-9 * np.abs(u).max()
    t.close()
