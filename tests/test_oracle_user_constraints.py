"""Pins the oracle's user constraints (oracle/user_constraints.py) and checks
the library's host plan against it, on the CPU.

- Dependent user values (a constant per node: each user row repeats an
  arithmetic row) are dropped by the orthogonalization: the same coarse
  dofs, and the preconditioner equals the arithmetic-constraint one (BDDC
  depends only on the span of each coarse node's rows).
- Values equal to the node's x coordinate: every edge along x gets exactly
  the span {1, x} per component (first moments along the edge), every other
  edge only {1} (x constant along it): counts and spans checked against that
  closed form.
- Effect and correctness: more constraints never raise the condition
  estimate, M equals the dense BDDC formula on the oracle's rows, and the
  solution equals the dense direct solve.
- The library's host plan (plan only) has the same coarse dofs as the oracle.
"""
import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import CubeProblem
from oracle.user_constraints import user_constraints
from tests.dense_ref import bddc_dense


def _setup(oracle_lib, p, wt=0):
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, wt)
    return o


def _values(p, kind):
    out = {}
    for d in p.subdomains():
        xyz = np.asarray(d["xyz"], dtype=float).reshape(d["nnods"], -1, order="F")
        if kind == "const":
            out[d["isub"]] = np.ones((1, d["nnods"]))
        elif kind == "x":
            out[d["isub"]] = xyz[:, 0][None, :]
        else:  # two consistent pseudo-random values per global node
            g = np.asarray(d["isngn"], dtype=float)
            out[d["isub"]] = np.vstack([np.sin(1.7 * g), np.cos(0.3 * g * g)])
    return out


def test_dependent_values_change_nothing(oracle_lib):
    p = CubeProblem(3, 2, kind="elasticity", rhs="random", seed=2)
    o1 = _setup(oracle_lib, p)
    o2 = _setup(oracle_lib, p)
    rep = user_constraints(o2, p, _values(p, "const"))
    assert rep["ndofc"] == o1.counts()["ndofc"]
    x = np.random.default_rng(0).standard_normal(o1.counts()["nig"])
    z1, z2 = o1.apply_pc(x), o2.apply_pc(x)
    assert np.abs(z1 - z2).max() <= 1e-11 * np.abs(z1).max()
    o1.close(), o2.close()


def test_first_moments_along_x(oracle_lib):
    p = CubeProblem(3, 3, kind="poisson", rhs="random", seed=2)
    o = _setup(oracle_lib, p)
    rep = user_constraints(o, p, _values(p, "x"))
    # coordinates of every global dof (Poisson: dof = node)
    xg = {}
    for d in p.subdomains():
        xyz = np.asarray(d["xyz"], dtype=float).reshape(d["nnods"], -1, order="F")
        for ln, g in enumerate(np.asarray(d["isvgvn"]) - 1):
            xg[int(g)] = xyz[ln]
    nedge = 0
    for e, q in rep["edge_rows"].items():
        pts = np.array([xg[int(g)] for g in rep["edge_gdofs"][e]])
        along_x = np.ptp(pts[:, 0]) > 0
        basis = [np.ones(len(pts))] + ([pts[:, 0]] if along_x else [])
        assert q.shape[0] == len(basis), e
        np.testing.assert_allclose(q @ q.T, np.eye(q.shape[0]), atol=1e-13)
        for v in basis:
            assert np.abs(v - q.T @ (q @ v)).max() <= 1e-12 * np.abs(v).max()
        nedge += along_x
    assert nedge > 0
    o.close()


@pytest.mark.parametrize("kind,values", [("poisson", "x"), ("elasticity", "random")])
def test_effect_and_solution(oracle_lib, kind, values):
    p = CubeProblem(3, 2, kind=kind, rhs="random", seed=3)
    o = _setup(oracle_lib, p)
    it0, r0, c0 = o.solve(0, 1e-10, 500, 50)
    rep = user_constraints(o, p, _values(p, values))
    assert rep["ndofc"] > 0
    it1, r1, c1 = o.solve(0, 1e-10, 500, 50)
    assert r0 == 0 and r1 == 0
    assert c1 <= c0 * (1 + 1e-10) and it1 <= it0
    Sg, Mg, _, _ = bddc_dense(o, p)
    x = np.random.default_rng(1).standard_normal(Sg.shape[0])
    z = o.apply_pc(x)
    assert np.abs(z - Mg @ x).max() <= 1e-10 * np.abs(z).max()
    A, f, fixed = p.global_dense()
    u = np.linalg.solve(A, f)
    assert np.abs(o.interface_solution() - u[o.interface_dofs()]).max() <= 1e-9 * np.abs(u).max()
    o.close()


@pytest.mark.parametrize("kind,values", [("poisson", "x"), ("elasticity", "random"),
                                         ("elasticity", "const")])
def test_library_plan_matches_oracle(oracle_lib, kind, values):
    p = CubeProblem(3, 2, kind=kind, rhs="random", seed=3)
    vals = _values(p, values)
    o = _setup(oracle_lib, p)
    rep = user_constraints(o, p, vals)
    o.close()
    b = bddcml_amd.Bddcml()
    try:
        b.init(2, [p.nsub, 1], numbase=p.numbase)
        b.set_option("plan_only", 1)
        for d in p.subdomains():
            d = dict(d)
            d["user_constraints"] = vals[d["isub"]]
            b.upload_subdomain_data(d)
        b.setup_preconditioner(1, 0, 1, 1, 1, 0, 1, 0)
        assert b.counts()["ndofc"] == rep["ndofc"]
    finally:
        b.finalize()
