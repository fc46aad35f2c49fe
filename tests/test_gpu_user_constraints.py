"""User constraints on edges (bddcml_upload_subdomain_data_c user_constraints,
bddcml_setup_preconditioner_c use_user_constraints_int = 1) on the GPU
against the oracle's restatement (oracle/user_constraints.py): the same
coarse dofs, M r within 1e-10 (the rows of an edge span the same space; the
preconditioner depends only on the spans), PCG iterations within +-1, the
condition estimate within 1e-6 and the interface solution within 1e-10.
Values: the node's x coordinate (first moments along edges) on Poisson, two
pseudo-random values per global node on elasticity, and on elasticity with
weights of type 1."""
import sys

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import CubeProblem
from oracle.user_constraints import user_constraints

pytestmark = pytest.mark.gpu


def _values(p, kind):
    out = {}
    for d in p.subdomains():
        xyz = np.asarray(d["xyz"], dtype=float).reshape(d["nnods"], -1, order="F")
        if kind == "x":
            out[d["isub"]] = xyz[:, 0][None, :]
        else:
            g = np.asarray(d["isngn"], dtype=float)
            out[d["isub"]] = np.vstack([np.sin(1.7 * g), np.cos(0.3 * g * g)])
    return out


@pytest.mark.parametrize("nel,ns,kind,values,wt", [(3, 3, "poisson", "x", 0),
                                                   (3, 2, "elasticity", "random", 0),
                                                   (3, 2, "elasticity", "x", 1)])
def test_user_constraints_match_oracle(oracle_lib, nel, ns, kind, values, wt):
    print(f"case user constraints {kind} {nel}x{ns} values={values} wt={wt}", flush=True)
    sys.stdout.flush()
    p = CubeProblem(nel, ns, kind=kind, rhs="random", seed=3)
    vals = _values(p, values)
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, wt)
    rep = user_constraints(o, p, vals)
    it_o, r_o, c_o = o.solve(0, 1e-10, 500, 50)
    b = bddcml_amd.Bddcml()
    try:
        b.init(2, [p.nsub, 1], numbase=p.numbase)
        for d in p.subdomains():
            d = dict(d)
            d["user_constraints"] = vals[d["isub"]]
            b.upload_subdomain_data(d)
        b.setup_preconditioner(1, 0, 1, 1, 1, 0, 1, wt)
        assert b.counts()["ndofc"] == rep["ndofc"] == o.counts()["ndofc"]
        x = np.random.default_rng(4).standard_normal(o.counts()["nig"])
        z, z_ref = b.apply_pc(x), o.apply_pc(x)
        assert np.abs(z - z_ref).max() <= 1e-10 * np.abs(z_ref).max()
        it, reason, cond = b.solve(0, 1e-10, 500, 50)
        assert reason == 0 == r_o
        assert abs(it - it_o) <= 1
        assert abs(cond - c_o) <= 1e-6 * c_o
        u_ref = o.interface_solution()
        assert np.abs(b.interface_solution() - u_ref).max() <= 1e-10 * np.abs(u_ref).max()
    finally:
        b.finalize()
        o.close()
