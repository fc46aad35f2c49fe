"""Global loading on the device: the whole mesh uploaded at once
(bddcml_upload_global_data_c), divided by the library, solved, and new
global data (bddcml_change_global_data_c + bddcml_setup_new_data_c),
against the oracle on the same division restated on the CPU
(oracle/global_loading.py):
- bisection division: iterations +-1, interface solution within 1e-10
  relative, the global solution (bddcml_download_global_solution_c) within
  1e-8 of a dense direct solve of the assembled global system;
- new right-hand side and Dirichlet values on the kept factorization: the
  same bars against the oracle on the new data;
- division file (partition_l1.ES): the cube's own division gives the same
  solve as uploading its subdomains one by one (iterations equal, solution
  within 1e-12).
"""
import os
import sys

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import CubeProblem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _direct(p, g):
    A = p.global_dense(eliminate=False)[0]  # full symmetric matrix
    fixed = np.asarray(g["ifix"]) > 0
    gv = np.asarray(g["fixv"])
    u = np.zeros(p.ndof)
    u[fixed] = gv[fixed]
    free = ~fixed
    rhs = g["rhs"][free] - A[np.ix_(free, fixed)] @ gv[fixed]
    u[free] = np.linalg.solve(A[np.ix_(free, free)], rhs)
    return u


@pytest.mark.parametrize("nel,ns,kind,nsub", [(3, 2, "poisson", 5), (2, 2, "elasticity", 6)])
def test_global_loading_matches_oracle(oracle_lib, tmp_path, nel, ns, kind, nsub):
    from oracle.global_loading import divide, localize, read_elm
    _say(f"case global_loading {kind} nel={nel} ns={ns} nsub={nsub}")
    p = CubeProblem(nel, ns, kind=kind, rhs="random", seed=7, dirichlet="random",
                    coef=lambda e: 1.0 + (np.asarray(e) % 4))
    path = str(tmp_path / "cube.ELM")
    p.write_elm(path)
    g = p.global_data()
    elm = read_elm(path, p.nelem)
    subs = localize(g, elm, divide(g, nsub), nsub)
    tol = 1e-10
    o = oracle_lib.Oracle(nsub, p.nnod, p.ndof, 3, 3, 1)
    for d in subs:
        o.upload(d)
    o.setup(1, 1, 0)
    it_o, r_o, _ = o.solve(0, tol, 300, 50)
    assert r_o == 0
    u_o = o.interface_solution()

    b = bddcml_amd.Bddcml()
    b.init(2, [nsub, 1], numbase=1)
    fd = os.open(path, os.O_RDONLY)
    b.upload_global_data(g, fd, neighbouring=1)
    os.close(fd)
    b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
    it, r, _ = b.solve(0, tol, 300, 50)
    _say(f"  device {it}, oracle {it_o} iterations")
    assert r == 0 and abs(it - it_o) <= 1
    assert _relmax(b.interface_solution(), u_o) < 1e-10
    assert _relmax(b.download_global_solution(p.ndof), _direct(p, g)) < 1e-8

    _say(f"case global_loading {kind} new global data")
    q = CubeProblem(nel, ns, kind=kind, rhs="random", seed=8, dirichlet="random")
    gq = q.global_data()
    b.change_global_data(g["ifix"], gq["fixv"], gq["rhs"], gq["sol"])
    b.L.bddcml_setup_new_data_c()
    it, r, _ = b.solve(0, tol, 300, 50)
    g2 = dict(g, fixv=gq["fixv"], rhs=gq["rhs"])
    o2 = oracle_lib.Oracle(nsub, p.nnod, p.ndof, 3, 3, 1)
    for d in localize(g2, elm, divide(g2, nsub), nsub):
        o2.upload(d)
    o2.setup(1, 1, 0)
    it_o, r_o, _ = o2.solve(0, tol, 300, 50)
    _say(f"  device {it}, oracle {it_o} iterations")
    assert r == 0 and r_o == 0 and abs(it - it_o) <= 1
    assert _relmax(b.interface_solution(), o2.interface_solution()) < 1e-10
    assert _relmax(b.download_global_solution(p.ndof), _direct(p, g2)) < 1e-8
    b.finalize()
    o.close()
    o2.close()


def test_division_file_equals_subdomain_loading(tmp_path):
    _say("case global_loading partition_l1.ES")
    p = CubeProblem(3, 2, kind="poisson", rhs="random", seed=9)
    p.write_elm(str(tmp_path / "cube.ELM"))
    nb, ns, m = p.nel, p.nsub_edge, p.nel_sub
    e = np.arange(p.nelem)
    iets = (e % nb) // m + ((e // nb) % nb) // m * ns + (e // (nb * nb)) // m * ns * ns
    np.savetxt(str(tmp_path / "partition_l1.ES"), iets + 1, fmt="%d")
    tol = 1e-10
    b, (it_l, r_l, _) = bddcml_amd.solve_cube(p, tol=tol, maxit=300)
    u_l = b.interface_solution()
    b.finalize()
    cwd = os.getcwd()
    os.chdir(str(tmp_path))
    try:
        b = bddcml_amd.Bddcml()
        b.init(2, [p.nsub, 1], numbase=1)
        fd = os.open("cube.ELM", os.O_RDONLY)
        b.upload_global_data(p.global_data(), fd, neighbouring=1, load_division=1)
        os.close(fd)
    finally:
        os.chdir(cwd)
    b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
    it, r, _ = b.solve(0, tol, 300, 50)
    _say(f"  global {it}, local {it_l} iterations")
    assert r == 0 and r_l == 0 and it == it_l
    assert _relmax(b.interface_solution(), u_l) < 1e-12
    b.finalize()
