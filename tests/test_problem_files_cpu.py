"""Problem files of the reference's file-driven drivers and its ParaView
export (bddcml_amd/problem_files.py), on the CPU.

- The list-directed reader follows the Fortran rules the reference's READ
  statements rely on (repeat counts, D exponents, slash, a READ starting on
  a new record), checked on hand-made files with known answers.
- Every format written by the library and read back gives the same arrays.
- A Fortran program (tests/drivers/problem_files_f.f90, amdflang) reads the
  library's files with the reference's READ sequence and writes them back
  with list-directed WRITE statements and an unformatted record; the
  library reads that output back to the same arrays.
- The VTU and PVD files equal, byte for byte, those of the reference's own
  module_paraview (compiled from the reference sources into oracle/_ref) for
  a subdomain of a cube with coordinates and values that exercise the E15.7
  format (negatives, tiny and huge magnitudes, rounding up to the next
  power of ten).
- The file-driven solve builds the same host plan as the oracle for the
  same cube (plan only).
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from bddcml_amd import problem_files as pf
from bddcml_amd.cube import CubeProblem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_list_directed_rules(tmp_path):
    p = tmp_path / "x.txt"
    p.write_text("1 2, 3*7\n 4 5 6\n8 /  9\n10 1.5D-03 -2.0d+2\n")
    r = pf.ListDirectedReader(str(p))
    assert list(r.read(4, kind=int)) == [1, 2, 7, 7]   # rest of line 1 discarded
    assert list(r.read(2, kind=int)) == [4, 5]
    # a slash ends the read: fewer values than asked is an error for a
    # required read and None for an optional one
    assert r.read(3, kind=int, optional=True) is None
    r2 = pf.ListDirectedReader(str(p))
    r2.pos = 3
    np.testing.assert_array_equal(r2.read(3), [10.0, 1.5e-3, -200.0])
    with pytest.raises(ValueError):
        r2.read(1)
    # PAR with and without the meshdim record
    (tmp_path / "a.PAR").write_text("3 8 64 125 125 0 512\n1.0D-08 500 50\n2\n")
    d = pf.read_par(str(tmp_path / "a.PAR"))
    assert d["ndim"] == 3 and d["linet"] == 512 and d["tol"] == 1e-8 and d["meshdim"] == 2
    (tmp_path / "b.PAR").write_text("3 8 64 125 125 0 512 1e-6 500 50 2\n")
    d = pf.read_par(str(tmp_path / "b.PAR"))
    assert d["meshdim"] == 3  # the value after ndecrmax is discarded with its record


def _problem():
    p = CubeProblem(3, 2, kind="elasticity", rhs="random", seed=4)
    g = p.global_data()
    rng = np.random.default_rng(1)
    g["fixv"] = np.where(np.asarray(g["ifix"]) != 0, rng.standard_normal(g["ndof"]), 0.0)
    return p, g


def _check_same(g, h):
    for k in ("inet", "nnet", "nndf", "ifix"):
        np.testing.assert_array_equal(np.asarray(g[k]), np.asarray(h[k]))
    for k in ("xyz", "fixv", "rhs"):
        np.testing.assert_array_equal(np.asarray(g[k], dtype=float), np.asarray(h[k], dtype=float))


def test_round_trip(tmp_path):
    p, g = _problem()
    pf.write_problem("cube", str(tmp_path), g, [p.nsub, 2, 1], tol=1e-7, maxit=321, ndecrmax=17)
    pr = pf.load_problem("cube", str(tmp_path))
    assert pr["nlevels"] == 3 and list(pr["nsublev"]) == [p.nsub, 2, 1]
    assert pr["par"]["tol"] == 1e-7 and pr["par"]["maxit"] == 321 and pr["par"]["ndecrmax"] == 17
    _check_same(g, pr["global_data"])
    # SOL, and an unformatted record split into subrecords
    sol, rea = np.arange(5.0), -np.arange(5.0)
    pf.write_sol(str(tmp_path / "c.SOL"), sol, rea, [1.0, 2.0, 3.0])
    s2, r2, sr = pf.read_sol(str(tmp_path / "c.SOL"), 5)
    np.testing.assert_array_equal(s2, sol), np.testing.assert_array_equal(r2, rea)
    np.testing.assert_array_equal(sr, [1.0, 2.0, 3.0])
    old = pf._SUB
    try:
        pf._SUB = 24
        pf.write_rhs(str(tmp_path / "s.RHS"), np.arange(10.0))
        assert len(pf.read_records(str(tmp_path / "s.RHS"))) == 1
        np.testing.assert_array_equal(pf.read_rhs(str(tmp_path / "s.RHS"), 10), np.arange(10.0))
    finally:
        pf._SUB = old


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not available")
def test_fortran_reads_and_writes_problem_files(tmp_path):
    exe = tmp_path / "pf_f"
    subprocess.check_call(["amdflang", "-O1", "-o", str(exe),
                           os.path.join(ROOT, "tests", "drivers", "problem_files_f.f90")], cwd=tmp_path)
    p, g = _problem()
    pf.write_problem("cube", str(tmp_path), g, [p.nsub, 1], tol=1e-9, maxit=77, ndecrmax=5)
    subprocess.check_call([str(exe), str(tmp_path / "cube"), str(tmp_path / "back")])
    par = pf.read_par(str(tmp_path / "back.PAR"))
    assert (par["nelem"], par["ndof"], par["nnod"], par["tol"], par["maxit"], par["ndecrmax"]) == \
        (g["nelem"], g["ndof"], g["nnod"], 1e-9, 77, 5)
    nl, nsublev = pf.read_ml(str(tmp_path / "back.ML"))
    assert nl == 2 and list(nsublev) == [p.nsub, 1]
    h = pf.read_gmis(str(tmp_path / "back.GMIS"), par["linet"], par["nelem"], par["nnod"], par["ndim"])
    h["ifix"], h["fixv"] = pf.read_fvs(str(tmp_path / "back.FVS"), par["ndof"])
    h["rhs"] = pf.read_rhs(str(tmp_path / "back.RHS"), par["ndof"])
    _check_same(g, h)


REF = os.path.join(ROOT, "oracle", "_ref", "paraview_ref")


@pytest.mark.skipif(not os.path.exists(REF), reason="reference ParaView writer not built (oracle/_ref)")
def test_vtu_matches_reference_writer(tmp_path):
    p = CubeProblem(3, 2, rhs="random", seed=5)
    d = [s for s in p.subdomains() if s["isub"] == 3][0]
    b = 1 if d["inet"].min() >= 1 else 0
    inet = np.asarray(d["inet"]) - b + 1
    nnet = np.asarray(d["nnet"])
    xyz = np.asarray(d["xyz"], dtype=float).reshape(d["nnods"], -1, order="F").copy()
    xyz[:, 0] -= 0.5  # negative coordinates
    rng = np.random.default_rng(2)
    sol = rng.standard_normal(d["nnods"]) * 10.0 ** rng.integers(-30, 30, d["nnods"])
    sol[:6] = [0.0, 0.99999996, -0.99999996, 1e-300, -1.5e150, 9.9999999e99]
    ref_dir, lib_dir = tmp_path / "ref", tmp_path / "lib"
    ref_dir.mkdir(), lib_dir.mkdir()
    (ref_dir / "sol").mkdir()
    with open(ref_dir / "in.txt", "w") as f:
        f.write("sol\n")
        f.write(f"{d['isub']} 8 {len(nnet)} {d['nnods']} 3 {xyz.shape[1]} {len(inet)}\n")
        for a in (inet, nnet):
            f.write(" ".join(str(int(v)) for v in a) + "\n")
        f.write(" ".join(repr(float(v)) for v in xyz.ravel(order="F")) + "\n")
        f.write(" ".join(repr(float(v)) for v in sol) + "\n")
    subprocess.run([REF, "in.txt"], cwd=ref_dir, check=True, capture_output=True)
    vtu = pf.write_subdomain_vtu("sol", d["isub"], inet, nnet, xyz, sol, meshdim=3, directory=str(lib_dir))
    pvd = pf.write_pvd("sol", 8, directory=str(lib_dir))
    for ours, theirs in ((vtu, ref_dir / pf.subdomain_vtu_name("sol", d["isub"])), (pvd, ref_dir / "sol.pvd")):
        a, b_ = open(ours).read(), open(theirs).read()
        assert a == b_, (ours, [(i, x, y) for i, (x, y) in enumerate(zip(a.splitlines(), b_.splitlines()))
                                if x != y][:5])


def test_plan_from_problem_files(oracle_lib, tmp_path):
    """The file-driven sequence builds the oracle's plan for the same cube
    (bisection division, plan only)."""
    from oracle.global_loading import divide, localize, read_elm
    p, g = _problem()
    nsub = 5
    pf.write_problem("cube", str(tmp_path), g, [nsub, 1])
    p.write_elm(str(tmp_path / "cube.ELM"))
    code = f"""
import sys, json
sys.path.insert(0, {ROOT!r})
from bddcml_amd import problem_files as pf
print(json.dumps(pf.solve_problem("cube", {str(tmp_path)!r}, plan_only=True)["counts"]))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    import json
    pc = json.loads(r.stdout.strip().splitlines()[-1])
    subs = localize(g, read_elm(str(tmp_path / "cube.ELM"), p.nelem), divide(g, nsub), nsub)
    o = oracle_lib.Oracle(nsub, p.nnod, p.ndof, 3, 3, 1)
    for d in subs:
        o.upload(d)
    o.setup(1, 1, 0)
    oc = o.counts()
    o.close()
    for k in ("ncorner", "nedge", "nface", "ndofc", "nig"):
        assert pc[k] == oc[k], k
