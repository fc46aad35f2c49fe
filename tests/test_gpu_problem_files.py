"""File-driven solve (bddcml_amd/problem_files.py, the sequence of the
reference's examples/bddcml_global.f90) on the GPU: an elasticity cube with
random Dirichlet values written as PAR/ML/GMIS/FVS/RHS/ELM files is solved
from the files; the global solution matches the dense direct solve within
1e-8, the reactions equal A u - f at the fixed dofs, the SOL file holds
what was returned, and the subdomain solutions export to VTU files that
ParaView's XML reader structure accepts (well-formed XML with the expected
point and cell counts)."""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd import problem_files as pf
from bddcml_amd.cube import CubeProblem

pytestmark = pytest.mark.gpu


def test_solve_from_problem_files(tmp_path):
    print("case problem files: elasticity cube 3x2, 2 levels", flush=True)
    sys.stdout.flush()
    p = CubeProblem(3, 2, kind="elasticity", rhs="random", seed=4, dirichlet="random")
    g = p.global_data()
    pf.write_problem("cube", str(tmp_path), g, [p.nsub, 1], tol=1e-12, maxit=300, ndecrmax=30)
    p.write_elm(str(tmp_path / "cube.ELM"))
    res = pf.solve_problem("cube", str(tmp_path), load_division=0, neighbouring=1)
    assert res["reason"] == 0
    A, f, fixed, _ = p.global_dense(eliminate=False)
    Ae, fe, _ = p.global_dense()
    u_ref = np.linalg.solve(Ae, fe)
    assert np.abs(res["sol"] - u_ref).max() <= 1e-8 * np.abs(u_ref).max()
    # reactions (src/module_dd.f90:1258-1380): (A u - f) at fixed dofs, the
    # couplings to fixed dofs elsewhere
    rea_ref = np.where(fixed, A @ u_ref - f, A[:, fixed] @ u_ref[fixed])
    assert np.abs(res["rea"] - rea_ref).max() <= 1e-8 * np.abs(rea_ref).max()
    sol, rea, sum_rea = pf.read_sol(str(tmp_path / "cube.SOL"), g["ndof"])
    np.testing.assert_array_equal(sol, res["sol"])
    np.testing.assert_array_equal(rea, res["rea"])
    np.testing.assert_array_equal(sum_rea, res["sum_rea"])


def test_subdomain_solutions_export(tmp_path):
    print("case VTU export: poisson cube 3x2", flush=True)
    p = CubeProblem(3, 2, rhs="random", seed=6)
    b = bddcml_amd.Bddcml()
    try:
        b.init(2, [p.nsub, 1], numbase=p.numbase)
        subs = list(p.subdomains())
        for d in subs:
            b.upload_subdomain_data(d)
        b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
        it, reason, cond = b.solve(0, 1e-10, 300, 30)
        assert reason == 0
        for d in subs:
            sols = b.download_local_solution(d["isub"], d["ndofs"])
            bb = 1 if d["inet"].min() >= 1 else 0
            xyz = np.asarray(d["xyz"], dtype=float).reshape(d["nnods"], -1, order="F")
            path = pf.write_subdomain_vtu("poisson", d["isub"], np.asarray(d["inet"]) - bb + 1, d["nnet"],
                                          xyz, sols, directory=str(tmp_path))
            piece = ET.parse(path).getroot().find("UnstructuredGrid/Piece")
            assert int(piece.get("NumberOfPoints")) == d["nnods"]
            assert int(piece.get("NumberOfCells")) == len(d["nnet"])
            vals = np.array(piece.find("PointData/DataArray").text.split(), dtype=float)
            assert np.abs(vals - sols).max() <= 1e-6 * np.abs(sols).max()
        root = ET.parse(pf.write_pvd("poisson", p.nsub, directory=str(tmp_path))).getroot()
        files = [ds.get("file") for ds in root.iter("DataSet")]
        assert len(files) == p.nsub and all(os.path.exists(tmp_path / f) for f in files)
    finally:
        b.finalize()
