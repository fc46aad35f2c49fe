"""The drop-in boundary from compiled callers: a Fortran program using the
module_bddcml interface (bddcml_amd/fortran/module_bddcml.f90, compiled with
amdflang) and a C program using include/bddcml_interface_c.h, both building
the Poisson cube themselves (tests/drivers/).  On the CPU they run the host
plan only and their counts must equal the oracle's; on the GPU they solve
and their global solutions must match the oracle's."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bddcml_amd.cube import CubeProblem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bddcml_amd")
FDIR = os.path.join(PKG, "fortran")
DRV = os.path.join(ROOT, "tests", "drivers")


def _tool(name, fallback):
    return shutil.which(name) or fallback


@pytest.fixture(scope="session")
def callers(tmp_path_factory):
    """Compiled Fortran and C drivers (the libraries must be built)."""
    for lib in (os.path.join(PKG, "libbddcml_amd.so"), os.path.join(FDIR, "libbddcml_amd_fortran.so"),
                os.path.join(FDIR, "include", "module_bddcml.mod")):
        if not os.path.exists(lib):
            raise RuntimeError(f"{lib} missing: build with `python -m bddcml_amd.build`")
    d = tmp_path_factory.mktemp("callers")
    fexe, cexe = str(d / "poisson_cube_f"), str(d / "poisson_cube_c")
    subprocess.run([_tool("amdflang", "/opt/rocm/bin/amdflang"), "-O1", "-I",
                    os.path.join(FDIR, "include"), os.path.join(DRV, "poisson_cube_f.f90"),
                    "-L" + FDIR, "-lbddcml_amd_fortran", "-L" + PKG, "-lbddcml_amd",
                    "-Wl,-rpath," + FDIR, "-Wl,-rpath," + PKG, "-o", fexe],
                   check=True, capture_output=True, text=True)
    subprocess.run([_tool("gcc", "/usr/bin/gcc"), "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(DRV, "poisson_cube_c.c"), "-L" + PKG, "-lbddcml_amd",
                    "-Wl,-rpath," + PKG, "-o", cexe], check=True, capture_output=True, text=True)
    return dict(fortran=fexe, c=cexe, dir=str(d))


def _run(exe, nes, ns, mode, out):
    r = subprocess.run([exe, str(nes), str(ns), str(mode), out], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    with open(out) as f:
        lines = f.read().split("\n")
    counts = [int(v) for v in lines[0].split()]
    res = None
    if mode == 1:
        it, reason, cond = lines[1].split()
        sol = np.array([float(v) for v in lines[2:] if v.strip()])
        res = (int(it), int(reason), float(cond), sol)
    return counts, res


@pytest.mark.parametrize("lang", ["fortran", "c"])
@pytest.mark.parametrize("nes,ns", [(2, 2), (3, 3), (2, 4)])
def test_compiled_caller_plan_matches_oracle(oracle_lib, callers, lang, nes, ns):
    out = os.path.join(callers["dir"], f"plan_{lang}_{nes}_{ns}.txt")
    counts, _ = _run(callers[lang], nes, ns, 0, out)
    p = CubeProblem(nes, ns, rhs="const")
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, 0)
    oc = o.counts()
    assert counts == [oc["ncorner"], oc["nedge"], oc["nface"], oc["ndofc"], oc["nig"]]
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("lang", ["fortran", "c"])
def test_compiled_caller_solve_matches_oracle(oracle_lib, callers, lang):
    print(f"case compiled {lang} caller: 12^3 Poisson, 27 subdomains, PCG to 1e-10", flush=True)
    out = os.path.join(callers["dir"], f"solve_{lang}.txt")
    counts, (it, reason, cond, sol) = _run(callers[lang], 4, 3, 1, out)
    p = CubeProblem(4, 3, rhs="const")
    o, (it_o, r_o, c_o) = oracle_lib.solve_cube(p, tol=1e-10, maxit=500)
    ref = np.zeros(p.ndof)
    for d in p.subdomains():
        ref[d["isvgvn"] - 1] = o.local_solution(d["isub"])
    o.close()
    assert reason == 0 == r_o
    assert abs(it - it_o) <= 1
    assert abs(cond - c_o) <= 1e-6 * c_o
    assert np.abs(sol - ref).max() <= 1e-10 * np.abs(ref).max()
