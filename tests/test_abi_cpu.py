"""CPU checks of the product library: the C ABI loads and exports every
declared symbol, and the host plan (neighbours, globs, corners, weights,
constraints, coarse numbering) agrees with the independent oracle
implementation.  No GPU call is made here (plan_only mode)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import CubeProblem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(bddcml_\w+)\s*\(", text)))


def test_headers_match_exported_symbols():
    L = bddcml_amd.lib()
    declared = _declared("bddcml_interface_c.h") + _declared("bddcml_amd.h")
    assert set(bddcml_amd.REFERENCE_SYMBOLS) <= set(declared)
    assert set(bddcml_amd.EXTRA_SYMBOLS) <= set(declared)
    for name in declared:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", bddcml_amd.LIB_PATH],
                         capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}\b", out), name


def test_reference_header_signatures_kept():
    """Every reference prototype keeps its name and parameter list."""
    ref = os.path.join(ROOT, "include", "bddcml_interface_c.h")
    text = re.sub(r"/\*.*?\*/", "", open(ref).read(), flags=re.S)
    protos = {m.group(1): re.sub(r"\s+", "", m.group(2))
              for m in re.finditer(r"void\s+(bddcml_\w+_c)\s*\(([^)]*)\)", text)}
    assert set(protos) == set(bddcml_amd.REFERENCE_SYMBOLS) | {"bddcml_finalize_c"} | set()
    assert protos["bddcml_solve_c"] == (
        "int*comm_all,int*method,double*tol,int*maxit,int*ndecrmax,int*recycling_int,"
        "int*max_number_of_stored_vectors,int*num_iter,int*converged_reason,double*condition_number")
    assert protos["bddcml_init_c"] == (
        "int*nl,int*nsublev,int*lnsublev,int*nsub_loc_1,int*comm_init,int*verbose_level,"
        "int*numbase,int*just_direct_solve_int")


def _plan_in_subprocess(nel, ns, kind, weights_type):
    """Run the product's host plan in a child process (the library keeps
    process-wide state) and return its counts and per-subdomain data."""
    code = f"""
import sys, json, numpy as np
sys.path.insert(0, {ROOT!r})
import bddcml_amd
from bddcml_amd.cube import CubeProblem
p = CubeProblem({nel}, {ns}, kind={kind!r}, rhs='random', seed=3)
b = bddcml_amd.Bddcml()
b.init(2, [p.nsub, 1], numbase=p.numbase)
b.set_option('plan_only', 1)
for d in p.subdomains():
    b.upload_subdomain_data(d)
b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, {weights_type})
out = dict(counts=b.counts(), ig=b.interface_dofs().tolist(), subs=[])
for s in range(p.nsub):
    g, c = b.sub_maps(s + 1)
    out['subs'].append(dict(g=g.tolist(), c=c.tolist()))
print(json.dumps(out))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    import json
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("nel,ns,kind,wt", [(2, 2, "poisson", 0), (3, 3, "poisson", 1),
                                            (4, 3, "poisson", 0), (2, 2, "elasticity", 0),
                                            (8, 4, "poisson", 0)])
def test_host_plan_matches_oracle(oracle_lib, nel, ns, kind, wt):
    p = CubeProblem(nel, ns, kind=kind, rhs="random", seed=3)
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, wt)
    prod = _plan_in_subprocess(nel, ns, kind, wt)
    oc = o.counts()
    for k in ("ncorner", "nedge", "nface", "ndofc", "nig"):
        assert prod["counts"][k] == oc[k], k
    assert prod["ig"] == o.interface_dofs().tolist()
    for s in range(p.nsub):
        isub = s + 1
        assert prod["subs"][s]["g"] == o.sub_interface_gdofs(isub).tolist()
        _, ind = o.sub_constraints(isub)
        assert sorted(prod["subs"][s]["c"]) == sorted(ind.tolist())
    o.close()
