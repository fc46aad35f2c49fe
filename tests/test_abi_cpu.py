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


def _planl_in_subprocess(nel, ns, kind, nsublev):
    """The product's host plan with the levels nsublev (plan_only); returns
    the counts of every level from the second on (None where absent)."""
    code = f"""
import sys, json
sys.path.insert(0, {ROOT!r})
import bddcml_amd
from bddcml_amd.cube import CubeProblem
p = CubeProblem({nel}, {ns}, kind={kind!r}, rhs='random', seed=3)
b = bddcml_amd.Bddcml()
b.init({len(nsublev)}, {list(nsublev)!r}, numbase=p.numbase)
b.set_option('plan_only', 1)
for d in p.subdomains():
    b.upload_subdomain_data(d)
b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
print(json.dumps([b.level_counts(l) for l in range(2, {len(nsublev)})]))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    import json
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("nel,ns,kind,nsublev", [(2, 4, "poisson", [64, 8, 2, 1]),
                                                 (2, 4, "poisson", [64, 16, 4, 2, 1]),
                                                 (2, 3, "elasticity", [27, 8, 2, 1])])
def test_deeper_level_plans_match_oracle(oracle_lib, nel, ns, kind, nsublev):
    """Four and five levels: the mesh, division, globs and corners of every
    level from the second on against the oracle chain's (counts, exact)."""
    from oracle.multilevel import MultiLevels
    p = CubeProblem(nel, ns, kind=kind, rhs="random", seed=3)
    t = MultiLevels(p.subdomains(), p.nsub, p.nnod, p.ndof, nsublev, p.numbase)
    prod = _planl_in_subprocess(nel, ns, kind, nsublev)
    assert len(prod) == len(t.levels) - 1
    for k, lc in enumerate(prod):
        oc = t.levels[k + 1].counts()
        assert lc is not None and lc["nsub"] == nsublev[k + 1] and lc["ndof"] == t.ndofcs[k]
        for key in ("ncorner", "nedge", "nface", "ndofc", "nig"):
            assert lc[key] == oc[key], (k + 2, key)
    t.close()


def _plan3_in_subprocess(nel, ns, kind, nsub2):
    """The product's host plan with three levels (plan_only: the second level
    is planned on identity element blocks, its topology does not depend on
    the values); returns both levels' counts."""
    code = f"""
import sys, json
sys.path.insert(0, {ROOT!r})
import bddcml_amd
from bddcml_amd.cube import CubeProblem
p = CubeProblem({nel}, {ns}, kind={kind!r}, rhs='random', seed=3)
b = bddcml_amd.Bddcml()
b.init(3, [p.nsub, {nsub2}, 1], numbase=p.numbase)
b.set_option('plan_only', 1)
for d in p.subdomains():
    b.upload_subdomain_data(d)
b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
print(json.dumps(dict(l1=b.counts(), l2=b.level2_counts())))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    import json
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("nel,ns,kind,nsub2", [(3, 4, "poisson", 8), (2, 4, "poisson", 16),
                                               (2, 4, "poisson", 5), (2, 3, "elasticity", 4)])
def test_second_level_plan_matches_oracle(oracle_lib, nel, ns, kind, nsub2):
    """Second-level mesh, division (recursive bisection), globs and corners of
    the product against the three-level oracle's (counts, exact)."""
    from oracle.multilevel import ThreeLevels
    p = CubeProblem(nel, ns, kind=kind, rhs="random", seed=3)
    t = ThreeLevels(p.subdomains(), p.nsub, p.nnod, p.ndof, nsub2, p.numbase)
    prod = _plan3_in_subprocess(nel, ns, kind, nsub2)
    o2 = t.o2.counts()
    assert prod["l2"] is not None
    assert prod["l2"]["nsub"] == nsub2 and prod["l2"]["ndof"] == t.ndofc
    for k in ("ncorner", "nedge", "nface", "ndofc", "nig"):
        assert prod["l2"][k] == o2[k], k
    t.close()


def test_level_arguments_rejected():
    """A level with more subdomains than the one before, a last level of more
    than one subdomain, and a second level larger than the first are rejected
    at init (reported, not aborted, with throw_on_error)."""
    code = f"""
import sys
sys.path.insert(0, {ROOT!r})
import ctypes as C
import bddcml_amd
L = bddcml_amd.lib()
L.bddcml_amd_set_option(b'throw_on_error', 1.0)
def init(nl, lev):
    arr = (C.c_int * len(lev))(*lev)
    ip = lambda v: C.byref(C.c_int(v))
    L.bddcml_init_c(ip(nl), arr, ip(len(lev)), C.byref(C.c_int(-1)), ip(0), ip(0), ip(1), ip(0))
    buf = C.create_string_buffer(512)
    L.bddcml_amd_last_error(buf, 512)
    return buf.value.decode()
print(repr(init(4, [8, 4, 8, 1])))
print(repr(init(3, [8, 2, 2])))
print(repr(init(3, [8, 16, 1])))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()[-3:]
    assert "level 3" in lines[0]
    assert "last level must be 1" in lines[1]
    assert "level 2" in lines[2]


def test_autotune_candidate_grid():
    from bddcml_amd.autotune import DEFAULT_GRID, candidate_sets
    c = candidate_sets()
    n = 1
    for v in DEFAULT_GRID.values():
        n *= len(v)
    assert len(c) == n and len({tuple(sorted(d.items())) for d in c}) == n
    assert candidate_sets({"nd_leaf": (64,), "task_order": (0, 1)}) == [
        {"nd_leaf": 64, "task_order": 0}, {"nd_leaf": 64, "task_order": 1}]


def _dotprod_in_subprocess(nel, ns, wt, before_setup=False):
    """bddcml_dotprod_subdomain_c of the product (plan_only setup, host
    weights) on seeded vectors of every subdomain; with before_setup the call
    is made before bddcml_setup_preconditioner_c and must fail."""
    code = f"""
import sys, json, numpy as np
sys.path.insert(0, {ROOT!r})
import bddcml_amd
from bddcml_amd.cube import BoxProblem, uniform01
p = BoxProblem({nel}, {ns}, seed=3, kind='poisson', layers=3, jump=100.0, thick=1)
b = bddcml_amd.Bddcml()
b.init(2, [p.nsub, 1], numbase=p.numbase)
b.set_option('plan_only', 1)
subs = list(p.subdomains())
for d in subs:
    b.upload_subdomain_data(d)
def vec(d, seed):
    return 2.0 * uniform01(seed, d['isvgvn'].astype(np.uint64)) - 1.0
if {before_setup!r}:
    try:
        b.dotprod_subdomain(1, vec(subs[0], 1), vec(subs[0], 2))
        print(json.dumps(dict(err='')))
    except RuntimeError as e:
        print(json.dumps(dict(err=str(e))))
    sys.exit(0)
b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, {wt})
out = [b.dotprod_subdomain(d['isub'], vec(d, 1), vec(d, 2)) for d in subs]
try:
    b.dotprod_subdomain(1, vec(subs[0], 1), vec(subs[0], 2)[:-1])
    mismatch = ''
except RuntimeError as e:
    mismatch = str(e)
print(json.dumps(dict(dots=out, mismatch=mismatch)))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    import json
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("wt", [0, 1])
def test_dotprod_subdomain_matches_oracle(oracle_lib, wt):
    """bddcml_dotprod_subdomain_c (reference src/module_bddcml.f90:731-758,
    src/module_dd.f90:12736-12788) against the oracle's weighted subdomain
    dot product, weights of type 0 and 1, and the sum over subdomains of a
    restricted global vector against the global dot product."""
    from bddcml_amd.cube import BoxProblem, uniform01
    p = BoxProblem(6, 5, seed=3, kind="poisson", layers=3, jump=100.0, thick=1)
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, wt)
    prod = _dotprod_in_subprocess(6, 5, wt)
    tot = 0.0
    for s, d in enumerate(p.subdomains()):
        v1 = 2.0 * uniform01(1, d["isvgvn"].astype(np.uint64)) - 1.0
        v2 = 2.0 * uniform01(2, d["isvgvn"].astype(np.uint64)) - 1.0
        ref = o.dotprod_subdomain(d["isub"], v1, v2)
        assert abs(prod["dots"][s] - ref) <= 1e-14 * max(abs(ref), 1.0)
        tot += prod["dots"][s]
    g1 = 2.0 * uniform01(1, np.arange(p.ndof, dtype=np.uint64) + 1) - 1.0
    g2 = 2.0 * uniform01(2, np.arange(p.ndof, dtype=np.uint64) + 1) - 1.0
    assert abs(tot - g1 @ g2) <= 1e-12 * np.abs(g1 * g2).sum()
    assert "Dimensions mismatch" in prod["mismatch"]
    o.close()


def test_dotprod_subdomain_needs_weights():
    """Before the preconditioner setup there are no weights: the call fails
    as the reference's does ('Weights not ready')."""
    prod = _dotprod_in_subprocess(6, 5, 0, before_setup=True)
    assert "Weights not ready" in prod["err"]


def test_autotune_rejects_unsupported_level_count():
    from bddcml_amd.autotune import autotune
    with pytest.raises(ValueError, match="nsublev"):
        autotune([], 8, levels=4, candidates=[{}])


def test_matrix_types_host_plan():
    """matrixtype 2 (symmetric, upper triangle) and matrixtype 0 (full
    storage, here the mirrored symmetric matrix in shuffled order) give the
    host plan of matrixtype 1; an unsymmetric full matrix and an unknown type
    are rejected with a message (reported to the caller), as the reference's
    upload rejects a lower-triangle entry for the symmetric types."""
    code = f"""
import sys, json, numpy as np
sys.path.insert(0, {ROOT!r})
import ctypes as C
import bddcml_amd
from bddcml_amd.cube import CubeProblem
p = CubeProblem(3, 2, kind='elasticity', rhs='random', seed=3)
subs = list(p.subdomains())
def full(d, skew=0.0):
    i, j, a = d['i_sparse'], d['j_sparse'], d['a_sparse']
    off = i != j
    I = np.concatenate([i, j[off]]); J = np.concatenate([j, i[off]])
    A = np.concatenate([a, a[off] * (1.0 + skew)])
    o = np.random.default_rng(1).permutation(len(I))
    return dict(d, i_sparse=I[o], j_sparse=J[o], a_sparse=A[o], matrixtype=0)
def plan(mt, make):
    b = bddcml_amd.Bddcml()
    b.init(2, [p.nsub, 1], numbase=p.numbase)
    b.set_option('plan_only', 1)
    try:
        for d in subs:
            b.upload_subdomain_data(make(d))
        b.setup_preconditioner(mt, 0, 1, 1, 1, 0, 0, 0)
        out = dict(err='', counts=b.counts())
    except RuntimeError as e:
        out = dict(err=str(e), counts=None)
    b.finalize()
    return out
res = dict(t1=plan(1, lambda d: d), t2=plan(2, lambda d: dict(d, matrixtype=2)),
           t0=plan(0, full), bad=plan(0, lambda d: full(d, 1e-6)),
           unknown=plan(3, lambda d: dict(d, matrixtype=3)))
print(json.dumps(res))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    import json
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["t1"]["err"] == "" and res["t1"]["counts"] is not None
    assert res["t2"]["err"] == "" and res["t2"]["counts"] == res["t1"]["counts"]
    assert res["t0"]["err"] == "" and res["t0"]["counts"] == res["t1"]["counts"]
    assert "not symmetric" in res["bad"]["err"]
    assert "unknown matrixtype" in res["unknown"]["err"]
