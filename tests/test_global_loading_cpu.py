"""Global loading on the CPU: the oracle's restatement pinned, and the
library's host plan for a globally loaded mesh against it (plan_only, no
GPU call).

Pins of the restatement (oracle/global_loading.py):
- the ELM records written by CubeProblem.write_elm read back as the element
  matrix's upper triangle by columns (a closed form: the reference element);
- with the cube's own division (element -> subdomain by position), the
  localized subdomains equal those CubeProblem.subdomain builds
  independently (nodes, elements, dofs, coordinates, boundary data, the
  assembled matrix);
- with the bisection division, the subdomain matrices sum to the global
  matrix assembled from the elements directly;
- the bisection's parts have equal sizes (+-1 element).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from bddcml_amd.cube import CubeProblem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _assembled(d, b=1):
    n = d["ndofs"]
    A = np.zeros((n, n))
    np.add.at(A, (d["i_sparse"] - b, d["j_sparse"] - b), d["a_sparse"])
    return A


def _natural_iets(p):
    """Element -> subdomain of the cube's own division (0-based)."""
    nb, ns, m = p.nel, p.nsub_edge, p.nel_sub
    e = np.arange(p.nelem)
    ei, ej, ek = e % nb, (e // nb) % nb, e // (nb * nb)
    return (ei // m) + (ej // m) * ns + (ek // m) * ns * ns


def test_elm_records_read_back(tmp_path):
    from oracle.global_loading import read_elm
    p = CubeProblem(2, 2, kind="elasticity", coef=lambda e: 1.0 + np.asarray(e) % 3)
    path = str(tmp_path / "cube.ELM")
    p.write_elm(path)
    elm = read_elm(path, p.nelem)
    ne = p.kref.shape[0]
    col = np.concatenate([p.kref[:j + 1, j] for j in range(ne)])
    assert len(elm) == p.nelem and len(elm[0]) == ne * (ne + 1) // 2
    for e in (0, 1, 2, p.nelem - 1):
        assert np.array_equal(elm[e], (1.0 + e % 3) * col)
    size = os.path.getsize(path)
    assert size == p.nelem * (12 + 8 * len(col))


@pytest.mark.parametrize("kind", ["poisson", "elasticity"])
def test_natural_division_equals_subdomain_loading(tmp_path, kind):
    from oracle.global_loading import divide, localize, read_elm
    p = CubeProblem(2, 2, kind=kind, rhs="random", seed=4, dirichlet="random")
    g = p.global_data()
    path = str(tmp_path / "cube.ELM")
    p.write_elm(path)
    subs = localize(g, read_elm(path, p.nelem), divide(g, p.nsub, iets=_natural_iets(p)), p.nsub)
    for s, d in zip(subs, p.subdomains()):
        for k in ("isngn", "inet", "isvgvn", "isegn", "nndf", "ifix"):
            assert np.array_equal(s[k], d[k]), k
        for k in ("xyz", "fixv", "rhs"):
            assert np.array_equal(np.asarray(s[k]), np.asarray(d[k])), k
        assert np.allclose(_assembled(s), _assembled(d), rtol=0, atol=1e-15)


def test_bisection_division_sums_to_global_matrix(tmp_path):
    from oracle.global_loading import divide, localize, read_elm
    p = CubeProblem(3, 2, kind="poisson", rhs="random", seed=5)
    g = p.global_data()
    path = str(tmp_path / "cube.ELM")
    p.write_elm(path)
    for nsub in (3, 5, 8):
        iets = divide(g, nsub)
        counts = np.bincount(iets, minlength=nsub)
        assert counts.max() - counts.min() <= 1
        subs = localize(g, read_elm(path, p.nelem), iets, nsub)
        A = np.zeros((p.ndof, p.ndof))
        for s in subs:
            gd = s["isvgvn"] - 1
            A[np.ix_(gd, gd)] += _assembled(s)
        A = A + np.triu(A, 1).T
        A_ref = p.global_dense(eliminate=False)[0]
        assert np.allclose(A, A_ref, rtol=0, atol=1e-14)


def _product_plan(nel, ns, kind, nsub, load_division, workdir):
    """The library's host plan of a globally loaded cube (plan_only)."""
    code = f"""
import os, sys, json
sys.path.insert(0, {ROOT!r})
import bddcml_amd
from bddcml_amd.cube import CubeProblem
p = CubeProblem({nel}, {ns}, kind={kind!r}, rhs='random', seed=3)
os.chdir({workdir!r})
b = bddcml_amd.Bddcml()
b.init(2, [{nsub}, 1], numbase=1)
b.set_option('plan_only', 1)
fd = os.open('cube.ELM', os.O_RDONLY)
b.upload_global_data(p.global_data(), fd, neighbouring=1, load_division={load_division})
os.close(fd)
b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
print(json.dumps(b.counts()))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    import json
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("nel,ns,kind,nsub", [(3, 2, "poisson", 5), (2, 3, "poisson", 8),
                                              (2, 2, "elasticity", 6)])
def test_global_plan_matches_oracle(oracle_lib, tmp_path, nel, ns, kind, nsub):
    """Bisection division: corners, edges, faces, coarse and interface dofs
    of the library's plan equal the oracle's on the restated subdomains."""
    from oracle.global_loading import divide, localize, read_elm
    p = CubeProblem(nel, ns, kind=kind, rhs="random", seed=3)
    p.write_elm(str(tmp_path / "cube.ELM"))
    g = p.global_data()
    subs = localize(g, read_elm(str(tmp_path / "cube.ELM"), p.nelem), divide(g, nsub), nsub)
    o = oracle_lib.Oracle(nsub, p.nnod, p.ndof, 3, 3, 1)
    for d in subs:
        o.upload(d)
    o.setup(1, 1, 0)
    oc = o.counts()
    o.close()
    pc = _product_plan(nel, ns, kind, nsub, 0, str(tmp_path))
    for k in ("ncorner", "nedge", "nface", "ndofc", "nig"):
        assert pc[k] == oc[k], k


def test_division_file_is_read(oracle_lib, tmp_path):
    """load_division: partition_l1.ES in the working directory gives the
    cube's own division, whose plan equals the per-subdomain upload's."""
    p = CubeProblem(2, 3, kind="poisson", rhs="random", seed=3)
    p.write_elm(str(tmp_path / "cube.ELM"))
    np.savetxt(str(tmp_path / "partition_l1.ES"), _natural_iets(p) + 1, fmt="%d")
    pc = _product_plan(2, 3, "poisson", p.nsub, 1, str(tmp_path))
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, 1)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, 0)
    oc = o.counts()
    o.close()
    for k in ("ncorner", "nedge", "nface", "ndofc", "nig"):
        assert pc[k] == oc[k], k


def test_global_loading_errors(tmp_path):
    """More subdomains than elements, a truncated ELM file and changing IFIX
    with new global data are reported (throw_on_error)."""
    code = f"""
import os, sys
sys.path.insert(0, {ROOT!r})
import numpy as np
import bddcml_amd
from bddcml_amd.cube import CubeProblem
p = CubeProblem(1, 2, kind='poisson')
os.chdir({str(tmp_path)!r})
p.write_elm('cube.ELM')
g = p.global_data()
def run(nsub, path):
    b = bddcml_amd.Bddcml()
    b.init(2, [nsub, 1])
    fd = os.open(path, os.O_RDONLY)
    try:
        b.upload_global_data(g, fd)
        return 'ok'
    except RuntimeError as e:
        return str(e)
    finally:
        os.close(fd)
print(repr(run(9, 'cube.ELM')))
open('short.ELM', 'wb').write(open('cube.ELM', 'rb').read()[:-20])
print(repr(run(2, 'short.ELM')))
b = bddcml_amd.Bddcml()
b.init(2, [2, 1])
b.set_option('plan_only', 1)
fd = os.open('cube.ELM', os.O_RDONLY)
b.upload_global_data(g, fd)
os.close(fd)
b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
ifix = g['ifix'].copy(); ifix[13] = 1 - ifix[13]
try:
    b.change_global_data(ifix, g['fixv'], g['rhs'], g['sol'])
    print(repr('ok'))
except RuntimeError as e:
    print(repr(str(e)))
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()[-3:]
    assert "must not be larger than number of elements" in lines[0]
    assert "truncated" in lines[1] or "ends before" in lines[1]
    assert "IFIX cannot be changed" in lines[2]
