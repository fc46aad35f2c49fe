"""bddcml_amd_element_gradient_noalloc (device buffers in the library's fixed
scratch array) against bddcml_amd_element_gradient (buffers allocated per
call), through the C ABI: bit-identical gradients on two calls in a row, the
second on a smaller problem, and the same values as numpy's
-lam_e^T A_e u_e."""
import ctypes as C
import sys

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import CubeProblem, uniform01

pytestmark = pytest.mark.gpu

dp = C.POINTER(C.c_double)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(p):
    b = bddcml_amd.Bddcml()
    b.init(2, [p.nsub, 1], numbase=p.numbase)
    for d in p.subdomains():
        b.upload_subdomain_data(d)
    b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, 0)
    return b


def _grad(b, fn, ae, nke, per_element, u, lam, nelem):
    out = np.full(nelem, np.nan)
    ae = np.ascontiguousarray(ae, dtype=np.float64)
    rc = getattr(b.L, fn)(ae.ctypes.data_as(dp), nke, per_element, u.ctypes.data_as(dp),
                          lam.ctypes.data_as(dp), out.ctypes.data_as(dp))
    assert rc == 0
    return out


def _numpy_grad(p, ae_of, u, lam):
    b = p.numbase
    g = np.zeros(p.nelem)
    for d in p.subdomains():
        inet = (d["inet"] - b).reshape(-1, 8)
        gd = d["isvgvn"] - b
        for e, nodes in zip(d["isegn"] - b, inet):
            g[e] = -lam[gd[nodes]] @ ae_of(e) @ u[gd[nodes]]
    return g


def test_noalloc_gradient_matches_allocating_entry_point():
    for call, (nel, ns) in enumerate([(4, 3), (2, 2)]):
        _say(f"case element_gradient_noalloc call {call + 1}: {nel * ns}^3 elements")
        p = CubeProblem(nel, ns, rhs="random", seed=3)
        b = _setup(p)
        try:
            u = 2.0 * uniform01(11 + call, np.arange(p.ndof, dtype=np.uint64)) - 1.0
            lam = 2.0 * uniform01(21 + call, np.arange(p.ndof, dtype=np.uint64)) - 1.0
            ae = p.kref
            g_a = _grad(b, "bddcml_amd_element_gradient", ae, 8, 0, u, lam, p.nelem)
            g_n = _grad(b, "bddcml_amd_element_gradient_noalloc", ae, 8, 0, u, lam, p.nelem)
            assert np.array_equal(g_a, g_n)
            g_ref = _numpy_grad(p, lambda e: ae, u, lam)
            assert np.abs(g_n - g_ref).max() <= 1e-13 * np.abs(g_ref).max()
            # one matrix per element
            scale = 1.0 + uniform01(31 + call, np.arange(p.nelem, dtype=np.uint64))
            aes = scale[:, None, None] * ae[None, :, :]
            g_a = _grad(b, "bddcml_amd_element_gradient", aes, 8, 1, u, lam, p.nelem)
            g_n = _grad(b, "bddcml_amd_element_gradient_noalloc", aes, 8, 1, u, lam, p.nelem)
            assert np.array_equal(g_a, g_n)
            g_ref = _numpy_grad(p, lambda e: aes[e], u, lam)
            assert np.abs(g_n - g_ref).max() <= 1e-13 * np.abs(g_ref).max()
        finally:
            b.finalize()


def test_noalloc_gradient_rejects_problems_beyond_scratch():
    """56^3 elements with one matrix per element need about 90 MiB: the
    entry point reports an error instead of writing past its array."""
    _say("case element_gradient_noalloc oversized: 56^3 elements, per-element matrices")
    p = CubeProblem(8, 7, rhs="random", seed=3)
    b = _setup(p)
    try:
        u = np.ones(p.ndof)
        aes = np.broadcast_to(p.kref, (p.nelem, 8, 8)).copy()
        out = np.zeros(p.nelem)
        with pytest.raises(RuntimeError, match="scratch"):
            b.L.bddcml_amd_element_gradient_noalloc(aes.ctypes.data_as(dp), 8, 1,
                                                    u.ctypes.data_as(dp), u.ctypes.data_as(dp),
                                                    out.ctypes.data_as(dp))
    finally:
        b.finalize()
