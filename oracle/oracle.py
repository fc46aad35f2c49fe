"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker.  The product (bddcml_amd) never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i = C.c_int
_d = C.c_double
_v = C.c_void_p


def _declare(L):
    L.orc_create.restype = _v
    L.orc_create.argtypes = [_i, _i, _i, _i, _i, _i]
    L.orc_destroy.argtypes = [_v]
    L.orc_upload_subdomain.restype = _i
    L.orc_upload_subdomain.argtypes = [
        _v, _i, _i, _i, _i, _ip, _i, _ip, _ip, _ip, _ip, _dp, _i, _i, _ip, _dp, _dp, _i,
        _dp, _i, _ip, _ip, _dp, _i, _i, _dp, _i, _i, _i, _i]
    L.orc_setup.restype = _i
    L.orc_setup.argtypes = [_v, _i, _i, _i]
    L.orc_solve.restype = _i
    L.orc_solve.argtypes = [_v, _i, _d, _i, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]
    L.orc_download_local_solution.restype = _i
    L.orc_download_local_solution.argtypes = [_v, _i, _dp, _i]
    L.orc_dotprod_subdomain.restype = _i
    L.orc_dotprod_subdomain.argtypes = [_v, _i, _dp, _dp, _i, C.POINTER(_d)]
    L.orc_global_counts.argtypes = [_v, _ip]
    L.orc_interface_dofs.argtypes = [_v, _ip]
    L.orc_corner_nodes.argtypes = [_v, _ip]
    L.orc_sub_sizes.argtypes = [_v, _i, _ip]
    L.orc_sub_interface_gdofs.argtypes = [_v, _i, _ip]
    for name in ("orc_sub_weights", "orc_sub_schur", "orc_sub_phi", "orc_sub_coarse"):
        getattr(L, name).argtypes = [_v, _i, _dp]
    L.orc_sub_constraints.argtypes = [_v, _i, _dp, _ip]
    L.orc_apply_schur.argtypes = [_v, _dp, _dp]
    L.orc_apply_pc.argtypes = [_v, _dp, _dp]
    L.orc_reduced_rhs.argtypes = [_v, _dp]
    L.orc_interface_solution.argtypes = [_v, _dp]
    L.orc_residual_history.restype = _i
    L.orc_residual_history.argtypes = [_v, C.c_void_p, _i]
    L.orc_coarse_nodes.argtypes = [_v, _ip, _dp]
    L.orc_sub_cnodes.restype = _i
    L.orc_sub_cnodes.argtypes = [_v, _i, C.c_void_p]
    L.orc_set_coarse_solver.argtypes = [_v, COARSE_FN, _v]
    L.orc_apply_level.restype = _i
    L.orc_apply_level.argtypes = [_v, _dp, _dp]
    L.orc_set_constraints.restype = _i
    L.orc_set_constraints.argtypes = [_v, _i, _i, _dp, _ip, _i]
    L.orc_refresh_coarse.restype = _i
    L.orc_refresh_coarse.argtypes = [_v]


# uc = f(ctx, rc) in place of the coarse Cholesky solve (three levels)
COARSE_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def _ia(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _da(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class Oracle:
    """Two-level BDDC on the CPU in FP64 following the reference algorithm."""

    def __init__(self, nsub, nnod, ndof, ndim=3, meshdim=3, numbase=1):
        self.L = lib()
        self.numbase = numbase
        self.h = self.L.orc_create(nsub, nnod, ndof, ndim, meshdim, numbase)
        self.nsub = nsub

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, d: dict, dof_data=None):
        xyz = _da(np.asarray(d["xyz"]).ravel(order="F"))
        dd = _da(dof_data if dof_data is not None else np.zeros(0))
        rc = self.L.orc_upload_subdomain(
            self.h, d["isub"], d["nelems"], d["nnods"], d["ndofs"], _ia(d["inet"]),
            len(d["inet"]), _ia(d["nnet"]), _ia(d["nndf"]), _ia(d["isngn"]), _ia(d["isvgvn"]),
            xyz, d["nnods"], 3, _ia(d["ifix"]), _da(d["fixv"]), _da(d["rhs"]),
            d["is_rhs_complete"], _da(d["sol"]), d["matrixtype"], _ia(d["i_sparse"]),
            _ia(d["j_sparse"]), _da(d["a_sparse"]), len(d["a_sparse"]), d["is_assembled"], dd,
            len(dd), d["find_components"], d["use_dual_mesh_graph"], d["neighbouring"])
        if rc != 0:
            raise RuntimeError(f"orc_upload_subdomain failed: {rc}")

    def setup(self, use_corner=1, use_arith=1, weights_type=0):
        rc = self.L.orc_setup(self.h, use_corner, use_arith, weights_type)
        if rc != 0:
            raise RuntimeError(f"orc_setup failed: {rc}")

    def solve(self, method=0, tol=1e-8, maxit=500, ndecrmax=50):
        it, reason, cond = C.c_int(), C.c_int(), C.c_double()
        rc = self.L.orc_solve(self.h, method, tol, maxit, ndecrmax, C.byref(it), C.byref(reason),
                              C.byref(cond))
        if rc != 0:
            raise RuntimeError(f"orc_solve failed: {rc}")
        return it.value, reason.value, cond.value

    def counts(self):
        out = np.zeros(5, dtype=np.int32)
        self.L.orc_global_counts(self.h, out)
        return dict(ncorner=int(out[0]), nedge=int(out[1]), nface=int(out[2]),
                    ndofc=int(out[3]), nig=int(out[4]))

    def interface_dofs(self):
        out = np.zeros(self.counts()["nig"], dtype=np.int32)
        self.L.orc_interface_dofs(self.h, out)
        return out

    def corner_nodes(self):
        out = np.zeros(self.counts()["ncorner"], dtype=np.int32)
        self.L.orc_corner_nodes(self.h, out)
        return out

    def sub_sizes(self, isub):
        out = np.zeros(8, dtype=np.int32)
        self.L.orc_sub_sizes(self.h, isub, out)
        keys = ("ndofs", "ndofi", "ndofo", "nconstr", "nnodi", "ncorner", "nedge", "nface")
        return dict(zip(keys, (int(v) for v in out)))

    def sub_interface_gdofs(self, isub):
        s = self.sub_sizes(isub)
        out = np.zeros(s["ndofi"], dtype=np.int32)
        self.L.orc_sub_interface_gdofs(self.h, isub, out)
        return out

    def sub_weights(self, isub):
        s = self.sub_sizes(isub)
        out = np.zeros(s["ndofi"])
        self.L.orc_sub_weights(self.h, isub, out)
        return out

    def sub_schur(self, isub):
        n = self.sub_sizes(isub)["ndofi"]
        out = np.zeros(n * n)
        self.L.orc_sub_schur(self.h, isub, out)
        return out.reshape(n, n)

    def sub_constraints(self, isub):
        s = self.sub_sizes(isub)
        C_ = np.zeros(s["nconstr"] * s["ndofi"])
        ind = np.zeros(s["nconstr"], dtype=np.int32)
        self.L.orc_sub_constraints(self.h, isub, C_, ind)
        return C_.reshape(s["nconstr"], s["ndofi"]), ind

    def sub_phi(self, isub):
        s = self.sub_sizes(isub)
        out = np.zeros(s["ndofi"] * s["nconstr"])
        self.L.orc_sub_phi(self.h, isub, out)
        return out.reshape(s["ndofi"], s["nconstr"])

    def sub_coarse(self, isub):
        m = self.sub_sizes(isub)["nconstr"]
        out = np.zeros(m * m)
        self.L.orc_sub_coarse(self.h, isub, out)
        return out.reshape(m, m)

    def apply_schur(self, x):
        y = np.zeros_like(_da(x))
        self.L.orc_apply_schur(self.h, _da(x), y)
        return y

    def apply_pc(self, r):
        z = np.zeros_like(_da(r))
        self.L.orc_apply_pc(self.h, _da(r), z)
        return z

    def reduced_rhs(self):
        g = np.zeros(self.counts()["nig"])
        self.L.orc_reduced_rhs(self.h, g)
        return g

    def interface_solution(self):
        u = np.zeros(self.counts()["nig"])
        self.L.orc_interface_solution(self.h, u)
        return u

    def residual_history(self):
        n = self.L.orc_residual_history(self.h, None, 0)
        out = np.zeros(n)
        self.L.orc_residual_history(self.h, out.ctypes.data, n)
        return out

    def local_solution(self, isub):
        n = self.sub_sizes(isub)["ndofs"]
        out = np.zeros(n)
        if self.L.orc_download_local_solution(self.h, isub, out, n):
            raise RuntimeError("no solution")
        return out

    def coarse_nodes(self):
        """(nndfc, xyzc) of the coarse nodes, xyzc of shape (nnodc, 3)."""
        c = self.counts()
        n = c["ncorner"] + c["nedge"] + c["nface"]
        nndf = np.zeros(max(n, 1), dtype=np.int32)
        xyz = np.zeros(max(n, 1) * 3)
        self.L.orc_coarse_nodes(self.h, nndf, xyz)
        return nndf[:n], xyz[:n * 3].reshape(3, n).T.copy() if n else np.zeros((0, 3))

    def sub_cnodes(self, isub):
        n = self.L.orc_sub_cnodes(self.h, isub, None)
        out = np.zeros(max(n, 1), dtype=np.int32)
        self.L.orc_sub_cnodes(self.h, isub, out.ctypes.data)
        return out[:n]

    def apply_level(self, f):
        x = np.zeros_like(_da(f))
        if self.L.orc_apply_level(self.h, _da(f), x):
            raise RuntimeError("orc_apply_level needs a level without Dirichlet data")
        return x

    def set_constraints(self, isub, C, indrowc, ndofc):
        """Replace subdomain isub's constraint rows by the dense rows C
        (nc x ndofi, the oracle's interface order) with global coarse dofs
        indrowc of a coarse problem of ndofc dofs."""
        C = np.ascontiguousarray(C, dtype=np.float64)
        rc = self.L.orc_set_constraints(self.h, isub, C.shape[0], C.ravel() if C.size else
                                        np.zeros(1), _ia(indrowc) if len(indrowc) else
                                        np.zeros(1, dtype=np.int32), int(ndofc))
        if rc != 0:
            raise RuntimeError(f"orc_set_constraints failed: {rc}")

    def refresh_coarse(self):
        """Coarse basis, coarse matrices and coarse factor for the current
        constraint rows."""
        rc = self.L.orc_refresh_coarse(self.h)
        if rc != 0:
            raise RuntimeError(f"orc_refresh_coarse failed: {rc}")

    def dotprod_subdomain(self, isub, v1, v2):
        out = C.c_double()
        self.L.orc_dotprod_subdomain(self.h, isub, _da(v1), _da(v2), len(v1), C.byref(out))
        return out.value


def solve_cube(problem, use_corner=1, use_arith=1, weights_type=0, method=0, tol=1e-8,
               maxit=500, ndecrmax=50):
    """Upload all subdomains of a CubeProblem, set up and solve; returns the
    Oracle (caller closes) and (num_iter, reason, cond)."""
    o = Oracle(problem.nsub, problem.nnod, problem.ndof, 3, 3, problem.numbase)
    for d in problem.subdomains():
        o.upload(d)
    o.setup(use_corner, use_arith, weights_type)
    res = o.solve(method, tol, maxit, ndecrmax)
    return o, res
