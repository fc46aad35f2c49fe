"""bddcml_amd -- MI355X BDDC library, Python front end.

The compute path is the in-tree shared library libbddcml_amd.so (HIP kernels
for gfx950 plus host C++), reached through its C ABI (include/*.h).  This
module mirrors the reference interface of src/module_bddcml.f90 (same
routine names and argument meaning) on top of ctypes, for tests and
benchmarks.  Importing it does not touch the GPU; loading fails loudly when
the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbddcml_amd.so")

_lib = None

# exported symbols of include/bddcml_interface_c.h (reference
# src/bddcml_interface_c.h:6-61) and include/bddcml_amd.h
REFERENCE_SYMBOLS = [
    "bddcml_init_c", "bddcml_upload_global_data_c", "bddcml_upload_subdomain_data_c",
    "bddcml_setup_preconditioner_c", "bddcml_solve_c", "bddcml_download_local_solution_c",
    "bddcml_download_global_solution_c", "bddcml_download_local_reactions_c",
    "bddcml_download_global_reactions_c", "bddcml_change_global_data_c",
    "bddcml_change_subdomain_data_c", "bddcml_setup_new_data_c", "bddcml_dotprod_subdomain_c",
    "bddcml_finalize_c",
]
EXTRA_SYMBOLS = [
    "bddcml_amd_counts", "bddcml_amd_interface_dofs", "bddcml_amd_apply_schur",
    "bddcml_amd_apply_pc", "bddcml_amd_reduced_rhs", "bddcml_amd_interface_solution",
    "bddcml_amd_residual_history", "bddcml_amd_timings", "bddcml_amd_set_option",
    "bddcml_amd_sub_info", "bddcml_amd_sub_maps", "bddcml_amd_sub_matrix",
    "bddcml_amd_sub_coarse", "bddcml_amd_last_error", "bddcml_amd_selftest_tile",
    "bddcml_amd_selftest_ginv", "bddcml_amd_selftest_symeig",
    "bddcml_amd_selftest_copy", "bddcml_amd_element_gradient", "bddcml_amd_recycling_basis",
    "bddcml_amd_level2_counts", "bddcml_amd_level_counts", "bddcml_amd_element_gradient_noalloc",
    "bddcml_amd_adaptive_pairs",
]


def lib():
    """The loaded shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not found: build it with `python -m bddcml_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.bddcml_amd_set_option.argtypes = [C.c_char_p, C.c_double]
        L.bddcml_amd_last_error.argtypes = [C.c_char_p, C.c_int]
        dp = C.POINTER(C.c_double)
        L.bddcml_amd_selftest_tile.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, C.c_double,
                                               C.c_double, dp]
        L.bddcml_amd_selftest_copy.argtypes = [C.c_longlong, dp]
        L.bddcml_amd_element_gradient.argtypes = [dp, C.c_int, C.c_int, dp, dp, dp]
        L.bddcml_amd_element_gradient_noalloc.argtypes = [dp, C.c_int, C.c_int, dp, dp, dp]
        _lib = L
    return _lib


class _Checked:
    """Calls into the library and raises RuntimeError with the library's
    message when the call reported an error (option throw_on_error)."""

    def __init__(self, L):
        self._L = L

    def __getattr__(self, name):
        f = getattr(self._L, name)

        def call(*args):
            r = f(*args)
            buf = C.create_string_buffer(2048)
            if self._L.bddcml_amd_last_error(buf, 2048):
                raise RuntimeError(buf.value.decode(errors="replace"))
            return r

        call.__name__ = name
        return call


def selftest_tile(op, A, B, C_, X=None, alpha=1.0, beta=1.0):
    """Run one tile task on the device (see include/bddcml_amd.h)."""
    L = _Checked(lib())
    L.bddcml_amd_set_option(b"throw_on_error", 1.0)
    A = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(-1, 64, 64).transpose(0, 2, 1))
    B = np.ascontiguousarray(np.asarray(B, dtype=np.float64).reshape(-1, 64, 64).transpose(0, 2, 1))
    Cm = np.ascontiguousarray(np.asarray(C_, dtype=np.float64).T)
    Xm = np.ascontiguousarray((np.zeros((64, 64)) if X is None else np.asarray(X, dtype=np.float64)).T)
    out = np.zeros(2 * 4096)
    dp = C.POINTER(C.c_double)
    L.bddcml_amd_selftest_tile(op, A.shape[0], A.ctypes.data_as(dp), B.ctypes.data_as(dp),
                               Cm.ctypes.data_as(dp), Xm.ctypes.data_as(dp), alpha, beta,
                               out.ctypes.data_as(dp))
    return out[:4096].reshape(64, 64).T.copy(), out[4096:].reshape(64, 64).T.copy()


def selftest_symeig(a):
    """Host symmetric eigendecomposition of the library (see
    include/bddcml_amd.h): eigenvalues (unordered) and eigenvectors as
    columns."""
    L = _Checked(lib())
    L.bddcml_amd_set_option(b"throw_on_error", 1.0)
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    n = a.shape[0]
    w, v = np.zeros(n), np.zeros((n, n))
    dp = C.POINTER(C.c_double)
    L.bddcml_amd_selftest_symeig(n, a.ctypes.data_as(dp), w.ctypes.data_as(dp), v.ctypes.data_as(dp))
    return w, v


def selftest_ginv(k, rel_tol=1e-12):
    """Host generalized inverse of a symmetric positive semidefinite matrix
    (see include/bddcml_amd.h): (g, numerical rank)."""
    L = _Checked(lib())
    L.bddcml_amd_set_option(b"throw_on_error", 1.0)
    k = np.ascontiguousarray(np.asarray(k, dtype=np.float64))
    n = k.shape[0]
    g = np.zeros((n, n))
    r = C.c_int(0)
    dp = C.POINTER(C.c_double)
    L.bddcml_amd_selftest_ginv(n, k.ctypes.data_as(dp), C.c_double(rel_tol), g.ctypes.data_as(dp),
                               C.byref(r))
    return g, r.value


def selftest_copy(nbytes):
    """Time a plain device copy of nbytes (reference point for HBM counters);
    returns milliseconds."""
    L = _Checked(lib())
    L.bddcml_amd_set_option(b"throw_on_error", 1.0)
    out = np.zeros(1)
    L.bddcml_amd_selftest_copy(int(nbytes), out.ctypes.data_as(C.POINTER(C.c_double)))
    return float(out[0])


def _ip(v):
    return C.byref(C.c_int(int(v)))


def _iarr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def _darr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


class Bddcml:
    """One solver instance per process, as in the reference (module state)."""

    def __init__(self, throw_on_error: bool = True):
        self.L = _Checked(lib())
        self.numbase = 1
        self.L.bddcml_amd_set_option(b"throw_on_error", 1.0 if throw_on_error else 0.0)

    # -- reference interface (src/module_bddcml.f90) ------------------------
    def init(self, nl, nsublev, numbase=1, verbose_level=0, just_direct_solve=0):
        self.numbase = numbase
        ns, pns = _iarr(nsublev)
        nsub_loc = C.c_int(-1)
        self.L.bddcml_init_c(_ip(nl), pns, _ip(len(ns)), C.byref(nsub_loc), _ip(0),
                             _ip(verbose_level), _ip(numbase), _ip(just_direct_solve))
        return nsub_loc.value

    def upload_subdomain_data(self, d: dict):
        keep = []

        def I(x):
            a, p = _iarr(x)
            keep.append(a)
            return p

        def D(x):
            a, p = _darr(x)
            keep.append(a)
            return p

        xyz = np.asfortranarray(d["xyz"], dtype=np.float64).ravel(order="F")
        empty = np.zeros(1)
        dof_data = d.get("dof_data", np.zeros(0))
        # user constraints: (rows x nnods) values per local node, or absent
        uc = d.get("user_constraints")
        if uc is None:
            uc, nuc1, nuc2 = empty, 0, 0
        else:
            uc = np.asarray(uc, dtype=np.float64)
            nuc1, nuc2 = uc.shape
            uc = uc.ravel(order="F")
        self.L.bddcml_upload_subdomain_data_c(
            _ip(d["nelem"]), _ip(d["nnod"]), _ip(d["ndof"]), _ip(d["ndim"]), _ip(d["meshdim"]),
            _ip(d["isub"]), _ip(d["nelems"]), _ip(d["nnods"]), _ip(d["ndofs"]),
            I(d["inet"]), _ip(len(d["inet"])), I(d["nnet"]), _ip(len(d["nnet"])),
            I(d["nndf"]), _ip(len(d["nndf"])), I(d["isngn"]), _ip(len(d["isngn"])),
            I(d["isvgvn"]), _ip(len(d["isvgvn"])), I(d["isegn"]), _ip(len(d["isegn"])),
            D(xyz), _ip(d["nnods"]), _ip(d["ndim"]),
            I(d["ifix"]), _ip(len(d["ifix"])), D(d["fixv"]), _ip(len(d["fixv"])),
            D(d["rhs"]), _ip(len(d["rhs"])), _ip(d["is_rhs_complete"]),
            D(d["sol"]), _ip(len(d["sol"])), _ip(d["matrixtype"]),
            I(d["i_sparse"]), I(d["j_sparse"]), D(d["a_sparse"]), _ip(len(d["a_sparse"])),
            _ip(d["is_assembled"]), D(uc), _ip(nuc1), _ip(nuc2), D(empty), _ip(0), _ip(0),
            D(dof_data if len(dof_data) else empty), _ip(len(dof_data)),
            _ip(d["find_components"]), _ip(d["use_dual_mesh_graph"]), _ip(d["neighbouring"]))

    def upload_global_data(self, g: dict, idelm: int, neighbouring: int = 1,
                           load_division: int = 0):
        """The whole mesh at once (bddcml_upload_global_data_c): g holds
        nelem, nnod, ndof, ndim, meshdim, inet, nnet, nndf, xyz (nnod x ndim),
        ifix, fixv, rhs, sol; idelm is an open file descriptor of the ELM file
        of element matrices (gfortran unformatted sequential records).  The
        library divides the mesh into this level's subdomains (the file
        partition_l1.ES in the working directory with load_division)."""
        keep = []

        def I(x):
            a, p = _iarr(x)
            keep.append(a)
            return p

        def D(x):
            a, p = _darr(x)
            keep.append(a)
            return p

        xyz = np.asfortranarray(g["xyz"], dtype=np.float64)
        self.L.bddcml_upload_global_data_c(
            _ip(g["nelem"]), _ip(g["nnod"]), _ip(g["ndof"]), _ip(g["ndim"]), _ip(g["meshdim"]),
            I(g["inet"]), _ip(len(g["inet"])), I(g["nnet"]), _ip(len(g["nnet"])),
            I(g["nndf"]), _ip(len(g["nndf"])), D(xyz.ravel(order="F")), _ip(xyz.shape[0]),
            _ip(xyz.shape[1]), I(g["ifix"]), _ip(len(g["ifix"])), D(g["fixv"]), _ip(len(g["fixv"])),
            D(g["rhs"]), _ip(len(g["rhs"])), D(g["sol"]), _ip(len(g["sol"])), _ip(idelm),
            _ip(neighbouring), _ip(load_division))

    def change_global_data(self, ifix, fixv, rhs, sol):
        """New global Dirichlet values, right-hand side and initial guess of
        a globally loaded problem, same fixed dofs
        (bddcml_change_global_data_c); then setup_new_data."""
        keep = []
        ia, ip_ = _iarr(ifix)
        fa, fp = _darr(fixv)
        ra, rp = _darr(rhs)
        sa, sp = _darr(sol)
        keep += [ia, fa, ra, sa]
        self.L.bddcml_change_global_data_c(ip_, _ip(len(ia)), fp, _ip(len(fa)), rp, _ip(len(ra)),
                                           sp, _ip(len(sa)))

    def change_subdomain_data(self, d: dict):
        """New Dirichlet values, right-hand side and initial guess of one
        subdomain (same fixed dofs), as bddcml_change_subdomain_data_c."""
        keep = []

        def I(x):
            a, p = _iarr(x)
            keep.append(a)
            return p

        def D(x):
            a, p = _darr(x)
            keep.append(a)
            return p

        n = len(d["ifix"])
        self.L.bddcml_change_subdomain_data_c(
            _ip(d["isub"]), I(d["ifix"]), _ip(n), D(d["fixv"]), _ip(n), D(d["rhs"]), _ip(n),
            _ip(d["is_rhs_complete"]), D(d["sol"]), _ip(n))

    def setup_new_data(self):
        self.L.bddcml_setup_new_data_c()

    def setup_preconditioner(self, matrixtype=1, use_defaults=0, parallel_division=1,
                             use_corner_constraints=1, use_arithmetic_constraints=1,
                             use_adaptive_constraints=0, use_user_constraints=0, weights_type=0):
        self.L.bddcml_setup_preconditioner_c(
            _ip(matrixtype), _ip(use_defaults), _ip(parallel_division), _ip(use_corner_constraints),
            _ip(use_arithmetic_constraints), _ip(use_adaptive_constraints),
            _ip(use_user_constraints), _ip(weights_type))

    def solve(self, method=0, tol=1e-8, maxit=500, ndecrmax=50, recycling=0,
              max_number_of_stored_vectors=50):
        it, reason, cond = C.c_int(), C.c_int(), C.c_double()
        self.L.bddcml_solve_c(_ip(0), _ip(method), C.byref(C.c_double(tol)), _ip(maxit),
                              _ip(ndecrmax), _ip(recycling), _ip(max_number_of_stored_vectors),
                              C.byref(it), C.byref(reason), C.byref(cond))
        return it.value, reason.value, cond.value

    def download_local_reactions(self, isub, ndofs):
        out = np.zeros(ndofs)
        self.L.bddcml_download_local_reactions_c(_ip(isub), out.ctypes.data_as(
            C.POINTER(C.c_double)), _ip(ndofs))
        return out

    def download_global_reactions(self, ndof):
        out = np.zeros(ndof)
        self.L.bddcml_download_global_reactions_c(out.ctypes.data_as(C.POINTER(C.c_double)),
                                                  _ip(ndof))
        return out

    def download_local_solution(self, isub, ndofs):
        out = np.zeros(ndofs)
        _, p = _darr(out)
        self.L.bddcml_download_local_solution_c(_ip(isub), out.ctypes.data_as(
            C.POINTER(C.c_double)), _ip(ndofs))
        return out

    def download_global_solution(self, ndof):
        out = np.zeros(ndof)
        self.L.bddcml_download_global_solution_c(out.ctypes.data_as(C.POINTER(C.c_double)),
                                                 _ip(ndof))
        return out

    def dotprod_subdomain(self, isub, v1, v2):
        a, pa = _darr(v1)
        b, pb = _darr(v2)
        out = C.c_double()
        self.L.bddcml_dotprod_subdomain_c(_ip(isub), pa, _ip(len(a)), pb, _ip(len(b)),
                                          C.byref(out))
        return out.value

    def finalize(self):
        self.L.bddcml_finalize_c()

    # -- extra interface (include/bddcml_amd.h) -----------------------------
    def set_option(self, key: str, value: float):
        return self.L.bddcml_amd_set_option(key.encode(), float(value))

    def counts(self):
        out = np.zeros(5, dtype=np.int32)
        rc = self.L.bddcml_amd_counts(out.ctypes.data_as(C.POINTER(C.c_int)))
        if rc:
            raise RuntimeError("not set up")
        return dict(ncorner=int(out[0]), nedge=int(out[1]), nface=int(out[2]),
                    ndofc=int(out[3]), nig=int(out[4]))

    def level2_counts(self):
        """Second level of a multilevel setup (None without one)."""
        return self.level_counts(2)

    def level_counts(self, ilevel):
        """Level ilevel >= 2 of a multilevel setup (None without one)."""
        out = np.zeros(7, dtype=np.int32)
        if self.L.bddcml_amd_level_counts(int(ilevel), out.ctypes.data_as(C.POINTER(C.c_int))):
            return None
        return dict(nsub=int(out[0]), ndof=int(out[1]), ncorner=int(out[2]), nedge=int(out[3]),
                    nface=int(out[4]), ndofc=int(out[5]), nig=int(out[6]))

    def adaptive_pairs(self):
        """Face pairs of an adaptive setup: list of dicts (face coarse node,
        isub, jsub, neig, nadaptive, iterations, eigenvalues largest first)."""
        n = self.L.bddcml_amd_adaptive_pairs(-1, None, None)
        out = []
        for k in range(max(n, 0)):
            o = np.zeros(6, dtype=np.int32)
            e = np.zeros(32)
            self.L.bddcml_amd_adaptive_pairs(k, o.ctypes.data_as(C.POINTER(C.c_int)),
                                             e.ctypes.data_as(C.POINTER(C.c_double)))
            out.append(dict(face=int(o[0]), isub=int(o[1]), jsub=int(o[2]), neig=int(o[3]),
                            nadaptive=int(o[4]), iterations=int(o[5]), eigenvalues=e[:o[3]].copy()))
        return out

    def interface_dofs(self):
        out = np.zeros(self.counts()["nig"], dtype=np.int32)
        self.L.bddcml_amd_interface_dofs(out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def _vec_call(self, fn, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        rc = fn(x.ctypes.data_as(C.POINTER(C.c_double)), y.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError(f"{fn.__name__} failed: {rc}")
        return y

    def apply_schur(self, x):
        return self._vec_call(self.L.bddcml_amd_apply_schur, x)

    def apply_pc(self, r):
        return self._vec_call(self.L.bddcml_amd_apply_pc, r)

    def reduced_rhs(self):
        g = np.zeros(self.counts()["nig"])
        if self.L.bddcml_amd_reduced_rhs(g.ctypes.data_as(C.POINTER(C.c_double))):
            raise RuntimeError("reduced rhs not available")
        return g

    def interface_solution(self):
        u = np.zeros(self.counts()["nig"])
        if self.L.bddcml_amd_interface_solution(u.ctypes.data_as(C.POINTER(C.c_double))):
            raise RuntimeError("no solution")
        return u

    def recycling_basis(self):
        """(number of stored recycling directions, ||V^T W - I||_F)"""
        n, orth = C.c_int(), C.c_double()
        self.L.bddcml_amd_recycling_basis(C.byref(n), C.byref(orth))
        return n.value, orth.value

    def residual_history(self):
        n = self.L.bddcml_amd_residual_history(None, 0)
        out = np.zeros(n)
        self.L.bddcml_amd_residual_history(out.ctypes.data_as(C.POINTER(C.c_double)), n)
        return out

    def timings(self):
        keys = ("topology", "plan", "factor", "krylov", "postprocess", "setup_total",
                "solve_total", "rhs", "ev_interior_ms", "ev_schur_ms", "ev_interface_ms",
                "ev_coarse_ms", "schur_flops", "interior_flops", "interface_flops",
                "ev_gemv_ms", "gemv_bytes", "iterations", "ev_kapply_ms", "kapply_bytes",
                "level2", "ev_gemm_ms", "gemm_flops", "gemm_launches", "refinements",
                "refine_relres", "refine", "ev_resid_ms", "resid_bytes")
        out = np.zeros(len(keys))
        self.L.bddcml_amd_timings(out.ctypes.data_as(C.POINTER(C.c_double)), len(keys))
        return dict(zip(keys, out.tolist()))

    def sub_info(self, isub):
        out = np.zeros(3, dtype=np.int32)
        if self.L.bddcml_amd_sub_info(int(isub), out.ctypes.data_as(C.POINTER(C.c_int))):
            raise RuntimeError("sub_info failed")
        return dict(ndofi=int(out[0]), nconstr=int(out[1]), ndofo=int(out[2]))

    def sub_maps(self, isub):
        info = self.sub_info(isub)
        g = np.zeros(info["ndofi"], dtype=np.int32)
        c = np.zeros(max(info["nconstr"], 1), dtype=np.int32)
        self.L.bddcml_amd_sub_maps(int(isub), g.ctypes.data_as(C.POINTER(C.c_int)),
                                   c.ctypes.data_as(C.POINTER(C.c_int)))
        return g, c[:info["nconstr"]]

    def sub_matrix(self, isub, which="S"):
        n = self.sub_info(isub)["ndofi"]
        out = np.zeros(n * n)
        rc = self.L.bddcml_amd_sub_matrix(int(isub), ord(which),
                                          out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError("sub_matrix failed")
        return out.reshape(n, n)

    def sub_coarse(self, isub):
        info = self.sub_info(isub)
        n, m = info["ndofi"], info["nconstr"]
        phi = np.zeros(max(n * m, 1))
        ac = np.zeros(max(m * m, 1))
        rc = self.L.bddcml_amd_sub_coarse(int(isub), phi.ctypes.data_as(C.POINTER(C.c_double)),
                                          ac.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError("sub_coarse failed")
        return phi[:n * m].reshape(n, m), ac[:m * m].reshape(m, m)


def solve_cube(problem, use_corner=1, use_arith=1, weights_type=0, method=0, tol=1e-8,
               maxit=500, ndecrmax=50, verbose=0, options=None, nsub2=0, nsublev=None):
    """Drive the drop-in interface on a CubeProblem the way the reference
    harness does (examples/poisson_on_cube.f90:373-560); returns the Bddcml
    handle and (num_iter, converged_reason, condition_number).  nsub2 > 0
    asks for three levels with nsub2 subdomains on the second level
    (nsublev = [nsub, nsub2, 1]); nsublev gives every level explicitly."""
    b = Bddcml()
    if nsublev is not None:
        b.init(len(nsublev), list(nsublev), numbase=problem.numbase, verbose_level=verbose)
    elif nsub2 > 0:
        b.init(3, [problem.nsub, nsub2, 1], numbase=problem.numbase, verbose_level=verbose)
    else:
        b.init(2, [problem.nsub, 1], numbase=problem.numbase, verbose_level=verbose)
    for k, v in (options or {}).items():
        if b.set_option(k, v) != 0:
            raise ValueError(f"option {k}={v} rejected")
    for d in problem.subdomains():
        b.upload_subdomain_data(d)
    b.setup_preconditioner(1, 0, 1, use_corner, use_arith, 0, 0, weights_type)
    res = b.solve(method, tol, maxit, ndecrmax)
    return b, res
