"""Global loading restated on the CPU: the whole mesh divided into subdomains.

TEST INFRASTRUCTURE ONLY: the checker for the library's global-loading path
(bddcml_upload_global_data_c / bddcml_change_global_data_c).  The product
(bddcml_amd) never imports this module.

Restated from the reference (paths relative to its root):
- the ELM file of element matrices: one unformatted sequential record per
  element, in element order, holding lelm and lelm doubles, the upper
  triangle column by column (sm_pmd_load / sm_pmd_make_element_numbering,
  src/module_sm.f90:170-214, 272-375; dd_load_matrix_from_file,
  src/module_dd.f90:1030-1060), with gfortran's 4-byte record markers;
- the division: partition_l1.ES when given (src/module_levels.f90:1568-1582);
  where the reference calls METIS, the library bisects the element
  centroids (node coordinates summed in INET order over NNET) recursively,
  restated with :func:`oracle.multilevel.bisect`;
- each subdomain as pp_create_submesh builds it (src/module_pp.f90:
  1443-1503): elements ascending, nodes the sorted distinct global nodes;
  IFIX/FIXV/SOL restricted, the right-hand side restricted and complete
  (levels_localize_new_data, src/module_levels.f90:3576-3625); components
  on the nodal graph (find_components 1, use_dual_mesh_graph 0).
"""
from __future__ import annotations

import numpy as np

from .multilevel import bisect


def read_elm(path: str, nelem: int) -> list:
    """Element matrices of an ELM file as a list of 1-D arrays (the stored
    upper triangle by columns)."""
    raw = np.fromfile(path, dtype=np.uint8)
    out, pos = [], 0
    for e in range(nelem):
        head, lelm = np.frombuffer(raw[pos:pos + 8].tobytes(), dtype="<i4")
        if head != 4 + 8 * lelm:
            raise ValueError(f"record {e + 1}: bad record marker")
        a = np.frombuffer(raw[pos + 8:pos + 8 + 8 * lelm].tobytes(), dtype="<f8").copy()
        (tail,) = np.frombuffer(raw[pos + 8 + 8 * lelm:pos + 12 + 8 * lelm].tobytes(), dtype="<i4")
        if tail != head:
            raise ValueError(f"record {e + 1}: bad closing marker")
        out.append(a)
        pos += 12 + 8 * lelm
    return out


def element_centroids(g: dict, numbase: int = 1) -> np.ndarray:
    """Mean of each element's node coordinates, summed in INET order."""
    inet = np.asarray(g["inet"]) - numbase
    nnet = np.asarray(g["nnet"])
    xyz = np.asarray(g["xyz"])
    cent = np.zeros((len(nnet), xyz.shape[1]))
    p = 0
    for e, n in enumerate(nnet):
        acc = np.zeros(xyz.shape[1])
        for k in range(n):
            acc = acc + xyz[inet[p + k]]
        cent[e] = acc / n
        p += n
    return cent


def divide(g: dict, nsub: int, numbase: int = 1, iets=None) -> np.ndarray:
    """Subdomain (0-based) of every element: iets when given (the content of
    partition_l1.ES minus one), else the bisection of the centroids."""
    if iets is not None:
        return np.asarray(iets, dtype=np.int64)
    return bisect(element_centroids(g, numbase), nsub)


def localize(g: dict, elm: list, iets: np.ndarray, nsub: int, numbase: int = 1,
             neighbouring: int = 1) -> list:
    """Subdomain upload dictionaries (numbase-based) of a divided global
    mesh, as the library builds them."""
    b = numbase
    inet = np.asarray(g["inet"]) - b
    nnet = np.asarray(g["nnet"])
    nndf = np.asarray(g["nndf"])
    xyz = np.asarray(g["xyz"])
    eptr = np.concatenate([[0], np.cumsum(nnet)])
    kdof = np.concatenate([[0], np.cumsum(nndf)])
    subs = []
    for s in range(nsub):
        el = np.flatnonzero(iets == s)
        if len(el) == 0:
            raise ValueError(f"Partition does not contain subdomain {s + b}")
        nodes = np.unique(np.concatenate([inet[eptr[e]:eptr[e + 1]] for e in el]))
        loc = {int(n): k for k, n in enumerate(nodes)}
        snndf = nndf[nodes]
        lkdof = np.concatenate([[0], np.cumsum(snndf)])
        gd = np.concatenate([kdof[n] + np.arange(nndf[n]) for n in nodes]).astype(np.int64)
        ii, jj, aa, sinet = [], [], [], []
        for e in el:
            kdofe = []
            for n in inet[eptr[e]:eptr[e + 1]]:
                ln = loc[int(n)]
                sinet.append(ln)
                kdofe += list(range(lkdof[ln], lkdof[ln] + snndf[ln]))
            nve = len(kdofe)
            a = elm[e]
            if len(a) != nve * (nve + 1) // 2:
                raise ValueError(f"element {e + 1}: not symmetric storage")
            p = 0
            for j in range(nve):
                for i in range(j + 1):
                    r, c = min(kdofe[i], kdofe[j]), max(kdofe[i], kdofe[j])
                    ii.append(r)
                    jj.append(c)
                    aa.append(a[p])
                    p += 1
        ndofs = int(lkdof[-1])
        subs.append(dict(
            nelem=int(g["nelem"]), nnod=int(g["nnod"]), ndof=int(g["ndof"]), ndim=int(g["ndim"]),
            meshdim=int(g["meshdim"]), isub=s + b, nelems=len(el), nnods=len(nodes), ndofs=ndofs,
            inet=np.asarray(sinet) + b, nnet=nnet[el], nndf=snndf, isngn=nodes + b,
            isvgvn=gd + b, isegn=el + b, xyz=np.asfortranarray(xyz[nodes]),
            ifix=np.asarray(g["ifix"])[gd], fixv=np.asarray(g["fixv"])[gd],
            rhs=np.asarray(g["rhs"])[gd], is_rhs_complete=1, sol=np.asarray(g["sol"])[gd],
            matrixtype=1, i_sparse=np.asarray(ii) + b, j_sparse=np.asarray(jj) + b,
            a_sparse=np.asarray(aa), is_assembled=0, find_components=1, use_dual_mesh_graph=0,
            neighbouring=neighbouring))
    return subs
