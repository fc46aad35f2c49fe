"""Problem files of the reference's file-driven drivers, and its ParaView
export.

Reading and writing (the reference's examples/bddcml_global.f90:223-320 with
src/module_pp.f90:543-1020; examples/bddcml_global.f90:452-476 for SOL):

  NAME.PAR   list-directed: ndim nsub nelem ndof nnod nnodc linet tol maxit
             ndecrmax, then (next record, optional) meshdim
  NAME.ML    list-directed: nlevels, then (next record) nsublev(1:nlevels)
  NAME.GMIS  list-directed, one read each: inet(linet), nnet(nelem),
             nndf(nnod), xyz(nnod, ndim) (column major: all x, then all y ...)
  NAME.FVS   list-directed: ifix(ndof), then fixv(ndof)
  NAME.RHS   unformatted sequential: one record rhs(ndof)
  NAME.ELM   unformatted sequential: per element one record lelm, elm(1:lelm)
             (read by the library itself from a file descriptor)
  NAME.SOL   unformatted sequential: sol(ndof); rea(ndof), sum of reactions
             per coordinate axis (3)

List-directed input follows the Fortran rules the reference's READ
statements rely on: values separated by blanks, commas or line ends; r*c
repeat counts; D exponents; a slash ends the read; and every READ statement
starts on a new record, discarding what is left of the record where the
previous one stopped.  Unformatted records carry 4-byte length markers
before and after (records above 2 GiB as subrecords with a negative leading
marker for all but the last, the gfortran convention).

solve_problem() runs the sequence of examples/bddcml_global.f90 on these
files through this library: init with the ML levels, upload_global_data
with the ELM file, setup_preconditioner, solve with the PAR tolerances,
download of the global solution and reactions, the SOL file.

The VTU / PVD writers reproduce src/module_paraview.f90:41-490 text for
text (ASCII XML; coordinates and float data in Fortran E15.7, integers in
I15; files PREFIX/PREFIX_nnnnn.vtu as getfname names them,
src/module_utils.f90:66-155), the way examples/poisson_on_cube.f90:925-965
exports each subdomain's solution.
"""
from __future__ import annotations

import os
import re
import struct

import numpy as np

__all__ = [
    "read_par", "write_par", "read_ml", "write_ml", "read_gmis", "write_gmis", "read_fvs",
    "write_fvs", "read_rhs", "write_rhs", "read_sol", "write_sol", "load_problem",
    "write_problem", "solve_problem", "fortran_e", "subdomain_vtu_name", "write_subdomain_vtu",
    "write_pvd",
]


# ----------------------------------------------------------- list-directed
class ListDirectedReader:
    """Sequential list-directed reads from a text file: read(n) returns the
    next n values of one READ statement, which starts on a new record."""

    _token = re.compile(r"[^\s,]+")

    def __init__(self, path: str):
        with open(path, "r") as f:
            self.lines = f.read().splitlines()
        self.pos = 0  # next record

    def read(self, n: int, kind=float, optional: bool = False):
        out = []
        line = self.pos
        while len(out) < n:
            if line >= len(self.lines):
                if optional:
                    return None
                raise ValueError(f"end of file after {len(out)} of {n} values")
            stop = False
            for tok in self._token.findall(self.lines[line]):
                if tok.startswith("/"):
                    stop = True
                    break
                rep, star, val = tok.partition("*")
                cnt = int(rep) if star else 1
                if not star:
                    val = rep
                if val == "":
                    raise ValueError("null values are not supported")
                for _ in range(cnt):
                    out.append(val)
                if len(out) >= n:
                    break
            line += 1
            if stop:
                break
        if len(out) < n:
            if optional:
                return None
            raise ValueError(f"read ended after {len(out)} of {n} values")
        self.pos = line
        vals = out[:n]
        try:
            if kind is int:
                return np.array([int(v) for v in vals], dtype=np.int64)
            return np.array([float(v.replace("D", "E").replace("d", "e")) for v in vals])
        except ValueError:
            if optional:
                return None
            raise


def _write_values(f, vals, per_line=8, kind=float):
    vals = np.asarray(vals).ravel()
    for k in range(0, max(len(vals), 1), per_line):
        chunk = vals[k:k + per_line]
        if kind is int:
            f.write(" ".join(str(int(v)) for v in chunk))
        else:
            f.write(" ".join(repr(float(v)) for v in chunk))
        f.write("\n")


# -------------------------------------------------------------- PAR, ML
def read_par(path: str) -> dict:
    """src/module_pp.f90:655-706 (meshdim defaults to ndim when its record is
    missing or unreadable)."""
    r = ListDirectedReader(path)
    head = r.read(10, kind=float)
    keys = ("ndim", "nsub", "nelem", "ndof", "nnod", "nnodc", "linet", "tol", "maxit", "ndecrmax")
    d = {k: (float(v) if k == "tol" else int(v)) for k, v in zip(keys, head)}
    for k in keys:
        if k != "tol" and float(d[k]) != head[keys.index(k)]:
            raise ValueError(f"{k} is not an integer")
    md = r.read(1, kind=int, optional=True)
    d["meshdim"] = int(md[0]) if md is not None else d["ndim"]
    return d


def write_par(path: str, ndim, nsub, nelem, ndof, nnod, linet, tol, maxit, ndecrmax, meshdim=None,
              nnodc=0):
    with open(path, "w") as f:
        f.write(f"{ndim} {nsub} {nelem} {ndof} {nnod} {nnodc} {linet} {float(tol)!r} {maxit} {ndecrmax}\n")
        if meshdim is not None:
            f.write(f"{meshdim}\n")


def read_ml(path: str):
    """examples/bddcml_global.f90:250-262: (nlevels, nsublev)."""
    r = ListDirectedReader(path)
    nlevels = int(r.read(1, kind=int)[0])
    return nlevels, r.read(nlevels, kind=int).astype(np.int32)


def write_ml(path: str, nsublev):
    nsublev = np.asarray(nsublev)
    with open(path, "w") as f:
        f.write(f"{len(nsublev)}\n")
        _write_values(f, nsublev, kind=int)


# ------------------------------------------------------------ GMIS, FVS
def read_gmis(path: str, linet: int, nelem: int, nnod: int, ndim: int) -> dict:
    """src/module_pp.f90:543-574, 956-987: inet, nnet, nndf, xyz (nnod x
    ndim)."""
    r = ListDirectedReader(path)
    inet = r.read(linet, kind=int).astype(np.int32)
    nnet = r.read(nelem, kind=int).astype(np.int32)
    nndf = r.read(nnod, kind=int).astype(np.int32)
    xyz = r.read(nnod * ndim).reshape(ndim, nnod).T.copy()
    return dict(inet=inet, nnet=nnet, nndf=nndf, xyz=xyz)


def write_gmis(path: str, inet, nnet, nndf, xyz):
    xyz = np.asarray(xyz, dtype=np.float64)
    with open(path, "w") as f:
        _write_values(f, inet, per_line=8, kind=int)
        _write_values(f, nnet, per_line=16, kind=int)
        _write_values(f, nndf, per_line=16, kind=int)
        _write_values(f, xyz.T.ravel(), per_line=4)


def read_fvs(path: str, ndof: int):
    """src/module_pp.f90:761-819: (ifix, fixv)."""
    r = ListDirectedReader(path)
    ifix = r.read(ndof, kind=int).astype(np.int32)
    fixv = r.read(ndof)
    return ifix, fixv


def write_fvs(path: str, ifix, fixv):
    with open(path, "w") as f:
        _write_values(f, ifix, per_line=16, kind=int)
        _write_values(f, fixv, per_line=4)


# ----------------------------------------------------- unformatted records
_SUB = 2147483639  # largest subrecord payload (bytes) of the gfortran format


def read_records(path: str) -> list:
    """All records of an unformatted sequential file as bytes objects."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        payload = b""
        while True:
            if p + 4 > len(data):
                raise ValueError("truncated record marker")
            (m,) = struct.unpack_from("<i", data, p)
            n = abs(m)
            if p + 8 + n > len(data):
                raise ValueError("truncated record")
            (m2,) = struct.unpack_from("<i", data, p + 4 + n)
            if abs(m2) != n:
                raise ValueError("record markers differ")
            payload += data[p + 4:p + 4 + n]
            p += 8 + n
            if m >= 0:
                break
        out.append(payload)
    return out


def _write_record(f, payload: bytes):
    if len(payload) <= _SUB:
        f.write(struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload)))
        return
    parts = [payload[k:k + _SUB] for k in range(0, len(payload), _SUB)]
    for i, part in enumerate(parts):
        n = len(part)
        lead = -n if i < len(parts) - 1 else n
        tail = -n if i > 0 else n
        f.write(struct.pack("<i", lead) + part + struct.pack("<i", tail))


def read_rhs(path: str, ndof: int) -> np.ndarray:
    """src/module_pp.f90:892-918: one record, its first ndof doubles."""
    rec = read_records(path)
    if not rec or len(rec[0]) < 8 * ndof:
        raise ValueError("RHS record shorter than ndof doubles")
    return np.frombuffer(rec[0], dtype="<f8", count=ndof).copy()


def write_rhs(path: str, rhs):
    with open(path, "wb") as f:
        _write_record(f, np.asarray(rhs, dtype="<f8").tobytes())


def read_sol(path: str, ndof: int):
    """(sol, rea, sum of reactions per axis) of a SOL file."""
    rec = read_records(path)
    sol = np.frombuffer(rec[0], dtype="<f8", count=ndof).copy()
    tail = np.frombuffer(rec[1], dtype="<f8")
    return sol, tail[:ndof].copy(), tail[ndof:ndof + 3].copy()


def write_sol(path: str, sol, rea, sum_rea):
    with open(path, "wb") as f:
        _write_record(f, np.asarray(sol, dtype="<f8").tobytes())
        _write_record(f, np.concatenate([np.asarray(rea, dtype="<f8"),
                                         np.asarray(sum_rea, dtype="<f8")[:3]]).tobytes())


# ------------------------------------------------------- whole problems
def load_problem(name: str, directory: str = ".") -> dict:
    """PAR, ML, GMIS, FVS and RHS of problem `name`: the PAR fields, nlevels,
    nsublev, and the arrays of bddcml_upload_global_data_c (sol zero)."""
    base = os.path.join(directory, name)
    par = read_par(base + ".PAR")
    nlevels, nsublev = read_ml(base + ".ML")
    if int(nsublev[0]) != par["nsub"]:
        raise ValueError("number of subdomains at first level mismatch")
    g = read_gmis(base + ".GMIS", par["linet"], par["nelem"], par["nnod"], par["ndim"])
    ifix, fixv = read_fvs(base + ".FVS", par["ndof"])
    rhs = read_rhs(base + ".RHS", par["ndof"])
    g.update(nelem=par["nelem"], nnod=par["nnod"], ndof=par["ndof"], ndim=par["ndim"],
             meshdim=par["meshdim"], ifix=ifix, fixv=fixv, rhs=rhs, sol=np.zeros(par["ndof"]))
    return dict(par=par, nlevels=nlevels, nsublev=nsublev, global_data=g, elm=base + ".ELM")


def write_problem(name: str, directory: str, g: dict, nsublev, tol=1e-8, maxit=500, ndecrmax=50):
    """PAR, ML, GMIS, FVS and RHS of a problem given as the arrays of
    bddcml_upload_global_data_c (the ELM file is written separately)."""
    base = os.path.join(directory, name)
    write_par(base + ".PAR", g["ndim"], int(nsublev[0]), g["nelem"], g["ndof"], g["nnod"],
              len(g["inet"]), tol, maxit, ndecrmax, meshdim=g.get("meshdim"))
    write_ml(base + ".ML", nsublev)
    write_gmis(base + ".GMIS", g["inet"], g["nnet"], g["nndf"], g["xyz"])
    write_fvs(base + ".FVS", g["ifix"], g["fixv"])
    write_rhs(base + ".RHS", g["rhs"])


def solve_problem(name: str, directory: str = ".", matrixtype=1, krylov_method=0,
                  corner=1, arithmetic=1, adaptive=0, weights_type=0, load_division=0,
                  neighbouring=1, options=None, write_sol_file=True, plan_only=False):
    """examples/bddcml_global.f90 on the problem files through this library
    (numbase 1; neighbouring as the reference driver's second command-line
    argument): returns dict(iterations, reason, condition_number, sol, rea,
    sum_rea) and writes NAME.SOL unless write_sol_file is False.  With
    plan_only the host plan is built and the counts are returned instead."""
    import bddcml_amd
    pr = load_problem(name, directory)
    par, g = pr["par"], pr["global_data"]
    b = bddcml_amd.Bddcml()
    try:
        b.init(pr["nlevels"], [int(x) for x in pr["nsublev"]], numbase=1)
        if plan_only:
            b.set_option("plan_only", 1)
        for k, v in (options or {}).items():
            b.set_option(k, v)
        fd = os.open(pr["elm"], os.O_RDONLY)
        try:
            b.upload_global_data(g, fd, neighbouring=neighbouring, load_division=load_division)
        finally:
            os.close(fd)
        b.setup_preconditioner(matrixtype, 0, 1, corner, arithmetic, adaptive, 0, weights_type)
        if plan_only:
            return dict(counts=b.counts())
        it, reason, cond = b.solve(krylov_method, par["tol"], par["maxit"], par["ndecrmax"])
        sol = b.download_global_solution(par["ndof"])
        rea = b.download_global_reactions(par["ndof"])
    finally:
        b.finalize()
    # reactions summed per coordinate direction over the nodes
    # (examples/bddcml_global.f90:444-451)
    sum_rea = np.zeros(3)
    k = 0
    for nd in g["nndf"]:
        for i in range(int(nd)):
            if i < 3:
                sum_rea[i] += rea[k]
            k += 1
    if write_sol_file:
        write_sol(os.path.join(directory, name + ".SOL"), sol, rea, sum_rea)
    return dict(iterations=it, reason=reason, condition_number=cond, sol=sol, rea=rea,
                sum_rea=sum_rea)


# ---------------------------------------------------------------- ParaView
def fortran_e(x: float, w: int = 15, d: int = 7) -> str:
    """Fortran Ew.d output of one value: 0.ddddddd mantissa, E+xx exponent
    (the E dropped for three-digit exponents), right-justified in w."""
    x = float(x)
    if x == 0.0:
        s = ("-" if np.signbit(x) else "") + "0." + "0" * d + "E+00"
        return s.rjust(w)
    t = "%.*E" % (d - 1, x)  # one digit, point, d-1 digits, E, exponent
    neg = t.startswith("-")
    t = t.lstrip("-")
    mant, ex = t.split("E")
    digits = mant.replace(".", "")
    e = int(ex) + 1
    if abs(e) <= 99:
        es = "E" + ("+" if e >= 0 else "-") + "%02d" % abs(e)
    else:
        es = ("+" if e >= 0 else "-") + "%03d" % abs(e)
    s = ("-" if neg else "") + "0." + digits + es
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def _fmt_rows(vals, per_line, fmt):
    out = []
    for k in range(0, len(vals), per_line):
        out.append("".join(fmt(v) for v in vals[k:k + per_line]))
    return out


def subdomain_vtu_name(prefix: str, isub: int) -> str:
    """PREFIX/PREFIX_nnnnn.vtu with isub zero-padded to five digits."""
    return f"{prefix}/{prefix}_{isub:05d}.vtu"


def _dataarray(lines, array, name, ncomp, is_int):
    typ = "Int32" if is_int else "Float64"
    lines.append(f'      <DataArray type="{typ}" NumberOfComponents="{ncomp}" Name="{name}" format="ascii">')
    a = np.asarray(array).ravel()
    npoint = len(a) // ncomp
    for i in range(npoint):
        vals = [a[c * npoint + i] for c in range(ncomp)]
        if is_int:
            lines.extend(_fmt_rows([int(v) for v in vals], 3, lambda v: "%15d" % v))
        else:
            lines.extend(_fmt_rows(vals, 3, fortran_e))
    lines.append("      </DataArray>")


def write_subdomain_vtu(prefix: str, isub: int, inet, nnet, xyz, solution, meshdim: int = 3,
                        directory: str = "."):
    """One subdomain's mesh (local 1-based inet) with its element numbers as
    cell data 'subdomain' and its local solution as point data 'Solution'
    (src/module_paraview.f90:150-249, 308-470; the sequence of
    examples/poisson_on_cube.f90:925-965).  Returns the file path."""
    inet = np.asarray(inet, dtype=np.int64)
    nnet = np.asarray(nnet, dtype=np.int64)
    xyz = np.asarray(xyz, dtype=np.float64)
    nnod, nelem = xyz.shape[0], len(nnet)
    L = ['<?xml version="1.0"?>', '<VTKFile type="UnstructuredGrid" byte_order="LittleEndian">',
         "  <UnstructuredGrid>",
         '  <Piece NumberOfPoints="%10d" NumberOfCells="%10d">' % (nnod, nelem), "    <Points>",
         '      <DataArray type="Float64" NumberOfComponents="%d" Name="Coordinates" format="ascii">'
         % xyz.shape[1]]
    for i in range(nnod):
        L.extend(_fmt_rows(list(xyz[i]), 3, fortran_e))
    L += ["      </DataArray>", "    </Points>", "    <Cells>",
          '      <DataArray type="Int32" NumberOfComponents="1" Name="connectivity" format="ascii">']
    k = 0
    for ie in range(nelem):
        nne = int(nnet[ie])
        L.extend(_fmt_rows([int(v) - 1 for v in inet[k:k + nne]], 10, lambda v: "%15d" % v))
        k += nne
    L += ["      </DataArray>",
          '      <DataArray type="Int32" NumberOfComponents="1" Name="offsets" format="ascii">']
    off = 0
    for ie in range(nelem):
        off += int(nnet[ie])
        L.append("%15d" % off)
    L += ["      </DataArray>",
          '      <DataArray type="Int32" NumberOfComponents="1" Name="types" format="ascii">']
    for ie in range(nelem):
        nne = int(nnet[ie])
        if nne == 8:
            t = 12
        elif nne == 4:
            t = 10 if meshdim == 3 else (9 if meshdim == 2 else 0)
        elif nne == 3:
            t = 5
        else:
            t = 41
        L.append("%15d" % t)
    L += ["      </DataArray>", "    </Cells>", "    <CellData>"]
    _dataarray(L, np.full(nelem, isub), "subdomain", 1, True)
    L += ["    </CellData>", "    <PointData>"]
    _dataarray(L, solution, "Solution", 1, False)
    L += ["    </PointData>", "  </Piece>", " </UnstructuredGrid>", "</VTKFile>"]
    path = os.path.join(directory, subdomain_vtu_name(prefix, isub))
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
    return path


def write_pvd(prefix: str, nsub: int, directory: str = "."):
    """The umbrella PVD file over the subdomain VTU files
    (src/module_paraview.f90:110-147).  Returns its path."""
    L = ['<?xml version="1.0"?>',
         '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">', "  <Collection>"]
    for isub in range(1, nsub + 1):
        L.append('    <DataSet part="%10d" file="./%s"/>' % (isub - 1, subdomain_vtu_name(prefix, isub)))
    L += ["  </Collection>", "</VTKFile>"]
    path = os.path.join(directory, prefix + ".pvd")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
    return path
