"""Structured cube problems, generated per subdomain exactly the way the
reference's example harness does it (examples/poisson_on_cube.f90:729-901 for
the mesh and numbering, :169-177 for the Q1 element matrix, :404-519 for the
upload loop), plus the seeded right-hand sides and coefficient fields of the
synthetic workloads.

Each generator yields one dict per subdomain whose keys are the argument
names of ``bddcml_upload_subdomain_data_c`` (reference
src/bddcml_interface_c.h:14-27).  Index arrays carry ``numbase``.

Random draws use a counter-based generator (splitmix64 of seed XOR index), so
the same values are produced by any language.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """splitmix64 finaliser applied elementwise to uint64 values."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, index):
    """u in [0,1): (splitmix64(seed ^ index) >> 11) * 2**-53."""
    idx = np.asarray(index, dtype=np.uint64)
    z = splitmix64(np.uint64(seed) ^ idx)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def lognormal(seed: int, index, mu: float = 0.0, sigma: float = 1.0):
    """exp(mu + sigma * z), z standard normal by Box-Muller from the pair of
    uniforms (2i, 2i+1) of the seed's stream."""
    i = np.asarray(index, dtype=np.uint64)
    u1 = 1.0 - uniform01(seed, np.uint64(2) * i)          # (0, 1]
    u2 = uniform01(seed, np.uint64(2) * i + np.uint64(1))
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return np.exp(mu + sigma * z)


def q1_laplace_reference_element() -> np.ndarray:
    """8x8 stiffness of -Laplace on the unit hexahedron with trilinear shape
    functions, node order of the reference harness (poisson_on_cube.f90:146-167):
    1/3 on the diagonal, 0 between nodes joined by an element edge, -1/12
    between nodes on a face diagonal or the body diagonal."""
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    k = np.zeros((8, 8))
    for a in range(8):
        for b in range(8):
            ndiff = int(np.sum(corners[a] != corners[b]))
            k[a, b] = 1.0 / 3.0 if ndiff == 0 else (0.0 if ndiff == 1 else -1.0 / 12.0)
    return k


def q1_elasticity_reference_element(E: float = 1.0, nu: float = 0.3) -> np.ndarray:
    """24x24 stiffness of isotropic linear elasticity on the unit hexahedron
    (trilinear, 2x2x2 Gauss), dof order node-major (u,v,w per node)."""
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    gp = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
    K = np.zeros((24, 24))
    for x in gp:
        for y in gp:
            for z in gp:
                dN = np.zeros((8, 3))
                for a, (cx, cy, cz) in enumerate(corners):
                    fx = x if cx else 1 - x
                    fy = y if cy else 1 - y
                    fz = z if cz else 1 - z
                    sx = 1 if cx else -1
                    sy = 1 if cy else -1
                    sz = 1 if cz else -1
                    dN[a] = [sx * fy * fz, fx * sy * fz, fx * fy * sz]
                B = np.zeros((6, 24))
                for a in range(8):
                    B[0, 3 * a] = dN[a, 0]
                    B[1, 3 * a + 1] = dN[a, 1]
                    B[2, 3 * a + 2] = dN[a, 2]
                    B[3, 3 * a] = dN[a, 1]
                    B[3, 3 * a + 1] = dN[a, 0]
                    B[4, 3 * a + 1] = dN[a, 2]
                    B[4, 3 * a + 2] = dN[a, 1]
                    B[5, 3 * a] = dN[a, 2]
                    B[5, 3 * a + 2] = dN[a, 0]
                K += B.T @ D @ B * 0.125
    return K


class CubeProblem:
    """Unit cube split into nsub_edge^3 cubic subdomains of nel_sub^3 Q1
    hexahedra (reference harness examples/poisson_on_cube.f90).

    kind: 'poisson' (1 dof per node) or 'elasticity' (3 dofs per node,
          all dofs on the boundary fixed).
    rhs:  'const'  -> h^3 per node (poisson_on_cube.f90:433-438),
          'random' -> uniform[-1,1] per global dof from splitmix64(seed ^ dof).
    coef: optional callable mapping global element index array -> per element
          coefficient multiplying the element matrix.
    dirichlet: 'zero' (homogeneous, as the reference example) or 'random'
          -> fixed values uniform[-1,1] per global dof from
          splitmix64((seed + 1000) ^ dof).
    """

    def __init__(self, nel_sub: int, nsub_edge: int, kind: str = "poisson",
                 rhs: str = "const", seed: int = 2, numbase: int = 1,
                 coef=None, E: float = 1.0, nu: float = 0.3, dirichlet: str = "zero"):
        self.nel_sub = nel_sub
        self.nsub_edge = nsub_edge
        self.kind = kind
        self.rhs_kind = rhs
        self.seed = seed
        self.numbase = numbase
        self.coef = coef
        self.dirichlet = dirichlet
        self.ndim = 3
        self.meshdim = 3
        self.ndf = 1 if kind == "poisson" else 3
        self.nel = nel_sub * nsub_edge
        self.h = 1.0 / self.nel
        self.nelem = self.nel ** 3
        self.nsub = nsub_edge ** 3
        self.nnod = (self.nel + 1) ** 3
        self.ndof = self.nnod * self.ndf
        if kind == "poisson":
            self.kref = q1_laplace_reference_element() * self.h
        else:
            self.kref = q1_elasticity_reference_element(E, nu) * self.h
        ne = self.kref.shape[0]
        self._iu, self._ju = np.triu_indices(ne)

    # -- geometry -----------------------------------------------------------
    def sub_indices(self, isub0: int):
        ns = self.nsub_edge
        return isub0 % ns, (isub0 // ns) % ns, isub0 // (ns * ns)

    def global_node(self, ig, jg, kg):
        n = self.nel + 1
        return ig + jg * n + kg * n * n

    def box(self, isub0: int):
        """Element box (x0, y0, z0, nx, ny, nz) of subdomain isub0."""
        nes = self.nel_sub
        sx, sy, sz = self.sub_indices(isub0)
        return sx * nes, sy * nes, sz * nes, nes, nes, nes

    def subdomain(self, isub0: int) -> dict:
        """Arguments of bddcml_upload_subdomain_data_c for subdomain isub0
        (0-based; the returned isub carries numbase)."""
        b = self.numbase
        x0, y0, z0, nx, ny, nz = self.box(isub0)
        mx, my, mz = nx + 1, ny + 1, nz + 1
        kk, jj, ii = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
        ig = (x0 + ii).ravel()
        jg = (y0 + jj).ravel()
        kg = (z0 + kk).ravel()
        isngn = self.global_node(ig, jg, kg)
        nnods = mx * my * mz
        ndf = self.ndf
        ndofs = nnods * ndf
        xyz = np.empty((nnods, 3))
        xyz[:, 0] = ig * self.h
        xyz[:, 1] = jg * self.h
        xyz[:, 2] = kg * self.h
        nb = self.nel
        onb = (ig == 0) | (ig == nb) | (jg == 0) | (jg == nb) | (kg == 0) | (kg == nb)
        ifix = np.repeat(onb.astype(np.int32), ndf)
        isvgvn = (np.repeat(isngn * ndf, ndf) + np.tile(np.arange(ndf), nnods))
        fixv = np.zeros(ndofs)
        if self.dirichlet == "random":
            fixv = np.where(ifix > 0, 2.0 * uniform01(self.seed + 1000, isvgvn) - 1.0, 0.0)
        # elements
        ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
        n1 = ei + ej * mx + ek * mx * my
        n2 = n1 + 1
        n3 = n2 + mx
        n4 = n3 - 1
        pl = mx * my
        inet = np.stack([n1, n2, n3, n4, n1 + pl, n2 + pl, n3 + pl, n4 + pl], axis=1)
        nelems = inet.shape[0]
        isegn = (x0 + ei) + (y0 + ej) * nb + (z0 + ek) * nb * nb
        # element dofs (node-major)
        edofs = (inet[:, :, None] * ndf + np.arange(ndf)[None, None, :]).reshape(nelems, -1)
        ii_, jj_ = self._iu, self._ju
        r = edofs[:, ii_]
        c = edofs[:, jj_]
        vals = np.broadcast_to(self.kref[ii_, jj_], r.shape).copy()
        if self.coef is not None:
            vals *= np.asarray(self.coef(isegn), dtype=float)[:, None]
        lo = np.minimum(r, c).ravel()
        hi = np.maximum(r, c).ravel()
        # rhs
        if self.rhs_kind == "const":
            rhs = np.full(ndofs, self.h ** 3)
        else:
            rhs = 2.0 * uniform01(self.seed, isvgvn) - 1.0
        rhs[ifix > 0] = 0.0
        return dict(
            nelem=self.nelem, nnod=self.nnod, ndof=self.ndof, ndim=3, meshdim=3,
            isub=isub0 + b, nelems=nelems, nnods=nnods, ndofs=ndofs,
            inet=(inet.ravel() + b).astype(np.int32),
            nnet=np.full(nelems, 8, dtype=np.int32),
            nndf=np.full(nnods, ndf, dtype=np.int32),
            isngn=(isngn + b).astype(np.int32),
            isvgvn=(isvgvn + b).astype(np.int32),
            isegn=(isegn + b).astype(np.int32),
            xyz=np.asfortranarray(xyz),
            ifix=ifix.astype(np.int32), fixv=fixv, rhs=rhs, is_rhs_complete=1,
            sol=np.zeros(ndofs), matrixtype=1,
            i_sparse=(lo + b).astype(np.int32), j_sparse=(hi + b).astype(np.int32),
            a_sparse=vals.ravel(), is_assembled=0,
            find_components=1, use_dual_mesh_graph=1, neighbouring=3,
        )

    def subdomains(self):
        for s in range(self.nsub):
            yield self.subdomain(s)

    # -- the whole mesh at once (global loading) -----------------------------
    def global_data(self) -> dict:
        """Arguments of bddcml_upload_global_data_c (numbase-based indices):
        the global mesh (elements numbered x fastest, nodes as
        global_node), Dirichlet data, right-hand side and a zero initial
        solution, equal to what the subdomains carry."""
        b = self.numbase
        ndf = self.ndf
        nb = self.nel
        n = nb + 1
        ek, ej, ei = np.meshgrid(np.arange(nb), np.arange(nb), np.arange(nb), indexing="ij")
        n1 = self.global_node(ei.ravel(), ej.ravel(), ek.ravel())
        n2 = n1 + 1
        n3 = n2 + n
        n4 = n3 - 1
        pl = n * n
        inet = np.stack([n1, n2, n3, n4, n1 + pl, n2 + pl, n3 + pl, n4 + pl], axis=1)
        kg, jg, ig = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        ig, jg, kg = ig.ravel(), jg.ravel(), kg.ravel()
        xyz = np.stack([ig * self.h, jg * self.h, kg * self.h], axis=1)
        onb = (ig == 0) | (ig == nb) | (jg == 0) | (jg == nb) | (kg == 0) | (kg == nb)
        ifix = np.repeat(onb.astype(np.int32), ndf)
        gdof = np.arange(self.ndof)
        fixv = np.zeros(self.ndof)
        if self.dirichlet == "random":
            fixv = np.where(ifix > 0, 2.0 * uniform01(self.seed + 1000, gdof) - 1.0, 0.0)
        if self.rhs_kind == "const":
            rhs = np.full(self.ndof, self.h ** 3)
        else:
            rhs = 2.0 * uniform01(self.seed, gdof) - 1.0
        rhs[ifix > 0] = 0.0
        return dict(nelem=self.nelem, nnod=self.nnod, ndof=self.ndof, ndim=3, meshdim=3,
                    inet=(inet.ravel() + b).astype(np.int32),
                    nnet=np.full(self.nelem, 8, dtype=np.int32),
                    nndf=np.full(self.nnod, ndf, dtype=np.int32),
                    xyz=np.asfortranarray(xyz), ifix=ifix, fixv=fixv, rhs=rhs,
                    sol=np.zeros(self.ndof))

    def write_elm(self, path: str):
        """The element matrices as the reference's ELM file: one gfortran
        unformatted sequential record per element (4-byte length markers
        around LELM and LELM doubles), the upper triangle column by column
        (src/module_sm.f90:170-214, 272-375)."""
        ne = self.kref.shape[0]
        jj, ii = np.triu_indices(ne)[::-1]  # column j, rows 0..j
        order = np.lexsort((ii, jj))
        ii, jj = ii[order], jj[order]
        col = self.kref[ii, jj]
        lelm = len(col)
        coef = np.ones(self.nelem)
        if self.coef is not None:
            coef = np.asarray(self.coef(np.arange(self.nelem)), dtype=float) * coef
        rec = np.zeros(self.nelem, dtype=[("head", "<i4"), ("lelm", "<i4"), ("a", "<f8", lelm),
                                          ("tail", "<i4")])
        rec["head"] = rec["tail"] = 4 + 8 * lelm
        rec["lelm"] = lelm
        rec["a"] = coef[:, None] * col[None, :]
        rec.tofile(path)

    # -- global assembly (tests only need this for small cubes) --------------
    def global_dense(self, eliminate: bool = True):
        """Global matrix (full, after Dirichlet elimination that keeps the
        original diagonal) and right-hand side, dense.  With eliminate=False:
        the assembled matrix and right-hand side as uploaded, and the fixed
        values."""
        n = self.ndof
        A = np.zeros((n, n))
        f = np.zeros(n)
        g = np.zeros(n)
        fixed = np.zeros(n, dtype=bool)
        b = self.numbase
        for d in self.subdomains():
            gi = d["isvgvn"] - b
            i = gi[d["i_sparse"] - b]
            j = gi[d["j_sparse"] - b]
            a = d["a_sparse"]
            np.add.at(A, (i, j), a)
            off = i != j
            np.add.at(A, (j[off], i[off]), a[off])
            f[gi] = d["rhs"]
            fixed[gi[d["ifix"] > 0]] = True
            g[gi] = d["fixv"]
        if not eliminate:
            return A, f, fixed, g
        # Dirichlet values move to the right-hand side; fixed rows keep the
        # diagonal, u_fixed = g
        dg = np.diag(A).copy()
        f -= A[:, fixed] @ g[fixed]
        A[fixed, :] = 0.0
        A[:, fixed] = 0.0
        A[fixed, fixed] = dg[fixed]
        f[fixed] = dg[fixed] * g[fixed]
        return A, f, fixed


class BoxProblem(CubeProblem):
    """Unit cube of n^3 Q1 hexahedra split into nsub boxes of uneven size by
    seeded random bisection (the box with the most elements is cut across its
    longest extent at a fraction uniform in [0.25, 0.75]), with `layers` stiff
    slabs: each is `thick` elements across, normal to a random axis at a
    random position, and multiplies the element stiffness by `jump`.  The
    default is 3D linear elasticity with every boundary dof fixed."""

    def __init__(self, n: int, nsub: int, seed: int = 3, kind: str = "elasticity",
                 layers: int = 27, jump: float = 1e4, thick: int = 0, rhs: str = "random",
                 numbase: int = 1, dirichlet: str = "zero", E: float = 1.0, nu: float = 0.3):
        self.layer_seed = seed
        self.nlayers = layers
        self.jump = jump
        self.thick = thick if thick > 0 else max(1, n // 48)
        super().__init__(n, 1, kind=kind, rhs=rhs, seed=seed, numbase=numbase,
                         coef=self._coef if layers > 0 else None, E=E, nu=nu,
                         dirichlet=dirichlet)
        self.nsub = nsub
        self.boxes = self._split(n, nsub, seed)
        self.nsub = len(self.boxes)

    @staticmethod
    def _split(n, nsub, seed):
        boxes = [(0, 0, 0, n, n, n)]
        k = 0
        while len(boxes) < nsub:
            i = max(range(len(boxes)), key=lambda j: boxes[j][3] * boxes[j][4] * boxes[j][5])
            x0, y0, z0, nx, ny, nz = boxes[i]
            dims = [nx, ny, nz]
            ax = int(np.argmax(dims))
            L = dims[ax]
            if L < 2:
                break
            u = float(uniform01(seed * 7919 + 17, k))
            k += 1
            cut = min(L - 1, max(1, int(round((0.25 + 0.5 * u) * L))))
            lo = list(boxes[i])
            hi = list(boxes[i])
            lo[3 + ax] = cut
            hi[ax] += cut
            hi[3 + ax] = L - cut
            boxes[i] = tuple(lo)
            boxes.append(tuple(hi))
        return boxes

    def _coef(self, e):
        n = self.nel
        e = np.asarray(e)
        idx = (e % n, (e // n) % n, e // (n * n))
        c = np.ones(e.shape)
        for l in range(self.nlayers):
            ax = int(3 * uniform01(self.layer_seed * 104729 + 3, 2 * l))
            pos = int(n * uniform01(self.layer_seed * 104729 + 3, 2 * l + 1))
            inl = (idx[ax] >= pos) & (idx[ax] < pos + self.thick)
            c[inl] = self.jump
        return c

    def box(self, isub0: int):
        return self.boxes[isub0]
