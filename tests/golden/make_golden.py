"""Writes the golden fixtures of this directory from the CPU oracle.

The oracle is pinned against dense algebra (tests/test_oracle_dense.py) and
an independent restatement of the glob detection (tests/test_oracle_globs.py);
these fixtures freeze its outputs on the cases the survey of the reference
names, so that the HIP path and later oracle changes are checked against
fixed data:

  rng.npz          uniform[0,1) streams splitmix64(seed ^ index) of the
                   synthetic inputs (seeds 2, 5, 6; indices 0..63 and 2^40..)
  cube2x2.npz      2x2x2 subdomains of 2^3 Q1 Poisson elements, random rhs
                   (seed 7): S_i, Phi_i, Ac_i, interface gdofs and coarse ids
                   of every subdomain, S x and M r for seeded x, r, the
                   interface solution, iterations and condition estimate
  elast2x2.npz     the same mesh with 3D elasticity (all boundary dofs fixed)
  poisson32.npz    the reference's poisson_on_cube case: 32^3 elements,
                   4^3 subdomains, rhs h^3, PCG to 1e-8: coarse counts,
                   iterations, condition estimate, interface solution

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bddcml_amd.cube import CubeProblem, uniform01  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def sub_data(o, p):
    out = {}
    for s in range(p.nsub):
        isub = s + 1
        cptr, ind = o.sub_constraints(isub)
        out[f"S_{isub}"] = o.sub_schur(isub)
        out[f"phi_{isub}"] = o.sub_phi(isub)
        out[f"ac_{isub}"] = o.sub_coarse(isub)
        out[f"gdofs_{isub}"] = o.sub_interface_gdofs(isub)
        out[f"indrowc_{isub}"] = np.asarray(ind)
    return out


def small_case(p, name):
    o, (it, reason, cond) = orc.solve_cube(p, tol=1e-12, maxit=300)
    assert reason == 0
    nig = len(o.interface_dofs())
    x = 2.0 * uniform01(101, np.arange(nig, dtype=np.uint64)) - 1.0
    r = 2.0 * uniform01(102, np.arange(nig, dtype=np.uint64)) - 1.0
    data = sub_data(o, p)
    data.update(interface_dofs=o.interface_dofs(), x=x, Sx=o.apply_schur(x), r=r,
                Mr=o.apply_pc(r), g=o.reduced_rhs(), u=o.interface_solution(),
                iterations=it, cond=cond, counts=np.array(list(o.counts().values())))
    o.close()
    np.savez_compressed(os.path.join(HERE, name), **data)
    print(f"{name}: {p.ndof} dofs, {it} iterations, cond {cond:.6f}")


def main():
    np.savez_compressed(os.path.join(HERE, "rng.npz"), **{
        f"seed{s}_{k}": uniform01(s, np.arange(64, dtype=np.uint64) + np.uint64(off))
        for s in (2, 5, 6) for k, off in (("lo", 0), ("hi", 1 << 40))})
    small_case(CubeProblem(2, 2, rhs="random", seed=7), "cube2x2.npz")
    small_case(CubeProblem(2, 2, kind="elasticity", rhs="random", seed=7), "elast2x2.npz")
    p = CubeProblem(8, 4, rhs="const")
    o, (it, reason, cond) = orc.solve_cube(p, tol=1e-8, maxit=500)
    assert reason == 0
    np.savez_compressed(os.path.join(HERE, "poisson32.npz"),
                        u=o.interface_solution(), interface_dofs=o.interface_dofs(),
                        iterations=it, cond=cond, counts=np.array(list(o.counts().values())))
    print(f"poisson32.npz: {p.ndof} dofs, {it} iterations, cond {cond:.6f}, counts {o.counts()}")
    o.close()


if __name__ == "__main__":
    main()
