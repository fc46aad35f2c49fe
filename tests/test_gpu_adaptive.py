"""Adaptive constraints on the device (batched LOBPCG over the face pairs,
bddcml_setup_preconditioner_c with use_adaptive_constraints_int = 1) against
the oracle's dense restatement (oracle/adaptivity.py): the same pairs; the
eigenvalues at or above half the threshold within 1e-6 relative, the
clustered ones far below it within 1e-3 of the pair's largest, none above its
exact value (Ritz values are lower bounds; Jacobi-preconditioned LOBPCG, 100
iterations); the same number of adaptive constraints per
face and coarse dofs, the preconditioner M r within 1e-4 (the face spans agree
to the eigenvector accuracy), PCG iterations within +-1 and the solution
within 1e-8.
Problems: elasticity with stiff channels through the faces (weights of
type 1, threshold 10), and layered Poisson boxes where many faces take the
full ten eigenvectors (threshold 2)."""
import sys

import numpy as np
import pytest

import bddcml_amd
from bddcml_amd.cube import BoxProblem, CubeProblem
from oracle.adaptivity import adaptive_constraints

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _channels(n, jump):
    def coef(e):
        e = np.asarray(e)
        y, z = (e // n) % n, e // (n * n)
        c = np.ones(e.shape)
        c[(y % 3 == 1) & (z % 3 == 1)] = jump
        return c
    return coef


CASES = {
    "channels_elasticity_3x3": (lambda: CubeProblem(3, 3, kind="elasticity", rhs="random", seed=2,
                                                    coef=_channels(9, 1e4)), 1, 10.0),
    "channels_elasticity_2x4": (lambda: CubeProblem(4, 2, kind="elasticity", rhs="random", seed=2,
                                                    coef=_channels(8, 1e4)), 1, 10.0),
    "layered_poisson_boxes": (lambda: BoxProblem(6, 5, seed=3, kind="poisson", layers=3, jump=1e3,
                                                 thick=1), 0, 2.0),
}


@pytest.mark.parametrize("case", list(CASES))
def test_adaptive_constraints_match_oracle(oracle_lib, case):
    make, wt, tau = CASES[case]
    p = make()
    _say(f"case adaptive {case}: {p.ndof} dofs, {p.nsub} subdomains, threshold {tau}")
    o = oracle_lib.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in p.subdomains():
        o.upload(d)
    o.setup(1, 1, wt)
    rep = adaptive_constraints(o, p, threshold=tau, weights_type=wt)
    it_o, r_o, c_o = o.solve(0, 1e-10, 500, 50)
    b = bddcml_amd.Bddcml()
    try:
        b.init(2, [p.nsub, 1], numbase=p.numbase)
        for k, v in (("adaptive_threshold", tau), ("adaptive_tol", 1e-11), ("adaptive_maxit", 100)):
            assert b.set_option(k, v) == 0
        for d in p.subdomains():
            b.upload_subdomain_data(d)
        b.setup_preconditioner(1, 0, 1, 1, 1, 1, 0, wt)
        pairs = b.adaptive_pairs()
        ref = {q["face"]: q for q in rep["pairs"]}
        assert sorted(q["face"] for q in pairs) == sorted(ref)
        for q in pairs:
            r = ref[q["face"]]
            assert (q["isub"], q["jsub"]) == (r["isub"], r["jsub"])
            assert q["neig"] == len(r["eigenvalues"])
            lam, lam_ref = q["eigenvalues"], r["eigenvalues"]
            # Ritz values never exceed the eigenvalues they approximate; those
            # that can decide a constraint (>= threshold / 2) are tight, the
            # clustered ones far below the threshold only roughly converged
            assert np.all(lam <= lam_ref * (1 + 1e-6) + 1e-12), (q["face"], lam, lam_ref)
            big = lam_ref >= 0.5 * tau
            assert np.all(np.abs(lam - lam_ref)[big] <= 1e-6 * lam_ref[big]), (q["face"], lam, lam_ref)
            assert np.abs(lam - lam_ref).max() <= 5e-2 * max(1.0, lam_ref.max()), (q["face"], lam, lam_ref)
            assert q["nadaptive"] == r["nadaptive"]
        assert b.counts()["ndofc"] == rep["ndofc"] == o.counts()["ndofc"]
        it, reason, cond = b.solve(0, 1e-10, 500, 50)
        assert reason == 0 == r_o
        assert abs(it - it_o) <= 1
        assert abs(cond - c_o) <= 1e-3 * c_o
        u_ref = o.interface_solution()
        assert np.abs(b.interface_solution() - u_ref).max() <= 1e-8 * np.abs(u_ref).max()
        x = np.random.default_rng(3).standard_normal(len(u_ref))
        z, z_ref = b.apply_pc(x), o.apply_pc(x)
        assert np.abs(z - z_ref).max() <= 1e-4 * np.abs(z_ref).max()
    finally:
        b.finalize()
        o.close()


def test_adaptive_reference_defaults_converge():
    """The reference's compiled LOBPCG settings (tolerance 1e-9 relative,
    15 iterations, threshold 10 here): the solve converges and the
    adaptive constraints lower the condition estimate of the arithmetic
    constraints."""
    p = CASES["channels_elasticity_3x3"][0]()
    _say("case adaptive defaults: channels elasticity 3x3")
    conds = {}
    for adaptive in (0, 1):
        b = bddcml_amd.Bddcml()
        try:
            b.init(2, [p.nsub, 1], numbase=p.numbase)
            b.set_option("adaptive_threshold", 10.0)
            for d in p.subdomains():
                b.upload_subdomain_data(d)
            b.setup_preconditioner(1, 0, 1, 1, 1, adaptive, 0, 1)
            it, reason, cond = b.solve(0, 1e-10, 500, 50)
            assert reason == 0
            conds[adaptive] = cond
        finally:
            b.finalize()
    assert conds[1] < conds[0] / 5
