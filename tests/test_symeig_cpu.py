"""Host symmetric eigensolver of the Rayleigh-Ritz steps of the adaptive
constraints' LOBPCG (Householder tridiagonalization + implicit QR with
Wilkinson shifts), through bddcml_amd_selftest_symeig: eigenvalues against
numpy.linalg.eigvalsh, eigenvector residuals A v - w v and orthonormality,
on random, clustered, rank-deficient (Gram matrices of dependent blocks),
diagonal, tridiagonal and zero matrices, including the order 30 of a
Rayleigh-Ritz step with ten eigenvectors."""
import numpy as np
import pytest

import bddcml_amd


def _cases():
    rng = np.random.default_rng(7)
    out = {}
    for n in (1, 2, 3, 7, 30, 50, 160):
        m = rng.standard_normal((n, n))
        out[f"random{n}"] = m + m.T
    q = np.linalg.qr(rng.standard_normal((40, 40)))[0]
    out["clustered"] = q @ np.diag(np.r_[np.ones(20), 1 + 1e-9 * np.arange(10), 5 * np.ones(10)]) @ q.T
    b = rng.standard_normal((90, 84))
    out["rank_deficient"] = b @ b.T  # 6-dimensional null space
    out["diagonal"] = np.diag(rng.standard_normal(30))
    t = np.diag(rng.standard_normal(25)) + np.diag(rng.standard_normal(24), 1)
    out["tridiagonal"] = t + t.T - np.diag(np.diag(t))
    out["zero"] = np.zeros((12, 12))
    out["identity"] = np.eye(9)
    return out


CASES = _cases()


@pytest.mark.parametrize("name", list(CASES))
def test_symeig_matches_numpy(name):
    a = CASES[name]
    n = a.shape[0]
    w, v = bddcml_amd.selftest_symeig(a)
    scale = max(np.abs(a).max(), 1e-300)
    ref = np.linalg.eigvalsh(a)
    assert np.abs(np.sort(w) - ref).max() <= 1e-12 * n * scale + 1e-300
    assert np.abs(a @ v - v * w).max() <= 1e-12 * n * scale + 1e-300
    assert np.abs(v.T @ v - np.eye(n)).max() <= 1e-12 * n
