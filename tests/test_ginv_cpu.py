"""Host generalized inverse of the adaptive constraints' pair coarse
problems (truncated pivoted Cholesky), through bddcml_amd_selftest_ginv:
k g k = k and g symmetric on singular positive semidefinite matrices (the
coarse matrix of a floating pair has rigid modes in its null space), the
rank equal to numpy's, and g = k^-1 on definite ones, including a badly
scaled one (coefficient jumps)."""
import numpy as np
import pytest

import bddcml_amd


def _cases():
    rng = np.random.default_rng(11)
    out = {}
    for n, r in ((1, 1), (2, 1), (5, 5), (40, 34), (160, 154), (200, 199)):
        b = rng.standard_normal((n, r))
        out[f"psd{n}_rank{r}"] = (b @ b.T, r)
    b = rng.standard_normal((60, 60))
    d = np.diag(10.0 ** rng.uniform(-6, 0, 60))
    out["definite_jumps"] = (d @ (b @ b.T + 60 * np.eye(60)) @ d, 60)
    out["zero"] = (np.zeros((8, 8)), 0)
    out["identity"] = (np.eye(7), 7)
    return out


CASES = _cases()


@pytest.mark.parametrize("name", list(CASES))
def test_generalized_inverse(name):
    k, rank = CASES[name]
    g, r = bddcml_amd.selftest_ginv(k)
    assert r == rank
    scale = max(np.abs(k).max(), 1e-300)
    assert np.abs(g - g.T).max() <= 1e-14 * max(np.abs(g).max(), 1e-300)
    assert np.abs(k @ g @ k - k).max() <= 1e-9 * scale
    if rank == k.shape[0]:
        ki = np.linalg.inv(k)
        assert np.abs(g - ki).max() <= 1e-8 * np.abs(ki).max()
