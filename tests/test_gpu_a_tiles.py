"""Unit checks of the batched 64x64 FP64 tile kernels (MFMA GEMM-sum with
its post operations, fused tile Cholesky + inverse) against numpy."""
import sys

import numpy as np
import pytest

import bddcml_amd

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(12)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("post", [0, 1, 2, 3])
def test_tile_gemm(kind, post):
    _say(f"case tile_gemm kind={kind} post={post}")
    npairs = 3
    A = rng.standard_normal((npairs, 64, 64))
    B = rng.standard_normal((npairs, 64, 64))
    C = rng.standard_normal((64, 64))
    X = rng.standard_normal((64, 64))
    alpha, beta = 0.75, -1.25
    out, _ = bddcml_amd.selftest_tile(post * 10 + kind, A, B, C, X, alpha, beta)
    if kind == 0:
        S = sum(A[p] @ B[p].T for p in range(npairs))
    elif kind == 1:
        S = sum(A[p] @ B[p] for p in range(npairs))
    else:
        S = sum(A[p].T @ B[p] for p in range(npairs))
    E = alpha * C + beta * S
    ref = [E, E @ X.T, X @ E, X.T @ E][post]
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


def test_tile_potrf_with_inverse():
    _say("case tile_potrf")
    M = rng.standard_normal((64, 64))
    C = M @ M.T + 64 * np.eye(64)
    A = 0.1 * rng.standard_normal((2, 64, 64))
    E = C - sum(A[p] @ A[p].T for p in range(2))
    L, Linv = bddcml_amd.selftest_tile(3, A, A, C)
    Lref = np.linalg.cholesky(E)
    assert np.abs(L - Lref).max() <= 1e-12 * np.abs(Lref).max()
    assert np.abs(np.triu(L, 1)).max() == 0.0
    assert np.abs(Linv @ Lref - np.eye(64)).max() <= 1e-12


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("post", [0, 1, 2, 3])
def test_tile_gemm_fp32(kind, post):
    """The FP32 instantiation (f32 16x16x4 MFMA, its own accumulator layout)
    against the FP64 numpy product, to FP32 accuracy."""
    _say(f"case tile_gemm fp32 kind={kind} post={post}")
    npairs = 3
    A = rng.standard_normal((npairs, 64, 64))
    B = rng.standard_normal((npairs, 64, 64))
    C = rng.standard_normal((64, 64))
    X = rng.standard_normal((64, 64))
    alpha, beta = 0.75, -1.25
    out, _ = bddcml_amd.selftest_tile(100 + post * 10 + kind, A, B, C, X, alpha, beta)
    S = sum((A[p] if kind < 2 else A[p].T) @ (B[p].T if kind == 0 else B[p]) for p in range(npairs))
    E = alpha * C + beta * S
    ref = [E, E @ X.T, X @ E, X.T @ E][post]
    assert np.abs(out - ref).max() <= 2e-5 * np.abs(ref).max()


def test_tile_potrf_with_inverse_fp32():
    _say("case tile_potrf fp32")
    M = rng.standard_normal((64, 64))
    C = M @ M.T + 64 * np.eye(64)
    A = 0.1 * rng.standard_normal((2, 64, 64))
    E = C - sum(A[p] @ A[p].T for p in range(2))
    L, Linv = bddcml_amd.selftest_tile(103, A, A, C)
    Lref = np.linalg.cholesky(E)
    assert np.abs(L - Lref).max() <= 1e-5 * np.abs(Lref).max()
    assert np.abs(np.triu(L, 1)).max() == 0.0
    assert np.abs(Linv @ Lref - np.eye(64)).max() <= 1e-5
