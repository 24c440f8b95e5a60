"""The fast mode's dense FP32 eigensolver (cooperative one-launch Householder
tridiagonalisation + rocSOLVER sstedc + blocked WY back-transformation)
against numpy's FP64 eigh, the same with rocSOLVER's sormtr as the
back-transformation, and the solve through rocSOLVER's ssyevd (the
fallback)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from pci_amd._lib import GpuContext
    return GpuContext(0)


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) * 0.05
    A = (M + M.T) / 2 + np.diag(np.linspace(-1.0, 2.0, n))
    return A.astype(np.float32)


def _check(A, w, Z, tol):
    A64 = A.astype(np.float64)
    wr = np.linalg.eigvalsh(A64)
    scale = np.abs(wr).max()
    assert np.abs(w.astype(np.float64) - wr).max() < tol * scale
    Z64 = Z.astype(np.float64)
    res = A64 @ Z64 - Z64 * w.astype(np.float64)
    assert np.abs(res).max() < 10 * tol * scale
    assert np.abs(Z64.T @ Z64 - np.eye(A.shape[0])).max() < 10 * tol


@pytest.mark.parametrize("n", [3, 4, 17, 255, 256, 257, 1000, 2995, 3072])
def test_cooperative_tridiagonalisation(ctx, n):
    print(f"[case] eigh32 n={n}", flush=True)
    A = _matrix(n, n)
    w, Z, path = ctx.eigh32(A)
    assert path == 1
    _check(A, w, Z, 2e-5)


@pytest.mark.parametrize("n", [257, 3073])
def test_rocsolver_path(ctx, n, monkeypatch):
    print(f"[case] eigh32 rocsolver n={n}", flush=True)
    if n <= 3072:
        monkeypatch.setenv("PCI_EIGH_ROCSOLVER", "1")
    A = _matrix(n, 7)
    w, Z, path = ctx.eigh32(A)
    assert path == 0
    _check(A, w, Z, 2e-5)


@pytest.mark.parametrize("n", [129, 1000])
def test_sormtr_back_transformation(ctx, n, monkeypatch):
    """rocSOLVER's sormtr in place of the blocked WY application of Q."""
    print(f"[case] eigh32 sormtr n={n}", flush=True)
    monkeypatch.setenv("PCI_EIGH_SORMTR", "1")
    A = _matrix(n, 11)
    w, Z, path = ctx.eigh32(A)
    assert path == 1
    _check(A, w, Z, 2e-5)


@pytest.mark.parametrize("n", [1000, 2995])
def test_repeat_is_bit_identical(ctx, n):
    """Two cooperative solves of one matrix give the same bits: the grid
    barrier's flags and the exchanged columns use relaxed agent-scope
    atomics and rely on the hardware's store completion before the flag
    store; a stale read would show up as a run-to-run difference."""
    print(f"[case] eigh32 repeat n={n}", flush=True)
    A = _matrix(n, 5)
    w1, Z1, p1 = ctx.eigh32(A)
    w2, Z2, p2 = ctx.eigh32(A)
    assert p1 == p2 == 1
    np.testing.assert_array_equal(w1, w2)
    np.testing.assert_array_equal(Z1, Z2)


def test_barrier_timeout_falls_back_to_ssyevd(ctx, monkeypatch):
    """A spin bound of 1 (PCI_TRD_SPIN, forced from the host: every wait of
    the kernel stays bounded and the kernel ends normally) makes the first
    barrier give up; the solve then comes from ssyevd on the untouched
    input (path 2) and is still correct."""
    print("[case] eigh32 barrier timeout", flush=True)
    monkeypatch.setenv("PCI_TRD_SPIN", "1")
    A = _matrix(600, 9)
    w, Z, path = ctx.eigh32(A)
    assert path == 2
    _check(A, w, Z, 2e-5)
