"""The CPU restatement of the learned correction (oracle/mlp.py), pinned
without the device: BF16 rounding against torch's conversion, and the
analytic gradient against central finite differences of the exact (FP64,
unrounded) network."""
import numpy as np
import pytest

from oracle import mlp as M


def test_bf16_matches_torch_rounding():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = (rng.standard_normal(200000) * np.logspace(-35, 35, 200000)).astype(np.float32)
    # exact ties of the BF16 grid must round to even
    ties = (np.arange(1, 5000, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    a = np.concatenate([a, ties, -ties, [0.0, -0.0, np.inf, -np.inf]]).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(a).view(np.uint32), ref.view(np.uint32))


def test_layout_constants():
    assert M.NPARAM == 128 * 32 + 128 + 128 * 128 + 128 + 128 + 1 + 64 * 16
    from pci_amd import _lib
    assert (_lib.MLP_NF, _lib.MLP_IN, _lib.MLP_H, _lib.MLP_NB, _lib.MLP_NE) == (M.NF, M.NIN, M.NH, M.NB, M.NE)


@pytest.mark.parametrize("seed", [1, 2])
def test_gradient_finite_differences(seed):
    rng = np.random.default_rng(seed)
    w = M.init_weights(seed, w3_scale=0.1).astype(np.float64)
    w[M.OFF_B1:M.OFF_W2] = rng.standard_normal(M.NH) * 0.1
    n = 48
    X = rng.standard_normal((n, M.NF))
    b = rng.integers(0, M.NB, n)
    r = rng.standard_normal(n)
    T = rng.standard_normal(n)
    _, g = M.loss_grad(w, X, b, r, T, bf=False, inv_b=0.5 / n, inv_n=0.5 / n)
    gs = np.abs(g).max()
    # every parameter block, including the embedding rows actually used
    blocks = [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
              (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.OFF_E)]
    idx = [rng.integers(lo, hi, 12) for lo, hi in blocks]
    idx.append(M.OFF_E + b[:12] * M.NE + rng.integers(0, M.NE, 12))
    for i in np.concatenate(idx):
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (M.loss_grad(wp, X, b, r, T, bf=False, inv_n=0.5 / n)[0]
              - M.loss_grad(wm, X, b, r, T, bf=False, inv_n=0.5 / n)[0]) / (2 * h)
        assert abs(g[i] - fd) <= 1e-7 * gs, (i, g[i], fd)


def test_zero_network_is_clipped_diagonal():
    """With w3 = b3 = 0 the gain is -f1 = -1/(u + sign(u) 0.02): the Dvdsn
    diagonal correction r / (theta - D) away from the clip."""
    w = M.init_weights(0)
    w[M.OFF_W3:M.OFF_E] = 0.0
    u = np.array([0.5, -0.7, 3.0, -2.0, 40.0])
    r = np.array([1.0, 2.0, -1.0, 0.5, 3.0])
    X = M.features(u, r, 1.0, np.zeros(5))
    t = M.forward(w, X, np.zeros(5, int), r, bf=False)
    np.testing.assert_allclose(t, r / (-(u + np.sign(u) * 0.02)), rtol=1e-6)
