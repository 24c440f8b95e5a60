"""Learned Davidson correction on the device: MFMA BF16 forward and
backward against the CPU restatement (oracle/mlp.py), training, and the
learned-vs-diagonal ablation inside the Davidson solver."""
import numpy as np
import pytest

from oracle import mlp as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from pci_amd._lib import GpuContext
    return GpuContext(0)


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(n) * 2.0
    r = rng.standard_normal(n)
    X = M.features(u, r, 1.0, rng.integers(0, 4, n))
    b = rng.integers(0, M.NB, n).astype(np.int32)
    return X, b, r.astype(np.float32)


@pytest.mark.parametrize("n", [1, 31, 1000, 4133])
def test_forward_matches_oracle(ctx, n):
    print(f"[case] forward n={n}", flush=True)
    w = M.init_weights(5, w3_scale=0.3)
    w[M.OFF_B1:M.OFF_W2] = np.random.default_rng(1).standard_normal(M.NH) * 0.1
    w[M.OFF_B3] = 0.25
    ctx.mlp_set_weights(w)
    X, b, r = _samples(n, 11 + n)
    t = ctx.mlp_forward(X, b, r)
    to = M.forward(w, X, b, r)
    # FP32 (MFMA) versus FP64 accumulation of the same BF16 products
    scale = np.abs(to).max()
    assert np.abs(t - to).max() <= 2e-4 * scale, np.abs(t - to).max() / scale
    assert np.median(np.abs(t - to) / (np.abs(to) + 1e-30)) < 1e-5


def test_gradient_matches_oracle(ctx):
    print("[case] gradient n=3000", flush=True)
    w = M.init_weights(7, w3_scale=0.2)
    w[M.OFF_B1:M.OFF_W2] = 0.05
    ctx.mlp_set_weights(w)
    n = 3000
    X, b, r = _samples(n, 3)
    T = np.random.default_rng(4).standard_normal(n).astype(np.float32)
    g, loss = ctx.mlp_grad(X, b, r, T)
    lo, go = M.loss_grad(w, X, b, r, T)
    assert abs(loss - lo) <= 1e-4 * lo
    blocks = {"W1": (M.OFF_W1, M.OFF_B1), "b1": (M.OFF_B1, M.OFF_W2), "W2": (M.OFF_W2, M.OFF_B2),
              "b2": (M.OFF_B2, M.OFF_W3), "w3": (M.OFF_W3, M.OFF_B3), "b3": (M.OFF_B3, M.OFF_E),
              "E": (M.OFF_E, M.NPARAM)}
    for name, (a, z) in blocks.items():
        err = np.linalg.norm(g[a:z] - go[a:z]) / np.linalg.norm(go[a:z])
        print(name, err)
        assert err < 2e-3, (name, err)


def test_training_fits_teacher(ctx):
    print("[case] training", flush=True)
    n = 8192
    X, b, r = _samples(n, 21)
    teacher = M.init_weights(9, w3_scale=0.2)
    T = M.forward(teacher, X, b, r).astype(np.float32)
    ctx.mlp_init(3)
    w0 = ctx.mlp_get_weights()
    l0 = M.loss_grad(w0, X, b, r, T)[0]
    hist = ctx.mlp_train(X, b, r, T, epochs=60, batch=1024, lr=2e-3)
    w1 = ctx.mlp_get_weights()
    l1 = M.loss_grad(w1, X, b, r, T)[0]
    print("loss", l0, "->", l1, hist[[0, -1]])
    assert abs(hist[0] - l0) <= 0.5 * l0 + 1e-12
    assert l1 < 0.1 * l0
    # the trained network still agrees with the oracle forward pass
    t = ctx.mlp_forward(X[:512], b[:512], r[:512])
    to = M.forward(w1, X[:512], b[:512], r[:512])
    assert np.abs(t - to).max() <= 2e-4 * np.abs(to).max()


def test_learned_vs_diagonal_ablation():
    """Same space, same H: the diagonal correction (PCI_DVD_LEARNED off) and
    the learned one trained on pairs collected on a sibling space.  Both must
    reach the same eigenvalues; iteration counts are reported."""
    from pci_amd import host
    from pci_amd._lib import PCI_DVD_COUPLED, PCI_DVD_FP32, PCI_DVD_LEARNED
    from tests.helpers import space_dir
    HARTREE_CM = 2 * 109737.315709
    print("[case] ablation train=t3b solve=t3", flush=True)
    sp_t, ctx_t = host.load(space_dir("t3b"))
    ctx_t.build_h()
    tk_t, _, v_t, _ = ctx_t.davidson(sp_t.Nlv, sp_t.Nd0, 1e-10, 400, 0)
    X, b, r, T = ctx_t.mlp_collect(sp_t.Nlv, sp_t.Nd0, sp_t.Crt4, sp_t.N_it, v_t, stride=1, cap=1 << 20)
    assert len(r) > 0
    ctx_t.mlp_init(1)
    hist = ctx_t.mlp_train(X, b, r, T, epochs=30, batch=2048, lr=1e-3)
    assert hist[-1] < hist[0]
    w = ctx_t.mlp_get_weights()
    del ctx_t
    sp, ctx2 = host.load(space_dir("t3"))
    ctx2.build_h()
    ctx2.mlp_set_weights(w)
    ref, _, _, _ = ctx2.davidson(sp.Nlv, sp.Nd0, 1e-10, 400, 0)
    for flags in (0, PCI_DVD_FP32):
        tk0, _, _, s0 = ctx2.davidson(sp.Nlv, sp.Nd0, sp.Crt4, sp.N_it, flags)
        tkc, _, _, sc = ctx2.davidson(sp.Nlv, sp.Nd0, sp.Crt4, sp.N_it, flags | PCI_DVD_COUPLED)
        tk1, _, _, s1 = ctx2.davidson(sp.Nlv, sp.Nd0, sp.Crt4, sp.N_it, flags | PCI_DVD_LEARNED)
        print(f"flags={flags} diagonal: {s0['iterations']} it; coupled: {sc['iterations']} it; "
              f"learned: {s1['iterations']} it (used {s1['learned_used']}, rejected {s1['learned_rejected']})")
        for tk, st in ((tk0, s0), (tkc, sc), (tk1, s1)):
            assert st["converged"] in (1, 2)
            assert np.abs(tk - ref).max() * HARTREE_CM < 0.01
        assert s1["learned_used"] + s1["learned_rejected"] > 0
        if s1["learned_used"] == 0:
            # every learned step was undone with all the state it started
            # from: the rest is the diagonal solve, one iteration later per
            # undone step
            assert s1["iterations"] == s0["iterations"] + s1["learned_rejected"]
            np.testing.assert_array_equal(tk1, tk0)
