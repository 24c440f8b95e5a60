"""Batched Diag4 (pci_gpu_davidson_batch): independent CI blocks solved at
once on their own streams give exactly what each block's own solve gives."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BLOCKS = ["t3", "t4", "c1", "t3b"]


@pytest.fixture(scope="module")
def blocks():
    from pci_amd import host
    from tests.helpers import space_dir
    out = []
    for name in BLOCKS:
        print(f"[case] batch block {name}", flush=True)
        sp, ctx = host.load(space_dir(name))
        ctx.build_h()
        out.append((name, sp, ctx))
    return out


@pytest.mark.parametrize("flags,eigh", [(0, "coop"), (1, "coop"), (1, "rocsolver")])
def test_batch_equals_single_solves(blocks, flags, eigh, monkeypatch):
    """The head eigensolves of a batch run one block at a time on an idle
    GPU, so the FP32 mode takes the same (cooperative or rocSOLVER) path as a
    single solve and every block's result is bit-identical."""
    from pci_amd import _lib
    if eigh == "rocsolver":
        monkeypatch.setenv("PCI_EIGH_ROCSOLVER", "1")
    params = [dict(nlv=sp.Nlv, nd0=sp.Nd0, crt4=sp.Crt4, n_it=sp.N_it, flags=flags) for _, sp, _ in blocks]
    single = [ctx.davidson(q["nlv"], q["nd0"], q["crt4"], q["n_it"], flags) for (_, _, ctx), q in zip(blocks, params)]
    batch = _lib.davidson_batch([ctx for _, _, ctx in blocks], params)
    for (name, _, _), (tk1, _, v1, s1), (tk2, _, v2, s2) in zip(blocks, single, batch):
        print(f"{name}: {s1['iterations']} / {s2['iterations']} iterations", flush=True)
        assert s2["converged"] in (1, 2), name
        assert s1["iterations"] == s2["iterations"], name
        np.testing.assert_array_equal(tk1, tk2)
        np.testing.assert_array_equal(v1, v2)


def test_batch_rejects_a_repeated_context(blocks):
    from pci_amd import _lib
    _, sp, ctx = blocks[0]
    q = dict(nlv=sp.Nlv, nd0=sp.Nd0, crt4=sp.Crt4, n_it=sp.N_it)
    with pytest.raises(_lib.PciError, match="twice"):
        _lib.davidson_batch([ctx, ctx], [q, q])


def test_batch_reports_a_failing_block_and_finishes_the_others(blocks):
    """A block that stops before its head eigensolve (Nd0 > Nd) must not hold
    the other blocks at the eigensolve rendezvous: the call returns, names
    the block, and the other blocks' results are complete."""
    from pci_amd import _lib
    params = [dict(nlv=sp.Nlv, nd0=sp.Nd0, crt4=sp.Crt4, n_it=sp.N_it, flags=1) for _, sp, _ in blocks]
    params[1]["nd0"] = blocks[1][1].Nd + 1
    with pytest.raises(_lib.PciError, match="block 1"):
        _lib.davidson_batch([ctx for _, _, ctx in blocks], params)
    # the contexts stay usable after the failed batch
    _, sp, ctx = blocks[0]
    tk, _, _, st = ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, sp.N_it, 1)
    assert st["converged"] in (1, 2)
