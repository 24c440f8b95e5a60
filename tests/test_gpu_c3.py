"""C3 at BASELINE's sizes: the twelve (J, parity) blocks of one 4-electron
atom (gen.c3_blocks: 36k to 376k determinants, 4.0e9 elements in all), each
built straight into residual blocks.  Per block: FormH rows spread over the
determinant list, read back from the blocks, against the CPU oracle bit for
bit; the batched solve (pci_gpu_davidson_batch, every block at once) against
each block's own solve bit for bit; and every block's roots within 0.01 cm^-1
of a tightly converged FP64 solve of the same block."""
import numpy as np
import pytest

from tests.helpers import spec_dir

pytestmark = pytest.mark.gpu

HARTREE_CM = 2 * 109737.315709


@pytest.fixture(scope="module")
def blocks():
    from pci_amd import _lib, gen, host
    out = []
    try:
        for spec in gen.c3_blocks():
            print(f"[case] C3 block {spec.name}", flush=True)
            d = spec_dir(spec)
            sp, c = host.load(d)
            out.append((spec, d, sp, c))
            c.build_h(_lib.PCI_BUILD_BLOCKS)
        yield out
    finally:
        for *_, c in out:
            c.close()


def test_sizes(blocks):
    nd = [sp.Nd for _, _, sp, _ in blocks]
    assert min(nd) > 20000 and max(nd) < 400000 and len(nd) == 12
    assert sum(c.build_info()["nnz_full"] for *_, c in blocks) > 3.5e9


def test_rows_against_oracle(blocks):
    from oracle.oracle import Oracle
    for spec, d, sp, c in blocks:
        o = Oracle.from_dir(d)
        assert o.Nd == sp.Nd
        for r in np.unique(np.linspace(1, sp.Nd, 5).round().astype(int)):
            rp, col, val, _, _ = o.formh(int(r), int(r))
            _, gc, gv = c.get_csr_rows(int(r), 1)
            np.testing.assert_array_equal(gc, col, err_msg=f"{spec.name} row {r}")
            np.testing.assert_array_equal(gv.view(np.int64), val.view(np.int64), err_msg=f"{spec.name} row {r}")


def test_batch_equals_single_and_roots(blocks):
    from pci_amd import _lib
    flags = _lib.PCI_DVD_FP32 | _lib.PCI_DVD_NOADMIX
    params = [dict(nlv=sp.Nlv, nd0=sp.Nd0, crt4=sp.Crt4, n_it=300, flags=flags) for _, _, sp, _ in blocks]
    single = [c.davidson(q["nlv"], q["nd0"], q["crt4"], q["n_it"], flags, vectors=False)
              for (*_, c), q in zip(blocks, params)]
    batch = _lib.davidson_batch([c for *_, c in blocks], params)
    for (spec, _, sp, c), s1, b1 in zip(blocks, single, batch):
        assert s1[3]["converged"] == 1 and b1[3]["converged"] == 1, spec.name
        assert b1[3]["iterations"] == s1[3]["iterations"], spec.name
        np.testing.assert_array_equal(b1[0], s1[0], err_msg=spec.name)
        tk_ref, _, _, st = c.davidson(sp.Nlv, sp.Nd0, 1e-10, 600, _lib.PCI_DVD_NOADMIX, vectors=False)
        assert st["max_residual"] < 1e-8, spec.name
        err = np.abs(s1[0] - tk_ref).max() * HARTREE_CM
        print(f"{spec.name}: Nd={sp.Nd} {s1[3]['iterations']} it, max |dE| {err:.2e} cm^-1", flush=True)
        assert err < 0.01, spec.name
