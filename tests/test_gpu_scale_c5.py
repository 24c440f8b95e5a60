"""The C5 workload at full size (4 valence electrons over s..g shells,
Nd = 2,332,805, ~1.24e10 elements in both triangles), held only as residual
blocks (PCI_BUILD_BLOCKS: its FP64 CSR would not fit beside them): FormH
rows spread over the whole determinant list, read back from the blocks,
equal the CPU oracle's fixture (tests/golden/make_scale_golden.py) bit for
bit, and the FP32 mode's 20 roots agree with the plain FP64 method and both
with a tightly converged FP64 solve.  A module of its own, so that the C4
context of test_gpu_scale.py is released before this one is built."""
import numpy as np
import pytest

from tests.helpers import space_dir
from tests.test_scale_fixtures import fixture_rows, load_scale

pytestmark = pytest.mark.gpu

HARTREE_CM = 2 * 109737.315709


@pytest.fixture(scope="module")
def c5():
    from pci_amd import _lib, host
    print("[case] scale space=c5 (residual blocks only)", flush=True)
    sp, ctx = host.load(space_dir("c5"))
    numh, numzero, hmin = ctx.build_h(_lib.PCI_BUILD_BLOCKS)
    yield sp, ctx, numh
    ctx.close()


def test_c5_rows_bit_exact(c5):
    sp, ctx, numh = c5
    g = load_scale("c5")
    assert sp.Nd == int(g["nd"]) and ctx.build_info()["nnz_full"] > 1.2e10
    for r, (col, bits) in fixture_rows(g).items():
        rp, c, v = ctx.get_csr_rows(r, 1)
        np.testing.assert_array_equal(c, col, err_msg=f"row {r}")
        np.testing.assert_array_equal(v.view(np.int64), bits, err_msg=f"row {r}")


def test_c5_fp32_roots_match_plain_fp64(c5):
    from pci_amd import _lib
    sp, ctx, _ = c5
    assert sp.Nlv == 20
    tk_ref, _, _, st_ref = ctx.davidson(sp.Nlv, sp.Nd0, 1e-10, 400, _lib.PCI_DVD_NOADMIX, vectors=False)
    assert st_ref["max_residual"] < 1e-8
    for mode, flags in (("plain", 0), ("fp32", _lib.PCI_DVD_FP32),
                        ("coupled_fp32", _lib.PCI_DVD_FP32 | _lib.PCI_DVD_COUPLED)):
        tk, _, _, st = ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, 400, flags | _lib.PCI_DVD_NOADMIX, vectors=False)
        assert st["converged"] == 1 and st["max_residual"] <= sp.Crt4, (mode, st["max_residual"])
        err = np.abs(tk - tk_ref) * HARTREE_CM
        print(f"C5 {mode}: {st['iterations']} it, max |dE| vs FP64 1e-10 {err.max():.2e} cm^-1", flush=True)
        assert err.max() < 0.01
