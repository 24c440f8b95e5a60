"""H built straight into the residual blocks (PCI_BUILD_BLOCKS, no CSR)
against the CSR build of the same space: the same statistics, the same
lower-triangle CSR read back bit for bit, the same SpMM as the block kernel
run on blocks made from the CSR, and Davidson solves that agree to the last
bit, iteration for iteration (same blocks, same Init4 block and diagonal,
same head-coupling elements).  The CSR build is itself pinned to the oracle
and the reference's fixtures (test_gpu_formh.py, test_golden.py)."""
import numpy as np
import pytest

from pci_amd import _lib

pytestmark = pytest.mark.gpu

SPACES = ["c1", "t3", "t3b", "t4", "t4b", "t4g", "t3sd"]
RBLK = "k_spmm_rblk<16 waves,2 chunks>"
RB_SHAPES = ["k_spmm_rblk<8 waves,4 chunks>", "k_spmm_rblk<16 waves,2 chunks>",
             "k_spmm_rbmf<8 waves,4 chunks>", "k_spmm_rbmf<16 waves,2 chunks>"]


def _variant_index(g, nvec, prec, name):
    n = g.spmm_set_variant(nvec, prec)
    for v in range(n):
        try:
            g.spmm_set_variant(nvec, prec, v)
        except _lib.PciError:
            continue
        if g.spmm_variant(nvec, prec) == name:
            return v
    raise AssertionError(name)


@pytest.fixture(scope="module", params=SPACES)
def pair(request):
    from pci_amd import host
    from tests.helpers import space_dir
    name = request.param
    print(f"[case] blocks-only build {name}", flush=True)
    sp, a = host.load(space_dir(name))
    _, b = host.load(space_dir(name))
    a.build_h()
    b.build_h(_lib.PCI_BUILD_BLOCKS)
    yield name, sp, a, b
    a.close()
    b.close()


def test_statistics_and_rows_equal(pair):
    name, sp, a, b = pair
    assert (b.numh, b.numzero) == (a.numh, a.numzero)
    assert b.hmin == a.hmin
    ia, ib = a.build_info(), b.build_info()
    assert ib["candidates"] == ia["candidates"] and ib["nnz_full"] == ia["nnz_full"]
    rpa, cola, vala = a.get_csr()
    rpb, colb, valb = b.get_csr()
    np.testing.assert_array_equal(rpb, rpa)
    np.testing.assert_array_equal(colb, cola)
    np.testing.assert_array_equal(valb.view(np.int64), vala.view(np.int64))
    r1, n = sp.Nd // 3 + 1, min(500, sp.Nd // 3)
    ra, ca, va = a.get_csr_rows(r1, n)
    rb, cb, vb = b.get_csr_rows(r1, n)
    np.testing.assert_array_equal(rb, ra)
    np.testing.assert_array_equal(cb, ca)
    np.testing.assert_array_equal(vb.view(np.int64), va.view(np.int64))


def _exact(a, nvec, prec):
    """The CSR context ran the block kernel too (its copy of the blocks was
    eligible): results must then agree bit for bit, else to rounding."""
    return a.spmm_variant(nvec, prec).startswith("k_spmm_rblk")


@pytest.mark.parametrize("nvec", [1, 5, 8, 20])
def test_spmm_equals_blocks_from_csr(pair, nvec):
    name, sp, a, b = pair
    x = np.random.default_rng(nvec).standard_normal((sp.Nd, nvec))
    for prec, tol in ((64, 1e-12), (32, 2e-5)):
        a.spmm_set_variant(nvec, prec, _variant_index(a, nvec, prec, RBLK))
        b.spmm_set_variant(nvec, prec, _variant_index(b, nvec, prec, RBLK))
        ya, _, _ = a.spmm(x, prec)
        yb, _, _ = b.spmm(x, prec)
        assert b.spmm_variant(nvec, prec) == RBLK, b.spmm_variant(nvec, prec)
        if _exact(a, nvec, prec):
            np.testing.assert_array_equal(yb, ya)
        else:
            assert np.abs(yb - ya).max() <= tol * np.abs(ya).max()


@pytest.mark.parametrize("flags", [0, _lib.PCI_DVD_FP32, _lib.PCI_DVD_FP32 | _lib.PCI_DVD_COUPLED,
                                   _lib.PCI_DVD_COUPLED])
def test_davidson_bit_identical(pair, flags):
    name, sp, a, b = pair
    nlv = sp.Nlv
    for prec in (64, 32):
        a.spmm_set_variant(nlv, prec, _variant_index(a, nlv, prec, RBLK))
        b.spmm_set_variant(nlv, prec, _variant_index(b, nlv, prec, RBLK))
    crt4 = max(sp.Crt4, 1e-7) if flags & _lib.PCI_DVD_FP32 else sp.Crt4
    fl = flags | _lib.PCI_DVD_NOADMIX
    tka, _, va, sa = a.davidson(nlv, sp.Nd0, crt4, 300, fl)
    tkb, _, vb, sb = b.davidson(nlv, sp.Nd0, crt4, 300, fl)
    print(f"{name} flags {flags}: {sa['iterations']} / {sb['iterations']} iterations", flush=True)
    assert sb["converged"] == sa["converged"] == 1
    if _exact(a, nlv, 64) and _exact(a, nlv, 32):
        assert sb["iterations"] == sa["iterations"]
        np.testing.assert_array_equal(tkb, tka)
        np.testing.assert_array_equal(vb, va)
    else:
        assert np.abs(tkb - tka).max() * 219474.63 < 1e-3


def test_only_block_shapes_run(pair):
    name, sp, a, b = pair
    with pytest.raises(_lib.PciError):
        b.spmm_set_variant(4, 64, 0)  # the CSR kernel k_spmm
    with pytest.raises(_lib.PciError):
        b.spmm(np.ones((sp.Nd, 30)), 64)  # wider than the block kernels


@pytest.mark.parametrize("nvec", [1, 3, 4, 7, 8, 11, 16, 17, 20, 24])
def test_block_shapes_agree(pair, nvec):
    """Every residual-block shape (vector ALU and matrix-core, both launch
    shapes) on the blocks-only matrix against the oracle's Mxmpy over the CSR
    read back from the blocks: FP64 to 1e-12 of the largest output, FP32 to
    2e-5."""
    from oracle.oracle import mxmpy
    name, sp, a, b = pair
    if name not in ("t4g", "t3b"):
        pytest.skip("two spaces suffice for the width sweep")
    x = np.random.default_rng(100 + nvec).standard_normal((sp.Nd, nvec))
    rp, col, val = b.get_csr()
    yo = np.stack([mxmpy(rp, col, val, b.hmin, x[:, j]) for j in range(nvec)], axis=1)
    for prec, tol in ((64, 1e-12), (32, 2e-5)):
        ys = {}
        for shape in RB_SHAPES:
            b.spmm_set_variant(nvec, prec, _variant_index(b, nvec, prec, shape))
            ys[shape], _, _ = b.spmm(x, prec)
        scale = np.abs(yo).max()
        for shape, y in ys.items():
            err = np.abs(y - yo).max() / scale
            assert err <= tol, (prec, shape, err)
