"""The sliced-copy SpMM kernels on a matrix wide enough for more than one
column panel of each kind (C2: Nd = 69,101 > 65,536), against the plain CSR
kernel k_spmm (which the shape test pins to the oracle's Mxmpy on the small
spaces).  Both sliced copies then also exercise their per-panel partial sums
and k_spmm_panel_sum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    from pci_amd import host
    from tests.helpers import space_dir
    print("[case] C2 sliced-copy SpMM", flush=True)
    sp, g = host.load(space_dir("c2"))
    g.build_h()
    yield sp, g
    g.close()


def _shapes(g, nvec):
    n = g.spmm_set_variant(nvec, 64)
    names = {}
    for v in range(n):
        g.spmm_set_variant(nvec, 64, v)
        names[g.spmm_variant(nvec, 64)] = v
    return names


def _check_sliced(g, x, nvec, precs=((64, 1e-12), (32, 2e-5))):
    names = _shapes(g, nvec)
    # LDS panels (3 shapes), the pipelined LDS panels (3), wide panels (2)
    sliced = [v for k, v in names.items() if k.startswith("k_spmm_sell")]
    assert len(sliced) == 8
    for prec, tol in precs:
        g.spmm_set_variant(nvec, prec, names["k_spmm"])
        ref, _, _ = g.spmm(x, prec)
        scale = np.abs(ref).max()
        for v in sliced:
            g.spmm_set_variant(nvec, prec, v)
            y, _, _ = g.spmm(x, prec)
            err = np.abs(y - ref).max() / scale
            print(f"{g.spmm_variant(nvec, prec)} fp{prec}: {err:.1e}", flush=True)
            assert err <= tol, (nvec, prec, g.spmm_variant(nvec, prec), err)


# widths in rising order: each wider width whose LDS panel is narrower than
# the copy's panels rebuilds the copy, which the narrower widths then share;
# 3, 7 and 11 take the FP32 tail reads of 3 columns and the FP64 scalar
# tail read of an odd width
@pytest.mark.parametrize("nvec", [1, 3, 7, 8, 10, 11, 16, 24])
def test_sliced_kernels_match_csr_kernel(c2, nvec):
    sp, g = c2
    x = np.random.default_rng(5).standard_normal((sp.Nd, nvec))
    _check_sliced(g, x, nvec)


def test_narrow_block_on_a_copy_built_for_a_wider_one(c2):
    """Davidson solves of different Nlv share one copy: a copy rebuilt for 24
    columns (narrower panels) then serves 8 columns."""
    sp, g = c2
    for nvec in (24, 8):
        x = np.random.default_rng(nvec).standard_normal((sp.Nd, nvec))
        _check_sliced(g, x, nvec)


def test_copy_that_does_not_fit_falls_back_to_csr_kernels():
    """A sliced copy that cannot be allocated (forced with PCI_SELL_OFF, the
    path a PCI_E_NOMEM takes) leaves the copy ineligible: the sliced shapes
    then run the CSR kernels and the results stay correct."""
    import os
    from pci_amd import host
    from tests.helpers import space_dir
    sp, g = host.load(space_dir("t3"))
    g.build_h()
    os.environ["PCI_SELL_OFF"] = "1"
    try:
        x = np.random.default_rng(3).standard_normal((sp.Nd, 6))
        names = _shapes(g, 6)
        for prec, tol in ((64, 1e-12), (32, 2e-5)):
            g.spmm_set_variant(6, prec, names["k_spmm"])
            ref, _, _ = g.spmm(x, prec)
            for k, v in names.items():
                if k.startswith("k_spmm_sell"):
                    g.spmm_set_variant(6, prec, v)
                    y, _, _ = g.spmm(x, prec)
                    assert np.abs(y - ref).max() <= tol * np.abs(ref).max(), (k, prec)
    finally:
        del os.environ["PCI_SELL_OFF"]
        g.close()
