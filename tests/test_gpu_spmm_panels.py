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


# widths in rising order: each wider width whose LDS panel is narrower than
# the copy's panels rebuilds the copy, which the narrower widths then share
@pytest.mark.parametrize("nvec", [1, 8, 10, 16, 24])
def test_sliced_kernels_match_csr_kernel(c2, nvec):
    sp, g = c2
    x = np.random.default_rng(5).standard_normal((sp.Nd, nvec))
    n = g.spmm_set_variant(nvec, 64)
    names = {}
    for v in range(n):
        g.spmm_set_variant(nvec, 64, v)
        names[g.spmm_variant(nvec, 64)] = v
    sliced = [v for k, v in names.items() if k.startswith("k_spmm_sell")]
    assert len(sliced) == 5
    for prec, tol in ((64, 1e-12), (32, 2e-5)):
        g.spmm_set_variant(nvec, prec, names["k_spmm"])
        ref, _, _ = g.spmm(x, prec)
        scale = np.abs(ref).max()
        for v in sliced:
            g.spmm_set_variant(nvec, prec, v)
            y, _, _ = g.spmm(x, prec)
            err = np.abs(y - ref).max() / scale
            print(f"{g.spmm_variant(nvec, prec)} fp{prec}: {err:.1e}", flush=True)
            assert err <= tol, (nvec, prec, g.spmm_variant(nvec, prec), err)
