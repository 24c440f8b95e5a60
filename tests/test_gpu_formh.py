"""Device H build, J^2 build and SpMM against the CPU oracle (bit-exact for
pattern, indices and values; SpMM within FP64/FP32 rounding)."""
import numpy as np
import pytest

from tests.helpers import gpu_context_like, oracle_for, space_dir

pytestmark = pytest.mark.gpu

# t3s / t3sd: MBPT one-electron effective integrals (Ksig = 1), without and
# with energy dependence; t3s2 / t3sd2: Ksig = 2 (screened R^k and the
# two-electron sigma of SCRC.CON); t3gj: Hmltn's Gj J^2 term (CONF.INP Gj = 0.05)
SPACES = ["c1", "t3", "t3b", "t4", "t4b", "t4g", "t3s", "t3sd", "t3s2", "t3sd2", "t3gj"]


@pytest.fixture(scope="module", params=SPACES)
def case(request):
    name = request.param
    print(f"[case] space={name}", flush=True)
    o = oracle_for(name)
    g = gpu_context_like(o, space_dir(name))
    return name, o, g


def test_build_h_bit_exact(case):
    name, o, g = case
    rp, col, val, nz, ncand = o.formh()
    numh, numzero, hmin = g.build_h()
    assert numh == rp[-1], name
    assert numzero == nz, name
    assert hmin == val.min()
    grp, gcol, gval = g.get_csr()
    np.testing.assert_array_equal(grp, rp)
    np.testing.assert_array_equal(gcol, col)
    # bit-exact FP64 values (same term order, no contraction)
    np.testing.assert_array_equal(gval.view(np.int64), val.view(np.int64))


def test_build_j_bit_exact(case):
    name, o, g = case
    rp, col, val = o.formj()
    numj = g.build_j()
    assert numj == rp[-1]
    grp, gcol, gval = g.get_jcsr()
    np.testing.assert_array_equal(grp, rp)
    np.testing.assert_array_equal(gcol, col)
    np.testing.assert_allclose(gval, val, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("nvec", [1, 3, 4, 10, 16, 24])
def test_spmm_matches_mxmpy(case, nvec):
    from oracle.oracle import mxmpy
    name, o, g = case
    rp, col, val, _, _ = o.formh()
    g.build_h()
    hmin = val.min()
    rng = np.random.default_rng(7)
    x = rng.standard_normal((o.Nd, nvec))
    ref = np.stack([mxmpy(rp, col, val, hmin, x[:, v]) for v in range(nvec)], axis=1)
    y64, _, _ = g.spmm(x, 64)
    scale = np.abs(ref).max()
    assert np.abs(y64 - ref).max() <= 1e-12 * scale
    y32, _, _ = g.spmm(x, 32)
    assert np.abs(y32 - ref).max() <= 2e-5 * scale


@pytest.mark.parametrize("nvec", [1, 5, 10, 16])
def test_every_spmm_kernel_shape(case, nvec):
    """Each kernel shape the autotuner can pick, pinned in turn, against Mxmpy."""
    from oracle.oracle import mxmpy
    name, o, g = case
    rp, col, val, _, _ = o.formh()
    g.build_h()
    hmin = val.min()
    x = np.random.default_rng(11).standard_normal((o.Nd, nvec))
    ref = np.stack([mxmpy(rp, col, val, hmin, x[:, v]) for v in range(nvec)], axis=1)
    scale = np.abs(ref).max()
    n = g.spmm_set_variant(nvec, 64)
    for v in range(n):
        for prec, tol in ((64, 1e-12), (32, 2e-5)):
            g.spmm_set_variant(nvec, prec, v)
            y, _, _ = g.spmm(x, prec)
            err = np.abs(y - ref).max() / scale
            assert err <= tol, (name, nvec, prec, v, g.spmm_variant(nvec, prec), err)


@pytest.mark.parametrize("prec,tol", [(64, 1e-12), (32, 2e-5)])
def test_residual_blocks_are_built_and_match(case, prec, tol):
    """The residual-block copy (dense blocks per residual list) is built on
    every space here (Ne = 2, 3, 4) and its products match Mxmpy."""
    from oracle.oracle import mxmpy
    name, o, g = case
    rp, col, val, _, _ = o.formh()
    g.build_h()
    hmin = val.min()
    x = np.random.default_rng(13).standard_normal((o.Nd, 7))
    ref = np.stack([mxmpy(rp, col, val, hmin, x[:, v]) for v in range(7)], axis=1)
    names = {}
    n = g.spmm_set_variant(7, prec)
    for v in range(n):
        g.spmm_set_variant(7, prec, v)
        names[g.spmm_variant(7, prec)] = v
    for k in ("k_spmm_rblk<8 waves,4 chunks>", "k_spmm_rblk<16 waves,2 chunks>"):
        g.spmm_set_variant(7, prec, names[k])
        y, _, _ = g.spmm(x, prec)
        assert g.spmm_variant(7, prec) == k, (name, "copy not built")
        assert np.abs(y - ref).max() <= tol * np.abs(ref).max(), (name, k)
