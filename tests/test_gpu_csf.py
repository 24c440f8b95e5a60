"""CSF basis (kCSF = 1) on the device against the CPU oracle:

* the basis: CSFs per class equal to jbasis's; each class's vectors span the
  same space as the oracle's (equal projectors C C^T), orthonormal;
* H in the CSF basis: pattern and every value bit for bit equal to the
  oracle's formh_sym run on the device's basis (the eigenvectors of a
  degenerate J are basis-dependent, so the basis is shared), NumH, numzero
  and Hmin equal;
* the Davidson solve in that basis: roots against the oracle's Diag4 on the
  oracle's own H_csf and against a dense solve (within 1e-3 cm^-1), plain
  FP64 and FP32 + FP64 refinement;
* unsym: bit for bit equal to the oracle's on the same vectors and basis.
"""
import numpy as np
import pytest

from oracle import csf
from oracle.oracle import Oracle, calc_nd0
from tests.helpers import full_from_lower, gpu_context_like, space_dir

pytestmark = pytest.mark.gpu

HARTREE_CM = 2 * 109737.315709
# t3s / t3sd / t3s2: MBPT Ksig = 1 (with Kdsig: E_k of the column determinant) and Ksig = 2
SPACES = ["c1", "t3", "t3b", "t4", "t4g", "t3s", "t3sd", "t3s2"]


@pytest.fixture(scope="module", params=SPACES)
def case(request):
    name = request.param
    print(f"[case] csf space={name}", flush=True)
    d = space_dir(name)
    o = Oracle.from_dir(d, kcsf=1)
    g = gpu_context_like(o, d)
    ndcs, ncsf, nccj = g.csf_basis(o.nc_neq, o.nconf_neq)
    b_o = csf.jbasis(o)
    b_g = csf.CsfBasis(o, ndcs, g.csf_get_basis())
    return name, o, g, b_o, b_g


def test_basis_matches_jbasis(case):
    name, o, g, b_o, b_g = case
    np.testing.assert_array_equal(b_g.ndcs, b_o.ndcs)
    assert g.ncsf == b_o.ncsf and g.nccj == b_o.nccj
    for q in range(o.nconf_neq):
        cg, co = b_g.block(q), b_o.block(q)
        if cg.size == 0:
            continue
        np.testing.assert_allclose(cg.T @ cg, np.eye(cg.shape[1]), atol=1e-12)
        np.testing.assert_allclose(cg @ cg.T, co @ co.T, atol=1e-12)


def test_h_csf_bit_exact(case):
    name, o, g, b_o, b_g = case
    rp, col, val, nz, hmin, _ = csf.formh_sym(o, b_g)
    numh, numzero, ghmin = g.build_h_csf()
    assert numh == rp[-1] and numzero == nz and ghmin == hmin, (name, numh, rp[-1], numzero, nz)
    grp, gcol, gval = g.get_csr()
    np.testing.assert_array_equal(grp, rp)
    np.testing.assert_array_equal(gcol, col)
    np.testing.assert_array_equal(gval.view(np.int64), val.view(np.int64))


@pytest.mark.parametrize("flags", [0, 1, 1 | 16 | 64])  # plain, FP32 + refinement, coupled FP32 (residual stop)
def test_davidson_in_csf_basis(case, flags):
    from oracle.davidson import diag4
    name, o, g, b_o, b_g = case
    g.build_h_csf()
    _, nd0 = calc_nd0(b_g.conf_ncsf, o.inp.Nc4, o.ci_in["MAXND0"])
    nlv = min(o.inp.Nlv, nd0)
    tk, tj, vecs, st = g.davidson(nlv, nd0, o.inp.Crt4, o.inp.N_it, flags)
    assert st["converged"] in (1, 2), st
    # the oracle's Diag4 on its own basis's H (the roots do not depend on the basis)
    rp, col, val, _, hmin, _ = csf.formh_sym(o, b_o)
    tk_o, _, _, _ = diag4(rp, col, val, hmin, nd0, nlv, o.inp.Crt4, o.inp.N_it)
    assert np.abs(tk - tk_o).max() * HARTREE_CM < 1e-3, (tk, tk_o)
    w = np.linalg.eigvalsh(full_from_lower(rp, col, val).toarray())
    assert np.abs(tk + w[:nlv]).max() * HARTREE_CM < 1e-3
    np.testing.assert_array_equal(tj, 0.0)  # no J^2 in this basis (the caller reports Jm)


def test_unsym_bit_exact(case):
    name, o, g, b_o, b_g = case
    x = np.random.default_rng(3).standard_normal((3, g.ncsf))
    y = g.csf_unsym(x)
    y_o = csf.unsym(o, b_g, x)
    np.testing.assert_array_equal(y.view(np.int64), y_o.view(np.int64))
