"""The CSF basis (kCSF = 1) on the device at the measured workloads' full
size (C2: Ncsf = 8,517 of Nd = 69,101; C4: Ncsf = 48,587 of Nd = 1,079,368),
loaded through the Fortran host as pconf_gpu does, against the CPU oracle's
fixtures (tests/golden/make_scale_golden.py csf_c2 csf_c4):

* the host's configuration classes and the device's CSF counts equal the
  oracle's; the device's vectors span the oracle's spaces (projectors);
* on the fixture's coefficients (pci_gpu_csf_set_basis) the H rows of the
  sampled configurations equal the oracle's formh_sym bit for bit (and at C2
  NumH, numzero, Hmin of the whole H);
* the roots of the Davidson solve in the device's own basis: at C2 within
  0.01 cm^-1 of the oracle's Lanczos on its H; at C4 the FP32 mode against
  the plain FP64 method, and the determinant-basis solve's J = Jm roots
  (J from J^2) against the lowest CSF-basis roots -- the same states found
  in the two bases.
"""
import os

import numpy as np
import pytest

from pci_amd import formats as F
from tests.helpers import csf_dir, space_dir
from tests.test_scale_fixtures import csf_fixture_rows, load_scale

pytestmark = pytest.mark.gpu

HARTREE_CM = 2 * 109737.315709


def _nd0(ndcs, nc_neq, nc4, maxnd0):
    """calcNd0 over CSF counts (src/determinants.f90:22-45)."""
    n1, nd0 = 0, 0
    for ic in range(min(nc4, nc_neq.size)):
        n1 += int(ndcs[nc_neq[ic] - 1])
        if n1 < maxnd0 + 1:
            nd0 = n1
    return nd0


@pytest.fixture(scope="module", params=["c2", "c4"])
def scale(request):
    from pci_amd import host
    name = request.param
    print(f"[case] csf scale space={name}", flush=True)
    g = load_scale(f"csf_{name}")
    sp, ctx = host.load(csf_dir(name))
    nc_neq, nneq = sp.csf_classes()
    yield name, g, sp, ctx, nc_neq, nneq
    ctx.close()


def test_classes_basis_and_projectors(scale):
    name, g, sp, ctx, nc_neq, nneq = scale
    np.testing.assert_array_equal(nc_neq, g["nc_neq"])
    assert nneq == int(g["nconf_neq"])
    ndcs, ncsf, nccj = ctx.csf_basis(nc_neq, nneq)
    np.testing.assert_array_equal(ndcs, g["ndcs"])
    assert ncsf == int(g["ncsf"]) and nccj == g["ccj"].size
    ccj = ctx.csf_get_basis()
    first = {}
    for ic, q in enumerate(nc_neq):
        first.setdefault(int(q) - 1, ic)
    ndc = sp.arrays()[1]
    off = 0
    for q in range(nneq):
        m = int(ndc[first[q]]) if q in first else 0
        k = int(ndcs[q])
        cg = ccj[off:off + m * k].reshape(k, m).T
        co = g["ccj"][off:off + m * k].reshape(k, m).T
        off += m * k
        if k:
            np.testing.assert_allclose(cg @ cg.T, co @ co.T, atol=1e-12)


def test_rows_bit_exact_on_the_fixture_basis(scale):
    name, g, sp, ctx, nc_neq, nneq = scale
    ctx.csf_set_basis(nc_neq, g["ndcs"], g["ccj"])
    numh, numzero, hmin = ctx.build_h_csf()
    info = ctx.build_info()
    print(f"{name}: H_csf NumH={numh} numzero={numzero} built in {info['build_ms']:.1f} ms", flush=True)
    if "numh" in g:
        assert numh == int(g["numh"]) and numzero == int(g["numzero"]) and hmin == float(g["hmin"])
    conf_ncsf = g["ndcs"][nc_neq - 1].astype(np.int64)
    mdcs = np.concatenate([[0], np.cumsum(conf_ncsf)])
    for ic, (cnt, col, bits) in csf_fixture_rows(g).items():
        rp, c, v = ctx.get_csr_rows(int(mdcs[ic - 1]) + 1, int(conf_ncsf[ic - 1]))
        np.testing.assert_array_equal(np.diff(rp), cnt, err_msg=f"configuration {ic}")
        np.testing.assert_array_equal(c, col, err_msg=f"configuration {ic}")
        np.testing.assert_array_equal(v.view(np.int64), bits, err_msg=f"configuration {ic}")


def test_roots_in_the_device_basis(scale):
    from pci_amd import _lib
    name, g, sp, ctx, nc_neq, nneq = scale
    ndcs, ncsf, _ = ctx.csf_basis(nc_neq, nneq)
    ctx.build_h_csf()
    d = csf_dir(name)
    nc4 = F.read_conf_inp(os.path.join(d, "CONF.INP")).Nc4
    nd0 = _nd0(ndcs, nc_neq, nc4, F.read_ci_in(os.path.join(d, "ci.in"))["MAXND0"])
    nlv = min(sp.Nlv, nd0)
    NOADMIX = _lib.PCI_DVD_NOADMIX
    tk64, _, _, st64 = ctx.davidson(nlv, nd0, sp.Crt4, 300, NOADMIX)
    tk32, _, _, st32 = ctx.davidson(nlv, nd0, sp.Crt4, 300, NOADMIX | _lib.PCI_DVD_FP32)
    assert st64["converged"] == 1 and st32["converged"] == 1
    assert np.abs(tk32 - tk64).max() * HARTREE_CM < 0.01
    if "eigvals" in g:
        err = np.abs(-tk64 - g["eigvals"][:nlv]) * HARTREE_CM
        print(f"{name} CSF: {st64['iterations']} it, max |dE| vs Lanczos {err.max():.2e} cm^-1", flush=True)
        assert err.max() < 0.01
    else:
        # the same states in the determinant basis: its roots with J = Jm
        from pci_amd import host
        spd, cd = host.load(space_dir(name))
        try:
            cd.build_h()
            cd.build_j()
            tkd, tjd, _, std = cd.davidson(spd.Nlv, spd.Nd0, spd.Crt4, 300, NOADMIX)
            assert std["converged"] == 1
        finally:
            cd.close()
        sel = np.abs(tjd - spd.params["Jm"]) < 0.1
        print(f"{name}: determinant-basis roots {tkd}, J {tjd}; CSF-basis roots {tk64}", flush=True)
        assert sel.any()
        m = int(sel.sum())
        assert np.abs(tkd[sel] - tk64[:m]).max() * HARTREE_CM < 0.01
