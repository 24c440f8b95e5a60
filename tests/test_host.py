"""The Fortran host (libpci_host.so): input parsing and determinant
enumeration against the CPU oracle; no GPU calls."""
import numpy as np
import pytest

from pci_amd import host
from tests.helpers import oracle_for, space_dir


@pytest.mark.parametrize("name", ["c1", "t3", "t3b", "t4", "tiny3d"])
def test_host_space_equals_oracle(name):
    from oracle.oracle import calc_nd0
    sp = host.Space(space_dir(name))
    o = oracle_for(name)
    iarr, ndc, jz, nh, ll, jj, kk = sp.arrays()
    assert sp.Nd == o.Nd and sp.Nc == o.Nc and sp.Nst == o.Nst
    np.testing.assert_array_equal(iarr, o.iarr)
    np.testing.assert_array_equal(ndc, o.ndc)
    np.testing.assert_array_equal(jz, o.jz)
    np.testing.assert_array_equal(nh, o.nh)
    np.testing.assert_array_equal(ll, o.shells[:, 1])
    np.testing.assert_array_equal(jj, o.shells[:, 3])
    np.testing.assert_array_equal(kk, o.shells[:, 2])
    _, nd0 = calc_nd0(o.ndc, o.inp.Nc4, o.ci_in["MAXND0"])
    assert sp.Nd0 == nd0
    assert sp.Crt4 == o.inp.Crt4 and sp.Nlv == min(o.inp.Nlv, nd0)


def test_host_reports_missing_files(tmp_path):
    with pytest.raises(RuntimeError, match="CONF.INP"):
        host.Space(str(tmp_path))


@pytest.mark.parametrize("name", ["c1", "t3", "t4"])
def test_host_csf_classes_equal_oracle(tmp_path, name):
    """kCSF = 1: the host's nonequiv_conf (classes, and the entry reordering
    seen through the determinant list) against the oracle's."""
    import shutil
    from pci_amd import formats as F
    from oracle.oracle import Oracle
    w = tmp_path / name
    shutil.copytree(space_dir(name), w)
    cio = F.read_ci_in(str(w / "ci.in"))
    F.write_ci_in(str(w / "ci.in"), Kl=cio["KL"], MaxNd0=cio["MAXND0"], kCSF=1, Ksig=cio["KSIG"],
                  Kdsig=cio["KDSIG"])
    sp = host.Space(str(w))
    o = Oracle.from_dir(str(w))
    assert o.kcsf == 1
    nc_neq, n = sp.csf_classes()
    assert n == o.nconf_neq
    np.testing.assert_array_equal(nc_neq, o.nc_neq)
    iarr, ndc = sp.arrays()[:2]
    np.testing.assert_array_equal(iarr, o.iarr)
    np.testing.assert_array_equal(ndc, o.ndc)
