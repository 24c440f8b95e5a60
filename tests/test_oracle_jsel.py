"""J selection (Kv = 3) on the CPU: the restated Jterm against the J^2
spectrum (closed form: eigenvalues jx (jx + 2) with the listed
multiplicities), the Fortran host's Jterm against it, and the Davidson
restatement with FormB0 filter, AvgDiag and Prj_J against a dense solve
restricted to the target J."""
import numpy as np
import pytest

from tests.helpers import full_from_lower, oracle_for, space_dir

HARTREE_CM = 2 * 109737.315709


@pytest.mark.parametrize("name", ["tiny2", "tiny3d", "t3j", "c1j"])
def test_jterms_match_j2_spectrum(name):
    o = oracle_for(name)
    jt, njt = o.jterms()
    rp, col, val = o.formj()
    lam = np.linalg.eigvalsh(full_from_lower(rp, col, val).toarray())
    twoj = np.round(np.sqrt(1 + np.maximum(lam, 0)) - 1).astype(int)
    vals, cnt = np.unique(twoj, return_counts=True)
    np.testing.assert_array_equal(jt, vals)
    np.testing.assert_array_equal(njt, cnt)


@pytest.mark.parametrize("name", ["t3j", "c1j"])
def test_host_jterms_equal_oracle(name):
    from pci_amd import host
    sp = host.Space(space_dir(name))
    assert sp.Kv == 3
    np.testing.assert_array_equal(sp.jterms(), oracle_for(name).jterms()[0])


@pytest.mark.parametrize("name", ["t3j", "c1j"])
def test_diag4_jsel_against_dense(name):
    from oracle.davidson import diag4
    from oracle.oracle import calc_nd0
    o = oracle_for(name)
    rp, col, val, _, _ = o.formh()
    jrp, jcol, jval = o.formj()
    jt, _ = o.jterms()
    _, nd0 = calc_nd0(o.ndc, o.inp.Nc4, o.ci_in["MAXND0"])
    js = dict(xj=o.inp.J, jt=jt, jrp=jrp, jcol=jcol, jval=jval, ndc=o.ndc)
    tk, V, it, tr = diag4(rp, col, val, val.min(), nd0, o.inp.Nlv, o.inp.Crt4, o.inp.N_it, jsel=js)
    H = full_from_lower(rp, col, val).toarray()
    w, U = np.linalg.eigh(H)
    J2 = full_from_lower(jrp, jcol, jval).toarray()
    jv = 0.5 * (np.sqrt(1 + np.maximum(np.einsum("ij,ij->j", U, J2 @ U), 0)) - 1)
    exact = w[np.abs(jv - o.inp.J) < 0.1][:len(tk)]
    assert len(tk) == o.inp.Nlv
    assert np.abs(-tk - exact).max() * HARTREE_CM < 1e-3
    # every returned vector has the target J
    from oracle.oracle import jav
    for i in range(len(tk)):
        assert abs(jav(jrp, jcol, jval, V[i]) - o.inp.J) < 1e-4
