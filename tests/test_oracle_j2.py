"""J^2 from the oracle's FormJ (formj2.f90:120-184) against the closed form.
F_J2 works in doubled angular momenta, so the stored matrix is 4 J^2: within
a space of fixed M_J every eigenvalue is 4 J(J+1) with J >= |M_J| and 2J of
the parity of the electron count, and J_av returns J = (sqrt(1 + x.J2.x) - 1)/2
(formj2.f90:381).  A wrong sign, index or factor in F_J2/Plj breaks this
quantisation."""
import numpy as np
import pytest

from tests.helpers import full_from_lower, oracle_for


@pytest.mark.parametrize("name", ["tiny2", "tiny3", "tiny3d", "c1"])
def test_j2_spectrum_is_quantised(name):
    o = oracle_for(name)
    rp, col, val = o.formj()
    J2 = full_from_lower(rp, col, val).toarray()
    lam = np.linalg.eigvalsh(J2)
    J = 0.5 * (np.sqrt(1.0 + np.maximum(lam, 0.0)) - 1.0)
    twoJ = 2.0 * J
    assert np.abs(twoJ - np.round(twoJ)).max() < 1e-9, twoJ
    parity = o.Ne % 2
    assert np.all(np.round(twoJ).astype(int) % 2 == parity)
    assert np.all(np.round(twoJ) >= o.Mj - 1e-9)
    # every allowed J between |M| and the largest one found occurs
    found = set(np.round(twoJ).astype(int).tolist())
    assert min(found) == o.Mj


def test_jav_of_eigenvectors():
    """J_av (formj2.f90:355-387) of J^2 eigenvectors gives their J."""
    from oracle.oracle import jav
    o = oracle_for("tiny3d")
    rp, col, val = o.formj()
    J2 = full_from_lower(rp, col, val).toarray()
    lam, vec = np.linalg.eigh(J2)
    for i in (0, len(lam) // 2, len(lam) - 1):
        assert abs(jav(rp, col, val, vec[:, i]) - 0.5 * (np.sqrt(1.0 + max(lam[i], 0.0)) - 1.0)) < 1e-10


def test_gj_term_is_gj_times_j2():
    """Hmltn's Gj term (src/pconf.F90:1653-1692): with CONF.INP Gj = 0.05 the
    oracle's H equals the Gj = 0 H of the same space plus Gj times the J^2
    operator the closed-form test above pins (t3gj is t3 with Gj set)."""
    o0, og = oracle_for("t3"), oracle_for("t3gj")
    assert og.inp.Gj == 0.05 and o0.inp.Gj == 0.0
    rp0, col0, val0, _, _ = o0.formh()
    rpg, colg, valg, _, _ = og.formh()
    H0 = full_from_lower(rp0, col0, val0)
    Hg = full_from_lower(rpg, colg, valg)
    J2 = full_from_lower(*o0.formj())
    d = (Hg - H0 - 0.05 * J2).toarray()
    assert np.abs(d).max() <= 1e-12 * abs(H0).max(), np.abs(d).max()
    assert abs(0.05 * J2).max() > 0.1  # the term is there
