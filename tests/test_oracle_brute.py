"""The CPU oracle against a brute-force second-quantisation Hamiltonian
(closed-form 3j angular factors, explicit fermion signs): pins Rspq phases and
index pairing, Gint direct/exchange, the Gaunt table symmetries and Hint."""
import numpy as np
import pytest

from oracle.brute import Brute
from pci_amd import formats as F
from tests.helpers import inputs, oracle_for, space_dir


@pytest.mark.parametrize("name", ["tiny2", "tiny3", "tiny3d", "tiny2g"])
def test_oracle_matches_second_quantisation(name):
    d = space_dir(name)
    o = oracle_for(name)
    ci, nlx, gidx, gnt = inputs(d)
    _, shells = F.read_conf_dat(d + "/CONF.DAT")
    b = Brute(shells, o.Nso, ci.rint1, ci.iint1, ci.rint2, ci.iint2, ci.iint3, ci.nx, o.jz, o.nh)
    Hb = b.matrix(o.iarr, o.Npos)
    assert np.abs(Hb - Hb.T).max() < 1e-14
    rp, col, val, nz, nc = o.formh()
    Ho = np.zeros((o.Nd, o.Nd))
    for n in range(o.Nd):
        Ho[n, col[rp[n]:rp[n + 1]] - 1] = val[rp[n]:rp[n + 1]]
    L = np.tril(Hb)
    scale = np.abs(Hb).max()
    assert np.abs(L - Ho).max() < 1e-14 * scale
    # every element the brute force finds nonzero is stored by the oracle
    assert not ((np.abs(L) > 1e-12 * scale) & (Ho == 0)).any()


def test_single_excitation_phase_planted_mistake():
    """A wrong sign convention (the bitstring variant's missing hole count,
    determinants.f90:1014-1025) disagrees with the brute force."""
    from oracle.oracle import rspq
    # det1 = {1, 2}, det2 = {2, 3}: one excitation 1 -> 3 across position 2
    nf, is_, i2, i3, j2, j3 = rspq([1, 2], [2, 3])
    assert (nf, i2, j2) == (1, 1, 3)
    # second quantisation: a+_3 a_1 |1 2> = a+_3 |2> * (+1) = |2 3> * (-1)
    sign_sq = -1
    assert is_ == sign_sq
    # the bitstring rule without holes between the orbitals gives +1
    assert is_ != +1
