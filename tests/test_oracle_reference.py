"""The CPU oracle against the reference's own Fortran modules (determinants,
integrals, conf_init, wigner) compiled from ud-pci/pCI's sources into
oracle/_ref.  Skipped where oracle/_ref was not built (reference sources absent)."""
import numpy as np
import pytest

from oracle import ref
from pci_amd import gen
from tests.helpers import oracle_for, space_dir

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")


def _space_and_formh(r, nrow):
    iarr, ndc, jz, nh = r.space()
    rp, col, val = r.formh(1, min(nrow, r.Nd))
    return iarr, ndc, jz, nh, rp, col, val


@pytest.mark.parametrize("name,nrow", [("c1", 1216), ("t3", 3819), ("t3b", 3819), ("t4", 2000), ("tiny3d", 305),
                                       ("tiny2g", 85), ("t4g", 3587)])
def test_enumeration_and_formh_bitwise(name, nrow):
    d = space_dir(name)
    o = oracle_for(name)
    iarr, ndc, jz, nh, rp, col, val = ref.run_in_child(d, _space_and_formh, nrow)
    np.testing.assert_array_equal(iarr, o.iarr)
    np.testing.assert_array_equal(ndc, o.ndc)
    np.testing.assert_array_equal(jz, o.jz[:jz.size])
    np.testing.assert_array_equal(nh, o.nh[:nh.size])
    orp, ocol, oval, _, _ = o.formh(1, min(nrow, o.Nd))
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(val.view(np.int64), oval.view(np.int64))


def _rspq_many(r, pairs):
    return [r.rspq(a, b) for a, b in pairs]


def test_rspq_random_pairs():
    d = space_dir("t4")
    o = oracle_for("t4")
    rng = np.random.default_rng(5)
    idx = rng.integers(0, o.Nd, size=(3000, 2))
    pairs = [(o.iarr[a], o.iarr[b]) for a, b in idx]
    from oracle.oracle import rspq
    mine = [rspq(a, b) for a, b in pairs]
    theirs = ref.run_in_child(d, _rspq_many, pairs)
    for m, t in zip(mine, theirs):
        if m[0] <= 2:
            assert m == t
        else:
            assert t[0] == 3


def _gaun_table(r, entries):
    return [r.L.ref_gaun(k, j1 / 2, m1 / 2, j2 / 2, m2 / 2) for (k, j1, m1, j2, m2) in entries]


def test_gaunt_table_against_reference_fj3():
    """CONF.GNT values of the generator (exact Racah) against breit.f90's Gaun
    evaluated with the reference FJ3."""
    nlx, counts, idx, val = gen.gaunt_table(5)
    _, ent = gen.gaunt_entries(5)
    entries = [(k, j1, m1, j2, m2) for (_, k, j1, m1, j2, m2, _) in ent]
    ref_vals = np.array(ref.run_in_child(space_dir("c1"), _gaun_table, entries))
    np.testing.assert_allclose(val, ref_vals, rtol=0, atol=1e-14)


def _jterms(r):
    return r.jterms()


@pytest.mark.parametrize("name", ["t3j", "c1j", "tiny3d"])
def test_jterms_against_reference(name):
    jt, njt = ref.run_in_child(space_dir(name), _jterms)
    ojt, onjt = oracle_for(name).jterms()
    np.testing.assert_array_equal(jt, ojt)
    np.testing.assert_array_equal(njt, onjt)
