"""The CPU oracle of the CSF-basis path (kCSF = 1) pinned by what the
mathematics fixes, since the reference ships no CSF fixtures:

* nonequiv_conf: classes are exactly the configurations with equal multisets
  of (l, 2j, occupation); the reordering makes every configuration's J^2
  block equal to its class representative's, element for element (without
  it the blocks differ), and leaves each configuration's determinant set
  unchanged;
* jbasis: a configuration's number of J = Jm states is N(M = Jm) - N(M =
  Jm + 1) (odometer counts); the kept vectors are orthonormal eigenvectors of
  the J^2 block (F_J2's doubled units, 4 J(J+1)) with eigenvalue
  2Jm (2Jm + 2);
* formh_sym: the stored pattern is every configuration pair within two
  electrons (brute force over shell multisets); the values equal C^T H C
  formed densely from the oracle's determinant-basis FormH; the spectrum of
  H in the CSF basis is the J = Jm part of the determinant-basis spectrum;
* unsym: it maps eigenvectors of H_csf to eigenvectors of H with the same
  eigenvalue and J = Jm, and is an isometry.
"""
from collections import Counter

import numpy as np
import pytest

from oracle import csf
from oracle.oracle import Oracle
from tests.helpers import full_from_lower, space_dir

SPACES = ["c1", "t3", "t4g"]


@pytest.fixture(scope="module", params=SPACES)
def case(request):
    d = space_dir(request.param)
    o = Oracle.from_dir(d, kcsf=1)
    o0 = Oracle.from_dir(d, kcsf=0)
    return request.param, o, o0, csf.jbasis(o)


def _conf_dets(o):
    md = np.concatenate([[0], np.cumsum(o.ndc)])
    return [o.iarr[md[i]:md[i + 1]] for i in range(o.Nc)]


def _signature(o, ic):
    """Multiset of (l, 2j, occupation) of configuration ic (0-based), from its
    first determinant's positions."""
    d = _conf_dets(o)[ic]
    if len(d) == 0:
        return None
    sh = Counter(int(o.nh[p - 1]) for p in d[0])
    return tuple(sorted((int(o.shells[s - 1, 1]), int(o.shells[s - 1, 3]), q) for s, q in sh.items()))


def test_classes_are_equal_shell_patterns(case):
    name, o, _, _ = case
    sig = [_signature(o, i) for i in range(o.Nc)]
    for i in range(o.Nc):
        for j in range(i):
            if sig[i] is None or sig[j] is None:
                continue
            assert (o.nc_neq[i] == o.nc_neq[j]) == (sig[i] == sig[j]), (name, i + 1, j + 1)
    assert o.nconf_neq == len(set(o.nc_neq.tolist()))


def test_reordering_keeps_determinant_sets(case):
    name, o, o0, _ = case
    np.testing.assert_array_equal(o.ndc, o0.ndc)
    for a, b in zip(_conf_dets(o), _conf_dets(o0)):
        assert sorted(map(tuple, a)) == sorted(map(tuple, b))


def test_class_members_share_the_j2_block(case):
    name, o, o0, _ = case
    first = csf.class_representatives(o)
    differs = False
    for ic in range(1, o.Nc + 1):
        rep = first[o.nc_neq[ic - 1] - 1]
        np.testing.assert_array_equal(csf.j2_block(o, ic), csf.j2_block(o, rep), err_msg=f"{name} conf {ic}")
        # without the reordering (pconf's order for kCSF = 0) the blocks differ
        o0.nc_neq, o0.nconf_neq = o.nc_neq, o.nconf_neq
        if not np.array_equal(csf.j2_block(o0, ic), csf.j2_block(o0, rep)):
            differs = True
    if name == "t3":  # (c1's two-electron blocks are symmetric under the swap)
        assert differs, "the reordering was not exercised"


def test_csf_counts_and_vectors(case):
    name, o, _, b = case
    for ic in range(1, o.Nc + 1):
        n_m = o.L.orc_count_m(o.c, ic, o.Mj)
        n_m1 = o.L.orc_count_m(o.c, ic, o.Mj + 2)
        assert b.conf_ncsf[ic - 1] == n_m - n_m1, (name, ic)
    jj = float(o.Mj * (o.Mj + 2))  # F_J2 works in doubled units: 4 J(J+1) = Mj (Mj + 2)
    first = csf.class_representatives(o)
    for q in range(o.nconf_neq):
        cb = b.block(q)
        if cb.size == 0:
            continue
        np.testing.assert_allclose(cb.T @ cb, np.eye(cb.shape[1]), atol=1e-12)
        j2 = csf.j2_block(o, first[q])
        np.testing.assert_allclose(j2 @ cb, jj * cb, atol=1e-10)


def _dense_c(o, b):
    md = np.concatenate([[0], np.cumsum(o.ndc)])
    cm = np.zeros((o.Nd, b.ncsf))
    for ic in range(o.Nc):
        q = b.nc_neq[ic] - 1
        if b.ndcs[q]:
            cm[md[ic]:md[ic + 1], b.mdcs[ic]:b.mdcs[ic] + b.ndcs[q]] = b.block(q)
    return cm


def test_formh_sym_pattern_values_and_spectrum(case):
    name, o, _, b = case
    rp, col, val, nz, hmin, npairs = csf.formh_sym(o, b)
    assert nz == int(np.sum(val == 0.0))
    assert hmin == val.min()
    # pattern: configuration pairs within two electrons, every CSF pair
    sig = []
    for i in range(o.Nc):
        d = _conf_dets(o)[i]
        sig.append(Counter(int(o.nh[p - 1]) for p in d[0]) if len(d) else None)
    conf_of = np.repeat(np.arange(o.Nc), b.conf_ncsf)
    rows = np.repeat(np.arange(b.ncsf), np.diff(rp))
    got = set(zip(rows.tolist(), (col - 1).tolist()))
    want = set()
    for r in range(b.ncsf):
        i = conf_of[r]
        for s in range(r + 1):
            j = conf_of[s]
            if sum((sig[i] - sig[j]).values()) <= 2:
                want.add((r, s))
    assert got == want
    # values: C^T H C from the determinant-basis FormH
    orp, ocol, oval, _, _ = o.formh()
    H = full_from_lower(orp, ocol, oval).toarray()
    cm = _dense_c(o, b)
    ref = cm.T @ H @ cm
    Hc = full_from_lower(rp, col, val).toarray()
    assert np.abs(Hc - ref).max() <= 1e-13 * np.abs(ref).max()
    # spectrum: the J = Jm eigenvalues of H
    w = np.linalg.eigvalsh(Hc)
    wd, vd = np.linalg.eigh(H)
    jrp, jcol, jval = o.formj()
    J2 = full_from_lower(jrp, jcol, jval).toarray()
    jv = np.einsum("ij,ij->j", vd, J2 @ vd)
    tj = 0.5 * (np.sqrt(1.0 + jv) - 1.0)
    sel = np.abs(tj - 0.5 * o.Mj) < 0.25
    assert sel.sum() == b.ncsf
    np.testing.assert_allclose(np.sort(wd[sel]), w, atol=1e-10)


def test_unsym_maps_eigenvectors(case):
    name, o, _, b = case
    rp, col, val, _, _, _ = csf.formh_sym(o, b)
    Hc = full_from_lower(rp, col, val).toarray()
    w, x = np.linalg.eigh(Hc)
    k = min(6, b.ncsf)
    y = csf.unsym(o, b, x[:, :k].T)
    orp, ocol, oval, _, _ = o.formh()
    H = full_from_lower(orp, ocol, oval)
    jrp, jcol, jval = o.formj()
    J2 = full_from_lower(jrp, jcol, jval)
    jj = float(o.Mj * (o.Mj + 2))
    for i in range(k):
        assert abs(np.linalg.norm(y[i]) - 1.0) < 1e-12
        assert np.abs(H @ y[i] - w[i] * y[i]).max() < 1e-10
        assert np.abs(J2 @ y[i] - jj * y[i]).max() < 1e-10
    # unsym is C x: the dense product agrees (different summation order)
    np.testing.assert_allclose(y, (_dense_c(o, b) @ x[:, :k]).T, atol=1e-14)
