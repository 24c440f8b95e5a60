"""MBPT effective integrals (Kl = 2, Ksig = 1 and 2): the CPU oracle's
HintS (src/integrals.f90:321-369), FormD (src/pconf.F90:852-880) and the
E_k = Diag(column) rule (src/pconf.F90:1275) against the reference's own
integrals module compiled into oracle/_ref, and, without energy dependence,
against the second-quantisation Hamiltonian with h + sigma."""
import numpy as np
import pytest

from oracle import ref
from oracle.brute import Brute
from pci_amd import formats as F
from tests.helpers import inputs, oracle_for, space_dir


def _formh(r, nrow):
    return r.formh(1, min(nrow, r.Nd))


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("name", ["tiny3s", "t3s", "t3sd"])
def test_sigma_formh_matches_reference(name):
    d = space_dir(name)
    o = oracle_for(name)
    assert o.ksig == 1
    rp, col, val = ref.run_in_child(d, _formh, o.Nd)
    orp, ocol, oval, _, _ = o.formh()
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(val.view(np.int64), oval.view(np.int64))


def test_sigma_changes_h_by_the_one_body_operator():
    d = space_dir("tiny3s")
    o = oracle_for("tiny3s")
    base = oracle_for("tiny3")          # same seed and basis, no SGC.CON
    np.testing.assert_array_equal(o.iarr, base.iarr)
    ci, nlx, gidx, gnt = inputs(d)
    _, shells = F.read_conf_dat(d + "/CONF.DAT")
    b = Brute(shells, o.Nso, ci.rint1, ci.iint1, ci.rint2, ci.iint2, ci.iint3, ci.nx, o.jz, o.nh)
    nmaxs, lmaxs, _, ent = F.read_sgc_con(d + "/SGC.CON")
    assert any(nb > na for na, nb, *_ in ent) and nmaxs < len(shells)   # off-diagonal pairs and a cut-off
    for na, nb, r, _, _ in ent:         # sigma is added to h for every listed pair
        b.h[na, nb] += r
        if nb != na:
            b.h[nb, na] += r
    Hb = b.matrix(o.iarr, o.Npos)
    rp, col, val, _, _ = o.formh()
    Ho = np.zeros((o.Nd, o.Nd))
    for n in range(o.Nd):
        Ho[n, col[rp[n]:rp[n + 1]] - 1] = val[rp[n]:rp[n + 1]]
    scale = np.abs(Hb).max()
    assert np.abs(np.tril(Hb) - Ho).max() < 1e-14 * scale
    # and sigma did change H (a mistake that drops it fails above only if this holds)
    brp, bcol, bval, _, _ = base.formh()
    Hbase = np.zeros((o.Nd, o.Nd))
    for n in range(o.Nd):
        Hbase[n, bcol[brp[n]:brp[n + 1]] - 1] = bval[brp[n]:brp[n + 1]]
    assert np.abs(Ho - Hbase).max() > 1e-5


def test_sgc_con_round_trip(tmp_path):
    ent = [(1, 1, np.float32(-1.25e-3), np.float32(2.5e-4), np.float32(-0.75)), (1, 4, np.float32(3.0e-5), 0.0, 0.5)]
    p = str(tmp_path / "SGC.CON")
    F.write_sgc_con(p, 7, 2, 1, [(a, b, float(r), float(dd), float(e)) for a, b, r, dd, e in ent])
    head = open(p).readline()
    # the (7X,I3,5X,I1,24X,I1) header: NmaxS in columns 8-10, LmaxS in 16, Khot in 41
    assert head[7:10] == "  7" and head[15] == "2" and head[40] == "1"
    nmaxs, lmaxs, khot, got = F.read_sgc_con(p)
    assert (nmaxs, lmaxs, khot) == (7, 2, 1)
    assert [(a, b) for a, b, *_ in got] == [(1, 1), (1, 4)]
    assert [g[2] for g in got] == [float(e[2]) for e in ent]


@pytest.mark.parametrize("name", ["t3s", "t3sd"])
def test_host_reads_sigma_tables(name):
    """The Fortran host's RintS/Eps/FormD restatement gives the oracle's tables."""
    from pci_amd import host
    d = space_dir(name)
    o = oracle_for(name)
    sp = host.Space(d)
    assert sp.Kl == 0                    # Kl = 2 runs start anew
    t = sp.sigma()
    sg = o.sigma
    assert (t["ksig"], t["kdsig"], t["nmaxs"], t["lmaxs"]) == (1, sg["kdsig"], sg["nmaxs"], sg["lmaxs"])
    np.testing.assert_array_equal(t["iint1s"], sg["ind"])
    for k, kk in (("rsig", "r"), ("dsig", "d"), ("esig", "e")):
        np.testing.assert_array_equal(t[k].view(np.int64), sg[kk].view(np.int64))
    np.testing.assert_array_equal(t["eps"][:o.Nsu], sg["eps"][:o.Nsu])
    if sg["kdsig"]:
        want = np.array([o.diags(n) for n in range(1, o.Nd + 1)])
        np.testing.assert_array_equal(t["diag"].view(np.int64), want.view(np.int64))
    else:
        assert t["diag"] is None


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("name", ["t3s2", "t3sd2"])
def test_screened_sigma_formh_matches_reference(name):
    """Ksig = 2: Gint's screening of R^k (src/integrals.f90:262-271, with the
    reference's test as it evaluates: the local shell nd and an unset Lmax)
    and GintS (371-493, Ksym = 0 with Kmax = 3 in t3s2, energy dependence in
    t3sd2) against the reference's own integrals module, bit for bit."""
    d = space_dir(name)
    o = oracle_for(name)
    assert o.ksig == 2
    rp, col, val = ref.run_in_child(d, _formh, o.Nd)
    orp, ocol, oval, _, _ = o.formh()
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(val.view(np.int64), oval.view(np.int64))


def test_screening_and_two_electron_sigma_change_h():
    """t3s2 is t3s (same seed, same one-electron sigma) with SCRC.CON and the
    screening coefficients added: H must change by more than rounding, and
    by less than its own scale."""
    rp1, col1, val1, _, _ = oracle_for("t3s").formh(1, 400)
    rp2, col2, val2, _, _ = oracle_for("t3s2").formh(1, 400)
    d1, d2 = {}, {}
    rows = np.repeat(np.arange(400), np.diff(rp1))
    d1 = dict(zip(zip(rows, col1), val1))
    rows = np.repeat(np.arange(400), np.diff(rp2))
    d2 = dict(zip(zip(rows, col2), val2))
    keys = set(d1) | set(d2)
    diff = max(abs(d1.get(k, 0.0) - d2.get(k, 0.0)) for k in keys)
    assert 1e-4 < diff < 0.5 * np.abs(val1).max(), diff
