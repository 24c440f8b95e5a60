"""The pconf_gpu program (Fortran host + HIP library) as a drop-in for pconf:
same input files, pCONF.HIJ / CONF.DET / CONF.XIJ outputs (and, with
PCI_CONF_HIJ=1, the legacy CONF.HIJ / CONF.JJJ triples) checked against the
CPU oracle; then Kl=1 reuses the written pCONF.HIJ."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from pci_amd import formats as F
from pci_amd.host import PROG_PATH
from tests.helpers import oracle_for, space_dir

pytestmark = pytest.mark.gpu


def _run(workdir, env_extra=None, stdin_text=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    r = subprocess.run([PROG_PATH], cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300, input=stdin_text if stdin_text is not None else "", env=env)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


@pytest.mark.parametrize("name", ["t3", "c1", "t3j", "t3sd", "t3sd2"])
def test_pconf_gpu_outputs(tmp_path, name):
    print(f"[case] pconf_gpu space={name}", flush=True)
    src = space_dir(name)
    w = tmp_path / name
    shutil.copytree(src, w)
    cio = F.read_ci_in(str(w / "ci.in"))
    maxnd0 = cio["MAXND0"]
    # Kl = 2 (MBPT, t3sd, t3sd2) is kept: Ksig = 1 or 2 with the valence energy on stdin
    F.write_ci_in(str(w / "ci.in"), Kl=cio["KL"], Kw=1, MaxNd0=maxnd0, Ksig=cio["KSIG"], Kdsig=cio["KDSIG"])
    e0 = open(w / "E0.in").read() if (w / "E0.in").exists() else None
    out = _run(str(w), {"PCI_CONF_HIJ": "1"}, e0)
    if cio["KSIG"]:
        assert "Give valence energy" in out
    assert "TIMING >>> FormH (GPU)" in out
    o = oracle_for(name)
    # CONF.DET: the determinant list, bit for bit
    iarr, nsu = F.read_conf_det(str(w / "CONF.DET"))
    np.testing.assert_array_equal(iarr, o.iarr)
    # pCONF.HIJ: global CSR of the lower triangle, bit for bit
    rp, col, val = F.read_hij(str(w / "pCONF.HIJ"))
    orp, ocol, oval, _, _ = o.formh()
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(val.view(np.int64), oval.view(np.int64))
    # pCONF.JJJ
    jrp, jcol, jval = F.read_hij(str(w / "pCONF.JJJ"))
    ojrp, ojcol, ojval = o.formj()
    np.testing.assert_array_equal(jrp, ojrp)
    np.testing.assert_array_equal(jcol, ojcol)
    # legacy CONF.HIJ / CONF.JJJ triples: the nonzero elements in row order
    for fname, (r_, c_, v_) in (("CONF.HIJ", (orp, ocol, oval)), ("CONF.JJJ", (ojrp, ojcol, ojval))):
        n, k, t = F.read_conf_hij(str(w / fname))
        rows = np.repeat(np.arange(1, r_.size, dtype=np.int32), np.diff(r_))
        nz = v_ != 0.0
        np.testing.assert_array_equal(n, rows[nz])
        np.testing.assert_array_equal(k, c_[nz])
        np.testing.assert_array_equal(t.view(np.int64), v_[nz].view(np.int64))
    # CONF.XIJ: energies of the CPU Davidson oracle
    from oracle.davidson import diag4
    tk, tj, vecs = F.read_xij(str(w / "CONF.XIJ"))
    from oracle.oracle import calc_nd0
    _, nd0 = calc_nd0(o.ndc, o.inp.Nc4, maxnd0)
    js = None
    if o.inp.Kv == 3:  # selection of states with J = XJ_av
        assert "Selection of states with J" in out
        js = dict(xj=o.inp.J, jt=o.jterms()[0], jrp=ojrp, jcol=ojcol, jval=ojval, ndc=o.ndc)
        np.testing.assert_allclose(tj, o.inp.J, atol=1e-4)
    tk_o, v_o, _, _ = diag4(orp, ocol, oval, oval.min(), nd0, tk.size, o.inp.Crt4, o.inp.N_it, jsel=js)
    assert np.abs(tk - tk_o).max() * 2 * 109737.315709 < 1e-3
    assert vecs.shape == (tk.size, o.Nd)
    # Kl=1: read pCONF.HIJ back instead of building H
    F.write_ci_in(str(w / "ci.in"), Kl=1, MaxNd0=maxnd0)
    out = _run(str(w))
    assert "H read from pCONF.HIJ" in out
    tk2, _, _ = F.read_xij(str(w / "CONF.XIJ"))
    assert np.abs(tk2 - tk).max() * 2 * 109737.315709 < 1e-4


@pytest.mark.parametrize("name", ["t3", "t4", "t3sd"])
def test_pconf_gpu_csf_basis(tmp_path, name):
    """kCSF = 1: CONF.DET in the order nonequiv_conf's reordering gives, the
    energies of the CSF-basis solve against the oracle's formh_sym and a
    dense solve, J = Jm, CONF.WFS (CSF basis) and CONF.XIJ (determinant
    basis, unsym) vectors: unit length, J^2 = Jm(Jm+1) and Rayleigh quotients
    equal to the energies."""
    from oracle import csf
    from oracle.oracle import Oracle
    from tests.helpers import full_from_lower
    print(f"[case] pconf_gpu kCSF=1 space={name}", flush=True)
    w = tmp_path / name
    shutil.copytree(space_dir(name), w)
    cio = F.read_ci_in(str(w / "ci.in"))
    F.write_ci_in(str(w / "ci.in"), Kl=cio["KL"], MaxNd0=cio["MAXND0"], kCSF=1, Ksig=cio["KSIG"],
                  Kdsig=cio["KDSIG"])
    e0 = open(w / "E0.in").read() if (w / "E0.in").exists() else None
    out = _run(str(w), None, e0)
    assert "Ncsf=" in out and "TIMING >>> FormH_sym (GPU)" in out and "Symmetry: CSF" in out
    assert "Number of nonequivalent configurations:" in out
    o = Oracle.from_dir(str(w))
    assert o.kcsf == 1
    iarr, _ = F.read_conf_det(str(w / "CONF.DET"))
    np.testing.assert_array_equal(iarr, o.iarr)
    b = csf.jbasis(o)
    # CSF.LOG: one line per class (class, first configuration, Ndi, Ncsfi, core)
    log = [ln.split() for ln in open(w / "CSF.LOG").read().splitlines()[1:]]
    assert len(log) == o.nconf_neq
    first = csf.class_representatives(o)
    for q, ic, ndi, nci, _ in (map(int, r) for r in log):
        assert ic == first[q - 1] and ndi == o.ndc[ic - 1] and nci == b.ndcs[q - 1]
    assert [int(r[2]) for r in log] == sorted((int(r[2]) for r in log), reverse=True)
    rp, col, val, _, _, _ = csf.formh_sym(o, b)
    ev = np.linalg.eigvalsh(full_from_lower(rp, col, val).toarray())
    tk, tj, vecs = F.read_xij(str(w / "CONF.XIJ"))
    nlv = tk.size
    assert np.abs(tk + ev[:nlv]).max() * 2 * 109737.315709 < 1e-3
    np.testing.assert_array_equal(tj, o.inp.Jm)
    tkw, tjw, xw = F.read_xij(str(w / "CONF.WFS"))
    np.testing.assert_array_equal(tkw, tk)
    assert xw.shape == (nlv, b.ncsf) and vecs.shape == (nlv, o.Nd)
    orp, ocol, oval, _, _ = o.formh()
    H = full_from_lower(orp, ocol, oval)
    jrp, jcol, jval = o.formj()
    J2 = full_from_lower(jrp, jcol, jval)
    for i in range(nlv):
        y = vecs[i]
        np.testing.assert_allclose(np.linalg.norm(y), np.linalg.norm(xw[i]), rtol=1e-12)
        assert abs(np.linalg.norm(y) - 1.0) < 1e-9
        assert np.abs(J2 @ y - o.Mj * (o.Mj + 2) * y).max() < 1e-9
        if not o.ci_in["KDSIG"]:  # (with Kdsig the determinant-basis H is not symmetric)
            assert abs(y @ (H @ y) + tk[i]) * 2 * 109737.315709 < 1e-3


def _run_fail(workdir, stdin_text=""):
    r = subprocess.run([PROG_PATH], cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300, input=stdin_text)
    print(r.stdout[-2000:])
    assert r.returncode != 0
    return r.stdout


def test_pconf_gpu_refusals(tmp_path):
    """pconf's refusals, as messages and a failed exit: Kv = 3 with kCSF = 1
    (src/pconf.F90:372-375) and Kl4 = 2 without the CONF.XIJ to start from."""
    print("[case] pconf_gpu refusals (t3j kCSF=1 Kv=3; t3 Kl4=2 without CONF.XIJ)", flush=True)
    w = tmp_path / "t3j"
    shutil.copytree(space_dir("t3j"), w)
    cio = F.read_ci_in(str(w / "ci.in"))
    F.write_ci_in(str(w / "ci.in"), Kl=0, MaxNd0=cio["MAXND0"], kCSF=1)
    out = _run_fail(str(w))
    assert "Kv=3 (J-projection) is not compatible with kCSF>0" in out
    w = tmp_path / "t3"
    shutil.copytree(space_dir("t3"), w)
    inp = F.read_conf_inp(str(w / "CONF.INP"))
    inp.Kl4 = 2
    F.write_conf_inp(str(w / "CONF.INP"), inp)
    out = _run_fail(str(w))
    assert "Kl4 = 2: CONF.XIJ is absent" in out
