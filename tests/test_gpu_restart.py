"""pconf's restart and periodic output around Diag4 (pci_gpu_davidson_ex):
the KXIJ checkpoint (FormB writes the current vectors every KXIJ
iterations, src/pconf.F90:2186-2211) and the |Kl4| = 2 start from the
vectors of an earlier run (FormB0, src/davidson.f90:136-154), through the
C ABI and through the pconf_gpu program."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from pci_amd import formats as F
from pci_amd.host import PROG_PATH
from tests.helpers import space_dir

pytestmark = pytest.mark.gpu

HARTREE_CM = 2 * 109737.315709


@pytest.fixture(scope="module")
def t4():
    from pci_amd import host
    print("[case] restart space=t4", flush=True)
    sp, ctx = host.load(space_dir("t4"))
    ctx.build_h()
    ctx.build_j()
    yield sp, ctx
    ctx.close()


def test_checkpoints_do_not_change_the_solve(t4):
    from pci_amd import _lib
    sp, ctx = t4
    seen = []
    tk, tj, vecs, st = ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, 300, _lib.PCI_DVD_NOADMIX)
    tk2, tj2, vecs2, st2 = ctx.davidson_ex(sp.Nlv, sp.Nd0, sp.Crt4, 300, _lib.PCI_DVD_NOADMIX, every=2,
                                           checkpoint=lambda it, a, b, v: seen.append((it, a, b, v)))
    np.testing.assert_array_equal(tk2, tk)
    np.testing.assert_array_equal(vecs2, vecs)
    its = [x[0] for x in seen]
    assert its == [0] + list(range(2, st["iterations"], 2)), its
    for it, a, b, v in seen:
        assert v.shape == (sp.Nlv, sp.Nd) and a.shape == (sp.Nlv,)
        # J_av of the checkpoint vectors, as FormB's caller computes it
        assert np.all(np.isfinite(b))
    # the later checkpoints approach the final roots
    assert np.abs(seen[-1][1] - tk).max() <= np.abs(seen[1][1] - tk).max()


@pytest.mark.parametrize("flags", [0, 1])
def test_restart_from_converged_vectors(t4, flags):
    from pci_amd import _lib
    sp, ctx = t4
    fl = _lib.PCI_DVD_NOADMIX | (_lib.PCI_DVD_FP32 if flags else 0)
    tk, tj, vecs, st = ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, 300, fl)
    tk2, _, vecs2, st2 = ctx.davidson_ex(sp.Nlv, sp.Nd0, sp.Crt4, 300, fl, tk0=tk, x0=vecs)
    assert st2["converged"] == 1 and st2["iterations"] <= 2, st2
    assert np.abs(tk2 - tk).max() * HARTREE_CM < 1e-3
    # from perturbed vectors (a restart after an interrupted run): the same
    # roots.  Dvdsn's projector then includes the leading Z1 columns (the
    # reference's rule for |Kl4| = 2), which can take more iterations than a
    # start from the Nd0 block does
    rng = np.random.default_rng(5)
    x0 = vecs + 1e-6 * rng.standard_normal(vecs.shape)  # ~1e-4 in norm
    tk3, _, _, st3 = ctx.davidson_ex(sp.Nlv, sp.Nd0, sp.Crt4, 300, fl, tk0=tk, x0=x0)
    print(f"restart from perturbed vectors: {st3['iterations']} it (fresh start {st['iterations']})", flush=True)
    assert st3["converged"] == 1, st3
    assert np.abs(tk3 - tk).max() * HARTREE_CM < 1e-3


def test_restart_trajectory_matches_the_oracle(t4):
    """The |Kl4| = 2 solve against the oracle's Diag4 from the same start
    vectors and energies: Dvdsn's Z1 projector then takes every Z1 column
    (n01 = 1, src/davidson.f90:518-519), not only those past Nlv; device and
    oracle agree iteration by iteration (max residual per iteration while it
    is far above rounding), in the iteration count and in the roots."""
    from oracle.davidson import diag4
    from oracle.oracle import Oracle
    from pci_amd import _lib
    sp, ctx = t4
    o = Oracle.from_dir(space_dir("t4"))
    rp, col, val, _, _ = o.formh()
    tk, _, vecs, _ = ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, 300, _lib.PCI_DVD_NOADMIX)
    rng = np.random.default_rng(11)
    x0 = vecs + 1e-5 * rng.standard_normal(vecs.shape)
    tk_d, _, _, st = ctx.davidson_ex(sp.Nlv, sp.Nd0, sp.Crt4, sp.N_it, _lib.PCI_DVD_TRACE, tk0=tk, x0=x0)
    tk_o, _, it_o, tr_o = diag4(rp, col, val, val.min(), sp.Nd0, sp.Nlv, sp.Crt4, sp.N_it, x0=x0, tk0=tk)
    n = min(st["iterations"], it_o)
    hist, tr = np.array(st["resid_hist"][:n]), np.array(tr_o[:n])
    print(f"restart: device {st['iterations']} it, oracle {it_o} it; histories {hist} {tr}", flush=True)
    assert abs(st["iterations"] - it_o) <= 2
    m = min(n, 10)
    np.testing.assert_allclose(hist[:m], tr[:m], rtol=1e-2)
    assert np.abs(tk_d - tk_o).max() * HARTREE_CM < 1e-3


def _run(workdir):
    r = subprocess.run([PROG_PATH], cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300, input="")
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


def test_pconf_gpu_kxij_and_kl4_restart(tmp_path):
    """KXIJ = 2 writes CONF.XIJ while iterating; a second run with Kl4 = 2 and
    Kl = 1 (pCONF.HIJ of the first) starts from the first run's CONF.XIJ and
    converges in a few iterations to the same energies."""
    print("[case] pconf_gpu KXIJ / Kl4=2 space=t3", flush=True)
    w = tmp_path / "t3"
    shutil.copytree(space_dir("t3"), w)
    cio = F.read_ci_in(str(w / "ci.in"))
    F.write_ci_in(str(w / "ci.in"), Kl=0, Kw=1, KXIJ=2, MaxNd0=cio["MAXND0"])
    out = _run(str(w))
    assert "FormB: Vectors saved at iteration 2" in out
    it1 = int(out.split("Davidson: ")[1].split()[0])
    # CONF.ENG: PrintEnergiesDvdsn's table at every checkpoint
    eng = open(w / "CONF.ENG").read()
    assert eng.count("Davidson Iteration #") == (it1 - 1) // 2 and "Energy levels (" in eng
    tk1, tj1, v1 = F.read_xij(str(w / "CONF.XIJ"))
    inp = F.read_conf_inp(str(w / "CONF.INP"))
    inp.Kl4 = 2
    F.write_conf_inp(str(w / "CONF.INP"), inp)
    F.write_ci_in(str(w / "ci.in"), Kl=1, KXIJ=2, MaxNd0=cio["MAXND0"])
    out = _run(str(w))
    assert "Vector length = " in out and "H read from pCONF.HIJ" in out
    it2 = int(out.split("Davidson: ")[1].split()[0])
    assert it2 < it1 and it2 <= 3, (it1, it2)
    tk2, tj2, v2 = F.read_xij(str(w / "CONF.XIJ"))
    assert np.abs(tk2 - tk1).max() * HARTREE_CM < 1e-3
