"""Device Davidson with J selection (Kv = 3: FormB0 J filter, AvgDiag,
Prj_J) against the Davidson restatement with the same selection and a
dense solve restricted to the target J."""
import numpy as np
import pytest

from tests.helpers import full_from_lower, oracle_for, space_dir

pytestmark = pytest.mark.gpu
HARTREE_CM = 2 * 109737.315709


@pytest.fixture(scope="module", params=["t3j", "c1j"])
def jcase(request):
    from oracle.davidson import diag4
    from oracle.oracle import calc_nd0
    from pci_amd import host
    name = request.param
    print(f"[case] jsel space={name}", flush=True)
    sp, ctx = host.load(space_dir(name))
    ctx.build_h()
    ctx.build_j()
    o = oracle_for(name)
    rp, col, val, _, _ = o.formh()
    jrp, jcol, jval = o.formj()
    jt, _ = o.jterms()
    _, nd0 = calc_nd0(o.ndc, o.inp.Nc4, o.ci_in["MAXND0"])
    js = dict(xj=o.inp.J, jt=jt, jrp=jrp, jcol=jcol, jval=jval, ndc=o.ndc)
    tk_o, v_o, it_o, _ = diag4(rp, col, val, val.min(), nd0, sp.Nlv, sp.Crt4, sp.N_it, jsel=js)
    return name, sp, ctx, o, tk_o


@pytest.mark.parametrize("fp32", [False, True])
def test_jsel_davidson(jcase, fp32):
    from pci_amd._lib import PCI_DVD_FP32, PCI_DVD_JSEL
    name, sp, ctx, o, tk_o = jcase
    flags = PCI_DVD_JSEL | (PCI_DVD_FP32 if fp32 else 0)
    tk, tj, vecs, st = ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, sp.N_it, flags)
    print({k: v for k, v in st.items() if k != "resid_hist"}, tj)
    assert st["converged"] in (1, 2), st
    assert np.abs(tk - tk_o).max() * HARTREE_CM < (1e-3 if not fp32 else 0.01)
    np.testing.assert_allclose(tj, o.inp.J, atol=1e-4)


def test_jsel_needs_terms():
    """Without J^2 and J terms the flag is refused with an error, not ignored."""
    from pci_amd._lib import GpuContext, PCI_DVD_JSEL, PciError
    from tests.helpers import gpu_context_like
    o = oracle_for("t3")
    g = gpu_context_like(o, space_dir("t3"))
    g.build_h()
    with pytest.raises(PciError):
        g.davidson(4, 400, 1e-8, 10, PCI_DVD_JSEL)
