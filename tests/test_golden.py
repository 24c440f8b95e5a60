"""Golden fixtures written by the reference's own compiled Fortran modules
(tests/golden/make_golden.py): determinant list, FormH rows (FP64 bits) and
dense eigenvalues.  The CPU oracle and the device path are both checked
against them; neither needs the reference at run time."""
import glob
import os

import numpy as np
import pytest

from tests.golden.make_golden import input_digest
from tests.helpers import oracle_for, space_dir

HERE = os.path.dirname(os.path.abspath(__file__))
# scale_*.npz: the CPU oracle's rows of the measured workloads (test_scale_fixtures.py)
FIXTURES = sorted(f for f in glob.glob(os.path.join(HERE, "golden", "*.npz"))
                  if not os.path.basename(f).startswith("scale_"))
NAMES = [os.path.basename(f)[:-4] for f in FIXTURES]


def _load(name):
    z = np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert {"tiny2", "tiny3", "tiny3d", "c1", "t3", "t3b", "t4"} <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference_fixture(name):
    g = _load(name)
    d = space_dir(name)
    assert input_digest(d) == str(g["digest"]), "generated inputs differ from the ones the fixture was made from"
    o = oracle_for(name)
    np.testing.assert_array_equal(o.iarr, g["iarr"])
    np.testing.assert_array_equal(o.ndc, g["ndc"])
    rp, col, val, _, _ = o.formh(1, int(g["last_row"]))
    np.testing.assert_array_equal(rp, g["rowptr"])
    np.testing.assert_array_equal(col, g["col"])
    np.testing.assert_array_equal(val.view(np.int64), g["val_bits"])


@pytest.mark.parametrize("name", [n for n in NAMES if n.startswith("tiny")])
def test_reference_eigenvalues_against_brute_force(name):
    """The reference's H on the tiny spaces against an independent second-
    quantisation build (oracle/brute.py, Coulomb), through the spectrum."""
    from oracle.brute import Brute
    from tests.helpers import inputs
    g = _load(name)
    o = oracle_for(name)
    ci, *_ = inputs(space_dir(name))
    if ci.kbrt:
        pytest.skip("Breit space")
    b = Brute(o.shells, o.Nso, ci.rint1, ci.iint1, ci.rint2, ci.iint2, ci.iint3, ci.nx, o.jz, o.nh)
    Hb = b.matrix(g["iarr"], o.Npos)
    np.testing.assert_allclose(np.linalg.eigvalsh(Hb), g["eigvals"], atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_device_matches_reference_fixture(name):
    from tests.helpers import gpu_context_like
    print(f"[case] golden {name}", flush=True)
    g = _load(name)
    o = oracle_for(name)
    ctx = gpu_context_like(o, space_dir(name))
    ctx.build_h()
    rp, col, val = ctx.get_csr()
    last = int(g["last_row"])
    n = int(g["rowptr"][-1])
    np.testing.assert_array_equal(rp[:last + 1], g["rowptr"])
    np.testing.assert_array_equal(col[:n], g["col"])
    np.testing.assert_array_equal(val[:n].view(np.int64), g["val_bits"])
