"""The CPU oracle against the committed fixtures of the measured workloads
(tests/golden/make_scale_golden.py): FormH rows spread over the whole C2, C4
and C5 determinant lists, FP64 bits, and the C2 eigenvalues from the oracle's
full FormH + Lanczos.  The device is checked against the same fixtures in
test_gpu_scale.py."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import input_digest
from tests.helpers import oracle_for, space_dir

HERE = os.path.dirname(os.path.abspath(__file__))


def load_scale(name):
    z = np.load(os.path.join(HERE, "golden", f"scale_{name}.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def fixture_rows(g):
    """{row: (1-based columns, FP64 bits)} of a scale fixture."""
    out, k = {}, 0
    for r, n in zip(g["rows"], g["counts"]):
        out[int(r)] = (g["col"][k:k + n], g["val_bits"][k:k + n])
        k += n
    return out


@pytest.mark.parametrize("name", ["c2", "c4", "c5"])
def test_oracle_rows_match_scale_fixture(name):
    g = load_scale(name)
    d = space_dir(name)
    assert input_digest(d) == str(g["digest"]), "generated inputs differ from the ones the fixture was made from"
    o = oracle_for(name)
    assert o.Nd == int(g["nd"])
    rows = fixture_rows(g)
    assert min(rows) == 1 and max(rows) == o.Nd and len(rows) >= 60
    for r, (col, bits) in rows.items():
        rp, c, v, _, _ = o.formh(r, r)
        np.testing.assert_array_equal(c, col)
        np.testing.assert_array_equal(v.view(np.int64), bits)


def test_c2_lanczos_eigenvalues_are_recorded():
    g = load_scale("c2")
    w = g["eigvals"]
    assert w.shape == (10,) and np.all(np.diff(w) > 0)
    assert int(g["numh"]) > 2e8 and w[-1] - w[0] < 1.0


def csf_fixture_rows(g):
    """{configuration: (row counts, 1-based columns, FP64 bits)} of a CSF scale fixture."""
    ndcs, nc_neq = g["ndcs"], g["nc_neq"]
    out, r, k = {}, 0, 0
    for ic in g["confs"]:
        n = int(ndcs[nc_neq[ic - 1] - 1])
        cnt = g["row_counts"][r:r + n]
        m = int(cnt.sum())
        out[int(ic)] = (cnt, g["col"][k:k + m], g["val_bits"][k:k + m])
        r += n
        k += m
    return out


@pytest.mark.parametrize("name", ["c2", "c4"])
def test_oracle_csf_rows_match_scale_fixture(name):
    """kCSF = 1 at the measured sizes: the oracle's classes and CSF counts, and
    formh_sym rows of a few sampled configurations on the fixture's basis."""
    from oracle import csf
    from oracle.oracle import Oracle
    g = load_scale(f"csf_{name}")
    d = space_dir(name)
    assert input_digest(d) == str(g["digest"])
    o = Oracle.from_dir(d, kcsf=1)
    np.testing.assert_array_equal(o.nc_neq, g["nc_neq"])
    b = csf.jbasis(o)
    np.testing.assert_array_equal(b.ndcs, g["ndcs"])
    bf = csf.CsfBasis(o, g["ndcs"], g["ccj"])
    rows = csf_fixture_rows(g)
    for ic in list(rows)[:3] + list(rows)[-1:]:
        cnt, col, bits = rows[ic]
        rp, c, v, _, _, _ = csf.formh_sym(o, bf, ic, ic)
        np.testing.assert_array_equal(np.diff(rp), cnt)
        np.testing.assert_array_equal(c, col)
        np.testing.assert_array_equal(v.view(np.int64), bits)
