"""The CPU Davidson restatement (oracle/davidson.py) against Lanczos on the
same H; pins the oracle the device Davidson is compared with."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from oracle.davidson import diag4
from oracle.oracle import calc_nd0
from tests.helpers import full_from_lower, oracle_for

HARTREE_CM = 2 * 109737.315709


@pytest.mark.parametrize("name", ["c1", "t3", "t4"])
def test_diag4_eigenvalues(name):
    o = oracle_for(name)
    rp, col, val, _, _ = o.formh()
    _, nd0 = calc_nd0(o.ndc, o.inp.Nc4, o.ci_in["MAXND0"])
    tk, vecs, it, trace = diag4(rp, col, val, val.min(), nd0, o.inp.Nlv, o.inp.Crt4, o.inp.N_it)
    H = full_from_lower(rp, col, val)
    w = np.sort(sla.eigsh(H, k=o.inp.Nlv, which="SA", tol=1e-14)[0])
    assert np.abs(-tk - w).max() * HARTREE_CM < 1e-3
    for i in range(o.inp.Nlv):
        r = H @ vecs[i] + tk[i] * vecs[i]
        assert np.linalg.norm(r) < 1e-5


def test_mxmpy_is_symmetric_scatter():
    from oracle.oracle import mxmpy
    o = oracle_for("t3")
    rp, col, val, _, _ = o.formh()
    hmin = val.min()
    x = np.random.default_rng(3).standard_normal(o.Nd)
    y, diag = mxmpy(rp, col, val, hmin, x, want_diag=True)
    H = full_from_lower(rp, col, val, hmin)
    np.testing.assert_allclose(y, H @ x, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(diag, H.diagonal(), rtol=0, atol=1e-15)


def test_rank_mxmpy_sums_to_the_whole_product():
    """Mxmpy split over ranks' row ranges (each scattering into a full y, as
    before pconf's Allreduce) sums to the single-rank product."""
    from oracle.oracle import mxmpy, mxmpy_rows
    o = oracle_for("t3")
    rp, col, val, _, _ = o.formh()
    hmin = val.min()
    x = np.random.default_rng(4).standard_normal(o.Nd)
    y = mxmpy(rp, col, val, hmin, x)
    edges = [1, 700, 1900, 2600, o.Nd + 1]
    ys = np.zeros(o.Nd)
    for a, b in zip(edges[:-1], edges[1:]):
        lrp = rp[a - 1:b] - rp[a - 1]
        ys += mxmpy_rows(o.Nd, a, lrp, col[rp[a - 1]:rp[b - 1]], val[rp[a - 1]:rp[b - 1]], hmin, x)
    np.testing.assert_allclose(ys, y, rtol=1e-12, atol=1e-12)


def test_oracle_ranks_match_the_single_process_run():
    """The oracle run as pconf's MPI ranks (oracle/ranks.py: interleaved row
    chunks, each rank's Mxmpy summed like the Allreduce) builds the same H
    (NumH, Hmin, numzero) and Diag4 on it finds the single-process roots in
    the same number of iterations."""
    from oracle.ranks import OracleRanks, rank_chunks
    from tests.helpers import space_dir
    o = oracle_for("t3")
    rp, col, val, nz, _ = o.formh()
    rp1, col1, val1, nz1, _ = o.formh_once()
    np.testing.assert_array_equal(rp1, rp)
    np.testing.assert_array_equal(col1, col)
    np.testing.assert_array_equal(val1.view(np.int64), val.view(np.int64))
    assert nz1 == nz
    rows = sorted(r for k in range(3) for a, b in rank_chunks(o.Nd, 3, k) for r in range(a, b + 1))
    assert rows == list(range(1, o.Nd + 1))
    _, nd0 = calc_nd0(o.ndc, o.inp.Nc4, o.ci_in["MAXND0"])
    tk, _, it, _ = diag4(rp, col, val, val.min(), nd0, o.inp.Nlv, o.inp.Crt4, o.inp.N_it)
    with OracleRanks(space_dir("t3"), o.Nd, 3, o.inp.Nlv) as R:
        _, numh, hmin, nzr = R.build()
        assert (numh, hmin, nzr) == (rp[-1], val.min(), nz)
        x = np.random.default_rng(5).standard_normal((o.Nd, 2))
        np.testing.assert_allclose(R.mx(x), full_from_lower(rp, col, val, hmin) @ x, rtol=1e-12, atol=1e-12)
        tk2, _, it2, _ = diag4(None, None, None, R.hmin, nd0, o.inp.Nlv, o.inp.Crt4, o.inp.N_it, ops=R)
    assert it2 == it
    assert np.abs(tk2 - tk).max() * HARTREE_CM < 1e-6


def test_diag4_restart_from_converged_vectors():
    """FormB0 with |Kl4| = 2 (start vectors and energies given; Dvdsn's Z1
    projector over every Z1 column): from converged vectors the restarted
    Diag4 stops within two iterations on the same roots; from slightly
    perturbed vectors it reaches the same roots.  (From vectors perturbed by
    1e-3 the reference's rule stalls: the leading Z1 columns, whose
    eigenvalues are close to theta, amplify the residual's head part.)"""
    o = oracle_for("t3")
    rp, col, val, _, _ = o.formh()
    _, nd0 = calc_nd0(o.ndc, o.inp.Nc4, o.ci_in["MAXND0"])
    tk, vecs, it, _ = diag4(rp, col, val, val.min(), nd0, o.inp.Nlv, o.inp.Crt4, o.inp.N_it)
    tk2, _, it2, _ = diag4(rp, col, val, val.min(), nd0, o.inp.Nlv, o.inp.Crt4, o.inp.N_it, x0=vecs, tk0=tk)
    assert it2 <= 2 and np.abs(tk2 - tk).max() * HARTREE_CM < 1e-6
    x0 = vecs + 1e-5 * np.random.default_rng(2).standard_normal(vecs.shape)
    tk3, v3, it3, _ = diag4(rp, col, val, val.min(), nd0, o.inp.Nlv, o.inp.Crt4, o.inp.N_it, x0=x0, tk0=tk)
    assert np.abs(tk3 - tk).max() * HARTREE_CM < 1e-3
    H = full_from_lower(rp, col, val)
    for i in range(o.inp.Nlv):
        assert np.linalg.norm(H @ v3[i] + tk3[i] * v3[i]) < 1e-5
