"""Legacy CONF.HIJ / CONF.JJJ triple files (the layout src/sort.f90:64-70
writes and src/pol.f90:766-798 / src/ine.f90 read), written by the Fortran
host from a lower-triangle CSR."""
import struct

import numpy as np

from pci_amd import formats as F
from pci_amd import host
from tests.helpers import oracle_for


def test_triples_known_bytes(tmp_path):
    # 3 x 3 lower triangle: rows 1: (1,1); 2: (2,1) (2,2); 3: (3,2) (3,3),
    # with an exact zero at (2,1) that the sort program leaves out
    rowptr = np.array([0, 1, 3, 5], np.int64)
    col = np.array([1, 1, 2, 2, 3], np.int32)
    val = np.array([-1.5, 0.0, 0.25, 1e-300, 3.0])
    p = str(tmp_path / "CONF.HIJ")
    host.write_conf_hij(p, rowptr, col, val)
    want = struct.pack("<q", 4)
    for n, k, t in [(1, 1, -1.5), (2, 2, 0.25), (3, 2, 1e-300), (3, 3, 3.0)]:
        want += struct.pack("<iid", n, k, t)
    assert open(p, "rb").read() == want


def test_triples_of_formh(tmp_path):
    o = oracle_for("t3")
    rp, col, val, _, _ = o.formh()
    val = val.copy()
    off = np.nonzero(np.diff(rp) > 1)[0][0]     # a row with an off-diagonal element
    val[rp[off]] = 0.0                           # planted exact zero
    p = str(tmp_path / "CONF.HIJ")
    host.write_conf_hij(p, rp, col, val)
    n, k, t = F.read_conf_hij(p)
    rows = np.repeat(np.arange(1, rp.size, dtype=np.int32), np.diff(rp))
    keep = val != 0.0
    assert n.size == int(keep.sum()) == rp[-1] - 1
    np.testing.assert_array_equal(n, rows[keep])
    np.testing.assert_array_equal(k, col[keep])
    np.testing.assert_array_equal(t.view(np.int64), val[keep].view(np.int64))
    # ordering and range as pol expects: by row, then column, lower triangle
    key = n.astype(np.int64) * (o.Nd + 1) + k
    assert np.all(np.diff(key) > 0)
    assert np.all((k >= 1) & (k <= n) & (n <= o.Nd))
    assert 8 + 16 * n.size == len(open(p, "rb").read())
    # pol's reader stops at column Nd + 1, which the full space never has
    assert F.read_conf_hij(p, nd=o.Nd)[0].size == n.size
