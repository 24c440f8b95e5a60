"""File-contract round trips (pconf formats)."""
import numpy as np
import pytest

from pci_amd import formats as F
from pci_amd import gen
from tests.helpers import space_dir


@pytest.mark.parametrize("n,l,kappa,q", [(4, 0, -1, 2), (12, 1, -2, 1), (4, 3, 3, 6), (9, 2, 2, 4), (14, 3, -4, 1)])
def test_shell_code_roundtrip(n, l, kappa, q):
    x = F.encode_shell(n, l, kappa, q)
    assert F.decode_shell(x) == (n, l, kappa, q)
    assert F.decode_shell(float(f"{x:7.4f}")) == (n, l, kappa, q)


def test_shell_code_examples():
    # docs/pci/pconf.rst: 1.2101 is 12p3/2^1, -0.4306 is 4f5/2^6
    assert F.decode_shell(1.2101) == (12, 1, -2, 1)
    assert F.decode_shell(-0.4306) == (4, 3, 3, 6)


def test_space_files_roundtrip(tmp_path):
    spec = gen.PRESETS["t3"]
    info = gen.write_space(spec, str(tmp_path))
    inp = F.read_conf_inp(str(tmp_path / "CONF.INP"))
    shells, configs = gen.build_space(spec)
    assert inp.Nc == len(configs) == info["Nc"]
    assert inp.Ne == spec.Ne and inp.Nlv == spec.Nlv and abs(inp.Jm - spec.Jm) < 1e-12
    for cfg, codes in zip(configs, inp.configs):
        assert [(shells[s][0], shells[s][1], shells[s][2], q) for s, q in cfg] == [F.decode_shell(x) for x in codes]
    Z, sh = F.read_conf_dat(str(tmp_path / "CONF.DAT"))
    assert Z == spec.Z and len(sh) == info["Ns"]
    ci = F.read_conf_int(str(tmp_path / "CONF.INT"))
    assert ci.iint2.size == info["Ngint"] and ci.rint1.size == info["Nhint"]
    nlx, counts, idx, gnt = F.read_conf_gnt(str(tmp_path / "CONF.GNT"))
    assert counts.sum() == idx.size == info["Ngaunt"]


def test_hij_xij_det_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    nd = 50
    rp = np.zeros(nd + 1, np.int64)
    rp[1:] = np.cumsum(rng.integers(1, 5, nd))
    col = rng.integers(1, nd, rp[-1]).astype(np.int32)
    val = rng.standard_normal(rp[-1])
    F.write_hij(str(tmp_path / "pCONF.HIJ"), rp, col, val)
    raw = (tmp_path / "pCONF.HIJ").read_bytes()
    assert len(raw) == 16 + 8 * (nd + 1) + 12 * rp[-1]
    r2, c2, v2 = F.read_hij(str(tmp_path / "pCONF.HIJ"))
    np.testing.assert_array_equal(r2, rp)
    np.testing.assert_array_equal(c2, col)
    np.testing.assert_array_equal(v2, val)
    vecs = rng.standard_normal((3, nd))
    F.write_xij(str(tmp_path / "CONF.XIJ"), [1.0, 2.0, 3.0], [0.5, 1.5, 2.5], vecs)
    tk, tj, v = F.read_xij(str(tmp_path / "CONF.XIJ"))
    np.testing.assert_array_equal(v, vecs)
    assert list(tk) == [1.0, 2.0, 3.0] and list(tj) == [0.5, 1.5, 2.5]
    # record layout: 4-byte marker, 8+8+4+8*nd payload, marker
    raw = (tmp_path / "CONF.XIJ").read_bytes()
    assert len(raw) == 3 * (8 + 20 + 8 * nd)
    iarr = rng.integers(1, 100, (nd, 3)).astype(np.int32)
    F.write_conf_det(str(tmp_path / "CONF.DET"), iarr, 7)
    a2, nsu = F.read_conf_det(str(tmp_path / "CONF.DET"))
    np.testing.assert_array_equal(a2, iarr)
    assert nsu == 7
