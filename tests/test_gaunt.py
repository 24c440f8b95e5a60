"""Angular factors: closed forms and the reference's table symmetries."""
import math

import numpy as np
import pytest

from pci_amd import gen
from tests.helpers import oracle_for


def test_threej_known_values():
    # (1/2 1/2 1; 1/2 -1/2 0) = 1/sqrt(6);  (1 1 0; 0 0 0) = -1/sqrt(3)
    s, v = gen.threej_sq(1, 1, 2, 1, -1, 0)
    assert abs(s * math.sqrt(v) - 1 / math.sqrt(6)) < 1e-15
    s, v = gen.threej_sq(2, 2, 0, 0, 0, 0)
    assert abs(s * math.sqrt(v) + 1 / math.sqrt(3)) < 1e-15


@pytest.mark.parametrize("j1,j2", [(1, 3), (3, 5), (5, 7), (7, 7)])
def test_threej_orthogonality(j1, j2):
    # sum_{m1,m2} (2k+1) (j1 j2 k; m1 m2 -M)^2 = 1 for every allowed k
    for k2 in range(abs(j1 - j2), j1 + j2 + 1, 2):
        tot = 0.0
        for m1 in range(-j1, j1 + 1, 2):
            for m2 in range(-j2, j2 + 1, 2):
                M = m1 + m2
                if abs(M) > k2:
                    continue
                s, v = gen.threej_sq(j1, j2, k2, m1, m2, -M)
                tot += float(v)
        # sum over m1, m2 of 3j^2 for fixed k equals 1 (summing all M)
        assert abs(tot - 1.0) < 1e-13


def test_gaunt_lut_symmetries_match_closed_form():
    """pconf's Gaunt (table for j1 <= j2, m1 > 0 plus swap/negation symmetries,
    integrals.f90:509-561) equals the closed form for every argument Gint can
    pass: k inside the triangle of (j1, j2) (Gint's kmin..kmax loop), j up to
    Nlx - 1/2 (see test_gaunt_lut_alias_at_top_j for j = Nlx + 1/2)."""
    o = oracle_for("c1")
    n = 0
    for j1 in range(1, 10, 2):
        for j2 in range(1, 10, 2):
            for k in range(abs(j1 - j2) // 2, (j1 + j2) // 2 + 1):
                for m1 in range(-j1, j1 + 1, 2):
                    for m2 in range(-j2, j2 + 1, 2):
                        if k == 0:
                            want = 1.0 if (j1 == j2 and m1 == m2) else 0.0
                        else:
                            want = gen.gaun(k, j1, m1, j2, m2)
                        got = o.gaunt(k, j1 / 2, m1 / 2, j2 / 2, m2 / 2)
                        assert abs(got - want) < 1e-14, (k, j1, m1, j2, m2, got, want)
                        n += 1
    assert n > 5000


def test_gaunt_lut_alias_at_top_j():
    """Reference quirk reproduced, not fixed: the LUT radix is 2*Nlx+1 while
    j2+m2 runs up to 2*j2 = 2*Nlx+1 for j2 = Nlx+1/2, so (j1+m1, j2+m2) and
    (j1+m1+1, 0) share one index and BuildGauntLUT keeps the entry written
    last (integrals.f90:544-546, 702-715; breit.f90:79-92)."""
    o = oracle_for("c1")
    nlx, counts, idx, val = gen.gaunt_table(5)
    # k=7, j1=3/2, m1=1/2, j2=11/2, m2=11/2 aliases k=7, j1=3/2, m1=3/2, j2=11/2, m2=-11/2
    a = gen.gaun(7, 3, 1, 11, 11)
    b = gen.gaun(7, 3, 3, 11, -11)
    assert a != b
    got = o.gaunt(7, 1.5, 0.5, 5.5, 5.5)
    assert got == b            # the later table entry wins
    assert abs(o.gaunt(7, 1.5, 1.5, 5.5, -5.5) - b) < 1e-15
