"""Brute-force second-quantisation Hamiltonian for tiny spaces (TEST INFRASTRUCTURE ONLY).

Independent of pconf's Rspq/Gint/Gaunt code paths: the matrix element
<Phi_n| H |Phi_k> is evaluated by applying

    H = sum_pq h_pq a+_p a_q + 1/2 sum_pqrs <pq|rs> a+_p a+_q a_s a_r

to |Phi_k> = a+_{k1} a+_{k2} ... |0> (positions ascending) with explicit
fermion signs.  The spin-orbital integrals are built from their definitions:

    h_pq    = h(shell p, shell q)          if m_p = m_q and kappa_p = kappa_q
    <pq|rs> = sum_k G^k(p, r) G^k(s, q) R^k(shell p, q, r, s)

with the angular factor G^k(j1 m1, j2 m2) evaluated directly from Racah's
closed form for every argument (no table, no symmetry relations), the
Coulomb rule k + l_p + l_r even and the triangle rule, and R^k looked up
through all 8 images of the (pq|rs) symmetry.  Coulomb integrals only (the
Breit storage convention is pinned by the compiled-reference comparison).
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np


@lru_cache(maxsize=None)
def gaunt_direct(k, j1x2, m1x2, j2x2, m2x2):
    """G^k(j1 m1, j2 m2) for any arguments (doubled integers)."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from pci_amd.gen import gaun
    return gaun(k, j1x2, m1x2, j2x2, m2x2)


class Brute:
    def __init__(self, shells, nso, rint1, iint1, rint2, iint2, iint3, nx, jz, nh):
        """shells: list of (n, l, kappa, 2j) per shell; jz, nh per position (0-based arrays)."""
        self.shells = [tuple(int(v) for v in s) for s in shells]
        self.jz = np.asarray(jz)
        self.nh = np.asarray(nh)
        ns = len(shells)
        self.h = np.zeros((ns + 1, ns + 1))
        for v, ind in zip(rint1, iint1):
            a = (int(ind) - 1) // nx + nso + 1
            b = (int(ind) - 1) % nx + nso + 1
            self.h[a, b] = v
            self.h[b, a] = v
        self.R = {}
        r2 = np.asarray(rint2).reshape(len(iint2), -1)[:, 0]
        for i in range(len(iint2)):
            iac, ibd = int(iint2[i]) - 1, int(iint3[i]) - 1
            k = iac // (nx * nx)
            a = (iac // nx) % nx + nso + 1
            c = iac % nx + nso + 1
            b = ibd // nx + nso + 1
            d = ibd % nx + nso + 1
            for img in ((a, b, c, d), (c, b, a, d), (a, d, c, b), (c, d, a, b),
                        (b, a, d, c), (d, a, b, c), (b, c, d, a), (d, c, b, a)):
                self.R.setdefault((k,) + img, float(r2[i]))

    def one(self, p, q):
        if self.jz[p - 1] != self.jz[q - 1]:
            return 0.0
        sp, sq = int(self.nh[p - 1]), int(self.nh[q - 1])
        if self.shells[sp - 1][2] != self.shells[sq - 1][2]:
            return 0.0
        return self.h[sp, sq]

    def two(self, p, q, r, s):
        """<pq|rs>: electron 1 goes p <- r, electron 2 goes q <- s (positions 1-based)."""
        mp, mq, mr, ms = (int(self.jz[x - 1]) for x in (p, q, r, s))
        if mp + mq != mr + ms:
            return 0.0
        sp, sq, sr, ss = (int(self.nh[x - 1]) for x in (p, q, r, s))
        (_, lp, _, jp), (_, lq, _, jq), (_, lr, _, jr), (_, ls, _, js) = (self.shells[x - 1] for x in (sp, sq, sr, ss))
        if (lp + lq + lr + ls) % 2:
            return 0.0
        e = 0.0
        for k in range(max(abs(jp - jr), abs(jq - js)) // 2, min(jp + jr, jq + js) // 2 + 1):
            if (k + lp + lr) % 2:
                continue
            g1 = gaunt_direct(k, jp, mp, jr, mr)
            g2 = gaunt_direct(k, js, ms, jq, mq)
            if g1 == 0.0 or g2 == 0.0:
                continue
            e += g1 * g2 * self.R[(k, sp, sq, sr, ss)]
        return e

    @staticmethod
    def _annihilate(det, p):
        """a_p |det>: (sign, new det) or None."""
        if p not in det:
            return None
        i = det.index(p)
        return (-1) ** i, det[:i] + det[i + 1:]

    @staticmethod
    def _create(det, p):
        """a+_p |det>: (sign, new det) or None; det stays sorted."""
        if p in det:
            return None
        i = 0
        while i < len(det) and det[i] < p:
            i += 1
        return (-1) ** i, det[:i] + (p,) + det[i:]

    def apply(self, det, npos):
        """H|det> as a dict {det: coefficient}; positions 1..npos."""
        out = {}
        det = tuple(det)
        for q in det:
            s1, d1 = self._annihilate(det, q)
            for p in range(1, npos + 1):
                h = self.one(p, q)
                if h == 0.0:
                    continue
                c = self._create(d1, p)
                if c is None:
                    continue
                out[c[1]] = out.get(c[1], 0.0) + s1 * c[0] * h
        for r in det:
            ar = self._annihilate(det, r)
            for s in det:
                if s == r:
                    continue
                as_ = self._annihilate(ar[1], s)
                base = ar[0] * as_[0]
                d2 = as_[1]
                for q in range(1, npos + 1):
                    cq = self._create(d2, q)
                    if cq is None:
                        continue
                    for p in range(1, npos + 1):
                        v = self.two(p, q, r, s)
                        if v == 0.0:
                            continue
                        cp = self._create(cq[1], p)
                        if cp is None:
                            continue
                        out[cp[1]] = out.get(cp[1], 0.0) + 0.5 * base * cq[0] * cp[0] * v
        return out

    def matrix(self, iarr, npos):
        """Dense H over the determinant list iarr (rows are 1-based sorted positions)."""
        index = {tuple(int(x) for x in d): i for i, d in enumerate(iarr)}
        nd = len(iarr)
        H = np.zeros((nd, nd))
        for k, d in enumerate(iarr):
            for det, v in self.apply(tuple(int(x) for x in d), npos).items():
                n = index.get(det)
                if n is not None:
                    H[n, k] += v
        return H
