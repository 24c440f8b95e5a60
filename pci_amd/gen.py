"""Seeded synthetic CI spaces written in pconf's own input formats.

A space is a relativistic orbital basis, a list of relativistic configurations
(single/double/triple excitations of reference configurations) and seeded
FP64 integrals.  ``write_space`` produces the five files pconf reads:
CONF.INP, ci.in, CONF.DAT, CONF.INT and CONF.GNT.

Conventions followed:
* shell codes and configuration lines of CONF.INP (src/conf_init.f90:271-313);
* the canonical two-electron integral set and its index encoding
  iac = Nx^2 k + Nx (na-Nso-1) + (nc-Nso), ibd = Nx (nb-Nso-1) + (nd-Nso),
  enumerated with the loop structure of pbasc's Rint (src/pbasc.f90:1007-1123);
* one-electron integrals for shell pairs of equal kappa,
  index Nx (na-Nso-1) + (nb-Nso), na <= nb (src/integrals.f90:122-151);
* the Gaunt table of src/breit.f90:48-128:
  index = jx^2 (jx^2 k + 2 jx j1 + 2 j2) + jx (j1+m1) + (j2+m2), jx = 2 Nlx + 1,
  value = (-1)^(m1+1/2) sqrt((2j1+1)(2j2+1)) (j1 j2 k; -m1 m2 m1-m2)(j1 j2 k; 1/2 -1/2 0),
  for j1 <= j2, m1 > 0, 1 <= k, |m2-m1| <= k.  The 3j symbols come from the
  Racah closed form evaluated in exact rational arithmetic.
"""
from __future__ import annotations

import dataclasses
import math
import os
from fractions import Fraction
from functools import lru_cache
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import formats as F

Shell = Tuple[int, int, int, int]  # n, l, kappa, 2j


# ----------------------------------------------------------------------------
# Basis and configurations
# ----------------------------------------------------------------------------

def make_basis(nmin: int, nmax: int, lmax: int, nmin_l: Dict[int, int] = None,
               nmax_l: Dict[int, int] = None) -> List[Shell]:
    """Shells ordered by n, then l, then j (j = l-1/2 before l+1/2); nmin_l
    and nmax_l narrow the n range of single l values."""
    out = []
    for n in range(nmin, nmax + 1):
        for l in range(0, min(lmax, n - 1) + 1):
            if nmin_l and n < nmin_l.get(l, 0):
                continue
            if nmax_l and n > nmax_l.get(l, nmax):
                continue
            if l > 0:
                out.append((n, l, l, 2 * l - 1))
            out.append((n, l, -(l + 1), 2 * l + 1))
    return out


def _parity(cfg, shells):
    return sum(shells[s][1] * q for s, q in cfg) % 2


def excite(refs: Sequence[Sequence[Tuple[int, int]]], shells: Sequence[Shell], level: int,
           parity: int) -> List[Tuple[Tuple[int, int], ...]]:
    """All configurations within ``level`` electron moves of any reference.

    A configuration is a tuple of (shell index, occupation) sorted by shell
    index.  References come first, then new configurations in order of
    discovery (excitation level, then lexicographic order of the moves).
    """
    cap = [s[3] + 1 for s in shells]
    ns = len(shells)
    seen = {}
    order = []

    def canon(occ):
        return tuple((i, q) for i, q in enumerate(occ) if q)

    def add(occ):
        c = canon(occ)
        if c not in seen and _parity(c, shells) == parity:
            seen[c] = len(order)
            order.append(c)

    frontier = []
    for r in refs:
        occ = [0] * ns
        for s, q in r:
            occ[s] += q
        if any(occ[i] > cap[i] for i in range(ns)):
            raise ValueError("reference configuration over-fills a shell")
        add(occ)
        frontier.append(tuple(occ))
    visited = set(frontier)
    for _ in range(level):
        nxt = []
        for occ in frontier:
            occ = list(occ)
            for a in range(ns):
                if occ[a] == 0:
                    continue
                occ[a] -= 1
                for b in range(ns):
                    if occ[b] < cap[b] and b != a:
                        occ[b] += 1
                        t = tuple(occ)
                        if t not in visited:
                            visited.add(t)
                            nxt.append(t)
                            add(occ)
                        occ[b] -= 1
                occ[a] += 1
        frontier = nxt
    return order


# ----------------------------------------------------------------------------
# Wigner 3j and Gaunt table
# ----------------------------------------------------------------------------

@lru_cache(maxsize=None)
def _fact(n: int) -> int:
    return math.factorial(n)


def threej_sq(j1x2, j2x2, j3x2, m1x2, m2x2, m3x2) -> Tuple[int, Fraction]:
    """Return (sign, value^2) of the 3j symbol, exactly (Racah's formula).

    Arguments are twice the angular momenta/projections.
    """
    if m1x2 + m2x2 + m3x2 != 0:
        return 0, Fraction(0)
    for j, m in ((j1x2, m1x2), (j2x2, m2x2), (j3x2, m3x2)):
        if abs(m) > j or (j + m) % 2:
            return 0, Fraction(0)
    if (j1x2 + j2x2 + j3x2) % 2:
        return 0, Fraction(0)
    a, b, c = (j1x2 + j2x2 - j3x2), (j1x2 - j2x2 + j3x2), (-j1x2 + j2x2 + j3x2)
    if a < 0 or b < 0 or c < 0:
        return 0, Fraction(0)
    a, b, c = a // 2, b // 2, c // 2
    tri = Fraction(_fact(a) * _fact(b) * _fact(c), _fact((j1x2 + j2x2 + j3x2) // 2 + 1))
    pre = (_fact((j1x2 + m1x2) // 2) * _fact((j1x2 - m1x2) // 2) * _fact((j2x2 + m2x2) // 2)
           * _fact((j2x2 - m2x2) // 2) * _fact((j3x2 + m3x2) // 2) * _fact((j3x2 - m3x2) // 2))
    # summation limits
    t1 = (j3x2 - j2x2 + m1x2) // 2
    t2 = (j3x2 - j1x2 - m2x2) // 2
    t3 = (j1x2 + j2x2 - j3x2) // 2
    t4 = (j1x2 - m1x2) // 2
    t5 = (j2x2 + m2x2) // 2
    s = Fraction(0)
    for t in range(max(0, -t1, -t2), min(t3, t4, t5) + 1):
        den = (_fact(t) * _fact(t1 + t) * _fact(t2 + t) * _fact(t3 - t) * _fact(t4 - t) * _fact(t5 - t))
        s += Fraction((-1) ** t, den)
    if s == 0:
        return 0, Fraction(0)
    phase = (j1x2 - j2x2 - m3x2) // 2
    sign = (-1) ** (phase % 2) * (1 if s > 0 else -1)
    return sign, tri * pre * s * s


def _sqrt_fraction(x: Fraction) -> float:
    """Correctly rounded sqrt of a non-negative rational."""
    if x == 0:
        return 0.0
    num, den = x.numerator, x.denominator
    shift = 128
    r = math.isqrt((num << (2 * shift)) // den)
    return math.ldexp(float(r), -shift) if r < (1 << 1000) else math.sqrt(num / den)


def gaun(k: int, j1x2: int, m1x2: int, j2x2: int, m2x2: int) -> float:
    """Angular factor of src/breit.f90:108-128, evaluated exactly then rounded."""
    s1, v1 = threej_sq(j1x2, j2x2, 2 * k, -m1x2, m2x2, m1x2 - m2x2)
    s2, v2 = threej_sq(j1x2, j2x2, 2 * k, 1, -1, 0)
    if s1 == 0 or s2 == 0:
        return 0.0
    is_ = -1 if ((m1x2 + 1) // 2) % 2 else 1
    sq = Fraction((j1x2 + 1) * (j2x2 + 1)) * v1 * v2
    return float(is_ * s1 * s2) * _sqrt_fraction(sq)


def gaunt_entries(nlx: int):
    """(index, k, 2j1, 2m1, 2j2, 2m2) of the CONF.GNT table in file order
    (loops of src/breit.f90:48-106)."""
    nlx = max(nlx, 5)
    lx = nlx + 1
    jx = 2 * lx - 1
    out = []
    for ij1 in range(1, lx + 1):
        j1x2 = 2 * ij1 - 1
        for ij2 in range(ij1, lx + 1):
            j2x2 = 2 * ij2 - 1
            kmin = max(ij2 - ij1 + 1, 2)
            kmax = ij1 + ij2
            if kmax < kmin:
                continue
            for kk in range(kmin, kmax + 1):
                k = kk - 1
                ind1 = jx * jx * k + jx * j1x2 + j2x2
                for im1 in range(1, ij1 + 1):
                    m1x2 = 2 * im1 - 1
                    for im2 in range(1, 2 * ij2 + 1):
                        m2x2 = 2 * im2 - j2x2 - 2
                        if abs(m2x2 - m1x2) > 2 * k:
                            continue
                        ind = jx * jx * ind1 + jx * (j1x2 + m1x2) // 2 + (j2x2 + m2x2) // 2
                        out.append((ind, k, j1x2, m1x2, j2x2, m2x2, ij1))
    return nlx, out


def gaunt_table(nlx: int):
    """Gaunt table in CONF.GNT order (src/breit.f90:48-106)."""
    nlx, ent = gaunt_entries(nlx)
    counts = np.zeros(nlx + 1, np.int32)
    idx, val = [], []
    for ind, k, j1x2, m1x2, j2x2, m2x2, ij1 in ent:
        idx.append(ind)
        val.append(gaun(k, j1x2, m1x2, j2x2, m2x2))
        counts[ij1 - 1] += 1
    return nlx, counts, np.array(idx, np.int32), np.array(val, np.float64)


# ----------------------------------------------------------------------------
# Integrals
# ----------------------------------------------------------------------------

def one_electron(shells: Sequence[Shell], nso: int, rng, nx=F.IPX, diag=(-0.5, 2.0), sigma=0.01):
    """h_ab for pairs of equal kappa (na <= nb); diagonal ascending by shell."""
    ns = len(shells)
    dvals = np.linspace(diag[0], diag[1], ns - nso) if ns > nso else np.zeros(0)
    r1, i1 = [], []
    for a in range(nso, ns):
        for b in range(a, ns):
            if shells[a][2] != shells[b][2]:
                continue
            v = dvals[a - nso] if a == b else rng.normal(0.0, sigma)
            r1.append(v)
            i1.append(nx * (a - nso) + (b + 1 - nso))
    return np.array(r1, np.float64), np.array(i1, np.int32)


def two_electron(shells: Sequence[Shell], nso: int, rng, kbrt: int = 0, nx=F.IPX,
                 sigma=0.05, decay=0.3, breit_sigma=0.002):
    """Canonical R^k set with the loop structure of pbasc's Rint.

    For each (n1 <= n2, n1 <= n3 <= n4) and the two orderings of (n3, n4) with
    nb <= nd, every k of the triangle range with k+la+lc even (all k when
    Breit is on) gets one value N(0, sigma) * exp(-decay (sum n - 4 n_min)).
    """
    ns = len(shells)
    nmin = min(s[0] for s in shells[nso:])
    nn = [s[0] for s in shells]
    ll = [s[1] for s in shells]
    jj = [s[3] for s in shells]
    keys2, keys3, scales, kparity = [], [], [], []
    for n1 in range(nso, ns):
        for n2 in range(n1, ns):
            for n3 in range(n1, ns):
                for n4 in range(n3, ns):
                    na, nb, nc, nd = n1, n2, n3, n4
                    for iis in range(2):
                        if nb <= nd:
                            la, lb, lc, ld = ll[na], ll[nb], ll[nc], ll[nd]
                            if (la + lb + lc + ld) % 2:
                                break
                            ja, jb, jc, jd = jj[na], jj[nb], jj[nc], jj[nd]
                            kmin = max(abs(ja - jc) // 2, abs(jb - jd) // 2)
                            kmax = min((ja + jc) // 2, (jb + jd) // 2)
                            sc = math.exp(-decay * (nn[na] + nn[nb] + nn[nc] + nn[nd] - 4 * nmin))
                            for k in range(kmin, kmax + 1):
                                par = (k + la + lc) % 2 == 0
                                if not par and kbrt == 0:
                                    continue
                                keys2.append(nx * nx * k + nx * (na - nso) + (nc + 1 - nso))
                                keys3.append(nx * (nb - nso) + (nd + 1 - nso))
                                scales.append(sc)
                                kparity.append(par)
                            if n1 == n2 or n3 == n4:
                                break
                            nc, nd = n4, n3
    scales = np.array(scales)
    par = np.array(kparity, bool)
    coul = rng.normal(0.0, sigma, scales.size) * scales
    coul[~par] = 0.0
    if kbrt:
        br = rng.normal(0.0, breit_sigma, scales.size) * scales
        rint2 = np.stack([coul, br], axis=1)
    else:
        rint2 = coul
    return rint2, np.array(keys2, np.int32), np.array(keys3, np.int32)


# ----------------------------------------------------------------------------
# Space specification and writer
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class SpaceSpec:
    name: str
    Ne: int
    Jm: float
    parity: int
    nmin: int
    nmax: int
    lmax: int
    refs: Sequence[Sequence[Tuple[int, int, int, int]]]   # (n, l, 2j, q)
    level: int = 2
    Nlv: int = 4
    Crt4: float = 1e-8
    N_it: int = 100
    Nc4: int = 0           # 0: all configurations
    MaxNd0: int = 3000
    Kbrt: int = 0
    seed: int = 1
    Z: float = 20.0
    nmin_l: Dict[int, int] = None
    nmax_l: Dict[int, int] = None
    Kv: int = 4            # 3: Davidson with selection of states of total J = J
    J: float = None        # target J for Kv = 3 (default: Jm)
    Ksig: int = 0          # 1: MBPT one-electron effective integrals (SGC.CON, Kl = 2)
    Kdsig: int = 0         # 1: energy-dependent sigma (valence energy E0 from stdin)
    E0: float = -1.5       # the valence energy a Kdsig = 1 run is given
    Gj: float = 0.0        # CONF.INP Gj: weight of the J^2 term Hmltn adds (src/pconf.F90:1653-1692)
    Ksym: int = 1          # Ksig = 2: SCRC.CON's Ksym (0: GintS assumes the approximate symmetry nb <= nd)
    KmaxS: int = 8         # Ksig = 2: SCRC.CON's Kmax (the largest multipole of the sigma integrals)
    LmaxS: int = 3         # SGC.CON covers shells with l <= LmaxS ...
    NmaxS: int = 0         # ... and index <= NmaxS (0: every valence shell)
    h_diag: Tuple[float, float] = (-0.5, 2.0)  # one-electron diagonal: first and last shell (Ha)
    r_sigma: float = 0.05  # scale of the Coulomb R^k draws (Ha)
    by_energy: bool = False  # list configurations by ascending one-electron energy


def c3_specs(nmax: int = 8, max_nd0: int = 3000) -> List[SpaceSpec]:
    """Batched-block workload: the twelve (J, parity) blocks of one 3-electron
    atom, M_J = J + 1/2 for J = 0..5 and both parities, s..f shells n = 2..nmax,
    SD excitations from three reference configurations per parity; Nlv = 6,
    Crt4 = 1e-7, per-block seed 300 + block, head blocks of up to 3000
    determinants.  Block sizes are uneven (they fall with M_J); at nmax = 8 they
    run from 1.5k to 17k determinants and every block converges."""
    odd = [[(2, 0, 1, 2), (2, 1, 1, 1)], [(2, 1, 1, 1), (2, 1, 3, 2)], [(2, 0, 1, 1), (2, 1, 3, 1), (3, 2, 5, 1)]]
    even = [[(2, 0, 1, 2), (3, 0, 1, 1)], [(2, 0, 1, 1), (2, 1, 1, 1), (2, 1, 3, 1)], [(2, 1, 1, 2), (3, 2, 3, 1)]]
    out = []
    for b, (par, refs) in enumerate([(p, r) for p, r in ((0, even), (1, odd)) for _ in range(6)]):
        J = b % 6
        out.append(SpaceSpec(name=f"C3_J{J}_{'odd' if par else 'even'}", Ne=3, Jm=J + 0.5, parity=par, nmin=2,
                             nmax=nmax, lmax=3, refs=refs, level=2, Nlv=6, Crt4=1e-7, N_it=300, seed=300 + b,
                             MaxNd0=max_nd0))
    return out


def c3_blocks() -> List[SpaceSpec]:
    """Batched-block workload at BASELINE's sizes: the twelve (J, parity)
    blocks M_J = J = 0..5, even and odd parity, of one 4-electron atom, s..f
    shells from n = 2 (n <= 6 for the even blocks, n <= 5 for the odd ones,
    so the sizes spread further), SD+T excitations from the first five C4
    references, configurations by one-electron energy (diagonal -0.5 ... 2.0
    Ha, R^k scale 0.05), Coulomb only; Nlv = 6, Crt4 = 1e-7, per-block seed
    300 + block.  Nd runs from 36,466 (J = 5, odd) to 375,569 (J = 0, even),
    4.0e9 elements in both triangles in all (BASELINE: 20,000 to 400,000 per
    block).  Two electrons (the C1 generator) cannot reach these sizes: a
    2-electron block of 4e5 determinants needs thousands of spin-orbitals.
    With the C4 scales (R^k 0.07, diagonal -0.5 ... 1.5 Ha) the J = 1 even
    block's pconf Davidson stalls near 6e-4."""
    out = []
    for b in range(12):
        par, J = b // 6, b % 6
        out.append(SpaceSpec(name=f"C3B_J{J}_{'odd' if par else 'even'}", Ne=4, Jm=float(J), parity=par, nmin=2,
                             nmax=5 if par else 6, lmax=3, refs=C4_REFS[:5], level=3, Nlv=6, Crt4=1e-7, N_it=300,
                             seed=300 + b, h_diag=(-0.5, 2.0), r_sigma=0.05, by_energy=True))
    return out


def _shell_index(shells, n, l, j2):
    for i, s in enumerate(shells):
        if s[0] == n and s[1] == l and s[3] == j2:
            return i
    raise KeyError((n, l, j2))


def build_space(spec: SpaceSpec):
    """Basis and configuration list.  With ``by_energy`` the configurations
    are listed by ascending sum of their shells' one-electron diagonal
    energies (ties in discovery order), as a pconf user lists the main
    configurations first: the Nd0 head block of Davidson's initial
    approximation then holds the lowest-lying determinants."""
    shells = make_basis(spec.nmin, spec.nmax, spec.lmax, spec.nmin_l, spec.nmax_l)
    refs = [[(_shell_index(shells, n, l, j2), q) for (n, l, j2, q) in r] for r in spec.refs]
    configs = excite(refs, shells, spec.level, spec.parity)
    if spec.by_energy:
        nsu = max(s for c in configs for s, _ in c) + 1
        e = np.linspace(spec.h_diag[0], spec.h_diag[1], nsu)  # the diagonal one_electron() writes
        configs = sorted(configs, key=lambda c: sum(q * e[s] for s, q in c))
    return shells, configs


def sigma_tables(shells, nso: int, nmaxs: int, lmaxs: int, rng, nmin: int):
    """Synthetic MBPT one-electron corrections in SGC.CON order: for every pair
    na <= nb (1-based, Nso < na <= nb <= NmaxS, l_a <= LmaxS, same l and j),
    R ~ N(0, 0.003) exp(-0.2 (n_a + n_b - 2 nmin)), D = R * U(-0.4, 0.6) (both
    signs of D/R), E = Eps(na) + N(0, 0.05); orbital energies Eps ascending
    from -0.8 to 1.5 Ha.  Values are kept at FP32 precision, as the file is
    read into single precision."""
    ns = len(shells)
    eps = np.linspace(-0.8, 1.5, ns)
    ent = []
    for na in range(nso + 1, nmaxs + 1):
        n_a, la, _, ja = shells[na - 1]
        if la > lmaxs:
            continue
        for nb in range(na, nmaxs + 1):
            n_b, lb, _, jb = shells[nb - 1]
            if lb != la or jb != ja:
                continue
            r = rng.normal(0.0, 0.003) * np.exp(-0.2 * (n_a + n_b - 2 * nmin))
            d = r * rng.uniform(-0.4, 0.6)
            e = eps[na - 1] + rng.normal(0.0, 0.05)
            ent.append((na, nb, float(np.float32(r)), float(np.float32(d)), float(np.float32(e))))
    return eps, ent


def gints_canonical(na: int, nb: int, nc: int, nd: int, ksym: int):
    """The shell order GintS looks its integrals up in (src/integrals.f90:
    413-436): the lowest shell first, ties broken as the reference does, and
    with Ksym = 0 the approximate symmetry nb <= nd."""
    if na > nc and nb >= nc and nd >= nc:
        na, nc, nb, nd = nc, na, nd, nb
    elif na > nb and nc > nb and nd >= nb:
        na, nb, nc, nd = nb, na, nd, nc
    elif na > nd and nb > nd and nc > nd:
        na, nd, nb, nc = nd, na, nc, nb
    if na == nb and nc > nd:
        nc, nd = nd, nc
    if na == nd and nb > nc:
        if na != nc:
            nc, nb = nb, nc
        else:
            nd, nb = nb, nd
    if na == nc and nb > nd:
        nb, nd = nd, nb
    if ksym == 0:
        if nb > nd:
            nb, nd = nd, nb
        if na == nb and nc > nd:
            nc, nd = nd, nc
    return na, nb, nc, nd


def sigma2_tables(shells, nso: int, nmaxs: int, lmaxs: int, kmax: int, ksym: int, nsum: int, eps, rng, nmin: int,
                  nx: int = 440):
    """Synthetic two-electron MBPT integrals (SCRC.CON) for every key GintS
    can ask for: all shell quadruples with l <= LmaxS, index <= NmaxS, index
    sum <= Nsum and even l sum, each in both of Gint's orders, k from the
    triangle rule of the quadruple as given up to Kmax.  R ~ N(0, 0.002)
    exp(-0.3 (sum n - 4 nmin)), D = R U(-0.4, 0.6) (both signs of D/R), E =
    Eps(a) + Eps(b) + N(0, 0.05) of the canonical quadruple's a, b; FP64."""
    ns = len(shells)
    val = [i for i in range(nso + 1, min(nmaxs, ns) + 1) if shells[i - 1][1] <= lmaxs]
    keys = {}
    for a in val:
        for b in val:
            for c in val:
                for d in val:
                    la, lb, lc, ld = (shells[x - 1][1] for x in (a, b, c, d))
                    if (la + lb + lc + ld) % 2 or a + b + c + d > nsum:
                        continue
                    ja, jb, jc, jd = (shells[x - 1][3] for x in (a, b, c, d))
                    na, nb, nc, nd = gints_canonical(a, b, c, d, ksym)
                    kmn = max(abs(ja - jc) // 2, abs(jb - jd) // 2)
                    kmx = min((ja + jc) // 2, (jb + jd) // 2, kmax)
                    for k in range(kmn, kmx + 1):
                        key = (nx * nx * k + nx * (na - nso - 1) + (nc - nso), nx * (nb - nso - 1) + (nd - nso))
                        if key not in keys:
                            keys[key] = (na, nb, nc, nd)
    order = sorted(keys)
    n = len(order)
    i2 = np.array([k[0] for k in order], np.int32)
    i3 = np.array([k[1] for k in order], np.int32)
    nsum_n = np.array([sum(shells[x - 1][0] for x in keys[k]) for k in order], np.float64)
    r = rng.normal(0.0, 0.002, n) * np.exp(-0.3 * (nsum_n - 4 * nmin))
    dd = r * rng.uniform(-0.4, 0.6, n)
    e = np.array([eps[keys[k][0] - 1] + eps[keys[k][1] - 1] for k in order]) + rng.normal(0.0, 0.05, n)
    return i2, i3, r, dd, e


def write_space(spec: SpaceSpec, outdir: str, Kw: int = 0, KXIJ: int = 0) -> dict:
    """Write CONF.INP, ci.in, CONF.DAT, CONF.INT, CONF.GNT (and, with Ksig,
    SGC.CON) for ``spec``."""
    os.makedirs(outdir, exist_ok=True)
    shells, configs = build_space(spec)
    used = sorted({s for c in configs for s, _ in c})
    nsu = used[-1] + 1
    shells = shells[:nsu]
    rng = np.random.default_rng(spec.seed)
    nso = 0
    codes = []
    for c in configs:
        codes.append([F.encode_shell(shells[s][0], shells[s][1], shells[s][2], q) for s, q in c])
    nc = len(configs)
    inp = F.ConfInp(name=spec.name, Z=spec.Z, Am=2 * spec.Z, J=spec.Jm if spec.J is None else spec.J,
                    Jm=spec.Jm, Nso=nso, Nc=nc, Kv=spec.Kv, Nlv=spec.Nlv, Ne=spec.Ne, Kl4=1, Nc4=spec.Nc4 or nc, Crt4=spec.Crt4,
                    N_it=spec.N_it, Kbrt=spec.Kbrt, Gj=spec.Gj, configs=codes)
    F.write_conf_inp(os.path.join(outdir, "CONF.INP"), inp)
    if spec.Ksig:
        # MBPT: Kl = 2 (new run with effective integrals), SGC.CON, energies in CONF.DAT
        srng = np.random.default_rng(spec.seed + 1000)
        nmaxs = spec.NmaxS or len(shells)
        eps, ent = sigma_tables(shells, nso, nmaxs, spec.LmaxS, srng, spec.nmin)
        F.write_ci_in(os.path.join(outdir, "ci.in"), Kl=2, Kw=Kw, KXIJ=KXIJ, MaxNd0=spec.MaxNd0, Ksig=spec.Ksig,
                      Kdsig=spec.Kdsig)
        F.write_conf_dat(os.path.join(outdir, "CONF.DAT"), spec.Z, shells, eps=eps)
        scr = None
        if spec.Ksig >= 2:
            # screening coefficients of Gint's R^k, k = 0..9, and the
            # two-electron sigma integrals of GintS (SCRC.CON)
            scr = np.round(np.linspace(0.72, 0.95, 10) + srng.uniform(-0.01, 0.01, 10), 5)
            nsum = 4 * nmaxs
            i2s, i3s, r2s, d2s, e2s = sigma2_tables(shells, nso, nmaxs, spec.LmaxS, spec.KmaxS, spec.Ksym, nsum,
                                                    eps, srng, spec.nmin)
            F.write_scrc_con(os.path.join(outdir, "SCRC.CON"), F.ScrcCon(
                ns=len(shells), nso=nso, khot=0, nsx1=nmaxs, nsx2=1, ksym=spec.Ksym, nmaxs=nmaxs, lmaxs=spec.LmaxS,
                kmax=spec.KmaxS, nav=nmaxs, kbox=0, nx=440, rint2s=r2s, iint2s=i2s, iint3s=i3s, dint2s=d2s,
                eint2s=e2s))
        F.write_sgc_con(os.path.join(outdir, "SGC.CON"), nmaxs, spec.LmaxS, 0, ent, scr=scr)
        with open(os.path.join(outdir, "E0.in"), "w") as f:  # the answer to pconf's valence-energy prompt
            f.write(f"{spec.E0!r}\n")
    else:
        F.write_ci_in(os.path.join(outdir, "ci.in"), Kw=Kw, KXIJ=KXIJ, MaxNd0=spec.MaxNd0)
        F.write_conf_dat(os.path.join(outdir, "CONF.DAT"), spec.Z, shells)
    r1, i1 = one_electron(shells, nso, rng, diag=spec.h_diag)
    r2, i2, i3 = two_electron(shells, nso, rng, kbrt=spec.Kbrt, sigma=spec.r_sigma)
    nsp = sum(len(c) for c in configs)
    ci = F.ConfInt(Ns=len(shells), Nso=nso, Nsp=nsp, Nsx=len(shells), Ecore=0.0,
                   shells=np.array([(n, k, l, j) for (n, l, k, j) in shells], np.int32),
                   nq_nip=np.zeros((nsp, 2), np.int32), kbrt=spec.Kbrt, rint1=r1, iint1=i1,
                   rint2=r2, iint2=i2, iint3=i3)
    F.write_conf_int(os.path.join(outdir, "CONF.INT"), ci)
    lmax = max(s[1] for s in shells)
    nlx, counts, idx, val = gaunt_table(lmax)
    F.write_conf_gnt(os.path.join(outdir, "CONF.GNT"), nlx, counts, idx, val)
    return dict(Ns=len(shells), Nc=nc, Ngint=int(i2.size), Nhint=int(i1.size), Ngaunt=int(idx.size))


# references of the C4 space: (n, l, 2j, occupation) per shell
C4_REFS = [[(2, 0, 1, 2), (2, 1, 1, 2)], [(3, 2, 5, 4)], [(4, 3, 7, 4)], [(4, 0, 1, 2), (4, 1, 3, 2)],
           [(5, 2, 5, 4)], [(5, 3, 7, 4)], [(6, 1, 3, 4)], [(6, 2, 5, 4)]]

# references of the C5 space: 4f7/2^4, 5g9/2^4, 5f7/2^4, 6g9/2^4, 6f7/2^4
C5_REFS = [[(4, 3, 7, 4)], [(5, 4, 9, 4)], [(5, 3, 7, 4)], [(6, 4, 9, 4)], [(6, 3, 7, 4)]]

# Workloads.  Sizes quoted are what the enumeration actually produces.
PRESETS: Dict[str, SpaceSpec] = {
    # 2 electrons, n=5..12, s/p/d, M_J=0, even parity, all 2e configurations
    # (SD from 5s^2): Nc=436, Nd=1216.
    "c1": SpaceSpec(name="C1_2e_spd", Ne=2, Jm=0.0, parity=0, nmin=5, nmax=12, lmax=2,
                    refs=[[(5, 0, 1, 2)]], level=2, Nlv=4, Crt4=1e-8, N_it=100, seed=1),
    # 3 electrons, s..f, n=2..14, M_J=1/2, odd parity, SD from three
    # reference configurations (2s2 2p1/2, 2p1/2 2p3/2^2, 2s 2p3/2 3d5/2):
    # Nc=7067, Nd=69101, NumH=202,977,105, Nd0=2995.
    "c2": SpaceSpec(name="C2_3e_spdf", Ne=3, Jm=0.5, parity=1, nmin=2, nmax=14, lmax=3,
                    refs=[[(2, 0, 1, 2), (2, 1, 1, 1)],
                          [(2, 1, 1, 1), (2, 1, 3, 2)],
                          [(2, 0, 1, 1), (2, 1, 3, 1), (3, 2, 5, 1)]],
                    level=2, Nlv=10, Crt4=1e-7, N_it=300, seed=2),
    # the same generator on n=3..14 with references 3s2 3p3/2, 3p3/2 3d5/2^2,
    # 3s 3p3/2 3d5/2: Nd=55259, NumH=149,162,987; the diagonal Davidson
    # correction converges slowly here (about 240 iterations to 1e-7)
    "c2a": SpaceSpec(name="C2A_3e_slow", Ne=3, Jm=0.5, parity=1, nmin=3, nmax=14, lmax=3,
                     refs=[[(3, 0, 1, 2), (3, 1, 3, 1)],
                           [(3, 1, 3, 1), (3, 2, 5, 2)],
                           [(3, 0, 1, 1), (3, 1, 3, 1), (3, 2, 5, 1)]],
                     level=2, Nlv=10, Crt4=1e-7, N_it=300, seed=2),
    # 4 electrons, s..f, n=2..7, M_J=0, even parity, SD+T from eight reference
    # configurations far apart from each other (2s2 2p1/2^2, 3d5/2^4, 4f7/2^4,
    # 4s2 4p3/2^2, 5d5/2^4, 5f7/2^4, 6p3/2^4, 6d5/2^4), Coulomb + Breit,
    # configurations listed by one-electron energy: Nc=29213, Nd=1079368,
    # NumH~2.09e9 (4.17e9 elements in both triangles), Nd0~2990.  SD+T of one
    # or two neighbouring references has ~1 % density, so BASELINE's element
    # count (~4e9) is reached at ~1.1e6 determinants, not ~2e6.  The
    # discovery-ordered list left low-lying configurations outside the Nd0
    # head and pconf's Davidson stalled near 1e-3.
    "c4": SpaceSpec(name="C4_4e_spdf_breit", Ne=4, Jm=0.0, parity=0, nmin=2, nmax=7, lmax=3,
                    refs=C4_REFS, level=3, Nlv=8, Crt4=1e-7, N_it=300, seed=4, Kbrt=1,
                    h_diag=(-0.5, 1.5), r_sigma=0.07, by_energy=True),
    # 4 electrons, s..g: s shells n=2..16, p, d, f n up to 7, g n=5..6,
    # M_J=0, even parity, SD+T from five reference configurations (4f7/2^4,
    # 5g9/2^4, 5f7/2^4, 6g9/2^4, 6f7/2^4), Coulomb + Breit, configurations by
    # one-electron energy (diagonal -0.5 ... 2.5 Ha), R^k scale 0.03, Nlv=20:
    # Nc=46033, Nd=2332805, ~1.22e10 elements in both triangles (~5300 per
    # row), residual lists of up to 1811 determinants.  With R^k scale 0.07
    # and a -0.5 ... 1.5 Ha diagonal the 20th root sits in a near-degenerate
    # pair and pconf's Davidson stalls near 1e-6 (plain) or 1e-3 (FP32).  Four electrons with SD+T do not reach BASELINE's
    # ~8e6 determinants at density 2e-4: every determinant shares a pair with
    # ~Nvirt^2 / (M spread) others, so rows hold thousands of elements once Nd
    # is in the millions; the element count (the SpMM's work and the resident
    # memory) is what this instance matches.
    "c5": SpaceSpec(name="C5_4e_spdfg_breit", Ne=4, Jm=0.0, parity=0, nmin=2, nmax=16, lmax=4,
                    nmax_l={1: 7, 2: 7, 3: 7, 4: 6}, refs=C5_REFS, level=3, Nlv=20, Crt4=1e-7, N_it=300,
                    seed=5, Kbrt=1, h_diag=(-0.5, 2.5), r_sigma=0.03, by_energy=True),
    # the C5 generator on n=2..5 (training pairs for the learned correction)
    "c5s": SpaceSpec(name="C5S_train", Ne=4, Jm=0.0, parity=0, nmin=2, nmax=5, lmax=4,
                     refs=C5_REFS[:3], level=3, Nlv=20, Crt4=1e-7, N_it=300, seed=15, Kbrt=1,
                     h_diag=(-0.5, 2.5), r_sigma=0.03, by_energy=True),
    # the same generator on n=2..5 (training pairs for the learned correction)
    "c4s": SpaceSpec(name="C4S_train", Ne=4, Jm=0.0, parity=0, nmin=2, nmax=5, lmax=3,
                     refs=C4_REFS[:5], level=3, Nlv=8, Crt4=1e-7, N_it=300, seed=14, Kbrt=1,
                     h_diag=(-0.5, 1.5), r_sigma=0.07, by_energy=True),
    # smaller siblings (n up to 8) that supply training pairs for the
    # learned correction: Nd=16250 (c2s), the c2a references (c2as)
    "c2s": SpaceSpec(name="C2S_3e_train", Ne=3, Jm=0.5, parity=1, nmin=2, nmax=8, lmax=3,
                     refs=[[(2, 0, 1, 2), (2, 1, 1, 1)],
                           [(2, 1, 1, 1), (2, 1, 3, 2)],
                           [(2, 0, 1, 1), (2, 1, 3, 1), (3, 2, 5, 1)]],
                     level=2, Nlv=10, Crt4=1e-7, N_it=300, seed=12),
    "c2as": SpaceSpec(name="C2AS_3e_train", Ne=3, Jm=0.5, parity=1, nmin=3, nmax=8, lmax=3,
                      refs=[[(3, 0, 1, 2), (3, 1, 3, 1)],
                            [(3, 1, 3, 1), (3, 2, 5, 2)],
                            [(3, 0, 1, 1), (3, 1, 3, 1), (3, 2, 5, 1)]],
                      level=2, Nlv=10, Crt4=1e-7, N_it=300, seed=12),
    # small 3-electron space for parity tests (all three excitation levels,
    # two common orbitals in single excitations)
    "t3": SpaceSpec(name="T3_3e_small", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=6, lmax=3,
                    refs=[[(4, 0, 1, 2), (4, 1, 1, 1)], [(4, 1, 3, 1), (4, 2, 5, 2)]],
                    level=2, Nlv=6, Crt4=1e-8, N_it=200, seed=3, MaxNd0=400),
    # the same with Breit integrals
    "t3b": SpaceSpec(name="T3B_breit", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=6, lmax=3,
                     refs=[[(4, 0, 1, 2), (4, 1, 1, 1)], [(4, 1, 3, 1), (4, 2, 5, 2)]],
                     level=2, Nlv=6, Crt4=1e-8, N_it=200, seed=13, Kbrt=1, MaxNd0=400),
    # t3 with Gj = 0.05: Hmltn adds Gj J^2 to the diagonal and double-difference elements
    "t3gj": SpaceSpec(name="T3GJ_j2", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=6, lmax=3,
                      refs=[[(4, 0, 1, 2), (4, 1, 1, 1)], [(4, 1, 3, 1), (4, 2, 5, 2)]],
                      level=2, Nlv=6, Crt4=1e-8, N_it=200, seed=3, MaxNd0=400, Gj=0.05),
    # J-selection runs (Kv = 3): t3 for the states with J = 3/2, c1 for J = 2
    "t3j": SpaceSpec(name="T3J_jsel", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=6, lmax=3,
                     refs=[[(4, 0, 1, 2), (4, 1, 1, 1)], [(4, 1, 3, 1), (4, 2, 5, 2)]],
                     level=2, Nlv=4, Crt4=1e-8, N_it=200, seed=3, MaxNd0=400, Kv=3, J=1.5),
    "c1j": SpaceSpec(name="C1J_jsel", Ne=2, Jm=0.0, parity=0, nmin=5, nmax=12, lmax=2,
                     refs=[[(5, 0, 1, 2)]], level=2, Nlv=4, Crt4=1e-8, N_it=100, seed=1, MaxNd0=600, Kv=3, J=2.0),
    # tiny spaces for the brute-force second-quantisation check
    "tiny2": SpaceSpec(name="TINY2", Ne=2, Jm=0.0, parity=0, nmin=4, nmax=5, lmax=2,
                       refs=[[(4, 0, 1, 2)]], level=2, Nlv=3, Crt4=1e-8, N_it=100, seed=21),
    "tiny3": SpaceSpec(name="TINY3", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=5, lmax=1,
                       refs=[[(4, 0, 1, 2), (4, 1, 1, 1)]], level=3, Nlv=3, Crt4=1e-8, N_it=100, seed=22),
    "tiny3d": SpaceSpec(name="TINY3D", Ne=3, Jm=1.5, parity=0, nmin=4, nmax=4, lmax=3,
                        refs=[[(4, 0, 1, 1), (4, 1, 3, 1), (4, 3, 7, 1)]], level=3, Nlv=3, Crt4=1e-8, N_it=100,
                        seed=23),
    # MBPT one-electron effective integrals (Kl = 2, Ksig = 1): t3 with SGC.CON
    # on shells up to l = 2, without (t3s) and with (t3sd) energy dependence,
    # and the tiny space for the second-quantisation check
    "t3s": SpaceSpec(name="T3S_sigma", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=6, lmax=3,
                     refs=[[(4, 0, 1, 2), (4, 1, 1, 1)], [(4, 1, 3, 1), (4, 2, 5, 2)]],
                     level=2, Nlv=6, Crt4=1e-8, N_it=200, seed=3, MaxNd0=400, Ksig=1, LmaxS=2),
    "t3sd": SpaceSpec(name="T3SD_sigma_e", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=6, lmax=3,
                      refs=[[(4, 0, 1, 2), (4, 1, 1, 1)], [(4, 1, 3, 1), (4, 2, 5, 2)]],
                      level=2, Nlv=6, Crt4=1e-8, N_it=200, seed=3, MaxNd0=400, Ksig=1, Kdsig=1, LmaxS=2,
                      E0=-1.7),
    # MBPT with screening and two-electron sigma (Ksig = 2: SCRC.CON, the
    # screening coefficients of SGC.CON): t3 without (t3s2, Ksym = 0, sigma
    # multipoles up to k = 3) and with (t3sd2) energy dependence
    "t3s2": SpaceSpec(name="T3S2_sigma2", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=6, lmax=3,
                      refs=[[(4, 0, 1, 2), (4, 1, 1, 1)], [(4, 1, 3, 1), (4, 2, 5, 2)]],
                      level=2, Nlv=6, Crt4=1e-8, N_it=200, seed=3, MaxNd0=400, Ksig=2, LmaxS=2, Ksym=0, KmaxS=3),
    "t3sd2": SpaceSpec(name="T3SD2_sigma2_e", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=6, lmax=3,
                       refs=[[(4, 0, 1, 2), (4, 1, 1, 1)], [(4, 1, 3, 1), (4, 2, 5, 2)]],
                       level=2, Nlv=6, Crt4=1e-8, N_it=200, seed=3, MaxNd0=400, Ksig=2, Kdsig=1, LmaxS=2,
                       E0=-1.7),
    "tiny3s": SpaceSpec(name="TINY3S", Ne=3, Jm=0.5, parity=1, nmin=4, nmax=5, lmax=1,
                        refs=[[(4, 0, 1, 2), (4, 1, 1, 1)]], level=3, Nlv=3, Crt4=1e-8, N_it=100, seed=22,
                        Ksig=1, LmaxS=1, NmaxS=5),
    # 4 electrons, s/p/d, SDT from 4s2 4p2
    "t4": SpaceSpec(name="T4_4e_small", Ne=4, Jm=0.0, parity=0, nmin=4, nmax=6, lmax=2,
                    refs=[[(4, 0, 1, 2), (4, 1, 1, 2)]], level=3, Nlv=5, Crt4=1e-8, N_it=200, seed=4,
                    MaxNd0=500),
    # the C4 combination at a size the oracle builds in seconds: 4 electrons,
    # s..f, SD+T, Coulomb + Breit (Nd = 9795)
    "t4b": SpaceSpec(name="T4B_4e_breit", Ne=4, Jm=0.0, parity=0, nmin=4, nmax=5, lmax=3,
                     refs=[[(4, 0, 1, 2), (4, 1, 1, 2)]], level=3, Nlv=5, Crt4=1e-8, N_it=200, seed=14,
                     Kbrt=1, MaxNd0=500),
    # g shells (l = 4, the C5 basis) at oracle size: two electrons over the
    # n = 5 shells for the second-quantisation check, and the C5 combination
    # (4 electrons, s..g, SD+T, Coulomb + Breit) on the n = 5 shells (Nd = 3587)
    "tiny2g": SpaceSpec(name="TINY2G", Ne=2, Jm=0.0, parity=0, nmin=5, nmax=5, lmax=4,
                        refs=[[(5, 0, 1, 2)]], level=2, Nlv=3, Crt4=1e-8, N_it=100, seed=24),
    "t4g": SpaceSpec(name="T4G_4e_spdfg_breit", Ne=4, Jm=0.0, parity=0, nmin=5, nmax=5, lmax=4,
                     refs=[[(5, 0, 1, 2), (5, 1, 1, 2)]], level=3, Nlv=5, Crt4=1e-8, N_it=200, seed=15,
                     Kbrt=1, MaxNd0=500),
}
