"""CPU oracle of pconf's CSF-basis path (kCSF = 1).

TEST INFRASTRUCTURE ONLY.  Imported by tests/ and by bench.py's cpu_baseline
leg as the checker; never by the product package.

    jbasis      src/csf.f90:17-407    J^2 blocks of the configuration classes,
                                      DSYEVD, the J = Jm eigenvectors (numpy's
                                      eigh is LAPACK's dsyevd as well)
    formh_sym   src/csf.f90:538-1006  H in the CSF basis (C: orc_formh_sym)
    unsym       src/csf.f90:1088-1134 CSF-basis vectors -> determinant basis

The configuration classes (nonequiv_conf, src/csf.f90:454-536) are formed by
Oracle.from_dir when ci.in has kCSF > 0 (Oracle.nonequiv).

The eigenvectors of a degenerate J = Jm subspace are not unique: any
orthonormal basis of it is as valid as LAPACK's, so a basis found elsewhere
(the device's) differs from this one by an orthogonal transformation inside
each class.  formh_sym and unsym take the basis as input, so they can be
compared bit for bit on any given basis; the basis itself is compared through
the projector C C^T of each class, which is unique.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .oracle import Oracle, _ptr


class CsfBasis:
    """jbasis's outputs: ndcs (CSFs per class), place (offset of each class's
    coefficients, iplace_cj), ccj (column-major ndi x ndcs blocks), and per
    configuration: the class (0-based), mdcs (first CSF) and the CSF count."""

    def __init__(self, o: Oracle, ndcs, ccj):
        self.nc_neq = np.ascontiguousarray(o.nc_neq, np.int32)
        self.nconf_neq = int(o.nconf_neq)
        self.ndcs = np.ascontiguousarray(ndcs, np.int32)
        first = class_representatives(o)
        ndc_neq = np.array([o.ndc[first[q] - 1] if first[q] else 0 for q in range(self.nconf_neq)], np.int64)
        self.ndc_neq = ndc_neq
        sizes = self.ndcs.astype(np.int64) * ndc_neq
        self.place = np.zeros(self.nconf_neq, np.int64)
        if self.nconf_neq > 1:
            np.cumsum(sizes[:-1], out=self.place[1:])
        self.nccj = int(sizes.sum())
        self.ccj = np.ascontiguousarray(ccj, np.float64)
        assert self.ccj.size == self.nccj
        self.conf_ncsf = self.ndcs[self.nc_neq - 1].astype(np.int64)
        self.mdcs = np.concatenate([[0], np.cumsum(self.conf_ncsf)[:-1]]).astype(np.int64)
        self.ncsf = int(self.conf_ncsf.sum())

    def block(self, q):
        """Coefficients of class q (0-based) as an ndi x ncsfi array."""
        ndi, nci = int(self.ndc_neq[q]), int(self.ndcs[q])
        return self.ccj[self.place[q]:self.place[q] + ndi * nci].reshape(nci, ndi).T


def class_representatives(o: Oracle):
    """First configuration (1-based) of each class, 0 for none."""
    first = [0] * o.nconf_neq
    for ic, q in enumerate(o.nc_neq):
        if first[q - 1] == 0:
            first[q - 1] = ic + 1
    return first


def j2_block(o: Oracle, ic: int) -> np.ndarray:
    """build_j2 (src/csf.f90:409-452) of configuration ic (1-based)."""
    ndi = int(o.ndc[ic - 1])
    zz = np.zeros(ndi * ndi, np.float64)
    if ndi:
        o.L.orc_j2_block(o.c, ic, zz)
    return zz.reshape(ndi, ndi)


def j_of(de):
    """jbasis's J of a J^2 eigenvalue: tj = (sqrt(1 + de) - 1) / 2 and the
    doubled J jtt = int(2 tj + 1e-4), truncated as Fortran's integer
    assignment does (src/csf.f90:175-182)."""
    tj = 0.5 * (np.sqrt(1.0 + de) - 1.0)
    jtt = np.trunc(2.0 * tj + 0.0001).astype(np.int64)
    return tj, jtt


def jbasis(o: Oracle) -> CsfBasis:
    """jbasis, npes = 1 (src/csf.f90:138-208, 290-381)."""
    first = class_representatives(o)
    ndcs = np.zeros(o.nconf_neq, np.int32)
    parts = []
    for q in range(o.nconf_neq):
        ic = first[q]
        if not ic or o.ndc[ic - 1] == 0:
            continue
        de, zz = np.linalg.eigh(j2_block(o, ic), UPLO="U")
        tj, jtt = j_of(de)
        bad = np.abs(2.0 * tj - jtt) > 1e-7
        if bad.any():
            raise ValueError(f"Value of j in jbasis is wrong: J={tj[bad][0]} configuration {ic}")
        sel = jtt == o.Mj
        ndcs[q] = int(sel.sum())
        if ndcs[q]:
            parts.append(zz[:, sel].T.reshape(-1))
    ccj = np.concatenate(parts) if parts else np.zeros(0)
    return CsfBasis(o, ndcs, ccj)


def formh_sym(o: Oracle, b: CsfBasis, ic_lo: int = 0, ic_hi: int = 0, stats: dict | None = None):
    """formh_sym on basis b -> (rowptr, col 1-based, val, numzero, hmin, npairs):
    the lower triangle of H in the CSF basis as pconf stores it (zeros kept).
    With ic_lo/ic_hi (1-based, inclusive) only the rows of those
    configurations: rowptr then runs over their rows (from 0).  stats, if
    given, receives the configuration and determinant pairs evaluated."""
    L = o.L
    counts = np.zeros(max(b.ncsf, 1), np.int64)
    nz, mn, npairs, ncand = C.c_int64(0), C.c_double(0.0), C.c_int64(0), C.c_int64(0)
    ccj = b.ccj if b.nccj else np.zeros(1)
    total = L.orc_formh_sym(o.c, b.nc_neq, b.ndcs, b.place, ccj, ic_lo, ic_hi, _ptr(counts), None, None, None,
                            C.byref(nz), C.byref(mn), C.byref(npairs), C.byref(ncand))
    rowptr = np.zeros(b.ncsf + 1, np.int64)
    np.cumsum(counts[:b.ncsf], out=rowptr[1:])
    col = np.zeros(max(total, 1), np.int32)
    val = np.zeros(max(total, 1), np.float64)
    L.orc_formh_sym(o.c, b.nc_neq, b.ndcs, b.place, ccj, ic_lo, ic_hi, None, _ptr(rowptr), _ptr(col), _ptr(val),
                    C.byref(nz), C.byref(mn), C.byref(npairs), C.byref(ncand))
    if L.orc_missing():
        raise KeyError("integral absent while building H")
    if stats is not None:
        stats.update(pairs=npairs.value, candidates=ncand.value)
    if ic_lo or ic_hi:
        r0 = int(b.mdcs[max(ic_lo, 1) - 1])
        r1 = int(b.mdcs[ic_hi - 1] + b.conf_ncsf[ic_hi - 1]) if ic_hi else b.ncsf
        rowptr = rowptr[r0:r1 + 1] - rowptr[r0]
    return rowptr, col[:total], val[:total], nz.value, mn.value, npairs.value


def unsym(o: Oracle, b: CsfBasis, x: np.ndarray) -> np.ndarray:
    """unsym (src/csf.f90:1096-1125): x[nvec, ncsf] -> determinant-basis
    vectors y[nvec, Nd]; cc(n) accumulates ccs(k) * ccj over the CSFs of n's
    configuration in order, from 0."""
    x = np.atleast_2d(np.asarray(x, np.float64))
    y = np.zeros((x.shape[0], o.Nd))
    md = np.concatenate([[0], np.cumsum(o.ndc)]).astype(np.int64)
    for ic in range(o.Nc):
        q = int(b.nc_neq[ic]) - 1
        nci = int(b.ndcs[q])
        if nci == 0:
            continue
        cb = b.block(q)
        n1, n2 = md[ic], md[ic + 1]
        for a in range(nci):
            y[:, n1:n2] = y[:, n1:n2] + x[:, b.mdcs[ic] + a][:, None] * cb[:, a][None, :]
    return y
