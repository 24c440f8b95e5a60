"""ctypes front end of the C oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg; never by the product package.

``Oracle.from_dir`` follows pconf's start-up sequence (src/pconf.F90:87-143):
read CONF.INP/CONF.DAT/CONF.INT/CONF.GNT, map configuration entries to basis
orbitals as Init does (src/pconf.F90:506-530), then Dinit + Det_List.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        L = C.CDLL(path)
        P = C.c_void_p
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        L.orc_create.restype = P
        L.orc_free.argtypes = [P]
        L.orc_set_space.argtypes = [P, C.c_int, C.c_int, i32p, i32p, i32p, i32p, C.c_int, i32p, i32p,
                                    C.c_int, C.c_int, C.c_double]
        L.orc_enumerate.argtypes = [P]
        L.orc_enumerate.restype = C.c_int64
        for f in ("orc_nst", "orc_npos", "orc_nc", "orc_nsu"):
            getattr(L, f).argtypes = [P]
        L.orc_get_space.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_dets.argtypes = [P, C.c_int, i32p, i32p]
        L.orc_set_integrals.argtypes = [P, C.c_int, C.c_int, f64p, i32p, C.c_int64, C.c_int, f64p, i32p, i32p]
        L.orc_set_gaunt.argtypes = [P, C.c_int, C.c_int, i32p, f64p]
        L.orc_gaunt.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_gaunt.restype = C.c_double
        L.orc_hint.argtypes = [P, C.c_int, C.c_int]
        L.orc_hint.restype = C.c_double
        L.orc_gint.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_gint.restype = C.c_double
        L.orc_missing.restype = C.c_int
        L.orc_rspq.argtypes = [C.c_int, i32p, i32p, i32p]
        L.orc_element.argtypes = [P, C.c_int, C.c_int, C.c_void_p]
        L.orc_element.restype = C.c_double
        L.orc_formh.argtypes = [P, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_formh.restype = C.c_int64
        L.orc_formh_grow.argtypes = [P, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.POINTER(C.c_int)),
                                     C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64)]
        L.orc_formh_grow.restype = C.c_int64
        L.orc_free_buf.argtypes = [C.c_void_p]
        L.orc_formj.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_formj.restype = C.c_int64
        L.orc_jterms.argtypes = [P, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_jterms.restype = C.c_int
        L.orc_jav.argtypes = [C.c_int, i64p, i32p, f64p, f64p, C.c_int]
        L.orc_jav.restype = C.c_double
        L.orc_mxmpy.argtypes = [C.c_int, i64p, i32p, f64p, C.c_double, f64p, f64p, C.c_void_p]
        L.orc_mxmpy_rows.argtypes = [C.c_int, C.c_int, C.c_int, i64p, i32p, f64p, C.c_double, f64p, f64p]
        L.orc_mxmpy_list.argtypes = [C.c_int, C.c_int, i32p, i64p, i32p, f64p, C.c_double, f64p, f64p]
        L.orc_set_sigma.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i32p, f64p, f64p, f64p, f64p,
                                    C.c_double]
        L.orc_set_sigma.restype = C.c_int
        L.orc_missing_sigma.restype = C.c_int
        L.orc_set_sigma2.argtypes = [P, C.c_int, C.c_int, C.c_int, f64p, C.c_int64, i32p, i32p, f64p, f64p, f64p]
        L.orc_set_sigma2.restype = C.c_int
        L.orc_diags.argtypes = [P, C.c_int]
        L.orc_diags.restype = C.c_double
        L.orc_nonequiv.argtypes = [P, i32p]
        L.orc_nonequiv.restype = C.c_int
        L.orc_count_m.argtypes = [P, C.c_int, C.c_int]
        L.orc_count_m.restype = C.c_int
        L.orc_j2_block.argtypes = [P, C.c_int, f64p]
        L.orc_formh_sym.argtypes = [P, i32p, i32p, i64p, f64p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_formh_sym.restype = C.c_int64
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def rspq(d1, d2):
    """Rspq_phase1+2 on two sorted 1-based position lists -> (nf, is, i2, i3, j2, j3)."""
    d1 = np.ascontiguousarray(d1, np.int32)
    d2 = np.ascontiguousarray(d2, np.int32)
    out = np.zeros(6, np.int32)
    lib().orc_rspq(d1.size, d1, d2, out)
    return tuple(int(x) for x in out)


def decode_entries(configs, core, shells):
    """Map CONF.INP shell codes to basis orbitals like Init (src/pconf.F90:506-530).

    shells: list of (n, l, kappa, 2j) in basis order.  Returns (nip, nq) arrays
    over all entries (core first), 1-based shell indices.
    """
    nip, nq = [], []
    for x in list(core) + [v for cfg in configs for v in cfg]:
        i = 1 if x >= 0 else -1
        d = abs(x) + 1e-14
        d = 10.0 * d
        nnj = int(d)
        d = 10.0 * (d - nnj)
        llj = int(d)
        jlj = 2 * llj + i
        kkj = -i * ((jlj + 1) // 2)
        d = 100.0 * (d - llj)
        q = int(d + 0.1)
        for ni, s in enumerate(shells):
            if s[0] == nnj and s[2] == kkj:
                break
        else:
            raise ValueError(f"no orbital for shell n={nnj} l={llj} kappa={kkj}")
        nip.append(ni + 1)
        nq.append(q)
    return np.array(nip, np.int32), np.array(nq, np.int32)


class Oracle:
    def __init__(self):
        self.L = lib()
        self.c = self.L.orc_create()

    def __del__(self):
        if getattr(self, "c", None):
            self.L.orc_free(self.c)
            self.c = None

    # -- set-up --------------------------------------------------------------
    def set_space(self, shells, nso, nip, nq, ne, mj, gj=0.0):
        sh = np.asarray(shells, np.int32).reshape(-1, 4)
        nn, ll, kk, jj = (np.ascontiguousarray(sh[:, i]) for i in range(4))
        self.shells = sh
        self.Ne, self.Mj, self.Nso = ne, mj, nso
        nc = self.L.orc_set_space(self.c, len(sh), nso, nn, ll, kk, jj, len(nip),
                                  np.ascontiguousarray(nip, np.int32), np.ascontiguousarray(nq, np.int32),
                                  ne, mj, gj)
        if nc < 0:
            raise ValueError("wrong number of electrons in a configuration")
        self.Nc = nc
        self.Nst = self.L.orc_nst(self.c)
        self.Npos = self.L.orc_npos(self.c)
        self.Nsu = self.L.orc_nsu(self.c)

    def enumerate(self):
        self.Nd = int(self.L.orc_enumerate(self.c))
        self.iarr = np.zeros((self.Nd, self.Ne), np.int32)
        self.ndc = np.zeros(self.Nc, np.int32)
        self.jz = np.zeros(self.Npos, np.int32)
        self.nh = np.zeros(self.Npos, np.int32)
        self.L.orc_get_space(self.c, _ptr(self.iarr), _ptr(self.ndc), _ptr(self.jz), _ptr(self.nh))
        return self.iarr, self.ndc

    def set_dets(self, iarr, ndc):
        self.iarr = np.ascontiguousarray(iarr, np.int32)
        self.ndc = np.ascontiguousarray(ndc, np.int32)
        self.Nd = self.iarr.shape[0]
        self.L.orc_set_dets(self.c, self.Nd, self.iarr.reshape(-1), self.ndc)

    def set_integrals(self, conf_int):
        ci = conf_int
        r2 = np.ascontiguousarray(np.asarray(ci.rint2, np.float64).reshape(-1))
        rc = self.L.orc_set_integrals(self.c, ci.nx, ci.rint1.size, np.ascontiguousarray(ci.rint1, np.float64),
                                      np.ascontiguousarray(ci.iint1, np.int32), ci.iint2.size, ci.kbrt, r2,
                                      np.ascontiguousarray(ci.iint2, np.int32),
                                      np.ascontiguousarray(ci.iint3, np.int32))
        if rc:
            raise ValueError("bad one-electron integral index")

    def set_gaunt(self, nlx, idx, gnt):
        self.L.orc_set_gaunt(self.c, nlx, idx.size, np.ascontiguousarray(idx, np.int32),
                             np.ascontiguousarray(gnt, np.float64))

    def nonequiv(self):
        """nonequiv_conf (src/csf.f90:454-536) on the configuration entries; call
        before enumerate().  Sets nc_neq (1-based class per configuration) and
        nconf_neq."""
        self.nc_neq = np.zeros(self.Nc, np.int32)
        self.nconf_neq = int(self.L.orc_nonequiv(self.c, self.nc_neq))
        return self.nc_neq, self.nconf_neq

    @classmethod
    def from_dir(cls, d, kcsf=None):
        """kcsf: None follows ci.in's kCSF; with kCSF > 0 the configuration
        classes are formed (and the entries reordered) before Dinit, as pconf
        does (src/pconf.F90:138)."""
        import sys
        sys.path.insert(0, os.path.dirname(_HERE))
        from pci_amd import formats as F
        inp = F.read_conf_inp(os.path.join(d, "CONF.INP"))
        _, shells = F.read_conf_dat(os.path.join(d, "CONF.DAT"))
        ci = F.read_conf_int(os.path.join(d, "CONF.INT"))
        nlx, _, idx, gnt = F.read_conf_gnt(os.path.join(d, "CONF.GNT"))
        cio = F.read_ci_in(os.path.join(d, "ci.in"))
        nip, nq = decode_entries(inp.configs, inp.core, shells)
        o = cls()
        o.inp, o.ci_in = inp, cio
        o.set_space(shells, inp.Nso, nip, nq, inp.Ne, int(2 * abs(inp.Jm) + 0.01), inp.Gj)
        o.kcsf = cio["KCSF"] if kcsf is None else kcsf
        if o.kcsf > 0:
            o.nonequiv()
        o.enumerate()
        o.set_integrals(ci)
        o.set_gaunt(nlx, idx, gnt)
        o.ksig = cio["KSIG"] if cio["KL"] == 2 else 0   # MBPT only with Kl = 2 (src/pconf.F90:345-356)
        if o.ksig:
            if o.ksig not in (1, 2):
                raise NotImplementedError(f"Ksig = {o.ksig}")
            sgc = F.read_sgc_con(os.path.join(d, "SGC.CON"), o.ksig)
            nmaxs, lmaxs, _, ent = sgc[:4]
            eps = F.read_conf_dat_eps(os.path.join(d, "CONF.DAT"), inp.Nso, o.Nsu)
            e0 = 0.0
            if cio["KDSIG"]:
                with open(os.path.join(d, "E0.in")) as f:
                    e0 = float(f.read().split()[0])
            o.set_sigma(1, cio["KDSIG"], nmaxs, lmaxs, ent, eps, e0, ci.nx, inp.Nso, len(shells))
            if o.ksig == 2:
                o.set_sigma2(sgc[4], F.read_scrc_con(os.path.join(d, "SCRC.CON")))
        return o

    def set_sigma(self, ksig, kdsig, nmaxs, lmaxs, entries, eps, e0, nx, nso, ns):
        """entries: (na, nb, R, D, E) in SGC.CON order; eps indexed by shell."""
        ind = np.array([nx * (na - nso - 1) + (nb - nso) for na, nb, _, _, _ in entries], np.int32)
        r = np.array([e[2] for e in entries], np.float64)
        dd = np.array([e[3] for e in entries], np.float64)
        ee = np.array([e[4] for e in entries], np.float64)
        ep = np.zeros(ns)
        ep[:min(ns, len(eps) - 1)] = np.asarray(eps, np.float64)[1:ns + 1]
        self.sigma = dict(ksig=ksig, kdsig=kdsig, nmaxs=nmaxs, lmaxs=lmaxs, ind=ind, r=r, d=dd, e=ee, eps=ep, e0=e0)
        rc = self.L.orc_set_sigma(self.c, ksig, kdsig, nmaxs, lmaxs, len(entries), ind, r, dd, ee, ep, e0)
        if rc != 0:
            raise ValueError("SGC.CON index out of range")

    def set_sigma2(self, scr, sc):
        """Ksig = 2: the screening coefficients (SGC.CON) and SCRC.CON."""
        self.sigma2 = dict(scr=np.asarray(scr, np.float64), ksym=sc.ksym, kmax=sc.kmax, nsum=sc.nsum,
                           i2=np.ascontiguousarray(sc.iint2s, np.int32), i3=np.ascontiguousarray(sc.iint3s, np.int32),
                           r=np.ascontiguousarray(sc.rint2s, np.float64), d=np.ascontiguousarray(sc.dint2s, np.float64),
                           e=np.ascontiguousarray(sc.eint2s, np.float64))
        g = self.sigma2
        self.L.orc_set_sigma2(self.c, sc.ksym, sc.kmax, sc.nsum, g["scr"], g["i2"].size, g["i2"], g["i3"], g["r"],
                              g["d"], g["e"])
        self.ksig = 2

    def diags(self, n):
        """FormD's Diag(n) (1-based) of a Kdsig run."""
        return self.L.orc_diags(self.c, n)

    # -- elements --------------------------------------------------------------
    def gaunt(self, k, j1, m1, j2, m2):
        return self.L.orc_gaunt(self.c, k, j1, m1, j2, m2)

    def hint(self, a, b):
        return self.L.orc_hint(self.c, a, b)

    def gint(self, a, b, c, d):
        v = self.L.orc_gint(self.c, a, b, c, d)
        m = self.L.orc_missing()
        if m:
            raise KeyError("integral absent" if m == 1 else "m-conservation violated")
        return v

    def element(self, n, k):
        nf = C.c_int(0)
        v = self.L.orc_element(self.c, n, k, C.byref(nf))
        return v, nf.value

    def formh(self, r0=1, r1=None):
        """FormH rows r0..r1 (1-based, inclusive) -> (rowptr, col, val, numzero, ncand)."""
        r1 = self.Nd if r1 is None else r1
        nrow = r1 - r0 + 1
        counts = np.zeros(nrow, np.int64)
        nz, ncand = C.c_int64(0), C.c_int64(0)
        total = self.L.orc_formh(self.c, r0, r1, _ptr(counts), None, None, C.byref(nz), C.byref(ncand))
        col = np.zeros(max(total, 1), np.int32)
        val = np.zeros(max(total, 1), np.float64)
        self.L.orc_formh(self.c, r0, r1, None, _ptr(col), _ptr(val), C.byref(nz), C.byref(ncand))
        if self.L.orc_missing():
            raise KeyError("integral absent while building H")
        if self.L.orc_missing_sigma():
            raise KeyError("HintS: integral is absent")
        rowptr = np.zeros(nrow + 1, np.int64)
        np.cumsum(counts, out=rowptr[1:])
        return rowptr, col[:total], val[:total], nz.value, ncand.value

    def formh_once(self, r0=1, r1=None):
        """formh evaluating each element once (growing buffers, like the
        reference's single pass) -> (rowptr, col, val, numzero, ncand)."""
        r1 = self.Nd if r1 is None else r1
        nrow = r1 - r0 + 1
        counts = np.zeros(nrow, np.int64)
        nz, ncand = C.c_int64(0), C.c_int64(0)
        cp, vp = C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
        total = self.L.orc_formh_grow(self.c, r0, r1, _ptr(counts), C.byref(cp), C.byref(vp),
                                      C.byref(nz), C.byref(ncand))
        try:
            col = np.ctypeslib.as_array(cp, (max(total, 1),))[:total].copy()
            val = np.ctypeslib.as_array(vp, (max(total, 1),))[:total].copy()
        finally:
            self.L.orc_free_buf(C.cast(cp, C.c_void_p))
            self.L.orc_free_buf(C.cast(vp, C.c_void_p))
        if self.L.orc_missing():
            raise KeyError("integral absent while building H")
        if self.L.orc_missing_sigma():
            raise KeyError("HintS: integral is absent")
        rowptr = np.zeros(nrow + 1, np.int64)
        np.cumsum(counts, out=rowptr[1:])
        return rowptr, col, val, nz.value, ncand.value

    def formj(self):
        counts = np.zeros(self.Nd, np.int64)
        total = self.L.orc_formj(self.c, _ptr(counts), None, None)
        col = np.zeros(max(total, 1), np.int32)
        val = np.zeros(max(total, 1), np.float64)
        self.L.orc_formj(self.c, None, _ptr(col), _ptr(val))
        rowptr = np.zeros(self.Nd + 1, np.int64)
        np.cumsum(counts, out=rowptr[1:])
        return rowptr, col[:total], val[:total]


def _jterms(self, cap=256):
    """Jterm: (Jt doubled J values ascending, Njt level counts)."""
    jt = np.zeros(cap, np.int32)
    njt = np.zeros(cap, np.int64)
    n = self.L.orc_jterms(self.c, _ptr(jt), _ptr(njt), cap)
    if n < 0:
        raise ValueError("too many J terms")
    return jt[:n].copy(), njt[:n].copy()


Oracle.jterms = _jterms


def calc_nd0(ndc, nc4, maxnd0):
    """calcNd0 (src/determinants.f90:22-45): (Nc1, Nd0) = longest prefix of the
    first Nc4 configurations holding at most MaxNd0 determinants."""
    n0 = maxnd0 + 1
    n1, nd0, ic1 = 0, 0, 0
    for ic in range(min(nc4, len(ndc))):
        n1 += int(ndc[ic])
        if n1 < n0:
            nd0, ic1 = n1, ic + 1
    return ic1, nd0


def jav(rowptr, col, val, x, nx=None):
    nd = rowptr.size - 1
    return lib().orc_jav(nd, np.ascontiguousarray(rowptr, np.int64), np.ascontiguousarray(col, np.int32),
                         np.ascontiguousarray(val, np.float64), np.ascontiguousarray(x, np.float64),
                         nd if nx is None else nx)


def mxmpy(rowptr, col, val, hmin, x, want_diag=False):
    nd = rowptr.size - 1
    y = np.zeros(nd)
    diag = np.zeros(nd) if want_diag else None
    lib().orc_mxmpy(nd, rowptr, col, val, hmin, np.ascontiguousarray(x, np.float64), y, _ptr(diag))
    return (y, diag) if want_diag else y


def mxmpy_rows(nd, r0, rowptr, col, val, hmin, x):
    """One rank's Mxmpy: rows r0 .. r0+len(rowptr)-2 (1-based) scattered into a
    full-length y (src/davidson.f90:617-634)."""
    y = np.zeros(nd)
    lib().orc_mxmpy_rows(nd, r0, rowptr.size - 1, np.ascontiguousarray(rowptr, np.int64),
                         np.ascontiguousarray(col, np.int32), np.ascontiguousarray(val, np.float64), hmin,
                         np.ascontiguousarray(x, np.float64), y)
    return y
