"""ctypes binding of libpci_gpu.so (the C ABI declared in include/pci_gpu.h).

There is no CPU fallback: if the library is missing or the GPU is absent the
calls raise.  Build the library with ``python -m pci_amd.build``.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpci_gpu.so")

PCI_DVD_FP32 = 1
PCI_DVD_LEARNED = 2
PCI_DVD_TRACE = 8
PCI_DVD_COUPLED = 16
PCI_DVD_JSEL = 32
PCI_DVD_NOADMIX = 64
PCI_BUILD_BLOCKS = 1

SYMBOLS = [
    "pci_gpu_create", "pci_gpu_destroy", "pci_gpu_last_error", "pci_gpu_set_space",
    "pci_gpu_set_integrals", "pci_gpu_set_sigma", "pci_gpu_set_sigma2", "pci_gpu_eigh32", "pci_gpu_build_h", "pci_gpu_build_info", "pci_gpu_get_csr", "pci_gpu_get_csr_rows", "pci_gpu_set_csr",
    "pci_gpu_build_j", "pci_gpu_get_jcsr", "pci_gpu_set_jterms", "pci_gpu_davidson", "pci_gpu_davidson_batch", "pci_gpu_spmm", "pci_gpu_spmm_noalloc", "pci_gpu_spmm_variant", "pci_gpu_spmm_set_variant",
    "pci_gpu_mlp_nparams", "pci_gpu_mlp_init", "pci_gpu_mlp_set_weights", "pci_gpu_mlp_get_weights",
    "pci_gpu_mlp_collect", "pci_gpu_mlp_train", "pci_gpu_mlp_grad", "pci_gpu_mlp_forward",
    "pci_gpu_mem_usage", "pci_gpu_csf_basis", "pci_gpu_csf_set_basis", "pci_gpu_csf_get_basis", "pci_gpu_build_h_csf", "pci_gpu_csf_unsym", "pci_gpu_davidson_ex",
]

# network layout (must match the flat parameter array of pci_gpu_mlp_*)
MLP_NF, MLP_IN, MLP_H, MLP_NB, MLP_NE = 8, 32, 128, 64, 16


class PciError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class DvdParams(C.Structure):
    _fields_ = [("nlv", C.c_int32), ("nd0", C.c_int32), ("crt4", C.c_double), ("n_it", C.c_int32),
                ("flags", C.c_int32), ("refine_every", C.c_int32), ("reserved", C.c_int32)]


class DvdStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("spmm_fp32", C.c_int32),
                ("spmm_fp64", C.c_int32), ("learned_used", C.c_int32), ("learned_rejected", C.c_int32),
                ("max_residual", C.c_double), ("t_total_ms", C.c_double), ("t_spmm_ms", C.c_double),
                ("t_init_ms", C.c_double), ("spmm_bytes", C.c_double), ("t_spmm32_ms", C.c_double),
                ("resid_hist", C.c_double * 512)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "resid_hist"}
        n = min(self.iterations, 512)
        d["resid_hist"] = list(self.resid_hist[:n])
        return d


# void (*)(void *user, int32 it, int32 nlv, int32 nd, const double *tk, const double *tj, const double *vecs)
CHECKPOINT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                            C.POINTER(C.c_double), C.POINTER(C.c_double))

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m pci_amd.build`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.pci_gpu_last_error.restype = C.c_char_p
        L.pci_gpu_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.pci_gpu_destroy.argtypes = [P]
        L.pci_gpu_set_space.argtypes = [P, C.c_int32, C.c_int32, P, P, P, C.c_int32, P, P, C.c_int32, C.c_int32,
                                        P, C.c_int32, P, C.c_int32, C.c_double]
        L.pci_gpu_set_integrals.argtypes = [P, C.c_int32, C.c_int32, P, P, C.c_int64, C.c_int32, P, P, P,
                                            C.c_int32, C.c_int32, P, P]
        L.pci_gpu_build_h.argtypes = [P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_double)]
        L.pci_gpu_get_csr.argtypes = [P, P, P, P]
        L.pci_gpu_build_info.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.pci_gpu_get_csr_rows.argtypes = [P, C.c_int32, C.c_int32, P, P, P, C.c_int64]
        L.pci_gpu_set_csr.argtypes = [P, C.c_int32, C.c_int64, P, P, P]
        L.pci_gpu_build_j.argtypes = [P, C.POINTER(C.c_int64)]
        L.pci_gpu_get_jcsr.argtypes = [P, P, P, P]
        L.pci_gpu_set_jterms.argtypes = [P, C.c_double, C.c_int32, P]
        L.pci_gpu_davidson.argtypes = [P, C.POINTER(DvdParams), P, P, P, C.POINTER(DvdStats)]
        L.pci_gpu_spmm.argtypes = [P, C.c_int32, C.c_int32, P, P, C.c_int32, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
        L.pci_gpu_spmm_noalloc.argtypes = [P, C.c_int32, C.c_int32, P, C.c_int32, P, C.c_int32]
        L.pci_gpu_mem_usage.argtypes = [P, C.POINTER(C.c_int64)]
        L.pci_gpu_set_sigma2.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, P, C.c_int64, P, P, P, P, P]
        L.pci_gpu_csf_basis.argtypes = [P, C.c_int32, P, P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.pci_gpu_csf_get_basis.argtypes = [P, P]
        L.pci_gpu_csf_set_basis.argtypes = [P, C.c_int32, P, P, P]
        L.pci_gpu_build_h_csf.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.pci_gpu_csf_unsym.argtypes = [P, C.c_int32, P, P]
        L.pci_gpu_davidson_ex.argtypes = [P, C.POINTER(DvdParams), P, P, C.c_int32, CHECKPOINT_FN, P, P, P, P,
                                          C.POINTER(DvdStats)]
        L.pci_gpu_spmm_variant.argtypes = [P, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
        L.pci_gpu_spmm_set_variant.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
        L.pci_gpu_mlp_nparams.argtypes = [C.POINTER(C.c_int64)]
        L.pci_gpu_mlp_init.argtypes = [P, C.c_uint64]
        L.pci_gpu_mlp_set_weights.argtypes = [P, P, C.c_int64]
        L.pci_gpu_mlp_get_weights.argtypes = [P, P, C.c_int64]
        L.pci_gpu_mlp_collect.argtypes = [P, C.POINTER(DvdParams), P, C.c_int32, C.c_int64, P, P, P, P,
                                          C.POINTER(C.c_int64)]
        L.pci_gpu_mlp_train.argtypes = [P, P, P, P, P, C.c_int64, C.c_int32, C.c_int32, C.c_float, P]
        L.pci_gpu_mlp_grad.argtypes = [P, P, P, P, P, C.c_int64, P, C.POINTER(C.c_float)]
        L.pci_gpu_mlp_forward.argtypes = [P, P, P, P, C.c_int64, P]
        for s in SYMBOLS:
            if s != "pci_gpu_last_error":
                getattr(L, s).restype = C.c_int
        _L = L
    return _L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise PciError(rc, lib().pci_gpu_last_error().decode(errors="replace"))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def mlp_nparams():
    n = C.c_int64()
    _check(lib().pci_gpu_mlp_nparams(C.byref(n)))
    return n.value


def davidson_batch(ctxs, params):
    """Diag4 of independent CI blocks at once (pci_gpu_davidson_batch).
    params: one dict per context with nlv, nd0, crt4, n_it[, flags].
    Returns one (tk, tj, vecs, stats) per block."""
    n = len(ctxs)
    prm = (DvdParams * max(n, 1))()
    st = (DvdStats * max(n, 1))()
    tks, tjs, vecs = [], [], []
    for b, (c, q) in enumerate(zip(ctxs, params)):
        prm[b] = DvdParams(nlv=q["nlv"], nd0=q["nd0"], crt4=q["crt4"], n_it=q["n_it"], flags=q.get("flags", 0),
                           refine_every=0)
        tks.append(np.zeros(q["nlv"]))
        tjs.append(np.zeros(q["nlv"]))
        vecs.append(np.zeros((q["nlv"], c.nd)))
    P = C.c_void_p
    hs = (P * max(n, 1))(*[c.h for c in ctxs])
    tk_p = (P * max(n, 1))(*[a.ctypes.data for a in tks])
    tj_p = (P * max(n, 1))(*[a.ctypes.data for a in tjs])
    v_p = (P * max(n, 1))(*[a.ctypes.data for a in vecs])
    _check(lib().pci_gpu_davidson_batch(n, hs, prm, tk_p, tj_p, v_p, st))
    return [(tks[b], tjs[b], vecs[b], st[b].as_dict()) for b in range(n)]


# Contexts still open when the interpreter exits are destroyed by an atexit
# handler, which runs before the HIP runtime, rocBLAS and any profiler
# attached to the process tear themselves down (their own exit handlers run
# after the interpreter's).  A finalizer that runs during interpreter
# shutdown does not release device resources: by then the order of
# teardown is no longer known.
_LIVE = weakref.WeakSet()


@atexit.register
def close_all():
    """Destroy every context that is still open (idempotent)."""
    for ctx in list(_LIVE):
        ctx.close()


class GpuContext:
    """One device context: space, integrals, resident H and Davidson.
    Close it with close() (or use it as a context manager); contexts left
    open are closed at interpreter exit."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        _check(lib().pci_gpu_create(device, C.byref(self.h)))
        self.nd = 0
        self.nd_space = 0
        _LIVE.add(self)

    def close(self):
        if self.h:
            lib().pci_gpu_destroy(self.h)
            self.h = C.c_void_p()
        _LIVE.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def set_space(self, ll, jj, kk, nso, jz, nh, iarr, ndc, mj, gj=0.0):
        ll, jj, kk, jz, nh, iarr, ndc = map(_i32, (ll, jj, kk, jz, nh, iarr, ndc))
        nd, ne = iarr.shape
        _check(lib().pci_gpu_set_space(self.h, ll.size, nso, _p(ll), _p(jj), _p(kk), jz.size, _p(jz), _p(nh),
                                       ne, nd, _p(iarr), ndc.size, _p(ndc), int(mj), float(gj)))
        self.nd = nd
        self.nd_space = nd

    def set_integrals(self, nx, rint1, iint1, kbrt, rint2, iint2, iint3, nlx, gidx, gnt):
        rint1, rint2, gnt = _f64(rint1), _f64(np.asarray(rint2).reshape(-1)), _f64(gnt)
        iint1, iint2, iint3, gidx = map(_i32, (iint1, iint2, iint3, gidx))
        _check(lib().pci_gpu_set_integrals(self.h, nx, rint1.size, _p(rint1), _p(iint1), iint2.size, kbrt,
                                           _p(rint2), _p(iint2), _p(iint3), nlx, gidx.size, _p(gidx), _p(gnt)))

    def build_h(self, flags=0):
        """FormH on the device; flags PCI_BUILD_BLOCKS builds the residual blocks only."""
        numh, nz, hmin = C.c_int64(), C.c_int64(), C.c_double()
        _check(lib().pci_gpu_build_h(self.h, flags, C.byref(numh), C.byref(nz), C.byref(hmin)))
        self.numh, self.numzero, self.hmin = numh.value, nz.value, hmin.value
        return self.numh, self.numzero, self.hmin

    def csf_basis(self, nc_neq, nconf_neq):
        """jbasis on the device -> (ndcs per class, ncsf, nccj)."""
        nc_neq = _i32(nc_neq)
        ndcs = np.zeros(nconf_neq, np.int32)
        ncsf, nccj = C.c_int32(), C.c_int64()
        _check(lib().pci_gpu_csf_basis(self.h, nconf_neq, _p(nc_neq), _p(ndcs), C.byref(ncsf), C.byref(nccj)))
        self.ncsf, self.nccj = ncsf.value, nccj.value
        return ndcs, self.ncsf, self.nccj

    def csf_set_basis(self, nc_neq, ndcs, ccj):
        """Install CSF coefficients computed elsewhere (ccj_module's layout)."""
        nc_neq, ndcs, ccj = _i32(nc_neq), _i32(ndcs), _f64(ccj)
        _check(lib().pci_gpu_csf_set_basis(self.h, ndcs.size, _p(nc_neq), _p(ndcs), _p(ccj)))
        conf_ncsf = ndcs[nc_neq - 1].astype(np.int64)
        self.ncsf, self.nccj = int(conf_ncsf.sum()), ccj.size

    def csf_get_basis(self):
        ccj = np.zeros(max(self.nccj, 1), np.float64)
        _check(lib().pci_gpu_csf_get_basis(self.h, _p(ccj)))
        return ccj[:self.nccj]

    def build_h_csf(self):
        """formh_sym on the device: H in the CSF basis becomes the resident H."""
        numh, nz, hmin = C.c_int64(), C.c_int64(), C.c_double()
        _check(lib().pci_gpu_build_h_csf(self.h, C.byref(numh), C.byref(nz), C.byref(hmin)))
        self.numh, self.numzero, self.hmin = numh.value, nz.value, hmin.value
        self.nd = self.ncsf
        return self.numh, self.numzero, self.hmin

    def csf_unsym(self, x):
        """CSF-basis vectors x[nvec, ncsf] -> determinant-basis vectors [nvec, nd]."""
        x = np.ascontiguousarray(np.atleast_2d(x), np.float64)
        y = np.zeros((x.shape[0], self.nd_space))
        _check(lib().pci_gpu_csf_unsym(self.h, x.shape[0], _p(x), _p(y)))
        return y

    def build_info(self):
        """Candidates evaluated, stored elements (both triangles) and wall ms of the last H build."""
        n, m, ms = C.c_int64(), C.c_int64(), C.c_double()
        _check(lib().pci_gpu_build_info(self.h, C.byref(n), C.byref(m), C.byref(ms)))
        return {"candidates": n.value, "nnz_full": m.value, "build_ms": ms.value}

    def get_csr(self):
        rp = np.zeros(self.nd + 1, np.int64)
        col = np.zeros(max(self.numh, 1), np.int32)
        val = np.zeros(max(self.numh, 1), np.float64)
        _check(lib().pci_gpu_get_csr(self.h, _p(rp), _p(col), _p(val)))
        return rp, col[:self.numh], val[:self.numh]

    def get_csr_rows(self, row1, nrows):
        """Lower-triangle rows row1 .. row1+nrows-1 (1-based): (rowptr from 0, col 1-based, val)."""
        rp = np.zeros(nrows + 1, np.int64)
        _check(lib().pci_gpu_get_csr_rows(self.h, row1, nrows, _p(rp), None, None, 0))
        n = int(rp[-1])
        col = np.zeros(max(n, 1), np.int32)
        val = np.zeros(max(n, 1), np.float64)
        _check(lib().pci_gpu_get_csr_rows(self.h, row1, nrows, _p(rp), _p(col), _p(val), n))
        return rp, col[:n], val[:n]

    def set_csr(self, rowptr, col, val):
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        col, val = _i32(col), _f64(val)
        nd = rowptr.size - 1
        _check(lib().pci_gpu_set_csr(self.h, nd, int(rowptr[-1]), _p(rowptr), _p(col), _p(val)))
        self.nd, self.numh = nd, int(rowptr[-1])

    def build_j(self):
        n = C.c_int64()
        _check(lib().pci_gpu_build_j(self.h, C.byref(n)))
        self.numj = n.value
        return self.numj

    def get_jcsr(self):
        rp = np.zeros(self.nd + 1, np.int64)
        col = np.zeros(max(self.numj, 1), np.int32)
        val = np.zeros(max(self.numj, 1), np.float64)
        _check(lib().pci_gpu_get_jcsr(self.h, _p(rp), _p(col), _p(val)))
        return rp, col[:self.numj], val[:self.numj]

    def set_sigma(self, ksig, kdsig, nmaxs, lmaxs, iint1s, rsig, dsig, esig, eps, diag=None):
        """MBPT one-electron effective integrals (Ksig = 1): SGC.CON tables,
        Eps(1..Ns) and, with kdsig, FormD's Diag(1..Nd)."""
        iint1s, rsig, dsig, esig, eps = _i32(iint1s), _f64(rsig), _f64(dsig), _f64(esig), _f64(eps)
        dg = _f64(diag) if diag is not None else None
        _check(lib().pci_gpu_set_sigma(self.h, int(ksig), int(kdsig), int(nmaxs), int(lmaxs), iint1s.size,
                                       _p(iint1s), _p(rsig), _p(dsig), _p(esig), _p(eps),
                                       _p(dg) if dg is not None else None))

    def set_sigma2(self, ksym, kmax, nsum, scr, iint2s, iint3s, rsig2, dsig2, esig2):
        """Ksig = 2: SGC.CON's screening coefficients and SCRC.CON's tables."""
        scr, i2, i3 = _f64(scr), _i32(iint2s), _i32(iint3s)
        r, d, e = _f64(rsig2), _f64(dsig2), _f64(esig2)
        _check(lib().pci_gpu_set_sigma2(self.h, int(ksym), int(kmax), int(nsum), _p(scr), i2.size, _p(i2), _p(i3),
                                        _p(r), _p(d), _p(e)))

    def eigh32(self, a):
        """Dense symmetric FP32 eigensolver of the fast mode (eigenvalues
        ascending, eigenvectors as columns, the path taken: 1 cooperative)."""
        a = np.array(a, np.float32, order="F", copy=True)
        n = a.shape[0]
        w = np.zeros(n, np.float32)
        path = C.c_int32(-1)
        _check(lib().pci_gpu_eigh32(self.h, n, a.ctypes.data_as(C.c_void_p), _p(w), C.byref(path)))
        return w, a, path.value

    def set_jterms(self, xj_av, jt):
        jt = _i32(jt)
        _check(lib().pci_gpu_set_jterms(self.h, float(xj_av), jt.size, _p(jt)))

    def spmm(self, x, precision=64, reps=1):
        x = _f64(x)
        if x.ndim == 1:
            x = x[:, None]
        nd, nvec = x.shape
        y = np.zeros_like(x)
        ms, nb = C.c_double(), C.c_double()
        _check(lib().pci_gpu_spmm(self.h, nvec, precision, _p(x), _p(y), reps, C.byref(ms), C.byref(nb)))
        return y, ms.value, nb.value

    def spmm_noalloc(self, nvec, precision, x_ptr, ldx, y_ptr, ldy):
        """pci_gpu_spmm_noalloc on device arrays given by address (row-major, padded)."""
        _check(lib().pci_gpu_spmm_noalloc(self.h, nvec, precision, C.c_void_p(x_ptr), ldx, C.c_void_p(y_ptr), ldy))

    def spmm_variant(self, nvec, precision):
        """Kernel shape the autotuner picked for (nvec, precision), '' if not yet tuned."""
        buf = C.create_string_buffer(64)
        _check(lib().pci_gpu_spmm_variant(self.h, nvec, precision, buf, 64))
        return buf.value.decode()

    def spmm_set_variant(self, nvec, precision, variant=-1):
        """Pin a kernel shape for (nvec, precision); returns the number of shapes."""
        n = C.c_int32()
        _check(lib().pci_gpu_spmm_set_variant(self.h, nvec, precision, variant, C.byref(n)))
        return n.value

    def davidson_ex(self, nlv, nd0, crt4, n_it, flags=0, tk0=None, x0=None, every=0, checkpoint=None):
        """Diag4 with a restart from x0[nlv, nd] (and energies tk0) and a
        checkpoint callable(it, tk, tj, vecs) every `every` iterations."""
        prm = DvdParams(nlv=nlv, nd0=nd0, crt4=crt4, n_it=n_it, flags=flags, refine_every=0)
        tk, tj = np.zeros(nlv), np.zeros(nlv)
        vecs = np.zeros((nlv, self.nd))
        st = DvdStats()
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, np.float64).reshape(nlv, self.nd)
        if tk0 is not None:
            tk0 = np.ascontiguousarray(tk0, np.float64)

        def _cb(user, it, n, nd, ptk, ptj, pv):
            if checkpoint is not None:
                checkpoint(it, np.ctypeslib.as_array(ptk, (n,)).copy(), np.ctypeslib.as_array(ptj, (n,)).copy(),
                           np.ctypeslib.as_array(pv, (n, nd)).copy())

        cb = CHECKPOINT_FN(_cb)
        _check(lib().pci_gpu_davidson_ex(self.h, C.byref(prm), _p(tk0), _p(x0), every, cb, None, _p(tk), _p(tj),
                                         _p(vecs), C.byref(st)))
        return tk, tj, vecs, st.as_dict()

    def davidson(self, nlv, nd0, crt4, n_it, flags=0, refine_every=0, vectors=True):
        """Diag4; with vectors=False the eigenvectors are not copied to the
        host (vecs is returned as None)."""
        prm = DvdParams(nlv=nlv, nd0=nd0, crt4=crt4, n_it=n_it, flags=flags, refine_every=refine_every)
        tk = np.zeros(nlv)
        tj = np.zeros(nlv)
        vecs = np.zeros((nlv, self.nd)) if vectors else None
        st = DvdStats()
        _check(lib().pci_gpu_davidson(self.h, C.byref(prm), _p(tk), _p(tj), _p(vecs), C.byref(st)))
        return tk, tj, vecs, st.as_dict()

    # ---- learned correction (MLP)
    def mlp_init(self, seed=0):
        _check(lib().pci_gpu_mlp_init(self.h, int(seed)))

    def mlp_set_weights(self, w):
        w = _f32(w)
        _check(lib().pci_gpu_mlp_set_weights(self.h, _p(w), w.size))

    def mlp_get_weights(self):
        w = np.zeros(mlp_nparams(), np.float32)
        _check(lib().pci_gpu_mlp_get_weights(self.h, _p(w), w.size))
        return w

    def mlp_collect(self, nlv, nd0, crt4, n_it, xref, stride=1, cap=1 << 20, flags=0):
        """Training samples from a Davidson run against converged vectors xref (nlv x nd)."""
        prm = DvdParams(nlv=nlv, nd0=nd0, crt4=crt4, n_it=n_it, flags=flags, refine_every=0)
        xref = _f64(xref)
        assert xref.shape == (nlv, self.nd)
        x = np.zeros((cap, MLP_NF), np.float32)
        b = np.zeros(cap, np.int32)
        r = np.zeros(cap, np.float32)
        t = np.zeros(cap, np.float32)
        n = C.c_int64()
        _check(lib().pci_gpu_mlp_collect(self.h, C.byref(prm), _p(xref), int(stride), int(cap), _p(x), _p(b),
                                         _p(r), _p(t), C.byref(n)))
        m = n.value
        return x[:m].copy(), b[:m].copy(), r[:m].copy(), t[:m].copy()

    def mlp_train(self, x, bucket, r, target, epochs=10, batch=4096, lr=1e-3):
        x, r, target = _f32(x), _f32(r), _f32(target)
        bucket = _i32(bucket)
        hist = np.zeros(epochs, np.float32)
        _check(lib().pci_gpu_mlp_train(self.h, _p(x), _p(bucket), _p(r), _p(target), r.size, int(epochs),
                                       int(batch), float(lr), _p(hist)))
        return hist

    def mlp_grad(self, x, bucket, r, target):
        x, r, target = _f32(x), _f32(r), _f32(target)
        bucket = _i32(bucket)
        g = np.zeros(mlp_nparams(), np.float32)
        loss = C.c_float()
        _check(lib().pci_gpu_mlp_grad(self.h, _p(x), _p(bucket), _p(r), _p(target), r.size, _p(g),
                                      C.byref(loss)))
        return g, loss.value

    def mlp_forward(self, x, bucket, r):
        x, r = _f32(x), _f32(r)
        bucket = _i32(bucket)
        t = np.zeros(r.size, np.float32)
        _check(lib().pci_gpu_mlp_forward(self.h, _p(x), _p(bucket), _p(r), r.size, _p(t)))
        return t

    def mem_usage(self):
        b = C.c_int64()
        _check(lib().pci_gpu_mem_usage(self.h, C.byref(b)))
        return b.value
