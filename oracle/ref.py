"""ctypes front end of oracle/_ref/libpconf_ref.so (TEST INFRASTRUCTURE ONLY).

The library holds the reference's own Fortran modules (determinants,
integrals, conf_init, wigner, ...) compiled from ud-pci/pCI's sources plus the
bind(C) driver in oracle/ref_driver.f90.  It keeps its state in Fortran module
globals, so one process can load one space: callers run it in a child process
(see ``run_in_child``).  It exists only where the reference sources were
present at build time; callers skip when it is absent.
"""
from __future__ import annotations

import ctypes as C
import multiprocessing as mp
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libpconf_ref.so")


def available() -> bool:
    return os.path.exists(PATH)


class Ref:
    def __init__(self, space_dir: str):
        self.L = C.CDLL(PATH)
        L = self.L
        L.ref_load.restype = C.c_int
        L.ref_gint.restype = C.c_double
        L.ref_gint.argtypes = [C.c_int] * 4
        L.ref_hint.restype = C.c_double
        L.ref_hint.argtypes = [C.c_int] * 2
        L.ref_gaunt.restype = C.c_double
        L.ref_gaunt.argtypes = [C.c_int] + [C.c_double] * 4
        L.ref_gaun.restype = C.c_double
        L.ref_gaun.argtypes = [C.c_int] + [C.c_double] * 4
        L.ref_formh.restype = C.c_int64
        L.ref_formh.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        cwd = os.getcwd()
        os.chdir(space_dir)
        try:
            rc = L.ref_load()
        finally:
            os.chdir(cwd)
        if rc:
            raise RuntimeError(f"ref_load failed with {rc}")
        dims = [C.c_int(0) for _ in range(5)]
        L.ref_dims(*[C.byref(d) for d in dims])
        self.Nd, self.Ne, self.Nc, self.Nst, self.Nsu = (d.value for d in dims)

    def jterms(self, cap=256):
        jt = np.zeros(cap, np.int32)
        njt = np.zeros(cap, np.int64)
        self.L.ref_jterms.restype = C.c_int
        self.L.ref_jterms.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        n = self.L.ref_jterms(jt.ctypes.data_as(C.c_void_p), njt.ctypes.data_as(C.c_void_p), cap)
        return jt[:n].copy(), njt[:n].copy()

    def space(self):
        iarr = np.zeros((self.Nd, self.Ne), np.int32)
        ndc = np.zeros(self.Nc, np.int32)
        jz = np.zeros(self.Nst, np.int32)
        nh = np.zeros(self.Nst, np.int32)
        self.L.ref_space(*(a.ctypes.data_as(C.c_void_p) for a in (iarr, ndc, jz, nh)))
        return iarr, ndc, jz, nh

    def rspq(self, d1, d2):
        d1 = np.ascontiguousarray(d1, np.int32)
        d2 = np.ascontiguousarray(d2, np.int32)
        out = np.zeros(6, np.int32)
        self.L.ref_rspq(*(a.ctypes.data_as(C.c_void_p) for a in (d1, d2, out)))
        return tuple(int(x) for x in out)

    def formh(self, r0, r1):
        counts = np.zeros(r1 - r0 + 1, np.int64)
        dummy_i = np.zeros(1, np.int32)
        dummy_f = np.zeros(1, np.float64)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        total = self.L.ref_formh(r0, r1, p(counts), 0, p(dummy_i), p(dummy_f))
        col = np.zeros(max(total, 1), np.int32)
        val = np.zeros(max(total, 1), np.float64)
        self.L.ref_formh(r0, r1, p(counts), total, p(col), p(val))
        rowptr = np.zeros(counts.size + 1, np.int64)
        np.cumsum(counts, out=rowptr[1:])
        return rowptr, col[:total], val[:total]


def _child(q, space_dir, fn, args):
    try:
        q.put(("ok", fn(Ref(space_dir), *args)))
    except BaseException as e:  # pragma: no cover - reported to parent
        q.put(("err", repr(e)))


def run_in_child(space_dir, fn, *args, timeout=600):
    """Run fn(Ref(space_dir), *args) in a fresh process and return its result."""
    ctx = mp.get_context("fork")
    q = ctx.Queue()
    p = ctx.Process(target=_child, args=(q, space_dir, fn, args))
    p.start()
    status, res = q.get(timeout=timeout)
    p.join(timeout=60)
    if status != "ok":
        raise RuntimeError(res)
    return res
