"""Python handle on the Fortran host (libpci_host.so, pci_amd/host/*.f90).

The Fortran module ``pconf_host`` reads pconf's inputs and enumerates the
determinant space exactly as pconf's Input/Init/Rint/Dinit/Det_List do; it
keeps the space in module state (one space per process, like pconf).  This
wrapper exposes it to Python for the benchmark and the tests and hands the
space to a device context through the Fortran ``upload`` routine, which is
the ISO_C_BINDING call path of the pconf_gpu program.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_PATH = os.path.join(HERE, "libpci_host.so")
PROG_PATH = os.path.join(HERE, "pconf_gpu")

_H = None

DIM_KEYS = ["Nd", "Ne", "Nc", "Nst", "Npos", "Ns", "Nso", "Nsu", "Nlv", "Nd0", "N_it", "Kl", "Kw", "Mj", "Kbrt",
            "Ngint", "Nhint", "Ngaunt", "KXIJ", "Kv", "Nc1", "Kl4"]
PAR_KEYS = ["Crt4", "Gj", "Z", "Jm", "Ecore", "J"]


def host():
    global _H
    if _H is None:
        if not os.path.exists(HOST_PATH):
            raise RuntimeError(f"{HOST_PATH} is missing: run `python -m pci_amd.build`")
        _lib.lib()  # the host library links against libpci_gpu.so
        L = C.CDLL(HOST_PATH)
        L.pconf_load.argtypes = [C.c_char_p]
        L.pconf_load.restype = C.c_int
        L.pconf_dims.argtypes = [C.c_void_p]
        L.pconf_params.argtypes = [C.c_void_p]
        L.pconf_space.argtypes = [C.c_void_p] * 7
        L.pconf_upload.argtypes = [C.c_void_p]
        L.pconf_upload.restype = C.c_int
        L.pconf_error.argtypes = [C.c_char_p, C.c_int]
        L.pconf_jterms.argtypes = [C.c_void_p, C.c_int]
        L.pconf_jterms.restype = C.c_int
        L.pconf_csf_classes.argtypes = [C.c_void_p]
        L.pconf_csf_classes.restype = C.c_int
        L.pconf_write_triples.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.pconf_write_triples.restype = None
        L.pconf_set_e0.argtypes = [C.c_double]
        L.pconf_set_e0.restype = None
        L.pconf_sigma_dims.argtypes = [C.c_void_p]
        L.pconf_sigma_tables.argtypes = [C.c_void_p] * 6
        _H = L
    return _H


def write_conf_hij(path: str, rowptr: np.ndarray, col: np.ndarray, val: np.ndarray) -> None:
    """Legacy CONF.HIJ / CONF.JJJ triple file of a lower-triangle CSR, written
    by the Fortran host (the writer pconf_gpu uses with PCI_CONF_HIJ=1)."""
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    nd = rowptr.size - 1
    if col.size < rowptr[-1] or val.size < rowptr[-1]:
        raise ValueError("col/val shorter than rowptr[-1]")
    host().pconf_write_triples(path.encode(), nd, rowptr.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p),
                               val.ctypes.data_as(C.c_void_p))


def _err():
    buf = C.create_string_buffer(512)
    host().pconf_error(buf, 512)
    return buf.value.decode(errors="replace")


class Space:
    """A determinant space loaded by the Fortran host from a pconf directory."""

    def __init__(self, directory: str, e0: float | None = None):
        """e0: the valence energy of a Kdsig = 1 run (pconf reads it from the
        terminal); by default the first number of the directory's E0.in."""
        L = host()
        e0f = os.path.join(directory, "E0.in")
        if e0 is None and os.path.exists(e0f):
            with open(e0f) as f:
                e0 = float(f.read().split()[0])
        if e0 is not None:
            L.pconf_set_e0(float(e0))
        rc = L.pconf_load(os.path.abspath(directory).encode())
        if rc != 0:
            raise RuntimeError(f"pconf_load({directory}): {_err()}")
        d = np.zeros(24, np.int64)
        L.pconf_dims(d.ctypes.data_as(C.c_void_p))
        self.dims = {k: int(v) for k, v in zip(DIM_KEYS, d)}
        r = np.zeros(8)
        L.pconf_params(r.ctypes.data_as(C.c_void_p))
        self.params = {k: float(v) for k, v in zip(PAR_KEYS, r)}
        for k, v in self.dims.items():
            setattr(self, k, v)
        self.Crt4 = self.params["Crt4"]
        self.directory = directory

    def sigma(self):
        """The MBPT tables the host read: dict(ksig, kdsig, nmaxs, lmaxs, iint1s,
        rsig, dsig, esig, eps (by shell, 1..Ns), diag (FormD, Kdsig only))."""
        L = host()
        dm = np.zeros(5, np.int32)
        L.pconf_sigma_dims(dm.ctypes.data_as(C.c_void_p))
        ksig, kdsig, nmaxs, lmaxs, nh = (int(v) for v in dm)
        ii = np.zeros(max(nh, 1), np.int32)
        r, dd, e = (np.zeros(max(nh, 1)) for _ in range(3))
        eps = np.zeros(self.Ns)
        diag = np.zeros(self.Nd)
        L.pconf_sigma_tables(*(a.ctypes.data_as(C.c_void_p) for a in (ii, r, dd, e, eps, diag)))
        return dict(ksig=ksig, kdsig=kdsig, nmaxs=nmaxs, lmaxs=lmaxs, iint1s=ii[:nh], rsig=r[:nh], dsig=dd[:nh],
                    esig=e[:nh], eps=eps, diag=diag if kdsig else None)

    def arrays(self):
        """(iarr (Nd, Ne), ndc, jz, nh, ll, jj, kk) as the host holds them."""
        iarr = np.zeros((self.Nd, self.Ne), np.int32)
        ndc = np.zeros(self.Nc, np.int32)
        jz = np.zeros(self.Npos, np.int32)
        nh = np.zeros(self.Npos, np.int32)
        ll = np.zeros(self.Ns, np.int32)
        jj = np.zeros(self.Ns, np.int32)
        kk = np.zeros(self.Ns, np.int32)
        host().pconf_space(*(a.ctypes.data_as(C.c_void_p) for a in (iarr, ndc, jz, nh, ll, jj, kk)))
        return iarr, ndc, jz, nh, ll, jj, kk

    def csf_classes(self):
        """(nc_neq, nconf_neq): the configuration classes of a kCSF > 0 space
        (nonequiv_conf), (None, 0) otherwise."""
        nc_neq = np.zeros(max(self.Nc, 1), np.int32)
        n = host().pconf_csf_classes(nc_neq.ctypes.data_as(C.c_void_p))
        return (nc_neq[:self.Nc].copy() if n > 0 else None), int(n)

    def jterms(self):
        """Doubled J values of the space, ascending (Jterm)."""
        jt = np.zeros(256, np.int32)
        n = host().pconf_jterms(jt.ctypes.data_as(C.c_void_p), 256)
        if n < 0:
            raise RuntimeError("more than 256 J terms")
        return jt[:n].copy()

    def upload(self, ctx: "_lib.GpuContext"):
        rc = host().pconf_upload(ctx.h)
        if rc != 0:
            raise RuntimeError(f"pconf_upload: {_err()}")
        ctx.nd = self.Nd
        if self.Kv == 3:
            ctx.set_jterms(self.params["J"], self.jterms())


def load(directory: str, device: int = 0):
    """Load a pconf directory through the Fortran host and upload it to a new
    device context.  Returns (space, context)."""
    sp = Space(directory)
    ctx = _lib.GpuContext(device)
    sp.upload(ctx)
    return sp, ctx
