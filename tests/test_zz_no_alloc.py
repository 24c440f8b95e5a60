"""The allocation-free SpMM entry point (pci_gpu_spmm_noalloc) against
pci_gpu_spmm on the same matrix and block: two calls in a row on one GPU,
the second with a smaller input (C2, then t3), both precisions.  The device
operands are allocated by the test through the HIP runtime the library
itself uses; the entry point allocates nothing and keeps its partial sums in
the library's scratch array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so.7")  # the runtime libpci_gpu.so is linked against (already loaded)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    for f in ("hipMalloc", "hipFree", "hipMemcpy"):
        getattr(hip, f).restype = C.c_int
    return hip


class DeviceBlock:
    """A zero-padded row-major nd x ld block on the device (test-side allocation)."""

    def __init__(self, hip, host):
        self.hip, self.host = hip, np.ascontiguousarray(host)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), self.host.nbytes) == 0
        assert hip.hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _check(ctx, nd, nvec, prec, seed):
    hip = _hip()
    x = np.random.default_rng(seed).standard_normal((nd, nvec))
    y_ref, _, _ = ctx.spmm(x, prec)               # existing entry point (chooses and builds the kernel)
    vw = 4 if prec == 32 else 2
    ld = (nvec + vw - 1) // vw * vw
    dt = np.float32 if prec == 32 else np.float64
    xh = np.zeros((nd, ld), dt)
    xh[:, :nvec] = x
    X = DeviceBlock(hip, xh)
    Y = DeviceBlock(hip, np.full((nd, ld), np.nan, dt))
    try:
        ctx.spmm_noalloc(nvec, prec, X.p.value, ld, Y.p.value, ld)
        y = Y.read()[:, :nvec].astype(np.float64)
    finally:
        X.free()
        Y.free()
    # the same kernel on the same operands: the same bits (pci_gpu_spmm rounds
    # an FP32 product to float before returning it as double)
    np.testing.assert_array_equal(y, y_ref)
    print(f"noalloc nd={nd} nvec={nvec} fp{prec}: equal ({ctx.spmm_variant(nvec, prec)})", flush=True)


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_matches_spmm_twice_second_smaller(prec):
    from pci_amd import host
    from tests.helpers import space_dir
    print(f"[case] noalloc fp{prec}: C2 then t3", flush=True)
    sp1, c1 = host.load(space_dir("c2"))
    sp2, c2 = host.load(space_dir("t3"))
    try:
        c1.build_h()
        c2.build_h()
        _check(c1, sp1.Nd, 10, prec, 1)
        _check(c2, sp2.Nd, 6, prec, 2)
    finally:
        c1.close()
        c2.close()


def test_noalloc_refuses_what_would_allocate():
    """Before any product of a width the kernel is not chosen (and its copy
    not built): the entry point returns PCI_E_ARG instead of tuning; FP32
    before the FP32 values exist likewise."""
    from pci_amd import _lib, host
    from tests.helpers import space_dir
    print("[case] noalloc refusals", flush=True)
    hip = _hip()
    sp, c = host.load(space_dir("t3"))
    X = Y = None
    try:
        c.build_h()
        X = DeviceBlock(hip, np.zeros((sp.Nd, 8)))
        Y = DeviceBlock(hip, np.zeros((sp.Nd, 8)))
        with pytest.raises(_lib.PciError) as e:
            c.spmm_noalloc(8, 64, X.p.value, 8, Y.p.value, 8)
        assert e.value.code == -1 and "chosen" in str(e.value)
        with pytest.raises(_lib.PciError) as e:
            c.spmm_noalloc(8, 32, X.p.value, 8, Y.p.value, 8)
        assert e.value.code == -1
    finally:
        for b in (X, Y):
            if b is not None:
                b.free()
        c.close()
