"""Build the native parts in-tree.

* ``pci_amd/libpci_gpu.so``  -- HIP library for gfx950 (the C ABI of include/pci_gpu.h)
* ``pci_amd/host/``          -- Fortran host: libpci_host.so and the pconf_gpu program

Run ``python -m pci_amd.build`` (or ``__graft_entry__.build()``).  Objects go
to ``pci_amd/_build/``; every call recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpci_gpu.so")
HOSTLIB = os.path.join(HERE, "libpci_host.so")
PROG = os.path.join(HERE, "pconf_gpu")

ARCH = os.environ.get("PCI_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("PCI_FC", "/opt/rocm/bin/amdflang")

# Per-file extra flags: the Slater-Condon elements must not be contracted
# into FMAs (pconf's exact-zero filtering depends on the rounded sums).
UNITS = {
    "context.hip": [],
    "formh.hip": ["-ffp-contract=off"],
    "spmm.hip": [],
    "davidson.hip": [],
    "eigh.hip": [],
    "mlp.hip": [],
    "csf.hip": ["-ffp-contract=off"],
}

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-Wno-unused-result"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout[-6000:]}")
    return r.stdout


def build_gpu(jobs: int = 8, incremental: bool = False) -> str:
    """Compile every unit (incremental: only units whose object is older than
    the unit or a header) and link libpci_gpu.so."""
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    headers.append(os.path.join(ROOT, "include", "pci_gpu.h"))
    newest_header = max(os.path.getmtime(h) for h in headers)

    def compile_unit(item):
        name, extra = item
        obj = os.path.join(BUILD, name.replace(".hip", ".o"))
        src = os.path.join(CSRC, name)
        if incremental and os.path.exists(obj) and \
                os.path.getmtime(obj) > max(os.path.getmtime(src), newest_header):
            return obj
        _run([HIPCC, *COMMON, *extra, "-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(jobs, len(UNITS))) as ex:
        objs = list(ex.map(compile_unit, UNITS.items()))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_host() -> str:
    """Fortran host (amdflang): module + shared library + pconf_gpu program."""
    if not os.path.isdir(HOST) or not any(f.endswith(".f90") for f in os.listdir(HOST)):
        return ""
    hb = os.path.join(BUILD, "host")
    os.makedirs(hb, exist_ok=True)
    srcs = ["pci_gpu_api.f90", "pconf_host.f90"]
    objs = []
    for s in srcs:
        o = os.path.join(hb, s.replace(".f90", ".o"))
        _run([FC, "-O2", "-fPIC", "-c", os.path.join(HOST, s), "-o", o, "-module-dir", hb], cwd=hb)
        objs.append(o)
    tmp = HOSTLIB + ".tmp"
    _run([FC, "-shared", "-o", tmp, *objs, LIB, f"-Wl,-rpath,{HERE}", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, HOSTLIB)
    mo = os.path.join(hb, "pconf_gpu.o")
    _run([FC, "-O2", "-c", os.path.join(HOST, "pconf_gpu.f90"), "-o", mo, "-module-dir", hb], cwd=hb)
    _run([FC, "-o", PROG, mo, *objs, LIB, f"-Wl,-rpath,{HERE}", "-Wl,-rpath,/opt/rocm/lib"])
    return HOSTLIB


def build_oracle() -> None:
    """The CPU oracle (test infrastructure) and, where the reference sources
    are present, oracle/_ref built from them."""
    od = os.path.join(ROOT, "oracle")
    _run(["make", "-C", od, "liboracle.so"])
    if os.path.isdir("/root/reference/src") and shutil.which("amdflang"):
        _run(["make", "-C", od, "ref"])


def build_all() -> None:
    build_gpu()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"gpu": build_gpu, "gpu-incremental": lambda: build_gpu(incremental=True), "host": build_host,
         "oracle": build_oracle, "all": build_all}[w]()
    print("built")
