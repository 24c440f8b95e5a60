"""Write the golden fixtures from the reference's own Fortran modules.

Runs in the build container only (needs oracle/_ref, built from the sources
under /root/reference by oracle/Makefile).  For each seeded space it stores
what the compiled reference computes: the determinant list (Iarr, Ndc), the
lower-triangle CSR of FormH over a row range (row pointers, 1-based columns,
FP64 values as raw bits), and, for spaces small enough, the eigenvalues of
that H from a dense LAPACK eigensolve.  Inputs are not stored: the spaces
are regenerated from pci_amd.gen presets, and the SHA-256 of the generated
files is recorded so that generator drift is detected.

    python tests/golden/make_golden.py [preset ...]   (default: every case)
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# (preset, last FormH row, dense eigenvalues?)
CASES = [("tiny2", None, True), ("tiny3", None, True), ("tiny3d", None, True), ("c1", 240, False),
         ("t3", 300, False), ("t3b", 300, False), ("t4", 200, False), ("tiny2g", None, True), ("t4g", 200, False)]
FILES = ["CONF.INP", "ci.in", "CONF.DAT", "CONF.INT", "CONF.GNT"]


def input_digest(d):
    h = hashlib.sha256()
    for f in FILES:
        with open(os.path.join(d, f), "rb") as fh:
            h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()


def _collect(r, last):
    iarr, ndc, jz, nh = r.space()
    rp, col, val = r.formh(1, last or r.Nd)
    return iarr, ndc, rp, col, val


def main():
    from oracle import ref
    from tests.helpers import full_from_lower, space_dir
    if not ref.available():
        raise SystemExit("oracle/_ref is not built (needs the reference sources)")
    only = set(sys.argv[1:])
    for name, last, dense in CASES:
        if only and name not in only:
            continue
        d = space_dir(name)
        iarr, ndc, rp, col, val = ref.run_in_child(d, _collect, last)
        out = dict(name=name, digest=input_digest(d), iarr=iarr, ndc=ndc, rowptr=rp, col=col,
                   val_bits=val.view(np.int64), last_row=np.int64(last or iarr.shape[0]))
        if dense:
            H = full_from_lower(rp, col, val).toarray()
            out["eigvals"] = np.linalg.eigvalsh(H)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: Nd={iarr.shape[0]} rows 1..{out['last_row']} NumH={rp[-1]} -> {os.path.getsize(path)} B")


if __name__ == "__main__":
    main()
