"""Write the CPU-oracle fixtures of the measured workloads (C2, C4, C5).

Runs in the build container (the CPU oracle, oracle/liboracle.so, is built
from this repository's C restatement of pconf's FormH).  For each space it
stores FormH rows spread evenly over the whole determinant list (first and
last row included): for each sampled row its lower-triangle columns
(1-based) and FP64 values as raw bits, as pconf's FormH computes them
(src/pconf.F90:1226-1356).  For C2 it also stores the 10 lowest eigenvalues
of H from the oracle's full FormH and a Lanczos solve (scipy eigsh on the
symmetric CSR), the eigenvalues the GPU Davidson is checked against.
Inputs are regenerated from the pci_amd.gen presets; the SHA-256 of the
generated files is recorded so that generator drift is detected.

With csf_<name> (csf_c2, csf_c4) it writes scale_csf_<name>.npz, the CSF
basis (kCSF = 1) of the space: the configuration classes and the oracle's
CSF coefficients (jbasis), and for configurations spread over the space
(first and last with CSFs included) the rows of H in that basis as
formh_sym stores them (columns, FP64 bits); for C2 also NumH, numzero, Hmin
of the whole H and its 10 lowest eigenvalues (Lanczos).

    python tests/golden/make_scale_golden.py [c2] [c4] [c5] [csf_c2] [csf_c4]
"""
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NROWS = 64          # sampled rows per space
LANCZOS = {"c2": 10}  # spaces whose lowest eigenvalues are written, and how many


def sample_rows(nd, n=NROWS):
    return np.unique(np.linspace(1, nd, n).round().astype(np.int64))


def _rows(d, r0, r1):
    from oracle.oracle import Oracle
    o = Oracle.from_dir(d)
    rp, col, val, nz, _ = o.formh(int(r0), int(r1))
    return rp, col, val, nz


def full_formh(d, nd, workers=8):
    """The whole lower-triangle FormH, row blocks in parallel processes."""
    edges = np.linspace(1, nd + 1, 8 * workers + 1).round().astype(np.int64)
    parts = []
    with ProcessPoolExecutor(workers) as ex:
        futs = [ex.submit(_rows, d, a, b - 1) for a, b in zip(edges[:-1], edges[1:]) if b > a]
        for f in futs:
            parts.append(f.result())
    counts = np.concatenate([np.diff(p[0]) for p in parts])
    rp = np.zeros(nd + 1, np.int64)
    np.cumsum(counts, out=rp[1:])
    col = np.concatenate([p[1] for p in parts])
    val = np.concatenate([p[2] for p in parts])
    return rp, col, val, sum(p[3] for p in parts)


def lowest_eigenvalues(rp, col, val, k):
    import scipy.sparse as sp
    from scipy.sparse.linalg import eigsh
    from tests.helpers import full_from_lower
    H = full_from_lower(rp, col, val)
    w = eigsh(H, k=k, which="SA", tol=1e-13, ncv=max(4 * k, 40), return_eigenvectors=False)
    return np.sort(w)


def make(name):
    from oracle.oracle import Oracle
    from tests.golden.make_golden import input_digest
    from tests.helpers import space_dir
    d = space_dir(name)
    o = Oracle.from_dir(d)
    nd = o.Nd
    rows = sample_rows(nd)
    cols, bits, counts = [], [], []
    t = time.time()
    for r in rows:
        rp, col, val, _, _ = o.formh(int(r), int(r))
        cols.append(col)
        bits.append(val.view(np.int64))
        counts.append(col.size)
    out = dict(name=name, digest=input_digest(d), nd=np.int64(nd), rows=rows,
               counts=np.array(counts, np.int64), col=np.concatenate(cols), val_bits=np.concatenate(bits))
    print(f"{name}: Nd={nd}, {rows.size} rows, {sum(counts)} elements in {time.time() - t:.1f} s", flush=True)
    if name in LANCZOS:
        t = time.time()
        rp, col, val, nz = full_formh(d, nd)
        print(f"{name}: full FormH NumH={rp[-1]} numzero={nz} in {time.time() - t:.1f} s", flush=True)
        t = time.time()
        w = lowest_eigenvalues(rp, col, val, LANCZOS[name])
        out.update(numh=np.int64(rp[-1]), numzero=np.int64(nz), hmin=np.float64(val.min()), eigvals=w)
        print(f"{name}: Lanczos {w} in {time.time() - t:.1f} s", flush=True)
    path = os.path.join(HERE, f"scale_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"-> {path} ({os.path.getsize(path)} B)", flush=True)


NCONF = 24  # sampled configurations per CSF fixture


def make_csf(name):
    from oracle import csf
    from oracle.oracle import Oracle
    from tests.golden.make_golden import input_digest
    from tests.helpers import full_from_lower, space_dir
    d = space_dir(name)
    t = time.time()
    o = Oracle.from_dir(d, kcsf=1)
    b = csf.jbasis(o)
    print(f"csf_{name}: Nc={o.Nc} classes={o.nconf_neq} Ncsf={b.ncsf} nccj={b.nccj} in {time.time() - t:.1f} s",
          flush=True)
    have = np.flatnonzero(b.conf_ncsf > 0) + 1
    confs = np.unique(have[np.linspace(0, have.size - 1, NCONF).round().astype(np.int64)])
    t = time.time()
    rps, cols, bits = [], [], []
    for ic in confs:
        rp, col, val, _, _, _ = csf.formh_sym(o, b, int(ic), int(ic))
        rps.append(np.diff(rp))
        cols.append(col)
        bits.append(val.view(np.int64))
    out = dict(name=name, digest=input_digest(d), nc_neq=o.nc_neq, nconf_neq=np.int64(o.nconf_neq), ndcs=b.ndcs,
               ccj=b.ccj, ncsf=np.int64(b.ncsf), confs=confs, row_counts=np.concatenate(rps),
               col=np.concatenate(cols), val_bits=np.concatenate(bits))
    print(f"csf_{name}: {confs.size} configurations, {out['col'].size} elements in {time.time() - t:.1f} s",
          flush=True)
    if name in LANCZOS:
        import scipy.sparse.linalg as sla
        t = time.time()
        rp, col, val, nz, hmin, npairs = csf.formh_sym(o, b)
        w = sla.eigsh(full_from_lower(rp, col, val), k=LANCZOS[name], which="SA", tol=1e-13,
                      ncv=max(4 * LANCZOS[name], 40), return_eigenvectors=False)
        out.update(numh=np.int64(rp[-1]), numzero=np.int64(nz), hmin=np.float64(hmin), eigvals=np.sort(w))
        print(f"csf_{name}: NumH={rp[-1]} numzero={nz} Lanczos {np.sort(w)} in {time.time() - t:.1f} s", flush=True)
    path = os.path.join(HERE, f"scale_csf_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"-> {path} ({os.path.getsize(path)} B)", flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or ["c2", "c4"]:
        if n.startswith("csf_"):
            make_csf(n[4:])
        else:
            make(n)
