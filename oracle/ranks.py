"""The CPU oracle run as pconf runs under MPI: P processes ("ranks"), each
owning a contiguous range of H's rows.

TEST INFRASTRUCTURE ONLY.  Used by bench.py's cpu_baseline leg (the complete
CPU run of a workload) and by tests/; never by the product package.

    FormH     src/pconf.F90:1160-1651      each rank forms the lower-triangle
                                           rows it owns, every element once
                                           (Oracle.formh_once)
    rows      src/pconf.F90:1357-1420      pconf's rank 0 hands out chunks of
                                           rows to the workers as they finish
                                           (and RedistributeHamCSR then moves
                                           the elements into contiguous row
                                           ranges, src/matrix_io.f90:318-340);
                                           here chunk c of CHUNK rows goes to
                                           rank c mod P, which spreads the
                                           build's and the product's work
                                           evenly without either step
    Mxmpy     src/davidson.f90:615-641     each rank scatters its rows into a
                                           full-length column, the columns of
                                           all ranks are summed (the
                                           Allreduce), rank order
    Diag4     oracle.davidson.diag4        in the parent, on the ranks'
                                           products through its ops hook

Columns travel through shared memory: the parent writes the block X (m
columns of Nd), every rank writes its partial products into its own slice
of Y, the parent sums the slices.
"""
from __future__ import annotations

import multiprocessing as mp
import time
from multiprocessing import shared_memory

import numpy as np


CHUNK = 64


def rank_chunks(nd: int, p: int, rank: int, chunk: int = CHUNK):
    """Row ranges (1-based, inclusive) of rank: chunks rank, rank + p, ..."""
    return [(a + 1, min(a + chunk, nd)) for a in range(rank * chunk, nd, p * chunk)]


def _rank_main(conn, d, kcsf, rank, nd, nmax, nranks, xname, yname):
    from .oracle import Oracle, lib
    o = Oracle.from_dir(d, kcsf=kcsf) if kcsf else Oracle.from_dir(d)
    conn.send("ready")
    conn.recv()                                   # "build"
    t = time.perf_counter()
    parts = [o.formh_once(r0, r1) for r0, r1 in rank_chunks(nd, nranks, rank)]
    t_build = time.perf_counter() - t
    rows = np.concatenate([np.arange(r0, r1 + 1, dtype=np.int32) for r0, r1 in rank_chunks(nd, nranks, rank)]
                          or [np.zeros(0, np.int32)])
    counts = np.concatenate([np.diff(q[0]) for q in parts] or [np.zeros(0, np.int64)])
    rp = np.zeros(rows.size + 1, np.int64)
    np.cumsum(counts, out=rp[1:])
    col = np.concatenate([q[1] for q in parts] or [np.zeros(0, np.int32)])
    val = np.concatenate([q[2] for q in parts] or [np.zeros(0)])
    nz = sum(q[3] for q in parts)
    del parts
    conn.send((int(rp[-1]), t_build, float(val.min()) if rp[-1] else np.inf, int(nz)))
    sx = shared_memory.SharedMemory(name=xname)
    sy = shared_memory.SharedMemory(name=yname)
    X = np.ndarray((nmax, nd), np.float64, sx.buf)
    Y = np.ndarray((nranks, nmax, nd), np.float64, sy.buf)
    hmin = None
    L = lib()
    nrows = rows.size
    try:
        while True:
            cmd = conn.recv()
            if cmd[0] == "hmin":
                hmin = cmd[1]
                conn.send("ok")
            elif cmd[0] == "mx":
                for c in range(cmd[1]):
                    L.orc_mxmpy_list(nd, nrows, rows, rp, col, val, hmin, X[c], Y[rank, c])
                conn.send("ok")
            elif cmd[0] == "diag":
                # columns ascend within a row and end at the diagonal
                dg = np.zeros(nd)
                last = rp[1:] - 1
                ok = (rp[1:] > rp[:-1])
                ok[ok] = col[last[ok]] == rows[ok]
                dg[rows[ok] - 1] = val[last[ok]] - hmin
                conn.send(dg)
            elif cmd[0] == "head":
                nd0 = cmd[1]
                sel = np.flatnonzero(rows <= nd0)
                blk = np.zeros((sel.size, nd0))
                for j, i in enumerate(sel):
                    a, b = rp[i], rp[i + 1]
                    c = col[a:b] - 1
                    m = c < nd0
                    blk[j, c[m]] = val[a:b][m]
                conn.send((rows[sel] - 1, blk))
            else:
                break
    finally:
        del X, Y
        sx.close()
        sy.close()
        conn.close()


class OracleRanks:
    """P oracle ranks over the space in directory d.  build() forms H;
    then nd, head(), mx(), diag() serve oracle.davidson.diag4 (ops)."""

    def __init__(self, d, nd, p, nmax, kcsf=0):
        self.nd, self.nmax = nd, nmax
        self.p = max(1, min(p, -(-nd // CHUNK)))
        nbytes = max(8 * nmax * nd, 8)
        self._sx = shared_memory.SharedMemory(create=True, size=nbytes)
        self._sy = shared_memory.SharedMemory(create=True, size=nbytes * self.p)
        self.X = np.ndarray((nmax, nd), np.float64, self._sx.buf)
        self.Y = np.ndarray((self.p, nmax, nd), np.float64, self._sy.buf)
        self.conns, self.procs = [], []
        ctx = mp.get_context("fork")
        for r in range(self.p):
            a, b = ctx.Pipe()
            pr = ctx.Process(target=_rank_main, args=(b, d, kcsf, r, nd, nmax, self.p,
                                                      self._sx.name, self._sy.name), daemon=True)
            pr.start()
            self.conns.append(a)
            self.procs.append(pr)
        for c in self.conns:
            assert c.recv() == "ready"
        self.hmin = None
        self.numh = 0

    def build(self):
        """FormH on every rank; returns (wall seconds, NumH, Hmin, numzero)."""
        t = time.perf_counter()
        for c in self.conns:
            c.send("build")
        rep = [c.recv() for c in self.conns]
        wall = time.perf_counter() - t
        self.numh = sum(r[0] for r in rep)
        self.hmin = min(r[2] for r in rep)    # Hamil%minval: the minimum over all ranks
        for c in self.conns:
            c.send(("hmin", self.hmin))
        for c in self.conns:
            c.recv()
        self.rank_build_s = [r[1] for r in rep]
        return wall, self.numh, self.hmin, sum(r[3] for r in rep)

    def head(self, nd0):
        for c in self.conns:
            c.send(("head", nd0))
        low = np.zeros((nd0, nd0))
        for c in self.conns:
            idx, blk = c.recv()
            low[idx] = blk
        return low + low.T - np.diag(np.diag(low))

    def diag(self):
        for c in self.conns:
            c.send(("diag",))
        return sum(c.recv() for c in self.conns)

    def mx(self, A):
        m = A.shape[1]
        assert m <= self.nmax
        self.X[:m] = A.T
        for c in self.conns:
            c.send(("mx", m))
        for c in self.conns:
            c.recv()
        return self.Y[:, :m].sum(axis=0).T

    def close(self):
        for c in self.conns:
            try:
                c.send(("stop",))
            except (BrokenPipeError, OSError):
                pass
        for pr in self.procs:
            pr.join(timeout=30)
        del self.X, self.Y
        for s in (self._sx, self._sy):
            s.close()
            s.unlink()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
