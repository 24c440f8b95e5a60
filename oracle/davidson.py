"""CPU restatement of pconf's Davidson procedure (TEST INFRASTRUCTURE ONLY).

Follows Diag4 (src/pconf.F90:1993-2264) step by step with npes=1:
  Init4          src/davidson.f90:20-73     Z1 = H[1:Nd0,1:Nd0] - Hmin I
  DiagInitApprox src/pconf.F90:1866-1891     DSYEVD(Z1)  (numpy.linalg.eigh)
  FormB0         src/davidson.f90:75-193     first Nlv eigenvectors (Kv=4: no J filter)
  Ortn           src/davidson.f90:649-741    classical Gram-Schmidt, <= 4 passes
  Mxmpy          src/davidson.f90:572-647    symmetric scatter (oracle.mxmpy, C)
  FormP          src/davidson.f90:300-369
  Dvdsn          src/davidson.f90:436-570    residual + Z1/diagonal correction
  DSYEV, admixture, FormB/FormBskip          src/pconf.F90:2130-2230, davidson.f90:195-267
With jsel (Kv = 3, K_prj = 1) the J selection of the reference is added:
  FormB0 J filter  src/davidson.f90:104-131   keep Z1 vectors with |J - XJ_av| <= 0.1
  AvgDiag          src/davidson.f90:371-434   diagonal averaged per configuration
  Prj_J            src/davidson.f90:743-884   (I - J^2/(jx(jx+2))) filters on the
                                              corrections, then Ortn again
                                              (src/pconf.F90:2084, 2109-2128)
For a J = 0 term the reference's factor 1/(jx(jx+2)) divides by zero; the
restatement (like the device code) uses the limit of the normalised filter,
x <- -J^2 x.
Returns the reported energies Tk = -(E + Hmin), the vectors and the
iteration trace.  LAPACK here is numpy's; eigenvector signs may differ from
another LAPACK, which changes the trajectory only by signs.
"""
from __future__ import annotations

import numpy as np

from .oracle import mxmpy


class OrthoFail(RuntimeError):
    pass


def _jsq_apply(jrp, jcol, jval, X):
    """J^2 X by symmetric scatter over the stored lower triangle (as Prj_J)."""
    nd = jrp.size - 1
    rows = np.repeat(np.arange(nd), np.diff(jrp))
    k = jcol - 1
    Y = np.zeros_like(X)
    np.add.at(Y, rows, jval[:, None] * X[k])
    off = rows != k
    np.add.at(Y, k[off], jval[off, None] * X[rows[off]])
    return Y


def diag4(rp, col, val, hmin, nd0, nlv, crt4, n_it, verbose=False, jsel=None, ops=None, x0=None, tk0=None):
    """jsel: None (Kv = 4) or dict(xj=XJ_av, jt=Jt, jrp/jcol/jval=J^2 lower CSR, ndc=Ndc).

    x0, tk0: FormB0 with |Kl4| = 2 (src/davidson.f90:135-153): the start
    vectors (x0[k, :ndpt], zero-padded to Nd; Nlv becomes their number) and
    energies are given instead of the Z1 eigenvectors, and Dvdsn's Z1
    projector then uses every Z1 column (n01 = 1, davidson.f90:518-519).

    ops: H held elsewhere (rp, col, val unused), e.g. as row ranges in
    separate processes like pconf's MPI ranks: an object with nd, head(nd0)
    (the dense symmetric H[1:nd0, 1:nd0]), mx(X) (Mxmpy of the columns of
    X, nd x m, the ranks' partial products summed as pconf's Allreduce
    does) and diag() (H(n, n) - Hmin)."""
    from .oracle import jav as _jav
    nd = ops.nd if ops is not None else rp.size - 1
    # Init4: leading block with the shifted diagonal
    if ops is not None:
        Z1 = ops.head(nd0) - hmin * np.eye(nd0)
    else:
        Z1 = np.zeros((nd0, nd0))
        for n in range(nd0):
            for j in range(rp[n], rp[n + 1]):
                k = col[j] - 1
                t = val[j] - (hmin if k == n else 0.0)
                if k < nd0:
                    Z1[n, k] = t
                    Z1[k, n] = t
    E1, Z1 = np.linalg.eigh(Z1)
    nlv = min(nlv, nd0)
    A = np.zeros((nd, 3 * nlv))
    if jsel is None:
        sel = list(range(nlv))
    else:
        # FormB0 with K_prj = 1: J_av over the first Nd0 rows, ierr if |J - XJ_av| > 0.1
        sel = []
        x = np.zeros(nd)
        for j in range(nd0):
            x[:nd0] = Z1[:, j]
            xj = _jav(jsel["jrp"], jsel["jcol"], jsel["jval"], x, nd0)
            if abs(xj - jsel["xj"]) <= 0.1:
                sel.append(j)
                if len(sel) >= nlv:
                    break
        nlv = len(sel)             # Nlv = num
        A = np.zeros((nd, 3 * nlv))
    A[:nd0, :nlv] = Z1[:, sel]
    E = -(E1[sel] + hmin)          # FormB0's E(num)
    n01 = nlv
    if x0 is not None:
        x0 = np.atleast_2d(np.asarray(x0, np.float64))
        nlv = x0.shape[0]
        A = np.zeros((nd, 3 * nlv))
        A[:x0.shape[1], :nlv] = x0.T
        E = np.asarray(tk0, np.float64)[:nlv].copy()
        n01 = 0
    Tk = E.copy()
    iconv = np.zeros(nlv, bool)
    diag = None
    ax, vmax, crit = 1.0, 0.0, 1e-6
    P = np.zeros((2 * nlv, 2 * nlv))
    trace = []
    it_done = 0

    def ortn(j):
        jm = j
        smax2 = 1e10
        it = 0
        while True:
            it += 1
            smax1 = 0.0
            if jm > 0:
                s = A[:, :jm].T @ A[:, j]
                for l in range(jm):
                    if l >= nlv and iconv[l - nlv]:
                        s[l] = 0.0
                    elif abs(s[l]) > smax1:
                        smax1 = abs(s[l])
                A[:, j] -= A[:, :jm] @ s
            if smax1 < 1e-6:
                break
            if smax1 > smax2 or it > 3:
                raise OrthoFail(f"Fail of orthogonalization of vector {j + 1}")
            smax2 = smax1
        s = float(A[:, j] @ A[:, j])
        if s < 1e-9:
            raise OrthoFail(f"Fail of normalization of vector {j + 1}")
        A[:, j] *= 1.0 / np.sqrt(s)

    def prj_j(js, trsd=1e-5):
        """Prj_J(Nlv+1, Nlv, 2Nlv+1, 1e-5) on the correction block."""
        jt = list(js["jt"])
        njd = len(jt)
        jav = int(2 * js["xj"] + 0.1)
        act = ~iconv
        cin = A[:, nlv:2 * nlv].copy()
        jx, f = None, 0.0
        for itr in range(1, njd + 2):
            itj = njd + 1 - itr
            if itj != 0:
                jx = jt[itj - 1]
                f = 1.0 / (jx * (jx + 2)) if jx != 0 else 0.0
            if jx != jav or itj == 0:
                cout = _jsq_apply(js["jrp"], js["jcol"], js["jval"], cin)
                cout[:, ~act] = 0.0
                A[:, 2 * nlv:] = cout
                err1 = 0.0
                for i in range(nlv):
                    if act[i]:
                        sj = float(cin[:, i] @ cout[:, i])
                        err1 = max(err1, abs(np.sqrt(1 + sj) - 1 - jav))
                if err1 < trsd:
                    break
                if itj != 0:
                    for i in range(nlv):
                        if act[i]:
                            a = (cin[:, i] - f * cout[:, i]) if jx != 0 else -cout[:, i]
                            cin[:, i] = a / np.sqrt(float(a @ a))
                            A[:, nlv + i] = cin[:, i]

    def mx(ip):
        nonlocal diag
        i1 = (ip - 1) * nlv
        i2 = ip * nlv
        if ops is not None:
            A[:, i2:i2 + nlv] = ops.mx(A[:, i1:i1 + nlv])
            if diag is None:
                diag = ops.diag()
            return
        for iv in range(nlv):
            if diag is None:
                y, diag = mxmpy(rp, col, val, hmin, A[:, i1 + iv], want_diag=True)
            else:
                y = mxmpy(rp, col, val, hmin, A[:, i1 + iv])
            A[:, i2 + iv] = y

    for it in range(1, n_it + 1):
        if not ax > crit:
            continue
        it_done = it
        for i in range(nlv):
            ortn(i)
        mx(1)
        if it == 1 and jsel is not None:
            # AvgDiag: the diagonal averaged over each relativistic configuration
            st = 0
            for n in jsel["ndc"]:
                n = int(n)
                if n:
                    diag[st:st + n] = diag[st:st + n].sum() / n
                st += n
        Pt = A[:, :nlv].T @ A[:, nlv:2 * nlv]
        for i in range(nlv):
            for k in range(i + 1):
                P[i, k] = P[k, i] = Pt[i, k]
        cnx = 0.0
        for j in range(nlv):
            j1 = j + nlv
            th = P[j, j]
            C = A[:, j1] - th * A[:, j]
            cn = np.sqrt(abs(float(C @ C)))
            iconv[j] = cn < crt4
            cnx = max(cnx, cn)
            if iconv[j]:
                A[:, j1] = 0.0
                continue
            B = np.zeros(nd)
            if nd0 - n01 > 0:
                sv = Z1[:, n01:].T @ C[:nd0]
                t = th - E1[n01:]
                sv = np.where(np.abs(t) < 1e-6, 0.0, sv / np.where(np.abs(t) < 1e-6, 1.0, t))
                B[:nd0] = Z1[:, n01:] @ sv
            B[nd0:] = C[nd0:] / (th - diag[nd0:])
            A[:, j1] = B / np.sqrt(B @ B)
            ortn(j1)
        if jsel is not None:
            prj_j(jsel)
            for i in range(nlv):
                if not iconv[i]:
                    ortn(nlv + i)
        trace.append(cnx)
        if verbose:
            print(f"it {it}: max |C| = {cnx:.3e}")
        if cnx > crt4:
            mx(2)
            pmax = max(P[i, i] for i in range(nlv))
            vmax = max(vmax, pmax)
            Pt = A[:, :2 * nlv].T @ A[:, 2 * nlv:3 * nlv]
            for i in range(nlv):
                for k in range(nlv, 2 * nlv):
                    P[i, k] = P[k, i] = Pt[i, k - nlv]
            for i in range(nlv, 2 * nlv):
                for k in range(nlv, i + 1):
                    P[i, k] = P[k, i] = Pt[i, k - nlv]
            for i in range(nlv):
                if iconv[i]:
                    i1 = i + nlv
                    P[i1, :] = 0.0
                    P[:, i1] = 0.0
                    P[i1, i1] = 2 * vmax - P[i, i] + 1.0
            Ev, Q = np.linalg.eigh(P)
            ax = 0.0
            vmax = -1e10
            for i in range(nlv):
                xx = np.abs(Q[nlv:, i]).max()
                ax = max(ax, xx)
                vmax = max(vmax, Ev[i])
            A[:, 2 * nlv:] = A[:, :2 * nlv] @ Q[:, :nlv]
            Tk = -(Ev[:nlv] + hmin)
            A[:, :nlv] = A[:, 2 * nlv:]
        else:
            break
    return Tk, A[:, :nlv].T.copy(), it_done, trace
