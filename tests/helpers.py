"""Shared test set-up: seeded spaces, the oracle, and device contexts."""
import os
import tempfile

import numpy as np

from pci_amd import formats as F
from pci_amd import gen

_CACHE = os.path.join(tempfile.gettempdir(), f"pci_spaces_{os.getuid()}")


def space_dir(name, **over):
    """Generate (once per machine) the preset space `name` and return its directory."""
    key = name + "".join(f"_{k}{v}" for k, v in sorted(over.items()))
    d = os.path.join(_CACHE, key)
    if not os.path.exists(os.path.join(d, "CONF.GNT")):
        import dataclasses
        spec = dataclasses.replace(gen.PRESETS[name], **over)
        tmp = d + f".tmp{os.getpid()}"
        gen.write_space(spec, tmp)
        os.makedirs(_CACHE, exist_ok=True)
        try:
            os.rename(tmp, d)
        except OSError:
            pass
    return d


def csf_dir(name, **over):
    """The preset space `name` with kCSF = 1 in its ci.in (the other input
    files linked from the space's own directory)."""
    d = space_dir(name, **over)
    c = d + "_kcsf1"
    if not os.path.exists(os.path.join(c, "ci.in")):
        tmp = c + f".tmp{os.getpid()}"
        os.makedirs(tmp, exist_ok=True)
        for f in os.listdir(d):
            if f != "ci.in":
                os.symlink(os.path.join(d, f), os.path.join(tmp, f))
        cio = F.read_ci_in(os.path.join(d, "ci.in"))
        F.write_ci_in(os.path.join(tmp, "ci.in"), Kl=cio["KL"], Kw=cio["KW"], KXIJ=cio["KXIJ"],
                      MaxNd0=cio["MAXND0"], kCSF=1, Ksig=cio["KSIG"], Kdsig=cio["KDSIG"])
        try:
            os.rename(tmp, c)
        except OSError:
            pass
    return c


def oracle_for(name, **over):
    from oracle.oracle import Oracle
    return Oracle.from_dir(space_dir(name, **over))


def inputs(d):
    ci = F.read_conf_int(os.path.join(d, "CONF.INT"))
    nlx, _, gidx, gnt = F.read_conf_gnt(os.path.join(d, "CONF.GNT"))
    return ci, nlx, gidx, gnt


def gpu_context_like(o, d):
    """Device context loaded with the same space/integrals the oracle holds."""
    from pci_amd._lib import GpuContext
    ci, nlx, gidx, gnt = inputs(d)
    g = GpuContext(0)
    sh = o.shells
    g.set_space(sh[:, 1], sh[:, 3], sh[:, 2], o.Nso, o.jz, o.nh, o.iarr, o.ndc, o.Mj, o.inp.Gj)
    g.set_integrals(ci.nx, ci.rint1, ci.iint1, ci.kbrt, ci.rint2, ci.iint2, ci.iint3, nlx, gidx, gnt)
    if getattr(o, "ksig", 0):
        sg = o.sigma
        diag = np.array([o.diags(n) for n in range(1, o.Nd + 1)]) if sg["kdsig"] else None
        g.set_sigma(o.ksig, sg["kdsig"], sg["nmaxs"], sg["lmaxs"], sg["ind"], sg["r"], sg["d"], sg["e"],
                    sg["eps"], diag)
        if o.ksig == 2:
            s2 = o.sigma2
            g.set_sigma2(s2["ksym"], s2["kmax"], s2["nsum"], s2["scr"], s2["i2"], s2["i3"], s2["r"], s2["d"],
                         s2["e"])
    return g


def full_from_lower(rp, col, val, hmin=0.0):
    """scipy CSR of the symmetric H - hmin I from a lower-triangle CSR (1-based cols)."""
    import scipy.sparse as sp
    nd = rp.size - 1
    rows = np.repeat(np.arange(nd), np.diff(rp))
    L = sp.csr_matrix((val, (rows, col - 1)), shape=(nd, nd))
    Hf = L + L.T - sp.diags(L.diagonal())
    return (Hf - hmin * sp.identity(nd)).tocsr()


def spec_dir(spec):
    """Generate (once per machine) the space of a SpaceSpec that is not a preset."""
    d = os.path.join(_CACHE, "spec_" + spec.name)
    if not os.path.exists(os.path.join(d, "CONF.GNT")):
        tmp = d + f".tmp{os.getpid()}"
        gen.write_space(spec, tmp)
        os.makedirs(_CACHE, exist_ok=True)
        try:
            os.rename(tmp, d)
        except OSError:
            import shutil
            shutil.rmtree(tmp, ignore_errors=True)
    return d
