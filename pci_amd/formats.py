"""Readers and writers for the pconf file contract.

Every layout here is the one pconf reads or writes, so files produced by this
module can be fed to the reference program and vice versa:

* ``CONF.INP``  text header + configuration list   (src/conf_init.f90:37-135, 215-313)
* ``ci.in``     text key=value job file             (src/conf_init.f90:137-213)
* ``CONF.DAT``  direct access, 470 doubles/record   (src/pconf.F90:443-503)
* ``CONF.INT``  sequential unformatted integrals    (src/integrals.f90:35-83, src/pbasc.f90:1169-1190)
* ``CONF.GNT``  sequential unformatted Gaunt table  (src/breit.f90:96-101, src/pconf.F90:383-406)
* ``CONF.DET``  sequential unformatted determinants (src/determinants.f90:448-462)
* ``pCONF.HIJ`` global CSR stream                   (src/matrix_io.f90:689-799)
* ``CONF.XIJ``  sequential unformatted eigenvectors (src/pconf.F90:2566-2602)

Sequential unformatted records use 4-byte little-endian length markers on both
sides of the payload (the layout gfortran, ifx and flang all write for records
below 2 GiB).
"""
from __future__ import annotations

import dataclasses
import os
import struct
from typing import Iterable, List, Sequence, Tuple

import numpy as np

IP6 = 470          # CONF.DAT record length in doubles (src/params.f90:33)
IPX = 440          # Nx, integral index radix (src/params.f90:30)
DPRY = 109737.315709  # Rydberg in cm^-1 (src/params.f90:18)
HARTREE_CM = 2.0 * DPRY


# ----------------------------------------------------------------------------
# Fortran sequential unformatted records
# ----------------------------------------------------------------------------

def write_records(path: str, records: Iterable[bytes]) -> None:
    with open(path, "wb") as f:
        for rec in records:
            n = len(rec)
            if n >= 2**31:
                raise ValueError("record larger than 2 GiB is not supported")
            f.write(struct.pack("<i", n))
            f.write(rec)
            f.write(struct.pack("<i", n))


def read_records(path: str) -> List[bytes]:
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        if n < 0:
            raise ValueError("split (>2 GiB) records are not supported")
        pos += 4
        out.append(data[pos:pos + n])
        pos += n
        (n2,) = struct.unpack_from("<i", data, pos)
        if n2 != n:
            raise ValueError(f"corrupt record markers in {path}")
        pos += 4
    return out


def _pack(*items) -> bytes:
    """Concatenate scalars/arrays into one record payload."""
    parts = []
    for it in items:
        if isinstance(it, np.ndarray):
            parts.append(np.ascontiguousarray(it).tobytes())
        elif isinstance(it, tuple):
            fmt, val = it
            parts.append(struct.pack("<" + fmt, val))
        else:
            raise TypeError(type(it))
    return b"".join(parts)


# ----------------------------------------------------------------------------
# Shell codes of CONF.INP:  +-n.lqq   (src/pconf.F90:506-516)
# ----------------------------------------------------------------------------

def encode_shell(n: int, l: int, kappa: int, q: int) -> float:
    """Return the CONF.INP number for shell (n, l, kappa) holding q electrons.

    The sign is + for j = l + 1/2 (kappa = -(l+1)) and - for j = l - 1/2
    (kappa = l), as decoded at src/pconf.F90:506-516.
    """
    if q < 0 or q > 99 or l > 9 or n > 99:
        raise ValueError("shell out of encodable range")
    mag = n / 10.0 + l / 100.0 + q / 10000.0
    return mag if kappa < 0 else -mag


def decode_shell(x: float) -> Tuple[int, int, int, int]:
    """Inverse of encode_shell, following pconf's own arithmetic."""
    i = 1 if x >= 0 else -1
    d = abs(x) + 1e-14
    d = 10.0 * d
    n = int(d)
    d = 10.0 * (d - n)
    l = int(d)
    jl = 2 * l + i
    kappa = -i * ((jl + 1) // 2)
    d = 100.0 * (d - l)
    q = int(d + 0.1)
    return n, l, kappa, q


def _fmt_f74(x: float) -> str:
    s = f"{x:7.4f}"
    if len(s) != 7:
        raise ValueError(f"value {x} does not fit F7.4")
    return s


# ----------------------------------------------------------------------------
# CONF.INP / ci.in
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class ConfInp:
    name: str
    Z: float
    Am: float
    J: float
    Jm: float
    Nso: int
    Nc: int
    Kv: int
    Nlv: int
    Ne: int
    Kl4: int
    Nc4: int
    Crt4: float
    N_it: int
    Kbrt: int = 0
    Gj: float = 0.0
    core: Sequence[float] = ()
    configs: Sequence[Sequence[float]] = ()  # each a list of shell codes


def write_conf_inp(path: str, inp: ConfInp) -> None:
    """Write CONF.INP in the layout read by ReadConfInp/ReadConfigurations.

    Keys are padded so the first five columns up to '=' match pconf's sample
    (docs/pci/pconf.rst:40-80).  Each configuration starts on its own line and
    holds at most six shells per line, format (6(4X,F7.4)).
    """
    lines = [" " + inp.name[:16]]
    kv = [("Z", f"{inp.Z:.1f}"), ("Am", f"{inp.Am:.1f}"), ("J", f"{inp.J:.1f}"),
          ("Jm", f"{inp.Jm:.1f}"), ("Nso", str(inp.Nso)), ("Nc", str(inp.Nc)),
          ("Kv", str(inp.Kv)), ("Nlv", str(inp.Nlv)), ("Ne", str(inp.Ne)),
          ("Kl4", str(inp.Kl4)), ("Nc4", str(inp.Nc4)), ("Crt4", f"{inp.Crt4:.3e}"),
          ("N_it", str(inp.N_it)), ("Kbrt", str(inp.Kbrt)), ("Gj", f"{inp.Gj:.4f}")]
    for k, v in kv:
        lines.append(f"{k:>4}= {v}")
    if len(inp.core) != inp.Nso:
        raise ValueError("core list length must equal Nso")
    if inp.Nso == 0:
        # ReadConfParams consumes the first line without '=' and calcNsp then
        # counts configurations from the next line (src/conf_init.f90:37-70,
        # 316-360); with no core shells that line must be blank.
        lines.append("")
    for i in range(0, len(inp.core), 6):
        lines.append("".join("    " + _fmt_f74(x) for x in inp.core[i:i + 6]))
    for ic, cfg in enumerate(inp.configs):
        for i in range(0, len(cfg), 6):
            chunk = cfg[i:i + 6]
            lab = f"{(ic + 1) % 10000:4d}" if i == 0 else "    "
            s = lab + _fmt_f74(chunk[0]) + "".join("    " + _fmt_f74(x) for x in chunk[1:])
            lines.append(s)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_conf_inp(path: str) -> ConfInp:
    """Parse CONF.INP the way ReadConfParams/ReadConfigurations do."""
    with open(path) as f:
        text = f.read().splitlines()
    name = text[0][1:17].rstrip()
    vals = {}
    i = 1
    while i < len(text) and "=" in text[i]:
        line = text[i]
        key, val = line.split("=", 1)
        val = val.split("#", 1)[0].strip()
        vals[key.strip().upper()[:4]] = val
        i += 1
    if i < len(text) and text[i].strip() == "":
        i += 1

    def num(key, default, typ):
        return typ(float(vals[key])) if key in vals else default

    Nso = num("NSO", 0, int)
    Nc = num("NC", 0, int)
    Ne = num("NE", 0, int)

    def entries(line):
        out = []
        for j in range(6):
            field = line[4 + 11 * j: 11 + 11 * j]
            out.append(float(field) if field.strip() else 0.0)
        return out

    core = []
    while len(core) < Nso:
        core.extend(x for x in entries(text[i]) if abs(x) > 1e-8)
        i += 1
    core = core[:Nso]
    configs = []
    for _ in range(Nc):
        ne0 = 0
        cfg = []
        while ne0 < Ne:
            for x in entries(text[i]):
                if abs(x) + 1e-9 < 1e-8:
                    break
                ax = abs(x) + 1e-9
                ne0 += int(10000 * ax) - 100 * int(100 * ax)
                cfg.append(x)
            i += 1
        configs.append(cfg)
    return ConfInp(name=name, Z=num("Z", 0.0, float), Am=num("AM", 0.0, float),
                   J=num("J", 0.0, float), Jm=num("JM", 0.0, float), Nso=Nso, Nc=Nc,
                   Kv=num("KV", 4, int), Nlv=num("NLV", 1, int), Ne=Ne,
                   Kl4=num("KL4", 1, int), Nc4=num("NC4", Nc, int),
                   Crt4=num("CRT4", 1e-4, float), N_it=num("N_IT", 20, int),
                   Kbrt=num("KBRT", 0, int), Gj=num("GJ", 0.0, float),
                   core=core, configs=configs)


def write_ci_in(path: str, Kl=0, Kw=0, KLSJ=0, KXIJ=10, MaxNd0=3000, kCSF=0, Ksig=0, Kdsig=0) -> None:
    with open(path, "w") as f:
        f.write(f"Kl={Kl}\nKw={Kw}\nKLSJ={KLSJ}\nKXIJ={KXIJ}\nMaxNd0={MaxNd0}\nkCSF={kCSF}\n")
        if Ksig:
            f.write(f"Ksig={Ksig}\nKdsig={Kdsig}\n")


def read_ci_in(path: str) -> dict:
    out = dict(KL=0, KSIG=0, KDSIG=0, KW=0, KLSJ=0, K_SMS=3, KXIJ=10, KWEIGHTS=0,
               MAXND0=3000, KCSF=0)
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            if "=" not in line or line.lstrip().startswith("#"):
                continue
            k, v = line.split("=", 1)
            v = v.split("#", 1)[0].strip()
            if k.strip().upper() in out:
                out[k.strip().upper()] = int(v)
    return out


# ----------------------------------------------------------------------------
# CONF.DAT (direct access; pconf reads records 1, 2, 5, 6 only)
# ----------------------------------------------------------------------------

def write_conf_dat(path: str, Z: float, shells: Sequence[Tuple[int, int, int, int]],
                   nrec: int = 8, eps: Sequence[float] | None = None) -> None:
    """Write the short-format CONF.DAT header pconf's Init reads.

    ``shells`` holds (n, l, kappa, 2j) per orbital.  pq(1)=Z, pq(2)=Ns and six
    doubles per orbital from pq(21) ordered (n, l, q, -, kappa, 2j); pq(1:1880)
    is the concatenation of records 1, 2, 5 and 6 (src/pconf.F90:443-503).
    With ``eps`` (one energy per orbital) the orbital records 2*ni+7 are
    written too, with p(ii+1) = -eps(ni) where ii = pq(3) = 0 grid points:
    what pconf reads for the MBPT energies (src/pconf.F90:663-667).
    """
    ns = len(shells)
    if 21 + 6 * ns > 4 * IP6:
        raise ValueError("too many orbitals for the short CONF.DAT header")
    pq = np.zeros(4 * IP6)
    pq[0] = Z
    pq[1] = ns
    pq[2] = 0.0
    for i, (n, l, kap, j2) in enumerate(shells):
        b = 20 + 6 * i
        pq[b:b + 6] = (n, l, 0.0, 0.0, kap, j2)
    if eps is not None:
        nrec = max(nrec, 2 * ns + 7)
    recs = np.zeros((max(nrec, 6), IP6))
    if eps is not None:
        for ni in range(1, ns + 1):
            recs[2 * ni + 6, 0] = -float(eps[ni - 1])
    recs[0] = pq[0:IP6]
    recs[1] = pq[IP6:2 * IP6]
    recs[4] = pq[2 * IP6:3 * IP6]
    recs[5] = pq[3 * IP6:4 * IP6]
    recs.astype("<f8").tofile(path)


def read_conf_dat_eps(path: str, nso: int, nsu: int) -> np.ndarray:
    """Orbital energies Eps(ni), ni = Nso+1..Nsu, of CONF.DAT (records 2*ni+7,
    Eps = -p(ii+1) with ii = pq(3), src/pconf.F90:470, 663-667).  Returns an
    array indexed by shell (index 0 and core shells 0)."""
    raw = np.fromfile(path, dtype="<f8").reshape(-1, IP6)
    ii = int(raw[0][2] + 0.01)
    eps = np.zeros(nsu + 1)
    for ni in range(nso + 1, nsu + 1):
        eps[ni] = -raw[2 * ni + 6][ii]
    return eps


def write_sgc_con(path: str, nmaxs: int, lmaxs: int, khot: int, entries, scr=None) -> None:
    """SGC.CON, the one-electron MBPT (sigma) file pconf reads with Ksig >= 1
    (src/pconf.F90:715-762): a header line read with (7X,I3,5X,I1,24X,I1)
    for NmaxS, LmaxS, Khot, then one list-directed line per shell pair
    (label, na, nb, R, D, E) in pconf's loop order: na = Nso+1..NmaxS with
    l_a <= LmaxS, nb = na..NmaxS with the same l and j.  R, D, E are read as
    single precision.  With scr (Ksig = 2, src/pconf.F90:764-775): a title
    line, then the ten screening coefficients, one per line as
    (5X,I2,5X,F8.5) with k = 0..9."""
    with open(path, "w") as f:
        f.write(f" NmaxS={nmaxs:3d} Lmx={lmaxs:1d}" + " " * 24 + f"{khot:1d}\n")
        for na, nb, r, d, e in entries:
            f.write(f" sig {na:4d} {nb:4d} {r: .8E} {d: .8E} {e: .8E}\n")
        if scr is not None:
            f.write(" screening coefficients\n")
            for k, x in enumerate(scr):
                f.write(f" k = {k:2d}  Scr{x:8.5f}\n")


def read_sgc_con(path: str, ksig: int = 1):
    """(NmaxS, LmaxS, Khot, [(na, nb, R, D, E)]) with R, D, E rounded to FP32
    as pconf's list-directed read into REAL variables does; with ksig = 2
    also the ten screening coefficients (as read by F8.5)."""
    with open(path) as f:
        head = f.readline()
        nmaxs, lmaxs, khot = int(head[7:10]), int(head[15:16]), int(head[40:41])
        out, scr = [], []
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        t = lines[i].split()
        if len(t) >= 6 and t[0] == "sig":
            out.append((int(t[1]), int(t[2]), float(np.float32(t[3])), float(np.float32(t[4])),
                        float(np.float32(t[5]))))
            i += 1
        else:
            break
    if ksig >= 2:
        for line in lines[i + 1:i + 11]:
            scr.append(float(line[12:20]))  # (5X,I2,5X,F8.5)
        return nmaxs, lmaxs, khot, out, np.array(scr)
    return nmaxs, lmaxs, khot, out


@dataclasses.dataclass
class ScrcCon:
    """SCRC.CON, the two-electron MBPT file of Ksig = 2 (read by RintS,
    src/pconf.F90:781-840).  Records: (Ns, Nso, Nsh, Khot, Nsx1, Nsx2, Ksym),
    (NmaxS, LmaxS, Kmax, Nav, Kbox), (NgintS, Nx^2), Rint2S, Iint2S, Iint3S,
    IntOrdS (Nx^2 int32), Dint2S, Eint2S; the radial values in FP64
    (USE_DP_INTEGRALS).  Keys as CONF.INT's: Iint2S = Nx^2 k + Nx (na-Nso-1)
    + nc-Nso, Iint3S = Nx (nb-Nso-1) + nd-Nso."""
    ns: int
    nso: int
    khot: int
    nsx1: int
    nsx2: int
    ksym: int
    nmaxs: int
    lmaxs: int
    kmax: int
    nav: int
    kbox: int
    nx: int
    rint2s: np.ndarray
    iint2s: np.ndarray
    iint3s: np.ndarray
    dint2s: np.ndarray
    eint2s: np.ndarray

    @property
    def nsum(self) -> int:
        """RintS: 4 Nav when Nsx2 = 0, else 4 NmaxS (src/pconf.F90:827-831)."""
        return 4 * self.nav if self.nsx2 == 0 else 4 * self.nmaxs


def write_scrc_con(path: str, sc: ScrcCon) -> None:
    n = int(sc.iint2s.size)
    write_records(path, [
        _pack(np.array([sc.ns, sc.nso, sc.nso + 1, sc.khot, sc.nsx1, sc.nsx2, sc.ksym], "<i4")),
        _pack(np.array([sc.nmaxs, sc.lmaxs, sc.kmax, sc.nav, sc.kbox], "<i4")),
        _pack(np.array([n, sc.nx * sc.nx], "<i4")),
        _pack(np.asarray(sc.rint2s, "<f8")),
        _pack(np.asarray(sc.iint2s, "<i4")),
        _pack(np.asarray(sc.iint3s, "<i4")),
        _pack(np.zeros(sc.nx * sc.nx, "<i4")),
        _pack(np.asarray(sc.dint2s, "<f8")),
        _pack(np.asarray(sc.eint2s, "<f8")),
    ])


def read_scrc_con(path: str) -> ScrcCon:
    recs = read_records(path)
    ns, nso, _, khot, nsx1, nsx2, ksym = struct.unpack_from("<7i", recs[0])
    nmaxs, lmaxs, kmax, nav, kbox = struct.unpack_from("<5i", recs[1])
    n, nrd = struct.unpack_from("<2i", recs[2])
    f8 = lambda r: np.frombuffer(r, "<f8").copy()
    return ScrcCon(ns=ns, nso=nso, khot=khot, nsx1=nsx1, nsx2=nsx2, ksym=ksym, nmaxs=nmaxs, lmaxs=lmaxs,
                   kmax=kmax, nav=nav, kbox=kbox, nx=int(round(np.sqrt(nrd))), rint2s=f8(recs[3]),
                   iint2s=np.frombuffer(recs[4], "<i4").copy(), iint3s=np.frombuffer(recs[5], "<i4").copy(),
                   dint2s=f8(recs[7]), eint2s=f8(recs[8]))


def read_conf_dat(path: str) -> Tuple[float, List[Tuple[int, int, int, int]]]:
    raw = np.fromfile(path, dtype="<f8")
    if raw.size % IP6 != 0:
        raise ValueError("CONF.DAT size is not a whole number of records")
    recs = raw.reshape(-1, IP6)
    pq = np.concatenate([recs[0], recs[1], recs[4], recs[5]])
    Z = float(pq[0])
    ns = int(pq[1] + 0.01)
    if abs(pq[19] - 0.98765) < 1e-6:
        iqn = pq.view("<i4")[40:40 + 4 * ns].reshape(ns, 4)
        shells = [(int(a), int(b), int(c), int(d)) for a, b, c, d in iqn]
    else:
        shells = []
        for i in range(ns):
            b = 20 + 6 * i
            c1 = 0.01
            kap = int(pq[b + 4] + np.copysign(c1, pq[b + 4]))
            j2 = int(pq[b + 5] + np.copysign(c1, pq[b + 5]))
            shells.append((int(pq[b] + c1), int(pq[b + 1] + c1), kap, j2))
    return Z, shells


# ----------------------------------------------------------------------------
# CONF.INT
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class ConfInt:
    Ns: int
    Nso: int
    Nsp: int
    Nsx: int
    Ecore: float
    shells: np.ndarray          # (Nsx, 4) int32: n, kappa, l, 2j   (record order)
    nq_nip: np.ndarray          # (Nsp, 2) int32
    kbrt: int
    rint1: np.ndarray           # (nhint,) f64
    iint1: np.ndarray           # (nhint,) i32
    rint2: np.ndarray           # (ngint,) or (ngint, 2) f64  (Coulomb[, Breit])
    iint2: np.ndarray           # (ngint,) i32
    iint3: np.ndarray           # (ngint,) i32
    nx: int = IPX


def write_conf_int(path: str, ci: ConfInt, rint2_dtype="<f8") -> None:
    """Write CONF.INT as pbasc does (src/pbasc.f90:1169-1190).

    Rint2 is written as FP64 (``USE_DP_INTEGRALS``) unless rint2_dtype='<f4'.
    """
    ngint = int(ci.iint2.size)
    recs = [
        _pack(np.array([ci.Ns, ci.Nso, ci.Nsp, ci.Nsx], "<i4"), np.array([ci.Ecore], "<f8")),
        _pack(np.asarray(ci.shells, "<i4").reshape(-1)),
        _pack(np.asarray(ci.nq_nip, "<i4").reshape(-1)),
        _pack(np.array([ci.rint1.size, ci.kbrt], "<i4")),
        _pack(np.asarray(ci.rint1, "<f8")),
        _pack(np.asarray(ci.iint1, "<i4")),
        _pack(np.array([ngint], "<i8"), np.array([0, ci.nx * ci.nx], "<i4")),
        _pack(np.asarray(ci.rint2, rint2_dtype).reshape(-1)),
        _pack(np.asarray(ci.iint2, "<i4")),
        _pack(np.asarray(ci.iint3, "<i4")),
        _pack(np.zeros(ci.nx * ci.nx, "<i8")),
    ]
    write_records(path, recs)


def read_conf_int(path: str, Nsx: int = None) -> ConfInt:
    recs = read_records(path)
    ns, nso, nsp, nsx = struct.unpack_from("<4i", recs[0])
    (ecore,) = struct.unpack_from("<d", recs[0], 16)
    shells = np.frombuffer(recs[1], "<i4").reshape(-1, 4).copy()
    nq_nip = np.frombuffer(recs[2], "<i4").reshape(-1, 2).copy()
    nhint, kbrt = struct.unpack_from("<2i", recs[3])
    rint1 = np.frombuffer(recs[4], "<f8").copy()
    iint1 = np.frombuffer(recs[5], "<i4").copy()
    (ngint,) = struct.unpack_from("<q", recs[6])
    _, nrd = struct.unpack_from("<2i", recs[6], 8)
    nvals = (2 if kbrt else 1) * ngint
    if len(recs[7]) == 8 * nvals:
        rint2 = np.frombuffer(recs[7], "<f8").copy()
    elif len(recs[7]) == 4 * nvals:
        rint2 = np.frombuffer(recs[7], "<f4").astype(np.float64)
    else:
        raise ValueError("Rint2 record size does not match ngint")
    if kbrt:
        rint2 = rint2.reshape(ngint, 2)
    iint2 = np.frombuffer(recs[8], "<i4").copy()
    iint3 = np.frombuffer(recs[9], "<i4").copy()
    nx = int(round(np.sqrt(nrd)))
    return ConfInt(Ns=ns, Nso=nso, Nsp=nsp, Nsx=nsx, Ecore=ecore, shells=shells,
                   nq_nip=nq_nip, kbrt=kbrt, rint1=rint1, iint1=iint1, rint2=rint2,
                   iint2=iint2, iint3=iint3, nx=nx)


# ----------------------------------------------------------------------------
# CONF.GNT
# ----------------------------------------------------------------------------

def write_conf_gnt(path: str, nlx: int, counts: np.ndarray, idx: np.ndarray, gnt: np.ndarray) -> None:
    write_records(path, [
        _pack(np.array([nlx, idx.size], "<i4")),
        _pack(np.asarray(counts, "<i4")),
        _pack(np.asarray(idx, "<i4")),
        _pack(np.asarray(gnt, "<f8")),
    ])


def read_conf_gnt(path: str):
    recs = read_records(path)
    nlx, ngaunt = struct.unpack_from("<2i", recs[0])
    counts = np.frombuffer(recs[1], "<i4").copy()
    idx = np.frombuffer(recs[2], "<i4").copy()
    gnt = np.frombuffer(recs[3], "<f8").copy()
    if idx.size != ngaunt or gnt.size != ngaunt:
        raise ValueError("CONF.GNT sizes inconsistent")
    return nlx, counts, idx, gnt


# ----------------------------------------------------------------------------
# CONF.DET, pCONF.HIJ, CONF.XIJ
# ----------------------------------------------------------------------------

def write_conf_det(path: str, iarr: np.ndarray, nsu: int) -> None:
    """iarr: (Nd, Ne) int32, 1-based sorted positions (src/determinants.f90:448-462)."""
    nd = iarr.shape[0]
    recs = [_pack(np.array([nd, nsu], "<i4"))]
    recs += [np.ascontiguousarray(iarr[i], "<i4").tobytes() for i in range(nd)]
    write_records(path, recs)


def read_conf_det(path: str) -> Tuple[np.ndarray, int]:
    recs = read_records(path)
    nd, nsu = struct.unpack_from("<2i", recs[0])
    iarr = np.stack([np.frombuffer(r, "<i4") for r in recs[1:1 + nd]]) if nd else np.zeros((0, 0), np.int32)
    return iarr.astype(np.int32), nsu


def write_hij(path: str, rowptr: np.ndarray, col: np.ndarray, val: np.ndarray) -> None:
    """pCONF.HIJ global CSR stream (src/matrix_io.f90:689-799).

    rowptr: int64 (Nd+1,) 0-based cumulative; col: int32 1-based; val: float64.
    """
    nd = rowptr.size - 1
    nnz = int(rowptr[-1])
    with open(path, "wb") as f:
        f.write(struct.pack("<qq", nd, nnz))
        f.write(np.asarray(rowptr, "<i8").tobytes())
        f.write(np.asarray(col[:nnz], "<i4").tobytes())
        f.write(np.asarray(val[:nnz], "<f8").tobytes())


def read_hij(path: str):
    with open(path, "rb") as f:
        nd, nnz = struct.unpack("<qq", f.read(16))
        rowptr = np.frombuffer(f.read(8 * (nd + 1)), "<i8").copy()
        col = np.frombuffer(f.read(4 * nnz), "<i4").copy()
        val = np.frombuffer(f.read(8 * nnz), "<f8").copy()
    return rowptr, col, val


def read_conf_hij(path: str, nd: int | None = None):
    """Legacy triple file CONF.HIJ / CONF.JJJ (src/sort.f90:64-70).

    Stream layout: int64 count, then count records (int32 row, int32 column,
    float64 value), 1-based, ordered by row then column, no exact zeros.  With
    nd given, stops like pol's ReadHIJ (src/pol.f90:786-793) at the first
    record whose column is nd + 1.  Returns (row, col, val) arrays.
    """
    rec = np.dtype([("n", "<i4"), ("k", "<i4"), ("t", "<f8")])
    with open(path, "rb") as f:
        (cnt,) = struct.unpack("<q", f.read(8))
        r = np.frombuffer(f.read(rec.itemsize * cnt), rec)
    if r.size != cnt:
        raise ValueError(f"{path}: {r.size} records for a count of {cnt}")
    if nd is not None:
        stop = np.nonzero(r["k"] == nd + 1)[0]
        if stop.size:
            r = r[: stop[0]]
    return r["n"].copy(), r["k"].copy(), r["t"].copy()


def write_xij(path: str, tk: Sequence[float], tj: Sequence[float], vecs: np.ndarray) -> None:
    """CONF.XIJ: one record per level: f64 Tk, f64 Tj, i32 Nd, f64 C(1:Nd).

    vecs has shape (Nlv, Nd).
    """
    nd = vecs.shape[1]
    recs = [_pack(np.array([tk[i], tj[i]], "<f8"), np.array([nd], "<i4"),
                  np.asarray(vecs[i], "<f8")) for i in range(vecs.shape[0])]
    write_records(path, recs)


def read_xij(path: str):
    recs = read_records(path)
    tk, tj, vecs = [], [], []
    for r in recs:
        a, b = struct.unpack_from("<2d", r)
        (nd,) = struct.unpack_from("<i", r, 16)
        tk.append(a)
        tj.append(b)
        vecs.append(np.frombuffer(r, "<f8", count=nd, offset=20).copy())
    return np.array(tk), np.array(tj), np.stack(vecs) if vecs else np.zeros((0, 0))
