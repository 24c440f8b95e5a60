#!/usr/bin/env python3
"""Benchmark of the MI355X pconf path on the C4 workload.

Metric: seconds from H resident on the device to all Nlv roots converged
(FP64 residual <= Crt4, every root within 0.01 cm^-1 of the FP64 oracle),
for the learned correction + FP32 variant next to the plain FP64 Davidson;
plus SpMM GB/s against the 8 TB/s HBM peak and H-build Mnnz/s.

A "step" is one complete Davidson solve of the workload (pci_gpu_davidson
from H resident on the device to the Nlv eigenvalues on the host; the
eigenvectors stay on the device in the timed solves of every variant).  The
JSON line printed last carries roofline and cpu_baseline objects, and
cpu_complete: C2 run to the end on the CPU oracle's ranks beside the device.

Workload C4: 4 valence electrons, s..f shells n=2..7, M_J=0, even parity,
SD+T excitations from eight reference configurations, Coulomb + Breit
(seeded synthetic integrals, generated in pconf's file formats):
Nd = 1,079,368, NumH ~ 2.09e9 stored lower-triangle elements (4.17e9 in both
triangles), Nlv = 8, Crt4 = 1e-7.

Variants timed on the same resident H, each a full solve to Crt4 on the FP64
residual (the admixture stop of pconf's Diag4 is off for all of them, so
every variant ends on the residual criterion):
  plain         FP64 H, FP64 SpMM, diagonal correction (the reference method)
  fp32          FP32-stored H, FP32 SpMM, FP64 refinement and FP64 final residual
  coupled_fp32  fp32 with the tail correction coupled to the head block,
                (r + H[tail, head] t_head) / (theta - D)
  learned_fp32  coupled_fp32 with the MLP gain in place of 1 / (theta - D),
                trained (untimed, before the timed region) on residual/
                correction pairs collected on a smaller space of the same
                generator (n=2..5)
`value` is the learned+FP32 time to roots; fp32 and coupled_fp32 are its
ablation (what the coupling and the network each change).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
HARTREE_CM = 2 * 109737.315709

# HBM bytes per launch of FP32 SpMM shapes (the kernel and its partial-sum
# or combine pass), from two rocprofv3 --pmc passes (FETCH_SIZE x 2.0, the
# factor a plain 1 GiB copy gives, then WRITE_SIZE x 1.0) of
# `bench.py --mode spmm`: keyed by (space, block width, kernel).
# roofline.traffic is null for any other combination.
SPMM32_TRAFFIC = {
    # residual blocks + the list-slot combine (profiles/spmm_c4_traffic.md)
    ("c4", 8, "k_spmm_rblk<16 waves,2 chunks>"): 19.815e9,
    ("c4", 8, "k_spmm_rblk<8 waves,4 chunks>"): 19.812e9,
    ("c4", 8, "k_spmm_rbmf<16 waves,2 chunks>"): 19.819e9,
    ("c4", 8, "k_spmm_rbmf<8 waves,4 chunks>"): 19.818e9,
    # C5: matrix-core block shapes + combine (profiles/spmm_c5_traffic.md)
    ("c5", 20, "k_spmm_rbmf<16 waves,2 chunks>"): 63.0798e9,
    ("c5", 20, "k_spmm_rbmf<8 waves,4 chunks>"): 63.0692e9,
}


# Workloads `--space` selects: the JSON line's config.workload, the smaller
# space of the same generator the learned correction is trained on, and how
# H is held (C5 is built straight into the residual blocks: its FP64 CSR
# beside the FP32 values and the blocks would not fit the 288 GB of HBM).
WORKLOADS = {
    "c4": ("C4: 4 valence e-, s..f n=2..7, M_J=0, even parity, SD+T from 8 refs, Breit", "c4s", "csr"),
    "c5": ("C5: 4 valence e-, s..g (s n=2..16, p..f n<=7, g n=5..6), M_J=0, even parity, SD+T from 5 refs, "
           "Breit, Nlv=20", "c5s", "blocks"),
    "c2": ("C2: 3 valence e-, s..f n=2..14, M_J=1/2, odd parity, SD from 3 refs", "c2s", "csr"),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def space_dir(name):
    from pci_amd import gen
    d = os.path.join(tempfile.gettempdir(), f"pci_bench_{name}_{os.getuid()}")
    if not os.path.exists(os.path.join(d, "CONF.GNT")):
        t = time.time()
        info = gen.write_space(gen.PRESETS[name], d + ".tmp")
        os.replace(d + ".tmp", d) if not os.path.exists(d) else None
        log(f"generated {name}: {info} in {time.time() - t:.1f} s")
    return d


def physical_cores():
    """Physical cores of this host (lscpu), and the CPUs this process may use."""
    import subprocess
    avail = len(os.sched_getaffinity(0))
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        cores = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
    except Exception:
        cores = 0
    return cores, avail


def _cpu_rank(d, r0, r1, hmin, reps):
    """One rank of the CPU oracle: FormH of its rows r0..r1 and its Mxmpy."""
    from oracle.oracle import Oracle, mxmpy_rows
    o = Oracle.from_dir(d)
    t = time.perf_counter()
    rp, col, val, _, _ = o.formh_once(r0, r1)
    t_formh = time.perf_counter() - t
    x = np.random.default_rng(r0).standard_normal(o.Nd)
    t = time.perf_counter()
    for _ in range(reps):
        mxmpy_rows(o.Nd, r0, rp, col, val, hmin, x)
    t_mx = (time.perf_counter() - t) / reps
    return int(rp[-1]), t_formh, t_mx


def cpu_baseline(a, ctx, sp, d, numh, iters_plain):
    """The CPU oracle (the C restatement of pconf's FormH and Mxmpy) on P
    processes, one per core, each owning a row range like a pconf MPI rank.

    Sample: each rank builds (FormH) and applies (Mxmpy, 3 columns) a block
    of rows taken at its own position in the determinant list; the Init4
    dense eigensolve of the Nd0 block runs in numpy/LAPACK on the same P
    threads.  Time to roots of the plain method is extrapolated as
    t_init + iterations x 2 Nlv x (NumH / aggregate Mxmpy rate), with the
    iteration count of the plain FP64 GPU solve; the H build as NumH over
    the aggregate FormH rate.  The Allreduce of each column is not counted."""
    from concurrent.futures import ProcessPoolExecutor
    from threadpoolctl import threadpool_limits
    cores, avail = physical_cores()
    P = max(1, min(a.cpu_workers, avail))
    nd, hmin = sp.Nd, ctx.hmin
    per = max(1000, int(a.cpu_rows))
    starts = [1 + (nd - per) * i // max(1, P - 1) for i in range(P)]
    t = time.perf_counter()
    with ProcessPoolExecutor(P) as ex:
        res = list(ex.map(_cpu_rank, [d] * P, starts, [s + per - 1 for s in starts], [hmin] * P, [3] * P))
    wall = time.perf_counter() - t
    el = sum(r[0] for r in res)
    formh_rate = sum(r[0] / r[1] for r in res)      # elements/s, all ranks
    mx_rate = sum(r[0] / r[2] for r in res)         # elements/s per column, all ranks
    # Init4: the Nd0 x Nd0 block of H - Hmin, dense eigensolve
    nd0 = sp.Nd0
    rp, col, val = ctx.get_csr_rows(1, nd0)
    Z = np.zeros((nd0, nd0))
    rows = np.repeat(np.arange(nd0), np.diff(rp))
    Z[rows, col - 1] = val
    Z = Z + Z.T - np.diag(np.diag(Z)) - hmin * np.eye(nd0)
    with threadpool_limits(P):
        t = time.perf_counter()
        np.linalg.eigh(Z)
        t_init = time.perf_counter() - t
    t_col = numh / mx_rate
    est = t_init + iters_plain * 2 * sp.Nlv * t_col
    return {"value": est, "unit": "s", "cores": P, "kind": "port",
            "host_physical_cores": cores, "host_cpus_available": avail,
            "formh_mnnz_s": formh_rate / 1e6, "hbuild_s_extrapolated": numh / formh_rate,
            "mxmpy_column_s": t_col, "init4_s": t_init,
            "sample": f"{P} oracle ranks, each FormH + 3 Mxmpy columns over {per} rows at its own position "
                      f"({el} stored elements in all, {wall:.1f} s wall); Init4 eigh of the Nd0={nd0} block on "
                      f"{P} threads {t_init:.2f} s; time to roots extrapolated to {iters_plain} iterations x 2 x "
                      f"Nlv={sp.Nlv} columns over NumH={numh} (Allreduce not counted)"}


def cpu_complete(a, name):
    """pconf's method run to the end on the CPU oracle, nothing extrapolated,
    beside the device on the same space.

    CPU: P oracle ranks (processes, one per core) each form the rows they own
    (FormH, every element once), then Diag4 (the oracle's restatement, pconf's
    stopping rule with the admixture test) runs in the parent with LAPACK/BLAS
    on P threads, every Mxmpy column applied by the ranks and summed as
    pconf's Allreduce does.  Device: H built from the same inputs, then the
    plain FP64 solve with the same stopping rule and the FP32 + coupled
    correction solve.  Returns the JSON object."""
    from threadpoolctl import threadpool_limits
    from oracle.davidson import diag4
    from oracle.ranks import OracleRanks
    from pci_amd import _lib, host
    d = space_dir(name)
    sp, ctx = host.load(d, a.device)
    nlv, nd0, crt4, n_it = sp.Nlv, sp.Nd0, sp.Crt4, max(sp.N_it, 300)
    try:
        t = time.perf_counter()
        numh_g, _, hmin_g = ctx.build_h(0)
        tb_g = time.perf_counter() - t
        gpu = {}
        for key, flags in (("plain", 0), ("coupled_fp32", _lib.PCI_DVD_FP32 | _lib.PCI_DVD_COUPLED)):
            ctx.davidson(nlv, nd0, crt4, n_it, flags, vectors=False)
            t = time.perf_counter()
            tk, _, _, st = ctx.davidson(nlv, nd0, crt4, n_it, flags, vectors=False)
            gpu[key] = {"s": time.perf_counter() - t, "iterations": st["iterations"], "tk": tk}
    finally:
        ctx.close()
    cores, avail = physical_cores()
    P = max(1, min(a.cpu_workers, avail))
    with OracleRanks(d, sp.Nd, P, nlv) as R:
        t_formh, numh, hmin, _ = R.build()
        with threadpool_limits(P):
            t = time.perf_counter()
            tk_c, _, it_c, _ = diag4(None, None, None, hmin, nd0, nlv, crt4, n_it, ops=R)
            t_diag = time.perf_counter() - t
        P = R.p
    cpu_s = t_formh + t_diag
    err = {k: float(np.abs(v["tk"] - tk_c).max() * HARTREE_CM) for k, v in gpu.items()}
    out = {"workload": WORKLOADS[name][0], "Nd": sp.Nd, "NumH": numh, "Nlv": nlv, "Nd0": nd0, "Crt4": crt4,
           "cpu": {"kind": "port", "cores": P, "host_physical_cores": cores, "host_cpus_available": avail,
                   "formh_s": t_formh, "formh_mnnz_s": numh / t_formh / 1e6, "diag4_s": t_diag,
                   "iterations": it_c, "total_s": cpu_s},
           "gpu": {"hbuild_s": tb_g, "plain_fp64_s": gpu["plain"]["s"],
                   "plain_iterations": gpu["plain"]["iterations"],
                   "coupled_fp32_s": gpu["coupled_fp32"]["s"],
                   "coupled_fp32_iterations": gpu["coupled_fp32"]["iterations"],
                   "total_plain_s": tb_g + gpu["plain"]["s"]},
           "same_h": bool(numh == numh_g and hmin == hmin_g),
           "max_root_diff_cm": err, "roots_agree_0.01cm": all(e < 0.01 for e in err.values()),
           "speedup_total_plain": cpu_s / (tb_g + gpu["plain"]["s"]),
           "speedup_diag4_plain": t_diag / gpu["plain"]["s"],
           "speedup_diag4_coupled_fp32": t_diag / gpu["coupled_fp32"]["s"],
           "stop": "pconf's rule on both sides: max residual <= Crt4 or admixture <= 1e-6",
           "sample": f"complete run: FormH of all {numh} stored elements on {P} oracle ranks, then Diag4 to "
                     f"convergence ({it_c} iterations) with every Mxmpy column applied by the ranks; "
                     f"nothing extrapolated"}
    log(f"complete {name} on the CPU ({P} ranks): FormH {t_formh:.1f} s + Diag4 {t_diag:.1f} s ({it_c} it); "
        f"device: build {tb_g:.3f} s + plain {gpu['plain']['s']:.3f} s ({gpu['plain']['iterations']} it); "
        f"roots differ by {max(err.values()):.1e} cm^-1")
    return out


def train_network(name, device, flags):
    """Collect residual/correction pairs on a smaller space (with the
    correction the network will refine: flags) and train the MLP there."""
    from pci_amd import _lib, host
    t = time.perf_counter()
    sp, c = host.load(space_dir(name), device)
    try:
        c.build_h()
        _, _, vref, _ = c.davidson(sp.Nlv, sp.Nd0, 1e-10, 1000, _lib.PCI_DVD_NOADMIX)
        X, b, r, T = c.mlp_collect(sp.Nlv, sp.Nd0, sp.Crt4, sp.N_it, vref, stride=1, cap=1 << 21, flags=flags)
        c.mlp_init(1)
        hist = c.mlp_train(X, b, r, T, epochs=20, batch=8192, lr=1e-3)
        w = c.mlp_get_weights()
    finally:
        c.close()
    info = {"train_space": name, "train_Nd": sp.Nd, "pairs": int(len(r)), "epochs": 20,
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
            "setup_s": time.perf_counter() - t}
    log(f"trained on {name}: {info}")
    return w, info


def copy_calibration(a):
    """A device-to-device copy of 1 GiB, a.spmm_reps times (bytes known exactly)."""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device=f"cuda:{a.device}")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.spmm_reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.spmm_reps
    print(json.dumps({"mode": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "ms": ms,
                      "gbs": 8 * n / (ms * 1e-3) / 1e9}), flush=True)


def batch_blocks(device, steps):
    """C3: the twelve (J, parity) blocks of one 4-electron atom, 36k to 376k
    determinants each (gen.c3_blocks), each held as residual blocks, FP32
    mode, solved one after the other and all at once
    (pci_gpu_davidson_batch).  H builds and a first (tuning) solve of every
    block are untimed."""
    from pci_amd import _lib, gen, host
    specs = gen.c3_blocks()
    blocks = []
    try:
        for sp_ in specs:
            d = os.path.join(tempfile.gettempdir(), f"pci_bench_{sp_.name}_{os.getuid()}")
            if not os.path.exists(os.path.join(d, "CONF.GNT")):
                gen.write_space(sp_, d + ".tmp")
                os.replace(d + ".tmp", d) if not os.path.exists(d) else None
            sp, c = host.load(d, device)
            blocks.append((sp, c))
            c.build_h(_lib.PCI_BUILD_BLOCKS)
        flags = _lib.PCI_DVD_FP32
        params = [dict(nlv=sp.Nlv, nd0=sp.Nd0, crt4=sp.Crt4, n_it=sp.N_it, flags=flags) for sp, _ in blocks]
        ctxs = [c for _, c in blocks]
        for (sp, c), q in zip(blocks, params):  # SpMM tuning per matrix, on a quiet GPU
            c.davidson(q["nlv"], q["nd0"], q["crt4"], q["n_it"], flags)
        _lib.davidson_batch(ctxs, params)
        seq, bat = [], []
        for _ in range(steps):
            t = time.perf_counter()
            out_seq = [c.davidson(q["nlv"], q["nd0"], q["crt4"], q["n_it"], flags) for c, q in zip(ctxs, params)]
            seq.append(time.perf_counter() - t)
            t = time.perf_counter()
            out = _lib.davidson_batch(ctxs, params)
            bat.append(time.perf_counter() - t)
        res = {"blocks": len(blocks), "Nd": [sp.Nd for sp, _ in blocks], "NumH": [c.numh for _, c in blocks],
               "Nlv": 6, "variant": "FP32 H (residual blocks), FP64 refinement",
               "sequential_s": float(np.mean(seq)),
               "batched_s": float(np.mean(bat)), "speedup": float(np.mean(seq) / np.mean(bat)),
               "iterations": [o[3]["iterations"] for o in out],
               "converged_batched": [o[3]["converged"] for o in out],
               "converged_single": [o[3]["converged"] for o in out_seq],
               "max_dE_cm_batch_vs_single": max(float(np.abs(p[0] - q[0]).max()) for p, q in zip(out_seq, out))
               * HARTREE_CM}
    finally:
        for _, c in blocks:
            c.close()
    log(f"C3 batch: {len(blocks)} blocks Nd {min(res['Nd'])}..{max(res['Nd'])}: sequential {res['sequential_s']:.3f} s, "
        f"batched {res['batched_s']:.3f} s ({res['speedup']:.2f}x)")
    return res


def spmm_mode(a, sp, ctx):
    widths = [int(w) for w in a.widths.split(",")] if a.widths else [sp.Nlv]
    for nv in widths:
        x = np.random.default_rng(1).standard_normal((sp.Nd, nv))
        out = {}
        for prec in (32, 64):
            _, ms, nb = ctx.spmm(x, prec, a.spmm_reps)
            out[f"fp{prec}"] = {"ms": ms, "bytes": nb, "gbs": nb / (ms * 1e-3) / 1e9,
                                "kernel": ctx.spmm_variant(nv, prec)}
        print(json.dumps({"mode": "spmm", "space": a.space, "nvec": nv, "reps": a.spmm_reps, **out}), flush=True)


def solve_mode(a, sp, ctx):
    """The `value` variant alone: the network trained (untimed), --warmup
    solves, then --steps timed solves, one JSON line."""
    from pci_amd import _lib
    w, _ = train_network(a.train_space, a.device, _lib.PCI_DVD_COUPLED)
    ctx.mlp_set_weights(w)
    fl = _lib.PCI_DVD_FP32 | _lib.PCI_DVD_COUPLED | _lib.PCI_DVD_LEARNED | _lib.PCI_DVD_NOADMIX
    n_it = max(sp.N_it, 300)
    for _ in range(a.warmup):
        ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, n_it, fl, vectors=False)
    times, its = [], []
    for _ in range(a.steps):
        t = time.perf_counter()
        _, _, _, st = ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, n_it, fl, vectors=False)
        times.append(time.perf_counter() - t)
        its.append(st["iterations"])
        log(f"solve: {times[-1]:.3f} s, {st['iterations']} it, residual {st['max_residual']:.2e}")
    print(json.dumps({"mode": "solve", "space": a.space, "s": times, "iterations": its}), flush=True)


def csf_space_dir(d):
    """The space directory d with kCSF = 1 in its ci.in (other files linked)."""
    from pci_amd import formats as F
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
        os.replace(tmp, c) if not os.path.exists(c) else None
    return c


def _csf_cpu_sample(d, nc_neq, ndcs, ccj, budget_s):
    """The oracle's formh_sym (one core) over configurations spread across the
    space until budget_s is spent: (seconds, determinant pairs, configurations)."""
    from oracle import csf
    from oracle.oracle import Oracle
    o = Oracle.from_dir(d, kcsf=1)
    b = csf.CsfBasis(o, ndcs, ccj)
    have = np.flatnonzero(b.conf_ncsf > 0) + 1
    order = have[np.random.default_rng(0).permutation(have.size)]
    t_all, cand, n = 0.0, 0, 0
    for ic in order:
        st = {}
        t = time.perf_counter()
        csf.formh_sym(o, b, int(ic), int(ic), stats=st)
        t_all += time.perf_counter() - t
        cand += st["candidates"]
        n += 1
        if t_all > budget_s:
            break
    return t_all, cand, n


def csf_mode(a):
    """kCSF = 1 on the workload's space, loaded through the Fortran host: the
    CSF basis (jbasis), H in it (formh_sym) and the time to the Nlv roots in
    that basis, plain FP64 and FP32 + FP64 refinement; the SpMM on H_csf
    against HBM; the oracle's formh_sym on a sample of configurations as the
    CPU baseline of the build.  One JSON line."""
    from pci_amd import _lib, host
    from pci_amd import formats as F
    d = csf_space_dir(space_dir(a.space))
    sp, ctx = host.load(d, a.device)
    try:
        nc_neq, nneq = sp.csf_classes()
        t = time.perf_counter()
        ndcs, ncsf, nccj = ctx.csf_basis(nc_neq, nneq)
        t_basis = time.perf_counter() - t
        t = time.perf_counter()
        numh, numzero, hmin = ctx.build_h_csf()
        t_h = time.perf_counter() - t
        info = ctx.build_info()
        log(f"CSF basis: {nneq} classes, Ncsf={ncsf} in {t_basis:.3f} s; H_csf NumH={numh} numzero={numzero} "
            f"in {t_h:.3f} s ({info['candidates']:.3e} determinant pairs)")
        nc4 = F.read_conf_inp(os.path.join(d, "CONF.INP")).Nc4
        maxnd0 = F.read_ci_in(os.path.join(d, "ci.in"))["MAXND0"]
        n1, nd0 = 0, 0
        for ic in range(min(nc4, nc_neq.size)):
            n1 += int(ndcs[nc_neq[ic] - 1])
            if n1 < maxnd0 + 1:
                nd0 = n1
        nlv, crt4 = min(sp.Nlv, nd0), sp.Crt4
        NOADMIX = _lib.PCI_DVD_NOADMIX
        tk_ref, _, _, st_ref = ctx.davidson(nlv, nd0, 1e-10, 1000, NOADMIX)
        runs = {}
        for name, flags in (("plain", 0), ("fp32", _lib.PCI_DVD_FP32),
                            ("coupled_fp32", _lib.PCI_DVD_FP32 | _lib.PCI_DVD_COUPLED)):
            for _ in range(a.warmup):
                ctx.davidson(nlv, nd0, crt4, 300, flags | NOADMIX, vectors=False)
            times, its, errs, ok = [], [], [], True
            for _ in range(a.steps):
                t = time.perf_counter()
                tk, _, _, st = ctx.davidson(nlv, nd0, crt4, 300, flags | NOADMIX, vectors=False)
                times.append(time.perf_counter() - t)
                its.append(st["iterations"])
                errs.append(float(np.abs(tk - tk_ref).max() * HARTREE_CM))
                ok = ok and st["converged"] == 1 and errs[-1] < 0.01
            runs[name] = {"s": float(np.mean(times)), "iterations": its, "max_err_cm": max(errs), "ok": ok,
                          "spmm_ms": st["t_spmm_ms"], "total_ms": st["t_total_ms"]}
            log(f"CSF {name}: {runs[name]['s']:.3f} s, iterations {its}, max err {max(errs):.2e} cm^-1")
        x = np.random.default_rng(1).standard_normal((ncsf, nlv))
        _, ms32, b32 = ctx.spmm(x, 32, a.spmm_reps)
        _, ms64, b64 = ctx.spmm(x, 64, a.spmm_reps)
        var32, var64 = ctx.spmm_variant(nlv, 32), ctx.spmm_variant(nlv, 64)
        gbs32 = b32 / (ms32 * 1e-3) / 1e9
        res = {
            "mode": "csf",
            "metric": "s to Nlv roots in the CSF basis (kCSF=1), FP32+FP64 refinement vs plain FP64",
            "value": runs["fp32"]["s"], "unit": "s", "higher_is_better": False, "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup,
            "config": {"workload": a.workload + " in the CSF basis (J = Jm)", "Nd": sp.Nd, "Nc": sp.Nc,
                       "classes": nneq, "Ncsf": ncsf, "nccj": nccj, "NumH_csf": numh, "numzero": numzero,
                       "Nlv": nlv, "Nd0": nd0, "Crt4": crt4},
            "basis_s": t_basis, "hbuild_s": t_h, "hbuild_kernel_ms": info["build_ms"],
            "hbuild_det_pairs": info["candidates"], "hbuild_det_pairs_per_s": info["candidates"] / t_h,
            "plain_fp64_s": runs["plain"]["s"], "speedup_fp32_vs_plain": runs["plain"]["s"] / runs["fp32"]["s"],
            "coupled_fp32_s": runs["coupled_fp32"]["s"],
            "runs": runs, "reference_iterations": st_ref["iterations"],
            "roofline": {"bound": "hbm", "achieved": gbs32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs32 / HBM_PEAK_GBS, "traffic": None, "kernel": var32,
                         "bytes_per_launch": b32, "avg_launch_ms": ms32,
                         "measured_in": f"{a.spmm_reps} back-to-back FP32 SpMM launches on H_csf"},
            "spmm_fp64_gbs": b64 / (ms64 * 1e-3) / 1e9, "spmm_fp64_kernel": var64,
        }
        if not a.no_cpu:
            ccj = ctx.csf_get_basis()
            t_s, cand_s, nconf_s = _csf_cpu_sample(d, nc_neq, ndcs, ccj, 12.0)
            est = t_s * info["candidates"] / max(cand_s, 1)
            res["cpu_baseline"] = {"value": est, "unit": "s (formh_sym, extrapolated)", "cores": 1, "kind": "port",
                                   "sample": f"the oracle's formh_sym on {nconf_s} configurations chosen at random "
                                             f"({cand_s:.3e} of {info['candidates']:.3e} determinant pairs) in "
                                             f"{t_s:.1f} s, scaled by the pair count",
                                   "gpu_speedup_build": est / t_h}
            log(f"cpu baseline (formh_sym): {est:.1f} s extrapolated vs GPU {t_h:.3f} s")
        print(json.dumps(res), flush=True)
    finally:
        ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--space", default="c4")
    ap.add_argument("--train-space", default=None, help="space the network is trained on (default: the workload's)")
    ap.add_argument("--build", choices=["auto", "csr", "blocks"], default="auto",
                    help="hold H as the CSR (+ SpMM copies) or only as residual blocks (auto: per workload)")
    ap.add_argument("--device", type=int, default=int(os.environ.get("LOCAL_RANK", "0")))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-blocks (C3) leg")
    ap.add_argument("--cpu-workers", type=int, default=16, help="CPU oracle ranks (capped at the CPUs available)")
    ap.add_argument("--cpu-rows", type=int, default=6000, help="rows per CPU oracle rank in the sample")
    ap.add_argument("--complete", default="c2",
                    help="space run to the end on the CPU oracle beside the device ('' to skip)")
    ap.add_argument("--spmm-reps", type=int, default=20)
    ap.add_argument("--widths", default="", help="--mode spmm: comma-separated block widths (default Nlv)")
    ap.add_argument("--mode", choices=["full", "spmm", "copy", "solve", "csf"], default="full",
                    help="spmm: only the resident-H SpMM launches (for counter runs); "
                         "copy: a plain device copy of known size (counter calibration); "
                         "solve: only the learned+FP32 solves (for kernel traces); "
                         "csf: the CSF-basis (kCSF=1) build and solve")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if a.mode == "copy":
        return copy_calibration(a)

    from pci_amd import _lib, host
    desc, train, how = WORKLOADS.get(a.space, (a.space, a.space + "s", "csr"))
    a.workload = desc
    a.train_space = a.train_space or train
    a.build = how if a.build == "auto" else a.build
    if a.mode == "csf":
        return csf_mode(a)
    d = space_dir(a.space)
    t = time.perf_counter()
    sp, ctx = host.load(d, a.device)
    try:
        log(f"space loaded: Nd={sp.Nd} Nc={sp.Nc} Nst={sp.Nst} Nd0={sp.Nd0} Nlv={sp.Nlv} "
            f"in {time.perf_counter() - t:.2f} s")
        mem0 = ctx.mem_usage()
        t = time.perf_counter()
        numh, numzero, hmin = ctx.build_h(_lib.PCI_BUILD_BLOCKS if a.build == "blocks" else 0)
        t_build = time.perf_counter() - t
        a.build_bytes = ctx.mem_usage() - mem0
        log(f"H built ({a.build}): NumH={numh} numzero={numzero} Hmin={hmin:.8f} in {t_build:.3f} s, "
            f"{a.build_bytes / 1e9:.1f} GB resident")
        if a.mode == "spmm":
            return spmm_mode(a, sp, ctx)
        if a.mode == "solve":
            return solve_mode(a, sp, ctx)
        run_full(a, sp, ctx, d, numh, numzero, t_build)
    finally:
        ctx.close()


def run_full(a, sp, ctx, d, numh, numzero, t_build):
    from pci_amd import _lib
    nlv, nd0, crt4, n_it = sp.Nlv, sp.Nd0, sp.Crt4, max(sp.N_it, 300)
    NOADMIX = _lib.PCI_DVD_NOADMIX
    hb = ctx.build_info()
    mem_h = ctx.mem_usage()
    # FP64 oracle eigenvalues: plain FP64 GPU Davidson converged to 1e-10
    tk_ref, _, _, st_ref = ctx.davidson(nlv, nd0, 1e-10, 1000, NOADMIX)
    log(f"FP64 oracle: {st_ref['iterations']} it, residual {st_ref['max_residual']:.2e}")
    # pconf's own stopping rule (admixture test on) for the record
    _, _, _, st_pc = ctx.davidson(nlv, nd0, crt4, n_it, 0)
    log(f"plain FP64 with pconf's admixture stop: {st_pc['iterations']} it, converged {st_pc['converged']}, "
        f"residual {st_pc['max_residual']:.2e}")

    # learned correction: train on a smaller space of the same generator (untimed set-up)
    w, train_info = train_network(a.train_space, a.device, _lib.PCI_DVD_COUPLED)
    ctx.mlp_set_weights(w)

    def timed(flags):
        for _ in range(a.warmup):
            ctx.davidson(nlv, nd0, crt4, n_it, flags | NOADMIX, vectors=False)
        times, errs, its, sts = [], [], [], []
        for _ in range(a.steps):
            t = time.perf_counter()
            tk, _, _, st = ctx.davidson(nlv, nd0, crt4, n_it, flags | NOADMIX, vectors=False)
            times.append(time.perf_counter() - t)
            errs.append((np.abs(tk - tk_ref) * HARTREE_CM).tolist())
            its.append(st["iterations"])
            sts.append(st)
        ok = all(max(e) < 0.01 for e in errs) and all(s["converged"] == 1 for s in sts)
        return {"s": float(np.mean(times)), "errs": errs, "its": its, "stats": sts, "ok": ok}

    runs = {}
    F32, CPL, LRN = _lib.PCI_DVD_FP32, _lib.PCI_DVD_COUPLED, _lib.PCI_DVD_LEARNED
    for name, flags in (("plain", 0), ("fp32", F32), ("coupled_fp32", F32 | CPL),
                        ("learned_fp32", F32 | CPL | LRN)):
        runs[name] = r = timed(flags)
        st = r["stats"][-1]
        log(f"{name}: {r['s']:.3f} s, iterations {r['its']}, max err {max(map(max, r['errs'])):.2e} cm^-1, "
            f"residual {st['max_residual']:.2e}, learned used {st['learned_used']} rejected {st['learned_rejected']}")
    fast = runs["learned_fp32"]
    # the FP32 SpMM launches inside the timed learned+FP32 solves (HIP events
    # on the context stream around each launch)
    n32_in = sum(q["spmm_fp32"] for q in fast["stats"])
    ms32_in = sum(q["t_spmm32_ms"] for q in fast["stats"]) / max(n32_in, 1)
    x = np.random.default_rng(1).standard_normal((sp.Nd, nlv))
    _, ms32, b32 = ctx.spmm(x, 32, a.spmm_reps)
    _, ms64, b64 = ctx.spmm(x, 64, a.spmm_reps)
    var32, var64 = ctx.spmm_variant(nlv, 32), ctx.spmm_variant(nlv, 64)
    gbs32_in = b32 / (ms32_in * 1e-3) / 1e9
    log(f"SpMM fp32: {ms32_in:.3f} ms {gbs32_in:.0f} GB/s in the timed solves ({n32_in} launches), "
        f"{ms32:.3f} ms {b32 / ms32 / 1e6:.0f} GB/s back to back ({var32}); "
        f"fp64: {ms64:.3f} ms {b64 / ms64 / 1e6:.0f} GB/s ({var64})")
    # the build's HBM write floor: the lower-triangle CSR it stores (column +
    # FP64 value per element, row pointers), or for a blocks-only build the
    # bytes it left resident (FP64 and FP32 blocks and their index)
    floor_bytes = numh * 12 + 8 * (sp.Nd + 1) if a.build == "csr" else a.build_bytes
    csr32 = (2 * numh - sp.Nd) * 8 + 8 * (sp.Nd + 1) + 2 * sp.Nd * nlv * 4
    res = {
        "metric": "s to Nlv roots at |dE|<0.01 cm^-1: learned+FP32 vs plain FP64; SpMM GB/s",
        "value": fast["s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": fast["s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 SpMM / f64",
        "data": "synthetic (seeded integrals in pconf formats)",
        "config": {"workload": a.workload, "H_storage": a.build,
                   "Nd": sp.Nd, "Nc": sp.Nc, "NumH": numh, "nnz_full": int(2 * numh - sp.Nd),
                   "Nlv": nlv, "Crt4": crt4, "Nd0": nd0,
                   "variant": "learned correction on the head-coupled residual + FP32 H, FP64 refinement",
                   "stop": "FP64 residual <= Crt4 (admixture stop off for every variant)",
                   "timed": "pci_gpu_davidson from H resident to the roots on the host; eigenvectors stay "
                            "on the device (not copied to the host) in every variant"},
        "roots_within_0.01cm": bool(fast["ok"]),
        "plain_fp64_s": runs["plain"]["s"],
        "fp32_s": runs["fp32"]["s"],
        "speedup_vs_plain_fp64": runs["plain"]["s"] / fast["s"],
        "speedup_fp32_only": runs["plain"]["s"] / runs["fp32"]["s"],
        "coupled_fp32_s": runs["coupled_fp32"]["s"],
        "speedup_coupled_fp32": runs["plain"]["s"] / runs["coupled_fp32"]["s"],
        "learned_vs_coupled": {"time_ratio": fast["s"] / runs["coupled_fp32"]["s"],
                               "iterations": [runs["coupled_fp32"]["its"][-1], fast["its"][-1]]},
        "all_variants_converged": all(r["ok"] for r in runs.values()),
        "iterations": {k: r["its"] for k, r in runs.items()},
        "pconf_admixture_stop": {"iterations": st_pc["iterations"], "converged": st_pc["converged"],
                                 "residual": st_pc["max_residual"]},
        "err_cm_per_root": {k: r["errs"][-1] for k, r in runs.items()},
        "learned": {"used": fast["stats"][-1]["learned_used"], "rejected": fast["stats"][-1]["learned_rejected"],
                    **train_info},
        "breakdown_ms": {k: {"total": r["stats"][-1]["t_total_ms"], "init": r["stats"][-1]["t_init_ms"],
                             "spmm": r["stats"][-1]["t_spmm_ms"], "spmm_fp32": r["stats"][-1]["spmm_fp32"],
                             "spmm_fp64": r["stats"][-1]["spmm_fp64"]} for k, r in runs.items()},
        "hbuild_s": t_build,
        "hbuild_mnnz_s": numh / t_build / 1e6,
        "hbuild_candidates": hb["candidates"],
        "hbuild_candidates_per_s": hb["candidates"] / t_build,
        "hbuild_kernel_ms": hb["build_ms"],
        "hbuild_write_floor_frac": floor_bytes / (t_build * HBM_PEAK_GBS * 1e9),
        "spmm_fp32_gbs_back_to_back": b32 / (ms32 * 1e-3) / 1e9,
        "spmm_fp64_gbs": b64 / (ms64 * 1e-3) / 1e9,
        "spmm_fp64_kernel": var64,
        "device_mem_gb": {"after_build": mem_h / 1e9, "after_solves": ctx.mem_usage() / 1e9},
        "roofline": {"bound": "hbm", "achieved": gbs32_in, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs32_in / HBM_PEAK_GBS,
                     "traffic": SPMM32_TRAFFIC.get((a.space, nlv, var32)),
                     "kernel": var32, "bytes_per_launch": b32, "avg_launch_ms": ms32_in, "launches": n32_in,
                     "measured_in": "the FP32 SpMM launches of the timed learned+FP32 solves",
                     "back_to_back_ms": ms32,
                     # the CSR stream the same product would need (value + 4-byte column per stored
                     # element, row pointers, X and Y once), over the same launch time
                     "csr_equivalent_bytes": csr32, "csr_equivalent_gbs": csr32 / (ms32_in * 1e-3) / 1e9},
    }
    if not a.no_batch:
        res["batch_c3"] = batch_blocks(a.device, a.steps)
    if not a.no_cpu:
        cb = cpu_baseline(a, ctx, sp, d, numh, runs["plain"]["its"][-1])
        res["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value']:.1f} s to roots (extrapolated, {cb['cores']} ranks), "
            f"FormH {cb['formh_mnnz_s']:.1f} Mnnz/s")
        if a.complete:
            res["cpu_complete"] = cpu_complete(a, a.complete)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
