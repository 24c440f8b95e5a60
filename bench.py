#!/usr/bin/env python3
"""Benchmark of the MI355X pconf path on the C2 workload.

Metric: seconds from H resident on the device to all Nlv roots converged
(FP64 residual <= Crt4, every root within 0.01 cm^-1 of the FP64 oracle), for
the fast variant (FP32-stored H with FP64 refinement) next to the plain FP64
Davidson; plus SpMM GB/s against the 8 TB/s HBM peak and H-build Mnnz/s.

A "step" is one complete Davidson solve of the workload.  The JSON line
printed last carries roofline and cpu_baseline objects.

Workload C2: 3 valence electrons, s..f shells n=2..14, M_J=1/2, odd parity,
SD excitations from three reference configurations (seeded synthetic
integrals, generated in pconf's file formats).

Variants timed on the same resident H, each a full solve to Crt4:
  plain   FP64 H, FP64 SpMM, diagonal correction (the reference method)
  fp32    FP32-stored H, FP32 SpMM, FP64 refinement and FP64 final residual
  coupled fp32 + the head-to-tail coupled correction (r + H[tail,head] t_head)
  learned coupled + the MLP gain, trained (untimed, before the timed region)
  on residual/correction pairs collected on a smaller sibling space (the
  same references on n=2..8).
`value` is the learned+FP32 time to roots.
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

# HBM bytes per launch of the FP32 SpMM shapes on C2 (Nlv = 10), measured with
# rocprofv3 on `bench.py --mode spmm`: reads = 128 RDREQ_128B + 64 RDREQ_64B +
# 32 RDREQ_32B (equal to 2 x FETCH_SIZE, the factor a plain 1 GiB copy gives),
# writes = WRITE_SIZE.  The k_spmm_sell figures include the panel-group
# partial-sum pass (k_spmm_panel_sum) of the same application.  Used for
# roofline.traffic when the autotuner picks one of them on C2; null otherwise.
C2_SPMM32_TRAFFIC = {
    "k_spmm_sell<16 waves>": 3.012e9,
    "k_spmm_sell<8 waves>": 3.008e9,
    "k_spmm_sell<16 waves,4|2 chunks/wave>": 3.113e9,
    "k_spmm_xlds<2 rows/wave>": 4.076e9,
    "k_spmm_xlds<2 rows/wave,U=2>": 4.022e9,
    "k_spmm_xlds<1 row/wave,U=2>": 4.009e9,
    "k_spmm_xlds<4 rows/wave>": 3.916e9,
    "k_spmm_lane<U=2,4 rows/block>": 3.747e9,
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


def cpu_baseline(ctx, sp, d, iters_plain, nd0):
    """CPU oracle (the C restatement of pconf, one thread) on a bounded sample.

    Sample: the Init4 dense eigensolve of the Nd0 block, plus three Mxmpy
    column applications over the full lower-triangle H; time-to-roots of the
    plain method is extrapolated as t_init + iterations x 2 Nlv x t_column,
    with the iteration count of the plain FP64 GPU run on the same space.
    """
    from threadpoolctl import threadpool_limits
    from oracle.oracle import mxmpy
    rp, col, val = ctx.get_csr()
    hmin = float(val.min())
    nlv = sp.Nlv
    with threadpool_limits(1):
        t = time.perf_counter()
        Z = np.zeros((nd0, nd0))
        for n in range(nd0):
            j0, j1 = rp[n], rp[n + 1]
            ks = col[j0:j1] - 1
            vs = val[j0:j1].copy()
            vs[ks == n] -= hmin
            m = ks < nd0
            Z[n, ks[m]] = vs[m]
            Z[ks[m], n] = vs[m]
        np.linalg.eigh(Z)
        t_init = time.perf_counter() - t
        x = np.random.default_rng(0).standard_normal(sp.Nd)
        t = time.perf_counter()
        ncol = 3
        for _ in range(ncol):
            mxmpy(rp, col, val, hmin, x)
        t_col = (time.perf_counter() - t) / ncol
    est = t_init + iters_plain * 2 * nlv * t_col
    return {"value": est, "unit": "s", "cores": 1, "kind": "port",
            "sample": f"oracle Init4 eigensolve (Nd0={nd0}) {t_init:.2f} s + {ncol} Mxmpy columns "
                      f"({t_col:.3f} s each over NumH={int(rp[-1])}); extrapolated to {iters_plain} "
                      f"iterations x 2 x Nlv={nlv} columns"}


def train_network(name, device):
    """Collect residual/correction pairs on a smaller space and train the MLP there."""
    from pci_amd import host
    t = time.perf_counter()
    sp, c = host.load(space_dir(name), device)
    c.build_h()
    _, _, vref, _ = c.davidson(sp.Nlv, sp.Nd0, 1e-10, 1000, 0)
    X, b, r, T = c.mlp_collect(sp.Nlv, sp.Nd0, sp.Crt4, sp.N_it, vref, stride=1, cap=1 << 21)
    c.mlp_init(1)
    hist = c.mlp_train(X, b, r, T, epochs=20, batch=8192, lr=1e-3)
    w = c.mlp_get_weights()
    c.close()
    info = {"train_space": name, "train_Nd": sp.Nd, "pairs": int(len(r)), "epochs": 20,
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
            "setup_s": time.perf_counter() - t}
    log(f"trained on {name}: {info}")
    return w, info


def cpu_formh_sample(d, nd, nrows=240):
    """Oracle FormH on evenly spaced row blocks; Mnnz/s of stored lower elements."""
    from oracle.oracle import Oracle
    o = Oracle.from_dir(d)
    blocks = 12
    per = nrows // blocks
    tot_el, tot_t = 0, 0.0
    for b in range(blocks):
        r0 = 1 + (nd - per) * b // (blocks - 1)
        t = time.perf_counter()
        rp, _, _, _, _ = o.formh(r0, r0 + per - 1)
        tot_t += time.perf_counter() - t
        tot_el += int(rp[-1])
    return tot_el / tot_t / 1e6, tot_el, tot_t


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
    """C3-style batch: the twelve (J, parity) blocks of one 3-electron atom
    (uneven sizes), FP32 mode, solved one after the other and all at once
    (pci_gpu_davidson_batch).  H builds and a first (tuning) solve of every
    block are untimed."""
    from pci_amd import _lib, gen, host
    specs = gen.c3_specs(nmax=8, max_nd0=3000)
    blocks = []
    for sp_ in specs:
        d = os.path.join(tempfile.gettempdir(), f"pci_bench_{sp_.name}_n8_m3000_{os.getuid()}")
        if not os.path.exists(os.path.join(d, "CONF.GNT")):
            gen.write_space(sp_, d + ".tmp")
            os.replace(d + ".tmp", d) if not os.path.exists(d) else None
        sp, c = host.load(d, device)
        c.build_h()
        blocks.append((sp, c))
    flags = _lib.PCI_DVD_FP32
    params = [dict(nlv=sp.Nlv, nd0=sp.Nd0, crt4=sp.Crt4, n_it=sp.N_it, flags=flags) for sp, _ in blocks]
    ctxs = [c for _, c in blocks]
    for (sp, c), q in zip(blocks, params):  # SpMM tuning per matrix, on a quiet GPU
        c.davidson(q["nlv"], q["nd0"], q["crt4"], q["n_it"], flags)
    _lib.davidson_batch(ctxs, params)
    seq, bat, its = [], [], None
    for _ in range(steps):
        t = time.perf_counter()
        out_seq = [c.davidson(q["nlv"], q["nd0"], q["crt4"], q["n_it"], flags) for c, q in zip(ctxs, params)]
        seq.append(time.perf_counter() - t)
        t = time.perf_counter()
        out = _lib.davidson_batch(ctxs, params)
        bat.append(time.perf_counter() - t)
        its = [o[3]["iterations"] for o in out]
        conv = [o[3]["converged"] for o in out]
        conv_seq = [o[3]["converged"] for o in out_seq]
        dmax = max(float(np.abs(a[0] - b[0]).max()) for a, b in zip(out_seq, out)) * HARTREE_CM
    res = {"blocks": len(blocks), "Nd": [sp.Nd for sp, _ in blocks], "NumH": [c.numh for _, c in blocks],
           "Nlv": 6, "variant": "FP32 H, FP64 refinement", "sequential_s": float(np.mean(seq)),
           "batched_s": float(np.mean(bat)), "speedup": float(np.mean(seq) / np.mean(bat)), "iterations": its,
           "converged_batched": conv, "converged_single": conv_seq,
           "max_dE_cm_batch_vs_single": dmax}
    for _, c in blocks:
        c.close()
    log(f"C3 batch: {len(blocks)} blocks Nd {min(res['Nd'])}..{max(res['Nd'])}: sequential {res['sequential_s']:.3f} s, "
        f"batched {res['batched_s']:.3f} s ({res['speedup']:.2f}x)")
    return res


def c4_throughput(a, sp, ctx, t_build, numh):
    """The C4 shape on one GPU: H build rate, SpMM GB/s per precision, and
    seconds per Davidson iteration for the plain FP64 method and the FP32
    mode over a fixed number of iterations (pconf's Davidson does not reach
    Crt4 on this synthetic spectrum, with either precision or on the CPU
    oracle's restatement, so time to roots is not defined here)."""
    from pci_amd import _lib
    x = np.random.default_rng(1).standard_normal((sp.Nd, sp.Nlv))
    spmm = {}
    for prec in (32, 64):
        _, ms, nb = ctx.spmm(x, prec, a.spmm_reps)
        spmm[f"fp{prec}"] = {"ms": ms, "gbs": nb / (ms * 1e-3) / 1e9, "bytes": nb,
                             "kernel": ctx.spmm_variant(sp.Nlv, prec)}
        log(f"C4 SpMM fp{prec}: {ms:.2f} ms {spmm[f'fp{prec}']['gbs']:.0f} GB/s ({spmm[f'fp{prec}']['kernel']})")
    n_it = a.c4_iterations
    per_it = {}
    for flags, name in ((0, "plain_fp64"), (_lib.PCI_DVD_FP32, "fp32")):
        ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, 2, flags)  # warm-up (autotuning of the precision)
        ts = []
        for _ in range(a.steps):
            t = time.perf_counter()
            tk, _, _, st = ctx.davidson(sp.Nlv, sp.Nd0, sp.Crt4, n_it, flags)
            ts.append(time.perf_counter() - t)
        per_it[name] = {"s_per_solve": float(np.mean(ts)), "iterations": st["iterations"],
                        "ms_per_iteration": float(np.mean(ts)) / st["iterations"] * 1e3,
                        "init_ms": st["t_init_ms"], "spmm_ms": st["t_spmm_ms"], "max_residual": st["max_residual"],
                        "E_lowest": float(-tk[0])}
        log(f"C4 {name}: {np.mean(ts):.3f} s for {st['iterations']} iterations "
            f"({per_it[name]['ms_per_iteration']:.1f} ms each), residual {st['max_residual']:.2e}")
    v = per_it["fp32"]["ms_per_iteration"]
    res = {"metric": "ms per Davidson iteration at the C4 shape (FP32 mode); H build and SpMM rates",
           "value": v, "unit": "ms", "higher_is_better": False, "n_gpus": 1, "steps": a.steps,
           "dtype": "f32 SpMM / f64", "data": "synthetic (seeded integrals in pconf formats)",
           "config": {"workload": "C4: 4 valence e-, s..f n=2..9, M_J=0, even parity, SD+T from 2 refs, Breit",
                      "Nd": sp.Nd, "NumH": numh, "nnz_full": 2 * numh - sp.Nd, "Nlv": sp.Nlv, "Nd0": sp.Nd0},
           "hbuild_s": t_build, "hbuild_mnnz_s": numh / t_build / 1e6, "spmm": spmm, "davidson": per_it,
           "speedup_fp32_per_iteration": per_it["plain_fp64"]["ms_per_iteration"] / v,
           "roofline": {"bound": "hbm", "achieved": spmm["fp32"]["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": spmm["fp32"]["gbs"] / HBM_PEAK_GBS, "traffic": None,
                        "kernel": spmm["fp32"]["kernel"]}}
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--space", default="c2")
    ap.add_argument("--train-space", default="c2s")
    ap.add_argument("--device", type=int, default=int(os.environ.get("LOCAL_RANK", "0")))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-blocks (C3) leg")
    ap.add_argument("--spmm-reps", type=int, default=20)
    ap.add_argument("--widths", default="", help="--mode spmm: comma-separated block widths (default Nlv)")
    ap.add_argument("--mode", choices=["full", "spmm", "copy", "c4"], default="full",
                    help="spmm: only the resident-H SpMM launches (for counter runs); "
                         "copy: a plain device copy of known size (counter calibration); "
                         "c4: throughput at the C4 shape (use with --space c4)")
    ap.add_argument("--c4-iterations", type=int, default=40,
                    help="Davidson iterations per timed solve in --mode c4")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        sys.exit(2)

    if a.mode == "copy":
        return copy_calibration(a)
    from pci_amd import host, _lib
    d = space_dir(a.space)
    t = time.perf_counter()
    sp, ctx = host.load(d, a.device)
    log(f"space loaded: Nd={sp.Nd} Nc={sp.Nc} Nst={sp.Nst} Nd0={sp.Nd0} Nlv={sp.Nlv} "
        f"in {time.perf_counter() - t:.2f} s")

    t = time.perf_counter()
    numh, numzero, hmin = ctx.build_h()
    t_build = time.perf_counter() - t
    log(f"H built: NumH={numh} numzero={numzero} Hmin={hmin:.8f} in {t_build:.3f} s")
    if a.mode == "c4":
        return c4_throughput(a, sp, ctx, t_build, numh)
    if a.mode == "spmm":
        widths = [int(w) for w in a.widths.split(",")] if a.widths else [sp.Nlv]
        for nv in widths:
            x = np.random.default_rng(1).standard_normal((sp.Nd, nv))
            out = {}
            for prec in (32, 64):
                _, ms, nb = ctx.spmm(x, prec, a.spmm_reps)
                out[f"fp{prec}"] = {"ms": ms, "bytes": nb, "gbs": nb / (ms * 1e-3) / 1e9,
                                    "kernel": ctx.spmm_variant(nv, prec)}
            print(json.dumps({"mode": "spmm", "space": a.space, "nvec": nv, "reps": a.spmm_reps, **out}), flush=True)
        return
    ctx.build_j()
    mem = ctx.mem_usage()

    nlv, nd0, crt4, n_it = sp.Nlv, sp.Nd0, sp.Crt4, max(sp.N_it, 300)
    # FP64 oracle eigenvalues: plain FP64 GPU Davidson converged to 1e-10
    tk_ref, _, _, st_ref = ctx.davidson(nlv, nd0, 1e-10, 1000, 0)
    log(f"FP64 oracle: {st_ref['iterations']} it, residual {st_ref['max_residual']:.2e}")

    # plain FP64 method
    t = time.perf_counter()
    tk64, _, _, st64 = ctx.davidson(nlv, nd0, crt4, n_it, 0)
    t_plain = time.perf_counter() - t
    err64 = np.abs(tk64 - tk_ref).max() * HARTREE_CM
    log(f"plain FP64: {t_plain:.3f} s, {st64['iterations']} it, max err {err64:.2e} cm^-1")

    # learned correction: train on a smaller sibling space (untimed set-up)
    w, train_info = train_network(a.train_space, a.device)
    ctx.mlp_set_weights(w)

    def timed(flags):
        for _ in range(a.warmup):
            ctx.davidson(nlv, nd0, crt4, n_it, flags)
        times, errs, its, sts = [], [], [], []
        for _ in range(a.steps):
            t = time.perf_counter()
            tk, _, _, st = ctx.davidson(nlv, nd0, crt4, n_it, flags)
            times.append(time.perf_counter() - t)
            errs.append((np.abs(tk - tk_ref) * HARTREE_CM).tolist())
            its.append(st["iterations"])
            sts.append(st)
        timed.stats = sts
        return float(np.mean(times)), errs, its, sts[-1]

    t_fp32, errs32, its32, st32 = timed(_lib.PCI_DVD_FP32)
    log(f"fp32: {t_fp32:.3f} s, iterations {its32}, max err {max(map(max, errs32)):.2e} cm^-1, "
        f"converged {st32['converged']}")
    t_cpl, errsc, itsc, stc = timed(_lib.PCI_DVD_FP32 | _lib.PCI_DVD_COUPLED)
    log(f"coupled+fp32: {t_cpl:.3f} s, iterations {itsc}, max err {max(map(max, errsc)):.2e} cm^-1, "
        f"converged {stc['converged']}")
    t_fast, errs, its, st = timed(_lib.PCI_DVD_FP32 | _lib.PCI_DVD_LEARNED)
    log(f"learned+fp32: {t_fast:.3f} s, iterations {its}, max err {max(map(max, errs)):.2e} cm^-1, "
        f"converged {st['converged']}, learned used {st['learned_used']} rejected {st['learned_rejected']}")
    ok = all(max(e) < 0.01 for e in errs) and st["converged"] in (1, 2)
    # the FP32 SpMM launches inside the timed learned+FP32 solves (HIP events
    # on the context stream around each launch)
    n32_in = sum(q["spmm_fp32"] for q in timed.stats)
    ms32_in = sum(q["t_spmm32_ms"] for q in timed.stats) / max(n32_in, 1)

    # SpMM roofline (dominant kernel): HIP events on the context stream
    x = np.random.default_rng(1).standard_normal((sp.Nd, nlv))
    _, ms32, b32 = ctx.spmm(x, 32, a.spmm_reps)
    _, ms64, b64 = ctx.spmm(x, 64, a.spmm_reps)
    var32, var64 = ctx.spmm_variant(nlv, 32), ctx.spmm_variant(nlv, 64)
    gbs32 = b32 / (ms32 * 1e-3) / 1e9
    gbs32_in = b32 / (ms32_in * 1e-3) / 1e9
    gbs64 = b64 / (ms64 * 1e-3) / 1e9
    log(f"SpMM fp32: {ms32_in:.3f} ms {gbs32_in:.0f} GB/s in the timed solves ({n32_in} launches), "
        f"{ms32:.3f} ms {gbs32:.0f} GB/s back to back ({var32}); fp64: {ms64:.3f} ms {gbs64:.0f} GB/s ({var64})")

    res = {
        "metric": "s to Nlv roots at |dE|<0.01 cm^-1: learned+FP32 vs plain FP64; SpMM GB/s",
        "value": t_fast,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t_fast * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 SpMM / f64",
        "data": "synthetic (seeded integrals in pconf formats)",
        "config": {"workload": "C2: 3 valence e-, s..f n=2..14, M_J=1/2, odd parity, SD from 3 refs",
                   "Nd": sp.Nd, "Nc": sp.Nc, "NumH": numh, "nnz_full": int(2 * numh - sp.Nd),
                   "Nlv": nlv, "Crt4": crt4, "Nd0": nd0, "variant": "learned correction + FP32 H, FP64 refinement"},
        "roots_within_0.01cm": bool(ok),
        "plain_fp64_s": t_plain,
        "fp32_s": t_fp32,
        "coupled_fp32_s": t_cpl,
        "speedup_vs_plain_fp64": t_plain / t_fast,
        "speedup_fp32_only": t_plain / t_fp32,
        "iterations": {"plain": st64["iterations"], "fp32": its32, "coupled_fp32": itsc, "learned_fp32": its},
        "err_cm_per_root_learned_fp32": errs[-1],
        "err_cm_per_root_fp32": errs32[-1],
        "max_err_cm_plain": float(err64),
        "learned": {"used": st["learned_used"], "rejected": st["learned_rejected"], **train_info},
        "breakdown_ms": {v: {"total": x["t_total_ms"], "init": x["t_init_ms"], "spmm": x["t_spmm_ms"],
                             "spmm_fp32": x["spmm_fp32"], "spmm_fp64": x["spmm_fp64"]}
                         for v, x in (("plain", st64), ("fp32", st32), ("coupled_fp32", stc), ("learned_fp32", st))},
        "hbuild_s": t_build,
        "hbuild_mnnz_s": numh / t_build / 1e6,
        "spmm_fp32_gbs": gbs32,
        "spmm_fp64_gbs": gbs64,
        "device_mem_gb": mem / 1e9,
        "roofline": {"bound": "hbm", "achieved": gbs32_in, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs32_in / HBM_PEAK_GBS,
                     "traffic": C2_SPMM32_TRAFFIC.get(var32) if a.space == "c2" and nlv == 10 else None,
                     "kernel": var32, "kernel_fp64": var64, "bytes_per_launch": b32,
                     "avg_launch_ms": ms32_in, "launches": n32_in,
                     "measured_in": "the FP32 SpMM launches of the timed learned+FP32 solves",
                     "back_to_back_ms": ms32, "back_to_back_gbs": gbs32},
    }
    if not a.no_batch:
        res["batch_c3"] = batch_blocks(a.device, a.steps)
    if not a.no_cpu:
        cb = cpu_baseline(ctx, sp, d, st64["iterations"], nd0)
        mn, el, tt = cpu_formh_sample(d, sp.Nd)
        cb["formh_mnnz_s"] = mn
        cb["formh_sample"] = f"oracle FormH on 240 rows in 12 evenly spaced blocks: {el} elements in {tt:.2f} s"
        res["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value']:.1f} s to roots (extrapolated), FormH {mn:.2f} Mnnz/s")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
