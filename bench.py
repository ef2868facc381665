"""Benchmark: H*psi band-applies/s of one full Gamma-point evaluation on one
MI355X, mixed precision beside FP64 in the same run, with the rooflines of
the dominant kernels and the CPU oracle timed on the host's cores beside it.

One step = ChargeDensity::update_density -> EnergyFunctional::update_vhxc ->
EnergyFunctional::energy(compute_hpsi=true, compute_forces=true), the
sequence of src/BOSampleStepper.cpp:1207-1217, on synthetic inputs already
resident in HBM.  value = bands / (seconds per step) of the mixed path
(FP32 band FFTs and projector GEMMs, FP64 reductions); the FP64 path runs
the same steps on the same inputs, and the mixed path's error against it is
reported per energy term (Ha/atom), for forces (Ha/bohr) and for H*psi.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]
                       [--prec mixed|fp64] [--cpu-seconds S | --no-cpu]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "qbox-public_amd"))

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and, in a run
# of its own, WRITE_SIZE (KiB per dispatch), scaled by what the counters read
# for a plain 1 GiB copy (--copy-probe 1024: FETCH_SIZE 0.5 GiB, WRITE_SIZE
# 1 GiB, so bytes = 1024 (2 FETCH_SIZE + WRITE_SIZE)); commands and the
# per-kernel figures are in profiles/c3_hbm_traffic.md.  Keyed by workload and
# precision; "gemm" is the mean of the three projector products of a step.
MEASURED_TRAFFIC = {
    ("c3: 256 atoms L=36 ecut=70Ry 512 bands", "bf16x2"): {
        "zbwd": 1527706296, "xbwd": 3053536864, "yvmul": 5114448256, "ydens": 2539485224,
        "xfwd": 3134759020, "zfwd": 2388999046,
        "gemm": (5263151974 + 3156065408 + 3500000128) // 3,
    },
}


def measured_traffic(ef, mode, kind):
    return MEASURED_TRAFFIC.get((ef.sysdef.name, mode), {}).get(kind)
# mixed modes: FP32 band FFTs with the projector GEMMs on FP32 MFMA, or as
# 3-part / 2-part BF16-split products; the bar of the headline metric
MIXED_MODES = ("fp32", "bf16x3", "bf16x2")
DE_TOL = 1e-6      # Ha/atom, every energy term
F_TOL = 1e-5       # Ha/bohr
FFT_KINDS = ("zbwd", "xbwd", "yvmul", "ydens", "xfwd", "zfwd")
# band pairs per pass launch (the library's batch)
PAIRS_PER_BATCH = 64
TERMS = ["ekin", "eps", "enl", "ehart", "exc", "esr", "etotal"]


def pair_bytes(ef, kind, sc):
    """Algorithmic HBM bytes of one band pair through a pass kernel.

    z-bwd reads two bands (2 ngw complex FP64) and writes the columns
    (nvec np2 elements of sc bytes); x-bwd reads the columns and writes the
    active rows A = 2 ntrans0 np0 np2; the y pass reads and writes A (v(r),
    shared by the pairs, is counted once per launch in fft_roofline); x-fwd
    and z-fwd mirror the backward passes, z-fwd read-modify-writes the two
    H*psi bands.
    """
    ngw, nvec, np2 = ef.ngw, ef.nvec, ef.np2
    cols = nvec * np2 * sc
    A = ef.nact * ef.np0 * ef.np2 * sc
    return {
        "zbwd": 2 * ngw * 16 + cols,
        "xbwd": cols + A,
        "yvmul": 2 * A,
        "ydens": A,
        "xfwd": A + cols,
        "zfwd": cols + 2 * 2 * ngw * 16,
    }[kind]


def fft_roofline(ef, kind, kt, steps, nst, sc, mode):
    tot_ms, nlaunch = kt[kind]
    launches_per_step = nlaunch / steps
    npairs = (nst + 1) // 2
    pairs_per_launch = npairs / -(-npairs // PAIRS_PER_BATCH)
    per_launch = pair_bytes(ef, kind, sc) * pairs_per_launch
    if kind == "ydens":
        per_launch += 2 * ef.N * 8                 # rho read-modify-write
    if kind == "yvmul":
        per_launch += ef.N * 8                     # v(r), read once per launch
    avg_ms = tot_ms / nlaunch
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return {"kernel": kind, "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(ef, mode, kind),
            "algorithmic_bytes_per_launch": int(per_launch),
            "avg_launch_ms": round(avg_ms, 4), "launches_per_step": launches_per_step}


def gemm_roofline(ef, kt, nst, mode):
    """Projector GEMM against the MFMA peak of the type it runs on.  achieved
    counts the algorithmic 2 M N K flops; in the BF16-split modes each
    product is `products` bf16 MFMA products, so the MFMA pipe does that
    many times the algorithmic work (mfma_busy_frac)."""
    tot_ms, nlaunch = kt["gemm"]
    flops = 2.0 * ef.nproj * 2 * ef.ngw * nst
    avg_ms = tot_ms / nlaunch
    achieved = flops / (avg_ms * 1e-3) / 1e12
    kind, products = {"fp64": ("fp64", 1), "fp32": ("fp32", 1), "bf16x3": ("bf16", 6),
                      "bf16x2": ("bf16", 3)}[mode]
    peak = MFMA_PEAK_TFLOPS[kind]
    return {"kernel": "gemm", "bound": "mfma", "achieved": round(achieved, 2), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "traffic": measured_traffic(ef, mode, "gemm"),
            "mfma_type": kind, "products_per_flop": products,
            "mfma_busy_frac": round(achieved * products / peak, 4),
            "algorithmic_flops_per_launch": flops, "avg_launch_ms": round(avg_ms, 4)}


def fp64_vs_oracle(config, r64, h64):
    """Relative error of this run's FP64 results against the CPU oracle's
    results for the same inputs, committed as tests/golden/oracle_<cfg>.npz
    (energies, fion, sampled H*psi); None when no such file exists."""
    name = {2: "c3", 4: "c5"}.get(config)
    path = ROOT / "tests" / "golden" / f"oracle_{name}.npz"
    if name is None or not path.exists():
        return None
    ref = np.load(path)
    e = dict(zip([str(x) for x in ref["names"]], ref["energies"]))
    terr = max(abs(r64[t] - v) / max(abs(v), 1e-12) for t, v in e.items())
    ferr = np.abs(r64["fion"] - ref["fion"]).max() / np.abs(ref["fion"]).max()
    hs = h64[ref["hpsi_band"], ref["hpsi_g"]]
    herr = np.abs(hs - ref["hpsi_val"]).max() / ref["hpsi_absmax"]
    return {"max_term_rel_err": float("%.3e" % terr), "fion_rel_err": float("%.3e" % ferr),
            "hpsi_sample_rel_err": float("%.3e" % herr),
            "source": f"tests/golden/{path.name} (oracle, same seeded inputs)"}


def cpu_baseline(sysdef, c, seconds):
    """Time the FP64 CPU oracle (OpenMP) on a bounded sample of the same
    workload: the full evaluation of the same system restricted to its first
    nb bands, nb doubled from 4 until one evaluation takes about a sixth of
    `seconds`; then the median of five timed evaluations after that one.
    The oracle is a plain restatement (its own loops for the projector
    products and its own FFT, no BLAS or FFTW), so the rate understates a
    BLAS/FFTW build of the reference on the same cores."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    nb = 4
    while True:
        sub = dataclasses.replace(sysdef, nst=nb)
        o = oracle.Oracle(sub)
        cs = np.ascontiguousarray(c[:nb])
        t0 = time.perf_counter()
        o.evaluate(cs, hpsi=True, forces=True)
        dt = time.perf_counter() - t0
        if dt > seconds / 6.0 or 2 * nb > sysdef.nst:
            break
        nb *= 2
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        o.evaluate(cs, hpsi=True, forces=True)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    return {"value": round(nb / dt, 3), "unit": "band-applies/s", "cores": cores, "kind": "port",
            "sample": f"one full FP64 evaluation (density, vhxc, energy, H*psi, forces) of the "
                      f"same system restricted to its first {nb} bands on the CPU oracle, "
                      f"median of 5 after a warm-up: {dt:.2f} s (runs "
                      + ", ".join(f"{t:.2f}" for t in times) + " s); the oracle's own loops "
                      "and FFT, no BLAS/FFTW, so below a BLAS/FFTW build of the reference"}


def run(ef, dc, hp, prec, steps, warmup, ctx):
    import qbg
    for _ in range(warmup):
        ef.evaluate(dc, hp, forces=True, prec=prec)
    ctx.sync()
    prof = qbg.Profiler(ctx)
    prof.reset()
    prof.enable(True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    res = None
    for _ in range(steps):
        res = ef.evaluate(dc, hp, forces=True, prec=prec)
    gpu_ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - t0
    prof.enable(False)
    return res, wall, gpu_ms, prof.read()


def bench_c4(args):
    """Twelve independent systems of uneven size (the uneven-batching
    workload), evaluated three ways: a loop over the systems (each waits for
    its results), one HIP graph holding the twelve evaluations one after the
    other, and one graph holding them as twelve parallel branches (one
    context and stream per system, joined by qbg_stream_join) so the small
    systems' launches run beside the large ones'."""
    import qbg
    from qbg import configs as cf
    prec = "bf16x2" if args.prec == "mixed" else "fp64"
    systems = cf.config4_systems()
    main = qbg.Context(0)
    ctxs, efs, dcs, hps = [], [], [], []
    for s in systems:
        ctx = qbg.Context(0)
        ef = qbg.EnergyFunctional(ctx, s)
        c = cf.gram_numpy(cf.random_coefficients(cf.BASE_SEED + 60, ef.basis(0)["g2"], s.nst))
        ctxs.append(ctx)
        efs.append(ef)
        dcs.append(ctx.to_device(c))
        hps.append(ctx.array(c.shape, np.complex128))
        print(f"[bench c4] {s.name}: grid {ef.np0}x{ef.np1}x{ef.np2} ngw {ef.ngw}",
              file=sys.stderr, flush=True)
    nb = sum(s.nst for s in systems)

    def loop():
        return [ef.evaluate(dc, hp, forces=True, prec=prec)["etotal"]
                for ef, dc, hp in zip(efs, dcs, hps)]

    for _ in range(max(1, args.warmup)):
        ref = loop()
    for ctx in ctxs:
        ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loop()
    t_loop = (time.perf_counter() - t0) / args.steps

    def graph(parallel):
        with qbg.Graph(main) as g:
            for ctx, ef, dc, hp in zip(ctxs, efs, dcs, hps):
                qbg.stream_join(main, ctx)
                ef.evaluate_async(dc, hp, forces=True, prec=prec)
                if not parallel:
                    qbg.stream_join(ctx, main)
            if parallel:
                for ctx in ctxs:
                    qbg.stream_join(ctx, main)
        return g

    out = {}
    for name, parallel in (("serial", False), ("parallel", True)):
        g = graph(parallel)
        g.launch()
        main.sync()
        got = [ef.results()["etotal"] for ef in efs]
        assert got == ref, f"graph ({name}) replay differs from the loop"
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.launch()
        main.sync()
        out[name] = (time.perf_counter() - t0) / args.steps
        g.release()
    t_graph = out["parallel"]
    print(json.dumps({
        "metric": "uneven batch: H*psi band-applies/s of twelve systems per evaluation round",
        "value": round(nb / t_graph, 2), "unit": "band-applies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * t_graph, 3),
        "higher_is_better": True, "dtype": "f32+f64" if prec != "fp64" else "f64",
        "data": "synthetic",
        "config": {"workload": "twelve Gamma-point systems, atoms in {8..128}, ecut 40-80 Ry",
                   "systems": [s.name for s in systems], "bands_total": nb,
                   "precision": prec},
        "loop_value": round(nb / t_loop, 2), "loop_ms_per_step": round(1e3 * t_loop, 3),
        "serial_graph_ms_per_step": round(1e3 * out["serial"], 3),
        "graph_over_loop": round(t_loop / t_graph, 3),
        "parallel_over_serial_graph": round(out["serial"] / t_graph, 3)}))


def bench_scf(args):
    """One SCF iteration of BOSampleStepper on the device (update_density,
    Anderson/Kerker charge mixing, update_vhxc, energy with H*psi, PSDA
    update with ortho_align), replayed from two captured graphs per
    iteration; the energy step in the chosen precision, the stepper in FP64."""
    import qbg
    from qbg import configs as cf
    from qbg.scf import ScfLoop
    sysdef = cf.config(args.config)
    ctx = qbg.Context(0)
    ef = qbg.EnergyFunctional(ctx, sysdef)
    b = ef.basis(0)
    c = cf.random_coefficients(cf.BASE_SEED + 50 + args.config, b["g2"], sysdef.nst)
    dc = ctx.to_device(c)
    qbg.gram(ctx, dc, ef.ngw, ef.ngw, sysdef.nst)
    hp = ctx.array(c.shape, np.complex128)
    prec = "bf16x2" if args.prec == "mixed" else "fp64"
    print(f"[bench scf] {sysdef.name} prec {prec}", file=sys.stderr, flush=True)
    # kernel breakdown from eager iterations with the profiler on
    loop = ScfLoop(ef, dc, hp, wf_dyn="psda", prec=prec, graphs=False)
    prof = qbg.Profiler(ctx)
    loop.run(1, 1)
    prof.reset()
    prof.enable(True)
    loop.run(2, 1)
    prof.enable(False)
    kt = prof.read()
    loop.release()
    # timed: graph replays after the first (eager) iteration
    dc.upload(c)
    qbg.gram(ctx, dc, ef.ngw, ef.ngw, sysdef.nst)
    loop = ScfLoop(ef, dc, hp, wf_dyn="psda", prec=prec, graphs=True)
    nit = 1 + args.warmup + args.steps
    hist = loop.run(nit, 1)
    secs = [h["seconds"] for h in hist[1 + args.warmup:]]
    t = float(np.mean(secs))
    loop.release()
    print(json.dumps({
        "metric": "SCF iterations/s on 1 GPU (density, charge mixing, vhxc, H*psi, PSDA "
                  "update with ortho_align), one graph pair per iteration",
        "value": round(1.0 / t, 3), "unit": "iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * t, 3),
        "higher_is_better": True, "dtype": "f32+f64" if prec != "fp64" else "f64",
        "data": "synthetic",
        "config": {"workload": sysdef.name, "bands": sysdef.nst, "precision": prec,
                   "stepper": "PSDA (FP64)", "mixer": "Anderson ndim 3, Kerker rc 10"},
        "band_updates_per_s": round(sysdef.nst / t, 2),
        "etotal_per_iteration": [h["etotal"] for h in hist],
        "eager_kernel_ms_per_iteration": {k: round(v[0] / 2, 4) for k, v in kt.items()}}))


def bench_kpoint(args):
    """One evaluation (density, update_vhxc, energy with H*psi and forces) of
    the workload at one k-point off Gamma (qbg_ef_create_k, the complex
    full-sphere basis), FP64 and BF16X2 on the same orthonormal random
    complex bands; the BF16X2 result is checked against the FP64 one."""
    import dataclasses
    import qbg
    from qbg import configs as cf
    if len(args.kpoint) % 3:
        print(json.dumps({"error": "--kpoint takes three numbers per k-point"}))
        sys.exit(2)
    kps = np.array(args.kpoint, dtype=float).reshape(-1, 3)
    nkp = kps.shape[0]
    ctx = qbg.Context(0)
    if nkp == 1:
        sysdef = dataclasses.replace(cf.config(args.config), kpoint=kps[0])
        ef = qbg.EnergyFunctional(ctx, sysdef)
    else:                           # a set with equal weights
        sysdef = dataclasses.replace(cf.config(args.config), kpoint=None)
        ef = qbg.EnergyFunctional(ctx, sysdef, kpoints=kps, weights=np.full(nkp, 1.0 / nkp))
    c = ef.kset_pack([cf.orthonormal_k(cf.random_coefficients_k(
        cf.BASE_SEED + 70 + args.config + ik, ef.kset_basis(ik)["kpg2"], sysdef.nst))
        for ik in range(nkp)])
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    del c
    nst, nat = sysdef.nst * nkp, sysdef.nat     # band-applies count every k-point's bands
    print(f"[bench kpoint] {sysdef.name} k {kps.tolist()}: ngw {ef.kngw} nproj {ef.nproj}",
          file=sys.stderr, flush=True)
    r64, wall64, _, kt64 = run(ef, dc, hp, "fp64", args.steps, args.warmup, ctx)
    h64 = hp.download()
    rbf, wallbf, _, ktbf = run(ef, dc, hp, "bf16x2", args.steps, args.warmup, ctx)
    hbf = hp.download()
    err = max(abs(rbf[t] - r64[t]) / nat for t in TERMS)
    ferr = float(np.abs(rbf["fion"] - r64["fion"]).max())
    herr = float(np.abs(hbf - h64).max() / np.abs(h64).max())
    t64, tbf = wall64 / args.steps, wallbf / args.steps
    print(json.dumps({
        "metric": "H*psi band-applies/s at k-points off Gamma on 1 GPU (complex basis; "
                  "density + update_vhxc + energy with H*psi and forces)",
        "value": round(nst / tbf, 2), "unit": "band-applies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * tbf, 3),
        "higher_is_better": True, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": sysdef.name, "kpoints": kps.tolist(), "ngw": ef.kngw,
                   "nproj": ef.nproj, "precision": "bf16x2", "bands_per_kpoint": sysdef.nst},
        "fp64_value": round(nst / t64, 2), "fp64_ms_per_step": round(1e3 * t64, 3),
        "bf16x2_vs_fp64": {"max_dE_per_atom": float("%.3e" % err),
                           "force_max_abs_err": float("%.3e" % ferr),
                           "hpsi_rel_err": float("%.3e" % herr),
                           "meets_accuracy": bool(err <= DE_TOL and ferr <= F_TOL)},
        "kernel_ms_per_step_fp64": {k: round(v[0] / args.steps, 3) for k, v in kt64.items()},
        "kernel_ms_per_step_bf16x2": {k: round(v[0] / args.steps, 3) for k, v in ktbf.items()},
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--prec", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c4", action="store_true",
                    help="only the uneven batch: twelve systems, per-system loop vs one graph")
    ap.add_argument("--scf", action="store_true",
                    help="only time whole SCF iterations (PSDA stepper, Anderson mixing)")
    ap.add_argument("--kpoint", type=float, nargs="+", default=None, metavar="K",
                    help="only time the evaluation at these k-points (three numbers each, "
                         "units of b; several: an equal-weight set) on the complex basis, "
                         "FP64 and BF16X2")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB three times (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import qbg
        ctx = qbg.Context(0)
        nb = args.copy_probe << 20
        src, dst = ctx.array(nb // 8), ctx.array(nb // 8)
        src.zero()
        for _ in range(3):
            ctx.check(ctx.lib.qbg_stream_copy(ctx.h, dst.p, src.p, nb))
        ctx.sync()
        print(json.dumps({"copy_probe_bytes_read": nb, "copy_probe_bytes_written": nb,
                          "launches": 3}))
        return
    if args.c4:
        return bench_c4(args)
    if args.scf:
        return bench_scf(args)
    if args.kpoint is not None:
        return bench_kpoint(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "note only (band sharding + RCCL reduction of rho), so "
                                   "--gpus must be 1"}))
        sys.exit(2)

    import qbg
    from qbg import configs as cf

    sysdef = cf.config(args.config)
    ctx = qbg.Context(0)
    ef = qbg.EnergyFunctional(ctx, sysdef)
    b = ef.basis(0)
    # the inputs of the committed oracle results (tests/golden/oracle_<cfg>.npz)
    c = cf.gram_numpy(cf.random_coefficients(cf.BASE_SEED + 50 + args.config, b["g2"],
                                             sysdef.nst))
    dc = ctx.to_device(c)
    hp = ctx.array(c.shape, np.complex128)
    del c
    ky = b["idx"][:, 1]
    ef.nact = 2 * (int(np.abs(ky).max()) + 1)            # active rows (byte model)
    nst, nat = sysdef.nst, sysdef.nat
    print(f"[bench] {sysdef.name}: grid {ef.np0}x{ef.np1}x{ef.np2} ngw {ef.ngw} "
          f"nproj {ef.nproj}", file=sys.stderr, flush=True)

    # FP64 path, then each mixed mode on the same inputs
    r64, wall64, gpu64, kt64 = run(ef, dc, hp, "fp64", args.steps, args.warmup, ctx)
    h64 = hp.download()
    fp64_check = fp64_vs_oracle(args.config, r64, h64)
    print(f"[bench] fp64 {1e3 * wall64 / args.steps:.2f} ms/step", file=sys.stderr, flush=True)
    modes = {}
    if args.prec == "mixed":
        for mode in MIXED_MODES:
            res_m, wall_m, gpu_m, kt_m = run(ef, dc, hp, mode, args.steps, args.warmup, ctx)
            hmx = hp.download()
            err = {t: abs(res_m[t] - r64[t]) / nat for t in TERMS}
            ferr = float(np.abs(res_m["fion"] - r64["fion"]).max())
            ok = max(err.values()) <= DE_TOL and ferr <= F_TOL
            modes[mode] = {
                "value": round(nst / (wall_m / args.steps), 2),
                "ms_per_step": round(1e3 * wall_m / args.steps, 3),
                "meets_accuracy": bool(ok),
                "dE_per_atom": {t: float("%.3e" % v) for t, v in err.items()},
                "max_dE_per_atom": float("%.3e" % max(err.values())),
                "force_max_abs_err": float("%.3e" % ferr),
                "hpsi_rel_err": float("%.3e" % (np.abs(hmx - h64).max() / np.abs(h64).max())),
                "_run": (res_m, wall_m, gpu_m, kt_m),
            }
            del hmx
            print(f"[bench] {mode} {1e3 * wall_m / args.steps:.2f} ms/step "
                  f"max dE/atom {max(err.values()):.2e} force err {ferr:.2e}",
                  file=sys.stderr, flush=True)
        passing = [m for m in MIXED_MODES if modes[m]["meets_accuracy"]]
        if not passing:
            print(json.dumps({"error": "no mixed mode meets the accuracy bar",
                              "modes": {m: {k: v for k, v in d.items() if k != "_run"}
                                        for m, d in modes.items()}}))
            sys.exit(1)
        best = max(passing, key=lambda m: modes[m]["value"])
        res, wall, gpu_ms, kt = modes[best]["_run"]
        out_prec = "mixed:" + best
        errors = {k: v for k, v in modes[best].items() if k not in ("_run", "value", "ms_per_step")}
        errors["reference"] = ("the FP64 GPU path of the same run; fp64_vs_oracle holds that "
                               "path's own error against the CPU oracle's committed results "
                               "for these inputs")
    else:
        res, wall, gpu_ms, kt = r64, wall64, gpu64, kt64
        errors = None
        best = "fp64"
        out_prec = "fp64"
    del h64

    # H*psi-only rate: energy with forces off, potential fixed
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ef.energy(dc, hp, forces=False, stress=False, prec=best)
    ctx.sync()
    hpsi_only = nst / ((time.perf_counter() - t0) / args.steps)

    ms_per_step = 1e3 * wall / args.steps
    value = nst / (wall / args.steps)
    value64 = nst / (wall64 / args.steps)
    sc = 8 if args.prec == "mixed" else 16
    fft_kind = max((k for k in FFT_KINDS if k in kt), key=lambda k: kt[k][0])
    froof = fft_roofline(ef, fft_kind, kt, args.steps, nst, sc, best)
    groof = gemm_roofline(ef, kt, nst, best)
    roof = froof if kt[fft_kind][0] >= kt["gemm"][0] else groof

    out = {
        "metric": "H*psi band-applies/s, mixed vs FP64 on 1 GPU at |dE|<=1e-6 Ha/atom "
                  "(density + update_vhxc + energy with H*psi and forces)",
        "value": round(value, 2),
        "unit": "band-applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.prec == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": sysdef.name, "atoms": nat, "bands": nst,
                   "ecut_ry": sysdef.ecut_ry, "grid": [ef.np0, ef.np1, ef.np2], "ngw": ef.ngw,
                   "nproj": ef.nproj, "precision": out_prec},
        "fp64_value": round(value64, 2),
        "fp64_ms_per_step": round(1e3 * wall64 / args.steps, 3),
        "mixed_over_fp64": round(value / value64, 3),
        "hpsi_only_value": round(hpsi_only, 2),
        "errors_vs_fp64": errors,
        "fp64_vs_oracle": fp64_check,
        "mixed_modes": {m: {k: v for k, v in d.items() if k != "_run"}
                        for m, d in modes.items()},
        "gpu_event_ms_per_step": round(gpu_ms / args.steps, 3),
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items()},
        "fp64_kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt64.items()},
        "etotal": res["etotal"],
        "roofline": roof,
        "fft_roofline": froof,
        "gemm_roofline": groof,
        "fp64_gemm_roofline": gemm_roofline(ef, kt64, nst, "fp64"),
        "fft_launch_tuning": {p: ef.band_grid().tuning(p) for p in ("fp64", best)},
        "gemm_launch_tuning": ef.gemm_plan(),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sysdef, dc.download(), args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
