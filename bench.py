#!/usr/bin/env python3
"""Benchmark of the batched RGF selected-inversion hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--energies E] [--workload energies|dominant|large]

Default workload ("energies"): one Hermitian block-tridiagonal H per block size
b in {128, 256, 384}, 64 blocks each, re/im ~ N(0, 1/sqrt(b)), couplings x0.5;
A(E_k) = (E_k + 1e-3 i) I - H for 256 energies uniform in [-2, 2]: 3 x 256 =
768 selected inversions (RGF forward + backward sweep, all diagonal and
off-diagonal blocks of G kept in HBM) per step.  A step runs the three
batched launches (one per block size) back to back on one stream.

Measured:
  * batched FP64 RGF (complex128 GEMMs on FP64 MFMA, FP64 inverses)
  * batched mixed RGF (complex64 GEMMs on FP32 MFMA, FP64 inverses) and its
    error against the FP64 result over every instance
  * batched int8x RGF (complex128 GEMMs emulated from 6 int8 digit planes per
    operand on the int8 matrix cores, FP64 inverses) and its error likewise
  * plain FP64 RGF, one energy per launch sequence (batch = 1), on a sample
  * the CPU baseline: the oracle's numpy restatement of rgf.py on the host
`value` is the solves/s of the fastest batched mode whose max relative error
(Frobenius over all selected blocks, per instance) is <= 1e-8.  The roofline
figure is the dominant kernel's: algorithmic flops over summed kernel time,
both from the library's launch profiler (hipEvents on the launch stream)
during one extra, instrumented step.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = ("solves/s + TFLOP/s: batched mixed-prec RGF fwd+bwd vs plain FP64 RGF, "
          "err≤1e-8")
PEAK_TFLOPS = {"gemm_f64": 78.6, "gemm_f32": 157.3, "gemm_i8x": 5000.0}  # int8: TOP/s
I8X_OPS_PER_FLOP = 15.75  # int8 MFMA ops per algorithmic flop (3 x 21 digit pairs x 2MNK / 8MNK)
HBM_TBPS = 8.0  # MI355X HBM3E
ERR_GATE = 1e-8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--energies", type=int, default=256)
    ap.add_argument("--workload", default="energies",
                    choices=["energies", "dominant", "large", "adjoint", "convert", "variants",
                             "gemm"])
    ap.add_argument("--chunk", type=int, default=128, help="energies per adjoint launch")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches instead of replays of one step captured into a HIP graph")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward+adjoint record of the default run")
    ap.add_argument("--no-plain", action="store_true", help="skip the batch=1 FP64 sample")
    ap.add_argument("--no-bcr", action="store_true",
                    help="skip BCR-serial on the single-instance workloads")
    ap.add_argument("--no-pcie", action="store_true",
                    help="skip the PCIe-inclusive figure (pinned copy rates, drop-in call)")
    ap.add_argument("--inverse", default="auto", choices=["auto", "pivoted"],
                    help="block inverse of the RGF plans: 'auto' takes the Newton-Schulz "
                         "iteration for latency-bound calls (at most 4 instances) on column "
                         "diagonally dominant matrices, the pivoted elimination otherwise")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the GEMM tile autotuner (untimed, before the timed steps)")
    ap.add_argument("--save-tune", default=None, help="write the tuned GEMM table (JSON) here")
    ap.add_argument("--load-tune", default=None,
                    help="install this tuned GEMM table instead of tuning")
    ap.add_argument("--traffic-pass", action="store_true",
                    help="run one FP64 step and a 2 GiB calibration copy (neg kernel), untimed, "
                         "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs; prints no bench line")
    return ap.parse_args()


def measured_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed PMC summary of this
    workload (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated on a copy kernel),
    or None when no summary is committed."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and f[0] == kernel:
            return float(f[4]) * 1e6
    return None


def _kind_summary(res, kind):
    """Launch-profiler totals of one kernel kind in a profiled step."""
    kms, kcnt, kfl = res["prof"].get(kind, (0.0, 0, 0.0))
    if not kcnt:
        return None
    tf = kfl / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    d = dict(ms=round(kms, 3), launches=int(kcnt), avg_launch_us=round(kms / kcnt * 1e3, 3))
    if kind == "i8x_split":
        gbps = res["bytes"].get(kind, 0.0) / (kms / 1e3) / 1e9 if kms > 0 else 0.0
        d["algorithmic_GBps"] = round(gbps, 1)
        d["hbm_frac"] = round(gbps / (HBM_TBPS * 1e3), 4)
        return d
    d["algorithmic_tflops"] = round(tf, 3)
    if kind == "gemm_i8x":
        d["int8_tops_executed"] = round(tf * I8X_OPS_PER_FLOP, 1)
        d["int8_peak_frac"] = round(tf * I8X_OPS_PER_FLOP / PEAK_TFLOPS["gemm_i8x"], 4)
        d["timed"] = "the product kernel alone (the digit split is the i8x_split kind)"
    return d


def rel_err_per_instance(G, R):
    """max over instances of ||sel(G)-sel(R)||_F / ||sel(R)||_F (torch tensors, dim 0 = batch)."""
    import torch
    num = torch.zeros(G[0].shape[0], dtype=torch.float64, device=G[0].device)
    den = torch.zeros_like(num)
    for g, r in zip(G, R):
        if g.numel() == 0:
            continue
        num += (g - r).abs().pow(2).flatten(1).sum(1)
        den += r.abs().pow(2).flatten(1).sum(1)
    return float((num / den).sqrt().max())


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(energies would shard across GPUs with no communication, one process per GPU, "
              "so the per-GPU figure is the one measured here)", file=sys.stderr)
        return 2
    import torch
    from sinv_amd import core, _native
    from sinv_amd.utils import configs

    dev = core.device()
    torch.cuda.synchronize()
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "convert":
        return bench_convert(args)
    if args.workload == "variants":
        return bench_variants(args)
    if args.workload == "gemm":
        return bench_gemm(args)

    # ------------------------------------------------------------ workload
    eta = 1e-3
    if args.workload == "energies":
        groups = []
        for gi, b in enumerate((128, 256, 384)):
            H = configs.hermitian_H(64, b, seed=10 + gi)
            E = configs.energies(args.energies, seed=20 + gi)
            groups.append(dict(b=b, N=64, H=H, shift=E + 1j * eta, dscale=-1.0, oscale=-1.0))
        wl = dict(workload="rgf_energies", blocks=64, block_sizes=[128, 256, 384],
                  energies_per_block_size=args.energies, eta=eta, solves_per_step=3 * args.energies,
                  matrix="A(E)=(E+i*eta)I-H, H Hermitian block-tridiagonal, c128")
    elif args.workload == "dominant":
        d = configs.dominant_blocks(16, 64, seed=1)
        groups = [dict(b=64, N=16, H=d, shift=None, dscale=1.0, oscale=1.0, batch=1)]
        wl = dict(workload="rgf_single_dominant", blocks=16, block_size=64, solves_per_step=1)
    else:
        d = configs.dominant_blocks(128, 512, seed=2)
        groups = [dict(b=512, N=128, H=d, shift=None, dscale=1.0, oscale=1.0, batch=1)]
        wl = dict(workload="rgf_single_dominant_large", blocks=128, block_size=512,
                  solves_per_step=1)
    solves = sum(len(g["shift"]) if g["shift"] is not None else 1 for g in groups)
    alg_flops = sum((len(g["shift"]) if g["shift"] is not None else 1)
                    * configs.rgf_flops(g["N"], g["b"]) for g in groups)

    for g in groups:
        g["dev"] = tuple(core.to_dev(x) for x in g["H"])
        g["out"] = None

    def make_plans(precision, batch_override=None, energies_slice=None):
        plans = []
        for g in groups:
            sh = g["shift"] if energies_slice is None or g["shift"] is None else g["shift"][energies_slice]
            p = core.RgfPlan(*g["dev"], shift=sh, dscale=g["dscale"], oscale=g["oscale"],
                             precision=precision, out=g["out"] if energies_slice is None else None,
                             inverse=args.inverse)
            if energies_slice is None:
                g["out"] = p.outputs()
                g["inverse"] = p.inverse[0]
            plans.append(p)
        return plans

    stream = torch.cuda.current_stream()

    if args.traffic_pass:
        # counters see every dispatch of the process: tune in an earlier run
        # (--save-tune) and load that table here, so only the step is counted
        plans = make_plans("fp64")
        if args.load_tune:
            from sinv_amd import autotune
            autotune.load(args.load_tune)
        elif not args.no_tune:
            from sinv_amd import autotune
            table = autotune.tune(lambda: [p.run() for p in plans], reps=2)
            if args.save_tune:
                autotune.save(table, args.save_tune)
        for p in plans:
            p.run()
        torch.cuda.synchronize()
        for p in plans:
            p.check()
        del plans
        if args.workload == "energies":
            # the int8x step too (its split and product kernels have their own names)
            for g in groups:
                g["out"] = None
            plans = make_plans("int8x")
            for p in plans:
                p.run()
            torch.cuda.synchronize()
            for p in plans:
                p.check()
            del plans
        n = 1 << 28  # float64 elements: 2 GiB read + 2 GiB written by one elementwise kernel
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print(json.dumps(dict(traffic_pass=wl["workload"], calib_kernel="neg", calib_read_bytes=8 * n,
                              calib_write_bytes=8 * n)), flush=True)
        return 0

    def timed(plans, steps, warmup, tag):
        # the block-size groups are independent: run them concurrently, one stream each
        pset = core.PlanSet(plans)
        run = pset.run
        if not args.eager:
            # one step captured into a HIP graph (the fork / join across the
            # group streams recorded as dependencies) and replayed
            graph = pset.capture(warmup=max(warmup, 1))
            torch.cuda.synchronize()
            pset.check()
            run = graph.replay
        for _ in range(warmup):
            run()
        torch.cuda.synchronize()
        pset.check()
        # roctx range: lets a rocprofv3 --marker-trace run select this region's dispatches
        torch.cuda.nvtx.range_push(f"sinv_bench_timed:{tag}")
        t0 = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            run()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        torch.cuda.nvtx.range_pop()
        return ev0.elapsed_time(ev1) / steps, wall / steps

    results = {}
    tuned = {}
    for prec in ("fp64", "mixed", "mixed_refined", "int8x"):
        log(f"[bench] {wl['workload']}: batched {prec}: allocating plans")
        plans = make_plans(prec)
        if prec == "int8x":
            pass  # one kernel shape for every product: nothing to tune
        elif args.load_tune:
            from sinv_amd import autotune
            autotune.load(args.load_tune)
        elif not args.no_tune:
            # GEMM tile configs per shape, timed on scratch operands (outside the timed region)
            from sinv_amd import autotune
            t0 = time.perf_counter()
            table = autotune.tune(lambda: [p.run() for p in plans], reps=2)
            if args.save_tune:
                autotune.save(table, args.save_tune + f".{prec}")
            tuned[prec] = dict(shapes=len(table), seconds=round(time.perf_counter() - t0, 2),
                               chosen={str(c): sum(1 for v in table.values() if v["best"] == c)
                                       for c, _ in autotune.configs()})
            log(f"[bench] autotuned {len(table)} GEMM shapes ({prec}) in {tuned[prec]['seconds']} s")
        ms, wall = timed(plans, args.steps, args.warmup, prec)
        log(f"[bench] batched {prec}: {ms:.1f} ms/step (wall {wall * 1e3:.1f}), "
            f"{solves / (ms / 1e3):.1f} solves/s, {alg_flops / (ms / 1e3) / 1e12:.2f} TFLOP/s")
        # kernel-level figures from one more step with the groups in sequence, so each
        # kernel's event-bracketed time is its own (roctx range for a rocprofv3 cross-check)
        torch.cuda.nvtx.range_push(f"sinv_bench_profiled:{prec}")
        with _native.LaunchProfiler() as prof:
            for p in plans:
                p.run()
            torch.cuda.synchronize()
        torch.cuda.nvtx.range_pop()
        results[prec] = dict(ms=ms, prof=prof.stats, bytes=prof.bytes)
        del plans
        torch.cuda.synchronize()

    # ------------------------------------------------------------ error of mixed vs FP64, all instances
    log("[bench] error of mixed, mixed_refined and int8x vs fp64 over every instance")
    err_mixed = 0.0
    err_ref = 0.0
    err_i8x = 0.0
    chunk = 16
    for gi, g in enumerate(groups):
        nE = len(g["shift"]) if g["shift"] is not None else 1
        for s0 in range(0, nE, chunk):
            sl = slice(s0, min(nE, s0 + chunk))
            sh = None if g["shift"] is None else g["shift"][sl]
            outs = []
            for prec in ("fp64", "mixed", "mixed_refined", "int8x"):
                p = core.RgfPlan(*g["dev"], shift=sh, dscale=g["dscale"], oscale=g["oscale"],
                                 precision=prec, batch=None, inverse=args.inverse)
                p.run()
                outs.append(p.outputs())
            err_mixed = max(err_mixed, rel_err_per_instance(outs[1], outs[0]))
            err_ref = max(err_ref, rel_err_per_instance(outs[2], outs[0]))
            err_i8x = max(err_i8x, rel_err_per_instance(outs[3], outs[0]))
            del outs, p
    torch.cuda.synchronize()
    log(f"[bench] max rel err vs fp64: mixed {err_mixed:.3e}, mixed_refined {err_ref:.3e}, "
        f"int8x {err_i8x:.3e}")

    # ------------------------------------------------------------ FP64 vs the CPU oracle (sample)
    err_fp64 = err_floor = err_sample = err_i8x_oracle = None
    if not args.no_cpu:
        from oracle import rgf_ref
        g = groups[0]
        sh = None if g["shift"] is None else g["shift"][:1]
        p = core.RgfPlan(*g["dev"], shift=sh, dscale=g["dscale"], oscale=g["oscale"],
                         inverse=args.inverse)
        p.run()
        p.check()
        got = tuple(x[0].cpu().numpy() for x in p.outputs())
        p8 = core.RgfPlan(*g["dev"], shift=sh, dscale=g["dscale"], oscale=g["oscale"],
                          precision="int8x", inverse=args.inverse)
        p8.run()
        p8.check()
        got8 = tuple(x[0].cpu().numpy() for x in p8.outputs())
        del p8
        if g["shift"] is not None:
            d, u, l = configs.shifted(g["H"], g["shift"][0])
        else:
            d, u, l = g["H"]
        ref = rgf_ref.rgf_blocks(d, u, l)
        num = sum(np.linalg.norm(a - r) ** 2 for a, r in zip(got, ref))
        den = sum(np.linalg.norm(r) ** 2 for r in ref)
        err_fp64 = float(np.sqrt(num / den))
        err_i8x_oracle = float(np.sqrt(sum(np.linalg.norm(a - r) ** 2
                                           for a, r in zip(got8, ref)) / den))
        # the instance's rounding floor: the oracle against itself with every
        # triple product associated the other way (an equally valid FP64 order)
        alt = rgf_ref.rgf_blocks(d, u, l, reassociate=True)
        err_floor = float(np.sqrt(sum(np.linalg.norm(a - r) ** 2 for a, r in zip(alt, ref)) / den))
        err_sample = (f"b = {g['b']} group, {'first energy' if g['shift'] is not None else 'the instance'}: "
                      "relative Frobenius distance over every selected block; the floor is the oracle "
                      "against itself with each triple product associated the other way")
        log(f"[bench] fp64 GPU vs CPU oracle (b={g['b']}, first energy): {err_fp64:.3e} "
            f"(oracle vs its reassociated self: {err_floor:.3e}); int8x: {err_i8x_oracle:.3e}")

    # ------------------------------------------------------------ plain FP64, batch = 1 (sample)
    plain = None
    if not args.no_plain:
        per_b = []
        nsamp = 3
        for g in groups:
            nE = len(g["shift"]) if g["shift"] is not None else 1
            ps = [core.RgfPlan(*g["dev"], shift=None if g["shift"] is None else g["shift"][k:k + 1],
                               dscale=g["dscale"], oscale=g["oscale"], inverse=args.inverse)
                  for k in range(min(nsamp, nE))]
            ps[0].run()
            torch.cuda.synchronize()

            def run_ps():
                for p in ps:
                    p.run()
            if not args.eager:  # the same launch mode as the batched steps
                side = torch.cuda.Stream()
                side.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(side):
                    run_ps()
                torch.cuda.current_stream().wait_stream(side)
                gp = torch.cuda.CUDAGraph()
                with torch.cuda.graph(gp):
                    run_ps()
                torch.cuda.synchronize()
                run_ps = gp.replay
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            run_ps()
            ev1.record(stream)
            torch.cuda.synchronize()
            t = ev0.elapsed_time(ev1) / len(ps)
            per_b.append((g["b"], nE, t))
            log(f"[bench] plain fp64 batch=1 b={g['b']}: {t:.2f} ms/solve")
            del ps, run_ps
        step_ms = sum(nE * t for _, nE, t in per_b)
        plain = dict(solves_per_s=solves / (step_ms / 1e3), ms_per_solve={str(b): t for b, _, t in per_b},
                     sample=f"{nsamp} energies per block size, batch=1 launch sequences",
                     launch="eager" if args.eager else "hip graph replay")

    # ------------------------------------------------------------ CPU baseline (oracle, sampled)
    # numpy restatement of rgf.py (the reference's own inv / @ calls), on the
    # host's cores: median of 5 solves after a warm-up per block size, all
    # BLAS threads; plus one single-thread solve per block size on at most 16
    # blocks, scaled to the block count by the flop count (linear in N)
    cpu = None
    if not args.no_cpu:
        from oracle import rgf_ref
        cores = len(os.sched_getaffinity(0))
        blas = {}
        limit1 = None
        try:
            from threadpoolctl import threadpool_info, threadpool_limits
            info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
            if info:
                cores = int(info[0]["num_threads"])
                blas = {k: info[0].get(k) for k in ("internal_api", "version", "threading_layer")}
            limit1 = threadpool_limits
        except Exception:
            pass
        t_cpu = t_cpu1 = 0.0
        parts = []
        for g in groups:
            nE = len(g["shift"]) if g["shift"] is not None else 1
            if g["shift"] is not None:
                d, u, l = configs.shifted(g["H"], g["shift"][0])
            else:
                d, u, l = g["H"]
            rgf_ref.rgf_blocks(d, u, l)  # warm-up
            runs = []
            for _ in range(5):
                t0 = time.perf_counter()
                rgf_ref.rgf_blocks(d, u, l)
                runs.append(time.perf_counter() - t0)
            t = float(np.median(runs))
            t_cpu += nE * t
            msg = f"b={g['b']}: {t:.3f} s/solve (median of 5, {cores} threads)"
            if limit1 is not None:
                Ns = min(g["N"], 16)
                with limit1(limits=1, user_api="blas"):
                    t0 = time.perf_counter()
                    rgf_ref.rgf_blocks(d[:Ns], u[:Ns - 1], l[:Ns - 1])
                    t1 = (time.perf_counter() - t0) * (9 * g["N"] - 8) / (9 * Ns - 8)
                t_cpu1 += nE * t1
                msg += f", {t1:.3f} s/solve on 1 thread ({Ns} of {g['N']} blocks, scaled)"
            parts.append(msg)
            log(f"[bench] cpu oracle {msg}")
        cpu = dict(value=solves / t_cpu, unit="solves/s", cores=cores, kind="port",
                   single_thread_value=(solves / t_cpu1 if t_cpu1 > 0 else None), blas=blas or None,
                   sample="numpy restatement of rgf.py (same inv / @ calls), one instance per "
                          "block size, scaled to the step's energy counts; " + "; ".join(parts))

    # ------------------------------------------------------------ pick the headline
    fp64_sps = solves / (results["fp64"]["ms"] / 1e3)
    mixed_sps = solves / (results["mixed"]["ms"] / 1e3)
    ref_sps = solves / (results["mixed_refined"]["ms"] / 1e3)
    i8x_sps = solves / (results["int8x"]["ms"] / 1e3)
    # the headline is the fastest mode inside the error gate (FP64 always is)
    cands = {"fp64": (fp64_sps, 0.0), "mixed": (mixed_sps, err_mixed),
             "mixed_refined": (ref_sps, err_ref), "int8x": (i8x_sps, err_i8x)}
    head = max((m for m, (_, e) in cands.items() if e <= ERR_GATE), key=lambda m: cands[m][0])
    ms = results[head]["ms"]

    # ------------------------------------------------------------ PCIe-inclusive rate (never `value`)
    pcie = None
    if not args.no_pcie:
        n = 1 << 28  # 2 GiB of float64 between pinned host memory and HBM
        hbuf = torch.empty(n, dtype=torch.float64, pin_memory=True)
        dbuf = torch.empty(n, dtype=torch.float64, device=dev)

        def copy_rate(fn, reps=3):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            return reps * 8 * n / (time.perf_counter() - t0) / 1e9

        h2d = copy_rate(lambda: dbuf.copy_(hbuf, non_blocking=True))
        d2h = copy_rate(lambda: hbuf.copy_(dbuf, non_blocking=True))
        del hbuf, dbuf
        blocks = lambda g: (3 * g["N"] - 2) * g["b"] ** 2 * 16  # complex128 band of one instance
        bytes_in = sum(blocks(g) for g in groups)  # H (or A) once; energies are a few bytes
        bytes_out = sum((len(g["shift"]) if g["shift"] is not None else 1) * blocks(g)
                        for g in groups)
        t_io = bytes_in / (h2d * 1e9) + bytes_out / (d2h * 1e9)
        pcie = dict(solves_per_s=round(solves / (ms / 1e3 + t_io), 3), h2d_GBps=round(h2d, 1),
                    d2h_GBps=round(d2h, 1), bytes_in=bytes_in, bytes_out=bytes_out,
                    method="headline step + the input band to the device + every selected block "
                           "back to pinned host memory, at the measured copy rates, not overlapped")
        if len(groups) == 1 and groups[0]["shift"] is None:
            # the drop-in itself: host container in, host container out (rgf.py's call surface)
            from sinv_amd.algorithms import rgf as rgf_dropin
            from sinv_amd.bsparse import BlockTridiag
            A_host = BlockTridiag(*groups[0]["H"])
            rgf_dropin.rgf(A_host)
            t0 = time.perf_counter()
            rgf_dropin.rgf(A_host)
            pcie["dropin_call_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
            del A_host
        log(f"[bench] PCIe-inclusive: {pcie}")

    # ------------------------------------------------------------ path-level roofline per mode
    # t_min = F_gemm / P_gemm + F_inv / P_64 over the work the step executes:
    # per solve N inverses (8 b^3 each) and 6 (N - 1) complex block products
    # (the forward sweep's A_{i+1,i} g_i is kept for the backward sweep, which
    # the reference's convention of 8 (N - 1) products does not do), each
    # executed as 3 real products in FP64 (3M: 6 b^3 on the FP64 matrix cores)
    # and as 4 in the complex64 mode (8 b^3 on the FP32 ones)
    nsolve = lambda g: (len(g["shift"]) if g["shift"] is not None else 1)
    f_inv = sum(nsolve(g) * 8.0 * g["b"] ** 3 * g["N"] for g in groups)
    n_prod = sum(nsolve(g) * 6.0 * (g["N"] - 1) * g["b"] ** 3 for g in groups)  # x b^3 units
    path_roof = {}
    for prec, p_gemm, per in (("fp64", PEAK_TFLOPS["gemm_f64"], 6.0),
                              ("mixed", PEAK_TFLOPS["gemm_f32"], 8.0)):
        t_min = (per * n_prod / (p_gemm * 1e12) + f_inv / (PEAK_TFLOPS["gemm_f64"] * 1e12)) * 1e3
        path_roof[prec] = dict(t_min_ms=round(t_min, 3), frac=round(t_min / results[prec]["ms"], 4))
    # int8x: 3 real products x 21 digit pairs x 2 b^3 int8 operations per block
    # product on the int8 matrix cores, the digit split's algorithmic bytes
    # (profiled step) at the HBM peak, the inverses in FP64
    i8_bytes = results["int8x"]["bytes"].get("i8x_split", 0.0)
    t_min = (I8X_OPS_PER_FLOP * 8.0 * n_prod / (PEAK_TFLOPS["gemm_i8x"] * 1e12)
             + i8_bytes / (HBM_TBPS * 1e12) + f_inv / (PEAK_TFLOPS["gemm_f64"] * 1e12)) * 1e3
    path_roof["int8x"] = dict(t_min_ms=round(t_min, 3), frac=round(t_min / results["int8x"]["ms"], 4),
                              split_bytes=i8_bytes)
    path_roof["work"] = dict(
        executed_gemm_flops_fp64=6.0 * n_prod, executed_gemm_flops_mixed=8.0 * n_prod,
        executed_int8_ops=I8X_OPS_PER_FLOP * 8.0 * n_prod,
        inverse_flops=f_inv, reference_convention_flops=alg_flops,
        note="executed work: 6 (N-1) block products per solve (3M in FP64), N inverses at 8 b^3; "
             "the line's tflops count the reference's convention (8 (N-1) products of 8 b^3)")

    # ------------------------------------------------------------ BCR-serial on the same instance
    # (single-instance workloads): GPU time, its distance to RGF, and the numpy
    # restatement of bcr_serial.py in the reference's dense storage where that fits
    bcr = None
    if len(groups) == 1 and groups[0]["shift"] is None and not args.no_bcr:
        g = groups[0]
        d, u, l = g["dev"]
        core.bcr_batched(d, u, l)
        torch.cuda.synchronize()
        reps = 5
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            Gb = core.bcr_batched(d, u, l, check=False)
        ev1.record(stream)
        torch.cuda.synchronize()
        t_gpu = ev0.elapsed_time(ev1) / reps
        Gr = core.rgf_batched(d, u, l)
        num = sum(float(torch.linalg.vector_norm(a - r)) ** 2 for a, r in zip(Gb, Gr))
        den = sum(float(torch.linalg.vector_norm(r)) ** 2 for r in Gr)
        from oracle import bcr_ref
        bcr = dict(ms_per_solve=round(t_gpu, 3), solves_per_s=round(1e3 / t_gpu, 3),
                   rel_err_vs_rgf=float(np.sqrt(num / den)),
                   executed_flops=8 * g["b"] ** 3 * sum(bcr_ref.bcr_flops(g["N"], g["b"])))
        del Gb, Gr
        n = g["N"] * g["b"]
        npad = (g["N"] + bcr_ref.distance_to_power_of_two(g["N"])) * g["b"]
        if not args.no_cpu and 4 * npad * npad * 16 <= 2 ** 33:  # four dense n_pad^2 arrays <= 8 GiB
            from sinv_amd.utils import matu
            A = matu.convertBlkTridiagToDense(*g["H"])
            bcr_ref.bcr_serial_dense(A, g["b"])  # warm-up
            runs = []
            for _ in range(5):
                t0 = time.perf_counter()
                bcr_ref.bcr_serial_dense(A, g["b"])
                runs.append(time.perf_counter() - t0)
            bcr["cpu_dense_ms_per_solve"] = round(float(np.median(runs)) * 1e3, 3)
            bcr["cpu_sample"] = (f"numpy restatement of bcr_serial.py on the dense n = {n} matrix "
                                 f"(padded to {npad}), median of 5, all BLAS threads")
        else:
            bcr["cpu_dense_ms_per_solve"] = None
            bcr["cpu_sample"] = (f"not run: the reference's dense storage needs four {npad}^2 "
                                 f"complex arrays ({4 * npad * npad * 16 / 1e9:.0f} GB)")
        log(f"[bench] bcr-serial: {bcr}")

    kind = {"fp64": "gemm_f64", "mixed": "gemm_f32", "mixed_refined": "gemm_f64",
            "int8x": "gemm_i8x"}[head]
    kms, kcnt, kfl = results[head]["prof"][kind]
    kname = {"fp64": "gemm_kernel<double>", "mixed": "gemm_kernel<float>",
             "mixed_refined": "gemm_kernel<double>", "int8x": "gemm_i8x_o2_kernel<6, true>"}[head]
    tkey = "gemm_i8x_o2_kernel" if head == "int8x" else kname  # name in the traffic summary
    achieved = kfl / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    if head == "int8x":
        # executed int8 operations: 3 real products x 21 digit pairs x 2 MNK = 15.75 x 8 MNK
        achieved *= I8X_OPS_PER_FLOP
    roof = dict(bound="mfma", achieved=round(achieved, 3), peak=PEAK_TFLOPS[kind],
                unit="TOP/s" if head == "int8x" else "TFLOP/s",
                frac=round(achieved / PEAK_TFLOPS[kind], 4),
                traffic=measured_traffic(wl["workload"], tkey),
                kernel=kname, launches=kcnt, avg_launch_us=round(kms / max(kcnt, 1) * 1e3, 3),
                flops_per_launch=kfl / max(kcnt, 1),
                algorithmic_bytes_per_launch=results[head]["bytes"][kind] / max(kcnt, 1),
                # FP64 complex products run as 3 real MFMA products (3M) where 8MNK counts 4:
                # the matrix cores execute 3/4 of the algorithmic flops
                mfma_executed_tflops=round(achieved * (0.75 if head == "fp64" else 1.0), 3),
                mfma_executed_frac=round(achieved * (0.75 if head == "fp64" else 1.0)
                                         / PEAK_TFLOPS[kind], 4),
                traffic_source=f"profiles/traffic_{wl['workload']}.txt (rocprofv3 --pmc "
                               "FETCH_SIZE + WRITE_SIZE, calibrated; null if absent)")
    if roof["traffic"] is not None and kcnt:
        # HBM side of the same kernel: measured bytes per launch over its launch time
        hbm = roof["traffic"] / (kms / kcnt / 1e3) / 1e9
        roof["hbm_GBps"] = round(hbm, 1)
        roof["hbm_frac"] = round(hbm / (HBM_TBPS * 1e3), 4)
    out = {
        "metric": METRIC,
        "value": round(solves / (ms / 1e3), 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "complex128", "mixed": "complex64-gemm/complex128-inverse",
                  "mixed_refined": "complex64-schur-gemm/complex128-refinement",
                  "int8x": "int8-digit-gemm/complex128"}[head],
        "data": "synthetic",
        "config": dict(wl, launch="eager" if args.eager else
                       "hip graph replay of one captured step (groups on their own streams)",
                       inverse={str(g["b"]): g.get("inverse") for g in groups}),
        "tflops": round(alg_flops / (ms / 1e3) / 1e12, 3),
        "headline_mode": head,
        "modes": {
            "batched_fp64": dict(solves_per_s=round(fp64_sps, 3), ms_per_step=round(results["fp64"]["ms"], 3),
                                 err_vs_cpu_oracle=err_fp64, oracle_rounding_floor=err_floor,
                                 err_sample=err_sample),
            "batched_mixed": dict(solves_per_s=round(mixed_sps, 3), ms_per_step=round(results["mixed"]["ms"], 3),
                                  max_rel_err_vs_fp64=err_mixed, passes_err_gate=err_mixed <= ERR_GATE),
            "batched_mixed_refined": dict(
                solves_per_s=round(ref_sps, 3), ms_per_step=round(results["mixed_refined"]["ms"], 3),
                max_rel_err_vs_fp64=err_ref, passes_err_gate=err_ref <= ERR_GATE,
                method="complex64 forward Schur GEMMs, FP64 inverses, one FP64 refinement step "
                       "of each g_i against M_i formed in FP64, FP64 backward sweep"),
            "batched_int8x": dict(solves_per_s=round(i8x_sps, 3), ms_per_step=round(results["int8x"]["ms"], 3),
                                  max_rel_err_vs_fp64=err_i8x, passes_err_gate=err_i8x <= ERR_GATE,
                                  err_vs_cpu_oracle=err_i8x_oracle, digit_planes=6,
                                  gemm_i8x=_kind_summary(results["int8x"], "gemm_i8x"),
                                  i8x_split=_kind_summary(results["int8x"], "i8x_split")),
            "plain_fp64_batch1": plain,
        },
        "speedup_vs_plain_fp64": (round(solves / (ms / 1e3) / plain["solves_per_s"], 2) if plain else None),
        "roofline": roof,
        "cpu_baseline": cpu,
        "path_roofline": path_roof,
        "pcie_inclusive": pcie,
        "bcr_serial": bcr,
        "gemm_autotune": tuned or None,
        "kernel_time": {k: dict(ms=round(v[0], 3), launches=v[1], tflops=round(v[2] / (v[0] / 1e3) / 1e12, 3)
                                if v[0] > 0 else 0.0) for k, v in results[head]["prof"].items() if v[1]},
    }
    if args.workload == "energies" and not args.no_adjoint:
        # the forward+adjoint workload (every mode, at most two timed steps after
        # one warm-up, so the default run stays within minutes) as a second record
        # the headline's outputs and inputs are no longer needed: free them so
        # the second workload has the device to itself
        for g in groups:
            g["out"] = None
            g["dev"] = None
        import gc
        gc.collect()
        torch.cuda.empty_cache()
        log(f"[bench] forward+adjoint workload (second record); "
            f"{torch.cuda.memory_allocated() / 1e9:.1f} GB still allocated")
        out["forward_adjoint"] = adjoint_record(args, modes=("fp64", "mixed", "int8x"),
                                                steps=min(args.steps, 2),
                                                warmup=min(args.warmup, 1),
                                                cpu_baseline=not args.no_cpu)
    print(json.dumps(out), flush=True)
    return 0


def adjoint_record(args, modes=("fp64", "mixed", "int8x"), steps=None, warmup=None,
                   cpu_baseline=True):
    """Batched forward + adjoint with checkpointing: 512 energies in [-2, 2],
    eta = 1e-3, one Hermitian H of 96 blocks whose sizes mix 256 and 416
    inside the instance; loss L = -(1/pi) sum_E sum_i Im tr G_ii(E); the
    gradient dL/dH on the block-tridiagonal pattern, summed over energies.
    A step is the loss and its full gradient over all energies, processed in
    energy chunks (left-connected blocks checkpointed every 8 blocks).
    Returns the record (a dict); modes other than fp64 are compared with it."""
    import torch
    from sinv_amd import core, _native
    from sinv_amd.utils import configs

    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    rng = np.random.default_rng(31)
    bs = [int(x) for x in rng.choice([256, 416], size=96)]
    H = configs.hermitian_H_ragged(bs, seed=32)
    nE = args.energies if args.energies != 256 else 512
    E = configs.energies(nE, seed=33) + 1e-3j
    Hd = [core.to_dev(x) for x in H[0]]
    Hu = [core.to_dev(x) for x in H[1]]
    Hl = [core.to_dev(x) for x in H[2]]
    gd = [(-1j / np.pi) * torch.eye(b, dtype=torch.complex128, device=Hd[0].device) for b in bs]
    zu = [torch.zeros_like(x) for x in Hu]
    zl = [torch.zeros_like(x) for x in Hl]
    every = 8

    def step(prec, limit=None):
        loss = torch.zeros((), dtype=torch.float64, device=Hd[0].device)
        grad = [torch.zeros_like(x) for x in Hd + Hu + Hl]
        for s0 in range(0, nE if limit is None else limit, args.chunk):
            sh = E[s0:s0 + args.chunk]
            # one adjoint call gives the gradient and, from its sweep over A^H, G_ii for the loss
            Ad, Au, Al, Gd = core.rgf_adjoint_ragged(Hd, Hu, Hl, gd, zu, zl, shift=sh, dscale=-1.0,
                                                     oscale=-1.0, precision=prec,
                                                     checkpoint_every=every, check=False,
                                                     return_g_diag=True, sum_batch=True)
            for g in Gd:
                loss -= torch.diagonal(g, dim1=-2, dim2=-1).imag.sum() / np.pi
            for acc, a in zip(grad, Ad + Au + Al):
                acc -= a                 # A = (E + i eta) I - H: dL/dH = -sum_E dL/dA
            del Gd, Ad, Au, Al
        return loss, grad

    if getattr(args, "traffic_pass", False):
        # PMC counters see every dispatch of the process: tune in an earlier
        # run (--save-tune) and load that table here, then count one chunk
        from sinv_amd import autotune
        if args.load_tune:
            autotune.load(args.load_tune)
        elif not args.no_tune:
            table = autotune.tune(lambda: step("fp64", limit=args.chunk), reps=2)
            if args.save_tune:
                autotune.save(table, args.save_tune)
        step("fp64", limit=args.chunk)
        torch.cuda.synchronize()
        n = 1 << 28  # float64 elements: 2 GiB read + 2 GiB written by one elementwise kernel
        src = torch.ones(n, dtype=torch.float64, device=Hd[0].device)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        return dict(traffic_pass="rgf_forward_adjoint", energies=args.chunk, calib_kernel="neg",
                    calib_read_bytes=8 * n, calib_write_bytes=8 * n)

    out = {}
    tuned = {}
    for prec in modes:
        if not args.no_tune and prec != "int8x":
            # every chunk issues the same GEMM shapes: tune on one (untimed)
            from sinv_amd import autotune
            t0 = time.perf_counter()
            table = autotune.tune(lambda: step(prec, limit=args.chunk), reps=2)
            tuned[prec] = dict(shapes=len(table), seconds=round(time.perf_counter() - t0, 2))
            log(f"[bench] autotuned {len(table)} GEMM shapes ({prec}) in {tuned[prec]['seconds']} s")
        log(f"[bench] adjoint workload {prec}: warmup")
        torch.cuda.synchronize()
        base_mem = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        for _ in range(warmup):
            step(prec)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            loss, grad = step(prec)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        # this workload's own peak (memory already held when it started excluded)
        peak = (torch.cuda.max_memory_allocated() - base_mem) / 1e9
        log(f"[bench] {prec}: {ms:.1f} ms/step, {nE / (ms / 1e3):.1f} solves/s (fwd+adj), "
            f"peak {peak:.1f} GB")
        # kernel-level figures from one energy chunk (same shapes as every chunk),
        # inside a roctx range that a rocprofv3 --marker-trace run can select
        torch.cuda.nvtx.range_push(f"sinv_bench_profiled_adjoint:{prec}")
        with _native.LaunchProfiler(max_launches=100000) as prof:
            step(prec, limit=args.chunk)
            torch.cuda.synchronize()
        torch.cuda.nvtx.range_pop()
        out[prec] = dict(ms=ms, loss=float(loss), grad=grad, prof=prof.stats, bytes=prof.bytes,
                         peak_gb=peak)

    def grad_err(prec):
        num = sum(float(((a - b).abs() ** 2).sum()) for a, b in zip(out[prec]["grad"], out["fp64"]["grad"]))
        den = sum(float((b.abs() ** 2).sum()) for b in out["fp64"]["grad"])
        return (float(np.sqrt(num / den)),
                abs(out[prec]["loss"] - out["fp64"]["loss"]) / abs(out["fp64"]["loss"]))

    # plain FP64: one energy per call (batch = 1), on a sample of energies
    nplain = 3
    for k in range(2):  # warm-up (first call allocates)
        core.rgf_adjoint_ragged(Hd, Hu, Hl, gd, zu, zl, shift=E[k:k + 1], dscale=-1.0,
                                oscale=-1.0, checkpoint_every=every, check=False,
                                return_g_diag=True)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(nplain):
        core.rgf_adjoint_ragged(Hd, Hu, Hl, gd, zu, zl, shift=E[k:k + 1], dscale=-1.0, oscale=-1.0,
                                checkpoint_every=every, check=False, return_g_diag=True)
    e1.record()
    torch.cuda.synchronize()
    plain_ms = e0.elapsed_time(e1) / nplain
    plain = dict(solves_per_s=round(1e3 / plain_ms, 3), ms_per_solve=round(plain_ms, 3),
                 sample=f"{nplain} energies, one forward+adjoint call each (batch = 1)")
    log(f"[bench] plain fp64 batch=1 forward+adjoint: {plain_ms:.1f} ms/solve")

    # CPU baseline: the reference has no adjoint, so its forward RGF (the
    # numpy restatement of rgf.py, same inv / @ calls) on sampled energies
    cpu = None
    if cpu_baseline:
        from oracle import rgf_ref
        cores = len(os.sched_getaffinity(0))
        try:
            from threadpoolctl import threadpool_info
            info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
            if info:
                cores = int(info[0]["num_threads"])
        except Exception:
            pass
        d, u, l = configs.shifted_lists(H, E[0])
        rgf_ref.rgf_lists(d[:16], u[:15], l[:15])  # warm-up
        runs = []
        for k in range(3):
            d, u, l = configs.shifted_lists(H, E[k])
            t0 = time.perf_counter()
            rgf_ref.rgf_lists(d, u, l)
            runs.append(time.perf_counter() - t0)
        t = float(np.median(runs))
        cpu = dict(value=round(1.0 / t, 4), unit="solves/s", cores=cores, kind="port",
                   sample=f"forward only (the reference has no adjoint): numpy restatement of "
                          f"rgf.py on this instance at 3 of the {nE} energies, median "
                          f"{t:.2f} s per solve, all BLAS threads")
        log(f"[bench] cpu oracle (forward only): {t:.2f} s/solve")

    fwd_flops = configs.rgf_flops_ragged(bs) * nE
    errs = {m: grad_err(m) for m in modes if m != "fp64"}
    # the headline mode: the fastest whose gradient is inside the gate (FP64 always is)
    head = max((m for m in modes if m == "fp64" or errs[m][0] <= ERR_GATE),
               key=lambda m: -out[m]["ms"])
    ms = out[head]["ms"]
    st = out[head]["prof"]
    # the profiled chunk is one of ceil(nE / chunk) identical chunks
    nchunks = -(-nE // args.chunk)
    executed = sum(v[2] for v in st.values() if v[1]) * nchunks
    kind = "gemm_i8x" if head == "int8x" else "gemm_f64"
    kms, kcnt, kfl = st[kind]
    achieved = kfl / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    if head == "int8x":
        achieved *= I8X_OPS_PER_FLOP  # executed int8 operations (3 x 21 digit pairs x 2 MNK)
        kname, tkey, exe = "gemm_i8x_o2_kernel<6, true>", "gemm_i8x_o2_kernel", 1.0
    else:
        # FP64 complex products run as 3 real MFMA products (3M): 3/4 of 8MNK executed
        kname, tkey, exe = "gemm_kernel<double>", "gemm_kernel<double>", 0.75
    peak = PEAK_TFLOPS[kind]
    roof = dict(bound="mfma", achieved=round(achieved, 3), peak=peak,
                unit="TOP/s" if head == "int8x" else "TFLOP/s", frac=round(achieved / peak, 4),
                traffic=measured_traffic("rgf_forward_adjoint", tkey), kernel=kname,
                launches=kcnt, avg_launch_us=round(kms / max(kcnt, 1) * 1e3, 3),
                flops_per_launch=kfl / max(kcnt, 1),
                algorithmic_bytes_per_launch=out[head]["bytes"][kind] / max(kcnt, 1),
                mfma_executed_tflops=round(achieved * exe, 3),
                mfma_executed_frac=round(achieved * exe / peak, 4),
                timed="launch profiler over one energy chunk (hipEvents on the launch stream)",
                traffic_source="profiles/traffic_rgf_forward_adjoint.txt (rocprofv3 --pmc "
                               "FETCH_SIZE + WRITE_SIZE, calibrated; null if absent)")
    if roof["traffic"] is not None and kcnt:
        hbm = roof["traffic"] / (kms / kcnt / 1e3) / 1e9
        roof["hbm_GBps"] = round(hbm, 1)
        roof["hbm_frac"] = round(hbm / (HBM_TBPS * 1e3), 4)
    mode_rec = {"fp64": dict(solves_per_s=round(nE / (out["fp64"]["ms"] / 1e3), 3),
                             ms_per_step=round(out["fp64"]["ms"], 3),
                             peak_hbm_gb=round(out["fp64"]["peak_gb"], 2))}
    for prec in modes:
        if prec == "fp64":
            continue
        gerr, lerr = errs[prec]
        log(f"[bench] gradient error vs fp64: {prec} {gerr:.3e}")
        mode_rec[prec] = dict(solves_per_s=round(nE / (out[prec]["ms"] / 1e3), 3),
                              ms_per_step=round(out[prec]["ms"], 3),
                              grad_rel_err_vs_fp64=gerr, loss_rel_err_vs_fp64=lerr,
                              passes_err_gate=gerr <= ERR_GATE,
                              peak_hbm_gb=round(out[prec]["peak_gb"], 2))
    mode_rec["plain_fp64_batch1"] = plain
    return {
        "metric": METRIC,
        "value": round(nE / (ms / 1e3), 3),
        "unit": "solves/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8-digit-gemm/complex128" if head == "int8x" else "complex128",
        "data": "synthetic", "headline_mode": head,
        "config": dict(workload="rgf_forward_adjoint_dos", blocks=96, block_sizes="256/416 mixed",
                       energies=nE, eta=1e-3, checkpoint_every=every, energy_chunk=args.chunk,
                       solve="DOS loss and dL/dH per energy from one adjoint sweep (G_ii "
                             "and the gradient come out of the same sweep over A^H)"),
        "executed_tflops": round(executed / (ms / 1e3) / 1e12, 3),
        "forward_algorithmic_tflops": round(fwd_flops / (ms / 1e3) / 1e12, 3),
        "peak_hbm_gb": round(out[head]["peak_gb"], 2),
        "gemm_autotune": tuned or None,
        "modes": mode_rec,
        "speedup_vs_plain_fp64": {m: round(nE / (out[m]["ms"] / 1e3) / plain["solves_per_s"], 2)
                                  for m in modes},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_time": {k: dict(ms=round(v[0], 3), launches=v[1]) for k, v in st.items() if v[1]},
    }


def bench_adjoint(args):
    print(json.dumps(adjoint_record(args, cpu_baseline=not args.no_cpu)), flush=True)
    return 0


def bench_gemm(args):
    """The headline's Schur-complement product shape on its own: 256 products
    b x b x b (complex128), the left operand per energy and the right one
    shared (stride 0), for b in {128, 256, 384}: the FP64 MFMA kernel (tuned
    config) against the int8x emulation (digit split of both operands plus the
    int8 product kernel, both inside the timed call), with the int8x error
    against the FP64 result.  Event-timed over `steps` calls after `warmup`."""
    import torch
    from sinv_amd import core, _native
    dev = core.device()
    gen = torch.Generator(device=dev).manual_seed(7)
    rows = []
    if args.traffic_pass:
        # one call per precision at b = 384 (default tile config) and a 2 GiB
        # calibration copy, untimed, for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs
        b, nb = 384, 256
        A = torch.complex(torch.randn(nb, b, b, dtype=torch.float64, device=dev, generator=gen),
                          torch.randn(nb, b, b, dtype=torch.float64, device=dev, generator=gen))
        H = torch.complex(torch.randn(1, b, b, dtype=torch.float64, device=dev, generator=gen),
                          torch.randn(1, b, b, dtype=torch.float64, device=dev, generator=gen))
        for prec in ("fp64", "int8x"):
            core.zgemm_batched(A, H, precision=prec)
        torch.cuda.synchronize()
        del A, H
        n = 1 << 28
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print(json.dumps(dict(traffic_pass="zgemm_headline_shapes", b=b, products=nb,
                              calib_kernel="neg", calib_read_bytes=8 * n,
                              calib_write_bytes=8 * n)), flush=True)
        return 0
    for b in (128, 256, 384):
        nb = 256
        A = torch.complex(torch.randn(nb, b, b, dtype=torch.float64, device=dev, generator=gen),
                          torch.randn(nb, b, b, dtype=torch.float64, device=dev, generator=gen))
        H = torch.complex(torch.randn(1, b, b, dtype=torch.float64, device=dev, generator=gen),
                          torch.randn(1, b, b, dtype=torch.float64, device=dev, generator=gen))
        row = dict(b=b, batch=nb)
        out = {}
        for prec in ("fp64", "int8x"):
            if prec == "fp64" and not args.no_tune:
                from sinv_amd import autotune
                autotune.tune(lambda: core.zgemm_batched(A, H, precision="fp64"), reps=2)
            for _ in range(max(args.warmup, 1)):
                out[prec] = core.zgemm_batched(A, H, precision=prec)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.steps):
                out[prec] = core.zgemm_batched(A, H, precision=prec)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / args.steps
            with _native.LaunchProfiler() as prof:
                core.zgemm_batched(A, H, precision=prec)
                torch.cuda.synchronize()
            kind = "gemm_f64" if prec == "fp64" else "gemm_i8x"
            call_ms = prof.stats[kind][0] + (prof.stats["i8x_split"][0] if prec == "int8x" else 0.0)
            row[prec] = dict(ms=round(ms, 4), tflops=round(8 * b ** 3 * nb / (ms / 1e3) / 1e12, 2),
                             profiled_call_ms=round(call_ms, 4))
            if prec == "int8x":
                row[prec]["product_ms"] = round(prof.stats["gemm_i8x"][0], 4)
                row[prec]["split_ms"] = round(prof.stats["i8x_split"][0], 4)
        ref = out["fp64"]
        row["int8x"]["rel_err_vs_fp64"] = float((out["int8x"] - ref).norm() / ref.norm())
        row["int8x_over_fp64_time"] = round(row["int8x"]["ms"] / row["fp64"]["ms"], 3)
        log(f"[bench] gemm b={b}: {row}")
        rows.append(row)
        del A, H, out, ref
    res = {
        "metric": "complex GEMM time, FP64 MFMA vs int8x emulation (headline Schur product shapes)",
        "value": rows[-1]["int8x"]["tflops"], "unit": "TFLOP/s (8MNK, int8x, b=384)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
        "dtype": "complex128", "data": "synthetic",
        "config": dict(workload="zgemm_headline_shapes", products=256, right_operand="shared"),
        "shapes": rows,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_convert(args):
    """Device dense <-> block-tridiagonal converters on a 32768 x 32768 complex128
    matrix with 512 x 512 blocks (17.2 GB dense, 64 blocks): gather of the
    3N - 2 band blocks, and the dense result (band blocks + zeros).  HBM-bound:
    algorithmic bytes = band read + band written (gather), band read + n^2
    written (dense); reported against 8 TB/s."""
    import torch
    from sinv_amd import core
    n, b = 32768, 512
    N = n // b
    dev = core.device()
    A = torch.empty((n, n), dtype=torch.complex128, device=dev)
    A.real.normal_()
    band = (3 * N - 2) * b * b * 16
    res = {}

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    blocks = core.dense_to_blocks(A, b)
    ms = timed(lambda: core.dense_to_blocks(A, b), max(args.steps, 1) * 10)
    res["dense_to_blocks"] = dict(ms=round(ms, 4), bytes=2 * band, gbps=round(2 * band / ms / 1e6, 1))
    del A
    torch.cuda.empty_cache()
    ms = timed(lambda: core.blocks_to_dense(*blocks), max(args.steps, 1) * 3)
    out_bytes = band + n * n * 16
    res["blocks_to_dense"] = dict(ms=round(ms, 4), bytes=out_bytes, gbps=round(out_bytes / ms / 1e6, 1))
    log(f"[bench] convert: {res}")
    head = res["blocks_to_dense"]
    print(json.dumps({
        "metric": "GB/s: device dense<->block-tridiagonal converters (matrices_utils.py:115-189)",
        "value": head["gbps"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128", "data": "synthetic",
        "config": dict(workload="dense_block_convert", n=n, blocksize=b, blocks=N),
        "roofline": dict(bound="hbm", achieved=head["gbps"], peak=8000.0, unit="GB/s",
                         frac=round(head["gbps"] / 8000.0, 4), traffic=None,
                         kernel="blocks_to_dense_kernel"),
        "cpu_baseline": None,
        "kernels": res}), flush=True)
    return 0



def bench_variants(args):
    """One large FP64 instance (128 blocks of 512, the diagonally dominant
    generator) through every selected-inversion method: RGF (block inverses by
    the Newton-Schulz iteration, which "auto" picks for this dominant instance,
    and by the pivoted elimination), BCR-serial, RGF-2sided, PSR and PDIV with
    their partitions batched on the GPU.  Per method: device ms per solve,
    executed TFLOP/s (launch profiler flops over the wall time of the solve)
    and relative error against RGF with the pivoted elimination."""
    import torch
    from sinv_amd import core, _native, autotune
    from sinv_amd.utils import configs
    N, b = 128, 512
    d, u, l = (core.to_dev(x) for x in configs.dominant_blocks(N, b, seed=2))
    ref = core.rgf_batched(d, u, l, inverse="pivoted")
    methods = [
        ("rgf", lambda: core.rgf_batched(d, u, l, check=False)),
        ("rgf_pivoted", lambda: core.rgf_batched(d, u, l, check=False, inverse="pivoted")),
        ("bcr_serial", lambda: core.bcr_batched(d, u, l, check=False)),
        ("rgf2sided", lambda: core.psr_batched(d, u, l, 2, first_rows=N // 2, check=False)),
        ("psr_p4", lambda: core.psr_batched(d, u, l, 4, check=False)),
        ("psr_p16", lambda: core.psr_batched(d, u, l, 16, check=False)),
        ("pdiv_p16", lambda: core.pdiv_batched(d, u, l, 16, check=False)),
        ("pdiv_p32", lambda: core.pdiv_batched(d, u, l, 32, check=False)),
    ]
    res = {}
    for name, fn in methods:
        if not args.no_tune:
            autotune.tune(fn, reps=2)
        out = fn()
        torch.cuda.synchronize()
        num = sum(float(((a - r).abs() ** 2).sum()) for a, r in zip(out, ref))
        den = sum(float((r.abs() ** 2).sum()) for r in ref)
        del out
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        reps = max(args.steps, 1)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        # one more solve under the launch profiler, inside a roctx range that a
        # rocprofv3 --marker-trace run can select (sinv_variant:<method>)
        torch.cuda.nvtx.range_push(f"sinv_variant:{name}")
        with _native.LaunchProfiler() as prof:
            fn()
            torch.cuda.synchronize()
        torch.cuda.nvtx.range_pop()
        fl = sum(v[2] for v in prof.stats.values())
        kt = {k: dict(ms=round(v[0], 3), launches=int(v[1])) for k, v in prof.stats.items() if v[1]}
        res[name] = dict(ms=round(ms, 3), executed_tflops=round(fl / (ms / 1e3) / 1e12, 3),
                         rel_err_vs_rgf=float(np.sqrt(num / den)), kernel_time=kt)
        log(f"[bench] {name}: {res[name]}")
    rgf_fl = configs.rgf_flops(N, b)
    head = res["bcr_serial"] if res["bcr_serial"]["ms"] < res["rgf"]["ms"] else res["rgf"]
    best = min(res, key=lambda k: res[k]["ms"])
    print(json.dumps({
        "metric": "ms per single large FP64 selected inversion (128 blocks of 512), every method",
        "value": res[best]["ms"], "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res[best]["ms"], "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "complex128", "data": "synthetic",
        "config": dict(workload="single_instance_all_methods", blocks=N, block_size=b,
                       fastest=best, rgf_algorithmic_tflop=round(rgf_fl / 1e12, 4)),
        "roofline": None, "cpu_baseline": None, "methods": res}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
