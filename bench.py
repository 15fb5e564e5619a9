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
PEAK_TFLOPS = {"gemm_f64": 78.6, "gemm_f32": 157.3}
ERR_GATE = 1e-8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--energies", type=int, default=256)
    ap.add_argument("--workload", default="energies", choices=["energies", "dominant", "large"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the batch=1 FP64 sample")
    return ap.parse_args()


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
              "(energies shard across GPUs with no communication; see DESIGN.md)", file=sys.stderr)
        return 2
    import torch
    from sinv_amd import core, _native
    from sinv_amd.utils import configs

    dev = core.device()
    torch.cuda.synchronize()

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
                             precision=precision, out=g["out"] if energies_slice is None else None)
            if energies_slice is None:
                g["out"] = p.outputs()
            plans.append(p)
        return plans

    stream = torch.cuda.current_stream()

    def timed(plans, steps, warmup, tag):
        for _ in range(warmup):
            for p in plans:
                p.run()
        torch.cuda.synchronize()
        for p in plans:
            p.check()
        # roctx range: lets a rocprofv3 --marker-trace run select this region's dispatches
        torch.cuda.nvtx.range_push(f"sinv_bench_timed:{tag}")
        t0 = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            for p in plans:
                p.run()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        torch.cuda.nvtx.range_pop()
        return ev0.elapsed_time(ev1) / steps, wall / steps

    results = {}
    for prec in ("fp64", "mixed"):
        log(f"[bench] {wl['workload']}: batched {prec}: allocating plans")
        plans = make_plans(prec)
        ms, wall = timed(plans, args.steps, args.warmup, prec)
        log(f"[bench] batched {prec}: {ms:.1f} ms/step (wall {wall * 1e3:.1f}), "
            f"{solves / (ms / 1e3):.1f} solves/s, {alg_flops / (ms / 1e3) / 1e12:.2f} TFLOP/s")
        with _native.LaunchProfiler() as prof:
            for p in plans:
                p.run()
            torch.cuda.synchronize()
        results[prec] = dict(ms=ms, prof=prof.stats)
        del plans
        torch.cuda.synchronize()

    # ------------------------------------------------------------ error of mixed vs FP64, all instances
    log("[bench] error of mixed vs fp64 over every instance")
    err_mixed = 0.0
    chunk = 16
    for gi, g in enumerate(groups):
        nE = len(g["shift"]) if g["shift"] is not None else 1
        for s0 in range(0, nE, chunk):
            sl = slice(s0, min(nE, s0 + chunk))
            sh = None if g["shift"] is None else g["shift"][sl]
            outs = []
            for prec in ("fp64", "mixed"):
                p = core.RgfPlan(*g["dev"], shift=sh, dscale=g["dscale"], oscale=g["oscale"],
                                 precision=prec, batch=None)
                p.run()
                outs.append(p.outputs())
            err_mixed = max(err_mixed, rel_err_per_instance(outs[1], outs[0]))
            del outs, p
    torch.cuda.synchronize()
    log(f"[bench] mixed max rel err vs fp64: {err_mixed:.3e}")

    # ------------------------------------------------------------ FP64 vs the CPU oracle (sample)
    err_fp64 = None
    if not args.no_cpu:
        from oracle import rgf_ref
        g = groups[0]
        sh = None if g["shift"] is None else g["shift"][:1]
        p = core.RgfPlan(*g["dev"], shift=sh, dscale=g["dscale"], oscale=g["oscale"])
        p.run()
        got = tuple(x[0].cpu().numpy() for x in p.outputs())
        if g["shift"] is not None:
            d, u, l = configs.shifted(g["H"], g["shift"][0])
        else:
            d, u, l = g["H"]
        ref = rgf_ref.rgf_blocks(d, u, l)
        num = sum(np.linalg.norm(a - r) ** 2 for a, r in zip(got, ref))
        den = sum(np.linalg.norm(r) ** 2 for r in ref)
        err_fp64 = float(np.sqrt(num / den))
        log(f"[bench] fp64 GPU vs CPU oracle (b={g['b']}, first energy): {err_fp64:.3e}")

    # ------------------------------------------------------------ plain FP64, batch = 1 (sample)
    plain = None
    if not args.no_plain:
        per_b = []
        nsamp = 3
        for g in groups:
            nE = len(g["shift"]) if g["shift"] is not None else 1
            ps = [core.RgfPlan(*g["dev"], shift=None if g["shift"] is None else g["shift"][k:k + 1],
                               dscale=g["dscale"], oscale=g["oscale"]) for k in range(min(nsamp, nE))]
            ps[0].run()
            torch.cuda.synchronize()
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for p in ps:
                p.run()
            ev1.record(stream)
            torch.cuda.synchronize()
            t = ev0.elapsed_time(ev1) / len(ps)
            per_b.append((g["b"], nE, t))
            log(f"[bench] plain fp64 batch=1 b={g['b']}: {t:.2f} ms/solve")
            del ps
        step_ms = sum(nE * t for _, nE, t in per_b)
        plain = dict(solves_per_s=solves / (step_ms / 1e3), ms_per_solve={str(b): t for b, _, t in per_b},
                     sample=f"{nsamp} energies per block size, batch=1 launch sequences")

    # ------------------------------------------------------------ CPU baseline (oracle, sampled)
    cpu = None
    if not args.no_cpu:
        from oracle import rgf_ref
        cores = len(os.sched_getaffinity(0))
        try:
            from threadpoolctl import threadpool_info
            blas = [i for i in threadpool_info() if i.get("user_api") == "blas"]
            if blas:
                cores = int(blas[0]["num_threads"])
        except Exception:
            pass
        t_cpu = 0.0
        parts = []
        for g in groups:
            nE = len(g["shift"]) if g["shift"] is not None else 1
            if g["shift"] is not None:
                d, u, l = configs.shifted(g["H"], g["shift"][0])
            else:
                d, u, l = g["H"]
            rgf_ref.rgf_blocks(d[:2], u[:1], l[:1])  # warm BLAS
            t0 = time.perf_counter()
            rgf_ref.rgf_blocks(d, u, l)
            t = time.perf_counter() - t0
            parts.append(f"b={g['b']}: {t:.2f} s/solve")
            t_cpu += nE * t
            log(f"[bench] cpu oracle b={g['b']}: {t:.2f} s/solve")
        cpu = dict(value=solves / t_cpu, unit="solves/s", cores=cores, kind="port",
                   sample="one solve per block size (numpy restatement of rgf.py, multithreaded "
                          "OpenBLAS), scaled to the step's energy counts; " + ", ".join(parts))

    # ------------------------------------------------------------ pick the headline
    fp64_sps = solves / (results["fp64"]["ms"] / 1e3)
    mixed_sps = solves / (results["mixed"]["ms"] / 1e3)
    head = "fp64"
    if err_mixed <= ERR_GATE and mixed_sps > fp64_sps:
        head = "mixed"
    ms = results[head]["ms"]
    kind = "gemm_f64" if head == "fp64" else "gemm_f32"
    kms, kcnt, kfl = results[head]["prof"][kind]
    achieved = kfl / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    roof = dict(bound="mfma", achieved=round(achieved, 3), peak=PEAK_TFLOPS[kind], unit="TFLOP/s",
                frac=round(achieved / PEAK_TFLOPS[kind], 4), traffic=None,
                kernel=f"gemm_kernel<{'double' if head == 'fp64' else 'float'}>",
                launches=kcnt, avg_launch_us=round(kms / max(kcnt, 1) * 1e3, 3),
                flops_per_launch=kfl / max(kcnt, 1))
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
        "dtype": "complex128" if head == "fp64" else "complex64-gemm/complex128-inverse",
        "data": "synthetic",
        "config": wl,
        "tflops": round(alg_flops / (ms / 1e3) / 1e12, 3),
        "headline_mode": head,
        "modes": {
            "batched_fp64": dict(solves_per_s=round(fp64_sps, 3), ms_per_step=round(results["fp64"]["ms"], 3),
                                 err_vs_cpu_oracle=err_fp64),
            "batched_mixed": dict(solves_per_s=round(mixed_sps, 3), ms_per_step=round(results["mixed"]["ms"], 3),
                                  max_rel_err_vs_fp64=err_mixed, passes_err_gate=err_mixed <= ERR_GATE),
            "plain_fp64_batch1": plain,
        },
        "speedup_vs_plain_fp64": (round(solves / (ms / 1e3) / plain["solves_per_s"], 2) if plain else None),
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_time": {k: dict(ms=round(v[0], 3), launches=v[1], tflops=round(v[2] / (v[0] / 1e3) / 1e12, 3)
                                if v[0] > 0 else 0.0) for k, v in results[head]["prof"].items() if v[1]},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
