"""GEMM tile-shape autotuner.

Every complex GEMM of the library runs one of a few tile configurations
(``configs()``: 64 x 64 with 4 waves, 32 x 32 with 4 or 1 waves, 64 x 32 and
32 x 64, 128 x 64 and 64 x 128 with 8 waves, and 32 x 32 / 16 x 16 tiles whose
waves split k among them).  Which one is fastest depends on the shape and the
batch: small or unbatched products need the small and split-k tiles to fill
256 CUs, large batched ones run best on the default tile.  The tuner

1. runs a workload once under the launch profiler to collect the exact GEMM
   shapes it issues (M, N, K, batch, operand ops, C0 present, precision),
2. times every configuration on each shape with freshly allocated scratch
   operands (never on the workload's own buffers: many of its GEMMs
   accumulate in place and cannot be repeated), and
3. installs the fastest configuration per shape in the library's table, which
   every later launch of that shape consults (including launches recorded into
   a HIP graph after tuning).

``tune(run)`` does all three; ``save``/``load`` keep the table as JSON.
"""

import ctypes
import json

import torch

from . import _native as nat
from . import core

_c128 = torch.complex128


def configs():
    """[(config id, (bm, bn))] of the library's GEMM tile configurations."""
    L = nat.lib()
    out = []
    for c in range(L.sinv_gemm_config_count()):
        bm, bn = ctypes.c_int(), ctypes.c_int()
        nat.check(L.sinv_gemm_config_tile(c, ctypes.byref(bm), ctypes.byref(bn)), "config_tile")
        out.append((c, (bm.value, bn.value)))
    return out


def shapes_of(run, max_launches=400000):
    """Distinct GEMM shape keys (M, N, K, batch, opA, opB, has_c0, precision) that
    one call of ``run()`` launches, with their launch counts."""
    with nat.LaunchProfiler(max_launches=max_launches) as prof:
        run()
        torch.cuda.synchronize()
    keys = {}
    for kind, d, _ in prof.records():
        if kind not in ("gemm_f64", "gemm_f32"):
            continue
        M, N, K, batch, flags = d
        key = (int(M), int(N), int(K), int(batch), int(flags & 3), int((flags >> 2) & 3),
               int((flags >> 4) & 1), nat.PREC["fp64" if kind == "gemm_f64" else "mixed"])
        keys[key] = keys.get(key, 0) + 1
    return keys


def _operands(key):
    M, N, K, batch, opA, opB, has_c0, _ = key
    dev = core.device()
    ar, ac = (M, K) if opA == 0 else (K, M)
    br, bc = (K, N) if opB == 0 else (N, K)
    g = torch.Generator(device=dev).manual_seed(1)
    A = torch.randn((batch, ar, ac), dtype=_c128, device=dev, generator=g)
    B = torch.randn((batch, br, bc), dtype=_c128, device=dev, generator=g)
    C = torch.empty((batch, M, N), dtype=_c128, device=dev)
    C0 = torch.randn((batch, M, N), dtype=_c128, device=dev, generator=g) if has_c0 else None
    return A, B, C, C0


def time_config(key, cfg, reps=3, operands=None):
    """Average ms of one launch of shape `key` with tile configuration `cfg`.

    The launches are timed as the library's drivers issue them: back to back
    on one stream with the host far ahead.  Short kernels are therefore timed
    from a captured HIP graph of at least ``reps`` launches (enough for about
    0.2 ms of GPU time), so that the host's per-launch cost, several times a
    small kernel's duration when launches come one by one from Python, does
    not decide the choice."""
    M, N, K, batch, opA, opB, has_c0, prec = key
    A, B, C, C0 = operands if operands is not None else _operands(key)
    L = nat.lib()
    one = nat.C128(1.0, 0.0)
    beta = nat.C128(1.0, 0.0) if has_c0 else nat.C128(0.0, 0.0)
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None

    def launch():
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        nat.check(L.sinv_zgemm_batched(
            M, N, K, opA, opB, one, p(A), A.shape[2], A.shape[1] * A.shape[2], p(B), B.shape[2],
            B.shape[1] * B.shape[2], beta, p(C0), N, M * N, p(C), N, M * N, None, batch, prec,
            stream), "zgemm")

    def elapsed(fn):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    nat.check(L.sinv_gemm_force_config(cfg), "force_config")
    try:
        launch()
        one_ms = elapsed(launch)
        n = max(reps, min(64, int(0.2 / max(one_ms, 1e-4)) + 1))
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=side):
            for _ in range(n):
                launch()
    finally:
        nat.check(L.sinv_gemm_force_config(-1), "force_config")
    graph.replay()
    ms = elapsed(graph.replay) / n
    del graph
    return ms


def tune_shapes(keys, reps=3, candidates=None, install=True):
    """Time every candidate configuration on each shape; install the fastest.
    Returns {key: {"best": cfg, "ms": {cfg: ms}}}."""
    cands = [c for c, _ in configs()] if candidates is None else list(candidates)
    L = nat.lib()
    table = {}
    for key in keys:
        ops = _operands(key)
        ms = {c: time_config(key, c, reps=reps, operands=ops) for c in cands}
        best = min(ms, key=ms.get)
        table[key] = {"best": best, "ms": ms}
        if install:
            M, N, K, batch, opA, opB, has_c0, prec = key
            nat.check(L.sinv_gemm_tune_set(M, N, K, batch, opA, opB, has_c0, prec, best),
                      "tune_set")
        del ops
    return table


def tune(run, reps=3, candidates=None):
    """Collect the GEMM shapes of ``run()`` and tune them (see module doc)."""
    return tune_shapes(shapes_of(run), reps=reps, candidates=candidates)


def clear():
    nat.check(nat.lib().sinv_gemm_tune_clear(), "tune_clear")


def install(table):
    L = nat.lib()
    for key, v in table.items():
        M, N, K, batch, opA, opB, has_c0, prec = key
        nat.check(L.sinv_gemm_tune_set(M, N, K, batch, opA, opB, has_c0, prec, int(v["best"])),
                  "tune_set")


def save(table, path):
    rows = [dict(key=list(k), best=int(v["best"]), ms={str(c): t for c, t in v["ms"].items()})
            for k, v in table.items()]
    with open(path, "w") as f:
        json.dump(dict(device=torch.cuda.get_device_name(core.device()), entries=rows), f,
                  indent=1)


def load(path, install_now=True):
    with open(path) as f:
        d = json.load(f)
    table = {tuple(r["key"]): {"best": r["best"], "ms": {int(c): t for c, t in r["ms"].items()}}
             for r in d["entries"]}
    if install_now:
        install(table)
    return table


def select(key):
    """Config id a launch of shape `key` would use now."""
    M, N, K, batch, opA, opB, has_c0, prec = key
    return nat.lib().sinv_gemm_select(M, N, K, batch, opA, opB, has_c0, prec)
