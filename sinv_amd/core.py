"""Device-side entry points of the selected-inversion hot path.

Every function here runs on the GPU through libsinv_hip (ctypes).  Inputs may
be numpy arrays (copied to the device) or torch tensors (used in place when
already complex128, contiguous and on the GPU); outputs are torch tensors on
the device.  Calls are ordered on torch's current stream.

Block layout (matrices_utils.py:141-151): diag (N,b,b), upper[i] = A_{i,i+1},
lower[i] = A_{i+1,i}; a leading batch dimension (B, ...) gives one instance
per batch entry, and a missing one means the blocks are shared by every
instance.
"""

import ctypes
import math

import numpy as np
import torch

from . import _native as nat

_c128 = torch.complex128


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("sinv_amd runs on a ROCm GPU; no GPU is visible (there is no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(x):
    """numpy / torch -> contiguous complex128 tensor on the current GPU."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    elif not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x)
    return x.to(device=device(), dtype=_c128).contiguous()


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _info_ptr(info):
    return ctypes.cast(ctypes.c_void_p(info.data_ptr()), ctypes.POINTER(ctypes.c_ulonglong))


def _workspace(nbytes):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device())


def _batched(x, nd):
    """Return (tensor, batch or None, element stride between instances)."""
    if x.dim() == nd + 1:
        return x, x.shape[0], x[0].numel() if x.shape[0] > 0 else 0
    if x.dim() == nd:
        return x, None, 0
    raise ValueError(f"expected a {nd}-d block stack or a batch of them, got shape {tuple(x.shape)}")


def _shape(x):
    """Shape of an array-like input, read without moving it to the device."""
    return tuple(x.shape) if hasattr(x, "shape") else np.shape(x)


def _check_stack(name, x, nblocks, rows, cols):
    """x must be a stack of nblocks rows x cols blocks, optionally batched."""
    s = _shape(x)
    if len(s) not in (3, 4) or s[-3:] != (nblocks, rows, cols):
        raise ValueError(f"{name} must hold {nblocks} blocks of {rows} x {cols} "
                         f"(optionally with a batch dimension), got shape {s}")


def _check_tridiag(diag, upper, lower, gbar=None):
    """Validate a block-tridiagonal stack (and its gradient blocks) before any
    launch: diag (N, b, b), upper / lower (N-1, b, b), each optionally batched.
    Returns (N, b)."""
    s = _shape(diag)
    if len(s) not in (3, 4) or s[-1] != s[-2]:
        raise ValueError(f"diag must be (N, b, b) or (B, N, b, b) with square blocks, got {s}")
    N, b = s[-3], s[-1]
    if N == 0 or b == 0:
        raise ValueError("need at least one block of size at least 1")
    if N > 1:
        _check_stack("upper", upper, N - 1, b, b)
        _check_stack("lower", lower, N - 1, b, b)
    if gbar is not None:
        _check_stack("gbar_diag", gbar[0], N, b, b)
        if N > 1:
            _check_stack("gbar_upper", gbar[1], N - 1, b, b)
            _check_stack("gbar_lower", gbar[2], N - 1, b, b)
    return N, b


def _resolve_batch(*cands):
    bs = {c for c in cands if c is not None}
    if len(bs) > 1:
        raise ValueError(f"inconsistent batch sizes {sorted(bs)}")
    return bs.pop() if bs else 1


# ---------------------------------------------------------------- inverse method
INVERSE_CHOICES = ("auto", "pivoted", "newton")
NEWTON_MAX_BATCH = 4      # include/sinv.h: launches of at most 4 blocks
NEWTON_ACCEPT = 1e-9      # acceptance bound on max |I - M X_k|
NEWTON_AUTO_MAX_ITER = 8  # "auto" takes the iteration only when this many steps suffice


def newton_iterations(rho):
    """Steps of the Newton-Schulz block inverse that are enough for blocks whose
    column dominance ratio is at most ``rho``: from the Jacobi start,
    ||I - M X_k||_1 <= rho^(2^k), so step k passes the acceptance bound once
    rho^(2^k) <= 1e-9 (max |.| <= ||.||_1).  None if rho >= 1."""
    rho = float(rho)
    if not rho < 1.0:
        return None
    if rho < 1e-300:
        return 1
    return max(1, math.ceil(math.log2(math.log(NEWTON_ACCEPT) / math.log(rho))))


def column_dominance(diag, upper, lower, *, shift=None, dscale=1.0, oscale=1.0):
    """Column dominance ratio max_j sum_{i != j} |a_ij| / |a_jj| of each
    block-tridiagonal instance (the RGF convention for diag / upper / lower,
    shift, dscale, oscale), as a float64 tensor on the device.  Below 1 the
    matrix is strictly column diagonally dominant, and so is every block the
    selected-inversion methods invert from it (ratio no larger)."""
    d, Bd, sd = _batched(to_dev(diag), 3)
    N = d.shape[-3]
    if N > 1:
        u, Bu, su = _batched(to_dev(upper), 3)
        l, Bl, sl = _batched(to_dev(lower), 3)
    else:
        u = l = None
        Bu = Bl = None
        su = sl = 0
    sh = to_dev(shift).reshape(-1) if shift is not None else None
    B = _resolve_batch(Bd, Bu, Bl, None if sh is None else sh.numel())
    return _dominance(d, sd, u, su, l, sl, sh, B, N, d.shape[-1], dscale, oscale)


def _dominance(d, sd, u, su, l, sl, sh, B, N, b, dscale, oscale):
    rho = torch.empty(B, dtype=torch.float64, device=device())
    desc = nat.RgfDesc(nblocks=N, blocksize=b, bsizes=None, batch=B, precision=0,
                       diag=_p(d), diag_stride=sd, upper=_p(u), upper_stride=su,
                       lower=_p(l), lower_stride=sl, dscale=nat.C128.of(dscale),
                       oscale=nat.C128.of(oscale), shift=_p(sh))
    nat.check(nat.lib().sinv_column_dominance(ctypes.byref(desc), _p(rho), _stream()),
              "column_dominance")
    return rho


def _choose_inverse(inverse, B, b, rho_fn):
    """(method, iteration limit) for a call of B instances of b-blocks."""
    if inverse not in INVERSE_CHOICES:
        raise ValueError(f"inverse must be one of {INVERSE_CHOICES}")
    if inverse == "pivoted":
        return "pivoted", 0
    if inverse == "newton":
        return "newton", 0
    # the dominance test synchronises, which a stream under graph capture
    # cannot: there "auto" takes the elimination (pass "newton" to capture
    # the iteration)
    if B > NEWTON_MAX_BATCH or b <= 64 or torch.cuda.is_current_stream_capturing():
        return "pivoted", 0
    k = newton_iterations(rho_fn().max().item())
    if k is None or k > NEWTON_AUTO_MAX_ITER:
        return "pivoted", 0
    return "newton", k


def _run_with_inverse(inverse, B, b, rho_fn, launch, info, check, what):
    """Run ``launch()`` (workspace query and library call) under the inverse
    method ``inverse`` resolves to; with "auto", a Newton-Schulz block inverse
    that misses its bound makes the call run again with the elimination."""
    method, iters = _choose_inverse(inverse, B, b, rho_fn)
    with nat.inverse_method(method, iters):
        launch()
    if check or (method == "newton" and inverse == "auto"):
        try:
            nat.check_info(info, what)
        except nat.NotConverged:
            if inverse != "auto":
                raise
            with nat.inverse_method("pivoted"):
                launch()
            if check:
                nat.check_info(info, what)
    return method


def rgf_batched(diag, upper, lower, *, sym_mat=False, save_off_diag=True, precision="fp64",
                shift=None, dscale=1.0, oscale=1.0, check=True, inverse="auto"):
    """RGF selected inversion of a batch of block-tridiagonal instances.

    Instance z is A_ii = dscale*diag[z][i] + shift[z]*I, A_{i,i+1} = oscale*upper[z][i],
    A_{i+1,i} = oscale*lower[z][i]; blocks without a batch dimension are shared.
    Returns (G_diag (B,N,b,b), G_upper (B,N-1,b,b), G_lower (B,N-1,b,b)).
    Follows rgf.py:40-62; with save_off_diag=False the off-diagonals are NaN
    (rgf.py:41,55), with sym_mat=True G_{i,i+1} = G_{i+1,i}^T (rgf.py:58).

    precision (the same choices for every RGF entry point; the adjoint takes
    all but "mixed_refined"):
      "fp64"           complex128 throughout, the reference's arithmetic;
      "int8x"          complex128 products formed from 6 int8 digit planes on the
                       int8 matrix cores, FP64 inverses: within 1e-9 of "fp64"
                       on the benchmark instances;
      "mixed"          complex64 products, FP64 inverses: about 1e-7 relative
                       rounding per product, amplified by the conditioning of
                       the Schur complements -- errors of 1e-3 relative near
                       the band at eta = 1e-3 are measured;
      "mixed_refined"  complex64 forward products, one FP64 refinement step of
                       each g_i, FP64 backward sweep: up to ~1e-7 on the same
                       instances.

    inverse: how the blocks M_i are inverted (rgf.py:44,48 call numpy.linalg.inv):
      "pivoted"  Gauss-Jordan elimination with partial pivoting over all rows,
                 numpy's semantics, for every input;
      "newton"   Newton-Schulz iteration on the FP64 GEMM for calls of at most
                 4 instances (blocks above 64), which converges for column
                 diagonally dominant blocks; raises NotConverged otherwise;
      "auto"     (default) "newton" when the call is latency-bound (at most 4
                 instances of blocks above 64) and the matrix is strictly
                 column diagonally dominant (``column_dominance`` below 1, one
                 device reduction and synchronisation), with the iteration
                 limit that ratio guarantees; "pivoted" otherwise, when an
                 iteration misses its bound, and on a stream under graph
                 capture (which cannot synchronise).
    """
    if precision not in nat.PREC:
        raise ValueError(f"precision must be one of {sorted(nat.PREC)}")
    _check_tridiag(diag, upper, lower)
    d, Bd, sd = _batched(to_dev(diag), 3)
    N, b = d.shape[-3], d.shape[-1]
    if N > 1:
        u, Bu, su = _batched(to_dev(upper), 3)
        l, Bl, sl = _batched(to_dev(lower), 3)
        if u.shape[-3:] != (N - 1, b, b) or l.shape[-3:] != (N - 1, b, b):
            raise ValueError("upper/lower must hold N-1 blocks of the diagonal block size")
    else:
        u = l = None
        Bu = Bl = None
        su = sl = 0
    sh = None
    Bs = None
    if shift is not None:
        sh = to_dev(shift).reshape(-1)
        Bs = sh.numel()
    B = _resolve_batch(Bd, Bu, Bl, Bs)
    dev = device()
    Gd = torch.empty((B, N, b, b), dtype=_c128, device=dev)
    Gu = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev)
    Gl = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev)
    info = torch.empty(1, dtype=torch.int64, device=dev)
    desc = nat.RgfDesc(
        nblocks=N, blocksize=b, bsizes=None, batch=B, sym_mat=int(bool(sym_mat)),
        save_off_diag=int(bool(save_off_diag)), precision=nat.PREC[precision],
        diag=_p(d), diag_stride=sd, upper=_p(u), upper_stride=su, lower=_p(l), lower_stride=sl,
        dscale=nat.C128.of(dscale), oscale=nat.C128.of(oscale), shift=_p(sh),
        g_diag=_p(Gd), g_upper=_p(Gu), g_lower=_p(Gl), info=_info_ptr(info))
    L = nat.lib()

    def launch():
        ws = _workspace(L.sinv_rgf_workspace_bytes(ctypes.byref(desc)))
        nat.check(L.sinv_rgf_batched(ctypes.byref(desc), _p(ws), ws.numel(), _stream()), "rgf")

    _run_with_inverse(inverse, B, b,
                      lambda: _dominance(d, sd, u, su, l, sl, sh, B, N, b, dscale, oscale),
                      launch, info, check, "rgf")
    return Gd, Gu, Gl


class RgfPlan:
    """Pre-allocated batched RGF: inputs, outputs, workspace and descriptor are
    bound once; ``run()`` only enqueues kernels on the current stream (nothing
    is allocated or synchronised).  Same semantics as ``rgf_batched``; the
    inverse method (default "pivoted") is fixed at construction, "auto"
    deciding on the inputs bound then, and ``check()`` raises NotConverged if
    a Newton-Schulz inverse missed its bound."""

    def __init__(self, diag, upper, lower, *, batch=None, sym_mat=False, save_off_diag=True,
                 precision="fp64", shift=None, dscale=1.0, oscale=1.0, out=None,
                 inverse="pivoted"):
        if precision not in nat.PREC:
            raise ValueError(f"precision must be one of {sorted(nat.PREC)}")
        _check_tridiag(diag, upper, lower)
        self.d, Bd, sd = _batched(to_dev(diag), 3)
        N, b = self.d.shape[-3], self.d.shape[-1]
        if N > 1:
            self.u, Bu, su = _batched(to_dev(upper), 3)
            self.l, Bl, sl = _batched(to_dev(lower), 3)
        else:
            self.u = self.l = None
            Bu = Bl = None
            su = sl = 0
        self.shift = to_dev(shift).reshape(-1) if shift is not None else None
        B = _resolve_batch(Bd, Bu, Bl, None if shift is None else self.shift.numel(), batch)
        dev = device()
        if out is None:
            out = (torch.empty((B, N, b, b), dtype=_c128, device=dev),
                   torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev),
                   torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev))
        self.Gd, self.Gu, self.Gl = out
        self.info = torch.empty(1, dtype=torch.int64, device=dev)
        self.desc = nat.RgfDesc(
            nblocks=N, blocksize=b, bsizes=None, batch=B, sym_mat=int(bool(sym_mat)),
            save_off_diag=int(bool(save_off_diag)), precision=nat.PREC[precision],
            diag=_p(self.d), diag_stride=sd, upper=_p(self.u), upper_stride=su,
            lower=_p(self.l), lower_stride=sl, dscale=nat.C128.of(dscale),
            oscale=nat.C128.of(oscale), shift=_p(self.shift), g_diag=_p(self.Gd),
            g_upper=_p(self.Gu), g_lower=_p(self.Gl), info=_info_ptr(self.info))
        self.lib = nat.lib()
        # the method is fixed here ("auto" decides on the inputs bound now; a
        # plan whose inputs change later may choose "pivoted" explicitly)
        self.inverse = _choose_inverse(
            inverse, B, b, lambda: _dominance(self.d, sd, self.u, su, self.l, sl, self.shift, B,
                                              N, b, dscale, oscale))
        with nat.inverse_method(*self.inverse):
            self.ws = _workspace(self.lib.sinv_rgf_workspace_bytes(ctypes.byref(self.desc)))
        self.N, self.b, self.batch = N, b, B

    def run(self):
        with nat.inverse_method(*self.inverse):
            nat.check(self.lib.sinv_rgf_batched(ctypes.byref(self.desc), _p(self.ws),
                                                 self.ws.numel(), _stream()), "rgf")

    def check(self):
        nat.check_info(self.info)

    def outputs(self):
        return self.Gd, self.Gu, self.Gl

    def capture(self, warmup=1):
        """Capture the whole launch sequence of one ``run()`` into a HIP graph
        (torch.cuda.CUDAGraph over the HIP runtime).  Returns the graph; its
        ``replay()`` re-runs every kernel of the sweep with one launch call.
        The sequence holds kernels only (the info word is cleared by a fill
        kernel, not a memset), so replays are exact repeats."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.run()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g


class PlanSet:
    """Independent plans (e.g. one ``RgfPlan`` per block size) run concurrently:
    ``run()`` forks from the current stream onto one stream per plan and joins
    back, so one plan's latency-bound launches (leaf inverses, small GEMM
    tails) overlap another's GEMMs.  Streams have the default priority.
    Nothing is allocated per ``run()``; results are ready once the current
    stream reaches the join."""

    def __init__(self, plans):
        self.plans = list(plans)
        self.streams = [torch.cuda.Stream(device=device()) for _ in self.plans]

    def run(self):
        cur = torch.cuda.current_stream()
        for st, p in zip(self.streams, self.plans):
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                p.run()
        for st in self.streams:
            cur.wait_stream(st)

    def check(self):
        for p in self.plans:
            p.check()

    def capture(self, warmup=1):
        """Capture one concurrent ``run()`` into a HIP graph; the fork/join
        across the plan streams is recorded as graph dependencies."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.run()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g


def bcr_batched(diag, upper, lower, *, shift=None, dscale=1.0, oscale=1.0, return_fill=False,
                check=True, npart=1, inverse="auto"):
    """BCR-serial selected inversion of a batch of instances (bcr_serial.py:19-466).

    Same instance convention as ``rgf_batched`` (uniform block size).  Returns
    (G_diag, G_upper, G_lower) and, with ``return_fill``, the produced blocks
    of levels >= 1 as (B, F, b, b) in the order documented in include/sinv.h.
    ``npart > 1`` issues every level operation per partition of the padded
    block rows, as BCR-parallel's ranks split them (sinv_bcr_parallel_batched;
    same bits).  ``inverse`` as for ``rgf_batched``; with "newton" the
    iteration serves the launches of at most 4 blocks (the upper levels and
    the root), the elimination the larger ones.
    """
    _check_tridiag(diag, upper, lower)
    d, Bd, sd = _batched(to_dev(diag), 3)
    N, b = d.shape[-3], d.shape[-1]
    if N > 1:
        u, Bu, su = _batched(to_dev(upper), 3)
        l, Bl, sl = _batched(to_dev(lower), 3)
    else:
        u = l = None
        Bu = Bl = None
        su = sl = 0
    sh = to_dev(shift).reshape(-1) if shift is not None else None
    B = _resolve_batch(Bd, Bu, Bl, None if sh is None else sh.numel())
    dev = device()
    L = nat.lib()
    Gd = torch.empty((B, N, b, b), dtype=_c128, device=dev)
    Gu = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev)
    Gl = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev)
    F = int(L.sinv_bcr_fill_blocks(N))
    Gf = torch.empty((B, F, b, b), dtype=_c128, device=dev) if (return_fill and F > 0) else None
    info = torch.empty(1, dtype=torch.int64, device=dev)
    desc = nat.BcrDesc(
        nblocks=N, blocksize=b, batch=B, diag=_p(d), diag_stride=sd, upper=_p(u), upper_stride=su,
        lower=_p(l), lower_stride=sl, dscale=nat.C128.of(dscale), oscale=nat.C128.of(oscale),
        shift=_p(sh), g_diag=_p(Gd), g_upper=_p(Gu), g_lower=_p(Gl), g_fill=_p(Gf),
        info=_info_ptr(info))
    def launch():
        ws = _workspace(L.sinv_bcr_workspace_bytes(ctypes.byref(desc)))
        if npart == 1:
            nat.check(L.sinv_bcr_batched(ctypes.byref(desc), _p(ws), ws.numel(), _stream()),
                      "bcr")
        else:
            nat.check(L.sinv_bcr_parallel_batched(ctypes.byref(desc), int(npart), _p(ws),
                                                  ws.numel(), _stream()), "bcr_parallel")

    _run_with_inverse(inverse, B, b,
                      lambda: _dominance(d, sd, u, su, l, sl, sh, B, N, b, dscale, oscale),
                      launch, info, check, "bcr")
    if return_fill:
        if Gf is None:
            Gf = torch.empty((B, 0, b, b), dtype=_c128, device=dev)
        return Gd, Gu, Gl, Gf
    return Gd, Gu, Gl


def rgf_adjoint_batched(diag, upper, lower, gbar_diag, gbar_upper, gbar_lower, *, shift=None,
                        dscale=1.0, oscale=1.0, precision="fp64", checkpoint_every=0,
                        check=True, return_g_diag=False):
    """Gradient of a real loss w.r.t. the blocks of A, given its gradient w.r.t.
    the selected blocks of G = A^-1: grad_A = -G^H grad_G G^H on the pattern.

    Instance convention as ``rgf_batched``; the grad_G blocks may be shared by
    all instances (no batch dimension).  Returns (grad_diag, grad_upper,
    grad_lower), each with a batch dimension.  ``checkpoint_every = k`` keeps
    the forward's left-connected blocks every k blocks only and recomputes the
    rest segment by segment.
    """
    if precision not in nat.PREC:
        raise ValueError(f"precision must be one of {sorted(nat.PREC)}")
    _check_tridiag(diag, upper, lower, gbar=(gbar_diag, gbar_upper, gbar_lower))
    d, Bd, sd = _batched(to_dev(diag), 3)
    N, b = d.shape[-3], d.shape[-1]
    gd, Bgd, sgd = _batched(to_dev(gbar_diag), 3)
    if N > 1:
        u, Bu, su = _batched(to_dev(upper), 3)
        l, Bl, sl = _batched(to_dev(lower), 3)
        gu, Bgu, sgu = _batched(to_dev(gbar_upper), 3)
        gl, Bgl, sgl = _batched(to_dev(gbar_lower), 3)
    else:
        u = l = gu = gl = None
        Bu = Bl = Bgu = Bgl = None
        su = sl = sgu = sgl = 0
    sh = to_dev(shift).reshape(-1) if shift is not None else None
    B = _resolve_batch(Bd, Bu, Bl, Bgd, Bgu, Bgl, None if sh is None else sh.numel())
    dev = device()
    Ad = torch.empty((B, N, b, b), dtype=_c128, device=dev)
    Au = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev)
    Al = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev)
    Gd = torch.empty((B, N, b, b), dtype=_c128, device=dev) if return_g_diag else None
    info = torch.empty(1, dtype=torch.int64, device=dev)
    desc = nat.AdjDesc(
        nblocks=N, blocksize=b, batch=B, precision=nat.PREC[precision],
        checkpoint_every=int(checkpoint_every),
        diag=_p(d), diag_stride=sd, upper=_p(u), upper_stride=su, lower=_p(l), lower_stride=sl,
        dscale=nat.C128.of(dscale), oscale=nat.C128.of(oscale), shift=_p(sh),
        gbar_diag=_p(gd), gbar_diag_stride=sgd, gbar_upper=_p(gu), gbar_upper_stride=sgu,
        gbar_lower=_p(gl), gbar_lower_stride=sgl,
        abar_diag=_p(Ad), abar_upper=_p(Au), abar_lower=_p(Al), g_diag=_p(Gd),
        info=_info_ptr(info))
    L = nat.lib()
    ws = _workspace(L.sinv_rgf_adjoint_workspace_bytes(ctypes.byref(desc)))
    nat.check(L.sinv_rgf_adjoint_batched(ctypes.byref(desc), _p(ws), ws.numel(), _stream()),
              "rgf_adjoint")
    if check:
        nat.check_info(info)
    if return_g_diag:
        return Ad, Au, Al, Gd
    return Ad, Au, Al


# ---------------------------------------------------------------------------
# ragged block sizes (per-block b_i): lists of blocks, packed back to back on
# the device in the layout documented with sinv_rgf_desc.bsizes


def _pack(blocks):
    """List of (b_r x b_c) or (B, b_r, b_c) blocks -> (packed tensor, batch or
    None, instance stride)."""
    ts = [to_dev(b) for b in blocks]
    Bs = {t.shape[0] for t in ts if t.dim() == 3}
    if len(Bs) > 1:
        raise ValueError(f"inconsistent batch sizes {sorted(Bs)}")
    if not ts:
        return None, None, 0
    if Bs:
        B = Bs.pop()
        flat = [(t if t.dim() == 3 else t.expand(B, *t.shape)).reshape(B, -1) for t in ts]
        packed = torch.cat(flat, dim=1).contiguous()
        return packed, B, packed.shape[1]
    return torch.cat([t.reshape(-1) for t in ts]).contiguous(), None, 0


def _unpack(packed, shapes):
    out, o = [], 0
    for r, c in shapes:
        out.append(packed[:, o:o + r * c].view(packed.shape[0], r, c))
        o += r * c
    return out


def _ragged_sizes(diag, upper, lower):
    bs = [int(np.shape(d)[-1]) for d in diag]
    N = len(bs)
    for i, d in enumerate(diag):
        if tuple(np.shape(d)[-2:]) != (bs[i], bs[i]):
            raise ValueError(f"diagonal block {i} is not square")
    if len(upper) != N - 1 or len(lower) != N - 1:
        raise ValueError("need N-1 upper and lower blocks")
    for i in range(N - 1):
        if tuple(np.shape(upper[i])[-2:]) != (bs[i], bs[i + 1]):
            raise ValueError(f"upper block {i} must be {bs[i]} x {bs[i + 1]}")
        if tuple(np.shape(lower[i])[-2:]) != (bs[i + 1], bs[i]):
            raise ValueError(f"lower block {i} must be {bs[i + 1]} x {bs[i]}")
    return bs


def rgf_ragged(diag, upper, lower, *, sym_mat=False, save_off_diag=True, precision="fp64",
               shift=None, dscale=1.0, oscale=1.0, check=True):
    """RGF with per-block sizes: diag[i] is b_i x b_i, upper[i] b_i x b_{i+1},
    lower[i] b_{i+1} x b_i (each optionally with a leading batch dimension).
    Returns lists of device tensors (B, ...) for the diagonal, upper and lower
    blocks of G."""
    if precision not in nat.PREC:
        raise ValueError(f"precision must be one of {sorted(nat.PREC)}")
    bs = _ragged_sizes(diag, upper, lower)
    N = len(bs)
    d, Bd, sd = _pack(diag)
    u, Bu, su = _pack(upper)
    l, Bl, sl = _pack(lower)
    sh = to_dev(shift).reshape(-1) if shift is not None else None
    B = _resolve_batch(Bd, Bu, Bl, None if sh is None else sh.numel())
    dsize = sum(b * b for b in bs)
    osize = sum(bs[i] * bs[i + 1] for i in range(N - 1))
    dev = device()
    Gd = torch.empty((B, dsize), dtype=_c128, device=dev)
    Gu = torch.empty((B, max(osize, 1)), dtype=_c128, device=dev)
    Gl = torch.empty((B, max(osize, 1)), dtype=_c128, device=dev)
    info = torch.empty(1, dtype=torch.int64, device=dev)
    bsz = (ctypes.c_int * N)(*bs)
    desc = nat.RgfDesc(
        nblocks=N, blocksize=0, bsizes=bsz, batch=B, sym_mat=int(bool(sym_mat)),
        save_off_diag=int(bool(save_off_diag)), precision=nat.PREC[precision],
        diag=_p(d), diag_stride=sd, upper=_p(u), upper_stride=su, lower=_p(l), lower_stride=sl,
        dscale=nat.C128.of(dscale), oscale=nat.C128.of(oscale), shift=_p(sh),
        g_diag=_p(Gd), g_upper=_p(Gu), g_lower=_p(Gl), info=_info_ptr(info))
    L = nat.lib()
    ws = _workspace(L.sinv_rgf_workspace_bytes(ctypes.byref(desc)))
    nat.check(L.sinv_rgf_batched(ctypes.byref(desc), _p(ws), ws.numel(), _stream()), "rgf")
    if check:
        nat.check_info(info)
    return (_unpack(Gd, [(b, b) for b in bs]),
            _unpack(Gu, [(bs[i], bs[i + 1]) for i in range(N - 1)]),
            _unpack(Gl, [(bs[i + 1], bs[i]) for i in range(N - 1)]))


def rgf_adjoint_ragged(diag, upper, lower, gbar_diag, gbar_upper, gbar_lower, *, shift=None,
                       dscale=1.0, oscale=1.0, precision="fp64", checkpoint_every=0, check=True,
                       return_g_diag=False, sum_batch=False):
    """``rgf_adjoint_batched`` with per-block sizes (lists of blocks, as
    ``rgf_ragged``).  Returns lists of gradient blocks (B, ...); with
    ``sum_batch=True`` the gradient blocks are summed over the batch (one
    reduction over each packed output, e.g. dL/dA summed over energies) and
    have no batch dimension (G_ii, when returned, keeps it)."""
    if precision not in nat.PREC:
        raise ValueError(f"precision must be one of {sorted(nat.PREC)}")
    bs = _ragged_sizes(diag, upper, lower)
    if _ragged_sizes(gbar_diag, gbar_upper, gbar_lower) != bs:
        raise ValueError("gbar blocks must have the sizes of the matrix blocks")
    N = len(bs)
    d, Bd, sd = _pack(diag)
    u, Bu, su = _pack(upper)
    l, Bl, sl = _pack(lower)
    gd, Bgd, sgd = _pack(gbar_diag)
    gu, Bgu, sgu = _pack(gbar_upper)
    gl, Bgl, sgl = _pack(gbar_lower)
    sh = to_dev(shift).reshape(-1) if shift is not None else None
    B = _resolve_batch(Bd, Bu, Bl, Bgd, Bgu, Bgl, None if sh is None else sh.numel())
    dsize = sum(b * b for b in bs)
    osize = sum(bs[i] * bs[i + 1] for i in range(N - 1))
    dev = device()
    Ad = torch.empty((B, dsize), dtype=_c128, device=dev)
    Au = torch.empty((B, max(osize, 1)), dtype=_c128, device=dev)
    Al = torch.empty((B, max(osize, 1)), dtype=_c128, device=dev)
    Gd = torch.empty((B, dsize), dtype=_c128, device=dev) if return_g_diag else None
    info = torch.empty(1, dtype=torch.int64, device=dev)
    bsz = (ctypes.c_int * N)(*bs)
    desc = nat.AdjDesc(
        nblocks=N, blocksize=0, bsizes=bsz, batch=B, precision=nat.PREC[precision],
        checkpoint_every=int(checkpoint_every),
        diag=_p(d), diag_stride=sd, upper=_p(u), upper_stride=su, lower=_p(l), lower_stride=sl,
        dscale=nat.C128.of(dscale), oscale=nat.C128.of(oscale), shift=_p(sh),
        gbar_diag=_p(gd), gbar_diag_stride=sgd, gbar_upper=_p(gu), gbar_upper_stride=sgu,
        gbar_lower=_p(gl), gbar_lower_stride=sgl,
        abar_diag=_p(Ad), abar_upper=_p(Au), abar_lower=_p(Al), g_diag=_p(Gd),
        info=_info_ptr(info))
    L = nat.lib()
    ws = _workspace(L.sinv_rgf_adjoint_workspace_bytes(ctypes.byref(desc)))
    nat.check(L.sinv_rgf_adjoint_batched(ctypes.byref(desc), _p(ws), ws.numel(), _stream()),
              "rgf_adjoint")
    if check:
        nat.check_info(info)
    if sum_batch:
        Ad, Au, Al = (x.sum(0, keepdim=True) for x in (Ad, Au, Al))
    out = (_unpack(Ad, [(b, b) for b in bs]),
           _unpack(Au, [(bs[i], bs[i + 1]) for i in range(N - 1)]),
           _unpack(Al, [(bs[i + 1], bs[i]) for i in range(N - 1)]))
    if sum_batch:
        out = tuple([x[0] for x in part] for part in out)
    if return_g_diag:
        return out + (_unpack(Gd, [(b, b) for b in bs]),)
    return out


def psr_batched(diag, upper, lower, npart, *, shift=None, dscale=1.0, oscale=1.0, check=True,
                first_rows=0, inverse="auto"):
    """PSR selected inversion (psr_seqsolve.py:20-72) with ``npart`` partitions
    on one GPU: partition-local Schur reduction and production batched over
    partitions, the 2(npart-1)-block reduced system solved by RGF.  Needs
    nblocks >= 3 * npart (psr_utils.check_input).  Same instance convention,
    outputs and ``inverse`` choices as ``rgf_batched`` (the iteration serves
    launches of at most 4 blocks)."""
    N, b = _check_tridiag(diag, upper, lower)
    if first_rows:
        if npart != 2 or not 1 <= first_rows < N:
            raise ValueError("an explicit first partition needs npart == 2 and 1 <= first_rows < N")
    elif N < 3 * npart or npart < 2:
        raise ValueError("The number of blocks is to low. There should be at least 3 blockrows "
                         "per partition (and at least 2 partitions)")
    d, Bd, sd = _batched(to_dev(diag), 3)
    u, Bu, su = _batched(to_dev(upper), 3)
    l, Bl, sl = _batched(to_dev(lower), 3)
    sh = to_dev(shift).reshape(-1) if shift is not None else None
    B = _resolve_batch(Bd, Bu, Bl, None if sh is None else sh.numel())
    dev = device()
    Gd = torch.empty((B, N, b, b), dtype=_c128, device=dev)
    Gu = torch.empty((B, N - 1, b, b), dtype=_c128, device=dev)
    Gl = torch.empty((B, N - 1, b, b), dtype=_c128, device=dev)
    info = torch.empty(1, dtype=torch.int64, device=dev)
    desc = nat.PsrDesc(
        nblocks=N, blocksize=b, batch=B, npart=int(npart), first_rows=int(first_rows),
        diag=_p(d), diag_stride=sd,
        upper=_p(u), upper_stride=su, lower=_p(l), lower_stride=sl, dscale=nat.C128.of(dscale),
        oscale=nat.C128.of(oscale), shift=_p(sh), g_diag=_p(Gd), g_upper=_p(Gu), g_lower=_p(Gl),
        info=_info_ptr(info))
    L = nat.lib()

    def launch():
        ws = _workspace(L.sinv_psr_workspace_bytes(ctypes.byref(desc)))
        nat.check(L.sinv_psr_batched(ctypes.byref(desc), _p(ws), ws.numel(), _stream()), "psr")

    _run_with_inverse(inverse, B, b,
                      lambda: _dominance(d, sd, u, su, l, sl, sh, B, N, b, dscale, oscale),
                      launch, info, check, "psr")
    return Gd, Gu, Gl


def pdiv_batched(diag, upper, lower, npart, *, check=True):
    """PDIV selected inversion (pdiv_localmap.py:27-90) with ``npart`` partitions
    on one GPU: dense partition inverses and their selected-block production
    batched over partitions, coupled by a Woodbury correction.  npart must be a
    power of 2 and at most nblocks (pdiv_utils.check_multiprocessing /
    check_input).  Same instance convention and outputs as ``rgf_batched``."""
    npart = int(npart)
    N, b = _check_tridiag(diag, upper, lower)
    if npart < 1 or npart & (npart - 1):
        raise ValueError("The number of processes must be a power of 2.")
    if N < npart:
        raise ValueError("The number of blocks must be greater than the number of processes.")
    d, Bd, sd = _batched(to_dev(diag), 3)
    if N > 1:
        u, Bu, su = _batched(to_dev(upper), 3)
        l, Bl, sl = _batched(to_dev(lower), 3)
    else:
        u = l = None
        Bu = Bl = None
        su = sl = 0
    B = _resolve_batch(Bd, Bu, Bl, None)
    dev = device()
    Gd = torch.empty((B, N, b, b), dtype=_c128, device=dev)
    Gu = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev)
    Gl = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=dev)
    info = torch.empty(1, dtype=torch.int64, device=dev)
    desc = nat.PdivDesc(
        nblocks=N, blocksize=b, batch=B, npart=npart, diag=_p(d), diag_stride=sd,
        upper=_p(u), upper_stride=su, lower=_p(l), lower_stride=sl, g_diag=_p(Gd),
        g_upper=_p(Gu) if N > 1 else None, g_lower=_p(Gl) if N > 1 else None,
        info=_info_ptr(info))
    L = nat.lib()
    ws = _workspace(L.sinv_pdiv_workspace_bytes(ctypes.byref(desc)))
    nat.check(L.sinv_pdiv_batched(ctypes.byref(desc), _p(ws), ws.numel(), _stream()), "pdiv")
    if check:
        nat.check_info(info)
    return Gd, Gu, Gl


def bcr_fill_positions(N):
    """(row, col) block positions of the blocks bcr_batched(return_fill=True) returns."""
    out = []
    s = 2
    while True:
        n = N // s
        if n < 1:
            break
        rows = [m * s + s - 1 for m in range(n)]
        out += [(rows[c], rows[c + 1]) for c in range(n - 1)]
        out += [(rows[c + 1], rows[c]) for c in range(n - 1)]
        s *= 2
    return out


def rgf_energies(H_diag, H_upper, H_lower, energies, eta, **kw):
    """RGF of A(E_k) = (E_k + i*eta) I - H for every energy, H shared.

    Only H is stored; A(E_k) is formed inside the kernels' epilogues."""
    e = torch.as_tensor(np.asarray(energies), dtype=torch.complex128) + 1j * float(eta)
    return rgf_batched(H_diag, H_upper, H_lower, shift=e, dscale=-1.0, oscale=-1.0, **kw)


def dense_to_blocks(A, blocksize):
    """Stacked tridiagonal blocks (diag (B,N,b,b), upper/lower (B,N-1,b,b)) of a
    dense (n, n) or (B, n, n) complex128 matrix, gathered on the device
    (matrices_utils.py:115-151).  n must be a multiple of blocksize."""
    A = to_dev(A)
    if A.dim() == 2:
        A = A[None]
    A = A.contiguous()
    B, n = A.shape[0], A.shape[-1]
    b = int(blocksize)
    if A.shape[-2] != n or b <= 0 or n % b != 0:
        raise ValueError(f"dense_to_blocks: square matrices with n a multiple of the block size "
                         f"(n={n}, b={b})")
    N = n // b
    d = torch.empty((B, N, b, b), dtype=_c128, device=A.device)
    u = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=A.device)
    l = torch.empty((B, max(N - 1, 0), b, b), dtype=_c128, device=A.device)
    nat.check(nat.lib().sinv_dense_to_blocks(_p(A), n, n, n * n, b, B, _p(d),
                                             _p(u) if N > 1 else None, _p(l) if N > 1 else None,
                                             _stream()), "dense_to_blocks")
    return d, u, l


def blocks_to_dense(diag, upper, lower):
    """Dense (B, N b, N b) matrices from stacked tridiagonal blocks, zeros
    elsewhere, written on the device (matrices_utils.py:155-189)."""
    d, _, _ = _batched(to_dev(diag), 3)
    if d.dim() == 3:
        d = d[None]
    B, N, b = d.shape[0], d.shape[1], d.shape[-1]
    d = d.contiguous()
    if N > 1:
        u = to_dev(upper).reshape(B, N - 1, b, b).contiguous()
        l = to_dev(lower).reshape(B, N - 1, b, b).contiguous()
    else:
        u = l = None
    n = N * b
    A = torch.empty((B, n, n), dtype=_c128, device=d.device)
    nat.check(nat.lib().sinv_blocks_to_dense(_p(d), _p(u), _p(l), N, b, B, _p(A), n, n * n,
                                             _stream()), "blocks_to_dense")
    return A


def scatter_blocks(A, blocks, positions):
    """Copy blocks (B, count, b, b) into the dense (B, n, n) A at the block
    positions [(row, col), ...] (in place, on the device)."""
    blocks = to_dev(blocks)
    if blocks.dim() == 3:
        blocks = blocks[None]
    blocks = blocks.contiguous()
    B, count, b = blocks.shape[0], blocks.shape[1], blocks.shape[-1]
    if count == 0:
        return A
    pos = torch.as_tensor(np.asarray(positions, dtype=np.int32).reshape(count, 2),
                          device=A.device)
    rows = pos[:, 0].contiguous()
    cols = pos[:, 1].contiguous()
    n = A.shape[-1]
    nat.check(nat.lib().sinv_scatter_blocks(_p(blocks), count * b * b, count, _p(rows), _p(cols),
                                            n // b, b, B, _p(A), n, n * n, _stream()),
              "scatter_blocks")
    return A


def zinv_batched(A, check=True, inverse="pivoted", max_iter=0):
    """In-place-style batched inverse of (..., n, n) complex128 blocks; returns a new tensor.

    inverse="pivoted": Gauss-Jordan elimination with partial pivoting (numpy's
    semantics); "newton": the Newton-Schulz iteration for batches of at most 4
    blocks above 64 (``max_iter`` steps at most, 0 = the library's limit),
    raising NotConverged when a block misses the acceptance bound."""
    if inverse not in ("pivoted", "newton"):
        raise ValueError('inverse must be "pivoted" or "newton"')
    X = to_dev(A).clone()
    n = X.shape[-1]
    batch = int(np.prod(X.shape[:-2])) if X.dim() > 2 else 1
    info = torch.empty(1, dtype=torch.int64, device=X.device)
    L = nat.lib()
    with nat.inverse_method(inverse, max_iter):
        ws = _workspace(L.sinv_zinv_workspace_bytes(n, batch))
        nat.check(L.sinv_zinv_batched(_p(X), n, n, n * n, batch, _p(ws), ws.numel(),
                                      _info_ptr(info), _stream()), "zinv")
    if check:
        nat.check_info(info)
    return X


def zgemm_batched(A, B, *, opA="N", opB="N", alpha=1.0, beta=0.0, C0=None, shift=None,
                  precision="fp64", slices=6):
    """C = alpha op(A) op(B) + beta C0 (+ shift I) over a batch (leading dim) of blocks.

    precision "int8x" forms the complex128 product from `slices` (5..7) balanced
    base-256 int8 digit planes per operand on the int8 matrix cores
    (sinv_zgemm_int8x_batched): error <= 4.1 (S + 4) 2^(-8S) K max|row| max|col|
    per real product."""
    if opA not in nat.OPS or opB not in nat.OPS:
        raise ValueError(f"opA / opB must be one of {sorted(nat.OPS)}")
    if precision not in ("fp64", "mixed", "int8x"):
        raise ValueError("precision must be fp64, mixed or int8x")
    sa, sb = _shape(A), _shape(B)
    if len(sa) not in (2, 3) or len(sb) not in (2, 3):
        raise ValueError(f"A and B must be (M, K) / (K, N) or batched, got {sa} and {sb}")
    ba = sa[0] if len(sa) == 3 else 1
    bb = sb[0] if len(sb) == 3 else 1
    batch = max(ba, bb)
    if ba not in (1, batch) or bb not in (1, batch):
        raise ValueError(f"batch sizes of A ({ba}) and B ({bb}) must be equal or 1")
    M, K = sa[-2:] if opA == "N" else sa[-2:][::-1]
    K2, N = sb[-2:] if opB == "N" else sb[-2:][::-1]
    if K != K2:
        raise ValueError("inner dimensions differ")
    if C0 is not None:
        sc = _shape(C0)
        if len(sc) not in (2, 3) or sc[-2:] != (M, N) or (len(sc) == 3 and sc[0] not in (1, batch)):
            raise ValueError(f"C0 must be ({M}, {N}) or ({batch}, {M}, {N}), got {sc}")
    if shift is not None and int(np.prod(_shape(shift))) != batch:
        raise ValueError(f"shift needs one entry per batch entry ({batch})")
    A = to_dev(A)
    B = to_dev(B)
    if A.dim() == 2:
        A = A[None]
    if B.dim() == 2:
        B = B[None]
    ac = A.shape[-1]
    bc = B.shape[-1]
    sA = A[0].numel() if A.shape[0] > 1 else 0
    sB = B[0].numel() if B.shape[0] > 1 else 0
    C = torch.empty((batch, M, N), dtype=_c128, device=A.device)
    if C0 is not None:
        C0 = to_dev(C0)
        if C0.dim() == 2:
            C0 = C0[None]
        sC0 = C0[0].numel() if C0.shape[0] > 1 else 0
    else:
        sC0 = 0
    sh = to_dev(shift).reshape(-1) if shift is not None else None
    L = nat.lib()
    if precision == "int8x":
        ws = _workspace(L.sinv_zgemm_int8x_workspace_bytes(M, N, K, sA, sB, batch, slices))
        nat.check(L.sinv_zgemm_int8x_batched(
            M, N, K, nat.OPS[opA], nat.OPS[opB], nat.C128.of(alpha), _p(A), ac, sA, _p(B), bc,
            sB, nat.C128.of(beta), _p(C0), N, sC0, _p(C), N, M * N, _p(sh), batch, slices,
            _p(ws), ws.numel(), _stream()), "zgemm_int8x")
        return C
    nat.check(L.sinv_zgemm_batched(
        M, N, K, nat.OPS[opA], nat.OPS[opB], nat.C128.of(alpha), _p(A), ac, sA, _p(B), bc, sB,
        nat.C128.of(beta), _p(C0), N, sC0, _p(C), N, M * N, _p(sh), batch, nat.PREC[precision],
        _stream()), "zgemm")
    return C
