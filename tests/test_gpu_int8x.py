"""GPU tests of the sliced-int8 complex GEMM (sinv_zgemm_int8x_batched) and of
RGF with precision="int8x".

The product is checked against numpy's complex128 matmul with the error bound
the digit splitting gives: per term, the dropped digit pairs (S - 1 of size
2^(-8S) at most, then geometrically smaller) and the two remainders beyond S
balanced base-256 digits give 4.1 (S + 4) 2^(-8S) max|row of op(A)|
max|column of op(B)| (the scales are below 2 (128/127) times the maxima); K
terms, and the 3M complex form adds at most three real products per part.  RGF in int8x mode is checked against the FP64 path and the
oracle at the 1e-8 gate.
"""

import numpy as np
import pytest

from oracle import dense, rgf_ref

pytestmark = pytest.mark.gpu


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _op(X, op):
    return X if op == "N" else (X.T if op == "T" else X.conj().T)


def _bound(A, B, opA, opB, S):
    """Elementwise bound (M x N) of |int8x - exact| for one batch entry."""
    a = np.abs(_op(A, opA))
    b = np.abs(_op(B, opB))
    # the three real parts of a row are bounded by 2 max|x|
    ra = 2 * a.max(axis=1)[:, None]
    cb = 2 * b.max(axis=0)[None, :]
    K = a.shape[1]
    return 3 * (4.1 * (S + 4) * 2.0 ** (-8 * S) + 1e-15) * K * ra * cb


# K <= 512: single-pass digit split; K > 512: row-maximum pass + digit pass
@pytest.mark.parametrize("M,N,K", [(64, 64, 64), (100, 70, 130), (1, 5, 3), (256, 128, 384),
                                   (416, 256, 416), (33, 200, 65), (64, 96, 512), (40, 48, 700)])
@pytest.mark.parametrize("ops", [("N", "N"), ("T", "N"), ("N", "C"), ("C", "T")])
def test_zgemm_int8x_against_fp64(built_lib, M, N, K, ops):
    import torch
    from sinv_amd import core
    opA, opB = ops
    print(f"int8x gemm M={M} N={N} K={K} ops={ops}", flush=True)
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    batch = 3
    A = _rand(rng, (batch,) + ((M, K) if opA == "N" else (K, M)))
    B = _rand(rng, (1,) + ((K, N) if opB == "N" else (N, K)))  # shared by the batch
    C0 = _rand(rng, (batch, M, N))
    shift = _rand(rng, (batch,))
    alpha, beta = 0.7 - 0.2j, -1.3 + 0.4j
    C = core.zgemm_batched(torch.from_numpy(A), torch.from_numpy(B), opA=opA, opB=opB,
                           alpha=alpha, beta=beta, C0=torch.from_numpy(C0),
                           shift=torch.from_numpy(shift), precision="int8x").cpu().numpy()
    for z in range(batch):
        ref = alpha * (_op(A[z], opA) @ _op(B[0], opB)) + beta * C0[z]
        ref[np.arange(min(M, N)), np.arange(min(M, N))] += shift[z]
        err = np.abs(C[z] - ref)
        bound = abs(alpha) * _bound(A[z], B[0], opA, opB, 6) + 1e-14 * np.abs(ref)
        assert np.all(err <= bound), (z, float((err / bound).max()))
        # observed scale (the elementwise bound above is the guarantee): the
        # row maxima, and with them the digits' quantum, grow slowly with K
        rel = np.linalg.norm(C[z] - ref) / np.linalg.norm(ref)
        assert rel < 2e-13, rel


@pytest.mark.parametrize("S", [5, 6, 7])
def test_zgemm_int8x_digits_set_the_error(built_lib, S):
    """Each extra digit plane cuts the error by about 2^8, within the bound."""
    import torch
    from sinv_amd import core
    rng = np.random.default_rng(11)
    M = N = K = 192
    A = _rand(rng, (2, M, K))
    B = _rand(rng, (2, K, N))
    C = core.zgemm_batched(torch.from_numpy(A), torch.from_numpy(B), precision="int8x",
                           slices=S).cpu().numpy()
    ref = A @ B
    for z in range(2):
        err = np.abs(C[z] - ref[z])
        assert np.all(err <= _bound(A[z], B[z], "N", "N", S))
    rel = np.linalg.norm(C - ref) / np.linalg.norm(ref)
    # observed scale of the error: 2^(-8S) times a modest factor
    assert rel < 2.0 ** (-8 * S) * 256, (S, rel)


def test_zgemm_int8x_row_scales(built_lib):
    """Rows and columns spanning 2^-40 .. 2^40 keep their relative accuracy
    (each gets its own power-of-two scale); zero rows give zero."""
    import torch
    from sinv_amd import core
    rng = np.random.default_rng(5)
    M, N, K = 96, 80, 112
    A = _rand(rng, (M, K)) * (2.0 ** rng.integers(-40, 40, size=(M, 1)))
    B = _rand(rng, (K, N)) * (2.0 ** rng.integers(-40, 40, size=(1, N)))
    A[7] = 0
    B[:, 3] = 0
    C = core.zgemm_batched(torch.from_numpy(A), torch.from_numpy(B),
                           precision="int8x").cpu().numpy()[0]
    ref = A @ B
    assert np.all(C[7] == 0) and np.all(C[:, 3] == 0)
    assert np.all(np.abs(C - ref) <= _bound(A, B, "N", "N", 6))
    rel_rows = np.linalg.norm(C - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
    assert rel_rows.max() < 1e-12


def test_zgemm_int8x_rejects_bad_arguments(built_lib):
    import torch
    from sinv_amd import core, _native as nat
    A = torch.zeros((1, 4, 4), dtype=torch.complex128)
    with pytest.raises(ValueError):
        core.zgemm_batched(A, A, precision="int8x", slices=4)
    with pytest.raises(ValueError):
        core.zgemm_batched(A, A, precision="int8x", slices=8)
    # the FP64/complex64 entry point does not take the int8x code (it needs a workspace)
    L = nat.lib()
    rc = L.sinv_zgemm_batched(4, 4, 4, 0, 0, nat.C128.of(1), None, 4, 0, None, 4, 0,
                              nat.C128.of(0), None, 4, 0, None, 4, 0, None, 1, 2, None)
    assert rc == nat.SINV_E_INVAL


def test_rgf_int8x_energies_meet_the_gate(built_lib):
    """Batched energies near the band (eta = 1e-3): int8x RGF within 1e-8 of
    the FP64 path and of the oracle, on every instance."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(12, 128, seed=3)
    E = configs.energies(16, seed=4)
    eta = 1e-3
    print("rgf int8x energies N=12 b=128 E=16", flush=True)
    G64 = core.rgf_energies(*H, E, eta)
    G8 = core.rgf_energies(*H, E, eta, precision="int8x")
    worst = 0.0
    for k in range(len(E)):
        a = tuple(x[k].cpu().numpy() for x in G8)
        b = tuple(x[k].cpu().numpy() for x in G64)
        worst = max(worst, dense.rel_err(a, b))
    assert worst < 1e-8, worst
    for k in (0, 7):
        d, u, l = configs.shifted(H, E[k] + 1j * eta)
        ref = rgf_ref.rgf_blocks(d, u, l)
        got = tuple(x[k].cpu().numpy() for x in G8)
        assert dense.rel_err(got, ref) < 1e-8


def test_rgf_int8x_symmetric_and_plane_reuse(built_lib):
    """The G_ii product reuses the digit planes of Q = g_i A_{i,i+1} from the
    G_{i,i+1} product just before it; with sym_mat there is no such product
    and Q is split for G_ii itself.  Both give FP64's blocks to 1e-10 (a
    symmetric instance: A = A^T), and the launch profiler's split bytes equal
    the count with exactly one Q split per backward step either way."""
    from sinv_amd import core, _native as nat
    N, b, nb = 5, 128, 2
    rng = np.random.default_rng(17)
    d = rng.standard_normal((N, b, b)) + 1j * rng.standard_normal((N, b, b))
    d = d + d.transpose(0, 2, 1) + 40 * np.eye(b)
    u = rng.standard_normal((N - 1, b, b)) + 1j * rng.standard_normal((N - 1, b, b))
    l = u.transpose(0, 2, 1).copy()
    sh = np.array([0.5 + 1e-2j, -1.0 + 1e-2j])
    print("rgf int8x sym / reuse N=5 b=128", flush=True)
    split = {}
    for sym in (False, True):
        G64 = core.rgf_batched(d, u, l, shift=sh, sym_mat=sym)
        with nat.LaunchProfiler() as p:
            G8 = core.rgf_batched(d, u, l, shift=sh, sym_mat=sym, precision="int8x")
        split[sym] = p.bytes["i8x_split"]
        for x, y in zip(G8, G64):
            x, y = x.cpu().numpy(), y.cpu().numpy()
            assert np.linalg.norm(x - y) <= 1e-10 * np.linalg.norm(y), sym
    # one split of a distinct b x b copy: 16 B read per element, 18 B of digits
    # written per element, 24 B of scales per row
    per_copy = 16 * b * b + 18 * b * b + 24 * b
    fwd = (N - 1) * ((1 + nb) + (nb + 1))          # L g (L shared), T U (U shared)
    bwd_saved = (N - 1) * (2 * nb + (nb + 1) + 2 * nb + nb)   # Xn, Q, G_{i,i+1}, G_ii (Q reused)
    bwd_sym = (N - 1) * (2 * nb + (nb + 1) + 2 * nb)          # Xn, Q, G_ii (G_{i,i+1} = G_{i+1,i}^T)
    assert split[False] == per_copy * (fwd + bwd_saved)
    assert split[True] == per_copy * (fwd + bwd_sym)


def test_rgf_int8x_blocks_above_the_single_pass_split(built_lib):
    """Blocks of 576 (K > 512: the row-maximum pass + digit pass split) and a
    ragged neighbour of 448 (single pass) in one RGF solve: int8x within 1e-10
    of FP64 on a well-conditioned instance."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    print("rgf int8x b=576/448", flush=True)
    bs = [576, 448, 576]
    H = configs.hermitian_H_ragged(bs, seed=41)
    d, u, l = configs.shifted_lists(H, 0.2 + 5e-2j)
    G64 = core.rgf_ragged(d, u, l)
    G8 = core.rgf_ragged(d, u, l, precision="int8x")
    for x, y in zip(G8, G64):
        for a, b in zip(x, y):
            a = a.cpu().numpy() if hasattr(a, "cpu") else a
            b = b.cpu().numpy() if hasattr(b, "cpu") else b
            assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b)


def test_rgf_int8x_ragged_and_unsaved(built_lib):
    """Per-block sizes (not multiples of 64) and save_off_diag=False in int8x mode."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    bs = [40, 96, 72, 130, 64, 17]
    H = configs.hermitian_H_ragged(bs, seed=9)
    d, u, l = configs.shifted_lists(H, 0.3 + 1e-2j)
    G64 = core.rgf_ragged(d, u, l)
    G8 = core.rgf_ragged(d, u, l, precision="int8x")
    for x, y in zip(G8, G64):
        for a, b in zip(x, y):
            a = a.cpu().numpy() if hasattr(a, "cpu") else a
            b = b.cpu().numpy() if hasattr(b, "cpu") else b
            assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b)
    dd, uu, ll = configs.dominant_blocks(6, 80, seed=2)
    Gd, Gu, Gl = core.rgf_batched(dd, uu, ll, save_off_diag=False, precision="int8x")
    ref = rgf_ref.rgf_blocks(dd, uu, ll)
    assert np.linalg.norm(Gd[0].cpu().numpy() - ref[0]) <= 1e-12 * np.linalg.norm(ref[0])
    assert np.isnan(Gu.cpu().numpy()).all() and np.isnan(Gl.cpu().numpy()).all()


def test_zgemm_int8x_product_variants_identical(built_lib):
    """Every product kernel (two LDS buffers, one workgroup per CU; one LDS
    buffer with padded rows, two workgroups per CU; the same with swizzled
    rows, the default) gives the same bits, for each digit-plane count."""
    import torch
    from sinv_amd import core, _native as nat
    rng = np.random.default_rng(21)
    L = nat.lib()
    for M, K, N, S in ((200, 150, 130, 6), (33, 700, 96, 6), (384, 384, 384, 6), (100, 96, 70, 5),
                       (100, 96, 70, 7)):
        print(f"int8x variants M={M} K={K} N={N} S={S}", flush=True)
        A = _rand(rng, (3, M, K))
        B = _rand(rng, (1, K, N))
        outs = []
        try:
            for v in (0, 1, 2):
                nat.check(L.sinv_int8x_force_variant(v), "force_variant")
                outs.append(core.zgemm_batched(torch.from_numpy(A), torch.from_numpy(B),
                                               precision="int8x", slices=S).cpu().numpy())
        finally:
            nat.check(L.sinv_int8x_force_variant(-1), "force_variant")
        assert np.array_equal(outs[0], outs[1])
        assert np.array_equal(outs[0], outs[2])
    assert L.sinv_int8x_force_variant(3) == nat.SINV_E_INVAL


def test_adjoint_int8x_against_dense_formula_and_fp64(built_lib):
    """The adjoint sweep with int8x GEMMs: within 1e-10 of the dense formula on
    a dominant case, and within 1e-8 of the FP64 adjoint on energies near the
    band (eta = 1e-3), ragged blocks, checkpointed."""
    from oracle import adjoint_ref
    from sinv_amd import core
    from sinv_amd.utils import configs
    rng = np.random.default_rng(17)
    N, b = 6, 80
    d, u, l = configs.dominant_blocks(N, b, seed=3)
    c = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    gd, gu, gl = c(N, b, b), c(N - 1, b, b), c(N - 1, b, b)
    got = [x[0].cpu().numpy() for x in
           core.rgf_adjoint_batched(d, u, l, gd, gu, gl, precision="int8x")]
    ref = adjoint_ref.grad_dense(d, u, l, gd, gu, gl)
    num = sum(np.linalg.norm(a - r) ** 2 for a, r in zip(got, ref))
    den = sum(np.linalg.norm(r) ** 2 for r in ref)
    assert np.sqrt(num / den) < 1e-10

    bs = [72, 130, 96, 64, 110, 80, 72]
    H = configs.hermitian_H_ragged(bs, seed=6)
    E = configs.energies(6, seed=8) + 1e-3j
    gdl = [(-1j / np.pi) * np.eye(x) for x in bs]
    zu = [np.zeros((bs[i], bs[i + 1]), complex) for i in range(len(bs) - 1)]
    zl = [np.zeros((bs[i + 1], bs[i]), complex) for i in range(len(bs) - 1)]
    outs = {}
    for prec in ("fp64", "int8x"):
        outs[prec] = core.rgf_adjoint_ragged(*H, gdl, zu, zl, shift=E, dscale=-1.0, oscale=-1.0,
                                             precision=prec, checkpoint_every=3)
    num = den = 0.0
    for xs, ys in zip(outs["int8x"], outs["fp64"]):
        for a, r in zip(xs, ys):
            num += float(((a - r).abs() ** 2).sum())
            den += float((r.abs() ** 2).sum())
    assert np.sqrt(num / den) < 1e-8, np.sqrt(num / den)


def test_rgf_int8x_graph_replays_match_eager(built_lib):
    """An int8x RGF plan captured into a HIP graph (digit splits and products
    from the plan's workspace, no allocation, no memset): every replay, with
    eager int8x work in between, reproduces the eager result bit for bit."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(7, 96, seed=14)
    E = configs.energies(5, seed=15) + 1e-3j
    plan = core.RgfPlan(*H, shift=E, dscale=-1.0, oscale=-1.0, precision="int8x")
    plan.run()
    torch.cuda.synchronize()
    ref = tuple(x.clone() for x in plan.outputs())
    g = plan.capture()
    for rep in range(3):
        for x in plan.outputs():
            x.zero_()
        other = core.rgf_batched(*H, shift=E[:2], dscale=-1.0, oscale=-1.0, precision="int8x")
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, r in zip(plan.outputs(), ref):
            assert torch.equal(a, r), rep
        plan.check()
        del other


def test_autograd_dos_gradient_int8x(built_lib):
    """dL/dH of the DOS loss through the autograd Function with int8x GEMMs,
    several energies near the band (eta = 1e-3): loss and gradient within the
    1e-8 gate of the FP64 path."""
    import torch
    from sinv_amd import autograd as sag
    from sinv_amd.utils import configs
    H = configs.hermitian_H(10, 96, seed=21)
    E = configs.energies(6, seed=22)
    dev = torch.device("cuda")
    out = {}
    for prec in ("fp64", "int8x"):
        Hd, Hu, Hl = (torch.tensor(x, device=dev, requires_grad=True) for x in H)
        Gd, Gu, Gl = sag.greens_function(Hd, Hu, Hl, E, 1e-3, precision=prec, checkpoint_every=4)
        L = sag.dos_loss(Gd)
        L.backward()
        out[prec] = (L.item(), [x.grad.detach() for x in (Hd, Hu, Hl)])
    (l64, g64), (l8, g8) = out["fp64"], out["int8x"]
    assert abs(l8 - l64) <= 1e-8 * abs(l64)
    num = sum(float(((a - b).abs() ** 2).sum()) for a, b in zip(g8, g64))
    den = sum(float((b.abs() ** 2).sum()) for b in g64)
    assert np.sqrt(num / den) < 1e-8, np.sqrt(num / den)


def test_planset_int8x_concurrent_matches_sequential(built_lib):
    """int8x plans of different block sizes run concurrently on their own
    streams (each with its own digit-plane workspace) give the same bits as
    the plans run one after another."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    plans = []
    for gi, b in enumerate((40, 64, 100)):
        H = configs.hermitian_H(6, b, seed=60 + gi)
        E = configs.energies(4, seed=70 + gi) + 1e-3j
        p = core.RgfPlan(*H, shift=E, dscale=-1.0, oscale=-1.0, precision="int8x")
        p.run()
        plans.append(p)
    torch.cuda.synchronize()
    refs = [tuple(x.clone() for x in p.outputs()) for p in plans]
    for p in plans:
        for x in p.outputs():
            x.zero_()
    core.PlanSet(plans).run()
    torch.cuda.synchronize()
    for p, ref in zip(plans, refs):
        for a, r in zip(p.outputs(), ref):
            assert torch.equal(a, r)
        p.check()
