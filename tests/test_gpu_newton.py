"""GPU parity of the Newton-Schulz block inverse (the latency-bound method for
column diagonally dominant blocks) and of RGF run with it, against
numpy.linalg.inv (the reference's call, rgf.py:44,48) and the RGF oracle.

Tolerances: the iteration accepts an instance once max |I - M X_k| <= 1e-9,
which bounds the next iterate's residual by (n 1e-9)^2 in norm, so its
relative error is at most (n 1e-9)^2 + rounding (2.6e-13 at n = 512); the
tests hold the inverse to 1e-12 relative (Frobenius) and RGF to the same
1e-12 it is held to with the pivoted inverse.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def _dominant(rng, n, shift):
    A = _rand(rng, n, n)
    return A + shift * n * np.eye(n)


def _rel(X, ref):
    return np.linalg.norm(X - ref) / np.linalg.norm(ref)


@pytest.mark.parametrize("n,batch", [(65, 1), (128, 3), (200, 2), (384, 1), (512, 4)])
def test_zinv_newton_dominant(built_lib, n, batch):
    from sinv_amd import core
    print(f"zinv newton n={n} batch={batch}", flush=True)
    rng = np.random.default_rng(1000 + n)
    # column ratios from ~0.2 to ~0.7 across the batch
    A = np.stack([_dominant(rng, n, s) for s in (3.0, 1.5, 1.2, 2.0)[:batch]])
    X = core.zinv_batched(A, inverse="newton").cpu().numpy()
    for z in range(batch):
        assert _rel(X[z], np.linalg.inv(A[z])) < 1e-12, (n, z)


def test_zinv_newton_steps_of_both_parities(built_lib):
    """Instances accepted at odd steps hold their result in the second buffer
    and are copied back; a batch mixing fast and slow convergence checks both
    paths and the per-instance early exit."""
    from sinv_amd import core
    print("zinv newton mixed convergence", flush=True)
    rng = np.random.default_rng(7)
    n = 256
    # dominance ratio ~0.01 (accepted at step 1) .. ~0.9 (step 4 or 5)
    A = np.stack([_dominant(rng, n, s) for s in (80.0, 8.0, 1.6, 0.95)])
    X = core.zinv_batched(A, inverse="newton", max_iter=8).cpu().numpy()
    for z in range(4):
        assert _rel(X[z], np.linalg.inv(A[z])) < 1e-12, z


def test_zinv_newton_reports_non_convergence(built_lib):
    """A permutation (zero diagonal) and a non-dominant block do not converge:
    NotConverged names the first instance; the pivoted elimination inverts both."""
    from sinv_amd import _native as nat, core
    print("zinv newton non-convergence", flush=True)
    rng = np.random.default_rng(3)
    n = 128
    P = np.roll(np.eye(n), 1, axis=0).astype(np.complex128)
    R = _rand(rng, n, n)
    A = np.stack([_dominant(rng, n, 3.0), P, R])
    with pytest.raises(nat.NotConverged) as e:
        core.zinv_batched(A, inverse="newton", max_iter=3)
    assert e.value.instance == 1
    X = core.zinv_batched(A, inverse="pivoted").cpu().numpy()
    assert np.array_equal(X[1], P.T)
    assert _rel(X[2], np.linalg.inv(R)) < 1e-15 * np.linalg.cond(R) * n


def test_zinv_newton_batch_above_limit_uses_elimination(built_lib):
    """More than 4 blocks per launch: the elimination runs (numpy's pivoting),
    so a permutation still inverts exactly."""
    from sinv_amd import core
    print("zinv newton batch 6 -> elimination", flush=True)
    n = 96
    P = np.roll(np.eye(n), 1, axis=0).astype(np.complex128)
    X = core.zinv_batched(np.stack([P] * 6), inverse="newton").cpu().numpy()
    for z in range(6):
        assert np.array_equal(X[z], P.T)


def test_column_dominance_matches_numpy(built_lib):
    from sinv_amd import core
    from sinv_amd.utils import configs
    print("column dominance", flush=True)
    d, u, l = configs.dominant_blocks(5, 96, seed=4)
    A = np.zeros((5 * 96, 5 * 96), complex)
    for i in range(5):
        A[i * 96:(i + 1) * 96, i * 96:(i + 1) * 96] = d[i]
        if i < 4:
            A[i * 96:(i + 1) * 96, (i + 1) * 96:(i + 2) * 96] = u[i]
            A[(i + 1) * 96:(i + 2) * 96, i * 96:(i + 1) * 96] = l[i]

    def ratio(M):
        a = np.abs(M)
        dg = np.diag(a)
        return ((a.sum(0) - dg) / dg).max()

    rho = core.column_dominance(d, u, l).cpu().numpy()
    assert abs(rho[0] - ratio(A)) < 1e-12
    # instances through shift / dscale / oscale: A = 2 D - 0.5 I, couplings x (1 + 1j)
    sh = np.array([-0.5, 40.0 + 3j])
    rho = core.column_dominance(d, u, l, shift=sh, dscale=2.0, oscale=1 + 1j).cpu().numpy()
    for z in range(2):
        Az = A.copy()
        for i in range(5):
            s = slice(i * 96, (i + 1) * 96)
            Az[s, s] = 2.0 * d[i] + sh[z] * np.eye(96)
            if i < 4:
                t = slice((i + 1) * 96, (i + 2) * 96)
                Az[s, t] = (1 + 1j) * u[i]
                Az[t, s] = (1 + 1j) * l[i]
        assert abs(rho[z] - ratio(Az)) < 1e-12, z


@pytest.mark.parametrize("N,b,batch", [(6, 128, 1), (5, 512, 1), (4, 200, 3)])
def test_rgf_newton_against_oracle(built_lib, N, b, batch):
    from oracle import dense, rgf_ref
    from sinv_amd import core
    from sinv_amd.utils import configs
    print(f"rgf newton N={N} b={b} batch={batch}", flush=True)
    d, u, l = configs.dominant_blocks(N, b, seed=N + b)
    sh = np.array([0.0, 5.0 + 2j, -3.0j])[:batch]
    Gn = core.rgf_batched(d, u, l, shift=sh, inverse="newton")
    Gp = core.rgf_batched(d, u, l, shift=sh, inverse="pivoted")
    for z in range(batch):
        dz = d + sh[z] * np.eye(b)
        ref = rgf_ref.rgf_blocks(dz, u, l)
        got = tuple(x[z].cpu().numpy() for x in Gn)
        piv = tuple(x[z].cpu().numpy() for x in Gp)
        assert dense.rel_err(got, ref) < 1e-12, z
        assert dense.rel_err(got, piv) < 1e-12, z


def test_rgf_auto_chooses_by_dominance(built_lib):
    """"auto" takes the iteration for a dominant single instance (its launches
    appear in the launch profiler) and the elimination for A(E) near the band."""
    from sinv_amd import _native as nat, core
    from sinv_amd.utils import configs
    print("rgf auto method choice", flush=True)
    d, u, l = configs.dominant_blocks(4, 160, seed=9)
    with nat.LaunchProfiler() as p:
        core.rgf_batched(d, u, l)
    assert p.stats["inv_newton"][1] > 0 and p.stats["inv_panel"][1] == 0
    H = configs.hermitian_H(4, 160, seed=9)
    with nat.LaunchProfiler() as p:
        core.rgf_batched(*H, shift=np.array([0.3 + 1e-3j]), dscale=-1.0, oscale=-1.0)
    assert p.stats["inv_newton"][1] == 0 and p.stats["inv_panel"][1] > 0


def test_rgf_auto_falls_back_when_iteration_misses(built_lib, monkeypatch):
    """With too few steps allowed the iteration misses its bound: an explicit
    "newton" plan reports NotConverged, and "auto" repeats the call with the
    elimination and returns the right blocks."""
    from oracle import dense, rgf_ref
    from sinv_amd import _native as nat, core
    from sinv_amd.utils import configs
    print("rgf auto fallback", flush=True)
    d, u, l = configs.dominant_blocks(5, 128, seed=12)
    plan = core.RgfPlan(d, u, l, inverse="newton")
    plan.inverse = ("newton", 1)
    plan.run()
    with pytest.raises(nat.NotConverged):
        plan.check()
    monkeypatch.setattr(core, "newton_iterations", lambda rho: 1)
    G = core.rgf_batched(d, u, l)
    got = tuple(x[0].cpu().numpy() for x in G)
    assert dense.rel_err(got, rgf_ref.rgf_blocks(d, u, l)) < 1e-12


def test_rgf_newton_plan_graph_replays_match_eager(built_lib):
    """The iteration's state (accepted step, reduction counters) is reset by a
    kernel at every run, so captured replays repeat the eager bits."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    print("rgf newton plan graph replay", flush=True)
    d, u, l = configs.dominant_blocks(6, 256, seed=21)
    plan = core.RgfPlan(d, u, l, inverse="auto")
    assert plan.inverse[0] == "newton"
    plan.run()
    plan.check()
    eager = [x.clone() for x in plan.outputs()]
    g = plan.capture()
    for _ in range(3):
        for x in plan.outputs():
            x.zero_()
        g.replay()
        torch.cuda.synchronize()
        plan.check()
        for x, y in zip(plan.outputs(), eager):
            assert torch.equal(x, y)


@pytest.mark.parametrize("method", ["bcr", "bcr_p3", "psr_p4", "rgf2sided"])
def test_other_methods_with_newton(built_lib, method):
    """BCR (serial and per-partition launches), PSR and RGF-2sided with the
    iteration serving their launches of at most 4 blocks: the same selected
    blocks as with the elimination (and as the RGF oracle) to 1e-12."""
    from oracle import dense, rgf_ref
    from sinv_amd import _native as nat, core
    from sinv_amd.utils import configs
    print(f"{method} newton", flush=True)
    N, b = 14, 160
    d, u, l = configs.dominant_blocks(N, b, seed=31)
    run = {
        "bcr": lambda inv: core.bcr_batched(d, u, l, inverse=inv),
        "bcr_p3": lambda inv: core.bcr_batched(d, u, l, npart=3, inverse=inv),
        "psr_p4": lambda inv: core.psr_batched(d, u, l, 4, inverse=inv),
        "rgf2sided": lambda inv: core.psr_batched(d, u, l, 2, first_rows=N // 2, inverse=inv),
    }[method]
    with nat.LaunchProfiler() as p:
        Gn = run("auto")
    assert p.stats["inv_newton"][1] > 0
    Gp = run("pivoted")
    ref = rgf_ref.rgf_blocks(d, u, l)
    got = tuple(x[0].cpu().numpy() for x in Gn)
    piv = tuple(x[0].cpu().numpy() for x in Gp)
    assert dense.rel_err(got, piv) < 1e-12
    assert dense.rel_err(got, ref) < 1e-12
