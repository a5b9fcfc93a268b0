"""The learned coarse-grid correction on the GPU (no counterpart in the reference).

* Forward pass and gradient of the BF16-MFMA kernels against the FP64 CPU
  oracle (oracle/cnn.py, itself pinned by finite differences) with the same
  BF16 roundings: a wrong lane layout, tap, transposition or skip path gives
  O(1) errors, BF16 rounding flips stay far below the tolerances.
* Training on pairs from the plain preconditioner lowers the loss and
  predicts held-out errors better than a zero prediction, and the ablation
  shows the stage cutting the preconditioner's error on fresh samples.
* The learned flexible GMRES stage converges to the same tolerance as the
  plain FP32-MG form (the solver's stopping test on the true residual), and
  falls back to the plain V-cycle when the network produces non-finite
  values.
"""
import numpy as np
import pytest

from lowmachmixtures_amd import problems as P
from lowmachmixtures_amd import stokes as S
from oracle import cnn as N

pytestmark = pytest.mark.gpu

SHAPES = N.layer_shapes(8, 16, 32, 64, 4)


def _inputs(nb, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nb,) + n + (8,)).astype(np.float32)
    t = (0.5 * rng.standard_normal((nb,) + n + (4,))).astype(np.float32)
    return x, t


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("nb,n", [(2, (8, 8, 8)), (1, (16, 8, 8))])
def test_forward_matches_oracle(nb, n):
    x, _ = _inputs(nb, n, 1)
    with S.Context(3, (16, 16, 16), 1.0) as ctx, S.Cnn(ctx, seed=3) as net:
        p = net.get_params().astype(np.float64)
        assert p.size == N.num_params(SHAPES)
        y = net.forward(x)
    want, _ = N.forward(p, x.astype(np.float64), SHAPES, bf=True)
    err = _rel(y, want)
    print(f"forward rel err {err:.2e}", flush=True)
    assert err < 1e-2


def test_gradient_matches_oracle():
    # 16^3 keeps at least 128 voxels on the coarsest level, so that a ReLU
    # mask flipped by a BF16 rounding difference moves no layer's sum by much
    nb, n = 2, (16, 16, 16)
    x, t = _inputs(nb, n, 2)
    with S.Context(3, (16, 16, 16), 1.0) as ctx, S.Cnn(ctx, seed=4) as net:
        p = net.get_params().astype(np.float64)
        loss, g = net.loss_grad(x, t)
    wl, wg = N.loss_and_grad(p, x.astype(np.float64), t.astype(np.float64), SHAPES, bf=True)
    print(f"loss {loss:.6e} oracle {wl:.6e}; grad rel err {_rel(g, wg):.2e}", flush=True)
    assert abs(loss - wl) <= 1e-2 * wl
    assert _rel(g, wg) < 2e-2
    # every layer, including the first (input) and last (linear) ones
    q, errs = 0, []
    for ci, co in SHAPES:
        m = 27 * ci * co
        errs.append(_rel(g[q:q + m], wg[q:q + m]))
        q += m
    print("per-layer grad rel err " + " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) < 5e-2


def test_set_params_round_trip():
    with S.Context(3, (16, 16, 16), 1.0) as ctx, S.Cnn(ctx, seed=5) as net:
        p = net.get_params()
        p2 = (0.5 * p).astype(np.float32)
        net.set_params(p2)
        assert np.array_equal(net.get_params(), p2)


def _problem(n, seed, contrast=100.0):
    return P.small_problem(3, n, visc_type=-1, seed=seed, contrast=contrast)


def test_training_lowers_loss_and_predicts_held_out_errors():
    n = (64, 64, 64)
    dx = _problem(n, 11).dx
    with S.Context(3, n, dx, visc_type=-1, precision=S.FP32_MG) as ctx, S.Cnn(ctx, seed=6, level=2) as net:
        for k, seed in enumerate((11, 12)):
            pb = _problem(n, seed)
            dp = S.DeviceProblem(ctx, pb, rhs=[np.zeros(pb.shape)] * 3)
            net.generate_pairs(dp, 96, seed=100 + k)
        held = 32
        total = net.num_pairs()
        assert total == 192
        before = net.evaluate(total - held, total)
        hist = net.train(steps=300, batch=16, lr=1e-3, seed=7, held_out=held)
        after = net.evaluate(total - held, total)
        seen = net.evaluate(0, held)
    first, last = float(np.mean(hist[:20])), float(np.mean(hist[-20:]))
    print(f"loss {first:.4e} -> {last:.4e}; held-out rel err {before:.3f} -> {after:.3f}; "
          f"training pairs {seen:.3f}", flush=True)
    assert np.all(np.isfinite(hist))
    assert last < 0.7 * first
    assert after < min(before, 0.95)


def _solve(ctx, pb, b_u):
    dp = S.DeviceProblem(ctx, pb, rhs=b_u)
    _, info = dp.solve()
    r_u, r_p = ctx.faces(), ctx.field(S.CELL)
    S.apply_matrix(ctx, r_u, r_p, dp.x_u, dp.x_p, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, pb.theta_alpha)
    Ax = [f.download() for f in r_u] + [r_p.download()]
    res = np.sqrt(sum(((a - b) ** 2).sum() for a, b in zip(Ax, b_u + [pb.b_p])))
    bn = np.sqrt(sum((b ** 2).sum() for b in b_u))
    return info, res / bn, dp.solution()


def test_learned_stage_gains_over_the_plain_preconditioner():
    """The ablation: a network trained on the preconditioner's slow error modes
    (64^3, contrast 100, correction on level 1) cuts the preconditioner's
    error on fresh samples and the flexible GMRES's iteration count, and the
    learned form reaches the same tolerance as the plain FP32-MG form with the
    same solution."""
    n = (64, 64, 64)
    pbs = [_problem(n, s) for s in (31, 32, 34)]
    test = _problem(n, 33)
    b_u = P.rhs_from_noise_host(test)
    with S.Context(3, n, test.dx, visc_type=-1, precision=S.FP32_MG) as ctx, \
            S.Cnn(ctx, seed=1, level=1, sample_steps=3) as net:
        for k, pb in enumerate(pbs):
            dp = S.DeviceProblem(ctx, pb, rhs=[np.zeros(pb.shape)] * 3)
            net.generate_pairs(dp, 128, seed=100 + k)
        net.train(steps=400, batch=8, lr=1e-3, seed=7)
        dpt = S.DeviceProblem(ctx, test, rhs=b_u)
        e_plain, e_learned = net.error_reduction(dpt, 6, seed=999)
        info_mg, res_mg, (u_mg, p_mg) = _solve(ctx, test, b_u)
        net.attach()
        ctx.set_params(precision=S.LEARNED)
        info_l, res_l, (u_l, p_l) = _solve(ctx, test, b_u)
    print(f"ablation: preconditioner error {e_plain:.4f} plain, {e_learned:.4f} learned; FGMRES iters "
          f"{info_mg['total_iter']} plain, {info_l['total_iter']} learned (applied {info_l['learned_applied']}, "
          f"fallback {info_l['learned_fallback']})", flush=True)
    assert e_learned < 0.85 * e_plain
    assert info_l["total_iter"] < info_mg["total_iter"]
    assert info_mg["status"] == 0 and info_l["status"] == 0
    assert res_mg < 2e-9 and res_l < 2e-9
    assert info_l["learned_applied"] >= 1
    err = np.linalg.norm(np.concatenate([(a - b).ravel() for a, b in zip(u_l, u_mg)])) / \
        np.linalg.norm(np.concatenate([b.ravel() for b in u_mg]))
    assert err < 1e-7


def test_learned_stage_falls_back_on_non_finite_output():
    n = (64, 64, 64)
    pb = _problem(n, 22)
    b_u = P.rhs_from_noise_host(pb)
    with S.Context(3, n, pb.dx, visc_type=-1, precision=S.FP32_MG) as ctx, S.Cnn(ctx, seed=10, level=2) as net:
        net.set_params(np.full(net.num_params(), 3.0e37, dtype=np.float32))
        info_mg, res_mg, _ = _solve(ctx, pb, b_u)
        net.attach()
        ctx.set_params(precision=S.LEARNED)
        info_l, res_l, _ = _solve(ctx, pb, b_u)
    print(f"fallback: iters {info_l['total_iter']} (plain {info_mg['total_iter']}) res {res_l:.2e}", flush=True)
    assert info_l["status"] == 0 and res_l < 2e-9
    assert info_l["learned_fallback"] == 1
    assert info_l["learned_applied"] == 1
    assert info_l["total_iter"] == info_mg["total_iter"]


def test_learned_needs_a_network():
    pb = _problem((16, 16, 16), 23)
    with S.Context(3, (16, 16, 16), pb.dx, visc_type=-1, precision=S.LEARNED) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=P.rhs_from_noise_host(pb))
        with pytest.raises(S.LmmError, match="network"):
            dp.solve()
