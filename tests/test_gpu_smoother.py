"""The fused red-black smoother (visc_type -1 and 1; 3D and 2D) on levels
whose extents allow it (x a multiple of 64, y of 8), against the oracle and against
the colour-by-colour kernels.  Both must agree bit for bit: every point is
computed with the reference's operation order.  Two forms of the fused sweep
exist (LMM_FUSED_KERNEL=march2, the default, block and march).
LMM_SMOOTHER=fused makes the small test levels use the fused sweep, which by
default only launches of at least 64 workgroups (32^3 and up) do; the
plane-marching forms choose their z-slab thickness for LMM_FUSED_TARGET
workgroups (768 by default: 2-plane slabs on these small levels, 1 for the
thickest slab a level allows)."""
import numpy as np
import pytest

from lowmachmixtures_amd import problems as P
from lowmachmixtures_amd import stokes as S
from oracle import oracle as O
from tests.helpers import bit_equal, oracle_args, relerr

pytestmark = pytest.mark.gpu

N3 = (64, 16, 32)


def _rand_faces(shape, seed, ncomp=3):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) for _ in range(ncomp)]


def _mg_solve(pb, kw, rhs, guess):
    with S.Context(pb.dim, pb.n, pb.dx, **kw) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=rhs)
        phi = ctx.faces(guess)
        ctx.profiling(True)
        S.stag_mg_solver(ctx, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, 0.9, phi, dp.b_u)
        fused = ctx.kernel_stats("stag_rb_fused")["launches"]
        return [f.download() for f in phi], fused


@pytest.mark.parametrize("kernel,target", [("march2", "1024"), ("march2", "1"), ("block", "1024"), ("march", "1")])
@pytest.mark.parametrize("visc", [-1, 1])
@pytest.mark.parametrize("vcycles,sweeps", [(1, (2, 2, 8)), (2, (2, 2, 8)), (1, (1, 3, 3))])
def test_fused_mg_solver_bit_exact(visc, vcycles, sweeps, kernel, target, monkeypatch):
    monkeypatch.setenv("LMM_SMOOTHER", "fused")
    monkeypatch.setenv("LMM_FUSED_KERNEL", kernel)
    monkeypatch.setenv("LMM_FUSED_TARGET", target)
    pb = P.small_problem(3, N3, visc_type=visc, seed=61)
    kw = dict(visc_type=visc, stag_mg_max_vcycles=vcycles, stag_mg_nsmooths_down=sweeps[0],
              stag_mg_nsmooths_up=sweeps[1], stag_mg_nsmooths_bottom=sweeps[2])
    g, p, cf = oracle_args(pb, **kw)
    rhs = _rand_faces(pb.shape, 7)
    guess = _rand_faces(pb.shape, 8)
    want = O.stag_mg_solver(g, p, cf, 0.9, rhs, phi=guess)
    got, fused = _mg_solve(pb, kw, rhs, guess)
    assert fused > 0, "the fused smoother did not run"
    assert bit_equal(got, want), relerr(got, want)


@pytest.mark.parametrize("visc", [-1, 1])
def test_fused_equals_colour_kernels(visc, monkeypatch):
    pb = P.small_problem(3, (128, 32, 16), visc_type=visc, seed=67)
    kw = dict(visc_type=visc)
    rhs = _rand_faces(pb.shape, 9)
    guess = _rand_faces(pb.shape, 10)
    monkeypatch.setenv("LMM_SMOOTHER", "colour")
    ref, fused0 = _mg_solve(pb, kw, rhs, guess)
    monkeypatch.setenv("LMM_SMOOTHER", "fused")
    got, fused1 = _mg_solve(pb, kw, rhs, guess)
    assert fused0 == 0 and fused1 > 0
    assert bit_equal(got, ref)


def test_fused_fp32_fgmres_equals_colour(monkeypatch):
    pb = P.small_problem(3, (64, 32, 32), visc_type=-1, seed=71, contrast=100.0)
    b_u = P.rhs_from_noise_host(pb)
    out = []
    for mode in ("colour", "fused"):
        monkeypatch.setenv("LMM_SMOOTHER", mode)
        with S.Context(3, pb.n, pb.dx, visc_type=-1, precision=S.FP32_MG) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            _, info = dp.solve()
            out.append((info["total_iter"], dp.solution()))
    assert out[0][0] == out[1][0]
    assert bit_equal(out[0][1][0], out[1][1][0]) and bit_equal(out[0][1][1], out[1][1][1])


def test_fused_gmres_parity(monkeypatch):
    monkeypatch.setenv("LMM_SMOOTHER", "fused")
    pb = P.small_problem(3, N3, visc_type=-1, seed=73, contrast=100.0)
    g, p, cf = oracle_args(pb)
    b_u = P.rhs_from_noise_host(pb)
    ou, op, oinfo = O.gmres(g, p, cf, pb.theta_alpha, b_u, pb.b_p)
    with S.Context(3, pb.n, pb.dx, **pb.solver) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=b_u)
        _, info = dp.solve()
        gu, gp = dp.solution()
    assert info["status"] == 0
    assert abs(info["total_iter"] - oinfo["total_iter"]) <= 1
    assert relerr(gu, ou) < 1e-10 and relerr(gp, op) < 1e-10


@pytest.mark.parametrize("kernel,target", [("march2", "1024"), ("march2", "1"), ("block", "1024"), ("march", "1")])
@pytest.mark.parametrize("n,vcycles,sweeps", [((64, 16, 32), 1, (2, 2)), ((128, 32, 16), 2, (3, 1)),
                                              ((64, 32, 32), 1, (1, 2))])
def test_fused_cc_multigrid_matches_oracle_and_colour(n, vcycles, sweeps, kernel, target, monkeypatch):
    monkeypatch.setenv("LMM_FUSED_KERNEL", kernel)
    monkeypatch.setenv("LMM_FUSED_TARGET", target)
    rng = np.random.default_rng(91)
    shape = tuple(reversed(n))
    ainv = [0.5 + rng.random(shape) for _ in range(3)]
    rhs = rng.standard_normal(shape)
    rhs -= rhs.mean()
    kw = dict(mg_max_vcycles=vcycles, mg_nsmooths_down=sweeps[0], mg_nsmooths_up=sweeps[1])
    want = O.macproject(O.geom(3, n, 0.5), O.params(**kw), ainv, rhs)
    out = {}
    for mode in ("colour", "fused"):
        monkeypatch.setenv("LMM_SMOOTHER", mode)
        with S.Context(3, n, 0.5, **kw) as ctx:
            a = ctx.faces(ainv)
            r = ctx.field(S.CELL, 1, rhs)
            phi = ctx.field(S.CELL)
            S.macproject(ctx, phi, None, a, r)
            out[mode] = phi.download()
    assert bit_equal(out["fused"], out["colour"])
    assert bit_equal(out["fused"], want)


@pytest.mark.parametrize("visc", [-1, 1])
@pytest.mark.parametrize("n,target", [((64, 16, 32), "1024"), ((128, 32, 16), "1"), ((64, 32, 32), "1"),
                                      ((32, 32, 16), "768")])
def test_three_components_in_one_launch_equal_separate(visc, n, target, monkeypatch):
    """The default fused sweep runs the three face components in one launch
    (64 x 8 tiles, or 32 x 16 on levels 32 points wide, as the plane-marching
    one-component launches do too); LMM_FUSED_X3=0 launches them one by one.
    Same point code, so the same bits."""
    monkeypatch.setenv("LMM_SMOOTHER", "fused")
    monkeypatch.setenv("LMM_FUSED_TARGET", target)
    pb = P.small_problem(3, n, visc_type=visc, seed=101)
    kw = dict(visc_type=visc, stag_mg_max_vcycles=2)
    rhs, guess = _rand_faces(pb.shape, 8), _rand_faces(pb.shape, 9)
    monkeypatch.setenv("LMM_FUSED_X3", "0")
    ref, n0 = _mg_solve(pb, kw, rhs, guess)
    monkeypatch.setenv("LMM_FUSED_X3", "1")
    got, n1 = _mg_solve(pb, kw, rhs, guess)
    assert n1 > 0
    if n[0] % 64 == 0 and n[0] // 2 % 64 == 0:  # no 32-wide fused level
        assert n0 == 3 * n1
    assert bit_equal(got, ref)


@pytest.mark.parametrize("visc", [-1, 1])
def test_narrow_tile_sweep_matches_oracle(visc, monkeypatch):
    """Levels 32 points wide run the three-component sweep on 32 x 16 tiles."""
    monkeypatch.setenv("LMM_SMOOTHER", "fused")
    pb = P.small_problem(3, (32, 32, 16), visc_type=visc, seed=107)
    kw = dict(visc_type=visc)
    g, p, cf = oracle_args(pb, **kw)
    rhs, guess = _rand_faces(pb.shape, 10), _rand_faces(pb.shape, 11)
    want = O.stag_mg_solver(g, p, cf, 0.9, rhs, phi=guess)
    got, fused = _mg_solve(pb, kw, rhs, guess)
    assert fused > 0
    assert bit_equal(got, want), relerr(got, want)


@pytest.mark.parametrize("n,coarse", [((64, 32, 2), "1"), ((128, 64, 4), "0"), ((128, 64, 4), "1")])
@pytest.mark.parametrize("prec", ["fp64", "fp32_mg"])
def test_first_sweep_from_zeros_equals_filled_guess(n, coarse, prec, monkeypatch):
    """A level whose first smoothing in a cycle is a fused sweep starts from
    its zero guess without the guess being written or read.  Covers a single
    level (z = 2 does not coarsen) and a bottom level smoothed in the level
    loop (LMM_COARSE=0), in both multigrids, against the colour kernels,
    which zero the guess in memory."""
    monkeypatch.setenv("LMM_COARSE", coarse)
    pb = P.small_problem(3, n, visc_type=-1, seed=109, contrast=30.0)
    b_u = P.rhs_from_noise_host(pb)
    precision = {"fp64": S.FP64, "fp32_mg": S.FP32_MG}[prec]
    out = {}
    for mode in ("fused", "colour"):
        monkeypatch.setenv("LMM_SMOOTHER", mode)
        with S.Context(3, pb.n, pb.dx, visc_type=-1, precision=precision, gmres_max_iter=12) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            _, info = dp.solve()
            out[mode] = (info["total_iter"], dp.solution())
    assert out["fused"][0] == out["colour"][0]
    assert bit_equal(out["fused"][1][0], out["colour"][1][0]) and bit_equal(out["fused"][1][1], out["colour"][1][1])


# 2D: both face components in one launch on 64 x 8 tiles


@pytest.mark.parametrize("visc", [-1, 1])
@pytest.mark.parametrize("n,vcycles,sweeps", [((128, 64), 1, (2, 2, 8)), ((256, 128), 2, (2, 2, 8)),
                                              ((128, 128), 1, (1, 3, 3))])
def test_fused_2d_mg_solver_bit_exact(n, visc, vcycles, sweeps, monkeypatch):
    monkeypatch.setenv("LMM_SMOOTHER", "fused")
    pb = P.small_problem(2, n, visc_type=visc, seed=113)
    kw = dict(visc_type=visc, stag_mg_max_vcycles=vcycles, stag_mg_nsmooths_down=sweeps[0],
              stag_mg_nsmooths_up=sweeps[1], stag_mg_nsmooths_bottom=sweeps[2])
    g, p, cf = oracle_args(pb, **kw)
    rhs = _rand_faces(pb.shape, 12, 2)
    guess = _rand_faces(pb.shape, 13, 2)
    want = O.stag_mg_solver(g, p, cf, 0.9, rhs, phi=guess)
    got, fused = _mg_solve(pb, kw, rhs, guess)
    assert fused > 0, "the fused smoother did not run"
    assert bit_equal(got, want), relerr(got, want)


@pytest.mark.parametrize("visc", [-1, 1])
def test_fused_2d_equals_colour_kernels(visc, monkeypatch):
    """At the default launch threshold (16 workgroups on 2D levels: 64^2 and up)."""
    pb = P.small_problem(2, (512, 256), visc_type=visc, seed=127)
    kw = dict(visc_type=visc, stag_mg_max_vcycles=2)
    rhs, guess = _rand_faces(pb.shape, 14, 2), _rand_faces(pb.shape, 15, 2)
    monkeypatch.setenv("LMM_SMOOTHER", "colour")
    ref, fused0 = _mg_solve(pb, kw, rhs, guess)
    monkeypatch.delenv("LMM_SMOOTHER")
    got, fused1 = _mg_solve(pb, kw, rhs, guess)
    assert fused0 == 0 and fused1 > 0
    assert bit_equal(got, ref)


@pytest.mark.parametrize("prec", ["fp64", "fp32_mg"])
@pytest.mark.parametrize("visc", [-1, 1])
def test_fused_2d_gmres_equals_colour(prec, visc, monkeypatch):
    """Whole solves: the V-cycles' first sweeps start from zeros that are not
    written (levels 1 and down, and the zero guess of each preconditioner
    application), in both multigrids."""
    pb = P.small_problem(2, (256, 128), visc_type=visc, seed=131, contrast=30.0)
    b_u = P.rhs_from_noise_host(pb)
    precision = {"fp64": S.FP64, "fp32_mg": S.FP32_MG}[prec]
    out = {}
    for mode in ("fused", "colour"):
        monkeypatch.setenv("LMM_SMOOTHER", mode)
        with S.Context(2, pb.n, pb.dx, visc_type=visc, precision=precision, gmres_max_iter=12) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            ctx.profiling(True)
            _, info = dp.solve()
            out[mode] = (info["total_iter"], dp.solution(), ctx.kernel_stats("stag_rb_fused")["launches"])
    assert out["fused"][2] > 0 and out["colour"][2] == 0
    assert out["fused"][0] == out["colour"][0]
    assert bit_equal(out["fused"][1][0], out["colour"][1][0]) and bit_equal(out["fused"][1][1], out["colour"][1][1])


@pytest.mark.parametrize("n,vcycles,sweeps,mode", [((128, 64), 1, (2, 2), "fused"), ((256, 128), 2, (3, 1), "fused"),
                                                   ((512, 256), 2, (2, 2), None)])
def test_fused_2d_cc_multigrid_matches_oracle_and_colour(n, vcycles, sweeps, mode, monkeypatch):
    """The 2D cell-centred red-black sweep in one launch per sweep (64 x 8
    tiles): with LMM_SMOOTHER=fused on every eligible level, and at the
    default threshold (16 workgroups on 2D levels: 8192 cells and up)."""
    rng = np.random.default_rng(93)
    shape = tuple(reversed(n))
    ainv = [0.5 + rng.random(shape) for _ in range(2)]
    rhs = rng.standard_normal(shape)
    rhs -= rhs.mean()
    kw = dict(mg_max_vcycles=vcycles, mg_nsmooths_down=sweeps[0], mg_nsmooths_up=sweeps[1])
    want = O.macproject(O.geom(2, n, 0.5), O.params(**kw), ainv, rhs)
    out = {}
    for m in ("colour", mode):
        if m is None:
            monkeypatch.delenv("LMM_SMOOTHER", raising=False)
        else:
            monkeypatch.setenv("LMM_SMOOTHER", m)
        with S.Context(2, n, 0.5, **kw) as ctx:
            a = ctx.faces(ainv)
            r = ctx.field(S.CELL, 1, rhs)
            phi = ctx.field(S.CELL)
            ctx.profiling(True)
            S.macproject(ctx, phi, None, a, r)
            out[m] = (phi.download(), ctx.kernel_stats("cc_smooth_l0")["launches"])
    assert out[mode][1] > 0 and out["colour"][1] == 0, "level 0 of the colour run is not fused; of the other run is"
    assert bit_equal(out[mode][0], out["colour"][0])
    assert bit_equal(out[mode][0], want)


# visc_type 2 and 3: the components couple, so a fused sweep runs one
# component per launch (red and black together), reading the other
# components in their state at that point of the colour order


@pytest.mark.parametrize("visc", [2, 3])
@pytest.mark.parametrize("dim,n,vcycles,sweeps", [(3, (64, 16, 32), 2, (2, 2, 8)), (3, (64, 32, 16), 1, (1, 3, 3)),
                                                  (2, (128, 64), 2, (2, 2, 8)), (2, (128, 128), 1, (1, 3, 3))])
def test_fused_coupled_visc_mg_solver_bit_exact(dim, n, vcycles, sweeps, visc, monkeypatch):
    monkeypatch.setenv("LMM_SMOOTHER", "fused")
    pb = P.small_problem(dim, n, visc_type=visc, seed=137)
    kw = dict(visc_type=visc, stag_mg_max_vcycles=vcycles, stag_mg_nsmooths_down=sweeps[0],
              stag_mg_nsmooths_up=sweeps[1], stag_mg_nsmooths_bottom=sweeps[2])
    g, p, cf = oracle_args(pb, **kw)
    rhs = _rand_faces(pb.shape, 16, dim)
    guess = _rand_faces(pb.shape, 17, dim)
    want = O.stag_mg_solver(g, p, cf, 0.9, rhs, phi=guess)
    got, fused = _mg_solve(pb, kw, rhs, guess)
    assert fused > 0, "the fused smoother did not run"
    assert bit_equal(got, want), relerr(got, want)


@pytest.mark.parametrize("prec", ["fp64", "fp32_mg"])
@pytest.mark.parametrize("dim,n,visc", [(3, (64, 32, 32), 2), (2, (256, 128), 2), (3, (64, 32, 32), 3)])
def test_fused_coupled_visc_gmres_equals_colour(dim, n, visc, prec, monkeypatch):
    """Whole solves at the default launch threshold, against the colour kernels."""
    pb = P.small_problem(dim, n, visc_type=visc, seed=139, contrast=1.0)
    b_u = P.rhs_from_noise_host(pb)
    precision = {"fp64": S.FP64, "fp32_mg": S.FP32_MG}[prec]
    out = {}
    for mode in ("colour", None):
        if mode is None:
            monkeypatch.delenv("LMM_SMOOTHER", raising=False)
        else:
            monkeypatch.setenv("LMM_SMOOTHER", mode)
        with S.Context(dim, pb.n, pb.dx, visc_type=visc, precision=precision, gmres_max_iter=12) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            ctx.profiling(True)
            _, info = dp.solve()
            out[mode] = (info["total_iter"], dp.solution(), ctx.kernel_stats("stag_rb_fused")["launches"])
    assert out[None][2] > 0 and out["colour"][2] == 0
    assert out[None][0] == out["colour"][0]
    assert bit_equal(out[None][1][0], out["colour"][1][0]) and bit_equal(out[None][1][1], out["colour"][1][1])


@pytest.mark.parametrize("prec", ["fp64", "fp32_mg"])
def test_autotune_keeps_the_bits(prec):
    """lmm_ctx_autotune picks the z-slab target among its candidates by timing
    the fine-level sweeps; a solve after it equals one before it bit for bit
    (the slab shape never changes the arithmetic), and the target stays in the
    context's kernel options."""
    pb = P.small_problem(3, (128, 64, 64), visc_type=-1, seed=149, contrast=10.0)
    b_u = P.rhs_from_noise_host(pb)
    precision = {"fp64": S.FP64, "fp32_mg": S.FP32_MG}[prec]
    with S.Context(3, pb.n, pb.dx, visc_type=-1, precision=precision, gmres_max_iter=8) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=b_u)
        _, i0 = dp.solve()
        before = dp.solution()
        t = ctx.autotune()
        assert t in (256, 384, 512, 768, 1024, 1536, 2048)
        assert ctx.kernel_options()["fused_target"] == t
        dp.zero_solution()
        _, i1 = dp.solve()
        after = dp.solution()
    assert i0["total_iter"] == i1["total_iter"]
    assert bit_equal(before[0], after[0]) and bit_equal(before[1], after[1])
