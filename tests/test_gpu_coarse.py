"""The single-workgroup coarse-level V-cycle against one launch per pass.

Levels with at most 4096 cells run the whole coarse part of both
multigrids (down-leg, bottom, up-leg) in one launch; LMM_COARSE=0 keeps one
launch per pass on every level.  Every point is computed by the same device
function in the same order of passes, so both must give the same bits."""
import numpy as np
import pytest

from lowmachmixtures_amd import problems as P
from lowmachmixtures_amd import stokes as S
from oracle import oracle as O
from tests.helpers import bit_equal, oracle_args, relerr

pytestmark = pytest.mark.gpu


def _faces(dim, shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) for _ in range(dim)]


def _mg(pb, kw, rhs, guess):
    with S.Context(pb.dim, pb.n, pb.dx, **kw) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=rhs)
        phi = ctx.faces(guess)
        ctx.profiling(True)
        S.stag_mg_solver(ctx, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, 0.7, phi, dp.b_u)
        n = ctx.kernel_stats("stag_coarse")["launches"]
        return [f.download() for f in phi], n


@pytest.mark.parametrize("dim,n,visc", [(3, (64, 64, 64), -1), (3, (64, 32, 32), -2), (3, (32, 32, 64), 3),
                                        (2, (256, 128), -1), (2, (128, 128), -3)])
def test_stag_coarse_launch_equals_per_pass(dim, n, visc, monkeypatch):
    pb = P.small_problem(dim, n, visc_type=visc, seed=81)
    kw = dict(visc_type=visc, stag_mg_max_vcycles=2)
    rhs, guess = _faces(dim, pb.shape, 1), _faces(dim, pb.shape, 2)
    monkeypatch.setenv("LMM_COARSE", "0")
    ref, n0 = _mg(pb, kw, rhs, guess)
    monkeypatch.setenv("LMM_COARSE", "1")
    got, n1 = _mg(pb, kw, rhs, guess)
    assert n0 == 0 and n1 > 0
    assert bit_equal(got, ref)


def test_stag_coarse_launch_matches_oracle(monkeypatch):
    monkeypatch.setenv("LMM_COARSE", "1")
    pb = P.small_problem(3, (64, 64, 32), visc_type=-1, seed=83)
    kw = dict(visc_type=-1)
    g, p, cf = oracle_args(pb, **kw)
    rhs, guess = _faces(3, pb.shape, 3), _faces(3, pb.shape, 4)
    want = O.stag_mg_solver(g, p, cf, 0.7, rhs, phi=guess)
    got, n1 = _mg(pb, kw, rhs, guess)
    assert n1 > 0
    assert bit_equal(got, want), relerr(got, want)


@pytest.mark.parametrize("dim,n", [(3, (64, 64, 64)), (2, (256, 256))])
def test_cc_coarse_launch_matches_oracle_and_per_pass(dim, n, monkeypatch):
    rng = np.random.default_rng(5)
    shape = tuple(reversed(n))
    ainv = [0.5 + rng.random(shape) for _ in range(dim)]
    rhs = rng.standard_normal(shape)
    rhs -= rhs.mean()
    g = O.geom(dim, n, 0.5)
    p = O.params(mg_max_vcycles=2)
    want = O.macproject(g, p, ainv, rhs)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("LMM_COARSE", mode)
        with S.Context(dim, n, 0.5, mg_max_vcycles=2) as ctx:
            a = ctx.faces(ainv)
            r = ctx.field(S.CELL, 1, rhs)
            phi = ctx.field(S.CELL)
            ctx.profiling(True)
            S.macproject(ctx, phi, None, a, r)
            out[mode] = (phi.download(), ctx.kernel_stats("cc_coarse")["launches"])
    assert out["0"][1] == 0 and out["1"][1] > 0
    assert bit_equal(out["1"][0], out["0"][0])
    assert bit_equal(out["1"][0], want)


def test_fp32_fgmres_coarse_equals_per_pass(monkeypatch):
    pb = P.small_problem(3, (64, 64, 64), visc_type=-1, seed=87, contrast=100.0)
    b_u = P.rhs_from_noise_host(pb)
    res = []
    for mode in ("0", "1"):
        monkeypatch.setenv("LMM_COARSE", mode)
        with S.Context(3, pb.n, pb.dx, visc_type=-1, precision=S.FP32_MG) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            _, info = dp.solve()
            res.append((info["total_iter"], dp.solution()))
    assert res[0][0] == res[1][0]
    assert bit_equal(res[0][1][0], res[1][1][0]) and bit_equal(res[0][1][1], res[1][1][1])
