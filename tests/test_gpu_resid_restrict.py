"""The fused residual + restriction of the staggered multigrid's down-leg
against the two separate passes (residual into a fine buffer, then the
full-weighting restriction) and against the oracle.  Levels with x a
multiple of 64 and y of 8 (3D, visc_type -1, 1, 2, 3) use the fused pass;
LMM_RESID_RESTRICT=0 keeps the two passes everywhere.  Every coarse value
sums the same fine residuals in the same order, so the bits agree."""
import numpy as np
import pytest

from lowmachmixtures_amd import problems as P
from lowmachmixtures_amd import stokes as S
from oracle import oracle as O
from tests.helpers import bit_equal, oracle_args, relerr

pytestmark = pytest.mark.gpu


def _faces(shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) for _ in range(3)]


def _mg(pb, kw, rhs, guess):
    with S.Context(3, pb.n, pb.dx, **kw) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=rhs)
        phi = ctx.faces(guess)
        ctx.profiling(True)
        S.stag_mg_solver(ctx, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, 0.8, phi, dp.b_u)
        return [f.download() for f in phi], ctx.kernel_stats("stag_rr_fused")["launches"]


@pytest.mark.parametrize("n,visc,target", [((128, 32, 32), -1, "1024"), ((128, 32, 32), 1, "1024"),
                                           ((64, 64, 64), -1, "1"), ((64, 16, 32), -1, "1"),
                                           ((64, 8, 16), 1, "1024"), ((128, 32, 32), 2, "1024"),
                                           ((64, 16, 32), 3, "1")])
def test_resid_restrict_fused_equals_two_passes(n, visc, target, monkeypatch):
    monkeypatch.setenv("LMM_FUSED_TARGET", target)
    pb = P.small_problem(3, n, visc_type=visc, seed=91)
    kw = dict(visc_type=visc, stag_mg_max_vcycles=2)
    rhs, guess = _faces(pb.shape, 5), _faces(pb.shape, 6)
    monkeypatch.setenv("LMM_RESID_RESTRICT", "0")
    ref, n0 = _mg(pb, kw, rhs, guess)
    monkeypatch.setenv("LMM_RESID_RESTRICT", "1")
    got, n1 = _mg(pb, kw, rhs, guess)
    assert n0 == 0 and n1 > 0
    assert bit_equal(got, ref)


@pytest.mark.parametrize("visc", [-1, 2, 3])
def test_resid_restrict_fused_matches_oracle(visc, monkeypatch):
    monkeypatch.setenv("LMM_RESID_RESTRICT", "1")
    pb = P.small_problem(3, (128, 16, 32), visc_type=visc, seed=93)
    kw = dict(visc_type=visc)
    g, p, cf = oracle_args(pb, **kw)
    rhs, guess = _faces(pb.shape, 7), _faces(pb.shape, 8)
    want = O.stag_mg_solver(g, p, cf, 0.8, rhs, phi=guess)
    got, n1 = _mg(pb, kw, rhs, guess)
    assert n1 > 0
    assert bit_equal(got, want), relerr(got, want)


def test_resid_restrict_fp32_fgmres_equal(monkeypatch):
    pb = P.small_problem(3, (128, 32, 32), visc_type=-1, seed=95, contrast=100.0)
    b_u = P.rhs_from_noise_host(pb)
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("LMM_RESID_RESTRICT", mode)
        with S.Context(3, pb.n, pb.dx, visc_type=-1, precision=S.FP32_MG) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            _, info = dp.solve()
            out.append((info["total_iter"], dp.solution()))
    assert out[0][0] == out[1][0]
    assert bit_equal(out[0][1][0], out[1][1][0]) and bit_equal(out[0][1][1], out[1][1][1])


@pytest.mark.parametrize("dim,n", [(2, (64, 32)), (3, (64, 32, 16)), (3, (32, 32, 32))])
def test_cc_resid_restrict_fused_equals_two_passes(dim, n, monkeypatch):
    """The cell-centred down-leg: residual kept in registers and averaged onto
    the coarse right-hand side, against residual-then-average.  The solution
    of the cell-centred multigrid agrees to the bit, and with the oracle."""
    pb = P.small_problem(dim, n, seed=97)
    g, p, cf = oracle_args(pb, mg_max_vcycles=3)
    ainv = [1.0 / a for a in pb.alpha]
    rhs = np.random.default_rng(11).standard_normal(pb.shape)
    rhs -= rhs.mean()
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("LMM_RESID_RESTRICT", mode)
        with S.Context(dim, pb.n, pb.dx, mg_max_vcycles=3) as ctx:
            ai = ctx.faces(ainv)
            r = ctx.field(S.CELL, 1, rhs)
            phi = ctx.field(S.CELL)
            S.macproject(ctx, phi, None, ai, r)
            out.append(phi.download())
    assert bit_equal(out[1], out[0])
    want = O.macproject(g, p, ainv, rhs)
    assert bit_equal(out[1], want), relerr(out[1], want)
