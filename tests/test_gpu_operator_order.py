"""The 3D visc -1 / 1 operator (apply_matrix.f90 through k_apply_matrix3)
numbers its work in 64 x 4 x 1 bricks, ordered so the workgroups in flight
cover a compact box, where x is a multiple of 64 and y of 4; elsewhere in row
order.  Every cell's arithmetic is the same in both orders, so the results
are the oracle's bits and the two orders agree to the bit (LMM_AM_BRICK=0
keeps row order; a positive value sets the slab depth in z)."""
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


@pytest.mark.parametrize("n", [(64, 8, 8), (128, 16, 32), (192, 16, 8)])
@pytest.mark.parametrize("visc", [-1, 1])
@pytest.mark.parametrize("depth", ["32", "16", "1"])
def test_brick_order_operator_matches_oracle(n, visc, depth, monkeypatch):
    monkeypatch.setenv("LMM_AM_BRICK", depth)
    pb = P.small_problem(3, n, visc_type=visc, seed=31)
    g, p, cf = oracle_args(pb)
    xu = _faces(pb.shape, 32)
    xp = np.random.default_rng(33).standard_normal(pb.shape)
    want_u, want_p = O.apply_matrix(g, p, cf, 1.3, xu, xp)
    with S.Context(3, pb.n, pb.dx, visc_type=visc) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=[np.zeros(pb.shape)] * 3)
        x_u, x_p = ctx.faces(xu), ctx.field(S.CELL, 1, xp)
        b_u, b_p = ctx.faces(), ctx.field(S.CELL)
        S.apply_matrix(ctx, b_u, b_p, x_u, x_p, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, 1.3)
        got_u, got_p = [f.download() for f in b_u], b_p.download()
    assert bit_equal(got_u, want_u), relerr(got_u, want_u)
    assert bit_equal(got_p, want_p), relerr(got_p, want_p)


@pytest.mark.parametrize("precision", [S.FP64, S.FP32_MG])
def test_brick_order_solve_equals_row_order(precision, monkeypatch):
    """Whole GMRES solves: the FP32-Z operator (FP32-MG) and the fused restart
    residual b - Ax go through the same kernel."""
    pb = P.small_problem(3, (128, 32, 32), visc_type=-1, seed=35, contrast=100.0)
    b_u = P.rhs_from_noise_host(pb)
    out = []
    for mode in ("0", "32"):
        monkeypatch.setenv("LMM_AM_BRICK", mode)
        with S.Context(3, pb.n, pb.dx, visc_type=-1, precision=precision) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            _, info = dp.solve()
            out.append((info["total_iter"], dp.solution()))
    assert out[0][0] == out[1][0]
    assert bit_equal(out[0][1][0], out[1][1][0]) and bit_equal(out[0][1][1], out[1][1][1])
