"""Kernel-shape switches that must not change a single bit.

Each of these picks another launch form for the same per-point arithmetic:
two point pairs per thread on 64 x 16 tiles in the FP32 three-component
sweep (LMM_SWEEP_PAIRS), the staggered prolongation from coarse values
staged in LDS (LMM_PROLONG_TILED), the coarse staggered V-cycle with one
workgroup per component (LMM_COARSE_SPLIT), and the residual + restriction
of the three components in one launch, on 64 x 8 and 32 x 16 tiles
(LMM_RR_THREE).  A whole GMRES solve with every switch at its default and
with the one switch off must give the same iteration count and the same
bits, FP64 and FP32-MG, on a box whose levels take every one of the forms
(x a multiple of 64, a 32-wide level, coarse levels below 16^3)."""
import pytest

from lowmachmixtures_amd import problems as P
from lowmachmixtures_amd import stokes as S
from tests.helpers import bit_equal

pytestmark = pytest.mark.gpu

SWITCHES = ["LMM_SWEEP_PAIRS=1", "LMM_PROLONG_TILED=0", "LMM_COARSE_SPLIT=0", "LMM_RR_THREE=0"]


def _solve(pb, b_u, precision, visc):
    with S.Context(3, pb.n, pb.dx, visc_type=visc, precision=precision) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=b_u)
        _, info = dp.solve()
        return info["total_iter"], dp.solution()


@pytest.mark.parametrize("switch", SWITCHES)
@pytest.mark.parametrize("precision,visc", [(S.FP32_MG, -1), (S.FP64, -1), (S.FP32_MG, 1), (S.FP64, 2)])
def test_switch_keeps_every_bit(switch, precision, visc, monkeypatch):
    pb = P.small_problem(3, (64, 64, 64), visc_type=visc, seed=91, contrast=100.0)
    b_u = P.rhs_from_noise_host(pb)
    for var in SWITCHES:
        monkeypatch.delenv(var.split("=")[0], raising=False)
    it0, (u0, p0) = _solve(pb, b_u, precision, visc)
    name, value = switch.split("=")
    monkeypatch.setenv(name, value)
    it1, (u1, p1) = _solve(pb, b_u, precision, visc)
    assert it0 == it1
    assert bit_equal(u0, u1) and bit_equal(p0, p1)
