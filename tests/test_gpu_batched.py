"""Batched ensembles (config 4: many independent 2D solves, FP32 V-cycles in
FP64 FGMRES): lmm_gmres_batched solves each member in its own context, on its
own stream, from a pool of host threads.  Every member must give exactly what
lmm_gmres gives for it alone."""
import numpy as np
import pytest

from lowmachmixtures_amd import problems as P
from lowmachmixtures_amd import stokes as S
from tests.helpers import bit_equal

pytestmark = pytest.mark.gpu


def _members(n, count, prec):
    out = []
    for m in range(count):
        pb = P.small_problem(2, n, visc_type=-1, seed=200 + m, contrast=10.0)
        b_u = P.rhs_from_noise_host(pb)
        ctx = S.Context(2, pb.n, pb.dx, visc_type=-1, precision=prec)
        out.append((ctx, S.DeviceProblem(ctx, pb, rhs=b_u)))
    return out


@pytest.mark.parametrize("prec,nthreads", [(S.FP32_MG, 0), (S.FP32_MG, 3), (S.FP64, 0)])
def test_batched_equals_one_at_a_time(prec, nthreads):
    members = _members((64, 64), 8, prec)
    try:
        seq = []
        for ctx, dp in members:
            _, info = dp.solve()
            seq.append((info["total_iter"], dp.solution()))
            dp.zero_solution()
        res = S.gmres_batched([dp for _, dp in members], nthreads=nthreads)
        for (status, _, info), (it, sol), (_, dp) in zip(res, seq, members):
            assert status == 0 and info["status"] == 0
            assert info["total_iter"] == it
            got = dp.solution()
            assert bit_equal(got[0], sol[0]) and bit_equal(got[1], sol[1])
    finally:
        for ctx, _ in members:
            ctx.close()


def test_batched_rejects_a_shared_context():
    members = _members((32, 32), 2, S.FP32_MG)
    try:
        dp = members[0][1]
        with pytest.raises(S.LmmError):
            S.gmres_batched([dp, dp])
    finally:
        for ctx, _ in members:
            ctx.close()
