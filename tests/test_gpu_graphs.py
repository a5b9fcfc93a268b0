"""HIP-graph replay of the FP32 preconditioner application (kernel option
precon_graph): the flexible GMRES captures each output slot's application
once and replays it.  A replay runs the same kernels with the same arguments,
so every solve must equal the eager one bit for bit -- checked over several
solves in a row on one context (replays after the first), with eager work on
the same stream between them (a different right-hand side, a changed
theta_alpha that forces a new capture, a changed kernel option that drops the
graphs), and over a batched ragged ensemble."""
import dataclasses

import pytest

from lowmachmixtures_amd import problems as P
from lowmachmixtures_amd import stokes as S
from tests.helpers import bit_equal

pytestmark = pytest.mark.gpu


def _solve(dp, theta=None):
    dp.zero_solution()
    if theta is not None:
        dp.pb = dataclasses.replace(dp.pb, theta_alpha=theta)
    _, info = dp.solve()
    u, p = dp.solution()
    return info["total_iter"], info["vcycles"], u, p


@pytest.mark.parametrize("dim,n", [(2, (64, 64)), (3, (32, 32, 32))])
def test_replayed_preconditioner_equals_eager(dim, n):
    pb = P.small_problem(dim, n, visc_type=-1, seed=11, contrast=10.0)
    b1 = P.rhs_from_noise_host(pb)
    b2 = [0.5 * b for b in P.rhs_from_noise_host(P.small_problem(dim, n, visc_type=-1, seed=12, contrast=10.0))]
    runs = {}
    for graph in (0, 1):
        with S.Context(dim, pb.n, pb.dx, visc_type=-1, precision=S.FP32_MG) as ctx:
            ctx.kernel_options(precon_graph=graph)
            dp = S.DeviceProblem(ctx, pb, rhs=b1)
            seq = [_solve(dp), _solve(dp)]                       # captures, then replays
            dp2 = S.DeviceProblem(ctx, pb, rhs=b2)               # eager uploads between replays
            seq.append(_solve(dp2))
            seq.append(_solve(dp, theta=2.0 * pb.theta_alpha))   # new theta: a new capture
            seq.append(_solve(dp, theta=2.0 * pb.theta_alpha))
            ctx.kernel_options(resid_restrict=0)                 # graphs dropped, recaptured
            seq.append(_solve(dp))
            seq.append(_solve(dp))
            runs[graph] = seq
    for k, (a, b) in enumerate(zip(runs[0], runs[1])):
        assert a[0] == b[0] and a[1] == b[1], (k, a[:2], b[:2])
        assert bit_equal(a[2], b[2]) and bit_equal(a[3], b[3]), k


def test_ensemble_with_graphs_equals_without():
    pbs = [P.config4_member(n, 300 + m) for m in range(3) for n in (64, 128)]
    got = {}
    for graph in (0, 1):
        members = []
        try:
            for pb in pbs:
                ctx = S.Context(2, pb.n, pb.dx, **pb.solver)
                ctx.kernel_options(precon_graph=graph)
                members.append((ctx, S.DeviceProblem(ctx, pb, rhs=P.rhs_from_noise_host(pb))))
            dps = [dp for _, dp in members]
            out = []
            for rep in range(2):   # the second batch replays graphs captured in the first
                for dp in dps:
                    dp.zero_solution()
                res = S.gmres_batched(dps)
                assert all(st == 0 and info["status"] == 0 for st, _, info in res)
                out.append([(info["total_iter"],) + tuple(dp.solution()) for (_, _, info), dp in zip(res, dps)])
            got[graph] = out
        finally:
            for ctx, _ in members:
                ctx.close()
    for rep in range(2):
        for a, b in zip(got[0][rep], got[1][rep]):
            assert a[0] == b[0]
            assert bit_equal(a[1], b[1]) and bit_equal(a[2], b[2])
