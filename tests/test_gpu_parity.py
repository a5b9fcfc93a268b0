"""Parity of the gfx950 path against the CPU oracle, through the C ABI.

Stencils, restrictions, prolongations, the staggered and cell-centred
V-cycles and the Stokes operator carry no reductions, use the reference's
operation order and are built without FP contraction: they must agree bit
for bit.  Anything with a global sum (inner products, pressure mean, GMRES)
differs only by summation order: relative 1e-10 on velocity and pressure and
the same GMRES iteration count (within +-1) are required.
"""
import numpy as np
import pytest

from lowmachmixtures_amd import problems as P
from lowmachmixtures_amd import stokes as S
from oracle import oracle as O
from tests.helpers import bit_equal, oracle_args, relerr

pytestmark = pytest.mark.gpu

VISC = [-1, 1, -2, 2, -3, 3]


def _rand_faces(dim, shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) for _ in range(dim)]


@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (3, (8, 4, 8))])
@pytest.mark.parametrize("visc", VISC)
def test_stag_applyop_bit_exact(dim, n, visc):
    pb = P.small_problem(dim, n, visc_type=visc, seed=11)
    g, p, cf = oracle_args(pb)
    phi = _rand_faces(dim, pb.shape, 5)
    with S.Context(dim, pb.n, pb.dx, visc_type=visc) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=[np.zeros(pb.shape)] * dim)
        ph = ctx.faces(phi)
        lp = ctx.faces()
        for color in range(0, 2 * dim + 1):
            want = O.stag_applyop(g, visc, cf, 1.7, phi, color=color)
            for f in lp:
                f.setval(0.0)
            S.stag_applyop_level(ctx, ph, lp, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, color=color, theta_alpha=1.7)
            got = [f.download() for f in lp]
            if color == 0:
                assert bit_equal(got, want), f"color {color}"
            else:
                a = (color - 1) // 2
                par = (color + 1) % 2
                idx = np.indices(pb.shape).sum(axis=0) % 2 == par
                assert np.array_equal(got[a][idx], want[a][idx]), f"color {color}"


@pytest.mark.parametrize("dim,n", [(2, (32, 16)), (3, (16, 8, 8))])
@pytest.mark.parametrize("visc", VISC)
@pytest.mark.parametrize("smoother,vcycles", [(1, 1), (1, 2), (0, 1)])
def test_stag_mg_solver_bit_exact(dim, n, visc, smoother, vcycles):
    pb = P.small_problem(dim, n, visc_type=visc, seed=13)
    kw = dict(visc_type=visc, stag_mg_smoother=smoother, stag_mg_max_vcycles=vcycles)
    g, p, cf = oracle_args(pb, **kw)
    rhs = _rand_faces(dim, pb.shape, 7)
    guess = _rand_faces(dim, pb.shape, 8)
    want = O.stag_mg_solver(g, p, cf, 0.9, rhs, phi=guess)
    with S.Context(dim, pb.n, pb.dx, **kw) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=rhs)
        phi = ctx.faces(guess)
        S.stag_mg_solver(ctx, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, 0.9, phi, dp.b_u)
        got = [f.download() for f in phi]
    assert bit_equal(got, want), relerr(got, want)


@pytest.mark.parametrize("dim,n", [(2, (32, 32)), (3, (16, 8, 8))])
def test_macproject_bit_exact(dim, n):
    pb = P.small_problem(dim, n, seed=17)
    g, p, cf = oracle_args(pb, mg_max_vcycles=2)
    ainv = [1.0 / a for a in pb.alpha]
    rhs = np.random.default_rng(3).standard_normal(pb.shape)
    rhs -= rhs.mean()
    want = O.macproject(g, p, ainv, rhs)
    with S.Context(dim, pb.n, pb.dx, mg_max_vcycles=2) as ctx:
        ai = ctx.faces(ainv)
        r = ctx.field(S.CELL, 1, rhs)
        phi = ctx.field(S.CELL)
        S.macproject(ctx, phi, None, ai, r)
        got = phi.download()
    assert bit_equal(got, want), relerr(got, want)


@pytest.mark.parametrize("dim,n", [(2, (16, 16)), (3, (8, 8, 8))])
@pytest.mark.parametrize("visc", VISC)
def test_apply_matrix_bit_exact(dim, n, visc):
    pb = P.small_problem(dim, n, visc_type=visc, seed=19)
    g, p, cf = oracle_args(pb)
    xu = _rand_faces(dim, pb.shape, 21)
    xp = np.random.default_rng(22).standard_normal(pb.shape)
    want_u, want_p = O.apply_matrix(g, p, cf, 1.3, xu, xp)
    with S.Context(dim, pb.n, pb.dx, visc_type=visc) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=[np.zeros(pb.shape)] * dim)
        x_u, x_p = ctx.faces(xu), ctx.field(S.CELL, 1, xp)
        b_u, b_p = ctx.faces(), ctx.field(S.CELL)
        S.apply_matrix(ctx, b_u, b_p, x_u, x_p, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, 1.3)
        got_u, got_p = [f.download() for f in b_u], b_p.download()
    assert bit_equal(got_u, want_u)
    assert bit_equal(got_p, want_p)


@pytest.mark.parametrize("dim,n", [(2, (32, 32)), (3, (16, 16, 8))])
@pytest.mark.parametrize("visc,theta", [(-1, 1.0), (-2, 1.0), (-3, 0.5), (1, 0.0)])
def test_apply_precon(dim, n, visc, theta):
    pb = P.small_problem(dim, n, visc_type=visc, seed=23, theta_alpha=theta)
    g, p, cf = oracle_args(pb)
    bu = _rand_faces(dim, pb.shape, 31)
    bp = np.random.default_rng(32).standard_normal(pb.shape)
    bp -= bp.mean()
    want_u, want_p = O.apply_precon(g, p, cf, theta, bu, bp)
    with S.Context(dim, pb.n, pb.dx, visc_type=visc) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=bu)
        b_p = ctx.field(S.CELL, 1, bp)
        S.apply_precon(ctx, dp.b_u, b_p, dp.x_u, dp.x_p, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, theta)
        got_u, got_p = dp.solution()
    if theta != 0.0:
        # velocity carries no global sum: exact
        assert bit_equal(got_u, want_u)
    else:
        assert relerr(got_u, want_u) < 1e-13
    # pressure: mean subtraction is a global sum (order differs)
    assert relerr(got_p, want_p) < 1e-13


@pytest.mark.parametrize("dim,n", [(2, (32, 32)), (3, (16, 16, 16))])
def test_inner_products(dim, n):
    shape = tuple(reversed(n))
    g = O.geom(dim, n)
    a, b = _rand_faces(dim, shape, 41), _rand_faces(dim, shape, 42)
    ca = np.random.default_rng(43).standard_normal(shape)
    cb = np.random.default_rng(44).standard_normal(shape)
    want = O.stag_inner_prod(g, a, b)
    wantc = O.cc_inner_prod(g, ca, cb)
    with S.Context(dim, n) as ctx:
        got = S.stag_inner_prod(ctx, ctx.faces(a), ctx.faces(b))
        gotc = S.cc_inner_prod(ctx, ctx.field(S.CELL, 1, ca), ctx.field(S.CELL, 1, cb))
    for x, y in zip(got, want):
        assert abs(x - y) <= 1e-12 * max(1.0, abs(y))
    assert abs(gotc - wantc) <= 1e-12 * max(1.0, abs(wantc))


def _solve_both(pb, **solver):
    g, p, cf = oracle_args(pb, **solver)
    b_u = P.rhs_from_noise_host(pb)
    ou, op, oinfo = O.gmres(g, p, cf, pb.theta_alpha, b_u, pb.b_p)
    kw = dict(pb.solver)
    kw.update(solver)
    with S.Context(pb.dim, pb.n, pb.dx, **kw) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=b_u)
        npr, info = dp.solve()
        gu, gp = dp.solution()
    return (ou, op, oinfo), (gu, gp, info, npr)


@pytest.mark.parametrize("dim,n,visc", [(2, (64, 64), -1), (2, (32, 64), -2), (3, (16, 16, 16), -1),
                                        (3, (16, 8, 16), -3), (3, (16, 16, 16), 1)])
def test_gmres_parity(dim, n, visc):
    pb = P.small_problem(dim, n, visc_type=visc, seed=29)
    (ou, op, oinfo), (gu, gp, info, npr) = _solve_both(pb, visc_type=visc)
    assert info["status"] == 0
    assert abs(info["total_iter"] - oinfo["total_iter"]) <= 1
    assert relerr(gu, ou) < 1e-10
    assert relerr(gp, op) < 1e-10
    assert abs(npr - oinfo["norm_pre_rhs"]) <= 1e-12 * oinfo["norm_pre_rhs"]


def test_gmres_config1_parity():
    """Plain FP64 2D periodic 128x128 solve (2 V-cycles per preconditioner call)."""
    pb = P.config1(128, seed=1)
    (ou, op, oinfo), (gu, gp, info, npr) = _solve_both(pb)
    assert info["status"] == 0
    assert abs(info["total_iter"] - oinfo["total_iter"]) <= 1
    assert relerr(gu, ou) < 1e-10 and relerr(gp, op) < 1e-10


def test_gmres_scale_factor_and_pnorm():
    pb = P.small_problem(2, (32, 32), visc_type=-1, seed=31)
    (ou, op, oinfo), (gu, gp, info, npr) = _solve_both(pb, scale_factor=2.5, p_norm_weight=0.5)
    assert abs(info["total_iter"] - oinfo["total_iter"]) <= 1
    assert relerr(gu, ou) < 1e-10 and relerr(gp, op) < 1e-10


def test_gmres_zero_rhs_returns_zero():
    pb = P.small_problem(2, (16, 16), seed=3)
    zero = [np.zeros(pb.shape)] * 2
    with S.Context(2, pb.n, pb.dx, visc_type=-1) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=zero)
        for f in dp.x_u:
            f.setval(3.0)
        npr, info = dp.solve()
        gu, gp = dp.solution()
    assert info["total_iter"] == 0
    assert all(np.all(u == 0) for u in gu) and np.all(gp == 0)


@pytest.mark.parametrize("dim,n", [(2, (64, 64)), (3, (32, 32, 32))])
def test_fp32_mg_fgmres_accuracy(dim, n):
    """FP32 V-cycles inside FP64 flexible GMRES: true residual 1e-9 and error vs the FP64 oracle."""
    pb = P.small_problem(dim, n, visc_type=-1, seed=37, contrast=100.0)
    g, p, cf = oracle_args(pb)
    b_u = P.rhs_from_noise_host(pb)
    ou, op, oinfo = O.gmres(g, p, cf, pb.theta_alpha, b_u, pb.b_p)
    with S.Context(dim, pb.n, pb.dx, visc_type=-1, precision=S.FP32_MG, gmres_max_inner=10) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=b_u)
        npr, info = dp.solve()
        gu, gp = dp.solution()
        # independent residual with the FP64 operator
        r_u, r_p = ctx.faces(), ctx.field(S.CELL)
        S.apply_matrix(ctx, r_u, r_p, dp.x_u, dp.x_p, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, pb.theta_alpha)
        Ax = [f.download() for f in r_u]
    assert info["status"] == 0
    res = np.sqrt(sum(((a - b) ** 2).sum() for a, b in zip(Ax, b_u)))
    bn = np.sqrt(sum((b ** 2).sum() for b in b_u))
    assert res / bn < 2e-9
    assert relerr(gu, ou) < 1e-7


@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (3, (8, 4, 8))])
@pytest.mark.parametrize("visc", VISC)
def test_diffusive_m_fluxdiv_bit_exact(dim, n, visc):
    """diffusive_m_fluxdiv.f90:16-73: m_update - (0 - L_eta)umac, bit for bit."""
    pb = P.small_problem(dim, n, visc_type=visc, seed=13)
    g = O.geom(dim, n, pb.dx)
    umac = _rand_faces(dim, pb.shape, 7)
    m0 = _rand_faces(dim, pb.shape, 8)
    kappa = 0.3 * pb.beta
    want = O.diffusive_m_fluxdiv(g, visc, m0, umac, pb.beta, pb.beta_ed, kappa)
    with S.Context(dim, pb.n, pb.dx, visc_type=visc) as ctx:
        m = ctx.faces(m0)
        S.diffusive_m_fluxdiv(ctx, m, ctx.faces(umac), ctx.field(S.CELL, 1, pb.beta), ctx.edges(pb.beta_ed),
                              ctx.field(S.CELL, 1, kappa))
        got = [f.download() for f in m]
    assert bit_equal(got, want)


@pytest.mark.parametrize("dim,n", [(2, (32, 16)), (3, (8, 16, 8))])
def test_stochastic_flux_bit_exact(dim, n):
    pb = P.small_problem(dim, n, seed=47)
    g = O.geom(dim, n, pb.dx)
    rng = np.random.default_rng(5)
    eta = 1.0 + rng.random(pb.shape)
    temp = 1.0 + rng.random(pb.shape)
    eta_ed = [P.avg_cc_to_node_2d(eta)] if dim == 2 else P.avg_cc_to_edge_3d(eta)
    temp_ed = [P.avg_cc_to_node_2d(temp)] if dim == 2 else P.avg_cc_to_edge_3d(temp)
    f0 = _rand_faces(dim, pb.shape, 6)
    want = O.stochastic_m_fluxdiv(g, pb.W_cc, pb.W_ed, eta, eta_ed, temp, temp_ed, 0.37, f0)
    with S.Context(dim, n, pb.dx) as ctx:
        force = ctx.faces(f0)
        S.stoch_m_force(ctx, force, ctx.field(S.CELL, dim, pb.W_cc), ctx.edges(pb.W_ed, ncomp=2),
                        ctx.field(S.CELL, 1, eta), ctx.edges(eta_ed), ctx.field(S.CELL, 1, temp),
                        ctx.edges(temp_ed), 0.37)
        got = [f.download() for f in force]
    assert bit_equal(got, want)


def test_stochastic_m_fluxdiv_module_noise_matches_explicit():
    dim, n = 3, (8, 8, 8)
    shape = (8, 8, 8)
    with S.Context(dim, n, 0.5, stoch_stress_form=0) as ctx:
        S.init_m_stochastic(ctx, 2)
        S.fill_m_stochastic(ctx, 99)
        ones = np.ones(shape)
        eta, ed = ctx.field(S.CELL, 1, ones), ctx.edges([ones] * 3)
        f1 = ctx.faces([np.zeros(shape)] * 3)
        S.stochastic_m_fluxdiv(ctx, f1, eta, ed, eta, ed, dt=0.1, weights=[1.0, 0.0], variance_coef_mom=1.0, k_B=1.0)
        got = [f.download() for f in f1]
        # the same stream through the explicit-noise entry point
        N = 8 ** 3
        z = S.fill_normal_host(99, 2 * (3 + 6) * N)[: (3 + 6) * N]
        Wcc = z[: 3 * N].reshape((3,) + shape)
        Wed = [z[3 * N + 2 * N * e: 3 * N + 2 * N * (e + 1)].reshape((2,) + shape) for e in range(3)]
        f2 = ctx.faces([np.zeros(shape)] * 3)
        variance = np.sqrt(1.0 * 2.0 * 1.0 / (0.5 ** 3 * 0.1))
        S.stoch_m_force(ctx, f2, ctx.field(S.CELL, 3, Wcc), ctx.edges(Wed, ncomp=2), eta, ed, eta, ed, variance)
        want = [f.download() for f in f2]
    assert bit_equal(got, want)


@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (3, (8, 16, 8))])
@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("width", [0, 1, 2, 4])
def test_stochastic_m_fluxdiv_forms_and_filtering_bit_exact(dim, n, form, width):
    """fill_m_stochastic (host MT19937 stream) in the non-symmetric and the
    symmetric stress form, two weighted stages, variable eta and T, every
    filtering width: the force equals the oracle's fill and
    stochastic_m_fluxdiv bit for bit."""
    shape = tuple(reversed(n))
    ne = 1 if dim == 2 else 3
    rng = np.random.default_rng(17)
    eta = 1.0 + rng.random(shape)
    temp = 1.0 + rng.random(shape)
    eta_ed = [P.avg_cc_to_node_2d(eta)] if dim == 2 else P.avg_cc_to_edge_3d(eta)
    temp_ed = [P.avg_cc_to_node_2d(temp)] if dim == 2 else P.avg_cc_to_edge_3d(temp)
    f0 = _rand_faces(dim, shape, 19)
    w = [0.7, -0.4]
    dt, dx = 0.05, 0.5
    variance = np.sqrt(1.3 * 2.0 * 0.8 / (dx ** dim * dt))
    stages = O.fill_m_stochastic(123, dim, shape, 2, symmetric=bool(form))
    # saxpy of the stages from zero (stochastic_m_fluxdiv.f90:114-123)
    Wc = 0.0 + w[0] * stages[0][0]
    Wc = Wc + w[1] * stages[1][0]
    We = [(0.0 + w[0] * stages[0][1][e]) + w[1] * stages[1][1][e] for e in range(ne)]
    g = O.geom(dim, n, dx)
    want = O.stochastic_m_fluxdiv(g, Wc, We, eta, eta_ed, temp, temp_ed, variance, f0, filtering_width=width)
    with S.Context(dim, n, dx, stoch_stress_form=form, filtering_width=width) as ctx:
        S.init_m_stochastic(ctx, 2)
        S.fill_m_stochastic(ctx, 123)
        force = ctx.faces(f0)
        S.stochastic_m_fluxdiv(ctx, force, ctx.field(S.CELL, 1, eta), ctx.edges(eta_ed), ctx.field(S.CELL, 1, temp),
                               ctx.edges(temp_ed), dt=dt, weights=w, variance_coef_mom=1.3, k_B=0.8)
        got = [f.download() for f in force]
    assert bit_equal(got, want)


def test_filtering_width_is_checked():
    with pytest.raises(S.LmmError):
        S.Context(3, (8, 8, 8), 0.5, filtering_width=3)


@pytest.mark.parametrize("dim,n,ns", [(2, (16, 8), 3), (3, (8, 8, 4), 4)])
def test_diffusive_mass_fluxdiv_bit_exact(dim, n, ns):
    g = O.geom(dim, n, 0.25)
    shape = tuple(reversed(n))
    rng = np.random.default_rng(9)
    x = rng.random((ns,) + shape)
    rw = rng.standard_normal((ns * ns,) + shape)
    gm = rng.standard_normal((ns * ns,) + shape)
    ft0 = [rng.standard_normal((ns,) + shape) for _ in range(dim)]
    rho = rng.random((ns,) + shape)
    rt = rho.sum(axis=0)
    want_div, want_ft = O.diffusive_mass_fluxdiv(g, ns, rho, rt, x, rw, gm, ft0, correct_flux=True)
    with S.Context(dim, n, 0.25) as ctx:
        ft = [ctx.field(S.FACE_X + d, ns, ft0[d]) for d in range(dim)]
        div = ctx.field(S.CELL, ns)
        S.diffusive_mass_fluxdiv(ctx, ns, ctx.field(S.CELL, ns, rho), ctx.field(S.CELL, 1, rt),
                                 ctx.field(S.CELL, ns, x), ctx.field(S.CELL, ns * ns, rw),
                                 ctx.field(S.CELL, ns * ns, gm), div, None, None, None, ft)
        got_div = div.download()
        got_ft = [f.download() for f in ft]
    assert bit_equal(got_div, want_div)
    assert bit_equal(got_ft, want_ft)


@pytest.mark.parametrize("dim", [2, 3])
def test_average_cc_bit_exact(dim):
    n = (8, 4, 8)[:dim]
    g = O.geom(dim, n)
    cc = np.random.default_rng(3).random(tuple(reversed(n)))
    with S.Context(dim, n) as ctx:
        c = ctx.field(S.CELL, 1, cc)
        fa = ctx.faces()
        S.average_cc_to_face(ctx, c, fa)
        assert bit_equal([f.download() for f in fa], O.average_cc_to_face(g, cc))
        ed = ctx.edges()
        if dim == 2:
            S.average_cc_to_node(ctx, c, ed[0])
            assert bit_equal(ed[0].download(), O.average_cc_to_node(g, cc))
        else:
            S.average_cc_to_edge(ctx, c, ed)
            assert bit_equal([e.download() for e in ed], O.average_cc_to_edge(g, cc))


def test_fab_roundtrip():
    """Reference fab layout (lo-ng, hi+1 faces, ghosts) in and out."""
    n = (8, 4, 4)
    ng = 1
    with S.Context(3, n) as ctx:
        f = ctx.field(S.FACE_X)
        shape = (n[0] + 1 + 2 * ng, n[1] + 2 * ng, n[2] + 2 * ng)
        dd = np.random.default_rng(1).standard_normal((n[2], n[1], n[0]))
        f.upload(dd)
        fab = f.download_fab(shape, 0, (0, 0, 0), (n[0], n[1] - 1, n[2] - 1), ng)
        # valid hi+1 face equals lo face, ghosts are periodic images
        assert np.array_equal(fab[ng + n[0], :, :], fab[ng, :, :])
        assert np.array_equal(fab[0, :, :], fab[n[0] - 1 + ng, :, :])
        g2 = ctx.field(S.FACE_X)
        g2.upload_fab(fab, 0, (0, 0, 0), (n[0], n[1] - 1, n[2] - 1), ng)
        assert np.array_equal(g2.download(), dd)


def test_host_noise_generator_matches_oracle():
    want = O.fill_normal(12345, 5000)
    got = S.fill_normal_host(12345, 5000)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("visc", [-1, 1, -2])
def test_fgmres_fp32_vectors_used_directly_equal_widened(visc, monkeypatch):
    """The flexible GMRES applies the FP64 operator to its FP32 preconditioned
    vectors and adds them to x without widening them into FP64 buffers first
    (LMM_Z32=0 widens first).  Widening is exact, so both give the same bits;
    visc_type -2 has no direct kernel and always widens."""
    pb = P.small_problem(3, (64, 32, 32), visc_type=visc, seed=89, contrast=30.0)
    b_u = P.rhs_from_noise_host(pb)
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("LMM_Z32", mode)
        with S.Context(3, pb.n, pb.dx, visc_type=visc, precision=S.FP32_MG) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            _, info = dp.solve()
            out.append((info["total_iter"], dp.solution()))
    assert out[0][0] == out[1][0]
    assert bit_equal(out[0][1][0], out[1][1][0]) and bit_equal(out[0][1][1], out[1][1][1])


@pytest.mark.parametrize("dim,n,prec", [(3, (64, 32, 32), "fp64"), (3, (64, 32, 32), "fp32_mg"),
                                        (2, (128, 64), "fp64"), (2, (128, 64), "fp32_mg")])
def test_fused_krylov_equals_separate_passes(dim, n, prec, monkeypatch):
    """GMRES fuses each modified Gram-Schmidt axpy with the next dot (or the
    norm) and the basis scaling with the FP32 conversion of the next
    preconditioner input; LMM_KRYLOV_FUSED=0 keeps the separate passes.  The
    arithmetic and the summation order are the same, so the solutions, the
    iteration counts and the residual estimates agree bit for bit."""
    pb = P.small_problem(dim, n, visc_type=-1, seed=97, contrast=30.0)
    b_u = P.rhs_from_noise_host(pb)
    precision = {"fp64": S.FP64, "fp32_mg": S.FP32_MG}[prec]
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("LMM_KRYLOV_FUSED", mode)
        with S.Context(dim, pb.n, pb.dx, visc_type=-1, precision=precision, gmres_max_inner=4) as ctx:
            dp = S.DeviceProblem(ctx, pb, rhs=b_u)
            _, info = dp.solve()
            out.append((info["total_iter"], info["norm_resid"], dp.solution()))
    assert out[0][0] == out[1][0] and out[0][0] > 4
    assert out[0][1] == out[1][1]
    assert bit_equal(out[0][2][0], out[1][2][0]) and bit_equal(out[0][2][1], out[1][2][1])


@pytest.mark.parametrize("dim,n", [(2, (32, 32)), (3, (64, 16, 16))])
@pytest.mark.parametrize("vcycles,tol", [(4, 0.5), (4, 1e-30), (3, 1e-3)])
def test_apply_precon_ignores_stale_output(dim, n, vcycles, tol):
    """apply_precon zeroes x_u before its staggered V-cycles
    (apply_precon.f90:88-90): the initial residual of stag_mg_solver's
    stopping test (stag_mg_solver.f90:261-310) is rhs, whatever x_u held.
    x_u is filled with random values first; several V-cycles with a loose,
    an unreachable and a middling tolerance must give the oracle's bits."""
    pb = P.small_problem(dim, n, visc_type=-1, seed=37)
    kw = dict(visc_type=-1, stag_mg_max_vcycles=vcycles, stag_mg_rel_tol=tol, mg_max_vcycles=vcycles,
              mg_rel_tol=tol)
    g, p, cf = oracle_args(pb, **kw)
    bu = _rand_faces(dim, pb.shape, 33)
    bp = np.random.default_rng(34).standard_normal(pb.shape)
    bp -= bp.mean()
    want_u, want_p = O.apply_precon(g, p, cf, 1.0, bu, bp)
    with S.Context(dim, pb.n, pb.dx, **kw) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=bu)
        b_p = ctx.field(S.CELL, 1, bp)
        for f, a in zip(dp.x_u, _rand_faces(dim, pb.shape, 35)):
            f.upload(a)
        dp.x_p.upload(np.random.default_rng(36).standard_normal(pb.shape))
        S.apply_precon(ctx, dp.b_u, b_p, dp.x_u, dp.x_p, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, 1.0)
        got_u, got_p = dp.solution()
    assert bit_equal(got_u, want_u)
    assert relerr(got_p, want_p) < 1e-13


def test_apply_precon_zero_rhs_clears_stale_output():
    """b = 0: every V-cycle returns at once (stag_mg_solver.f90:276-306) and
    the result is zero even where x_u held other values."""
    pb = P.small_problem(3, (16, 16, 16), visc_type=-1, seed=38)
    with S.Context(3, pb.n, pb.dx, visc_type=-1, stag_mg_max_vcycles=3, mg_max_vcycles=3) as ctx:
        dp = S.DeviceProblem(ctx, pb, rhs=[np.zeros(pb.shape)] * 3)
        b_p = ctx.field(S.CELL, 1, np.zeros(pb.shape))
        for f in dp.x_u:
            f.setval(5.0)
        dp.x_p.setval(-2.0)
        S.apply_precon(ctx, dp.b_u, b_p, dp.x_u, dp.x_p, dp.alpha, dp.beta, dp.beta_ed, dp.gamma, 1.0)
        got_u, got_p = dp.solution()
    assert all(np.all(u == 0) for u in got_u) and np.all(got_p == 0)


@pytest.mark.parametrize("dim,n", [(2, (32, 32)), (3, (64, 16, 16))])
@pytest.mark.parametrize("tol", [0.9, 1e-2, 1e-30])
def test_macproject_stopping_test_bit_exact(dim, n, tol):
    """Up to 6 cell-centred V-cycles with the stopping test max|r| <=
    mg_rel_tol*max|rhs| (macproject.f90:402-425): the GPU stops after the
    same cycle as the oracle and gives its bits."""
    pb = P.small_problem(dim, n, seed=39)
    g, p, cf = oracle_args(pb, mg_max_vcycles=6, mg_rel_tol=tol)
    ainv = [1.0 / a for a in pb.alpha]
    rhs = np.random.default_rng(40).standard_normal(pb.shape)
    rhs -= rhs.mean()
    want = O.macproject(g, p, ainv, rhs)
    with S.Context(dim, pb.n, pb.dx, mg_max_vcycles=6, mg_rel_tol=tol) as ctx:
        ai = ctx.faces(ainv)
        r = ctx.field(S.CELL, 1, rhs)
        phi = ctx.field(S.CELL)
        phi.setval(7.0)
        S.macproject(ctx, phi, None, ai, r)
        got = phi.download()
    assert bit_equal(got, want), relerr(got, want)
