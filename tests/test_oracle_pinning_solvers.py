"""Pin the oracle's assembled solvers and the pieces the stencil tests leave open.

Each check is fixed by the mathematics or by the reference's loop structure,
not by the oracle's own formulas, and each is chosen so that a plausible
mistake fails it:

* the staggered V-cycle (stag_mg_solver.f90:323-811) reduces the residual at
  least tenfold per cycle on periodic constant-coefficient problems (the
  rate StokesPreconditioners.pdf p.13 reports for 2 smoothing sweeps), also
  with theta*alpha large, where coarse levels that miss theta fail; many
  cycles converge to the dense solve of the assembled operator;
* the projection preconditioner P1 (apply_precon.f90:103-195,434-460) with
  near-exact sub-solves equals the dense block formula
  x_u = A^-1 b_u - alpha^-1 G phi,  phi = (D alpha^-1 G)^+ (D A^-1 b_u + b_p),
  x_p = theta_alpha phi - c mac_rhs  (c = beta, 2 beta, 4/3 beta + gamma),
  and makes GMRES converge in one iteration on periodic constant-coefficient
  problems (paper p.13, Appendix B);
* average_cc_to_face / _node / _edge (convert_stag.f90:229-393,462-516,
  681-863) reproduce linear cell fields at the face, node and edge positions;
* the 12-point restriction (stag_mg_solver.f90:1133-1171) returns f(2I)
  times the transverse mean for f linear along the normal, which fixes the
  2I-1 / 2I+1 offsets;
* sum_umac_press (norm_inner_product.f90:579-721) against a brute-force walk
  over the reference's tiles;
* the stopping tests of both multigrids.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import relerr


def shape_of(dim, n):
    return tuple(reversed(n[:dim]))


def flat(u):
    return np.concatenate([np.ravel(a) for a in u])


def unflat(v, dim, n):
    shp = shape_of(dim, n)
    m = int(np.prod(shp))
    return [v[d * m:(d + 1) * m].reshape(shp).copy() for d in range(dim)]


def coefs(dim, n, seed, const=False, gamma=0.0):
    rng = np.random.default_rng(seed)
    shp = shape_of(dim, n)
    ne = 1 if dim == 2 else 3
    if const:
        a0, b0 = 0.5 + rng.random(), 0.7 + rng.random()
        return ([np.full(shp, a0) for _ in range(dim)], np.full(shp, b0), [np.full(shp, b0) for _ in range(ne)],
                np.full(shp, gamma))
    alpha = [1.0 + rng.random(shp) for _ in range(dim)]
    beta = 0.5 + rng.random(shp)
    bed = [0.5 + rng.random(shp) for _ in range(ne)]
    return alpha, beta, bed, gamma * (0.5 + rng.random(shp))


def assemble_velocity(g, visc, cf, theta_alpha, dim, n):
    m = dim * int(np.prod(shape_of(dim, n)))
    A = np.zeros((m, m))
    for q in range(m):
        e = np.zeros(m)
        e[q] = 1.0
        A[:, q] = flat(O.stag_applyop(g, visc, cf, theta_alpha, unflat(e, dim, n)))
    return A


def assemble_grad_div(g, dim, n):
    shp = shape_of(dim, n)
    m = int(np.prod(shp))
    G = np.zeros((dim * m, m))
    for q in range(m):
        e = np.zeros(m)
        e[q] = 1.0
        G[:, q] = flat(O.compute_grad(g, e.reshape(shp)))
    D = np.zeros((m, dim * m))
    for q in range(dim * m):
        e = np.zeros(dim * m)
        e[q] = 1.0
        D[:, q] = O.compute_div(g, unflat(e, dim, n)).ravel()
    return G, D


def stag_residual(g, visc, cf, theta_alpha, phi, rhs):
    Lp = O.stag_applyop(g, visc, cf, theta_alpha, phi)
    return [r - l for r, l in zip(rhs, Lp)]


# ------------------------------------------------------------ staggered V-cycle

@pytest.mark.parametrize("dim,n,visc,theta,dx,meanfree,bound", [
    (2, (64, 64), 1, 1.0, 1 / 8, False, 0.1), (2, (64, 64), -1, 50.0, 1 / 8, False, 0.1),
    (2, (64, 64), 1, 1.0, 1 / 64, True, 0.1), (2, (64, 32), -2, 1.0, 1 / 8, False, 0.1),
    (3, (32, 32, 32), 1, 1.0, 1 / 8, False, 0.1), (3, (32, 32, 32), -1, 50.0, 1 / 8, False, 0.1),
    (3, (16, 32, 16), -3, 1.0, 1 / 4, False, 0.15),
])
def test_stag_vcycle_reduces_residual_tenfold(dim, n, visc, theta, dx, meanfree, bound):
    """One V-cycle per call with the previous result as guess; every cycle
    must cut the residual at least tenfold (paper p.13: >= 10x per cycle with
    2 smoothing sweeps; the full-stress form with a bulk term, whose
    Gauss-Seidel couples the components, within 0.15).  A wrong residual
    sign diverges, coarse levels without theta*alpha (theta = 50), shifted
    transfer stencils or a missing bottom solve stall.

    On a periodic box the constant velocity modes have the eigenvalue
    theta*alpha alone; where that is tiny against beta/dx^2 only the 8
    bottom sweeps reduce them and no V-cycle reaches 10x (the paper's figure
    is for no-slip walls, which have no such modes).  The cases keep
    theta*alpha*dx^2/beta of order 1e-2 or larger, or take a right-hand side
    without those modes."""
    g = O.geom(dim, n, dx)
    c = coefs(dim, n, 3, const=True, gamma=0.4 if abs(visc) == 3 else 0.0)
    cf = O.coef(*c)
    p = O.params(visc_type=visc, stag_mg_max_vcycles=1)
    rng = np.random.default_rng(4)
    rhs = [rng.standard_normal(shape_of(dim, n)) for _ in range(dim)]
    if meanfree:
        rhs = [r - r.mean() for r in rhs]
    phi = [np.zeros_like(r) for r in rhs]
    r0 = np.linalg.norm(flat(rhs))
    for cyc in range(4):
        phi = O.stag_mg_solver(g, p, cf, theta, rhs, phi=phi)
        r1 = np.linalg.norm(flat(stag_residual(g, visc, cf, theta, phi, rhs)))
        assert r1 < bound * r0, f"cycle {cyc}: residual ratio {r1 / r0:.3f}"
        r0 = r1


@pytest.mark.parametrize("dim,n,visc,theta", [(2, (16, 16), -1, 1.0), (2, (16, 8), -3, 0.5),
                                              (3, (8, 8, 8), -1, 1.0), (3, (8, 8, 4), -2, 2.0)])
def test_stag_vcycles_converge_to_dense_solution(dim, n, visc, theta):
    g = O.geom(dim, n, 0.5)
    c = coefs(dim, n, 5, gamma=0.3 if abs(visc) == 3 else 0.0)
    cf = O.coef(*c)
    A = assemble_velocity(g, visc, cf, theta, dim, n)
    rng = np.random.default_rng(6)
    rhs = [rng.standard_normal(shape_of(dim, n)) for _ in range(dim)]
    want = np.linalg.solve(A, flat(rhs))
    p = O.params(visc_type=visc, stag_mg_max_vcycles=80, stag_mg_rel_tol=1e-15)
    got = O.stag_mg_solver(g, p, cf, theta, rhs)
    assert relerr(flat(got), want) < 1e-11


def test_stag_stopping_test_counts_cycles():
    """stag_mg_rel_tol 0.9 stops after the first cycle (any cycle reduces
    max|r| below 0.9 of the initial), 1e-30 never stops early."""
    dim, n = 2, (32, 32)
    g = O.geom(dim, n, 1.0 / 32)
    cf = O.coef(*coefs(dim, n, 7))
    rng = np.random.default_rng(8)
    rhs = [rng.standard_normal(shape_of(dim, n)) for _ in range(dim)]

    def run(k, tol):
        return O.stag_mg_solver(g, O.params(visc_type=-1, stag_mg_max_vcycles=k, stag_mg_rel_tol=tol), cf, 1.0, rhs)

    one = run(1, 1e-9)
    assert all(np.array_equal(a, b) for a, b in zip(run(5, 0.9), one))
    five = run(5, 1e-30)
    assert all(np.array_equal(a, b) for a, b in zip(run(5, 1e-30), five))
    assert not all(np.array_equal(a, b) for a, b in zip(five, one))


def test_macproject_stopping_test_counts_cycles():
    dim, n = 2, (32, 32)
    g = O.geom(dim, n, 0.5)
    rng = np.random.default_rng(9)
    ainv = [0.5 + rng.random(shape_of(dim, n)) for _ in range(dim)]
    rhs = rng.standard_normal(shape_of(dim, n))
    rhs -= rhs.mean()
    one = O.macproject(g, O.params(mg_max_vcycles=1), ainv, rhs)
    assert np.array_equal(O.macproject(g, O.params(mg_max_vcycles=6, mg_rel_tol=0.9), ainv, rhs), one)
    six = O.macproject(g, O.params(mg_max_vcycles=6, mg_rel_tol=1e-30), ainv, rhs)
    assert not np.array_equal(six, one)
    # a zero right-hand side passes the test before the first cycle: phi = 0
    z = O.macproject(g, O.params(mg_max_vcycles=3), ainv, np.zeros_like(rhs))
    assert np.all(z == 0.0)


# ------------------------------------------------------------ P1 preconditioner

def schur_constant(visc, beta, gamma):
    if abs(visc) == 1:
        return beta
    if abs(visc) == 2:
        return 2.0 * beta
    return 4.0 / 3.0 * beta + gamma


@pytest.mark.parametrize("dim,n,visc,theta", [(2, (8, 8), -1, 1.0), (2, (8, 8), -2, 0.5), (2, (8, 4), -3, 2.0),
                                              (3, (4, 4, 4), -1, 0.5), (3, (4, 4, 4), -3, 1.0)])
def test_apply_precon_matches_dense_block_formula(dim, n, visc, theta):
    g = O.geom(dim, n, 0.5)
    c = coefs(dim, n, 10, gamma=0.3 if abs(visc) == 3 else 0.0)
    alpha, beta, bed, gamma = c
    cf = O.coef(*c)
    shp = shape_of(dim, n)
    m = int(np.prod(shp))
    rng = np.random.default_rng(11)
    bu = [rng.standard_normal(shp) for _ in range(dim)]
    bp = rng.standard_normal(shp)
    bp -= bp.mean()
    p = O.params(visc_type=visc, stag_mg_max_vcycles=100, stag_mg_rel_tol=1e-15, mg_max_vcycles=100,
                 mg_rel_tol=1e-15)
    got_u, got_p = O.apply_precon(g, p, cf, theta, bu, bp)

    A = assemble_velocity(g, visc, cf, theta, dim, n)
    G, D = assemble_grad_div(g, dim, n)
    ainv = 1.0 / flat(alpha)  # 1/alpha without theta (apply_precon.f90:74-76)
    xs = np.linalg.solve(A, flat(bu))
    mac_rhs = D @ xs + bp.ravel()
    Lp = D @ (ainv[:, None] * G)
    phi = np.linalg.lstsq(Lp, mac_rhs, rcond=None)[0]
    want_u = xs - ainv * (G @ phi)
    want_p = theta * phi - schur_constant(visc, beta.ravel(), gamma.ravel()) * mac_rhs
    want_p -= want_p.mean()
    assert relerr(flat(got_u), want_u) < 1e-10
    assert relerr(got_p.ravel(), want_p) < 1e-10
    # the velocity it returns is discretely divergence-free up to -b_p
    assert np.max(np.abs(D @ flat(got_u) + bp.ravel())) < 1e-9 * np.max(np.abs(bp))


@pytest.mark.parametrize("dim,n,visc", [(2, (16, 16), 1), (2, (16, 16), 2), (2, (16, 8), 3), (3, (8, 8, 8), 1),
                                        (3, (8, 8, 8), -2)])
def test_p1_exact_subsolves_converge_in_one_iteration(dim, n, visc):
    """Periodic constant coefficients: P1 with exact sub-solves is the exact
    inverse, so GMRES stops after one iteration (paper p.13).  dx = 1/4
    keeps theta*alpha*dx^2/beta large enough that 100 V-cycles also converge
    the constant velocity modes, which only the bottom sweeps reduce."""
    g = O.geom(dim, n, 0.25)
    c = coefs(dim, n, 12, const=True, gamma=0.6 if abs(visc) == 3 else 0.0)
    cf = O.coef(*c)
    rng = np.random.default_rng(13)
    shp = shape_of(dim, n)
    bu = [rng.standard_normal(shp) for _ in range(dim)]
    bp = rng.standard_normal(shp)
    bp -= bp.mean()
    p = O.params(visc_type=visc, stag_mg_max_vcycles=100, stag_mg_rel_tol=1e-15, mg_max_vcycles=100,
                 mg_rel_tol=1e-15, gmres_rel_tol=1e-9)
    _, _, info = O.gmres(g, p, cf, 1.0, bu, bp)
    assert info["status"] == 0
    assert info["total_iter"] == 1
    # one V-cycle each (inexact sub-solves) needs several iterations
    _, _, info1 = O.gmres(g, O.params(visc_type=visc, gmres_rel_tol=1e-9), cf, 1.0, bu, bp)
    assert info1["total_iter"] > 1


# ------------------------------------------------------------ coefficient averaging

def linear_cells(dim, n, dx, coef):
    """a + sum_d b_d x_d at the cell centres x_d = (i_d + 1/2) dx."""
    shp = shape_of(dim, n)
    idx = np.indices(shp)  # axis order (z, y, x)
    v = np.full(shp, coef[0])
    for d in range(dim):
        v = v + coef[1 + d] * (idx[dim - 1 - d] + 0.5) * dx
    return v


def linear_at(dim, shp, dx, coef, nodal):
    """The same linear function at positions nodal in the listed directions."""
    idx = np.indices(shp)
    v = np.full(shp, coef[0])
    for d in range(dim):
        off = 0.0 if d in nodal else 0.5
        v = v + coef[1 + d] * (idx[dim - 1 - d] + off) * dx
    return v


def interior(dim, shp, nodal):
    """Positions whose averaging stencil does not wrap: index >= 1 in nodal directions."""
    sl = [slice(None)] * dim
    for d in nodal:
        sl[dim - 1 - d] = slice(1, None)
    return tuple(sl)


@pytest.mark.parametrize("dim,n", [(2, (8, 6)), (3, (6, 4, 8))])
def test_average_cc_to_face_reproduces_linear_fields(dim, n):
    dx = 0.25
    g = O.geom(dim, n, dx)
    coef = [0.3, 1.7, -2.2, 0.9][: dim + 1]
    cc = linear_cells(dim, n, dx, coef)
    faces = O.average_cc_to_face(g, cc)
    shp = shape_of(dim, n)
    for d in range(dim):
        want = linear_at(dim, shp, dx, coef, {d})
        sl = interior(dim, shp, {d})
        assert np.allclose(faces[d][sl], want[sl], rtol=0, atol=1e-13)


def test_average_cc_to_node_reproduces_linear_fields():
    dim, n, dx = 2, (8, 6), 0.5
    g = O.geom(dim, n, dx)
    coef = [1.1, -0.7, 2.3]
    node = O.average_cc_to_node(g, linear_cells(dim, n, dx, coef))
    shp = shape_of(dim, n)
    want = linear_at(dim, shp, dx, coef, {0, 1})
    sl = interior(dim, shp, {0, 1})
    assert np.allclose(node[sl], want[sl], rtol=0, atol=1e-13)


def test_average_cc_to_edge_reproduces_linear_fields():
    dim, n, dx = 3, (6, 4, 8), 0.5
    g = O.geom(dim, n, dx)
    coef = [0.4, 1.3, -0.6, 2.1]
    edges = O.average_cc_to_edge(g, linear_cells(dim, n, dx, coef))
    shp = shape_of(dim, n)
    for e, nodal in enumerate(({0, 1}, {0, 2}, {1, 2})):  # xy, xz, yz
        want = linear_at(dim, shp, dx, coef, nodal)
        sl = interior(dim, shp, nodal)
        assert np.allclose(edges[e][sl], want[sl], rtol=0, atol=1e-13), e


# ------------------------------------------------------------ 12-point restriction

@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (3, (8, 8, 4))])
def test_full_restriction_on_fields_linear_along_the_normal(dim, n):
    g = O.geom(dim, n, 1.0)
    shp = shape_of(dim, n)
    nc = [v // 2 for v in n[:dim]]
    cshp = shape_of(dim, nc)
    rng = np.random.default_rng(14)
    idx = np.indices(shp)
    u, t = [], []
    for d in range(dim):
        f = 0.8 + 1.9 * idx[dim - 1 - d]          # linear along the normal
        td = rng.standard_normal(shp)
        td = np.mean(td, axis=dim - 1 - d, keepdims=True) + 0 * td  # constant along the normal
        u.append(f * td)
        t.append(td)
    got = O.stag_restriction(g, u, simple=False)
    cidx = np.indices(cshp)
    for d in range(dim):
        ax = dim - 1 - d
        # transverse mean of the 2^(dim-1) children of each coarse face
        tm = t[d]
        for e in range(dim):
            if e != d:
                a = dim - 1 - e
                tm = 0.5 * (np.take(tm, range(0, shp[a], 2), axis=a) + np.take(tm, range(1, shp[a], 2), axis=a))
        tm = np.take(tm, range(0, shp[ax], 2), axis=ax)
        want = (0.8 + 1.9 * 2 * cidx[ax]) * tm
        sl = [slice(None)] * dim
        sl[ax] = slice(1, None)  # coarse I = 0 reads the periodic image n-1 for 2I-1
        assert np.allclose(got[d][tuple(sl)], want[tuple(sl)], rtol=1e-13, atol=1e-13), d


# ------------------------------------------------------------ sum_umac_press

TILE = (1024000, 8, 8)  # BoxLib's default tile size (F_BaseLib layout)


def brute_sum_umac(dim, n, u_faces):
    """The reference's loops (norm_inner_product.f90:640-721) on its own face
    layout (faces 0..n, face n the periodic copy of face 0), tile by tile."""
    sums = []
    for d in range(dim):
        ref = np.concatenate([u_faces[d], np.take(u_faces[d], [0], axis=dim - 1 - d)], axis=dim - 1 - d)
        tiles = [list(range(0, n[e], TILE[e])) for e in range(dim)]
        s = 0.0
        for lo in itertools.product(*tiles):
            tlo = list(lo)
            thi = [min(lo[e] + TILE[e], n[e]) - 1 for e in range(dim)]
            nlo, nhi = list(tlo), list(thi)
            if thi[d] == n[d] - 1:
                nhi[d] = n[d]          # the nodal tile box reaches hi+1 at the box end
            rng_ = [range(nlo[e], nhi[e] + 1) for e in range(dim)]
            for pt in itertools.product(*rng_):
                s += ref[tuple(reversed(pt))]
            if nlo[d] == 0 and nhi[d] == n[d]:
                tr = [rng_[e] for e in range(dim)]
                for pt in itertools.product(*[tr[e] if e != d else [0] for e in range(dim)]):
                    q = list(pt)
                    s += 0.5 * ref[tuple(reversed(q))]
                    q[d] = n[d]
                    s += 0.5 * ref[tuple(reversed(q))]
        sums.append(s)
    return sums


@pytest.mark.parametrize("dim,n", [(2, (6, 4)), (2, (12, 20)), (3, (6, 4, 12)), (3, (4, 10, 6))])
def test_sum_umac_press_brute_force(dim, n):
    g = O.geom(dim, n, 1.0)
    rng = np.random.default_rng(15)
    shp = shape_of(dim, n)
    u = [rng.standard_normal(shp) for _ in range(dim)]
    p = rng.standard_normal(shp)
    av_p, av_u = O.sum_umac_press(g, p, u)
    vol = float(np.prod(shp))
    assert av_p == pytest.approx(float(p.sum()) / vol, rel=1e-13, abs=1e-15)
    for d, s in enumerate(brute_sum_umac(dim, n, u)):
        assert av_u[d] == pytest.approx(s / vol, rel=1e-12, abs=1e-14), d
