"""Pin the CPU oracle against what the mathematics fixes.

None of these checks retypes an oracle formula: each one uses a property of
the discrete operators that a plausible mistake (dropped term, wrong sign,
shifted index, transposed operand) would break:

* constant-coefficient forms of the variable-coefficient operators coincide
  (stag_applyop.f90: visc_type -1/-2/-3 against 1/2/3);
* the Fourier symbol of the constant-coefficient Laplacian;
* the MAC-grid identities div(curl) = 0 and grad/div/Laplacian commuting,
  which fix how the full-stress forms act on solenoidal and gradient fields;
* symmetry and positive definiteness of the assembled operator, zero row sum
  without the alpha term;
* the colour-by-colour Gauss-Seidel and the damped Jacobi smoothers against
  the same sweeps done with the assembled matrix;
* restriction preserves constants and commutes with the divergence, the
  12-point form reduces to the face average on fields constant along the
  normal, prolongation is the tensor product of 1-D linear interpolations;
* inner products against brute-force sums (norm_inner_product.f90:115-168
  weights the box-boundary faces by 1.5);
* Givens least squares against numpy's lstsq; the CC multigrid and GMRES
  against dense direct solves; flux divergences against conservation and
  the species-sum constraint.
"""
import numpy as np
import pytest

from lowmachmixtures_amd import problems as P
from oracle import oracle as O
from tests.helpers import relerr


def shape_of(dim, n):
    return tuple(reversed(n[:dim]))


def roll(a, shift, d, dim):
    """Shift along direction d (0 = x) of a (.., y, x) array: out[i] = a[i - shift]."""
    return np.roll(a, shift, axis=dim - 1 - d)


def coefs(dim, n, seed, const=False, gamma=0.0):
    rng = np.random.default_rng(seed)
    shp = shape_of(dim, n)
    alpha = [1.0 + rng.random(shp) for _ in range(dim)]
    if const:
        b0 = 0.7 + rng.random()
        beta = np.full(shp, b0)
        bed = [np.full(shp, b0) for _ in range(1 if dim == 2 else 3)]
        gam = np.full(shp, gamma)
    else:
        beta = 0.5 + rng.random(shp)
        bed = [0.5 + rng.random(shp) for _ in range(1 if dim == 2 else 3)]
        gam = gamma * (0.5 + rng.random(shp))
    return alpha, beta, bed, gam


def cf_of(c):
    return O.coef(*c)


def rand_faces(dim, n, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape_of(dim, n)) for _ in range(dim)]


def flat(u):
    return np.concatenate([np.ravel(a) for a in u])


def unflat(v, dim, n):
    shp = shape_of(dim, n)
    m = int(np.prod(shp))
    return [v[d * m:(d + 1) * m].reshape(shp).copy() for d in range(dim)]


def assemble(g, visc, cf, theta_alpha, dim, n):
    m = dim * int(np.prod(shape_of(dim, n)))
    A = np.zeros((m, m))
    for q in range(m):
        e = np.zeros(m)
        e[q] = 1.0
        A[:, q] = flat(O.stag_applyop(g, visc, cf, theta_alpha, unflat(e, dim, n)))
    return A


GEOMS = [(2, (8, 8), 1.0), (2, (6, 4), 0.5), (3, (4, 4, 4), 0.5)]


# ---------------------------------------------------------------- levels

@pytest.mark.parametrize("n,minwidth,want", [
    ((256, 256, 1), 2, 8), ((128, 128, 128), 2, 7), ((64, 32, 1), 2, 5), ((12, 12, 1), 2, 3),
    ((2, 2, 1), 2, 1), ((256, 256, 256), 4, 7), ((24, 24, 24), 2, 4),
])
def test_compute_nlevs(n, minwidth, want):
    # n = minwidth * 2^k coarsens k times: k + 1 levels; an odd extent stops coarsening
    dim = 2 if n[2] == 1 else 3
    assert O.compute_nlevs_mg(O.geom(dim, n[:dim]), minwidth) == want


# ---------------------------------------------------------------- applyop

@pytest.mark.parametrize("dim,n,dx", GEOMS)
@pytest.mark.parametrize("vt", [1, 2, 3])
@pytest.mark.parametrize("color", [0, 1, 2])
def test_constant_coefficient_forms_agree(dim, n, dx, vt, color):
    g = O.geom(dim, n, dx)
    c = coefs(dim, n, 11, const=True, gamma=0.3 if vt == 3 else 0.0)
    cf = cf_of(c)
    u = rand_faces(dim, n, 5)
    L0 = [np.full_like(a, 7.0) for a in u]  # colour sweeps leave other points alone
    a = O.stag_applyop(g, vt, cf, 0.8, u, color=color, Lphi=L0)
    b = O.stag_applyop(g, -vt, cf, 0.8, u, color=color, Lphi=L0)
    assert relerr(a, b) < 1e-14


@pytest.mark.parametrize("dim,n,dx", GEOMS)
def test_laplacian_fourier_symbol(dim, n, dx):
    g = O.geom(dim, n, dx)
    c = coefs(dim, n, 2, const=True)
    alpha0 = 1.3
    c = ([np.full_like(a, alpha0) for a in c[0]],) + c[1:]
    b0 = c[1].flat[0]
    idx = np.meshgrid(*[np.arange(s) for s in shape_of(dim, n)], indexing="ij")
    kvec = [1, 2, 1][:dim]  # wavenumbers (x, y, z)
    theta = [2 * np.pi * kvec[d] / n[d] for d in range(dim)]
    phase = sum(theta[d] * idx[dim - 1 - d] for d in range(dim))
    u = [np.cos(phase + 0.3 * d) for d in range(dim)]
    lam = 0.5 * alpha0 + b0 * sum(2 - 2 * np.cos(t) for t in theta) / dx ** 2
    Lu = O.stag_applyop(g, 1, cf_of(c), 0.5, u)
    assert relerr(Lu, [lam * a for a in u]) < 1e-13


@pytest.mark.parametrize("dim,n,dx", GEOMS)
@pytest.mark.parametrize("visc", [-1, 1, -2])
def test_diffusive_m_fluxdiv_adds_the_viscous_laplacian(dim, n, dx, visc):
    """diffusive_m_fluxdiv.f90:16-73 on a Fourier mode with constant eta:
    m_update gains +eta*Lap_h(u) = -eta*lambda_h*u (lambda_h the symbol of
    -Lap_h), whatever alpha the caller's problem has.  For visc_type 2 the
    mode is a gradient field (u = grad p), on which the full stress acts as
    2*eta*Lap_h."""
    g = O.geom(dim, n, dx)
    c = coefs(dim, n, 4, const=True)
    b0 = c[1].flat[0]
    idx = np.meshgrid(*[np.arange(s) for s in shape_of(dim, n)], indexing="ij")
    kvec = [1, 1, 1][:dim]
    theta = [2 * np.pi * kvec[d] / n[d] for d in range(dim)]
    lam = sum(2 - 2 * np.cos(t) for t in theta) / dx ** 2
    if abs(visc) == 2:
        phase = sum(theta[d] * idx[dim - 1 - d] for d in range(dim))
        p = np.cos(phase)
        u = mac_grad(p, dim, dx)
        factor = 2.0
    else:
        phase = sum(theta[d] * idx[dim - 1 - d] for d in range(dim))
        u = [np.cos(phase + 0.4 * d) for d in range(dim)]
        factor = 1.0
    m0 = rand_faces(dim, n, 9)
    got = O.diffusive_m_fluxdiv(g, visc, m0, u, c[1], c[2], c[3])
    want = [m - factor * b0 * lam * a for m, a in zip(m0, u)]
    assert relerr(got, want) < 1e-12


def mac_grad(p, dim, dx):
    return [(p - roll(p, 1, d, dim)) / dx for d in range(dim)]


def solenoidal(dim, n, dx, seed):
    """u = curl(psi e_z) with psi on nodes (lower-left corner of each cell)."""
    psi = np.random.default_rng(seed).standard_normal(shape_of(dim, n))
    ux = (roll(psi, -1, 1, dim) - psi) / dx
    uy = -(roll(psi, -1, 0, dim) - psi) / dx
    return [ux, uy] + ([np.zeros_like(psi)] if dim == 3 else [])


@pytest.mark.parametrize("dim,n,dx", GEOMS)
@pytest.mark.parametrize("sign", [1, -1])
def test_full_stress_on_solenoidal_and_gradient_fields(dim, n, dx, sign):
    g = O.geom(dim, n, dx)
    gam = 0.4
    c = coefs(dim, n, 4, const=True, gamma=gam)
    cf = cf_of(c)
    b0 = c[1].flat[0]
    u = solenoidal(dim, n, dx, 9)
    assert np.max(np.abs(O.compute_div(g, u))) < 1e-12
    L1 = O.stag_applyop(g, sign * 1, cf, 0.0, u)
    for vt in (2, 3):
        assert relerr(O.stag_applyop(g, sign * vt, cf, 0.0, u), L1) < 1e-13
    p = np.random.default_rng(3).standard_normal(shape_of(dim, n))
    gp = mac_grad(p, dim, dx)
    L1 = O.stag_applyop(g, sign * 1, cf, 0.0, gp)
    assert relerr(O.stag_applyop(g, sign * 2, cf, 0.0, gp), [2 * a for a in L1]) < 1e-12
    fac = 4.0 / 3.0 + gam / b0
    assert relerr(O.stag_applyop(g, sign * 3, cf, 0.0, gp), [fac * a for a in L1]) < 1e-12


@pytest.mark.parametrize("dim,n,dx", GEOMS)
@pytest.mark.parametrize("vt", [-1, -2, -3, 1, 2, 3])
def test_operator_symmetric_positive_definite(dim, n, dx, vt):
    g = O.geom(dim, n, dx)
    c = coefs(dim, n, 21, const=vt > 0, gamma=0.2 if abs(vt) == 3 else 0.0)
    A = assemble(g, vt, cf_of(c), 0.9, dim, n)
    assert np.max(np.abs(A - A.T)) <= 1e-12 * np.max(np.abs(A))
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    # the alpha term is the diagonal alpha * theta_alpha
    A0 = assemble(g, vt, cf_of(c), 0.0, dim, n)
    assert np.allclose(A - A0, np.diag(0.9 * flat(c[0])), rtol=0, atol=1e-12 * np.max(np.abs(A)))


@pytest.mark.parametrize("dim,n,dx", GEOMS)
@pytest.mark.parametrize("vt", [-1, -2, -3, 1, 2, 3])
def test_constants_in_null_space_without_alpha(dim, n, dx, vt):
    g = O.geom(dim, n, dx)
    c = coefs(dim, n, 8, const=vt > 0, gamma=0.3)
    u = [np.full(shape_of(dim, n), 1.0 + d) for d in range(dim)]
    Lu = O.stag_applyop(g, vt, cf_of(c), 0.0, u)
    assert np.max(np.abs(flat(Lu))) < 1e-12


# ---------------------------------------------------------------- smoothers

def colour_points(dim, n, color):
    """Component and mask of colour c: component ceil(c/2), (i+j+k) % 2 == (c+1) % 2."""
    d = (color - 1) // 2
    idx = np.indices(shape_of(dim, n)).sum(axis=0)
    return d, (idx % 2) == ((color + 1) % 2)


@pytest.mark.parametrize("dim,n,dx", [(2, (8, 8), 1.0), (3, (4, 4, 4), 0.5)])
@pytest.mark.parametrize("vt", [-1, -2, -3, 1, 3])
def test_gauss_seidel_colour_sweep_matches_assembled(dim, n, dx, vt):
    g = O.geom(dim, n, dx)
    c = coefs(dim, n, 13, const=vt > 0, gamma=0.25)
    cf = cf_of(c)
    A = assemble(g, vt, cf, 1.0, dim, n)
    phi = rand_faces(dim, n, 1)
    rhs = rand_faces(dim, n, 2)
    p = O.params(visc_type=vt, stag_mg_smoother=1, stag_mg_omega=1.0)
    got = O.stag_mg_smooth(g, p, cf, 1.0, phi, rhs, 2)
    x = flat(phi)
    b = flat(rhs)
    dg = np.diag(A)
    m = int(np.prod(shape_of(dim, n)))
    for _ in range(2):
        for color in range(1, 2 * dim + 1):
            d, mask = colour_points(dim, n, color)
            sel = np.zeros(dim * m, bool)
            sel[d * m:(d + 1) * m] = mask.ravel()
            r = b - A @ x
            x[sel] += r[sel] / dg[sel]
    assert relerr(got, unflat(x, dim, n)) < 1e-12
    # the last colour's points have zero residual after the sweep
    d, mask = colour_points(dim, n, 2 * dim)
    r = unflat(b - A @ flat(got), dim, n)[d]
    assert np.max(np.abs(r[mask])) < 1e-10 * np.max(np.abs(b))


@pytest.mark.parametrize("vt", [-1, -3, 2])
def test_damped_jacobi_matches_assembled(vt):
    dim, n, dx = 2, (8, 8), 1.0
    g = O.geom(dim, n, dx)
    c = coefs(dim, n, 17, const=vt > 0, gamma=0.25)
    cf = cf_of(c)
    A = assemble(g, vt, cf, 1.0, dim, n)
    phi, rhs = rand_faces(dim, n, 3), rand_faces(dim, n, 4)
    p = O.params(visc_type=vt, stag_mg_smoother=0, stag_mg_omega=0.6)
    got = O.stag_mg_smooth(g, p, cf, 1.0, phi, rhs, 3)
    x, b = flat(phi), flat(rhs)
    for _ in range(3):
        x = x + 0.6 * (b - A @ x) / np.diag(A)
    assert relerr(got, unflat(x, dim, n)) < 1e-12


# ---------------------------------------------------------------- transfer

@pytest.mark.parametrize("dim,n", [(2, (8, 8)), (2, (8, 4)), (3, (4, 8, 4))])
def test_restrictions_preserve_constants(dim, n):
    g = O.geom(dim, n, 1.0)
    shp = shape_of(dim, n)
    u = [np.full(shp, 2.0 + d) for d in range(dim)]
    for simple in (True, False):
        for d, a in enumerate(O.stag_restriction(g, u, simple=simple)):
            assert np.allclose(a, 2.0 + d, rtol=0, atol=1e-15)
    assert np.allclose(O.cc_restriction(g, np.full(shp, 3.5)), 3.5, rtol=0, atol=1e-15)
    if dim == 3:
        for e, a in enumerate(O.edge_restriction(g, [np.full(shp, 1.0 + e) for e in range(3)])):
            assert np.allclose(a, 1.0 + e, rtol=0, atol=1e-15)


@pytest.mark.parametrize("dim,n", [(2, (8, 8)), (2, (16, 8)), (3, (4, 8, 4))])
def test_face_average_restriction_commutes_with_divergence(dim, n):
    gf = O.geom(dim, n, 0.5)
    gc = O.geom(dim, [v // 2 for v in n[:dim]], 1.0)
    u = rand_faces(dim, n, 7)
    lhs = O.compute_div(gc, O.stag_restriction(gf, u, simple=True))
    rhs = O.cc_restriction(gf, O.compute_div(gf, u))
    assert relerr(lhs, rhs) < 1e-13


@pytest.mark.parametrize("dim,n", [(2, (8, 8)), (3, (4, 8, 4))])
def test_full_restriction_reduces_to_face_average_when_constant_along_normal(dim, n):
    g = O.geom(dim, n, 1.0)
    rng = np.random.default_rng(5)
    shp = shape_of(dim, n)
    u = []
    for d in range(dim):
        ax = dim - 1 - d
        line_shape = list(shp)
        line_shape[ax] = 1
        u.append(np.broadcast_to(rng.standard_normal(line_shape), shp).copy())
    full = O.stag_restriction(g, u, simple=False)
    simple = O.stag_restriction(g, u, simple=True)
    assert relerr(full, simple) < 1e-15
    # and a general field differs (the 12-point form is not the face average)
    v = rand_faces(dim, n, 6)
    assert relerr(O.stag_restriction(g, v, simple=False), O.stag_restriction(g, v, simple=True)) > 1e-3


def interp_node(nc):
    """1-D linear interpolation between coarse nodes: fine 2I = c(I), 2I+1 = mean(c(I), c(I+1))."""
    M = np.zeros((2 * nc, nc))
    for I in range(nc):
        M[2 * I, I] = 1.0
        M[2 * I + 1, I] = 0.5
        M[2 * I + 1, (I + 1) % nc] = 0.5
    return M


def interp_cell(nc):
    """1-D linear interpolation of cell averages to the two half cells (3/4, 1/4)."""
    M = np.zeros((2 * nc, nc))
    for I in range(nc):
        M[2 * I, I] += 0.75
        M[2 * I, (I - 1) % nc] += 0.25
        M[2 * I + 1, I] += 0.75
        M[2 * I + 1, (I + 1) % nc] += 0.25
    return M


@pytest.mark.parametrize("dim,n", [(2, (8, 8)), (2, (16, 8)), (3, (8, 4, 8))])
def test_prolongation_is_tensor_product_linear_interpolation(dim, n):
    gf = O.geom(dim, n, 1.0)
    nc = [v // 2 for v in n[:dim]]
    coarse = rand_faces(dim, nc, 8)
    base = rand_faces(dim, n, 9)
    got = O.stag_prolongation(gf, base, coarse)
    for d in range(dim):
        want = coarse[d]
        for e in range(dim):  # axis of direction e is dim-1-e
            M = interp_node(nc[e]) if e == d else interp_cell(nc[e])
            want = np.moveaxis(np.tensordot(M, want, axes=([1], [dim - 1 - e])), 0, dim - 1 - e)
        assert relerr(got[d] - base[d], want) < 1e-14


def test_nodal_and_edge_restriction_are_injection_and_edge_mean():
    g2 = O.geom(2, (8, 8), 1.0)
    f = np.random.default_rng(1).standard_normal((8, 8))
    assert np.array_equal(O.nodal_restriction(g2, f), f[::2, ::2])
    g3 = O.geom(3, (4, 8, 4), 1.0)
    fe = rand_faces(3, (4, 8, 4), 2)
    ce = O.edge_restriction(g3, fe)
    # edge 0 (xy) lies along z, edge 1 (xz) along y, edge 2 (yz) along x
    assert np.allclose(ce[0], 0.5 * (fe[0][0::2, ::2, ::2] + fe[0][1::2, ::2, ::2]), rtol=0, atol=1e-15)
    assert np.allclose(ce[1], 0.5 * (fe[1][::2, 0::2, ::2] + fe[1][::2, 1::2, ::2]), rtol=0, atol=1e-15)
    assert np.allclose(ce[2], 0.5 * (fe[2][::2, ::2, 0::2] + fe[2][::2, ::2, 1::2]), rtol=0, atol=1e-15)


# ---------------------------------------------------------------- div / grad / inner products

@pytest.mark.parametrize("dim,n,dx", GEOMS)
def test_div_grad_mac_definitions_and_adjointness(dim, n, dx):
    g = O.geom(dim, n, dx)
    p = np.random.default_rng(1).standard_normal(shape_of(dim, n))
    u = rand_faces(dim, n, 2)
    gp = O.compute_grad(g, p)
    assert relerr(gp, mac_grad(p, dim, dx)) < 1e-15
    div = O.compute_div(g, u)
    want = sum((roll(u[d], -1, d, dim) - u[d]) / dx for d in range(dim))
    assert relerr(div, want) < 1e-15
    assert abs(sum(np.vdot(a, b) for a, b in zip(u, gp)) + np.vdot(p, div)) < 1e-11 * np.abs(div).sum()
    inc = O.compute_div(g, u, div=np.ones_like(p), increment=True)
    assert relerr(inc, div + 1.0) < 1e-15


@pytest.mark.parametrize("dim,n", [(2, (8, 4)), (3, (4, 6, 8))])
def test_inner_products_brute_force(dim, n):
    g = O.geom(dim, n, 1.0)
    m1, m2 = rand_faces(dim, n, 3), rand_faces(dim, n, 4)
    got = O.stag_inner_prod(g, m1, m2)
    for d in range(dim):
        w = np.ones(shape_of(dim, n))
        sl = [slice(None)] * dim
        sl[dim - 1 - d] = 0
        w[tuple(sl)] = 3.0  # face lo and its periodic image hi+1, each 1 + 0.5
        assert got[d] == pytest.approx(float(np.sum(w * m1[d] * m2[d])), rel=1e-13)
    a, b = m1[0], m2[0]
    assert O.cc_inner_prod(g, a, b) == pytest.approx(float(np.sum(a * b)), rel=1e-13)


# ---------------------------------------------------------------- GMRES pieces

@pytest.mark.parametrize("m", [1, 3, 8])
def test_givens_least_squares_matches_lstsq(m):
    rng = np.random.default_rng(m)
    H = np.asfortranarray(np.triu(rng.standard_normal((m + 1, m)), -1))
    H0 = H.copy()
    beta = 2.5
    s = np.zeros(m + 1)
    s[0] = beta
    cs, sn = np.zeros(m), np.zeros(m)
    for i in range(1, m + 1):
        O.least_squares(i, m, H, cs, sn, s)
    y = O.solve_utriangular(m - 1, m, H, s)
    e1 = np.zeros(m + 1)
    e1[0] = beta
    want, res, *_ = np.linalg.lstsq(H0, e1, rcond=None)
    assert relerr(y, want) < 1e-12
    assert abs(s[m]) == pytest.approx(np.linalg.norm(H0 @ want - e1), rel=1e-10, abs=1e-14)


def mac_cc_operator(ainv, dim, dx):
    """Dense div(a grad .) on a periodic cell grid, faces at the low side of each cell."""
    shp = ainv[0].shape
    N = int(np.prod(shp))
    L = np.zeros((N, N))
    for q in range(N):
        e = np.zeros(N)
        e[q] = 1.0
        p = e.reshape(shp)
        flux = [ainv[d] * (p - roll(p, 1, d, dim)) / dx for d in range(dim)]
        L[:, q] = sum((roll(flux[d], -1, d, dim) - flux[d]) / dx for d in range(dim)).ravel()
    return L


@pytest.mark.parametrize("dim,n", [(2, (16, 16)), (3, (8, 8, 8))])
def test_cc_multigrid_converges_to_dense_solution(dim, n):
    dx = 0.25
    g = O.geom(dim, n, dx)
    rng = np.random.default_rng(4)
    shp = shape_of(dim, n)
    ainv = [0.5 + rng.random(shp) for _ in range(dim)]
    rhs = rng.standard_normal(shp)
    rhs -= rhs.mean()
    p = O.params(mg_max_vcycles=60)
    phi = O.macproject(g, p, ainv, rhs)
    L = mac_cc_operator(ainv, dim, dx)
    want = np.linalg.lstsq(L, rhs.ravel(), rcond=None)[0]
    got = phi.ravel() - phi.mean()
    want = want - want.mean()
    assert relerr(got, want) < 1e-8
    # one V-cycle reduces the residual by a solid factor
    phi1 = O.macproject(g, O.params(mg_max_vcycles=1), ainv, rhs)
    assert np.linalg.norm(rhs.ravel() - L @ phi1.ravel()) < 0.3 * np.linalg.norm(rhs)


def assemble_stokes(g, p, cf, theta_alpha, dim, n):
    shp = shape_of(dim, n)
    m = int(np.prod(shp))
    M = np.zeros(((dim + 1) * m, (dim + 1) * m))
    for q in range((dim + 1) * m):
        e = np.zeros((dim + 1) * m)
        e[q] = 1.0
        xu = [e[d * m:(d + 1) * m].reshape(shp) for d in range(dim)]
        xp = e[dim * m:].reshape(shp)
        bu, bp = O.apply_matrix(g, p, cf, theta_alpha, xu, xp)
        M[:, q] = np.concatenate([flat(bu), bp.ravel()])
    return M


@pytest.mark.parametrize("dim,n,vt", [(2, (8, 8), -1), (2, (8, 8), -3), (3, (4, 4, 4), -2)])
def test_apply_matrix_block_structure(dim, n, vt):
    g = O.geom(dim, n, 0.5)
    c = coefs(dim, n, 30, gamma=0.2)
    cf = cf_of(c)
    p = O.params(visc_type=vt)
    M = assemble_stokes(g, p, cf, 1.0, dim, n)
    m = int(np.prod(shape_of(dim, n)))
    A = assemble(g, vt, cf, 1.0, dim, n)
    k = dim * m
    assert np.allclose(M[:k, :k], A, rtol=0, atol=1e-12)
    G = M[:k, k:]
    D = M[k:, :k]
    # b_u = A x_u + grad x_p and b_p = -div x_u; on the MAC grid grad = -div^T, so D = G^T
    e = np.zeros(m)
    e[3] = 1.0
    assert np.allclose(G[:, 3], flat(mac_grad(e.reshape(shape_of(dim, n)), dim, 0.5)), rtol=0, atol=1e-12)
    assert np.allclose(D, G.T, rtol=0, atol=1e-12)
    assert np.max(np.abs(M[k:, k:])) == 0.0


@pytest.mark.parametrize("dim,n,vt,theta", [(2, (8, 8), -1, 1.0), (2, (8, 8), -2, 0.5),
                                            (3, (4, 4, 4), -1, 1.0)])
def test_gmres_converges_to_dense_solution(dim, n, vt, theta):
    g = O.geom(dim, n, 0.5)
    c = coefs(dim, n, 40, gamma=0.0)
    cf = cf_of(c)
    p = O.params(visc_type=vt, gmres_rel_tol=1e-12, gmres_max_iter=400, gmres_max_outer=80,
                 gmres_max_inner=10, stag_mg_max_vcycles=1, mg_max_vcycles=1)
    M = assemble_stokes(g, p, cf, theta, dim, n)
    rng = np.random.default_rng(41)
    shp = shape_of(dim, n)
    b_u = [rng.standard_normal(shp) for _ in range(dim)]
    b_p = rng.standard_normal(shp)
    b_p -= b_p.mean()
    x_u, x_p, info = O.gmres(g, p, cf, theta, b_u, b_p)
    assert info["status"] == 0
    want = np.linalg.lstsq(M, np.concatenate([flat(b_u), b_p.ravel()]), rcond=None)[0]
    m = int(np.prod(shp))
    wu, wp = want[:dim * m], want[dim * m:]
    assert relerr(flat(x_u), wu) < 1e-9
    assert relerr(x_p.ravel() - x_p.mean(), wp - wp.mean()) < 1e-8


# ---------------------------------------------------------------- fluxes

@pytest.mark.parametrize("dim,n", [(2, (8, 6)), (3, (4, 6, 4))])
def test_stochastic_flux_divergence_conservative_and_linear(dim, n):
    g = O.geom(dim, n, 0.5)
    shp = shape_of(dim, n)
    rng = np.random.default_rng(6)
    ne = 1 if dim == 2 else 3
    eta = 0.5 + rng.random(shp)
    temp = 0.5 + rng.random(shp)
    eta_ed = [0.5 + rng.random(shp) for _ in range(ne)]
    temp_ed = [0.5 + rng.random(shp) for _ in range(ne)]

    def run(Wc, We):
        force = [np.zeros(shp) for _ in range(dim)]
        return O.stochastic_m_fluxdiv(g, Wc, We, eta, eta_ed, temp, temp_ed, 0.7, force)

    def draw(seed):
        r = np.random.default_rng(seed)
        return r.standard_normal((dim,) + shp), [r.standard_normal((2,) + shp) for _ in range(ne)]

    W1, W2 = draw(1), draw(2)
    f1, f2 = run(*W1), run(*W2)
    f12 = run(W1[0] + 2 * W2[0], [a + 2 * b for a, b in zip(W1[1], W2[1])])
    assert relerr(f12, [a + 2 * b for a, b in zip(f1, f2)]) < 1e-13
    for a in f1:  # a discrete divergence sums to zero over the periodic box
        assert abs(a.sum()) < 1e-12 * np.abs(a).sum()
    # uniform noise with uniform eta, T has zero divergence
    eta[:] = 1.0
    temp[:] = 2.0
    eta_ed = [np.ones(shp) for _ in range(ne)]
    temp_ed = [np.full(shp, 2.0) for _ in range(ne)]
    Wc = np.ones((dim,) + shp)
    We = [np.full((2,) + shp, 0.3) for _ in range(ne)]
    force = [np.zeros(shp) for _ in range(dim)]
    out = O.stochastic_m_fluxdiv(g, Wc, We, eta, eta_ed, temp, temp_ed, 0.7, force)
    assert np.max(np.abs(flat(out))) < 1e-13


@pytest.mark.parametrize("dim,n,ns", [(2, (8, 6), 3), (3, (4, 4, 6), 4)])
def test_diffusive_mass_flux_conservation(dim, n, ns):
    g = O.geom(dim, n, 0.5)
    shp = shape_of(dim, n)
    rng = np.random.default_rng(9)
    x = rng.random((ns,) + shp)
    rhoWchi = rng.standard_normal((ns, ns) + shp)
    Gama = rng.standard_normal((ns, ns) + shp)
    rho = rng.random((ns,) + shp)
    rhotot = rho.sum(axis=0)

    def run(molar, correct):
        ft = [np.zeros((ns,) + shp) for _ in range(dim)]
        return O.diffusive_mass_fluxdiv(g, ns, rho, rhotot, molar, rhoWchi, Gama, ft, correct)

    div, flux = run(x, True)
    for d in range(dim):  # the correction makes the species fluxes sum to zero
        assert np.max(np.abs(flux[d].sum(axis=0))) < 1e-13
    assert np.max(np.abs(div.sum(axis=0))) < 1e-12
    for s in range(ns):  # conservative: every species' divergence sums to zero
        assert abs(div[s].sum()) < 1e-11 * np.abs(div[s]).sum()
    # no gradient, no flux
    div0, flux0 = run(np.broadcast_to(x[:, :1, :1] if dim == 2 else x[:, :1, :1, :1], x.shape).copy(), False)
    assert np.max(np.abs(div0)) == 0.0 and all(np.max(np.abs(f)) == 0.0 for f in flux0)
    # linear in the composition without the correction
    da, _ = run(x, False)
    db, _ = run(2.0 * x, False)
    assert relerr(db, 2.0 * da) < 1e-14


# ------------------------------------------------- stochastic stress forms

def _filter_gain(width, th):
    """Fourier gain of multifab_filter's 1-D stencil (multifab_filter.f90:67-81)."""
    c = np.cos
    if width == 1:
        return 0.5 + 0.5 * c(th)
    if width == 2:
        return 5 / 8 + 0.5 * c(th) - 0.125 * c(2 * th)
    return 93 / 128 + 7 / 16 * c(th) - 7 / 32 * c(2 * th) + 1 / 16 * c(3 * th) - 1 / 128 * c(4 * th)


@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (3, (8, 16, 8))])
@pytest.mark.parametrize("width", [1, 2, 4])
def test_filtered_stochastic_flux_is_the_filtered_divergence(dim, n, width):
    """With eta = T = 1 the filter and the divergence are both linear and
    translation invariant, so a noise field made of one Fourier mode comes out
    as the unfiltered force times the product of the 1-D filter gains of its
    wavenumbers (a wrong coefficient, sign or offset in any pass breaks it)."""
    g = O.geom(dim, n, 0.5)
    shp = shape_of(dim, n)
    idx = np.meshgrid(*[np.arange(s) for s in shp], indexing="ij")
    # theta = pi/2 along x and z, pi/4 along y: distinct gains per direction
    kv = [n[0] // 4, n[1] // 8 if n[1] >= 8 else 1, (n[2] // 4) if dim == 3 else 0][:dim]
    th = [2 * np.pi * kv[d] / n[d] for d in range(dim)]
    phase = sum(th[d] * idx[dim - 1 - d] for d in range(dim))
    Wc = np.stack([np.cos(phase + 0.3 * d) for d in range(dim)])
    ne = 1 if dim == 2 else 3
    We = [np.stack([np.sin(phase + 0.2 * e), np.cos(phase - 0.1 * e)]) for e in range(ne)]
    one = np.ones(shp)
    ones_ed = [one.copy() for _ in range(ne)]
    zero = [np.zeros(shp) for _ in range(dim)]
    f0 = O.stochastic_m_fluxdiv(g, Wc, We, one, ones_ed, one, ones_ed, 0.9, zero)
    fw = O.stochastic_m_fluxdiv(g, Wc, We, one, ones_ed, one, ones_ed, 0.9, zero, filtering_width=width)
    gain = np.prod([_filter_gain(width, t) for t in th])
    assert relerr(fw, [gain * a for a in f0]) < 1e-12
    assert abs(gain - 1.0) > 0.05  # the mode is really damped


@pytest.mark.parametrize("dim,n", [(2, (4, 6)), (3, (4, 4, 4))])
def test_symmetric_stress_noise_satisfies_fluctuation_dissipation(dim, n):
    """Fluctuation-dissipation balance of the symmetric stress
    (stochastic_m_fluxdiv.f90:569-604 with eta = T = 1, variance 1): the
    covariance of the force D.W_sym, W_sym having variance-2 diagonal entries
    and one shared draw for both off-diagonal components of a node/edge, is
    exactly minus the full-stress operator (visc_type 2 with alpha = 0,
    stag_applyop.f90); the non-symmetric noise gives minus the Laplacian
    (visc_type 1).  Built by brute force over every noise degree of freedom."""
    g = O.geom(dim, n, 0.5)
    shp = shape_of(dim, n)
    N = int(np.prod(shp))
    ne = 1 if dim == 2 else 3
    one = np.ones(shp)
    ones_ed = [one.copy() for _ in range(ne)]
    zero = [np.zeros(shp) for _ in range(dim)]

    def force(Wc, We):
        return flat(O.stochastic_m_fluxdiv(g, Wc, We, one, ones_ed, one, ones_ed, 1.0, zero))

    def covariance(symmetric):
        cols = []
        for d in range(dim):
            for q in range(N):
                Wc = np.zeros((dim,) + shp)
                Wc.reshape(dim, -1)[d, q] = np.sqrt(2.0) if symmetric else 1.0
                cols.append(force(Wc, [np.zeros((2,) + shp) for _ in range(ne)]))
        for e in range(ne):
            for comp in ([0] if symmetric else [0, 1]):
                for q in range(N):
                    We = [np.zeros((2,) + shp) for _ in range(ne)]
                    We[e].reshape(2, -1)[comp, q] = 1.0
                    if symmetric:
                        We[e].reshape(2, -1)[1, q] = 1.0
                    cols.append(force(np.zeros((dim,) + shp), We))
        Fm = np.array(cols).T
        return Fm @ Fm.T

    def operator(visc):
        alpha = [np.zeros(shp) for _ in range(dim)]
        cf = O.coef(alpha, one, ones_ed, np.zeros(shp))
        return assemble(g, visc, cf, 1.0, dim, n)

    C2 = covariance(True)
    A2 = operator(2)
    assert np.abs(C2 - A2).max() < 1e-10 * np.abs(A2).max()
    C1 = covariance(False)
    A1 = operator(1)
    assert np.abs(C1 - A1).max() < 1e-10 * np.abs(A1).max()
    # the two forms really differ
    assert np.abs(C2 - C1).max() > 0.1 * np.abs(A1).max()


def test_fill_m_stochastic_forms():
    """Structure of the restated fill: the symmetric form scales W_cc by
    sqrt(2) of the same draws and consumes one component per node/edge
    array, copied to the second."""
    shp = (4, 6, 8)
    ns = O.fill_m_stochastic(5, 3, shp, 2, symmetric=False)
    sy = O.fill_m_stochastic(5, 3, shp, 2, symmetric=True)
    N = int(np.prod(shp))
    raw = O.fill_normal(5, 2 * 6 * N)
    assert np.array_equal(ns[0][0].ravel(), raw[:3 * N])
    assert np.array_equal(sy[0][0].ravel(), np.sqrt(2.0) * raw[:3 * N])
    for e in range(3):
        assert np.array_equal(sy[0][1][e][0], sy[0][1][e][1])
        assert np.array_equal(sy[0][1][e][0].ravel(), raw[(3 + e) * N:(4 + e) * N])
    # stage 2 of the symmetric form starts after 6 N values
    assert np.array_equal(sy[1][0].ravel(), np.sqrt(2.0) * raw[6 * N:9 * N])
