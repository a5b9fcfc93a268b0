"""Pins of the species mass-flux oracle (oracle/orc_species.c) against what
the mathematics fixes, independent of the oracle's own loops:

* convert_cons_to_prim: mole fractions sum to one; equal molar masses give
  mole fraction = mass fraction (mass_flux_utilities.f90:279-313);
* compute_mixture_properties: the Dbar / H_offdiag input order is the
  strict lower triangle by rows (compute_mixture_properties.f90:209-230);
* compute_Gama: ideal mixtures give the identity exactly; otherwise
  I + (diag(x) - x x^T) H (mass_flux_utilities.f90:564-599);
* compute_chi: the Dbar2chi_iterative series (matrix_utilities.f90:21-156)
  converges to the closed form (Lambda + alpha W W^T)^-1 - 1/alpha of
  compute_chi_lapack (:833-927); chi W = 0; a binary mixture gives Fick's
  law w_1 (chi_11 - chi_12) = D_12; the 10-term series equals the same
  series written as numpy matrix products;
* compute_Lonsager: the Cholesky factor reproduces
  molmtot rho W chi W / k_B (numpy), its leading block equals
  numpy.linalg.cholesky;
* the assembled diffusive flux of a binary mixture of equal molar masses at
  constant total density is -rho D times the discrete Laplacian of the mole
  fraction; the temperature piece vanishes when Dtherm is uniform;
* the stochastic mass flux conserves mass, and its covariance at uniform
  composition is variance^2 times the Onsager matrix (fluctuation-
  dissipation), which fails for a transposed Cholesky factor.
"""
import numpy as np
import pytest

from oracle import oracle as O

RNG = np.random.default_rng(20261203)


def _dbar(ns, vals):
    D = np.zeros((ns, ns))
    r, c = np.tril_indices(ns, -1)
    D[r, c] = vals
    D[c, r] = vals
    return D


def _random_mixture(ns, rng):
    m = rng.uniform(0.5, 3.0, ns)
    dvals = rng.uniform(0.5, 2.0, ns * (ns - 1) // 2)
    rho = rng.uniform(0.2, 1.0, ns)
    return m, dvals, rho


def _closed_form_chi(rho, m, D):
    W = rho / rho.sum()
    mt = 1.0 / np.sum(W / m)
    x = mt * W / m
    ns = len(rho)
    Lam = -np.outer(x, x) / np.where(D > 0, D, 1.0)
    np.fill_diagonal(Lam, 0.0)
    np.fill_diagonal(Lam, -Lam.sum(axis=1))
    a = np.trace(Lam)
    return np.linalg.inv(Lam + a * np.outer(W, W)) - 1.0 / a, x, W, mt, Lam


@pytest.mark.parametrize("ns", [2, 3, 4, 6, 8])
def test_chi_series_converges_to_closed_form(ns):
    m, dvals, rho = _random_mixture(ns, RNG)
    D = _dbar(ns, dvals)
    chi_ref, x, W, _, _ = _closed_form_chi(rho, m, D)
    sp = O.species(ns, molmass=m, Dbar=np.pad(dvals, (0, 28 - len(dvals))), chi_iterations=300)
    chi, _ = O.sp_cell_chi(sp, rho, x, D)
    scale = np.abs(chi_ref).max()
    assert np.abs(chi - chi_ref).max() <= 1e-11 * scale
    assert np.abs(chi @ W).max() <= 1e-12 * scale
    for inverse_type in (1, 2):
        spl = O.species(ns, molmass=m, Dbar=np.pad(dvals, (0, 28 - len(dvals))), use_lapack=1,
                        inverse_type=inverse_type)
        chil, _ = O.sp_cell_chi(spl, rho, x, D)
        assert np.abs(chil - chi_ref).max() <= 1e-11 * scale, inverse_type


@pytest.mark.parametrize("ns", [3, 5])
def test_chi_ten_terms_equal_matrix_series(ns):
    m, dvals, rho = _random_mixture(ns, RNG)
    D = _dbar(ns, dvals)
    _, x, _, _, Lam = _closed_form_chi(rho, m, D)
    sp = O.species(ns, molmass=m, Dbar=np.pad(dvals, (0, 28 - len(dvals))))
    chi, _ = O.sp_cell_chi(sp, rho, x, D)
    # chi_{k+1} = P J chi_k + P M^-1 P^T from chi_0 = P M^-1 P^T (Giovangigli's
    # iteration as matrix products)
    Y = m * x / np.sum(m * x)
    Dd = np.where(D > 0, D, np.inf)
    Di = (1.0 - Y) / np.array([np.sum(np.delete(x / Dd[i], i)) for i in range(ns)])
    Mi = Di / x
    P = np.eye(ns) - np.outer(np.ones(ns), Y)
    Z = -Lam + np.diag(x / Di)
    PJ = P @ (Mi[:, None] * Z)
    M2 = P @ np.diag(Mi) @ P.T
    c = M2.copy()
    for _ in range(10):
        c = PJ @ c + M2
    assert np.abs(chi - c).max() <= 1e-13 * np.abs(c).max()


@pytest.mark.parametrize("w1", [0.15, 0.5, 0.8])
def test_binary_chi_is_ficks_law(w1):
    D = 1.7
    sp = O.species(2, Dbar=D, chi_iterations=60)
    rho = np.array([w1, 1 - w1]) * 1.3
    x = rho / rho.sum()
    chi, _ = O.sp_cell_chi(sp, rho, x, np.array([[0, D], [D, 0]]))
    assert abs(w1 * (chi[0, 0] - chi[0, 1]) - D) <= 1e-13 * D


def test_zeta_by_temp():
    ns = 4
    m, dvals, rho = _random_mixture(ns, RNG)
    D = _dbar(ns, dvals)
    _, x, _, _, Lam = _closed_form_chi(rho, m, D)
    Dt = RNG.uniform(-1, 1, ns)
    T = 1.9
    sp = O.species(ns, molmass=m, Dbar=np.pad(dvals, (0, 28 - len(dvals))))
    _, z = O.sp_cell_chi(sp, rho, x, D, Temp=T, D_therm=Dt)
    off = Lam - np.diag(np.diag(Lam))
    want = (off * (Dt[:, None] - Dt[None, :])).sum(axis=1) / T
    assert np.abs(z - want).max() <= 1e-14 * max(1.0, np.abs(want).max())


def test_gama():
    ns = 4
    x = RNG.uniform(0.1, 1, ns)
    x /= x.sum()
    H = RNG.normal(size=(ns, ns))
    H = H + H.T
    G = O.sp_cell_gama(O.species(ns), x, H)
    assert np.array_equal(G, np.eye(ns))
    G = O.sp_cell_gama(O.species(ns, is_ideal_mixture=0), x, H)
    want = np.eye(ns) + (np.diag(x) - np.outer(x, x)) @ H
    assert np.abs(G - want).max() <= 1e-14 * np.abs(want).max()


def test_mixture_properties_order():
    ns = 4
    g = O.geom(2, (4, 4))
    dv = np.arange(1, 7, dtype=float)
    hv = np.arange(11, 17, dtype=float)
    sp = O.species(ns, Dbar=np.pad(dv, (0, 22)), H_offdiag=np.pad(hv, (0, 22)), H_diag=[5, 6, 7, 8, 0, 0, 0, 0],
                   Dtherm=[1, 2, 3, 4, 0, 0, 0, 0], is_ideal_mixture=0)
    Db, Dt, H = O.compute_mixture_properties(g, sp, (4, 4))
    want_D = _dbar(ns, dv)
    want_H = _dbar(ns, hv) + np.diag([5, 6, 7, 8])
    assert np.array_equal(Db[:, 1, 2].reshape(ns, ns).T, want_D)
    assert np.array_equal(H[:, 3, 0].reshape(ns, ns).T, want_H)
    assert np.array_equal(Dt[:, 0, 0], [1, 2, 3, 4])
    _, _, H0 = O.compute_mixture_properties(g, O.species(ns), (4, 4))
    assert not H0.any()


def test_cons_to_prim_and_correction():
    ns = 3
    g = O.geom(2, (8, 4))
    rho = RNG.uniform(0.1, 1.0, (ns, 4, 8))
    rho[1, 2, 3] = 0.0
    m = np.array([1.0, 2.0, 0.5])
    sp = O.species(ns, molmass=m, fraction_tolerance=1e-3)
    rc, drho = O.correct_rho_with_drho(g, sp, rho)
    tot = rho.sum(axis=0)
    assert rc[1, 2, 3] == 1e-3 * tot[2, 3] and drho[1, 2, 3] == rc[1, 2, 3]
    keep = np.ones_like(rho, dtype=bool)
    keep[1, 2, 3] = False
    assert np.array_equal(rc[keep], rho[keep]) and not drho[keep].any()
    rhotot, x, mt = O.convert_cons_to_prim(g, sp, rho)
    assert np.array_equal(rhotot, tot)
    assert np.abs(x.sum(axis=0) - 1).max() <= 4e-16
    W = rho / tot
    assert np.abs(mt - 1 / (W / m[:, None, None]).sum(axis=0)).max() <= 1e-15 * mt.max()
    _, x1, _ = O.convert_cons_to_prim(g, O.species(ns), rho)
    assert np.abs(x1 - W).max() <= 2e-16


@pytest.mark.parametrize("use_lapack", [0, 1])
@pytest.mark.parametrize("ns", [2, 3, 5])
def test_lonsager_cholesky(ns, use_lapack):
    m, dvals, rho = _random_mixture(ns, RNG)
    D = _dbar(ns, dvals)
    chi, x, W, mt, _ = _closed_form_chi(rho, m, D)
    kB = 0.37
    sp = O.species(ns, molmass=m, k_B=kB, use_lapack=use_lapack)
    Lc = O.sp_cell_lonsager(sp, rho, rho.sum(), mt, chi)
    L = mt * rho.sum() * np.outer(W, W) * chi / kB
    assert np.array_equal(np.triu(Lc, 1), np.zeros_like(Lc))
    assert np.abs(Lc @ Lc.T - L).max() <= 1e-12 * np.abs(L).max()
    lead = np.linalg.cholesky(L[:-1, :-1])
    assert np.abs(Lc[:-1, :-1] - lead).max() <= 1e-12 * np.abs(lead).max()
    if use_lapack:
        assert Lc[-1, -1] == 0.0


def test_invert_and_pinv():
    for ns in (2, 4, 7):
        A = RNG.normal(size=(ns, ns)) + ns * np.eye(ns)
        assert np.abs(O.sp_invert(A) - np.linalg.inv(A)).max() <= 1e-13 * np.abs(np.linalg.inv(A)).max()
        U, _ = np.linalg.qr(RNG.normal(size=(ns, ns)))
        V, _ = np.linalg.qr(RNG.normal(size=(ns, ns)))
        s = np.linspace(2.0, 1.0, ns)
        s[-1] = 1e-18
        B = U @ np.diag(s) @ V.T
        want = np.linalg.pinv(B, rcond=1e-10)
        assert np.abs(O.sp_pinv(B, 1e-13) - want).max() <= 1e-12 * np.abs(want).max()


def _lap(f, dx):
    out = np.zeros_like(f)
    for ax in range(f.ndim):
        out += (np.roll(f, -1, ax) - 2 * f + np.roll(f, 1, ax)) / dx ** 2
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_binary_diffusive_flux_is_laplacian(dim):
    n = (16, 12, 8)[:dim]
    shape = tuple(reversed(n))
    g = O.geom(dim, n, dx=0.5)
    Dv = 1.3
    rhot = 2.0
    grid = np.meshgrid(*[np.arange(k) for k in shape], indexing="ij")
    w1 = 0.4 + 0.1 * np.sin(2 * np.pi * grid[-1] / shape[-1]) * np.cos(2 * np.pi * grid[0] / shape[0])
    rho = np.stack([rhot * w1, rhot * (1 - w1)])
    sp = O.species(2, Dbar=Dv, chi_iterations=80, is_nonisothermal=0)
    rhotot, x, mt = O.convert_cons_to_prim(g, sp, rho)
    Db, Dt, H = O.compute_mixture_properties(g, sp, shape)
    Gama = O.compute_Gama(g, sp, x, H)
    chi, _ = O.compute_chi(g, sp, rho, rhotot, x, Db, Dt, np.ones(shape))
    rw = O.compute_rhoWchi(g, sp, rho, chi)
    ft = [np.zeros((2,) + shape) for _ in range(dim)]
    div, ft2 = O.diffusive_mass_flux_full(g, sp, rho, rhotot, x, rw, Gama, ft)
    want = -rhot * Dv * _lap(x[0], 0.5)
    assert np.abs(div[0] - want).max() <= 1e-12 * np.abs(want).max()
    assert np.abs(div.sum(axis=0)).max() <= 1e-13 * np.abs(want).max()
    # the isothermal restatement of the same flux
    div_old, _ = O.diffusive_mass_fluxdiv(g, 2, rho, rhotot, x, rw, Gama, ft)
    assert np.array_equal(div_old, div)
    # uniform Dtherm: the temperature piece adds nothing
    spT = O.species(2, Dbar=Dv, chi_iterations=80, Dtherm=0.7)
    T = 1.0 + 0.2 * np.cos(2 * np.pi * grid[-1] / shape[-1])
    chiT, zeta = O.compute_chi(g, spT, rho, rhotot, x, Db, np.full((2,) + shape, 0.7), T)
    assert not zeta.any()
    divT, _ = O.diffusive_mass_flux_full(g, spT, rho, rhotot, x, rw, Gama, ft, Temp=T, zeta_by_Temp=zeta)
    assert np.array_equal(divT, div)


def test_thermodiffusion_piece_direction():
    """Uniform composition, varying T: the flux is rhoWchi zeta dT/dx on faces."""
    ns, n = 3, (16, 8)
    shape = (8, 16)
    g = O.geom(2, n)
    rho = np.ones((ns,) + shape) * np.array([0.3, 0.5, 0.2])[:, None, None]
    m = np.array([1.0, 2.0, 3.0])
    Dt = np.array([0.5, -0.2, 0.1])
    sp = O.species(ns, molmass=m, Dtherm=np.pad(Dt, (0, 5)), Dbar=[1.0, 1.5, 0.8] + [0] * 25)
    rhotot, x, mt = O.convert_cons_to_prim(g, sp, rho)
    Db, Dtf, H = O.compute_mixture_properties(g, sp, shape)
    T = 1.0 + 0.3 * np.sin(2 * np.pi * np.arange(16) / 16)[None, :] * np.ones(shape)
    chi, zeta = O.compute_chi(g, sp, rho, rhotot, x, Db, Dtf, T)
    rw = O.compute_rhoWchi(g, sp, rho, chi)
    Gama = O.compute_Gama(g, sp, x, H)
    ft = [np.zeros((ns,) + shape) for _ in range(2)]
    div, ft = O.diffusive_mass_flux_full(g, sp, rho, rhotot, x, rw, Gama, ft, Temp=T, zeta_by_Temp=zeta)
    # uniform composition: rhoWchi is the same in every cell, zeta = S / T
    R = rw[:, 0, 0].reshape(ns, ns).T
    gT = T - np.roll(T, 1, axis=1)
    want = np.einsum("rc,cyx->ryx", R, 0.5 * (zeta + np.roll(zeta, 1, axis=2)) * gT)
    assert np.abs(ft[0] - want).max() <= 1e-13 * np.abs(want).max()
    assert np.abs(ft[0].sum(axis=0)).max() <= 1e-15
    assert not ft[1].any()


@pytest.mark.parametrize("ns", [2, 4])
def test_stochastic_mass_flux_fluctuation_dissipation(ns):
    dim, n = 2, (64, 64)
    shape = (64, 64)
    g = O.geom(dim, n, dx=0.5)
    m, dvals, rho0 = _random_mixture(ns, RNG)
    rho = np.ones((ns,) + shape) * rho0[:, None, None]
    sp = O.species(ns, molmass=m, Dbar=np.pad(dvals, (0, 28 - len(dvals))), k_B=0.8, variance_coef_mass=1.5,
                   chi_iterations=200, is_nonisothermal=0)
    rhotot, x, mt = O.convert_cons_to_prim(g, sp, rho)
    Db, Dt, H = O.compute_mixture_properties(g, sp, shape)
    chi, _ = O.compute_chi(g, sp, rho, rhotot, x, Db, Dt, np.ones(shape))
    Gama = O.compute_Gama(g, sp, x, H)
    dt = 0.01
    variance = np.sqrt(2 * 0.8 * 1.5 / (0.5 * 0.5 * dt))
    chi_c, _, W0, mtc, _ = _closed_form_chi(rho0, m, _dbar(ns, dvals))
    L = mtc * rho0.sum() * np.outer(W0, W0) * chi_c / 0.8
    samples = []
    for rep in range(6):
        Wn = RNG.normal(size=(1, dim, ns) + shape)
        ft = [np.zeros((ns,) + shape) for _ in range(dim)]
        div, ft = O.stochastic_mass_fluxdiv(g, sp, rho, rhotot, x, mt, chi, Gama, ft, dt, Wn, [1.0])
        assert np.abs(div.sum(axis=0)).max() <= 1e-12 * np.abs(div).max()
        assert np.abs(div.sum(axis=(1, 2))).max() <= 1e-10 * np.abs(div).max()
        samples.append(ft[0].reshape(ns, -1))
    F = np.concatenate(samples, axis=1)
    cov = F @ F.T / F.shape[1]
    want = variance ** 2 * L
    err = np.abs(cov - want).max() / np.abs(want).max()
    assert err < 0.03, err
    # the covariance test tells the factor from its transpose
    if ns > 2:
        Lfac = O.sp_cell_lonsager(sp, rho0, rho0.sum(), mtc, chi_c)
        assert np.abs(Lfac.T @ Lfac - L).max() > 0.1 * np.abs(L).max()


def test_stochastic_mass_flux_linear_in_weights():
    ns, dim = 3, 3
    n = (8, 6, 4)
    shape = (4, 6, 8)
    g = O.geom(dim, n)
    rho = RNG.uniform(0.2, 1.0, (ns,) + shape)
    sp = O.species(ns, Dbar=[1.0, 0.7, 1.3] + [0] * 25)
    rhotot, x, mt = O.convert_cons_to_prim(g, sp, rho)
    Db, Dt, H = O.compute_mixture_properties(g, sp, shape)
    chi, _ = O.compute_chi(g, sp, rho, rhotot, x, Db, Dt, np.ones(shape))
    Gama = O.compute_Gama(g, sp, x, H)
    W = RNG.normal(size=(2, dim, ns) + shape)
    ft = [np.zeros((ns,) + shape) for _ in range(dim)]
    a, _ = O.stochastic_mass_fluxdiv(g, sp, rho, rhotot, x, mt, chi, Gama, ft, 0.1, W, [0.5, -1.5])
    b, _ = O.stochastic_mass_fluxdiv(g, sp, rho, rhotot, x, mt, chi, Gama, ft, 0.1, W[:1], [1.0])
    c, _ = O.stochastic_mass_fluxdiv(g, sp, rho, rhotot, x, mt, chi, Gama, ft, 0.1, W[1:], [1.0])
    assert np.abs(a - (0.5 * b - 1.5 * c)).max() <= 1e-12 * np.abs(a).max()


def test_compute_mass_fluxdiv_composes_the_chain():
    ns, dim = 3, 2
    n = (12, 8)
    shape = (8, 12)
    g = O.geom(dim, n)
    rho = RNG.uniform(0.2, 1.0, (ns,) + shape)
    rho[2, 3, 4] = 1e-20
    T = RNG.uniform(0.9, 1.1, shape)
    sp = O.species(ns, Dbar=[1.0, 0.7, 1.3] + [0] * 25, Dtherm=[0.1, 0.0, -0.2, 0, 0, 0, 0, 0],
                   molmass=[1.0, 2.0, 1.5, 1, 1, 1, 1, 1], fraction_tolerance=1e-6)
    W = RNG.normal(size=(1, dim, ns) + shape)
    rho_after, diff, stoch, ft = O.compute_mass_fluxdiv(g, sp, rho, T, 0.05, W, [1.0])
    assert np.abs(rho_after - rho).max() <= 1e-15
    rc, _ = O.correct_rho_with_drho(g, sp, rho)
    rhotot, x, mt = O.convert_cons_to_prim(g, sp, rc)
    Db, Dt, H = O.compute_mixture_properties(g, sp, shape)
    G = O.compute_Gama(g, sp, x, H)
    chi, zeta = O.compute_chi(g, sp, rc, rhotot, x, Db, Dt, T)
    rw = O.compute_rhoWchi(g, sp, rc, chi)
    z0 = [np.zeros((ns,) + shape) for _ in range(dim)]
    d1, f1 = O.diffusive_mass_flux_full(g, sp, rc, rhotot, x, rw, G, z0, Temp=T, zeta_by_Temp=zeta)
    s1, f2 = O.stochastic_mass_fluxdiv(g, sp, rc, rhotot, x, mt, chi, G, f1, 0.05, W, [1.0])
    assert np.array_equal(diff, d1) and np.array_equal(stoch, s1)
    for d in range(dim):
        assert np.array_equal(ft[d], f2[d])
    # the Onsager factor is a function of the composition: no mass leaves a cell
    assert np.abs(diff.sum(axis=0)).max() <= 1e-12 * np.abs(diff).max()
    assert np.abs(stoch.sum(axis=0)).max() <= 1e-12 * np.abs(stoch).max()
