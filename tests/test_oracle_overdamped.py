"""Pins of the overdamped-step oracle (oracle.advance_timestep_overdamped,
advance_timestep_overdamped.f90:60-700) against what the mathematics fixes:

* a fluid at rest stays at rest;
* with no forcing the delta-form solve (theta_alpha = 0, A_0 v in the
  right-hand side) takes the non-uniform part of any velocity to zero in one
  step, leaving a uniform velocity (the null mode of the viscous operator on
  a periodic box, which the solve does not fix): dropping A_0 v, or a nonzero
  theta_alpha, fails this;
* a uniform flow U0 carries a composition wave sin(kx) by the midpoint
  scheme's amplification G = 1 + i a dt - (a dt)^2 / 2, a = -U0 sin(k dx)/dx:
  a full predictor step, or a corrector that advects the old faces, misses
  it by (a dt)^2 / 2;
* the step's final velocity is the Stokes solution for the corrector's
  forcing alone: for algorithm_type 2 the noise (W1 + W2)/sqrt(2) at dt; a
  corrector that kept the predictor's weights (1, 0) or its dt/2 fails this;
* species mass is conserved, with noise, for algorithm_type 1 and 2;
* noise is drawn at the start of every step but istep 1 and restart+1.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

RNG = np.random.default_rng(20261205)


def _shape(n):
    return tuple(reversed(n))


def _state(dim, n, ns, rng, uniform=False, u_scale=0.0):
    shape = _shape(n)
    g = O.geom(dim, n, 1.0 / n[0])
    if uniform:
        w = np.linspace(1.0, 2.0, ns)
        rho = np.stack([np.full(shape, w[c] / w.sum()) for c in range(ns)])
    else:
        rho = np.stack([0.5 + rng.random(shape) for _ in range(ns)]) / ns
    rhotot = O.compute_rhotot(rho)
    eta = np.full(shape, 0.8)
    temp = np.ones(shape)
    st = dict(umac=[u_scale * rng.standard_normal(shape) for _ in range(dim)], rho_old=rho, rhotot_old=rhotot,
              pi=np.zeros(shape), eta=eta, eta_ed=O._edges(g, eta), kappa=np.ones(shape), Temp=temp,
              Temp_ed=O._edges(g, temp), diff_mass_fluxdiv=np.zeros((ns,) + shape),
              stoch_mass_fluxdiv=np.zeros((ns,) + shape), gradp_baro=None)
    return g, st


def _species(ns, **kw):
    d = dict(molmass=np.linspace(1.0, 2.0, ns), rhobar=np.linspace(1.0, 1.5, ns),
             Dbar=np.linspace(0.5, 1.0, ns * (ns - 1) // 2), k_B=1e-4, variance_coef_mass=0.0)
    d.update(kw)
    return O.species(ns, **d)


def _params(visc_type=2):
    p = O.params(visc_type=visc_type)
    p.gmres_rel_tol = 1e-12
    return p


@pytest.mark.parametrize("alg", [1, 2])
def test_rest_state_stays_at_rest(alg):
    g, st = _state(2, (16, 16), 3, RNG, uniform=True)
    out, i1, i2 = O.advance_timestep_overdamped(g, _params(), _species(3), st, 0.01, 0.0, 1, alg,
                                                variance_coef_mom=0.0, visc_coef=0.8)
    assert all(np.max(np.abs(u)) == 0.0 for u in out["umac"])
    assert np.max(np.abs(out["rho_new"] - st["rho_old"])) < 1e-15
    assert np.max(np.abs(out["pi"])) < 1e-14


@pytest.mark.parametrize("dim,n,visc", [(2, (16, 16), 2), (2, (16, 16), -1), (3, (8, 8, 8), 1)])
def test_unforced_velocity_relaxes_in_one_step(dim, n, visc):
    g, st = _state(dim, n, 2, RNG, uniform=True, u_scale=0.1)
    U = [0.3, -0.2, 0.1][:dim]
    # a uniform part on top of a random field (taken to zero)
    st["umac"] = [u + U[d] for d, u in enumerate(st["umac"])]
    out, i1, i2 = O.advance_timestep_overdamped(g, _params(visc), _species(2), st, 0.01, 0.0, 1, 1,
                                                variance_coef_mom=0.0, visc_coef=0.8)
    assert i1["status"] == 0 and i2["status"] == 0
    for d in range(dim):
        u = out["umac"][d]
        assert np.max(u) - np.min(u) < 1e-9, d
        assert abs(np.mean(u) - U[d]) < 0.05, d
    # the composition is uniform and stays so
    assert np.max(np.abs(out["rho_new"] - st["rho_old"])) < 1e-15


def test_uniform_flow_advects_by_midpoint_amplification():
    n = (32, 8)
    shape = _shape(n)
    g = O.geom(2, n, 1.0 / n[0])
    k = 2 * np.pi
    x = (np.arange(n[0]) + 0.5) / n[0] * np.ones((n[1], 1))
    amp = 0.1
    r1 = 0.5 + amp * np.sin(k * x)
    rho = np.stack([r1, 1.0 - r1])
    U0, dt = 0.5, 0.02
    small = 1e-12
    st = dict(umac=[np.full(shape, U0), np.zeros(shape)], rho_old=rho, rhotot_old=O.compute_rhotot(rho),
              pi=np.zeros(shape), eta=np.ones(shape), eta_ed=O._edges(g, np.ones(shape)), kappa=np.ones(shape),
              Temp=np.ones(shape), Temp_ed=O._edges(g, np.ones(shape)), diff_mass_fluxdiv=np.zeros((2,) + shape),
              stoch_mass_fluxdiv=np.zeros((2,) + shape), gradp_baro=None)
    sp = _species(2, rhobar=[1.0, 1.0], Dbar=[small])
    out, _, _ = O.advance_timestep_overdamped(g, _params(-1), sp, st, dt, 0.0, 1, 1, variance_coef_mom=0.0,
                                              visc_coef=1.0)
    a = -U0 * math.sin(k * g.dx) / g.dx
    G = 1 + 1j * a * dt - (a * dt) ** 2 / 2
    want = 0.5 + amp * (G.real * np.sin(k * x) + G.imag * np.cos(k * x))
    assert np.max(np.abs(out["rho_new"][0] - want)) < 1e-9
    # the tempting variants are far off
    for Gx in (1 + 1j * a * dt, 1 + 1j * a * dt - (a * dt) ** 2):
        bad = 0.5 + amp * (Gx.real * np.sin(k * x) + Gx.imag * np.cos(k * x))
        assert np.max(np.abs(bad - want)) > 1e-5
    assert np.max(np.abs(out["umac"][0] - U0)) < 1e-12 and np.max(np.abs(out["umac"][1])) < 1e-12


@pytest.mark.parametrize("dim,n", [(2, (16, 16)), (3, (8, 8, 8))])
def test_final_velocity_is_stokes_solution_of_corrector_noise(dim, n):
    g, st = _state(dim, n, 2, RNG, uniform=True)
    shape = _shape(n)
    dt, eta, var = 0.01, 0.8, 1.0
    sp = _species(2, k_B=1e-3)
    W = O.fill_m_stochastic(21, dim, shape, 2, symmetric=True)
    p = _params(2)
    out, i1, i2 = O.advance_timestep_overdamped(g, p, sp, st, dt, 0.0, 1, 2, variance_coef_mom=var,
                                                visc_coef=eta, W_m=W)
    dv = g.dx ** dim
    w = 1.0 / math.sqrt(2.0)
    wcc = (0.0 + w * W[0][0]) + w * W[1][0]
    wed = [(0.0 + w * a) + w * b for a, b in zip(W[0][1], W[1][1])]
    zf = [np.zeros(shape) for _ in range(dim)]
    f = O.stochastic_m_fluxdiv(g, wcc, wed, st["eta"], st["eta_ed"], st["Temp"], st["Temp_ed"],
                               math.sqrt(var * 2.0 * sp.k_B / (dv * dt)), zf)
    cf = O.coef(O.average_cc_to_face(g, st["rhotot_old"]), st["eta"], st["eta_ed"], st["kappa"])
    u, _, _ = O.gmres(g, p, cf, 0.0, f, np.zeros(shape))
    scale = max(np.max(np.abs(x)) for x in u)
    assert scale > 0
    for d in range(dim):
        assert np.max(np.abs(out["umac"][d] - u[d])) < 1e-8 * scale, d
    # the predictor's weights (1, 0) at dt/2 give another velocity
    f2 = O.stochastic_m_fluxdiv(g, 0.0 + 1.0 * W[0][0], [0.0 + 1.0 * a for a in W[0][1]], st["eta"],
                                st["eta_ed"], st["Temp"], st["Temp_ed"],
                                math.sqrt(var * 2.0 * sp.k_B / (dv * 0.5 * dt)), zf)
    u2, _, _ = O.gmres(g, p, cf, 0.0, f2, np.zeros(shape))
    assert max(np.max(np.abs(a - b)) for a, b in zip(u2, u)) > 0.1 * scale


@pytest.mark.parametrize("alg", [1, 2])
@pytest.mark.parametrize("dim,n,ns", [(2, (16, 16), 3), (3, (8, 8, 8), 2)])
def test_mass_is_conserved_with_noise(alg, dim, n, ns):
    g, st = _state(dim, n, ns, RNG, u_scale=0.05)
    shape = _shape(n)
    sp = _species(ns, variance_coef_mass=1e-8, k_B=1e-9)
    dt = 2e-4
    N = int(np.prod(shape))
    wm = O.fill_m_stochastic(31, dim, shape, alg, symmetric=True)
    wmass = O.fill_normal(32, alg * dim * ns * N).reshape((alg, dim, ns) + shape)
    out, i1, i2 = O.advance_timestep_overdamped(g, _params(), sp, st, dt, 0.0, 1, alg, variance_coef_mom=1.0,
                                                visc_coef=0.8, W_m=wm, W_mass=wmass)
    assert i1["status"] == 0 and i2["status"] == 0
    for c in range(ns):
        m0, m1 = np.sum(st["rho_old"][c]), np.sum(out["rho_new"][c])
        assert abs(m1 - m0) < 1e-12 * abs(m0)
    assert np.max(np.abs(out["rho_new"] - st["rho_old"])) > 0
    assert all(np.isfinite(u).all() for u in out["umac"])


def test_noise_is_drawn_except_at_first_step_and_restart():
    g, st = _state(2, (8, 8), 2, RNG, uniform=True)
    shape = _shape((8, 8))
    old = O.fill_m_stochastic(41, 2, shape, 1, symmetric=True)
    new = O.fill_m_stochastic(42, 2, shape, 1, symmetric=True)
    kw = dict(variance_coef_mom=1.0, visc_coef=0.8, W_m=old, W_m_new=new)
    sp = _species(2, k_B=1e-6)
    for istep, restart, drawn in ((1, -1, False), (2, -1, True), (5, 4, False), (6, 4, True)):
        out, _, _ = O.advance_timestep_overdamped(g, _params(), sp, st, 0.01, 0.0, istep, 1, restart=restart, **kw)
        assert out["W_m"] is (new if drawn else old), (istep, restart)
