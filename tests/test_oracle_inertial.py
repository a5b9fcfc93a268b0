"""Pins of the inertial-step oracle (oracle/orc_inertial.c and
oracle.advance_timestep_inertial) against what the mathematics fixes:

* mk_advective_s_fluxdiv (mk_advective_s_fluxdiv.f90:14-166): a uniform
  velocity u0 along direction d gives the centred difference
  -u0 (s_{i+1} - s_{i-1}) / (2 dx) of the face-averaged field; the total over
  the box is zero (conservation); a constant field in a discretely
  divergence-free velocity (from a stream function) is not changed;
* mk_advective_m_fluxdiv (mk_advective_m_fluxdiv.f90:16-331): a uniform
  velocity gives the centred difference of each momentum component along it;
  the total of each component is zero; for a divergence-free velocity
  advecting itself the kinetic energy sum_faces u . div(u u) vanishes, which
  fails if any transverse flux takes the wrong neighbour;
* mk_grav_force (mk_grav_force.f90:17-161): 0.5 g (s_old + s_new);
* the assembled step (advance_timestep_inertial.f90:58-990): a fluid at rest
  stays at rest; a uniform fluid under gravity moves with u = g dt; a shear
  wave u_x(y) decays by the Crank-Nicolson factor (1 - dt nu lam/2) / (1 +
  dt nu lam/2) of its discrete Laplacian eigenvalue lam; species mass is
  conserved; without gravity the total momentum sum rho_f u is conserved to
  the solver's tolerance.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

RNG = np.random.default_rng(20261204)


def _faces(dim, shape, rng, scale=1.0):
    return [scale * rng.standard_normal(shape) for _ in range(dim)]


def _shape(n):
    return tuple(reversed(n))


def _roll(a, d, s):
    """a at index i + s along direction d (x = last numpy axis)."""
    return np.roll(a, -s, axis=a.ndim - 1 - d)


@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (3, (8, 6, 4))])
def test_advective_s_uniform_velocity_is_centred_difference(dim, n):
    g = O.geom(dim, n, 0.25)
    shape = _shape(n)
    s = RNG.standard_normal((2,) + shape)
    s_fc = O._faces_comps(g, s)
    for d in range(dim):
        u0 = 0.7 + 0.1 * d
        umac = [np.zeros(shape) for _ in range(dim)]
        umac[d][...] = u0
        got = O.mk_advective_s_fluxdiv(g, umac, s_fc, np.zeros((2,) + shape))
        want = -u0 * (_roll(s, d, 1) - _roll(s, d, -1)) / (2 * g.dx)
        assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("dim,n", [(2, (12, 10)), (3, (6, 8, 4))])
def test_advective_s_conserves_and_keeps_constants(dim, n):
    g = O.geom(dim, n, 0.5)
    shape = _shape(n)
    umac = _faces(dim, shape, RNG)
    s_fc = [RNG.standard_normal((3,) + shape) for _ in range(dim)]
    upd = RNG.standard_normal((3,) + shape)
    got = O.mk_advective_s_fluxdiv(g, umac, s_fc, upd)
    for c in range(3):
        assert abs(np.sum(got[c] - upd[c])) < 1e-12 * np.sum(np.abs(got[c] - upd[c]))
    # constant field, velocity from a stream function (2D) / vector potential (3D)
    if dim == 2:
        psi = RNG.standard_normal(shape)  # on nodes
        u = (_roll(psi, 1, 1) - psi) / g.dx
        v = -(_roll(psi, 0, 1) - psi) / g.dx
        umac = [u, v]
    else:
        A = [RNG.standard_normal(shape) for _ in range(3)]  # A_x on yz edges, A_y on xz, A_z on xy
        u = ((_roll(A[2], 1, 1) - A[2]) - (_roll(A[1], 2, 1) - A[1])) / g.dx
        v = ((_roll(A[0], 2, 1) - A[0]) - (_roll(A[2], 0, 1) - A[2])) / g.dx
        w = ((_roll(A[1], 0, 1) - A[1]) - (_roll(A[0], 1, 1) - A[0])) / g.dx
        umac = [u, v, w]
    assert np.max(np.abs(O.compute_div(g, umac))) < 1e-13
    const = [np.full((1,) + shape, 2.5) for _ in range(dim)]
    got = O.mk_advective_s_fluxdiv(g, umac, const, np.zeros((1,) + shape))
    assert np.max(np.abs(got)) < 1e-12


@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (3, (8, 6, 4))])
def test_advective_m_uniform_velocity_is_centred_difference(dim, n):
    g = O.geom(dim, n, 0.5)
    shape = _shape(n)
    m = _faces(dim, shape, RNG)
    for d in range(dim):
        u0 = 1.3 - 0.2 * d
        umac = [np.zeros(shape) for _ in range(dim)]
        umac[d][...] = u0
        got = O.mk_advective_m_fluxdiv(g, umac, m, [np.zeros(shape)] * dim)
        for a in range(dim):
            want = -u0 * (_roll(m[a], d, 1) - _roll(m[a], d, -1)) / (2 * g.dx)
            assert np.max(np.abs(got[a] - want)) < 1e-12 * np.max(np.abs(want)), (d, a)


@pytest.mark.parametrize("dim,n", [(2, (12, 10)), (3, (6, 8, 4))])
def test_advective_m_conserves_momentum_and_energy(dim, n):
    g = O.geom(dim, n, 0.5)
    shape = _shape(n)
    umac, m = _faces(dim, shape, RNG), _faces(dim, shape, RNG)
    got = O.mk_advective_m_fluxdiv(g, umac, m, [np.zeros(shape)] * dim)
    for a in range(dim):
        assert abs(np.sum(got[a])) < 1e-12 * np.sum(np.abs(got[a]))
    # divergence-free velocity advecting itself: sum u . div(u u) = 0
    if dim == 2:
        psi = RNG.standard_normal(shape)
        umac = [(_roll(psi, 1, 1) - psi) / g.dx, -(_roll(psi, 0, 1) - psi) / g.dx]
    else:
        A = [RNG.standard_normal(shape) for _ in range(3)]
        umac = [((_roll(A[2], 1, 1) - A[2]) - (_roll(A[1], 2, 1) - A[1])) / g.dx,
                ((_roll(A[0], 2, 1) - A[0]) - (_roll(A[2], 0, 1) - A[2])) / g.dx,
                ((_roll(A[1], 0, 1) - A[1]) - (_roll(A[0], 1, 1) - A[0])) / g.dx]
    fd = O.mk_advective_m_fluxdiv(g, umac, umac, [np.zeros(shape)] * dim)
    e = sum(np.sum(u * f) for u, f in zip(umac, fd))
    scale = sum(np.sum(np.abs(u * f)) for u, f in zip(umac, fd))
    assert abs(e) < 1e-12 * scale
    # ... and with a velocity that is not divergence-free it does not vanish
    umac = _faces(dim, shape, RNG)
    fd = O.mk_advective_m_fluxdiv(g, umac, umac, [np.zeros(shape)] * dim)
    e = sum(np.sum(u * f) for u, f in zip(umac, fd))
    assert abs(e) > 1e-6 * sum(np.sum(np.abs(u * f)) for u, f in zip(umac, fd))


def test_grav_force_closed_form():
    g = O.geom(3, (4, 4, 4), 1.0)
    shape = (4, 4, 4)
    so, sn, f = _faces(3, shape, RNG), _faces(3, shape, RNG), _faces(3, shape, RNG)
    grav = (0.0, -9.8, 2.0)
    got = O.mk_grav_force(g, grav, so, sn, f)
    for d in range(3):
        assert np.allclose(got[d], f[d] + 0.5 * grav[d] * (so[d] + sn[d]), rtol=0, atol=1e-14)


# ---------------------------------------------------------------------------
# the assembled step

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
    ne = 1 if dim == 2 else 3
    st = dict(umac=_faces(dim, shape, rng, u_scale), rho_old=rho, rhotot_old=rhotot, pi=np.zeros(shape),
              eta=eta, eta_ed=O._edges(g, eta), kappa=np.ones(shape), Temp=temp, Temp_ed=O._edges(g, temp),
              diff_mass_fluxdiv=np.zeros((ns,) + shape), stoch_mass_fluxdiv=np.zeros((ns,) + shape),
              gradp_baro=None)
    return g, st


def _species(ns, **kw):
    d = dict(molmass=np.linspace(1.0, 2.0, ns), rhobar=np.linspace(1.0, 1.5, ns),
             Dbar=np.linspace(0.5, 1.0, ns * (ns - 1) // 2), k_B=1e-4, variance_coef_mass=0.0)
    d.update(kw)
    return O.species(ns, **d)


def test_rest_state_stays_at_rest():
    g, st = _state(2, (16, 16), 3, RNG, uniform=True)
    p = O.params(visc_type=2)
    out, i1, i2 = O.advance_timestep_inertial(g, p, _species(3), st, 0.01, 0.0, variance_coef_mom=0.0, visc_coef=0.8)
    assert all(np.max(np.abs(u)) == 0.0 for u in out["umac"])
    assert np.max(np.abs(out["rho_new"] - st["rho_old"])) < 1e-15
    assert np.max(np.abs(out["pi"])) < 1e-14


@pytest.mark.parametrize("dim,n", [(2, (16, 16)), (3, (8, 8, 8))])
def test_uniform_fluid_under_gravity_moves_with_g_dt(dim, n):
    g, st = _state(dim, n, 2, RNG, uniform=True)
    grav = (0.3, -0.2, 0.5)[:dim] + (0.0,) * (3 - dim)
    dt = 0.02
    out, _, _ = O.advance_timestep_inertial(g, O.params(visc_type=2), _species(2), st, dt, 0.0, grav=grav,
                                            variance_coef_mom=0.0, visc_coef=0.8)
    for d in range(dim):
        assert np.max(np.abs(out["umac"][d] - grav[d] * dt)) < 1e-9 * abs(grav[d] * dt), d


@pytest.mark.parametrize("visc_type", [2, -1])
def test_shear_wave_decays_by_crank_nicolson_factor(visc_type):
    n = (16, 32)
    g, st = _state(2, n, 2, RNG, uniform=True)
    shape = _shape(n)
    j = np.arange(n[1])[:, None] * np.ones((1, n[0]))
    st["umac"] = [0.1 * np.sin(2 * np.pi * j / n[1]), np.zeros(shape)]
    dt, eta = 0.01, 0.8
    out, i1, i2 = O.advance_timestep_inertial(g, O.params(visc_type=visc_type), _species(2), st, dt, 0.0,
                                              variance_coef_mom=0.0, visc_coef=eta)
    rho = float(np.mean(st["rhotot_old"]))
    lam = 4.0 / g.dx ** 2 * math.sin(math.pi / n[1]) ** 2
    R = (rho / dt - eta * lam / 2) / (rho / dt + eta * lam / 2)
    assert np.max(np.abs(out["umac"][0] - R * st["umac"][0])) < 1e-8 * 0.1
    assert np.max(np.abs(out["umac"][1])) < 1e-10
    # an explicit viscous half (factor 1 - dt nu lam) is far off
    assert abs(R - (1 - dt * eta * lam / rho)) > 1e-4


@pytest.mark.parametrize("dim,n,ns", [(2, (16, 16), 3), (3, (8, 8, 8), 2)])
def test_mass_and_momentum_are_conserved(dim, n, ns):
    g, st = _state(dim, n, ns, RNG, u_scale=0.05)
    shape = _shape(n)
    sp = _species(ns, variance_coef_mass=1e-8, k_B=1e-9)  # noise small against the mean state
    dt = 2e-4  # explicit mass diffusion: D dt / dx^2 well below 1/2
    wm = O.fill_m_stochastic(11, dim, shape, 1, symmetric=True)[0]
    wm2 = O.fill_m_stochastic(12, dim, shape, 1, symmetric=True)[0]
    wmass = O.fill_normal(13, dim * ns * np.prod(shape)).reshape((1, dim, ns) + shape)
    wmass2 = O.fill_normal(14, dim * ns * np.prod(shape)).reshape((1, dim, ns) + shape)
    out, i1, i2 = O.advance_timestep_inertial(g, O.params(visc_type=2), sp, st, dt, 0.0, variance_coef_mom=1.0,
                                              visc_coef=0.8, W_m=wm, W_m_new=wm2, W_mass=wmass, W_mass_new=wmass2)
    assert i1["status"] == 0 and i2["status"] == 0
    for c in range(ns):
        m0, m1 = np.sum(st["rho_old"][c]), np.sum(out["rho_new"][c])
        assert abs(m1 - m0) < 1e-12 * abs(m0)
    rf0 = O.average_cc_to_face(g, st["rhotot_old"])
    rf1 = O.average_cc_to_face(g, out["rhotot_new"])
    for d in range(dim):
        p0 = np.sum(rf0[d] * st["umac"][d])
        p1 = np.sum(rf1[d] * out["umac"][d])
        scale = np.sum(np.abs(rf1[d] * out["umac"][d]))
        assert abs(p1 - p0) < 1e-7 * scale, (d, p0, p1, scale)
    # the velocity changed (noise and flow), and stayed finite
    assert all(np.isfinite(u).all() for u in out["umac"])
    assert max(np.max(np.abs(u - u0)) for u, u0 in zip(out["umac"], st["umac"])) > 1e-6


def test_uniform_flow_advects_composition_and_momentum():
    """A uniform flow U0 along x carries a composition wave and a transverse
    velocity wave (both functions of x only, equal rhobar: no volume change,
    vanishing diffusion and viscosity): over one step each changes by
    -dt U0 times its centred difference, to the trapezoidal rule's
    O(dt U0^2 k^2) -- a dropped 1/2 on either advective average, or a wrong
    sign, is off by order one."""
    n = (32, 16)
    shape = _shape(n)
    g = O.geom(2, n, 1.0 / n[0])
    x = (np.arange(n[0]) + 0.5) / n[0] * np.ones((n[1], 1))
    xf = np.arange(n[0]) / n[0] * np.ones((n[1], 1))
    r1 = 0.5 + 0.1 * np.sin(2 * np.pi * x)
    rho = np.stack([r1, 1.0 - r1])
    U0, dt = 0.5, 1e-3
    st = dict(umac=[np.full(shape, U0), 0.01 * np.sin(2 * np.pi * xf)], rho_old=rho,
              rhotot_old=O.compute_rhotot(rho), pi=np.zeros(shape), eta=np.full(shape, 1e-8),
              eta_ed=O._edges(g, np.full(shape, 1e-8)), kappa=np.ones(shape), Temp=np.ones(shape),
              Temp_ed=O._edges(g, np.ones(shape)), diff_mass_fluxdiv=np.zeros((2,) + shape),
              stoch_mass_fluxdiv=np.zeros((2,) + shape), gradp_baro=None)
    sp = _species(2, rhobar=[1.0, 1.0], Dbar=[1e-8])
    out, _, _ = O.advance_timestep_inertial(g, O.params(visc_type=-1), sp, st, dt, 0.0, variance_coef_mom=0.0,
                                            visc_coef=1e-8)
    cd = lambda a: (_roll(a, 0, 1) - _roll(a, 0, -1)) / (2 * g.dx)
    want = -U0 * cd(r1)
    got = (out["rho_new"][0] - r1) / dt
    assert np.max(np.abs(got - want)) < 2e-2 * np.max(np.abs(want))
    v0 = st["umac"][1]
    want = -U0 * cd(v0)
    got = (out["umac"][1] - v0) / dt
    assert np.max(np.abs(got - want)) < 2e-2 * np.max(np.abs(want))
    assert np.max(np.abs(out["umac"][0] - U0)) < 1e-9
