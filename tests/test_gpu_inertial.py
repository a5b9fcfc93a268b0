"""The inertial time step (advance_timestep_inertial.f90:58-990) and its
advection and gravity routines on the GPU against the oracle, through the C
ABI.  The building blocks are bit-exact; the whole step is compared field by
field over two consecutive steps, with the host noise stream (the oracle draws
the same numbers) so that the momentum and mass noise enter both sides
identically, to 1e-11 relative: the gmres inner products (and so the
solutions) round differently from the oracle's OpenMP sums.  A larger 3D step
with the device noise checks conservation."""
import numpy as np
import pytest

from lowmachmixtures_amd import stokes as S
from oracle import oracle as O
from tests.helpers import bit_equal, relerr

pytestmark = pytest.mark.gpu


def _shape(n):
    return tuple(reversed(n))


@pytest.mark.parametrize("dim,n", [(2, (32, 16)), (3, (8, 16, 8))])
def test_advective_s_fluxdiv_bit_exact(dim, n):
    rng = np.random.default_rng(201)
    shape = _shape(n)
    g = O.geom(dim, n, 0.25)
    umac = [rng.standard_normal(shape) for _ in range(dim)]
    s_fc = [rng.standard_normal((4,) + shape) for _ in range(dim)]
    upd = rng.standard_normal((4,) + shape)
    want = upd.copy()
    want[1:3] = O.mk_advective_s_fluxdiv(g, umac, [s[1:3] for s in s_fc], upd[1:3])
    with S.Context(dim, n, 0.25) as ctx:
        u = ctx.faces(umac)
        sf = [ctx.field(S.FACE_X + d, 4, s_fc[d]) for d in range(dim)]
        su = ctx.field(S.CELL, 4, upd)
        S.mk_advective_s_fluxdiv(ctx, u, sf, su, start_comp=1, num_comp=2)
        got = su.download()
    assert bit_equal(got, want), relerr(got, want)


@pytest.mark.parametrize("dim,n", [(2, (32, 16)), (3, (8, 16, 8))])
def test_advective_m_fluxdiv_and_grav_bit_exact(dim, n):
    rng = np.random.default_rng(203)
    shape = _shape(n)
    g = O.geom(dim, n, 0.5)
    umac, m, upd = ([rng.standard_normal(shape) for _ in range(dim)] for _ in range(3))
    want = O.mk_advective_m_fluxdiv(g, umac, m, upd)
    so, sn = ([1.0 + rng.random(shape) for _ in range(dim)] for _ in range(2))
    grav = (0.4, -9.8, 1.5)
    want_g = O.mk_grav_force(g, grav, so, sn, want)
    with S.Context(dim, n, 0.5) as ctx:
        u, mm, up = ctx.faces(umac), ctx.faces(m), ctx.faces(upd)
        S.mk_advective_m_fluxdiv(ctx, u, mm, up)
        got = [f.download() for f in up]
        S.mk_grav_force(ctx, up, ctx.faces(so), ctx.faces(sn), grav)
        got_g = [f.download() for f in up]
    assert bit_equal(got, want), relerr(got, want)
    assert bit_equal(got_g, want_g), relerr(got_g, want_g)


def _knobs(ns):
    return dict(molmass=list(np.linspace(1.0, 2.5, ns)), rhobar=list(np.linspace(1.0, 1.4, ns)),
                Dbar=list(np.linspace(0.4, 1.0, ns * (ns - 1) // 2)), k_B=1e-9, variance_coef_mass=1e-8)


STEP_CASES = [
    dict(dim=2, n=(32, 16), ns=3, visc=2, eta_model=0, grav=(0.0, -2.0, 0.0)),
    dict(dim=2, n=(16, 16), ns=2, visc=-1, eta_model=1, grav=(0.5, 0.0, 0.0)),
    dict(dim=3, n=(16, 8, 8), ns=4, visc=2, eta_model=0, grav=(0.0, 0.0, -1.0)),
    dict(dim=3, n=(8, 8, 16), ns=2, visc=-2, eta_model=1, grav=(0.0, 0.0, 0.0)),
    # barodiffusion (the lagged grad pi in the mass fluxes): k_B on the scale of
    # the molar masses, and a state near equilibrium; the pressure, the least
    # well determined output of the solves, now feeds the next fluxes, so the
    # second step agrees to the solver's tolerance rather than to 1e-11
    dict(dim=2, n=(16, 32), ns=3, visc=-1, eta_model=0, grav=(0.0, -1.0, 0.0), baro=2, kn=dict(k_B=1.0, variance_coef_mass=1e-17), vmom=1e-9, amp=0.01, u=1e-4, tol=1e-9),
]


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
def test_inertial_step_matches_oracle(case):
    c = STEP_CASES[case]
    dim, n, ns = c["dim"], c["n"], c["ns"]
    shape = _shape(n)
    rng = np.random.default_rng(300 + case)
    dx = 1.0 / n[0]
    g = O.geom(dim, n, dx)
    kn = dict(_knobs(ns), barodiffusion_type=c.get("baro", 0), **c.get("kn", {}))
    sp = O.species(ns, **kn)
    p = O.params(visc_type=c["visc"])
    vmom, amp, us = c.get("vmom", 1.0), c.get("amp", 0.5), c.get("u", 0.05)
    rho = np.stack([0.5 + amp * rng.random(shape) for _ in range(ns)]) / ns
    if c["eta_model"] == 1:
        rho = rho / O.compute_rhotot(rho) * (0.9 + 0.05 * rng.random(shape))  # eta = -5 + 6 rhotot > 0
    rhotot = O.compute_rhotot(rho)
    temp = np.ones(shape)
    eta = np.full(shape, 0.8) if c["eta_model"] == 0 else 0.8 * (-5.0 + 6.0 * rhotot)
    ed = O._edges(g, eta)
    st = dict(umac=[us * rng.standard_normal(shape) for _ in range(dim)], rho_old=rho, rhotot_old=rhotot,
              pi=0.1 * rng.standard_normal(shape), eta=eta, eta_ed=ed, kappa=np.ones(shape), Temp=temp,
              Temp_ed=O._edges(g, temp), diff_mass_fluxdiv=0.01 * rng.standard_normal((ns,) + shape),
              stoch_mass_fluxdiv=0.01 * rng.standard_normal((ns,) + shape), gradp_baro=None)
    dt, seed = 2e-4, 40
    ip = dict(grav=c["grav"], variance_coef_mom=vmom, visc_coef=0.8, eta_model=c["eta_model"])
    N = int(np.prod(shape))
    mom = lambda s_: O.fill_m_stochastic(s_, dim, shape, 1, symmetric=True)[0]
    mass = lambda s_: O.fill_normal(s_, dim * ns * N).reshape((1, dim, ns) + shape)
    # the oracle over two steps
    W_m, W_mass = mom(7), mass(8)
    want = []
    s_o = dict(st)
    for istep in range(2):
        out, i1, i2 = O.advance_timestep_inertial(
            g, p, sp, s_o, dt, istep * dt, grav=ip["grav"], variance_coef_mom=vmom, visc_coef=0.8,
            eta_model=c["eta_model"], W_m=W_m, W_m_new=mom(seed + 2 * istep), W_mass=W_mass,
            W_mass_new=mass(seed + 2 * istep + 1))
        want.append((out, i1, i2))
        W_m, W_mass = out["W_m"], out["W_mass"]
        s_o = dict(st, umac=out["umac"], rho_old=out["rho_new"], rhotot_old=out["rhotot_new"], pi=out["pi"],
                   eta=out["eta"], eta_ed=out["eta_ed"], kappa=out["kappa"],
                   diff_mass_fluxdiv=out["diff_mass_fluxdiv"], stoch_mass_fluxdiv=out["stoch_mass_fluxdiv"])
    # the GPU over the same two steps
    with S.Context(dim, n, dx, visc_type=c["visc"], stoch_stress_form=1) as ctx:
        S.set_species(ctx, ns, **kn)
        S.init_m_stochastic(ctx, 1)
        S.fill_m_stochastic(ctx, 7)
        S.init_mass_stochastic(ctx, 1)
        S.fill_mass_stochastic(ctx, 8)
        umac = ctx.faces(st["umac"])
        r_old, r_new = ctx.field(S.CELL, ns, rho), ctx.field(S.CELL, ns)
        rt_old, rt_new = ctx.field(S.CELL, 1, rhotot), ctx.field(S.CELL, 1)
        f_pi, f_eta, f_kap = ctx.field(S.CELL, 1, st["pi"]), ctx.field(S.CELL, 1, eta), ctx.field(S.CELL, 1, st["kappa"])
        f_ed, f_T, f_Ted = ctx.edges(ed), ctx.field(S.CELL, 1, temp), ctx.edges(st["Temp_ed"])
        f_diff = ctx.field(S.CELL, ns, st["diff_mass_fluxdiv"])
        f_stoch = ctx.field(S.CELL, ns, st["stoch_mass_fluxdiv"])
        gpb = ctx.faces() if c.get("baro") else None
        prm = S.inertial_params(noise=1, noise_seed=seed, **ip)
        for istep in range(2):
            i1, i2 = S.advance_timestep_inertial(ctx, umac, r_old, r_new, rt_old, rt_new, gpb, f_pi, f_eta, f_ed,
                                                 f_kap, f_T, f_Ted, f_diff, f_stoch, dt, istep * dt, istep, prm)
            out, o1, o2 = want[istep]
            assert i1["status"] == 0 and i2["status"] == 0
            assert (i1["total_iter"], i2["total_iter"]) == (o1["total_iter"], o2["total_iter"]), istep
            got = dict(umac=[f.download() for f in umac], rho_new=r_new.download(), rhotot_new=rt_new.download(),
                       pi=f_pi.download(), eta=f_eta.download(), diff_mass_fluxdiv=f_diff.download(),
                       stoch_mass_fluxdiv=f_stoch.download())
            if gpb is not None:
                got["gradp_baro"] = [f.download() for f in gpb]
            # the two solves' inner products sum in another order than the
            # oracle's, so the step agrees to rounding, not bit for bit
            for k, v in got.items():
                assert relerr(v, out[k]) < c.get("tol", 1e-11), (istep, k, relerr(v, out[k]))
            # the new state becomes the old one
            for src, dst in ((r_new, r_old), (rt_new, rt_old)):
                S.field_copy(dst, src)


def test_inertial_step_64cubed_device_noise_conserves_mass():
    """A 3D box of 64^3 cells, 4 species, the device noise stream: species
    mass is conserved to rounding, both solves converge, the state stays
    finite and moves."""
    dim, n, ns = 3, (64, 64, 64), 4
    shape = _shape(n)
    rng = np.random.default_rng(404)
    z = (np.arange(n[2]) + 0.5)[:, None, None] / n[2] * np.ones(shape)
    c1 = 0.5 + 0.4 * np.tanh((z - 0.5) / 0.05)
    frac = np.stack([c1 * 0.6, c1 * 0.4, (1 - c1) * 0.5, (1 - c1) * 0.5])
    rho = frac * 1.0
    rhotot = rho.sum(axis=0)
    kn = _knobs(ns)
    with S.Context(dim, n, 1.0 / n[0], visc_type=2, stoch_stress_form=1) as ctx:
        S.set_species(ctx, ns, **kn)
        S.init_m_stochastic(ctx, 1)
        S.fill_m_stochastic_device(ctx, 3, 0)
        S.init_mass_stochastic(ctx, 1)
        S.fill_mass_stochastic_device(ctx, 3, 1)
        umac = ctx.faces([0.01 * rng.standard_normal(shape) for _ in range(dim)])
        r_old, r_new = ctx.field(S.CELL, ns, rho), ctx.field(S.CELL, ns)
        rt_old, rt_new = ctx.field(S.CELL, 1, rhotot), ctx.field(S.CELL, 1)
        ones = np.ones(shape)
        f_pi, f_eta, f_kap = ctx.field(S.CELL, 1), ctx.field(S.CELL, 1, 0.8 * ones), ctx.field(S.CELL, 1, ones)
        f_ed, f_T, f_Ted = ctx.edges([0.8 * ones] * 3), ctx.field(S.CELL, 1, ones), ctx.edges([ones] * 3)
        f_diff, f_stoch = ctx.field(S.CELL, ns), ctx.field(S.CELL, ns)
        prm = S.inertial_params(noise=0, noise_seed=3, grav=(0.0, 0.0, -1.0), visc_coef=0.8)
        i1, i2 = S.advance_timestep_inertial(ctx, umac, r_old, r_new, rt_old, rt_new, None, f_pi, f_eta, f_ed, f_kap,
                                             f_T, f_Ted, f_diff, f_stoch, 1e-4, 0.0, 0, prm)
        out = r_new.download()
        u = [f.download() for f in umac]
    assert i1["status"] == 0 and i2["status"] == 0
    for c in range(ns):
        assert abs(out[c].sum() - rho[c].sum()) < 1e-11 * rho[c].sum()
    assert all(np.isfinite(x).all() for x in u) and np.isfinite(out).all()
    assert np.max(np.abs(out - rho)) > 0


OD_CASES = [
    dict(dim=2, n=(32, 16), ns=3, visc=2, eta_model=0, alg=1),
    dict(dim=2, n=(16, 16), ns=2, visc=-1, eta_model=1, alg=2),
    dict(dim=3, n=(16, 8, 8), ns=4, visc=1, eta_model=0, alg=2),
    dict(dim=3, n=(8, 8, 16), ns=2, visc=-2, eta_model=1, alg=1),
    # barodiffusion is 2D only (compute_baro_coef, mass_flux_utilities.f90:1291-1340);
    # with theta_alpha = 0 the lagged grad pi feeds back through S without
    # inertia to damp it, so the coupling (1/k_B) is kept weak
    dict(dim=2, n=(32, 16), ns=3, visc=2, eta_model=0, alg=2, baro=2, kn=dict(k_B=100.0, variance_coef_mass=1e-19),
         vmom=1e-11, amp=0.01, u=1e-4, tol=1e-9),
]


@pytest.mark.parametrize("case", range(len(OD_CASES)))
def test_overdamped_step_matches_oracle(case):
    """advance_timestep_overdamped.f90:60-700 over istep 1 (the noise of the
    first fill) and istep 2 (new noise drawn at the start), host noise on
    both sides.  No gravity: a uniform body force has no steady Stokes
    solution on a periodic box."""
    c = OD_CASES[case]
    dim, n, ns, alg = c["dim"], c["n"], c["ns"], c["alg"]
    shape = _shape(n)
    rng = np.random.default_rng(500 + case)
    dx = 1.0 / n[0]
    g = O.geom(dim, n, dx)
    kn = dict(_knobs(ns), barodiffusion_type=c.get("baro", 0), **c.get("kn", {}))
    sp = O.species(ns, **kn)
    p = O.params(visc_type=c["visc"])
    vmom, amp, us = c.get("vmom", 1.0), c.get("amp", 0.5), c.get("u", 0.05)
    rho = np.stack([0.5 + amp * rng.random(shape) for _ in range(ns)]) / ns
    if c["eta_model"] == 1:
        rho = rho / O.compute_rhotot(rho) * (0.9 + 0.05 * rng.random(shape))
    rhotot = O.compute_rhotot(rho)
    temp = np.ones(shape)
    eta = np.full(shape, 0.8) if c["eta_model"] == 0 else 0.8 * (-5.0 + 6.0 * rhotot)
    ed = O._edges(g, eta)
    st = dict(umac=[us * rng.standard_normal(shape) for _ in range(dim)], rho_old=rho, rhotot_old=rhotot,
              pi=0.1 * rng.standard_normal(shape), eta=eta, eta_ed=ed, kappa=np.ones(shape), Temp=temp,
              Temp_ed=O._edges(g, temp), diff_mass_fluxdiv=np.zeros((ns,) + shape),
              stoch_mass_fluxdiv=np.zeros((ns,) + shape), gradp_baro=None)
    dt, seed = 2e-4, 60
    N = int(np.prod(shape))
    mom = lambda s_: O.fill_m_stochastic(s_, dim, shape, alg, symmetric=True)
    mass = lambda s_: O.fill_normal(s_, alg * dim * ns * N).reshape((alg, dim, ns) + shape)
    W_m, W_mass = mom(9), mass(10)
    want = []
    s_o = dict(st)
    for istep in (1, 2):
        out, i1, i2 = O.advance_timestep_overdamped(
            g, p, sp, s_o, dt, (istep - 1) * dt, istep, alg, variance_coef_mom=vmom, visc_coef=0.8,
            eta_model=c["eta_model"], W_m=W_m, W_mass=W_mass, W_m_new=mom(seed + 2 * istep),
            W_mass_new=mass(seed + 2 * istep + 1))
        want.append((out, i1, i2))
        W_m, W_mass = out["W_m"], out["W_mass"]
        s_o = dict(st, umac=out["umac"], rho_old=out["rho_new"], rhotot_old=out["rhotot_new"], pi=out["pi"],
                   eta=out["eta"], eta_ed=out["eta_ed"], kappa=out["kappa"])
    with S.Context(dim, n, dx, visc_type=c["visc"], stoch_stress_form=1) as ctx:
        S.set_species(ctx, ns, **kn)
        S.init_m_stochastic(ctx, alg)
        S.fill_m_stochastic(ctx, 9)
        S.init_mass_stochastic(ctx, alg)
        S.fill_mass_stochastic(ctx, 10)
        umac = ctx.faces(st["umac"])
        r_old, r_new = ctx.field(S.CELL, ns, rho), ctx.field(S.CELL, ns)
        rt_old, rt_new = ctx.field(S.CELL, 1, rhotot), ctx.field(S.CELL, 1)
        f_pi, f_eta, f_kap = ctx.field(S.CELL, 1, st["pi"]), ctx.field(S.CELL, 1, eta), ctx.field(S.CELL, 1, st["kappa"])
        f_ed, f_T, f_Ted = ctx.edges(ed), ctx.field(S.CELL, 1, temp), ctx.edges(st["Temp_ed"])
        f_diff, f_stoch = ctx.field(S.CELL, ns), ctx.field(S.CELL, ns)
        gpb = ctx.faces() if c.get("baro") else None
        prm = S.inertial_params(noise=1, noise_seed=seed, algorithm_type=alg, variance_coef_mom=vmom, visc_coef=0.8,
                                eta_model=c["eta_model"])
        for k, istep in enumerate((1, 2)):
            i1, i2 = S.advance_timestep_overdamped(ctx, umac, r_old, r_new, rt_old, rt_new, gpb, f_pi, f_eta, f_ed,
                                                   f_kap, f_T, f_Ted, f_diff, f_stoch, dt, (istep - 1) * dt, istep,
                                                   prm)
            out, o1, o2 = want[k]
            assert i1["status"] == 0 and i2["status"] == 0
            assert (i1["total_iter"], i2["total_iter"]) == (o1["total_iter"], o2["total_iter"]), istep
            got = dict(umac=[f.download() for f in umac], rho_old=r_old.download(), rho_new=r_new.download(),
                       rhotot_new=rt_new.download(), pi=f_pi.download(), eta=f_eta.download(),
                       diff_mass_fluxdiv=f_diff.download(), stoch_mass_fluxdiv=f_stoch.download())
            if gpb is not None:
                got["gradp_baro"] = [f.download() for f in gpb]
            for key, v in got.items():
                assert relerr(v, out[key]) < c.get("tol", 1e-11), (istep, key, relerr(v, out[key]))
            for src, dst in ((r_new, r_old), (rt_new, rt_old)):
                S.field_copy(dst, src)


def test_overdamped_step_checks_its_arguments():
    with S.Context(2, (16, 16), 1.0 / 16) as ctx:
        S.set_species(ctx, 2, **_knobs(2))
        S.init_m_stochastic(ctx, 1)
        S.fill_m_stochastic(ctx, 1)
        S.init_mass_stochastic(ctx, 1)
        S.fill_mass_stochastic(ctx, 2)
        z = np.zeros((16, 16))
        umac = ctx.faces([z, z])
        fs = [ctx.field(S.CELL, 2, np.full((2, 16, 16), 0.5)), ctx.field(S.CELL, 2), ctx.field(S.CELL, 1, z + 1.0),
              ctx.field(S.CELL, 1)]
        rest = [None, ctx.field(S.CELL, 1), ctx.field(S.CELL, 1, z + 1.0), ctx.edges([z + 1.0]),
                ctx.field(S.CELL, 1, z + 1.0), ctx.field(S.CELL, 1, z + 1.0), ctx.edges([z + 1.0]),
                ctx.field(S.CELL, 2), ctx.field(S.CELL, 2)]
        for alg in (0, 3):
            with pytest.raises(S.LmmError):
                S.advance_timestep_overdamped(ctx, umac, *fs, *rest, 1e-3, 0.0, 1,
                                              S.inertial_params(algorithm_type=alg))
        # two noise stages wanted, one initialised
        with pytest.raises(S.LmmError):
            S.advance_timestep_overdamped(ctx, umac, *fs, *rest, 1e-3, 0.0, 1, S.inertial_params(algorithm_type=2))
        i1, i2 = S.advance_timestep_overdamped(ctx, umac, *fs, *rest, 1e-3, 0.0, 1,
                                               S.inertial_params(algorithm_type=1))
        assert i1["status"] == 0 and i2["status"] == 0
