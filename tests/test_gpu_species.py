"""Species mass-flux chain on the GPU against the CPU oracle, through the C ABI.

The per-cell routines (correct_rho_with_drho, convert_cons_to_prim,
compute_mixture_properties, compute_Gama, compute_chi, compute_rhoWchi,
compute_Lonsager), the diffusive flux with its temperature and
barodiffusion pieces, the stochastic flux and the whole compute_mass_fluxdiv
chain must equal the oracle bit for bit on the default path
(Dbar2chi_iterative, choldc).  The use_lapack path restates LAPACK's results
by other algorithms and is held to 1e-12 relative.  At a larger size the
chain is checked through properties that do not need the oracle: species
fluxes sum to zero (correct_flux), divergences sum to zero over the periodic
box, rho comes back as rho_c - drho.
"""
import numpy as np
import pytest

from lowmachmixtures_amd import stokes as S
from oracle import oracle as O
from tests.helpers import bit_equal

pytestmark = pytest.mark.gpu


def _knobs(ns, rng, **kw):
    nd = ns * (ns - 1) // 2
    d = dict(molmass=list(rng.uniform(0.5, 3.0, ns)) + [1.0] * (8 - ns),
             Dbar=list(rng.uniform(0.5, 2.0, nd)) + [1.0] * (28 - nd))
    d.update(kw)
    return d


def _rho(ns, shape, rng, tiny=True):
    rho = rng.uniform(0.2, 1.0, (ns,) + shape)
    if tiny:
        rho[ns - 1].flat[3] = 1e-17  # below fraction_tolerance * rhotot: corrected
    return rho


def _ctx(dim, n, dx, ns, **knobs):
    ctx = S.Context(dim, n, dx)
    S.set_species(ctx, ns, **knobs)
    return ctx


CASES = [(2, (16, 8), 2), (2, (8, 8), 3), (3, (8, 4, 4), 4), (3, (4, 4, 4), 8)]


@pytest.mark.parametrize("dim,n,ns", CASES)
@pytest.mark.parametrize("ideal", [1, 0])
def test_cell_routines_bit_exact(dim, n, ns, ideal):
    rng = np.random.default_rng(100 + ns + 10 * ideal)
    shape = tuple(reversed(n))
    kn = _knobs(ns, rng, is_ideal_mixture=ideal, fraction_tolerance=1e-9,
                H_offdiag=list(rng.normal(size=28)), H_diag=list(rng.uniform(0.5, 1.0, 8)),
                Dtherm=list(rng.uniform(-1, 1, 8)))
    sp = O.species(ns, **kn)
    g = O.geom(dim, n, 0.5)
    rho = _rho(ns, shape, rng)
    T = rng.uniform(0.8, 1.2, shape)
    rc_w, drho_w = O.correct_rho_with_drho(g, sp, rho)
    rt_w, x_w, mt_w = O.convert_cons_to_prim(g, sp, rc_w)
    Db_w, Dt_w, H_w = O.compute_mixture_properties(g, sp, shape)
    G_w = O.compute_Gama(g, sp, x_w, H_w)
    chi_w, z_w = O.compute_chi(g, sp, rc_w, rt_w, x_w, Db_w, Dt_w, T)
    rw_w = O.compute_rhoWchi(g, sp, rc_w, chi_w)
    L_w = O.compute_Lonsager(g, sp, rc_w, rt_w, mt_w, chi_w)
    with _ctx(dim, n, 0.5, ns, **kn) as ctx:
        f_rho = ctx.field(S.CELL, ns, rho)
        f_drho = ctx.field(S.CELL, ns)
        S.correct_rho_with_drho(ctx, f_rho, f_drho)
        assert bit_equal(f_rho.download(), rc_w) and bit_equal(f_drho.download(), drho_w)
        f_rt, f_x, f_mt = ctx.field(S.CELL), ctx.field(S.CELL, ns), ctx.field(S.CELL)
        S.convert_cons_to_prim(ctx, f_rho, f_rt, f_x, f_mt)
        assert bit_equal(f_rt.download(), rt_w) and bit_equal(f_x.download(), x_w)
        assert bit_equal(f_mt.download(), mt_w)
        f_Db, f_Dt, f_H = ctx.field(S.CELL, ns * ns), ctx.field(S.CELL, ns), ctx.field(S.CELL, ns * ns)
        S.compute_mixture_properties(ctx, f_rho, f_rt, f_Db, f_Dt, f_H, None)
        assert bit_equal(f_Db.download(), Db_w) and bit_equal(f_Dt.download(), Dt_w)
        assert bit_equal(f_H.download(), H_w)
        f_G = ctx.field(S.CELL, ns * ns)
        S.compute_Gama(ctx, f_rho, f_rt, f_x, f_mt, f_H, f_G)
        assert bit_equal(f_G.download(), G_w)
        f_chi, f_z, f_T = ctx.field(S.CELL, ns * ns), ctx.field(S.CELL, ns), ctx.field(S.CELL, 1, T)
        S.compute_chi(ctx, f_rho, f_rt, f_x, f_chi, f_Db, f_Dt, f_T, f_z)
        assert bit_equal(f_chi.download(), chi_w) and bit_equal(f_z.download(), z_w)
        f_rw = ctx.field(S.CELL, ns * ns)
        S.compute_rhoWchi(ctx, f_rho, f_rt, f_chi, f_rw)
        assert bit_equal(f_rw.download(), rw_w)
        f_L = ctx.field(S.CELL, ns * ns)
        S.compute_Lonsager(ctx, f_rho, f_rt, f_x, f_mt, f_chi, f_G, f_L)
        assert bit_equal(f_L.download(), L_w)


@pytest.mark.parametrize("inverse_type", [1, 2])
@pytest.mark.parametrize("ns", [3, 6])
def test_lapack_path_close(ns, inverse_type):
    rng = np.random.default_rng(7 + ns + inverse_type)
    dim, n = 2, (8, 8)
    shape = (8, 8)
    kn = _knobs(ns, rng, use_lapack=1, inverse_type=inverse_type)
    sp = O.species(ns, **kn)
    g = O.geom(dim, n)
    rho = _rho(ns, shape, rng, tiny=False)
    T = np.ones(shape)
    rt_w, x_w, mt_w = O.convert_cons_to_prim(g, sp, rho)
    Db_w, Dt_w, _ = O.compute_mixture_properties(g, sp, shape)
    chi_w, _ = O.compute_chi(g, sp, rho, rt_w, x_w, Db_w, Dt_w, T)
    L_w = O.compute_Lonsager(g, sp, rho, rt_w, mt_w, chi_w)
    with _ctx(dim, n, 1.0, ns, **kn) as ctx:
        f_rho = ctx.field(S.CELL, ns, rho)
        f_rt, f_x, f_mt = ctx.field(S.CELL, 1, rt_w), ctx.field(S.CELL, ns, x_w), ctx.field(S.CELL, 1, mt_w)
        f_Db, f_Dt = ctx.field(S.CELL, ns * ns, Db_w), ctx.field(S.CELL, ns, Dt_w)
        f_chi, f_z = ctx.field(S.CELL, ns * ns), ctx.field(S.CELL, ns)
        S.compute_chi(ctx, f_rho, f_rt, f_x, f_chi, f_Db, f_Dt, ctx.field(S.CELL, 1, T), f_z)
        chi = f_chi.download()
        f_L = ctx.field(S.CELL, ns * ns)
        S.compute_Lonsager(ctx, f_rho, f_rt, f_x, f_mt, ctx.field(S.CELL, ns * ns, chi_w), None, f_L)
        L = f_L.download()
    assert np.abs(chi - chi_w).max() <= 1e-12 * np.abs(chi_w).max()
    assert np.abs(L - L_w).max() <= 1e-12 * np.abs(L_w).max()


@pytest.mark.parametrize("dim,n,ns", [(2, (16, 8), 3), (3, (8, 8, 4), 4)])
def test_diffusive_flux_temperature_piece_bit_exact(dim, n, ns):
    rng = np.random.default_rng(31 + dim)
    shape = tuple(reversed(n))
    kn = _knobs(ns, rng, Dtherm=list(rng.uniform(-1, 1, 8)), is_ideal_mixture=0,
                H_offdiag=list(0.1 * rng.normal(size=28)), H_diag=list(rng.uniform(0.1, 0.3, 8)))
    sp = O.species(ns, **kn)
    g = O.geom(dim, n, 0.25)
    rho = _rho(ns, shape, rng, tiny=False)
    T = rng.uniform(0.8, 1.2, shape)
    rt, x, mt = O.convert_cons_to_prim(g, sp, rho)
    Db, Dt, H = O.compute_mixture_properties(g, sp, shape)
    G = O.compute_Gama(g, sp, x, H)
    chi, z = O.compute_chi(g, sp, rho, rt, x, Db, Dt, T)
    rw = O.compute_rhoWchi(g, sp, rho, chi)
    ft0 = [rng.standard_normal((ns,) + shape) for _ in range(dim)]
    want_div, want_ft = O.diffusive_mass_flux_full(g, sp, rho, rt, x, rw, G, ft0, Temp=T, zeta_by_Temp=z)
    with _ctx(dim, n, 0.25, ns, **kn) as ctx:
        ft = [ctx.field(S.FACE_X + d, ns, ft0[d]) for d in range(dim)]
        div = ctx.field(S.CELL, ns)
        S.diffusive_mass_fluxdiv(ctx, ns, ctx.field(S.CELL, ns, rho), ctx.field(S.CELL, 1, rt),
                                 ctx.field(S.CELL, ns, x), ctx.field(S.CELL, ns * ns, rw),
                                 ctx.field(S.CELL, ns * ns, G), div, ctx.field(S.CELL, 1, T),
                                 ctx.field(S.CELL, ns, z), None, ft)
        assert bit_equal(div.download(), want_div)
        assert bit_equal([f.download() for f in ft], want_ft)


def test_diffusive_flux_barodiffusion_bit_exact():
    rng = np.random.default_rng(41)
    dim, n, ns = 2, (16, 8), 3
    shape = (8, 16)
    kn = _knobs(ns, rng, barodiffusion_type=1, rhobar=list(rng.uniform(0.8, 1.5, 8)), k_B=0.7)
    sp = O.species(ns, **kn)
    g = O.geom(dim, n)
    rho = _rho(ns, shape, rng, tiny=False)
    T = rng.uniform(0.8, 1.2, shape)
    gpb = [rng.standard_normal(shape) for _ in range(dim)]
    rt, x, mt = O.convert_cons_to_prim(g, sp, rho)
    Db, Dt, H = O.compute_mixture_properties(g, sp, shape)
    G = O.compute_Gama(g, sp, x, H)
    chi, z = O.compute_chi(g, sp, rho, rt, x, Db, Dt, T)
    rw = O.compute_rhoWchi(g, sp, rho, chi)
    ft0 = [np.zeros((ns,) + shape) for _ in range(dim)]
    want_div, want_ft = O.diffusive_mass_flux_full(g, sp, rho, rt, x, rw, G, ft0, Temp=T, zeta_by_Temp=z,
                                                   gradp_baro=gpb)
    with _ctx(dim, n, 1.0, ns, **kn) as ctx:
        ft = [ctx.field(S.FACE_X + d, ns, ft0[d]) for d in range(dim)]
        div = ctx.field(S.CELL, ns)
        S.diffusive_mass_fluxdiv(ctx, ns, ctx.field(S.CELL, ns, rho), ctx.field(S.CELL, 1, rt),
                                 ctx.field(S.CELL, ns, x), ctx.field(S.CELL, ns * ns, rw),
                                 ctx.field(S.CELL, ns * ns, G), div, ctx.field(S.CELL, 1, T),
                                 ctx.field(S.CELL, ns, z), [ctx.field(S.FACE_X + d, 1, gpb[d]) for d in range(dim)], ft)
        assert bit_equal(div.download(), want_div)
        assert bit_equal([f.download() for f in ft], want_ft)
    # compute_baro_coef_3d is not written in the reference: rejected
    with _ctx(3, (4, 4, 4), 1.0, ns, **kn) as ctx:
        z3 = ctx.field(S.CELL, ns)
        with pytest.raises(S.LmmError):
            S.diffusive_mass_fluxdiv(ctx, ns, ctx.field(S.CELL, ns), ctx.field(S.CELL), ctx.field(S.CELL, ns),
                                     ctx.field(S.CELL, ns * ns), ctx.field(S.CELL, ns * ns), ctx.field(S.CELL, ns),
                                     ctx.field(S.CELL), z3, ctx.faces(), [ctx.field(S.FACE_X + d, ns)
                                                                         for d in range(3)])


def _mass_noise(seed, n_rngs, dim, ns, shape):
    cnt = n_rngs * dim * ns * int(np.prod(shape))
    return O.fill_normal(seed, cnt).reshape((n_rngs, dim, ns) + shape)


@pytest.mark.parametrize("dim,n,ns,n_rngs", [(2, (16, 8), 3, 2), (3, (8, 4, 8), 4, 1)])
@pytest.mark.parametrize("increment", [False, True])
def test_stochastic_mass_fluxdiv_bit_exact(dim, n, ns, n_rngs, increment):
    rng = np.random.default_rng(51 + dim)
    shape = tuple(reversed(n))
    kn = _knobs(ns, rng, k_B=0.3, variance_coef_mass=2.0)
    sp = O.species(ns, **kn)
    g = O.geom(dim, n, 0.5)
    rho = _rho(ns, shape, rng, tiny=False)
    rt, x, mt = O.convert_cons_to_prim(g, sp, rho)
    Db, Dt, H = O.compute_mixture_properties(g, sp, shape)
    G = O.compute_Gama(g, sp, x, H)
    chi, _ = O.compute_chi(g, sp, rho, rt, x, Db, Dt, np.ones(shape))
    W = _mass_noise(77, n_rngs, dim, ns, shape)
    weights = [0.75, -0.5][:n_rngs]
    ft0 = [rng.standard_normal((ns,) + shape) for _ in range(dim)]
    d0 = rng.standard_normal((ns,) + shape)
    want, want_ft = O.stochastic_mass_fluxdiv(g, sp, rho, rt, x, mt, chi, G, ft0, 0.01, W, weights,
                                              stoch_fluxdiv=d0, increment=increment)
    with _ctx(dim, n, 0.5, ns, **kn) as ctx:
        S.init_mass_stochastic(ctx, n_rngs)
        S.fill_mass_stochastic(ctx, 77)
        ft = [ctx.field(S.FACE_X + d, ns, ft0[d]) for d in range(dim)]
        div = ctx.field(S.CELL, ns, d0)
        S.stochastic_mass_fluxdiv(ctx, ctx.field(S.CELL, ns, rho), ctx.field(S.CELL, 1, rt), ctx.field(S.CELL, ns, x),
                                  ctx.field(S.CELL, 1, mt), ctx.field(S.CELL, ns * ns, chi),
                                  ctx.field(S.CELL, ns * ns, G), div, ft, 0.01, weights, increment=increment)
        assert bit_equal(div.download(), want)
        assert bit_equal([f.download() for f in ft], want_ft)


MASS_CASES = [
    dict(dim=2, n=(16, 8), ns=3, kn=dict()),
    dict(dim=3, n=(8, 8, 4), ns=4, kn=dict(is_ideal_mixture=0, Dtherm=[0.3, -0.1, 0.2, 0.0, 0, 0, 0, 0],
                                          H_offdiag=[0.05] * 28, H_diag=[0.2] * 8)),
    dict(dim=2, n=(8, 8), ns=2, kn=dict(barodiffusion_type=1, is_nonisothermal=0)),
    dict(dim=3, n=(4, 4, 8), ns=8, kn=dict(fraction_tolerance=1e-8)),
    dict(dim=2, n=(8, 16), ns=4, kn=dict(variance_coef_mass=0.0)),
]


@pytest.mark.parametrize("case", range(len(MASS_CASES)))
def test_compute_mass_fluxdiv_bit_exact(case):
    c = MASS_CASES[case]
    dim, n, ns = c["dim"], c["n"], c["ns"]
    rng = np.random.default_rng(61 + case)
    shape = tuple(reversed(n))
    kn = _knobs(ns, rng, k_B=0.5, **c["kn"])
    sp = O.species(ns, **kn)
    g = O.geom(dim, n, 0.5)
    rho = _rho(ns, shape, rng)
    T = rng.uniform(0.9, 1.1, shape)
    gpb = [rng.standard_normal(shape) for _ in range(dim)] if kn.get("barodiffusion_type") else None
    W = _mass_noise(5, 1, dim, ns, shape)
    want_rho, want_diff, want_stoch, want_ft = O.compute_mass_fluxdiv(g, sp, rho, T, 0.02, W, [1.0], gradp_baro=gpb)
    with _ctx(dim, n, 0.5, ns, **kn) as ctx:
        S.init_mass_stochastic(ctx, 1)
        S.fill_mass_stochastic(ctx, 5)
        f_rho = ctx.field(S.CELL, ns, rho)
        diff, stoch = ctx.field(S.CELL, ns), ctx.field(S.CELL, ns, np.full((ns,) + shape, 7.0))
        ft = [ctx.field(S.FACE_X + d, ns, np.full((ns,) + shape, 3.0)) for d in range(dim)]
        gp = [ctx.field(S.FACE_X + d, 1, gpb[d]) for d in range(dim)] if gpb is not None else None
        S.compute_mass_fluxdiv(ctx, f_rho, gp, diff, stoch, ctx.field(S.CELL, 1, T), ft, 0.02, 0.0, [1.0])
        assert bit_equal(f_rho.download(), want_rho)
        assert bit_equal(diff.download(), want_diff)
        assert bit_equal(stoch.download(), want_stoch)
        assert bit_equal([f.download() for f in ft], want_ft)
    if kn.get("variance_coef_mass", 1.0) == 0.0:
        assert not want_stoch.any()


def test_device_mass_noise_matches_counter_stream():
    dim, n, ns, n_rngs = 3, (8, 4, 4), 3, 2
    shape = (4, 4, 8)
    rng = np.random.default_rng(3)
    kn = _knobs(ns, rng)
    sp = O.species(ns, **kn)
    g = O.geom(dim, n)
    rho = _rho(ns, shape, rng, tiny=False)
    rt, x, mt = O.convert_cons_to_prim(g, sp, rho)
    Db, Dt, H = O.compute_mixture_properties(g, sp, shape)
    chi, _ = O.compute_chi(g, sp, rho, rt, x, Db, Dt, np.ones(shape))
    G = O.compute_Gama(g, sp, x, H)
    cnt = n_rngs * dim * ns * int(np.prod(shape))
    W = O.fill_normal_philox(11, 4, 0, cnt).reshape((n_rngs, dim, ns) + shape)
    ft0 = [np.zeros((ns,) + shape) for _ in range(dim)]
    want, _ = O.stochastic_mass_fluxdiv(g, sp, rho, rt, x, mt, chi, G, ft0, 0.1, W, [1.0, 0.5])
    with _ctx(dim, n, 1.0, ns, **kn) as ctx:
        S.init_mass_stochastic(ctx, n_rngs)
        S.fill_mass_stochastic_device(ctx, 11, 4)
        div = ctx.field(S.CELL, ns)
        S.stochastic_mass_fluxdiv(ctx, ctx.field(S.CELL, ns, rho), ctx.field(S.CELL, 1, rt), ctx.field(S.CELL, ns, x),
                                  ctx.field(S.CELL, 1, mt), ctx.field(S.CELL, ns * ns, chi),
                                  ctx.field(S.CELL, ns * ns, G), div, [ctx.field(S.FACE_X + d, ns) for d in range(dim)],
                                  0.1, [1.0, 0.5])
        got = div.download()
    # the device's log / cos / sin in Box-Muller may differ from the host's in
    # the last bits (the stream itself is checked to 1e-14 in test_gpu_rng.py);
    # everything after the noise is linear
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_chain_errors():
    with S.Context(2, (8, 8)) as ctx:
        rho = ctx.field(S.CELL, 3)
        with pytest.raises(S.LmmError):  # no species knobs yet
            S.correct_rho_with_drho(ctx, rho, ctx.field(S.CELL, 3))
        S.set_species(ctx, 3)
        out = [ctx.field(S.CELL, 3) for _ in range(2)]
        ft = [ctx.field(S.FACE_X + d, 3) for d in range(2)]
        with pytest.raises(S.LmmError):  # mass noise not initialised
            S.compute_mass_fluxdiv(ctx, rho, None, out[0], out[1], ctx.field(S.CELL), ft, 0.1)
        with pytest.raises(S.LmmError):  # nspecies outside 2..8
            S.set_species(ctx, 9)


@pytest.mark.parametrize("dim,n,ns", [(3, (128, 128, 128), 4), (2, (512, 512), 8)])
def test_compute_mass_fluxdiv_conservation_full_size(dim, n, ns):
    """Mass conservation of the whole chain at full size (no oracle)."""
    rng = np.random.default_rng(71)
    shape = tuple(reversed(n))
    kn = _knobs(ns, rng, Dtherm=list(rng.uniform(-0.5, 0.5, 8)))
    grid = np.meshgrid(*[np.arange(k) for k in shape], indexing="ij")
    base = 1.0 + 0.3 * np.sin(2 * np.pi * grid[-1] / shape[-1])
    rho = np.stack([base * (0.5 + 0.4 * np.cos(2 * np.pi * (s + 1) * grid[0] / shape[0])) for s in range(ns)])
    rho[0].flat[::97] = 0.0  # corrected to fraction_tolerance * rhotot, then restored
    T = 1.0 + 0.1 * np.cos(2 * np.pi * grid[-1] / shape[-1])
    with _ctx(dim, n, 1.0 / n[0], ns, **kn) as ctx:
        S.init_mass_stochastic(ctx, 2)
        S.fill_mass_stochastic_device(ctx, 2024, 0)
        f_rho = ctx.field(S.CELL, ns, rho)
        diff, stoch = ctx.field(S.CELL, ns), ctx.field(S.CELL, ns)
        ft = [ctx.field(S.FACE_X + d, ns) for d in range(dim)]
        S.compute_mass_fluxdiv(ctx, f_rho, None, diff, stoch, ctx.field(S.CELL, 1, T), ft, 1e-3, 0.0, [1.0, 0.5])
        back = f_rho.download()
        dv, sv = diff.download(), stoch.download()
        fts = [f.download() for f in ft]
    # rho_c - drho gives rho back to rounding; cells that were not corrected exactly
    assert np.abs(back - rho).max() <= 1e-15
    assert np.array_equal(back[1:], rho[1:])
    for v in (dv, sv):
        scale = np.abs(v).max()
        assert scale > 0 and np.isfinite(v).all()
        assert np.abs(v.sum(axis=0)).max() <= 1e-12 * scale          # species sum (correct_flux)
        assert np.abs(v.reshape(ns, -1).sum(axis=1)).max() <= 1e-9 * scale * v[0].size ** 0.5  # periodic box
    for f in fts:
        assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).max()
