"""ctypes binding of the CPU oracle (``liblmm_oracle.so``).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker or the CPU baseline; the product package never does.

The C sources restate the reference's Fortran (see the header comment of
each ``orc_*.c`` file for the file:line each function follows).  Arrays
crossing this binding are deduplicated periodic fields: numpy float64 arrays
of shape ``(nz, ny, nx)`` (``(ny, nx)`` in 2D), x fastest.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

DPTR = C.POINTER(C.c_double)


class Params(C.Structure):
    _fields_ = [
        ("visc_type", C.c_int),
        ("precon_type", C.c_int),
        ("p_norm_weight", C.c_double),
        ("scale_factor", C.c_double),
        ("mg_max_vcycles", C.c_int),
        ("mg_minwidth", C.c_int),
        ("mg_nsmooths_down", C.c_int),
        ("mg_nsmooths_up", C.c_int),
        ("mg_nsmooths_bottom", C.c_int),
        ("mg_rel_tol", C.c_double),
        ("stag_mg_max_vcycles", C.c_int),
        ("stag_mg_minwidth", C.c_int),
        ("stag_mg_nsmooths_down", C.c_int),
        ("stag_mg_nsmooths_up", C.c_int),
        ("stag_mg_nsmooths_bottom", C.c_int),
        ("stag_mg_smoother", C.c_int),
        ("stag_mg_omega", C.c_double),
        ("stag_mg_rel_tol", C.c_double),
        ("gmres_rel_tol", C.c_double),
        ("gmres_abs_tol", C.c_double),
        ("gmres_max_outer", C.c_int),
        ("gmres_max_inner", C.c_int),
        ("gmres_max_iter", C.c_int),
        ("gmres_min_iter", C.c_int),
        ("verbose", C.c_int),
    ]


class Geom(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("dx", C.c_double)]


class Coef(C.Structure):
    _fields_ = [
        ("alpha", DPTR * 3),
        ("beta", DPTR),
        ("beta_ed", DPTR * 3),
        ("gamma", DPTR),
    ]


class GmresInfo(C.Structure):
    _fields_ = [
        ("total_iter", C.c_int),
        ("outer_iter", C.c_int),
        ("norm_pre_rhs", C.c_double),
        ("norm_resid", C.c_double),
        ("norm_b", C.c_double),
        ("norm_resid_true", C.c_double),
        ("status", C.c_int),
        ("dup_mismatch", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# reference defaults: src_gmres/probin_gmres.f90:119-157
DEFAULTS = dict(
    visc_type=1, precon_type=1, p_norm_weight=1.0, scale_factor=1.0,
    mg_max_vcycles=1, mg_minwidth=2, mg_nsmooths_down=2, mg_nsmooths_up=2,
    mg_nsmooths_bottom=8, mg_rel_tol=1e-9,
    stag_mg_max_vcycles=1, stag_mg_minwidth=2, stag_mg_nsmooths_down=2,
    stag_mg_nsmooths_up=2, stag_mg_nsmooths_bottom=8, stag_mg_smoother=1,
    stag_mg_omega=1.0, stag_mg_rel_tol=1e-9,
    gmres_rel_tol=1e-9, gmres_abs_tol=0.0, gmres_max_outer=20, gmres_max_inner=5,
    gmres_max_iter=100, gmres_min_iter=1, verbose=0,
)


def build():
    """Compile liblmm_oracle.so (and the reference RNG when the tree is mounted)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir("/root/reference/src_random"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblmm_oracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _LIB.orc_cc_inner_prod.restype = C.c_double
        _LIB.orc_inv_norm_dist.restype = C.c_double
        _LIB.orc_mt_next.restype = C.c_uint32
        _LIB.orc_mt_normal.restype = C.c_double
        _LIB.orc_mt_uniform.restype = C.c_double
    return _LIB


def ref_rng():
    """The reference's own RNGs.c compiled into oracle/_ref, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "librngs_ref.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
        _REF.genrandbase.restype = C.c_ulong
        _REF.InvNormDist.restype = C.c_double
        _REF.srandgen.argtypes = [C.c_ulong]
        _REF.InvNormDist.argtypes = [C.c_double]
    return _REF


def params(**kw) -> Params:
    d = dict(DEFAULTS)
    d.update(kw)
    p = Params()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def geom(dim, n, dx=1.0) -> Geom:
    g = Geom()
    g.dim = dim
    nn = list(n) + [1] * (3 - len(n))
    for d in range(3):
        g.n[d] = nn[d]
    g.dx = dx
    return g


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "oracle needs C-contiguous float64"
    return a.ctypes.data_as(DPTR)


def _arr3(lst):
    arr = (DPTR * 3)()
    for i, a in enumerate(lst):
        if a is not None:
            arr[i] = _p(a)
    return arr


def coef(alpha, beta, beta_ed, gamma):
    c = Coef()
    for i, a in enumerate(alpha):
        c.alpha[i] = _p(a)
    c.beta = _p(beta)
    for i, a in enumerate(beta_ed):
        c.beta_ed[i] = _p(a)
    c.gamma = _p(gamma)
    c._keep = (alpha, beta, beta_ed, gamma)
    return c


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def _copy(lst):
    return [np.ascontiguousarray(a, dtype=np.float64).copy() for a in lst]


def stag_applyop(g, visc_type, cf, theta_alpha, phi, color=0, Lphi=None):
    out = _copy(Lphi) if Lphi is not None else [np.zeros_like(a) for a in phi]
    phi = _copy(phi)
    _check(lib().orc_stag_applyop(C.byref(g), visc_type, C.byref(cf), C.c_double(theta_alpha),
                                  _arr3(phi), _arr3(out), color), "stag_applyop")
    return out


def diffusive_m_fluxdiv(g, visc_type, m_update, umac, eta, eta_ed, kappa):
    """diffusive_m_fluxdiv.f90:16-73: m_update - stag_applyop_level(umac) with alpha_fc = 0
    (multifab_sub_sub_c of :58-61)."""
    dim = len(umac)
    cf = coef([np.zeros_like(u) for u in umac], eta, eta_ed, kappa)
    lphi = stag_applyop(g, visc_type, cf, 1.0, umac)
    return [np.asarray(m, dtype=np.float64) - l for m, l in zip(m_update[:dim], lphi)]


def stag_mg_smooth(g, p, cf, theta_alpha, phi, rhs, nsweeps):
    phi = _copy(phi)
    rhs = _copy(rhs)
    _check(lib().orc_stag_mg_smooth(C.byref(g), C.byref(p), C.byref(cf), C.c_double(theta_alpha),
                                    _arr3(phi), _arr3(rhs), nsweeps), "stag_mg_smooth")
    return phi


def _coarse_shape(a):
    return tuple(s // 2 for s in a.shape)


def stag_restriction(g, fine, simple=False):
    fine = _copy(fine)
    coarse = [np.zeros(_coarse_shape(a)) for a in fine]
    _check(lib().orc_stag_restriction(C.byref(g), _arr3(fine), _arr3(coarse), int(simple)),
           "stag_restriction")
    return coarse


def stag_prolongation(g, fine, coarse):
    fine = _copy(fine)
    coarse = _copy(coarse)
    _check(lib().orc_stag_prolongation(C.byref(g), _arr3(fine), _arr3(coarse)), "stag_prolongation")
    return fine


def cc_restriction(g, fine):
    fine = np.ascontiguousarray(fine, dtype=np.float64)
    coarse = np.zeros(_coarse_shape(fine))
    _check(lib().orc_cc_restriction(C.byref(g), _p(fine), _p(coarse)), "cc_restriction")
    return coarse


def edge_restriction(g, fine):
    fine = _copy(fine)
    coarse = [np.zeros(_coarse_shape(a)) for a in fine]
    _check(lib().orc_edge_restriction(C.byref(g), _arr3(fine), _arr3(coarse)), "edge_restriction")
    return coarse


def nodal_restriction(g, fine):
    fine = np.ascontiguousarray(fine, dtype=np.float64)
    coarse = np.zeros(_coarse_shape(fine))
    _check(lib().orc_nodal_restriction(C.byref(g), _p(fine), _p(coarse)), "nodal_restriction")
    return coarse


def compute_div(g, u, div=None, increment=False):
    u = _copy(u)
    out = np.zeros_like(u[0]) if div is None else np.ascontiguousarray(div, dtype=np.float64).copy()
    _check(lib().orc_compute_div(C.byref(g), _arr3(u), _p(out), int(increment)), "compute_div")
    return out


def compute_grad(g, p_cc):
    p_cc = np.ascontiguousarray(p_cc, dtype=np.float64)
    out = [np.zeros_like(p_cc) for _ in range(g.dim)]
    _check(lib().orc_compute_grad(C.byref(g), _p(p_cc), _arr3(out)), "compute_grad")
    return out


def stag_inner_prod(g, m1, m2):
    m1, m2 = _copy(m1), _copy(m2)
    out = (C.c_double * 3)()
    _check(lib().orc_stag_inner_prod(C.byref(g), _arr3(m1), _arr3(m2), out), "stag_inner_prod")
    return [out[d] for d in range(g.dim)]


def cc_inner_prod(g, a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_cc_inner_prod(C.byref(g), _p(a), _p(b))


def sum_umac_press(g, p_cc, u):
    p_cc = np.ascontiguousarray(p_cc, dtype=np.float64)
    u = _copy(u)
    avp = C.c_double()
    avu = (C.c_double * 3)()
    _check(lib().orc_sum_umac_press(C.byref(g), _p(p_cc), _arr3(u), C.byref(avp), avu), "sum_umac_press")
    return avp.value, [avu[d] for d in range(g.dim)]


def stag_mg_solver(g, p, cf, theta_alpha, rhs, phi=None):
    rhs = _copy(rhs)
    phi = _copy(phi) if phi is not None else [np.zeros_like(a) for a in rhs]
    _check(lib().orc_stag_mg_solver(C.byref(g), C.byref(p), C.byref(cf), C.c_double(theta_alpha),
                                    _arr3(phi), _arr3(rhs)), "stag_mg_solver")
    return phi


def macproject(g, p, alphainv, mac_rhs):
    alphainv = _copy(alphainv)
    mac_rhs = np.ascontiguousarray(mac_rhs, dtype=np.float64)
    phi = np.zeros_like(mac_rhs)
    _check(lib().orc_macproject(C.byref(g), C.byref(p), _arr3(alphainv), _p(mac_rhs), _p(phi)),
           "macproject")
    return phi


def apply_matrix(g, p, cf, theta_alpha, x_u, x_p):
    x_u = _copy(x_u)
    x_p = np.ascontiguousarray(x_p, dtype=np.float64)
    b_u = [np.zeros_like(a) for a in x_u]
    b_p = np.zeros_like(x_p)
    _check(lib().orc_apply_matrix(C.byref(g), C.byref(p), C.byref(cf), C.c_double(theta_alpha),
                                  _arr3(x_u), _p(x_p), _arr3(b_u), _p(b_p)), "apply_matrix")
    return b_u, b_p


def apply_precon(g, p, cf, theta_alpha, b_u, b_p):
    b_u = _copy(b_u)
    b_p = np.ascontiguousarray(b_p, dtype=np.float64)
    x_u = [np.zeros_like(a) for a in b_u]
    x_p = np.zeros_like(b_p)
    _check(lib().orc_apply_precon(C.byref(g), C.byref(p), C.byref(cf), C.c_double(theta_alpha),
                                  _arr3(b_u), _p(b_p), _arr3(x_u), _p(x_p)), "apply_precon")
    return x_u, x_p


def gmres(g, p, cf, theta_alpha, b_u, b_p, x_u=None, x_p=None):
    b_u = _copy(b_u)
    b_p = np.ascontiguousarray(b_p, dtype=np.float64).copy()
    x_u = _copy(x_u) if x_u is not None else [np.zeros_like(a) for a in b_u]
    x_p = np.ascontiguousarray(x_p, dtype=np.float64).copy() if x_p is not None else np.zeros_like(b_p)
    info = GmresInfo()
    lib().orc_gmres(C.byref(g), C.byref(p), C.byref(cf), C.c_double(theta_alpha), _arr3(b_u),
                    _p(b_p), _arr3(x_u), _p(x_p), C.byref(info))
    return x_u, x_p, info.as_dict()


def least_squares(i, m, H, cs, sn, s):
    """In-place Givens update of column i (1-based) of the (m+1) x m matrix H."""
    for a in (H, cs, sn, s):
        assert a.dtype == np.float64
    lib().orc_least_squares(i, m, H.ctypes.data_as(DPTR), cs.ctypes.data_as(DPTR),
                            sn.ctypes.data_as(DPTR), s.ctypes.data_as(DPTR))


def solve_utriangular(k, m, H, s):
    y = np.zeros(m)
    lib().orc_solve_utriangular(k, m, H.ctypes.data_as(DPTR), s.ctypes.data_as(DPTR),
                                y.ctypes.data_as(DPTR))
    return y


def average_cc_to_face(g, cc):
    cc = np.ascontiguousarray(cc, dtype=np.float64)
    out = [np.zeros_like(cc) for _ in range(g.dim)]
    _check(lib().orc_average_cc_to_face(C.byref(g), _p(cc), _arr3(out)), "average_cc_to_face")
    return out


def average_cc_to_node(g, cc):
    cc = np.ascontiguousarray(cc, dtype=np.float64)
    out = np.zeros_like(cc)
    _check(lib().orc_average_cc_to_node(C.byref(g), _p(cc), _p(out)), "average_cc_to_node")
    return out


def average_cc_to_edge(g, cc):
    cc = np.ascontiguousarray(cc, dtype=np.float64)
    out = [np.zeros_like(cc) for _ in range(3)]
    _check(lib().orc_average_cc_to_edge(C.byref(g), _p(cc), _arr3(out)), "average_cc_to_edge")
    return out


def stochastic_m_fluxdiv(g, W_cc, W_ed, eta, eta_ed, temp, temp_ed, variance, force, filtering_width=0):
    """W_cc: (dim, *shape); W_ed: list of (2, *shape) (1 node array in 2D, 3 edges in 3D)."""
    W_cc = np.ascontiguousarray(W_cc, dtype=np.float64)
    W_ed = _copy(W_ed)
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    temp = np.ascontiguousarray(temp, dtype=np.float64)
    eta_ed, temp_ed = _copy(eta_ed), _copy(temp_ed)
    force = _copy(force)
    _check(lib().orc_stochastic_m_fluxdiv(C.byref(g), _p(W_cc), _arr3(W_ed), _p(eta), _arr3(eta_ed),
                                          _p(temp), _arr3(temp_ed), C.c_double(variance),
                                          int(filtering_width), _arr3(force)),
           "stochastic_m_fluxdiv")
    return force


def diffusive_mass_fluxdiv(g, ns, rho, rhotot, molarconc, rhoWchi, Gama, flux_total, correct_flux=True):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    rhotot = np.ascontiguousarray(rhotot, dtype=np.float64)
    molarconc = np.ascontiguousarray(molarconc, dtype=np.float64)
    rhoWchi = np.ascontiguousarray(rhoWchi, dtype=np.float64)
    Gama = np.ascontiguousarray(Gama, dtype=np.float64)
    flux_total = _copy(flux_total)
    div = np.zeros_like(molarconc)
    _check(lib().orc_diffusive_mass_fluxdiv(C.byref(g), ns, _p(rho), _p(rhotot), _p(molarconc), _p(rhoWchi),
                                            _p(Gama), _p(div), _arr3(flux_total), int(correct_flux)),
           "diffusive_mass_fluxdiv")
    return div, flux_total


def fill_normal(seed, count):
    out = np.empty(int(count), dtype=np.float64)
    lib().orc_fill_normal(C.c_uint32(seed), C.c_int64(int(count)), out.ctypes.data_as(DPTR))
    return out


def philox4x32_10(ctr, key):
    """One Philox4x32-10 block (orc_rng.c): four 32-bit words."""
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def fill_m_stochastic(seed, dim, shape, n_rngs, symmetric, philox_step=None):
    """fill_m_stochastic (stochastic_m_fluxdiv.f90:556-609) on the deduplicated
    periodic layout, per stage: W_cc (dim components), then each node/edge
    array.  Non-symmetric: both components of a node/edge array drawn.
    Symmetric: W_cc drawn and scaled by sqrt(2) (multifab_fill_random with
    variance 2, multifab_fill_random.f90:58-60), component 1 of each node/edge
    array drawn and copied to component 2 (multifab_copy_c).  Values come from
    the host MT19937/Acklam stream (fill_normal), or the counter-based stream
    of fill_normal_philox(seed, philox_step, ...) when philox_step is given.
    Returns a list over stages of (W_cc (dim, *shape), [W_ed (2, *shape)])."""
    N = int(np.prod(shape))
    ne = 1 if dim == 2 else 3
    edraw = 1 if symmetric else 2
    per = (dim + edraw * ne) * N
    if philox_step is None:
        vals = fill_normal(seed, per * n_rngs)
    else:
        vals = fill_normal_philox(seed, philox_step, 0, per * n_rngs)
    out = []
    s2 = np.sqrt(2.0)
    for r in range(n_rngs):
        v = vals[r * per:(r + 1) * per]
        wcc = v[:dim * N].reshape((dim,) + tuple(shape)).copy()
        if symmetric:
            wcc = s2 * wcc
        wed = []
        for e in range(ne):
            base = dim * N + e * edraw * N
            c1 = v[base:base + N].reshape(shape)
            c2 = c1 if symmetric else v[base + N:base + 2 * N].reshape(shape)
            wed.append(np.stack([c1, c2]).copy())
        out.append((wcc, wed))
    return out


def fill_normal_philox(seed, step, v0, count):
    """Values v0 .. v0+count-1 of the counter-based Gaussian stream (orc_rng.c)."""
    out = np.empty(int(count), dtype=np.float64)
    lib().orc_fill_normal_philox(C.c_uint64(int(seed)), C.c_uint64(int(step)), C.c_int64(int(v0)),
                                 C.c_int64(int(count)), out.ctypes.data_as(DPTR))
    return out


def inv_norm_dist(p):
    return lib().orc_inv_norm_dist(C.c_double(p))


def compute_nlevs_mg(g, minwidth=2):
    return lib().orc_compute_nlevs_mg(C.byref(g), minwidth)


# ---------------- species mass-flux chain (orc_species.c) ----------------

MAX_SPECIES = 8


class Species(C.Structure):
    _fields_ = [
        ("nspecies", C.c_int),
        ("molmass", C.c_double * 8), ("rhobar", C.c_double * 8),
        ("Dbar", C.c_double * 28), ("Dtherm", C.c_double * 8),
        ("H_offdiag", C.c_double * 28), ("H_diag", C.c_double * 8),
        ("k_B", C.c_double), ("variance_coef_mass", C.c_double), ("fraction_tolerance", C.c_double),
        ("is_ideal_mixture", C.c_int), ("is_nonisothermal", C.c_int), ("use_lapack", C.c_int),
        ("inverse_type", C.c_int), ("chi_iterations", C.c_int),
        ("correct_flux", C.c_int), ("barodiffusion_type", C.c_int),
    ]


# probin_multispecies.f90:70-90 and probin_common.f90:210-220 defaults
SPECIES_DEFAULTS = dict(
    molmass=1.0, rhobar=1.0, Dbar=1.0, Dtherm=0.0, H_offdiag=0.0, H_diag=0.0,
    k_B=1.0, variance_coef_mass=1.0, fraction_tolerance=1e-13,
    is_ideal_mixture=1, is_nonisothermal=1, use_lapack=0, inverse_type=1, chi_iterations=10,
    correct_flux=1, barodiffusion_type=0,
)


def species(nspecies, **kw) -> Species:
    """Species knobs; array knobs take a scalar (every entry) or a sequence."""
    d = dict(SPECIES_DEFAULTS)
    d.update(kw)
    s = Species()
    s.nspecies = int(nspecies)
    for k, v in d.items():
        field = getattr(s, k)
        if isinstance(field, C.Array):
            vals = np.broadcast_to(np.asarray(v, dtype=np.float64), (len(field),)) if np.ndim(v) == 0 else v
            for i, x in enumerate(vals):
                field[i] = float(x)
        else:
            setattr(s, k, v)
    return s


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def sp_cell_chi(sp, rho, molarconc, D_bar, Temp=1.0, D_therm=None):
    """compute_chi_local on one cell: returns (chi (ns, ns), zeta_by_Temp)."""
    ns = sp.nspecies
    rho, x = _c(rho), _c(molarconc)
    Db = _c(np.asarray(D_bar).T)  # column-major
    Dt = _c(D_therm if D_therm is not None else np.zeros(ns))
    chi = np.zeros(ns * ns)
    z = np.zeros(ns)
    lib().orc_sp_chi(C.byref(sp), _p(rho), C.c_double(float(rho.sum())), _p(x), _p(Db), C.c_double(Temp),
                     _p(Dt), _p(chi), _p(z))
    return chi.reshape(ns, ns).T.copy(), z


def sp_cell_lonsager(sp, rho, rhotot, molmtot, chi):
    ns = sp.nspecies
    L = np.zeros(ns * ns)
    lib().orc_sp_lonsager(C.byref(sp), _p(_c(rho)), C.c_double(rhotot), C.c_double(molmtot),
                          _p(_c(np.asarray(chi).T)), _p(L))
    return L.reshape(ns, ns).T.copy()


def sp_cell_gama(sp, molarconc, H):
    ns = sp.nspecies
    G = np.zeros(ns * ns)
    lib().orc_sp_gama(C.byref(sp), _p(_c(molarconc)), _p(_c(np.asarray(H).T)), _p(G))
    return G.reshape(ns, ns).T.copy()


def sp_invert(A):
    ns = A.shape[0]
    a = _c(np.asarray(A).T).copy()
    _check(lib().orc_sp_invert(ns, _p(a)), "sp_invert")
    return a.reshape(ns, ns).T.copy()


def sp_pinv(A, tol):
    ns = A.shape[0]
    out = np.zeros(ns * ns)
    lib().orc_sp_pinv(ns, _p(_c(np.asarray(A).T)), C.c_double(tol), _p(out))
    return out.reshape(ns, ns).T.copy()


def correct_rho_with_drho(g, sp, rho):
    rho = _c(rho).copy()
    drho = np.zeros_like(rho)
    _check(lib().orc_correct_rho_with_drho(C.byref(g), C.byref(sp), _p(rho), _p(drho)), "correct_rho_with_drho")
    return rho, drho


def convert_cons_to_prim(g, sp, rho):
    rho = _c(rho)
    rhotot = np.zeros(rho.shape[1:])
    molarconc = np.zeros_like(rho)
    molmtot = np.zeros(rho.shape[1:])
    _check(lib().orc_convert_cons_to_prim(C.byref(g), C.byref(sp), _p(rho), _p(rhotot), _p(molarconc),
                                          _p(molmtot)), "convert_cons_to_prim")
    return rhotot, molarconc, molmtot


def compute_mixture_properties(g, sp, shape):
    ns = sp.nspecies
    D_bar = np.zeros((ns * ns,) + tuple(shape))
    D_therm = np.zeros((ns,) + tuple(shape))
    H = np.zeros((ns * ns,) + tuple(shape))
    _check(lib().orc_compute_mixture_properties(C.byref(g), C.byref(sp), _p(D_bar), _p(D_therm), _p(H)),
           "compute_mixture_properties")
    return D_bar, D_therm, H


def compute_Gama(g, sp, molarconc, Hessian):
    molarconc, Hessian = _c(molarconc), _c(Hessian)
    Gama = np.zeros_like(Hessian)
    _check(lib().orc_compute_Gama(C.byref(g), C.byref(sp), _p(molarconc), _p(Hessian), _p(Gama)), "compute_Gama")
    return Gama


def compute_chi(g, sp, rho, rhotot, molarconc, D_bar, D_therm, Temp):
    rho, rhotot, molarconc, D_bar, D_therm, Temp = map(_c, (rho, rhotot, molarconc, D_bar, D_therm, Temp))
    chi = np.zeros_like(D_bar)
    zeta = np.zeros_like(rho)
    _check(lib().orc_compute_chi(C.byref(g), C.byref(sp), _p(rho), _p(rhotot), _p(molarconc), _p(chi), _p(D_bar),
                                 _p(D_therm), _p(Temp), _p(zeta)), "compute_chi")
    return chi, zeta


def compute_rhoWchi(g, sp, rho, chi):
    rho, chi = _c(rho), _c(chi)
    out = np.zeros_like(chi)
    _check(lib().orc_compute_rhoWchi(C.byref(g), C.byref(sp), _p(rho), _p(chi), _p(out)), "compute_rhoWchi")
    return out


def compute_Lonsager(g, sp, rho, rhotot, molmtot, chi):
    rho, rhotot, molmtot, chi = map(_c, (rho, rhotot, molmtot, chi))
    out = np.zeros_like(chi)
    _check(lib().orc_compute_Lonsager(C.byref(g), C.byref(sp), _p(rho), _p(rhotot), _p(molmtot), _p(chi), _p(out)),
           "compute_Lonsager")
    return out


def diffusive_mass_flux_full(g, sp, rho, rhotot, molarconc, rhoWchi, Gama, flux_total, Temp=None,
                             zeta_by_Temp=None, gradp_baro=None):
    rho, rhotot, molarconc, rhoWchi, Gama = map(_c, (rho, rhotot, molarconc, rhoWchi, Gama))
    flux_total = _copy(flux_total)
    div = np.zeros_like(molarconc)
    T = _c(Temp) if Temp is not None else None
    z = _c(zeta_by_Temp) if zeta_by_Temp is not None else None
    gb = _copy(gradp_baro) if gradp_baro is not None else None
    _check(lib().orc_diffusive_mass_flux_full(C.byref(g), C.byref(sp), _p(rho), _p(rhotot), _p(molarconc),
                                              _p(rhoWchi), _p(Gama), _p(div), _p(T) if T is not None else None,
                                              _p(z) if z is not None else None,
                                              _arr3(gb) if gb is not None else None, _arr3(flux_total)),
           "diffusive_mass_flux_full")
    return div, flux_total


def stochastic_mass_fluxdiv(g, sp, rho, rhotot, molarconc, molmtot, chi, Gama, flux_total, dt, W, weights,
                            stoch_fluxdiv=None, increment=False):
    """W: array (n_rngs, dim, ns, *shape)."""
    rho, rhotot, molarconc, molmtot, chi, Gama, W = map(_c, (rho, rhotot, molarconc, molmtot, chi, Gama, W))
    w = _c(weights)
    flux_total = _copy(flux_total)
    div = _c(stoch_fluxdiv).copy() if stoch_fluxdiv is not None else np.zeros_like(rho)
    _check(lib().orc_stochastic_mass_fluxdiv(C.byref(g), C.byref(sp), _p(rho), _p(rhotot), _p(molarconc),
                                             _p(molmtot), _p(chi), _p(Gama), _p(div), _arr3(flux_total),
                                             C.c_double(dt), _p(W), int(W.shape[0]), _p(w), int(increment)),
           "stochastic_mass_fluxdiv")
    return div, flux_total


def compute_mass_fluxdiv(g, sp, rho, Temp, dt, W, weights, gradp_baro=None):
    """compute_mass_fluxdiv.f90:24-147; returns (rho after the call, diff_fluxdiv,
    stoch_fluxdiv, flux_total)."""
    rho = _c(rho).copy()
    T = _c(Temp)
    W = _c(W)
    w = _c(weights)
    shape = rho.shape[1:]
    ns = sp.nspecies
    flux_total = [np.zeros((ns,) + tuple(shape)) for _ in range(g.dim)]
    diff = np.zeros_like(rho)
    stoch = np.zeros_like(rho)
    gb = _copy(gradp_baro) if gradp_baro is not None else None
    _check(lib().orc_compute_mass_fluxdiv(C.byref(g), C.byref(sp), _p(rho), _arr3(gb) if gb is not None else None,
                                          _p(diff), _p(stoch), _p(T), _arr3(flux_total), C.c_double(dt), _p(W),
                                          int(W.shape[0]), _p(w)), "compute_mass_fluxdiv")
    return rho, diff, stoch, flux_total


# ---------------------------------------------------------------------------
# Inertial time step (advance_timestep_inertial.f90:58-990) composed from the
# oracle's routines.  Periodic box, centred advection (advection_type 0),
# barodiffusion_type 0 or 2; the wall, reservoir and ghost-cell calls of the
# reference do nothing on a periodic box.


def mk_advective_s_fluxdiv(g, umac, s_fc, s_update):
    """mk_advective_s_fluxdiv.f90:14-166: s_update - div(umac s_fc); s_fc[d] is
    (ncomp, *shape), s_update (ncomp, *shape)."""
    s_update = _c(s_update).copy()
    ncomp = s_update.shape[0]
    u = _copy(umac)
    s = _copy(s_fc)
    _check(lib().orc_mk_advective_s_fluxdiv(C.byref(g), ncomp, _arr3(u), _arr3(s), _p(s_update)),
           "mk_advective_s_fluxdiv")
    return s_update


def mk_advective_m_fluxdiv(g, umac, m, m_update):
    """mk_advective_m_fluxdiv.f90:16-331: m_update + (-div(m u)) per face component."""
    u, mm, out = _copy(umac), _copy(m), _copy(m_update)
    _check(lib().orc_mk_advective_m_fluxdiv(C.byref(g), _arr3(u), _arr3(mm), _arr3(out)),
           "mk_advective_m_fluxdiv")
    return out


def mk_grav_force(g, grav, s_old, s_new, m_force):
    """mk_grav_force.f90:17-161: m_force + 0.5 grav_d (s_old + s_new)."""
    gv = (C.c_double * 3)(*[float(x) for x in (list(grav) + [0.0, 0.0, 0.0])[:3]])
    so, sn, out = _copy(s_old), _copy(s_new), _copy(m_force)
    _check(lib().orc_mk_grav_force(C.byref(g), gv, _arr3(so), _arr3(sn), _arr3(out)), "mk_grav_force")
    return out


def _faces_comps(g, cc):
    """average_cc_to_face of every component of cc (ncomp, *shape) -> list over d of (ncomp, *shape)."""
    per = [average_cc_to_face(g, c) for c in cc]
    return [np.stack([per[c][d] for c in range(len(cc))]) for d in range(g.dim)]


def _edges(g, cc):
    return [average_cc_to_node(g, cc)] if g.dim == 2 else average_cc_to_edge(g, cc)


def compute_rhotot(rho):
    """mass_flux_utilities.f90:376-420: rhotot = sum over species, from 0."""
    t = np.zeros_like(rho[0])
    for c in range(rho.shape[0]):
        t = t + rho[c]
    return t


def compute_eta_kappa(g, rho, rhotot, visc_coef, eta_model):
    """compute_mixture_properties.f90:235-298 with compute_eta_2d/3d: eta =
    visc_coef (the default prob_type branch) or visc_coef*(-5 + 6 rhotot)
    (prob_type -12); kappa = 1; eta on nodes (2D) / edges (3D)."""
    if eta_model == 0:
        eta = np.full_like(rhotot, float(visc_coef))
    else:
        eta = visc_coef * (-5.0 + 6.0 * rhotot)
    return eta, _edges(g, eta), np.ones_like(rhotot)


def advance_timestep_inertial(g, p, sp, st, dt, time, grav=(0.0, 0.0, 0.0), variance_coef_mom=1.0,
                              visc_coef=1.0, eta_model=0, W_m=None, W_m_new=None, W_mass=None,
                              W_mass_new=None, filtering_width=0):
    """One inertial step (advance_timestep_inertial.f90:58-990), steps 2-7 in
    the reference's order of operations.  st holds the state: umac (dim
    faces), rho_old (ns, *shape), rhotot_old, pi, eta, eta_ed, kappa, Temp,
    Temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, gradp_baro (dim faces).
    W_m / W_mass: the momentum noise (W_cc, [W_ed]) and mass noise (1, dim, ns,
    *shape) the step starts with; W_m_new / W_mass_new: what fill_m_stochastic
    and fill_mass_stochastic draw in step 7.  Returns the new state and the
    two solves' info dicts."""
    dim = g.dim
    ns = sp.nspecies
    var_mass = sp.variance_coef_mass
    baro = sp.barodiffusion_type
    umac = _copy(st["umac"])
    rho_old = _c(st["rho_old"])
    rhotot_old = _c(st["rhotot_old"])
    pi = _c(st["pi"]).copy()
    eta, eta_ed, kappa = _c(st["eta"]), _copy(st["eta_ed"]), _c(st["kappa"])
    Temp, Temp_ed = _c(st["Temp"]), _copy(st["Temp_ed"])
    diff, stoch = _c(st["diff_mass_fluxdiv"]).copy(), _c(st["stoch_mass_fluxdiv"]).copy()
    gradp_baro = _copy(st["gradp_baro"]) if st.get("gradp_baro") is not None else [np.zeros_like(pi)
                                                                                     for _ in range(dim)]
    shape = pi.shape
    zf = [np.zeros(shape) for _ in range(dim)]
    theta_alpha = 1.0 / dt
    dv = 1.0
    for _ in range(dim):
        dv *= g.dx
    variance = math.sqrt(variance_coef_mom * 2.0 * sp.k_B / (dv * dt))

    def stoch_m(force, e, e_ed, W):
        wcc = 0.0 + 1.0 * W[0]
        wed = [0.0 + 1.0 * w for w in W[1]]
        return stochastic_m_fluxdiv(g, wcc, wed, e, e_ed, Temp, Temp_ed, variance, force,
                                    filtering_width=filtering_width)

    def mass_flux(rho):
        W = W_mass if W_mass is not None else np.zeros((1, dim, ns) + shape)
        return compute_mass_fluxdiv(g, sp, rho, Temp, dt, W, [1.0], gradp_baro if baro != 0 else None)

    # ---- step 2: predictor Euler step for rho (:169-257)
    rho_fc = _faces_comps(g, rho_old)
    rhotot_fc_old = average_cc_to_face(g, rhotot_old)
    ru = 0.0 + diff
    if var_mass != 0.0:
        ru = ru + stoch
    ru = mk_advective_s_fluxdiv(g, umac, rho_fc, ru)
    ru = ru * dt
    rho_new = rho_old + ru
    rhotot_new = compute_rhotot(rho_new)
    rho_new = (rho_new / rhotot_new) * rhotot_new
    rho_fc = _faces_comps(g, rho_new)
    rhotot_fc = average_cc_to_face(g, rhotot_new)

    # ---- step 4: predictor Crank-Nicolson step (:264-496)
    mold = [u * r for u, r in zip(umac, rhotot_fc_old)]
    v = [m * (1.0 / dt) for m in mold]
    gradpi = compute_grad(g, pi)
    if baro == 2:
        gradp_baro = _copy(gradpi)
    v = [a - b for a, b in zip(v, gradpi)]
    m_a = mk_advective_m_fluxdiv(g, umac, mold, zf)
    v = [a + b for a, b in zip(v, m_a)]
    m_d = diffusive_m_fluxdiv(g, p.visc_type, zf, umac, eta, eta_ed, kappa)
    m_d = [a * 0.5 for a in m_d]
    v = [a + b for a, b in zip(v, m_d)]
    m_s = _copy(zf)
    if variance_coef_mom != 0.0:
        m_s = stoch_m(m_s, eta, eta_ed, W_m)
    v = [a + b for a, b in zip(v, m_s)]
    if any(x != 0.0 for x in grav[:dim]):
        v = mk_grav_force(g, grav, rhotot_fc_old, rhotot_fc_old, v)
    eta, eta_ed, kappa = compute_eta_kappa(g, rho_new, rhotot_new, visc_coef, eta_model)
    rho_new, diff, stoch, _ = mass_flux(rho_new)
    diff = diff * -1.0
    if var_mass != 0.0:
        stoch = stoch * -1.0
    bp = np.zeros(shape)
    for c in range(ns):
        a = -1.0 / sp.rhobar[c]
        bp = bp + a * diff[c]
        if var_mass != 0.0:
            bp = bp + a * stoch[c]
    mt = [u * r for u, r in zip(umac, rhotot_fc)]
    mt = [m * (1.0 / dt) for m in mt]
    v = [a - b for a, b in zip(v, mt)]
    mt = diffusive_m_fluxdiv(g, p.visc_type, zf, umac, eta, eta_ed, kappa)
    mt = [a * 0.5 for a in mt]
    v = [a + b for a, b in zip(v, mt)]
    ru = 0.0 + diff
    if var_mass != 0.0:
        ru = ru + stoch
    divu = compute_div(g, umac)
    bp = bp + divu
    cf = coef(rhotot_fc, eta * 0.5, [e * 0.5 for e in eta_ed], kappa * 0.5)
    du, dpi, info1 = gmres(g, p, cf, theta_alpha, v, bp)
    p2 = Params.from_buffer_copy(p)
    p2.gmres_abs_tol = max(p.gmres_abs_tol, info1["norm_pre_rhs"] * p.gmres_rel_tol)
    umac = [u + d for u, d in zip(umac, du)]
    pi = pi + dpi
    mtemp = [u * r for u, r in zip(umac, rhotot_fc)]

    # ---- step 5: trapezoidal scalar corrector (:585-606)
    ru = mk_advective_s_fluxdiv(g, umac, rho_fc, ru)
    rho_new = rho_new + rho_old
    rho_new = rho_new * 0.5
    ru = ru * (dt / 2.0)
    rho_new = rho_new + ru
    rhotot_new = compute_rhotot(rho_new)
    rho_new = (rho_new / rhotot_new) * rhotot_new
    rho_fc = _faces_comps(g, rho_new)
    rhotot_fc = average_cc_to_face(g, rhotot_new)
    eta, eta_ed, kappa = compute_eta_kappa(g, rho_new, rhotot_new, visc_coef, eta_model)

    # ---- step 7: corrector Crank-Nicolson step (:641-957)
    v = [m * (1.0 / dt) for m in mold]
    gradpi = compute_grad(g, pi)
    if baro == 2:
        gradp_baro = _copy(gradpi)
    v = [a - b for a, b in zip(v, gradpi)]
    m_a = mk_advective_m_fluxdiv(g, umac, mtemp, m_a)
    m_a = [a * 0.5 for a in m_a]
    v = [a + b for a, b in zip(v, m_a)]
    v = [a + b for a, b in zip(v, m_d)]
    if variance_coef_mom != 0.0:
        m_s = stoch_m(m_s, eta, eta_ed, W_m)
        m_s = [a * 0.5 for a in m_s]
    v = [a + b for a, b in zip(v, m_s)]
    if any(x != 0.0 for x in grav[:dim]):
        v = mk_grav_force(g, grav, rhotot_fc_old, rhotot_fc, v)
    W_m, W_mass = W_m_new, W_mass_new
    rho_new, diff, stoch, _ = mass_flux(rho_new)
    diff = diff * -1.0
    if var_mass != 0.0:
        stoch = stoch * -1.0
    bp = np.zeros(shape)
    for c in range(ns):
        a = -1.0 / sp.rhobar[c]
        bp = bp + a * diff[c]
        if var_mass != 0.0:
            bp = bp + a * stoch[c]
    mt = [u * r for u, r in zip(umac, rhotot_fc)]
    mt = [m * (1.0 / dt) for m in mt]
    v = [a - b for a, b in zip(v, mt)]
    m_d = diffusive_m_fluxdiv(g, p.visc_type, zf, umac, eta, eta_ed, kappa)
    m_d = [a * 0.5 for a in m_d]
    v = [a + b for a, b in zip(v, m_d)]
    divu = compute_div(g, umac)
    bp = bp + divu
    cf = coef(rhotot_fc, eta * 0.5, [e * 0.5 for e in eta_ed], kappa * 0.5)
    du, dpi, info2 = gmres(g, p2, cf, theta_alpha, v, bp)
    umac = [u + d for u, d in zip(umac, du)]
    pi = pi + dpi
    out = dict(umac=umac, rho_new=rho_new, rhotot_new=rhotot_new, pi=pi, eta=eta, eta_ed=eta_ed, kappa=kappa,
               diff_mass_fluxdiv=diff, stoch_mass_fluxdiv=stoch, gradp_baro=gradp_baro, W_m=W_m, W_mass=W_mass)
    return out, info1, info2


def advance_timestep_overdamped(g, p, sp, st, dt, time, istep, algorithm_type, grav=(0.0, 0.0, 0.0),
                                variance_coef_mom=1.0, visc_coef=1.0, eta_model=0, restart=-1, W_m=None,
                                W_mass=None, W_m_new=None, W_mass_new=None):
    """One overdamped midpoint step (advance_timestep_overdamped.f90:60-700)
    in the reference's order of operations, algorithm_type 1 or 2.  st holds
    the state as for advance_timestep_inertial (its diff/stoch_mass_fluxdiv
    are not read).  W_m: the momentum noise held on entry, a list over stages
    of (W_cc, [W_ed]); W_mass: the mass noise (n_rngs, dim, ns, *shape);
    W_m_new / W_mass_new: what the fills at the start of the step draw
    (:158-165), used unless istep is 1 or restart+1.  Returns the new state
    (with the noise held at the end) and the two solves' info dicts."""
    alg = int(algorithm_type)
    if alg not in (1, 2):
        raise ValueError("algorithm_type is 1 or 2")
    dim = g.dim
    ns = sp.nspecies
    var_mass = sp.variance_coef_mass
    baro = sp.barodiffusion_type
    umac = _copy(st["umac"])
    rho_old = _c(st["rho_old"]).copy()
    rhotot_old = _c(st["rhotot_old"])
    pi = _c(st["pi"]).copy()
    eta, eta_ed, kappa = _c(st["eta"]), _copy(st["eta_ed"]), _c(st["kappa"])
    Temp, Temp_ed = _c(st["Temp"]), _copy(st["Temp_ed"])
    gradp_baro = _copy(st["gradp_baro"]) if st.get("gradp_baro") is not None else [np.zeros_like(pi)
                                                                                     for _ in range(dim)]
    shape = pi.shape
    if istep != 1 and istep != restart + 1:
        W_m, W_mass = W_m_new, W_mass_new
    if W_mass is None:
        W_mass = np.zeros((alg, dim, ns) + shape)
    dv = 1.0
    for _ in range(dim):
        dv *= g.dx
    weights = [1.0] + [0.0] * (alg - 1)

    def stoch_m(force, dt_s, w):
        variance = math.sqrt(variance_coef_mom * 2.0 * sp.k_B / (dv * dt_s))
        wcc = 0.0 + w[0] * W_m[0][0]
        wed = [0.0 + w[0] * x for x in W_m[0][1]]
        for r in range(1, len(w)):
            wcc = wcc + w[r] * W_m[r][0]
            wed = [a + w[r] * x for a, x in zip(wed, W_m[r][1])]
        return stochastic_m_fluxdiv(g, wcc, wed, eta, eta_ed, Temp, Temp_ed, variance, force)

    def stage(rho_s, dt_s, w, pp):
        nonlocal umac, pi, gradp_baro
        gradpi = compute_grad(g, pi)
        if baro == 2:
            gradp_baro = _copy(gradpi)
        v = [0.0 - x for x in gradpi]
        if variance_coef_mom != 0.0:
            v = stoch_m(v, dt_s, w)
        if any(x != 0.0 for x in grav[:dim]):
            v = mk_grav_force(g, grav, rf, rf, v)
        rho_s, diff, stoch, _ = compute_mass_fluxdiv(g, sp, rho_s, Temp, dt_s, W_mass, w,
                                                     gradp_baro if baro != 0 else None)
        diff = diff * -1.0
        if var_mass != 0.0:
            stoch = stoch * -1.0
        bp = np.zeros(shape)
        for c in range(ns):
            a = -1.0 / sp.rhobar[c]
            bp = bp + a * diff[c]
            if var_mass != 0.0:
                bp = bp + a * stoch[c]
        v = diffusive_m_fluxdiv(g, p.visc_type, v, umac, eta, eta_ed, kappa)
        bp = bp + compute_div(g, umac)
        du, dpi, info = gmres(g, pp, coef(rf, eta, eta_ed, kappa), 0.0, v, bp)
        umac = [u + d for u, d in zip(umac, du)]
        pi = pi + dpi
        return rho_s, diff, stoch, info

    def scalar(diff, stoch, dt_s):
        ru = 0.0 + diff
        if var_mass != 0.0:
            ru = ru + stoch
        ru = mk_advective_s_fluxdiv(g, umac, rho_fc, ru)
        ru = ru * dt_s
        rn = rho_old + ru
        rt = compute_rhotot(rn)
        return (rn / rt) * rt, rt

    rho_fc = _faces_comps(g, rho_old)
    rf = average_cc_to_face(g, rhotot_old)
    # ---- steps 1-2: predictor fluxes and Stokes solve (:166-336)
    rho_old, diff, stoch, info1 = stage(rho_old, dt if alg == 1 else 0.5 * dt, weights, p)
    p2 = Params.from_buffer_copy(p)
    p2.gmres_abs_tol = max(p.gmres_abs_tol, info1["norm_pre_rhs"] * p.gmres_rel_tol)
    # ---- step 3: scalar predictor (:340-418)
    rho_new, rhotot_new = scalar(diff, stoch, 0.5 * dt)
    rho_fc = _faces_comps(g, rho_new)
    rf = average_cc_to_face(g, rhotot_new)
    eta, eta_ed, kappa = compute_eta_kappa(g, rho_new, rhotot_new, visc_coef, eta_model)
    # ---- steps 4-5: corrector fluxes and Stokes solve (:420-582)
    if alg == 2:
        weights = [1.0 / math.sqrt(2.0)] * 2
    rho_new, diff, stoch, info2 = stage(rho_new, dt, weights, p2)
    # ---- step 6: midpoint scalar corrector (:586-648)
    rho_new, rhotot_new = scalar(diff, stoch, dt)
    eta, eta_ed, kappa = compute_eta_kappa(g, rho_new, rhotot_new, visc_coef, eta_model)
    out = dict(umac=umac, rho_old=rho_old, rho_new=rho_new, rhotot_new=rhotot_new, pi=pi, eta=eta, eta_ed=eta_ed,
               kappa=kappa, diff_mass_fluxdiv=diff, stoch_mass_fluxdiv=stoch, gradp_baro=gradp_baro, W_m=W_m,
               W_mass=W_mass)
    return out, info1, info2
