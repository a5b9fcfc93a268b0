"""Host-side mirror of the reference's solver interface over the gfx950 C ABI.

The functions keep the names, argument order and meaning of the reference's
Fortran routines (``gmres``, ``apply_matrix``, ``apply_precon``,
``stag_mg_solver``, ``macproject``, ``stag_applyop_level``,
``stochastic_m_fluxdiv``, ``diffusive_mass_fluxdiv``, ``stag_inner_prod``,
``cc_inner_prod``); a multifab becomes a :class:`Field` resident in HBM and
the layout / bc_tower / probin globals become a :class:`Context`.  Every
call goes to ``liblmm_hip.so``; there is no CPU fallback, and loading fails
with an error when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblmm_hip.so")
_lib = None

CELL, FACE_X, FACE_Y, FACE_Z, NODE, EDGE_XY, EDGE_XZ, EDGE_YZ = range(8)
FP64, FP32_MG, LEARNED = 0, 1, 2
STATUS = {0: "ok", 1: "bad argument", 2: "unsupported", 3: "breakdown", 4: "device error",
          5: "out of memory", 6: "not converged"}


class LmmError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("visc_type", C.c_int), ("precon_type", C.c_int), ("p_norm_weight", C.c_double),
        ("scale_factor", C.c_double), ("mg_max_vcycles", C.c_int), ("mg_minwidth", C.c_int),
        ("mg_nsmooths_down", C.c_int), ("mg_nsmooths_up", C.c_int), ("mg_nsmooths_bottom", C.c_int),
        ("mg_rel_tol", C.c_double), ("stag_mg_max_vcycles", C.c_int), ("stag_mg_minwidth", C.c_int),
        ("stag_mg_nsmooths_down", C.c_int), ("stag_mg_nsmooths_up", C.c_int),
        ("stag_mg_nsmooths_bottom", C.c_int), ("stag_mg_smoother", C.c_int),
        ("stag_mg_omega", C.c_double), ("stag_mg_rel_tol", C.c_double), ("gmres_rel_tol", C.c_double),
        ("gmres_abs_tol", C.c_double), ("gmres_max_outer", C.c_int), ("gmres_max_inner", C.c_int),
        ("gmres_max_iter", C.c_int), ("gmres_min_iter", C.c_int), ("gmres_verbose", C.c_int),
        ("precision", C.c_int), ("stoch_stress_form", C.c_int), ("filtering_width", C.c_int),
    ]


class GmresInfo(C.Structure):
    _fields_ = [
        ("total_iter", C.c_int), ("outer_iter", C.c_int), ("norm_pre_rhs", C.c_double),
        ("norm_resid", C.c_double), ("norm_b", C.c_double), ("norm_resid_true", C.c_double),
        ("status", C.c_int), ("vcycles", C.c_int), ("seconds", C.c_double),
        ("learned_applied", C.c_int), ("learned_fallback", C.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelOptions(C.Structure):
    """lmm_kernel_options: kernel forms of a context (every choice but flex_cgs gives
    the same bits; flex_cgs selects the flexible GMRES's Gram-Schmidt form)."""
    _fields_ = [
        ("fused_sweeps", C.c_int), ("fused_min_tiles", C.c_int), ("fused_kernel", C.c_int),
        ("fused_target", C.c_int), ("fused_x3", C.c_int), ("resid_restrict", C.c_int),
        ("coarse_one_launch", C.c_int), ("am_brick", C.c_int), ("z32_direct", C.c_int),
        ("krylov_fused", C.c_int), ("flex_cgs", C.c_int), ("sweep_prefetch", C.c_int),
        ("precon_graph", C.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SpeciesParams(C.Structure):
    """lmm_species_params: probin_multispecies / probin_common species knobs."""
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


class InertialParams(C.Structure):
    """lmm_inertial_params: the probin_common knobs of the inertial step."""
    _fields_ = [
        ("grav", C.c_double * 3), ("variance_coef_mom", C.c_double), ("visc_coef", C.c_double),
        ("eta_model", C.c_int), ("advection_type", C.c_int), ("noise", C.c_int), ("algorithm_type", C.c_int),
        ("noise_seed", C.c_uint64), ("restart", C.c_int), ("reserved", C.c_int),
    ]


class GmresMember(C.Structure):
    """lmm_gmres_member: one solve of a batch."""
    _fields_ = [
        ("ctx", C.c_void_p), ("b_u", C.c_void_p * 3), ("b_p", C.c_void_p), ("x_u", C.c_void_p * 3),
        ("x_p", C.c_void_p), ("alpha_fc", C.c_void_p * 3), ("beta", C.c_void_p), ("beta_ed", C.c_void_p * 3),
        ("gamma", C.c_void_p), ("theta_alpha", C.c_double), ("norm_pre_rhs", C.c_double), ("info", GmresInfo),
        ("status", C.c_int),
    ]


# (name, restype, argtypes); every entry point of include/lmm.h
_FP = C.c_void_p
_SIGS = {
    "lmm_params_default": (None, [C.POINTER(Params)]),
    "lmm_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.c_double, C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "lmm_ctx_destroy": (C.c_int, [C.c_void_p]),
    "lmm_ctx_set_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "lmm_ctx_get_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "lmm_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "lmm_kernel_options_default": (None, [C.POINTER(KernelOptions)]),
    "lmm_ctx_get_kernel_options": (C.c_int, [C.c_void_p, C.POINTER(KernelOptions)]),
    "lmm_ctx_set_kernel_options": (C.c_int, [C.c_void_p, C.POINTER(KernelOptions)]),
    "lmm_ctx_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "lmm_ctx_nlevels": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "lmm_ctx_last_error": (C.c_char_p, [C.c_void_p]),
    "lmm_ctx_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "lmm_ctx_kernel_stats": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
    "lmm_ctx_reset_stats": (C.c_int, [C.c_void_p]),
    "lmm_field_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "lmm_field_destroy": (C.c_int, [C.c_void_p]),
    "lmm_field_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "lmm_field_upload": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "lmm_field_download": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "lmm_field_upload_fab": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.c_int]),
    "lmm_field_download_fab": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.c_int]),
    "lmm_field_setval": (C.c_int, [C.c_void_p, C.c_double]),
    "lmm_field_copy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lmm_field_device_ptr": (C.c_void_p, [C.c_void_p]),
    "lmm_gmres": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, _FP, C.c_double,
                            C.POINTER(C.c_double), C.POINTER(GmresInfo)]),
    "lmm_apply_matrix": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, _FP, C.c_double]),
    "lmm_apply_precon": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, _FP, C.c_double]),
    "lmm_stag_mg_solver": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP * 3, _FP, C.c_double, _FP * 3, _FP * 3]),
    "lmm_macproject": (C.c_int, [C.c_void_p, _FP, _FP * 3, _FP * 3, _FP]),
    "lmm_stag_applyop": (C.c_int, [C.c_void_p, _FP * 3, _FP * 3, _FP * 3, _FP, _FP * 3, _FP, C.c_double, C.c_int]),
    "lmm_diffusive_m_fluxdiv": (C.c_int, [C.c_void_p, _FP * 3, _FP * 3, _FP, _FP * 3, _FP]),
    "lmm_stag_inner_prod": (C.c_int, [C.c_void_p, _FP * 3, _FP * 3, C.POINTER(C.c_double)]),
    "lmm_cc_inner_prod": (C.c_int, [C.c_void_p, _FP, _FP, C.POINTER(C.c_double)]),
    "lmm_compute_div": (C.c_int, [C.c_void_p, _FP * 3, _FP, C.c_int]),
    "lmm_compute_grad": (C.c_int, [C.c_void_p, _FP, _FP * 3]),
    "lmm_compute_div_comps": (C.c_int, [C.c_void_p, _FP * 3, _FP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "lmm_compute_grad_comps": (C.c_int, [C.c_void_p, _FP, _FP * 3, C.c_int, C.c_int, C.c_int]),
    "lmm_average_cc_to_face_comps": (C.c_int, [C.c_void_p, _FP, _FP * 3, C.c_int, C.c_int]),
    "lmm_average_cc_to_face": (C.c_int, [C.c_void_p, _FP, _FP * 3]),
    "lmm_average_cc_to_node": (C.c_int, [C.c_void_p, _FP, _FP]),
    "lmm_average_cc_to_edge": (C.c_int, [C.c_void_p, _FP, _FP * 3]),
    "lmm_stochastic_m_fluxdiv": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, C.c_double,
                                           C.POINTER(C.c_double), C.c_double, C.c_double]),
    "lmm_init_m_stochastic": (C.c_int, [C.c_void_p, C.c_int]),
    "lmm_fill_m_stochastic": (C.c_int, [C.c_void_p, C.c_uint32]),
    "lmm_gmres_batched": (C.c_int, [C.c_int, C.POINTER(GmresMember), C.c_int]),
    "lmm_fill_m_stochastic_device": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "lmm_field_fill_normal": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int64]),
    "lmm_stoch_m_force": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, _FP, _FP * 3, C.c_double]),
    "lmm_diffusive_mass_fluxdiv": (C.c_int, [C.c_void_p, C.c_int, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP * 3,
                                             _FP * 3, C.c_int]),
    "lmm_diffusive_mass_fluxdiv_noalloc": (C.c_int, [C.c_void_p, C.c_int, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP,
                                                     _FP * 3, _FP * 3, C.c_int]),
    "lmm_fill_normal_host": (C.c_int, [C.c_uint32, C.c_int64, C.POINTER(C.c_double)]),
    "lmm_cnn_config_default": (None, [C.c_void_p]),
    "lmm_cnn_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "lmm_cnn_destroy": (C.c_int, [C.c_void_p]),
    "lmm_cnn_num_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "lmm_cnn_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "lmm_cnn_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "lmm_cnn_forward_host": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]),
    "lmm_cnn_loss_grad_host": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float),
                                         C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    "lmm_cnn_generate_pairs": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP * 3, _FP, C.c_double, C.c_int, C.c_uint64]),
    "lmm_cnn_num_pairs": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "lmm_cnn_error_reduction": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP * 3, _FP, C.c_double, C.c_int, C.c_uint64,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "lmm_cnn_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int64,
                                C.POINTER(C.c_float)]),
    "lmm_cnn_evaluate": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_double)]),
    "lmm_ctx_set_cnn": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lmm_species_params_default": (None, [C.POINTER(SpeciesParams), C.c_int]),
    "lmm_ctx_set_species": (C.c_int, [C.c_void_p, C.POINTER(SpeciesParams)]),
    "lmm_ctx_get_species": (C.c_int, [C.c_void_p, C.POINTER(SpeciesParams)]),
    "lmm_correct_rho_with_drho": (C.c_int, [C.c_void_p, _FP, _FP]),
    "lmm_convert_cons_to_prim": (C.c_int, [C.c_void_p, _FP, _FP, _FP, _FP]),
    "lmm_compute_mixture_properties": (C.c_int, [C.c_void_p, _FP, _FP, _FP, _FP, _FP, _FP]),
    "lmm_compute_Gama": (C.c_int, [C.c_void_p, _FP, _FP, _FP, _FP, _FP, _FP]),
    "lmm_compute_chi": (C.c_int, [C.c_void_p, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP]),
    "lmm_compute_rhoWchi": (C.c_int, [C.c_void_p, _FP, _FP, _FP, _FP]),
    "lmm_compute_Lonsager": (C.c_int, [C.c_void_p, _FP, _FP, _FP, _FP, _FP, _FP, _FP]),
    "lmm_init_mass_stochastic": (C.c_int, [C.c_void_p, C.c_int]),
    "lmm_fill_mass_stochastic": (C.c_int, [C.c_void_p, C.c_uint32]),
    "lmm_fill_mass_stochastic_device": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "lmm_stochastic_mass_fluxdiv": (C.c_int, [C.c_void_p, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP * 3, C.c_double,
                                              C.POINTER(C.c_double), C.c_int]),
    "lmm_compute_mass_fluxdiv": (C.c_int, [C.c_void_p, _FP, _FP * 3, _FP, _FP, _FP, _FP * 3, C.c_double, C.c_double,
                                           C.POINTER(C.c_double)]),
    "lmm_inertial_params_default": (None, [C.POINTER(InertialParams)]),
    "lmm_advance_timestep_inertial": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP, _FP, _FP, _FP * 3, _FP, _FP, _FP * 3,
                                                _FP, _FP, _FP * 3, _FP, _FP, C.c_double, C.c_double, C.c_int64,
                                                C.POINTER(InertialParams), C.POINTER(GmresInfo * 2)]),
    "lmm_advance_timestep_overdamped": (C.c_int, [C.c_void_p, _FP * 3, _FP, _FP, _FP, _FP, _FP * 3, _FP, _FP,
                                                  _FP * 3, _FP, _FP, _FP * 3, _FP, _FP, C.c_double, C.c_double,
                                                  C.c_int64, C.POINTER(InertialParams), C.POINTER(GmresInfo * 2)]),
    "lmm_mk_advective_s_fluxdiv": (C.c_int, [C.c_void_p, _FP * 3, _FP * 3, _FP, C.c_int, C.c_int]),
    "lmm_mk_advective_m_fluxdiv": (C.c_int, [C.c_void_p, _FP * 3, _FP * 3, _FP * 3]),
    "lmm_mk_grav_force": (C.c_int, [C.c_void_p, _FP * 3, _FP * 3, _FP * 3, C.POINTER(C.c_double)]),
}


def exported_names():
    return list(_SIGS)


def lib():
    """Load liblmm_hip.so; raises LmmError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LmmError(f"{LIB_PATH} is missing: build it with `python -m lowmachmixtures_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def default_params(**kw) -> Params:
    p = Params()
    lib().lmm_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Context:
    """ml_layout + bc_tower + dx + probin_gmres knobs for one periodic box."""

    def __init__(self, dim, n, dx=1.0, **params):
        self.dim = dim
        self.n = tuple(n)[:dim] + (1,) * (3 - dim)
        self.dx = dx
        self.params = default_params(**params)
        arr = (C.c_int * 3)(*self.n)
        h = C.c_void_p()
        rc = lib().lmm_ctx_create(dim, arr, C.c_double(dx), C.byref(self.params), C.byref(h))
        if rc != 0:
            raise LmmError(f"lmm_ctx_create failed: {STATUS.get(rc, rc)}")
        self.h = h
        self._fields = []

    @property
    def shape(self):
        return tuple(reversed(self.n[: self.dim]))

    def check(self, rc, what):
        if rc != 0:
            msg = lib().lmm_ctx_last_error(self.h)
            raise LmmError(f"{what}: {STATUS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self.check(lib().lmm_ctx_set_params(self.h, C.byref(self.params)), "set_params")

    def kernel_options(self, **kw):
        """Current kernel forms of this context; keyword arguments change them first."""
        o = KernelOptions()
        self.check(lib().lmm_ctx_get_kernel_options(self.h, C.byref(o)), "get_kernel_options")
        if kw:
            for k, v in kw.items():
                if not hasattr(o, k):
                    raise KeyError(k)
                setattr(o, k, int(v))
            self.check(lib().lmm_ctx_set_kernel_options(self.h, C.byref(o)), "set_kernel_options")
        return o.as_dict()

    def autotune(self):
        """lmm_ctx_autotune: times the fine-level fused sweeps for a set of z-slab
        targets on this GPU and keeps the fastest; returns it."""
        t = C.c_int()
        self.check(lib().lmm_ctx_autotune(self.h, C.byref(t)), "autotune")
        return t.value

    def nlevels(self):
        a, b = C.c_int(), C.c_int()
        lib().lmm_ctx_nlevels(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def synchronize(self):
        self.check(lib().lmm_ctx_synchronize(self.h), "synchronize")

    def profiling(self, on=True):
        """on: False/0 off, True/1 every kernel family, 2 only the fused sweeps
        (kernel-exact events, negligible cost)."""
        self.check(lib().lmm_ctx_set_profiling(self.h, int(on)), "lmm_ctx_set_profiling")

    def reset_stats(self):
        lib().lmm_ctx_reset_stats(self.h)

    def kernel_stats(self, family):
        n, ms, by = C.c_int64(), C.c_double(), C.c_double()
        lib().lmm_ctx_kernel_stats(self.h, family.encode(), C.byref(n), C.byref(ms), C.byref(by))
        return {"launches": n.value, "ms": ms.value, "bytes": by.value}

    def field(self, kind, ncomp=1, data=None):
        f = Field(self, kind, ncomp)
        if data is not None:
            f.upload(data)
        return f

    def faces(self, data=None):
        return [self.field(FACE_X + d, 1, None if data is None else data[d]) for d in range(self.dim)]

    def edges(self, data=None, ncomp=1):
        kinds = [NODE] if self.dim == 2 else [EDGE_XY, EDGE_XZ, EDGE_YZ]
        return [self.field(k, ncomp, None if data is None else data[e]) for e, k in enumerate(kinds)]

    def close(self):
        for f in self._fields:
            f.free()
        self._fields = []
        if self.h:
            lib().lmm_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Field:
    """A device-resident multifab of one periodic box (deduplicated layout)."""

    def __init__(self, ctx: Context, kind: int, ncomp: int = 1):
        self.ctx = ctx
        self.kind = kind
        self.ncomp = ncomp
        h = C.c_void_p()
        ctx.check(lib().lmm_field_create(ctx.h, kind, ncomp, C.byref(h)), "field_create")
        self.h = h
        ctx._fields.append(self)

    @property
    def shape(self):
        s = self.ctx.shape
        return s if self.ncomp == 1 else (self.ncomp,) + s

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != int(np.prod(self.shape)):
            raise ValueError(f"size {a.size} does not match field shape {self.shape}")
        self.ctx.check(lib().lmm_field_upload(self.h, a.ctypes.data_as(C.POINTER(C.c_double))), "upload")
        return self

    def download(self):
        out = np.empty(self.shape, dtype=np.float64)
        self.ctx.check(lib().lmm_field_download(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "download")
        return out

    def fill_normal(self, seed, step=0, v0=0):
        """Values v0, v0+1, ... of the counter-based Gaussian stream (seed, step)
        in storage order."""
        self.ctx.check(lib().lmm_field_fill_normal(self.h, C.c_uint64(seed), C.c_uint64(step), C.c_int64(v0)),
                       "field_fill_normal")
        return self

    def upload_fab(self, a, comp, lo, hi, ng):
        a = np.asfortranarray(a, dtype=np.float64)
        lo3 = (C.c_int * 3)(*(list(lo) + [0] * (3 - len(lo))))
        hi3 = (C.c_int * 3)(*(list(hi) + [0] * (3 - len(hi))))
        self.ctx.check(lib().lmm_field_upload_fab(self.h, comp, a.ctypes.data_as(C.POINTER(C.c_double)), lo3, hi3, ng),
                       "upload_fab")

    def download_fab(self, shape, comp, lo, hi, ng):
        out = np.zeros(shape, dtype=np.float64, order="F")
        lo3 = (C.c_int * 3)(*(list(lo) + [0] * (3 - len(lo))))
        hi3 = (C.c_int * 3)(*(list(hi) + [0] * (3 - len(hi))))
        self.ctx.check(lib().lmm_field_download_fab(self.h, comp, out.ctypes.data_as(C.POINTER(C.c_double)), lo3,
                                                    hi3, ng), "download_fab")
        return out

    def setval(self, v):
        self.ctx.check(lib().lmm_field_setval(self.h, C.c_double(v)), "setval")
        return self

    def device_ptr(self):
        return lib().lmm_field_device_ptr(self.h)

    def free(self):
        if self.h:
            lib().lmm_field_destroy(self.h)
            self.h = None


def _a3(fields):
    arr = (_FP * 3)()
    if fields is None:
        return arr
    for i, f in enumerate(fields):
        if f is not None:
            arr[i] = f.h
    return arr


def _h(f):
    return None if f is None else f.h


# ---------------------------------------------------------------------------
# reference entry points
# ---------------------------------------------------------------------------

def gmres(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha):
    """gmres.f90:30-45.  Returns (norm_pre_rhs, info dict); x_u, x_p updated in place."""
    info = GmresInfo()
    npr = C.c_double()
    rc = lib().lmm_gmres(ctx.h, _a3(b_u), _h(b_p), _a3(x_u), _h(x_p), _a3(alpha_fc), _h(beta), _a3(beta_ed),
                         _h(gamma), C.c_double(theta_alpha), C.byref(npr), C.byref(info))
    ctx.check(rc, "gmres")
    return npr.value, info.as_dict()


def gmres_batched(problems, nthreads=0):
    """Independent solves of several DeviceProblems (each with its own
    context) at once: lmm_gmres_batched.  Returns [(status, norm_pre_rhs,
    info dict)] in order; raises LmmError when any member failed."""
    n = len(problems)
    arr = (GmresMember * n)()
    for m, dp in zip(arr, problems):
        m.ctx = dp.ctx.h
        for d in range(3):
            m.b_u[d] = dp.b_u[d].h if d < len(dp.b_u) else None
            m.x_u[d] = dp.x_u[d].h if d < len(dp.x_u) else None
            m.alpha_fc[d] = dp.alpha[d].h if d < len(dp.alpha) else None
            m.beta_ed[d] = dp.beta_ed[d].h if d < len(dp.beta_ed) else None
        m.b_p, m.x_p, m.beta = dp.b_p.h, dp.x_p.h, dp.beta.h
        m.gamma = _h(dp.gamma)
        m.theta_alpha = dp.pb.theta_alpha
    rc = lib().lmm_gmres_batched(n, arr, int(nthreads))
    out = [(m.status, m.norm_pre_rhs, m.info.as_dict()) for m in arr]
    if rc != 0:
        bad = [i for i, m in enumerate(arr) if m.status != 0]
        if not bad:  # rejected before any member ran (e.g. a shared context)
            raise LmmError(f"gmres_batched: {STATUS.get(rc, rc)}")
        problems[bad[0]].ctx.check(arr[bad[0]].status, f"gmres_batched member {bad[0]}")
    return out


def apply_matrix(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha):
    """apply_matrix.f90:23-39."""
    ctx.check(lib().lmm_apply_matrix(ctx.h, _a3(b_u), _h(b_p), _a3(x_u), _h(x_p), _a3(alpha_fc), _h(beta),
                                     _a3(beta_ed), _h(gamma), C.c_double(theta_alpha)), "apply_matrix")


def apply_precon(ctx, b_u, b_p, x_u, x_p, alpha_fc, beta, beta_ed, gamma, theta_alpha):
    """apply_precon.f90:29-43."""
    ctx.check(lib().lmm_apply_precon(ctx.h, _a3(b_u), _h(b_p), _a3(x_u), _h(x_p), _a3(alpha_fc), _h(beta),
                                     _a3(beta_ed), _h(gamma), C.c_double(theta_alpha)), "apply_precon")


def stag_mg_solver(ctx, alpha_fc, beta_cc, beta_ed, gamma_cc, theta_alpha, phi_fc, rhs_fc):
    """stag_mg_solver.f90:37-50."""
    ctx.check(lib().lmm_stag_mg_solver(ctx.h, _a3(alpha_fc), _h(beta_cc), _a3(beta_ed), _h(gamma_cc),
                                       C.c_double(theta_alpha), _a3(phi_fc), _a3(rhs_fc)), "stag_mg_solver")


def macproject(ctx, phi, umac, alphainv_fc, mac_rhs):
    """macproject.f90:34-45."""
    ctx.check(lib().lmm_macproject(ctx.h, _h(phi), _a3(umac), _a3(alphainv_fc), _h(mac_rhs)), "macproject")


def stag_applyop_level(ctx, phi_fc, Lphi_fc, alpha_fc, beta_cc, beta_ed, gamma_cc, color=0, theta_alpha=1.0):
    """stag_applyop.f90:70-81 (alpha_fc multiplied by theta_alpha first)."""
    ctx.check(lib().lmm_stag_applyop(ctx.h, _a3(phi_fc), _a3(Lphi_fc), _a3(alpha_fc), _h(beta_cc), _a3(beta_ed),
                                     _h(gamma_cc), C.c_double(theta_alpha), color), "stag_applyop")


def diffusive_m_fluxdiv(ctx, m_update, umac, eta, eta_ed, kappa):
    """diffusive_m_fluxdiv.f90:16-73: m_update += L_eta(umac) (stag_applyop with alpha = 0, subtracted)."""
    ctx.check(lib().lmm_diffusive_m_fluxdiv(ctx.h, _a3(m_update), _a3(umac), _h(eta), _a3(eta_ed), _h(kappa)),
              "diffusive_m_fluxdiv")


def stag_inner_prod(ctx, m1, m2):
    out = (C.c_double * 3)()
    ctx.check(lib().lmm_stag_inner_prod(ctx.h, _a3(m1), _a3(m2), out), "stag_inner_prod")
    return [out[d] for d in range(ctx.dim)]


def cc_inner_prod(ctx, m1, m2):
    out = C.c_double()
    ctx.check(lib().lmm_cc_inner_prod(ctx.h, _h(m1), _h(m2), C.byref(out)), "cc_inner_prod")
    return out.value


def compute_div(ctx, phi_fc, div, increment=False):
    ctx.check(lib().lmm_compute_div(ctx.h, _a3(phi_fc), _h(div), int(increment)), "compute_div")


def compute_grad(ctx, phi, gradp):
    ctx.check(lib().lmm_compute_grad(ctx.h, _h(phi), _a3(gradp)), "compute_grad")


def compute_div_comps(ctx, phi_fc, div, start_incomp, start_outcomp, num_comp, increment=False):
    """div_and_grad.f90:304 with component arguments (0-based)."""
    ctx.check(lib().lmm_compute_div_comps(ctx.h, _a3(phi_fc), _h(div), start_incomp, start_outcomp, num_comp,
                                          int(increment)), "compute_div_comps")


def compute_grad_comps(ctx, phi, gradp, start_incomp, start_outcomp, num_comp):
    """div_and_grad.f90:16 with component arguments (0-based)."""
    ctx.check(lib().lmm_compute_grad_comps(ctx.h, _h(phi), _a3(gradp), start_incomp, start_outcomp, num_comp),
              "compute_grad_comps")


def average_cc_to_face_comps(ctx, cc, face, start_comp, num_comp):
    """convert_stag.f90:132 with component arguments (0-based)."""
    ctx.check(lib().lmm_average_cc_to_face_comps(ctx.h, _h(cc), _a3(face), start_comp, num_comp),
              "average_cc_to_face_comps")


def average_cc_to_face(ctx, cc, face):
    ctx.check(lib().lmm_average_cc_to_face(ctx.h, _h(cc), _a3(face)), "average_cc_to_face")


def average_cc_to_node(ctx, cc, node):
    ctx.check(lib().lmm_average_cc_to_node(ctx.h, _h(cc), _h(node)), "average_cc_to_node")


def average_cc_to_edge(ctx, cc, edge):
    ctx.check(lib().lmm_average_cc_to_edge(ctx.h, _h(cc), _a3(edge)), "average_cc_to_edge")


def init_m_stochastic(ctx, n_rngs):
    ctx.check(lib().lmm_init_m_stochastic(ctx.h, n_rngs), "init_m_stochastic")


def fill_m_stochastic(ctx, seed):
    ctx.check(lib().lmm_fill_m_stochastic(ctx.h, C.c_uint32(seed)), "fill_m_stochastic")


def fill_m_stochastic_device(ctx, seed, step=0):
    """Device fill of the context's noise from the counter-based stream (no
    counterpart in the reference, whose noise is the host MT19937 stream)."""
    ctx.check(lib().lmm_fill_m_stochastic_device(ctx.h, C.c_uint64(seed), C.c_uint64(step)),
              "fill_m_stochastic_device")


def stochastic_m_fluxdiv(ctx, stoch_m_force, eta, eta_ed, temperature, temperature_ed, dt, weights,
                         variance_coef_mom=1.0, k_B=1.0):
    """stochastic_m_fluxdiv.f90:41-51 (increments stoch_m_force)."""
    w = (C.c_double * len(weights))(*weights)
    ctx.check(lib().lmm_stochastic_m_fluxdiv(ctx.h, _a3(stoch_m_force), _h(eta), _a3(eta_ed), _h(temperature),
                                             _a3(temperature_ed), C.c_double(dt), w, C.c_double(variance_coef_mom),
                                             C.c_double(k_B)), "stochastic_m_fluxdiv")


def stoch_m_force(ctx, force, W_cc, W_ed, eta, eta_ed, temperature, temperature_ed, variance):
    ctx.check(lib().lmm_stoch_m_force(ctx.h, _a3(force), _h(W_cc), _a3(W_ed), _h(eta), _a3(eta_ed), _h(temperature),
                                      _a3(temperature_ed), C.c_double(variance)), "stoch_m_force")


def species_params(nspecies, **kw) -> SpeciesParams:
    """Reference defaults for nspecies species; array knobs take a scalar or a sequence."""
    p = SpeciesParams()
    lib().lmm_species_params_default(C.byref(p), int(nspecies))
    for k, v in kw.items():
        field = getattr(p, k)
        if isinstance(field, C.Array):
            vals = [v] * len(field) if np.ndim(v) == 0 else list(v)
            for i, x in enumerate(vals):
                field[i] = float(x)
        else:
            setattr(p, k, v)
    return p


def set_species(ctx, nspecies, **kw):
    """Sets the context's species knobs (lmm_ctx_set_species); returns them."""
    p = species_params(nspecies, **kw)
    ctx.check(lib().lmm_ctx_set_species(ctx.h, C.byref(p)), "ctx_set_species")
    return p


def correct_rho_with_drho(ctx, rho, drho):
    """mass_flux_utilities.f90:26-150."""
    ctx.check(lib().lmm_correct_rho_with_drho(ctx.h, _h(rho), _h(drho)), "correct_rho_with_drho")


def convert_cons_to_prim(ctx, rho, rhotot, molarconc, molmtot):
    """mass_flux_utilities.f90:152-313."""
    ctx.check(lib().lmm_convert_cons_to_prim(ctx.h, _h(rho), _h(rhotot), _h(molarconc), _h(molmtot)),
              "convert_cons_to_prim")


def compute_mixture_properties(ctx, rho, rhotot, D_bar, D_therm, Hessian, Temp):
    """compute_mixture_properties.f90:24-233 (constant coefficients)."""
    ctx.check(lib().lmm_compute_mixture_properties(ctx.h, _h(rho), _h(rhotot), _h(D_bar), _h(D_therm), _h(Hessian),
                                                   _h(Temp)), "compute_mixture_properties")


def compute_Gama(ctx, rho, rhotot, molarconc, molmtot, Hessian, Gama):
    """mass_flux_utilities.f90:428-599."""
    ctx.check(lib().lmm_compute_Gama(ctx.h, _h(rho), _h(rhotot), _h(molarconc), _h(molmtot), _h(Hessian), _h(Gama)),
              "compute_Gama")


def compute_chi(ctx, rho, rhotot, molarconc, chi, D_bar, D_therm, Temp, zeta_by_Temp):
    """mass_flux_utilities.f90:601-927."""
    ctx.check(lib().lmm_compute_chi(ctx.h, _h(rho), _h(rhotot), _h(molarconc), _h(chi), _h(D_bar), _h(D_therm),
                                    _h(Temp), _h(zeta_by_Temp)), "compute_chi")


def compute_rhoWchi(ctx, rho, rhotot, chi, rhoWchi):
    """mass_flux_utilities.f90:1120-1273."""
    ctx.check(lib().lmm_compute_rhoWchi(ctx.h, _h(rho), _h(rhotot), _h(chi), _h(rhoWchi)), "compute_rhoWchi")


def compute_Lonsager(ctx, rho, rhotot, molarconc, molmtot, chi, Gama, Lonsager):
    """mass_flux_utilities.f90:929-1118."""
    ctx.check(lib().lmm_compute_Lonsager(ctx.h, _h(rho), _h(rhotot), _h(molarconc), _h(molmtot), _h(chi), _h(Gama),
                                         _h(Lonsager)), "compute_Lonsager")


def init_mass_stochastic(ctx, n_rngs):
    ctx.check(lib().lmm_init_mass_stochastic(ctx.h, int(n_rngs)), "init_mass_stochastic")


def fill_mass_stochastic(ctx, seed):
    ctx.check(lib().lmm_fill_mass_stochastic(ctx.h, C.c_uint32(seed)), "fill_mass_stochastic")


def fill_mass_stochastic_device(ctx, seed, step=0):
    ctx.check(lib().lmm_fill_mass_stochastic_device(ctx.h, C.c_uint64(seed), C.c_uint64(step)),
              "fill_mass_stochastic_device")


def stochastic_mass_fluxdiv(ctx, rho, rhotot, molarconc, molmtot, chi, Gama, stoch_fluxdiv, flux_total, dt, weights,
                            increment=False):
    """stochastic_mass_fluxdiv.f90:32-154."""
    w = (C.c_double * 8)(*[float(x) for x in weights])
    ctx.check(lib().lmm_stochastic_mass_fluxdiv(ctx.h, _h(rho), _h(rhotot), _h(molarconc), _h(molmtot), _h(chi),
                                                _h(Gama), _h(stoch_fluxdiv), _a3(flux_total), C.c_double(dt), w,
                                                int(increment)), "stochastic_mass_fluxdiv")


def compute_mass_fluxdiv(ctx, rho, gradp_baro, diff_fluxdiv, stoch_fluxdiv, Temp, flux_total, dt, stage_time=0.0,
                         weights=(1.0,)):
    """compute_mass_fluxdiv.f90:24-147."""
    w = (C.c_double * 8)(*[float(x) for x in weights])
    ctx.check(lib().lmm_compute_mass_fluxdiv(ctx.h, _h(rho), _a3(gradp_baro), _h(diff_fluxdiv), _h(stoch_fluxdiv),
                                             _h(Temp), _a3(flux_total), C.c_double(dt), C.c_double(stage_time), w),
              "compute_mass_fluxdiv")


def diffusive_mass_fluxdiv(ctx, nspecies, rho, rhotot, molarconc, rhoWchi, Gama, diff_fluxdiv, Temp,
                           zeta_by_Temp, gradp_baro, flux_total, correct_flux=True):
    """diffusive_mass_fluxdiv.f90:26-42 (temperature piece with Temp and
    zeta_by_Temp, barodiffusion per the context's species knobs)."""
    ctx.check(lib().lmm_diffusive_mass_fluxdiv(ctx.h, nspecies, _h(rho), _h(rhotot), _h(molarconc), _h(rhoWchi),
                                               _h(Gama), _h(diff_fluxdiv), _h(Temp), _h(zeta_by_Temp),
                                               _a3(gradp_baro), _a3(flux_total), int(correct_flux)),
              "diffusive_mass_fluxdiv")


def fill_normal_host(seed, count):
    out = np.empty(int(count))
    rc = lib().lmm_fill_normal_host(C.c_uint32(seed), C.c_int64(int(count)), out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise LmmError("fill_normal_host failed")
    return out


class DeviceProblem:
    """A StokesProblem uploaded to a context: coefficients, RHS and solution fields."""

    def __init__(self, ctx: Context, pb, rhs=None):
        self.ctx = ctx
        self.pb = pb
        self.alpha = ctx.faces(pb.alpha)
        self.beta = ctx.field(CELL, 1, pb.beta)
        self.gamma = ctx.field(CELL, 1, pb.gamma)
        self.beta_ed = ctx.edges(pb.beta_ed)
        self.b_u = ctx.faces()
        for f in self.b_u:
            f.setval(0.0)
        if rhs is not None:
            for f, a in zip(self.b_u, rhs):
                f.upload(a)
        else:
            self.rhs_from_noise()
        self.b_p = ctx.field(CELL, 1, pb.b_p)
        self.x_u = ctx.faces()
        self.x_p = ctx.field(CELL)
        self.zero_solution()

    def rhs_from_noise(self):
        """b_u = stoch_m_force(W) with eta = T = 1 and variance 1, on the device."""
        ctx, pb = self.ctx, self.pb
        ones = np.ones(ctx.shape)
        eta = ctx.field(CELL, 1, ones)
        ed = ctx.edges([ones] * (1 if ctx.dim == 2 else 3))
        Wcc = ctx.field(CELL, ctx.dim, pb.W_cc)
        Wed = ctx.edges(pb.W_ed, ncomp=2)
        for f in self.b_u:
            f.setval(0.0)
        stoch_m_force(ctx, self.b_u, Wcc, Wed, eta, ed, eta, ed, 1.0)
        for f in [eta, Wcc] + ed + Wed:
            f.free()

    def zero_solution(self):
        for f in self.x_u:
            f.setval(0.0)
        self.x_p.setval(0.0)

    def solve(self):
        return gmres(self.ctx, self.b_u, self.b_p, self.x_u, self.x_p, self.alpha, self.beta, self.beta_ed,
                     self.gamma, self.pb.theta_alpha)

    def solution(self):
        return [f.download() for f in self.x_u], self.x_p.download()


# ---------------------------------------------------------------- learned correction

class CnnConfig(C.Structure):
    """lmm_cnn_config: channels of the U-Net and the multigrid level it acts on."""
    _fields_ = [("cin", C.c_int), ("c1", C.c_int), ("c2", C.c_int), ("c3", C.c_int), ("cout", C.c_int),
                ("level", C.c_int), ("sample_steps", C.c_int)]


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Cnn:
    """The learned coarse-grid correction network of a context (lmm_cnn)."""

    def __init__(self, ctx, seed=0, **cfg):
        c = CnnConfig()
        lib().lmm_cnn_config_default(C.byref(c))
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, int(v))
        self.cfg = c
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().lmm_cnn_create(ctx.h, C.byref(c), C.c_uint64(seed), C.byref(h)), "cnn_create")
        self.h = h

    def close(self):
        if self.h:
            lib().lmm_cnn_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def num_params(self):
        n = C.c_int64()
        self.ctx.check(lib().lmm_cnn_num_params(self.h, C.byref(n)), "cnn_num_params")
        return n.value

    def get_params(self):
        out = np.zeros(self.num_params(), dtype=np.float32)
        self.ctx.check(lib().lmm_cnn_get_params(self.h, _fptr(out)), "cnn_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params()
        self.ctx.check(lib().lmm_cnn_set_params(self.h, _fptr(p)), "cnn_set_params")

    def forward(self, x):
        """x: (nb, nz, ny, nx, cin) -> (nb, nz, ny, nx, cout)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        nb, nz, ny, nx, _ = x.shape
        y = np.zeros((nb, nz, ny, nx, self.cfg.cout), dtype=np.float32)
        n = (C.c_int * 3)(nx, ny, nz)
        self.ctx.check(lib().lmm_cnn_forward_host(self.h, nb, n, _fptr(x), _fptr(y)), "cnn_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        nb, nz, ny, nx, _ = x.shape
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = C.c_double()
        n = (C.c_int * 3)(nx, ny, nz)
        self.ctx.check(lib().lmm_cnn_loss_grad_host(self.h, nb, n, _fptr(x), _fptr(t), C.byref(loss), _fptr(g)),
                       "cnn_loss_grad")
        return loss.value, g

    def generate_pairs(self, dp, npairs, seed):
        """Training pairs from the plain V-cycle on a DeviceProblem's coefficients."""
        self.ctx.check(lib().lmm_cnn_generate_pairs(self.h, _a3(dp.alpha), _h(dp.beta), _a3(dp.beta_ed), _h(dp.gamma),
                                                    C.c_double(dp.pb.theta_alpha), int(npairs), C.c_uint64(seed)),
                       "cnn_generate_pairs")

    def error_reduction(self, dp, npairs, seed):
        """Mean |phi* - z| / |phi*| after the plain preconditioner and after it plus the learned stage."""
        a, b = C.c_double(), C.c_double()
        self.ctx.check(lib().lmm_cnn_error_reduction(self.h, _a3(dp.alpha), _h(dp.beta), _a3(dp.beta_ed),
                                                     _h(dp.gamma), C.c_double(dp.pb.theta_alpha), int(npairs),
                                                     C.c_uint64(seed), C.byref(a), C.byref(b)),
                       "cnn_error_reduction")
        return a.value, b.value

    def num_pairs(self):
        n = C.c_int64()
        self.ctx.check(lib().lmm_cnn_num_pairs(self.h, C.byref(n)), "cnn_num_pairs")
        return n.value

    def train(self, steps, batch, lr, seed=0, held_out=0):
        hist = np.zeros(steps, dtype=np.float32)
        self.ctx.check(lib().lmm_cnn_train(self.h, int(steps), int(batch), C.c_double(lr), C.c_uint64(seed),
                                           C.c_int64(held_out), _fptr(hist)), "cnn_train")
        return hist

    def evaluate(self, first, last):
        r = C.c_double()
        self.ctx.check(lib().lmm_cnn_evaluate(self.h, C.c_int64(first), C.c_int64(last), C.byref(r)), "cnn_evaluate")
        return r.value

    def attach(self):
        """Use this network for precision LEARNED solves of its context."""
        self.ctx.check(lib().lmm_ctx_set_cnn(self.ctx.h, self.h), "ctx_set_cnn")

    def detach(self):
        self.ctx.check(lib().lmm_ctx_set_cnn(self.ctx.h, None), "ctx_set_cnn")


# ---------------------------------------------------------------------------
# inertial time step (advance_timestep_inertial.f90)

def inertial_params(**kw) -> InertialParams:
    """probin_common defaults (lmm_inertial_params_default), then kw."""
    p = InertialParams()
    lib().lmm_inertial_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "grav":
            for i, x in enumerate(list(v) + [0.0] * (3 - len(v))):
                p.grav[i] = float(x)
        else:
            setattr(p, k, v)
    return p


def advance_timestep_inertial(ctx, umac, rho_old, rho_new, rhotot_old, rhotot_new, gradp_baro, pi, eta, eta_ed, kappa,
                              Temp, Temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, time, istep, params=None):
    """advance_timestep_inertial.f90:58-79; returns the two solves' info dicts."""
    p = params if params is not None else inertial_params()
    info = (GmresInfo * 2)()
    ctx.check(lib().lmm_advance_timestep_inertial(
        ctx.h, _a3(umac), _h(rho_old), _h(rho_new), _h(rhotot_old), _h(rhotot_new), _a3(gradp_baro), _h(pi), _h(eta),
        _a3(eta_ed), _h(kappa), _h(Temp), _a3(Temp_ed), _h(diff_mass_fluxdiv), _h(stoch_mass_fluxdiv),
        C.c_double(dt), C.c_double(time), C.c_int64(int(istep)), C.byref(p), C.byref(info)),
        "advance_timestep_inertial")
    return info[0].as_dict(), info[1].as_dict()


def advance_timestep_overdamped(ctx, umac, rho_old, rho_new, rhotot_old, rhotot_new, gradp_baro, pi, eta, eta_ed,
                                kappa, Temp, Temp_ed, diff_mass_fluxdiv, stoch_mass_fluxdiv, dt, time, istep,
                                params=None):
    """advance_timestep_overdamped.f90:60-66; returns the two solves' info dicts."""
    p = params if params is not None else inertial_params(algorithm_type=1)
    info = (GmresInfo * 2)()
    ctx.check(lib().lmm_advance_timestep_overdamped(
        ctx.h, _a3(umac), _h(rho_old), _h(rho_new), _h(rhotot_old), _h(rhotot_new), _a3(gradp_baro), _h(pi), _h(eta),
        _a3(eta_ed), _h(kappa), _h(Temp), _a3(Temp_ed), _h(diff_mass_fluxdiv), _h(stoch_mass_fluxdiv),
        C.c_double(dt), C.c_double(time), C.c_int64(int(istep)), C.byref(p), C.byref(info)),
        "advance_timestep_overdamped")
    return info[0].as_dict(), info[1].as_dict()


def mk_advective_s_fluxdiv(ctx, umac, s_fc, s_update, start_comp=0, num_comp=1):
    """mk_advective_s_fluxdiv.f90:14-21 (0-based components)."""
    ctx.check(lib().lmm_mk_advective_s_fluxdiv(ctx.h, _a3(umac), _a3(s_fc), _h(s_update), int(start_comp),
                                               int(num_comp)), "mk_advective_s_fluxdiv")


def mk_advective_m_fluxdiv(ctx, umac, m, m_update):
    """mk_advective_m_fluxdiv.f90:16-22."""
    ctx.check(lib().lmm_mk_advective_m_fluxdiv(ctx.h, _a3(umac), _a3(m), _a3(m_update)), "mk_advective_m_fluxdiv")


def mk_grav_force(ctx, m_force, s_fc_old, s_fc_new, grav):
    """mk_grav_force.f90:17-22."""
    g = (C.c_double * 3)(*[float(x) for x in (list(grav) + [0.0, 0.0, 0.0])[:3]])
    ctx.check(lib().lmm_mk_grav_force(ctx.h, _a3(m_force), _a3(s_fc_old), _a3(s_fc_new), g), "mk_grav_force")


def field_copy(dst, src):
    """lmm_field_copy: dst = src (same context, kind size and component count)."""
    dst.ctx.check(lib().lmm_field_copy(dst.h, src.h), "field_copy")
