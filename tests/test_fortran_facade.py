"""The Fortran side of the boundary (lowmachmixtures_amd/fortran/lmm_facade.f90)
compiled with amdflang and linked against liblmm_hip.so; a Fortran program
calls the C ABI through iso_c_binding without a GPU (defaults, argument
checks that return before the device is touched, the host generator, struct
layouts).  Skipped where no Fortran compiler is installed."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from lowmachmixtures_amd import stokes as S
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "lowmachmixtures_amd", "fortran", "lmm_facade.f90")
CHECK = os.path.join(ROOT, "tests", "fortran", "facade_check.f90")


def _flang():
    for cand in ("/opt/rocm/llvm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


@pytest.fixture(scope="module")
def run_check(tmp_path_factory):
    fc = _flang()
    if fc is None:
        pytest.skip("no Fortran compiler")
    S.lib()  # the library is built
    libdir = os.path.dirname(S.LIB_PATH)
    d = tmp_path_factory.mktemp("facade")
    exe = str(d / "facade_check")
    subprocess.run([fc, "-c", FACADE, "-J", str(d), "-o", str(d / "lmm_facade.o")], check=True, cwd=d)
    subprocess.run([fc, CHECK, str(d / "lmm_facade.o"), "-I", str(d), "-o", exe, f"-L{libdir}", "-llmm_hip",
                    f"-Wl,-rpath,{libdir}"], check=True, cwd=d)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
    vals = {}
    for line in out.splitlines():
        w = line.split()
        if w:
            vals[w[0]] = w[1:]
    return vals


def test_defaults_through_fortran(run_check):
    v = run_check
    assert v["visc_type"] == ["1"] and v["precon_type"] == ["1"]
    assert v["gmres_max_inner"] == ["5"] and v["gmres_max_outer"] == ["20"]
    assert v["stag_mg_nsmooths_bottom"] == ["8"]
    assert float(v["gmres_rel_tol"][0]) == 1e-9
    assert v["stoch_stress_form"] == ["1"]


def test_struct_layouts_match_the_c_abi(run_check):
    assert int(run_check["sizeof_params"][0]) == C.sizeof(S.Params)
    assert int(run_check["sizeof_info"][0]) == C.sizeof(S.GmresInfo)
    assert int(run_check["sizeof_species"][0]) == C.sizeof(S.SpeciesParams)
    assert int(run_check["sizeof_inertial"][0]) == C.sizeof(S.InertialParams)


def test_species_knobs_through_fortran(run_check):
    v = run_check
    assert v["species_n"] == ["4"] and v["chi_iterations"] == ["10"]
    assert float(v["fraction_tolerance"][0]) == 1e-13
    assert v["species_flags"] == ["1", "1", "0"]   # ideal, non-isothermal, iterative chi
    assert v["set_species_null_ctx"] == ["1"] and v["mass_fluxdiv_null_ctx"] == ["1"]


def test_argument_rejections_through_fortran(run_check):
    v = run_check
    assert v["create_dim4"] == ["1"]      # LMM_ERR_ARG
    assert v["create_odd"] == ["2"]       # LMM_ERR_UNSUPPORTED
    assert v["create_visc4"] == ["1"]
    assert v["create_filter3"] == ["1"]
    assert v["gmres_null_ctx"] == ["1"]


def test_host_generator_through_fortran(run_check):
    assert run_check["fill_rc"] == ["0"]
    got = np.array([float(x) for x in run_check["normals"]])
    assert np.array_equal(got, O.fill_normal(99, 4))


def test_inertial_knobs_through_fortran(run_check):
    v = run_check
    assert [float(x) for x in v["inertial_coefs"]] == [1.0, 1.0]  # variance_coef_mom, visc_coef
    assert v["inertial_flags"] == ["0", "0", "0"]  # constant eta, centred advection, device noise
    assert v["inertial_null_ctx"] == ["1"]
    assert v["overdamped_knobs"] == ["0", "-1"]  # algorithm_type, restart (probin_common.f90:205, :218)
    assert v["overdamped_null_ctx"] == ["1"]
