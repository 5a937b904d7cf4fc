"""Host-side checks of the C ABI that need no GPU: the library loads, exports
every entry point include/chui_treepm.h declares, and its host-only logic
(option defaults, .param.txt keys, validation, split diagnostics, KDK factors,
argument errors) behaves like the reference's.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cosmosim._lib as L

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "chui_treepm.h"
FIX = json.loads((Path(__file__).parent / "golden" / "reference_fixtures.json").read_text())


def declared_entry_points():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(chui_treepm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    return L.load()


def test_library_exports_every_declared_symbol(lib):
    names = declared_entry_points()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], check=True, capture_output=True,
                         text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)
    # the Python binding declares a signature for every entry point
    assert sorted(L.SIGNATURES) == names


def test_abi_version(lib):
    m = re.search(r"#define\s+CHUI_TREEPM_ABI_VERSION\s+(\d+)", HEADER.read_text())
    assert lib.chui_treepm_abi_version() == int(m.group(1))


def test_default_options_mirror_reference_structs():
    # TreeGravityOptions (tree_gravity.hpp:30-43), PmSolveOptions
    # (pm_solver.hpp:223-254), split cells as config.hpp:783-794.
    o = L.default_options()
    assert o.struct_size == C.sizeof(L.Options)
    assert o.opening_theta == 0.7
    assert o.opening_criterion == 1  # COM distance
    assert o.multipole_order == 2
    assert o.max_leaf_size == 16
    assert o.relative_force_tolerance == 0.005
    assert o.relative_force_acceleration_floor == 1e-30
    assert o.gravitational_constant == 1.0 and o.scale_factor == 1.0
    assert o.assignment_scheme == 0 and o.window_deconvolution == 0
    assert o.asmth_cells == 1.25 and o.rcut_cells == 6.25
    assert o.residual_block_size == 64
    assert o.precision == L.PRECISION_FP64


def test_options_set_accepts_param_keys(lib):
    o = L.default_options()
    for key, value in (("numerics.treepm_asmth_cells", "1.5"), ("numerics.treepm_rcut_cells", "5.0"),
                       ("numerics.treepm_tree_opening_theta", "0.5"),
                       ("numerics.treepm_tree_opening_criterion", "geometric"),
                       ("numerics.treepm_tree_max_leaf_size", "8"), ("numerics.treepm_assignment_scheme", "tsc"),
                       ("numerics.treepm_enable_window_deconvolution", "true"), ("box_size", "100"),
                       ("gravity_softening", "20 kpc"), ("treepm_precision", "mixed")):
        assert lib.chui_treepm_options_set(C.byref(o), key.encode(), value.encode()) == L.CHUI_OK
    assert o.asmth_cells == 1.5 and o.rcut_cells == 5.0 and o.opening_theta == 0.5
    assert o.opening_criterion == 0 and o.max_leaf_size == 8 and o.assignment_scheme == 1
    assert o.window_deconvolution == 1 and list(o.box_size) == [100.0] * 3
    assert o.softening_epsilon == pytest.approx(0.02)  # kpc -> Mpc
    assert o.precision == L.PRECISION_MIXED


def test_options_set_lengths_take_units():
    # config.cpp:300-362: "<magnitude> [mpc|kpc]", a bare magnitude in the
    # default length unit (Mpc)
    o = L.default_options()
    for key, value, field, want in (("gravity_softening", "0.015", "softening_epsilon", 0.015),
                                    ("gravity_softening", "15 KPC", "softening_epsilon", 0.015),
                                    ("box_size", "14761.89 kpc", "box_size", 14.76189),
                                    ("box_size_y", "12 mpc", "box_size", 12.0)):
        assert L.load().chui_treepm_options_set(C.byref(o), key.encode(), value.encode()) == L.CHUI_OK
        got = o.softening_epsilon if field == "softening_epsilon" else o.box_size[1]
        assert got == pytest.approx(want, rel=1e-14)


@pytest.mark.parametrize("key,value,status", [
    ("gravity_softening", "1 pc", L.CHUI_ERR_INVALID_ARGUMENT),
    ("box_size", "10 mpc extra", L.CHUI_ERR_INVALID_ARGUMENT),
    ("box_size", "inf", L.CHUI_ERR_INVALID_ARGUMENT),
    ("numerics.treepm_no_such_key", "1", L.CHUI_ERR_INVALID_ARGUMENT),
    ("numerics.treepm_asmth_cells", "abc", L.CHUI_ERR_INVALID_ARGUMENT),
    ("numerics.treepm_enable_window_deconvolution", "maybe", L.CHUI_ERR_INVALID_ARGUMENT),
    ("numerics.treepm_assignment_scheme", "pcs", L.CHUI_ERR_INVALID_ARGUMENT),
    ("numerics.treepm_pm_grid", "256", L.CHUI_ERR_INVALID_ARGUMENT),
    ("numerics.treepm_pm_force_gradient", "finite_difference", L.CHUI_ERR_UNSUPPORTED),
])
def test_options_set_rejects(lib, key, value, status):
    o = L.default_options()
    assert lib.chui_treepm_options_set(C.byref(o), key.encode(), value.encode()) == status
    assert lib.chui_treepm_last_error().decode()


def valid_options():
    return L.default_options(box_size=100.0, softening_epsilon=0.01, rcut_cells=5.625, asmth_cells=1.25)


MESH = (C.c_int32 * 3)(64, 64, 64)


@pytest.mark.parametrize("field,value,status", [
    (None, None, L.CHUI_OK),
    ("softening_epsilon", 0.5, L.CHUI_ERR_INVALID_ARGUMENT),   # eps > 0.2 r_s (config.cpp:3669)
    ("rcut_cells", 40.0, L.CHUI_ERR_INVALID_ARGUMENT),         # r_cut >= L/2
    ("opening_theta", 0.0, L.CHUI_ERR_INVALID_ARGUMENT),
    ("max_leaf_size", 0, L.CHUI_ERR_INVALID_ARGUMENT),
    ("residual_block_size", 48, L.CHUI_ERR_INVALID_ARGUMENT),
    ("gravitational_constant", -1.0, L.CHUI_ERR_OVERFLOW),
    ("assignment_scheme", 7, L.CHUI_ERR_INVALID_ARGUMENT),
    ("multipole_order", 1, L.CHUI_ERR_INVALID_ARGUMENT),
    ("acceptance_policy", 1, L.CHUI_OK),                        # adaptive (FP64)
    ("acceptance_policy", 2, L.CHUI_ERR_INVALID_ARGUMENT),
    ("target_mapping", L.TARGETS_PER_WAVE, L.CHUI_OK),
    ("target_mapping", 3, L.CHUI_ERR_INVALID_ARGUMENT),
    ("target_mapping", -1, L.CHUI_ERR_INVALID_ARGUMENT),
    ("struct_size", 8, L.CHUI_ERR_INVALID_ARGUMENT),
])
def test_options_validate(lib, field, value, status):
    o = valid_options()
    if field is not None:
        setattr(o, field, value)
    assert lib.chui_treepm_options_validate(C.byref(o), MESH) == status


def test_split_diagnostics_known_answers(lib):
    f = FIX["split_mesh_semantics"]
    d = L.SplitDiagnostics()
    assert lib.chui_treepm_split_diagnostics(f["asmth_cells"], f["rcut_cells"], f["mesh_spacing"], C.byref(d)) == 0
    assert abs(d.split_scale_comoving - f["asmth_cells"] * f["mesh_spacing"]) < f["tolerance"]
    assert abs(d.cutoff_radius_comoving - f["rcut_cells"] * f["mesh_spacing"]) < f["tolerance"]
    g = FIX["split_diagnostics_continuity"]
    assert lib.chui_treepm_split_diagnostics(g["asmth_cells"], g["rcut_cells"], g["mesh_spacing"], C.byref(d)) == 0
    assert d.short_range_factor_at_split > 0 and d.long_range_factor_at_split > 0
    assert 0 <= d.short_range_factor_at_cutoff < 0.5 and d.long_range_factor_at_cutoff > 0.5
    assert d.composition_error_at_split < g["composition_tolerance"]
    assert d.max_relative_composition_error < g["composition_tolerance"]
    # S(r_s) from the closed form erfc(1/2) + exp(-1/4)/sqrt(pi)
    assert abs(d.short_range_factor_at_split - (math.erfc(0.5) + math.exp(-0.25) / math.sqrt(math.pi))) < 1e-15
    assert lib.chui_treepm_split_diagnostics(0.0, 4.5, 1.0, C.byref(d)) == L.CHUI_ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("a0,a1", [(0.05, 0.095), (0.3, 0.5), (0.9, 1.0)])
def test_kdk_factors_against_quadrature(lib, a0, a1):
    from scipy.integrate import quad

    H0, om, ol = 100.0, 0.3, 0.7
    H = lambda a: H0 * math.sqrt(om / a**3 + ol)  # noqa: E731
    f = L.KdkFactors()
    assert lib.chui_treepm_kdk_factors(a0, a1, H0, om, ol, C.byref(f)) == 0
    drift = quad(lambda a: 1 / (a * a * H(a)), a0, a1, epsabs=0, epsrel=1e-13)[0]
    assert abs(f.drift_factor / drift - 1) < 1e-4  # 64-sample midpoint rule
    am = f.scale_factor_midpoint
    t1 = quad(lambda a: 1 / (a * H(a)), a0, am, epsabs=0, epsrel=1e-13)[0]
    t2 = quad(lambda a: 1 / (a * H(a)), am, a1, epsabs=0, epsrel=1e-13)[0]
    assert abs(t1 - t2) < 1e-8 * (t1 + t2)  # am halves the cosmic time
    k1 = quad(lambda a: 1 / (a * a * H(a)), a0, am, epsabs=0, epsrel=1e-13)[0] / am
    k2 = quad(lambda a: 1 / (a * a * H(a)), am, a1, epsabs=0, epsrel=1e-13)[0] / a1
    assert abs(f.first_kick_factor / k1 - 1) < 1e-4 and abs(f.second_kick_factor / k2 - 1) < 1e-4
    assert f.first_hubble_drag == a0 / am and f.second_hubble_drag == am / a1


def test_kdk_factors_reject_bad_arguments(lib):
    f = L.KdkFactors()
    assert lib.chui_treepm_kdk_factors(0.5, 0.4, 100.0, 0.3, 0.7, C.byref(f)) == L.CHUI_ERR_INVALID_ARGUMENT
    assert lib.chui_treepm_kdk_factors(0.0, 0.4, 100.0, 0.3, 0.7, C.byref(f)) == L.CHUI_ERR_INVALID_ARGUMENT


def test_context_argument_errors_before_any_device_call(lib):
    assert lib.chui_treepm_destroy(None) == L.CHUI_OK
    ctx = C.c_void_p()
    for bad_axis in (40, 45, 3):  # a factor 5, odd, odd: not an even 2^a 3^b mesh
        bad = (C.c_int32 * 3)(bad_axis, 64, 64)
        assert lib.chui_treepm_create(bad, 0, C.byref(ctx)) == L.CHUI_ERR_UNSUPPORTED
    zero = (C.c_int32 * 3)(0, 64, 64)
    assert lib.chui_treepm_create(zero, 0, C.byref(ctx)) == L.CHUI_ERR_INVALID_ARGUMENT
    o = valid_options()
    st = lib.chui_treepm_solve_device(None, None, None, None, None, 0, None, 0, None, None, None, None, None, None,
                                      None, None, None, C.byref(o), 1, None)
    assert st == L.CHUI_ERR_INVALID_ARGUMENT
    assert "context" in lib.chui_treepm_last_error().decode().lower()


def test_python_mirror_raises_reference_exception_types():
    with pytest.raises(ValueError):
        L.check(L.CHUI_ERR_INVALID_ARGUMENT)
    with pytest.raises(OverflowError):
        L.check(L.CHUI_ERR_OVERFLOW)
    with pytest.raises(NotImplementedError):
        L.check(L.CHUI_ERR_UNSUPPORTED)
    assert np.dtype(np.float64).itemsize == 8


def test_adaptive_policy_validates_in_both_precisions(lib):
    o = valid_options()
    o.acceptance_policy = 1
    for precision in (L.PRECISION_FP64, L.PRECISION_MIXED):
        o.precision = precision
        assert lib.chui_treepm_options_validate(C.byref(o), MESH) == L.CHUI_OK


def test_adaptive_options_mirror_reference_defaults(lib):
    # TreePmOptions (tree_pm_coupling.hpp:248-270)
    assert C.sizeof(L.AdaptiveOptions) == 48 and C.sizeof(L.WalkCounters) == 128
    ad = L.default_adaptive_options()
    assert ad.struct_size == C.sizeof(L.AdaptiveOptions)
    assert ad.maximum_opening_angle == 0.25
    assert ad.error_estimator == L.ESTIMATOR_RESIDUAL_DERIVATIVE
    assert ad.cutoff_node_policy == L.CUTOFF_CONTAIN_NODE
    assert ad.gaussian_pair_lookup == 0 and ad.radial_coefficients == 0 and ad.group_walk == 0
    assert ad.group_target_size == 0 and ad.group_max_extent_cutoff_fraction == 0.5


def test_adaptive_options_accept_the_fast_deck_keys(lib):
    # configs/treepm_evolved_snapshot_fast.param.txt:63-73
    ad = L.default_adaptive_options()
    for key, value in [("treepm_adaptive_maximum_opening_angle", "1.0"),
                       ("treepm_adaptive_error_estimator", "relative_monopole"),
                       ("numerics.treepm_cutoff_node_policy", "com_distance"),
                       ("treepm_group_walk_enabled", "true"),
                       ("treepm_gaussian_pair_lookup_enabled", "true"),
                       ("treepm_radial_coefficients_enabled", "false"),
                       ("treepm_group_walk_target_size", "32")]:
        L.set_adaptive_param(ad, key, value)
    assert ad.maximum_opening_angle == 1.0 and ad.error_estimator == L.ESTIMATOR_RELATIVE_MONOPOLE
    assert ad.cutoff_node_policy == L.CUTOFF_COM_DISTANCE and ad.group_walk == 1
    assert ad.gaussian_pair_lookup == 1 and ad.radial_coefficients == 0 and ad.group_target_size == 32
    o = valid_options()
    L.set_param(o, "treepm_adaptive_acceptance_enabled", True)
    assert o.acceptance_policy == 1
    for key, value in [("treepm_adaptive_error_estimator", "monopole"), ("treepm_cutoff_node_policy", "x"),
                       ("treepm_group_walk_target_size", "1"), ("treepm_group_walk_enabled", "maybe"),
                       ("treepm_unknown", "1")]:
        assert lib.chui_treepm_adaptive_options_set(C.byref(ad), key.encode(), value.encode()) == \
            L.CHUI_ERR_INVALID_ARGUMENT


def test_adaptive_entry_point_argument_errors(lib):
    o = valid_options()
    o.acceptance_policy = 1
    ad = L.default_adaptive_options()
    args = [None] * 16
    st = lib.chui_treepm_solve_adaptive_device(None, *args[:4], 0, None, 0, *args[:10], None, C.byref(o),
                                               C.byref(ad), 1, None)
    assert st == L.CHUI_ERR_INVALID_ARGUMENT
