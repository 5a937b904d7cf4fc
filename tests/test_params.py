"""The `.param.txt` deck bridge (cosmosim.params) on the CPU.

Syntax, units and defaults follow the reference's reader (config.cpp:109-189,
300-362, 1723-1766, 4855-4889; units.cpp:20-176). Pins:
- G in code units: exactly the SI value for m / kg / m_s, and the published
  4.3009e-9 Mpc (km/s)^2 / Msun for mpc / msun / km_s (x1000 for kpc);
- lengths with and without units, in mpc and kpc code units;
- a deck with the fast TreePM deck's settings fills every option field;
- an empty deck takes the registry defaults, not the option structs';
- the reader's errors (duplicates, partial box or mesh axes, unknown units and
  keys, values the option setters reject, force-changing keys not built here).
"""
from __future__ import annotations

import ctypes as C

import pytest

from cosmosim import _lib as L
from cosmosim import params as P

FAST = """
schema_version = 1
[units]
length_unit = mpc   # code length
mass_unit = msun
velocity_unit = km_s
[cosmology]
omega_matter = 0.3099
box_size = 14.761890702961235 mpc
[numerics]
gravity_softening = 15 kpc
treepm_pm_grid = 128
treepm_asmth_cells = 1.25
treepm_rcut_cells = 6.25
treepm_assignment_scheme = tsc
treepm_enable_window_deconvolution = true
treepm_tree_opening_criterion = com_distance
treepm_tree_opening_theta = 0.7
treepm_tree_relative_force_tolerance = 0.0025
treepm_adaptive_acceptance_enabled = true
treepm_adaptive_maximum_opening_angle = 1.0
treepm_adaptive_error_estimator = relative_monopole
treepm_cutoff_node_policy = com_distance
treepm_group_walk_enabled = true
treepm_full_mac_diagnostics = false
treepm_gaussian_pair_lookup_enabled = true
treepm_full_residual_lookup_enabled = false
treepm_leaf_simd_enabled = false
treepm_radial_coefficients_enabled = false
treepm_identical_source_tree_reuse_enabled = false
treepm_topology_refit_enabled = false
treepm_residual_block_size = 64
treepm_tree_max_leaf_size = 16
"""


def test_syntax_sections_comments_quotes_and_duplicates():
    e = P.parse_entries('[Numerics]\nA = 1 # c\nb 2 ; c\nc = "x # y" // c\nd =\ne = a#b\n[mode]\nnumerics.z = 3\n')
    assert e["numerics.a"][0] == "1" and e["numerics.b"][0] == "2"
    assert e["numerics.c"][0] == '"x # y"' and e["numerics.d"][0] == ""
    assert e["numerics.e"][0] == "a#b"  # no whitespace before '#': not a comment
    assert e["numerics.z"][0] == "3"    # dotted keys keep their own section
    with pytest.raises(P.ParamError, match="line 3: duplicate key 'numerics.a' \\(first at line 2\\)"):
        P.parse_entries("[numerics]\na = 1\na = 2\n")
    with pytest.raises(P.ParamError, match="empty section"):
        P.parse_entries("[ ]\n")
    with pytest.raises(P.ParamError, match="expected key and value"):
        P.parse_entries("[numerics]\nlonely\n")


def test_gravitational_constant_in_code_units():
    assert P.gravitational_constant_code("m", "kg", "m_s") == 6.67430e-11
    g = P.gravitational_constant_code("mpc", "msun", "km_s")
    assert abs(g / 4.3009e-9 - 1.0) < 1e-4
    assert abs(P.gravitational_constant_code("kpc", "msun", "km_s") / (1e3 * g) - 1.0) < 1e-12
    with pytest.raises(P.ParamError):
        P.gravitational_constant_code("pc", "msun", "km_s")


def test_lengths_with_units():
    assert P.length_mpc("15 kpc", "mpc", "k") == pytest.approx(0.015, rel=1e-15)
    assert P.length_mpc("2.5", "kpc", "k") == pytest.approx(0.0025, rel=1e-15)
    assert P.length_mpc("2.5 MPC", "kpc", "k") == 2.5
    for bad in ("15 pc", "abc kpc", "1 kpc extra", "inf"):
        with pytest.raises(P.ParamError):
            P.length_mpc(bad, "mpc", "k")


def test_fast_deck_fills_every_field():
    d = P.treepm_deck(FAST, precision=L.PRECISION_FP64)
    o, a = d.options, d.adaptive
    assert d.mesh == (128, 128, 128)
    assert tuple(o.box_size) == (14.761890702961235,) * 3
    assert o.softening_epsilon == pytest.approx(0.015, rel=1e-15)
    assert o.gravitational_constant == d.gravitational_constant == pytest.approx(4.3009e-9, rel=1e-4)
    assert (o.asmth_cells, o.rcut_cells, o.opening_theta) == (1.25, 6.25, 0.7)
    assert o.assignment_scheme == 1 and o.window_deconvolution == 1  # TSC with deconvolution
    assert o.opening_criterion == 1  # com_distance
    assert o.relative_force_tolerance == 0.0025 and o.acceptance_policy == 1
    assert (o.residual_block_size, o.max_leaf_size) == (64, 16)
    assert a.maximum_opening_angle == 1.0 and a.error_estimator == L.ESTIMATOR_RELATIVE_MONOPOLE
    assert a.cutoff_node_policy == L.CUTOFF_COM_DISTANCE and a.group_walk == 1
    assert a.gaussian_pair_lookup == 1 and a.radial_coefficients == 0
    for k in ("numerics.treepm_full_mac_diagnostics", "numerics.treepm_leaf_simd_enabled",
              "numerics.treepm_topology_refit_enabled"):
        assert k in d.ignored
    assert "cosmology.omega_matter" not in d.applied + d.ignored
    lib = L.load()
    mesh = (C.c_int32 * 3)(*d.mesh)
    assert lib.chui_treepm_options_validate(C.byref(o), mesh) == L.CHUI_OK


def test_deck_file(tmp_path):
    f = tmp_path / "fast.param.txt"
    f.write_text(FAST)
    d = P.load_treepm_deck(f)
    assert d.mesh == (128, 128, 128) and d.adaptive.group_walk == 1


def test_empty_deck_takes_the_registry_defaults():
    d = P.treepm_deck("[numerics]\n", precision=L.PRECISION_MIXED)
    o = d.options
    assert d.mesh == (16, 16, 16) and tuple(o.box_size) == (50.0, 50.0, 50.0)
    assert o.softening_epsilon == pytest.approx(1.0e-3, rel=1e-15)
    assert (o.rcut_cells, o.opening_theta, o.relative_force_tolerance) == (6.25, 0.7, 0.005)
    assert o.assignment_scheme == 1 and o.window_deconvolution == 1 and o.acceptance_policy == 0
    assert o.precision == L.PRECISION_MIXED
    defaults = L.default_options()  # the option struct's own defaults differ
    assert (defaults.rcut_cells, defaults.assignment_scheme) != (o.rcut_cells, o.assignment_scheme)


def test_kpc_code_units_and_rectangular_box():
    d = P.treepm_deck("[units]\nlength_unit = kpc\n[cosmology]\nbox_size_x = 8000\nbox_size_y = 12 mpc\n"
                      "box_size_z = 16000 kpc\n[numerics]\ngravity_softening = 5\ntreepm_pm_grid_nx = 32\n"
                      "treepm_pm_grid_ny = 48\ntreepm_pm_grid_nz = 64\n")
    assert tuple(d.options.box_size) == pytest.approx((8.0, 12.0, 16.0), rel=1e-15)
    assert d.options.softening_epsilon == pytest.approx(0.005, rel=1e-15)
    assert d.mesh == (32, 48, 64)
    assert d.gravitational_constant == pytest.approx(P.gravitational_constant_code("mpc", "msun", "km_s") * 1e3)


@pytest.mark.parametrize("text,exc", [
    ("[cosmology]\nbox_size_x = 10\n", P.ParamError),
    ("[numerics]\ntreepm_pm_grid_nx = 32\n", P.ParamError),
    ("[numerics]\ntreepm_pm_grid = -4\n", P.ParamError),
    ("[numerics]\ngravity_softening = 1 pc\n", P.ParamError),
    ("[numerics]\ntreepm_not_a_key = 1\n", P.ParamError),
    ("[numerics]\ntreepm_topology_refit_enabled = true\n", L.ChuiUnsupported),
    ("[numerics]\ntreepm_full_residual_lookup_enabled = true\n", L.ChuiUnsupported),
    ("[numerics]\ntreepm_pm_force_gradient = finite_difference_4\n", L.ChuiUnsupported),
    ("[numerics]\ntreepm_adaptive_error_estimator = magic\n", L.ChuiInvalidArgument),
    ("[numerics]\ntreepm_assignment_scheme = pcs\n", L.ChuiInvalidArgument),
])
def test_deck_errors(text, exc):
    with pytest.raises(exc):
        P.treepm_deck(text)


def test_pm_reuse_policy_and_cadence():
    # config.cpp:3410-3434: cadence 1 under 'exact'; the stale window needs a
    # cadence in [2, 64] with global KDK and turns on the stale-field option
    # (gravity_runtime.cpp:564-565); the multirate cycle needs hierarchical KDK
    d = P.treepm_deck("[numerics]\n")
    assert (d.pm_reuse_policy, d.update_cadence_steps) == ("exact", 1)
    assert d.options.allow_stale_long_range_field == 0
    d = P.treepm_deck("[numerics]\ntreepm_pm_reuse_policy = stale_within_cadence\ntreepm_update_cadence_steps = 4\n")
    assert (d.pm_reuse_policy, d.update_cadence_steps) == ("stale_within_cadence", 4)
    assert d.options.allow_stale_long_range_field == 1
    assert "numerics.treepm_update_cadence_steps" in d.applied
    d = P.treepm_deck("[numerics]\ntreepm_pm_reuse_policy = multirate_outer_kick\ntreepm_update_cadence_steps = 2\n"
                      "hierarchical_max_rung = 3\n")
    assert d.options.allow_stale_long_range_field == 0 and d.update_cadence_steps == 2


@pytest.mark.parametrize("text", [
    "[numerics]\ntreepm_update_cadence_steps = 2\n",
    "[numerics]\ntreepm_pm_reuse_policy = stale_within_cadence\n",
    "[numerics]\ntreepm_pm_reuse_policy = stale_within_cadence\ntreepm_update_cadence_steps = 65\n",
    "[numerics]\ntreepm_pm_reuse_policy = stale_within_cadence\ntreepm_update_cadence_steps = 4\n"
    "hierarchical_max_rung = 2\n",
    "[numerics]\ntreepm_pm_reuse_policy = multirate_outer_kick\ntreepm_update_cadence_steps = 4\n",
    "[numerics]\ntreepm_pm_reuse_policy = sometimes\n",
    "[numerics]\ntreepm_update_cadence_steps = two\n",
])
def test_pm_cadence_rejected(text):
    with pytest.raises(P.ParamError):
        P.treepm_deck(text)


def test_options_set_reuse_policy():
    o = L.default_options()
    assert o.allow_stale_long_range_field == 0 and o.force_epoch_scale_factor == 1.0
    L.set_param(o, "treepm_pm_reuse_policy", "stale_within_cadence")
    assert o.allow_stale_long_range_field == 1
    L.set_param(o, "numerics.treepm_pm_reuse_policy", "exact")
    assert o.allow_stale_long_range_field == 0
    with pytest.raises(L.ChuiInvalidArgument):
        L.set_param(o, "treepm_pm_reuse_policy", "never")
