"""The reference's `.param.txt` decks -> chui_treepm options, adaptive options
and mesh shape.

Follows the reference's reading of a deck as far as the TreePM force path
needs it:
- syntax (config.cpp:109-189): `[section]` headers, `key = value` (or
  `key value`) lines, comments from `#`, `;` or `//` at a whitespace boundary
  outside quotes, lower-cased keys prefixed by their section, duplicate keys
  rejected with both line numbers;
- units (units.cpp:20-63, 149-176): code length, mass and velocity units give
  G in code units, G_SI M_unit T_unit^2 / L_unit^3 with T_unit = L_unit /
  V_unit;
- lengths (config.cpp:300-362): `magnitude [unit]` with unit mpc or kpc,
  defaulting to units.length_unit; the box is in Mpc (cosmology.box_size, or
  box_size_x/y/z together, :4855-4889); the softening is converted to kpc
  and handed to the force path in Mpc (gravity_runtime.cpp:566-572);
- the mesh is numerics.treepm_pm_grid, or treepm_pm_grid_nx/ny/nz together;
- unset TreePM keys take the deck defaults of the reference's key registry
  (config.cpp:1723-1766), which differ from the option structs' defaults
  (for example r_cut 6.25 cells, theta 0.7, TSC with deconvolution);
- the remaining numerics.treepm_* keys go through chui_treepm_options_set and
  chui_treepm_adaptive_options_set, which apply the reference's value checks.

Keys that only affect diagnostics or CPU-side performance are reported as
ignored; keys that would change the forces in a way this library does not
build raise ChuiUnsupported when set away from their defaults.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

from . import _lib


class ParamError(ValueError):
    """A deck the reference's reader rejects (its ConfigError)."""


# ---- syntax --------------------------------------------------------------------
def _strip_comment(line: str) -> str:
    single = double = False
    for pos, ch in enumerate(line):
        if ch == '"' and not single:
            double = not double
            continue
        if ch == "'" and not double:
            single = not single
            continue
        if single or double:
            continue
        boundary = pos == 0 or line[pos - 1].isspace()
        if boundary and (ch in "#;" or line.startswith("//", pos)):
            return line[:pos]
    return line


def parse_entries(text: str) -> dict:
    """{section.key: (value, line number)} in the reference's syntax."""
    entries: dict = {}
    section = ""
    for number, raw in enumerate(text.splitlines(), start=1):
        line = _strip_comment(raw).strip()
        if not line:
            continue
        if line.startswith("[") and line.endswith("]"):
            section = line[1:-1].strip().lower()
            if not section:
                raise ParamError(f"line {number}: empty section name")
            continue
        if "=" in line:
            key, value = (part.strip() for part in line.split("=", 1))
            allow_empty = True
        else:
            parts = line.split(None, 1)
            key, value = parts[0], (parts[1].strip() if len(parts) > 1 else "")
            allow_empty = False
        key = key.strip().lower()
        if not key or (not value and not allow_empty):
            raise ParamError(f"line {number}: expected key and value")
        if section and "." not in key:
            key = f"{section}.{key}"
        if key in entries:
            raise ParamError(f"line {number}: duplicate key '{key}' (first at line {entries[key][1]})")
        entries[key] = (value, number)
    return entries


# ---- units -----------------------------------------------------------------------
# constants.hpp: SI values of the units the reference's decks name
NEWTON_G_SI = 6.67430e-11
PARSEC_SI = 3.0856775814913673e16
LENGTH_SI = {"m": 1.0, "cm": 1.0e-2, "kpc": 1.0e3 * PARSEC_SI, "mpc": 1.0e6 * PARSEC_SI}
MASS_SI = {"kg": 1.0, "g": 1.0e-3, "msun": 1.98847e30}
VELOCITY_SI = {"m_s": 1.0, "cm_s": 1.0e-2, "km_s": 1.0e3}


def gravitational_constant_code(length_unit: str, mass_unit: str, velocity_unit: str) -> float:
    """G in code units (newtonGravitationalConstantCode)."""
    try:
        l_si = LENGTH_SI[length_unit.lower()]
        m_si = MASS_SI[mass_unit.lower()]
        v_si = VELOCITY_SI[velocity_unit.lower()]
    except KeyError as e:
        raise ParamError(f"unsupported unit: {e.args[0]}") from None
    t_si = l_si / v_si
    return NEWTON_G_SI * m_si * t_si * t_si / (l_si * l_si * l_si)


def split_magnitude_unit(value: str, key: str) -> tuple:
    parts = value.split()
    try:
        magnitude = float(parts[0]) if parts else float("nan")
    except ValueError:
        magnitude = float("nan")
    if not parts or magnitude != magnitude or magnitude in (float("inf"), float("-inf")):
        raise ParamError(f"key '{key}': invalid value '{value}': expected finite numeric magnitude")
    if len(parts) > 2:
        raise ParamError(f"key '{key}': invalid value '{value}': unexpected trailing token(s) after magnitude/unit")
    return magnitude, (parts[1].lower() if len(parts) == 2 else "")


def length_mpc(value: str, default_unit: str, key: str) -> float:
    magnitude, unit = split_magnitude_unit(value, key)
    unit = unit or default_unit.lower()
    if unit == "mpc":
        return magnitude
    if unit == "kpc":
        return magnitude * 1.0e-3
    raise ParamError(f"key '{key}': unsupported unit in value '{value}'")


# ---- the TreePM deck ---------------------------------------------------------------
# config.cpp:1723-1766: deck defaults of the keys the force path reads
DECK_DEFAULTS = {
    "numerics.treepm_asmth_cells": "1.25",
    "numerics.treepm_rcut_cells": "6.25",
    "numerics.treepm_tree_opening_criterion": "com_distance",
    "numerics.treepm_tree_opening_theta": "0.7",
    "numerics.treepm_tree_relative_force_tolerance": "0.005",
    "numerics.treepm_tree_relative_force_acceleration_floor": "1.0e-30",
    "numerics.treepm_residual_block_size": "64",
    "numerics.treepm_tree_max_leaf_size": "16",
    "numerics.treepm_adaptive_acceptance_enabled": "false",
    "numerics.treepm_assignment_scheme": "tsc",
    "numerics.treepm_enable_window_deconvolution": "true",
    "numerics.treepm_pm_force_gradient": "spectral",
}
ADAPTIVE_DEFAULTS = {
    "numerics.treepm_adaptive_maximum_opening_angle": "0.25",
    "numerics.treepm_adaptive_error_estimator": "residual_derivative",
    "numerics.treepm_cutoff_node_policy": "contain",
    "numerics.treepm_group_walk_enabled": "false",
    "numerics.treepm_group_walk_target_size": "0",
    "numerics.treepm_gaussian_pair_lookup_enabled": "false",
    "numerics.treepm_radial_coefficients_enabled": "false",
}
# keys that leave the forces unchanged here (diagnostics, CPU scheduling,
# distributed exchange sizes, caller-side cadence)
IGNORED = {
    "numerics.treepm_full_mac_diagnostics", "numerics.treepm_leaf_simd_enabled",
    "numerics.treepm_identical_source_tree_reuse_enabled", "numerics.treepm_multirate_rollback_memory_policy",
    "numerics.treepm_allow_diagnostic_naive_dft", "numerics.treepm_pm_decomposition_mode",
    "numerics.treepm_tree_exchange_batch_bytes", "numerics.treepm_pm_exchange_batch_bytes",
}
# keys that change the forces in a form not built here: allowed at their defaults only
UNSUPPORTED = {
    "numerics.treepm_full_residual_lookup_enabled": "false",
    "numerics.treepm_topology_refit_enabled": "false",
}


@dataclass
class TreePmDeck:
    options: _lib.Options
    adaptive: _lib.AdaptiveOptions
    mesh: tuple
    gravitational_constant: float
    applied: list = field(default_factory=list)   # deck keys that set a field
    ignored: list = field(default_factory=list)   # numerics.treepm_* keys with no effect on the forces
    # PM cadence of the caller's integrator (numerics.treepm_update_cadence_steps)
    # and the reuse policy; stale_within_cadence sets
    # options.allow_stale_long_range_field
    update_cadence_steps: int = 1
    pm_reuse_policy: str = "exact"


def _mesh(entries) -> tuple:
    axes = [f"numerics.treepm_pm_grid_n{a}" for a in "xyz"]
    given = [k in entries for k in axes]
    if any(given) and not all(given):
        raise ParamError("numerics.treepm_pm_grid_nx, _ny and _nz must be specified together")
    keys = axes if all(given) else ["numerics.treepm_pm_grid"] * 3
    out = []
    for k in keys:
        value = entries.get(k, ("16", 0))[0]
        try:
            n = int(value)
        except ValueError:
            raise ParamError(f"key '{k}': expected an integer, got '{value}'") from None
        if n <= 0:
            raise ParamError("numerics.treepm_pm_grid_n{xyz} must all be > 0")
        out.append(n)
    return tuple(out)


def _box_mpc(entries, length_unit: str) -> tuple:
    axes = [f"cosmology.box_size_{a}" for a in "xyz"]
    given = [k in entries for k in axes]
    if any(given) and not all(given):
        raise ParamError("cosmology.box_size_x, cosmology.box_size_y, and cosmology.box_size_z must be "
                         "specified together")
    if all(given):
        return tuple(length_mpc(entries[k][0], length_unit, k) for k in axes)
    v = length_mpc(entries.get("cosmology.box_size", ("50.0", 0))[0], length_unit, "cosmology.box_size")
    return (v, v, v)


def _int_key(entries, key: str, default: str) -> int:
    value = entries.get(key, (default, 0))[0]
    try:
        return int(value)
    except ValueError:
        raise ParamError(f"key '{key}': expected an integer, got '{value}'") from None


def _pm_cadence(entries) -> tuple:
    """(policy, cadence) with the reference's checks (config.cpp:987-996, 3410-3434)."""
    policy = entries.get("numerics.treepm_pm_reuse_policy", ("exact", 0))[0].strip().lower()
    if policy not in ("exact", "stale_within_cadence", "multirate_outer_kick"):
        raise ParamError(f"key 'numerics.treepm_pm_reuse_policy': invalid value '{policy}' "
                         "(supported: exact, stale_within_cadence, multirate_outer_kick)")
    cadence = _int_key(entries, "numerics.treepm_update_cadence_steps", "1")
    max_rung = _int_key(entries, "numerics.hierarchical_max_rung", "0")
    if policy == "exact" and cadence != 1:
        raise ParamError("numerics.treepm_update_cadence_steps must be exactly 1 with treepm_pm_reuse_policy=exact; "
                         "select treepm_pm_reuse_policy=stale_within_cadence for experimental PM reuse")
    if policy == "stale_within_cadence" and (cadence < 2 or cadence > 64 or max_rung != 0):
        raise ParamError("numerics.treepm_pm_reuse_policy=stale_within_cadence requires treepm_update_cadence_steps "
                         "in [2,64] and global KDK (hierarchical_max_rung=0; hierarchical KDK already applies PM at "
                         "coarse endpoints)")
    if policy == "multirate_outer_kick" and (cadence < 2 or cadence > 64 or max_rung <= 0):
        raise ParamError("numerics.treepm_pm_reuse_policy=multirate_outer_kick requires treepm_update_cadence_steps "
                         "in [2,64] (coarse blocks per PM cycle) and hierarchical KDK (hierarchical_max_rung > 0)")
    return policy, cadence


def load_treepm_deck(path, precision: int = _lib.PRECISION_FP64) -> TreePmDeck:
    """Options of the TreePM force path from a deck file."""
    return treepm_deck(Path(path).read_text(), precision)


def treepm_deck(text: str, precision: int = _lib.PRECISION_FP64) -> TreePmDeck:
    """Options of the TreePM force path from the text of a deck."""
    entries = parse_entries(text)
    units = {u: entries.get(f"units.{u}_unit", (d, 0))[0]
             for u, d in (("length", "mpc"), ("mass", "msun"), ("velocity", "km_s"))}
    g_code = gravitational_constant_code(units["length"], units["mass"], units["velocity"])
    opts = _lib.default_options()
    ad = _lib.default_adaptive_options()
    deck = TreePmDeck(opts, ad, _mesh(entries), g_code)
    for k in ("units.length_unit", "units.mass_unit", "units.velocity_unit"):
        if k in entries:
            deck.applied.append(k)
    deck.applied += [k for k in entries if k.startswith("numerics.treepm_pm_grid")]
    box = _box_mpc(entries, units["length"])
    opts.box_size = (type(opts.box_size))(*box)
    deck.applied += [k for k in entries if k.startswith("cosmology.box_size")]
    opts.gravitational_constant = g_code
    soft_kpc = 1.0e3 * length_mpc(entries.get("numerics.gravity_softening", ("1.0 kpc", 0))[0], units["length"],
                                  "numerics.gravity_softening")
    _lib.set_param(opts, "gravity_softening_mpc", repr(soft_kpc * 1.0e-3))
    if "numerics.gravity_softening" in entries:
        deck.applied.append("numerics.gravity_softening")
    for key, default in DECK_DEFAULTS.items():
        value = entries[key][0] if key in entries else default
        _lib.set_param(opts, key, value)
        if key in entries:
            deck.applied.append(key)
    for key, default in ADAPTIVE_DEFAULTS.items():
        value = entries[key][0] if key in entries else default
        _lib.set_adaptive_param(ad, key, value)
        if key in entries:
            deck.applied.append(key)
    deck.pm_reuse_policy, deck.update_cadence_steps = _pm_cadence(entries)
    _lib.set_param(opts, "treepm_pm_reuse_policy", deck.pm_reuse_policy)
    deck.applied += [k for k in ("numerics.treepm_pm_reuse_policy", "numerics.treepm_update_cadence_steps")
                     if k in entries]
    for key, default in UNSUPPORTED.items():
        if key in entries and entries[key][0].lower() != default:
            raise _lib.ChuiUnsupported(_lib.CHUI_ERR_UNSUPPORTED,
                                       f"{key} = {entries[key][0]} changes the forces in a form not built here")
    for key in entries:
        if key.startswith("numerics.treepm_") and key not in deck.applied:
            deck.ignored.append(key)
            if key not in IGNORED and key not in UNSUPPORTED:
                raise ParamError(f"unknown TreePM deck key: {key}")
    opts.precision = precision
    return deck


__all__ = ["ParamError", "TreePmDeck", "gravitational_constant_code", "length_mpc", "load_treepm_deck",
           "parse_entries", "treepm_deck"]
