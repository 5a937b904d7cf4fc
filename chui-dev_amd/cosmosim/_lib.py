"""ctypes binding of libchui_treepm.so (the C ABI in include/chui_treepm.h).

The library is loaded from the package directory; there is no fallback: a
missing or unloadable library raises ImportError at first use so that no code
path can silently run without the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PACKAGE_DIR / "libchui_treepm.so"

CHUI_OK = 0
CHUI_ERR_INVALID_ARGUMENT = 1
CHUI_ERR_RUNTIME = 2
CHUI_ERR_OVERFLOW = 3
CHUI_ERR_LOGIC = 4
CHUI_ERR_DEVICE = 5
CHUI_ERR_OUT_OF_MEMORY = 6
CHUI_ERR_UNSUPPORTED = 7

PRECISION_FP64 = 0
PRECISION_MIXED = 1

TARGETS_AUTO = 0
TARGETS_PER_LANE = 1
TARGETS_PER_WAVE = 2


class ChuiError(RuntimeError):
    """Base class; .status carries the chui_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class ChuiInvalidArgument(ChuiError, ValueError):
    pass


class ChuiOverflow(ChuiError, OverflowError):
    pass


class ChuiUnsupported(ChuiError, NotImplementedError):
    pass


class ChuiDeviceError(ChuiError):
    pass


class ChuiRuntimeError(ChuiError):
    """std::runtime_error of the reference (cadence / field identity violations)."""


class ChuiLogicError(ChuiError):
    """std::logic_error of the reference (a broken internal invariant)."""


class ChuiOutOfMemory(ChuiError, MemoryError):
    pass


_ERRORS = {
    CHUI_ERR_INVALID_ARGUMENT: ChuiInvalidArgument,
    CHUI_ERR_OVERFLOW: ChuiOverflow,
    CHUI_ERR_UNSUPPORTED: ChuiUnsupported,
    CHUI_ERR_RUNTIME: ChuiRuntimeError,
    CHUI_ERR_LOGIC: ChuiLogicError,
    CHUI_ERR_DEVICE: ChuiDeviceError,
    CHUI_ERR_OUT_OF_MEMORY: ChuiOutOfMemory,
}


class Options(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("box_size", C.c_double * 3),
        ("scale_factor", C.c_double),
        ("gravitational_constant", C.c_double),
        ("assignment_scheme", C.c_int32),
        ("window_deconvolution", C.c_int32),
        ("asmth_cells", C.c_double),
        ("rcut_cells", C.c_double),
        ("opening_theta", C.c_double),
        ("opening_criterion", C.c_int32),
        ("multipole_order", C.c_int32),
        ("max_leaf_size", C.c_int32),
        ("target_mapping", C.c_int32),
        ("softening_epsilon", C.c_double),
        ("relative_force_tolerance", C.c_double),
        ("relative_force_acceleration_floor", C.c_double),
        ("acceptance_policy", C.c_int32),
        ("residual_block_size", C.c_int32),
        ("precision", C.c_int32),
        ("short_range_only", C.c_int32),
        ("source_generation", C.c_uint64),
        ("pm_field_version", C.c_uint64),
        ("force_epoch_sequence", C.c_uint64),
        ("force_epoch_scale_factor", C.c_double),
        ("allow_stale_long_range_field", C.c_int32),
        ("reserved1", C.c_int32),
    ]


class Diagnostics(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("reserved0", C.c_uint32),
        ("node_count", C.c_uint64),
        ("leaf_count", C.c_uint64),
        ("max_depth", C.c_uint64),
        ("direct_pair_evaluations", C.c_uint64),
        ("accepted_internal_multipoles", C.c_uint64),
        ("accepted_leaves", C.c_uint64),
        ("cutoff_pruned_nodes", C.c_uint64),
        ("visited_nodes", C.c_uint64),
        ("split_scale_comoving", C.c_double),
        ("cutoff_radius_comoving", C.c_double),
        ("mesh_spacing_comoving", C.c_double),
        ("tree_root_half_size_comoving", C.c_double),
        ("unwrap_anchor", C.c_double * 3),
        ("ms_total", C.c_double),
        ("ms_unwrap_keys_sort", C.c_double),
        ("ms_tree_build", C.c_double),
        ("ms_multipoles", C.c_double),
        ("ms_pm_deposit", C.c_double),
        ("ms_pm_fft_poisson", C.c_double),
        ("ms_pm_interpolate", C.c_double),
        ("ms_walk", C.c_double),
        ("wave_node_visits", C.c_uint64),
        ("wave_leaf_source_slots", C.c_uint64),
        ("wave_sum_flushes", C.c_uint64),
        ("pair_set_checksum", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            if name.startswith("reserved") or name == "struct_size":
                continue
            v = getattr(self, name)
            out[name] = list(v) if name == "unwrap_anchor" else v
        return out


ESTIMATOR_RESIDUAL_DERIVATIVE = 0
ESTIMATOR_RELATIVE_MONOPOLE = 1
CUTOFF_CONTAIN_NODE = 0
CUTOFF_COM_DISTANCE = 1
GROUP_WALK_STATUS = {0: "not_requested", 1: "active", 2: "no_local_tree", 3: "incompatible_policy",
                     6: "cutoff_outside_table", 8: "lookup_scaling_unrepresentable"}


class AdaptiveOptions(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("error_estimator", C.c_int32),
        ("maximum_opening_angle", C.c_double),
        ("cutoff_node_policy", C.c_int32),
        ("gaussian_pair_lookup", C.c_int32),
        ("radial_coefficients", C.c_int32),
        ("group_walk", C.c_int32),
        ("group_target_size", C.c_uint32),
        ("reserved0", C.c_uint32),
        ("group_max_extent_cutoff_fraction", C.c_double),
    ]


class WalkCounters(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("group_walk_status", C.c_uint32)] + [(n, C.c_uint64) for n in (
        "visited_nodes", "cutoff_pruned_nodes", "direct_pair_evaluations", "accepted_internal_multipoles",
        "accepted_leaves", "group_walk_groups", "group_walk_grouped_targets", "group_walk_target_group_size_max",
        "group_walk_fallback_targets", "group_walk_fallback_history", "group_walk_fallback_geometry",
        "group_walk_candidate_leaves", "group_walk_culled_leaves", "group_walk_candidate_particles",
        "group_walk_list_entries_max")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "struct_size"}


class Box(C.Structure):
    """chui_treepm_box: one box of a batched ensemble solve (device pointers)."""
    _fields_ = [("pos_x", C.c_void_p), ("pos_y", C.c_void_p), ("pos_z", C.c_void_p), ("mass", C.c_void_p),
                ("n", C.c_uint64), ("accel_x", C.c_void_p), ("accel_y", C.c_void_p), ("accel_z", C.c_void_p),
                ("options", C.POINTER(Options))]


class KdkFactors(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "scale_factor_begin", "scale_factor_midpoint", "scale_factor_end", "drift_factor",
        "first_kick_factor", "second_kick_factor", "first_hubble_drag", "second_hubble_drag")]


class SplitDiagnostics(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "split_scale_comoving", "cutoff_radius_comoving", "short_range_factor_at_split",
        "long_range_factor_at_split", "short_range_factor_at_cutoff", "long_range_factor_at_cutoff",
        "composition_error_at_split", "max_relative_composition_error")]


_P = C.c_void_p
_D = C.c_double
_U64 = C.c_uint64
_I32 = C.c_int32

# name -> argtypes (restype is chui_status = int unless listed in _RESTYPES)
SIGNATURES = {
    "chui_treepm_default_options": [C.POINTER(Options)],
    "chui_treepm_options_set": [C.POINTER(Options), C.c_char_p, C.c_char_p],
    "chui_treepm_options_validate": [C.POINTER(Options), C.POINTER(_I32)],
    "chui_treepm_create": [C.POINTER(_I32), _I32, C.POINTER(_P)],
    "chui_treepm_destroy": [_P],
    "chui_treepm_set_stream": [_P, _P],
    "chui_treepm_synchronize": [_P],
    "chui_treepm_solve": [_P, _P, _P, _P, _P, _U64, _P, _U64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                          C.POINTER(Options), _I32, C.POINTER(Diagnostics)],
    "chui_treepm_solve_device": [_P, _P, _P, _P, _P, _U64, _P, _U64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                 C.POINTER(Options), _I32, C.POINTER(Diagnostics)],
    "chui_treepm_default_adaptive_options": [C.POINTER(AdaptiveOptions)],
    "chui_treepm_adaptive_options_set": [C.POINTER(AdaptiveOptions), C.c_char_p, C.c_char_p],
    "chui_treepm_solve_adaptive_device": [_P, _P, _P, _P, _P, _U64, _P, _U64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                          _P, _P, C.POINTER(Options), C.POINTER(AdaptiveOptions), _I32,
                                          C.POINTER(WalkCounters)],
    "chui_treepm_ensemble_create": [_I32, C.POINTER(_I32), _I32, _I32, C.POINTER(_P)],
    "chui_treepm_ensemble_destroy": [_P],
    "chui_treepm_ensemble_box": [_P, _I32, C.POINTER(_P)],
    "chui_treepm_ensemble_solve_device": [_P, C.POINTER(Box), _I32, _I32],
    "chui_treepm_ensemble_set_capture_safe": [_P, _I32],
    "chui_treepm_ensemble_graph_begin": [_P],
    "chui_treepm_ensemble_graph_end": [_P, C.POINTER(_P)],
    "chui_treepm_set_capture_safe": [_P, _I32],
    "chui_treepm_fault_check": [_P],
    "chui_treepm_set_launch_shape": [_P, _I32, _I32],
    "chui_treepm_graph_begin": [_P],
    "chui_treepm_graph_end": [_P, C.POINTER(_P)],
    "chui_treepm_graph_launch": [_P],
    "chui_treepm_graph_destroy": [_P],
    "chui_treepm_profiling": [_P, _I32],
    "chui_treepm_profile_read": [_P, C.POINTER(_D), C.POINTER(_U64)],
    "chui_treepm_kdk_factors": [_D, _D, _D, _D, _D, C.POINTER(KdkFactors)],
    "chui_treepm_kick_device": [_P, _P, _P, _P, _P, _P, _P, _I32, _P, _U64, _D, _D],
    "chui_treepm_drift_device": [_P, _P, _P, _P, _P, _P, _P, _U64, _D, C.POINTER(_D)],
    "chui_treepm_mass_momentum_device": [_P, _P, _P, _P, _P, _U64, C.POINTER(_D)],
    "chui_treepm_mass_momentum_noalloc_device": [_P, _P, _P, _P, _P, _U64, _P],
    "chui_treepm_force_vjp_device": [_P, _P, _P, _P, _P, _U64, _P, _P, _P, _P, _P, _P, C.POINTER(Options)],
    "chui_treepm_cic_density_device": [_P, _P, _P, _P, _U64, _D, _I32, _P],
    "chui_treepm_cic_density_vjp_device": [_P, _P, _P, _P, _U64, _D, _I32, _P, _P, _P, _P],
    "chui_treepm_unwrap_device": [_P, _P, _U64, _D, _P, C.POINTER(_D)],
    "chui_treepm_morton_device": [_P, _P, _P, _P, _U64, _P, _P],
    "chui_treepm_tree_size": [_P, C.POINTER(_U64), C.POINTER(_U64)],
    "chui_treepm_tree_copy": [_P, _P, _P, _P, _P, _P, _P],
    "chui_treepm_ewald_device": [_P, _P, _P, _P, _P, _U64, _P, _U64, C.POINTER(_D), _D, _D, _I32, _I32, _D,
                                 _P, _P, _P],
    "chui_treepm_split_diagnostics": [_D, _D, _D, C.POINTER(SplitDiagnostics)],
    "chui_treepm_debug_inject_walk_fault": [_P],
    "chui_treepm_last_error": [],
    "chui_treepm_abi_version": [],
}
_RESTYPES = {"chui_treepm_last_error": C.c_char_p, "chui_treepm_abi_version": C.c_int32}

_lib = None


def load() -> C.CDLL:
    """Load libchui_treepm.so once; raises ImportError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("CHUI_TREEPM_LIB", str(LIB_PATH)))
    if not path.exists():
        raise ImportError(f"libchui_treepm.so not built at {path}; run __graft_entry__.build()")
    lib = C.CDLL(str(path))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(status: int) -> None:
    if status == CHUI_OK:
        return
    msg = load().chui_treepm_last_error().decode(errors="replace")
    raise _ERRORS.get(status, ChuiError)(status, msg)


def default_options(**overrides) -> Options:
    opts = Options()
    check(load().chui_treepm_default_options(C.byref(opts)))
    for key, value in overrides.items():
        if key == "box_size":
            if isinstance(value, (int, float)):
                value = (value, value, value)
            opts.box_size = (C.c_double * 3)(*value)
        else:
            setattr(opts, key, value)
    return opts


def default_adaptive_options(**overrides) -> AdaptiveOptions:
    ad = AdaptiveOptions()
    check(load().chui_treepm_default_adaptive_options(C.byref(ad)))
    for key, value in overrides.items():
        setattr(ad, key, value)
    return ad


def set_adaptive_param(ad: AdaptiveOptions, key: str, value) -> None:
    if isinstance(value, bool):
        value = "true" if value else "false"
    check(load().chui_treepm_adaptive_options_set(C.byref(ad), key.encode(), str(value).encode()))


def set_param(opts: Options, key: str, value) -> None:
    if isinstance(value, bool):
        value = "true" if value else "false"
    check(load().chui_treepm_options_set(C.byref(opts), key.encode(), str(value).encode()))
