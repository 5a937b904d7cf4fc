"""Python mirror of gravity::TreePmCoordinator over the C ABI.

TreePmCoordinator(pm_shape) and solve_active_set_with_pm_cadence(...) follow
include/cosmosim/gravity/tree_pm_coupling.hpp:545-584: positions/masses are
caller-owned, the coordinator owns mesh, tree and scratch across calls. Device
arrays (torch tensors on a HIP device) go through chui_treepm_solve_device with
no host copies; numpy arrays go through the host-pointer drop-in
chui_treepm_solve.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(a.data_ptr())  # torch tensor


@dataclass
class ForceAccumulator:
    """TreePmForceAccumulatorView (tree_pm_coupling.hpp:127-165): active rows,
    total accelerations and the optional PM-only / tree-only capture lanes."""

    accel: tuple
    active: object = None
    long_range: tuple | None = None
    short_range: tuple | None = None
    # previous_acceleration_magnitude_code: |a_prev| per active slot (adaptive policy)
    previous_acceleration: object = None
    # short_range_nearest_pair_r2_comoving: source-row indexed encounter lane
    nearest_pair_r2: object = None


@dataclass
class SolveResult:
    diagnostics: dict = field(default_factory=dict)


class Graph:
    """A recorded HIP graph of capture-safe solves (chui_treepm_graph)."""

    def __init__(self, handle: C.c_void_p):
        self._lib = _lib.load()
        self._h = handle

    def launch(self):
        """Replay on the recording stream after the members' earlier work; the
        members' later work (and synchronize) waits for it."""
        _lib.check(self._lib.chui_treepm_graph_launch(self._h))

    def close(self):
        if getattr(self, "_h", None):
            _lib.check(self._lib.chui_treepm_graph_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TreePmCoordinator:
    def __init__(self, pm_shape, device: int = 0):
        lib = _lib.load()
        if isinstance(pm_shape, int):
            pm_shape = (pm_shape, pm_shape, pm_shape)
        self.pm_shape = tuple(int(v) for v in pm_shape)
        mesh = (C.c_int32 * 3)(*self.pm_shape)
        handle = C.c_void_p()
        _lib.check(lib.chui_treepm_create(mesh, int(device), C.byref(handle)))
        self._h = handle
        self._lib = lib
        self._owned = True

    @classmethod
    def borrowed(cls, handle: C.c_void_p, pm_shape) -> "TreePmCoordinator":
        """A view of a context another owner (an ensemble) creates and destroys."""
        self = cls.__new__(cls)
        self.pm_shape = tuple(int(v) for v in pm_shape)
        self._h = handle
        self._lib = _lib.load()
        self._owned = False
        return self

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                _lib.check(self._lib.chui_treepm_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        _lib.check(self._lib.chui_treepm_synchronize(self._h))

    def set_stream(self, stream_handle: int | None):
        _lib.check(self._lib.chui_treepm_set_stream(self._h, C.c_void_p(stream_handle) if stream_handle else None))

    def set_launch_shape(self, fft_lines_per_block: int = 0, walk_waves_per_block: int = 0):
        """Lines per block of the PM line passes and waves per block of the
        timed mixed walk (0 = built-in; results do not depend on them)."""
        _lib.check(self._lib.chui_treepm_set_launch_shape(self._h, int(fft_lines_per_block),
                                                          int(walk_waves_per_block)))

    # -- capture-safe solves and HIP graphs (no reference counterpart) ---------
    def set_capture_safe(self, enable: bool = True):
        """Capture-safe mode: after one solve of a source count, solves of that
        count read nothing back and allocate nothing (chui_treepm_set_capture_safe)."""
        _lib.check(self._lib.chui_treepm_set_capture_safe(self._h, 1 if enable else 0))

    def fault_check(self):
        """Faults of solves replayed from a graph the caller recorded on this
        context's stream (chui_treepm_fault_check)."""
        _lib.check(self._lib.chui_treepm_fault_check(self._h))

    def graph_begin(self):
        """Start recording this context's stream into a HIP graph."""
        _lib.check(self._lib.chui_treepm_graph_begin(self._h))

    def graph_end(self) -> "Graph":
        out = C.c_void_p()
        _lib.check(self._lib.chui_treepm_graph_end(self._h, C.byref(out)))
        return Graph(out)

    def solve_active_set_with_pm_cadence(self, pos_x, pos_y, pos_z, mass, accumulator: ForceAccumulator,
                                         options: _lib.Options, refresh_long_range_field: bool = True,
                                         want_diagnostics: bool = False,
                                         adaptive_options: _lib.AdaptiveOptions | None = None) -> SolveResult:
        n = int(pos_x.shape[0])
        active = accumulator.active
        n_active = int(active.shape[0]) if active is not None else n
        lr = accumulator.long_range or (None, None, None)
        sr = accumulator.short_range or (None, None, None)
        diag = None
        if want_diagnostics:
            diag = _lib.Diagnostics()
            diag.struct_size = C.sizeof(_lib.Diagnostics)
        on_device = not isinstance(pos_x, np.ndarray)
        if adaptive_options is None and accumulator.nearest_pair_r2 is not None:
            adaptive_options = _lib.default_adaptive_options()
        if adaptive_options is not None:
            # adaptive acceptance / group walk (chui_treepm_solve_adaptive_device)
            if not on_device:
                raise ValueError("the adaptive entry point takes device arrays")
            counters = _lib.WalkCounters()
            counters.struct_size = C.sizeof(_lib.WalkCounters)
            _lib.check(self._lib.chui_treepm_solve_adaptive_device(
                self._h, _ptr(pos_x), _ptr(pos_y), _ptr(pos_z), _ptr(mass), n, _ptr(active), n_active,
                _ptr(accumulator.accel[0]), _ptr(accumulator.accel[1]), _ptr(accumulator.accel[2]),
                _ptr(lr[0]), _ptr(lr[1]), _ptr(lr[2]), _ptr(sr[0]), _ptr(sr[1]), _ptr(sr[2]),
                _ptr(accumulator.nearest_pair_r2), _ptr(accumulator.previous_acceleration), C.byref(options),
                C.byref(adaptive_options),
                1 if refresh_long_range_field else 0, C.byref(counters) if want_diagnostics else None))
            return SolveResult(diagnostics=counters.as_dict() if want_diagnostics else {})
        fn = self._lib.chui_treepm_solve_device if on_device else self._lib.chui_treepm_solve
        _lib.check(fn(self._h, _ptr(pos_x), _ptr(pos_y), _ptr(pos_z), _ptr(mass), n, _ptr(active), n_active,
                      _ptr(accumulator.accel[0]), _ptr(accumulator.accel[1]), _ptr(accumulator.accel[2]),
                      _ptr(lr[0]), _ptr(lr[1]), _ptr(lr[2]), _ptr(sr[0]), _ptr(sr[1]), _ptr(sr[2]),
                      C.byref(options), 1 if refresh_long_range_field else 0,
                      C.byref(diag) if diag is not None else None))
        return SolveResult(diagnostics=diag.as_dict() if diag is not None else {})

    def solve_active_set(self, pos_x, pos_y, pos_z, mass, accumulator, options, want_diagnostics=False,
                         adaptive_options=None):
        return self.solve_active_set_with_pm_cadence(pos_x, pos_y, pos_z, mass, accumulator, options, True,
                                                     want_diagnostics, adaptive_options)

    def profiling(self, enable: bool):
        _lib.check(self._lib.chui_treepm_profiling(self._h, 1 if enable else 0))

    def profile_read(self) -> dict:
        ms = (C.c_double * 8)()
        solves = C.c_uint64()
        _lib.check(self._lib.chui_treepm_profile_read(self._h, ms, C.byref(solves)))
        names = ("total", "tree_build", "gather", "pm", "targets", "walk")
        return {"solves": solves.value, **{f"{k}_ms": ms[i] for i, k in enumerate(names)}}

    # ---- inspection / parity entry points --------------------------------
    def tree_arrays(self):
        nodes = C.c_uint64()
        parts = C.c_uint64()
        _lib.check(self._lib.chui_treepm_tree_size(self._h, C.byref(nodes), C.byref(parts)))
        nn, npart = nodes.value, parts.value
        geometry = np.zeros((nn, 4))
        moments = np.zeros((nn, 11))
        begin = np.zeros(nn, np.uint32)
        count = np.zeros(nn, np.uint32)
        child = np.zeros((nn, 8), np.uint32)
        order = np.zeros(npart, np.uint32)
        _lib.check(self._lib.chui_treepm_tree_copy(self._h, _ptr(geometry), _ptr(moments), _ptr(begin),
                                                   _ptr(count), _ptr(child), _ptr(order)))
        return dict(geometry=geometry, moments=moments, begin=begin, count=count, child_index=child,
                    order=order)

    def morton(self, x, y, z, keys, order):
        _lib.check(self._lib.chui_treepm_morton_device(self._h, _ptr(x), _ptr(y), _ptr(z), int(x.shape[0]),
                                                       _ptr(keys), _ptr(order)))

    def unwrap(self, coord, box_length, out) -> float:
        anchor = C.c_double()
        _lib.check(self._lib.chui_treepm_unwrap_device(self._h, _ptr(coord), int(coord.shape[0]),
                                                       float(box_length), _ptr(out), C.byref(anchor)))
        return anchor.value

    def ewald(self, x, y, z, m, targets, box, G, alpha, real_images, modes, eps, ax, ay, az):
        L = (C.c_double * 3)(*box)
        _lib.check(self._lib.chui_treepm_ewald_device(self._h, _ptr(x), _ptr(y), _ptr(z), _ptr(m),
                                                      int(x.shape[0]), _ptr(targets), int(targets.shape[0]), L,
                                                      float(G), float(alpha), int(real_images), int(modes),
                                                      float(eps), _ptr(ax), _ptr(ay), _ptr(az)))

    def kick(self, vx, vy, vz, ax, ay, az, kick, drag, active=None):
        is_float = 1 if str(ax.dtype).endswith("float32") else 0
        n = int(active.shape[0]) if active is not None else int(vx.shape[0])
        _lib.check(self._lib.chui_treepm_kick_device(self._h, _ptr(vx), _ptr(vy), _ptr(vz), _ptr(ax), _ptr(ay),
                                                     _ptr(az), is_float, _ptr(active), n, float(kick),
                                                     float(drag)))

    def drift(self, x, y, z, vx, vy, vz, drift, box):
        L = (C.c_double * 3)(*box)
        _lib.check(self._lib.chui_treepm_drift_device(self._h, _ptr(x), _ptr(y), _ptr(z), _ptr(vx), _ptr(vy),
                                                      _ptr(vz), int(x.shape[0]), float(drift), L))

    def force_vjp(self, x, y, z, m, lam, out, options):
        """out_k = sum_i (d a_i / d x_k)^T lam_i for every source (FP64 device
        arrays in caller order; options must be FP64 with the strict acceptance
        policy, multipole_order 0 or 2)."""
        _lib.check(self._lib.chui_treepm_force_vjp_device(
            self._h, _ptr(x), _ptr(y), _ptr(z), _ptr(m), int(x.shape[0]), _ptr(lam[0]), _ptr(lam[1]),
            _ptr(lam[2]), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), C.byref(options)))

    def cic_density(self, x, y, z, box, mesh, delta):
        """delta = CIC count * mesh^3 / n - 1 on a mesh^3 grid (row-major)."""
        _lib.check(self._lib.chui_treepm_cic_density_device(self._h, _ptr(x), _ptr(y), _ptr(z), int(x.shape[0]),
                                                            float(box), int(mesh), _ptr(delta)))

    def cic_density_vjp(self, x, y, z, box, mesh, grad_delta, out):
        _lib.check(self._lib.chui_treepm_cic_density_vjp_device(
            self._h, _ptr(x), _ptr(y), _ptr(z), int(x.shape[0]), float(box), int(mesh), _ptr(grad_delta),
            _ptr(out[0]), _ptr(out[1]), _ptr(out[2])))

    def mass_momentum(self, m, vx, vy, vz):
        out = (C.c_double * 4)()
        _lib.check(self._lib.chui_treepm_mass_momentum_device(self._h, _ptr(m), _ptr(vx), _ptr(vy), _ptr(vz),
                                                              int(m.shape[0]), out))
        return list(out)


def kdk_factors(a0: float, a1: float, hubble0: float, omega_m: float, omega_lambda: float) -> dict:
    out = _lib.KdkFactors()
    _lib.check(_lib.load().chui_treepm_kdk_factors(a0, a1, hubble0, omega_m, omega_lambda, C.byref(out)))
    return {name: getattr(out, name) for name, _ in out._fields_}


def split_diagnostics(asmth_cells: float, rcut_cells: float, mesh_spacing: float) -> dict:
    out = _lib.SplitDiagnostics()
    _lib.check(_lib.load().chui_treepm_split_diagnostics(asmth_cells, rcut_cells, mesh_spacing, C.byref(out)))
    return {name: getattr(out, name) for name, _ in out._fields_}
