"""ctypes binding of oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

The CPU FP64 restatement of the reference path (treepm_oracle.cpp). Imported by
tests/, the smoke check and the CPU-baseline leg of bench.py; never by the
product (chui-dev_amd/).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.oracle_last_error.restype = C.c_char_p
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        msg = load().oracle_last_error().decode()
        raise {1: ValueError, 3: OverflowError}.get(rc, RuntimeError)(msg)


def set_threads(n: int) -> int:
    return load().oracle_set_threads(C.c_int(n))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def unwrap_axis(coord, box_length, canonical=None):
    coord = _f64(coord)
    if canonical is None:
        canonical = bool(coord.size == 0 or (coord.min() >= 0.0 and coord.max() < box_length))
    out = np.empty_like(coord)
    anchor = C.c_double()
    _check(load().oracle_unwrap_axis(_p(coord), C.c_uint64(coord.size), C.c_double(box_length),
                                     C.c_int(1 if canonical else 0), _p(out), C.byref(anchor)))
    return out, anchor.value


def morton_keys(x, y, z):
    x, y, z = _f64(x), _f64(y), _f64(z)
    keys = np.empty(x.size, np.uint64)
    _check(load().oracle_morton_keys(_p(x), _p(y), _p(z), C.c_uint64(x.size), _p(keys)))
    return keys


def morton_order(x, y, z):
    x, y, z = _f64(x), _f64(y), _f64(z)
    keys = np.empty(x.size, np.uint64)
    order = np.empty(x.size, np.uint32)
    _check(load().oracle_morton_order(_p(x), _p(y), _p(z), C.c_uint64(x.size), _p(keys), _p(order)))
    return keys, order


def build_tree(x, y, z, m, leaf=16, quad=True, eps=0.0):
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    nodes = C.c_uint64()
    _check(load().oracle_build_tree(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), C.c_int(leaf),
                                    C.c_int(1 if quad else 0), C.c_double(eps), C.byref(nodes)))
    return last_tree(nodes.value, x.size)


def last_tree(node_count, n):
    geometry = np.zeros((node_count, 4))
    moments = np.zeros((node_count, 11))
    begin = np.zeros(node_count, np.uint32)
    count = np.zeros(node_count, np.uint32)
    child = np.zeros((node_count, 8), np.uint32)
    order = np.zeros(n, np.uint32)
    subtree = np.zeros(node_count, np.uint32)
    _check(load().oracle_tree_copy(_p(geometry), _p(moments), _p(begin), _p(count), _p(child), _p(order),
                                   _p(subtree)))
    return dict(geometry=geometry, moments=moments, begin=begin, count=count, child_index=child, order=order,
                subtree=subtree)


def split_factors(r, a, eps):
    r = _f64(r)
    S, L, R = np.empty_like(r), np.empty_like(r), np.empty_like(r)
    _check(load().oracle_split_factors(_p(r), C.c_uint64(r.size), C.c_double(a), _p(S), _p(L), _p(R),
                                       C.c_double(eps)))
    return S, L, R


def _mesh(mesh):
    if isinstance(mesh, int):
        mesh = (mesh, mesh, mesh)
    return (C.c_int32 * 3)(*mesh)


def pm_accel(x, y, z, m, options, mesh, want_density=False):
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    a = [np.empty(x.size) for _ in range(3)]
    msh = _mesh(mesh)
    dens = np.empty(msh[0] * msh[1] * msh[2]) if want_density else None
    _check(load().oracle_pm_accel(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), C.byref(options), msh,
                                  _p(a[0]), _p(a[1]), _p(a[2]), _p(dens)))
    return (a, dens) if want_density else a


def short_range(x, y, z, m, options, mesh):
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    a = [np.empty(x.size) for _ in range(3)]
    counters = np.zeros(6, np.uint64)
    anchors = np.zeros(3)
    _check(load().oracle_short_range(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), C.byref(options),
                                     _mesh(mesh), _p(a[0]), _p(a[1]), _p(a[2]), _p(counters), _p(anchors)))
    names = ("visited", "pruned", "pairs", "multipoles", "leaves", "skipped_pairs")
    return a, dict(zip(names, counters.tolist())), anchors


def solve(x, y, z, m, options, mesh):
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    a = [np.empty(x.size) for _ in range(3)]
    counters = np.zeros(6, np.uint64)
    _check(load().oracle_solve(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), C.byref(options), _mesh(mesh),
                               _p(a[0]), _p(a[1]), _p(a[2]), _p(counters)))
    return a


def ewald(x, y, z, m, targets, box, G, alpha, real_images, modes, eps=0.0):
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    targets = np.ascontiguousarray(targets, dtype=np.uint32)
    a = [np.empty(targets.size) for _ in range(3)]
    L = (C.c_double * 3)(*box)
    _check(load().oracle_ewald(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), _p(targets),
                               C.c_uint64(targets.size), L, C.c_double(G), C.c_double(alpha), C.c_int(real_images),
                               C.c_int(modes), C.c_double(eps), _p(a[0]), _p(a[1]), _p(a[2])))
    return a


def kick(vx, vy, vz, ax, ay, az, kick_factor, drag):
    load().oracle_kick(_p(vx), _p(vy), _p(vz), _p(_f64(ax)), _p(_f64(ay)), _p(_f64(az)), C.c_uint64(vx.size),
                       C.c_double(kick_factor), C.c_double(drag))


def drift(x, y, z, vx, vy, vz, drift_factor, box):
    L = (C.c_double * 3)(*box)
    load().oracle_drift(_p(x), _p(y), _p(z), _p(vx), _p(vy), _p(vz), C.c_uint64(x.size),
                        C.c_double(drift_factor), L)


def pm_mesh(density, options, mesh, split_scale):
    """Mesh-level solve of a given density (shape mesh, row-major) -> (fx, fy, fz) meshes."""
    density = np.ascontiguousarray(density, dtype=np.float64)
    msh = _mesh(mesh)
    out = [np.empty(density.shape) for _ in range(3)]
    _check(load().oracle_pm_mesh(_p(density), C.byref(options), msh, C.c_double(split_scale),
                                 _p(out[0]), _p(out[1]), _p(out[2])))
    return out


def pm_interpolate(fields, x, y, z, options, mesh):
    x, y, z = _f64(x), _f64(y), _f64(z)
    f = [np.ascontiguousarray(a, dtype=np.float64) for a in fields]
    a = [np.empty(x.size) for _ in range(3)]
    _check(load().oracle_pm_interpolate(_p(f[0]), _p(f[1]), _p(f[2]), _mesh(mesh), _p(x), _p(y), _p(z),
                                        C.c_uint64(x.size), C.byref(options), _p(a[0]), _p(a[1]), _p(a[2])))
    return a


def long_range_inv_r3(r, a):
    r = _f64(r)
    out = np.empty_like(r)
    _check(load().oracle_long_range_inv_r3(_p(r), C.c_uint64(r.size), C.c_double(a), _p(out)))
    return out


def residual_inv_r3(r2, eps, a):
    r2 = _f64(r2)
    out = np.empty_like(r2)
    _check(load().oracle_residual_inv_r3(_p(r2), C.c_uint64(r2.size), C.c_double(eps), C.c_double(a), _p(out)))
    return out


def cluster_node_accel(x, y, z, m, target, G, eps, a, quad=True):
    """Residual acceleration of a source cluster evaluated as one accepted node."""
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    t = (C.c_double * 3)(*target)
    out = (C.c_double * 3)()
    _check(load().oracle_cluster_node_accel(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), t, C.c_double(G),
                                            C.c_double(eps), C.c_double(a), C.c_int(1 if quad else 0), out))
    return np.array(out[:])


_WALK_COUNTER_NAMES = ("visited_nodes", "cutoff_pruned_nodes", "direct_pair_evaluations",
                       "accepted_internal_multipoles", "accepted_leaves", "group_walk_groups",
                       "group_walk_grouped_targets", "group_walk_target_group_size_max",
                       "group_walk_fallback_targets", "group_walk_fallback_history", "group_walk_fallback_geometry",
                       "group_walk_candidate_leaves", "group_walk_culled_leaves", "group_walk_candidate_particles",
                       "group_walk_list_entries_max")


def short_range_adaptive(x, y, z, m, options, adaptive, mesh, active=None, previous=None):
    """Adaptive-policy residual per active slot -> ([ax, ay, az], counters dict).

    options / adaptive are ctypes structures with the chui_treepm_options and
    chui_treepm_adaptive_options layouts; previous = |a_prev| per active slot."""
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    act = None if active is None else np.ascontiguousarray(active, dtype=np.uint32)
    na = x.size if act is None else act.size
    prev = None if previous is None else _f64(previous)
    if prev is not None and prev.size != na:
        raise ValueError("previous acceleration lane must have one entry per active slot")
    a = [np.empty(na) for _ in range(3)]
    raw = np.zeros(16, np.uint64)
    _check(load().oracle_short_range_adaptive(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), _p(act),
                                              C.c_uint64(na), _p(prev), C.byref(options), C.byref(adaptive),
                                              _mesh(mesh), _p(a[0]), _p(a[1]), _p(a[2]), _p(raw)))
    counters = dict(zip(_WALK_COUNTER_NAMES, raw[1:].tolist()))
    counters["group_walk_status"] = int(raw[0] >> np.uint64(32))
    return a, counters


def gauss_table(t, want_knots=False):
    """(H(t) as dimensionlessCoefficient, clamped Hermite of the group kernel[, knots (2, 4097)])."""
    t = _f64(t)
    coef, herm = np.empty_like(t), np.empty_like(t)
    knots = np.empty((2, 4097)) if want_knots else None
    _check(load().oracle_gauss_table(_p(t), C.c_uint64(t.size), _p(coef), _p(herm), _p(knots)))
    return (coef, herm, knots) if want_knots else (coef, herm)


def residual_radial(x, eps, a):
    """(value, f'(r)/r, (f'' - f'/r)/r^2) of the residual kernel at x = r^2, shape (n, 3)."""
    x = _f64(x)
    out = np.empty((x.size, 3))
    _check(load().oracle_residual_radial(_p(x), C.c_uint64(x.size), C.c_double(eps), C.c_double(a), _p(out)))
    return out


def cluster_node_accel_radial(x, y, z, m, target, G, eps, a, quad=True, radial=True):
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    t = (C.c_double * 3)(*target)
    out = (C.c_double * 3)()
    _check(load().oracle_cluster_node_accel_radial(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), t,
                                                   C.c_double(G), C.c_double(eps), C.c_double(a),
                                                   C.c_int(1 if quad else 0), C.c_int(1 if radial else 0), out))
    return np.array(out[:])


def group_list(x, y, z, m, rows, previous, options, adaptive, mesh, cap=1 << 16):
    """Interaction list of one target group -> (accepted nodes, listed leaves, (centre, half)).

    The tree it was built on is available from group_tree()."""
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    prev = _f64(previous)
    nodes = np.zeros(cap, np.uint32)
    leaves = np.zeros(cap, np.uint32)
    sizes = np.zeros(2, np.uint64)
    box = np.zeros(6)
    _check(load().oracle_group_list(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), _p(rows), C.c_uint64(rows.size),
                                    _p(prev), C.byref(options), C.byref(adaptive), _mesh(mesh), C.c_uint64(cap),
                                    _p(nodes), _p(leaves), _p(sizes), _p(box)))
    nn, nl = int(sizes[0]), int(sizes[1])
    if max(nn, nl) > cap:
        raise ValueError("interaction list exceeds cap")
    return nodes[:nn].copy(), leaves[:nl].copy(), (box[:3].copy(), box[3:].copy())


def group_tree(n):
    """Tree built by the last group_list call (node_count from its geometry)."""
    count = C.c_uint64()
    _check(load().oracle_last_tree_size(C.byref(count)))
    return last_tree(count.value, n)


def probe_walk(x, y, z, m, options, mesh, target, cap=1 << 16):
    """One strict-policy walk for the point target (not a source) over the
    tree of the given sources: residual acceleration and the visit-order
    lists of accepted internal nodes, opened internal nodes, summed leaves and
    pruned nodes (preorder node numbers of last_tree())."""
    x, y, z, m = (_f64(a) for a in (x, y, z, m))
    acc = np.zeros(3)
    lists = [np.zeros(cap, np.uint32) for _ in range(4)]
    sizes = np.zeros(4, np.uint64)
    _check(load().oracle_probe_walk(_p(x), _p(y), _p(z), _p(m), C.c_uint64(x.size), C.byref(options), _mesh(mesh),
                                    C.c_double(target[0]), C.c_double(target[1]), C.c_double(target[2]), _p(acc),
                                    _p(lists[0]), _p(lists[1]), _p(lists[2]), _p(lists[3]), C.c_uint64(cap),
                                    _p(sizes)))
    if int(sizes.max()) > cap:
        raise RuntimeError("probe walk lists exceed their capacity")
    names = ("accepted", "opened", "leaves", "pruned")
    return acc, {k: lists[i][:int(sizes[i])].copy() for i, k in enumerate(names)}


def residual_proxy(h, offset, mass, r2, eps, a, G):
    """residualAccuracyProxy (tree_pm_acceptance.hpp:14-51) elementwise."""
    h, offset, mass, r2 = np.broadcast_arrays(*(np.asarray(v, np.float64) for v in (h, offset, mass, r2)))
    h, offset, mass, r2 = (np.ascontiguousarray(v) for v in (h, offset, mass, r2))
    out = np.zeros(h.size)
    _check(load().oracle_residual_proxy(_p(h), _p(offset), _p(mass), _p(r2), C.c_uint64(h.size), C.c_double(eps),
                                        C.c_double(a), C.c_double(G), _p(out)))
    return out.reshape(h.shape)
