"""Shared helpers for the parity tests (options, device transfer, error norms)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import cosmosim._lib as L
from cosmosim import ics


def options_struct(box, G=ics.G_CODE, eps=0.0, theta=0.5, asmth=1.25, rcut_cells=4.5 * 1.25, leaf=16,
                   quad=True, assignment=0, deconv=0, criterion=1, precision=0) -> L.Options:
    """Options filled without touching the GPU library (so CPU tests can build them)."""
    o = L.Options()
    o.struct_size = C.sizeof(L.Options)
    if isinstance(box, (int, float)):
        box = (box, box, box)
    o.box_size = (C.c_double * 3)(*box)
    o.scale_factor = 1.0
    o.gravitational_constant = G
    o.assignment_scheme = assignment
    o.window_deconvolution = deconv
    o.asmth_cells = asmth
    o.rcut_cells = rcut_cells
    o.opening_theta = theta
    o.opening_criterion = criterion
    o.multipole_order = 2 if quad else 0
    o.max_leaf_size = leaf
    o.softening_epsilon = eps
    o.relative_force_tolerance = 0.005
    o.relative_force_acceleration_floor = 1e-30
    o.acceptance_policy = 0
    o.residual_block_size = 64
    o.precision = precision
    return o


def adaptive_struct(estimator=L.ESTIMATOR_RELATIVE_MONOPOLE, max_angle=0.25, com_cut=False, lookup=True,
                    radial=False, group=True, group_size=0, extent_fraction=0.5) -> L.AdaptiveOptions:
    """Adaptive/group-walk options (tree_pm_coupling.hpp:248-276) without the GPU library."""
    ad = L.AdaptiveOptions()
    ad.struct_size = C.sizeof(L.AdaptiveOptions)
    ad.error_estimator = estimator
    ad.maximum_opening_angle = max_angle
    ad.cutoff_node_policy = L.CUTOFF_COM_DISTANCE if com_cut else L.CUTOFF_CONTAIN_NODE
    ad.gaussian_pair_lookup = 1 if lookup else 0
    ad.radial_coefficients = 1 if radial else 0
    ad.group_walk = 1 if group else 0
    ad.group_target_size = group_size
    ad.group_max_extent_cutoff_fraction = extent_fraction
    return ad


def seam_clusters(per_cluster=96):
    """Particles of the reference's group-walk culling test
    (tests/integration/test_tree_pm_coupling_periodic.cpp:2148-2176): four
    clusters in a unit box, one straddling a face, one a corner, alternating a
    dense core (spread 0.02) and a halo (0.3) wider than r_cut, from the same
    64-bit LCG. Returns x, y, z, m (unit masses)."""
    centres = ((0.01, 0.5, 0.5), (0.995, 0.99, 0.005), (0.5, 0.5, 0.5), (0.3, 0.7, 0.2))
    n = len(centres) * per_cluster
    state = 0x9E3779B97F4A7C15
    mask = (1 << 64) - 1

    def uniform():
        nonlocal state
        state = (state * 6364136223846793005 + 1442695040888963407) & mask
        return (state >> 11) * (1.0 / 9007199254740992.0)

    def wrap(v):
        v -= math.floor(v)
        return 0.0 if v >= 1.0 else v

    pos = np.zeros((3, n))
    for i in range(n):
        c = centres[i // per_cluster]
        spread = 0.02 if i % 2 == 0 else 0.3
        for a in range(3):
            pos[a, i] = wrap(c[a] + spread * (uniform() - 0.5))
    return pos[0].copy(), pos[1].copy(), pos[2].copy(), np.ones(n)


def seam_cluster_options(max_angle=1.0e-6, group=True, group_size=0, alpha=1.0e-3, **kw):
    """Options of that test: eps 0.002, leaf 8, alpha 1e-3, mesh 32 (h = 1/32),
    r_s = 1.25 h, r_cut = 6.25 h, relative_monopole with the Gaussian lookup."""
    o = options_struct(1.0, G=1.0, eps=0.002, leaf=8, asmth=1.25, rcut_cells=6.25, quad=True)
    o.relative_force_tolerance = alpha
    o.acceptance_policy = 1
    ad = adaptive_struct(max_angle=max_angle, group=group, group_size=group_size, **kw)
    return o, ad


def rel_l2(value, ref):
    v = np.stack(value) if isinstance(value, (list, tuple)) else value
    r = np.stack(ref) if isinstance(ref, (list, tuple)) else ref
    return float(np.sqrt(np.sum((v - r) ** 2) / max(np.sum(r**2), 1e-300)))


def error_metrics(value, ref, floor_fraction=1e-3):
    """test_tree_pm_ewald_accuracy.cpp:722-770: relative L2 and nearest-rank
    percentiles of |da| / max(|a_ref|, floor * rms_ref)."""
    v = np.stack(value)
    r = np.stack(ref)
    err = np.sqrt(np.sum((v - r) ** 2, axis=0))
    rn = np.sqrt(np.sum(r**2, axis=0))
    rms_ref = np.sqrt(np.mean(rn**2))
    norm = err / np.maximum(rn, floor_fraction * rms_ref)

    def pct(p):
        s = np.sort(norm)
        rank = int(np.ceil(p * s.size))
        return float(s[min(max(rank, 1) - 1, s.size - 1)])

    return dict(relative_l2=rel_l2(value, ref), p50=pct(0.5), p90=pct(0.9), p99=pct(0.99), max=float(norm.max()))


def to_dev(*arrays, dtype=None):
    import torch

    out = []
    for a in arrays:
        t = torch.from_numpy(np.ascontiguousarray(a))
        if dtype is not None:
            t = t.to(dtype)
        out.append(t.to("cuda:0"))
    return out


def empty_dev(n, dtype):
    import torch

    return torch.zeros(n, dtype=dtype, device="cuda:0")
