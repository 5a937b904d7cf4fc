"""Shared helpers for the parity tests (options, device transfer, error norms)."""
from __future__ import annotations

import ctypes as C

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
