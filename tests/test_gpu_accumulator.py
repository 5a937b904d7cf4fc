"""GPU checks of the accumulator lanes (TreePmForceAccumulatorView,
tree_pm_coupling.hpp:127-165 and tree_pm_coupling.cpp:1957-1976).

The short-range lanes and the nearest-pair (encounter) lane are indexed by
source row: active rows are written, the others left untouched. The
nearest-pair lane holds r^2 of the nearest distinct source inside r_cut, over
minimum images, or +infinity when none. It is checked against a brute-force
search: exactly for the FP64 per-target walk and the group walk, and to FP32
rounding for mixed precision. Coincident distinct sources give 0.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import empty_dev, options_struct, seam_cluster_options, to_dev

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import ForceAccumulator, TreePmCoordinator  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from test_oracle_pins import clustered  # noqa: E402


def brute_nearest(x, y, z, box, rcut, rows):
    pos = np.stack([x, y, z])
    out = np.full(rows.size, np.inf)
    for k, i in enumerate(rows):
        d = pos - pos[:, i:i + 1]
        d -= box * np.round(d / box)
        r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        r2[i] = np.inf
        r2[r2 > rcut * rcut] = np.inf
        out[k] = r2.min()
    return out


def solve_nearest(x, y, z, m, o, mesh, active=None, adaptive=None, previous=None, precision=0):
    dt = torch.float64 if precision == 0 else torch.float32
    o.precision = precision
    dx, dy, dz, dm = to_dev(x, y, z, m)
    na = x.size if active is None else len(active)
    acc = tuple(empty_dev(na, dt) for _ in range(3))
    sr = tuple(empty_dev(x.size, dt) for _ in range(3))
    near = torch.full((x.size,), -1.0, dtype=dt, device="cuda:0")  # marks untouched rows
    act = to_dev(np.asarray(active, np.uint32).view(np.int32))[0] if active is not None else None
    prev = to_dev(np.asarray(previous, np.float64))[0] if previous is not None else None
    torch.cuda.synchronize()
    with TreePmCoordinator(mesh) as coord:
        coord.solve_active_set(dx, dy, dz, dm, ForceAccumulator(acc, act, None, sr, prev, near), o,
                               adaptive_options=adaptive)
        coord.synchronize()
    return near.double().cpu().numpy(), [s.double().cpu().numpy() for s in sr]


@pytest.mark.parametrize("precision,rtol", [(0, 0.0), (1, 2e-6)])
def test_nearest_pair_lane_strict_walk(precision, rtol):
    box = 10.0
    x, y, z, m = clustered(1500, 4, box)
    x[7], y[7], z[7] = x[3], y[3], z[3]  # a coincident distinct pair
    o = options_struct(box, eps=0.01, leaf=8, quad=False)  # monopoles: every source inside r_cut is direct-summed
    near, _ = solve_nearest(x, y, z, m, o, 16, precision=precision)
    rcut = 4.5 * 1.25 * box / 16
    ref = brute_nearest(x, y, z, box, rcut, np.arange(x.size))
    assert near[3] == 0.0 and near[7] == 0.0
    if precision == 0:
        assert np.array_equal(near, ref)
    else:  # away from the cutoff, where FP32 rounding cannot move a pair across it
        sel = ref < 0.99 * rcut * rcut
        assert np.all(np.abs(near[sel] - ref[sel]) <= rtol * ref[sel] + 1e-12)


def test_nearest_pair_lane_isolated_target_is_infinite():
    x = np.array([1.0, 1.2, 6.0])
    y = np.array([1.0, 1.0, 6.0])
    z = np.array([1.0, 1.0, 6.0])
    o = options_struct(10.0, eps=0.01, quad=False)
    near, _ = solve_nearest(x, y, z, np.ones(3), o, 16)
    assert np.isinf(near[2]) and abs(near[0] - 0.04) < 1e-15 and abs(near[1] - 0.04) < 1e-15


def test_active_rows_only_are_written():
    box = 10.0
    x, y, z, m = clustered(900, 6, box)
    active = np.random.default_rng(2).choice(x.size, 300, replace=False).astype(np.uint32)
    o = options_struct(box, eps=0.01, quad=False)
    near, sr = solve_nearest(x, y, z, m, o, 16, active=active)
    inactive = np.setdiff1d(np.arange(x.size), active)
    assert np.all(near[inactive] == -1.0) and np.all(sr[0][inactive] == 0.0)
    rcut = 4.5 * 1.25 * box / 16
    ref = brute_nearest(x, y, z, box, rcut, active)
    assert np.array_equal(near[active], ref)


def test_nearest_pair_lane_group_walk():
    x, y, z, m = clustered(2000, 8, 1.0)
    o, ad = seam_cluster_options(max_angle=1.0e-6, group=True)  # no accepted node: all pairs direct
    prev = np.full(x.size, 5000.0)
    near, _ = solve_nearest(x, y, z, m, o, 32, adaptive=ad, previous=prev)
    ref = brute_nearest(x, y, z, 1.0, 6.25 / 32, np.arange(x.size))
    assert np.array_equal(near, ref)
