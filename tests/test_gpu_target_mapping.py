"""GPU parity of the two target mappings of the strict walks.

CHUI_TARGETS_PER_WAVE (one target per wave, walkmix_target_kernel) must give
the pairs and multipoles of CHUI_TARGETS_PER_LANE (64 targets per wave,
walkmix_kernel) for every target: the two differ only in summation order, so
they agree to FP32 rounding of the sums (relative L2 below 1e-5), and the
per-wave lanes keep the mixed error budget (relative L2 below 1e-4, p99
below 1e-3) against the oracle's FP64 short-range residual
(tree_pm_coupling.cpp:4026-4272). Cases: uniform and clustered sources,
quadrupole acceptance (accepted multipoles), both opening criteria, a
rectangular box, per-particle masses, coincident sources, leaves of up to
200 sources (several 64-source chunks, full source lists), sparse and dense active
sets, the PM interpolation fused in (TSC) and short-range-only solves. The
FP64 strict walk's per-wave kernel (walk64_target_kernel) matches its per-lane
walk and the oracle to 1e-12.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import empty_dev, error_metrics, options_struct, rel_l2, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import ForceAccumulator, TreePmCoordinator  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from test_oracle_pins import clustered  # noqa: E402


def solve_mixed(x, y, z, m, o, mesh, mapping, active=None, sr_only=False, precision=L.PRECISION_MIXED):
    """Solve -> (totals per active slot, short-range lanes per source row)."""
    o.precision = precision
    o.target_mapping = mapping
    o.short_range_only = 1 if sr_only else 0
    dx, dy, dz, dm = to_dev(x, y, z, m)
    na = x.size if active is None else len(active)
    dt = torch.float32 if precision == L.PRECISION_MIXED else torch.float64
    acc = tuple(empty_dev(na, dt) for _ in range(3))
    sr = tuple(torch.zeros(x.size, dtype=dt, device="cuda:0") for _ in range(3))
    act = to_dev(np.asarray(active, np.uint32).view(np.int32))[0] if active is not None else None
    torch.cuda.synchronize()
    with TreePmCoordinator(mesh) as coord:
        coord.solve_active_set_with_pm_cadence(dx, dy, dz, dm, ForceAccumulator(acc, act, None, sr), o,
                                               not sr_only)
        coord.synchronize()
    o.short_range_only = 0
    o.precision = L.PRECISION_MIXED
    return np.stack([a.double().cpu().numpy() for a in acc]), np.stack([s.double().cpu().numpy() for s in sr])


def cases():
    rng = np.random.default_rng(11)
    u = rng.uniform(0.0, 20.0, (3, 6000))
    yield "uniform", (u[0], u[1], u[2], np.ones(6000)), options_struct(20.0, eps=0.01), 32
    x, y, z, m = clustered(8000, 8, 20.0)
    yield "clustered_quad", (x, y, z, m), options_struct(20.0, eps=0.005, theta=0.7), 32
    yield "clustered_geometric", (x, y, z, m), options_struct(20.0, eps=0.005, criterion=0), 32
    box = (16.0, 24.0, 12.0)
    r = rng.uniform(0.0, 1.0, (3, 5000)) * np.array(box)[:, None]
    yield "rectangular_masses_tsc", (r[0], r[1], r[2], rng.uniform(0.5, 2.0, 5000)), \
        options_struct(box, eps=0.01, assignment=1), (32, 48, 24)
    d = rng.uniform(0.0, 10.0, (3, 6000))
    d[:, 100:110] = d[:, 100:101]  # 10 coincident sources
    # leaves of up to 200 sources: several 64-source chunks per leaf and more
    # listed sources than one evaluation holds
    yield "large_leaves", (d[0], d[1], d[2], np.ones(6000)), options_struct(10.0, eps=0.02, leaf=200), 16


CASES = list(cases())


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_box_walk_matches_counted_lane_walk(case):
    # The timed mixed walk classifies nodes against each wave's target box and
    # walks lane by lane only the subtrees some lane may accept; the counting
    # (diagnostic) walk goes lane by lane everywhere. Same pairs and
    # multipoles, other summation order: FP32 rounding of the sums apart, the
    # short-range lanes agree; a missed or doubled multipole or leaf would
    # show at the 1e-3 level.
    _, (x, y, z, m), o, mesh = case
    o.precision = L.PRECISION_MIXED
    o.target_mapping = L.TARGETS_PER_LANE
    o.short_range_only = 1
    dx, dy, dz, dm = to_dev(x, y, z, m)
    lanes = []
    for diagnostics in (True, False):
        acc = tuple(empty_dev(x.size, torch.float32) for _ in range(3))
        sr = tuple(torch.zeros(x.size, dtype=torch.float32, device="cuda:0") for _ in range(3))
        torch.cuda.synchronize()
        with TreePmCoordinator(mesh) as coord:
            res = coord.solve_active_set_with_pm_cadence(dx, dy, dz, dm, ForceAccumulator(acc, None, None, sr), o,
                                                         False, want_diagnostics=diagnostics)
            coord.synchronize()
        if diagnostics:
            print(case[0], "multipoles", res.diagnostics["accepted_internal_multipoles"],
                  "pairs", res.diagnostics["direct_pair_evaluations"])
        lanes.append(np.stack([s_.double().cpu().numpy() for s_ in sr]))
    o.short_range_only = 0
    o.target_mapping = L.TARGETS_AUTO
    counted, boxed = lanes
    assert rel_l2(boxed, counted) < 2e-6
    per = np.linalg.norm(boxed - counted, axis=0) / np.maximum(np.linalg.norm(counted, axis=0), 1e-30)
    assert per.max() < 1e-4


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("fraction", [0.02, 0.3, 1.0])
def test_per_wave_matches_per_lane(case, fraction):
    _, (x, y, z, m), o, mesh = case
    n = x.size
    active = None
    if fraction < 1.0:
        active = np.sort(np.random.default_rng(5).choice(n, int(fraction * n), replace=False)).astype(np.uint32)
    wave_acc, wave_sr = solve_mixed(x, y, z, m, o, mesh, L.TARGETS_PER_WAVE, active)
    lane_acc, lane_sr = solve_mixed(x, y, z, m, o, mesh, L.TARGETS_PER_LANE, active)
    rows = np.arange(n) if active is None else active.astype(np.int64)
    assert rel_l2(wave_acc, lane_acc) < 1e-5
    assert rel_l2(wave_sr[:, rows], lane_sr[:, rows]) < 1e-5
    inactive = np.setdiff1d(np.arange(n), rows)
    assert np.all(wave_sr[:, inactive] == 0.0)
    ref, _, _ = O.short_range(x, y, z, m, o, mesh)
    e = error_metrics(wave_sr[:, rows], np.stack(ref)[:, rows])
    assert e["relative_l2"] < 1e-4 and e["p99"] < 1e-3


@pytest.mark.parametrize("case", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_auto_mapping_short_range_only(case):
    """Auto picks one target per wave when at most 1/64 of the sources are
    active (bit-identical to the forced per-wave solve) and 64 targets per
    wave otherwise."""
    _, (x, y, z, m), o, mesh = case
    sparse = np.arange(0, x.size, 100, dtype=np.uint32)
    dense = np.arange(0, x.size, 3, dtype=np.uint32)
    for active, mapping in ((sparse, L.TARGETS_PER_WAVE), (dense, L.TARGETS_PER_LANE)):
        auto_acc, auto_sr = solve_mixed(x, y, z, m, o, mesh, L.TARGETS_AUTO, active, sr_only=True)
        ref_acc, _ = solve_mixed(x, y, z, m, o, mesh, mapping, active, sr_only=True)
        assert np.array_equal(auto_acc, ref_acc)
        # short-range only: the totals are the short-range lanes of the active rows
        assert np.array_equal(auto_acc, auto_sr[:, active.astype(np.int64)])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("fraction", [0.02, 1.0])
def test_fp64_per_wave_matches_per_lane_and_oracle(case, fraction):
    """The FP64 strict walk with one target per wave (walk64_target_kernel):
    the same pairs and multipoles as the per-lane walk (the box-classified
    walk with the polynomial H table when no counters are asked for), summed
    in another order, so both agree with each other and with the oracle to
    1e-12."""
    _, (x, y, z, m), o, mesh = case
    n = x.size
    active = None
    if fraction < 1.0:
        active = np.sort(np.random.default_rng(6).choice(n, int(fraction * n), replace=False)).astype(np.uint32)
    fp64 = L.PRECISION_FP64
    wave_acc, wave_sr = solve_mixed(x, y, z, m, o, mesh, L.TARGETS_PER_WAVE, active, precision=fp64)
    lane_acc, lane_sr = solve_mixed(x, y, z, m, o, mesh, L.TARGETS_PER_LANE, active, precision=fp64)
    rows = np.arange(n) if active is None else active.astype(np.int64)
    assert rel_l2(wave_sr[:, rows], lane_sr[:, rows]) < 1e-12
    assert rel_l2(wave_acc, lane_acc) < 1e-12
    ref, _, _ = O.short_range(x, y, z, m, o, mesh)
    assert rel_l2(wave_sr[:, rows], np.stack(ref)[:, rows]) < 1e-12
    assert rel_l2(lane_sr[:, rows], np.stack(ref)[:, rows]) < 1e-12


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp64_box_walk_matches_counted_walk(case):
    """The timed FP64 walk (box classification, Chebyshev H table, v_rsq_f64
    with Newton steps) against the counting FP64 walk (the reference's
    arithmetic: erf/exp residual, exact minimum image): same pairs and
    multipoles, so they agree to 1e-12 relative L2 and 1e-10 per particle."""
    _, (x, y, z, m), o, mesh = case
    o.precision = L.PRECISION_FP64
    o.target_mapping = L.TARGETS_PER_LANE
    o.short_range_only = 1
    dx, dy, dz, dm = to_dev(x, y, z, m)
    lanes = []
    for diagnostics in (True, False):
        acc = tuple(empty_dev(x.size, torch.float64) for _ in range(3))
        sr = tuple(torch.zeros(x.size, dtype=torch.float64, device="cuda:0") for _ in range(3))
        torch.cuda.synchronize()
        with TreePmCoordinator(mesh) as coord:
            coord.solve_active_set_with_pm_cadence(dx, dy, dz, dm, ForceAccumulator(acc, None, None, sr), o, False,
                                                   want_diagnostics=diagnostics)
            coord.synchronize()
        lanes.append(np.stack([s_.cpu().numpy() for s_ in sr]))
    o.short_range_only = 0
    o.target_mapping = L.TARGETS_AUTO
    o.precision = L.PRECISION_MIXED
    counted, boxed = lanes
    assert rel_l2(boxed, counted) < 1e-12
    per = np.linalg.norm(boxed - counted, axis=0) / np.maximum(np.linalg.norm(counted, axis=0), 1e-300)
    assert per.max() < 1e-10


@pytest.mark.parametrize("precision", [L.PRECISION_FP64, L.PRECISION_MIXED])
@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_per_wave_tiny_sets(precision, n):
    """Trees of one leaf or a few nodes: the per-wave walk's window and chain
    handle a node count below one wave."""
    rng = np.random.default_rng(40 + n)
    x, y, z = rng.uniform(0.0, 10.0, (3, n))
    x[0], y[0], z[0] = 9.99, 0.01, 5.0  # a pair across the periodic seam when n > 1
    if n > 1:
        x[1], y[1], z[1] = 0.02, 9.98, 5.1
    m = rng.uniform(0.5, 1.5, n)
    o = options_struct(10.0, eps=0.02)
    wave_acc, wave_sr = solve_mixed(x, y, z, m, o, 16, L.TARGETS_PER_WAVE, precision=precision)
    lane_acc, lane_sr = solve_mixed(x, y, z, m, o, 16, L.TARGETS_PER_LANE, precision=precision)
    tol = 1e-12 if precision == L.PRECISION_FP64 else 1e-6
    assert np.all(np.isfinite(wave_sr)) and rel_l2(wave_sr, lane_sr) <= tol
    assert rel_l2(wave_acc, lane_acc) <= (1e-12 if precision == L.PRECISION_FP64 else 1e-5)
