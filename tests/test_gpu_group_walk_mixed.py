"""GPU checks of the adaptive policy in mixed precision.

The mixed group walk keeps the FP64 group walk's groups, lists and every
acceptance decision and evaluates only the listed-source pairs in FP32; the
per-target adaptive walk (fallback targets, or every target without the group
walk) runs in FP64 and gets the FP32 PM part afterwards. So against the
oracle:
  - every traversal counter equals the oracle's, except the in-cutoff pair
    count, where an FP32 separation may land on the other side of r_cut
    (allowed: 1e-4 of the pairs, at least 2);
  - the short-range accelerations agree to 1e-5 relative L2 and p99 1e-4
    (FP32 pair arithmetic; the FP64 lanes agree to 1e-12);
  - the PM lanes agree with the FP64 solve's to 1e-5 relative L2 (the FP32
    mesh), and each total equals its PM plus short-range lanes in FP32.
The cases are those of the FP64 group-walk tests: the reference's seam
clusters (tests/integration/test_tree_pm_coupling_periodic.cpp:2140-2230),
accepted-node-rich clustered sets with both cutoff node policies and both node
expansions, histories with gaps, active subsets, groups above 64 targets and
inputs outside the bounded metric.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import empty_dev, error_metrics, rel_l2, seam_cluster_options, seam_clusters, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - GPU-only module
    pytest.skip("no HIP device", allow_module_level=True)

from cosmosim import ForceAccumulator, TreePmCoordinator  # noqa: E402
from cosmosim import _lib as L  # noqa: E402
from test_oracle_pins import clustered  # noqa: E402

MESH = 32
EXACT = ("visited_nodes", "cutoff_pruned_nodes", "accepted_internal_multipoles", "accepted_leaves",
         "group_walk_groups", "group_walk_grouped_targets", "group_walk_target_group_size_max",
         "group_walk_fallback_targets", "group_walk_fallback_history", "group_walk_fallback_geometry",
         "group_walk_candidate_leaves", "group_walk_culled_leaves", "group_walk_candidate_particles",
         "group_walk_list_entries_max", "group_walk_status")


def solve(x, y, z, m, o, ad, precision, previous=None, active=None, nearest=False):
    o.precision = precision
    dt = torch.float32 if precision == L.PRECISION_MIXED else torch.float64
    dx, dy, dz, dm = to_dev(x, y, z, m)
    na = x.size if active is None else len(active)
    acc = [empty_dev(na, dt) for _ in range(3)]
    pm = [empty_dev(na, dt) for _ in range(3)]
    sr = [empty_dev(x.size, dt) for _ in range(3)]  # source-row indexed
    near = empty_dev(x.size, dt) if nearest else None
    act = to_dev(np.asarray(active, np.uint32).view(np.int32))[0] if active is not None else None
    prev = to_dev(np.asarray(previous, np.float64))[0] if previous is not None else None
    with TreePmCoordinator(MESH) as coord:
        res = coord.solve_active_set(dx, dy, dz, dm,
                                     ForceAccumulator(tuple(acc), act, tuple(pm), tuple(sr), prev, near),
                                     o, want_diagnostics=True, adaptive_options=ad)
        torch.cuda.synchronize()
    host = lambda ts: [t.cpu().numpy().astype(np.float64) for t in ts]  # noqa: E731
    rows = np.arange(x.size) if active is None else np.asarray(active)
    out = dict(acc=host(acc), pm=host(pm), sr=[a[rows] for a in host(sr)], c=res.diagnostics)
    if nearest:
        out["near"] = near.cpu().numpy().astype(np.float64)[rows]
    return out


def check_mixed(x, y, z, m, o, ad, previous=None, active=None):
    mix = solve(x, y, z, m, o, ad, L.PRECISION_MIXED, previous, active)
    f64 = solve(x, y, z, m, o, ad, L.PRECISION_FP64, previous, active)
    ref, rc = O.short_range_adaptive(x, y, z, m, o, ad, MESH, active=active, previous=previous)
    c = mix["c"]
    mismatched = {k: (c[k], rc[k]) for k in EXACT if c[k] != rc[k]}
    assert not mismatched, mismatched
    pairs, rpairs = c["direct_pair_evaluations"], rc["direct_pair_evaluations"]
    assert abs(pairs - rpairs) <= max(2, 1e-4 * rpairs), (pairs, rpairs)
    err = error_metrics(mix["sr"], ref)
    print(f"mixed short-range vs oracle: {err}; pairs {pairs} (oracle {rpairs})", flush=True)
    assert err["relative_l2"] < 1e-5 and err["p99"] < 1e-4, err
    assert rel_l2(f64["sr"], ref) < 1e-12
    assert rel_l2(mix["pm"], f64["pm"]) < 1e-5
    for a, p, s in zip(mix["acc"], mix["pm"], mix["sr"]):
        assert np.allclose(a, p + s, rtol=0, atol=4e-7 * np.abs(a).max())
    return mix, err


@pytest.fixture(scope="module")
def seam():
    return seam_clusters()


@pytest.fixture(scope="module")
def dense():
    x, y, z, m = clustered(3000, 8, 1.0)
    prev = np.random.default_rng(2).uniform(2500.0, 10000.0, x.size)
    return x, y, z, m, prev


@pytest.mark.parametrize("size", [0, 32])
def test_seam_clusters_mixed_group_walk(seam, size):
    x, y, z, m = seam
    o, ad = seam_cluster_options(group=True, group_size=size)
    mix, _ = check_mixed(x, y, z, m, o, ad, previous=np.ones(x.size))
    assert mix["c"]["group_walk_status"] == 1 and mix["c"]["group_walk_grouped_targets"] == x.size


@pytest.mark.parametrize("com_cut", [False, True])
@pytest.mark.parametrize("radial", [False, True])
def test_accepting_mixed_group_walk(dense, com_cut, radial):
    x, y, z, m, prev = dense
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True, com_cut=com_cut, radial=radial)
    o.max_leaf_size = 4
    prev = prev.copy()
    prev[::97] = np.nan  # a few groups without history fall back to the FP64 per-target walk
    mix, _ = check_mixed(x, y, z, m, o, ad, previous=prev)
    c = mix["c"]
    assert c["accepted_internal_multipoles"] > 0 and c["group_walk_fallback_history"] > 0


@pytest.mark.parametrize("estimator", [L.ESTIMATOR_RESIDUAL_DERIVATIVE, L.ESTIMATOR_RELATIVE_MONOPOLE])
def test_mixed_per_target_adaptive_walk(dense, estimator):
    x, y, z, m, prev = dense
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=False, estimator=estimator, lookup=True)
    o.max_leaf_size = 4
    prev = prev.copy()
    prev[::50] = np.nan
    mix, _ = check_mixed(x, y, z, m, o, ad, previous=prev)
    assert mix["c"]["accepted_internal_multipoles"] > 0


def test_mixed_missing_history_and_active_subset(seam, dense):
    x, y, z, m = seam
    o, ad = seam_cluster_options(group=True, group_size=32)
    mix, _ = check_mixed(x, y, z, m, o, ad, previous=None)
    assert mix["c"]["group_walk_fallback_history"] == x.size and mix["c"]["group_walk_groups"] == 0
    x, y, z, m, prev = dense
    active = np.sort(np.random.default_rng(11).choice(x.size, 1100, replace=False))[::-1].copy()
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True, group_size=16)
    o.residual_block_size = 32
    check_mixed(x, y, z, m, o, ad, previous=prev[active], active=active)


def test_mixed_groups_above_one_wave_and_unbounded_inputs(dense):
    x, y, z, m, prev = dense
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True, group_size=0)
    o.max_leaf_size = 200
    o.residual_block_size = 256
    mix, _ = check_mixed(x, y, z, m, o, ad, previous=prev)
    assert mix["c"]["group_walk_target_group_size_max"] > 64
    x = x.copy()
    x[:40] += 1.0
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True)
    mix, _ = check_mixed(x, y, z, m, o, ad, previous=prev)
    assert mix["c"]["group_walk_fallback_geometry"] > 0


def test_mixed_nearest_pair_lane(dense):
    # the group walk's FP32 separations and the FP64 fallback walk both fill it
    x, y, z, m, prev = dense
    o, ad = seam_cluster_options(max_angle=1.0, alpha=0.01, group=True)
    o.max_leaf_size = 4
    prev = prev.copy()
    prev[::97] = np.nan
    mix = solve(x, y, z, m, o, ad, L.PRECISION_MIXED, prev, nearest=True)
    f64 = solve(x, y, z, m, o, ad, L.PRECISION_FP64, prev, nearest=True)
    finite = np.isfinite(f64["near"])
    assert np.array_equal(finite, np.isfinite(mix["near"]))
    assert np.allclose(mix["near"][finite], f64["near"][finite], rtol=1e-5, atol=0)
